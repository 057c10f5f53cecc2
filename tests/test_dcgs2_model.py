"""The delayed CGS2 of the device solve keeps raw vectors in the basis slots
and carries the final basis as V_final = V_raw R (R upper triangular per
column, built by k_dcgs2_fix); the update then writes only the new vector.
This checks the algebra of that form on the CPU: a numpy restatement of the
device's step (raw dots -> final coefficients through R, R's new column, the
update's coefficients on the raw slots, the lazy scale of the newest vector)
against the explicit form that overwrites each slot with its final vector,
and both against the Arnoldi properties they must satisfy: the final basis
orthonormal and A Z = V H (flexible GMRES, fgmres.hpp:300-540).  Real and
complex operators."""
import numpy as np
import pytest


def _arnoldi(A, M, b, m, implicit):
    n = A.shape[0]
    dt = A.dtype
    V = np.zeros((m + 2, n), dtype=dt)
    Z = np.zeros((m + 1, n), dtype=dt)
    H = np.zeros((m + 1, m), dtype=dt)
    R = np.zeros((m + 1, m + 1), dtype=dt)
    beta = np.linalg.norm(b)
    V[0] = b / beta
    sg = 1.0  # pending scale of slot p
    hist = []
    for p in range(m):
        Z[p] = M @ (sg * V[p])
        w = A @ Z[p]
        q = V[p].copy()
        ds = V[:p].conj() @ q  # raw dots, as k_dcgs2_dots
        da = V[:p].conj() @ w
        qq = sg * sg * np.vdot(q, q).real
        qw = sg * np.vdot(q, w)
        if implicit:
            Rp = R[:p, :p]
            s = sg * (Rp.conj().T @ ds)
            a = Rp.conj().T @ da
        else:
            s, a = sg * ds, da
        nu = 1.0
        if p > 0:
            nu = np.sqrt(max(qq - np.sum(np.abs(s) ** 2), 0.0))
            hq = H[p, p - 1]
            H[:p, p - 1] += hq * s
            H[p, p - 1] = hq * nu
        inv = 1.0 / nu
        cp = (qw - np.vdot(s, a)) * inv
        H[:p, p] = a
        H[p, p] = cp
        gm = cp * inv
        if implicit:
            Rp = R[:p, :p]
            R[:p, p] = -(Rp @ s) * inv
            R[p, p] = sg * inv
            e = np.concatenate([Rp @ (a - gm * s), [sg * gm]])
            un = w - e @ V[:p + 1]
        else:
            g = a - gm * s
            V[p] = (q - (s / sg) @ V[:p]) * (sg * inv)
            un = w - g @ V[:p] - gm * sg * q
        V[p + 1] = un
        t = np.linalg.norm(un)
        H[p + 1, p] = t
        sg = 1.0 / t
        if p > 0:
            e1 = np.zeros(p + 1, dtype=dt)
            e1[0] = beta
            y = np.linalg.lstsq(H[:p + 1, :p], e1, rcond=None)[0]
            hist.append(np.linalg.norm(e1 - H[:p + 1, :p] @ y) / beta)
    if implicit:
        Vf = (R[:m, :m].T @ V[:m])  # row j = sum_i R[i, j] slot_i
    else:
        Vf = V[:m]
    return np.array(hist), Vf, Z[:m], H


def _problem(cx, n=240, seed=0):
    rng = np.random.default_rng(seed)
    A = np.eye(n) * 4 + rng.standard_normal((n, n)) * 0.3
    if cx:
        A = A + 1j * rng.standard_normal((n, n)) * 0.1
    M = np.linalg.inv(A + rng.standard_normal((n, n)) * 0.1)  # an inexact preconditioner
    b = rng.standard_normal(n) + (1j * rng.standard_normal(n) if cx else 0)
    return A, M, b


@pytest.mark.parametrize("cx", [False, True], ids=["real", "complex"])
def test_raw_slots_with_triangular_factor_match_explicit_form(cx):
    A, M, b = _problem(cx)
    m = 24
    h_exp, Vf_exp, Z_exp, H_exp = _arnoldi(A, M, b, m, implicit=False)
    h_imp, Vf_imp, Z_imp, H_imp = _arnoldi(A, M, b, m, implicit=True)
    assert 1e-12 < h_exp[-1] < 0.05 * h_exp[0]  # the iteration makes progress, no breakdown
    assert np.all(np.abs(h_imp - h_exp) <= 1e-10 * h_exp)
    assert np.linalg.norm(Vf_imp - Vf_exp) < 1e-10
    for Vf, Z, H in ((Vf_exp, Z_exp, H_exp), (Vf_imp, Z_imp, H_imp)):
        assert np.linalg.norm(Vf.conj() @ Vf.T - np.eye(m)) < 1e-12
        # A Z = V H on the finalized columns 0..m-2 (column m-1 waits for
        # the closing pass)
        k = m - 1
        lhs = A @ Z[:k].T
        rhs = Vf[:k + 1].T @ H[:k + 1, :k]
        assert np.linalg.norm(lhs - rhs) < 1e-10 * np.linalg.norm(lhs)


def test_triangular_factor_with_a_wrong_column_fails():
    """A planted mistake in R's new column (its diagonal without the newest
    vector's pending scale) leaves a basis that is not orthonormal (or
    breaks the run down), so the checks above would catch it."""
    A, M, b = _problem(False, seed=3)
    m = 16
    _, Vf, _, _ = _arnoldi(A, M, b, m, implicit=True)
    assert np.linalg.norm(Vf @ Vf.T - np.eye(m)) < 1e-12
    # the same Arnoldi run with R[p, p] = 1 / nu
    n = A.shape[0]
    V = np.zeros((m + 2, n))
    R = np.zeros((m + 1, m + 1))
    V[0] = b / np.linalg.norm(b)
    sg = 1.0
    with np.errstate(all="ignore"):
        for p in range(m):
            w = A @ (M @ (sg * V[p]))
            q = V[p].copy()
            Rp = R[:p, :p]
            s = sg * (Rp.T @ (V[:p] @ q))
            a = Rp.T @ (V[:p] @ w)
            nu = np.sqrt(max(sg * sg * (q @ q) - s @ s, 0.0)) if p else 1.0
            inv = 1.0 / nu
            gm = (sg * (q @ w) - s @ a) * inv * inv
            R[:p, p] = -(Rp @ s) * inv
            R[p, p] = inv
            e = np.concatenate([Rp @ (a - gm * s), [sg * gm]])
            V[p + 1] = w - e @ V[:p + 1]
            sg = 1.0 / np.linalg.norm(V[p + 1])
        Vbad = R[:m, :m].T @ V[:m]
        err = np.linalg.norm(Vbad @ Vbad.T - np.eye(m))
    assert not (np.isfinite(err) and err < 1e-6)
