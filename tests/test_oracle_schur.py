"""Pins the oracle's inner Schur iteration (the solve phase of SolveLevel with
PARGEMSLR_IO_SCHUR_ENABLE, gemslr.cpp:3556-3563) against dense linear
algebra: exact block factors make S = C - E B^{-1} F exact, so K inner steps
are a dense right-preconditioned GMRES on S, and enough steps give A^{-1}.
Also checks the library's host apply against the oracle.  CPU only."""
import numpy as np
import pytest

import pargemslr_amd as pg
from oracle import oracle as orc
from tests import problems as P


def _csr_dense(t, shape):
    ia, ja, a = t
    M = np.zeros(shape, dtype=a.dtype if len(a) else np.float64)
    for i in range(shape[0]):
        M[i, ja[ia[i]:ia[i + 1]]] = a[ia[i]:ia[i + 1]]
    return M


def _blocks(A, G):
    perm = G.perm()
    Ad = P.dense(A)[np.ix_(perm, perm)]
    lv = G.levels()
    e = lv[0]["nloc"]
    return perm, Ad, lv, e


def _dense_gmres(S, M, b, steps, tol=0.0):
    """Right-preconditioned GMRES from x0 = 0 (MGS Arnoldi, least squares per
    step), stopping after `steps` steps or when the residual drops to
    tol * ||b||; returns (x, residual history)."""
    beta = np.linalg.norm(b)
    n = len(b)
    V = np.zeros((n, steps + 1), dtype=np.result_type(S, b))
    Z = np.zeros((n, steps), dtype=V.dtype)
    H = np.zeros((steps + 1, steps), dtype=V.dtype)
    V[:, 0] = b / beta
    hist = [beta]
    y = None
    k = 0
    for j in range(steps):
        Z[:, j] = M @ V[:, j]
        w = S @ Z[:, j]
        for i in range(j + 1):
            H[i, j] = np.vdot(V[:, i], w)
            w = w - H[i, j] * V[:, i]
        H[j + 1, j] = np.linalg.norm(w)
        V[:, j + 1] = w / H[j + 1, j]
        e1 = np.zeros(j + 2, dtype=V.dtype)
        e1[0] = beta
        y = np.linalg.lstsq(H[:j + 2, :j + 1], e1, rcond=None)[0]
        hist.append(np.linalg.norm(e1 - H[:j + 2, :j + 1] @ y))
        k = j + 1
        if hist[-1] <= tol * beta:
            break
    return Z[:, :k] @ y, np.array(hist)


def _expected_apply(Ad, e, v, perm, inner):
    """Dense GeMSLR apply with an exact level-0 block solve and x_l = inner(z_l)."""
    B, F, E = Ad[:e, :e], Ad[:e, e:], Ad[e:, :e]
    bp = v[perm]
    zu = np.linalg.solve(B, bp[:e])
    zl = bp[e:] - E @ zu
    xl = inner(zl)
    xu = np.linalg.solve(B, bp[:e] - F @ xl)
    x = np.empty_like(bp)
    x[perm] = np.concatenate([xu, xl])
    return x


def _schur_parts(Ad, e, lv0):
    B, F, E, Cm = Ad[:e, :e], Ad[:e, e:], Ad[e:, :e], Ad[e:, e:]
    S = Cm - E @ np.linalg.solve(B, F)
    return S, Cm


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("nlev", [2, 3])
def test_inner_iteration_enough_steps_is_inverse(nlev, cplx):
    """Exact ILU, no low-rank term: the inner FGMRES solves S x_l = z_l to
    1e-14, so the apply is A^{-1} (without it, the apply is not: C is used
    in place of S)."""
    A = pg.laplacian(10, 10, 1, shift=(0.2 + 0.02j) if cplx else 0.0, alphas=(0.1, -0.2, 0.0), complex_=cplx)
    m_guess = A.n
    G = P.setup(A, nlev, 4, 0.0, A.n, 0, schur=(min(32, m_guess), 1e-14))
    lv = G.levels()
    assert lv[0]["C"] is not None and lv[0]["nrem"] <= 32
    b = pg.rng_normal(6, 2 * A.n).view(np.complex128) if cplx else pg.rng_normal(6, A.n)
    z = P.oracle_of(G).apply(b)
    ref = np.linalg.solve(P.dense(A), b)
    assert np.linalg.norm(z - ref) <= 1e-10 * np.linalg.norm(ref)
    G0 = P.setup(A, nlev, 4, 0.0, A.n, 0)
    z0 = P.oracle_of(G0).apply(b)
    assert np.linalg.norm(z0 - ref) > 1e-6 * np.linalg.norm(ref)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("K", [1, 3])
def test_inner_steps_equal_dense_gmres(K, cplx):
    """Two levels, exact factors, rank-r low-rank term: K inner steps (no
    early stop) equal a dense GMRES on S = C - E B^{-1} F with the
    preconditioner C^{-1} (I + W Hk W^H)."""
    A = pg.laplacian(12, 12, 1, shift=(0.3 + 0.05j) if cplx else 0.3, alphas=(0.15, -0.1, 0.0), complex_=cplx)
    G = P.setup(A, 2, 4, 0.0, A.n, 4, lr_tol_eig=1e-8, schur=(K, 0.0))
    perm, Ad, lv, e = _blocks(A, G)
    S, Cm = _schur_parts(Ad, e, lv[0])
    W, WH = lv[0]["W"], lv[0]["WH"]
    assert lv[0]["rank"] > 0
    M = np.linalg.solve(Cm, np.eye(Cm.shape[0]) + WH @ W.conj().T)
    b = pg.rng_normal(8, 2 * A.n).view(np.complex128) if cplx else pg.rng_normal(8, A.n)
    ref = _expected_apply(Ad, e, b, perm, lambda zl: _dense_gmres(S, M, zl, K)[0])
    z = P.oracle_of(G).apply(b)
    assert np.linalg.norm(z - ref) <= 1e-11 * np.linalg.norm(ref)
    # a plausible mistake: the inner operator taken as C (no E B^{-1} F term)
    wrong = _expected_apply(Ad, e, b, perm, lambda zl: _dense_gmres(Cm, M, zl, K)[0])
    assert np.linalg.norm(z - wrong) > 1e-6 * np.linalg.norm(ref)


def test_inner_early_stop_at_tolerance():
    """The inner FGMRES stops at the first step whose Givens residual is at
    or below tol * ||z_l||: with tol between the step-1 and step-2 residuals
    the apply equals the 2-step GMRES result."""
    A = pg.laplacian(12, 12, 1, shift=0.3, alphas=(0.15, -0.1, 0.0))
    G5 = P.setup(A, 2, 4, 0.0, A.n, 0, schur=(5, 0.0))
    perm, Ad, lv, e = _blocks(A, G5)
    S, Cm = _schur_parts(Ad, e, lv[0])
    M = np.linalg.inv(Cm)
    b = pg.rng_normal(9, A.n)
    bp = b[perm]
    zl = bp[e:] - Ad[e:, :e] @ np.linalg.solve(Ad[:e, :e], bp[:e])
    _, hist = _dense_gmres(S, M, zl, 5)
    assert hist[1] > hist[2] > hist[3]
    tol = 0.5 * (hist[1] + hist[2]) / hist[0]
    G = P.setup(A, 2, 4, 0.0, A.n, 0, schur=(5, tol))
    ref = _expected_apply(Ad, e, b, perm, lambda z: _dense_gmres(S, M, z, 2)[0])
    z = P.oracle_of(G).apply(b)
    assert np.linalg.norm(z - ref) <= 1e-11 * np.linalg.norm(ref)
    ref3 = _expected_apply(Ad, e, b, perm, lambda z: _dense_gmres(S, M, z, 3)[0])
    assert np.linalg.norm(z - ref3) > 1e-8 * np.linalg.norm(ref)


def test_inner_zero_rhs_gives_zero():
    A = pg.laplacian(12, 12, 1)
    G = P.setup(A, 2, 4, 1e-2, 10, 0, schur=(5, 1e-2))
    z = P.oracle_of(G).apply(np.zeros(A.n))
    assert not z.any()


@pytest.mark.parametrize("cplx", [False, True])
def test_host_apply_matches_oracle_with_inner_iteration(cplx):
    A = pg.laplacian(16, 16, 16, shift=(0.2 + 0.02j) if cplx else 0.0, alphas=(0.05, -0.02, 0.01), complex_=cplx)
    G = P.setup(A, 3, 8, 1e-3, 20, 16, schur=(5, 1e-2))
    b = pg.rng_normal(11, 2 * A.n).view(np.complex128) if cplx else pg.rng_normal(11, A.n)
    zh = G.host_apply(b)
    zo = P.oracle_of(G).apply(b)
    assert np.linalg.norm(zh - zo) <= 1e-13 * np.linalg.norm(zo)


def test_outer_fgmres_with_inner_iteration():
    """The reference's default (inner FGMRES(5), 1e-2) on the shifted
    indefinite Laplacian: the outer FGMRES reaches the tolerance in fewer
    iterations than without the inner iteration."""
    A = pg.laplacian(20, 20, 20, shift=0.3)
    b = pg.rng_normal(12, A.n)
    its = {}
    for schur in (None, (5, 1e-2)):
        G = P.setup(A, 3, 8, 1e-3, 20, 16, schur=schur)
        x, st, it, rr, _ = orc.fgmres(A, b, prec=P.oracle_of(G), kdim=50, maxits=500, tol=1e-8)
        assert st == 0 and np.linalg.norm(b - A.matvec(x)) <= 1e-8 * np.linalg.norm(b)
        its[schur] = it
    assert its[(5, 1e-2)] < its[None]


def test_setup_rejects_bad_inner_settings():
    A = pg.laplacian(8, 8, 1)
    for schur in ((0, 1e-2), (33, 1e-2), (5, -1.0)):
        with pytest.raises(pg.PgsError) as e:
            P.setup(A, 2, 4, 1e-2, 10, 0, schur=schur)
        assert e.value.code == 100
