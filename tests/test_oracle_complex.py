"""Pins the complex (complex128) oracle functions the way tests/test_oracle.py
pins the real ones: closed forms, dense numpy linear algebra, an independent
complex GMRES, the reference's complex test matrices, and consistency with
the real oracle on a complex operator whose imaginary parts are zero.  CPU
only."""
import numpy as np
import pytest

import pargemslr_amd as pg
from oracle import oracle as orc
from tests import problems as P
from tests.test_oracle import _nbr_count

SHIFT = 0.2 + 0.02j  # the serving operator's shift: diag 5.8 - 0.02i


def _crandn(seed, n):
    return pg.rng_normal(seed, 2 * n).view(np.complex128)


def _lower_upper(lv, n):
    Li, Lj, La = lv["L"]
    Ui, Uj, Ua = lv["U"]
    L = np.eye(n, dtype=np.complex128)
    U = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        L[i, Lj[Li[i]:Li[i + 1]]] = La[Li[i]:Li[i + 1]]
        U[i, Uj[Ui[i]:Ui[i + 1]]] = Ua[Ui[i]:Ui[i + 1]]
    return L, U


def test_complex_generator_row_sums():
    """A*1 = (6 - shift) - #neighbours for the complex-shifted 3D Laplacian."""
    grid = (5, 4, 3)
    A = pg.laplacian(*grid, shift=SHIFT, complex_=True)
    assert A.a.dtype == np.complex128
    y = orc.csr_matvec(A.ia, A.ja, A.a, np.ones(A.n, np.complex128))
    assert np.allclose(y, (6.0 - SHIFT) - _nbr_count(*grid), rtol=0, atol=1e-15)


def test_complex_matvec_vs_scipy():
    A = pg.laplacian(6, 5, 4, shift=SHIFT, alphas=(0.1, -0.2, 0.05), complex_=True)
    x, y0 = _crandn(1, A.n), _crandn(2, A.n)
    y = orc.csr_matvec(A.ia, A.ja, A.a, x, alpha=-0.7, beta=1.3, y=y0)
    ref = -0.7 * (A.to_scipy() @ x) + 1.3 * y0
    assert np.allclose(y, ref, rtol=1e-14, atol=1e-13)


def test_complex_ilu_exact_special_case():
    A = pg.laplacian(6, 5, 1, shift=SHIFT, alphas=(0.2, -0.1, 0.0), complex_=True)
    G = P.setup(A, 1, 2, 0.0, A.n, 0)
    lv = G.levels()[0]
    assert lv["Dinv"].dtype == np.complex128
    perm = G.perm()
    Ad = P.dense(A)[np.ix_(perm, perm)]
    L, U = _lower_upper(lv, A.n)
    assert np.allclose(L @ (np.diag(1.0 / lv["Dinv"]) + U), Ad, atol=1e-12)
    rhs = _crandn(3, A.n)
    x = orc.ilu_solve(lv["L"], lv["U"], lv["Dinv"], rhs)
    assert np.allclose(x, np.linalg.solve(Ad, rhs), rtol=1e-12, atol=1e-12)


def test_complex_ilu_residual_invariant_with_dropping():
    A = pg.laplacian(9, 9, 1, shift=SHIFT, complex_=True)
    G = P.setup(A, 1, 2, 1e-2, 4, 0)
    lv = G.levels()[0]
    L, U = _lower_upper(lv, A.n)
    rhs = _crandn(4, A.n)
    x = orc.ilu_solve(lv["L"], lv["U"], lv["Dinv"], rhs)
    assert np.linalg.norm(L @ ((np.diag(1.0 / lv["Dinv"]) + U) @ x) - rhs) <= 1e-13 * np.linalg.norm(rhs)


def test_complex_lowrank_apply_conjugates_w():
    """x = WH (W^H z): the product with W is conjugated (a plain transpose fails)."""
    rng = np.random.default_rng(4)
    W = rng.standard_normal((50, 7)) + 1j * rng.standard_normal((50, 7))
    WH = rng.standard_normal((50, 7)) + 1j * rng.standard_normal((50, 7))
    z = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    x = orc.lowrank_apply(W, WH, z)
    assert np.allclose(x, WH @ (W.conj().T @ z), rtol=1e-13, atol=1e-13)
    assert not np.allclose(x, WH @ (W.T @ z), rtol=1e-3)


def test_complex_gemslr_two_level_block_formula():
    A = pg.laplacian(12, 12, 1, shift=SHIFT, alphas=(0.15, -0.25, 0.0), complex_=True)
    G = P.setup(A, 2, 4, 0.0, A.n, 0)
    lv = G.levels()
    perm = G.perm()
    Ad = P.dense(A)[np.ix_(perm, perm)]
    e = lv[0]["nloc"]
    B, F, E, Cm = Ad[:e, :e], Ad[:e, e:], Ad[e:, :e], Ad[e:, e:]
    b = _crandn(5, A.n)
    bp = b[perm]
    zu = np.linalg.solve(B, bp[:e])
    xl = np.linalg.solve(Cm, bp[e:] - E @ zu)
    xu = np.linalg.solve(B, bp[:e] - F @ xl)
    x = np.empty(A.n, np.complex128)
    x[perm] = np.concatenate([xu, xl])
    assert np.allclose(P.oracle_of(G).apply(b), x, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("nlev", [2, 3])
def test_complex_gemslr_full_rank_is_inverse(nlev):
    """Exact ILU + full-rank complex low-rank correction: M^{-1} = A^{-1}."""
    A = pg.laplacian(10, 10, 1, shift=SHIFT, alphas=(0.1, -0.2, 0.0), complex_=True)
    G = P.setup(A, nlev, 4, 0.0, A.n, A.n, lr_tol_eig=1e-8)
    lv = G.levels()
    for l in range(nlev - 1):
        assert lv[l]["rank"] == lv[l]["nrem"], (l, lv[l]["rank"], lv[l]["nrem"])
    b = _crandn(6, A.n)
    z = P.oracle_of(G).apply(b)
    assert np.allclose(z, np.linalg.solve(P.dense(A), b), rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("cgs2", [False, True])
def test_complex_orthogonalization(cgs2):
    rng = np.random.default_rng(7)
    n, k = 40, 6
    Q, _ = np.linalg.qr(rng.standard_normal((n, k + 1)) + 1j * rng.standard_normal((n, k + 1)))
    V = np.zeros((n, k + 2), dtype=np.complex128, order="F")
    V[:, :k + 1] = Q
    w = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    V[:, k + 1] = w
    H = np.zeros((k + 3, k + 2), dtype=np.complex128, order="F")
    t = orc.mgs(V, k, H, cgs2=cgs2)
    assert np.allclose(V.conj().T @ V, np.eye(k + 2), atol=1e-13)
    assert np.allclose(V[:, :k + 2] @ H[:k + 2, k], w, atol=1e-13)
    assert t == pytest.approx(np.linalg.norm(w - Q @ (Q.conj().T @ w)), rel=1e-12)


def _numpy_gmres_history(A, b, steps):
    """Independent complex GMRES: Arnoldi (two-pass Gram-Schmidt with V^H) and
    a least-squares solve per step."""
    n = len(b)
    beta = np.linalg.norm(b)
    V = np.zeros((n, steps + 1), np.complex128)
    H = np.zeros((steps + 1, steps), np.complex128)
    V[:, 0] = b / beta
    hist = [1.0]
    for j in range(steps):
        w = A @ V[:, j]
        for _ in range(2):
            h = V[:, :j + 1].conj().T @ w
            w = w - V[:, :j + 1] @ h
            H[:j + 1, j] += h
        H[j + 1, j] = np.linalg.norm(w)
        V[:, j + 1] = w / H[j + 1, j]
        e1 = np.zeros(j + 2, np.complex128)
        e1[0] = beta
        y = np.linalg.lstsq(H[:j + 2, :j + 1], e1, rcond=None)[0]
        hist.append(np.linalg.norm(e1 - H[:j + 2, :j + 1] @ y) / beta)
    return np.array(hist)


@pytest.mark.parametrize("orth", [0, 1])
def test_complex_fgmres_unpreconditioned_matches_gmres(orth):
    A = pg.laplacian(6, 6, 6, shift=SHIFT, alphas=(0.3, 0.1, -0.2), complex_=True)
    b = _crandn(3, A.n)
    x, st, it, rr, hist = orc.fgmres(A, b, prec=None, kdim=200, maxits=200, tol=1e-10, orth=orth)
    assert st == 0
    ref = _numpy_gmres_history(A.to_scipy(), b, it)
    k = min(it, 30)
    assert np.allclose(hist[:k + 1], ref[:k + 1], rtol=1e-7, atol=1e-14)
    true = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
    assert true == pytest.approx(rr, rel=1e-4, abs=1e-13)


def test_complex_fgmres_exact_preconditioner_one_iteration():
    A = pg.laplacian(8, 8, 1, shift=SHIFT, alphas=(0.1, 0.0, 0.0), complex_=True)
    G = P.setup(A, 1, 2, 0.0, A.n, 0)
    b = _crandn(9, A.n)
    x, st, it, rr, _ = orc.fgmres(A, b, prec=P.oracle_of(G), kdim=10, maxits=10, tol=1e-10)
    assert st == 0 and it == 1
    assert np.allclose(x, np.linalg.solve(P.dense(A), b), rtol=1e-10)


def test_complex_with_zero_imaginary_parts_equals_real():
    """A complex operator with zero imaginary parts and a real right-hand side
    follows the real path step for step (apply, residual history, solution)."""
    Ar = pg.laplacian(16, 16, 1, alphas=(0.1, -0.05, 0.0))
    Az = pg.CsrHost(Ar.n, Ar.ia, Ar.ja, Ar.a.astype(np.complex128))
    cfg = (2, 4, 1e-2, 10, 8)
    Gr, Gz = P.setup(Ar, *cfg), P.setup(Az, *cfg)
    b = pg.rng_uniform(2, Ar.n)
    zr = P.oracle_of(Gr).apply(b)
    zz = P.oracle_of(Gz).apply(b.astype(np.complex128))
    assert np.allclose(zz.real, zr, rtol=1e-12, atol=1e-13) and np.abs(zz.imag).max() <= 1e-12
    xr, _, itr, _, hr = orc.fgmres(Ar, b, prec=P.oracle_of(Gr), kdim=20, maxits=100, tol=1e-8)
    xz, _, itz, _, hz = orc.fgmres(Az, b.astype(np.complex128), prec=P.oracle_of(Gz), kdim=20, maxits=100, tol=1e-8)
    assert itr == itz
    assert np.allclose(hz, hr, rtol=1e-8, atol=1e-15)
    assert np.allclose(xz.real, xr, rtol=1e-9, atol=1e-12)


def test_complex_host_apply_matches_oracle():
    A = pg.laplacian(16, 16, 16, shift=SHIFT, complex_=True)
    G = P.setup(A, 3, 8, 1e-3, 20, 16)
    b = _crandn(11, A.n)
    z_host = G.host_apply(b)
    z_orc = P.oracle_of(G).apply(b)
    assert np.linalg.norm(z_host - z_orc) <= 1e-13 * np.linalg.norm(z_orc)


def test_golden_test_smallz_exact():
    """Reference complex test matrix test_smallz with its right-hand side."""
    A = P.csr_from_scipy(P.read_mtx("test_smallz.mtx"))
    assert A.is_complex
    b = np.asarray(P.read_mtx("test_smallz_rhs.mtx").toarray(), dtype=np.complex128).ravel()
    G = P.setup(A, 1, 2, 0.0, A.n, 0)
    z = P.oracle_of(G).apply(b)
    assert np.allclose(z, np.linalg.solve(P.dense(A), b), rtol=1e-13)


def test_golden_hel20():
    """Reference complex Helmholtz matrix hel20 (n = 8000) with the reference
    driver's default right-hand side b = A * ones: GeMSLR-FGMRES reaches the
    tolerance and the solution is ones."""
    A = P.csr_from_scipy(P.read_mtx("hel20.mtx.gz"))
    assert A.is_complex and A.n == 8000
    b = A.matvec(np.ones(A.n, np.complex128))
    G = P.setup(A, 3, 4, 1e-3, 50, 32)
    x, st, it, rr, _ = orc.fgmres(A, b, prec=P.oracle_of(G), kdim=50, maxits=200, tol=1e-8)
    assert st == 0
    assert np.linalg.norm(b - A.matvec(x)) <= 1e-8 * np.linalg.norm(b)
    assert np.abs(x - 1).max() <= 1e-6
