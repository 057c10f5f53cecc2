"""Pins the oracle (oracle/oracle.c) against closed forms, dense linear algebra
and an independent GMRES, and checks the library's host setup against the
oracle.  CPU only."""
import numpy as np
import pytest

import pargemslr_amd as pg
from oracle import oracle as orc
from tests import problems as P


def _nbr_count(nx, ny, nz):
    cnt = np.zeros(nx * ny * nz)
    for k in range(nz):
        for i in range(ny):
            for j in range(nx):
                c = (j > 0) + (j < nx - 1) + (i > 0) + (i < ny - 1) + (k > 0) + (k < nz - 1)
                cnt[k * nx * ny + i * nx + j] = c
    return cnt


@pytest.mark.parametrize("grid,shift", [((7, 5, 1), 0.0), ((5, 4, 3), 0.3)])
def test_matvec_laplacian_row_sums(grid, shift):
    """A*1 = diag - #neighbours for the Dirichlet Laplacian (closed form)."""
    A = pg.laplacian(*grid, shift=shift)
    diag = (4.0 if 1 in grid else 6.0) - shift
    y = orc.csr_matvec(A.ia, A.ja, A.a, np.ones(A.n))
    assert np.array_equal(y, diag - _nbr_count(*grid))


def test_matvec_alpha_beta_vs_scipy():
    A = pg.convdiff(6, 5, 4, 0.3)
    rng = np.random.default_rng(1)
    x, y0 = rng.standard_normal(A.n), rng.standard_normal(A.n)
    y = orc.csr_matvec(A.ia, A.ja, A.a, x, alpha=-0.7, beta=1.3, y=y0)
    ref = -0.7 * (A.to_scipy() @ x) + 1.3 * y0
    assert np.allclose(y, ref, rtol=1e-14, atol=1e-13)


def test_convdiff_generator_stencil():
    """Upwind convection-diffusion: diag 6+3c, upstream -1-c, downstream -1."""
    c = 0.3
    A = P.dense(pg.convdiff(4, 4, 4, c))
    n = 4
    idx = lambda x, y, z: z * n * n + y * n + x
    r = idx(1, 2, 1)
    assert A[r, r] == pytest.approx(6 + 3 * c)
    assert A[r, idx(0, 2, 1)] == pytest.approx(-1 - c)
    assert A[r, idx(2, 2, 1)] == pytest.approx(-1)
    assert A[r, idx(1, 1, 1)] == pytest.approx(-1 - c)
    assert A[r, idx(1, 2, 0)] == pytest.approx(-1 - c)
    assert A[r, idx(1, 2, 2)] == pytest.approx(-1)
    assert np.count_nonzero(A[r]) == 7


def _exact_single_level(A):
    """nlev=1 with no dropping: GeMSLR is one exact LU of P A P^T."""
    return P.setup(A, 1, 2, 0.0, A.n, 0)


def test_ilu_exact_special_case():
    A = pg.laplacian(6, 5, 1, alphas=(0.2, -0.1, 0.0))
    G = _exact_single_level(A)
    lv = G.levels()[0]
    perm = G.perm()
    Ad = P.dense(A)[np.ix_(perm, perm)]
    n = A.n
    Li, Lj, La = lv["L"]
    Ui, Uj, Ua = lv["U"]
    L = np.eye(n)
    U = np.zeros((n, n))
    for i in range(n):
        L[i, Lj[Li[i]:Li[i + 1]]] = La[Li[i]:Li[i + 1]]
        U[i, Uj[Ui[i]:Ui[i + 1]]] = Ua[Ui[i]:Ui[i + 1]]
    D = np.diag(1.0 / lv["Dinv"])
    # factor convention of IluClass: A = L (D + U)
    assert np.allclose(L @ (D + U), Ad, atol=1e-12)
    rhs = np.random.default_rng(2).standard_normal(n)
    x = orc.ilu_solve(lv["L"], lv["U"], lv["Dinv"], rhs)
    assert np.allclose(x, np.linalg.solve(Ad, rhs), rtol=1e-12, atol=1e-12)


def test_ilu_residual_invariant_with_dropping():
    A = pg.laplacian(9, 9, 1)
    G = P.setup(A, 1, 2, 1e-2, 4, 0)
    lv = G.levels()[0]
    n = A.n
    Li, Lj, La = lv["L"]
    Ui, Uj, Ua = lv["U"]
    L = np.eye(n)
    U = np.zeros((n, n))
    for i in range(n):
        L[i, Lj[Li[i]:Li[i + 1]]] = La[Li[i]:Li[i + 1]]
        U[i, Uj[Ui[i]:Ui[i + 1]]] = Ua[Ui[i]:Ui[i + 1]]
    rhs = np.random.default_rng(3).standard_normal(n)
    x = orc.ilu_solve(lv["L"], lv["U"], lv["Dinv"], rhs)
    assert np.linalg.norm(L @ ((np.diag(1.0 / lv["Dinv"]) + U) @ x) - rhs) <= 1e-13 * np.linalg.norm(rhs)
    # fill limit: at most 4 entries per row of L and of U
    assert np.diff(Li).max() <= 4 and np.diff(Ui).max() <= 4


def test_lowrank_apply_vs_numpy():
    rng = np.random.default_rng(4)
    W = rng.standard_normal((50, 7))
    WH = rng.standard_normal((50, 7))
    z = rng.standard_normal(50)
    assert np.allclose(orc.lowrank_apply(W, WH, z), WH @ (W.T @ z), rtol=1e-13, atol=1e-13)


def _blocks(A, G):
    perm = G.perm()
    Ad = P.dense(A)[np.ix_(perm, perm)]
    return perm, Ad


def test_gemslr_two_level_block_formula():
    """k = 0, exact block factors: z_u = B^-1 b_u, x_l = C^-1 (b_l - E z_u),
    x_u = B^-1 (b_u - F x_l) (gemslr.cpp:3474-3625) with dense B, C, E, F."""
    A = pg.laplacian(12, 12, 1, alphas=(0.15, -0.25, 0.0))  # nonsymmetric: E != F^T
    G = P.setup(A, 2, 4, 0.0, A.n, 0)
    lv = G.levels()
    perm, Ad = _blocks(A, G)
    e = lv[0]["nloc"]
    B, F, E, Cm = Ad[:e, :e], Ad[:e, e:], Ad[e:, :e], Ad[e:, e:]
    b = np.random.default_rng(5).standard_normal(A.n)
    bp = b[perm]
    zu = np.linalg.solve(B, bp[:e])
    xl = np.linalg.solve(Cm, bp[e:] - E @ zu)
    xu = np.linalg.solve(B, bp[:e] - F @ xl)
    x = np.empty(A.n)
    x[perm] = np.concatenate([xu, xl])
    O = P.oracle_of(G)
    assert np.allclose(O.apply(b), x, rtol=1e-11, atol=1e-11)
    # and it is not A^{-1} (C is used, not the Schur complement)
    assert not np.allclose(x, np.linalg.solve(P.dense(A), b), rtol=1e-6)


@pytest.mark.parametrize("nlev", [2, 3])
def test_gemslr_full_rank_is_inverse(nlev):
    """Exact ILU + full-rank low-rank correction on every level: M^{-1} = A^{-1}."""
    A = pg.laplacian(10, 10, 1, alphas=(0.1, -0.2, 0.0))
    G = P.setup(A, nlev, 4, 0.0, A.n, A.n, lr_tol_eig=1e-8)
    lv = G.levels()
    for l in range(nlev - 1):
        assert lv[l]["rank"] == lv[l]["nrem"], (l, lv[l]["rank"], lv[l]["nrem"])
    b = np.random.default_rng(6).standard_normal(A.n)
    z = P.oracle_of(G).apply(b)
    assert np.allclose(z, np.linalg.solve(P.dense(A), b), rtol=1e-9, atol=1e-9)


def test_gemslr_dropping_low_rank_term_breaks_inverse():
    """Plausible mistake planted in the oracle's input: zero the low-rank term
    and the full-rank identity M^{-1} = A^{-1} must fail."""
    A = pg.laplacian(10, 10, 1, alphas=(0.1, -0.2, 0.0))
    G = P.setup(A, 2, 4, 0.0, A.n, A.n, lr_tol_eig=1e-8)
    lv = G.levels()
    lv[0]["WH"] = np.zeros_like(lv[0]["WH"])
    from oracle import oracle as o2
    z = o2.GemslrOracle(A.n, G.perm(), lv).apply(np.ones(A.n))
    assert not np.allclose(z, np.linalg.solve(P.dense(A), np.ones(A.n)), rtol=1e-6)


@pytest.mark.parametrize("cgs2", [False, True])
def test_orthogonalization(cgs2):
    rng = np.random.default_rng(7)
    n, k = 40, 6
    Q, _ = np.linalg.qr(rng.standard_normal((n, k + 1)))
    V = np.zeros((n, k + 2), order="F")
    V[:, :k + 1] = Q
    w = rng.standard_normal(n)
    V[:, k + 1] = w
    H = np.zeros((k + 3, k + 2), order="F")
    t = orc.mgs(V, k, H, cgs2=cgs2)
    assert np.allclose(V.T @ V, np.eye(k + 2), atol=1e-13)
    assert np.allclose(V[:, :k + 2] @ H[:k + 2, k], w, atol=1e-13)
    assert t == pytest.approx(np.linalg.norm(w - Q @ (Q.T @ w)), rel=1e-12)


def _numpy_gmres_history(A, b, steps):
    """Independent GMRES: Arnoldi (two-pass Gram-Schmidt) + least squares per step."""
    n = len(b)
    beta = np.linalg.norm(b)
    V = np.zeros((n, steps + 1))
    H = np.zeros((steps + 1, steps))
    V[:, 0] = b / beta
    hist = [1.0]
    for j in range(steps):
        w = A @ V[:, j]
        for _ in range(2):
            h = V[:, :j + 1].T @ w
            w = w - V[:, :j + 1] @ h
            H[:j + 1, j] += h
        H[j + 1, j] = np.linalg.norm(w)
        V[:, j + 1] = w / H[j + 1, j]
        e1 = np.zeros(j + 2)
        e1[0] = beta
        y = np.linalg.lstsq(H[:j + 2, :j + 1], e1, rcond=None)[0]
        hist.append(np.linalg.norm(e1 - H[:j + 2, :j + 1] @ y) / beta)
    return np.array(hist)


@pytest.mark.parametrize("orth", [0, 1])
def test_fgmres_unpreconditioned_matches_gmres(orth):
    A = pg.convdiff(6, 6, 6, 0.3)
    b = pg.rng_normal(3, A.n)
    x, st, it, rr, hist = orc.fgmres(A, b, prec=None, kdim=200, maxits=200, tol=1e-10, orth=orth)
    assert st == 0
    ref = _numpy_gmres_history(A.to_scipy(), b, it)
    k = min(it, 30)
    assert np.allclose(hist[:k + 1], ref[:k + 1], rtol=1e-7, atol=1e-14)
    # no restart: the Givens estimate equals the true residual
    true = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
    assert true == pytest.approx(rr, rel=1e-4, abs=1e-13)


def test_fgmres_exact_preconditioner_one_iteration():
    A = pg.laplacian(8, 8, 1, alphas=(0.1, 0.0, 0.0))
    G = _exact_single_level(A)
    b = pg.rng_uniform(9, A.n)
    x, st, it, rr, _ = orc.fgmres(A, b, prec=P.oracle_of(G), kdim=10, maxits=10, tol=1e-10)
    assert st == 0 and it == 1
    assert np.allclose(x, np.linalg.solve(P.dense(A), b), rtol=1e-10)


def test_fgmres_zero_rhs():
    A = pg.laplacian(5, 5, 1)
    x, st, it, rr, _ = orc.fgmres(A, np.zeros(A.n), x0=np.ones(A.n), kdim=5, maxits=5)
    assert st == 0 and it == 0 and not x.any()


def test_fgmres_restart_keeps_givens_estimate():
    """With restarts the true residual ends at or below the tolerance."""
    A, G = P.build(P.C1)
    xs = pg.rng_uniform(0, A.n)
    b = A.matvec(xs)
    x, st, it, rr, hist = orc.fgmres(A, b, prec=P.oracle_of(G), kdim=30, maxits=300, tol=1e-8)
    assert st == 0 and it > 30  # at least one restart with FGMRES(30)
    assert np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b) <= 1e-8
    assert np.all(np.diff(hist[1:31]) <= 1e-14)  # monotone within a cycle


def test_golden_gr_30_30():
    """Reference test matrix gr_30_30 with its test right-hand side."""
    A = P.csr_from_scipy(P.read_mtx("gr_30_30.mtx"))
    b = np.asarray(P.read_mtx("gr_30_30_b.mtx").toarray()).ravel()
    G = P.setup(A, 3, 2, 1e-3, 200, 32)
    x, st, it, rr, _ = orc.fgmres(A, b, prec=P.oracle_of(G), kdim=50, maxits=200, tol=1e-8)
    assert st == 0
    assert np.linalg.norm(b - A.matvec(x)) <= 1e-8 * np.linalg.norm(b)
    import scipy.sparse.linalg as spl
    xe = spl.spsolve(A.to_scipy().tocsc(), b)
    assert np.linalg.norm(x - xe) <= 1e-6 * np.linalg.norm(xe)


def test_golden_test_small_exact():
    A = P.csr_from_scipy(P.read_mtx("test_small.mtx"))
    G = _exact_single_level(A)
    b = np.arange(1.0, A.n + 1)
    z = P.oracle_of(G).apply(b)
    assert np.allclose(z, np.linalg.solve(P.dense(A), b), rtol=1e-13)


# ---------------------------------------------------------------- host setup vs oracle

def test_setup_structure():
    A, G = P.build(P.C2_SMALL, alphas=(0.05, -0.02, 0.01))
    perm = G.perm()
    assert np.array_equal(np.sort(perm), np.arange(A.n))
    lv = G.levels()
    Ad = A.to_scipy()[perm][:, perm].tocsr()
    for l, d in enumerate(lv):
        assert d["lev_start"] + d["nloc"] + d["nrem"] == A.n
        assert d["dom_ptr"][0] == 0 and d["dom_ptr"][-1] == d["nloc"]
        if l < len(lv) - 1:
            s, e = d["lev_start"], d["lev_start"] + d["nloc"]
            Ei, Ej, Ea = d["E"]
            E = np.zeros((d["nrem"], d["nloc"]))
            for i in range(d["nrem"]):
                E[i, Ej[Ei[i]:Ei[i + 1]]] = Ea[Ei[i]:Ei[i + 1]]
            assert np.array_equal(E, Ad[e:, s:e].toarray())
            Fi, Fj, Fa = d["F"]
            F = np.zeros((d["nloc"], d["nrem"]))
            for i in range(d["nloc"]):
                F[i, Fj[Fi[i]:Fi[i + 1]]] = Fa[Fi[i]:Fi[i + 1]]
            assert np.array_equal(F, Ad[s:e, e:].toarray())
            # diagonal blocks are decoupled: no entries between different blocks of a level
            B = Ad[s:e, s:e].tocoo()
            blk = np.searchsorted(d["dom_ptr"], np.arange(d["nloc"]), side="right") - 1
            assert np.all(blk[B.row] == blk[B.col])
    assert lv[-1]["ncomp"] == 1


def test_host_apply_matches_oracle():
    A, G = P.build(P.C2_SMALL, alphas=(0.05, -0.02, 0.01))
    b = pg.rng_normal(11, A.n)
    z_host = G.host_apply(b)
    z_orc = P.oracle_of(G).apply(b)
    assert np.linalg.norm(z_host - z_orc) <= 1e-13 * np.linalg.norm(z_orc)


def test_low_rank_is_invariant_subspace():
    """The setup's W spans an approximate invariant subspace of G = I - S C~^{-1}:
    W is orthonormal and the correction (I-H)^{-1} - I is finite."""
    A, G = P.build(P.C2_SMALL)
    for d in G.levels()[:-1]:
        if d["rank"] == 0:
            continue
        W = d["W"]
        assert np.allclose(W.T @ W, np.eye(d["rank"]), atol=1e-10)
        assert np.all(np.isfinite(d["WH"]))


def test_setup_error_codes():
    A = pg.laplacian(6, 6, 1)
    p = pg.config_params(2, 4, 1e-2, 10, 0)
    p[pg.capi.IO["SCHUR_ENABLE"]] = 1
    p[pg.capi.IO["SCHUR_MAXITS"]] = 0  # the inner iteration needs at least one step
    with pytest.raises(pg.PgsError) as e:
        pg.Gemslr(A, p)
    assert e.value.code == 100
    bad = pg.CsrHost(A.n, A.ia, A.ja.copy(), A.a)
    bad.ja[3] = A.n + 5
    with pytest.raises(pg.PgsError) as e:
        pg.Gemslr(bad, pg.config_params(2, 4, 1e-2, 10, 0))
    assert e.value.code == 101
    # an all-zero row: ILUT reports PARGEMSLR_ERROR_ILU_EMPTY_ROW
    a = A.a.copy()
    a[A.ia[7]:A.ia[8]] = 0.0
    with pytest.raises(pg.PgsError) as e:
        pg.Gemslr(pg.CsrHost(A.n, A.ia, A.ja, a), pg.config_params(1, 2, 1e-2, 10, 0))
    assert e.value.code == 103


def test_setup_deterministic():
    A = pg.laplacian(20, 20, 1)
    p = pg.config_params(3, 4, 1e-3, 20, 8)
    G1, G2 = pg.Gemslr(A, p, seed=3), pg.Gemslr(A, p, seed=3)
    assert np.array_equal(G1.perm(), G2.perm())
    b = np.ones(A.n)
    assert np.array_equal(G1.host_apply(b), G2.host_apply(b))


def test_threaded_block_solves_give_the_same_bits():
    """The CPU baseline's OpenMP-over-blocks variant (set_threads > 1) solves
    each diagonal block as the serial oracle does: the same bits."""
    from tests import problems as P
    A, G = P.build(P.C2_SMALL)
    O = orc.GemslrOracle(A.n, G.perm(), G.levels())
    v = pg.rng_normal(1, A.n)
    z1 = O.apply(v)
    try:
        orc.set_threads(4)
        z4 = O.apply(v)
    finally:
        orc.set_threads(1)
    assert np.array_equal(z1, z4)
