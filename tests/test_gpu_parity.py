"""Parity of the HIP solve path (through the C ABI) with the oracle."""
import numpy as np
import pytest

import pargemslr_amd as pg
from oracle import oracle as orc
from tests import problems as P

pytestmark = pytest.mark.gpu


def _dev(x, dev):
    return pg.DeviceArray.from_numpy(np.ascontiguousarray(x, dtype=np.float64), device=dev)


def _host(t):
    pg.synchronize(t.device)
    return t.numpy()


@pytest.fixture(scope="module")
def small(gpu):
    A, G = P.build(P.C2_SMALL, alphas=(0.05, -0.02, 0.01))
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=8, kdim=50)
    return A, G, D, W


@pytest.fixture(scope="module")
def c1(gpu):
    A, G = P.build(P.C1)
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=4, kdim=30)
    return A, G, D, W


def test_spmv_parity(small, gpu):
    A, G, D, W = small
    x = pg.rng_normal(1, A.n)
    y0 = pg.rng_normal(2, A.n)
    ref = orc.csr_matvec(A.ia, A.ja, A.a, x, alpha=-0.5, beta=2.0, y=y0)
    xd, yd = _dev(x, gpu), _dev(y0, gpu)
    W.spmv(xd, yd, alpha=-0.5, beta=2.0)
    y = _host(yd)
    assert np.linalg.norm(y - ref) <= 1e-14 * np.linalg.norm(ref)


@pytest.mark.parametrize("which", ["c1", "small"])
def test_apply_fp64_parity(which, request, gpu):
    A, G, D, W = request.getfixturevalue(which)
    O = P.oracle_of(G)
    for seed in (3, 4):
        v = pg.rng_normal(seed, A.n)
        ref = O.apply(v)
        zd = _dev(np.zeros(A.n), gpu)
        W.apply(_dev(v, gpu), zd, prec=pg.PGS_PREC_FP64)
        z = _host(zd)
        assert np.linalg.norm(z - ref) <= 1e-12 * np.linalg.norm(ref), (which, seed)


def test_apply_fp32_close(small, gpu):
    A, G, D, W = small
    v = pg.rng_normal(5, A.n)
    ref = P.oracle_of(G).apply(v)
    zd = _dev(np.zeros(A.n), gpu)
    W.apply(_dev(v, gpu), zd, prec=pg.PGS_PREC_FP32)
    z = _host(zd)
    err = np.linalg.norm(z - ref) / np.linalg.norm(ref)
    print(f"[parity] FP32 apply vs the oracle's FP64 apply: {err:.2e}", flush=True)
    # the mixed form's bar for the apply: 1e-5 relative to the FP64 apply
    assert err <= 1e-5


def test_apply_multi_rhs_matches_columns(small, gpu):
    A, G, D, W = small
    nr = 8
    V = np.stack([pg.rng_normal(10 + c, A.n) for c in range(nr)], axis=1)
    Zd = _dev(np.zeros((A.n, nr)), gpu)
    W.apply(_dev(V, gpu), Zd, nrhs=nr, prec=pg.PGS_PREC_FP64)
    Z = _host(Zd)
    O = P.oracle_of(G)
    for c in range(nr):
        ref = O.apply(V[:, c])
        assert np.linalg.norm(Z[:, c] - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("which,orth", [("c1", pg.PGS_ORTH_MGS), ("c1", pg.PGS_ORTH_CGS2),
                                        ("small", pg.PGS_ORTH_CGS2), ("c1", pg.PGS_ORTH_DCGS2),
                                        ("small", pg.PGS_ORTH_DCGS2)])
def test_fgmres_fp64_parity(which, orth, request, gpu):
    A, G, D, W = request.getfixturevalue(which)
    kdim = 30 if which == "c1" else 50
    xs = pg.rng_uniform(0, A.n)
    b = A.matvec(xs)
    xo, st, it_o, rr_o, _ = orc.fgmres(A, b, prec=P.oracle_of(G), kdim=kdim, maxits=500, tol=1e-8, orth=0)
    assert st == 0
    xd = _dev(np.zeros(A.n), gpu)
    (status, it, rr), = W.fgmres(_dev(b, gpu), xd, kdim=kdim, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP64, orth=orth)
    x = _host(xd)
    assert status == 0
    assert abs(it - it_o) <= 1, (it, it_o)
    true = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
    assert true <= 1e-8
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


def test_fgmres_mixed_precision(small, gpu):
    A, G, D, W = small
    b = pg.rng_normal(21, A.n)
    xo, st, it_o, _, _ = orc.fgmres(A, b, prec=P.oracle_of(G), kdim=50, maxits=500, tol=1e-10)
    xd = _dev(np.zeros(A.n), gpu)
    (status, it, rr), = W.fgmres(_dev(b, gpu), xd, kdim=50, maxits=500, tol=1e-10, prec=pg.PGS_PREC_FP32)
    x = _host(xd)
    assert status == 0
    assert np.linalg.norm(b - A.matvec(x)) <= 1e-10 * np.linalg.norm(b)
    assert np.linalg.norm(x - xo) <= 1e-7 * np.linalg.norm(xo)
    assert it <= it_o + 3


def test_fgmres_multi_rhs_columns_match_single(small, gpu):
    A, G, D, W = small
    nr = 4
    B = np.stack([pg.rng_normal(30 + c, A.n) for c in range(nr)], axis=1)
    B[:, 2] = 0.0  # zero column: x = 0, no iterations
    Xd = _dev(np.zeros((A.n, nr)), gpu)
    res = W.fgmres(_dev(B, gpu), Xd, nrhs=nr, kdim=50, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP32)
    X = _host(Xd)
    for c in range(nr):
        st, it, rr = res[c]
        assert st == 0
        if c == 2:
            assert it == 0 and not X[:, c].any()
            continue
        x1 = _dev(np.zeros(A.n), gpu)
        (st1, it1, _), = W.fgmres(_dev(B[:, c], gpu), x1, kdim=50, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP32)
        assert abs(it1 - it) <= 1
        xs1 = _host(x1)
        assert np.linalg.norm(B[:, c] - A.matvec(X[:, c])) <= 1e-8 * np.linalg.norm(B[:, c])
        assert np.linalg.norm(X[:, c] - xs1) <= 1e-6 * np.linalg.norm(xs1)


def test_fgmres_unpreconditioned_matches_oracle(c1, gpu):
    A, G, D, W = c1
    b = pg.rng_normal(41, A.n)
    xo, st, it_o, rr_o, _ = orc.fgmres(A, b, prec=None, kdim=30, maxits=60, tol=1e-8, orth=0)
    xd = _dev(np.zeros(A.n), gpu)
    (status, it, rr), = W.fgmres(_dev(b, gpu), xd, kdim=30, maxits=60, tol=1e-8, use_precond=False,
                                 orth=pg.PGS_ORTH_MGS)
    assert it == it_o == 60 and status == pg.capi.PGS_ERR_NOT_CONVERGED
    assert rr == pytest.approx(rr_o, rel=1e-6)


@pytest.mark.parametrize("env", [{}, {"PGS_TRSV_CG": "1"}, {"PGS_TRSV_CG": "2"},
                                 {"PGS_RING_GEOM": "64x4"}, {"PGS_RING_GEOM": "64x4", "PGS_TRSV_CG": "1"},
                                 {"PGS_RING_GEOM": "64x2", "PGS_TRSV_CG": "2"}, {"PGS_RING_GEOM": "64x2", "PGS_TRSV_CG": "1"},
                                 {"PGS_RING_GEOM": "64x2", "PGS_TRSV_CG": "4"},
                                 {"PGS_TRSV_NT": "256"}, {"PGS_TRSV_VARIANT": "1"}, {"PGS_TRSV_GROUPS": "1"},
                                 {"PGS_TRSV_GROUPS": "2"}, {"PGS_TRSV_CG": "8", "PGS_RING_GEOM": "128x4"},
                                 {"PGS_TRSV_CG": "8", "PGS_RING_GEOM": "64x4"}],
                         ids=["ring4", "ring1", "ring2", "wave4x4", "wave4x1", "wave2x2", "wave2x1", "wave2x4", "pipelined",
                              "direct", "ring4-one-group", "ring4-paired", "ring8", "wave4x8"])
def test_apply_trsv_kernels(env, small, gpu, monkeypatch):
    """Every triangular-solve kernel (ring with 128 threads and 1, 2 or 4
    columns per group, two 4-column groups per workgroup for eight columns
    when the paired grid fills the device, else one; 8 columns per group for
    FP32 values, with 128 threads or one wave; single-wave ring on half
    steps with 4 lanes per row and 1 or 4 columns, or whole steps with 2
    lanes per row and 1, 2 or 4 columns;
    the pipelined kernel; the direct kernel) gives the oracle's apply, for one
    and for eight right-hand sides."""
    A, G, _, _ = small
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=8, kdim=2)
    O = P.oracle_of(G)
    nr = 8
    V = np.stack([pg.rng_normal(60 + c, A.n) for c in range(nr)], axis=1)
    refs = [O.apply(V[:, c]) for c in range(nr)]
    for prec, tol in ((pg.PGS_PREC_FP64, 1e-12), (pg.PGS_PREC_FP32, 1e-5)):
        Zd = _dev(np.zeros((A.n, nr)), gpu)
        W.apply(_dev(V, gpu), Zd, nrhs=nr, prec=prec)
        Z = _host(Zd)
        for c in range(nr):
            assert np.linalg.norm(Z[:, c] - refs[c]) <= tol * np.linalg.norm(refs[c]), (env, prec, c)
        zd = _dev(np.zeros(A.n), gpu)
        W.apply(_dev(V[:, 0], gpu), zd, prec=prec)
        assert np.linalg.norm(_host(zd) - refs[0]) <= tol * np.linalg.norm(refs[0]), (env, prec)


def test_apply_wide_rows(gpu):
    """Exact block factors (no dropping) have rows wider than the pipelined
    kernels take: the direct kernel solves them."""
    A = pg.laplacian(10, 10, 10)
    G = P.setup(A, 2, 4, 0.0, 10000, 0)
    assert max(int(np.diff(d["L"][0]).max()) for d in G.levels()) > 20
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=2, kdim=2)
    O = P.oracle_of(G)
    V = np.stack([pg.rng_normal(70 + c, A.n) for c in range(2)], axis=1)
    Zd = _dev(np.zeros((A.n, 2)), gpu)
    W.apply(_dev(V, gpu), Zd, nrhs=2, prec=pg.PGS_PREC_FP64)
    Z = _host(Zd)
    for c in range(2):
        ref = O.apply(V[:, c])
        assert np.linalg.norm(Z[:, c] - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("nr,orth,kdim", [(4, pg.PGS_ORTH_CGS2, 50), (64, pg.PGS_ORTH_CGS2, 50),
                                          (4, pg.PGS_ORTH_DCGS2, 50), (64, pg.PGS_ORTH_DCGS2, 50),
                                          (8, pg.PGS_ORTH_DCGS2, 80), (64, pg.PGS_ORTH_DCGS2, 80)])
def test_fgmres_cgs2_long_basis(nr, orth, kdim, c1, gpu):
    """Unpreconditioned FGMRES(kdim) with CGS2 or delayed CGS2: the basis
    grows past 32 and 64 vectors' worth of register sweeps (two lanes per
    element; at 64 columns a wave then covers half a row's columns); every
    column's Givens residual after 60 (or kdim + 10) iterations equals the
    oracle's (GMRES residuals are unique)."""
    A = c1[0]
    D = pg.DevOp(c1[1])
    W = pg.Workspace(D, max_nrhs=nr, kdim=kdim)
    its = 60 if kdim <= 50 else kdim + 10
    B = np.stack([pg.rng_normal(80 + c, A.n) for c in range(nr)], axis=1)
    Xd = _dev(np.zeros((A.n, nr)), gpu)
    res = W.fgmres(_dev(B, gpu), Xd, nrhs=nr, kdim=kdim, maxits=its, tol=1e-12, use_precond=False, orth=orth)
    for c in (0, 1, nr - 1):
        _, st, it_o, rr_o, _ = orc.fgmres(A, B[:, c], prec=None, kdim=kdim, maxits=its, tol=1e-12, orth=1)
        st_g, it_g, rr_g = res[c]
        assert it_g == it_o == its
        assert rr_g == pytest.approx(rr_o, rel=1e-6), c


@pytest.mark.parametrize("prec,orth", [(pg.PGS_PREC_FP64, pg.PGS_ORTH_CGS2), (pg.PGS_PREC_FP32, pg.PGS_ORTH_CGS2),
                                       (pg.PGS_PREC_FP32, pg.PGS_ORTH_DCGS2)])
def test_fgmres_graph_replay_matches_direct(prec, orth, small, gpu):
    """Inner steps captured as HIP graphs and replayed (FGMRES(10): every
    graph is replayed once per restart cycle, and the whole solve twice with
    direct-launch solves in between) give the same bits as direct launches."""
    A, G, D, W = small
    nr = 4
    B = np.stack([pg.rng_normal(90 + c, A.n) for c in range(nr)], axis=1)
    outs = []
    for use_graph in (False, True, False, True):
        Xd = _dev(np.zeros((A.n, nr)), gpu)
        res = W.fgmres(_dev(B, gpu), Xd, nrhs=nr, kdim=10, maxits=500, tol=1e-8, prec=prec, use_graph=use_graph,
                       orth=orth)
        outs.append((_host(Xd), [r[:2] for r in res]))
    for X, res in outs[1:]:
        assert res == outs[0][1]
        assert np.array_equal(X, outs[0][0])
    assert all(st == 0 for st, _ in outs[0][1]) and max(it for _, it in outs[0][1]) > 10


def test_fgmres_dcgs2_is_its_own_orthogonalization(small, gpu):
    """Delayed CGS2 runs its own sweeps: iteration count within one of CGS2's,
    solutions within the tolerance of each other but not the same bits (a
    silent fallback to CGS2 would give identical solutions)."""
    A, G, D, W = small
    b = pg.rng_normal(44, A.n)
    out = {}
    for orth in (pg.PGS_ORTH_CGS2, pg.PGS_ORTH_DCGS2):
        xd = _dev(np.zeros(A.n), gpu)
        (st, it, _), = W.fgmres(_dev(b, gpu), xd, kdim=50, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP64, orth=orth)
        assert st == 0
        out[orth] = (_host(xd), it)
    (x1, it1), (x2, it2) = out[pg.PGS_ORTH_CGS2], out[pg.PGS_ORTH_DCGS2]
    assert abs(it1 - it2) <= 1
    assert not np.array_equal(x1, x2)
    assert np.linalg.norm(x1 - x2) <= 1e-6 * np.linalg.norm(x1)


def test_fgmres_dcgs2_batched_columns_match_single(small, gpu):
    """Delayed CGS2, mixed form, 4 right-hand sides (one zero): every column
    reaches the tolerance and matches its single solve; iterations within one
    of the oracle's (modified Gram-Schmidt)."""
    A, G, D, W = small
    nr = 4
    B = np.stack([pg.rng_normal(100 + c, A.n) for c in range(nr)], axis=1)
    B[:, 3] = 0.0
    Xd = _dev(np.zeros((A.n, nr)), gpu)
    res = W.fgmres(_dev(B, gpu), Xd, nrhs=nr, kdim=50, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP32,
                   orth=pg.PGS_ORTH_DCGS2)
    X = _host(Xd)
    O = P.oracle_of(G)
    for c in range(nr):
        st, it, rr = res[c]
        assert st == 0
        if c == 3:
            assert it == 0 and not X[:, c].any()
            continue
        assert np.linalg.norm(B[:, c] - A.matvec(X[:, c])) <= 1e-8 * np.linalg.norm(B[:, c])
        _, _, it_o, _, _ = orc.fgmres(A, B[:, c], prec=O, kdim=50, maxits=500, tol=1e-8, orth=0)
        assert abs(it - it_o) <= 1, (c, it, it_o)
        x1 = _dev(np.zeros(A.n), gpu)
        (st1, it1, _), = W.fgmres(_dev(B[:, c], gpu), x1, kdim=50, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP32,
                                  orth=pg.PGS_ORTH_DCGS2)
        assert st1 == 0 and abs(it1 - it) <= 1
        assert np.linalg.norm(X[:, c] - _host(x1)) <= 1e-6 * np.linalg.norm(X[:, c])


def test_fgmres_dcgs2_restarts(c1, gpu):
    """Delayed CGS2 with FGMRES(10): several restart cycles, each closed by
    the final dot pass; iterations within one of the oracle's."""
    A, G, D, W = c1
    b = pg.rng_normal(110, A.n)
    _, st, it_o, _, _ = orc.fgmres(A, b, prec=P.oracle_of(G), kdim=10, maxits=500, tol=1e-10, orth=0)
    xd = _dev(np.zeros(A.n), gpu)
    (status, it, rr), = W.fgmres(_dev(b, gpu), xd, kdim=10, maxits=500, tol=1e-10, prec=pg.PGS_PREC_FP64,
                                 orth=pg.PGS_ORTH_DCGS2)
    x = _host(xd)
    assert status == 0 and it > 20 and abs(it - it_o) <= 1, (it, it_o)
    assert np.linalg.norm(b - A.matvec(x)) <= 1e-10 * np.linalg.norm(b)
