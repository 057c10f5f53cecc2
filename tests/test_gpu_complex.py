"""Parity of the complex HIP path (complex128 operator and Krylov vectors;
complex128 or complex64 preconditioner apply) with the complex oracle,
through the C ABI."""
import numpy as np
import pytest

import pargemslr_amd as pg
from oracle import oracle as orc
from tests import problems as P

pytestmark = pytest.mark.gpu

SHIFT = 0.2 + 0.02j


def _crandn(seed, n):
    return pg.rng_normal(seed, 2 * n).view(np.complex128)


def _dev(x, dev):
    return pg.DeviceArray.from_numpy(np.ascontiguousarray(x, dtype=np.complex128), device=dev)


def _host(t):
    pg.synchronize(t.device)
    return t.numpy()


@pytest.fixture(scope="module")
def cplx(gpu):
    """Complex-shifted 3D Laplacian 24^3 with convection terms (nonsymmetric),
    3 levels, 8 blocks, rank 16."""
    A = pg.laplacian(24, 24, 24, shift=SHIFT, alphas=(0.05, -0.02, 0.01), complex_=True)
    G = P.setup(A, 3, 8, 1e-3, 20, 16)
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=8, kdim=50)
    return A, G, D, W


def test_complex_spmv_parity(cplx, gpu):
    A, G, D, W = cplx
    x, y0 = _crandn(1, A.n), _crandn(2, A.n)
    alpha, beta = -0.5 + 0.25j, 2.0 - 1.0j
    ref = alpha * (A.to_scipy() @ x) + beta * y0
    xd, yd = _dev(x, gpu), _dev(y0, gpu)
    W.spmv(xd, yd, alpha=alpha, beta=beta)
    y = _host(yd)
    assert np.linalg.norm(y - ref) <= 1e-14 * np.linalg.norm(ref)
    nr = 4
    X = np.stack([_crandn(3 + c, A.n) for c in range(nr)], axis=1)
    Yd = _dev(np.zeros((A.n, nr)), gpu)
    W.spmv(_dev(X, gpu), Yd, nrhs=nr)
    Y = _host(Yd)
    for c in range(nr):
        r = A.to_scipy() @ X[:, c]
        assert np.linalg.norm(Y[:, c] - r) <= 1e-14 * np.linalg.norm(r)


@pytest.mark.parametrize("nr", [1, 8])
def test_complex_apply_parity(nr, cplx, gpu):
    """complex128 apply within 1e-12 of the oracle; complex64 within 1e-4."""
    A, G, D, W = cplx
    O = P.oracle_of(G)
    V = np.stack([_crandn(10 + c, A.n) for c in range(nr)], axis=1)
    refs = [O.apply(V[:, c]) for c in range(nr)]
    for prec, tol in ((pg.PGS_PREC_FP64, 1e-12), (pg.PGS_PREC_FP32, 1e-4)):
        Zd = _dev(np.zeros((A.n, nr)), gpu)
        W.apply(_dev(V, gpu), Zd, nrhs=nr, prec=prec)
        Z = _host(Zd)
        for c in range(nr):
            assert np.linalg.norm(Z[:, c] - refs[c]) <= tol * np.linalg.norm(refs[c]), (prec, c)


@pytest.mark.parametrize("geom", ["64x4", "64x2"])
def test_complex_apply_single_wave_ring(geom, cplx, gpu, monkeypatch):
    """The single-wave ring geometries on complex values (4 right-hand sides)."""
    A, G, _, _ = cplx
    monkeypatch.setenv("PGS_RING_GEOM", geom)
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=4, kdim=2)
    O = P.oracle_of(G)
    V = np.stack([_crandn(90 + c, A.n) for c in range(4)], axis=1)
    for prec, tol in ((pg.PGS_PREC_FP64, 1e-12), (pg.PGS_PREC_FP32, 1e-4)):
        Zd = _dev(np.zeros((A.n, 4)), gpu)
        W.apply(_dev(V, gpu), Zd, nrhs=4, prec=prec)
        Z = _host(Zd)
        for c in range(4):
            ref = O.apply(V[:, c])
            assert np.linalg.norm(Z[:, c] - ref) <= tol * np.linalg.norm(ref), (geom, prec, c)


def test_complex_apply_host_agrees(cplx, gpu):
    """The library's host apply and the device apply agree (the same factors
    run on the CPU and on the GPU)."""
    A, G, D, W = cplx
    v = _crandn(20, A.n)
    zd = _dev(np.zeros(A.n), gpu)
    W.apply(_dev(v, gpu), zd, prec=pg.PGS_PREC_FP64)
    ref = G.host_apply(v)
    assert np.linalg.norm(_host(zd) - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("orth", [pg.PGS_ORTH_MGS, pg.PGS_ORTH_CGS2, pg.PGS_ORTH_DCGS2])
def test_complex_fgmres_fp64_parity(orth, cplx, gpu):
    A, G, D, W = cplx
    b = _crandn(21, A.n)
    xo, st, it_o, rr_o, _ = orc.fgmres(A, b, prec=P.oracle_of(G), kdim=50, maxits=500, tol=1e-8, orth=0)
    assert st == 0
    xd = _dev(np.zeros(A.n), gpu)
    (status, it, rr), = W.fgmres(_dev(b, gpu), xd, kdim=50, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP64, orth=orth)
    x = _host(xd)
    assert status == 0
    assert abs(it - it_o) <= 1, (it, it_o)
    assert np.linalg.norm(b - A.matvec(x)) <= 1e-8 * np.linalg.norm(b)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


def test_complex_fgmres_mixed_batched(cplx, gpu):
    """complex64 preconditioner inside complex128 FGMRES, 4 right-hand sides
    (one zero): every column reaches the tolerance and equals its single solve."""
    A, G, D, W = cplx
    nr = 4
    B = np.stack([_crandn(30 + c, A.n) for c in range(nr)], axis=1)
    B[:, 1] = 0.0
    Xd = _dev(np.zeros((A.n, nr)), gpu)
    res = W.fgmres(_dev(B, gpu), Xd, nrhs=nr, kdim=50, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP32)
    X = _host(Xd)
    _, _, it_o, _, _ = orc.fgmres(A, B[:, 0], prec=P.oracle_of(G), kdim=50, maxits=500, tol=1e-8)
    for c in range(nr):
        st, it, rr = res[c]
        assert st == 0
        if c == 1:
            assert it == 0 and not X[:, c].any()
            continue
        assert np.linalg.norm(B[:, c] - A.matvec(X[:, c])) <= 1e-8 * np.linalg.norm(B[:, c])
        x1 = _dev(np.zeros(A.n), gpu)
        (st1, it1, _), = W.fgmres(_dev(B[:, c], gpu), x1, kdim=50, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP32)
        assert st1 == 0 and abs(it1 - it) <= 1
        assert np.linalg.norm(X[:, c] - _host(x1)) <= 1e-6 * np.linalg.norm(X[:, c])
    assert res[0][1] <= it_o + 3


def test_complex_fgmres_unpreconditioned_matches_oracle(cplx, gpu):
    """No preconditioner, 40 iterations: the Givens residual estimate equals
    the oracle's (GMRES residuals are unique; conj(c) in the rotations)."""
    A, G, D, W = cplx
    b = _crandn(41, A.n)
    _, st, it_o, rr_o, _ = orc.fgmres(A, b, prec=None, kdim=20, maxits=40, tol=1e-14, orth=0)
    xd = _dev(np.zeros(A.n), gpu)
    (status, it, rr), = W.fgmres(_dev(b, gpu), xd, kdim=20, maxits=40, tol=1e-14, use_precond=False,
                                 orth=pg.PGS_ORTH_CGS2)
    assert it == it_o == 40
    assert rr == pytest.approx(rr_o, rel=1e-6)


@pytest.mark.parametrize("kdim", [20, 40, 64])
def test_complex_dcgs2_unpreconditioned_matches_oracle(kdim, cplx, gpu):
    """Delayed CGS2 on a complex basis (conjugated dots, the lagged
    normalization and Hessenberg correction, complex Givens): without a
    preconditioner the Givens residual estimates equal the oracle's MGS ones
    over 2 cycles (GMRES residuals are unique), for bases that take the
    one-lane and two-lane update kernels."""
    A, G, D, W = cplx
    b = _crandn(42, A.n)
    W2 = pg.Workspace(D, max_nrhs=2, kdim=kdim)
    _, st, it_o, rr_o, _ = orc.fgmres(A, b, prec=None, kdim=kdim, maxits=2 * kdim, tol=1e-14, orth=0)
    xd = _dev(np.zeros(A.n), gpu)
    (status, it, rr), = W2.fgmres(_dev(b, gpu), xd, kdim=kdim, maxits=2 * kdim, tol=1e-14, use_precond=False,
                                  orth=pg.PGS_ORTH_DCGS2)
    assert it == it_o == 2 * kdim
    assert rr == pytest.approx(rr_o, rel=1e-6)


def test_complex_dcgs2_is_its_own_orthogonalization(cplx, gpu):
    """Delayed CGS2 on a complex basis runs its own sweeps: same iteration
    count as CGS2 (within one) and the same tolerance, but not the same bits
    (a silent fallback to CGS2 would give identical solutions)."""
    A, G, D, W = cplx
    b = _crandn(43, A.n)
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


def test_complex_dcgs2_batched_equals_single(cplx, gpu):
    """Two complex columns with delayed CGS2 and the complex64 apply: each
    column reaches the tolerance and equals its single solve."""
    A, G, D, W = cplx
    nr = 2
    B = np.stack([_crandn(60 + c, A.n) for c in range(nr)], axis=1)
    Xd = _dev(np.zeros((A.n, nr)), gpu)
    res = W.fgmres(_dev(B, gpu), Xd, nrhs=nr, kdim=30, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP32,
                   orth=pg.PGS_ORTH_DCGS2)
    X = _host(Xd)
    for c in range(nr):
        st, it, _ = res[c]
        assert st == 0
        assert np.linalg.norm(B[:, c] - A.matvec(X[:, c])) <= 1e-8 * np.linalg.norm(B[:, c])
        x1 = _dev(np.zeros(A.n), gpu)
        (st1, it1, _), = W.fgmres(_dev(B[:, c], gpu), x1, kdim=30, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP32,
                                  orth=pg.PGS_ORTH_DCGS2)
        assert st1 == 0 and abs(it1 - it) <= 1
        assert np.linalg.norm(X[:, c] - _host(x1)) <= 1e-6 * np.linalg.norm(X[:, c])


@pytest.mark.parametrize("orth", [pg.PGS_ORTH_CGS2, pg.PGS_ORTH_DCGS2])
def test_complex_graph_replay_matches_direct(orth, cplx, gpu):
    A, G, D, W = cplx
    nr = 2
    B = np.stack([_crandn(50 + c, A.n) for c in range(nr)], axis=1)
    outs = []
    for use_graph in (False, True, True):
        Xd = _dev(np.zeros((A.n, nr)), gpu)
        res = W.fgmres(_dev(B, gpu), Xd, nrhs=nr, kdim=10, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP32,
                       use_graph=use_graph, orth=orth)
        outs.append((_host(Xd), [r[:2] for r in res]))
    for X, res in outs[1:]:
        assert res == outs[0][1]
        assert np.array_equal(X, outs[0][0])


def test_golden_hel20_on_device(gpu):
    """Reference complex Helmholtz matrix hel20 with b = A * ones: complex128
    and mixed solves reach the tolerance with x = ones, iterations within one
    of the oracle's."""
    A = P.csr_from_scipy(P.read_mtx("hel20.mtx.gz"))
    b = A.matvec(np.ones(A.n, np.complex128))
    G = P.setup(A, 3, 4, 1e-3, 50, 32)
    _, st, it_o, _, _ = orc.fgmres(A, b, prec=P.oracle_of(G), kdim=50, maxits=200, tol=1e-8)
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=1, kdim=50)
    for prec in (pg.PGS_PREC_FP64, pg.PGS_PREC_FP32):
        xd = _dev(np.zeros(A.n), gpu)
        (status, it, rr), = W.fgmres(_dev(b, gpu), xd, kdim=50, maxits=200, tol=1e-8, prec=prec)
        x = _host(xd)
        assert status == 0
        assert np.linalg.norm(b - A.matvec(x)) <= 1e-8 * np.linalg.norm(b)
        assert np.abs(x - 1).max() <= 1e-6
        if prec == pg.PGS_PREC_FP64:
            assert abs(it - it_o) <= 1, (it, it_o)


def test_server_complex_operator(cplx, gpu):
    """A complex and a real operator in one server: every request equals the
    same right-hand sides solved alone."""
    A, G, D, W = cplx
    Ar = pg.laplacian(32, 32, 1)
    Dr = pg.DevOp(P.setup(Ar, 2, 4, 1e-2, 10, 0))
    S = pg.Server(max_batch=4, slots=2, mem_reserve=1 << 30, kdim=50, tol=1e-8)
    iz, ir = S.add_operator(D), S.add_operator(Dr)
    plan = [(iz, np.stack([_crandn(60 + c, A.n) for c in range(2)], axis=1)),
            (ir, np.stack([pg.rng_normal(70 + c, Ar.n) for c in range(3)], axis=1)),
            (iz, _crandn(62, A.n)[:, None]),
            (ir, pg.rng_normal(73, Ar.n)[:, None])]
    reqs = [S.submit(i, B) for i, B in plan]
    results = [S.wait(r) for r in reqs]
    S.close()
    for (i, B), (X, its, info) in zip(plan, results):
        Aop, Dop = (A, D) if i == iz else (Ar, Dr)
        assert info["status"] == 0
        assert X.dtype == (np.complex128 if i == iz else np.float64)
        W1 = pg.Workspace(Dop, max_nrhs=1, kdim=50)
        for c in range(B.shape[1]):
            b = B[:, c]
            assert np.linalg.norm(b - Aop.matvec(X[:, c])) <= 1e-8 * np.linalg.norm(b)
            dt = X.dtype
            x1 = pg.DeviceArray(Aop.n, dtype=dt)
            (st, it, _), = W1.fgmres(pg.DeviceArray.from_numpy(np.ascontiguousarray(b, dtype=dt)), x1, kdim=50,
                                     maxits=1000, tol=1e-8, prec=pg.PGS_PREC_FP32)
            pg.synchronize()
            assert st == 0 and abs(int(its[c]) - it) <= 1
            assert np.linalg.norm(X[:, c] - x1.numpy()) <= 1e-6 * np.linalg.norm(X[:, c])
