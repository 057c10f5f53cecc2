"""Parity of the device inner Schur iteration (PARGEMSLR_IO_SCHUR_ENABLE:
inner FGMRES on the level-0 Schur complement in every apply) with the
oracle, through the C ABI."""
import numpy as np
import pytest

import pargemslr_amd as pg
from oracle import oracle as orc
from tests import problems as P

pytestmark = pytest.mark.gpu


def _dev(x, dev):
    dt = np.complex128 if np.iscomplexobj(x) else np.float64
    return pg.DeviceArray.from_numpy(np.ascontiguousarray(x, dtype=dt), device=dev)


def _host(t):
    pg.synchronize(t.device)
    return t.numpy()


def _rhs(A, seed):
    if A.is_complex:
        return pg.rng_normal(seed, 2 * A.n).view(np.complex128)
    return pg.rng_normal(seed, A.n)


def _op(cplx, schur):
    A = pg.laplacian(24, 24, 24, shift=(0.2 + 0.02j) if cplx else 0.3, alphas=(0.05, -0.02, 0.01), complex_=cplx)
    G = P.setup(A, 3, 8, 1e-3, 20, 16, schur=schur)
    return A, G


@pytest.fixture(scope="module")
def real5(gpu):
    """Shifted (indefinite) 24^3 Laplacian, the reference's default inner
    iteration: FGMRES(5) to 1e-2."""
    A, G = _op(False, (5, 1e-2))
    D = pg.DevOp(G)
    return A, G, D, pg.Workspace(D, max_nrhs=8, kdim=50)


@pytest.mark.parametrize("nr", [1, 8])
def test_schur_apply_fp64_parity(nr, real5, gpu):
    A, G, D, W = real5
    O = P.oracle_of(G)
    V = np.stack([_rhs(A, 10 + c) for c in range(nr)], axis=1)
    Zd = _dev(np.zeros((A.n, nr)), gpu)
    W.apply(_dev(V, gpu), Zd, nrhs=nr, prec=pg.PGS_PREC_FP64)
    Z = _host(Zd)
    for c in range(nr):
        ref = O.apply(V[:, c])
        assert np.linalg.norm(Z[:, c] - ref) <= 1e-12 * np.linalg.norm(ref), c


@pytest.mark.parametrize("cplx", [False, True])
def test_schur_apply_fixed_steps(cplx, gpu):
    """Tolerance 0 (always K = 3 inner steps, so the result is a smooth
    function of the data): complex128 / FP64 within 1e-12 of the oracle,
    complex64 / FP32 within 1e-4."""
    A, G = _op(cplx, (3, 0.0))
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=4, kdim=2)
    O = P.oracle_of(G)
    V = np.stack([_rhs(A, 20 + c) for c in range(4)], axis=1)
    refs = [O.apply(V[:, c]) for c in range(4)]
    for prec, tol in ((pg.PGS_PREC_FP64, 1e-12), (pg.PGS_PREC_FP32, 1e-4)):
        Zd = _dev(np.zeros((A.n, 4), dtype=V.dtype), gpu)
        W.apply(_dev(V, gpu), Zd, nrhs=4, prec=prec)
        Z = _host(Zd)
        for c in range(4):
            assert np.linalg.norm(Z[:, c] - refs[c]) <= tol * np.linalg.norm(refs[c]), (prec, c)


def test_schur_apply_host_agrees(real5, gpu):
    A, G, D, W = real5
    v = _rhs(A, 30)
    zd = _dev(np.zeros(A.n), gpu)
    W.apply(_dev(v, gpu), zd, prec=pg.PGS_PREC_FP64)
    ref = G.host_apply(v)
    assert np.linalg.norm(_host(zd) - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("cplx", [False, True])
def test_schur_fgmres_fp64_parity(cplx, gpu):
    """Outer FGMRES with the inner iteration in every apply: iterations
    within one of the oracle's, true residual at the tolerance."""
    A, G = _op(cplx, (5, 1e-2))
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=1, kdim=50)
    b = _rhs(A, 40)
    xo, st, it_o, _, _ = orc.fgmres(A, b, prec=P.oracle_of(G), kdim=50, maxits=500, tol=1e-8)
    assert st == 0
    xd = _dev(np.zeros(A.n, dtype=b.dtype), gpu)
    (status, it, rr), = W.fgmres(_dev(b, gpu), xd, kdim=50, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP64,
                                 orth=pg.PGS_ORTH_MGS)
    x = _host(xd)
    assert status == 0 and abs(it - it_o) <= 1, (it, it_o)
    assert np.linalg.norm(b - A.matvec(x)) <= 1e-8 * np.linalg.norm(b)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


def test_schur_fgmres_mixed_batched(real5, gpu):
    """FP32 apply (FP32 inner vectors, double inner scalars) inside FP64
    FGMRES, 4 right-hand sides solved together: every column reaches the
    tolerance and equals its single solve."""
    A, G, D, W = real5
    nr = 4
    B = np.stack([_rhs(A, 50 + c) for c in range(nr)], axis=1)
    Xd = _dev(np.zeros((A.n, nr)), gpu)
    res = W.fgmres(_dev(B, gpu), Xd, nrhs=nr, kdim=50, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP32)
    X = _host(Xd)
    for c in range(nr):
        st, it, _ = res[c]
        assert st == 0
        assert np.linalg.norm(B[:, c] - A.matvec(X[:, c])) <= 1e-8 * np.linalg.norm(B[:, c])
        x1 = _dev(np.zeros(A.n), gpu)
        (st1, it1, _), = W.fgmres(_dev(B[:, c], gpu), x1, kdim=50, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP32)
        assert st1 == 0 and abs(it1 - it) <= 1
        assert np.linalg.norm(X[:, c] - _host(x1)) <= 1e-6 * np.linalg.norm(X[:, c])


def test_schur_graph_replay_matches_direct(real5, gpu):
    """The inner iteration launches nothing that waits for the host, so the
    captured inner FGMRES steps replay to the same bits."""
    A, G, D, W = real5
    nr = 2
    B = np.stack([_rhs(A, 60 + c) for c in range(nr)], axis=1)
    outs = []
    for use_graph in (False, True, True):
        Xd = _dev(np.zeros((A.n, nr)), gpu)
        res = W.fgmres(_dev(B, gpu), Xd, nrhs=nr, kdim=10, maxits=200, tol=1e-8, prec=pg.PGS_PREC_FP32,
                       use_graph=use_graph)
        outs.append((_host(Xd), [r[:2] for r in res]))
    for X, res in outs[1:]:
        assert res == outs[0][1]
        assert np.array_equal(X, outs[0][0])
