"""The shifted-indefinite and convection-diffusion workloads at sizes the
oracle finishes in seconds: the HIP path (through the C ABI) against the
oracle on the same operator."""
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
def shifted(gpu):
    """3D Laplacian 32^3 shifted by 0.1: indefinite (the unshifted operator's
    eigenvalues start at about 0.027; a few dozen fall below the shift),
    3 levels, 8 subdomains, rank 32."""
    A = pg.laplacian(32, 32, 32, shift=0.1)
    G = P.setup(A, 3, 8, 1e-3, 20, 32)
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=4, kdim=50)
    return A, G, D, W


@pytest.fixture(scope="module")
def convdiff(gpu):
    """Upwind convection-diffusion 24^3, c = 0.3 (nonsymmetric)."""
    A = pg.convdiff(24, 24, 24, 0.3)
    G = P.setup(A, 3, 8, 1e-3, 20, 16)
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=64, kdim=50)
    return A, G, D, W


def test_shifted_fp64_matches_oracle(shifted, gpu):
    A, G, D, W = shifted
    b = pg.rng_normal(0, A.n)
    O = P.oracle_of(G)
    v = pg.rng_normal(1, A.n)
    zd = _dev(np.zeros(A.n), gpu)
    W.apply(_dev(v, gpu), zd, prec=pg.PGS_PREC_FP64)
    ref = O.apply(v)
    assert np.linalg.norm(_host(zd) - ref) <= 1e-12 * np.linalg.norm(ref)
    xo, st, it_o, _, _ = orc.fgmres(A, b, prec=O, kdim=50, maxits=1000, tol=1e-10, orth=0)
    assert st == 0
    xd = _dev(np.zeros(A.n), gpu)
    (status, it, _), = W.fgmres(_dev(b, gpu), xd, kdim=50, maxits=1000, tol=1e-10, prec=pg.PGS_PREC_FP64,
                                orth=pg.PGS_ORTH_MGS)
    x = _host(xd)
    assert status == 0 and abs(it - it_o) <= 1, (it, it_o)
    assert np.linalg.norm(b - A.matvec(x)) <= 1e-10 * np.linalg.norm(b)


def test_shifted_mixed_vs_fp64_oracle(shifted, gpu):
    """FP32 apply, FP64 Krylov, tol 1e-10: true residual at tolerance and the
    error against the oracle's FP64 solution."""
    A, G, D, W = shifted
    nr = 4
    B = np.stack([pg.rng_normal(10 + c, A.n) for c in range(nr)], axis=1)
    Xd = _dev(np.zeros((A.n, nr)), gpu)
    res = W.fgmres(_dev(B, gpu), Xd, nrhs=nr, kdim=50, maxits=1000, tol=1e-10, prec=pg.PGS_PREC_FP32)
    X = _host(Xd)
    O = P.oracle_of(G)
    for c in range(nr):
        assert res[c][0] == 0
        assert np.linalg.norm(B[:, c] - A.matvec(X[:, c])) <= 1e-10 * np.linalg.norm(B[:, c])
    xo, st, it_o, _, _ = orc.fgmres(A, B[:, 0], prec=O, kdim=50, maxits=1000, tol=1e-10)
    assert st == 0
    assert np.linalg.norm(X[:, 0] - xo) <= 1e-7 * np.linalg.norm(xo)
    assert res[0][1] <= it_o + 3


def test_convdiff_64_rhs_block(convdiff, gpu):
    """64 right-hand sides solved as one block: every column converges, and
    sampled columns equal the oracle's FP64 solution and the column solved alone."""
    A, G, D, W = convdiff
    nr = 64
    B = np.stack([pg.rng_normal(100 + c, A.n) for c in range(nr)], axis=1)
    Xd = _dev(np.zeros((A.n, nr)), gpu)
    res = W.fgmres(_dev(B, gpu), Xd, nrhs=nr, kdim=50, maxits=1000, tol=1e-8, prec=pg.PGS_PREC_FP32)
    X = _host(Xd)
    R = B - np.stack([A.matvec(X[:, c]) for c in range(nr)], axis=1)
    rel = np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)
    assert all(r[0] == 0 for r in res)
    assert rel.max() <= 1e-8
    O = P.oracle_of(G)
    for c in (0, 37, 63):
        xo, st, _, _, _ = orc.fgmres(A, B[:, c], prec=O, kdim=50, maxits=1000, tol=1e-8)
        assert st == 0
        assert np.linalg.norm(X[:, c] - xo) <= 1e-5 * np.linalg.norm(xo)
        x1 = _dev(np.zeros(A.n), gpu)
        (st1, it1, _), = W.fgmres(_dev(B[:, c], gpu), x1, kdim=50, maxits=1000, tol=1e-8, prec=pg.PGS_PREC_FP32)
        assert st1 == 0 and abs(it1 - res[c][1]) <= 1
