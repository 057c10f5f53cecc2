"""pgs_apply_noalloc (the apply with its buffers in the library's static
device scratch array, no allocation) against pgs_apply (workspace buffers),
through the C ABI: the same bits on two calls in a row, the second with a
smaller input, in both precisions; the oracle's apply to the FP64 bar; sizes
beyond the array's capacity and operators with the inner Schur iteration are
refused with their own codes."""
import numpy as np
import pytest

import pargemslr_amd as pg
from tests import problems as P

pytestmark = pytest.mark.gpu


def _dev(x):
    return pg.DeviceArray.from_numpy(np.ascontiguousarray(x, dtype=np.float64))


def _pair(D, W, V, nrhs, prec):
    n = D.n
    za, zb = pg.DeviceArray((n, nrhs)), pg.DeviceArray((n, nrhs))
    vd = _dev(V)
    W.apply(vd, za, nrhs=nrhs, prec=prec)
    D.apply_noalloc(vd, zb, nrhs=nrhs, prec=prec)
    pg.synchronize()
    return za.numpy(), zb.numpy()


@pytest.fixture(scope="module")
def ops(gpu):
    out = []
    for cfg in (P.C2_SMALL, P.C1):
        A, G = P.build(cfg)
        D = pg.DevOp(G)
        W = pg.Workspace(D, max_nrhs=8, kdim=10)
        out.append((A, G, D, W))
    return out


def test_noalloc_equals_workspace_apply(ops, gpu):
    (A, G, D, W), (A1, G1, D1, W1) = ops
    O = P.oracle_of(G)
    for prec in (pg.PGS_PREC_FP64, pg.PGS_PREC_FP32):
        # first call: 8 columns of the 32^3 operator
        V = np.stack([pg.rng_normal(70 + c, A.n) for c in range(8)], axis=1)
        za, zb = _pair(D, W, V, 8, prec)
        print(f"[noalloc] prec {prec} 32^3 x 8: max |diff| {np.abs(za - zb).max():.3e}", flush=True)
        assert np.array_equal(za, zb)
        if prec == pg.PGS_PREC_FP64:
            ref = O.apply(V[:, 0])
            assert np.linalg.norm(zb.reshape(A.n, 8)[:, 0] - ref) <= 1e-12 * np.linalg.norm(ref)
        # second call, smaller: one column of the same operator, then the 64^2 operator
        v = pg.rng_normal(90, A.n)
        za, zb = _pair(D, W, v, 1, prec)
        assert np.array_equal(za, zb)
        v1 = pg.rng_normal(91, A1.n)
        za, zb = _pair(D1, W1, v1, 1, prec)
        assert np.array_equal(za, zb)


def test_noalloc_refuses_what_does_not_fit(gpu):
    A, G = P.build(dict(grid=(48, 48, 48), nlev=2, ncomp=8, droptol=1e-2, rownnz=10, rank=0, kdim=10))
    D = pg.DevOp(G, prec_mask=pg.PGS_PREC_FP64)
    n, nr = A.n, 64  # 48^3 x 64 doubles: 56.6 MB per vector block, three of them
    v, z = pg.DeviceArray((n, nr)), pg.DeviceArray((n, nr))
    with pytest.raises(pg.PgsError) as e:
        D.apply_noalloc(v, z, nrhs=nr)
    assert e.value.code == pg.capi.PGS_ERR_SCRATCH_LIMIT
    # a size that fits on the same operator works
    W = pg.Workspace(D, max_nrhs=4, kdim=10)
    V = np.stack([pg.rng_normal(5 + c, n) for c in range(4)], axis=1)
    za, zb = _pair(D, W, V, 4, pg.PGS_PREC_FP64)
    assert np.array_equal(za, zb)
    As, Gs = P.build(P.C1)
    Gs2 = P.setup(As, 2, 4, 1e-2, 10, 0, schur=(5, 1e-2))
    Ds = pg.DevOp(Gs2)
    with pytest.raises(pg.PgsError) as e:
        Ds.apply_noalloc(pg.DeviceArray(As.n), pg.DeviceArray(As.n))
    assert e.value.code == 205
