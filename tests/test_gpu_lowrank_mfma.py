"""The low-rank correction z += WH (W^T z) of FP32 blocks of 16, 32 and 64
columns on the matrix cores (k_lr_dot_mfma / k_lr_update_mfma,
gemslr.cpp:3681-3719) against the same apply with the VALU kernels
(PGS_LR_MFMA=0 when the operator is uploaded) and against the oracle's FP64
apply, through the C ABI; ranks 16 and 32 (one and two 16-row tiles of the
rank index), achieved ranks below the requested one included."""
import os

import numpy as np
import pytest

import pargemslr_amd as pg
from tests import problems as P

pytestmark = pytest.mark.gpu


def _dev(x):
    return pg.DeviceArray.from_numpy(np.ascontiguousarray(x, dtype=np.float64))


def _upload(G, mfma):
    old = os.environ.get("PGS_LR_MFMA")
    os.environ["PGS_LR_MFMA"] = "1" if mfma else "0"
    try:
        return pg.DevOp(G)
    finally:
        if old is None:
            del os.environ["PGS_LR_MFMA"]
        else:
            os.environ["PGS_LR_MFMA"] = old


@pytest.fixture(scope="module", params=[16, 32], ids=["rank16", "rank32"])
def op(request, gpu):
    A = pg.convdiff(24, 24, 24, 0.3)
    G = P.setup(A, 3, 8, 1e-3, 20, request.param)
    ranks = [lv["rank"] for lv in G.levels()[:-1]]
    print(f"[lr-mfma] requested rank {request.param}, achieved {ranks}", flush=True)
    assert max(ranks) > 0
    Dm, Dv = _upload(G, True), _upload(G, False)
    return A, G, Dm, Dv


@pytest.mark.parametrize("nrhs", [16, 32, 64])
def test_mfma_lowrank_matches_valu_and_oracle(op, nrhs, gpu):
    A, G, Dm, Dv = op
    V = np.stack([pg.rng_normal(300 + c, A.n) for c in range(nrhs)], axis=1)
    out = []
    for D in (Dm, Dv):
        W = pg.Workspace(D, max_nrhs=nrhs, kdim=10)
        z = pg.DeviceArray((A.n, nrhs))
        W.apply(_dev(V), z, nrhs=nrhs, prec=pg.PGS_PREC_FP32)
        pg.synchronize()
        out.append(z.numpy().reshape(A.n, nrhs))
        W.close()
    zm, zv = out
    rel = np.linalg.norm(zm - zv) / np.linalg.norm(zv)
    O = P.oracle_of(G)
    worst = 0.0
    for c in (0, nrhs // 2, nrhs - 1):
        ref = O.apply(V[:, c])
        worst = max(worst, np.linalg.norm(zm[:, c] - ref) / np.linalg.norm(ref))
    print(f"[lr-mfma] nrhs {nrhs}: MFMA vs VALU {rel:.2e}, vs oracle FP64 {worst:.2e}", flush=True)
    assert not np.array_equal(zm, zv), "the matrix-core path was not taken"
    assert rel <= 1e-6
    assert worst <= 1e-5


def test_mfma_lowrank_in_a_batched_solve(op, gpu):
    """A 32-column mixed FGMRES with the matrix-core low-rank term converges
    every column to 1e-8 in the VALU path's iteration counts (+-1)."""
    A, G, Dm, Dv = op
    nr = 32
    B = np.stack([pg.rng_normal(500 + c, A.n) for c in range(nr)], axis=1)
    its = []
    for D in (Dm, Dv):
        W = pg.Workspace(D, max_nrhs=nr, kdim=50)
        X = pg.DeviceArray((A.n, nr))
        res = W.fgmres(_dev(B), X, nrhs=nr, kdim=50, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP32,
                       orth=pg.PGS_ORTH_DCGS2)
        pg.synchronize()
        Xh = X.numpy().reshape(A.n, nr)
        assert all(r[0] == 0 for r in res)
        for c in range(nr):
            assert np.linalg.norm(B[:, c] - A.matvec(Xh[:, c])) <= 1e-8 * np.linalg.norm(B[:, c])
        its.append([r[1] for r in res])
        W.close()
    assert max(abs(a - b) for a, b in zip(*its)) <= 1, its
