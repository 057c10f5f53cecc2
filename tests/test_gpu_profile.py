"""The workspace profiler's kernel families (pgs_ws_profile_families): launch
counts of one preconditioner apply follow the level recursion of SolveLevel
(gemslr.cpp:3474-3625), the algorithmic bytes follow the header's formulas,
and a batched FGMRES solve records every family of its path with a positive
time.  Profiling does not change results."""
import numpy as np
import pytest

import pargemslr_amd as pg
from tests import problems as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def op(gpu):
    A, G = P.build(P.C2_SMALL)
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=8, kdim=20)
    return A, G, D, W


def test_apply_families(op, gpu):
    A, G, D, W = op
    nr = 4
    lv = G.levels()
    nlev = len(lv)
    V = pg.DeviceArray.from_numpy(np.stack([pg.rng_normal(20 + c, A.n) for c in range(nr)], axis=1))
    Z0, Z1 = pg.DeviceArray((A.n, nr)), pg.DeviceArray((A.n, nr))
    W.apply(V, Z0, nrhs=nr, prec=pg.PGS_PREC_FP32)
    W.profile(True)
    W.apply(V, Z1, nrhs=nr, prec=pg.PGS_PREC_FP32)
    pg.synchronize()
    f = W.profile_families()
    W.profile(False)
    assert np.array_equal(Z0.numpy(), Z1.numpy())
    # two block solves on every level but the last, one there
    assert f["trsv level 0"]["launches"] == 2
    assert f["trsv levels 1+"]["launches"] == 2 * (nlev - 2) + 1
    assert f["spmv E/F"]["launches"] == 2 * (nlev - 1)
    assert f["permute"]["launches"] == 2
    assert f["low-rank"]["launches"] == sum(1 for d in lv[:-1] if d["rank"] > 0)
    # gather FP64 -> FP32 and scatter FP32 -> FP64 of n x nr values, plus the permutation
    assert f["permute"]["total_bytes"] == 2 * (A.n * nr * (8 + 4) + 4 * A.n)
    for name in ("trsv level 0", "trsv levels 1+", "spmv E/F", "permute"):
        assert f[name]["total_ms"] > 0, name
    print({k: v for k, v in f.items() if v["launches"]}, flush=True)


def test_fgmres_families(op, gpu):
    A, G, D, W = op
    nr = 8
    B = pg.DeviceArray.from_numpy(np.stack([pg.rng_normal(40 + c, A.n) for c in range(nr)], axis=1))
    X = pg.DeviceArray((A.n, nr))
    W.profile(True)
    res = W.fgmres(B, X, nrhs=nr, kdim=20, maxits=200, tol=1e-8, prec=pg.PGS_PREC_FP32, orth=pg.PGS_ORTH_DCGS2)
    pg.synchronize()
    f = W.profile_families()
    W.profile(False)
    its = max(r[1] for r in res)
    assert all(r[0] == 0 for r in res)
    # one product with A per step plus the initial residual and one per restart
    assert its + 1 <= f["spmv A"]["launches"] <= its + 2 + 2 * (its // 20 + 1)
    assert f["trsv level 0"]["launches"] >= 2 * its
    for name in ("spmv A", "orth dots", "orth update", "x update", "state"):
        assert f[name]["launches"] > 0 and f[name]["total_ms"] > 0, name
    # delayed CGS2 leaves new basis vectors unnormalized for their readers to
    # scale: no column-scaling sweep
    assert f["scale"]["launches"] == 0
