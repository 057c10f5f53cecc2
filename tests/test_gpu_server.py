"""Request server (continuous batching over resident operators): every
request's result equals the same right-hand sides solved alone."""
import threading
import time

import numpy as np
import pytest

import pargemslr_amd as pg
from tests import problems as P

pytestmark = pytest.mark.gpu

TOL = 1e-8


@pytest.fixture(scope="module")
def operators(gpu):
    """Two resident operators: 2D Laplacian 64^2 and upwind convection-diffusion 20^3."""
    out = []
    for A, cfg in ((pg.laplacian(64, 64, 1), (2, 4, 1e-2, 10, 0)), (pg.convdiff(20, 20, 20, 0.3), (3, 8, 1e-3, 20, 16))):
        G = P.setup(A, *cfg)
        out.append((A, G, pg.DevOp(G)))
    return out


def _alone(A, D, b):
    W = pg.Workspace(D, max_nrhs=1, kdim=50)
    xd = pg.DeviceArray(A.n)
    (st, it, _), = W.fgmres(pg.DeviceArray.from_numpy(b), xd, kdim=50, maxits=1000, tol=TOL, prec=pg.PGS_PREC_FP32)
    pg.synchronize()
    return st, it, xd.numpy()


@pytest.mark.parametrize("use_graph", [False, True], ids=["launches", "graphs"])
def test_server_mix_equals_alone(use_graph, operators, gpu):
    S = pg.Server(max_batch=8, slots=2, mem_reserve=1 << 30, kdim=50, tol=TOL, use_graph=use_graph)
    ids = [S.add_operator(D) for _, _, D in operators]
    rng = np.random.default_rng(0)
    plan = []  # (op, B)
    for k in range(14):
        op = int(rng.integers(0, len(ids)))
        nr = int(rng.integers(1, 5))
        A = operators[op][0]
        B = np.stack([pg.rng_normal(1000 + 10 * k + c, A.n) for c in range(nr)], axis=1)
        plan.append((op, B))
    reqs = [None] * len(plan)

    # half submitted at once, half from another thread while the first ones run
    def late():
        for k in range(len(plan) // 2, len(plan)):
            time.sleep(0.01)
            reqs[k] = S.submit(ids[plan[k][0]], plan[k][1])

    for k in range(len(plan) // 2):
        reqs[k] = S.submit(ids[plan[k][0]], plan[k][1])
    t = threading.Thread(target=late)
    t.start()
    t.join()
    results = [S.wait(r) for r in reqs]
    S.close()
    grouped = False
    for (op, B), (X, its, info) in zip(plan, results):
        A, _, D = operators[op]
        assert info["status"] == 0 and info["nrhs"] == B.shape[1]
        assert info["total_ms"] >= info["solve_ms"] > 0
        grouped |= info["batch_nrhs"] > B.shape[1]
        for c in range(B.shape[1]):
            b = B[:, c]
            assert np.linalg.norm(b - A.matvec(X[:, c])) <= TOL * np.linalg.norm(b)
            st, it, x1 = _alone(A, D, b)
            assert st == 0 and abs(int(its[c]) - it) <= 1, (its[c], it)
            assert np.linalg.norm(X[:, c] - x1) <= 1e-6 * np.linalg.norm(x1)
    assert grouped, "no request shared a batch"


def test_server_rejects_bad_requests(operators, gpu):
    S = pg.Server(max_batch=4, slots=1, mem_reserve=1 << 30)
    i = S.add_operator(operators[0][2])
    n = operators[0][0].n
    with pytest.raises(pg.PgsError):
        S.submit(i, np.ones((n, 5)))  # more columns than max_batch
    with pytest.raises(pg.PgsError):
        S.submit(i + 7, np.ones(n))   # no such operator
    r = S.submit(i, np.zeros(n))      # zero right-hand side: x = 0, no iterations
    X, its, info = S.wait(r)
    assert info["status"] == 0 and not X.any() and its[0] == 0
    S.close()


def test_request_joins_a_running_batch_at_a_restart(gpu):
    """One worker, FGMRES(4) on a 2D Laplacian 256^2 with a one-level ILUT
    (tens of restart cycles): a request submitted while another batch of the
    same operator runs joins that batch at a restart (joined_cycle > 0)
    instead of waiting for it to finish, and both requests' columns equal the
    same columns solved alone (iterations within two: the server restarts
    from the true residual's norm, one uninterrupted solve from the Givens
    estimate)."""
    A = pg.laplacian(256, 256, 1)
    G = P.setup(A, 1, 1, 1e-2, 5, 0)
    D = pg.DevOp(G, prec_mask=pg.PGS_PREC_FP32)
    kd = 4
    S = pg.Server(max_batch=8, slots=1, mem_reserve=1 << 30, kdim=kd, maxits=4000, tol=TOL)
    i = S.add_operator(D)
    B1 = np.stack([pg.rng_normal(7000 + c, A.n) for c in range(2)], axis=1)
    B2 = np.stack([pg.rng_normal(7100 + c, A.n) for c in range(3)], axis=1)
    r1 = S.submit(i, B1)
    time.sleep(0.01)  # the worker has taken the first request and runs its first cycles
    r2 = S.submit(i, B2)
    X2, its2, info2 = S.wait(r2)
    X1, its1, info1 = S.wait(r1)
    S.close()
    print(f"[server] first {info1}, iters {list(its1)}; joiner {info2}, iters {list(its2)}", flush=True)
    assert info1["status"] == 0 and info2["status"] == 0
    assert info1["joined_cycle"] == 0
    assert info2["joined_cycle"] > 0, "the second request did not join the running batch"
    W = pg.Workspace(D, max_nrhs=1, kdim=kd)
    for B, X, its in ((B1, X1, its1), (B2, X2, its2)):
        for c in range(B.shape[1]):
            b = B[:, c]
            assert np.linalg.norm(b - A.matvec(X[:, c])) <= TOL * np.linalg.norm(b)
            xd = pg.DeviceArray(A.n)
            (st, it, _), = W.fgmres(pg.DeviceArray.from_numpy(b), xd, kdim=kd, maxits=4000, tol=TOL,
                                    prec=pg.PGS_PREC_FP32)
            pg.synchronize()
            assert st == 0 and abs(int(its[c]) - it) <= 2, (its[c], it)
            assert np.linalg.norm(X[:, c] - xd.numpy()) <= 1e-6 * np.linalg.norm(xd.numpy())


def test_nonfinite_request_fails_alone(operators, gpu):
    """A request with a NaN in its right-hand side, batched with finite
    requests for the same operator: it returns PGS_ERR_NONFINITE after its
    first cycle, and the other requests' columns converge with the iterations
    and solutions they have without it (to rounding)."""
    A, G, D = operators[1]
    S = pg.Server(max_batch=8, slots=1, mem_reserve=1 << 30, kdim=20, tol=TOL)
    i = S.add_operator(D)
    B1 = np.stack([pg.rng_normal(8000 + c, A.n) for c in range(2)], axis=1)
    B2 = np.stack([pg.rng_normal(8100 + c, A.n) for c in range(2)], axis=1)
    Bbad = np.stack([pg.rng_normal(8200 + c, A.n) for c in range(2)], axis=1)
    Bbad[A.n // 2, 1] = np.nan
    outs = []
    for with_bad in (False, True):
        reqs = [S.submit(i, B1)] + ([S.submit(i, Bbad)] if with_bad else []) + [S.submit(i, B2)]
        outs.append([S.wait(r) for r in reqs])
    S.close()
    (g1, g2), (b1, bad, b2) = outs
    print(f"[server] bad request {bad[2]}, iterations {list(bad[1])}", flush=True)
    assert bad[2]["status"] == pg.PGS_ERR_NONFINITE
    assert int(bad[1][1]) == 0  # the NaN column stops before its first step
    for (Xa, ia, fa), (Xb, ib, fb) in ((g1, b1), (g2, b2)):
        assert fa["status"] == 0 and fb["status"] == 0
        # the batch widths differ (the reductions' partial sums with them): rounding only
        assert max(abs(int(a) - int(b)) for a, b in zip(ia, ib)) <= 1
        assert np.linalg.norm(Xa - Xb) <= 1e-6 * np.linalg.norm(Xa)
