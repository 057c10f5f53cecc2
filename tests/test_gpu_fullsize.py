"""The benchmark workloads at their own sizes, through the C ABI, against the
oracle on the same uploaded operator.

- 3D Laplacian 128^3 (n = 2,097,152), 3 levels, 64 subdomains, ILUT
  1e-3/20, rank 16: FP64 apply within 1e-12 of the oracle, FP64 FGMRES(50)
  to 1e-8 within one iteration of the oracle's (fgmres.hpp:263-558), and a
  32-column mixed batch with every column at 1e-8 and column 0 within 1e-7
  of the oracle's FP64 solution.
- The same Laplacian shifted by 0.3 (indefinite), rank 32: FGMRES(50)
  stagnates; 100 FP64 iterations report "not converged" with the Givens
  residual estimate equal to the oracle's at the same iteration.
- Upwind convection-diffusion 96^3 (c = 0.3), 64 right-hand sides in one
  block, FP32 apply: every column at 1e-8, three sampled columns against the
  oracle's FP64 solution.
- The four operators of the serving stream resident in one server with
  mixed-size requests: each request's columns equal the same columns solved
  alone.

Setup of each operator runs once per module (about 10 s at 128^3)."""
import time

import numpy as np
import pytest

import pargemslr_amd as pg
from oracle import oracle as orc
from tests import problems as P

pytestmark = pytest.mark.gpu


def _dev(x, dev):
    x = np.ascontiguousarray(x)
    if not np.iscomplexobj(x):
        x = x.astype(np.float64)
    return pg.DeviceArray.from_numpy(x, device=dev)


def _host(t):
    pg.synchronize(t.device)
    return t.numpy()


def _log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def lap128(gpu):
    t0 = time.time()
    A = pg.laplacian(128, 128, 128)
    G = P.setup(A, 3, 64, 1e-3, 20, 16)
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=32, kdim=50)
    _log(f"[fullsize] laplacian 128^3 setup+upload {time.time() - t0:.1f}s")
    yield A, G, D, W
    W.close()
    D.close()


def test_c2_apply_fp64_matches_oracle(lap128, gpu):
    A, G, D, W = lap128
    O = P.oracle_of(G)
    v = pg.rng_normal(3, A.n)
    zd = _dev(np.zeros(A.n), gpu)
    W.apply(_dev(v, gpu), zd, prec=pg.PGS_PREC_FP64)
    ref = O.apply(v)
    err = np.linalg.norm(_host(zd) - ref) / np.linalg.norm(ref)
    _log(f"[fullsize] C2 FP64 apply rel err {err:.2e}")
    assert err <= 1e-12


def test_c2_fgmres_fp64_iterations_match_oracle(lap128, gpu):
    A, G, D, W = lap128
    O = P.oracle_of(G)
    b = A.matvec(pg.rng_uniform(0, A.n))
    xo, st_o, it_o, _, _ = orc.fgmres(A, b, prec=O, kdim=50, maxits=300, tol=1e-8)
    assert st_o == 0
    for orth in (pg.PGS_ORTH_MGS, pg.PGS_ORTH_DCGS2):
        xd = _dev(np.zeros(A.n), gpu)
        (st, it, _), = W.fgmres(_dev(b, gpu), xd, kdim=50, maxits=300, tol=1e-8, prec=pg.PGS_PREC_FP64, orth=orth)
        x = _host(xd)
        true = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
        _log(f"[fullsize] C2 FP64 FGMRES orth={orth}: iters {it} (oracle {it_o}), true rel res {true:.2e}")
        assert st == 0 and abs(it - it_o) <= 1, (orth, it, it_o)
        assert true <= 1e-8
        assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


def test_c2_mixed_batch_32(lap128, gpu):
    A, G, D, W = lap128
    nr = 32
    B = np.stack([A.matvec(pg.rng_uniform(c, A.n)) for c in range(nr)], axis=1)
    Xd = _dev(np.zeros((A.n, nr)), gpu)
    res = W.fgmres(_dev(B, gpu), Xd, nrhs=nr, kdim=50, maxits=300, tol=1e-8, prec=pg.PGS_PREC_FP32,
                   orth=pg.PGS_ORTH_DCGS2)
    X = _host(Xd)
    rel = np.array([np.linalg.norm(B[:, c] - A.matvec(X[:, c])) / np.linalg.norm(B[:, c]) for c in range(nr)])
    its = [r[1] for r in res]
    _log(f"[fullsize] C2 mixed 32 RHS: iters {min(its)}-{max(its)}, max true rel res {rel.max():.2e}")
    assert all(r[0] == 0 for r in res)
    assert rel.max() <= 1e-8
    xo, st_o, _, _, _ = orc.fgmres(A, B[:, 0], prec=P.oracle_of(G), kdim=50, maxits=300, tol=1e-8)
    assert st_o == 0
    err = np.linalg.norm(X[:, 0] - xo) / np.linalg.norm(xo)
    _log(f"[fullsize] C2 mixed column 0 vs oracle FP64 solution {err:.2e}")
    assert err <= 1e-7


def test_c3_shifted_stagnation_matches_oracle(gpu):
    """sigma = 0.3 puts about 6,000 eigenvalues of the 128^3 Laplacian below
    zero; FGMRES(50) with rank-32 corrections stagnates.  The device and the
    oracle (both MGS) report the same Givens estimate after 100 iterations,
    and the device reports the column as not converged."""
    t0 = time.time()
    A = pg.laplacian(128, 128, 128, shift=0.3)
    G = P.setup(A, 3, 64, 1e-3, 20, 32)
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=1, kdim=50)
    _log(f"[fullsize] shifted 128^3 setup+upload {time.time() - t0:.1f}s")
    b = pg.rng_normal(0, A.n)
    O = P.oracle_of(G)
    _, st_o, it_o, rr_o, hist = orc.fgmres(A, b, prec=O, kdim=50, maxits=100, tol=1e-10)
    xd = _dev(np.zeros(A.n), gpu)
    (st, it, rr), = W.fgmres(_dev(b, gpu), xd, kdim=50, maxits=100, tol=1e-10, prec=pg.PGS_PREC_FP64,
                             orth=pg.PGS_ORTH_MGS)
    x = _host(xd)
    true = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
    _log(f"[fullsize] C3 100 its: device status {st} rel res {rr:.6e} (true {true:.3e}); "
         f"oracle status {st_o} rel res {rr_o:.6e}; oracle history at 25/50/75: {hist[25]:.4e} {hist[50]:.4e} "
         f"{hist[75]:.4e}")
    assert st_o != 0 and it_o == 100
    assert st == pg.PGS_ERR_NOT_CONVERGED and it == 100, (st, it)
    assert abs(rr - rr_o) <= 1e-6 * rr_o
    assert rr > 1e-3  # stagnated far above the tolerance
    W.close()
    D.close()


def test_c4_convdiff_96_64_rhs(gpu):
    t0 = time.time()
    A = pg.convdiff(96, 96, 96, 0.3)
    G = P.setup(A, 3, 64, 1e-3, 20, 16)
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=64, kdim=50)
    _log(f"[fullsize] convdiff 96^3 setup+upload {time.time() - t0:.1f}s")
    nr = 64
    B = np.stack([pg.rng_normal(100 + c, A.n) for c in range(nr)], axis=1)
    Xd = _dev(np.zeros((A.n, nr)), gpu)
    res = W.fgmres(_dev(B, gpu), Xd, nrhs=nr, kdim=50, maxits=1000, tol=1e-8, prec=pg.PGS_PREC_FP32,
                   orth=pg.PGS_ORTH_DCGS2)
    X = _host(Xd)
    rel = np.array([np.linalg.norm(B[:, c] - A.matvec(X[:, c])) / np.linalg.norm(B[:, c]) for c in range(nr)])
    its = [r[1] for r in res]
    _log(f"[fullsize] C4 64 RHS: iters {min(its)}-{max(its)}, max true rel res {rel.max():.2e}")
    assert all(r[0] == 0 for r in res)
    assert rel.max() <= 1e-8
    O = P.oracle_of(G)
    for c in (0, 37, 63):
        xo, st, it_o, _, _ = orc.fgmres(A, B[:, c], prec=O, kdim=50, maxits=1000, tol=1e-8)
        err = np.linalg.norm(X[:, c] - xo) / np.linalg.norm(xo)
        _log(f"[fullsize] C4 column {c}: iters {its[c]} (oracle FP64 {it_o}), error vs oracle {err:.2e}")
        assert st == 0
        assert err <= 1e-5
        assert abs(its[c] - it_o) <= 3
    W.close()
    D.close()


def _rhs(seed, A):
    if A.is_complex:
        return pg.rng_normal(seed, 2 * A.n).view(np.complex128)
    return pg.rng_normal(seed, A.n)


def test_c5_four_operators_server_mix_equals_alone(gpu):
    """2D Laplacian 1024^2, 3D Laplacian 64^3, convection-diffusion 96^3 and
    the complex-shifted Laplacian 80^3 (shift 0.2+0.02i) resident in one
    server; ten requests of 1-16 columns, half submitted while the first
    ones run.  Each request's columns equal the same columns solved as a
    batch of their own (iterations within one, solutions within 1e-6)."""
    mats = [pg.laplacian(1024, 1024, 1), pg.laplacian(64, 64, 64), pg.convdiff(96, 96, 96, 0.3),
            pg.laplacian(80, 80, 80, shift=0.2 + 0.02j, complex_=True)]
    params = pg.config_params(3, 64, 1e-3, 20, 16, kdim=50)
    ops = []
    for A in mats:
        t0 = time.time()
        G = pg.Gemslr(A, params, seed=0)
        D = pg.DevOp(G, prec_mask=pg.PGS_PREC_FP32)
        ops.append((A, G, D))
        _log(f"[fullsize] C5 operator n={A.n} complex={A.is_complex} setup {time.time() - t0:.1f}s")
    S = pg.Server(max_batch=16, slots=3, mem_reserve=8 << 30, kdim=50, tol=1e-8, orth=pg.PGS_ORTH_DCGS2)
    ids = [S.add_operator(D) for _, _, D in ops]
    rng = np.random.default_rng(5)
    plan = []
    for k in range(10):
        op = k % 4 if k < 4 else int(rng.integers(0, 4))
        nr = int(rng.integers(1, 17))
        A = ops[op][0]
        plan.append((op, np.stack([_rhs(5000 + 32 * k + c, A) for c in range(nr)], axis=1)))
    t0 = time.time()
    reqs = [S.submit(ids[op], B) for op, B in plan[:5]]
    time.sleep(0.2)
    reqs += [S.submit(ids[op], B) for op, B in plan[5:]]
    out = [S.wait(r) for r in reqs]
    S.close()
    _log(f"[fullsize] C5 {len(plan)} requests served in {time.time() - t0:.1f}s, batch sizes "
         f"{[o[2]['batch_nrhs'] for o in out]}")
    for (op, B), (X, its, info) in zip(plan, out):
        A, G, D = ops[op]
        assert info["status"] == 0 and info["nrhs"] == B.shape[1]
        # alone: the request's columns as a block of their own, padded with
        # zero columns to the next power of two (a zero column takes no
        # iterations and changes no other column)
        nr = B.shape[1]
        npad = 1 << (nr - 1).bit_length()
        Bp = np.zeros((B.shape[0], npad), dtype=B.dtype)
        Bp[:, :nr] = B
        W = pg.Workspace(D, max_nrhs=npad, kdim=50)
        Xa = _dev(np.zeros_like(Bp), gpu)
        res = W.fgmres(_dev(Bp, gpu), Xa, nrhs=npad, kdim=50, maxits=1000, tol=1e-8, prec=pg.PGS_PREC_FP32,
                       orth=pg.PGS_ORTH_DCGS2)
        Xa = _host(Xa)[:, :nr]
        W.close()
        for c in range(B.shape[1]):
            b = B[:, c]
            true = np.linalg.norm(b - A.matvec(X[:, c])) / np.linalg.norm(b)
            assert true <= 1e-8, (op, c, true)
            assert res[c][0] == 0 and abs(int(its[c]) - res[c][1]) <= 1, (op, c, its[c], res[c])
            assert np.linalg.norm(X[:, c] - Xa[:, c]) <= 1e-6 * np.linalg.norm(Xa[:, c]), (op, c)
        _log(f"[fullsize] C5 request op {op} nrhs {B.shape[1]}: iters {list(map(int, its))} alone "
             f"{[r[1] for r in res]}, total {info['total_ms']:.0f} ms")
