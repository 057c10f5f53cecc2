"""Operators read from an operator file and matrices read from MatrixMarket
files on the device: a reloaded operator uploads to the same device
operator (apply and FGMRES bit-identical to the one set up in this process),
and the reference's general-matrix driver flow solves the reference's own
matrix files through the C++ class surface."""
import gzip
import os
import subprocess

import numpy as np
import pytest

import pargemslr_amd as pg
from oracle import oracle as orc
from tests import problems as P

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "driver_gen")


def _dev(x, dev):
    dt = np.complex128 if np.iscomplexobj(x) else np.float64
    return pg.DeviceArray.from_numpy(np.ascontiguousarray(x, dtype=dt), device=dev)


def _host(t):
    pg.synchronize(t.device)
    return t.numpy()


def _rhs(A, seed, nr):
    if A.is_complex:
        return np.stack([pg.rng_normal(seed + c, 2 * A.n).view(np.complex128) for c in range(nr)], axis=1)
    return np.stack([pg.rng_normal(seed + c, A.n) for c in range(nr)], axis=1)


@pytest.mark.parametrize("cplx", [False, True])
def test_reloaded_operator_solves_bit_identically(tmp_path, cplx, gpu):
    A = pg.laplacian(24, 24, 24, shift=(0.2 + 0.02j) if cplx else 0.3, alphas=(0.05, -0.02, 0.01), complex_=cplx)
    G = P.setup(A, 3, 8, 1e-3, 20, 16, schur=(5, 1e-2))
    f = str(tmp_path / "op.pgs")
    G.save(f)
    H = pg.Gemslr.load(f)
    nr = 4
    B = _rhs(A, 70, nr)
    outs = []
    for op in (G, H):
        D = pg.DevOp(op)
        W = pg.Workspace(D, max_nrhs=nr, kdim=50)
        Z = _dev(np.zeros_like(B), gpu)
        W.apply(_dev(B, gpu), Z, nrhs=nr, prec=pg.PGS_PREC_FP32)
        X = _dev(np.zeros_like(B), gpu)
        res = W.fgmres(_dev(B, gpu), X, nrhs=nr, kdim=50, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP32)
        outs.append((_host(Z), _host(X), [r[:2] for r in res]))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert outs[0][2] == outs[1][2] and all(st == 0 for st, _ in outs[0][2])


def test_file_matrix_fgmres_parity(gpu):
    """gr_30_30 and its right-hand side from the reference's test files: FP64
    device FGMRES within one iteration of the oracle, true residual at the
    tolerance."""
    A = pg.read_mm(os.path.join(P.GOLDEN, "gr_30_30.mtx"))
    b = pg.read_mm_vector(os.path.join(P.GOLDEN, "gr_30_30_b.mtx"))
    G = P.setup(A, 2, 4, 1e-3, 20, 8)
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=1, kdim=30)
    xo, st, it_o, _, _ = orc.fgmres(A, b, prec=P.oracle_of(G), kdim=30, maxits=300, tol=1e-8)
    assert st == 0
    xd = _dev(np.zeros(A.n), gpu)
    (status, it, _), = W.fgmres(_dev(b, gpu), xd, kdim=30, maxits=300, tol=1e-8, prec=pg.PGS_PREC_FP64,
                                orth=pg.PGS_ORTH_MGS)
    x = _host(xd)
    assert status == 0 and abs(it - it_o) <= 1, (it, it_o)
    assert np.linalg.norm(b - A.matvec(x)) <= 1e-8 * np.linalg.norm(b)


def _driver(args):
    assert os.path.exists(DRIVER), "tests/cpp/driver_gen is missing: run __graft_entry__.build()"
    out = subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    return out.stdout


def _line(out, key):
    return next(l for l in out.splitlines() if l.startswith(key))


def test_driver_gen_real_with_rhs_file_save_then_load(tmp_path, gpu):
    """driver_gen on gr_30_30 with its vector file; the second run loads the
    operator file the first one wrote and reports the same iterations and
    the same Givens residual to the last digit."""
    mat = os.path.join(P.GOLDEN, "gr_30_30.mtx")
    rhs = os.path.join(P.GOLDEN, "gr_30_30_b.mtx")
    opf = str(tmp_path / "gr.pgs")
    common = [mat, rhs, "0", "2", "4", "1e-3", "20", "8", "30", "1e-8", "1"]
    first = _driver(common + ["save", opf])
    second = _driver(common + ["load", opf])
    for key in ("Number of iterations", "Final rel res (Givens)"):
        assert _line(first, key) == _line(second, key)


def test_driver_gen_complex_hel20(tmp_path, gpu):
    """The reference's complex general driver input (hel20, x* = 1), complex64
    apply inside complex128 FGMRES."""
    path = tmp_path / "hel20.mtx"
    path.write_bytes(gzip.open(os.path.join(P.GOLDEN, "hel20.mtx.gz")).read())
    _driver([str(path), ".", "1", "3", "4", "1e-3", "20", "16", "50", "1e-8", "2"])
