"""The U variant of SolveLevel (PARGEMSLR_IO_ADVANCED_GLOBAL_SOLVE = 1,
kGemslrUSolve, gemslr.cpp:3527-3532) on the device, through the C ABI,
against the oracle: FP64 applies to 1e-12 (real and complex, with and
without the inner Schur iteration), FP32 applies to 1e-5, FP64 FGMRES in
the oracle's iteration count (+-1)."""
import numpy as np
import pytest

import pargemslr_amd as pg
from oracle import oracle as orc
from tests import problems as P

pytestmark = pytest.mark.gpu


def _dev(x):
    dt = np.complex128 if np.iscomplexobj(x) else np.float64
    return pg.DeviceArray.from_numpy(np.ascontiguousarray(x, dtype=dt))


def _rhs(A, seed):
    if A.is_complex:
        return pg.rng_normal(seed, 2 * A.n).view(np.complex128)
    return pg.rng_normal(seed, A.n)


def _op(cplx, schur=None):
    A = pg.laplacian(24, 24, 24, shift=(0.2 + 0.02j) if cplx else 0.3, alphas=(0.05, -0.02, 0.01), complex_=cplx)
    G = P.setup(A, 3, 8, 1e-3, 20, 16, usolve=True, schur=schur)
    assert G.usolve()
    return A, G


@pytest.mark.parametrize("cplx,schur", [(False, None), (True, None), (False, (5, 1e-2))],
                         ids=["real", "complex", "real-schur"])
def test_apply_parity(cplx, schur, gpu):
    A, G = _op(cplx, schur)
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=4, kdim=10)
    O = P.oracle_of(G)
    for seed in (3, 4):
        v = _rhs(A, seed)
        ref = O.apply(v)
        z = pg.DeviceArray.from_numpy(np.zeros_like(v))
        W.apply(_dev(v), z, prec=pg.PGS_PREC_FP64)
        pg.synchronize()
        assert np.linalg.norm(z.numpy() - ref) <= 1e-12 * np.linalg.norm(ref), seed
    if not cplx:
        nr = 4
        V = np.stack([_rhs(A, 10 + c) for c in range(nr)], axis=1)
        Z = pg.DeviceArray((A.n, nr))
        W.apply(_dev(V), Z, nrhs=nr, prec=pg.PGS_PREC_FP32)
        pg.synchronize()
        Zh = Z.numpy().reshape(A.n, nr)
        for c in range(nr):
            ref = O.apply(V[:, c])
            assert np.linalg.norm(Zh[:, c] - ref) <= 1e-5 * np.linalg.norm(ref), c
    W.close()


def test_fgmres_parity(gpu):
    """On the convection-diffusion operator (the shifted Laplacian above is
    indefinite, and the U variant leaves FGMRES(50) stagnating on it)."""
    A = pg.convdiff(24, 24, 24, 0.3)
    G = P.setup(A, 3, 8, 1e-3, 20, 16, usolve=True)
    D = pg.DevOp(G)
    W = pg.Workspace(D, max_nrhs=1, kdim=50)
    b = _rhs(A, 21)
    xo, st, it_o, _, _ = orc.fgmres(A, b, prec=P.oracle_of(G), kdim=50, maxits=500, tol=1e-8)
    assert st == 0
    xd = _dev(np.zeros(A.n))
    (status, it, rr), = W.fgmres(_dev(b), xd, kdim=50, maxits=500, tol=1e-8, prec=pg.PGS_PREC_FP64,
                                 orth=pg.PGS_ORTH_DCGS2)
    pg.synchronize()
    x = xd.numpy()
    print(f"[usolve] fgmres iterations {it} (oracle {it_o})", flush=True)
    assert status == 0
    assert abs(it - it_o) <= 1, (it, it_o)
    assert np.linalg.norm(b - A.matvec(x)) <= 1e-8 * np.linalg.norm(b)
    W.close()
