"""The hypre-style preconditioner object (pgs_hypre_gemslr_*) on the device:
created from a hypre-layout inputs file, set up on a CSR matrix, its solve is
one GeMSLR apply (the oracle's to 1e-12 in FP64, 1e-4 in FP32), and as the
preconditioner of a flexible GMRES on the host (what hypre's FlexGMRES does
with HYPRE_GEMSLRSolve) it takes the oracle FGMRES's iteration count."""
import os

import numpy as np
import pytest

import pargemslr_amd as pg
from oracle import oracle as orc
from tests import problems as P

pytestmark = pytest.mark.gpu


def _inputs(tmp_path):
    """The reference's hypre inputs with the options this library implements:
    ILUT block solves, 3 levels, 8 subdomains, ILUT 1e-3/20, rank 8; no inner
    Schur iteration (its early stop would let the FP32 apply take other
    steps than the FP64 one)."""
    lines = open(os.path.join(P.GOLDEN, "inputs_hypre")).read().splitlines()
    lines[2] = " ILUT  3. B_precond_top"
    lines[4] = " ILUT  5. B_precond_lower"
    lines[7] = " 3 2   8. nlev"
    lines[8] = " 8 2   9. ncomp"
    lines[11] = " 1e-3 1e-3 1e-3 1e-3  12. ilu_droptol"
    lines[12] = " 20 20 20  13. ilu_max_row_nnz"
    lines[17] = " 8 8  18. lr_rank"
    lines[18] = " 1e-3 1e-3  19. lr_tol_eig"
    lines[19] = " 1.0 1.0  20. lr_rank_factor"
    lines[20] = " 2.0 2.0  21. lr_arnoldi_factor"
    lines[21] = " 10 10  22. lr_maxits"
    lines[22] = " 0  23. inner_iteration"
    f = tmp_path / "inputs"
    f.write_text("\n".join(lines) + "\n")
    return str(f)


def _fgmres(A, b, prec, m, tol):
    """Right-preconditioned flexible GMRES from x0 = 0, MGS, no restart."""
    beta = np.linalg.norm(b)
    V = np.zeros((A.n, m + 1))
    Z = np.zeros((A.n, m))
    H = np.zeros((m + 1, m))
    V[:, 0] = b / beta
    for j in range(m):
        Z[:, j] = prec(V[:, j])
        w = A.matvec(Z[:, j])
        for i in range(j + 1):
            H[i, j] = V[:, i] @ w
            w -= H[i, j] * V[:, i]
        H[j + 1, j] = np.linalg.norm(w)
        V[:, j + 1] = w / H[j + 1, j]
        e1 = np.zeros(j + 2)
        e1[0] = beta
        y = np.linalg.lstsq(H[:j + 2, :j + 1], e1, rcond=None)[0]
        if np.linalg.norm(e1 - H[:j + 2, :j + 1] @ y) <= tol * beta:
            return Z[:, :j + 1] @ y, j + 1
    return Z @ y, m


def test_hypre_object_apply_and_krylov(tmp_path, gpu):
    A = pg.laplacian(16, 16, 16, alphas=(0.05, -0.02, 0.01))
    path = _inputs(tmp_path)
    h = pg.HypreGemslr(path)
    p = h.params()
    assert p[0] == 2 and p[20] == 3 and p[22] == 8 and p[74] == 8
    h.setup(A)
    G = pg.Gemslr(A, p)
    O = P.oracle_of(G)
    v = pg.rng_normal(7, A.n)
    ref = O.apply(v)
    assert np.linalg.norm(h.solve(v) - ref) <= 1e-12 * np.linalg.norm(ref)
    h.set_precision(pg.PGS_PREC_FP32)
    assert np.linalg.norm(h.solve(v) - ref) <= 1e-4 * np.linalg.norm(ref)
    h.set_precision(pg.PGS_PREC_FP64)
    b = A.matvec(pg.rng_uniform(8, A.n))
    x, it = _fgmres(A, b, h.solve, 100, 1e-8)
    assert np.linalg.norm(b - A.matvec(x)) <= 1e-8 * np.linalg.norm(b) * 1.01
    _, st, it_o, _, _ = orc.fgmres(A, b, prec=O, kdim=100, maxits=100, tol=1e-8)
    assert st == 0 and abs(it - it_o) <= 1, (it, it_o)
    h.close()
