"""The U variant of SolveLevel (PARGEMSLR_IO_ADVANCED_GLOBAL_SOLVE = 1,
kGemslrUSolve, gemslr.cpp:3527-3532): in the solve phase the block L solve
on the way down is skipped, so with exact block factors and no low-rank term
one apply is the block upper-triangular solve [[B, F], [0, C]]^{-1}.  Pins the
oracle's U variant against that closed form and the library's host apply
against the oracle.  CPU only."""
import numpy as np
import pytest

import pargemslr_amd as pg
from tests import problems as P


def _two_level(usolve):
    A = pg.laplacian(12, 12, 1, alphas=(0.15, -0.25, 0.0))  # nonsymmetric: E != F^T
    return A, P.setup(A, 2, 4, 0.0, A.n, 0, usolve=usolve)


def test_u_variant_is_the_block_upper_solve():
    A, G = _two_level(True)
    assert G.usolve()
    lv = G.levels()
    perm = G.perm()
    Ad = P.dense(A)[np.ix_(perm, perm)]
    e = lv[0]["nloc"]
    B, F, Cm = Ad[:e, :e], Ad[:e, e:], Ad[e:, e:]
    b = np.random.default_rng(5).standard_normal(A.n)
    bp = b[perm]
    xl = np.linalg.solve(Cm, bp[e:])
    xu = np.linalg.solve(B, bp[:e] - F @ xl)
    x = np.empty(A.n)
    x[perm] = np.concatenate([xu, xl])
    z = P.oracle_of(G).apply(b)
    assert np.allclose(z, x, rtol=1e-11, atol=1e-11)
    # the LU variant on the same factors is a different operator (E enters it)
    _, G_lu = _two_level(False)
    assert not G_lu.usolve()
    assert not np.allclose(P.oracle_of(G_lu).apply(b), x, rtol=1e-6)


def test_u_variant_without_coupling_is_block_jacobi():
    """Dropping F as well leaves diag(B, C)^{-1}: a planted mistake in the
    oracle's input (F zeroed) must change the apply, so F is used."""
    A, G = _two_level(True)
    lv = G.levels()
    b = np.random.default_rng(6).standard_normal(A.n)
    z = P.oracle_of(G).apply(b)
    lv[0]["F"] = (lv[0]["F"][0], lv[0]["F"][1], np.zeros_like(lv[0]["F"][2]))
    from oracle import oracle as orc
    z0 = orc.GemslrOracle(A.n, G.perm(), lv, usolve=True).apply(b)
    assert not np.allclose(z, z0, rtol=1e-6)


def test_setup_factors_do_not_depend_on_the_variant():
    """The option only changes the solve phase: the setup (and the low-rank
    build, which applies the LU variant) gives the same bits."""
    A = pg.convdiff(16, 16, 16, 0.3)
    G0 = P.setup(A, 3, 8, 1e-3, 20, 16, usolve=False)
    G1 = P.setup(A, 3, 8, 1e-3, 20, 16, usolve=True)
    for a, b in zip(G0.levels(), G1.levels()):
        assert a["rank"] == b["rank"]
        if a["rank"]:
            assert np.array_equal(a["W"], b["W"]) and np.array_equal(a["WH"], b["WH"])


@pytest.mark.parametrize("schur", [None, (5, 1e-2)], ids=["plain", "schur"])
def test_host_apply_matches_oracle(schur):
    A = pg.convdiff(16, 16, 16, 0.3)
    G = P.setup(A, 3, 8, 1e-3, 20, 16, usolve=True, schur=schur)
    b = pg.rng_normal(11, A.n)
    z_host = G.host_apply(b)
    z_orc = P.oracle_of(G).apply(b)
    assert np.linalg.norm(z_host - z_orc) <= 1e-13 * np.linalg.norm(z_orc)


def test_u_variant_preconditions_fgmres():
    from oracle import oracle as orc
    A = pg.convdiff(16, 16, 16, 0.3)
    G = P.setup(A, 3, 8, 1e-3, 20, 16, usolve=True)
    b = pg.rng_normal(12, A.n)
    x, st, its, rr, _ = orc.fgmres(A, b, prec=P.oracle_of(G), kdim=50, maxits=300, tol=1e-8)
    assert st == 0
    assert np.linalg.norm(b - A.matvec(x)) <= 1e-8 * np.linalg.norm(b)
