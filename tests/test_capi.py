"""The C-ABI library loads and exports every entry point the header declares
(no device calls).  CPU only."""
import os
import re
import subprocess

import pargemslr_amd as pg
from pargemslr_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", hdr)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(pgs_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    pg.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (pgs_[a-z0-9_]+)", out.stdout))
    declared = _declared()
    assert declared, "no declarations parsed"
    assert declared <= exported, sorted(declared - exported)
    assert set(capi.EXPORTED) == declared


def test_params_default_matches_reference_defaults():
    p = pg.default_params()
    I = capi.IO
    assert p[I["SOLVER_KDIM"]] == 50 and p[I["SOLVER_MAXITS"]] == 100 and p[I["SOLVER_TOL"]] == 1e-6
    assert p[I["NLEV"]] == 5 and p[I["NCOMP"]] == 4 and p[I["MINSEP"]] == 10
    assert p[I["DROPTOL_B"]] == 1e-2 and p[I["DROPTOL_C"]] == 1e-3
    assert p[I["ROWNNZ_B"]] == 100 and p[I["ROWNNZ_C"]] == 200
    assert p[I["LR_RANK1"]] == 100 and p[I["LR_RANK2"]] == 50 and p[I["LR_MAXITS1"]] == 3
    assert p[I["SCHUR_ENABLE"]] == 1


def test_strerror():
    assert pg.capi.strerror(0) == "success"
    assert "empty row" in pg.capi.strerror(103)


def test_generators_and_rng_deterministic():
    a = pg.rng_uniform(5, 1000)
    assert a.min() >= -1 and a.max() < 1 and abs(a.mean()) < 0.1
    assert (pg.rng_uniform(5, 1000) == a).all()
    z = pg.rng_normal(5, 20000)
    assert abs(z.mean()) < 0.05 and abs(z.std() - 1) < 0.05
    A = pg.laplacian(128, 128, 128)
    assert A.n == 2097152 and A.nnz == 14581760
    C = pg.convdiff(96, 96, 96, 0.3)
    assert C.n == 884736 and C.nnz == 6137856
