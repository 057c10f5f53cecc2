"""Parameter files of the reference's drivers (pgs_params_read_inputs) and the
hypre-style preconditioner object's host side.  The expected values are read
off the reference's own `inputs` files (tests/golden/inputs_sequential from
TESTS/sequential, tests/golden/inputs_hypre from TESTS/hypre) column by
column; the indices are the PARGEMSLR_IO_* ones of ParGeMSLR/SRC/utils/
utils.hpp.  CPU only."""
import os

import numpy as np
import pytest

import pargemslr_amd as pg
from tests import problems as P

SEQ = os.path.join(P.GOLDEN, "inputs_sequential")
HYP = os.path.join(P.GOLDEN, "inputs_hypre")


def test_sequential_inputs_file():
    p = pg.read_inputs(SEQ, pg.PGS_INPUTS_SEQUENTIAL)
    d = pg.default_params()
    want = {
        1: 0, 2: 1, 3: 0,                      # ILUT, 1 level, ILUT
        11: 50, 12: 200, 13: 1e-8, 14: 0,      # kdim, maxits, tol, absolute
        20: 5, 21: 2, 22: 2, 23: 2,            # nlev0/1, ncomp0/1
        24: 0, 25: 0,                          # ND, ND
        40: 1e-3, 41: 1e-3, 44: 1e-3, 45: 1e-3,
        48: 200, 49: 200, 50: 200, 51: 200,
        52: 1, 53: 200, 54: 1, 55: 200,        # lfil; the LOCAL twins take ROWNNZ (the reference's copy)
        58: 1, 59: 1,                          # RCM
        70: 1, 72: 1, 71: 1, 73: 1,            # TR, TR
        74: 32, 76: 32, 75: 32, 77: 32,
        86: 1e-3, 88: 1e-3, 87: 1e-3, 89: 1e-3,
        78: 1.0, 80: 1.0, 79: 1.0, 81: 1.0,
        82: 2.0, 84: 2.0, 83: 2.0, 85: 2.0,
        90: 10, 92: 10, 91: 10, 93: 10,
        130: 1, 131: 1e-2, 132: 5, 142: 0.0,
        141: 0,                                # LU
        150: 1,
    }
    for i, v in want.items():
        assert p[i] == v, (i, p[i], v)
    untouched = [i for i in range(180) if i not in want]
    assert np.array_equal(p[untouched], d[untouched])


def test_hypre_inputs_file():
    p = pg.read_inputs(HYP, pg.PGS_INPUTS_HYPRE)
    want = {
        0: 2, 34: 0,                           # GEMSLR, no global partition
        1: 1, 2: 1, 3: 1, 4: 1, 160: 1,        # ILUK, 1, ILUK, C ILUK, C_iter
        20: 5, 21: 2, 22: 16, 23: 2,
        24: 1, 25: 1,                          # RKWAY, RKWAY
        40: 1e-2, 42: 1e-2, 46: 1e-2, 44: 1e-2, 41: 1e-2, 43: 1e-2, 47: 1e-2, 45: 1e-2,
        48: 200, 56: 200, 50: 200, 49: 200, 57: 200, 51: 200,
        52: 0, 54: 0, 53: 0, 55: 0,
        58: 1, 59: 1, 70: 1, 72: 1, 71: 1, 73: 1,
        74: 50, 75: 50, 76: 50, 77: 50,
        86: 1e-9, 87: 1e-9, 88: 1e-9, 89: 1e-9,
        78: 1.1, 79: 1.1, 80: 1.1, 81: 1.1,
        82: 3.0, 83: 3.0, 84: 3.0, 85: 3.0,
        90: 5, 91: 5, 92: 5, 93: 5,
        130: 1, 131: 1e-2, 132: 5, 142: 0, 141: 0, 143: 0, 144: 0, 5: 0, 150: 1,
    }
    for i, v in want.items():
        assert p[i] == v, (i, p[i], v)


def test_keyword_case_and_partial_numbers(tmp_path):
    """Keywords are upper-cased; a numeric line keeps the values it could
    read and the old values after the first field that is not a number."""
    lines = open(HYP).read().splitlines()
    lines[0] = "   gemslr   1. global_precond"
    lines[7] = " 7   x   8. nlev"   # nlev global read, local kept
    f = tmp_path / "inputs"
    f.write_text("\n".join(lines) + "\n")
    d = pg.default_params()
    p = pg.read_inputs(str(f), pg.PGS_INPUTS_HYPRE)
    assert p[0] == 2 and p[20] == 7 and p[21] == d[21]


@pytest.mark.parametrize("fmt, line", [(pg.PGS_INPUTS_SEQUENTIAL, 11), (pg.PGS_INPUTS_HYPRE, 9)])
def test_unknown_keyword_is_refused(tmp_path, fmt, line):
    src = SEQ if fmt == pg.PGS_INPUTS_SEQUENTIAL else HYP
    lines = open(src).read().splitlines()
    lines[line] = " METIS   bad partition option"
    f = tmp_path / "inputs"
    f.write_text("\n".join(lines) + "\n")
    with pytest.raises(pg.PgsError) as e:
        pg.read_inputs(str(f), fmt)
    assert e.value.code == 102


def test_missing_file():
    with pytest.raises(pg.PgsError) as e:
        pg.read_inputs("/nonexistent/inputs")
    assert e.value.code == 102


def test_setup_refuses_options_it_does_not_implement():
    """The reference's sample files ask for ND partitions (sequential) or
    ILUK block solves (hypre): the setup refuses them instead of computing a
    different preconditioner."""
    A = pg.laplacian(8, 8, 8)
    for path, fmt in ((SEQ, pg.PGS_INPUTS_SEQUENTIAL), (HYP, pg.PGS_INPUTS_HYPRE)):
        p = pg.read_inputs(path, fmt)
        with pytest.raises(pg.PgsError) as e:
            pg.Gemslr(A, p)
        assert e.value.code == 100
    base = pg.config_params(2, 4, 1e-2, 10, 4)
    for idx, val in ((1, 1), (3, 2), (141, 2), (58, 2), (70, 0)):
        p = base.copy()
        p[idx] = val
        with pytest.raises(pg.PgsError) as e:
            pg.Gemslr(A, p)
        assert e.value.code == 100, idx


def test_hypre_object_reads_defaults_without_a_file(tmp_path, monkeypatch):
    """Like HYPRE_GEMSLRCreate: no file named means `inputs` in the working
    directory, and a missing file leaves the defaults."""
    monkeypatch.chdir(tmp_path)
    h = pg.HypreGemslr()
    assert np.array_equal(h.params(), pg.default_params())
    (tmp_path / "inputs").write_text(open(HYP).read())
    h2 = pg.HypreGemslr()
    assert h2.params()[22] == 16
    with pytest.raises(pg.PgsError) as e:
        h2.solve(np.zeros(4))      # solve before setup
    assert e.value.code == 106
