"""MatrixMarket ingest (pgs_read_mm_csr / pgs_read_mm_vector) against the
reference's reader rules (CooMatrixReadFromFile, matrixops.cpp:11134-11453;
CooMatrixClass::ToCsr, coo_matrix.cpp:479-512; SequentialVectorReadFromFile,
vectorops.cpp:1116-1336).  Values are pinned by scipy.io.mmread, an
independent reader, on the reference's own matrix files; the row order and
the mirroring rules by small files written here.  CPU only."""
import gzip
import os

import numpy as np
import pytest
import scipy.io as sio

import pargemslr_amd as pg
from tests import problems as P

GOLDEN = P.GOLDEN


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def _rows(A):
    return [(list(A.ja[A.ia[i]:A.ia[i + 1]]), list(A.a[A.ia[i]:A.ia[i + 1]])) for i in range(A.n)]


@pytest.mark.parametrize("name", ["gr_30_30.mtx", "test_small.mtx", "test_nd.mtx", "test_sym_small.mtx"])
def test_reference_matrix_files_match_scipy(name):
    path = os.path.join(GOLDEN, name)
    A = pg.read_mm(path)
    M = sio.mmread(path).toarray()
    assert A.n == M.shape[0] and not A.is_complex
    assert np.array_equal(A.to_scipy().toarray(), M)
    # the complexd reader gives the same values with zero imaginary parts
    Z = pg.read_mm(path, complex_=True)
    assert Z.is_complex and np.array_equal(Z.ia, A.ia) and np.array_equal(Z.ja, A.ja)
    assert np.array_equal(Z.a, A.a.astype(np.complex128))


def test_gr_30_30_symmetric_expansion_count():
    """gr_30_30 is stored as its lower triangle (4322 entries, 900 on the
    diagonal): the expansion mirrors the 3422 off-diagonal ones."""
    A = pg.read_mm(os.path.join(GOLDEN, "gr_30_30.mtx"))
    assert A.nnz == 900 + 2 * 3422


@pytest.mark.parametrize("name", ["test_smallz.mtx"])
def test_reference_complex_files_match_scipy(name):
    path = os.path.join(GOLDEN, name)
    A = pg.read_mm(path, complex_=True)
    assert np.array_equal(A.to_scipy().toarray(), sio.mmread(path).toarray())
    with pytest.raises(pg.PgsError) as e:
        pg.read_mm(path)  # the double reader takes real and integer files only
    assert e.value.code == 102


def test_hel20_complex_general(tmp_path):
    raw = gzip.open(os.path.join(GOLDEN, "hel20.mtx.gz")).read()
    path = tmp_path / "hel20.mtx"
    path.write_bytes(raw)
    A = pg.read_mm(str(path), complex_=True)
    M = sio.mmread(str(path)).tocsr()
    assert A.n == 8000 and A.nnz == 53600
    assert abs(A.to_scipy() - M).max() == 0


def test_rows_keep_file_order_and_duplicates(tmp_path):
    """ToCsr buckets COO entries by row without sorting or summing."""
    f = _write(tmp_path, "g.mtx", "%%MatrixMarket matrix coordinate real general\n% c\n3 3 5\n"
               "2 3 1.5\n2 1 -2\n1 1 4\n3 3 1\n2 3 0.25\n")
    A = pg.read_mm(f)
    assert _rows(A) == [([0], [4.0]), ([2, 0, 2], [1.5, -2.0, 0.25]), ([2], [1.0])]


def test_symmetric_mirror_follows_its_entry(tmp_path):
    f = _write(tmp_path, "s.mtx", "%%MatrixMarket matrix coordinate real symmetric\n3 3 3\n"
               "1 1 2\n3 1 -1\n2 2 5\n")
    A = pg.read_mm(f)
    assert _rows(A) == [([0, 2], [2.0, -1.0]), ([1], [5.0]), ([0], [-1.0])]


def test_complex_storage_rules(tmp_path):
    body = "3 3 2\n1 1 1 1\n3 2 2 -3\n"
    want = {"general": None, "symmetric": 2 - 3j, "hermitian": 2 + 3j, "skew-symmetric": -2 + 3j}
    for storage, mirror in want.items():
        f = _write(tmp_path, storage + ".mtx", f"%%MatrixMarket matrix coordinate complex {storage}\n" + body)
        D = pg.read_mm(f, complex_=True).to_scipy().toarray()
        assert D[0, 0] == 1 + 1j and D[2, 1] == 2 - 3j
        assert D[1, 2] == (0 if mirror is None else mirror), storage


def test_real_reader_mirrors_symmetric_only(tmp_path):
    """The double reader mirrors only for "symmetric"; a skew-symmetric real
    file keeps its stored triangle (the complexd reader mirrors it as -a)."""
    f = _write(tmp_path, "k.mtx", "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 3\n")
    assert pg.read_mm(f).to_scipy().toarray().tolist() == [[0, 0], [3, 0]]
    assert pg.read_mm(f, complex_=True).to_scipy().toarray().tolist() == [[0, -3], [3, 0]]


def test_zero_based_indices(tmp_path):
    f = _write(tmp_path, "z.mtx", "%%MatrixMarket matrix coordinate integer general\n2 2 2\n0 0 1\n1 0 7\n")
    A = pg.read_mm(f, idxin=0)
    assert A.to_scipy().toarray().tolist() == [[1, 0], [7, 0]]
    with pytest.raises(pg.PgsError):
        pg.read_mm(f)  # 1-based reading sees index 0


def test_banner_is_case_insensitive_after_the_first_token(tmp_path):
    f = _write(tmp_path, "c.mtx", "%%MatrixMarket MATRIX Coordinate REAL General\n\n% late comment\n1 1 1\n1 1 2.5\n")
    assert pg.read_mm(f).a.tolist() == [2.5]


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix coordinate pattern general\n2 2 1\n1 1\n",
    "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
    "%%MatrixMarket vector coordinate real general\n2 2 1\n1 1 1\n",
    "%%Matrixmarket matrix coordinate real general\n2 2 1\n1 1 1\n",
    "%%MatrixMarket matrix coordinate real\n2 2 1\n1 1 1\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 x 1\n",
    "%%MatrixMarket matrix coordinate real general\n% only a comment\n",
], ids=["pattern", "array", "object", "banner-case", "four-tokens", "short-list", "row-range", "bad-number",
        "no-size"])
def test_rejected_files(tmp_path, text):
    f = _write(tmp_path, "bad.mtx", text)
    for cplx in (False, True):
        with pytest.raises(pg.PgsError) as e:
            pg.read_mm(f, complex_=cplx)
        assert e.value.code == 102


def test_missing_file(tmp_path):
    with pytest.raises(pg.PgsError) as e:
        pg.read_mm(str(tmp_path / "none.mtx"))
    assert e.value.code == 102


def test_vector_files():
    b = pg.read_mm_vector(os.path.join(GOLDEN, "gr_30_30_b.mtx"))
    assert np.array_equal(b, sio.mmread(os.path.join(GOLDEN, "gr_30_30_b.mtx")).toarray().ravel())
    z = pg.read_mm_vector(os.path.join(GOLDEN, "test_smallz_rhs.mtx"), complex_=True)
    assert np.array_equal(z, sio.mmread(os.path.join(GOLDEN, "test_smallz_rhs.mtx")).toarray().ravel())


def test_vector_rules(tmp_path):
    """v[I] = value, later entries overwrite, J ignored by the double reader;
    the complexd reader requires one column."""
    f = _write(tmp_path, "v.mtx", "%%MatrixMarket matrix coordinate real general\n4 2 3\n2 1 5\n4 2 1\n2 2 -1\n")
    assert pg.read_mm_vector(f).tolist() == [0, -1, 0, 1]
    with pytest.raises(pg.PgsError):
        pg.read_mm_vector(f, complex_=True)


def test_read_matrix_then_setup_and_host_apply():
    """A file-read operator goes through the setup like a generated one."""
    A = pg.read_mm(os.path.join(GOLDEN, "gr_30_30.mtx"))
    G = P.setup(A, 2, 4, 1e-3, 20, 8)
    v = pg.rng_normal(3, A.n)
    z = G.host_apply(v)
    ref = P.oracle_of(G).apply(v)
    assert np.linalg.norm(z - ref) <= 1e-13 * np.linalg.norm(ref)
