"""Operator files (pgs_gemslr_save / pgs_gemslr_load): a saved and reloaded
operator equals the one that was set up, bit for bit, and a damaged file is
refused.  CPU only (the device upload of a reloaded operator is covered by
the GPU tests)."""
import os
import struct

import numpy as np
import pytest

import pargemslr_amd as pg
from tests import problems as P

# fixed-size parts of the file layout: magic + version + kind, the parameter
# block (20 int32, 11 double, 2 uint8, 1 uint64), then n, nlev, solve_phase
# and the four setup times
HEADER = 8 + 4 + 4
PARAMS = 20 * 4 + 11 * 8 + 2 + 8
SCALARS = 4 + 4 + 1 + 4 * 8


def _fnv1a(data):
    h = 1469598103934665603
    for b in data:
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _op(cplx, schur):
    A = pg.laplacian(12, 12, 12, shift=(0.2 + 0.02j) if cplx else 0.3, alphas=(0.05, -0.02, 0.01), complex_=cplx)
    return A, P.setup(A, 3, 4, 1e-3, 20, 8, schur=schur)


def _same_levels(G, H):
    for lg, lh in zip(G.levels(), H.levels()):
        assert lg.keys() == lh.keys()
        for k in lg:
            a, b = lg[k], lh[k]
            if isinstance(a, np.ndarray):
                assert a.dtype == b.dtype and np.array_equal(a, b), k
            elif isinstance(a, tuple):
                assert all(np.array_equal(x, y) for x, y in zip(a, b)), k
            else:
                assert a == b, k


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("schur", [None, (5, 1e-2)])
def test_round_trip_is_bit_identical(tmp_path, cplx, schur):
    A, G = _op(cplx, schur)
    f = str(tmp_path / "op.pgs")
    G.save(f)
    assert not os.path.exists(f + ".tmp")
    H = pg.Gemslr.load(f)
    assert np.array_equal(H.A.ia, A.ia) and np.array_equal(H.A.ja, A.ja) and np.array_equal(H.A.a, A.a)
    assert np.array_equal(H.perm(), G.perm())
    gi, hi = G.info(), H.info()
    for k in ("n", "nlev", "is_complex", "nnz_a", "nnz_ilu", "nnz_lr", "schur_maxits", "schur_tol"):
        assert getattr(gi, k) == getattr(hi, k), k
    assert list(gi.setup_seconds) == list(hi.setup_seconds)
    assert H.schur() == G.schur()
    _same_levels(G, H)
    v = pg.rng_normal(4, 2 * A.n).view(np.complex128) if cplx else pg.rng_normal(4, A.n)
    assert np.array_equal(G.host_apply(v), H.host_apply(v))
    # the reloaded operator saves to the same bytes
    f2 = str(tmp_path / "op2.pgs")
    H.save(f2)
    assert open(f, "rb").read() == open(f2, "rb").read()


def test_file_checksum_and_layout(tmp_path):
    A, G = _op(False, None)
    f = str(tmp_path / "op.pgs")
    G.save(f)
    data = open(f, "rb").read()
    assert data[:8] == b"PGSGMSLR" and struct.unpack_from("<II", data, 8) == (1, 0)
    assert struct.unpack_from("<Q", data, len(data) - 8)[0] == _fnv1a(data[:-8])
    n, nlev = struct.unpack_from("<ii", data, HEADER + PARAMS)
    assert n == A.n and nlev == G.info().nlev


def _corrupt(tmp_path, data, name):
    p = tmp_path / name
    p.write_bytes(data)
    with pytest.raises(pg.PgsError) as e:
        pg.Gemslr.load(str(p))
    assert e.value.code == 102


def test_damaged_files_are_refused(tmp_path):
    A, G = _op(False, (5, 1e-2))
    f = str(tmp_path / "op.pgs")
    G.save(f)
    data = bytearray(open(f, "rb").read())
    flipped = bytearray(data)
    flipped[len(data) // 2] ^= 0x10
    _corrupt(tmp_path, bytes(flipped), "flip.pgs")
    _corrupt(tmp_path, bytes(data[:-9]), "short.pgs")
    _corrupt(tmp_path, bytes(data) + b"\0", "long.pgs")
    _corrupt(tmp_path, b"XGSGMSLR" + bytes(data[8:]), "magic.pgs")
    version = bytearray(data)
    struct.pack_into("<I", version, 8, 2)
    _corrupt(tmp_path, bytes(version), "version.pgs")
    _corrupt(tmp_path, b"", "empty.pgs")
    with pytest.raises(pg.PgsError) as e:
        pg.Gemslr.load(str(tmp_path / "missing.pgs"))
    assert e.value.code == 102


def _resum(data):
    data[-8:] = struct.pack("<Q", _fnv1a(bytes(data[:-8])))
    return bytes(data)


def test_structurally_invalid_files_are_refused(tmp_path):
    """A file whose checksum is right but whose content would send a kernel
    out of bounds is refused: a repeated permutation entry, a column index
    outside the matrix."""
    A, G = _op(False, None)
    f = str(tmp_path / "op.pgs")
    G.save(f)
    data = bytearray(open(f, "rb").read())
    n, nnz = A.n, A.nnz
    off_a = HEADER + PARAMS + SCALARS          # csr A: nrows, ncols, vec ia, vec ja, vec a
    off_ja = off_a + 8 + 8 + 4 * (n + 1) + 8  # first ja entry
    off_perm = off_a + 8 + (8 + 4 * (n + 1)) + (8 + 4 * nnz) + (8 + 8 * nnz) + 8
    assert struct.unpack_from("<i", data, off_ja)[0] == A.ja[0]
    assert struct.unpack_from("<i", data, off_perm)[0] == G.perm()[0]
    bad_perm = bytearray(data)
    struct.pack_into("<i", bad_perm, off_perm + 4, G.perm()[0])
    _corrupt(tmp_path, _resum(bad_perm), "perm.pgs")
    bad_col = bytearray(data)
    struct.pack_into("<i", bad_col, off_ja, n)
    _corrupt(tmp_path, _resum(bad_col), "col.pgs")
    # the untouched file with a recomputed checksum still loads
    pg.Gemslr.load(str(tmp_path / "op.pgs"))


def test_file_from_a_read_matrix(tmp_path):
    """MatrixMarket in, operator file out, operator file in: the apply of the
    reloaded operator equals the oracle on the original setup."""
    A = pg.read_mm(os.path.join(P.GOLDEN, "gr_30_30.mtx"))
    G = P.setup(A, 2, 4, 1e-3, 20, 8)
    f = str(tmp_path / "gr.pgs")
    G.save(f)
    H = pg.Gemslr.load(f)
    v = pg.rng_normal(5, A.n)
    ref = P.oracle_of(G).apply(v)
    z = H.host_apply(v)
    assert np.linalg.norm(z - ref) <= 1e-13 * np.linalg.norm(ref)


def test_save_to_unwritable_path(tmp_path):
    A, G = _op(False, None)
    with pytest.raises(pg.PgsError) as e:
        G.save(str(tmp_path / "no" / "such" / "dir" / "op.pgs"))
    assert e.value.code == 102
