"""Single-vector operations and the stand-alone device CSR of the class
surface (pgs_vec_*, pgs_dev_csr_*), through the C ABI, against numpy and the
oracle's CSR product: Axpy / Scale / Fill exact, Dot and Norm2 to rounding
(vectorops.cpp:93-146, 633-680), MatVec with alpha and beta as
CsrMatrixClass::MatVec (csr_matrix.cpp:1340-1365).  Real and complex."""
import ctypes as C

import numpy as np
import pytest

import pargemslr_amd as pg
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _pair(v):
    v = complex(v)
    return (C.c_double * 2)(v.real, v.imag)


def _vec(rng, n, cplx):
    x = rng.standard_normal(n)
    return x + 1j * rng.standard_normal(n) if cplx else x


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
@pytest.mark.parametrize("n", [0, 1, 1000, 262144 + 17])
def test_vector_ops(gpu, cplx, n):
    L = pg.lib()
    rng = np.random.default_rng(n + 3)
    x, y = _vec(rng, n, cplx), _vec(rng, n, cplx)
    xd, yd = pg.DeviceArray.from_numpy(x), pg.DeviceArray.from_numpy(y)
    alpha = (0.7 - 0.2j) if cplx else -1.25
    rc = L.pgs_vec_axpy(0, n, int(cplx), _pair(alpha), xd.ptr, yd.ptr, None)
    assert rc == 0, (rc, L.pgs_last_hip_error())
    pg.synchronize()
    assert np.array_equal(yd.numpy(), y + alpha * x) or np.allclose(yd.numpy(), y + alpha * x, rtol=0, atol=1e-15)
    y = y + alpha * x
    r = (C.c_double * 2)()
    assert L.pgs_vec_dot(0, n, int(cplx), xd.ptr, yd.ptr, r, None) == 0
    ref = np.vdot(x, y) if n else 0.0
    got = complex(r[0], r[1]) if cplx else r[0]
    assert abs(got - ref) <= 1e-12 * max(1.0, np.linalg.norm(x) * np.linalg.norm(y))
    r2 = (C.c_double * 2)()
    assert L.pgs_vec_dot(0, n, int(cplx), xd.ptr, yd.ptr, r2, None) == 0
    assert r2[0] == r[0] and r2[1] == r[1]  # fixed summation order
    assert L.pgs_vec_scale(0, n, int(cplx), _pair(alpha), xd.ptr, None) == 0
    pg.synchronize()
    assert np.allclose(xd.numpy(), alpha * x, rtol=0, atol=1e-15)
    assert L.pgs_vec_fill(0, n, int(cplx), _pair(alpha), xd.ptr, None) == 0
    pg.synchronize()
    assert np.all(xd.numpy() == alpha)


def test_vector_ops_reject_bad_arguments(gpu):
    L = pg.lib()
    assert L.pgs_vec_axpy(0, 10, 0, _pair(1.0), None, None, None) == pg.capi.PGS_ERROR_INVALID_PARAM
    assert L.pgs_vec_dot(0, -1, 0, None, None, (C.c_double * 2)(), None) == pg.capi.PGS_ERROR_INVALID_PARAM


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_device_csr_matvec(gpu, cplx):
    L = pg.lib()
    A = pg.laplacian(20, 20, 20, shift=(0.2 + 0.02j) if cplx else 0.3, alphas=(0.05, -0.02, 0.01), complex_=cplx)
    rng = np.random.default_rng(1)
    x, y0 = _vec(rng, A.n, cplx), _vec(rng, A.n, cplx)
    a = A.a.view(np.float64) if cplx else A.a
    h = C.c_void_p()
    assert L.pgs_dev_csr_create(0, A.n, pg.capi._iptr(A.ia), pg.capi._iptr(A.ja),
                                a.ctypes.data_as(C.POINTER(C.c_double)), int(cplx), C.byref(h)) == 0
    try:
        xd, yd = pg.DeviceArray.from_numpy(x), pg.DeviceArray.from_numpy(y0)
        alpha, beta = (-0.5 + 0.25j, 2.0 - 1j) if cplx else (-0.5, 2.0)
        assert L.pgs_dev_csr_spmv(h, _pair(alpha), xd.ptr, _pair(beta), yd.ptr, None) == 0
        pg.synchronize()
        ref = orc.csr_matvec(A.ia, A.ja, A.a, x, alpha=alpha, beta=beta, y=y0) if not cplx else \
            beta * y0 + alpha * A.matvec(x)
        assert np.linalg.norm(yd.numpy() - ref) <= 1e-14 * np.linalg.norm(ref)
    finally:
        L.pgs_dev_csr_destroy(h)
    bad = np.array(A.ja, dtype=np.int32)
    bad[5] = A.n
    h2 = C.c_void_p()
    assert L.pgs_dev_csr_create(0, A.n, pg.capi._iptr(A.ia), pg.capi._iptr(bad),
                                a.ctypes.data_as(C.POINTER(C.c_double)), int(cplx), C.byref(h2)) == \
        pg.capi.PGS_ERROR_INVALID_PARAM


def test_device_array_copies_on_a_stream(gpu):
    """DeviceArray's upload, zero and numpy take a Stream object as well as a
    raw stream handle."""
    s = pg.Stream()
    try:
        x = np.arange(1000, dtype=np.float64)
        d = pg.DeviceArray((1000,))
        d.upload(x, stream=s)
        s.synchronize()
        assert np.array_equal(d.numpy(stream=s), x)
        d.zero(stream=s.handle)
        s.synchronize()
        assert not d.numpy().any()
    finally:
        s.close()
