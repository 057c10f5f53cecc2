"""ctypes bindings of the pargemslr_amd C ABI (include/pargemslr_amd.h).

Host arrays are numpy; device arrays (DeviceArray) and streams (Stream) come
from the library's own HIP runtime (pgs_dev_malloc, pgs_stream_create), so
no second HIP runtime shares the process.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpargemslr_amd.so")

PGS_IO_SIZE = 180
PGS_PREC_FP64 = 1
PGS_PREC_FP32 = 2
PGS_ORTH_MGS = 0
PGS_ORTH_CGS2 = 1
PGS_ORTH_DCGS2 = 2
PGS_ERROR_INVALID_PARAM = 101
PGS_ERR_HIP = 200
PGS_ERR_BREAKDOWN = 203
PGS_ERR_NOT_CONVERGED = 204
PGS_ERR_NONFINITE = 206
PGS_ERR_SCRATCH_LIMIT = 207
PGS_APPLY_NOALLOC_SCRATCH = 64 * 1024 * 1024

# parameter indices (reference PARGEMSLR_IO_* numbering)
IO = dict(
    SOLVER_KDIM=11, SOLVER_MAXITS=12, SOLVER_TOL=13, SOLVER_ATOL=14,
    NLEV=20, NCOMP=22, PARTITION=24, MINSEP=26, KMIN=27, KFACTOR=29, VTXSEP=31,
    DROPTOL_B=40, DROPTOL_C=44, ROWNNZ_B=48, ROWNNZ_C=50, PERM_OPTION=58,
    LR_OPTION1=70, LR_OPTION2=71, LR_RANK1=74, LR_RANK2=75, LR_RANK_FACTOR1=78, LR_RANK_FACTOR2=79,
    LR_ARNOLDI_FACTOR1=82, LR_ARNOLDI_FACTOR2=83, LR_TOL_EIG1=86, LR_TOL_EIG2=87,
    LR_MAXITS1=90, LR_MAXITS2=91, LR_TR_FACTOR=94, LR_TOL_ORTH=95, LR_RAND_INIT=97,
    SCHUR_ENABLE=130, SCHUR_ITER_TOL=131, SCHUR_MAXITS=132, GLOBAL_SOLVE=141,
)


class PgsError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        super().__init__(f"{what}: {strerror(code)} ({code})")


class GemslrInfo(C.Structure):
    _fields_ = [("n", C.c_int), ("nlev", C.c_int), ("is_complex", C.c_int), ("nnz_a", C.c_int64),
                ("nnz_ilu", C.c_int64), ("nnz_lr", C.c_int64), ("setup_seconds", C.c_double * 4),
                ("schur_maxits", C.c_int), ("schur_tol", C.c_double), ("global_solve", C.c_int)]


class LevelView(C.Structure):
    _fields_ = [("lev_start", C.c_int), ("nloc", C.c_int), ("nrem", C.c_int), ("ncomp", C.c_int), ("rank", C.c_int),
                ("dom_ptr", C.POINTER(C.c_int)),
                ("L_ia", C.POINTER(C.c_int)), ("L_ja", C.POINTER(C.c_int)),
                ("U_ia", C.POINTER(C.c_int)), ("U_ja", C.POINTER(C.c_int)),
                ("L_a", C.POINTER(C.c_double)), ("U_a", C.POINTER(C.c_double)), ("Dinv", C.POINTER(C.c_double)),
                ("E_ia", C.POINTER(C.c_int)), ("E_ja", C.POINTER(C.c_int)),
                ("F_ia", C.POINTER(C.c_int)), ("F_ja", C.POINTER(C.c_int)),
                ("E_a", C.POINTER(C.c_double)), ("F_a", C.POINTER(C.c_double)),
                ("W", C.POINTER(C.c_double)), ("WH", C.POINTER(C.c_double)),
                ("C_ia", C.POINTER(C.c_int)), ("C_ja", C.POINTER(C.c_int)), ("C_a", C.POINTER(C.c_double))]


class FgmresOpts(C.Structure):
    _fields_ = [("kdim", C.c_int), ("maxits", C.c_int), ("tol", C.c_double), ("absolute_tol", C.c_int),
                ("prec", C.c_int), ("orth", C.c_int), ("use_precond", C.c_int), ("check_every", C.c_int),
                ("use_graph", C.c_int)]


class FgmresResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iters", C.c_int), ("rel_res", C.c_double)]


class KernelProfile(C.Structure):
    _fields_ = [("launches", C.c_int64), ("total_ms", C.c_double), ("total_bytes", C.c_double)]


class ServerOpts(C.Structure):
    _fields_ = [("max_batch", C.c_int), ("slots", C.c_int), ("mem_reserve", C.c_int64), ("kdim", C.c_int),
                ("maxits", C.c_int), ("tol", C.c_double), ("prec", C.c_int), ("orth", C.c_int),
                ("check_every", C.c_int), ("use_graph", C.c_int)]


class RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("nrhs", C.c_int), ("batch_nrhs", C.c_int), ("queue_ms", C.c_double),
                ("solve_ms", C.c_double), ("total_ms", C.c_double), ("joined_cycle", C.c_int)]


# every extern "C" entry point of include/pargemslr_amd.h
EXPORTED = [
    "pgs_params_default", "pgs_gemslr_setup", "pgs_gemslr_setup_device", "pgs_gemslr_destroy", "pgs_gemslr_info", "pgs_gemslr_level_view",
    "pgs_gemslr_perm", "pgs_gemslr_host_apply", "pgs_device_count", "pgs_dev_op_create", "pgs_dev_op_destroy",
    "pgs_dev_op_bytes", "pgs_ws_create", "pgs_ws_destroy", "pgs_spmv", "pgs_apply", "pgs_fgmres", "pgs_strerror",
    "pgs_last_hip_error", "pgs_gen_laplacian", "pgs_gen_convdiff", "pgs_free", "pgs_rng_uniform", "pgs_rng_normal",
    "pgs_ws_profile", "pgs_ws_profile_read", "pgs_dev_malloc", "pgs_dev_free", "pgs_memcpy_htod", "pgs_memcpy_dtoh",
    "pgs_memset_zero", "pgs_stream_create", "pgs_stream_destroy", "pgs_stream_synchronize", "pgs_device_synchronize",
    "pgs_memcpy_dtod", "pgs_dev_op_dims", "pgs_ws_estimate", "pgs_server_create", "pgs_server_add_operator", "pgs_server_submit",
    "pgs_server_wait", "pgs_server_destroy", "pgs_gemslr_matrix", "pgs_gemslr_save", "pgs_gemslr_load",
    "pgs_read_mm_csr", "pgs_read_mm_vector", "pgs_params_read_inputs", "pgs_hypre_gemslr_create",
    "pgs_hypre_gemslr_params", "pgs_hypre_gemslr_set_precision", "pgs_hypre_gemslr_setup", "pgs_hypre_gemslr_solve",
    "pgs_hypre_gemslr_destroy", "pgs_vec_axpy", "pgs_vec_scale", "pgs_vec_fill", "pgs_vec_dot",
    "pgs_dev_csr_create", "pgs_dev_csr_destroy", "pgs_dev_csr_spmv", "pgs_apply_noalloc", "pgs_ws_profile_families",
]

# kernel families of pgs_ws_profile_families, by index
PROFILE_FAMILIES = ["trsv level 0", "trsv levels 1+", "spmv A", "spmv E/F", "low-rank", "permute", "orth dots",
                    "orth update", "scale", "x update", "state"]

_lib = None


def lib():
    """Loads the in-tree shared library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m pargemslr_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    L.pgs_params_default.argtypes = [dp]
    L.pgs_gemslr_setup.argtypes = [C.c_int, ip, ip, dp, C.c_int, dp, C.c_uint64, C.POINTER(vp)]
    L.pgs_gemslr_setup_device.argtypes = [C.c_int, ip, ip, dp, C.c_int, dp, C.c_uint64, C.c_int, C.POINTER(vp)]
    L.pgs_gemslr_destroy.argtypes = [vp]
    L.pgs_gemslr_info.argtypes = [vp, C.POINTER(GemslrInfo)]
    L.pgs_gemslr_level_view.argtypes = [vp, C.c_int, C.POINTER(LevelView)]
    L.pgs_gemslr_perm.argtypes = [vp, C.POINTER(ip)]
    L.pgs_gemslr_host_apply.argtypes = [vp, dp, dp]
    L.pgs_gemslr_matrix.argtypes = [vp, ip, C.POINTER(ip), C.POINTER(ip), C.POINTER(dp)]
    L.pgs_gemslr_save.argtypes = [vp, C.c_char_p]
    L.pgs_gemslr_load.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.pgs_read_mm_csr.argtypes = [C.c_char_p, C.c_int, C.c_int, ip, ip, C.POINTER(C.c_int64), C.POINTER(ip),
                                  C.POINTER(ip), C.POINTER(dp)]
    L.pgs_read_mm_vector.argtypes = [C.c_char_p, C.c_int, C.c_int, ip, C.POINTER(dp)]
    L.pgs_params_read_inputs.argtypes = [C.c_char_p, C.c_int, dp]
    L.pgs_hypre_gemslr_create.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.pgs_hypre_gemslr_params.argtypes = [vp, dp]
    L.pgs_hypre_gemslr_set_precision.argtypes = [vp, C.c_int]
    L.pgs_hypre_gemslr_setup.argtypes = [vp, C.c_int, ip, ip, dp, C.c_int]
    L.pgs_hypre_gemslr_solve.argtypes = [vp, dp, dp]
    L.pgs_hypre_gemslr_destroy.argtypes = [vp]
    L.pgs_device_count.argtypes = [ip]
    L.pgs_dev_op_create.argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp)]
    L.pgs_dev_op_destroy.argtypes = [vp]
    L.pgs_dev_op_bytes.argtypes = [vp, C.POINTER(C.c_int64)]
    L.pgs_ws_create.argtypes = [vp, C.c_int, C.c_int, C.POINTER(vp)]
    L.pgs_ws_destroy.argtypes = [vp]
    L.pgs_spmv.argtypes = [vp, vp, dp, vp, dp, vp, C.c_int, vp]
    L.pgs_apply.argtypes = [vp, vp, vp, vp, C.c_int, C.c_int, vp]
    L.pgs_fgmres.argtypes = [vp, vp, vp, vp, C.c_int, C.POINTER(FgmresOpts), C.POINTER(FgmresResult), vp]
    L.pgs_strerror.argtypes = [C.c_int]
    L.pgs_strerror.restype = C.c_char_p
    L.pgs_last_hip_error.restype = C.c_char_p
    L.pgs_gen_laplacian.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_int, ip, C.POINTER(C.c_int64), C.POINTER(ip), C.POINTER(ip),
                                    C.POINTER(dp)]
    L.pgs_gen_convdiff.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, ip, C.POINTER(C.c_int64), C.POINTER(ip),
                                   C.POINTER(ip), C.POINTER(dp)]
    L.pgs_free.argtypes = [vp]
    L.pgs_rng_uniform.argtypes = [C.c_uint64, C.c_int64, C.c_double, C.c_double, dp]
    L.pgs_rng_normal.argtypes = [C.c_uint64, C.c_int64, dp]
    L.pgs_ws_profile.argtypes = [vp, C.c_int]
    L.pgs_ws_profile_read.argtypes = [vp, C.POINTER(KernelProfile)]
    L.pgs_dev_malloc.argtypes = [C.c_int, C.c_int64, C.POINTER(vp)]
    L.pgs_dev_free.argtypes = [C.c_int, vp]
    L.pgs_memcpy_htod.argtypes = [C.c_int, vp, vp, C.c_int64, vp]
    L.pgs_memcpy_dtoh.argtypes = [C.c_int, vp, vp, C.c_int64, vp]
    L.pgs_memcpy_dtod.argtypes = [C.c_int, vp, vp, C.c_int64, vp]
    L.pgs_memset_zero.argtypes = [C.c_int, vp, C.c_int64, vp]
    L.pgs_stream_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.pgs_stream_destroy.argtypes = [C.c_int, vp]
    L.pgs_stream_synchronize.argtypes = [C.c_int, vp]
    L.pgs_device_synchronize.argtypes = [C.c_int]
    L.pgs_dev_op_dims.argtypes = [vp, ip, ip, ip]
    L.pgs_ws_estimate.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int64)]
    L.pgs_server_create.argtypes = [C.c_int, C.POINTER(ServerOpts), C.POINTER(vp)]
    L.pgs_server_add_operator.argtypes = [vp, vp, ip]
    L.pgs_server_submit.argtypes = [vp, C.c_int, C.c_int, dp, dp, ip, C.POINTER(C.c_int64)]
    L.pgs_server_wait.argtypes = [vp, C.c_int64, C.POINTER(RequestInfo)]
    L.pgs_server_destroy.argtypes = [vp]
    L.pgs_vec_axpy.argtypes = [C.c_int, C.c_int64, C.c_int, dp, vp, vp, vp]
    L.pgs_vec_scale.argtypes = [C.c_int, C.c_int64, C.c_int, dp, vp, vp]
    L.pgs_vec_fill.argtypes = [C.c_int, C.c_int64, C.c_int, dp, vp, vp]
    L.pgs_vec_dot.argtypes = [C.c_int, C.c_int64, C.c_int, vp, vp, dp, vp]
    L.pgs_dev_csr_create.argtypes = [C.c_int, C.c_int, ip, ip, dp, C.c_int, C.POINTER(vp)]
    L.pgs_dev_csr_destroy.argtypes = [vp]
    L.pgs_dev_csr_spmv.argtypes = [vp, dp, vp, dp, vp, vp]
    L.pgs_apply_noalloc.argtypes = [vp, vp, vp, C.c_int, C.c_int, vp]
    L.pgs_ws_profile_families.argtypes = [vp, C.POINTER(KernelProfile)]
    _lib = L
    return L


def strerror(code):
    try:
        return lib().pgs_strerror(code).decode()
    except Exception:
        return "unknown"


def _chk(code, what):
    if code != 0:
        if code == PGS_ERR_HIP:
            raise PgsError(code, f"{what} [{lib().pgs_last_hip_error().decode(errors='replace')}]")
        raise PgsError(code, what)


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_params():
    p = np.zeros(PGS_IO_SIZE, dtype=np.float64)
    _chk(lib().pgs_params_default(_dptr(p)), "pgs_params_default")
    return p


# ---------------------------------------------------------------- generators
class CsrHost:
    """0-based CSR on the host (numpy)."""

    def __init__(self, n, ia, ja, a):
        self.n = int(n)
        self.ia = np.ascontiguousarray(ia, dtype=np.int32)
        self.ja = np.ascontiguousarray(ja, dtype=np.int32)
        self.a = np.ascontiguousarray(a)
        self.is_complex = np.iscomplexobj(self.a)

    @property
    def nnz(self):
        return int(self.ia[-1])

    def matvec(self, x):
        """y = A x (numpy, for residual checks)."""
        rows = np.repeat(np.arange(self.n), np.diff(self.ia))
        y = np.zeros(x.shape, dtype=np.result_type(self.a, x))
        prod = self.a[:, None] * x[self.ja] if x.ndim == 2 else self.a * x[self.ja]
        np.add.at(y, rows, prod)
        return y

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.a, self.ja, self.ia), shape=(self.n, self.n))


def _take_csr(n, nnz, ia, ja, a, is_complex):
    L = lib()
    n = n.value
    nnz = nnz.value
    ia_np = np.ctypeslib.as_array(ia, shape=(n + 1,)).copy()
    ja_np = np.ctypeslib.as_array(ja, shape=(max(nnz, 1),)).copy()[:nnz]
    if is_complex:
        a_np = np.ctypeslib.as_array(a, shape=(max(2 * nnz, 2),)).copy()[: 2 * nnz].view(np.complex128)
    else:
        a_np = np.ctypeslib.as_array(a, shape=(max(nnz, 1),)).copy()[:nnz]
    for p in (ia, ja, a):
        L.pgs_free(C.cast(p, C.c_void_p))
    return CsrHost(n, ia_np, ja_np, a_np)


def read_mm(path, complex_=False, idxin=1):
    """MatrixMarket coordinate file into a square CsrHost
    (CsrMatrixClass::ReadFromMMFile rules; complex_ selects the complexd
    reader)."""
    n, m, nnz = C.c_int(), C.c_int(), C.c_int64()
    ia, ja, a = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_double)()
    _chk(lib().pgs_read_mm_csr(os.fsencode(path), idxin, int(complex_), C.byref(n), C.byref(m), C.byref(nnz),
                               C.byref(ia), C.byref(ja), C.byref(a)), "pgs_read_mm_csr")
    if n.value != m.value:
        for p in (ia, ja, a):
            lib().pgs_free(C.cast(p, C.c_void_p))
        raise PgsError(101, f"read_mm: {path} is {n.value} x {m.value}, not square")
    return _take_csr(n, nnz, ia, ja, a, complex_)


PGS_INPUTS_SEQUENTIAL = 0
PGS_INPUTS_HYPRE = 1


def read_inputs(path, fmt=PGS_INPUTS_SEQUENTIAL, params=None):
    """Parameter array from one of the reference drivers' `inputs` files
    (pgs_params_read_inputs), starting from `params` or the defaults."""
    p = default_params() if params is None else np.array(params, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    _chk(lib().pgs_params_read_inputs(os.fsencode(path), fmt, _dptr(p)), "pgs_params_read_inputs")
    return p


class HypreGemslr:
    """The hypre-style preconditioner object (pgs_hypre_gemslr_*): create from
    an inputs file, setup on a CSR matrix, solve = one apply on host arrays."""

    def __init__(self, inputs_file=None):
        self._h = C.c_void_p()
        _chk(lib().pgs_hypre_gemslr_create(None if inputs_file is None else os.fsencode(inputs_file),
                                           C.byref(self._h)), "pgs_hypre_gemslr_create")

    def params(self):
        p = np.zeros(180)
        _chk(lib().pgs_hypre_gemslr_params(self._h, _dptr(p)), "pgs_hypre_gemslr_params")
        return p

    def set_precision(self, prec):
        _chk(lib().pgs_hypre_gemslr_set_precision(self._h, prec), "pgs_hypre_gemslr_set_precision")

    def setup(self, A, device=0):
        self.n = A.n
        _chk(lib().pgs_hypre_gemslr_setup(self._h, A.n, _iptr(A.ia), _iptr(A.ja), _dptr(np.ascontiguousarray(A.a)),
                                          device), "pgs_hypre_gemslr_setup")

    def solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        _chk(lib().pgs_hypre_gemslr_solve(self._h, _dptr(b), _dptr(x)), "pgs_hypre_gemslr_solve")
        return x

    def close(self):
        if self._h:
            lib().pgs_hypre_gemslr_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_mm_vector(path, complex_=False, idxin=1):
    """MatrixMarket coordinate vector file into a dense numpy vector
    (SequentialVectorClass::ReadFromMMFile rules)."""
    n = C.c_int()
    v = C.POINTER(C.c_double)()
    _chk(lib().pgs_read_mm_vector(os.fsencode(path), idxin, int(complex_), C.byref(n), C.byref(v)),
         "pgs_read_mm_vector")
    nd = 2 if complex_ else 1
    out = np.ctypeslib.as_array(v, shape=(max(n.value, 1) * nd,)).copy()[: n.value * nd]
    lib().pgs_free(C.cast(v, C.c_void_p))
    return out.view(np.complex128) if complex_ else out


def laplacian(nx, ny, nz, shift=0.0, alphas=(0.0, 0.0, 0.0), complex_=False):
    n = C.c_int()
    nnz = C.c_int64()
    ia = C.POINTER(C.c_int)()
    ja = C.POINTER(C.c_int)()
    a = C.POINTER(C.c_double)()
    sh = complex(shift)
    _chk(lib().pgs_gen_laplacian(nx, ny, nz, sh.real, sh.imag, alphas[0], alphas[1], alphas[2], int(complex_),
                                 C.byref(n), C.byref(nnz), C.byref(ia), C.byref(ja), C.byref(a)), "pgs_gen_laplacian")
    return _take_csr(n, nnz, ia, ja, a, complex_)


def convdiff(nx, ny, nz, c):
    n = C.c_int()
    nnz = C.c_int64()
    ia = C.POINTER(C.c_int)()
    ja = C.POINTER(C.c_int)()
    a = C.POINTER(C.c_double)()
    _chk(lib().pgs_gen_convdiff(nx, ny, nz, c, C.byref(n), C.byref(nnz), C.byref(ia), C.byref(ja), C.byref(a)),
         "pgs_gen_convdiff")
    return _take_csr(n, nnz, ia, ja, a, False)


def rng_uniform(seed, n, lo=-1.0, hi=1.0):
    out = np.empty(n, dtype=np.float64)
    _chk(lib().pgs_rng_uniform(seed, n, lo, hi, _dptr(out)), "pgs_rng_uniform")
    return out


def rng_normal(seed, n):
    out = np.empty(n, dtype=np.float64)
    _chk(lib().pgs_rng_normal(seed, n, _dptr(out)), "pgs_rng_normal")
    return out


# ---------------------------------------------------------------- host setup
class Gemslr:
    """Host-resident GeMSLR operator (GemslrClass after Setup)."""

    def __init__(self, A, params, seed=0, lowrank_device=None):
        """lowrank_device: None builds the low-rank corrections on the host,
        a device index builds them on that device (pgs_gemslr_setup_device)."""
        self.A = A
        self._h = C.c_void_p()
        params = np.ascontiguousarray(params, dtype=np.float64)
        self.params = params.copy()
        a = A.a.view(np.float64) if A.is_complex else A.a
        a = np.ascontiguousarray(a)
        if lowrank_device is None:
            _chk(lib().pgs_gemslr_setup(A.n, _iptr(A.ia), _iptr(A.ja), _dptr(a), int(A.is_complex),
                                        _dptr(params), C.c_uint64(seed), C.byref(self._h)), "pgs_gemslr_setup")
        else:
            _chk(lib().pgs_gemslr_setup_device(A.n, _iptr(A.ia), _iptr(A.ja), _dptr(a), int(A.is_complex),
                                               _dptr(params), C.c_uint64(seed), int(lowrank_device),
                                               C.byref(self._h)), "pgs_gemslr_setup_device")

    @classmethod
    def load(cls, path):
        """Operator read from a file written by save() (no setup is run); its
        matrix comes from the file too."""
        self = cls.__new__(cls)
        self._h = C.c_void_p()
        self.params = None
        _chk(lib().pgs_gemslr_load(os.fsencode(path), C.byref(self._h)), "pgs_gemslr_load")
        n = C.c_int()
        ia, ja, a = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_double)()
        _chk(lib().pgs_gemslr_matrix(self._h, C.byref(n), C.byref(ia), C.byref(ja), C.byref(a)), "pgs_gemslr_matrix")
        n = n.value
        ia_np = np.ctypeslib.as_array(ia, shape=(n + 1,)).copy()
        nnz = int(ia_np[-1])
        ja_np = np.ctypeslib.as_array(ja, shape=(max(nnz, 1),)).copy()[:nnz]
        cplx = bool(self.info().is_complex)
        nd = 2 if cplx else 1
        a_np = np.ctypeslib.as_array(a, shape=(max(nnz, 1) * nd,)).copy()[: nnz * nd]
        self.A = CsrHost(n, ia_np, ja_np, a_np.view(np.complex128) if cplx else a_np)
        return self

    def save(self, path):
        """Writes the operator (the whole setup output) to path."""
        _chk(lib().pgs_gemslr_save(self._h, os.fsencode(path)), "pgs_gemslr_save")

    def schur(self):
        """(maxits, tol) of the inner Schur iteration, or None when it is off."""
        inf = self.info()
        if inf.schur_maxits == 0:
            return None
        return int(inf.schur_maxits), float(inf.schur_tol)

    def usolve(self):
        """True when the apply uses the U variant of SolveLevel (no block L
        solve on the way down; PARGEMSLR_IO_ADVANCED_GLOBAL_SOLVE = 1)."""
        return self.info().global_solve == 1

    def close(self):
        if self._h:
            lib().pgs_gemslr_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def info(self):
        inf = GemslrInfo()
        _chk(lib().pgs_gemslr_info(self._h, C.byref(inf)), "pgs_gemslr_info")
        return inf

    def level_view(self, l):
        v = LevelView()
        _chk(lib().pgs_gemslr_level_view(self._h, l, C.byref(v)), "pgs_gemslr_level_view")
        return v

    def perm(self):
        p = C.POINTER(C.c_int)()
        _chk(lib().pgs_gemslr_perm(self._h, C.byref(p)), "pgs_gemslr_perm")
        return np.ctypeslib.as_array(p, shape=(self.A.n,)).copy()

    def levels(self):
        """Level structure as numpy arrays (copies; complex128 values for a
        complex operator)."""
        out = []
        cplx = bool(self.A.is_complex)
        nd = 2 if cplx else 1

        def vals(p, count):
            a = np.ctypeslib.as_array(p, shape=(max(count, 1) * nd,)).copy()[: count * nd]
            return a.view(np.complex128) if cplx else a

        nlev = self.info().nlev
        for l in range(nlev):
            v = self.level_view(l)
            d = dict(lev_start=v.lev_start, nloc=v.nloc, nrem=v.nrem, ncomp=v.ncomp, rank=v.rank)
            d["dom_ptr"] = np.ctypeslib.as_array(v.dom_ptr, shape=(v.ncomp + 1,)).copy()
            for nm, nr in (("L", v.nloc), ("U", v.nloc), ("E", v.nrem), ("F", v.nloc)):
                ia = getattr(v, nm + "_ia")
                if nm in ("E", "F") and l == nlev - 1:
                    d[nm] = None
                    continue
                ia_np = np.ctypeslib.as_array(ia, shape=(nr + 1,)).copy()
                nnz = int(ia_np[-1])
                ja_np = np.ctypeslib.as_array(getattr(v, nm + "_ja"), shape=(max(nnz, 1),)).copy()[:nnz] if nnz else np.zeros(0, np.int32)
                a_np = vals(getattr(v, nm + "_a"), nnz) if nnz else np.zeros(0, np.complex128 if cplx else np.float64)
                d[nm] = (ia_np, ja_np, a_np)
            d["Dinv"] = vals(v.Dinv, v.nloc)
            if bool(v.C_ia):
                cia = np.ctypeslib.as_array(v.C_ia, shape=(v.nrem + 1,)).copy()
                cnnz = int(cia[-1])
                cja = np.ctypeslib.as_array(v.C_ja, shape=(max(cnnz, 1),)).copy()[:cnnz]
                d["C"] = (cia, cja, vals(v.C_a, cnnz))
            else:
                d["C"] = None
            if v.rank > 0:
                d["W"] = vals(v.W, v.rank * v.nrem).reshape(v.rank, v.nrem).T.copy()
                d["WH"] = vals(v.WH, v.rank * v.nrem).reshape(v.rank, v.nrem).T.copy()
            out.append(d)
        return out

    def host_apply(self, v):
        dt = np.complex128 if self.A.is_complex else np.float64
        v = np.ascontiguousarray(v, dtype=dt)
        z = np.empty_like(v)
        _chk(lib().pgs_gemslr_host_apply(self._h, _dptr(v.view(np.float64)), _dptr(z.view(np.float64))),
             "pgs_gemslr_host_apply")
        return z


# ---------------------------------------------------------------- device
def device_count():
    c = C.c_int()
    rc = lib().pgs_device_count(C.byref(c))
    return c.value if rc == 0 else 0


def synchronize(device=0):
    _chk(lib().pgs_device_synchronize(device), "pgs_device_synchronize")


class DeviceArray:
    """A dense float64/float32 array in device memory (row-major, n x nrhs
    blocks are row-interleaved as the C ABI expects)."""

    def __init__(self, shape, dtype=np.float64, device=0, zero=True):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.device = device
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _chk(lib().pgs_dev_malloc(device, self.nbytes, C.byref(p)), "pgs_dev_malloc")
        self._p = p
        if zero:
            self.zero()

    @classmethod
    def from_numpy(cls, a, device=0):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, device, zero=False)
        d.upload(a)
        return d

    @property
    def ptr(self):
        return self._p

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        _chk(lib().pgs_memcpy_htod(self.device, self._p, C.c_void_p(a.ctypes.data), self.nbytes, _sh(stream)),
             "pgs_memcpy_htod")

    def numpy(self, stream=None):
        out = np.empty(self.shape, dtype=self.dtype)
        _chk(lib().pgs_memcpy_dtoh(self.device, C.c_void_p(out.ctypes.data), self._p, self.nbytes, _sh(stream)),
             "pgs_memcpy_dtoh")
        return out

    def zero(self, stream=None):
        _chk(lib().pgs_memset_zero(self.device, self._p, self.nbytes, _sh(stream)), "pgs_memset_zero")

    def close(self):
        if self._p:
            lib().pgs_dev_free(self.device, self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    def __init__(self, device=0):
        self.device = device
        s = C.c_void_p()
        _chk(lib().pgs_stream_create(device, C.byref(s)), "pgs_stream_create")
        self.handle = s

    def synchronize(self):
        _chk(lib().pgs_stream_synchronize(self.device, self.handle), "pgs_stream_synchronize")

    def close(self):
        if self.handle:
            lib().pgs_stream_destroy(self.device, self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _tptr(t):
    """Device pointer of a DeviceArray (or of any object exposing data_ptr())."""
    if isinstance(t, DeviceArray):
        return t.ptr
    return C.c_void_p(t.data_ptr())


def _sh(stream):
    if stream is None:
        return None
    return stream.handle if isinstance(stream, Stream) else stream


class DevOp:
    """Operator resident in device memory."""

    def __init__(self, gem, device=0, prec_mask=PGS_PREC_FP64 | PGS_PREC_FP32):
        self.gem = gem
        self.n = gem.A.n
        self.is_complex = bool(gem.A.is_complex)
        self.dtype = np.dtype(np.complex128 if self.is_complex else np.float64)
        self.device = device
        self._h = C.c_void_p()
        _chk(lib().pgs_dev_op_create(gem.handle, device, prec_mask, C.byref(self._h)), "pgs_dev_op_create")

    def bytes(self):
        b = C.c_int64()
        _chk(lib().pgs_dev_op_bytes(self._h, C.byref(b)), "pgs_dev_op_bytes")
        return b.value

    def apply_noalloc(self, v, z, nrhs=1, prec=PGS_PREC_FP64, stream=None):
        """z = M^{-1} v with no device allocation and no workspace
        (pgs_apply_noalloc; calls on one device must not overlap)."""
        _chk(lib().pgs_apply_noalloc(self._h, _tptr(v), _tptr(z), nrhs, prec, _sh(stream)), "pgs_apply_noalloc")

    def close(self):
        if self._h:
            lib().pgs_dev_op_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Workspace:
    def __init__(self, dop, max_nrhs=1, kdim=50):
        self.dop = dop
        self._h = C.c_void_p()
        _chk(lib().pgs_ws_create(dop._h, max_nrhs, kdim, C.byref(self._h)), "pgs_ws_create")

    def close(self):
        if self._h:
            lib().pgs_ws_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def profile(self, enable=True):
        _chk(lib().pgs_ws_profile(self._h, int(enable)), "pgs_ws_profile")

    def profile_read(self):
        p = KernelProfile()
        _chk(lib().pgs_ws_profile_read(self._h, C.byref(p)), "pgs_ws_profile_read")
        return dict(launches=p.launches, total_ms=p.total_ms, total_bytes=p.total_bytes)

    def profile_families(self):
        """Per kernel family (PROFILE_FAMILIES): launches, ms and algorithmic
        bytes recorded since profile(True)."""
        arr = (KernelProfile * len(PROFILE_FAMILIES))()
        _chk(lib().pgs_ws_profile_families(self._h, arr), "pgs_ws_profile_families")
        return {name: dict(launches=p.launches, total_ms=p.total_ms, total_bytes=p.total_bytes)
                for name, p in zip(PROFILE_FAMILIES, arr)}

    def spmv(self, x, y, alpha=1.0, beta=0.0, nrhs=1, stream=None):
        """y = alpha A x + beta y (alpha, beta complex for a complex operator)."""
        alpha, beta = complex(alpha), complex(beta)
        a = (C.c_double * 2)(alpha.real, alpha.imag)
        b = (C.c_double * 2)(beta.real, beta.imag)
        _chk(lib().pgs_spmv(self.dop._h, self._h, a, _tptr(x), b, _tptr(y), nrhs, _sh(stream)), "pgs_spmv")

    def apply(self, v, z, nrhs=1, prec=PGS_PREC_FP64, stream=None):
        _chk(lib().pgs_apply(self.dop._h, self._h, _tptr(v), _tptr(z), nrhs, prec, _sh(stream)), "pgs_apply")

    def fgmres(self, b, x, nrhs=1, kdim=50, maxits=1000, tol=1e-8, prec=PGS_PREC_FP64, orth=PGS_ORTH_CGS2,
               use_precond=True, absolute=False, check_every=1, use_graph=False, stream=None):
        o = FgmresOpts(kdim, maxits, tol, int(absolute), prec, orth, int(use_precond), check_every, int(use_graph))
        res = (FgmresResult * nrhs)()
        _chk(lib().pgs_fgmres(self.dop._h, self._h, _tptr(b), _tptr(x), nrhs, C.byref(o), res, _sh(stream)),
             "pgs_fgmres")
        return [(r.status, r.iters, r.rel_res) for r in res]


class Server:
    """Request server over resident operators (pgs_server_*): continuous
    batching of solve requests; see include/pargemslr_amd.h."""

    def __init__(self, device=0, max_batch=16, slots=3, mem_reserve=8 << 30, kdim=50, maxits=1000, tol=1e-8,
                 prec=PGS_PREC_FP32, orth=PGS_ORTH_CGS2, check_every=1, use_graph=False):
        self.device = device
        o = ServerOpts(max_batch, slots, mem_reserve, kdim, maxits, tol, prec, orth, check_every, int(use_graph))
        self._h = C.c_void_p()
        self._ops = []
        self._live = {}  # request id -> (b, x, iters) kept alive until waited for
        _chk(lib().pgs_server_create(device, C.byref(o), C.byref(self._h)), "pgs_server_create")

    def add_operator(self, dop):
        i = C.c_int()
        _chk(lib().pgs_server_add_operator(self._h, dop._h, C.byref(i)), "pgs_server_add_operator")
        self._ops.append(dop)
        return i.value

    def submit(self, op_id, B):
        """B: n x nrhs (or n) host array (complex128 for a complex operator);
        returns a request id."""
        dt = self._ops[op_id].dtype if 0 <= op_id < len(self._ops) else np.dtype(np.float64)
        B = np.asarray(B, dtype=dt)
        if B.ndim == 1:
            B = B[:, None]
        n, nr = B.shape
        Bc = np.asfortranarray(B)
        X = np.zeros((n, nr), dtype=dt, order="F")
        it = np.zeros(nr, dtype=np.int32)
        r = C.c_int64()
        _chk(lib().pgs_server_submit(self._h, op_id, nr, Bc.ctypes.data_as(C.POINTER(C.c_double)),
                                     X.ctypes.data_as(C.POINTER(C.c_double)), _iptr(it), C.byref(r)),
             "pgs_server_submit")
        self._live[r.value] = (Bc, X, it)
        return r.value

    def wait(self, req):
        """Returns (X, iterations per column, info dict)."""
        info = RequestInfo()
        _chk(lib().pgs_server_wait(self._h, req, C.byref(info)), "pgs_server_wait")
        _, X, it = self._live.pop(req)
        return X, it, dict(status=info.status, nrhs=info.nrhs, batch_nrhs=info.batch_nrhs, queue_ms=info.queue_ms,
                           solve_ms=info.solve_ms, total_ms=info.total_ms, joined_cycle=info.joined_cycle)

    def close(self):
        if self._h:
            lib().pgs_server_destroy(self._h)
            self._h = C.c_void_p()
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def config_params(nlev, ncomp, droptol, rownnz, rank, kdim=50, maxits=1000, tol=1e-8, lr_tol_eig=1e-3,
                  lr_maxits=10, droptol_c=None, rownnz_c=None, perm=1, schur=None, usolve=False):
    """Reference parameter array for a GeMSLR configuration.  schur: None
    (inner Schur iteration off) or (maxits, tol) of the inner FGMRES on the
    level-0 Schur complement (the reference's defaults are (5, 1e-2))."""
    p = default_params()
    p[IO["NLEV"]] = nlev
    p[IO["NCOMP"]] = ncomp
    p[IO["PARTITION"]] = 1
    p[IO["DROPTOL_B"]] = droptol
    p[IO["DROPTOL_C"]] = droptol if droptol_c is None else droptol_c
    p[IO["ROWNNZ_B"]] = rownnz
    p[IO["ROWNNZ_C"]] = rownnz if rownnz_c is None else rownnz_c
    p[IO["PERM_OPTION"]] = perm
    p[IO["LR_RANK1"]] = rank
    p[IO["LR_RANK2"]] = rank
    p[IO["LR_TOL_EIG1"]] = lr_tol_eig
    p[IO["LR_TOL_EIG2"]] = lr_tol_eig
    p[IO["LR_MAXITS1"]] = lr_maxits
    p[IO["LR_MAXITS2"]] = lr_maxits
    p[IO["SCHUR_ENABLE"]] = 0 if schur is None else 1
    p[IO["GLOBAL_SOLVE"]] = 1 if usolve else 0
    if schur is not None:
        p[IO["SCHUR_MAXITS"]] = schur[0]
        p[IO["SCHUR_ITER_TOL"]] = schur[1]
    p[IO["SOLVER_KDIM"]] = kdim
    p[IO["SOLVER_MAXITS"]] = maxits
    p[IO["SOLVER_TOL"]] = tol
    return p
