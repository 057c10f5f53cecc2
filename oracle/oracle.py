"""ctypes wrapper of the oracle (oracle/oracle.c) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg;
never by the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OrcLevel(C.Structure):
    _fields_ = [("lev_start", C.c_int), ("nloc", C.c_int), ("nrem", C.c_int), ("rank", C.c_int),
                ("L_ia", C.c_void_p), ("L_ja", C.c_void_p), ("U_ia", C.c_void_p), ("U_ja", C.c_void_p),
                ("L_a", C.c_void_p), ("U_a", C.c_void_p), ("Dinv", C.c_void_p),
                ("E_ia", C.c_void_p), ("E_ja", C.c_void_p), ("F_ia", C.c_void_p), ("F_ja", C.c_void_p),
                ("E_a", C.c_void_p), ("F_a", C.c_void_p), ("W", C.c_void_p), ("WH", C.c_void_p)]


class OrcGemslr(C.Structure):
    _fields_ = [("n", C.c_int), ("nlev", C.c_int), ("perm", C.c_void_p), ("levels", C.POINTER(OrcLevel))]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_csr_matvec.argtypes = [C.c_int, vp, vp, vp, C.c_double, vp, C.c_double, vp]
        L.orc_ilu_solve.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.orc_lowrank_apply.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp]
        L.orc_gemslr_work_size.argtypes = [C.POINTER(OrcGemslr)]
        L.orc_gemslr_work_size.restype = C.c_long
        L.orc_gemslr_apply.argtypes = [C.POINTER(OrcGemslr), vp, vp, vp]
        L.orc_mgs.argtypes = [C.c_int, vp, C.c_int, vp, C.c_int, C.POINTER(C.c_double)]
        L.orc_cgs2.argtypes = [C.c_int, vp, C.c_int, vp, C.c_int, C.POINTER(C.c_double)]
        L.orc_fgmres.argtypes = [C.c_int, vp, vp, vp, C.POINTER(OrcGemslr), vp, vp, C.c_int, C.c_int, C.c_double,
                                 C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double), vp]
        _lib = L
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def csr_matvec(ia, ja, a, x, alpha=1.0, beta=0.0, y=None):
    ia, ja, a, x = _i32(ia), _i32(ja), _f64(a), _f64(x)
    n = len(ia) - 1
    y = np.zeros(n) if y is None else _f64(y).copy()
    lib().orc_csr_matvec(n, _p(ia), _p(ja), _p(a), alpha, _p(x), beta, _p(y))
    return y


def ilu_solve(L, U, Dinv, rhs):
    (Li, Lj, La), (Ui, Uj, Ua) = L, U
    n = len(Li) - 1
    arrs = [_i32(Li), _i32(Lj), _f64(La), _i32(Ui), _i32(Uj), _f64(Ua), _f64(Dinv), _f64(rhs)]
    x = np.zeros(n)
    lib().orc_ilu_solve(n, *[_p(a) for a in arrs], _p(x))
    return x


def lowrank_apply(W, WH, z):
    m, k = W.shape
    Wc = np.asfortranarray(W, dtype=np.float64)
    WHc = np.asfortranarray(WH, dtype=np.float64)
    z = _f64(z)
    x = np.zeros(m)
    t = np.zeros(max(k, 1))
    lib().orc_lowrank_apply(m, k, C.c_void_p(Wc.ctypes.data), C.c_void_p(WHc.ctypes.data), _p(z), _p(x), _p(t))
    return x


class GemslrOracle:
    """Oracle view of a factored operator given as the level dicts of
    pargemslr_amd.Gemslr.levels() plus the permutation."""

    def __init__(self, n, perm, levels):
        self.n = n
        self._keep = []
        self.perm = self._k(_i32(perm))
        lv = (OrcLevel * len(levels))()
        for i, d in enumerate(levels):
            o = lv[i]
            o.lev_start, o.nloc, o.nrem, o.rank = d["lev_start"], d["nloc"], d["nrem"], d["rank"]
            for nm in ("L", "U", "E", "F"):
                if d.get(nm) is None:
                    continue
                ia, ja, a = d[nm]
                setattr(o, nm + "_ia", _p(self._k(_i32(ia))))
                setattr(o, nm + "_ja", _p(self._k(_i32(ja))))
                setattr(o, nm + "_a", _p(self._k(_f64(a))))
            o.Dinv = _p(self._k(_f64(d["Dinv"])))
            if d["rank"] > 0:
                o.W = C.c_void_p(self._k(np.asfortranarray(d["W"], dtype=np.float64)).ctypes.data)
                o.WH = C.c_void_p(self._k(np.asfortranarray(d["WH"], dtype=np.float64)).ctypes.data)
        self._levels = lv
        self.s = OrcGemslr(n, len(levels), _p(self.perm), lv)
        self.work = np.zeros(lib().orc_gemslr_work_size(C.byref(self.s)))

    def _k(self, a):
        self._keep.append(a)
        return a

    def apply(self, v):
        v = _f64(v)
        z = np.zeros(self.n)
        lib().orc_gemslr_apply(C.byref(self.s), _p(v), _p(z), _p(self.work))
        return z


def fgmres(A, b, x0=None, prec=None, kdim=50, maxits=1000, tol=1e-8, absolute=False, orth=0):
    """Oracle FGMRES; A is a pargemslr_amd.CsrHost (real), prec a GemslrOracle or None.
    Returns (x, status, iters, rel_res, hist)."""
    ia, ja, a = _i32(A.ia), _i32(A.ja), _f64(A.a)
    b = _f64(b)
    x = np.zeros(A.n) if x0 is None else _f64(x0).copy()
    it = C.c_int()
    rr = C.c_double()
    hist = np.full(maxits + 1, np.nan)
    st = lib().orc_fgmres(A.n, _p(ia), _p(ja), _p(a), C.byref(prec.s) if prec is not None else None, _p(b), _p(x),
                          kdim, maxits, tol, int(absolute), orth, C.byref(it), C.byref(rr), _p(hist))
    return x, st, it.value, rr.value, hist[: it.value + 1]


def mgs(V, k, H, cgs2=False):
    """Orthogonalize column k+1 of V (n x (k+2), Fortran order) in place; fills H[:, k]."""
    n = V.shape[0]
    t = C.c_double()
    f = lib().orc_cgs2 if cgs2 else lib().orc_mgs
    f(n, C.c_void_p(V.ctypes.data), k, C.c_void_p(H.ctypes.data), H.shape[0], C.byref(t))
    return t.value
