"""ctypes wrapper of the oracle (oracle/oracle.c) -- TEST INFRASTRUCTURE ONLY.

Every function takes real (float64) or complex (complex128) data and calls
the matching C instantiation (orc_* or orc_*_z).

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
                ("E_a", C.c_void_p), ("F_a", C.c_void_p), ("W", C.c_void_p), ("WH", C.c_void_p),
                ("C_ia", C.c_void_p), ("C_ja", C.c_void_p), ("C_a", C.c_void_p),
                ("ncomp", C.c_int), ("dom_ptr", C.c_void_p)]


class OrcGemslr(C.Structure):
    _fields_ = [("n", C.c_int), ("nlev", C.c_int), ("perm", C.c_void_p), ("levels", C.POINTER(OrcLevel)),
                ("inner_maxits", C.c_int), ("inner_tol", C.c_double), ("usolve", C.c_int)]


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
        L.orc_gemslr_work_size.argtypes = [C.POINTER(OrcGemslr)]
        L.orc_gemslr_work_size.restype = C.c_long
        L.orc_set_threads.argtypes = [C.c_int]
        for sfx in ("", "_z"):
            getattr(L, "orc_csr_matvec" + sfx).argtypes = [C.c_int, vp, vp, vp, C.c_double, vp, C.c_double, vp]
            getattr(L, "orc_ilu_solve" + sfx).argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp]
            getattr(L, "orc_lowrank_apply" + sfx).argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp]
            getattr(L, "orc_gemslr_apply" + sfx).argtypes = [C.POINTER(OrcGemslr), vp, vp, vp]
            getattr(L, "orc_mgs" + sfx).argtypes = [C.c_int, vp, C.c_int, vp, C.c_int, C.POINTER(C.c_double)]
            getattr(L, "orc_cgs2" + sfx).argtypes = [C.c_int, vp, C.c_int, vp, C.c_int, C.POINTER(C.c_double)]
            getattr(L, "orc_fgmres" + sfx).argtypes = [C.c_int, vp, vp, vp, C.POINTER(OrcGemslr), vp, vp, C.c_int,
                                                       C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_int),
                                                       C.POINTER(C.c_double), vp]
        _lib = L
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _cplx(*arrs):
    return any(np.iscomplexobj(a) for a in arrs if a is not None)


def _val(a, cplx):
    return np.ascontiguousarray(a, dtype=np.complex128 if cplx else np.float64)


def _fn(name, cplx):
    return getattr(lib(), name + ("_z" if cplx else ""))


def csr_matvec(ia, ja, a, x, alpha=1.0, beta=0.0, y=None):
    cz = _cplx(a, x, y)
    ia, ja, a, x = _i32(ia), _i32(ja), _val(a, cz), _val(x, cz)
    n = len(ia) - 1
    y = np.zeros(n, dtype=a.dtype) if y is None else _val(y, cz).copy()
    _fn("orc_csr_matvec", cz)(n, _p(ia), _p(ja), _p(a), alpha, _p(x), beta, _p(y))
    return y


def ilu_solve(L, U, Dinv, rhs):
    (Li, Lj, La), (Ui, Uj, Ua) = L, U
    cz = _cplx(La, Ua, Dinv, rhs)
    n = len(Li) - 1
    arrs = [_i32(Li), _i32(Lj), _val(La, cz), _i32(Ui), _i32(Uj), _val(Ua, cz), _val(Dinv, cz), _val(rhs, cz)]
    x = np.zeros(n, dtype=arrs[-1].dtype)
    _fn("orc_ilu_solve", cz)(n, *[_p(a) for a in arrs], _p(x))
    return x


def lowrank_apply(W, WH, z):
    cz = _cplx(W, WH, z)
    dt = np.complex128 if cz else np.float64
    m, k = W.shape
    Wc = np.asfortranarray(W, dtype=dt)
    WHc = np.asfortranarray(WH, dtype=dt)
    z = _val(z, cz)
    x = np.zeros(m, dtype=dt)
    t = np.zeros(max(k, 1), dtype=dt)
    _fn("orc_lowrank_apply", cz)(m, k, C.c_void_p(Wc.ctypes.data), C.c_void_p(WHc.ctypes.data), _p(z), _p(x), _p(t))
    return x


class GemslrOracle:
    """Oracle view of a factored operator given as the level dicts of
    pargemslr_amd.Gemslr.levels() plus the permutation.  schur = (maxits,
    tol) turns on the inner FGMRES on the level-0 Schur complement (level 0
    must carry its block "C"); usolve selects the U variant of SolveLevel."""

    def __init__(self, n, perm, levels, schur=None, usolve=False):
        self.n = n
        self._keep = []
        self.is_complex = cz = _cplx(*[d["Dinv"] for d in levels])
        self.dtype = np.dtype(np.complex128 if cz else np.float64)
        self.perm = self._k(_i32(perm))
        lv = (OrcLevel * len(levels))()
        for i, d in enumerate(levels):
            o = lv[i]
            o.lev_start, o.nloc, o.nrem, o.rank = d["lev_start"], d["nloc"], d["nrem"], d["rank"]
            for nm in ("L", "U", "E", "F", "C"):
                if d.get(nm) is None:
                    continue
                ia, ja, a = d[nm]
                setattr(o, nm + "_ia", _p(self._k(_i32(ia))))
                setattr(o, nm + "_ja", _p(self._k(_i32(ja))))
                setattr(o, nm + "_a", _p(self._k(_val(a, cz))))
            o.Dinv = _p(self._k(_val(d["Dinv"], cz)))
            if d.get("dom_ptr") is not None:
                o.ncomp = int(d["ncomp"])
                o.dom_ptr = _p(self._k(_i32(d["dom_ptr"])))
            if d["rank"] > 0:
                o.W = C.c_void_p(self._k(np.asfortranarray(d["W"], dtype=self.dtype)).ctypes.data)
                o.WH = C.c_void_p(self._k(np.asfortranarray(d["WH"], dtype=self.dtype)).ctypes.data)
        self._levels = lv
        inner = (0, 0.0)
        if schur is not None and len(levels) > 1:
            if levels[0].get("C") is None:
                raise ValueError("the inner iteration needs level 0's block C")
            inner = (int(schur[0]), float(schur[1]))
        self.s = OrcGemslr(n, len(levels), _p(self.perm), lv, inner[0], inner[1], 1 if usolve else 0)
        self.work = np.zeros(lib().orc_gemslr_work_size(C.byref(self.s)), dtype=self.dtype)

    def _k(self, a):
        self._keep.append(a)
        return a

    def apply(self, v):
        v = _val(v, self.is_complex)
        z = np.zeros(self.n, dtype=self.dtype)
        _fn("orc_gemslr_apply", self.is_complex)(C.byref(self.s), _p(v), _p(z), _p(self.work))
        return z


def fgmres(A, b, x0=None, prec=None, kdim=50, maxits=1000, tol=1e-8, absolute=False, orth=0):
    """Oracle FGMRES; A is a pargemslr_amd.CsrHost (real or complex), prec a
    GemslrOracle of the same value type or None.  Returns (x, status, iters,
    rel_res, hist)."""
    cz = _cplx(A.a)
    if prec is not None and prec.is_complex != cz:
        raise ValueError("operator and preconditioner value types differ")
    ia, ja, a = _i32(A.ia), _i32(A.ja), _val(A.a, cz)
    b = _val(b, cz)
    x = np.zeros(A.n, dtype=b.dtype) if x0 is None else _val(x0, cz).copy()
    it = C.c_int()
    rr = C.c_double()
    hist = np.full(maxits + 1, np.nan)
    st = _fn("orc_fgmres", cz)(A.n, _p(ia), _p(ja), _p(a), C.byref(prec.s) if prec is not None else None, _p(b),
                               _p(x), kdim, maxits, tol, int(absolute), orth, C.byref(it), C.byref(rr), _p(hist))
    return x, st, it.value, rr.value, hist[: it.value + 1]


def set_threads(t):
    """Threads over which the oracle solves a level's diagonal blocks (the
    CPU baseline's OpenMP-over-blocks variant; the results are the same
    bits); 1 = serial, the default."""
    lib().orc_set_threads(int(t))


def mgs(V, k, H, cgs2=False):
    """Orthogonalize column k+1 of V (n x (k+2), Fortran order) in place; fills H[:, k]."""
    n = V.shape[0]
    t = C.c_double()
    cz = np.iscomplexobj(V)
    assert np.iscomplexobj(H) == cz
    f = _fn("orc_cgs2" if cgs2 else "orc_mgs", cz)
    f(n, C.c_void_p(V.ctypes.data), k, C.c_void_p(H.ctypes.data), H.shape[0], C.byref(t))
    return t.value
