"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and to the compiled
reference (oracle/_ref/libgsks_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker and by bench.py's cpu_baseline leg.  The product (ks_amd) never imports
this module.

The oracle restates dgsks_ref (frame/dgsks_ref.c:94-535); see gsks_oracle.h for
the per-function citations and for what pins each function.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libgsks_ref.so")
REF_SRC = "/root/reference"

# ks_type order of the reference (include/ks.h:49-58)
GAUSSIAN, POLYNOMIAL, LAPLACE, VAR_BANDWIDTH, TANH, QUARTIC, MULTIQUADRATIC, EPANECHNIKOV = range(8)
TYPE_NAMES = ["Gaussian", "Polynomial", "Laplace", "Var_bandwidth", "Tanh", "Quartic",
              "Multiquadratic", "Epanechnikov"]

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


class OrcKernel(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("powe", ctypes.c_double), ("scal", ctypes.c_double),
                ("cons", ctypes.c_double), ("hi", _dp), ("hj", _dp), ("clamp", ctypes.c_int)]


class RefKsT(ctypes.Structure):
    """The reference's struct kernel_s (include/ks.h:60-69)."""
    _fields_ = [("type", ctypes.c_int), ("powe", ctypes.c_double), ("scal", ctypes.c_double),
                ("cons", ctypes.c_double), ("hi", _dp), ("hj", _dp), ("h", _dp)]


_lock = threading.Lock()
_lib = None
_ref = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with gcc (oracle/Makefile)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")] +
                       (["-B"] if force else []), check=True)
    return LIB_PATH


def build_ref(force: bool = False) -> str | None:
    """Compile the reference from its sources into oracle/_ref (only where the
    reference tree is present)."""
    if not os.path.isdir(REF_SRC):
        return REF_PATH if os.path.exists(REF_PATH) else None
    if force or not os.path.exists(REF_PATH):
        subprocess.run(["make", "-s", "-j8", "-f", os.path.join(HERE, "Makefile.ref"),
                        f"REF={REF_SRC}"] + (["-B"] if force else []), check=True)
    return REF_PATH


def lib():
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                build()
            L = ctypes.CDLL(LIB_PATH)
            L.orc_laplace_params.argtypes = [ctypes.c_int, _dp, _dp]
            L.orc_sqnorm.argtypes = [_dp, ctypes.c_long, ctypes.c_int, _dp]
            L.orc_pair.argtypes = [ctypes.POINTER(OrcKernel), ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double]
            L.orc_pair.restype = ctypes.c_double
            L.orc_dgsks_rows.argtypes = [ctypes.POINTER(OrcKernel), ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, _dp, _ip, _dp, _dp, _ip, _dp, _dp, _ip,
                                         _dp, _ip, ctypes.c_int, _ip, _dp, ctypes.c_int]
            L.orc_dgsks_ref.argtypes = [ctypes.POINTER(OrcKernel), ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, _dp, _ip, _dp, _dp, _ip, _dp, _dp, _ip,
                                        _dp, _ip, ctypes.c_int]
            L.orc_dgsks_list.argtypes = [ctypes.c_int, ctypes.POINTER(OrcKernel), ctypes.c_int,
                                         _dp, ctypes.c_long, _lp, _ip, _ip, _lp, _ip, _ip,
                                         _dp, _dp, _dp, _dp, _dp, ctypes.c_int]
            L.orc_gauss_grad.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                         _dp, _dp, ctypes.c_double, ctypes.c_int, _ip, _dp,
                                         ctypes.c_int, _ip, _dp, _dp, _dp, ctypes.c_int]
            L.orc_ks_grad.argtypes = [ctypes.POINTER(OrcKernel), ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, _dp, _ip, _dp, _ip, _dp, _ip, _dp, _ip,
                                      _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int]
            _lib = L
    return _lib


def ref_lib():
    """The compiled reference library, or None where it has not been built."""
    global _ref
    with _lock:
        if _ref is None:
            if not os.path.exists(REF_PATH):
                return None
            R = ctypes.CDLL(REF_PATH)
            sig = [ctypes.POINTER(RefKsT), ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _ip,
                   _dp, _dp, _ip, _dp, _dp, _ip, _dp, _ip]
            R.dgsks.argtypes = sig
            R.dgsks_ref.argtypes = sig
            _ref = R
    return _ref


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _nthreads(n):
    if n:
        return int(n)
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(avail, 16))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def laplace_params(k: int):
    p, s = ctypes.c_double(), ctypes.c_double()
    lib().orc_laplace_params(int(k), ctypes.byref(p), ctypes.byref(s))
    return p.value, s.value


def sqnorm(X):
    X = _f64(X)
    out = np.empty(X.shape[0], dtype=np.float64)
    lib().orc_sqnorm(_d(X), X.shape[0], X.shape[1], _d(out))
    return out


def _kernel(ktype, scal=0.0, cons=0.0, powe=0.0, hi=None, hj=None, clamp=False):
    kr = OrcKernel(int(ktype), float(powe), float(scal), float(cons), None, None, int(clamp))
    keep = []
    if hi is not None:
        hi = _f64(hi)
        keep.append(hi)
        kr.hi = _d(hi)
    if hj is not None:
        hj = _f64(hj)
        keep.append(hj)
        kr.hj = _d(hj)
    return kr, keep


def pair(ktype, c, aa, bb, hi=0.0, hj=0.0, *, scal=0.0, cons=0.0, powe=0.0, clamp=False):
    """K of one pair from c = a.b, |a|^2, |b|^2 (frame/dgsks_ref.c:253-424);
    clamp=True: d^2 clamped at 0 as the fused micro-kernels do
    (micro_kernel/x86_64/haswell/sq2nrm_int_d8x6.h:64-78)."""
    kr, _ = _kernel(ktype, scal, cons, powe, clamp=clamp)
    return lib().orc_pair(ctypes.byref(kr), float(c), float(aa), float(bb), float(hi), float(hj))


def dgsks_rows(ktype, XA, XA2, amap, XB, XB2, bmap, w, wmap, u, umap, rows=None, *,
               scal=0.0, cons=0.0, powe=0.0, hi=None, hj=None, nthreads=None,
               apply_laplace=True, clamp=False):
    """out[r] = u[umap[rows[r]]] + sum_j K(rows[r], j) w[wmap[j]] (rows=None: all).
    clamp=True evaluates K with dgsks()'s d^2 >= 0 clamp instead of dgsks_ref's
    unclamped form."""
    XA, XA2, XB, XB2, w, u = map(_f64, (XA, XA2, XB, XB2, w, u))
    amap, bmap, wmap, umap = map(_i32, (amap, bmap, wmap, umap))
    m, n = amap.size, bmap.size
    k = XA.shape[1]
    if ktype == LAPLACE and apply_laplace:
        powe, scal = laplace_params(k)
    rows = np.arange(m, dtype=np.int32) if rows is None else _i32(rows)
    out = np.empty(rows.size, dtype=np.float64)
    kr, keep = _kernel(ktype, scal, cons, powe, hi, hj, clamp)
    lib().orc_dgsks_rows(ctypes.byref(kr), m, n, k, _d(u), _i(umap), _d(XA), _d(XA2), _i(amap),
                         _d(XB), _d(XB2), _i(bmap), _d(w), _i(wmap), rows.size, _i(rows),
                         _d(out), _nthreads(nthreads))
    return out


def dgsks_ref(ktype, u, umap, XA, XA2, amap, XB, XB2, bmap, w, wmap, *, scal=0.0, cons=0.0,
              powe=0.0, hi=None, hj=None, nthreads=None):
    """In-place u update with dgsks_ref's semantics; returns the (powe, scal) the
    descriptor holds afterwards (Laplace writes them)."""
    assert u.dtype == np.float64 and u.flags.c_contiguous
    XA, XA2, XB, XB2, w = map(_f64, (XA, XA2, XB, XB2, w))
    amap, bmap, wmap, umap = map(_i32, (amap, bmap, wmap, umap))
    kr, keep = _kernel(ktype, scal, cons, powe, hi, hj)
    lib().orc_dgsks_ref(ctypes.byref(kr), amap.size, bmap.size, XA.shape[1], _d(u), _i(umap),
                        _d(XA), _d(XA2), _i(amap), _d(XB), _d(XB2), _i(bmap), _d(w), _i(wmap),
                        _nthreads(nthreads))
    return kr.powe, kr.scal


def dgsks_list(kernels, k, u, off_a, umap, amap, off_b, bmap, wmap, XA, XA2, XB, XB2, w,
               nthreads=None):
    """kernels: list of dicts {type, scal, cons, powe, hi, hj}; u updated in place."""
    ntask = len(kernels)
    arr = (OrcKernel * max(ntask, 1))()
    keep = []
    for t, kd in enumerate(kernels):
        kr, kp = _kernel(kd["type"], kd.get("scal", 0.0), kd.get("cons", 0.0),
                         kd.get("powe", 0.0), kd.get("hi"), kd.get("hj"))
        keep += kp
        arr[t] = kr
    off_a = np.ascontiguousarray(off_a, dtype=np.int64)
    off_b = np.ascontiguousarray(off_b, dtype=np.int64)
    umap, amap, bmap, wmap = map(_i32, (umap, amap, bmap, wmap))
    XA, XA2, XB, XB2, w = map(_f64, (XA, XA2, XB, XB2, w))
    lib().orc_dgsks_list(ntask, arr, int(k), _d(u), u.size, off_a.ctypes.data_as(_lp), _i(umap),
                         _i(amap), off_b.ctypes.data_as(_lp), _i(bmap), _i(wmap), _d(XA),
                         _d(XA2), _d(XB), _d(XB2), _d(w), _nthreads(nthreads))


def gauss_grad(XA, XB, w, g, h, rows_a=None, rows_b=None, want_dh=True, nthreads=None):
    XA, XB, w, g = map(_f64, (XA, XB, w, g))
    m, k = XA.shape
    n = XB.shape[0]
    ra = None if rows_a is None else _i32(rows_a)
    rb = None if rows_b is None else _i32(rows_b)
    na = m if ra is None else ra.size
    nb = n if rb is None else rb.size
    dXA = np.zeros((na, k))
    dXB = np.zeros((nb, k))
    dw = np.zeros(nb)
    dh = ctypes.c_double(0.0)
    lib().orc_gauss_grad(m, n, k, _d(XA), _d(XB), _d(w), _d(g), float(h), na,
                         None if ra is None else _i(ra), _d(dXA), nb,
                         None if rb is None else _i(rb), _d(dXB), _d(dw),
                         ctypes.byref(dh) if want_dh else None, _nthreads(nthreads))
    return dXA, dXB, dw, (dh.value if want_dh else None)


def ks_grad(ktype, g, umap, XA, amap, XB, bmap, w, wmap, *, scal=0.0, hi=None, hj=None,
            nthreads=None):
    """Gradient tables of L = g . du for the mapped Gaussian or variable-bandwidth
    summation (orc_ks_grad): dict with dXA, dXB, dw (shaped like XA, XB, w),
    dscal (Gaussian) or dhi, dhj (shaped like hi, hj; variable bandwidth),
    accumulated from zero."""
    XA, XB, w, g = map(_f64, (XA, XB, w, g))
    umap, amap, bmap, wmap = map(_i32, (umap, amap, bmap, wmap))
    k = XA.shape[1]
    out = {"dXA": np.zeros_like(XA), "dXB": np.zeros_like(XB), "dw": np.zeros_like(w)}
    var = ktype == VAR_BANDWIDTH
    ds = ctypes.c_double(0.0)
    if var:
        out["dhi"] = np.zeros(np.asarray(hi).shape)
        out["dhj"] = np.zeros(np.asarray(hj).shape)
    kr, keep = _kernel(ktype, scal, 0.0, 0.0, hi, hj, False)
    lib().orc_ks_grad(ctypes.byref(kr), amap.size, bmap.size, k, _d(g), _i(umap), _d(XA),
                      _i(amap), _d(XB), _i(bmap), _d(w), _i(wmap), _d(out["dXA"]),
                      _d(out["dXB"]), _d(out["dw"]), None if var else ctypes.byref(ds),
                      _d(out["dhi"]) if var else None, _d(out["dhj"]) if var else None,
                      _nthreads(nthreads))
    if not var:
        out["dscal"] = ds.value
    return out


def ref_call(fn_name, ktype, u, umap, XA, XA2, amap, XB, XB2, bmap, w, wmap, *, scal=0.0,
             cons=0.0, powe=0.0, hi=None, hj=None):
    """Call the compiled reference's dgsks or dgsks_ref in place on u."""
    R = ref_lib()
    if R is None:
        raise RuntimeError("oracle/_ref/libgsks_ref.so is not built")
    assert u.dtype == np.float64 and u.flags.c_contiguous
    XA, XA2, XB, XB2, w = map(_f64, (XA, XA2, XB, XB2, w))
    amap, bmap, wmap, umap = map(_i32, (amap, bmap, wmap, umap))
    kt = RefKsT(int(ktype), float(powe), float(scal), float(cons), None, None, None)
    keep = []
    if hi is not None:
        hi = _f64(hi)
        keep.append(hi)
        kt.hi = _d(hi)
    if hj is not None:
        hj = _f64(hj)
        keep.append(hj)
        kt.hj = _d(hj)
    getattr(R, fn_name)(ctypes.byref(kt), amap.size, bmap.size, XA.shape[1], _d(u), _i(umap),
                        _d(XA), _d(XA2), _i(amap), _d(XB), _d(XB2), _i(bmap), _d(w), _i(wmap))
    return kt.powe, kt.scal


def rel_l2(x, ref):
    """The reference's error metric (test/test_dgsks.c:43-79)."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.sqrt(np.sum(ref * ref))
    num = np.sqrt(np.sum((x - ref) ** 2))
    return float(num / den) if den > 0 else float(num)
