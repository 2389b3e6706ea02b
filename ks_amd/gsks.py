"""Python interface of the fused kernel summation, mirroring the reference's
python/gsks.py (python/gsks.py:18-199): same functions, same positional
arguments and option names, same shape/type assertions.

    dgsks(d, u, mu, XA, XA2, alpha, XB, XB2, beta, w, omega, **options)
    dgsks_ref(d, u, mu, XA, XA2, alpha, XB, XB2, beta, w, omega, **options)
    sgsks(d, u, mu, XA, XA2, alpha, XB, XB2, beta, w, omega, **options)   # float32

u[mu[i]] += sum_j K(XA[alpha[i]], XB[beta[j]]) * w[omega[j]], in place.
options: type in {"Gaussian", "Polynomial", "Laplace", "Variable_Bandwidth",
"Tanh", "Quartic", "Multiquadratic", "Epanechnikov"}, scal, cons, powe, h.

The arrays may be numpy arrays (host memory: the library stages them through
HBM, like the reference's ctypes call of dgsks_wrapper) or torch tensors on the
GPU (device memory: the call is enqueued on torch's current stream through
dgsks_device, with no copies).  Differences from the reference wrapper:
  * cons is passed as cons (the reference passes it in the scal slot,
    python/gsks.py:53-54,65-66,74, and the C wrapper never sets it);
  * precision= selects the mixed-precision forms for device tensors.
Every summation runs on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import KsT

TYPES = {
    "Gaussian": _lib.KS_GAUSSIAN,
    "Polynomial": _lib.KS_POLYNOMIAL,
    "Laplace": _lib.KS_LAPLACE,
    "Variable_Bandwidth": _lib.KS_GAUSSIAN_VAR_BANDWIDTH,
    "Tanh": _lib.KS_TANH,
    "Quartic": _lib.KS_QUARTIC,
    "Multiquadratic": _lib.KS_MULTIQUADRATIC,
    "Epanechnikov": _lib.KS_EPANECHNIKOV,
}
PRECISIONS = {"fp64": _lib.KS_PREC_FP64, "mixed_fp32": _lib.KS_PREC_MIXED_FP32,
              "mixed_bf16x3": _lib.KS_PREC_MIXED_BF16X3}


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _require(cond: bool, msg: str):
    # python/gsks.py:19-36 uses assert; keep AssertionError but independent of -O
    if not cond:
        raise AssertionError(msg)


def _size(x) -> int:
    return int(x.numel()) if _is_torch(x) else int(np.asarray(x).size)


def _check_args(d, u, mu, XA, XA2, alpha, XB, XB2, beta, w, omega, single=False):
    """The range and type checks of python/gsks.py:19-36 (float32 arrays and
    optional norms for the single-precision call)."""
    M, N = XA.shape[0], XB.shape[0]
    m, n = _size(alpha), _size(beta)
    _require(XA.shape[1] == XB.shape[1], "XA.shape[1] == XB.shape[1]")
    _require(M >= m, "M >= m")
    _require(N >= n, "N >= n")
    _require((single and XA2 is None) or _size(XA2) == M, "XA2.size == M")
    _require((single and XB2 is None) or _size(XB2) == N, "XB2.size == N")
    real = [("XA", XA), ("XB", XB), ("XA2", XA2), ("XB2", XB2), ("u", u), ("w", w)]
    ints = [("alpha", alpha), ("beta", beta), ("mu", mu), ("omega", omega)]
    if _is_torch(XA):
        import torch
        fdt = torch.float32 if single else torch.float64
        for (name, t), dt in [(r, fdt) for r in real] + [(i, torch.int32) for i in ints]:
            if t is None:
                continue
            _require(t.dtype == dt, f"{name}.dtype == {dt}")
            _require(t.is_contiguous(), f"{name} is contiguous")
    else:
        fdt = np.float32 if single else np.float64
        for (name, a), dt in [(r, fdt) for r in real] + [(i, np.int32) for i in ints]:
            if a is None:
                continue
            _require(a.dtype == dt, f"{name}.dtype == {np.dtype(dt).name}")
            _require(a.flags.c_contiguous, f"{name} is C-contiguous")
    _require(_size(mu) == m, "mu.size == alpha.size")
    _require(_size(omega) == n, "omega.size == beta.size")
    _require(int(d) == XA.shape[1], "d == XA.shape[1]")
    return m, n


def _ptr(x):
    if x is None:
        return None
    return ctypes.c_void_p(x.data_ptr() if _is_torch(x) else x.ctypes.data)


def make_kernel(options) -> KsT | None:
    """ks_t from the option dictionary (python/gsks.py:38-79, with cons fixed)."""
    name = options.get("type")
    if name not in TYPES:
        return None
    kt = KsT()
    kt.type = TYPES[name]
    # defaults of python/gsks.py:38-41
    kt.scal = float(options.get("scal", 1.0))
    kt.cons = float(options.get("cons", 0.0))
    kt.powe = float(options.get("powe", 1.0))
    if name == "Variable_Bandwidth":
        h = options.get("h")
        _require(h is not None, "Variable_Bandwidth needs h")
        kt.hi = kt.hj = _ptr(h).value
    return kt


def _call(fn_name, d, u, mu, XA, XA2, alpha, XB, XB2, beta, w, omega, options):
    single = fn_name == "sgsks"
    m, n = _check_args(d, u, mu, XA, XA2, alpha, XB, XB2, beta, w, omega, single)
    kt = make_kernel(options)
    if kt is None:
        # python/gsks.py:80-82
        print("This kernel option is not suppurted.")
        return
    L = _lib.lib()
    args = [_ptr(a) for a in (u, mu, XA, XA2, alpha, XB, XB2, beta, w, omega)]
    given = options.get("precision")
    precision = PRECISIONS[given or "fp64"]
    if _is_torch(u) and u.is_cuda:
        import torch
        stream = ctypes.c_void_p(torch.cuda.current_stream(u.device).cuda_stream)
        if fn_name == "dgsks" or (single and given is not None):
            rc = getattr(L, fn_name + "_device")(ctypes.byref(kt), m, n, int(d), *args,
                                                 precision, stream)
            _lib.check(rc, fn_name + "_device")
            return
        torch.cuda.current_stream(u.device).synchronize()
    if single:  # the host sgsks() picks its precision itself
        _require(given is None, "sgsks: precision= needs device tensors")
    _require(given is None or precision == _lib.KS_PREC_FP64 or fn_name == "dgsks_ref",
             "an explicit mixed precision needs device tensors")
    getattr(L, fn_name)(ctypes.byref(kt), m, n, int(d), *args)
    code = L.ks_last_error()
    if code == _lib.KS_ERR_LAPLACE_K:  # the reference prints and carries on
        import warnings
        warnings.warn("laplace kernel only supports k > 2")
        return
    _lib.check(code, fn_name)


def dgsks(d, u, mu, XA, XA2, alpha, XB, XB2, beta, w, omega, **options):
    """Fused summation on the GPU (python/gsks.py:112-199)."""
    _call("dgsks", d, u, mu, XA, XA2, alpha, XB, XB2, beta, w, omega, options)


def dgsks_ref(d, u, mu, XA, XA2, alpha, XB, XB2, beta, w, omega, **options):
    """The reference algorithm (unfused: cross-term GEMM panel, elementwise
    kernel, GEMV) on the GPU (python/gsks.py:18-105)."""
    _call("dgsks_ref", d, u, mu, XA, XA2, alpha, XB, XB2, beta, w, omega, options)


def sgsks(d, u, mu, XA, XA2, alpha, XB, XB2, beta, w, omega, **options):
    """Single precision: float32 arrays (XA2 / XB2 may be None: norms computed
    in FP64), summed on the GPU after an exact widening to FP64.  Without
    precision= the sum is FP64; precision= (device tensors) selects a mixed
    form, whose accuracy depends on the data (1e-6 where near pairs carry the
    sums, about 1e-5 for rows made only of far pairs)."""
    _call("sgsks", d, u, mu, XA, XA2, alpha, XB, XB2, beta, w, omega, options)


def sqnorm(X):
    """Squared row norms of a (n, k) float64 CUDA tensor, computed on the GPU."""
    import torch
    _require(X.is_cuda and X.dtype == torch.float64 and X.is_contiguous(), "X: contiguous f64 cuda")
    out = torch.empty(X.shape[0], dtype=torch.float64, device=X.device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(X.device).cuda_stream)
    rc = _lib.lib().ks_sqnorm_device(_ptr(X), X.shape[0], X.shape[1], _ptr(out), stream)
    _lib.check(rc, "ks_sqnorm_device")
    return out
