"""ctypes binding of ks_amd/libgsks.so (the HIP kernels behind the C ABI of
include/ks.h).  Loading fails loudly when the library has not been built:
there is no CPU fallback anywhere in this package."""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
# KS_LIBGSKS names another build of the same library (the host-sanitized one
# that `python -m ks_amd.build --sanitize` writes), for the host-layer checks.
LIB_PATH = os.environ.get("KS_LIBGSKS") or os.path.join(HERE, "libgsks.so")

# include/ks.h enums
KS_GAUSSIAN, KS_POLYNOMIAL, KS_LAPLACE, KS_GAUSSIAN_VAR_BANDWIDTH, KS_TANH, KS_QUARTIC, \
    KS_MULTIQUADRATIC, KS_EPANECHNIKOV = range(8)
KS_OK, KS_ERR_INVALID_TYPE, KS_ERR_INVALID_ARG, KS_ERR_NO_BANDWIDTH, KS_ERR_LAPLACE_K, \
    KS_ERR_DEVICE, KS_ERR_ALLOC, KS_ERR_MIXED_POINTERS, KS_ERR_LIST = range(9)
KS_ERRMODE_EXIT, KS_ERRMODE_RETURN = 0, 1
KS_PREC_FP64, KS_PREC_MIXED_FP32, KS_PREC_MIXED_BF16X3 = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_vp = ctypes.c_void_p


class KsT(ctypes.Structure):
    """struct kernel_s of include/ks.h (reference include/ks.h:60-69)."""
    _fields_ = [("type", ctypes.c_int), ("powe", ctypes.c_double), ("scal", ctypes.c_double),
                ("cons", ctypes.c_double), ("hi", _vp), ("hj", _vp), ("h", _vp)]


_lock = threading.Lock()
_lib = None

_SIGS = {
    "dgsks": (None, [ctypes.POINTER(KsT), ctypes.c_int, ctypes.c_int, ctypes.c_int] + [_vp] * 10),
    "dgsks_ref": (None, [ctypes.POINTER(KsT), ctypes.c_int, ctypes.c_int, ctypes.c_int] + [_vp] * 10),
    "ks_malloc_aligned": (_vp, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "dgsks_wrapper": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [_vp] * 10 +
                      [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _vp]),
    "dgsks_ref_wrapper": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [_vp] * 10 +
                          [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _vp]),
    "ks_last_error": (ctypes.c_int, []),
    "ks_error_string": (ctypes.c_char_p, [ctypes.c_int]),
    "ks_error_detail": (ctypes.c_char_p, []),
    "ks_set_error_mode": (None, [ctypes.c_int]),
    "ks_get_error_mode": (ctypes.c_int, []),
    "dgsks_device": (ctypes.c_int, [ctypes.POINTER(KsT), ctypes.c_int, ctypes.c_int, ctypes.c_int] +
                     [_vp] * 10 + [ctypes.c_int, _vp]),
    "dgsks_device_noalloc": (ctypes.c_int, [ctypes.POINTER(KsT), ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int] + [_vp] * 10 + [ctypes.c_int, _vp]),
    "sgsks_device": (ctypes.c_int, [ctypes.POINTER(KsT), ctypes.c_int, ctypes.c_int, ctypes.c_int] +
                     [_vp] * 10 + [ctypes.c_int, _vp]),
    "sgsks": (None, [ctypes.POINTER(KsT), ctypes.c_int, ctypes.c_int, ctypes.c_int] + [_vp] * 10),
    "sgsks_wrapper": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [_vp] * 10 +
                      [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _vp]),
    "ks_sqnorm_device": (ctypes.c_int, [_vp, ctypes.c_long, ctypes.c_int, _vp, _vp]),
    "dgsks_list": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(KsT), ctypes.c_int, _vp,
                                  ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                  _vp, _vp, ctypes.c_int]),
    "dgsks_multirhs_device": (ctypes.c_int, [ctypes.POINTER(KsT), ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int] + [_vp] * 10 + [_vp]),
    "dgsks_multirhs_mixed_device": (ctypes.c_int, [ctypes.POINTER(KsT), ctypes.c_int,
                                                   ctypes.c_int, ctypes.c_int, ctypes.c_int]
                                    + [_vp] * 10 + [_vp]),
    "ks_gauss_grad_device": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp,
                                            _vp, _vp, ctypes.c_double, _vp, _vp, _vp, _vp,
                                            ctypes.c_int, _vp]),
    "ks_gauss_grad_forward_device": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                    _vp, _vp, _vp, ctypes.c_double, _vp, _vp,
                                                    ctypes.c_int, _vp]),
    "ks_gauss_grad_backward_device": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                     _vp, _vp, _vp, _vp, ctypes.c_double, _vp,
                                                     _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "ks_grad_device": (ctypes.c_int, [ctypes.POINTER(KsT), ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                      _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "ks_list_plan_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(KsT),
                                           ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                           ctypes.c_int]),
    "ks_list_plan_execute": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ks_list_plan_destroy": (None, [_vp]),
    "ks_list_plan_symmetric_tasks": (ctypes.c_int, [_vp]),
    "ks_list_autotune": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
}


def _bind_process_hip_runtime():
    """Load torch's HIP runtime before libgsks.so when torch is installed.

    torch ships its own libamdhip64 with the same SONAME as ROCm's.  If
    libgsks.so were loaded first, a later `import torch` would bring a second
    HIP runtime into the process and the two would compete for the device;
    loading torch first makes libgsks.so bind to the runtime already loaded."""
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def lib():
    """Load libgsks.so (once).  Raises when it is missing."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is not built: run `python -m ks_amd.build` "
                    "(ks_amd has no CPU fallback)")
            _bind_process_hip_runtime()
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _SIGS.items():
                fn = getattr(L, name, None)
                if fn is None:
                    continue
                fn.restype = res
                fn.argtypes = args
            # Python callers get exceptions instead of the reference's exit(1)
            L.ks_set_error_mode(KS_ERRMODE_RETURN)
            _lib = L
    return _lib


def error_string(code: int) -> str:
    return lib().ks_error_string(int(code)).decode()


class KsError(RuntimeError):
    def __init__(self, code: int, where: str = ""):
        self.code = code
        detail = lib().ks_error_detail().decode() if code == KS_ERR_DEVICE else ""
        super().__init__(f"{where}: {error_string(code)} (code {code})" +
                         (f" [{detail}]" if detail else ""))


def check(code: int, where: str):
    if code != KS_OK:
        raise KsError(code, where)
