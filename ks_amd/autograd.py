"""torch.autograd.Function for the Gaussian kernel summation

    u = K(XA, XB; h) w,   K_ij = exp(-|a_i - b_j|^2 / (2 h^2)),

forward through the fused HIP kernel (dgsks_device, FP64 drop-in or a mixed
form), backward through the hand-written gradient kernels
(ks_gauss_grad_device): dL/dw is the transposed summation, dL/dXA, dL/dXB
and dL/dh come from two multi-right-hand-side summations that recompute the
kernel tiles instead of storing the m x n matrix.  All tensors are float64 on
the GPU; there is no CPU path.

Precisions: "fp64" runs both directions in FP64.  "mixed_fp32" runs the
forward (one right-hand side) and both gradient summations through the
mixed multi-right-hand-side kernel (dgsks_multirhs_mixed_device: FP32 matrix
cores on centred points wherever a per-block error bound allows, FP64
elsewhere), the form for unclustered points, whose pairs are mostly near.
"mixed_bf16x3" runs the forward through dgsks_device's split-BF16 form (for
clustered points, where most blocks are far) and the backward as
"mixed_fp32".
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import KsT

PRECISIONS = {"fp64": _lib.KS_PREC_FP64, "mixed_fp32": _lib.KS_PREC_MIXED_FP32,
              "mixed_bf16x3": _lib.KS_PREC_MIXED_BF16X3}


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _check(t, name, ndim):
    if not (t.is_cuda and t.dtype == torch.float64 and t.dim() == ndim):
        raise AssertionError(f"{name}: float64 CUDA tensor with {ndim} dims expected")


def gaussian_summation_forward(XA, XB, w, h: float, precision: str = "fp64"):
    """u = K(XA, XB; h) w with identity maps (no autograd)."""
    L = _lib.lib()
    XA = XA.contiguous()
    XB = XB.contiguous()
    w = w.contiguous()
    m, k = XA.shape
    n = XB.shape[0]
    dev = XA.device
    XA2 = torch.empty(m, dtype=torch.float64, device=dev)
    XB2 = torch.empty(n, dtype=torch.float64, device=dev)
    s = _stream(dev)
    _lib.check(L.ks_sqnorm_device(_p(XA), m, k, _p(XA2), s), "ks_sqnorm_device")
    _lib.check(L.ks_sqnorm_device(_p(XB), n, k, _p(XB2), s), "ks_sqnorm_device")
    ia = torch.arange(m, dtype=torch.int32, device=dev)
    ib = torch.arange(n, dtype=torch.int32, device=dev)
    u = torch.zeros(m, dtype=torch.float64, device=dev)
    kt = KsT()
    kt.type = _lib.KS_GAUSSIAN
    kt.scal = -1.0 / (2.0 * h * h)
    if precision == "mixed_fp32":
        rc = L.dgsks_multirhs_mixed_device(ctypes.byref(kt), m, n, k, 1, _p(u), _p(ia), _p(XA),
                                           _p(XA2), _p(ia), _p(XB), _p(XB2), _p(ib), _p(w),
                                           _p(ib), s)
        _lib.check(rc, "dgsks_multirhs_mixed_device")
        return u
    rc = L.dgsks_device(ctypes.byref(kt), m, n, k, _p(u), _p(ia), _p(XA), _p(XA2), _p(ia),
                        _p(XB), _p(XB2), _p(ib), _p(w), _p(ib), PRECISIONS[precision], s)
    _lib.check(rc, "dgsks_device")
    return u


def gaussian_summation_grad(XA, XB, w, g, h: float, precision: str = "fp64"):
    """(dXA, dXB, dw, dh) of L = g . u through the hand-written HIP gradient
    (FP64, or the mixed form for either mixed precision)."""
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
    L = _lib.lib()
    XA, XB, w, g = (t.contiguous() for t in (XA, XB, w, g))
    m, k = XA.shape
    n = XB.shape[0]
    dev = XA.device
    dXA = torch.empty_like(XA)
    dXB = torch.empty_like(XB)
    dw = torch.empty(n, dtype=torch.float64, device=dev)
    dh = torch.empty(1, dtype=torch.float64, device=dev)
    code = _lib.KS_PREC_FP64 if precision == "fp64" else _lib.KS_PREC_MIXED_FP32
    rc = L.ks_gauss_grad_device(m, n, k, _p(XA), _p(XB), _p(w), _p(g), float(h), _p(dXA),
                                _p(dXB), _p(dw), _p(dh), code, _stream(dev))
    _lib.check(rc, "ks_gauss_grad_device")
    return dXA, dXB, dw, dh


class GaussianSummation(torch.autograd.Function):
    """u = K(XA, XB; h) w.  `h` is a 0-d or 1-element float64 tensor (it may
    require grad); `precision` selects the forward and backward forms."""

    @staticmethod
    def forward(ctx, XA, XB, w, h, precision="fp64"):
        _check(XA, "XA", 2)
        _check(XB, "XB", 2)
        _check(w, "w", 1)
        hv = float(h.detach().reshape(-1)[0].item()) if torch.is_tensor(h) else float(h)
        ctx.save_for_backward(XA, XB, w)
        ctx.h = hv
        ctx.h_is_tensor = torch.is_tensor(h)
        ctx.h_shape = h.shape if torch.is_tensor(h) else None
        ctx.precision = precision
        return gaussian_summation_forward(XA, XB, w, hv, precision)

    @staticmethod
    def backward(ctx, g):
        XA, XB, w = ctx.saved_tensors
        dXA, dXB, dw, dh = gaussian_summation_grad(XA, XB, w, g.to(torch.float64), ctx.h,
                                                   ctx.precision)
        dh_out = dh.reshape(ctx.h_shape) if ctx.h_is_tensor else None
        return dXA, dXB, dw, dh_out, None


def gaussian_summation(XA, XB, w, h, precision="fp64"):
    return GaussianSummation.apply(XA, XB, w, h, precision)
