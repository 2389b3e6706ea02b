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

When XA or h needs a gradient (and grad mode is on), the "fp64" and
"mixed_fp32" forwards run the gradient's first summation instead
(ks_gauss_grad_forward_device: k + 1 right-hand sides [w b, w] -> [V, u]) and
keep V, so the backward runs the second summation only: a training step costs
two summations, not three.  Otherwise the forward is the plain summation and
the backward, which then needs no V, runs the second summation only.
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


def _prec_code(precision):
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
    return _lib.KS_PREC_FP64 if precision == "fp64" else _lib.KS_PREC_MIXED_FP32


def gaussian_summation_grad(XA, XB, w, g, h: float, precision: str = "fp64"):
    """(dXA, dXB, dw, dh) of L = g . u through the hand-written HIP gradient
    (FP64, or the mixed form for either mixed precision): both summations."""
    code = _prec_code(precision)
    L = _lib.lib()
    XA, XB, w, g = (t.contiguous() for t in (XA, XB, w, g))
    m, k = XA.shape
    n = XB.shape[0]
    dev = XA.device
    dXA = torch.empty_like(XA)
    dXB = torch.empty_like(XB)
    dw = torch.empty(n, dtype=torch.float64, device=dev)
    dh = torch.empty(1, dtype=torch.float64, device=dev)
    rc = L.ks_gauss_grad_device(m, n, k, _p(XA), _p(XB), _p(w), _p(g), float(h), _p(dXA),
                                _p(dXB), _p(dw), _p(dh), code, _stream(dev))
    _lib.check(rc, "ks_gauss_grad_device")
    return dXA, dXB, dw, dh


def gaussian_summation_forward_state(XA, XB, w, h: float, precision: str = "fp64"):
    """(u, state): u = K w and the gradient state ([V | u] rows and the frame
    origin, m (k + 1) + k doubles) from the gradient's first summation."""
    code = _prec_code(precision)
    L = _lib.lib()
    XA, XB, w = (t.contiguous() for t in (XA, XB, w))
    m, k = XA.shape
    n = XB.shape[0]
    dev = XA.device
    u = torch.empty(m, dtype=torch.float64, device=dev)
    state = torch.empty(m * (k + 1) + k, dtype=torch.float64, device=dev)
    rc = L.ks_gauss_grad_forward_device(m, n, k, _p(XA), _p(XB), _p(w), float(h), _p(u),
                                        _p(state), code, _stream(dev))
    _lib.check(rc, "ks_gauss_grad_forward_device")
    return u, state


def gaussian_summation_backward(XA, XB, w, g, h: float, state=None, precision: str = "fp64",
                                want=(True, True, True, True)):
    """The gradient's second summation and finishing kernels: (dXA, dXB, dw,
    dh), None where `want` is False.  dXA and dh need the state that
    gaussian_summation_forward_state returned for the same inputs."""
    code = _prec_code(precision)
    L = _lib.lib()
    XA, XB, w, g = (t.contiguous() for t in (XA, XB, w, g))
    m, k = XA.shape
    n = XB.shape[0]
    dev = XA.device
    if (want[0] or want[3]) and state is None:
        raise ValueError("dXA and dh need the forward state")
    out = [torch.empty_like(XA) if want[0] else None,
           torch.empty_like(XB) if want[1] else None,
           torch.empty(n, dtype=torch.float64, device=dev) if want[2] else None,
           torch.empty(1, dtype=torch.float64, device=dev) if want[3] else None]
    ptr = [_p(t) if t is not None else None for t in out]
    rc = L.ks_gauss_grad_backward_device(m, n, k, _p(XA), _p(XB), _p(w), _p(g), float(h),
                                         _p(state) if state is not None else None, *ptr, code,
                                         _stream(dev))
    _lib.check(rc, "ks_gauss_grad_backward_device")
    return tuple(out)


class GaussianSummation(torch.autograd.Function):
    """u = K(XA, XB; h) w.  `h` is a 0-d or 1-element float64 tensor (it may
    require grad); `precision` selects the forward and backward forms;
    `for_grad` (grad mode was on) lets the forward keep the gradient state."""

    @staticmethod
    def forward(ctx, XA, XB, w, h, precision="fp64", for_grad=True):
        _check(XA, "XA", 2)
        _check(XB, "XB", 2)
        _check(w, "w", 1)
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        hv = float(h.detach().reshape(-1)[0].item()) if torch.is_tensor(h) else float(h)
        need = ctx.needs_input_grad
        ctx.h = hv
        ctx.h_is_tensor = torch.is_tensor(h)
        ctx.h_shape = h.shape if torch.is_tensor(h) else None
        ctx.precision = precision
        ctx.with_state = bool(for_grad and (need[0] or need[3]) and precision != "mixed_bf16x3")
        if ctx.with_state:
            u, state = gaussian_summation_forward_state(XA, XB, w, hv, precision)
            ctx.save_for_backward(XA, XB, w, state)
            return u
        ctx.save_for_backward(XA, XB, w)
        return gaussian_summation_forward(XA, XB, w, hv, precision)

    @staticmethod
    def backward(ctx, g):
        need = ctx.needs_input_grad
        g = g.to(torch.float64)
        if ctx.with_state:
            XA, XB, w, state = ctx.saved_tensors
            dXA, dXB, dw, dh = gaussian_summation_backward(
                XA, XB, w, g, ctx.h, state, ctx.precision, want=tuple(need[:4]))
        elif need[0] or need[3]:  # mixed_bf16x3: its forward kept no state
            XA, XB, w = ctx.saved_tensors
            dXA, dXB, dw, dh = gaussian_summation_grad(XA, XB, w, g, ctx.h, ctx.precision)
        else:
            XA, XB, w = ctx.saved_tensors
            dXA, dXB, dw, dh = gaussian_summation_backward(
                XA, XB, w, g, ctx.h, None, ctx.precision, want=(False, need[1], need[2], False))
        dh_out = dh.reshape(ctx.h_shape) if (ctx.h_is_tensor and dh is not None) else None
        return dXA, dXB, dw, dh_out, None, None


def gaussian_summation(XA, XB, w, h, precision="fp64"):
    return GaussianSummation.apply(XA, XB, w, h, precision, torch.is_grad_enabled())
