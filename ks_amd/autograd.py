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

MappedSummation covers index maps (XA[amap[i]], XB[bmap[j]], w[wmap[j]]) and
the variable-bandwidth Gaussian with its per-point bandwidths: forward through
dgsks_device, backward through ks_grad_device, which adds each row's gradient
into the tables at the mapped entries.
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


# ------------------------------------------------------------ mapped summations

def _i32(t, name):
    if not (t.is_cuda and t.dtype == torch.int32 and t.dim() == 1):
        raise AssertionError(f"{name}: int32 CUDA vector expected")
    return t.contiguous()


def mapped_summation_grad(ktype, g, umap, XA, amap, XB, bmap, w, wmap, *, scal=0.0, hi=None,
                          hj=None, precision="fp64", want=("dXA", "dXB", "dw", "dparam")):
    """Gradient tables of L = g . du, du[umap[i]] += sum_j K(XA[amap[i]],
    XB[bmap[j]]) w[wmap[j]] (ks_grad_device), accumulated from zero: a dict
    with dXA, dXB, dw and dscal (KS_GAUSSIAN, a 1-element tensor) or dhi, dhj
    (KS_GAUSSIAN_VAR_BANDWIDTH) for the names in `want` ("dparam" is dscal or
    dhi and dhj)."""
    L = _lib.lib()
    var = ktype == _lib.KS_GAUSSIAN_VAR_BANDWIDTH
    XA, XB, w, g = (t.contiguous() for t in (XA, XB, w, g))
    umap, amap, bmap, wmap = (_i32(t, nm) for t, nm in
                              ((umap, "umap"), (amap, "amap"), (bmap, "bmap"), (wmap, "wmap")))
    k = XA.shape[1]
    dev = XA.device
    out = {}
    if "dXA" in want:
        out["dXA"] = torch.zeros_like(XA)
    if "dXB" in want:
        out["dXB"] = torch.zeros_like(XB)
    if "dw" in want:
        out["dw"] = torch.zeros_like(w)
    kt = KsT()
    kt.type = ktype
    keep = []
    if var:
        hi, hj = hi.contiguous(), hj.contiguous()
        keep += [hi, hj]
        kt.hi, kt.hj = hi.data_ptr(), hj.data_ptr()
        if "dparam" in want:
            out["dhi"] = torch.zeros_like(hi)
            out["dhj"] = torch.zeros_like(hj)
    else:
        kt.scal = float(scal)
        if "dparam" in want:
            out["dscal"] = torch.zeros(1, dtype=torch.float64, device=dev)
    ptr = lambda nm: _p(out[nm]) if nm in out else None  # noqa: E731
    rc = L.ks_grad_device(ctypes.byref(kt), amap.numel(), bmap.numel(), k, _p(g), _p(umap),
                          _p(XA), _p(amap), _p(XB), _p(bmap), _p(w), _p(wmap), ptr("dXA"),
                          ptr("dXB"), ptr("dw"), ptr("dscal"), ptr("dhi"), ptr("dhj"),
                          _prec_code(precision), _stream(dev))
    _lib.check(rc, "ks_grad_device")
    return out


def mapped_summation_forward(ktype, XA, amap, XB, bmap, w, wmap, *, scal=0.0, hi=None,
                             hj=None, precision="fp64"):
    """u_i = sum_j K(XA[amap[i]], XB[bmap[j]]) w[wmap[j]], one row per amap
    entry (dgsks_device with an identity umap; no autograd)."""
    L = _lib.lib()
    XA, XB, w = (t.contiguous() for t in (XA, XB, w))
    amap, bmap, wmap = (_i32(t, nm) for t, nm in ((amap, "amap"), (bmap, "bmap"), (wmap, "wmap")))
    m, n, k = amap.numel(), bmap.numel(), XA.shape[1]
    dev = XA.device
    s = _stream(dev)
    XA2 = torch.empty(XA.shape[0], dtype=torch.float64, device=dev)
    XB2 = torch.empty(XB.shape[0], dtype=torch.float64, device=dev)
    _lib.check(L.ks_sqnorm_device(_p(XA), XA.shape[0], k, _p(XA2), s), "ks_sqnorm_device")
    _lib.check(L.ks_sqnorm_device(_p(XB), XB.shape[0], k, _p(XB2), s), "ks_sqnorm_device")
    iu = torch.arange(m, dtype=torch.int32, device=dev)
    u = torch.zeros(m, dtype=torch.float64, device=dev)
    kt = KsT()
    kt.type = ktype
    keep = []
    if ktype == _lib.KS_GAUSSIAN_VAR_BANDWIDTH:
        hi, hj = hi.contiguous(), hj.contiguous()
        keep += [hi, hj]
        kt.hi, kt.hj = hi.data_ptr(), hj.data_ptr()
    else:
        kt.scal = float(scal)
    rc = L.dgsks_device(ctypes.byref(kt), m, n, k, _p(u), _p(iu), _p(XA), _p(XA2), _p(amap),
                        _p(XB), _p(XB2), _p(bmap), _p(w), _p(wmap), PRECISIONS[precision], s)
    _lib.check(rc, "dgsks_device")
    return u


class MappedSummation(torch.autograd.Function):
    """u_i = sum_j K(XA[amap[i]], XB[bmap[j]]) w[wmap[j]] for the Gaussian
    (`param` = scal, a 1-element float64 tensor) or the variable-bandwidth
    Gaussian (`param` = hi, `param_b` = hj, per-point bandwidth vectors).  The
    same tensor may be passed as XA and XB, or as hi and hj: autograd then sums
    both sides' gradients.  Backward through ks_grad_device ("fp64"; the
    Gaussian's mixed precisions through the mixed gradient, the variable
    bandwidth's in FP64 whatever the forward's precision)."""

    @staticmethod
    def forward(ctx, XA, XB, w, param, param_b, amap, bmap, wmap, ktype, precision):
        _check(XA, "XA", 2)
        _check(XB, "XB", 2)
        _check(w, "w", 1)
        var = ktype == _lib.KS_GAUSSIAN_VAR_BANDWIDTH
        ctx.ktype, ctx.var = ktype, var
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        # the mixed gradient covers the Gaussian; the variable bandwidth's is FP64
        ctx.precision = "fp64" if var else precision
        if var:
            u = mapped_summation_forward(ktype, XA, amap, XB, bmap, w, wmap, hi=param,
                                         hj=param_b, precision=precision)
        else:
            ctx.scal = float(param.detach().reshape(-1)[0].item())
            u = mapped_summation_forward(ktype, XA, amap, XB, bmap, w, wmap, scal=ctx.scal,
                                         precision=precision)
        ctx.save_for_backward(XA, XB, w, param, param_b, amap, bmap, wmap)
        return u

    @staticmethod
    def backward(ctx, g):
        XA, XB, w, param, param_b, amap, bmap, wmap = ctx.saved_tensors
        need = ctx.needs_input_grad
        want = tuple(nm for nm, f in (("dXA", need[0]), ("dXB", need[1]), ("dw", need[2]),
                                      ("dparam", need[3] or need[4])) if f)
        iu = torch.arange(amap.numel(), dtype=torch.int32, device=g.device)
        bw = {"hi": param, "hj": param_b} if ctx.var else {"scal": ctx.scal}
        G = mapped_summation_grad(ctx.ktype, g.to(torch.float64).contiguous(), iu, XA, amap, XB,
                                  bmap, w, wmap, precision=ctx.precision, want=want, **bw)
        if ctx.var:
            dp, dpb = G.get("dhi"), G.get("dhj")
        else:
            dp = G["dscal"].reshape(param.shape) if "dscal" in G else None
            dpb = None
        return (G.get("dXA"), G.get("dXB"), G.get("dw"), dp, dpb, None, None, None, None, None)


def mapped_gaussian_summation(XA, XB, w, scal, amap, bmap, wmap, precision="fp64"):
    """Gaussian summation over index maps, differentiable in XA, XB, w and the
    1-element scal tensor."""
    return MappedSummation.apply(XA, XB, w, scal, None, amap, bmap, wmap, _lib.KS_GAUSSIAN,
                                 precision)


def mapped_var_bandwidth_summation(XA, XB, w, hi, hj, amap, bmap, wmap, precision="fp64"):
    """Variable-bandwidth Gaussian summation over index maps, differentiable in
    XA, XB, w and the bandwidth vectors hi, hj."""
    return MappedSummation.apply(XA, XB, w, hi, hj, amap, bmap, wmap,
                                 _lib.KS_GAUSSIAN_VAR_BANDWIDTH, precision)
