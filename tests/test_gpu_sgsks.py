"""GPU tests of the single-precision entry points (sgsks_device, sgsks,
sgsks_wrapper and the Python sgsks): float32 points, norms, weights and
potentials.  The oracle runs in FP64 on the float inputs widened to double,
and the expected float result is float(u + sum), so the bars are:
  * FP64 summation: relative L2 <= 1e-7 (one float rounding per target,
    2^-24 relative, on top of the FP64 drop-in's 1e-12);
  * mixed-precision summation: relative L2 <= 1e-6 + 2^-24 (the mixed form's
    bar plus that rounding).
The error metric is the reference's (test/test_dgsks.c:43-79)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_F64 = 1e-7
TOL_MIXED = 1e-6 + 2.0 ** -24
TYPES = {0: dict(scal=-20.0), 1: dict(scal=0.5, cons=1.0, powe=3.0), 2: dict(),
         3: dict(), 4: dict(scal=0.5, cons=0.1), 5: dict(), 6: dict(cons=0.5), 7: dict()}


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


@pytest.fixture(scope="module")
def L():
    from ks_amd import _lib
    return _lib.lib()


def _f32_problem(M, N, m, n, k, seed, dup_umap=False):
    rng = np.random.default_rng(seed)
    s = 0.6 / np.sqrt(k) ** 0.5
    XA = (rng.random((M, k)) * s).astype(np.float32)
    XB = (rng.random((N, k)) * s).astype(np.float32)
    w = rng.standard_normal(n + 13).astype(np.float32)
    amap = rng.choice(M, m, replace=False).astype(np.int32)
    bmap = rng.choice(N, n, replace=False).astype(np.int32)
    if dup_umap:
        umap = rng.integers(0, M, m).astype(np.int32)
    else:
        umap = rng.permutation(M)[:m].astype(np.int32)
    wmap = rng.choice(w.size, n, replace=False).astype(np.int32)
    hA = 1.0 / rng.uniform(0.2, 0.4, M)
    hB = 1.0 / rng.uniform(0.2, 0.4, N)
    u0 = rng.standard_normal(M).astype(np.float32)
    return XA, XB, w, amap, bmap, umap, wmap, hA, hB, u0


def _f32_norms(X):
    return np.sum(X.astype(np.float64) ** 2, axis=1).astype(np.float32)


def _kt(t, hi=0, hj=0):
    from ks_amd import _lib
    kt = _lib.KsT()
    kt.type = t
    for key, val in TYPES[t].items():
        setattr(kt, key, val)
    if t == 3:
        kt.hi, kt.hj = hi, hj
    return kt


def _expected(oracle, t, XA, XA2, amap, XB, XB2, bmap, w, wmap, u0, umap, hA=None, hB=None,
              rows=None):
    """float(u0[umap[i]] + sum_j K w) for the selected rows, from the FP64 oracle
    on the widened inputs; None norms are FP64 norms of the widened points.  The
    oracle runs in its clamp mode (d^2 >= 0, dgsks()'s semantics): rounded float
    norms make the d^2 of coincident pairs slightly negative, where the clamp
    matters."""
    XAd, XBd = XA.astype(np.float64), XB.astype(np.float64)
    XA2d = oracle.sqnorm(XAd) if XA2 is None else XA2.astype(np.float64)
    XB2d = oracle.sqnorm(XBd) if XB2 is None else XB2.astype(np.float64)
    opts = dict(TYPES[t])
    if t == 3:
        opts.update(hi=hA, hj=hB)
    ref = oracle.dgsks_rows(t, XAd, XA2d, amap, XBd, XB2d, bmap, w.astype(np.float64), wmap,
                            u0.astype(np.float64), umap, rows=rows, clamp=True, **opts)
    return ref.astype(np.float32).astype(np.float64)


def _device_call(L, torch, kt, k, u, umap, XA, XA2, amap, XB, XB2, bmap, w, wmap, prec):
    P = lambda x: None if x is None else ctypes.c_void_p(x.data_ptr())  # noqa: E731
    rc = L.sgsks_device(ctypes.byref(kt), amap.numel(), bmap.numel(), k, P(u), P(umap), P(XA),
                        P(XA2), P(amap), P(XB), P(XB2), P(bmap), P(w), P(wmap), prec,
                        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc


@pytest.mark.parametrize("k", [3, 16])
@pytest.mark.parametrize("t", list(TYPES))
def test_sgsks_device_fp64_all_types(L, torch_cuda, oracle, t, k):
    torch = torch_cuda
    XA, XB, w, amap, bmap, umap, wmap, hA, hB, u0 = _f32_problem(3300, 2900, 3000, 2500, k,
                                                                 900 + 10 * t + k)
    XA2, XB2 = _f32_norms(XA), _f32_norms(XB)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    dXA, dXB, dXA2, dXB2, dw, dam, dbm, dum, dwm, dhA, dhB, du = map(
        dev, (XA, XB, XA2, XB2, w, amap, bmap, umap, wmap, hA, hB, u0))
    kt = _kt(t, dhA.data_ptr(), dhB.data_ptr())
    rc = _device_call(L, torch, kt, k, du, dum, dXA, dXA2, dam, dXB, dXB2, dbm, dw, dwm, 0)
    assert rc == 0
    torch.cuda.synchronize()
    ref = _expected(oracle, t, XA, XA2, amap, XB, XB2, bmap, w, wmap, u0, umap, hA, hB)
    got = du.cpu().numpy()
    err = oracle.rel_l2(got[umap].astype(np.float64), ref)
    print(f"sgsks_device fp64 type {t} k={k}: rel L2 = {err:.3e}")
    assert err <= TOL_F64
    # rows that no umap entry names are untouched
    rest = np.setdiff1d(np.arange(u0.size), umap)
    assert np.array_equal(got[rest], u0[rest])


@pytest.mark.parametrize("t", [0, 2, 3])
def test_sgsks_device_null_norms(L, torch_cuda, oracle, t):
    """XA2 / XB2 NULL: the norms are computed in FP64 from the float points."""
    torch = torch_cuda
    k = 8
    XA, XB, w, amap, bmap, umap, wmap, hA, hB, u0 = _f32_problem(2100, 2200, 2000, 1900, k, 31 + t)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    dXA, dXB, dw, dam, dbm, dum, dwm, dhA, dhB, du = map(
        dev, (XA, XB, w, amap, bmap, umap, wmap, hA, hB, u0))
    kt = _kt(t, dhA.data_ptr(), dhB.data_ptr())
    assert _device_call(L, torch, kt, k, du, dum, dXA, None, dam, dXB, None, dbm, dw, dwm, 0) == 0
    torch.cuda.synchronize()
    ref = _expected(oracle, t, XA, None, amap, XB, None, bmap, w, wmap, u0, umap, hA, hB)
    err = oracle.rel_l2(du.cpu().numpy()[umap].astype(np.float64), ref)
    print(f"sgsks_device NULL norms type {t}: rel L2 = {err:.3e}")
    assert err <= TOL_F64


@pytest.mark.parametrize("prec", [2, 1])
@pytest.mark.parametrize("t", [0, 3])
def test_sgsks_device_mixed_k64(L, torch_cuda, oracle, t, prec):
    """k = 64 Gaussian mixture, m = n = 20000 (the pivot-bucketed mixed path),
    float inputs; 160 sampled rows against the FP64 oracle."""
    torch = torch_cuda
    from ks_amd import workloads as W
    d = W.mixture_k64(N=20000)
    X = d["X"].astype(np.float32)
    idx = d["amap"]
    w = d["w"].astype(np.float32)
    h = d["hinv"] if t == 3 else None
    X2 = _f32_norms(X)
    u0 = np.zeros(X.shape[0], dtype=np.float32)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    dX, dX2, dw, didx, du = map(dev, (X, X2, w, idx, u0))
    dh = dev(h) if h is not None else None
    kt = _kt(t, dh.data_ptr() if dh is not None else 0, dh.data_ptr() if dh is not None else 0)
    kt.scal = d["kernel"]["scal"]
    assert _device_call(L, torch, kt, 64, du, didx, dX, dX2, didx, dX, dX2, didx, dw, didx,
                        prec) == 0
    torch.cuda.synchronize()
    rows = np.random.default_rng(5).choice(idx.size, 160, replace=False).astype(np.int32)
    Xd = X.astype(np.float64)
    opts = dict(scal=d["kernel"]["scal"])
    if t == 3:
        opts.update(hi=h, hj=h)
    ref = oracle.dgsks_rows(t, Xd, X2.astype(np.float64), idx, Xd, X2.astype(np.float64), idx,
                            w.astype(np.float64), idx, np.zeros(X.shape[0]), idx, rows=rows,
                            clamp=True, **opts)
    got = du.cpu().numpy()[idx[rows]].astype(np.float64)
    err = oracle.rel_l2(got, ref.astype(np.float32).astype(np.float64))
    print(f"sgsks_device mixed prec {prec} type {t} k=64: rel L2 = {err:.3e}")
    assert err <= TOL_MIXED


@pytest.mark.parametrize("t", [0, 2])
def test_sgsks_device_symmetric(L, torch_cuda, oracle, t):
    """One map and one table for both sides: the widened tables are shared and
    the symmetric FP64 kernel runs."""
    torch = torch_cuda
    k, N, M = 3, 5000, 5200
    rng = np.random.default_rng(61 + t)
    XA = (rng.random((M, k)) * 0.45).astype(np.float32)
    w = rng.standard_normal(N + 7).astype(np.float32)
    amap = rng.permutation(M)[:N].astype(np.int32)
    umap = rng.permutation(M)[:N].astype(np.int32)
    wmap = rng.choice(w.size, N, replace=False).astype(np.int32)
    u0 = rng.standard_normal(M).astype(np.float32)
    XA2 = _f32_norms(XA)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    dX, dX2, dw, dam, dum, dwm, du = map(dev, (XA, XA2, w, amap, umap, wmap, u0))
    kt = _kt(t)
    assert _device_call(L, torch, kt, k, du, dum, dX, dX2, dam, dX, dX2, dam, dw, dwm, 0) == 0
    torch.cuda.synchronize()
    ref = _expected(oracle, t, XA, XA2, amap, XA, XA2, amap, w, wmap, u0, umap)
    err = oracle.rel_l2(du.cpu().numpy()[umap].astype(np.float64), ref)
    print(f"sgsks_device symmetric type {t}: rel L2 = {err:.3e}")
    assert err <= TOL_F64


def test_sgsks_host_repeated_umap_and_auto_precision(L, oracle):
    """Host arrays through sgsks(): a repeated umap entry keeps the value of its
    last row (as dgsks does); Polynomial k = 3 and Gaussian k = 16 at these
    sizes sum in FP64 (the automatic choice)."""
    for t, k, tol in ((1, 3, TOL_F64), (0, 16, TOL_F64)):
        XA, XB, w, amap, bmap, umap, wmap, hA, hB, u0 = _f32_problem(2600, 2400, 2500, 2300, k,
                                                                     7 + k, dup_umap=True)
        XA2, XB2 = _f32_norms(XA), _f32_norms(XB)
        u = u0.copy()
        kt = _kt(t)
        p = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
        L.sgsks(ctypes.byref(kt), amap.size, bmap.size, k, p(u), p(umap), p(XA), p(XA2), p(amap),
                p(XB), p(XB2), p(bmap), p(w), p(wmap))
        assert L.ks_last_error() == 0
        rows = _expected(oracle, t, XA, XA2, amap, XB, XB2, bmap, w, wmap, u0, umap)
        want = u0.astype(np.float64).copy()
        for i in range(amap.size):  # row order: the last row of a repeated target wins
            want[umap[i]] = rows[i]
        err = oracle.rel_l2(u.astype(np.float64), want)
        print(f"sgsks host type {t} k={k}: rel L2 = {err:.3e}")
        assert err <= tol


def test_sgsks_wrapper_and_python_api(L, torch_cuda, oracle):
    torch = torch_cuda
    from ks_amd import gsks
    k = 4
    XA, XB, w, amap, bmap, umap, wmap, hA, hB, u0 = _f32_problem(1500, 1700, 1400, 1600, k, 19)
    XA2, XB2 = _f32_norms(XA), _f32_norms(XB)
    ref = _expected(oracle, 0, XA, XA2, amap, XB, XB2, bmap, w, wmap, u0, umap)
    # the wrapper (host arrays)
    u = u0.copy()
    p = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    L.sgsks_wrapper(amap.size, bmap.size, k, p(u), p(umap), p(XA), p(XA2), p(amap), p(XB),
                    p(XB2), p(bmap), p(w), p(wmap), 0, TYPES[0]["scal"], 0.0, 0.0, None)
    assert L.ks_last_error() == 0
    assert oracle.rel_l2(u[umap].astype(np.float64), ref) <= TOL_F64
    # Python, torch tensors on the GPU, norms left to the library
    ref_nn = _expected(oracle, 0, XA, None, amap, XB, None, bmap, w, wmap, u0, umap)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    du = dev(u0)
    gsks.sgsks(k, du, dev(umap), dev(XA), None, dev(amap), dev(XB), None, dev(bmap), dev(w),
               dev(wmap), type="Gaussian", scal=TYPES[0]["scal"], precision="fp64")
    torch.cuda.synchronize()
    assert oracle.rel_l2(du.cpu().numpy()[umap].astype(np.float64), ref_nn) <= TOL_F64
    # numpy arrays, automatic precision
    u = u0.copy()
    gsks.sgsks(k, u, umap, XA, XA2, amap, XB, XB2, bmap, w, wmap, type="Gaussian",
               scal=TYPES[0]["scal"])
    assert oracle.rel_l2(u[umap].astype(np.float64), ref) <= TOL_F64


def test_sgsks_device_rejects_mixed_for_other_kernels(L, torch_cuda):
    torch = torch_cuda
    from ks_amd import _lib
    k = 4
    XA, XB, w, amap, bmap, umap, wmap, hA, hB, u0 = _f32_problem(100, 100, 64, 64, k, 1)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    dXA, dXB, dw, dam, dbm, dum, dwm, du = map(dev, (XA, XB, w, amap, bmap, umap, wmap, u0))
    kt = _kt(1)
    rc = _device_call(L, torch, kt, k, du, dum, dXA, None, dam, dXB, None, dbm, dw, dwm, 2)
    assert rc == _lib.KS_ERR_INVALID_ARG
    assert np.array_equal(du.cpu().numpy(), u0)
