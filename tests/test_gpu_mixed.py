"""GPU accuracy of the mixed-precision forms (BF16x3 and FP32 cross term,
FP32 kernel values for far pairs, FP64 near-pair recompute and FP64
accumulation) against the FP64 oracle.  The bar is relative L2 <= 1e-6 on the
rows checked (the bar set for the mixed form); the measured error is printed
per kernel type."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-6
PRECS = {"mixed_bf16x3": 2, "mixed_fp32": 1}


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available()
    return torch


def _run(torch, t, prec, X, X2, amap, bmap, w, scal=0.0, h=None, umap=None, wmap=None):
    from ks_amd import _lib
    L = _lib.lib()
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    umap = amap if umap is None else umap
    wmap = bmap if wmap is None else wmap
    dX, dX2, dam, dum, dbm, dwm, dw = map(dev, (X, X2, amap, umap, bmap, wmap, w))
    du = torch.zeros(X.shape[0], dtype=torch.float64, device="cuda")
    kt = _lib.KsT()
    kt.type = t
    kt.scal = scal
    dh = None
    if h is not None:
        dh = dev(h)
        kt.hi = kt.hj = dh.data_ptr()
    P = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
    rc = L.dgsks_device(ctypes.byref(kt), amap.size, bmap.size, X.shape[1], P(du), P(dum), P(dX),
                        P(dX2), P(dam), P(dX), P(dX2), P(dbm), P(dw), P(dwm), prec,
                        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, _lib.error_string(rc)
    torch.cuda.synchronize()
    return du.cpu().numpy()


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("variable", [False, True])
def test_mixture_k64(oracle, torch_cuda, prec, variable):
    """m = n = 131072, k = 64 Gaussian mixture; 192 sampled rows vs the FP64 oracle."""
    from ks_amd import workloads as W
    d = W.mixture_k64()
    X, X2, w, idx = d["X"], d["X2"], d["w"], d["amap"]
    t = 3 if variable else 0
    h = d["hinv"] if variable else None
    u = _run(torch_cuda, t, PRECS[prec], X, X2, idx, idx, w, scal=d["kernel"]["scal"], h=h)
    rows = np.random.default_rng(21).choice(X.shape[0], 192, replace=False).astype(np.int32)
    ref = oracle.dgsks_rows(t, X, X2, idx, X, X2, idx, w, idx, np.zeros(X.shape[0]), idx,
                            rows=rows, scal=d["kernel"]["scal"], hi=h, hj=h)
    err = oracle.rel_l2(u[rows], ref)
    print(f"mixture_k64 {prec} {'var_bandwidth' if variable else 'gaussian'}: rel L2 = {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("k,h", [(5, 0.3), (33, 0.6), (64, 5.0), (100, 0.8), (16, 0.02)])
def test_random_shapes_and_regimes(oracle, torch_cuda, prec, k, h):
    """Ragged shapes, random maps; h = 5 makes every pair near (queue-heavy),
    h = 0.02 makes nearly every pair far."""
    rng = np.random.default_rng(k)
    nx, m, n = 5000, 3001, 4093
    X = rng.random((nx, k)) * (1.0 / np.sqrt(k)) * 2
    X2 = oracle.sqnorm(X)
    amap = rng.permutation(nx)[:m].astype(np.int32)
    umap = rng.permutation(nx)[:m].astype(np.int32)
    bmap = rng.permutation(nx)[:n].astype(np.int32)
    wmap = rng.permutation(nx)[:n].astype(np.int32)
    w = rng.standard_normal(nx)
    scal = -1.0 / (2 * h * h)
    u = _run(torch_cuda, 0, PRECS[prec], X, X2, amap, bmap, w, scal=scal, umap=umap, wmap=wmap)
    ref = oracle.dgsks_rows(0, X, X2, amap, X, X2, bmap, w, wmap, np.zeros(nx), umap, scal=scal)
    err = oracle.rel_l2(u[umap], ref)
    print(f"random k={k} h={h} {prec}: rel L2 = {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("x2", [-40.0, -200.0])
def test_far_only_rows_and_skip_cutoff(oracle, torch_cuda, prec, x2):
    """Targets in one cluster, sources in another, every kernel value about
    2^x2.  Rows made only of far pairs take K in FP32 from an FP32 exponent
    (about 1e-4 relative), so the bar here is 1e-3; a block skipped in error
    would lose its rows' whole value.  At 2^-200 the BF16x3 Gaussian skip tier
    may drop every block: each target then stays within the documented cutoff
    n 2^-50 max|w|."""
    rng = np.random.default_rng(int(-x2))
    k, m, n, sig, scal = 64, 20000, 18000, 0.05, -1.0  # m, n >= 16384: bucketed
    d2 = x2 / (scal * np.log2(np.e))  # squared distance that gives K = 2^x2
    D = np.sqrt(d2 - 2 * k * sig * sig)
    ca = np.full(k, 0.3)
    cb = ca.copy()
    cb[0] += D
    X = np.vstack([ca + sig * rng.standard_normal((m, k)), cb + sig * rng.standard_normal((n, k))])
    X2 = oracle.sqnorm(X)
    amap = np.arange(m, dtype=np.int32)
    bmap = np.arange(m, m + n, dtype=np.int32)
    w = rng.standard_normal(m + n)
    u = _run(torch_cuda, 0, PRECS[prec], X, X2, amap, bmap, w, scal=scal)
    rows = np.random.default_rng(3).choice(m, 160, replace=False).astype(np.int32)
    ref = oracle.dgsks_rows(0, X, X2, amap, X, X2, bmap, w, bmap, np.zeros(m + n), amap,
                            rows=rows, scal=scal)
    if x2 > -50:
        err = oracle.rel_l2(u[rows], ref)
        print(f"far-only rows 2^{x2:g} {prec}: rel L2 = {err:.3e}")
        assert err <= 1e-3
    else:
        bound = n * 2.0 ** -50 * np.abs(w).max()
        print(f"deep-far rows 2^{x2:g} {prec}: max |u| = {np.abs(u[:m]).max():.3e} "
              f"(cutoff bound {bound:.3e}, FP64 {np.abs(ref).max():.3e})")
        assert np.abs(u[:m]).max() <= bound


def test_unsupported_type_reports(torch_cuda):
    """The mixed forms cover the Gaussian kernels only: tanh at mixed precision
    returns KS_ERR_INVALID_ARG (error mode RETURN), records it for
    ks_last_error() and leaves u untouched."""
    from ks_amd import _lib
    torch = torch_cuda
    L = _lib.lib()
    rng = np.random.default_rng(0)
    X = rng.random((100, 8))
    X2 = (X * X).sum(1)
    idx = np.arange(100, dtype=np.int32)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    dX, dX2, di, dw = map(dev, (X, X2, idx, np.ones(100)))
    du = torch.full((100,), 7.0, dtype=torch.float64, device="cuda")
    kt = _lib.KsT()
    kt.type = 4
    kt.scal = 0.1
    P = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
    mode = L.ks_get_error_mode()
    L.ks_set_error_mode(_lib.KS_ERRMODE_RETURN)
    try:
        L.ks_last_error()  # clear
        rc = L.dgsks_device(ctypes.byref(kt), 100, 100, 8, P(du), P(di), P(dX), P(dX2), P(di),
                            P(dX), P(dX2), P(di), P(dw), P(di), 2,
                            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert rc == _lib.KS_ERR_INVALID_ARG
        assert L.ks_last_error() == _lib.KS_ERR_INVALID_ARG
    finally:
        L.ks_set_error_mode(mode)
    torch.cuda.synchronize()
    assert torch.all(du == 7.0)


def test_many_row_blocks(oracle, torch_cuda):
    """270000 clustered targets (8438 row blocks of 32 per slice: the dispatch
    order by radix sort instead of by rank in LDS), 20000 sources, k = 8,
    BF16x3 Gaussian with the pruning bounds: 192 sampled rows against the FP64
    oracle, and the same bits on a second call."""
    rng = np.random.default_rng(33)
    m, n, k = 270000, 20000, 8
    C = rng.random((16, k))
    X = C[rng.integers(0, 16, m + n)] + 0.05 * rng.standard_normal((m + n, k))
    X2 = np.einsum("ij,ij->i", X, X)
    w = rng.standard_normal(m + n)
    amap = np.arange(m, dtype=np.int32)
    bmap = np.arange(m, m + n, dtype=np.int32)
    scal = -1.0 / (2 * 0.1 ** 2)
    u = _run(torch_cuda, 0, 2, X, X2, amap, bmap, w, scal=scal)
    u2 = _run(torch_cuda, 0, 2, X, X2, amap, bmap, w, scal=scal)
    assert np.array_equal(u, u2)
    rows = np.sort(rng.choice(m, 192, replace=False)).astype(np.int32)
    ref = oracle.dgsks_rows(0, X, X2, amap, X, X2, bmap, w, bmap, np.zeros(m + n), amap,
                            rows=rows, scal=scal)
    err = oracle.rel_l2(u[rows], ref)
    print(f"many row blocks m={m} n={n} k={k}: rel L2 = {err:.3e}")
    assert err <= TOL
