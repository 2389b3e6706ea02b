"""GPU tests of the symmetric path (sources are the targets: m == n, one map for
amap and bmap, one point table and one norm vector), which evaluates each pair
of different row blocks once for both of its terms.  It must agree with the
oracle to relative L2 <= 1e-12 (the reference's metric, test/test_dgsks.c:43-79)
for every kernel type, with any umap / wmap, and with the general path, which
runs when the same values come in distinct arrays."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12
TYPES = {0: dict(scal=-40.0), 1: dict(scal=0.5, cons=1.0, powe=3.0), 2: dict(),
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


def _problem(N, k, seed):
    rng = np.random.default_rng(seed)
    X = rng.random((N + 37, k)) * (0.6 / np.sqrt(k) ** 0.5)
    X2 = np.zeros(X.shape[0])
    for p in range(k):
        X2 += X[:, p] * X[:, p]
    w = rng.standard_normal(N + 11)
    amap = rng.permutation(X.shape[0])[:N].astype(np.int32)
    umap = rng.permutation(N).astype(np.int32)            # any target order
    wmap = rng.choice(w.size, N, replace=False).astype(np.int32)  # any weight order
    h = 1.0 / rng.uniform(0.2, 0.4, X.shape[0])
    return X, X2, w, amap, umap, wmap, h


def _kt(t, h_dev=None):
    from ks_amd import _lib
    kt = _lib.KsT()
    kt.type = t
    for key, val in TYPES[t].items():
        setattr(kt, key, val)
    if t == 3:
        kt.hi = kt.hj = h_dev.data_ptr()
    return kt


def _run(L, torch, kt, N, k, u, umap, X, X2, amap, bmap, w, wmap):
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    s = torch.cuda.current_stream()
    rc = L.dgsks_device(ctypes.byref(kt), N, N, k, P(u), P(umap), P(X), P(X2), P(amap), P(X),
                        P(X2), P(bmap), P(w), P(wmap), 0, ctypes.c_void_p(s.cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()


@pytest.mark.parametrize("k", [2, 3, 4, 8, 16, 40, 64])
@pytest.mark.parametrize("t", list(TYPES))
def test_symmetric_all_types(L, torch_cuda, oracle, t, k):
    """k <= 8 runs the vector kernel, 9 <= k <= 64 the matrix-core one (not
    Laplace, which takes the general path there)."""
    if t == 2 and k < 3:
        pytest.skip("Laplace is defined for k >= 3 (frame/dgsks.c:445-447 divides by k - 2)")
    torch = torch_cuda
    N = 5000  # not a multiple of the 1024-row blocks
    X, X2, w, amap, umap, wmap, h = _problem(N, k, 700 + 10 * t + k)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    dX, dX2, dw, dam, dum, dwm, dh = map(dev, (X, X2, w, amap, umap, wmap, h))
    kt = _kt(t, dh)
    u0 = np.random.default_rng(3).standard_normal(N)
    du = dev(u0)
    _run(L, torch, kt, N, k, du, dum, dX, dX2, dam, dam, dw, dwm)  # symmetric path
    opts = dict(TYPES[t])
    if t == 3:
        opts.update(hi=h, hj=h)
    ref = oracle.dgsks_rows(t, X, X2, amap, X, X2, amap, w, wmap, u0, umap, **opts)
    got = du.cpu().numpy()[umap]
    err = oracle.rel_l2(got, ref)
    print(f"symmetric type {t} k={k}: rel L2 = {err:.3e}")
    assert err <= TOL
    # the general path (bmap a separate copy of amap) agrees
    du2 = dev(u0)
    _run(L, torch, kt, N, k, du2, dum, dX, dX2, dam, dam.clone(), dw, dwm)
    assert oracle.rel_l2(du.cpu().numpy(), du2.cpu().numpy()) <= TOL


def test_symmetric_deterministic(L, torch_cuda):
    torch = torch_cuda
    N, k = 9000, 3
    X, X2, w, amap, umap, wmap, h = _problem(N, k, 77)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    dX, dX2, dw, dam, dum, dwm = map(dev, (X, X2, w, amap, umap, wmap))
    kt = _kt(0)
    outs = []
    for _ in range(2):
        du = torch.zeros(N, dtype=torch.float64, device="cuda")
        _run(L, torch, kt, N, k, du, dum, dX, dX2, dam, dam, dw, dwm)
        outs.append(du.cpu())
    assert torch.equal(outs[0], outs[1])


def test_symmetric_host_path_and_coincident_laplace(L, torch_cuda, oracle):
    """Host arrays with one map object for amap and bmap (the staging keeps one
    device copy, so the symmetric path runs), Laplace with duplicated points:
    coincident pairs (d^2 = 0) are dropped on both sides of the symmetry."""
    from ks_amd import _lib
    N, k = 4100, 3
    X, X2, w, amap, umap, wmap, h = _problem(N, k, 91)
    amap[2000:2010] = amap[0:10]  # ten points appear twice, in different row blocks
    kt = _lib.KsT()
    kt.type = 2
    u = np.zeros(X.shape[0])
    p = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    L.dgsks(ctypes.byref(kt), N, N, k, p(u), p(umap), p(X), p(X2), p(amap), p(X), p(X2), p(amap),
            p(w), p(wmap))
    assert L.ks_last_error() == 0
    ref = oracle.dgsks_rows(2, X, X2, amap, X, X2, amap, w, wmap, np.zeros(X.shape[0]), umap)
    err = oracle.rel_l2(u[umap], ref)
    print(f"symmetric host Laplace with duplicates: rel L2 = {err:.3e}")
    assert err <= TOL
