"""GPU gradient of the Gaussian summation (ks_gauss_grad_device behind
torch.autograd.Function): checked against central finite differences of the
GPU forward (eps = 1e-6, FP64) on random coordinates, and against the FP64
oracle gradient (oracle/gsks_oracle.c orc_gauss_grad) at m = n = 32768,
k = 16, h = 0.4.  Also checks the multi-RHS summation it is built on."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available()
    return torch


def test_central_differences(torch_cuda):
    torch = torch_cuda
    from ks_amd.autograd import gaussian_summation, gaussian_summation_forward
    rng = np.random.default_rng(5)
    m, n, k, h = 300, 257, 16, 0.4
    XA = torch.tensor(rng.random((m, k)), device="cuda", requires_grad=True)
    XB = torch.tensor(rng.random((n, k)), device="cuda", requires_grad=True)
    w = torch.tensor(rng.standard_normal(n), device="cuda", requires_grad=True)
    hh = torch.tensor(h, dtype=torch.float64, device="cuda", requires_grad=True)
    g = torch.tensor(rng.standard_normal(m), device="cuda")
    u = gaussian_summation(XA, XB, w, hh)
    (u * g).sum().backward()

    def L(A, B, ww, hv):
        return float((gaussian_summation_forward(A, B, ww, hv) * g).sum())

    eps = 1e-6
    picks = [(int(i), int(p)) for i, p in zip(rng.integers(0, m, 32), rng.integers(0, k, 32))]
    with torch.no_grad():
        for (i, p) in picks:
            Ap, Am = XA.detach().clone(), XA.detach().clone()
            Ap[i, p] += eps
            Am[i, p] -= eps
            fd = (L(Ap, XB.detach(), w.detach(), h) - L(Am, XB.detach(), w.detach(), h)) / (2 * eps)
            assert XA.grad[i, p].item() == pytest.approx(fd, rel=1e-6, abs=1e-9)
        for (i, p) in picks:
            j = i % n
            Bp, Bm = XB.detach().clone(), XB.detach().clone()
            Bp[j, p] += eps
            Bm[j, p] -= eps
            fd = (L(XA.detach(), Bp, w.detach(), h) - L(XA.detach(), Bm, w.detach(), h)) / (2 * eps)
            assert XB.grad[j, p].item() == pytest.approx(fd, rel=1e-6, abs=1e-9)
        for j in rng.integers(0, n, 8):
            wp, wm = w.detach().clone(), w.detach().clone()
            wp[j] += eps
            wm[j] -= eps
            fd = (L(XA.detach(), XB.detach(), wp, h) - L(XA.detach(), XB.detach(), wm, h)) / (2 * eps)
            assert w.grad[j].item() == pytest.approx(fd, rel=1e-7, abs=1e-10)
        fd = (L(XA.detach(), XB.detach(), w.detach(), h + eps) -
              L(XA.detach(), XB.detach(), w.detach(), h - eps)) / (2 * eps)
        assert hh.grad.item() == pytest.approx(fd, rel=1e-6)


def test_reference_gradient_full_shape(torch_cuda, oracle):
    """m = n = 32768, k = 16, h = 0.4: GPU gradient vs the FP64 oracle gradient
    on 64 sampled rows of dXA and 64 sampled sources of dXB and dw, and dh
    over all 1.07e9 pairs."""
    torch = torch_cuda
    from ks_amd import workloads as W
    from ks_amd.autograd import gaussian_summation_grad
    d = W.grad_gauss16()
    XA, XB, w, g, h = d["XA"], d["XB"], d["w"], d["g"], d["h"]
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    dXA, dXB, dw, dh = gaussian_summation_grad(dev(XA), dev(XB), dev(w), dev(g), h)
    torch.cuda.synchronize()
    rng = np.random.default_rng(6)
    ra = rng.choice(XA.shape[0], 64, replace=False).astype(np.int32)
    rb = rng.choice(XB.shape[0], 64, replace=False).astype(np.int32)
    oA, oB, ow, odh = oracle.gauss_grad(XA, XB, w, g, h, rows_a=ra, rows_b=rb, want_dh=True)
    eA = oracle.rel_l2(dXA.cpu().numpy()[ra], oA)
    eB = oracle.rel_l2(dXB.cpu().numpy()[rb], oB)
    ew = oracle.rel_l2(dw.cpu().numpy()[rb], ow)
    eh = abs(dh.item() - odh) / abs(odh)
    print(f"grad_gauss16 rel errors: dXA {eA:.2e} dXB {eB:.2e} dw {ew:.2e} dh {eh:.2e}")
    assert eA <= 1e-11 and eB <= 1e-11 and ew <= 1e-12 and eh <= 1e-10


def test_multirhs_matches_single_rhs(torch_cuda, oracle):
    """Each column of the multi-RHS result equals a single-RHS summation."""
    torch = torch_cuda
    from ks_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(8)
    nx, m, n, k, R = 2000, 900, 1500, 5, 7
    X = rng.random((nx, k)) * 0.5
    X2 = oracle.sqnorm(X)
    amap = rng.permutation(nx)[:m].astype(np.int32)
    umap = rng.permutation(nx)[:m].astype(np.int32)
    bmap = rng.permutation(nx)[:n].astype(np.int32)
    wmap = rng.permutation(nx)[:n].astype(np.int32)
    Wm = rng.standard_normal((nx, R))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    dX, dX2, da, du_, db, dwm, dW = map(dev, (X, X2, amap, umap, bmap, wmap, Wm))
    U = torch.zeros((nx, R), dtype=torch.float64, device="cuda")
    kt = _lib.KsT()
    kt.type = 2  # Laplace (exact-order cross term path)
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    rc = L.dgsks_multirhs_device(ctypes.byref(kt), m, n, k, R, P(U), P(du_), P(dX), P(dX2), P(da),
                                 P(dX), P(dX2), P(db), P(dW), P(dwm),
                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    Uh = U.cpu().numpy()
    for r in range(R):
        ref = oracle.dgsks_rows(2, X, X2, amap, X, X2, bmap, Wm[:, r].copy(), wmap, np.zeros(nx),
                                umap)
        assert oracle.rel_l2(Uh[umap, r], ref) <= 1e-12, r


@pytest.mark.parametrize("t,k,R", [(0, 16, 7), (1, 40, 16), (3, 64, 23), (4, 32, 32), (5, 20, 3),
                                   (6, 16, 17), (3, 33, 17), (0, 64, 17), (7, 64, 9),
                                   (0, 16, 40)])
def test_multirhs_matrix_core_form(torch_cuda, oracle, t, k, R):
    """The matrix-core multi-RHS form (9 <= k <= 64, not Laplace, R <= 32: the
    sources are the A operand of the cross term, so the kernel values land in
    the A-operand layout of the right-hand-side product) against the oracle,
    column by column, with random maps and ragged sizes; R = 17 adds the last
    right-hand side on the VALU, R = 40 runs the vector form."""
    torch = torch_cuda
    from ks_amd import _lib
    L = _lib.lib()
    prm = {0: dict(scal=-1.5), 1: dict(scal=0.3, cons=0.5, powe=3.0), 3: dict(),
           4: dict(scal=0.2, cons=0.1), 5: dict(), 6: dict(cons=0.5), 7: dict()}[t]
    rng = np.random.default_rng(100 + t + k + R)
    nx, m, n = 2500, 1111, 1777
    X = rng.random((nx, k)) * (1.2 / np.sqrt(k))
    X2 = oracle.sqnorm(X)
    amap = rng.permutation(nx)[:m].astype(np.int32)
    umap = rng.permutation(nx)[:m].astype(np.int32)
    bmap = rng.permutation(nx)[:n].astype(np.int32)
    wmap = rng.permutation(nx)[:n].astype(np.int32)
    Wm = rng.standard_normal((nx, R))
    hv = rng.uniform(1.0, 3.0, nx)
    U0 = rng.standard_normal((nx, R))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    dX, dX2, da, du_, db, dwm, dW, dh, U = map(dev, (X, X2, amap, umap, bmap, wmap, Wm, hv, U0))
    kt = _lib.KsT()
    kt.type = t
    for key, val in prm.items():
        setattr(kt, key, val)
    if t == 3:
        kt.hi = kt.hj = dh.data_ptr()
    P = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
    rc = L.dgsks_multirhs_device(ctypes.byref(kt), m, n, k, R, P(U), P(du_), P(dX), P(dX2), P(da),
                                 P(dX), P(dX2), P(db), P(dW), P(dwm),
                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    Uh = U.cpu().numpy()
    h = hv if t == 3 else None
    worst = 0.0
    for r in range(R):
        ref = oracle.dgsks_rows(t, X, X2, amap, X, X2, bmap, Wm[:, r].copy(), wmap, U0[:, r].copy(),
                                umap, hi=h, hj=h, **prm)
        worst = max(worst, oracle.rel_l2(Uh[umap, r], ref))
    print(f"multirhs type {t} k={k} R={R}: worst column rel L2 = {worst:.3e}")
    assert worst <= 1e-12
    # targets no umap entry names keep their initial values
    rest = np.setdiff1d(np.arange(nx), umap)
    assert np.array_equal(Uh[rest], U0[rest])
