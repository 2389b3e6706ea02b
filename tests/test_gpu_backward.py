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


def test_central_differences_full_shape(torch_cuda):
    """m = n = 32768, k = 16, h = 0.4 (the gradient workload's points, weights
    and g): the autograd gradient, whose forward keeps the state of the first
    summation, against central differences of the FP64 forward (eps = 1e-6) on
    64 seeded coordinates: 24 of XA, 24 of XB, 15 of w and h.  Each difference
    is g . (u+ - u-), the unchanged rows cancelling exactly (the forward is
    deterministic), so its rounding stays far below the 1e-6 bar."""
    torch = torch_cuda
    from ks_amd import workloads as W
    from ks_amd.autograd import gaussian_summation, gaussian_summation_forward
    d = W.grad_gauss16()
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    XA, XB, w, g = dev(d["XA"]), dev(d["XB"]), dev(d["w"]), dev(d["g"])
    h, (m, k), n = d["h"], d["XA"].shape, d["XB"].shape[0]
    XA.requires_grad_(True)
    XB.requires_grad_(True)
    w.requires_grad_(True)
    hh = torch.tensor(h, dtype=torch.float64, device="cuda", requires_grad=True)
    u = gaussian_summation(XA, XB, w, hh)
    (u * g).sum().backward()
    grads = {"XA": XA.grad.clone(), "XB": XB.grad.clone(), "w": w.grad.clone(),
             "h": hh.grad.reshape(1).clone()}
    A, B, ww = XA.detach().clone(), XB.detach().clone(), w.detach().clone()
    u0 = gaussian_summation_forward(A, B, ww, h)
    assert float(torch.linalg.norm(u.detach() - u0) / torch.linalg.norm(u0)) <= 1e-14

    rng = np.random.default_rng(21)
    eps = 1e-6
    worst = {}

    def diff(t, idx, hv=h):
        x = t[idx].item()
        t[idx] = x + eps
        up = gaussian_summation_forward(A, B, ww, hv)
        t[idx] = x - eps
        um = gaussian_summation_forward(A, B, ww, hv)
        t[idx] = x
        return float(torch.dot(g, up - um)) / (2 * eps)

    picks = ([("XA", (int(i), int(p))) for i, p in zip(rng.integers(0, m, 24),
                                                        rng.integers(0, k, 24))] +
             [("XB", (int(j), int(p))) for j, p in zip(rng.integers(0, n, 24),
                                                        rng.integers(0, k, 24))] +
             [("w", (int(j),)) for j in rng.integers(0, n, 15)])
    with torch.no_grad():
        for name, idx in picks:
            fd = diff({"XA": A, "XB": B, "w": ww}[name], idx)
            an = grads[name][idx].item()
            scale = float(grads[name].abs().max())
            err = abs(an - fd) / max(abs(fd), 1e-3 * scale)
            worst[name] = max(worst.get(name, 0.0), err)
            assert err <= 1e-6, (name, idx, an, fd)
        fd = (float(torch.dot(g, gaussian_summation_forward(A, B, ww, h + eps))) -
              float(torch.dot(g, gaussian_summation_forward(A, B, ww, h - eps)))) / (2 * eps)
        worst["h"] = abs(grads["h"].item() - fd) / abs(fd)
        assert worst["h"] <= 1e-6, (grads["h"].item(), fd)
    print("full-shape central differences, worst relative error:",
          {a: f"{b:.1e}" for a, b in worst.items()})


def test_gradient_halves(torch_cuda):
    """The forward half (u and the state) and the backward half against the
    one-shot gradient, which runs the same two summations: identical bits;
    u against the plain summation; the backward without a state (dXB and dw
    only) identical too; autograd with only XB and w requiring grad."""
    torch = torch_cuda
    from ks_amd import autograd as AG
    rng = np.random.default_rng(31)
    m, n, k, h = 1500, 1237, 12, 0.5
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    XA, XB = dev(rng.random((m, k))), dev(rng.random((n, k)))
    w, g = dev(rng.standard_normal(n)), dev(rng.standard_normal(m))
    for prec in ("fp64", "mixed_fp32"):
        u, state = AG.gaussian_summation_forward_state(XA, XB, w, h, prec)
        ref_u = AG.gaussian_summation_forward(XA, XB, w, h, "fp64")
        tol = 1e-14 if prec == "fp64" else 1e-6
        assert float(torch.linalg.norm(u - ref_u) / torch.linalg.norm(ref_u)) <= tol, prec
        assert torch.equal(state[m * (k + 1) - 1::k + 1][:1], u[-1:])
        halves = AG.gaussian_summation_backward(XA, XB, w, g, h, state, prec)
        whole = AG.gaussian_summation_grad(XA, XB, w, g, h, prec)
        for a, b in zip(halves, whole):
            assert torch.equal(a, b), prec
        part = AG.gaussian_summation_backward(XA, XB, w, g, h, None, prec,
                                              want=(False, True, True, False))
        assert part[0] is None and part[3] is None
        assert torch.equal(part[1], whole[1]) and torch.equal(part[2], whole[2]), prec
        with pytest.raises(ValueError):
            AG.gaussian_summation_backward(XA, XB, w, g, h, None, prec)
    XBr, wr = XB.clone().requires_grad_(True), w.clone().requires_grad_(True)
    (AG.gaussian_summation(XA, XBr, wr, h) * g).sum().backward()
    whole = AG.gaussian_summation_grad(XA, XB, w, g, h, "fp64")
    assert torch.equal(XBr.grad, whole[1]) and torch.equal(wr.grad, whole[2])


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
                                   (0, 16, 40), (0, 20, 18), (3, 16, 18), (3, 48, 19),
                                   (0, 16, 19), (1, 16, 19)])
def test_multirhs_matrix_core_form(torch_cuda, oracle, t, k, R):
    """The matrix-core multi-RHS form (9 <= k <= 64, not Laplace, R <= 32: the
    sources are the A operand of the cross term, so the kernel values land in
    the A-operand layout of the right-hand-side product) against the oracle,
    column by column, with random maps and ragged sizes; R = 17 adds the last
    right-hand side on the VALU (R = 18 and 19 the last two and three for the
    Gaussian kernels; a polynomial at 19 takes two tiles), R = 40 runs the
    vector form."""
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
