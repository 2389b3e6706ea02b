"""GPU gradient of the mapped Gaussian and variable-bandwidth summations
(ks_grad_device, ks_amd.autograd.MappedSummation) against the FP64 oracle
gradient (oracle/gsks_oracle.c orc_ks_grad, itself pinned against central
differences of the oracle forward) on index maps with repeated entries and
one point table and bandwidth vector on both sides, against the identity-map
gradient (ks_gauss_grad_device), and against central differences of the GPU
forward at m = n = 32768, k = 16."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available()
    return torch


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _problem(seed, var, k, nx=2000, m=1500, n=1700):
    rng = np.random.default_rng(seed)
    X = rng.random((nx, k))
    d = {
        "X": X,
        "h": (0.5 + rng.random(nx)) * np.sqrt(6.0 / k),
        "w": rng.standard_normal(nx),
        "g": rng.standard_normal(nx),
        "amap": rng.integers(0, nx, m).astype(np.int32),  # repeated entries on purpose
        "bmap": rng.integers(0, nx, n).astype(np.int32),
        "wmap": rng.integers(0, nx, n).astype(np.int32),
        "umap": rng.integers(0, nx, m).astype(np.int32),
        "scal": -3.0 / k,
        "var": var,
    }
    return d


def _gpu_grad(torch, d, precision=0, shared_out=True):
    """ks_grad_device with XA == XB and hi == hj; shared_out adds both sides
    into one table (dXA == dXB, dhi == dhj)."""
    from ks_amd import _lib
    L = _lib.lib()
    X, h, w, g = (_dev(torch, d[nm]) for nm in ("X", "h", "w", "g"))
    am, bm, wm, um = (_dev(torch, d[nm]) for nm in ("amap", "bmap", "wmap", "umap"))
    dXA = torch.zeros_like(X)
    dXB = dXA if shared_out else torch.zeros_like(X)
    dw = torch.zeros_like(w)
    kt = _lib.KsT()
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    out = {"dXA": dXA, "dXB": dXB, "dw": dw}
    if d["var"]:
        kt.type = _lib.KS_GAUSSIAN_VAR_BANDWIDTH
        kt.hi = kt.hj = h.data_ptr()
        dhi = torch.zeros_like(h)
        dhj = dhi if shared_out else torch.zeros_like(h)
        out.update(dhi=dhi, dhj=dhj)
        extra = (None, P(dhi), P(dhj))
    else:
        kt.type = _lib.KS_GAUSSIAN
        kt.scal = d["scal"]
        dscal = torch.zeros(1, dtype=torch.float64, device="cuda")
        out["dscal"] = dscal
        extra = (P(dscal), None, None)
    rc = L.ks_grad_device(ctypes.byref(kt), d["amap"].size, d["bmap"].size, X.shape[1], P(g),
                          P(um), P(X), P(am), P(X), P(bm), P(w), P(wm), P(dXA), P(dXB), P(dw),
                          *extra, precision,
                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, _lib.error_string(rc)
    torch.cuda.synchronize()
    return {nm: t.cpu().numpy() for nm, t in out.items()}


def _oracle_grad(oracle, d):
    T = oracle.VAR_BANDWIDTH if d["var"] else oracle.GAUSSIAN
    return oracle.ks_grad(T, d["g"], d["umap"], d["X"], d["amap"], d["X"], d["bmap"], d["w"],
                          d["wmap"], scal=d["scal"], hi=d["h"] if d["var"] else None,
                          hj=d["h"] if d["var"] else None)


@pytest.mark.parametrize("var", [False, True], ids=["gaussian", "var_bandwidth"])
@pytest.mark.parametrize("k", [3, 16, 40])
def test_mapped_grad_vs_oracle(torch_cuda, oracle, var, k):
    """Repeated map entries, one table on both sides, both sides' terms added
    into one output table: every entry against the oracle at 1e-11."""
    d = _problem(k + 100 * var, var, k)
    got = _gpu_grad(torch_cuda, d)
    ref = _oracle_grad(oracle, d)
    errs = {"dX": oracle.rel_l2(got["dXA"], ref["dXA"] + ref["dXB"]),
            "dw": oracle.rel_l2(got["dw"], ref["dw"])}
    if var:
        errs["dh"] = oracle.rel_l2(got["dhi"], ref["dhi"] + ref["dhj"])
    else:
        errs["dscal"] = abs(got["dscal"][0] - ref["dscal"]) / abs(ref["dscal"])
    print(f"mapped grad {'var' if var else 'gauss'} k={k}: " +
          " ".join(f"{a} {e:.2e}" for a, e in errs.items()))
    assert max(errs.values()) <= 1e-11
    # entries no map selects stay exactly zero
    unused = np.setdiff1d(np.arange(d["X"].shape[0]), np.union1d(d["amap"], d["bmap"]))
    assert not np.any(got["dXA"][unused])


@pytest.mark.parametrize("var", [False, True], ids=["gaussian", "var_bandwidth"])
def test_mapped_grad_separate_tables_and_determinism(torch_cuda, oracle, var):
    """Separate output tables receive each side's terms alone; two calls give
    the same bits (stable map sort, fixed-order sums)."""
    d = _problem(7 + var, var, 16)
    a = _gpu_grad(torch_cuda, d, shared_out=False)
    b = _gpu_grad(torch_cuda, d, shared_out=False)
    for nm in a:
        assert np.array_equal(a[nm], b[nm]), nm
    ref = _oracle_grad(oracle, d)
    for nm in ("dXA", "dXB", "dw") + (("dhi", "dhj") if var else ()):
        assert oracle.rel_l2(a[nm], ref[nm]) <= 1e-11, nm


def test_mapped_grad_mixed(torch_cuda, oracle):
    """KS_PREC_MIXED_FP32 (Gaussian, k = 16) against the oracle at 1e-6."""
    d = _problem(31, False, 16, nx=6000, m=5000, n=5500)
    got = _gpu_grad(torch_cuda, d, precision=1)
    ref = _oracle_grad(oracle, d)
    e = max(oracle.rel_l2(got["dXA"], ref["dXA"] + ref["dXB"]),
            oracle.rel_l2(got["dw"], ref["dw"]),
            abs(got["dscal"][0] - ref["dscal"]) / abs(ref["dscal"]))
    print(f"mapped grad mixed_fp32 k=16: rel {e:.2e}")
    assert e <= 1e-6


def test_identity_maps_match_fixed_bandwidth_gradient(torch_cuda):
    """Identity maps, scal = -1/(2h^2): dXA, dXB, dw as ks_gauss_grad_device
    gives them, and dh = dscal / h^3."""
    torch = torch_cuda
    from ks_amd.autograd import gaussian_summation_grad, mapped_summation_grad
    from ks_amd import _lib
    rng = np.random.default_rng(3)
    m, n, k, h = 3000, 2500, 16, 0.4
    XA, XB = _dev(torch, rng.random((m, k))), _dev(torch, rng.random((n, k)))
    w, g = _dev(torch, rng.standard_normal(n)), _dev(torch, rng.standard_normal(m))
    ia = torch.arange(m, dtype=torch.int32, device="cuda")
    ib = torch.arange(n, dtype=torch.int32, device="cuda")
    G = mapped_summation_grad(_lib.KS_GAUSSIAN, g, ia, XA, ia, XB, ib, w, ib,
                              scal=-1.0 / (2 * h * h))
    dXA, dXB, dw, dh = gaussian_summation_grad(XA, XB, w, g, h)
    rel = lambda a, b: float(torch.linalg.norm(a - b) / torch.linalg.norm(b))  # noqa: E731
    assert rel(G["dXA"], dXA) <= 1e-12 and rel(G["dXB"], dXB) <= 1e-12
    assert rel(G["dw"], dw) <= 1e-13
    assert float(G["dscal"][0]) / h ** 3 == pytest.approx(float(dh[0]), rel=1e-12)


@pytest.mark.parametrize("var", [False, True], ids=["gaussian", "var_bandwidth"])
def test_autograd_central_differences_full_shape(torch_cuda, var):
    """m = n = 32768 over a 40000-point pool, k = 16, random maps with repeats,
    one table and (variable bandwidth) one bandwidth vector on both sides:
    MappedSummation's gradient against central differences of the FP64 GPU
    forward (eps = 1e-6) on 8 coordinates, 8 weights and 8 bandwidths (or
    scal).  Each difference is g . (u+ - u-) from a deterministic forward."""
    torch = torch_cuda
    from ks_amd import _lib
    from ks_amd.autograd import (mapped_gaussian_summation, mapped_summation_forward,
                                 mapped_var_bandwidth_summation)
    rng = np.random.default_rng(41 + var)
    nx, m, n, k = 40000, 32768, 32768, 16
    X = _dev(torch, rng.random((nx, k))).requires_grad_(True)
    w = _dev(torch, rng.standard_normal(nx)).requires_grad_(True)
    h = _dev(torch, (0.5 + rng.random(nx)) * 1.2).requires_grad_(True)
    scal = torch.tensor([-1.0 / (2 * 0.4 ** 2)], dtype=torch.float64,
                        device="cuda").requires_grad_(True)
    am = _dev(torch, rng.integers(0, nx, m).astype(np.int32))
    bm = _dev(torch, rng.integers(0, nx, n).astype(np.int32))
    wm = _dev(torch, rng.integers(0, nx, n).astype(np.int32))
    g = _dev(torch, rng.standard_normal(m))
    if var:
        u = mapped_var_bandwidth_summation(X, X, w, h, h, am, bm, wm)
    else:
        u = mapped_gaussian_summation(X, X, w, scal, am, bm, wm)
    (u * g).sum().backward()
    T = _lib.KS_GAUSSIAN_VAR_BANDWIDTH if var else _lib.KS_GAUSSIAN
    Xd, wd, hd = X.detach().clone(), w.detach().clone(), h.detach().clone()
    s0 = float(scal.detach()[0])

    def fwd(sv=s0):
        with torch.no_grad():
            return mapped_summation_forward(T, Xd, am, Xd, bm, wd, wm, scal=sv, hi=hd, hj=hd)

    eps = 1e-6
    worst = 0.0

    def check(tab, idx, grad):
        nonlocal worst
        x = tab[idx].item()
        tab[idx] = x + eps
        up = fwd()
        tab[idx] = x - eps
        um = fwd()
        tab[idx] = x
        fd = float((g * (up - um)).sum()) / (2 * eps)
        err = abs(grad - fd) / max(abs(fd), 1e-3)
        worst = max(worst, err)
        assert grad == pytest.approx(fd, rel=1e-6, abs=1e-6), (idx, grad, fd)

    used = np.union1d(am.cpu().numpy(), bm.cpu().numpy())
    for x in rng.choice(used, 8, replace=False):
        p = int(rng.integers(0, k))
        check(Xd, (int(x), p), X.grad[int(x), p].item())
    for x in rng.choice(np.unique(wm.cpu().numpy()), 8, replace=False):
        check(wd, int(x), w.grad[int(x)].item())
    if var:
        for x in rng.choice(used, 8, replace=False):
            check(hd, int(x), h.grad[int(x)].item())
    else:
        fd = float((g * (fwd(s0 + eps) - fwd(s0 - eps))).sum()) / (2 * eps)
        assert scal.grad[0].item() == pytest.approx(fd, rel=1e-6)
    print(f"mapped autograd {'var' if var else 'gauss'} 32768^2 k=16: worst rel {worst:.2e}")
