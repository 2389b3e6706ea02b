"""The batched summation at its full workload: 4096 tasks over a pool of
1,048,576 points, k in {3, 8, 32}, m and n log-uniform in [64, 8192], target
and source lists random subsets of the pool, each task's kernel drawn from
{Gaussian h = 0.2, Laplace, polynomial of degree 2} (ks_amd.workloads.list_pool;
the reference's batched path is frame/omp_dgsks_list.cpp:297-420, its driver
test/test_dgsks_list.cpp:193-247).  About 1.2e10 kernel evaluations per k.

The oracle runs over every task on the host's cores (module fixture, once per
k) from a nonzero u, so accumulation is checked too.  Paths: a batch plan run
(ks_list_plan_execute), the same plan captured into a HIP graph and replayed,
the one-shot dgsks_list on host arrays, and the mixed-precision batch.

Bars:
* FP64: relative L2 <= 1e-12 over u (test/test_dgsks.c:43-79), and per entry
  the reference list driver's 1e-10 relative (test/test_dgsks_list.cpp:241-247),
  measured on the batch's contribution c = u - u0 with an absolute floor of
  1e-14 max|c| for entries whose terms cancel (weights are N(0,1) here, where
  the reference driver's are 1).
* The graph replays and the one-shot call lay out the same tiles as the eager
  plan run and must equal it bit for bit.
* Mixed precision: relative L2 <= 1e-6 against the oracle."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12
POOL_TIMEOUT = 900  # the oracle's ~1.2e10 evaluations, on the host's cores


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


@pytest.fixture(scope="module")
def L():
    from ks_amd import _lib
    return _lib.lib()


_CACHE = {}


def _pool(k, oracle):
    """The workload and the oracle over all its tasks (cached per k)."""
    if k not in _CACHE:
        import time
        from ks_amd import workloads as W
        d = W.list_pool(k)
        u0 = np.random.default_rng(40 + k).standard_normal(d["X"].shape[0])
        ref = u0.copy()
        t0 = time.perf_counter()
        oracle.dgsks_list(d["kernels"], k, ref, d["a_off"], d["amap"], d["amap"], d["b_off"],
                          d["bmap"], d["bmap"], d["X"], d["X2"], d["X"], d["X2"], d["w"])
        evals = float(np.sum(np.diff(d["a_off"]) * np.diff(d["b_off"])))
        print(f"list_pool k={k}: oracle over {len(d['kernels'])} tasks, {evals:.3e} evaluations, "
              f"{time.perf_counter() - t0:.1f} s", flush=True)
        _CACHE[k] = (d, u0, ref)
    return _CACHE[k]


def _kernels(d):
    from ks_amd import _lib
    nt = len(d["kernels"])
    arr = (_lib.KsT * nt)()
    for t, kd in enumerate(d["kernels"]):
        arr[t].type = kd["type"]
        arr[t].scal = kd.get("scal") or 0.0
        arr[t].cons = kd.get("cons", 0.0)
        arr[t].powe = kd.get("powe", 0.0)
    return arr


def _check_fp64(got, ref, u0, what):
    c = ref - u0
    err = float(np.sqrt(np.sum((got - ref) ** 2) / np.sum(ref * ref)))
    dev = np.abs((got - u0) - c)
    floor = 1e-14 * float(np.abs(c).max())
    bad = int(np.sum(dev > 1e-10 * np.abs(c) + floor))
    floored = int(np.sum((dev > 1e-10 * np.abs(c)) & (dev <= 1e-10 * np.abs(c) + floor)))
    print(f"{what}: rel L2 = {err:.3e}; entries beyond 1e-10 relative: {bad} "
          f"(+{floored} within the cancellation floor)", flush=True)
    assert err <= TOL
    assert bad == 0


class _DevPlan:
    def __init__(self, L, torch, d, precision):
        self.L = L
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
        self.amap, self.bmap = dev(d["amap"]), dev(d["bmap"])
        self.X, self.X2, self.w = dev(d["X"]), dev(d["X2"]), dev(d["w"])
        self.kern = _kernels(d)
        self.h = ctypes.c_void_p()
        P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        H = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
        rc = L.ks_list_plan_create(ctypes.byref(self.h), len(d["kernels"]), self.kern, d["k"],
                                   H(d["a_off"]), P(self.amap), P(self.amap), H(d["b_off"]),
                                   P(self.bmap), P(self.bmap), precision)
        assert rc == 0, rc

    def run(self, u, stream):
        P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        return self.L.ks_list_plan_execute(self.h, P(u), P(self.X), P(self.X2), P(self.X),
                                           P(self.X2), P(self.w), ctypes.c_void_p(stream))

    def close(self):
        self.L.ks_list_plan_destroy(self.h)


@pytest.mark.timeout(POOL_TIMEOUT)
@pytest.mark.parametrize("k", [3, 8, 32])
def test_list_pool_fp64_plan_graph_and_oneshot(L, torch_cuda, oracle, k):
    torch = torch_cuda
    d, u0, ref = _pool(k, oracle)
    plan = _DevPlan(L, torch, d, 0)
    try:
        s = torch.cuda.Stream()
        u = torch.from_numpy(u0.copy()).cuda()
        with torch.cuda.stream(s):
            assert plan.run(u, s.cuda_stream) == 0
        torch.cuda.synchronize()
        eager = u.cpu().numpy()
        _check_fp64(eager, ref, u0, f"list_pool k={k} FP64 plan")
        # captured once, replayed twice with eager work between
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            assert plan.run(u, s.cuda_stream) == 0
        torch.cuda.synchronize()
        junk = torch.ones(1 << 20, dtype=torch.float64, device="cuda")
        u0_dev = torch.from_numpy(u0).cuda()
        for r in range(2):
            u.copy_(u0_dev)
            junk.mul_(1.0001)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert np.array_equal(u.cpu().numpy(), eager), f"replay {r} differs"
        del g
    finally:
        plan.close()
    # the one-shot call on host arrays (staged through HBM) lays out the same tiles
    u1 = u0.copy()
    kern = _kernels(d)
    H = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    rc = L.dgsks_list(len(d["kernels"]), kern, k, H(u1), u1.size, H(d["a_off"]), H(d["amap"]),
                      H(d["amap"]), H(d["b_off"]), H(d["bmap"]), H(d["bmap"]), H(d["X"]),
                      H(d["X2"]), H(d["X"]), H(d["X2"]), H(d["w"]), 0)
    assert rc == 0
    assert np.array_equal(u1, eager), "one-shot dgsks_list differs from the plan run"


@pytest.mark.timeout(POOL_TIMEOUT)
@pytest.mark.parametrize("k", [3, 8, 32])
def test_list_pool_mixed_precision(L, torch_cuda, oracle, k):
    torch = torch_cuda
    d, u0, ref = _pool(k, oracle)
    plan = _DevPlan(L, torch, d, 2)
    try:
        u = torch.from_numpy(u0.copy()).cuda()
        s = torch.cuda.current_stream()
        assert plan.run(u, s.cuda_stream) == 0
        torch.cuda.synchronize()
        got = u.cpu().numpy()
    finally:
        plan.close()
    err = float(np.sqrt(np.sum((got - ref) ** 2) / np.sum(ref * ref)))
    c = ref - u0
    errc = float(np.sqrt(np.sum(((got - u0) - c) ** 2) / np.sum(c * c)))
    print(f"list_pool k={k} mixed BF16x3 batch: rel L2 = {err:.3e} "
          f"(of the batch's contribution: {errc:.3e})", flush=True)
    assert err <= 1e-6 and errc <= 1e-6
