"""Benchmark of the fused kernel summation on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-suite]

A step is one summation u += K(XA, XB) w over the whole workload, enqueued
through the C ABI (dgsks_device) on torch's current stream with every input
already resident in HBM.  The headline workload is the k = 64 Gaussian mixture
(m = n = 131072, h = 0.3) on which the FP64-vs-mixed-precision metric is
quoted: the mixed BF16x3 form is timed beside the FP64 drop-in, its relative
L2 error against the drop-in's result is measured in the same run, and `value`
is the mixed form's evals/s when that error is <= 1e-6 (the FP64 drop-in's
otherwise).  The line carries the roofline position of the dominant kernel
(HIP events around each launch on the launch stream) and a CPU baseline: the
reference's own dgsks (compiled from its sources into oracle/_ref) timed on
this host's cores over a bounded sample of target rows of the same workload.

Unless --no-suite (or an explicit --workload) is given, the line also carries
`subrecords`, one per other benchmark shape, each with its own roofline and
CPU baseline:
  uniform_gauss16        Gaussian h = 0.5, m = n = 8192, k = 16: per call and
                         replayed from a HIP graph of 32 calls
  nbody_k3:<type>        m = n = 262144, k = 3, permutation maps, four kernels
  mixture_k64_varbw      variable-bandwidth Gaussian, mixed beside FP64
  list_pool_k{3,8,32}    4096 tasks over a 1,048,576-point pool, FP64 and mixed
  list_sym_k3            the same with every task's sources its targets: symmetric
                         layout beside the full one
  grad_gauss16           forward + gradient step, m = n = 32768, k = 16
  grad_varbw16_maps      variable-bandwidth forward + gradient over index maps, 32768^2, k = 16
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "kernel evals/s (m·n/s), FP64 drop-in vs mixed precision at rel err ≤1e-6"
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector = FP64 matrix peak
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 matrix peak
CPU_THREADS_CAP = 16  # one GPU's share of the host's CPUs
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE + WRITE_SIZE, each
# scaled by what the counter reads for a 1 GiB device copy (FETCH_SIZE x 2,
# WRITE_SIZE x 1), keyed by (workload, precision), with the profile summary
# that holds the counts and the commands that produced them.
TRAFFIC = {
    ("mixture_k64", 0): (2.344e9, "profiles/mixture_k64_fp64_and_mixed.md: "
                                  "fused_f64_mfma_kernel FETCH_SIZE + WRITE_SIZE"),
    ("mixture_k64", 2): (3.54e9, "profiles/mixture_k64_fp64_and_mixed.md: mixed_kernel + "
                                 "pre-pass kernels FETCH_SIZE + WRITE_SIZE"),
}
# flops per kernel evaluation in the reference's model, test/test_dgsks.c:254-283
FLOP_C = {0: 37, 1: 6, 2: 60, 3: 35, 4: 89, 5: 8, 6: 6, 7: 7}
TYPE_NAMES = ["gaussian", "polynomial", "laplace", "var_bandwidth", "tanh", "quartic",
              "multiquadric", "epanechnikov"]
PREC_NAMES = {0: "fp64", 1: "mixed_fp32", 2: "mixed_bf16x3"}


def build_workload(name):
    from ks_amd import workloads as W
    if name == "mixture_k64":
        d = W.mixture_k64()
        d["kernel"] = dict(d["kernel"])
        d["label"] = "mixture_k64 gaussian h=0.3 m=n=131072 k=64"
    elif name == "mixture_k64_sym":
        # targets are the sources: one map for both, so the FP64 drop-in takes
        # its symmetric path (each pair of different row blocks once)
        d = W.mixture_k64()
        d["kernel"] = dict(d["kernel"])
        d["bmap"] = d["amap"]
        d["label"] = "mixture_k64 gaussian h=0.3 m=n=131072 k=64, one map (symmetric)"
    elif name == "mixture_k64_varbw":
        d = W.mixture_k64()
        d["kernel"] = dict(type=W.VAR_BANDWIDTH)
        d["label"] = "mixture_k64 variable-bandwidth gaussian h~U[0.2,0.4) m=n=131072 k=64"
    elif name == "uniform_gauss16":
        d = W.uniform_gauss16()
        d["label"] = "uniform_gauss16 gaussian h=0.5 m=n=8192 k=16"
    elif name.startswith("list_task_gauss_k"):
        # the largest Gaussian task of the batched workload alone: m = n = 8192
        # points uniform in [0,1)^k, h = 0.2, the mixed form beside FP64
        k = int(name.split("_k")[1])
        d = W.uniform_gauss16(k=k, seed=4)
        d["kernel"] = dict(type=W.GAUSSIAN, scal=W.gaussian_scal(0.2))
        d["label"] = f"list task gaussian h=0.2 m=n=8192 k={k} uniform [0,1)^k"
    elif name.startswith("nbody_k3"):
        kind = name.split(":")[1] if ":" in name else "gaussian"
        d = W.nbody_k3()
        d["kernel"] = dict(W.NBODY_KERNELS[kind])
        d["label"] = f"nbody_k3 {kind} m=n=262144 k=3 permutation maps"
    elif name == "grad_gauss16":
        d = W.grad_gauss16()
        d["label"] = ("grad_gauss16 gaussian h=0.4 m=n=32768 k=16: forward u = K w and the "
                      "gradient of g.u in dXA, dXB, dw, dh")
    elif name == "grad_varbw16_maps":
        d = W.grad_varbw16_maps()
        d["label"] = ("grad_varbw16_maps variable bandwidth m=n=32768 k=16 over a 40000-point "
                      "table, maps with repeats: forward u and the gradient of g.u in the "
                      "tables dX, dw, dh")
    elif name.startswith("list_pool") or name.startswith("list_sym"):
        k = int(name.split("_k")[1]) if "_k" in name else 3
        sym = name.startswith("list_sym")
        d = W.list_pool(k=k, symmetric=sym)
        d["symmetric"] = sym
        d["label"] = (f"list_sym 4096 tasks, sources = targets, m = n log-U[64,8192] k={k} "
                      "pool=1048576 kernels {gaussian h=0.2, laplace, polynomial deg 2}" if sym else
                      f"list_pool 4096 tasks m,n log-U[64,8192] k={k} pool=1048576 "
                      "kernels {gaussian h=0.2, laplace, polynomial deg 2}")
    else:
        raise SystemExit(f"unknown workload {name}")
    return d


# ----------------------------------------------------------------- CPU side

def host_info():
    """nproc, the CPUs this process may run on, the threads the baseline uses
    and the CPU model."""
    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = nproc
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": nproc, "cpus_available": avail,
            "threads": max(1, min(avail, CPU_THREADS_CAP)),
            "threads_cap": CPU_THREADS_CAP, "cap_reason": "one GPU's share of the host CPUs",
            "cpu_model": model}


def _parallel(jobs, threads):
    """Run the callables `jobs` on `threads` Python threads (the reference's
    C calls release the GIL); returns the wall time in seconds."""
    lock = threading.Lock()
    it = iter(jobs)

    def worker():
        while True:
            with lock:
                job = next(it, None)
            if job is None:
                return
            job()

    ths = [threading.Thread(target=worker) for _ in range(threads)]
    t0 = time.perf_counter()
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    return time.perf_counter() - t0


def _ref_kernel(O, kd, hv=None):
    kt = O.RefKsT(int(kd["type"]), float(kd.get("powe") or 0.0), float(kd.get("scal") or 0.0),
                  float(kd.get("cons") or 0.0), None, None, None)
    if hv is not None:
        kt.hi = kt.hj = hv.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return kt


def cpu_baseline(d, budget_s=8.0):
    """The reference dgsks (oracle/_ref) on this host over a sample of target
    rows: first on one core (calibration and the one-core rate), then on every
    thread of the host's share, rows split over threads.  Each thread calls
    the reference's sequential dgsks on its own slice of rows (the reference's
    unit of parallel work, one dgsks per task, frame/omp_dgsks_list.cpp:366-397)."""
    from oracle import oracle as O
    R = O.ref_lib()
    if R is None:
        return None
    info = host_info()
    threads = info["threads"]
    X, X2, w = d["X"], d["X2"], d["w"]
    amap, bmap = d["amap"], d["bmap"]
    kd = d["kernel"]
    n, k = bmap.size, d["k"]
    hv = d.get("hinv")
    out = np.zeros(X.shape[0])

    def run(rows_lo, rows_hi):
        am = np.ascontiguousarray(amap[rows_lo:rows_hi])
        kt = _ref_kernel(O, kd, hv)
        R.dgsks(ctypes.byref(kt), am.size, n, k, O._d(out), O._i(am), O._d(X), O._d(X2),
                O._i(am), O._d(X), O._d(X2), O._i(bmap), O._d(w), O._i(bmap))

    # a warm-up call sizes the one-core run, which lasts about a quarter of the budget
    rows0 = max(1, min(amap.size, int(2e6 // max(n, 1)) + 1))
    t0 = time.perf_counter()
    run(0, rows0)
    rate0 = rows0 * n / max(time.perf_counter() - t0, 1e-9)
    rows1 = int(min(amap.size, max(1, budget_s / 4 * rate0 / n)))
    t0 = time.perf_counter()
    run(0, rows1)
    dt1 = max(time.perf_counter() - t0, 1e-9)
    rate1 = rows1 * n / dt1
    rows = int(min(amap.size, max(threads, budget_s * rate1 * threads / n)))
    per = max(1, rows // threads)
    rows = per * threads
    dt = _parallel([lambda i=i: run(i * per, (i + 1) * per) for i in range(threads)], threads)
    res = {"value": rows * n / dt, "unit": "evals/s", "cores": threads, "kind": "reference",
           "sample": f"reference dgsks (oracle/_ref, sandybridge AVX build) on {rows} of "
                     f"{amap.size} target rows x all {n} sources, {threads} threads x {per} "
                     f"rows, {dt:.2f} s; one core: {rows1} rows, {dt1:.2f} s",
           "one_core_value": rate1}
    res.update(info)
    return res


def cpu_baseline_list(d, budget_s=4.0):
    """The reference dgsks (oracle/_ref) over a seeded sample of the batch's
    tasks, one sequential dgsks per task with the task's own kernel (the
    reference's batched path runs one dgsks per task on each worker,
    frame/omp_dgsks_list.cpp:366-397): one core first, then every thread of the
    host's share with tasks handed out dynamically."""
    from oracle import oracle as O
    R = O.ref_lib()
    if R is None:
        return None
    info = host_info()
    threads = info["threads"]
    X, X2, w, k = d["X"], d["X2"], d["w"], d["k"]
    a_off, b_off, amap, bmap = d["a_off"], d["b_off"], d["amap"], d["bmap"]
    ntask = len(d["kernels"])
    order = np.random.default_rng(7).permutation(ntask)
    out = np.zeros(X.shape[0])

    def run(t):
        am = np.ascontiguousarray(amap[a_off[t]:a_off[t + 1]])
        bm = np.ascontiguousarray(bmap[b_off[t]:b_off[t + 1]])
        kt = _ref_kernel(O, d["kernels"][t])
        R.dgsks(ctypes.byref(kt), am.size, bm.size, k, O._d(out), O._i(am), O._d(X), O._d(X2),
                O._i(am), O._d(X), O._d(X2), O._i(bm), O._d(w), O._i(bm))

    def evals(ts):
        return float(sum(float(a_off[t + 1] - a_off[t]) * float(b_off[t + 1] - b_off[t])
                         for t in ts))

    # one core: tasks in the sampled order until about a quarter of the budget
    t0 = time.perf_counter()
    i = 0
    while i < ntask and (i < 4 or time.perf_counter() - t0 < budget_s / 4):
        run(int(order[i]))
        i += 1
    one = order[:i]
    rate1 = evals(one) / (time.perf_counter() - t0)
    # every thread: the next tasks of the order, sized for the budget
    want = budget_s * rate1 * threads
    j, acc = i, 0.0
    while j < ntask and acc < want:
        t = int(order[j])
        acc += float(a_off[t + 1] - a_off[t]) * float(b_off[t + 1] - b_off[t])
        j += 1
    many = [int(t) for t in order[i:j]] or [int(t) for t in one]
    dt = _parallel([lambda t=t: run(t) for t in many], threads)
    res = {"value": evals(many) / dt, "unit": "evals/s", "cores": threads, "kind": "reference",
           "sample": f"reference dgsks (oracle/_ref) on {len(many)} of {ntask} seeded tasks "
                     f"({evals(many):.3e} evaluations), {threads} threads, {dt:.2f} s; one "
                     f"core: {len(one)} tasks",
           "one_core_value": rate1}
    res.update(info)
    return res


def cpu_baseline_grad(d, budget_s=4.0):
    """The reference has no backward: the CPU figure is the oracle's FP64
    gradient (a port, oracle/gsks_oracle.c orc_gauss_grad) over seeded rows,
    dXA for R target rows and dXB, dw for R source rows, timed on one core and
    on every thread of the host's share.  A full step's rate is extrapolated
    from rows (cost linear in rows): pairs of the step / (time x m / R)."""
    from oracle import oracle as O
    info = host_info()
    threads = info["threads"]
    XA, XB, w, g, h = d["XA"], d["XB"], d["w"], d["g"], d["h"]
    m, n = XA.shape[0], XB.shape[0]
    rng = np.random.default_rng(11)

    def timed(rows, nth):
        ra = np.sort(rng.choice(m, rows, replace=False)).astype(np.int32)
        rb = np.sort(rng.choice(n, rows, replace=False)).astype(np.int32)
        t0 = time.perf_counter()
        O.gauss_grad(XA, XB, w, g, h, rows_a=ra, rows_b=rb, want_dh=False, nthreads=nth)
        return time.perf_counter() - t0

    t0 = timed(4, 1)  # warm-up, sizes the one-core run (about a quarter of the budget)
    r1 = int(min(m, max(4, budget_s / 4 * 4 / t0)))
    t1 = timed(r1, 1)
    rate1 = r1 * n / t1  # pairs of a full step per second
    rows = int(min(m, max(threads, budget_s * rate1 * threads / n)))
    dt = timed(rows, threads)
    res = {"value": rows * n / dt, "unit": "evals/s", "cores": threads, "kind": "port",
           "sample": f"the reference has no backward; oracle FP64 gradient (dXA over {rows} "
                     f"target rows, dXB and dw over {rows} source rows, dh left out), {threads} "
                     f"threads, {dt:.2f} s, rate = rows x n / time; one core: {r1} rows",
           "one_core_value": rate1}
    res.update(info)
    return res


# ----------------------------------------------------------------- GPU side

def _events_timed(stream, step, steps, warmup):
    """Run `warmup` then `steps` calls of step(); returns (wall ms per step,
    mean HIP-event ms per step on `stream`)."""
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return wall, float(np.mean([a.elapsed_time(b) for a, b in ev]))


def run_list(args, L, name, precision, steps, warmup, with_cpu):
    """Batched workload: a batch plan built once (host tile layout, device
    descriptors), then one plan run per step with the pool resident in HBM,
    timed with HIP events on the run's stream.  The one-shot dgsks_list wall
    time (layout + upload + run + synchronisation) is reported beside it, and
    for the mixed form its relative L2 error against the FP64 plan's result.
    For the symmetric batch (every task's sources are its targets) the same plan
    is also timed with a copy of the point table as the sources, which runs the
    full layout: the difference is what evaluating each pair once saves."""
    import torch
    from ks_amd import _lib
    t_start = time.perf_counter()
    d = build_workload(name)
    X = torch.from_numpy(d["X"]).cuda()
    X2 = torch.from_numpy(d["X2"]).cuda()
    w = torch.from_numpy(d["w"]).cuda()
    amap = torch.from_numpy(d["amap"]).cuda()
    bmap = torch.from_numpy(d["bmap"]).cuda()
    u = torch.zeros(X.shape[0], dtype=torch.float64, device="cuda")
    ntask = len(d["kernels"])
    kernels = (_lib.KsT * ntask)()
    for t, kd in enumerate(d["kernels"]):
        kernels[t].type = kd["type"]
        kernels[t].scal = kd.get("scal") or 0.0
        kernels[t].cons = kd.get("cons", 0.0)
        kernels[t].powe = kd.get("powe", 0.0)
    a_off, b_off = d["a_off"], d["b_off"]
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    H = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    k = d["k"]
    prec = {"auto": 0, "fp64": 0, "mixed_fp32": 1, "mixed_bf16x3": 2}[precision]
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)

    def make_plan(pr):
        plan = ctypes.c_void_p()
        rc = L.ks_list_plan_create(ctypes.byref(plan), ntask, kernels, k, H(a_off), P(amap),
                                   P(amap), H(b_off), P(bmap), P(bmap), pr)
        if rc != 0:
            raise RuntimeError(f"ks_list_plan_create failed: {_lib.error_string(rc)}")
        return plan

    def run_plan(plan):
        rc = L.ks_list_plan_execute(plan, P(u), P(X), P(X2), P(X), P(X2), P(w), sp)
        if rc != 0:
            raise RuntimeError(f"ks_list_plan_execute failed: {_lib.error_string(rc)}")

    t0 = time.perf_counter()
    plan = make_plan(prec)
    plan_ms = (time.perf_counter() - t0) * 1e3
    ms_per_step, kernel_ms = _events_timed(stream, lambda: run_plan(plan), steps, warmup)
    extra = {}
    sym = bool(d.get("symmetric"))
    if sym:
        nsym = L.ks_list_plan_symmetric_tasks(plan)
        u.zero_()
        run_plan(plan)
        torch.cuda.synchronize()
        u_sym = u.clone()
        Xc, X2c = X.clone(), X2.clone()

        def run_full():
            rc = L.ks_list_plan_execute(plan, P(u), P(X), P(X2), P(Xc), P(X2c), P(w), sp)
            if rc != 0:
                raise RuntimeError(f"ks_list_plan_execute failed: {_lib.error_string(rc)}")
        full_ms, full_kms = _events_timed(stream, run_full, steps, warmup)
        u.zero_()
        run_full()
        torch.cuda.synchronize()
        extra["symmetric"] = {
            "symmetric_tasks": nsym,
            "full_layout_ms_per_step": full_ms, "full_layout_kernel_ms": full_kms,
            "speedup_vs_full_layout": full_ms / ms_per_step,
            "rel_diff_vs_full_layout": float(torch.linalg.norm(u - u_sym) /
                                             torch.linalg.norm(u_sym))}
        del Xc, X2c
    if prec != 0:
        # the mixed batch's result against the FP64 batch's, from u = 0
        p64 = make_plan(0)
        u.zero_()
        run_plan(p64)
        torch.cuda.synchronize()
        u64 = u.clone()
        u.zero_()
        run_plan(plan)
        torch.cuda.synchronize()
        extra["rel_err_vs_fp64"] = float(torch.linalg.norm(u - u64) / torch.linalg.norm(u64))
        f_ms, f_kms = _events_timed(stream, lambda: run_plan(p64), max(3, steps // 2), 1)
        extra["fp64_kernel_ms"] = f_kms
        L.ks_list_plan_destroy(p64)
    L.ks_list_plan_destroy(plan)

    # the one-shot call, for comparison (includes the host layout each time)
    def oneshot():
        rc = L.dgsks_list(ntask, kernels, k, P(u), u.numel(), H(a_off), P(amap), P(amap),
                          H(b_off), P(bmap), P(bmap), P(X), P(X2), P(X), P(X2), P(w), prec)
        if rc != 0:
            raise RuntimeError(f"dgsks_list failed: {_lib.error_string(rc)}")
    oneshot()
    t0 = time.perf_counter()
    for _ in range(2):
        oneshot()
    oneshot_ms = (time.perf_counter() - t0) * 1e3 / 2

    evals = float(np.sum(np.diff(a_off) * np.diff(b_off)))
    types = [kd["type"] for kd in d["kernels"]]
    # tasks a mixed-precision batch sends to the mixed kernels: Gaussian,
    # k >= 16 and m, n >= 16384 (the routing rule of the batched launch)
    n_mixed = 0 if prec == 0 else int(np.sum([
        t in (0, 3) and k >= 16 and ma >= 16384 and nb >= 16384
        for t, ma, nb in zip(types, np.diff(a_off), np.diff(b_off))]))
    def evaluated(ma, nb):
        # a symmetric task (m = n >= 128 at k in {3, 8} or 9..64) evaluates row
        # block I (64 rows) against sources 64 I .. n only
        if not (sym and ma == nb and ma >= 128 and (k in (3, 8) or 9 <= k <= 64)):
            return float(ma) * float(nb)
        return float(sum(min(64, ma - r0) * (nb - r0) for r0 in range(0, int(ma), 64)))
    per = float(np.sum([(2 * k + FLOP_C[t]) * evaluated(ma, nb) for t, ma, nb in
                        zip(types, np.diff(a_off), np.diff(b_off))]))
    achieved = per / (kernel_ms / 1e3) / 1e12
    out = {
        "metric": METRIC, "value": evals / (ms_per_step / 1e3), "unit": "evals/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if n_mixed == 0 else "mixed", "data": "synthetic (seeded numpy)",
        "config": {"workload": d["label"], "ntask": ntask, "evals": evals, "k": k,
                   "precision": precision, "tasks_on_mixed_kernels": n_mixed},
        "roofline": {"bound": "mfma" if 9 <= k <= 64 else "valu", "pipe": "fp64",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "traffic_source": None,
                     "flop_model": "2k + c_type per evaluation, test/test_dgsks.c:254-283",
                     "evaluated_pairs": float(np.sum([evaluated(ma, nb) for ma, nb in
                                                      zip(np.diff(a_off), np.diff(b_off))])),
                     "kernel_ms": kernel_ms,
                     "kernel": "gsks::list_kernel (+ list_rowsum_kernel)"},
        "plan_create_ms": plan_ms,
        "oneshot_dgsks_list_ms": oneshot_ms,
    }
    out.update(extra)
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline_list(d, budget_s=args.sub_cpu_seconds)
    out["bench_s"] = time.perf_counter() - t_start
    return out


def run_grad(args, name, precision, steps, warmup, with_cpu):
    """One step = the forward u = K w and the whole gradient of L = g.u (dXA,
    dXB, dw, dh): two multi-right-hand-side summations that recompute K, the
    first of which (k + 1 right-hand sides [w b, w]) also yields u, as the
    autograd forward runs it when XA or h needs a gradient.
    Precision auto / mixed_fp32: the mixed form (FP32 matrix cores on centred
    points where the block bound allows, FP64 elsewhere) beside the FP64 form,
    with the largest relative L2 error of u, dXA, dXB, dw and dh against the
    FP64 step; fp64: the FP64 form alone.  Flops per pair by the reference's
    model for the forward (2k + 37) plus 4k + 4 for the backward; HIP events on
    the stream time each step."""
    import torch
    from ks_amd import autograd as AG
    t_start = time.perf_counter()
    d = build_workload(name)
    XA = torch.from_numpy(d["XA"]).cuda()
    XB = torch.from_numpy(d["XB"]).cuda()
    w = torch.from_numpy(d["w"]).cuda()
    g = torch.from_numpy(d["g"]).cuda()
    h, k = d["h"], d["k"]
    m, n = XA.shape[0], XB.shape[0]
    stream = torch.cuda.current_stream()
    want = "mixed_fp32" if precision in ("auto", "mixed_bf16x3") else precision

    def step(prec):
        u, state = AG.gaussian_summation_forward_state(XA, XB, w, h, prec)
        return (u,) + AG.gaussian_summation_backward(XA, XB, w, g, h, state, prec)

    pairs = float(m) * float(n)
    fpe = (2 * k + FLOP_C[0]) + (4 * k + 4)
    extra = {}
    if want != "fp64":
        f_ms, f_kms = _events_timed(stream, lambda: step("fp64"), max(3, steps // 2), 1)
        ref = step("fp64")
        got = step(want)
        errs = [float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) for a, b in zip(got, ref)]
        extra = {"fp64_value": pairs / (f_ms / 1e3), "fp64_ms_per_step": f_ms,
                 "fp64_kernel_ms": f_kms, "rel_err_vs_fp64": max(errs),
                 "rel_err_vs_fp64_each": dict(zip(["u", "dXA", "dXB", "dw", "dh"], errs))}
    wall, kms = _events_timed(stream, lambda: step(want), steps, warmup)
    value = pairs / (wall / 1e3)
    if want != "fp64":
        extra["speedup_vs_fp64"] = value / extra["fp64_value"]
        if extra["rel_err_vs_fp64"] > 1e-6:
            extra["note"] = "mixed form missed the 1e-6 bar; value reports the FP64 form"
            value, wall = extra["fp64_value"], extra["fp64_ms_per_step"]
            kms = f_kms
            want = "fp64"
    achieved = pairs * fpe / (kms / 1e3) / 1e12
    mixed = want != "fp64"
    out = {"metric": METRIC, "value": value, "unit": "evals/s", "n_gpus": 1,
           "steps": steps, "warmup": warmup, "ms_per_step": wall,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64" if mixed else "f64",
           "data": "synthetic (seeded numpy, regenerated each run)",
           "config": {"workload": d["label"], "precision": want, "m": m, "n": n, "k": k,
                      "kernel": "gaussian"},
           "roofline": {"bound": "mfma",
                        "pipe": ("fp32 matrix (FP64 norms, exponents and sums)" if mixed
                                 else "fp64 (matrix and vector peaks are equal)"),
                        "achieved": achieved,
                        "peak": FP32_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS,
                        "unit": "TFLOP/s",
                        "frac": achieved / (FP32_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS),
                        "traffic": None, "traffic_source": None, "flops_per_eval": fpe,
                        "flop_model": "2k + 37 forward (test/test_dgsks.c:254-283) + 4k + 4 "
                                      "backward", "kernel_ms": kms,
                        "kernel": ("gsks::multirhs_mixed_kernel (2 passes of k + 1 "
                                   "right-hand sides)" if mixed
                                   else "gsks::multirhs_mfma_kernel (2 passes of k + 1 "
                                        "right-hand sides)")}}
    out.update(extra)
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline_grad(d, budget_s=args.sub_cpu_seconds)
    out["bench_s"] = time.perf_counter() - t_start
    return out


def cpu_baseline_grad_maps(d, budget_s=4.0):
    """The oracle's mapped FP64 gradient (a port, oracle/gsks_oracle.c
    orc_ks_grad) on the first R target rows against all n sources, on one core
    and on every thread of the host's share; rate = R x n pairs / time."""
    from oracle import oracle as O
    info = host_info()
    threads = info["threads"]
    n = d["bmap"].size

    def timed(rows, nth):
        t0 = time.perf_counter()
        O.ks_grad(O.VAR_BANDWIDTH, d["g"], np.arange(rows, dtype=np.int32), d["X"],
                  d["amap"][:rows], d["X"], d["bmap"], d["w"], d["wmap"], hi=d["h"], hj=d["h"],
                  nthreads=nth)
        return time.perf_counter() - t0

    m = d["amap"].size
    t0 = timed(4, 1)
    r1 = int(min(m, max(4, budget_s / 4 * 4 / t0)))
    t1 = timed(r1, 1)
    rate1 = r1 * n / t1
    rows = int(min(m, max(threads, budget_s * rate1 * threads / n)))
    dt = timed(rows, threads)
    res = {"value": rows * n / dt, "unit": "evals/s", "cores": threads, "kind": "port",
           "sample": f"the reference has no backward; oracle FP64 mapped gradient over the first "
                     f"{rows} target rows and all {n} sources, {threads} threads, {dt:.2f} s; "
                     f"one core: {r1} rows",
           "one_core_value": rate1}
    res.update(info)
    return res


def run_grad_maps(args, name, steps, warmup, with_cpu):
    """One step = the mapped variable-bandwidth forward u (dgsks_device) and
    the gradient of L = g.u added into the tables dX, dw, dh (ks_grad_device:
    two multi-right-hand-side passes of k + 2 and k + 3 columns, then the
    ordered scatter into the tables).  FP64; HIP events on the stream.  Flops
    per pair: 2k + 35 for the forward (the reference's model) and 2k + 35 +
    2R for each pass of R right-hand sides."""
    import torch
    from ks_amd import _lib
    from ks_amd.autograd import mapped_summation_forward, mapped_summation_grad
    t_start = time.perf_counter()
    d = build_workload(name)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    X, h, w, g = dev(d["X"]), dev(d["h"]), dev(d["w"]), dev(d["g"])
    am, bm, wm = dev(d["amap"]), dev(d["bmap"]), dev(d["wmap"])
    iu = torch.arange(am.numel(), dtype=torch.int32, device="cuda")
    T = _lib.KS_GAUSSIAN_VAR_BANDWIDTH
    k, m, n = d["k"], am.numel(), bm.numel()
    stream = torch.cuda.current_stream()

    def step():
        u = mapped_summation_forward(T, X, am, X, bm, w, wm, hi=h, hj=h)
        G = mapped_summation_grad(T, g, iu, X, am, X, bm, w, wm, hi=h, hj=h)
        return u, G

    wall, kms = _events_timed(stream, step, steps, warmup)
    pairs = float(m) * float(n)
    fpe = 3 * (2 * k + FLOP_C[3]) + 2 * (k + 2) + 2 * (k + 3)
    achieved = pairs * fpe / (kms / 1e3) / 1e12
    out = {"metric": METRIC, "value": pairs / (wall / 1e3), "unit": "evals/s", "n_gpus": 1,
           "steps": steps, "warmup": warmup, "ms_per_step": wall,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded numpy, regenerated each run)",
           "config": {"workload": d["label"], "precision": "fp64", "m": m, "n": n, "k": k,
                      "kernel": "var_bandwidth"},
           "roofline": {"bound": "mfma", "pipe": "fp64 (matrix and vector peaks are equal)",
                        "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                        "traffic_source": None, "flops_per_eval": fpe,
                        "flop_model": "3 (2k + 35) (forward and two passes, "
                                      "test/test_dgsks.c:254-283) + 2 (k + 2) + 2 (k + 3)",
                        "kernel_ms": kms,
                        "kernel": "fused forward + gsks::multirhs_mfma_kernel (2 passes)"}}
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline_grad_maps(d, budget_s=args.sub_cpu_seconds)
    out["bench_s"] = time.perf_counter() - t_start
    return out


def bench_single(L, kt, k, amap, bmap, X, w, stream, steps, u64):
    """The single-precision front end (sgsks_device) on float copies of the same
    inputs, norms left to the library: in the precision sgsks() picks (FP64
    sums) and in the mixed BF16x3 form a caller may ask for (Gaussian kernels).
    Per-call time (HIP events, widening and scatter included) and the relative
    difference from the FP64 result on the double inputs."""
    import torch
    from ks_amd import _lib
    Xf, wf = X.float(), w.float()
    uf = torch.zeros(X.shape[0], dtype=torch.float32, device="cuda")
    m, n = amap.numel(), bmap.numel()
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    out = []
    for prec in ((0, 2) if kt.type in (0, 3) else (0,)):
        argv = (ctypes.byref(kt), m, n, k, P(uf), P(amap), P(Xf), None, P(amap), P(Xf), None,
                P(bmap), P(wf), P(bmap), prec, ctypes.c_void_p(stream.cuda_stream))

        def step():
            rc = L.sgsks_device(*argv)
            if rc != 0:
                raise RuntimeError(f"sgsks_device failed: {_lib.error_string(rc)}")

        uf.zero_()
        step()
        torch.cuda.synchronize()
        rel = float(torch.linalg.norm(uf.double() - u64) / torch.linalg.norm(u64))
        _, ms = _events_timed(stream, step, max(3, steps // 2), 0)
        out.append({"entry": "sgsks_device", "precision": PREC_NAMES[prec],
                    "default_of_sgsks": prec == 0, "ms_per_call": ms,
                    "evals_per_s": float(m) * float(n) / (ms / 1e3),
                    "rel_diff_vs_fp64_on_double_inputs": rel})
    return out


def run_single(args, L, name, precision, steps, warmup, with_cpu, cpu_s, graph=False,
               with_sgsks=False):
    """One summation per step through dgsks_device.  Precision auto: the mixed
    BF16x3 form beside the FP64 drop-in for the Gaussian kernels on the
    clustered k = 64 workloads, the FP64 drop-in otherwise.  graph: also the
    per-call rate of 32 calls captured into one HIP graph and replayed."""
    import torch
    from ks_amd import _lib
    t_start = time.perf_counter()
    d = build_workload(name)
    kd = d["kernel"]
    X = torch.from_numpy(d["X"]).cuda()
    X2 = torch.from_numpy(d["X2"]).cuda()
    w = torch.from_numpy(d["w"]).cuda()
    amap = torch.from_numpy(d["amap"]).cuda()
    # one map for targets and sources (the N-body shape) stays one device array,
    # so the library sees a symmetric sum
    bmap = amap if d["bmap"] is d["amap"] else torch.from_numpy(d["bmap"]).cuda()
    u = torch.zeros(X.shape[0], dtype=torch.float64, device="cuda")
    hdev = torch.from_numpy(d["hinv"]).cuda() if "hinv" in d else None
    m, n, k = amap.numel(), bmap.numel(), d["k"]
    kt = _lib.KsT()
    kt.type = kd["type"]
    kt.scal = kd.get("scal", 0.0)
    kt.cons = kd.get("cons", 0.0)
    kt.powe = kd.get("powe", 0.0)
    if kd["type"] == 3:
        kt.hi = kt.hj = hdev.data_ptr()
    stream = torch.cuda.current_stream()
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731

    def call(precision_code, s):
        return L.dgsks_device(ctypes.byref(kt), m, n, k, P(u), P(amap), P(X), P(X2), P(amap),
                              P(X), P(X2), P(bmap), P(w), P(bmap), precision_code,
                              ctypes.c_void_p(s.cuda_stream))

    def timed(precision_code, steps_, warmup_):
        def step():
            rc = call(precision_code, stream)
            if rc != 0:
                raise RuntimeError(f"dgsks_device failed: {_lib.error_string(rc)}")
        return _events_timed(stream, step, steps_, warmup_)

    def one_result(precision_code):
        u.zero_()
        assert call(precision_code, stream) == 0
        torch.cuda.synchronize()
        return u.clone()

    evals = float(m) * float(n)
    flops_per_eval = 2 * k + FLOP_C[kd["type"]]
    # A symmetric sum (one map and one table for targets and sources) evaluates
    # each pair of different row blocks once for both terms -- row blocks of
    # 1024 on the vector units (k in {2, 3, 4, 8}), of 128 on the matrix cores
    # (9 <= k <= 64, not Laplace): the roofline counts the evaluations the
    # kernel performs.
    one_map = bmap is amap and m == n
    sym_valu = one_map and k in (2, 3, 4, 8) and m >= 2048
    sym_mfma = one_map and 9 <= k <= 64 and kd["type"] != 2 and m >= 256
    symmetric = sym_valu or sym_mfma
    B = 1024 if sym_valu else 128
    evaluated = (sum(min(B, m - i) * (m - i) for i in range(0, m, B)) if symmetric
                 else evals)

    def traffic(precision_code):
        t = TRAFFIC.get((name, precision_code))
        return (t[0], t[1]) if t else (None, None)

    def fp64_roofline(kernel_ms):
        achieved = evaluated * flops_per_eval / (kernel_ms / 1e3) / 1e12
        tr, src = traffic(0)
        return {"bound": "mfma", "pipe": "fp64 (matrix and vector peaks are equal)",
                "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_PEAK_TFLOPS, "traffic": tr, "traffic_source": src,
                "flops_per_eval": flops_per_eval, "evaluated_pairs": evaluated,
                "flop_model": "2k + c_type, test/test_dgsks.c:254-283", "kernel_ms": kernel_ms,
                "kernel": ("gsks::fused_f64_sym_kernel (+ sym_colsum_kernel)" if sym_valu
                           else "gsks::fused_f64_sym_mfma_kernel (+ sym_colsum_kernel)" if sym_mfma
                           else "gsks::fused_f64_mfma_kernel" if 9 <= k <= 64 and kd["type"] != 2
                           else "gsks::fused_f64_kernel")}

    def mixed_roofline(kernel_ms, precision_code):
        # The reference's flop model (2k + c per evaluation), as for the FP64
        # form, against the peak of the pipe the cross term runs on.  The
        # split-BF16 form nominally issues 3 BF16 products of 2k flops per pair;
        # its bounds skip most 16 x 16 blocks before or after the first
        # product, so the kernel executes fewer matrix flops than that count.
        peak = BF16_PEAK_TFLOPS if precision_code == 2 else FP32_PEAK_TFLOPS
        achieved = evals * flops_per_eval / (kernel_ms / 1e3) / 1e12
        tr, src = traffic(precision_code)
        return {"bound": "mfma",
                "pipe": "bf16 dense matrix" if precision_code == 2 else "fp32 matrix",
                "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": tr, "traffic_source": src, "flops_per_eval": flops_per_eval,
                "flop_model": "2k + c_type, test/test_dgsks.c:254-283",
                "nominal_matrix_flops_per_eval": (3 if precision_code == 2 else 1) * 2 * k,
                "kernel_ms": kernel_ms,
                "kernel": "gsks::mixed_kernel (+ pre-pass: bucketing, radix sorts, bounds, "
                          "need masks, packs)"}

    mixed_ok = kd["type"] in (0, 3)
    want = precision
    if want == "auto":
        # the metric compares the two forms on the clustered k = 64 workloads
        # (and on the batched workload's Gaussian task shape)
        want = ("mixed_bf16x3" if (mixed_ok and (name.startswith("mixture") or
                                                 name.startswith("list_task")))
                else "fp64")
    prec = {"fp64": 0, "mixed_fp32": 1, "mixed_bf16x3": 2}[want]
    extra = {}
    if prec != 0:
        # the FP64 drop-in beside it, and the mixed form's error against it
        f_ms, f_kms = timed(0, max(3, steps // 2), 1)
        u64 = one_result(0)
        umx = one_result(prec)
        rel = float(torch.linalg.norm(umx - u64) / torch.linalg.norm(u64))
        extra = {"fp64_value": evals / (f_ms / 1e3), "fp64_kernel_ms": f_kms,
                 "rel_err_vs_fp64": rel, "roofline_fp64": fp64_roofline(f_kms)}
    ms_per_step, kernel_ms = timed(prec, steps, warmup)
    value = evals / (ms_per_step / 1e3)
    if prec != 0:
        extra["speedup_vs_fp64"] = value / extra["fp64_value"]
        if extra["rel_err_vs_fp64"] > 1e-6:
            extra["note"] = "mixed form missed the 1e-6 bar; value reports the FP64 drop-in"
            value, ms_per_step, kernel_ms = extra["fp64_value"], evals / extra["fp64_value"] * 1e3, \
                extra["fp64_kernel_ms"]
            prec = 0
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == 0 else ("f32+f64" if prec == 1 else "bf16x3+f32+f64"),
        "data": "synthetic (seeded numpy, regenerated each run)",
        "config": {"workload": d["label"], "precision": PREC_NAMES[prec],
                   "m": m, "n": n, "k": k, "kernel": TYPE_NAMES[kd["type"]]},
        "roofline": fp64_roofline(kernel_ms) if prec == 0 else mixed_roofline(kernel_ms, prec),
        "kernel_evals_per_s": evals / (kernel_ms / 1e3),
    }
    out.update(extra)
    if graph:
        out["graph_replay"] = graph_replay(L, torch, call, u, prec, evals, evaluated,
                                           flops_per_eval)
    if with_sgsks:
        out["single"] = bench_single(L, kt, k, amap, bmap, X, w, stream, steps, one_result(0))
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline(d, budget_s=cpu_s)
    out["bench_s"] = time.perf_counter() - t_start
    return out


def graph_replay(L, torch, call, u, prec, evals, evaluated, flops_per_eval, calls=32, reps=10):
    """`calls` summations captured into one HIP graph on a side stream (after
    an eager run there), replayed `reps` times: per-call time from HIP events
    around the replays, and whether one replay equals the same calls made
    eagerly, bit for bit."""
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        u.zero_()
        for _ in range(calls):
            assert call(prec, s) == 0
    torch.cuda.synchronize()
    eager = u.clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        for _ in range(calls):
            assert call(prec, s) == 0
    torch.cuda.synchronize()
    u.zero_()
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    same = bool(torch.equal(u, eager))
    cur = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(cur)
    for _ in range(reps):
        g.replay()
    b.record(cur)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / (reps * calls)
    achieved = evaluated * flops_per_eval / (ms / 1e3) / 1e12
    peak = FP64_PEAK_TFLOPS if prec == 0 else (BF16_PEAK_TFLOPS if prec == 2
                                                else FP32_PEAK_TFLOPS)
    return {"calls_per_graph": calls, "replays": reps, "ms_per_call": ms,
            "evals_per_s": evals / (ms / 1e3), "achieved_tflops": achieved,
            "frac_of_peak": achieved / peak, "replay_equals_eager": same}


def suite(args, L):
    """Sub-records of every other benchmark shape (each with its own roofline
    and CPU baseline), in a bounded number of steps each."""
    steps = max(3, min(args.steps, 10))
    warm = max(1, min(args.warmup, 2))
    cpu = not args.no_cpu_baseline
    cs = args.sub_cpu_seconds
    subs = {}

    def guard(key, fn):
        try:
            subs[key] = fn()
        except Exception as e:  # one shape failing must not hide the others
            subs[key] = {"error": f"{type(e).__name__}: {e}"}
        print(f"[bench] {key}: done", file=sys.stderr, flush=True)

    guard("uniform_gauss16", lambda: run_single(args, L, "uniform_gauss16", "fp64", 20, 3, cpu,
                                                cs, graph=True))
    for kind in ("gaussian", "laplace", "polynomial", "tanh"):
        guard(f"nbody_k3:{kind}", lambda kind=kind: run_single(
            args, L, f"nbody_k3:{kind}", "fp64", 3, 1, cpu, cs))
    guard("mixture_k64_varbw", lambda: run_single(args, L, "mixture_k64_varbw", "auto", steps,
                                                  warm, cpu, cs))
    for k in (3, 8, 32):
        guard(f"list_pool_k{k}", lambda k=k: run_list(args, L, f"list_pool_k{k}", "fp64", steps,
                                                      warm, cpu))
        guard(f"list_pool_k{k}_mixed", lambda k=k: run_list(
            args, L, f"list_pool_k{k}", "mixed_bf16x3", steps, warm, False))
    guard("list_sym_k3", lambda: run_list(args, L, "list_sym_k3", "fp64", steps, warm, cpu))
    guard("grad_gauss16", lambda: run_grad(args, "grad_gauss16", "auto", steps, warm, cpu))
    guard("grad_varbw16_maps", lambda: run_grad_maps(args, "grad_varbw16_maps", steps, warm, cpu))
    return subs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=None,
                    help="one workload alone (default: the k = 64 mixture headline with the "
                         "other shapes as sub-records)")
    ap.add_argument("--precision", default="auto",
                    choices=["auto", "fp64", "mixed_fp32", "mixed_bf16x3"],
                    help="auto: the mixed BF16x3 form beside the FP64 drop-in for the "
                         "Gaussian kernels, the FP64 drop-in otherwise")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-suite", action="store_true",
                    help="the headline workload without the sub-records")
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--sub-cpu-seconds", type=float, default=2.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        nb = args.copy_calibration << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nb, "copy_bytes_written": nb}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the fused summation is a single-GPU kernel "
              "here (it would shard by target rows and source slices; not run)",
              file=sys.stderr)
        sys.exit(2)

    import torch  # noqa: F401
    from ks_amd import _lib

    L = _lib.lib()
    cpu = not args.no_cpu_baseline
    name = args.workload
    if name is not None:
        if name.startswith("list_pool") or name.startswith("list_sym"):
            out = run_list(args, L, name, args.precision, args.steps, args.warmup, cpu)
        elif name == "grad_gauss16":
            out = run_grad(args, name, args.precision, args.steps, args.warmup, cpu)
        elif name == "grad_varbw16_maps":
            out = run_grad_maps(args, name, args.steps, args.warmup, cpu)
        else:
            out = run_single(args, L, name, args.precision, args.steps, args.warmup, cpu,
                             args.cpu_seconds, graph=name == "uniform_gauss16",
                             with_sgsks=name.startswith("mixture"))
        print(json.dumps(out), flush=True)
        return
    out = run_single(args, L, "mixture_k64", args.precision, args.steps, args.warmup, cpu,
                     args.cpu_seconds, with_sgsks=True)
    print(f"[bench] mixture_k64: done, value {out['value']:.4e} evals/s", file=sys.stderr,
          flush=True)
    if not args.no_suite:
        out["subrecords"] = suite(args, L)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
