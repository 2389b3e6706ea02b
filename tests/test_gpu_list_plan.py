"""GPU tests of the reusable batch plans (ks_list_plan_*), HIP graph capture of
the stream-ordered entry points, the tile-shape autotuner and the routing of
mixed-precision batches.  FP64 results: relative L2 <= 1e-12 against the
oracle (the reference's error metric, test/test_dgsks.c:43-79); graph replays
must equal the eager run bit for bit."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


@pytest.fixture(scope="module")
def L():
    from ks_amd import _lib
    return _lib.lib()


def _batch(k, seed, pool=9000, ntask=36, big_gauss=0, big=4200, sep_maps=False):
    """Tasks over one pool: Gaussian / Laplace / polynomial kernels, shared
    targets, an empty task, sizes across the tile slices; `big_gauss` extra
    Gaussian tasks of big x big (16384 x 16384 reaches the mixed kernels in a
    mixed-precision batch).  sep_maps: umap and wmap are other permutations of
    the pool than amap and bmap (umap[i] = pu[amap[i]], wmap[j] = pw[bmap[j]]),
    so a gather that reads the wrong map, or a scatter through amap, shows."""
    from ks_amd import _lib
    rng = np.random.default_rng(seed)
    X = rng.random((pool, k)) * (1.5 / np.sqrt(k))
    X2 = np.zeros(pool)
    for p in range(k):
        X2 += X[:, p] * X[:, p]
    w = rng.standard_normal(pool)
    ms = list(rng.integers(1, 300, ntask)) + [big] * big_gauss
    ns = list(rng.integers(1, 2600, ntask)) + [big] * big_gauss
    ms[2] = 0
    # Laplace needs k >= 3 (frame/dgsks.c:445-447 divides by k - 2): a Gaussian of
    # another width takes its place below that
    lap = dict(type=2) if k >= 3 else dict(type=0, scal=-0.7)
    kinds = [dict(type=0, scal=-1.5), lap, dict(type=1, scal=1.0 / k, cons=1.0, powe=2.0)]
    kernels_py = [kinds[t % 3] for t in range(ntask)] + [dict(type=0, scal=-40.0)] * big_gauss
    amap = np.concatenate([rng.choice(pool, int(s), replace=False).astype(np.int32) for s in ms])
    bmap = np.concatenate([rng.choice(pool, int(s), replace=False).astype(np.int32) for s in ns])
    a_off = np.concatenate([[0], np.cumsum(ms)]).astype(np.int64)
    b_off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    nt = len(kernels_py)
    kernels = (_lib.KsT * nt)()
    for t, kd in enumerate(kernels_py):
        kernels[t].type = kd["type"]
        kernels[t].scal = kd.get("scal", 0.0)
        kernels[t].cons = kd.get("cons", 0.0)
        kernels[t].powe = kd.get("powe", 0.0)
    umap, wmap = amap, bmap
    if sep_maps:
        pu = rng.permutation(pool).astype(np.int32)
        pw = rng.permutation(pool).astype(np.int32)
        umap, wmap = pu[amap], pw[bmap]
    return dict(k=k, X=X, X2=X2, w=w, amap=amap, bmap=bmap, umap=umap, wmap=wmap, a_off=a_off,
                b_off=b_off, kernels=kernels, kernels_py=kernels_py, ntask=nt, pool=pool)


class _Plan:
    """A plan over device copies of the batch's maps."""

    def __init__(self, L, torch, B, precision=0):
        self.L, self.torch = L, torch
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
        self.amap, self.bmap = dev(B["amap"]), dev(B["bmap"])
        self.umap, self.wmap = dev(B["umap"]), dev(B["wmap"])
        self.X, self.X2, self.w = dev(B["X"]), dev(B["X2"]), dev(B["w"])
        self.h = ctypes.c_void_p()
        P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        hp = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
        rc = L.ks_list_plan_create(ctypes.byref(self.h), B["ntask"], B["kernels"], B["k"],
                                   hp(B["a_off"]), P(self.umap), P(self.amap), hp(B["b_off"]),
                                   P(self.bmap), P(self.wmap), precision)
        assert rc == 0, rc

    def run(self, u, stream):
        P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
        return self.L.ks_list_plan_execute(self.h, P(u), P(self.X), P(self.X2), P(self.X),
                                           P(self.X2), P(self.w), ctypes.c_void_p(stream.cuda_stream))

    def close(self):
        self.L.ks_list_plan_destroy(self.h)


def _oracle_list(oracle, B, u0):
    ref = u0.copy()
    oracle.dgsks_list(B["kernels_py"], B["k"], ref, B["a_off"], B["umap"], B["amap"], B["b_off"],
                      B["bmap"], B["wmap"], B["X"], B["X2"], B["X"], B["X2"], B["w"])
    return ref


@pytest.mark.parametrize("k,sep_maps", [(3, False), (20, False), (1, True), (2, True),
                                        (3, True), (5, True), (7, True), (20, True)])
def test_plan_matches_oracle_and_dgsks_list(L, torch_cuda, oracle, k, sep_maps):
    """k < 8 runs the compact source pool (templated gathers for k = 1..7),
    k = 20 the padded row tiles or matrix-core tiles; with sep_maps the target
    and weight maps differ from the point maps."""
    torch = torch_cuda
    B = _batch(k, 400 + k, sep_maps=sep_maps)
    rng = np.random.default_rng(1)
    u0 = rng.standard_normal(B["pool"])
    ref = _oracle_list(oracle, B, u0)
    plan = _Plan(L, torch, B)
    try:
        s = torch.cuda.current_stream()
        u = torch.from_numpy(u0.copy()).cuda()
        assert plan.run(u, s) == 0
        torch.cuda.synchronize()
        got = u.cpu().numpy()
        err = oracle.rel_l2(got, ref)
        print(f"plan k={k} sep_maps={sep_maps}: rel L2 = {err:.3e}")
        assert err <= TOL
        # the one-shot API lays out the same tiles, so it agrees bit for bit
        u1 = u0.copy()
        p = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
        kern = B["kernels"]
        assert L.dgsks_list(B["ntask"], kern, k, p(u1), u1.size, p(B["a_off"]), p(B["umap"]),
                            p(B["amap"]), p(B["b_off"]), p(B["bmap"]), p(B["wmap"]), p(B["X"]),
                            p(B["X2"]), p(B["X"]), p(B["X2"]), p(B["w"]), 0) == 0
        assert np.array_equal(u1, got)
    finally:
        plan.close()


def _replay_check(torch, run_on, u, reps=4):
    """Eager run, then a captured run replayed `reps` times with eager work and
    synchronisations between replays; every replay must equal the eager result."""
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        u.zero_()
        assert run_on(s) == 0  # eager (also sizes any scratch on this stream)
    torch.cuda.synchronize()
    eager = u.clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        assert run_on(s) == 0
    torch.cuda.synchronize()
    junk = torch.ones(1 << 20, dtype=torch.float64, device="cuda")
    for r in range(reps):
        u.zero_()
        junk.mul_(1.0001).add_(0.5)  # unrelated eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(u, eager), f"replay {r} differs from the eager run"
    return eager


@pytest.mark.parametrize("k", [3, 8, 32])
def test_plan_graph_capture_replays(L, torch_cuda, oracle, k):
    """k = 3 captures the compact-pool gather too; separate target and weight
    maps."""
    torch = torch_cuda
    B = _batch(k, 500 + k, ntask=24, sep_maps=True)
    plan = _Plan(L, torch, B)
    try:
        u = torch.zeros(B["pool"], dtype=torch.float64, device="cuda")
        eager = _replay_check(torch, lambda s: plan.run(u, s), u)
        ref = _oracle_list(oracle, B, np.zeros(B["pool"]))
        assert oracle.rel_l2(eager.cpu().numpy(), ref) <= TOL
    finally:
        plan.close()


@pytest.mark.parametrize("precision", [0, 2])
def test_dgsks_device_graph_capture_replays(L, torch_cuda, oracle, precision):
    """The stream-ordered single summation, FP64 and mixed (its bucketing sort
    and packing included), captured after one eager run on the same stream."""
    from ks_amd import _lib, workloads
    torch = torch_cuda
    d = workloads.mixture_k64(N=8192, k=64)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    X, X2, w, ids = dev(d["X"]), dev(d["X2"]), dev(d["w"]), dev(d["amap"])
    kt = _lib.KsT()
    kt.type = 0
    kt.scal = d["kernel"]["scal"]
    u = torch.zeros(8192, dtype=torch.float64, device="cuda")
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731

    def run_on(s):
        return L.dgsks_device(ctypes.byref(kt), 8192, 8192, 64, P(u), P(ids), P(X), P(X2), P(ids),
                              P(X), P(X2), P(ids), P(w), P(ids), precision,
                              ctypes.c_void_p(s.cuda_stream))

    eager = _replay_check(torch, run_on, u)
    rows = np.arange(0, 8192, 37).astype(np.int32)
    ref = oracle.dgsks_rows(0, d["X"], d["X2"], d["amap"], d["X"], d["X2"], d["amap"], d["w"],
                            d["amap"], np.zeros(8192), d["amap"], rows, scal=d["kernel"]["scal"])
    err = oracle.rel_l2(eager.cpu().numpy()[rows], ref)
    assert err <= (TOL if precision == 0 else 1e-6)


def test_autotune_choices(L, torch_cuda):
    out = []
    for t, k in [(0, 3), (2, 3), (0, 32), (2, 32), (1, 20)]:
        sl, tl = ctypes.c_int(), ctypes.c_int()
        assert L.ks_list_autotune(t, k, 0, ctypes.byref(sl), ctypes.byref(tl)) == 0
        assert sl.value in (256, 512, 1024, 2048, 4096)
        assert tl.value in (0, 1)
        if k < 9:
            assert tl.value == 0  # k <= 8 has no matrix-core tile
        sl2, tl2 = ctypes.c_int(), ctypes.c_int()
        assert L.ks_list_autotune(t, k, 0, ctypes.byref(sl2), ctypes.byref(tl2)) == 0
        assert (sl2.value, tl2.value) == (sl.value, tl.value)  # cached
        out.append((t, k, sl.value, tl.value))
    print("tuned shapes:", out)


def test_laplace_matrix_core_tiles_exact_drops(tmp_path, torch_cuda, oracle):
    """Laplace tasks at k = 32 on the matrix-core tiles (KS_AUTOTUNE=0 gives the
    default matrix-core shape, in a process of its own) over points with exact
    duplicates and near-coincident pairs whose d^2 straddles the reference's
    1e-15 drop threshold (frame/dgsks_ref.c:343-345): the per-pair bound must
    send those pairs through the reference's sequential sum, so every drop
    decision and every kernel value of a kept near pair (up to ~1e225) is the
    reference's, and the batch matches the oracle to 1e-12."""
    rng = np.random.default_rng(2024)
    k, pool = 32, 2400
    X = rng.random((pool, k)) * (1.5 / np.sqrt(k))
    # points 1200.. are copies of points 0.. moved by |delta| in 0 .. 6e-8
    deltas = np.concatenate([[0.0] * 20, 10.0 ** rng.uniform(-12, -7.2, 180)])
    for i, dl in enumerate(deltas):
        d = rng.standard_normal(k)
        X[1200 + i] = X[i] + dl * d / np.linalg.norm(d)
    X2 = (X * X).sum(axis=1)
    w = rng.standard_normal(pool)
    ntask = 6
    ms = rng.integers(150, 400, ntask)
    ns = rng.integers(600, 1800, ntask)
    near = np.arange(200)
    amap = np.concatenate([np.concatenate([near, rng.choice(np.arange(200, pool), m - 200,
                                                            replace=False)])
                           for m in ms]).astype(np.int32)
    bmap = np.concatenate([np.concatenate([near + 1200, near, rng.choice(np.arange(200, pool),
                                                                         n - 400, replace=False)])
                           for n in ns]).astype(np.int32)
    a_off = np.concatenate([[0], np.cumsum(ms)]).astype(np.int64)
    b_off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    u0 = rng.standard_normal(pool)
    data = tmp_path / "batch.npz"
    np.savez(data, X=X, X2=X2, w=w, amap=amap, bmap=bmap, a_off=a_off, b_off=b_off, u0=u0)
    out = tmp_path / "u.npy"
    code = f"""
import ctypes, sys
import numpy as np
sys.path.insert(0, {ROOT!r})
from ks_amd import _lib
L = _lib.lib()
d = np.load({str(data)!r})
k, nt = {k}, {ntask}
kern = (_lib.KsT * nt)()
for t in range(nt):
    kern[t].type = _lib.KS_LAPLACE
sl, tl = ctypes.c_int(), ctypes.c_int()
assert L.ks_list_autotune(_lib.KS_LAPLACE, k, 0, ctypes.byref(sl), ctypes.byref(tl)) == 0
assert tl.value == 1, "the default shape is the matrix-core tile"
p = lambda a: ctypes.c_void_p(a.ctypes.data)
u = d["u0"].copy()
arr = {{n: np.ascontiguousarray(d[n]) for n in d.files}}
rc = L.dgsks_list(nt, kern, k, p(u), u.size, p(arr["a_off"]), p(arr["amap"]), p(arr["amap"]),
                  p(arr["b_off"]), p(arr["bmap"]), p(arr["bmap"]), p(arr["X"]), p(arr["X2"]),
                  p(arr["X"]), p(arr["X2"]), p(arr["w"]), 0)
assert rc == 0, rc
np.save({str(out)!r}, u)
"""
    env = dict(os.environ, KS_AUTOTUNE="0")
    env.pop("KS_TUNE_CACHE", None)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    got = np.load(out)
    ref = u0.copy()
    oracle.dgsks_list([dict(type=2)] * ntask, k, ref, a_off, amap, amap, b_off, bmap, bmap, X, X2,
                      X, X2, w)
    # scaled by the largest entry first: kept near pairs reach ~1e227, whose
    # squares would overflow
    sc = float(np.max(np.abs(ref)))
    assert np.all(np.isfinite(got))
    err = oracle.rel_l2(got / sc, ref / sc)
    big = np.abs(ref) > 1e100
    err_big = oracle.rel_l2(got[big] / sc, ref[big] / sc)
    print(f"Laplace matrix-core tiles, near-coincident pairs: rel L2 = {err:.3e}; "
          f"{int(big.sum())} targets carry a kept near pair (rel L2 {err_big:.3e})")
    assert err <= TOL and err_big <= TOL


@pytest.mark.parametrize("gauss,lap,poly", [(512, 1024, 2048), (256, 4096, 512),
                                            (4096, 2048, 256)])
def test_mixed_slices_per_kernel_type(tmp_path, torch_cuda, oracle, gauss, lap, poly):
    """Regression: tasks of different kernel types with different tuned slice
    widths (forced through a KS_TUNE_CACHE file, in a process of its own) put
    split and unsplit row blocks of every size next to each other in the tile
    order.  With lane-0-only branches in the persistent loop, these layouts
    once gave wrong sums that changed from run to run (lanes 1..63 of a wave
    started an iteration without lane 0 after a row block's last tile).  Every
    run must match the oracle, and repeated plan runs must agree bit for bit."""
    cache = tmp_path / "tiles.txt"
    cache.write_text(f"gfx950 0 3 0 {gauss} 0\ngfx950 2 3 0 {lap} 0\ngfx950 1 3 0 {poly} 0\n")
    out = tmp_path / "u.npz"
    code = f"""
import ctypes, sys
import numpy as np
import torch
sys.path.insert(0, {ROOT!r}); sys.path.insert(0, {HERE!r})
from ks_amd import _lib
import test_gpu_list_plan as T
L = _lib.lib()
sl, tl = ctypes.c_int(), ctypes.c_int()
for t, want in ((0, {gauss}), (2, {lap}), (1, {poly})):
    assert L.ks_list_autotune(t, 3, 0, ctypes.byref(sl), ctypes.byref(tl)) == 0
    assert (sl.value, tl.value) == (want, 0), (t, sl.value, tl.value)
B = T._batch(3, 403)
u0 = np.random.default_rng(1).standard_normal(B["pool"])
p = lambda a: ctypes.c_void_p(a.ctypes.data)
outs = {{}}
for rep in range(2):
    u = u0.copy()
    rc = L.dgsks_list(B["ntask"], B["kernels"], 3, p(u), u.size, p(B["a_off"]), p(B["amap"]),
                      p(B["amap"]), p(B["b_off"]), p(B["bmap"]), p(B["bmap"]), p(B["X"]),
                      p(B["X2"]), p(B["X"]), p(B["X2"]), p(B["w"]), 0)
    assert rc == 0, rc
    outs[f"oneshot{{rep}}"] = u
plan = T._Plan(L, torch, B)
for rep in range(3):
    u = torch.from_numpy(u0.copy()).cuda()
    assert plan.run(u, torch.cuda.current_stream()) == 0
    torch.cuda.synchronize()
    outs[f"plan{{rep}}"] = u.cpu().numpy()
plan.close()
np.savez({str(out)!r}, u0=u0, **outs)
"""
    env = dict(os.environ, KS_TUNE_CACHE=str(cache))
    env.pop("KS_AUTOTUNE", None)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    d = np.load(out)
    B = _batch(3, 403)
    ref = _oracle_list(oracle, B, d["u0"])
    runs = [n for n in d.files if n != "u0"]
    errs = {n: oracle.rel_l2(d[n], ref) for n in runs}
    print(f"slices gauss {gauss} laplace {lap} poly {poly}: " +
          " ".join(f"{n} {e:.2e}" for n, e in errs.items()))
    assert max(errs.values()) <= TOL
    assert all(np.array_equal(d["plan0"], d[n]) for n in runs)


def test_autotune_cache_file(tmp_path, torch_cuda):
    """KS_TUNE_CACHE: a fresh process tunes and appends its choice under the
    device's architecture; a second process reads it back without tuning
    (KS_AUTOTUNE=0 would otherwise give the default shape).  Lines of another
    architecture or of the older five-field format are ignored."""
    cache = tmp_path / "tiles.txt"
    code = ("import ctypes,sys; sys.path.insert(0, %r); from ks_amd import _lib; L=_lib.lib(); "
            "a,b=ctypes.c_int(),ctypes.c_int(); "
            "assert L.ks_list_autotune(4, 12, 0, ctypes.byref(a), ctypes.byref(b))==0; "
            "print(a.value, b.value)") % ROOT
    env = dict(os.environ, KS_TUNE_CACHE=str(cache))
    r1 = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                        timeout=300)
    assert r1.returncode == 0, r1.stderr
    lines = cache.read_text().split("\n")
    assert lines[0].split()[:4] == ["gfx950", "4", "12", "0"]
    env2 = dict(env, KS_AUTOTUNE="0")
    r2 = subprocess.run([sys.executable, "-c", code], env=env2, capture_output=True, text=True,
                        timeout=300)
    assert r2.returncode == 0, r2.stderr
    assert r2.stdout.split() == r1.stdout.split()
    # a cache of another architecture and of the five-field format gives the
    # default shape (slice 1024), not the 256 those lines name
    foreign = tmp_path / "foreign.txt"
    foreign.write_text("gfx942 4 12 0 256 0\n4 12 0 256 0\n")
    env3 = dict(os.environ, KS_TUNE_CACHE=str(foreign), KS_AUTOTUNE="0")
    r3 = subprocess.run([sys.executable, "-c", code], env=env3, capture_output=True, text=True,
                        timeout=300)
    assert r3.returncode == 0, r3.stderr
    assert r3.stdout.split()[0] == "1024"


def test_mixed_batch_routing(L, torch_cuda, oracle):
    """A mixed-precision batch: Laplace and polynomial tasks and small Gaussian
    tasks run in the persistent FP64 launch, the large Gaussian tasks
    (16384 x 16384, k = 16) in the mixed kernels; the whole batch stays within
    1e-6 of the oracle."""
    B = _batch(16, 600, pool=20000, ntask=30, big_gauss=2, big=16384)
    u = np.zeros(B["pool"])
    ref = _oracle_list(oracle, B, u)
    p = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    rc = L.dgsks_list(B["ntask"], B["kernels"], 16, p(u), u.size, p(B["a_off"]), p(B["amap"]),
                      p(B["amap"]), p(B["b_off"]), p(B["bmap"]), p(B["bmap"]), p(B["X"]),
                      p(B["X2"]), p(B["X"]), p(B["X2"]), p(B["w"]), 2)
    assert rc == 0
    err = oracle.rel_l2(u, ref)
    print(f"mixed batch: rel L2 = {err:.3e}")
    assert err <= 1e-6


def test_scratch_growth_after_capture_keeps_graph_valid(L, torch_cuda, oracle):
    """A graph captured from dgsks_device keeps the addresses of its stream's
    scratch.  A later, larger call on the same stream grows that scratch: the
    captured buffers must stay valid (retired, not freed), so replays still give
    the eager result, and memory allocated after the growth is not written by
    the replays."""
    from ks_amd import _lib
    torch = torch_cuda
    rng = np.random.default_rng(77)
    k = 16
    N1, N2 = 3000, 60000  # the larger call needs more split partials
    X = rng.random((N2, k))
    X2 = oracle.sqnorm(X)
    w = rng.standard_normal(N2)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    dX, dX2, dw = dev(X), dev(X2), dev(w)
    ids = dev(np.arange(N2, dtype=np.int32))
    kt = _lib.KsT()
    kt.type = 0
    kt.scal = -2.0
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    u1 = torch.zeros(N2, dtype=torch.float64, device="cuda")
    u2 = torch.zeros(N2, dtype=torch.float64, device="cuda")

    def call(n, u, s):
        return L.dgsks_device(ctypes.byref(kt), n, n, k, P(u), P(ids), P(dX), P(dX2), P(ids),
                              P(dX), P(dX2), P(ids), P(dw), P(ids), 0,
                              ctypes.c_void_p(s.cuda_stream))

    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        assert call(N1, u1, s) == 0  # sizes the scratch for N1
    torch.cuda.synchronize()
    eager = u1.clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        assert call(N1, u1, s) == 0
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        assert call(N2, u2, s) == 0  # grows the scratch of stream s
    torch.cuda.synchronize()
    rows = np.arange(0, N2, 97).astype(np.int32)
    ref2 = oracle.dgsks_rows(0, X, X2, np.arange(N2, dtype=np.int32), X, X2,
                             np.arange(N2, dtype=np.int32), w, np.arange(N2, dtype=np.int32),
                             np.zeros(N2), np.arange(N2, dtype=np.int32), rows, scal=-2.0)
    assert oracle.rel_l2(u2.cpu().numpy()[rows], ref2) <= TOL
    fresh = [torch.full((1 << 20,), 3.0, dtype=torch.float64, device="cuda") for _ in range(8)]
    for r in range(3):
        u1.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(u1, eager), f"replay {r} after the scratch grew differs"
    assert all(bool(torch.all(f == 3.0)) for f in fresh)


def test_mixed_plan_capture(L, torch_cuda, oracle):
    """A mixed-precision plan with tasks on the mixed kernels: capturing it on a
    stream whose scratch has not grown yet is refused with KS_ERR_ALLOC (no
    allocation or synchronisation under capture); after one eager run on that
    stream it captures, and every replay equals the eager run."""
    from ks_amd import _lib
    torch = torch_cuda
    B = _batch(16, 610, pool=20000, ntask=12, big_gauss=1, big=16384, sep_maps=True)
    plan = _Plan(L, torch, B, precision=2)
    try:
        u = torch.zeros(B["pool"], dtype=torch.float64, device="cuda")
        s0 = torch.cuda.Stream()
        g0 = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g0, stream=s0):
            rc = plan.run(u, s0)
        assert rc == _lib.KS_ERR_ALLOC, rc
        del g0
        eager = _replay_check(torch, lambda s: plan.run(u, s), u)
        ref = _oracle_list(oracle, B, np.zeros(B["pool"]))
        err = oracle.rel_l2(eager.cpu().numpy(), ref)
        print(f"mixed plan (captured): rel L2 = {err:.3e}")
        assert err <= 1e-6
    finally:
        plan.close()


def _sym_batch(k, seed, pool=12000, ntask=18):
    """Tasks whose source list is their target list (m == n, 128 <= m <= 2600:
    every kernel type, variable bandwidth with hi == hj), a task of the same
    size whose sources are its targets reordered (not symmetric), and plain
    tasks; umap and wmap are other permutations of the pool."""
    rng = np.random.default_rng(seed)
    X = rng.random((pool, k)) * (1.5 / np.sqrt(k))
    X2 = np.zeros(pool)
    for p in range(k):
        X2 += X[:, p] * X[:, p]
    w = rng.standard_normal(pool)
    h = rng.uniform(1.5, 3.0, pool)  # 1/h of the variable-bandwidth form
    kinds = [dict(type=0, scal=-1.5), dict(type=2), dict(type=1, scal=1.0 / k, cons=1.0, powe=2.0),
             dict(type=3), dict(type=4, scal=0.3, cons=0.1), dict(type=5), dict(type=6, cons=0.5),
             dict(type=7), dict(type=1, scal=0.5 / k, cons=0.5, powe=3.0)]
    amaps, bmaps, kernels_py, nsym = [], [], [], 0
    for t in range(ntask):
        kd = dict(kinds[t % len(kinds)])
        m = int(rng.integers(128, 2600))
        a = rng.choice(pool, m, replace=False).astype(np.int32)
        if t % 6 == 5:  # same set, other order: evaluated in full
            b = rng.permutation(a).astype(np.int32)
        elif t % 6 == 4:  # plain
            b = rng.choice(pool, int(rng.integers(1, 3000)), replace=False).astype(np.int32)
        else:
            b = a.copy()
            nsym += 1
        if kd["type"] == 2 and k < 3:
            kd = dict(type=0, scal=-0.7)
        amaps.append(a)
        bmaps.append(b)
        kernels_py.append(kd)
    ms = [len(a) for a in amaps]
    ns = [len(b) for b in bmaps]
    amap, bmap = np.concatenate(amaps), np.concatenate(bmaps)
    pu = rng.permutation(pool).astype(np.int32)
    pw = rng.permutation(pool).astype(np.int32)
    return dict(k=k, X=X, X2=X2, w=w, h=h, amap=amap, bmap=bmap, umap=pu[amap], wmap=pw[bmap],
                a_off=np.concatenate([[0], np.cumsum(ms)]).astype(np.int64),
                b_off=np.concatenate([[0], np.cumsum(ns)]).astype(np.int64),
                kernels_py=kernels_py, ntask=ntask, pool=pool, nsym=nsym)


def _kernel_array(B, hptr):
    from ks_amd import _lib
    arr = (_lib.KsT * B["ntask"])()
    for t, kd in enumerate(B["kernels_py"]):
        arr[t].type = kd["type"]
        arr[t].scal = kd.get("scal", 0.0)
        arr[t].cons = kd.get("cons", 0.0)
        arr[t].powe = kd.get("powe", 0.0)
        if kd["type"] == 3:
            arr[t].hi = arr[t].hj = hptr
    return arr


@pytest.mark.parametrize("k", [3, 8, 20, 40])
def test_symmetric_list_tasks(L, torch_cuda, oracle, k):
    """Tasks whose source list is their target list evaluate each pair of
    different row blocks once at k = 3 and 8 (ks_list_plan_symmetric_tasks counts
    them; at k = 20 and 40 none are, the full matrix-core layout being faster): a
    plan run with one table for targets and sources, its HIP graph replays, a run
    with a separate copy of the table (the full layout the plan keeps), and the
    one-shot dgsks_list on host arrays all match the oracle to 1e-12."""
    torch = torch_cuda
    B = _sym_batch(k, 900 + k)
    kpy = [dict(kd, hi=B["h"], hj=B["h"]) if kd["type"] == 3 else kd for kd in B["kernels_py"]]
    u0 = np.random.default_rng(2).standard_normal(B["pool"])
    ref = u0.copy()
    oracle.dgsks_list(kpy, k, ref, B["a_off"], B["umap"], B["amap"], B["b_off"], B["bmap"],
                      B["wmap"], B["X"], B["X2"], B["X"], B["X2"], B["w"])
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    H = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    dh = dev(B["h"])
    kern = _kernel_array(B, dh.data_ptr())
    amap, bmap, umap, wmap = map(dev, (B["amap"], B["bmap"], B["umap"], B["wmap"]))
    X, X2, w = dev(B["X"]), dev(B["X2"]), dev(B["w"])
    Xc, X2c = X.clone(), X2.clone()
    h = ctypes.c_void_p()
    assert L.ks_list_plan_create(ctypes.byref(h), B["ntask"], kern, k, H(B["a_off"]), P(umap),
                                 P(amap), H(B["b_off"]), P(bmap), P(wmap), 0) == 0
    try:
        nsym = L.ks_list_plan_symmetric_tasks(h)
        assert nsym == (B["nsym"] if k in (3, 8) else 0), (nsym, B["nsym"])
        s = torch.cuda.Stream()

        def run(xb, xb2):
            return L.ks_list_plan_execute(h, P(u), P(X), P(X2), P(xb), P(xb2), P(w),
                                          ctypes.c_void_p(s.cuda_stream))
        u = dev(u0.copy())
        with torch.cuda.stream(s):
            assert run(X, X2) == 0
        torch.cuda.synchronize()
        got = u.cpu().numpy()
        err = oracle.rel_l2(got, ref)
        print(f"symmetric list k={k} ({nsym} symmetric tasks): rel L2 = {err:.3e}")
        assert err <= TOL
        # replays of a captured run equal the eager run
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            assert run(X, X2) == 0
        torch.cuda.synchronize()
        u0d = dev(u0)
        for r in range(3):
            u.copy_(u0d)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert np.array_equal(u.cpu().numpy(), got), f"replay {r} differs"
        del g
        # another table for the sources (same values): the full layout
        u.copy_(u0d)
        with torch.cuda.stream(s):
            assert run(Xc, X2c) == 0
        torch.cuda.synchronize()
        err_full = oracle.rel_l2(u.cpu().numpy(), ref)
        print(f"  full layout: rel L2 = {err_full:.3e}")
        assert err_full <= TOL
    finally:
        L.ks_list_plan_destroy(h)
    # the one-shot call on host arrays, one table
    kern_h = _kernel_array(B, B["h"].ctypes.data)
    u1 = u0.copy()
    rc = L.dgsks_list(B["ntask"], kern_h, k, H(u1), u1.size, H(B["a_off"]), H(B["umap"]),
                      H(B["amap"]), H(B["b_off"]), H(B["bmap"]), H(B["wmap"]), H(B["X"]),
                      H(B["X2"]), H(B["X"]), H(B["X2"]), H(B["w"]), 0)
    assert rc == 0
    err1 = oracle.rel_l2(u1, ref)
    print(f"  dgsks_list: rel L2 = {err1:.3e}")
    assert err1 <= TOL
