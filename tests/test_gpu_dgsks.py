"""GPU parity of the FP64 fused summation (dgsks through the C ABI) against the
oracle and the reference's golden outputs.  Tolerance: relative L2 <= 1e-12
over the updated rows (the reference's error metric, test/test_dgsks.c:43-79;
the accuracy bar of the FP64 drop-in)."""
import ctypes

import numpy as np
import pytest

from conftest import golden_cases

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


@pytest.fixture(scope="module")
def G():
    from ks_amd import gsks
    gsks._lib.lib()
    return gsks


def _opts(t, scal=0.0, cons=0.0, powe=0.0, h=None):
    from ks_amd.gsks import TYPES
    name = [k for k, v in TYPES.items() if v == t][0]
    o = dict(type=name, scal=scal, cons=cons, powe=powe)
    if h is not None:
        o["h"] = h
    return o


def _dev(torch, *arrs):
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


# ------------------------------------------------------------------ golden fixtures

@pytest.mark.parametrize("path", ["host", "device", "reference_algorithm"])
def test_golden(G, golden, torch_cuda, path):
    torch = torch_cuda
    for key, d in golden_cases(golden):
        t, m, n, k = (int(x) for x in d["meta"])
        scal, cons, powe = d["par"][:3]
        h = d.get("h")
        X, X2 = d["X"], d["X2"]
        amap, umap, bmap, wmap = d["amap"], d["umap"], d["bmap"], d["wmap"]
        if path == "host":
            u = d["u0"].copy()
            G.dgsks(k, u, umap, X, X2, amap, X, X2, bmap, d["w"], wmap, **_opts(t, scal, cons, powe, h))
        elif path == "reference_algorithm":
            u = d["u0"].copy()
            G.dgsks_ref(k, u, umap, X, X2, amap, X, X2, bmap, d["w"], wmap,
                        **_opts(t, scal, cons, powe, h))
        else:
            du, dumap, dX, dX2, damap, dbmap, dw, dwmap = _dev(
                torch, d["u0"], umap, X, X2, amap, bmap, d["w"], wmap)
            dh = _dev(torch, h)[0] if h is not None else None
            G.dgsks(k, du, dumap, dX, dX2, damap, dX, dX2, dbmap, dw, dwmap,
                    **_opts(t, scal, cons, powe, dh))
            torch.cuda.synchronize()
            u = du.cpu().numpy()
        ref = d["ref_u"]
        rows = umap[:m]
        err = np.linalg.norm(u[rows] - ref[rows]) / np.linalg.norm(ref[rows])
        tol = 1e-14 if path == "reference_algorithm" else TOL
        assert err <= tol, (key, t, path, err)
        untouched = np.setdiff1d(np.arange(u.size), rows)
        assert np.array_equal(u[untouched], d["u0"][untouched]), key


# ------------------------------------------------------------------ all types x k

PARAMS = {0: dict(scal=-1.3), 1: dict(scal=0.45, cons=0.25, powe=3.0), 2: dict(), 3: dict(),
          4: dict(scal=0.3, cons=0.1), 5: dict(), 6: dict(cons=0.5), 7: dict()}


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 8, 16, 32, 64])
def test_all_types_random_maps(G, oracle, torch_cuda, k):
    torch = torch_cuda
    rng = np.random.default_rng(1000 + k)
    nx = 3000
    m, n = 1237, 2011  # ragged: not multiples of any tile
    X = rng.random((nx, k)) * (1.2 / np.sqrt(k))
    X2 = oracle.sqnorm(X)
    amap = rng.permutation(nx)[:m].astype(np.int32)
    umap = rng.permutation(nx)[:m].astype(np.int32)
    bmap = rng.permutation(nx)[:n].astype(np.int32)
    wmap = rng.permutation(nx)[:n].astype(np.int32)
    w = rng.standard_normal(nx)
    u0 = rng.standard_normal(nx)
    hv = rng.uniform(1.0, 3.0, nx)
    dX, dX2, damap, dumap, dbmap, dwmap, dw, dh = _dev(torch, X, X2, amap, umap, bmap, wmap, w, hv)
    for t in range(8):
        if t == 2 and k < 3:
            continue
        prm = PARAMS[t]
        h = hv if t == 3 else None
        du = _dev(torch, u0)[0]
        G.dgsks(k, du, dumap, dX, dX2, damap, dX, dX2, dbmap, dw, dwmap,
                **_opts(t, h=dh if t == 3 else None, **prm))
        torch.cuda.synchronize()
        got = du.cpu().numpy()[umap]
        ref = oracle.dgsks_rows(t, X, X2, amap, X, X2, bmap, w, wmap, u0, umap, hi=h, hj=h, **prm)
        err = oracle.rel_l2(got, ref)
        assert err <= TOL, (t, k, err)


def test_distinct_tables_and_bandwidths(G, oracle, torch_cuda):
    """XA != XB, hi != hj, umap != amap."""
    torch = torch_cuda
    rng = np.random.default_rng(77)
    k, na, nb, m, n = 16, 900, 1500, 700, 1400
    A, B = rng.random((na, k)) * 0.4, rng.random((nb, k)) * 0.4
    A2, B2 = oracle.sqnorm(A), oracle.sqnorm(B)
    amap = rng.integers(0, na, m).astype(np.int32)
    umap = rng.permutation(2000)[:m].astype(np.int32)
    bmap = rng.integers(0, nb, n).astype(np.int32)
    wmap = rng.integers(0, 1800, n).astype(np.int32)
    w = rng.standard_normal(1800)
    u0 = rng.standard_normal(2000)
    hi, hj = rng.uniform(0.5, 2, na), rng.uniform(0.5, 2, nb)
    from ks_amd import _lib
    kt = _lib.KsT()
    kt.type = 3
    dA, dA2, dB, dB2, damap, dumap, dbmap, dwmap, dw, du, dhi, dhj = _dev(
        torch, A, A2, B, B2, amap, umap, bmap, wmap, w, u0, hi, hj)
    kt.hi, kt.hj = dhi.data_ptr(), dhj.data_ptr()
    p = lambda x: ctypes.c_void_p(x.data_ptr())  # noqa: E731
    rc = _lib.lib().dgsks_device(ctypes.byref(kt), m, n, k, p(du), p(dumap), p(dA), p(dA2),
                                 p(damap), p(dB), p(dB2), p(dbmap), p(dw), p(dwmap), 0,
                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    ref = oracle.dgsks_rows(3, A, A2, amap, B, B2, bmap, w, wmap, u0, umap, hi=hi, hj=hj)
    assert oracle.rel_l2(du.cpu().numpy()[umap], ref) <= TOL


@pytest.mark.parametrize("k", [3, 16])
def test_gaussian_exp_ulps(G, torch_cuda, k):
    """The FP64 kernels' exp (exp_tab, a table-driven exp of their own) against numpy's (correctly rounded in practice)
    over [-745.5, 0] and near its edges: one source at the origin with weight 1
    and targets whose norms are x, so u_i = exp(-x_i) exactly (d^2 = x_i; the
    coordinates are zero).  k = 3 runs the vector kernel, k = 16 the matrix-core
    one.  Bar: 2 ulps in the normal range, 1 ulp of the smallest subnormal
    below it, exact 1 at x = 0 and 0 past the underflow."""
    torch = torch_cuda
    rng = np.random.default_rng(11 + k)
    x = np.concatenate([[0.0, 5e-324, 1e-300, 1e-17, 0.5, 1.0, np.log(2), 708.0, 708.4, 709.0,
                         744.0, 745.1, 745.2, 800.0, 1e6],
                        rng.uniform(0, 1, 4000), rng.uniform(0, 50, 4000),
                        rng.uniform(0, 745.2, 4000)])
    m = x.size
    XA = np.zeros((m, k))
    XB = np.zeros((1, k))
    dXA, dXA2, dXB, dXB2 = _dev(torch, XA, x, XB, np.zeros(1))
    idx = np.arange(m, dtype=np.int32)
    dm, dz, dw = _dev(torch, idx, np.zeros(1, dtype=np.int32), np.ones(1))
    du = torch.zeros(m, dtype=torch.float64, device="cuda")
    G.dgsks(k, du, dm, dXA, dXA2, dm, dXB, dXB2, dz, dw, dz, type="Gaussian", scal=-1.0)
    torch.cuda.synchronize()
    got = du.cpu().numpy()
    want = np.exp(-x)
    assert got[0] == 1.0 and got[x > 745.2].max() == 0.0
    tiny = np.finfo(float).tiny
    ulp = np.spacing(np.maximum(np.abs(want), tiny))
    err = np.abs(got - want) / ulp
    normal = want >= tiny
    print(f"k={k}: max ulp error {err[normal].max():.2f} (normal), "
          f"{np.abs(got - want)[~normal].max() / 5e-324:.2f} x 2^-1074 (subnormal)")
    assert err[normal].max() <= 2.0
    assert np.abs(got - want)[~normal].max() <= 5e-324


@pytest.mark.parametrize("k", [3, 16])
def test_tanh_ulps(G, torch_cuda, k):
    """The FP64 kernels' tanh (-e / (2 + e), e = expm1(-2|x|), sign restored)
    against numpy's over both signs, tiny to saturating arguments and the
    special values: one source e_1 with weight 1 and targets (x_i, 0, ...) with
    scal = 1, cons = 0, so c = x_i exactly and u_i = tanh(x_i).  k = 3 runs the
    vector kernel, k = 16 the matrix-core one.  Bar: 4 ulps (3 in a CPU
    restatement with libm's expm1), 0 and +-1 exact, NaN kept (u starts at 0,
    so tanh(-0) arrives as +0)."""
    torch = torch_cuda
    rng = np.random.default_rng(21 + k)
    x = np.concatenate([[0.0, -0.0, 5e-324, 1e-300, 1e-17, 1e-8, 0.25, 0.5493, 0.55, 1.0, 10.0,
                         19.0, 20.0, 40.0, 1e300, np.inf, -np.inf, np.nan],
                        10 ** rng.uniform(-300, -8, 2000), 10 ** rng.uniform(-8, 0, 4000),
                        rng.uniform(0, 1.2, 4000), rng.uniform(0.5, 25, 4000)])
    x = np.concatenate([x, -x[2:]])
    m = x.size
    XA = np.zeros((m, k))
    XA[:, 0] = x
    XB = np.zeros((1, k))
    XB[0, 0] = 1.0
    dXA, dXA2, dXB, dXB2 = _dev(torch, XA, np.zeros(m), XB, np.ones(1))
    idx = np.arange(m, dtype=np.int32)
    dm, dz, dw = _dev(torch, idx, np.zeros(1, dtype=np.int32), np.ones(1))
    du = torch.zeros(m, dtype=torch.float64, device="cuda")
    G.dgsks(k, du, dm, dXA, dXA2, dm, dXB, dXB2, dz, dw, dz, type="Tanh", scal=1.0, cons=0.0)
    torch.cuda.synchronize()
    got = du.cpu().numpy()
    want = np.tanh(x)
    nan = np.isnan(want)
    assert np.all(np.isnan(got[nan]))
    exact = (want == 0) | (np.abs(want) == 1)
    assert np.array_equal(got[exact], want[exact])
    rest = ~nan & ~exact
    err = np.abs(got[rest] - want[rest]) / np.spacing(np.abs(want[rest]))
    print(f"k={k}: tanh max ulp error {err.max():.2f}, mean {err.mean():.3f}")
    assert err.max() <= 4.0


@pytest.mark.parametrize("k", [3, 16, 64])
def test_clamp_semantics_with_inconsistent_norms(G, oracle, torch_cuda, k):
    """Norms lowered by 1e-7 (relative) and every target also a source: d^2 of
    each coincident pair is negative.  dgsks clamps it at 0 like the reference's
    fused micro-kernels (the oracle's clamp mode), dgsks_ref does not (its
    default mode); both to 1e-12, on the general, symmetric (k = 3, one map
    tensor) and matrix-core paths, and the mixed form to 1e-6."""
    torch = torch_cuda
    rng = np.random.default_rng(500 + k)
    nx, m = 3000, 2500
    X = rng.random((nx, k)) * (1.2 / np.sqrt(k))
    X2 = oracle.sqnorm(X) * (1.0 - 1e-7)
    amap = rng.permutation(nx)[:m].astype(np.int32)
    umap = rng.permutation(nx)[:m].astype(np.int32)
    wmap = rng.permutation(nx)[:m].astype(np.int32)
    w = rng.standard_normal(nx)
    u0 = rng.standard_normal(nx)
    hv = rng.uniform(1.0, 3.0, nx)
    dX, dX2, damap, dumap, dwmap, dw, dh = _dev(torch, X, X2, amap, umap, wmap, w, hv)
    for t in (0, 3, 5, 6, 7):
        prm = PARAMS[t]
        h = hv if t == 3 else None
        ref = {c: oracle.dgsks_rows(t, X, X2, amap, X, X2, amap, w, wmap, u0, umap, hi=h, hj=h,
                                    clamp=c, **prm) for c in (True, False)}
        assert oracle.rel_l2(ref[False], ref[True]) > 1e-10  # the case discriminates
        runs = [("dgsks", damap, "fp64", True, TOL), ("dgsks", damap.clone(), "fp64", True, TOL),
                ("dgsks_ref", damap, "fp64", False, TOL)]
        if t in (0, 3):
            runs.append(("dgsks", damap, "mixed_bf16x3", True, 1e-6))
        for fn, dbmap, prec, clamp, tol in runs:
            du = _dev(torch, u0)[0]
            getattr(G, fn)(k, du, dumap, dX, dX2, damap, dX, dX2, dbmap, dw, dwmap,
                           precision=prec, **_opts(t, h=dh if t == 3 else None, **prm))
            torch.cuda.synchronize()
            err = oracle.rel_l2(du.cpu().numpy()[umap], ref[clamp])
            assert err <= tol, (t, k, fn, prec, err)


# ------------------------------------------------------------------ benchmark shapes

def test_uniform_gauss16_full(G, oracle, torch_cuda):
    """m = n = 8192, k = 16, Gaussian h = 0.5: full oracle, every row."""
    from ks_amd import workloads as W
    torch = torch_cuda
    d = W.uniform_gauss16()
    X, X2, w, idx = d["X"], d["X2"], d["w"], d["amap"]
    u = np.zeros(X.shape[0])
    dX, dX2, dw, didx, du = _dev(torch, X, X2, w, idx, u)
    G.dgsks(16, du, didx, dX, dX2, didx, dX, dX2, didx, dw, didx, type="Gaussian",
            scal=d["kernel"]["scal"])
    torch.cuda.synchronize()
    ref = oracle.dgsks_rows(0, X, X2, idx, X, X2, idx, w, idx, u, idx, scal=d["kernel"]["scal"])
    err = oracle.rel_l2(du.cpu().numpy(), ref)
    print(f"uniform_gauss16 rel L2 = {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("name", ["gaussian", "laplace", "polynomial", "tanh"])
def test_nbody_k3_sampled_rows(G, oracle, torch_cuda, name):
    """m = n = 262144, k = 3, permutation maps; oracle on 256 sampled rows."""
    from ks_amd import workloads as W
    torch = torch_cuda
    d = W.nbody_k3()
    kd = W.NBODY_KERNELS[name]
    X, X2, w, perm = d["X"], d["X2"], d["w"], d["amap"]
    N = X.shape[0]
    dX, dX2, dw, dperm = _dev(torch, X, X2, w, perm)
    du = torch.zeros(N, dtype=torch.float64, device="cuda")
    opts = _opts(kd["type"], kd.get("scal", 0.0), kd.get("cons", 0.0), kd.get("powe", 0.0))
    G.dgsks(3, du, dperm, dX, dX2, dperm, dX, dX2, dperm, dw, dperm, **opts)
    torch.cuda.synchronize()
    rows = np.random.default_rng(12).choice(N, 256, replace=False).astype(np.int32)
    ref = oracle.dgsks_rows(kd["type"], X, X2, perm, X, X2, perm, w, perm, np.zeros(N), perm,
                            rows=rows, scal=kd.get("scal", 0.0), cons=kd.get("cons", 0.0),
                            powe=kd.get("powe", 0.0))
    got = du.cpu().numpy()[perm[rows]]
    err = oracle.rel_l2(got, ref)
    print(f"nbody_k3 {name} rel L2 = {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("variable", [False, True])
def test_mixture_k64_fp64_sampled_rows(G, oracle, torch_cuda, variable):
    from ks_amd import workloads as W
    torch = torch_cuda
    d = W.mixture_k64()
    X, X2, w, idx = d["X"], d["X2"], d["w"], d["amap"]
    N = X.shape[0]
    dX, dX2, dw, didx, dh = _dev(torch, X, X2, w, idx, d["hinv"])
    du = torch.zeros(N, dtype=torch.float64, device="cuda")
    if variable:
        G.dgsks(64, du, didx, dX, dX2, didx, dX, dX2, didx, dw, didx, type="Variable_Bandwidth",
                h=dh)
    else:
        G.dgsks(64, du, didx, dX, dX2, didx, dX, dX2, didx, dw, didx, type="Gaussian",
                scal=d["kernel"]["scal"])
    torch.cuda.synchronize()
    rows = np.random.default_rng(13).choice(N, 128, replace=False).astype(np.int32)
    t = 3 if variable else 0
    ref = oracle.dgsks_rows(t, X, X2, idx, X, X2, idx, w, idx, np.zeros(N), idx, rows=rows,
                            scal=d["kernel"]["scal"], hi=d["hinv"] if variable else None,
                            hj=d["hinv"] if variable else None)
    err = oracle.rel_l2(du.cpu().numpy()[rows], ref)
    print(f"mixture_k64 variable={variable} rel L2 = {err:.3e}")
    assert err <= TOL


# ------------------------------------------------------------------ edge cases

def test_edges(G, oracle, torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(5)
    X = rng.random((50, 4))
    X2 = oracle.sqnorm(X)
    w = rng.standard_normal(50)
    idx = np.arange(50, dtype=np.int32)
    # m = 1, n = 1
    u = np.zeros(50)
    G.dgsks(4, u, idx[:1], X, X2, idx[3:4], X, X2, idx[7:8], w, idx[9:10], type="Gaussian",
            scal=-0.5)
    assert u[0] == pytest.approx(np.exp(-0.5 * np.sum((X[3] - X[7]) ** 2)) * w[9], rel=1e-14)
    assert np.all(u[1:] == 0)
    # empty call is a no-op
    u2 = u.copy()
    G.dgsks(4, u2, idx[:0], X, X2, idx[:0], X, X2, idx[:5], w, idx[:5], type="Gaussian", scal=-0.5)
    assert np.array_equal(u, u2)
    # accumulation: two calls add twice
    dX, dX2, dw, didx = _dev(torch, X, X2, w, idx)
    du = torch.zeros(50, dtype=torch.float64, device="cuda")
    for _ in range(2):
        G.dgsks(4, du, didx, dX, dX2, didx, dX, dX2, didx, dw, didx, type="Gaussian", scal=-0.5)
    torch.cuda.synchronize()
    one = oracle.dgsks_rows(0, X, X2, idx, X, X2, idx, w, idx, np.zeros(50), idx, scal=-0.5)
    assert oracle.rel_l2(du.cpu().numpy(), 2 * one) <= TOL


def test_deterministic(G, torch_cuda):
    """Same inputs, same bits: the split-source reduction has a fixed order."""
    from ks_amd import workloads as W
    torch = torch_cuda
    d = W.uniform_gauss16(m=4096, n=8192)
    X, X2, w = _dev(torch, d["X"], d["X2"], d["w"])
    ia, ib = _dev(torch, d["amap"], d["bmap"])
    outs = []
    for _ in range(3):
        du = torch.zeros(X.shape[0], dtype=torch.float64, device="cuda")
        G.dgsks(16, du, ia, X, X2, ia, X, X2, ib, w, ib, type="Gaussian", scal=-2.0)
        outs.append(du.cpu().numpy())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])


def test_error_paths_leave_u_untouched(G, torch_cuda):
    from ks_amd import _lib
    L = _lib.lib()
    L.ks_set_error_mode(_lib.KS_ERRMODE_RETURN)
    X = np.random.default_rng(1).random((10, 3))
    X2 = (X * X).sum(1)
    idx = np.arange(10, dtype=np.int32)
    u = np.ones(10)
    w = np.ones(10)
    kt = _lib.KsT()
    kt.type = 11  # illegal
    p = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    L.dgsks(ctypes.byref(kt), 10, 10, 3, p(u), p(idx), p(X), p(X2), p(idx), p(X), p(X2), p(idx),
            p(w), p(idx))
    assert L.ks_last_error() == _lib.KS_ERR_INVALID_TYPE and np.all(u == 1)
    kt.type = 3  # variable bandwidth without hi / hj
    L.dgsks(ctypes.byref(kt), 10, 10, 3, p(u), p(idx), p(X), p(X2), p(idx), p(X), p(X2), p(idx),
            p(w), p(idx))
    assert L.ks_last_error() == _lib.KS_ERR_NO_BANDWIDTH and np.all(u == 1)


@pytest.mark.parametrize("k", [12, 32, 64])
def test_laplace_matrix_cores_near_coincident_pairs(G, oracle, torch_cuda, k):
    """Laplace at 9 <= k <= 64 runs on the FP64 matrix cores.  Exact duplicates
    and near-coincident pairs whose d^2 straddles the reference's 1e-15 drop
    threshold (frame/dgsks_ref.c:343-345) must be recomputed with the
    reference's sequential sum, so every drop decision and every kernel value
    of a kept near pair (up to ~1e100 and beyond) is the reference's.  Compared
    after scaling by the largest entry, whose square would overflow."""
    torch = torch_cuda
    rng = np.random.default_rng(3000 + k)
    nx, m, n = 2400, 900, 1500
    X = rng.random((nx, k)) * (1.5 / np.sqrt(k))
    deltas = np.concatenate([[0.0] * 20, 10.0 ** rng.uniform(-12, -7.2, 180)])
    for i, dl in enumerate(deltas):
        d = rng.standard_normal(k)
        X[1200 + i] = X[i] + dl * d / np.linalg.norm(d)
    X2 = oracle.sqnorm(X)
    near = np.arange(200)
    amap = np.concatenate([near, rng.choice(np.arange(200, nx), m - 200, replace=False)])
    bmap = np.concatenate([near + 1200, near, rng.choice(np.arange(200, nx), n - 400,
                                                        replace=False)])
    amap, bmap = amap.astype(np.int32), bmap.astype(np.int32)
    umap = rng.permutation(nx)[:m].astype(np.int32)
    wmap = rng.permutation(nx)[:n].astype(np.int32)
    w = rng.standard_normal(nx)
    u0 = rng.standard_normal(nx)
    dX, dX2, damap, dumap, dbmap, dwmap, dw, du = _dev(torch, X, X2, amap, umap, bmap, wmap, w, u0)
    G.dgsks(k, du, dumap, dX, dX2, damap, dX, dX2, dbmap, dw, dwmap, **_opts(2))
    torch.cuda.synchronize()
    got = du.cpu().numpy()[umap]
    ref = oracle.dgsks_rows(2, X, X2, amap, X, X2, bmap, w, wmap, u0, umap)
    # at k = 64 a kept pair's d^-62 can overflow: those rows must overflow alike
    fin = np.isfinite(ref)
    assert np.array_equal(got[~fin], ref[~fin], equal_nan=True)
    sc = float(np.max(np.abs(ref[fin])))
    err = oracle.rel_l2(got[fin] / sc, ref[fin] / sc)
    big = int(np.sum(np.abs(ref[fin]) > 1e50))
    print(f"Laplace k={k} near-coincident pairs: rel L2 {err:.3e}, {big} rows with a kept near pair")
    assert err <= TOL


def test_laplace_descriptor_side_effect(G):
    from ks_amd import _lib
    L = _lib.lib()
    X = np.random.default_rng(2).random((20, 5))
    X2 = (X * X).sum(1)
    idx = np.arange(20, dtype=np.int32)
    u, w = np.zeros(20), np.ones(20)
    kt = _lib.KsT()
    kt.type = 2
    p = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    L.dgsks(ctypes.byref(kt), 20, 20, 5, p(u), p(idx), p(X), p(X2), p(idx), p(X), p(X2), p(idx),
            p(w), p(idx))
    assert kt.powe == -1.5
    import math
    assert kt.scal == pytest.approx(math.gamma(3.5) / (5 * 3 * math.pi ** 2.5), rel=1e-15)
    assert np.all(np.isfinite(u)) and np.all(u > 0)


def test_sqnorm_device(G, oracle, torch_cuda):
    torch = torch_cuda
    X = np.random.default_rng(3).random((1000, 37))
    got = G.sqnorm(torch.from_numpy(X).cuda()).cpu().numpy()
    assert np.array_equal(got, oracle.sqnorm(X))


# ------------------------------------------------------------------ batched list

def test_list_golden(G, golden, torch_cuda):
    from ks_amd import _lib
    L = _lib.lib()
    g = golden
    ntask = len(g["L_aoff"]) - 1
    kernels = (_lib.KsT * ntask)()
    for t in range(ntask):
        kernels[t].type = 0
        kernels[t].scal = -0.5
    u = np.zeros_like(g["L_ref_u"])
    p = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    X, X2, w = g["L_X"], g["L_X2"], g["L_w"]
    rc = L.dgsks_list(ntask, kernels, X.shape[1], p(u), u.size, p(g["L_aoff"]), p(g["L_amap"]),
                      p(g["L_amap"]), p(g["L_boff"]), p(g["L_bmap"]), p(g["L_bmap"]), p(X), p(X2),
                      p(X), p(X2), p(w), 0)
    assert rc == 0
    ref = g["L_ref_u"]
    assert np.linalg.norm(u - ref) / np.linalg.norm(ref) <= TOL


def test_list_mixed_kernels(G, oracle, torch_cuda):
    from ks_amd import _lib
    from ks_amd import workloads as W
    L = _lib.lib()
    d = W.list_pool(k=8, ntask=64, pool=20000, lo=8, hi=600, seed=9)
    ntask = len(d["kernels"])
    kernels = (_lib.KsT * ntask)()
    for t, kd in enumerate(d["kernels"]):
        kernels[t].type = kd["type"]
        kernels[t].scal = kd.get("scal", 0.0) or 0.0
        kernels[t].cons = kd.get("cons", 0.0)
        kernels[t].powe = kd.get("powe", 0.0)
    X, X2, w = d["X"], d["X2"], d["w"]
    u = np.zeros(X.shape[0])
    p = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    rc = L.dgsks_list(ntask, kernels, 8, p(u), u.size, p(d["a_off"]), p(d["amap"]), p(d["amap"]),
                      p(d["b_off"]), p(d["bmap"]), p(d["bmap"]), p(X), p(X2), p(X), p(X2), p(w), 0)
    assert rc == 0
    ref = np.zeros_like(u)
    oracle.dgsks_list(d["kernels"], 8, ref, d["a_off"], d["amap"], d["amap"], d["b_off"],
                      d["bmap"], d["bmap"], X, X2, X, X2, w)
    assert oracle.rel_l2(u, ref) <= TOL


@pytest.mark.parametrize("k", [3, 8, 32, 20])
def test_list_persistent_shared_targets(G, oracle, torch_cuda, k):
    """Persistent list launch: mixed kernel types (Gaussian, Laplace, polynomial),
    tasks sharing target points, some tasks with more sources than one tile
    slice, empty tasks; every k path (scalar-load, matrix-core, generic)."""
    from ks_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(300 + k)
    pool, ntask = 6000, 40
    X = rng.random((pool, k)) * (1.5 / np.sqrt(k))
    X2 = oracle.sqnorm(X)
    w = rng.standard_normal(pool)
    ms = rng.integers(1, 300, ntask)
    ns = rng.integers(1, 2600, ntask)
    ms[3] = 0
    ns[5] = 0
    amaps = [rng.choice(pool, int(s), replace=False).astype(np.int32) for s in ms]
    bmaps = [rng.choice(pool, int(s), replace=False).astype(np.int32) for s in ns]
    kinds = [dict(type=0, scal=-1.5), dict(type=2), dict(type=1, scal=1.0 / k, cons=1.0, powe=2.0)]
    kernels_py = [kinds[t % 3] for t in range(ntask)]
    a_off = np.concatenate([[0], np.cumsum(ms)]).astype(np.int64)
    b_off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    amap = np.concatenate(amaps)
    bmap = np.concatenate(bmaps)
    kernels = (_lib.KsT * ntask)()
    for t, kd in enumerate(kernels_py):
        kernels[t].type = kd["type"]
        kernels[t].scal = kd.get("scal", 0.0)
        kernels[t].cons = kd.get("cons", 0.0)
        kernels[t].powe = kd.get("powe", 0.0)
    u = rng.standard_normal(pool)
    ref = u.copy()
    p = lambda a: ctypes.c_void_p(a.ctypes.data)  # noqa: E731
    rc = L.dgsks_list(ntask, kernels, k, p(u), u.size, p(a_off), p(amap), p(amap), p(b_off),
                      p(bmap), p(bmap), p(X), p(X2), p(X), p(X2), p(w), 0)
    assert rc == 0
    oracle.dgsks_list(kernels_py, k, ref, a_off, amap, amap, b_off, bmap, bmap, X, X2, X, X2, w)
    err = oracle.rel_l2(u, ref)
    print(f"list k={k}: rel L2 = {err:.3e}")
    assert err <= TOL
    if k >= 3:  # Laplace descriptor side effect per task
        assert kernels[1].powe == 0.5 * (2 - k)
