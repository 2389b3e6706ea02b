"""Pins the CPU oracle (oracle/gsks_oracle.c, a restatement of dgsks_ref,
frame/dgsks_ref.c:94-535) before anything is checked against it:
  * the golden fixtures written from the compiled reference (tests/golden/);
  * the compiled reference itself where oracle/_ref is built;
  * an independent numpy restatement;
  * mpmath brute force at 50 digits on tiny inputs;
  * closed forms and invariants chosen so that a dropped term, a wrong sign or
    index or a transposed operand fails at least one of them.
CPU only."""
import math

import numpy as np
import pytest

import np_restate as NR
from conftest import golden_cases

TYPES = range(8)


def _params(t, k):
    return {0: dict(scal=-0.8), 1: dict(scal=0.4, cons=0.3, powe=3.0), 2: dict(),
            3: dict(), 4: dict(scal=0.35, cons=-0.2), 5: dict(), 6: dict(cons=0.7),
            7: dict()}[t]


def _problem(rng, m, n, k, nx=None, scale=0.5):
    nx = nx or max(m, n) + 5
    X = rng.random((nx, k)) * scale
    amap = rng.permutation(nx)[:m].astype(np.int32)
    umap = rng.permutation(nx)[:m].astype(np.int32)
    bmap = rng.permutation(nx)[:n].astype(np.int32)
    wmap = rng.permutation(nx)[:n].astype(np.int32)
    w = rng.standard_normal(nx)
    u = rng.standard_normal(nx)
    h = rng.uniform(-2.0, -0.5, nx)
    return X, amap, umap, bmap, wmap, w, u, h


# ------------------------------------------------------------------ golden fixtures

def test_golden_fixture_has_every_type(golden):
    types = {int(d["meta"][0]) for _, d in golden_cases(golden)}
    assert types == set(range(8))


def test_oracle_matches_golden_reference_outputs(oracle, golden):
    """oracle dgsks_ref vs the compiled reference's dgsks_ref outputs."""
    for key, d in golden_cases(golden):
        t, m, n, k = (int(x) for x in d["meta"])
        scal, cons, powe, ref_powe, ref_scal = d["par"]
        h = d.get("h")
        u = d["u0"].copy()
        p_after, s_after = oracle.dgsks_ref(t, u, d["umap"], d["X"], d["X2"], d["amap"], d["X"],
                                            d["X2"], d["bmap"], d["w"], d["wmap"], scal=scal,
                                            cons=cons, powe=powe, hi=h, hj=h)
        ref = d["ref_u"]
        rel = np.abs(u - ref) / np.maximum(np.abs(ref), 1e-300)
        assert rel.max() <= 1e-14, (key, t, rel.max())
        if t == oracle.LAPLACE:  # the descriptor side effect (frame/dgsks_ref.c:146-148)
            assert p_after == ref_powe and s_after == pytest.approx(ref_scal, rel=1e-15)
        # the reference's fast path agrees with its oracle at its own 1e-13 bar
        # (test/test_dgsks.c:41), which sanity-checks the fixture itself
        assert oracle.rel_l2(d["fast_u"], ref) <= 1e-13, key


def test_oracle_list_matches_golden(oracle, golden):
    g = golden
    ntask = len(g["L_aoff"]) - 1
    kernels = [dict(type=oracle.GAUSSIAN, scal=-0.5)] * ntask
    u = np.zeros_like(g["L_ref_u"])
    oracle.dgsks_list(kernels, g["L_X"].shape[1], u, g["L_aoff"], g["L_amap"], g["L_amap"],
                      g["L_boff"], g["L_bmap"], g["L_bmap"], g["L_X"], g["L_X2"], g["L_X"],
                      g["L_X2"], g["L_w"])
    ref = g["L_ref_u"]
    nz = ref != 0
    assert np.max(np.abs(u[nz] - ref[nz]) / np.abs(ref[nz])) <= 1e-10  # test_dgsks_list.cpp:241
    assert oracle.rel_l2(u, ref) <= 1e-14


@pytest.mark.parametrize("shape", [(45, 33, 3), (17, 80, 16), (64, 64, 64), (9, 5, 1)])
def test_oracle_bitexact_vs_compiled_reference(oracle, shape):
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(sum(shape))
    m, n, k = shape
    X, amap, umap, bmap, wmap, w, u0, h = _problem(rng, m, n, k)
    X2 = oracle.sqnorm(X)
    for t in TYPES:
        if t == oracle.LAPLACE and k < 3:
            continue
        prm = _params(t, k)
        hh = h if t == oracle.VAR_BANDWIDTH else None
        a = u0.copy()
        oracle.dgsks_ref(t, a, umap, X, X2, amap, X, X2, bmap, w, wmap, hi=hh, hj=hh, **prm)
        b = u0.copy()
        oracle.ref_call("dgsks_ref", t, b, umap, X, X2, amap, X, X2, bmap, w, wmap, hi=hh, hj=hh,
                        **prm)
        assert np.array_equal(a, b), (t, shape)


@pytest.mark.parametrize("t", [0, 3, 5, 6, 7])
def test_clamp_mode_vs_compiled_reference_dgsks(oracle, t):
    """Norms lowered by 1e-7 (relative) make d^2 of every coincident pair
    negative.  The fused dgsks clamps d^2 at 0 before these kernels
    (micro_kernel/x86_64/sandybridge/ks_gaussian_int_d8x4.c:169 and its
    siblings), dgsks_ref does not: the oracle's clamp mode must follow the
    compiled dgsks, its default mode must visibly differ from it."""
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(40 + t)
    m, k = 40, 8
    X, amap, umap, _, wmap, w, u0, h = _problem(rng, m, m, k)
    bmap = amap.copy()  # every target meets itself
    X2 = oracle.sqnorm(X) * (1.0 - 1e-7)
    prm = _params(t, k)
    hh = h if t == oracle.VAR_BANDWIDTH else None
    ref = u0.copy()
    oracle.ref_call("dgsks", t, ref, umap, X, X2, amap, X, X2, bmap, w, wmap, hi=hh, hj=hh, **prm)
    rows = {c: oracle.dgsks_rows(t, X, X2, amap, X, X2, bmap, w, wmap, u0, umap, hi=hh, hj=hh,
                                 clamp=c, **prm) for c in (True, False)}
    assert oracle.rel_l2(rows[True], ref[umap]) <= 1e-13
    assert oracle.rel_l2(rows[False], ref[umap]) > 1e-9


def test_clamp_mode_closed_forms(oracle):
    """K at a negative d^2 under the clamp is K at d^2 = 0, for every type it
    applies to; without it the unclamped formula is evaluated; Laplace is the
    same either way (its d^2 < 1e-15 branch)."""
    aa = bb = 0.5
    c = 0.5 + 1e-6  # d^2 = aa + bb - 2c = -2e-6
    P = oracle.pair
    assert P(0, c, aa, bb, scal=-3.0, clamp=True) == 1.0
    assert P(0, c, aa, bb, scal=-3.0) == math.exp(-3.0 * (aa + bb - 2 * c))
    assert P(3, c, aa, bb, 2.0, 1.5, clamp=True) == 1.0
    assert P(5, c, aa, bb, clamp=True) == 15.0 / 16.0
    assert P(5, c, aa, bb) > 15.0 / 16.0
    assert P(6, c, aa, bb, cons=0.25, clamp=True) == 0.25
    assert P(6, c, aa, bb, cons=0.25) < 0.25
    assert P(7, c, aa, bb, clamp=True) == 0.75
    assert P(7, c, aa, bb) > 0.75
    powe, scal = oracle.laplace_params(3)
    assert P(2, c, aa, bb, powe=powe, scal=scal, clamp=True) == P(2, c, aa, bb, powe=powe,
                                                                 scal=scal)
    # non-negative d^2: the clamp changes nothing
    for t in (0, 5, 6, 7):
        assert P(t, 0.1, aa, bb, scal=-3.0, cons=0.25, clamp=True) == \
            P(t, 0.1, aa, bb, scal=-3.0, cons=0.25)


# ------------------------------------------------------------------ independent restatements

@pytest.mark.parametrize("t", TYPES)
def test_oracle_vs_numpy_restatement(oracle, t):
    rng = np.random.default_rng(100 + t)
    m, n, k = 57, 43, 6
    X, amap, umap, bmap, wmap, w, u0, h = _problem(rng, m, n, k)
    Y = rng.random((X.shape[0], k)) * 0.5  # distinct source table
    X2, Y2 = oracle.sqnorm(X), oracle.sqnorm(Y)
    prm = _params(t, k)
    hh = h if t == oracle.VAR_BANDWIDTH else None
    a = u0.copy()
    oracle.dgsks_ref(t, a, umap, X, X2, amap, Y, Y2, bmap, w, wmap, hi=hh, hj=hh, **prm)
    b = NR.dgsks(t, u0, umap, X, X2, amap, Y, Y2, bmap, w, wmap, hi=hh, hj=hh, **prm)
    assert oracle.rel_l2(a, b) <= 1e-13


def _mp_kernel(mp, t, a, b, k, prm, hi=None, hj=None):
    c = mp.fsum(mp.mpf(x) * mp.mpf(y) for x, y in zip(a, b))
    aa = mp.fsum(mp.mpf(x) ** 2 for x in a)
    bb = mp.fsum(mp.mpf(y) ** 2 for y in b)
    d2 = aa + bb - 2 * c
    if t == 0:
        return mp.exp(mp.mpf(prm["scal"]) * d2)
    if t == 1:
        return (mp.mpf(prm["scal"]) * c + mp.mpf(prm["cons"])) ** int(prm["powe"])
    if t == 2:
        s = mp.gamma(mp.mpf(k) / 2 + 1) / (k * (k - 2) * mp.pi ** (mp.mpf(k) / 2))
        return s * d2 ** (mp.mpf(2 - k) / 2)
    if t == 3:
        return mp.exp(-d2 * mp.mpf(hi) * mp.mpf(hj) / 2)
    if t == 4:
        return mp.tanh(mp.mpf(prm["scal"]) * c + mp.mpf(prm["cons"]))
    if t == 5:
        return mp.mpf(15) / 16 * (1 - d2) ** 2 if d2 < 1 else mp.mpf(0)
    if t == 6:
        return d2 + mp.mpf(prm["cons"])
    return mp.mpf(3) / 4 * (1 - d2) if d2 < 1 else mp.mpf(0)


@pytest.mark.parametrize("t", TYPES)
def test_oracle_vs_mpmath_bruteforce(oracle, t):
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 50
    rng = np.random.default_rng(7 + t)
    m, n, k = 7, 11, 4
    X = rng.random((16, k)) * 0.5
    X2 = oracle.sqnorm(X)
    amap = rng.permutation(16)[:m].astype(np.int32)
    bmap = rng.permutation(16)[:n].astype(np.int32)
    umap = amap.copy()
    wmap = bmap.copy()
    w = rng.standard_normal(16)
    h = rng.uniform(-2.0, -0.5, 16)
    prm = _params(t, k)
    hh = h if t == 3 else None
    u = np.zeros(16)
    oracle.dgsks_ref(t, u, umap, X, X2, amap, X, X2, bmap, w, wmap, hi=hh, hj=hh, **prm)
    exact = []
    for i in range(m):
        s = mp.fsum(_mp_kernel(mp, t, X[amap[i]], X[bmap[j]], k, prm,
                               None if hh is None else h[amap[i]],
                               None if hh is None else h[bmap[j]]) * mp.mpf(w[wmap[j]])
                    for j in range(n) if not (t == 2 and amap[i] == bmap[j]))
        exact.append(float(s))
    got = u[umap[:m]]
    assert oracle.rel_l2(got, np.array(exact)) <= 1e-13, t


# ------------------------------------------------------------------ closed forms

def test_closed_form_multiquadric(oracle):
    """u_i = (|a_i|^2 + c) sum w + sum_j w_j |b_j|^2 - 2 a_i . sum_j w_j b_j."""
    rng = np.random.default_rng(3)
    m, n, k = 300, 500, 7
    A, B = rng.random((m, k)), rng.random((n, k))
    w = rng.standard_normal(n)
    A2, B2 = oracle.sqnorm(A), oracle.sqnorm(B)
    ia, ib = np.arange(m, dtype=np.int32), np.arange(n, dtype=np.int32)
    got = oracle.dgsks_rows(oracle.MULTIQUADRATIC, A, A2, ia, B, B2, ib, w, ib, np.zeros(m), ia,
                            cons=0.25)
    closed = (A2 + 0.25) * w.sum() + (w * B2).sum() - 2.0 * A @ (B.T @ w)
    assert oracle.rel_l2(got, closed) <= 1e-12


def test_closed_form_polynomial_square(oracle):
    """(s a.b + c)^2 summed with weights = s^2 a^T (sum w b b^T) a + 2 s c a.(sum w b) + c^2 sum w."""
    rng = np.random.default_rng(4)
    m, n, k = 200, 300, 5
    A, B = rng.random((m, k)), rng.random((n, k))
    w = rng.standard_normal(n)
    s, c = 0.7, -0.3
    ia, ib = np.arange(m, dtype=np.int32), np.arange(n, dtype=np.int32)
    for powe in (2.0, 3.0):
        got = oracle.dgsks_rows(oracle.POLYNOMIAL, A, oracle.sqnorm(A), ia, B, oracle.sqnorm(B),
                                ib, w, ib, np.zeros(m), ia, scal=s, cons=c, powe=powe)
        if powe == 2.0:
            M2 = (B * w[:, None]).T @ B
            closed = s * s * np.einsum("ip,pq,iq->i", A, M2, A) + 2 * s * c * A @ (B.T @ w) \
                + c * c * w.sum()
        else:  # third moment expansion
            M1 = B.T @ w
            M2 = (B * w[:, None]).T @ B
            M3 = np.einsum("j,jp,jq,jr->pqr", w, B, B, B)
            closed = (s ** 3 * np.einsum("ip,iq,ir,pqr->i", A, A, A, M3)
                      + 3 * s * s * c * np.einsum("ip,pq,iq->i", A, M2, A)
                      + 3 * s * c * c * A @ M1 + c ** 3 * w.sum())
        assert oracle.rel_l2(got, closed) <= 1e-12, powe


def test_gaussian_scal_zero_is_weight_sum(oracle):
    rng = np.random.default_rng(5)
    X = rng.random((100, 9))
    w = rng.standard_normal(100)
    idx = np.arange(100, dtype=np.int32)
    got = oracle.dgsks_rows(oracle.GAUSSIAN, X, oracle.sqnorm(X), idx, X, oracle.sqnorm(X), idx, w,
                            idx, np.zeros(100), idx, scal=0.0)
    assert np.allclose(got, w.sum(), rtol=1e-13, atol=0)


def test_laplace_parameters(oracle):
    p, s = oracle.laplace_params(3)
    assert p == -0.5 and s == pytest.approx(1.0 / (4.0 * math.pi), rel=1e-15)
    p, s = oracle.laplace_params(4)  # Gamma(3) / (4*2*pi^2) = 1 / (4 pi^2)
    assert p == -1.0 and s == pytest.approx(1.0 / (4.0 * math.pi ** 2), rel=1e-15)


def test_laplace_coincident_points_vanish(oracle):
    """d^2 < 1e-15 is replaced by 1.79e308 (frame/dgsks_ref.c:343-345)."""
    X = np.array([[0.1, 0.2, 0.3], [0.1, 0.2, 0.3], [0.5, 0.5, 0.5]])
    idx = np.arange(3, dtype=np.int32)
    w = np.array([1.0, 1.0, 0.0])
    got = oracle.dgsks_rows(oracle.LAPLACE, X, oracle.sqnorm(X), idx[:1], X, oracle.sqnorm(X),
                            idx, w, idx, np.zeros(3), idx[:1])
    assert 0.0 < got[0] < 1e-150


# ------------------------------------------------------------------ invariants

@pytest.mark.parametrize("t", TYPES)
def test_invariants(oracle, t):
    rng = np.random.default_rng(50 + t)
    m, n, k = 40, 60, 4
    X, amap, umap, bmap, wmap, w, u0, h = _problem(rng, m, n, k, nx=80)
    X2 = oracle.sqnorm(X)
    prm = _params(t, k)
    hh = h if t == 3 else None
    zero = np.zeros_like(u0)
    f = lambda ww, am=amap, um=umap, bm=bmap, wm=wmap: oracle.dgsks_rows(  # noqa: E731
        t, X, X2, am, X, X2, bm, ww, wm, zero, um, hi=hh, hj=hh, **prm)
    base = f(w)
    # linearity in w
    w2 = rng.standard_normal(w.size)
    assert oracle.rel_l2(f(2.0 * w - 3.0 * w2), 2.0 * base - 3.0 * f(w2)) <= 1e-12
    # permutation covariance of the target maps
    perm = rng.permutation(m)
    assert np.allclose(f(w, amap[perm], umap[perm]), base[perm], rtol=1e-15, atol=0)
    # order of the source points only changes rounding
    pb = rng.permutation(n)
    assert oracle.rel_l2(f(w, bm=bmap[pb], wm=wmap[pb]), base) <= 1e-13
    # accumulation into a nonzero u
    got = oracle.dgsks_rows(t, X, X2, amap, X, X2, bmap, w, wmap, u0, umap, hi=hh, hj=hh, **prm)
    assert np.allclose(got, u0[umap] + base, rtol=1e-13, atol=1e-13)


def test_symmetry_and_translation(oracle):
    rng = np.random.default_rng(9)
    X = rng.random((30, 5))
    for t in (0, 5, 6, 7):
        prm = _params(t, 5)
        K = np.array([[oracle.pair(t, X[i] @ X[j], X[i] @ X[i], X[j] @ X[j], **prm)
                       for j in range(30)] for i in range(30)])
        assert np.allclose(K, K.T, rtol=1e-13, atol=1e-15), t
        Y = X + 0.37
        KY = NR.kernel_matrix(t, Y, Y, (Y * Y).sum(1), (Y * Y).sum(1), **prm)
        assert np.allclose(K, KY, rtol=1e-11, atol=1e-12), t


def test_compact_support(oracle):
    X = np.array([[0.0, 0.0], [0.6, 0.0], [1.5, 0.0]])
    idx = np.arange(3, dtype=np.int32)
    for t, inside in ((5, 15 / 16 * (1 - 0.36) ** 2), (7, 0.75 * (1 - 0.36))):
        w = np.array([0.0, 1.0, 1.0])
        got = oracle.dgsks_rows(t, X, oracle.sqnorm(X), idx[:1], X, oracle.sqnorm(X), idx, w, idx,
                                np.zeros(3), idx[:1])
        assert got[0] == pytest.approx(inside, rel=1e-14)


# ------------------------------------------------------------------ gradient oracle

def test_gauss_grad_vs_central_differences(oracle):
    rng = np.random.default_rng(11)
    m, n, k, h = 20, 25, 3, 0.4
    A, B = rng.random((m, k)), rng.random((n, k))
    w, g = rng.standard_normal(n), rng.standard_normal(m)
    ia, ib = np.arange(m, dtype=np.int32), np.arange(n, dtype=np.int32)

    def loss(A, B, w, h):
        u = oracle.dgsks_rows(oracle.GAUSSIAN, A, oracle.sqnorm(A), ia, B, oracle.sqnorm(B), ib, w,
                              ib, np.zeros(m), ia, scal=-1.0 / (2 * h * h))
        return float(g @ u)

    dA, dB, dw, dh = oracle.gauss_grad(A, B, w, g, h)
    eps = 1e-6
    for (i, p) in [(0, 0), (5, 2), (19, 1)]:
        Ap, Am = A.copy(), A.copy()
        Ap[i, p] += eps
        Am[i, p] -= eps
        assert dA[i, p] == pytest.approx((loss(Ap, B, w, h) - loss(Am, B, w, h)) / (2 * eps),
                                         rel=1e-6, abs=1e-8)
        Bp, Bm = B.copy(), B.copy()
        Bp[i, p] += eps
        Bm[i, p] -= eps
        assert dB[i, p] == pytest.approx((loss(A, Bp, w, h) - loss(A, Bm, w, h)) / (2 * eps),
                                         rel=1e-6, abs=1e-8)
    wp, wm = w.copy(), w.copy()
    wp[3] += eps
    wm[3] -= eps
    assert dw[3] == pytest.approx((loss(A, B, wp, h) - loss(A, B, wm, h)) / (2 * eps), rel=1e-7)
    assert dh == pytest.approx((loss(A, B, w, h + eps) - loss(A, B, w, h - eps)) / (2 * eps),
                               rel=1e-6)
    # subsets index the same values
    dA2, dB2, dw2, _ = oracle.gauss_grad(A, B, w, g, h, rows_a=[3, 7], rows_b=[1, 4],
                                         want_dh=False)
    assert np.array_equal(dA2, dA[[3, 7]]) and np.array_equal(dB2, dB[[1, 4]])
    assert np.array_equal(dw2, dw[[1, 4]])


@pytest.mark.parametrize("ktype", ["gaussian", "var_bandwidth"])
def test_mapped_grad_vs_central_differences(oracle, ktype):
    """orc_ks_grad (index maps with repeated entries, one point table and one
    bandwidth vector on both sides, as the reference's symmetric calls use them)
    against central differences of the oracle's own mapped forward.  A point
    that both maps select receives dXA + dXB, a bandwidth dhi + dhj; a repeated
    wmap entry sums its sources' terms."""
    rng = np.random.default_rng(7)
    var = ktype == "var_bandwidth"
    T = oracle.VAR_BANDWIDTH if var else oracle.GAUSSIAN
    nx, k, m, n = 30, 4, 25, 28
    X = rng.random((nx, k))
    hv = 1.0 + rng.random(nx) * 4.0
    w = rng.standard_normal(nx)
    g = rng.standard_normal(nx)
    amap = rng.integers(0, nx, m).astype(np.int32)  # repeats on purpose
    bmap = rng.integers(0, nx, n).astype(np.int32)
    wmap = rng.integers(0, nx, n).astype(np.int32)
    umap = rng.integers(0, nx, m).astype(np.int32)
    scal0 = -1.7

    def loss(X, w, hv, scal):
        X2 = oracle.sqnorm(X)
        out = oracle.dgsks_rows(T, X, X2, amap, X, X2, bmap, w, wmap, np.zeros(nx), umap,
                                scal=scal, hi=hv if var else None, hj=hv if var else None)
        return float(g[umap] @ out)

    G = oracle.ks_grad(T, g, umap, X, amap, X, bmap, w, wmap, scal=scal0,
                       hi=hv if var else None, hj=hv if var else None)
    dX = G["dXA"] + G["dXB"]
    eps = 1e-6
    for x in sorted(set(amap[:4]) | set(bmap[:4])):
        for p in (0, k - 1):
            Xp, Xm = X.copy(), X.copy()
            Xp[x, p] += eps
            Xm[x, p] -= eps
            fd = (loss(Xp, w, hv, scal0) - loss(Xm, w, hv, scal0)) / (2 * eps)
            assert dX[x, p] == pytest.approx(fd, rel=1e-6, abs=1e-8)
    unused = sorted(set(range(nx)) - set(amap) - set(bmap))
    assert not np.any(dX[unused])
    for x in sorted(set(wmap[:5])):
        wp, wm = w.copy(), w.copy()
        wp[x] += eps
        wm[x] -= eps
        fd = (loss(X, wp, hv, scal0) - loss(X, wm, hv, scal0)) / (2 * eps)
        assert G["dw"][x] == pytest.approx(fd, rel=1e-7, abs=1e-9)
    if var:
        dh = G["dhi"] + G["dhj"]
        for x in sorted(set(amap[:3]) | set(bmap[:3])):
            hp, hm = hv.copy(), hv.copy()
            hp[x] += eps
            hm[x] -= eps
            fd = (loss(X, w, hp, scal0) - loss(X, w, hm, scal0)) / (2 * eps)
            assert dh[x] == pytest.approx(fd, rel=1e-6, abs=1e-8)
    else:
        fd = (loss(X, w, hv, scal0 + eps) - loss(X, w, hv, scal0 - eps)) / (2 * eps)
        assert G["dscal"] == pytest.approx(fd, rel=1e-6)


def test_mapped_grad_reduces_to_identity_gradient(oracle):
    """With identity maps and scal = -1/(2h^2) the mapped gradient is the
    fixed-bandwidth one: dXA, dXB, dw agree with orc_gauss_grad, and dh =
    dscal * dscal/dh = dscal / h^3."""
    rng = np.random.default_rng(12)
    m, n, k, h = 17, 23, 5, 0.6
    A, B = rng.random((m, k)), rng.random((n, k))
    w, g = rng.standard_normal(n), rng.standard_normal(m)
    ia, ib = np.arange(m, dtype=np.int32), np.arange(n, dtype=np.int32)
    G = oracle.ks_grad(oracle.GAUSSIAN, g, ia, A, ia, B, ib, w, ib, scal=-1.0 / (2 * h * h))
    dA, dB, dw, dh = oracle.gauss_grad(A, B, w, g, h)
    assert np.allclose(G["dXA"], dA, rtol=1e-12, atol=1e-14)
    assert np.allclose(G["dXB"], dB, rtol=1e-12, atol=1e-14)
    assert np.allclose(G["dw"], dw, rtol=1e-13, atol=1e-15)
    assert G["dscal"] / h ** 3 == pytest.approx(dh, rel=1e-12)

