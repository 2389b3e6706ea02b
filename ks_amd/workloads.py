"""Synthetic inputs of the benchmark shapes (seeded numpy generators).

    uniform_gauss16   m = n = 8192, k = 16, uniform [0,1), Gaussian h = 0.5
    nbody_k3          m = n = 262144, k = 3, uniform cube, random permutation maps,
                      Laplace / Gaussian h = 0.05 / polynomial (a.b/3 + 1)^3 /
                      tanh(a.b/3 + 0.1)
    mixture_k64       m = n = 131072, k = 64, 32-component Gaussian mixture
                      (sigma 0.05), Gaussian h = 0.3 and variable bandwidth
                      h_i ~ U[0.2, 0.4)
    list_pool         4096 tasks over a pool of 1,048,576 points, k in {3, 8, 32},
                      m, n log-uniform in [64, 8192], kernel in
                      {Gaussian h = 0.2, Laplace, polynomial degree 2}; with
                      symmetric=True every task's sources are its targets
    grad_gauss16      m = n = 32768, k = 16, Gaussian h = 0.4, upstream gradient

Squared norms are summed left to right in FP64 like the reference's norm
precompute (frame/omp_dgsks_list.cpp:75-81).  Variable bandwidth follows the
reference's product form exp(-0.5 d^2 hi hj) (frame/dgsks_ref.c:271-288) with
hi = hj = 1/h, so K = exp(-d^2 / (2 h_i h_j)).
"""
from __future__ import annotations

import numpy as np

GAUSSIAN, POLYNOMIAL, LAPLACE, VAR_BANDWIDTH, TANH = 0, 1, 2, 3, 4


def sqnorm(X: np.ndarray) -> np.ndarray:
    """Left-to-right FP64 sum of squares per row."""
    out = np.zeros(X.shape[0])
    for p in range(X.shape[1]):
        out += X[:, p] * X[:, p]
    return out


def gaussian_scal(h: float) -> float:
    return -1.0 / (2.0 * h * h)


def uniform_gauss16(m=8192, n=8192, k=16, seed=1):
    rng = np.random.default_rng(seed)
    N = max(m, n)
    X = rng.random((N, k))
    w = rng.standard_normal(N)
    return dict(X=X, X2=sqnorm(X), w=w, amap=np.arange(m, dtype=np.int32),
                bmap=np.arange(n, dtype=np.int32), k=k,
                kernel=dict(type=GAUSSIAN, scal=gaussian_scal(0.5)))


NBODY_KERNELS = {
    "laplace": dict(type=LAPLACE),
    "gaussian": dict(type=GAUSSIAN, scal=gaussian_scal(0.05)),
    "polynomial": dict(type=POLYNOMIAL, scal=1.0 / 3.0, cons=1.0, powe=3.0),
    "tanh": dict(type=TANH, scal=1.0 / 3.0, cons=0.1),
}


def nbody_k3(N=262144, k=3, seed=2):
    rng = np.random.default_rng(seed)
    X = rng.random((N, k))
    w = rng.random(N)
    perm = rng.permutation(N).astype(np.int32)
    return dict(X=X, X2=sqnorm(X), w=w, amap=perm, bmap=perm, k=k)


def mixture_k64(N=131072, k=64, ncomp=32, sigma=0.05, seed=3):
    rng = np.random.default_rng(seed)
    C = rng.random((ncomp, k))
    lab = rng.integers(0, ncomp, N)
    X = C[lab] + sigma * rng.standard_normal((N, k))
    w = rng.standard_normal(N)
    h = rng.uniform(0.2, 0.4, N)
    return dict(X=X, X2=sqnorm(X), w=w, amap=np.arange(N, dtype=np.int32),
                bmap=np.arange(N, dtype=np.int32), k=k, hinv=1.0 / h,
                kernel=dict(type=GAUSSIAN, scal=gaussian_scal(0.3)))


LIST_KERNELS = [dict(type=GAUSSIAN, scal=gaussian_scal(0.2)), dict(type=LAPLACE),
                dict(type=POLYNOMIAL, scal=None, cons=1.0, powe=2.0)]


def list_pool(k=3, ntask=4096, pool=1 << 20, lo=64, hi=8192, seed=4, symmetric=False):
    """Tasks as CSR offsets into concatenated maps; umap = amap, wmap = bmap.
    symmetric: every task's source list is its target list (the near-field
    self-interaction of a leaf: m = n, the same points in the same order)."""
    rng = np.random.default_rng(seed)
    P = rng.random((pool, k))
    w = rng.standard_normal(pool)
    ms = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), ntask))).astype(np.int64)
    ns = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), ntask))).astype(np.int64)
    if symmetric:
        ns = ms.copy()
    amaps, bmaps, kernels = [], [], []
    for t in range(ntask):
        amaps.append(rng.choice(pool, int(ms[t]), replace=False).astype(np.int32))
        if symmetric:
            bmaps.append(amaps[-1].copy())
        else:
            bmaps.append(rng.choice(pool, int(ns[t]), replace=False).astype(np.int32))
        kd = dict(LIST_KERNELS[int(rng.integers(0, 3))])
        if kd["type"] == POLYNOMIAL:
            kd["scal"] = 1.0 / k
        kernels.append(kd)
    a_off = np.concatenate([[0], np.cumsum(ms)]).astype(np.int64)
    b_off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    return dict(X=P, X2=sqnorm(P), w=w, k=k, kernels=kernels, a_off=a_off, b_off=b_off,
                amap=np.concatenate(amaps), bmap=np.concatenate(bmaps))


def grad_varbw16_maps(m=32768, n=32768, k=16, nx=40000, seed=9):
    """Variable-bandwidth gradient over index maps: one table of nx points in
    [0,1)^k and one bandwidth vector on both sides (hi = hj, per-point values
    in [0.6, 1.8)), maps drawn with repeats."""
    rng = np.random.default_rng(seed)
    return dict(X=rng.random((nx, k)), h=(0.5 + rng.random(nx)) * 1.2,
                w=rng.standard_normal(nx), g=rng.standard_normal(m),
                amap=rng.integers(0, nx, m).astype(np.int32),
                bmap=rng.integers(0, nx, n).astype(np.int32),
                wmap=rng.integers(0, nx, n).astype(np.int32), k=k)


def grad_gauss16(m=32768, n=32768, k=16, h=0.4, seed=5):
    rng = np.random.default_rng(seed)
    X = rng.random((max(m, n), k))
    w = rng.standard_normal(n)
    g = rng.standard_normal(m)
    return dict(XA=X[:m].copy(), XB=X[:n].copy(), w=w, g=g, h=h, k=k)
