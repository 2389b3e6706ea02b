"""Writes tests/golden/gsks_golden.npz: small input/output vectors of the
reference, produced by calling the compiled reference library
(oracle/_ref/libgsks_ref.so, built from the reference's own sources by
oracle/Makefile.ref) -- both its oracle dgsks_ref (frame/dgsks_ref.c:94) and its
fused dgsks (frame/dgsks.c:367) -- on seeded inputs.

Run in a container that holds the reference tree:
    python tests/golden/make_golden.py
The fixture is data only (inputs and the reference's outputs).

Cases:
  * every kernel type at three ragged shapes (m, n not multiples of the
    reference's 8x4 register tile), random index maps that are subsets of a
    larger point table, a nonzero starting u (the call accumulates);
  * the data convention of the reference's own driver test/test_dgsks.c:144-207
    (coordinates (r % 100) / 1000, w = (r % 1000) / 1000, XB = XA, identity maps,
    variable bandwidth hi = hj = -1 / (2 (1 + 0.5 / (1 + exp(-|x|^2)))^2));
  * a batch of tasks as test/test_dgsks_list.cpp:75-230 builds them (random
    permutation prefixes, w = 1, Gaussian scal = -0.5), reference = the
    per-task dgsks_ref loop of that test.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))
from oracle import oracle as O  # noqa: E402

PARAMS = {
    O.GAUSSIAN: dict(scal=-0.7),
    O.POLYNOMIAL: dict(scal=0.5, cons=0.2, powe=3.0),
    O.LAPLACE: dict(),
    O.VAR_BANDWIDTH: dict(),
    O.TANH: dict(scal=0.3, cons=0.1),
    O.QUARTIC: dict(),
    O.MULTIQUADRATIC: dict(cons=1.0),
    O.EPANECHNIKOV: dict(),
}
SHAPES = [(37, 29, 3), (70, 45, 16), (33, 61, 5)]


def main():
    if O.build_ref() is None:
        raise SystemExit("oracle/_ref is not built and the reference tree is absent")
    O.build()
    out = {}
    rng = np.random.default_rng(20251203)
    case = 0
    for (m, n, k) in SHAPES:
        for t in range(8):
            nx = max(m, n) + 7
            X = rng.random((nx, k)) * 0.6
            X2 = O.sqnorm(X)
            amap = rng.permutation(nx)[:m].astype(np.int32)
            umap = rng.permutation(nx)[:m].astype(np.int32)
            bmap = rng.permutation(nx)[:n].astype(np.int32)
            wmap = rng.permutation(nx)[:n].astype(np.int32)
            w = rng.standard_normal(nx)
            u0 = rng.standard_normal(nx)
            prm = dict(PARAMS[t])
            if t == O.POLYNOMIAL:
                prm["powe"] = [2.0, 3.0, 4.0][case % 3]
            h = rng.uniform(-2.0, -0.5, nx) if t == O.VAR_BANDWIDTH else None
            ref_u = u0.copy()
            powe, scal = O.ref_call("dgsks_ref", t, ref_u, umap, X, X2, amap, X, X2, bmap, w,
                                    wmap, hi=h, hj=h, **prm)
            fast_u = u0.copy()
            O.ref_call("dgsks", t, fast_u, umap, X, X2, amap, X, X2, bmap, w, wmap, hi=h,
                       hj=h, **prm)
            key = f"c{case:02d}"
            out[key + "_meta"] = np.array([t, m, n, k], dtype=np.int64)
            out[key + "_par"] = np.array([prm.get("scal", 0.0), prm.get("cons", 0.0),
                                          prm.get("powe", 0.0), powe, scal])
            for name, arr in (("X", X), ("X2", X2), ("amap", amap), ("umap", umap),
                              ("bmap", bmap), ("wmap", wmap), ("w", w), ("u0", u0),
                              ("ref_u", ref_u), ("fast_u", fast_u)):
                out[f"{key}_{name}"] = arr
            if h is not None:
                out[key + "_h"] = h
            case += 1

    # the input convention of the test program test/test_dgsks.c (glibc rand() replaced by a seeded
    # integer stream with the same value ranges)
    rs = np.random.default_rng(1)
    for t in range(8):
        m, n, k = 53, 41, 8
        nx = 64
        X = (rs.integers(0, 2**31 - 1, (nx, k)) % 100) / 1000.0
        X2 = O.sqnorm(X)
        w = (rs.integers(0, 2**31 - 1, nx) % 1000) / 1000.0
        ident_m = np.arange(m, dtype=np.int32)
        ident_n = np.arange(n, dtype=np.int32)
        prm = {O.GAUSSIAN: dict(scal=-0.5), O.POLYNOMIAL: dict(powe=4.0, scal=0.1, cons=0.1),
               O.TANH: dict(scal=0.1, cons=0.1), O.MULTIQUADRATIC: dict(cons=1.0)}.get(t, {})
        h = None
        if t == O.VAR_BANDWIDTH:
            h = 1.0 + 0.5 / (1.0 + np.exp(-1.0 * X2))
            h = -1.0 / (2.0 * h * h)
        ref_u = np.zeros(nx)
        powe, scal = O.ref_call("dgsks_ref", t, ref_u, ident_m, X, X2, ident_m, X, X2, ident_n,
                                w, ident_n, hi=h, hj=h, **prm)
        fast_u = np.zeros(nx)
        O.ref_call("dgsks", t, fast_u, ident_m, X, X2, ident_m, X, X2, ident_n, w, ident_n,
                   hi=h, hj=h, **prm)
        key = f"d{t}"
        out[key + "_meta"] = np.array([t, m, n, k], dtype=np.int64)
        out[key + "_par"] = np.array([prm.get("scal", 0.0), prm.get("cons", 0.0),
                                      prm.get("powe", 0.0), powe, scal])
        for name, arr in (("X", X), ("X2", X2), ("amap", ident_m), ("umap", ident_m),
                          ("bmap", ident_n), ("wmap", ident_n), ("w", w),
                          ("u0", np.zeros(nx)), ("ref_u", ref_u), ("fast_u", fast_u)):
            out[f"{key}_{name}"] = arr
        if h is not None:
            out[key + "_h"] = h

    # list: tasks of test/test_dgsks_list.cpp, one shared point table, w = 1
    rl = np.random.default_rng(7)
    nx, k, ntask = 300, 4, 12
    X = (rl.integers(0, 2**31 - 1, (nx, k)) % 100) / 1000.0
    X2 = O.sqnorm(X)
    w = np.ones(nx)
    a_sizes = rl.integers(1, 40, ntask) + 5
    b_sizes = rl.integers(1, 30, ntask) + 5
    amaps = [rl.permutation(nx)[:s].astype(np.int32) for s in a_sizes]
    bmaps = [rl.permutation(nx)[:s].astype(np.int32) for s in b_sizes]
    ref_u = np.zeros(nx)
    for t in range(ntask):
        O.ref_call("dgsks_ref", O.GAUSSIAN, ref_u, amaps[t], X, X2, amaps[t], X, X2, bmaps[t], w,
                   bmaps[t], scal=-0.5)
    out["L_X"] = X
    out["L_X2"] = X2
    out["L_w"] = w
    out["L_aoff"] = np.concatenate([[0], np.cumsum(a_sizes)]).astype(np.int64)
    out["L_boff"] = np.concatenate([[0], np.cumsum(b_sizes)]).astype(np.int64)
    out["L_amap"] = np.concatenate(amaps)
    out["L_bmap"] = np.concatenate(bmaps)
    out["L_ref_u"] = ref_u
    path = os.path.join(HERE, "gsks_golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {case} + 8 + list cases)")


if __name__ == "__main__":
    main()
