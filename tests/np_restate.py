"""An independent numpy restatement of the reference formulas
(frame/dgsks_ref.c:253-424), vectorised over a full m x n block -- a second,
differently written implementation the C oracle is checked against.  Small
sizes only (it materialises the kernel matrix)."""
import math

import numpy as np

GAUSSIAN, POLYNOMIAL, LAPLACE, VAR_BANDWIDTH, TANH, QUARTIC, MULTIQUADRATIC, EPANECHNIKOV = range(8)


def kernel_matrix(t, A, B, A2, B2, scal=0.0, cons=0.0, powe=0.0, hi=None, hj=None):
    """K[i, j] for rows A (m x k) and B (n x k); hi/hj per row/column."""
    G = A @ B.T
    D = (A2[:, None] - 2.0 * G) + B2[None, :]
    if t == GAUSSIAN:
        return np.exp(scal * D)
    if t == VAR_BANDWIDTH:
        return np.exp(-0.5 * D * hi[:, None] * hj[None, :])
    if t == POLYNOMIAL:
        return np.power(scal * G + cons, powe)
    if t == LAPLACE:
        k = A.shape[1]
        p = 0.5 * (2.0 - k)
        s = math.gamma(0.5 * k + 1.0) / (k * (k - 2) * math.pi ** (0.5 * k))
        Dc = np.where(D < 1e-15, 1.79e308, D)
        return s * np.power(Dc, p)
    if t == TANH:
        return np.tanh(scal * G + cons)
    if t == QUARTIC:
        return np.where(D < 1.0, (15.0 / 16.0) * (1.0 - D) ** 2, 0.0)
    if t == MULTIQUADRATIC:
        return D + cons
    if t == EPANECHNIKOV:
        return np.where(D < 1.0, 0.75 * (1.0 - D), 0.0)
    raise ValueError(t)


def dgsks(t, u, umap, X, X2, amap, Y, Y2, bmap, w, wmap, **kw):
    hi = kw.pop("hi", None)
    hj = kw.pop("hj", None)
    K = kernel_matrix(t, X[amap], Y[bmap], X2[amap], Y2[bmap],
                      hi=None if hi is None else hi[amap], hj=None if hj is None else hj[bmap],
                      **kw)
    out = u.copy()
    out[umap] = u[umap] + K @ w[wmap]
    return out
