// Device-side definitions shared by the fused summation kernels: the problem
// descriptor handed to a launch and the per-type kernel functions in FP64.
//
// Kernel-function semantics follow the reference's fused micro-kernels
// (micro_kernel/x86_64/sandybridge/ks_*_int_d8x4.c): the squared distance is
// assembled as |a|^2 - 2 a.b + |b|^2 and clamped at 0 before the kernel
// function; Laplace replaces d^2 < 1e-15 by 1.79e308 so that coincident
// points contribute ~0 (its d^2 is formed bit-identically to the reference's);
// quartic and Epanechnikov clamp d^2 into [0, 1].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gsks {

enum KType : int {
  GAUSSIAN = 0,
  POLYNOMIAL = 1,
  LAPLACE = 2,
  VAR_BANDWIDTH = 3,
  TANH = 4,
  QUARTIC = 5,
  MULTIQUADRATIC = 6,
  EPANECHNIKOV = 7
};

// How the polynomial / Laplace power is evaluated.
enum PowMode : int {
  POW_GENERAL = 0,  // pow(x, powe)
  POW_SQUARE = 1,   // x^2   (reference fast path, frame/dgsks_ref.c:297-305)
  POW_FOURTH = 2,   // (x^2)^2 (frame/dgsks_ref.c:306-319)
  POW_INT = 3       // x^ipow by repeated squaring, |ipow| <= 64
};

struct KerParams {
  double scal;
  double cons;
  double powe;
  int ipow;   // integer exponent (POW_INT), or k-2 for Laplace
  int pmode;  // PowMode
};

// Internal kernel types: the polynomial with its power fixed at compile time,
// so that a kernel's inner loop carries no per-pair branch on the power mode.
// Same products as powe_f64 (x^2, (x^2) x, (x^2)^2).
constexpr int POLY_SQUARE_T = 8;
constexpr int POLY_CUBE_T = 9;
constexpr int POLY_FOURTH_T = 10;
__host__ __device__ inline int poly_type(const KerParams &P) {
  if (P.pmode == POW_SQUARE || (P.pmode == POW_INT && P.ipow == 2)) return POLY_SQUARE_T;
  if (P.pmode == POW_INT && P.ipow == 3) return POLY_CUBE_T;
  if (P.pmode == POW_FOURTH || (P.pmode == POW_INT && P.ipow == 4)) return POLY_FOURTH_T;
  return POLYNOMIAL;
}

// One summation, all pointers on the device.
struct Problem {
  int m, n, k;
  int type;
  KerParams P;
  double *u;
  const int *umap;
  const double *XA, *XA2;
  const int *amap;
  const double *XB, *XB2;
  const int *bmap;
  const double *w;
  const int *wmap;
  const double *hi, *hj;
};

// ---------------------------------------------------------------- FP64 math

// x^e for integer e (|e| <= 64) by binary powering; a negative e inverts at the end.
__device__ __forceinline__ double ipow_f64(double x, int e) {
  int ae = e < 0 ? -e : e;
  double r = 1.0, b = x;
#pragma unroll 1
  while (ae) {
    if (ae & 1) r *= b;
    b *= b;
    ae >>= 1;
  }
  return e < 0 ? 1.0 / r : r;
}

// x[i]^e for N values at once (e wave-uniform): one squaring ladder for all N,
// so their multiplications interleave instead of running N dependent loops.
// The same products in the same order as ipow_f64, so the results are equal.
template <int N>
__device__ __forceinline__ void ipow_n_f64(double (&x)[N], int e) {
  int ae = e < 0 ? -e : e;
  double r[N], b[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    r[i] = 1.0;
    b[i] = x[i];
  }
#pragma unroll 1
  while (ae) {
    if (ae & 1) {
#pragma unroll
      for (int i = 0; i < N; ++i) r[i] *= b[i];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) b[i] *= b[i];
    ae >>= 1;
  }
#pragma unroll
  for (int i = 0; i < N; ++i) x[i] = e < 0 ? 1.0 / r[i] : r[i];
}

__device__ __forceinline__ double powe_f64(double x, const KerParams &P) {
  switch (P.pmode) {
    case POW_SQUARE: return x * x;
    case POW_FOURTH: { double t = x * x; return t * t; }
    case POW_INT: return ipow_f64(x, P.ipow);
    default: return pow(x, P.powe);
  }
}

// Laplace: scal * d2^((2-k)/2) = scal * rsqrt(d2)^(k-2).  P.ipow holds k-2.
__device__ __forceinline__ double laplace_f64(double d2, const KerParams &P) {
  d2 = (d2 < 1e-15) ? 1.79e308 : d2;  // frame/dgsks_ref.c:343-345
  double s = rsqrt(d2);
  double r = (P.ipow == 1) ? s : ipow_f64(s, P.ipow);
  return P.scal * r;
}

// laplace_f64 of N squared distances at once (the power ladder shared), in place.
template <int N>
__device__ __forceinline__ void laplace_n_f64(double (&d2)[N], const KerParams &P) {
#pragma unroll
  for (int i = 0; i < N; ++i) d2[i] = rsqrt((d2[i] < 1e-15) ? 1.79e308 : d2[i]);
  if (P.ipow != 1) ipow_n_f64<N>(d2, P.ipow);
#pragma unroll
  for (int i = 0; i < N; ++i) d2[i] = P.scal * d2[i];
}

// Pointer into the constant address space: loads through it are known not to be
// clobbered by the kernel, so wave-uniform loads become scalar (SMEM) loads.
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;

template <typename T>
__device__ __forceinline__ cptr<T> to_const(const T *p) {
  return (cptr<T>)(p);
}

// The reference's sequential rank-k sum for a runtime k through plain pointers
// (rare exact recomputes inside matrix-core kernels).
__device__ __forceinline__ double dot_exact_n(const double *a, const double *b, int k) {
#pragma clang fp contract(off)
  double c = a[0] * b[0];
  for (int p = 1; p < k; ++p) c = c + a[p] * b[p];
  return c;
}

// Rank-k sum taken exactly as the reference's loops do (a rounded product added
// to a rounded running sum, p = 0..k-1, no FMA).  Used where a threshold on d^2
// makes the result discontinuous (Laplace: d^2 < 1e-15 drops the pair,
// frame/dgsks_ref.c:343-345): with the same sum and the reference's assembly
// order, d^2 is bit-identical to the reference's, so the same pairs are dropped.
template <int K>
__device__ __forceinline__ double dot_exact(const double *a, cptr<double> b) {
#pragma clang fp contract(off)
  double c = a[0] * b[0];
#pragma unroll
  for (int p = 1; p < K; ++p) c = c + a[p] * b[p];
  return c;
}

__device__ __forceinline__ double dot_exact_rt(const double *a, cptr<double> b, int k) {
#pragma clang fp contract(off)
  double c = a[0] * b[0];
  for (int p = 1; p < k; ++p) c = c + a[p] * b[p];
  return c;
}

// FP64 matrix-core operand fragment of one point (v_mfma_f64_16x16x4f64, lane
// group g = lane / 16 supplies the k index): coordinates c*16 + 4g .. +3 for each 16-chunk c.
// EXACT (k == 16*KC, 16-byte aligned rows): two unguarded 16-byte loads per chunk.
template <int KC, bool EXACT>
__device__ __forceinline__ void load_frag(const double *__restrict__ X, long base, int k, int g,
                                          double (&f)[KC][4]) {
#pragma unroll
  for (int c = 0; c < KC; ++c) {
    if constexpr (EXACT) {
      const double2 *q = reinterpret_cast<const double2 *>(X + base + c * 16 + 4 * g);
      const double2 x0 = q[0], x1 = q[1];
      f[c][0] = x0.x;
      f[c][1] = x0.y;
      f[c][2] = x1.x;
      f[c][3] = x1.y;
    } else {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        int idx = c * 16 + 4 * g + t;
        f[c][t] = (idx < k) ? X[base + idx] : 0.0;
      }
    }
  }
}

// tanh x = -e / (2 + e) with e = expm1(-2|x|) in (-1, 0], sign restored: no
// cancellation anywhere (expm1 is accurate for tiny arguments), at most ~3 ulps,
// and tanh(+-0) = +-0, tanh(+-inf) = +-1, NaN in NaN out.  A quarter of the
// FP64 operations of the device library's tanh, which the tanh kernel's time
// is made of.
__device__ __forceinline__ double tanh_f64(double x) {
  const double e = expm1(-2.0 * fabs(x));
  return copysign(-e / (2.0 + e), x);
}

// Kernel value from the rank-k sum c = a.b, the norms aa = |a|^2, bb = |b|^2 and
// (variable bandwidth) the per-point factors hi, hj.
template <int TYPE>
__device__ __forceinline__ double kval_f64(double c, double aa, double bb, double hi,
                                           double hj, const KerParams &P) {
  if constexpr (TYPE == POLYNOMIAL) {
    return powe_f64(c * P.scal + P.cons, P);
  } else if constexpr (TYPE == POLY_SQUARE_T) {
    const double x = c * P.scal + P.cons;
    return x * x;
  } else if constexpr (TYPE == POLY_CUBE_T) {
    const double x = c * P.scal + P.cons;
    return (x * x) * x;
  } else if constexpr (TYPE == POLY_FOURTH_T) {
    const double x = c * P.scal + P.cons;
    const double t = x * x;
    return t * t;
  } else if constexpr (TYPE == TANH) {
    return tanh_f64(c * P.scal + P.cons);
  } else {
    double d2 = fma(-2.0, c, aa) + bb;  // |a|^2 - 2 a.b + |b|^2, reference order
    if constexpr (TYPE == GAUSSIAN) {
      return exp(fmax(d2, 0.0) * P.scal);
    } else if constexpr (TYPE == VAR_BANDWIDTH) {
      // exp(-0.5 d2 hi hj); hi already carries the -0.5 (exact power-of-two scaling)
      return exp((fmax(d2, 0.0) * hi) * hj);
    } else if constexpr (TYPE == LAPLACE) {
      return laplace_f64(d2, P);
    } else if constexpr (TYPE == QUARTIC) {
      double t = 1.0 - fmin(fmax(d2, 0.0), 1.0);
      return (15.0 / 16.0) * t * t;
    } else if constexpr (TYPE == MULTIQUADRATIC) {
      return fmax(d2, 0.0) + P.cons;
    } else {  // EPANECHNIKOV
      return 0.75 * (1.0 - fmin(fmax(d2, 0.0), 1.0));
    }
  }
}

}  // namespace gsks
