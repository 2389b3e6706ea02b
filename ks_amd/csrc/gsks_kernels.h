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

// exp(x) by Tang's table-driven method (P. T. P. Tang, "Table-driven
// implementation of the exponential function in IEEE floating-point
// arithmetic", ACM TOMS 15(2), 1989) with 64 table points:
//   n = rint(64 x / ln 2)        (the low word of x 64/ln2 + 1.5 2^52),
//   r = x - n ln2/64             (ln2/64 in two parts, the first with 36
//                                 significant bits, so n times it is exact),
//   exp(x) = 2^(n >> 6) 2^((n & 63)/64) exp(r),
// with exp(r) - 1 as its Taylor series to r^5 (|r| <= ln2/128, truncation
// 3.4e-17 relative) and 2^(j/64) as a head and a tail.  Within 0.76 ulp of the
// exact value over [-745.2, 709.7] (checked against 50-digit arithmetic), 0
// below the subnormal range, inf above 709.78, exactly 1 at 0.  About 15
// vector instructions against about 22 in the device library's exp, whose
// kernel values the FP64 kernels spend most of their vector time on.  The
// table is read from LDS: a kernel copies kExpTable there first.
constexpr double kExpC = 0x1.71547652b82fep+6;    // 64 / ln 2
constexpr double kExpL1 = 0x1.62e42fefa0000p-7;   // ln2/64, 36 significant bits
constexpr double kExpL2 = 0x1.cf79abc9e3b3ap-46;  // ln2/64 - kExpL1
static __constant__ double2 kExpTable[64] = {  // 2^(j/64) = head + tail, correctly rounded head
    {0x1.0000000000000p+0, 0x0.0p+0}, {0x1.02c9a3e778061p+0, -0x1.19083535b085dp-56},
    {0x1.059b0d3158574p+0, 0x1.d73e2a475b465p-55}, {0x1.0874518759bc8p+0, 0x1.186be4bb284ffp-57},
    {0x1.0b5586cf9890fp+0, 0x1.8a62e4adc610bp-54}, {0x1.0e3ec32d3d1a2p+0, 0x1.03a1727c57b53p-59},
    {0x1.11301d0125b51p+0, -0x1.6c51039449b3ap-54}, {0x1.1429aaea92de0p+0, -0x1.32fbf9af1369ep-54},
    {0x1.172b83c7d517bp+0, -0x1.19041b9d78a76p-55}, {0x1.1a35beb6fcb75p+0, 0x1.e5b4c7b4968e4p-55},
    {0x1.1d4873168b9aap+0, 0x1.e016e00a2643cp-54}, {0x1.2063b88628cd6p+0, 0x1.dc775814a8495p-55},
    {0x1.2387a6e756238p+0, 0x1.9b07eb6c70573p-54}, {0x1.26b4565e27cddp+0, 0x1.2bd339940e9d9p-55},
    {0x1.29e9df51fdee1p+0, 0x1.612e8afad1255p-55}, {0x1.2d285a6e4030bp+0, 0x1.0024754db41d5p-54},
    {0x1.306fe0a31b715p+0, 0x1.6f46ad23182e4p-55}, {0x1.33c08b26416ffp+0, 0x1.32721843659a6p-54},
    {0x1.371a7373aa9cbp+0, -0x1.63aeabf42eae2p-54}, {0x1.3a7db34e59ff7p+0, -0x1.5e436d661f5e3p-56},
    {0x1.3dea64c123422p+0, 0x1.ada0911f09ebcp-55}, {0x1.4160a21f72e2ap+0, -0x1.ef3691c309278p-58},
    {0x1.44e086061892dp+0, 0x1.89b7a04ef80d0p-59}, {0x1.486a2b5c13cd0p+0, 0x1.3c1a3b69062f0p-56},
    {0x1.4bfdad5362a27p+0, 0x1.d4397afec42e2p-56}, {0x1.4f9b2769d2ca7p+0, -0x1.4b309d25957e3p-54},
    {0x1.5342b569d4f82p+0, -0x1.07abe1db13cadp-55}, {0x1.56f4736b527dap+0, 0x1.9bb2c011d93adp-54},
    {0x1.5ab07dd485429p+0, 0x1.6324c054647adp-54}, {0x1.5e76f15ad2148p+0, 0x1.ba6f93080e65ep-54},
    {0x1.6247eb03a5585p+0, -0x1.383c17e40b497p-54}, {0x1.6623882552225p+0, -0x1.bb60987591c34p-54},
    {0x1.6a09e667f3bcdp+0, -0x1.bdd3413b26456p-54}, {0x1.6dfb23c651a2fp+0, -0x1.bbe3a683c88abp-57},
    {0x1.71f75e8ec5f74p+0, -0x1.16e4786887a99p-55}, {0x1.75feb564267c9p+0, -0x1.0245957316dd3p-54},
    {0x1.7a11473eb0187p+0, -0x1.41577ee04992fp-55}, {0x1.7e2f336cf4e62p+0, 0x1.05d02ba15797ep-56},
    {0x1.82589994cce13p+0, -0x1.d4c1dd41532d8p-54}, {0x1.868d99b4492edp+0, -0x1.fc6f89bd4f6bap-54},
    {0x1.8ace5422aa0dbp+0, 0x1.6e9f156864b27p-54}, {0x1.8f1ae99157736p+0, 0x1.5cc13a2e3976cp-55},
    {0x1.93737b0cdc5e5p+0, -0x1.75fc781b57ebcp-57}, {0x1.97d829fde4e50p+0, -0x1.d185b7c1b85d1p-54},
    {0x1.9c49182a3f090p+0, 0x1.c7c46b071f2bep-56}, {0x1.a0c667b5de565p+0, -0x1.359495d1cd533p-54},
    {0x1.a5503b23e255dp+0, -0x1.d2f6edb8d41e1p-54}, {0x1.a9e6b5579fdbfp+0, 0x1.0fac90ef7fd31p-54},
    {0x1.ae89f995ad3adp+0, 0x1.7a1cd345dcc81p-54}, {0x1.b33a2b84f15fbp+0, -0x1.2805e3084d708p-57},
    {0x1.b7f76f2fb5e47p+0, -0x1.5584f7e54ac3bp-56}, {0x1.bcc1e904bc1d2p+0, 0x1.23dd07a2d9e84p-55},
    {0x1.c199bdd85529cp+0, 0x1.11065895048ddp-55}, {0x1.c67f12e57d14bp+0, 0x1.2884dff483cadp-54},
    {0x1.cb720dcef9069p+0, 0x1.503cbd1e949dbp-56}, {0x1.d072d4a07897cp+0, -0x1.cbc3743797a9cp-54},
    {0x1.d5818dcfba487p+0, 0x1.2ed02d75b3707p-55}, {0x1.da9e603db3285p+0, 0x1.c2300696db532p-54},
    {0x1.dfc97337b9b5fp+0, -0x1.1a5cd4f184b5cp-54}, {0x1.e502ee78b3ff6p+0, 0x1.39e8980a9cc8fp-55},
    {0x1.ea4afa2a490dap+0, -0x1.e9c23179c2893p-54}, {0x1.efa1bee615a27p+0, 0x1.dc7f486a4b6b0p-54},
    {0x1.f50765b6e4540p+0, 0x1.9d3e12dd8a18bp-54}, {0x1.fa7c1819e90d8p+0, 0x1.74853f3a5931ep-55},
};

__device__ __forceinline__ double exp_tab(double x, const double2 *__restrict__ tab) {
  x = fmin(fmax(x, -746.0), 710.0);  // n stays far inside int; exp(-746) = 0, exp(710) = inf
  const double s = fma(x, kExpC, 0x1.8p52);
  const int n = (int)__double2loint(s);
  const double nf = s - 0x1.8p52;
  double r = fma(-nf, kExpL1, x);
  r = fma(-nf, kExpL2, r);
  double q = fma(r, 1.0 / 120.0, 1.0 / 24.0);
  q = fma(q, r, 1.0 / 6.0);
  q = fma(q, r, 0.5);
  q *= r;
  const double p = fma(q, r, r);  // exp(r) - 1
  const double2 t = tab[n & 63];
  return ldexp(t.x + fma(t.x, p, t.y), n >> 6);
}

// One DPP lane exchange of a double (two 32-bit moves), all lanes valid.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  int2 v = __builtin_bit_cast(int2, x);
  v.x = __builtin_amdgcn_mov_dpp(v.x, CTRL, 0xF, 0xF, true);
  v.y = __builtin_amdgcn_mov_dpp(v.y, CTRL, 0xF, 0xF, true);
  return __builtin_bit_cast(double, v);
}

// Sum over the 16 lanes of each DPP row; every lane of a row gets its row's
// value.  Butterflies: swap neighbours, swap pairs, mirror the 8-lane halves,
// mirror the row (a fixed order: the same lanes give the same sum every run).
__device__ __forceinline__ double row_sum16(double x) {
  x += dpp_f64<0xB1>(x);   // quad_perm [1,0,3,2]
  x += dpp_f64<0x4E>(x);   // quad_perm [2,3,0,1]
  x += dpp_f64<0x141>(x);  // row_half_mirror
  x += dpp_f64<0x140>(x);  // row_mirror
  return x;
}

// A kernel's LDS copy of kExpTable, filled by its first 64 lanes (the caller
// synchronises the workgroup before the first exp_tab).
__device__ __forceinline__ void load_exp_table(double2 *tab) {
  if (threadIdx.x < 64) tab[threadIdx.x] = kExpTable[threadIdx.x];
}

// Whether TYPE's FP64 kernel values call exp (and so take exp_tab).
template <int TYPE>
constexpr bool uses_exp = TYPE == GAUSSIAN || TYPE == VAR_BANDWIDTH;

// Kernel value from the rank-k sum c = a.b, the norms aa = |a|^2, bb = |b|^2 and
// (variable bandwidth) the per-point factors hi, hj.
// With ETAB, the Gaussian kernels take exp_tab over `etab` (an LDS copy of
// kExpTable).
template <int TYPE, bool ETAB = false>
__device__ __forceinline__ double kval_f64(double c, double aa, double bb, double hi,
                                           double hj, const KerParams &P,
                                           const double2 *etab = nullptr) {
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
      const double x = fmax(d2, 0.0) * P.scal;
      if constexpr (ETAB) return exp_tab(x, etab);
      return exp(x);
    } else if constexpr (TYPE == VAR_BANDWIDTH) {
      // exp(-0.5 d2 hi hj); hi already carries the -0.5 (exact power-of-two scaling)
      const double x = (fmax(d2, 0.0) * hi) * hj;
      if constexpr (ETAB) return exp_tab(x, etab);
      return exp(x);
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
