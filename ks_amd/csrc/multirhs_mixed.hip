// Mixed-precision multi-right-hand-side Gaussian summation
//
//   U[umap[i]][r] += sum_j exp(scal |a_i - b_j|^2) W[wmap[j]][r],  r < nrhs <= 33,
//
// for points whose pairs are mostly "near" (exponents of a few units, as in a
// training step over unclustered data), where the tiers of the mixed
// single-right-hand-side kernel would recompute every block in FP64.  What
// makes FP32 unsafe for a pair is not a large kernel value but a large
// |scal| |a| |b|: the exponent's error is absolute, |scal| times the error of
// |a|^2 + |b|^2 - 2 a.b, and FP32 arithmetic bounds that by a multiple of
// 2^-24 |a| |b|.  Translating every point by one centre (the mean of the
// targets and sources; distances do not change) makes |a| and |b| as small as
// the data allows, and a rigorous bound then decides, per 16 x 16 block,
// whether FP32 is accurate enough:
//
//   FP32 tier:  a32 = fl32(a - centre), b32 likewise; |a32|^2 and |b32|^2 in
//     FP64 (exact squares, FP64 sums); c = a32 . b32 on the FP32 matrix cores
//     (v_mfma_f32_16x16x4f32).  The exponent in log2 units,
//        t = xs (|a|^2 + |b|^2 - 2c) = alpha_i + beta_j + gamma c,
//     xs = log2(e) scal, gamma = -2 xs, has its per-point parts split once in
//     FP64 into an integer and an FP32 fraction: alpha_i = n_i + f_i,
//     beta_j = n_j + f_j, |f| <= 1/2.  Per pair only fma(gamma, c, f_i) + f_j
//     and one v_exp_f32 remain (3 VALU instructions): 2^n_j is folded into
//     source j's FP32 weights (exact scaling) and 2^n_i into target i's FP64
//     row sums at the end.  U += K W runs on the FP32 matrix cores, flushed
//     into FP64 accumulators every 32 sources.  The exponent's error, in
//     natural-log units, is at most
//        |scal| ((2 (kp + 2) + 6) u A B + 2.01 u (A + B)^2) + 4u,   u = 2^-24,
//     where kp is k rounded up to 16, A bounds |a32| over the wave's 32 targets
//     and B bounds |b32| over the block's 16 sources: 2 (kp + 2) u A B covers
//     the FP32 dot product, 6 u A B the FP32 roundings of gamma, gamma c and
//     the two additions, 2.01 u (A + B)^2 the rounding of both points to FP32
//     and 4u the fractions.  A block takes this tier only when the bound is at
//     most kTau = 2^-15, so every kernel value has a relative error of at most
//     2^-15 + 2^-22 (exp2 within 1 ulp, the FP32 roundings of K and W); the
//     errors are roundings of either sign and the summed results are observed
//     near 3e-7 of the FP64 values.  d^2 is not clamped at 0 here: a computed
//     d^2 below 0 is within the bound of the true one.
//   FP64 tier (every other block): the FP64 matrix-core block of the FP64
//     multi-right-hand-side kernel, from the caller's FP64 coordinates and
//     norms, exactly as the drop-in computes it.
// The FP64 tier's results (rows g + 4v of the FP64 output layout) are added
// into an LDS tile of the workgroup's rows as each block finishes; the FP32
// tier's (rows 4g + v of the FP32 layout) stay in registers and are added to
// the tile at the end, so both tiers may serve one wave.
//
// Packing (pre-pass, in the order of the maps): block b of 16 points, chunk c of
// 16 coordinates is 64 lanes x 16 B, lane 16 g + p holding point 16 b + p,
// coordinates 16 c + 4 g .. + 3 (FP32, centred, zero past k).  Per point: the
// fraction f (FP32), the scale 2^n (FP64) and a bound on its norm (FP32).
// Weight tile t of block b is 64 lanes x 16 B, lane 16 g + p holding
// 2^n_j W[16 b + 4 g + v][16 t + p], j = 16 b + 4 g + v, for v = 0..3 (the B
// operand of the right-hand-side product).  Right-hand side nrhs - 1 is kept
// apart (one float per source, also scaled) and summed on the VALU.
#include <hip/hip_runtime.h>

#include "gsks_kernels.h"
#include "gsks_launch.h"
#include "host_util.h"

namespace gsks {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int MX_WG = 256;                // 4 waves
constexpr int MX_RB = 2;                  // 16-target groups per wave
constexpr int MX_ROWS = 4 * MX_RB * 16;   // targets per workgroup
constexpr float kTau = 1.0f / 32768.0f;   // largest exponent error of the FP32 tier
constexpr double kU32 = 1.0 / 16777216.0; // 2^-24
constexpr float kNormUp = 1.0f + 1.0f / 1048576.0f;  // covers the FP32 roundings of a norm bound
constexpr double kLog2eD = 1.4426950408889634;

struct MxArgs {
  Problem p;  // FP64 problem (FP64 tier): maps, coordinates, norms, W [.][nrhs]
  int nrhs;
  int cols_per_split, nsplit;  // cols_per_split is a multiple of 32
  double *partial;             // [nsplit][m][nrhs]
  const f4 *T;                 // targets, packed (FP32 centred)
  const float *Tf;             // [m_pad] fraction f_i of xs |a32|^2
  const double *Ts;            // [m_pad] 2^n_i
  const float *Tn;             // [m_pad] bound on |a32|
  const f4 *S;                 // sources, packed
  const float *Sf;             // [n_pad] fraction f_j
  const float *Smax;           // [n_pad / 16] bound on |b32| over each block
  const f4 *Wt;                // weight tiles (scaled by 2^n_j)
  const float *Wx;             // [n_pad] right-hand side nrhs - 1 (scaled)
  float gam;                   // -2 log2(e) scal
  float c1, c2, c0;            // exponent error <= c1 A B + c2 (A + B)^2 + c0
};

constexpr int kCentreRows = 256;  // rows per workgroup of the centre's first pass

// Centre, first pass: workgroup w sums coordinates 0..k-1 over rows
// [w kCentreRows, (w + 1) kCentreRows) of the list "targets, then sources"
// (through the maps); lanes q and q + 64, q + 128, q + 192 take coordinate q of
// every fourth row, so each row is read by consecutive lanes, and each lane
// keeps four rows' loads in flight.
__global__ __launch_bounds__(256) void mx_centre_partial_kernel(Problem p, double *part) {
  __shared__ double s[256];
  const int q = threadIdx.x & 63, r4 = threadIdx.x >> 6, k = p.k;
  const long r0 = (long)blockIdx.x * kCentreRows, total = (long)p.m + p.n;
  const long r1 = r0 + kCentreRows < total ? r0 + kCentreRows : total;
  auto row = [&](long r) {
    return r < p.m ? p.XA + (long)p.amap[r] * k : p.XB + (long)p.bmap[r - p.m] * k;
  };
  double acc = 0.0;
  if (q < k) {
    long r = r0 + r4;
    for (; r + 12 < r1; r += 16) {
      const double *x0 = row(r), *x1 = row(r + 4), *x2 = row(r + 8), *x3 = row(r + 12);
      const double v0 = x0[q], v1 = x1[q], v2 = x2[q], v3 = x3[q];
      acc += (v0 + v1) + (v2 + v3);
    }
    for (; r < r1; r += 4) acc += row(r)[q];
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < 64)
    part[(long)blockIdx.x * 64 + q] = (s[q] + s[q + 64]) + (s[q + 128] + s[q + 192]);
}

// Centre, second pass: lanes q, q + 64, q + 128, q + 192 sum every fourth
// partial of coordinate q, then the four are added: a fixed order.
__global__ __launch_bounds__(256) void mx_centre_final_kernel(const double *part, int nparts,
                                                              double inv, int k, double *centre) {
  __shared__ double s[256];
  const int q = threadIdx.x & 63, r4 = threadIdx.x >> 6;
  double acc = 0.0;
  if (q < k)
    for (int w = r4; w < nparts; w += 4) acc += part[(long)w * 64 + q];
  s[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < 64 && q < k)
    centre[q] = ((s[q] + s[q + 64]) + (s[q + 128] + s[q + 192])) * inv;
}

int centre_parts(long rows) { return (int)((rows + kCentreRows - 1) / kCentreRows); }

hipError_t launch_centre(const Problem &p, double *part, double *centre, hipStream_t s) {
  const int np = centre_parts((long)p.m + p.n);
  hipLaunchKernelGGL(mx_centre_partial_kernel, dim3(np), dim3(256), 0, s, p, part);
  hipLaunchKernelGGL(mx_centre_final_kernel, dim3(1), dim3(256), 0, s, part, np,
                     1.0 / ((double)p.m + (double)p.n), p.k, centre);
  return hipGetLastError();
}

// Centred FP32 coordinates of the mapped points, packed; per point the FP32
// fraction f and FP64 scale 2^n of xs |x32|^2 = n + f (|x32|^2 in FP64, the
// exponent clamped at -1000, where every FP32 value is 0) and a bound on |x32|.
template <int KC>
__global__ __launch_bounds__(256) void mx_pack_points_kernel(const double *__restrict__ X,
                                                             const int *__restrict__ map, int rows,
                                                             int rows_pad, int k,
                                                             const double *__restrict__ centre,
                                                             double xs, f4 *__restrict__ P,
                                                             float *__restrict__ F,
                                                             double *__restrict__ Sc,
                                                             float *__restrict__ Nb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows_pad) return;
  const bool ok = i < rows;
  const double *x = X + (ok ? (long)map[i] * k : 0l);
  const long b = i >> 4;
  const int p16 = i & 15;
  double nn = 0.0;
#pragma unroll
  for (int c = 0; c < KC; ++c)
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      f4 v;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int q = 16 * c + 4 * g + t;
        const float f = (ok && q < k) ? (float)(x[q] - centre[q]) : 0.0f;
        v[t] = f;
        nn = fma((double)f, (double)f, nn);
      }
      P[(b * KC + c) * 64 + 16 * g + p16] = v;
    }
  const double al = fmax(xs * nn, -1000.0);
  const double ni = rint(al);
  F[i] = (float)(al - ni);
  Sc[i] = ldexp(1.0, (int)ni);
  Nb[i] = sqrtf((float)nn) * kNormUp;
}

// Per source: the block norm bound, the separate right-hand side and the
// weight tiles, scaled by 2^n_j.  n_pad is a multiple of 16, so whole 16-lane
// groups are present.
template <int NT>
__global__ __launch_bounds__(256) void mx_pack_weights_kernel(const double *__restrict__ W,
                                                              const int *__restrict__ wmap, int n,
                                                              int n_pad, int nr,
                                                              const float *__restrict__ Nb,
                                                              const double *__restrict__ Sc,
                                                              float *__restrict__ Smax,
                                                              float *__restrict__ Wt,
                                                              float *__restrict__ Wx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_pad) return;
  const bool ok = j < n;
  const long wrow = ok ? (long)wmap[j] * nr : 0l;
  const double sc = Sc[j];
  float bn = Nb[j];
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) bn = fmaxf(bn, __shfl_xor(bn, o));
  if ((j & 15) == 0) Smax[j >> 4] = bn;
  Wx[j] = ok ? (float)(W[wrow + nr - 1] * sc) : 0.0f;
  const long b = j >> 4;
  const int g = (j >> 2) & 3, v = j & 3;
#pragma unroll
  for (int t = 0; t < NT; ++t)
    for (int pp = 0; pp < 16; ++pp) {
      const int r = 16 * t + pp;
      Wt[((b * NT + t) * 64 + 16 * g + pp) * 4 + v] =
          (ok && r < nr - 1) ? (float)(W[wrow + r] * sc) : 0.0f;
    }
}

// k <= 16 with up to 17 right-hand sides fits 3 waves per SIMD (168 VGPRs)
template <int KC, int NT>
__global__ __launch_bounds__(MX_WG, (KC <= 1 && NT <= 1) ? 3 : 2) void multirhs_mixed_kernel(
    MxArgs A) {
  constexpr int NTA = NT > 0 ? NT : 1;  // array extent
  constexpr int C = 16 * NT + 1;        // LDS columns: the tiles, then right-hand side nrhs - 1
  __shared__ double Us[MX_ROWS][C];
  const Problem &p = A.p;
  const int m = p.m, k = p.k, nr = A.nrhs;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int p16 = lane & 15, g = lane >> 4;
  const int rblk = blockIdx.x, split = blockIdx.y;
  const int wr0 = wave * (MX_RB * 16);  // the wave's first row in the workgroup
  const int row0 = rblk * MX_ROWS + wr0;

  // targets (B operand of the cross term): lane (p16, g) -> target row0 + 16 rb + p16
  f4 tf[MX_RB][KC];
  float fi[MX_RB];
  float amax = 0.0f;
#pragma unroll
  for (int rb = 0; rb < MX_RB; ++rb) {
    const long tb = (row0 + 16 * rb) >> 4;
#pragma unroll
    for (int c = 0; c < KC; ++c) tf[rb][c] = A.T[(tb * KC + c) * 64 + lane];
    fi[rb] = A.Tf[row0 + 16 * rb + p16];
    amax = fmaxf(amax, A.Tn[row0 + 16 * rb + p16]);
  }
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) amax = fmaxf(amax, __shfl_xor(amax, o));
  const float Ab = amax;

  // FP32 tier: FP64 sums (rows 4g + v, before the 2^n_i scaling) in registers;
  // FP64 tier: added straight into this wave's rows of Us (zeroed here)
  d4 Uf[MX_RB][NTA];
  double uxf[MX_RB];  // right-hand side nrhs - 1, target p16
#pragma unroll
  for (int rb = 0; rb < MX_RB; ++rb) {
    uxf[rb] = 0.0;
#pragma unroll
    for (int t = 0; t < NTA; ++t) Uf[rb][t] = d4{0.0, 0.0, 0.0, 0.0};
  }
  for (int e = lane; e < MX_RB * 16 * C; e += 64) Us[wr0 + e / C][e % C] = 0.0;

  const int j0 = split * A.cols_per_split;
  const int j1 = min(p.n, j0 + A.cols_per_split);
  const KerParams P = p.P;
  const float gam = A.gam;

  struct SB {
    f4 s[KC];
    f4 w[NTA];
    f4 wx, fj;
    float bmax;
  };
  auto fetch = [&](int jb, SB &b) {
    const long bi = jb >> 4;
#pragma unroll
    for (int c = 0; c < KC; ++c) b.s[c] = A.S[(bi * KC + c) * 64 + lane];
#pragma unroll
    for (int t = 0; t < NT; ++t) b.w[t] = A.Wt[(bi * NT + t) * 64 + lane];
    b.wx = *reinterpret_cast<const f4 *>(A.Wx + jb + 4 * g);
    b.fj = *reinterpret_cast<const f4 *>(A.Sf + jb + 4 * g);
    b.bmax = A.Smax[bi];
  };
  auto fp32_ok = [&](float bmax) {
    const float s = Ab + bmax;
    return __builtin_amdgcn_readfirstlane(
               (int)(A.c1 * Ab * bmax + A.c2 * s * s + A.c0 <= kTau)) != 0;
  };

  f4 Ua[MX_RB][NTA];  // FP32 partials of the current 32 sources
  float uxa[MX_RB];
  auto block32 = [&](const SB &b) {
    // the two target groups' cross terms as two interleaved MFMA chains
    f4 c[MX_RB];
#pragma unroll
    for (int rb = 0; rb < MX_RB; ++rb) c[rb] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int cc = 0; cc < KC; ++cc)
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int rb = 0; rb < MX_RB; ++rb)
          c[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(b.s[cc][t], tf[rb][cc][t], c[rb], 0, 0, 0);
    // slot v: source jb + 4g + v, target p16; K / (2^n_i 2^n_j)
    float Kv[MX_RB][4];
#pragma unroll
    for (int rb = 0; rb < MX_RB; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        Kv[rb][v] = __builtin_amdgcn_exp2f(fmaf(gam, c[rb][v], fi[rb]) + b.fj[v]);
#pragma unroll
    for (int rb = 0; rb < MX_RB; ++rb) {
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          Ua[rb][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(Kv[rb][v], b.w[t][v], Ua[rb][t], 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; ++v) uxa[rb] = fmaf(Kv[rb][v], b.wx[v], uxa[rb]);
    }
  };
  // FP64 tier: sources jb + 4v + g as the A operand's rows (FP64 output layout)
  auto block64 = [&](int jb) {
    double sf[KC][4];
    const int col = jb + p16;
    load_frag<KC, false>(p.XB, (long)p.bmap[col < j1 ? col : j1 - 1] * k, k, g, sf);
    double bb[4], wv[4][NTA], wx[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int src = jb + 4 * v + g;
      const bool ok = src < j1;
      bb[v] = p.XB2[p.bmap[ok ? src : j1 - 1]];
      const long wrow = ok ? (long)p.wmap[src] * nr : -1l;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int r = 16 * t + p16;
        wv[v][t] = (wrow >= 0 && r < nr - 1) ? p.w[wrow + r] : 0.0;
      }
      wx[v] = wrow >= 0 ? p.w[wrow + nr - 1] : 0.0;
    }
#pragma unroll
    for (int rb = 0; rb < MX_RB; ++rb) {
      const int r = row0 + 16 * rb + p16;
      const int ai = p.amap[r < m ? r : m - 1];
      double af[KC][4];
      load_frag<KC, false>(p.XA, (long)ai * k, k, g, af);
      const double a2 = p.XA2[ai];
      d4 c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int cc = 0; cc < KC; ++cc)
#pragma unroll
        for (int t = 0; t < 4; ++t)
          c = __builtin_amdgcn_mfma_f64_16x16x4f64(sf[cc][t], af[cc][t], c, 0, 0, 0);
      double Kd[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) Kd[v] = kval_f64<GAUSSIAN>(c[v], a2, bb[v], 0.0, 0.0, P);
      // this lane's cells of the wave's LDS rows (rows g + 4v of the FP64 layout)
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        d4 u = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int v = 0; v < 4; ++v)
          u = __builtin_amdgcn_mfma_f64_16x16x4f64(Kd[v], wv[v][t], u, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) Us[wr0 + 16 * rb + g + 4 * v][16 * t + p16] += u[v];
      }
      double x = 0.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) x = fma(Kd[v], wx[v], x);
      x += __shfl_xor(x, 16);
      x += __shfl_xor(x, 32);
      if (g == 0) Us[wr0 + 16 * rb + p16][16 * NT] += x;
    }
  };

  SB b0, b1;
  if (j0 < j1) fetch(j0, b0);
  for (int jb = j0; jb < j1; jb += 32) {
    const bool two = jb + 16 < j1;
    if (two) fetch(jb + 16, b1);
#pragma unroll
    for (int rb = 0; rb < MX_RB; ++rb) {
      uxa[rb] = 0.0f;
#pragma unroll
      for (int t = 0; t < NTA; ++t) Ua[rb][t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    if (fp32_ok(b0.bmax))
      block32(b0);
    else
      block64(jb);
    if (two) {
      const float bm1 = b1.bmax;
      if (jb + 32 < j1) fetch(jb + 32, b0);
      if (fp32_ok(bm1))
        block32(b1);
      else
        block64(jb + 16);
    }
#pragma unroll
    for (int rb = 0; rb < MX_RB; ++rb) {
      uxf[rb] += (double)uxa[rb];
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) Uf[rb][t][v] += (double)Ua[rb][t][v];
    }
  }

  // ---- add the FP32 tier, with its targets' 2^n_i, to the FP64 tier's sums in
  // this wave's rows of Us (rows wr0 .. wr0 + 31)
  __syncthreads();
#pragma unroll
  for (int rb = 0; rb < MX_RB; ++rb) {
    const int rt = row0 + 16 * rb;  // first target of the group
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double sc = A.Ts[rt + 4 * g + v];
#pragma unroll
      for (int t = 0; t < NT; ++t) Us[wr0 + 16 * rb + 4 * g + v][16 * t + p16] += Uf[rb][t][v] * sc;
    }
    double xf = uxf[rb];
    xf += __shfl_xor(xf, 16);
    xf += __shfl_xor(xf, 32);
    if (g == 0) Us[wr0 + 16 * rb + p16][16 * NT] += xf * A.Ts[rt + p16];
  }
  __syncthreads();

  const int base = rblk * MX_ROWS;
  if (A.nsplit == 1) {
    for (int e = threadIdx.x; e < MX_ROWS * C; e += MX_WG) {
      const int rr = e / C, cc = e - rr * C, i = base + rr;
      const int r = cc == 16 * NT ? nr - 1 : cc;
      if (i < m && (cc == 16 * NT || cc < nr - 1)) p.u[(long)p.umap[i] * nr + r] += Us[rr][cc];
    }
    return;
  }
  for (int e = threadIdx.x; e < MX_ROWS * C; e += MX_WG) {
    const int rr = e / C, cc = e - rr * C, i = base + rr;
    const int r = cc == 16 * NT ? nr - 1 : cc;
    if (i < m && (cc == 16 * NT || cc < nr - 1))
      A.partial[((long)split * m + i) * nr + r] = Us[rr][cc];
  }
  // the splits' partials are added in split order by a reduction kernel
  // launched next (no device-scope fence here: on gfx950 it would write back
  // and invalidate this XCD's L2)
}

inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <int KC>
hipError_t mx_pack_points(const double *X, const int *map, int rows, int rows_pad, int k,
                          const double *centre, double xs, f4 *P, float *F, double *Sc, float *Nb,
                          hipStream_t s) {
  hipLaunchKernelGGL(mx_pack_points_kernel<KC>, dim3(blocks_for(rows_pad, 256)), dim3(256), 0, s, X,
                     map, rows, rows_pad, k, centre, xs, P, F, Sc, Nb);
  return hipGetLastError();
}

template <int NT>
hipError_t mx_pack_weights(const Problem &p, int nr, int n_pad, const float *Nb, const double *Sc,
                           float *Smax, float *Wt, float *Wx, hipStream_t s) {
  hipLaunchKernelGGL(mx_pack_weights_kernel<NT>, dim3(blocks_for(n_pad, 256)), dim3(256), 0, s,
                     p.w, p.wmap, p.n, n_pad, nr, Nb, Sc, Smax, Wt, Wx);
  return hipGetLastError();
}

template <int KC, int NT>
hipError_t mx_go(const MxArgs &A, int rblocks, hipStream_t s) {
  hipLaunchKernelGGL((multirhs_mixed_kernel<KC, NT>), dim3(rblocks, A.nsplit), dim3(MX_WG), 0, s, A);
  return hipGetLastError();
}

template <int KC>
hipError_t mx_nt(const MxArgs &A, int nt, int rblocks, hipStream_t s) {
  switch (nt) {
    case 0: return mx_go<KC, 0>(A, rblocks, s);
    case 1: return mx_go<KC, 1>(A, rblocks, s);
    default: return mx_go<KC, 2>(A, rblocks, s);
  }
}

}  // namespace

int enqueue_multirhs_mixed(const Problem &p, int nrhs, const double *centre, hipStream_t s) {
  if (p.type != GAUSSIAN || p.k < 1 || p.k > kMixedMrhsMaxK || nrhs < 1 ||
      nrhs > kMixedMrhsMaxRhs)
    return KS_ERR_INVALID_ARG;
  const int m = p.m, n = p.n, k = p.k;
  const int KC = (k + 15) / 16, NT = (nrhs - 1 + 15) / 16;
  FusedPlan plan = plan_split(m, n, MX_ROWS, 2048, 1024);
  const int cps = (plan.cols_per_split + 31) / 32 * 32;
  const int nsplit = (n + cps - 1) / cps;
  const long m_pad = (long)plan.row_blocks * MX_ROWS;
  const long n_pad = ((long)n + 15) / 16 * 16;

  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t off = 0;
  const size_t o_T = off; off += al((size_t)m_pad * 16 * KC * 4);
  const size_t o_Tf = off; off += al((size_t)m_pad * 4);
  const size_t o_Ts = off; off += al((size_t)m_pad * 8);
  const size_t o_Tn = off; off += al((size_t)m_pad * 4);
  const size_t o_S = off; off += al((size_t)n_pad * 16 * KC * 4);
  const size_t o_Sf = off; off += al((size_t)n_pad * 4);
  const size_t o_Ss = off; off += al((size_t)n_pad * 8);
  const size_t o_Sn = off; off += al((size_t)n_pad * 4);
  const size_t o_Sm = off; off += al((size_t)(n_pad / 16) * 4);
  const size_t o_Wt = off; off += al((size_t)n_pad * 16 * NT * 4);
  const size_t o_Wx = off; off += al((size_t)n_pad * 4);
  const size_t o_c = off; off += al(64 * 8);
  const size_t o_cp = off; off += al((size_t)centre_parts((long)m + n) * 64 * 8);
  char *base = nullptr;
  ScratchLease lease(s);
  int rc = lease.aux(off, &base);
  if (rc != KS_OK) return rc;
  hipError_t e = hipSuccess;
  if (!centre) {
    e = launch_centre(p, (double *)(base + o_cp), (double *)(base + o_c), s);
    if (e != hipSuccess) return hip_failure(e, "mixed multi-RHS centre");
    centre = (const double *)(base + o_c);
  }
  const double xs = p.P.scal * kLog2eD;
  f4 *T = (f4 *)(base + o_T), *S = (f4 *)(base + o_S);
  float *Tf = (float *)(base + o_Tf), *Tn = (float *)(base + o_Tn);
  float *Sf = (float *)(base + o_Sf), *Sn = (float *)(base + o_Sn);
  double *Ts = (double *)(base + o_Ts), *Ss = (double *)(base + o_Ss);
  switch (KC) {
    case 1:
      e = mx_pack_points<1>(p.XA, p.amap, m, (int)m_pad, k, centre, xs, T, Tf, Ts, Tn, s);
      if (e == hipSuccess)
        e = mx_pack_points<1>(p.XB, p.bmap, n, (int)n_pad, k, centre, xs, S, Sf, Ss, Sn, s);
      break;
    case 2:
      e = mx_pack_points<2>(p.XA, p.amap, m, (int)m_pad, k, centre, xs, T, Tf, Ts, Tn, s);
      if (e == hipSuccess)
        e = mx_pack_points<2>(p.XB, p.bmap, n, (int)n_pad, k, centre, xs, S, Sf, Ss, Sn, s);
      break;
    case 3:
      e = mx_pack_points<3>(p.XA, p.amap, m, (int)m_pad, k, centre, xs, T, Tf, Ts, Tn, s);
      if (e == hipSuccess)
        e = mx_pack_points<3>(p.XB, p.bmap, n, (int)n_pad, k, centre, xs, S, Sf, Ss, Sn, s);
      break;
    default:
      e = mx_pack_points<4>(p.XA, p.amap, m, (int)m_pad, k, centre, xs, T, Tf, Ts, Tn, s);
      if (e == hipSuccess)
        e = mx_pack_points<4>(p.XB, p.bmap, n, (int)n_pad, k, centre, xs, S, Sf, Ss, Sn, s);
      break;
  }
  if (e != hipSuccess) return hip_failure(e, "mixed multi-RHS pack");
  float *Sm = (float *)(base + o_Sm), *Wt = (float *)(base + o_Wt), *Wx = (float *)(base + o_Wx);
  switch (NT) {
    case 0: e = mx_pack_weights<0>(p, nrhs, (int)n_pad, Sn, Ss, Sm, Wt, Wx, s); break;
    case 1: e = mx_pack_weights<1>(p, nrhs, (int)n_pad, Sn, Ss, Sm, Wt, Wx, s); break;
    default: e = mx_pack_weights<2>(p, nrhs, (int)n_pad, Sn, Ss, Sm, Wt, Wx, s); break;
  }
  if (e != hipSuccess) return hip_failure(e, "mixed multi-RHS weight pack");
  MxArgs A;
  A.p = p;
  A.nrhs = nrhs;
  A.cols_per_split = cps;
  A.nsplit = nsplit;
  A.partial = nullptr;
  A.T = T;
  A.Tf = Tf;
  A.Ts = Ts;
  A.Tn = Tn;
  A.S = S;
  A.Sf = Sf;
  A.Smax = Sm;
  A.Wt = (const f4 *)Wt;
  A.Wx = Wx;
  const double as = fabs(p.P.scal);
  A.gam = (float)(-2.0 * xs);
  A.c1 = (float)(as * (2.0 * (16 * KC + 2) + 6.0) * kU32);
  A.c2 = (float)(as * 2.01 * kU32);
  A.c0 = (float)(4.0 * kU32);
  if (nsplit > 1) {
    rc = lease.partials((size_t)nsplit * m * nrhs, &A.partial);
    if (rc != KS_OK) return rc;
  }
  switch (KC) {
    case 1: e = mx_nt<1>(A, NT, plan.row_blocks, s); break;
    case 2: e = mx_nt<2>(A, NT, plan.row_blocks, s); break;
    case 3: e = mx_nt<3>(A, NT, plan.row_blocks, s); break;
    default: e = mx_nt<4>(A, NT, plan.row_blocks, s); break;
  }
  if (e == hipSuccess && nsplit > 1)
    e = launch_split_reduce_rhs(A.partial, nsplit, m, nrhs, p.u, p.umap, s);
  return e == hipSuccess ? KS_OK : hip_failure(e, "mixed multi-RHS launch");
}

// Translate n rows of X [n][k] by -centre into Xc (FP64).
__global__ void mx_shift_kernel(const double *X, long total, int k, const double *centre,
                                double *Xc) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < total) Xc[t] = X[t] - centre[t % k];
}

size_t centre_work_elems(long rows) { return (size_t)centre_parts(rows) * 64; }

int enqueue_centre(const double *XA, int m, const double *XB, int n, int k, const int *ident,
                   double *work, double *centre, hipStream_t s) {
  Problem q{};
  q.m = m;
  q.n = n;
  q.k = k;
  q.XA = XA;
  q.XB = XB;
  q.amap = ident;
  q.bmap = ident;
  hipError_t e = launch_centre(q, work, centre, s);
  return e == hipSuccess ? KS_OK : hip_failure(e, "centre launch");
}

int enqueue_shift(const double *X, int rows, int k, const double *centre, double *Xc,
                  hipStream_t s) {
  const long total = (long)rows * k;
  hipLaunchKernelGGL(mx_shift_kernel, dim3(blocks_for(total, 256)), dim3(256), 0, s, X, total, k,
                     centre, Xc);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? KS_OK : hip_failure(e, "shift launch");
}

}  // namespace gsks

using namespace gsks;

extern "C" int dgsks_multirhs_mixed_device(const ks_t *kernel, int m, int n, int k, int nrhs,
                                           double *U, const int *umap, const double *XA,
                                           const double *XA2, const int *amap, const double *XB,
                                           const double *XB2, const int *bmap, const double *W,
                                           const int *wmap, void *stream) {
  if (!kernel || m < 0 || n < 0 || k < 0 || nrhs < 1 || nrhs > kMixedMrhsMaxRhs ||
      k > kMixedMrhsMaxK || ker_type(kernel) != KS_GAUSSIAN)
    return raise_error(KS_ERR_INVALID_ARG, false,
                       "dgsks_multirhs_mixed_device(): Gaussian kernel, k <= %d and nrhs <= %d "
                       "expected\n", kMixedMrhsMaxK, kMixedMrhsMaxRhs);
  if (m == 0 || n == 0 || k == 0) return KS_OK;
  if (ensure_device() != KS_OK)
    return raise_error(KS_ERR_DEVICE, false, "dgsks_multirhs_mixed_device(): no GPU available\n");
  Problem p;
  int rc = make_params(kernel, k, p.P, true);
  if (rc != KS_OK) return raise_error(rc, false, "dgsks_multirhs_mixed_device(): %s\n", err_string(rc));
  const void *arrays[] = {U, umap, XA, XA2, amap, XB, XB2, bmap, W, wmap};
  for (const void *a : arrays)
    if (!a || !device_ptr_cached(a))
      return raise_error(KS_ERR_MIXED_POINTERS, false,
                         "dgsks_multirhs_mixed_device(): every array must be device memory\n");
  p.m = m; p.n = n; p.k = k; p.type = GAUSSIAN;
  p.u = U; p.umap = umap; p.XA = XA; p.XA2 = XA2; p.amap = amap;
  p.XB = XB; p.XB2 = XB2; p.bmap = bmap; p.w = W; p.wmap = wmap;
  p.hi = p.hj = nullptr;
  rc = enqueue_multirhs_mixed(p, nrhs, nullptr, (hipStream_t)stream);
  return rc == KS_OK ? KS_OK
                     : raise_error(rc, false, "dgsks_multirhs_mixed_device(): %s\n", err_string(rc));
}
