// Mixed-precision fused summation (Gaussian and variable-bandwidth Gaussian):
//   u[umap[i]] += sum_j exp(x_ij) w[wmap[j]],  x = scal*d^2  or  -0.5 d^2 hi hj.
//
// Every pair that can matter is evaluated; what changes against the FP64
// drop-in is where:
//   1. an estimate of every cross term a.b runs on the BF16 matrix cores as a
//      split-BF16 three-product (a ~= ah + al, b ~= bh + bl, bf16 each;
//      a.b ~= ah.bh + ah.bl + al.bh accumulated in FP32 by
//      v_mfma_f32_16x16x32_bf16, 3 instructions per 32 coordinates of a 16 x 16
//      pair block: KS_PREC_MIXED_BF16X3), or on the FP32 matrix cores
//      (v_mfma_f32_16x16x4f32, 8 per 32 coordinates: KS_PREC_MIXED_FP32);
//   2. BF16x3 Gaussian only: the first product ah.bh (one instruction per 32
//      coordinates) bounds the block before the other two run.  With a
//      rigorous margin for what the split and FP32 accumulation can change
//      (2^-7 |a||b| from the BF16 roundings of both operands plus the FP32
//      sums, using the largest |a| of the wave's rows and each source's |b|),
//      a block whose every exponent is below kSkipT2 (K < 2^-50 for every
//      pair) is skipped: a target loses at most n 2^-50 max|w| in total --
//      below the far tier's own error bound for a single pair (about 1e-4 of
//      K < 1.1e-7);
//   3. a 16 x 16 block whose estimated exponents all lie below -kFarT
//      (K < ~1e-7 for every pair) is "far": its kernel values are taken in FP32
//      (v_exp_f32) from the estimate, whose error is bounded by the estimate's
//      error times K, far below 1e-6 of u;
//   4. any other block is "near": its cross terms are recomputed exactly on the
//      FP64 matrix cores (v_mfma_f64_16x16x4f64 from FP64 coordinates) and its
//      kernel values evaluated in FP64 with FP64 norms, so the |a|^2 + |b|^2 -
//      2 a.b cancellation of close pairs never meets an FP32 rounding.
// u accumulates in FP64 across source blocks (FP32 only within kFlush far
// blocks of one row).
//
// Locality.  Near pairs only fill whole 16 x 16 blocks when close points sit
// next to each other in the block order.  A pre-pass therefore buckets the
// targets and the sources by their nearest of up to kPivots pivot points (FP32
// distances; the pivots are targets spaced evenly through amap) and orders both
// by bucket with a stable radix sort, so that the points of one cluster form
// contiguous runs.  The order only decides which pairs share a block; each
// target's sum is the same set of terms, accumulated in source-block order.
//
// Packing: the pre-pass writes the MFMA operands in sorted order, in blocks of
// 16 points.  BF16 / FP32 operand: block b, chunk c (32 coordinates), part h is
// 64 lanes x 16 B, lane 16 g + p holding point 16 b + p, coordinates 32 c + 8 g
// .. + 7 (BF16x3: part 0 the 8 hi, part 1 the 8 lo bf16; FP32: part h the
// floats 4h .. 4h + 3).  FP64 source operand: block b, chunk c (16
// coordinates) is 64 lanes x 32 B, lane 16 g + p holding point 16 b + p,
// coordinates 16 c + 4 g .. + 3.  One fragment load is then contiguous across
// the wave.  Per-point FP32 and FP64 factors (norms, bandwidths, weights) go
// beside them; points past m / n are zero with a -inf exponent and weight 0,
// so the kernel never clamps an index.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "gsks_kernels.h"
#include "gsks_launch.h"
#include "host_util.h"

namespace gsks {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr float kFarT = 16.0f;          // far: estimated x < -16 (natural-log units)
constexpr float kSkipT2 = -50.0f;       // skipped: every exponent < -50 (log2 units)
constexpr float kLog2e = 1.4426950408889634f;
constexpr int kPivots = 128;            // bucketing pivots (at most)
constexpr int kSortMin = 16384;         // bucket only when m and n are both at least this
constexpr int kXcds = 8;                // L2 domains; workgroups go to them round robin
constexpr int kCUs = 256;               // compute units (32 per L2 domain)
constexpr long kSliceHiBytes = 2l << 20;  // BF16 hi halves of one source slice: 2 MiB
                                        // (its sorts cost ~15 launches per call)
constexpr int kMixRB = 2;               // 16-row blocks per wave

// The skip tier runs for the BF16x3 Gaussian only: with a bandwidth per point,
// one small hi hj in a 16 x 16 block keeps it from being skipped, and the
// two-stage test then only adds work.
template <int PREC, int TYPE>
constexpr bool kSkipTier = PREC == KS_PREC_MIXED_BF16X3 && TYPE == GAUSSIAN;

// Waves per workgroup of the main kernel.  The skip-tier kernel takes one: its
// waves' work differs widely (their rows' clusters in the slice or not), and a
// workgroup holds its CU slot until its slowest wave ends (four waves per
// workgroup: 2.85 ms against 2.35 at mixture_k64).  The others take four: their
// waves walk the same source blocks at the same pace and share them through L1
// (one wave per workgroup made the variable bandwidth 5% slower).
constexpr int mix_waves(int precision, int type) {
  return precision == KS_PREC_MIXED_BF16X3 && type == GAUSSIAN ? 1 : 4;
}
constexpr int mix_rows_per_wg(int precision, int type) {
  return mix_waves(precision, type) * kMixRB * 16;
}

// ------------------------------------------------------------------ pre-pass

// Pruning bounds of the skip tier (mix_bounds_kernel): Vt[T][i] and Vs[i][S]
// per pivot i and 16-point target block T / source block S over the block's
// points in its first bucket, Vt1 and Vs1 over its other points (+inf in a
// block of one bucket), the bucket range (first | last << 16) of every block,
// and the pivot distances pp[A][B] = |p_A - p_B|^2 rounded up.
struct MixBounds {
  const float *Vt, *Vs, *Vt1, *Vs1, *pp;
  const unsigned *rt, *rs;
  int ntb, nsb;  // target and source blocks in the tables
  float thr_d2;  // a block pair whose d^2 lower bound exceeds this is negligible
};

struct MixPack {
  // Which source blocks each wave's row group needs (bucketed BF16x3 Gaussian
  // with a decaying kernel; null: every block): word (g nsplit + split) ngw + gi
  // has bit b set when block 64 gi + b of slice `split` may hold a pair with
  // K >= 2^-50 for row group g (32 sorted targets), mix_need_kernel.
  const unsigned long long *need;
  int ngw;          // mask words per (row group, slice)
  uint4 *Ap;        // BF16 / FP32 target operand [mp/16][nc][2 parts][64 lanes]
  double *Ad;       // FP64 targets in sorted order [mp][32 nc], zero past k and m
  uint4 *Bp;        // BF16 / FP32 source operand [np/16][nc][2 parts][64 lanes]
  double4 *Bd;      // FP64 source operand [np/16][2 nc][64 lanes]
  float2 *ra;       // per target: {-|a|^2 / 2, -hi log2(e) / 2 (VarBW)}
  float4 *cb;       // per source: {xs |b|^2, |b - p_own|, w, |b|} (Gaussian) or {|b|^2, hj, w, |b|}
  float2 *cs;       // per source, the skip test's two factors: {xs |b|^2, |b|} (Gaussian)
  double2 *rd;      // per target: {|a|^2, -hi / 2}
  double4 *cd;      // per source: {|b|^2, w, hj, 0}
  const int *prow;  // sorted target position -> target index of the call
  const int *pcol;  // sorted source position -> source index of the call
  // The centred near tier (bucketed BF16x3 Gaussian with pruning bounds; Ex is
  // null otherwise): the MixTabs tables and operands, and the bucket range of
  // every 16-point target / source block (first | last << 16).
  const float *Ex, *Qx;
  const uint4 *Ac, *Bc;
  const float *na, *nb;
  const unsigned *rt, *rs;
  long mp, np;
  // per slice, its row blocks by descending needed-block count (null: in order)
  const int *order;
};

// Coordinates q0 .. q0 + 7 of row x (zero past k, or all zero for a null row),
// as two 32-byte loads where vec says the rows are 32-byte aligned (k a
// multiple of 4 and the table aligned).
__device__ __forceinline__ void load_run8(const double *x, int k, int q0, bool vec,
                                          double (&v)[8]) {
  if (x && vec && q0 + 8 <= k) {
    const double4 *xv = reinterpret_cast<const double4 *>(x + q0);
    const double4 u0 = xv[0], u1 = xv[1];
    v[0] = u0.x;
    v[1] = u0.y;
    v[2] = u0.z;
    v[3] = u0.w;
    v[4] = u1.x;
    v[5] = u1.y;
    v[6] = u1.z;
    v[7] = u1.w;
  } else {
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = (x && q0 + t < k) ? x[q0 + t] : 0.0;
  }
}
__device__ __forceinline__ bool rows_vec_aligned(const double *tab, int k) {
  return (k & 3) == 0 && ((uintptr_t)tab & 31) == 0;
}

// The coordinate that lane group g supplies at step st of the bounds kernel's
// FP64 dot products: 32 (st / 8) + 8 g + st % 8.  A lane's coordinates are then
// those of its lane in the BF16 split operands (lane 16 g + p holds 32 cc + 8 g
// .. + 7), so the split reuses the values loaded for the products.
__device__ __forceinline__ int bnd_coord(int st, int g) {
  return 32 * (st >> 3) + 8 * g + (st & 7);
}

// FP32 copies of the pivots: pivot i is target (i * m) / P, zero past k; also
// in the B-operand order of v_mfma_f32_16x16x4f32 for mix_bucket_kernel (pivB:
// tile t, step st, lane 16 g + c holding coordinate bnd_coord(st, g) of pivot
// 16 t + c).
// With pivA, also as FP64 in the A-operand order of v_mfma_f64_16x16x4f64 for
// mix_bounds_kernel: tile t (pivots 16 t ..), step st, lane 16 g + c holding
// coordinate bnd_coord(st, g) of pivot 16 t + f64_a_row(c) -- the
// row order that lands pivot 16 t + 4 g + v in accumulator slot v of lane group
// g -- so that one fragment load is 512 contiguous bytes.  f64_a_row is its own
// inverse (a 4 x 4 transpose of the row index).
__global__ void mix_pivot_kernel(const Problem p, int P, int kp, float *piv, float *pivB,
                                 double *pivA) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * kp) return;
  const int i = t / kp, q = t - i * kp;
  const int r = (int)(((long)i * p.m) / P);
  const float v = q < p.k ? (float)p.XA[(long)p.amap[r] * p.k + q] : 0.0f;
  piv[t] = v;
  const int st = 8 * (q >> 5) + (q & 7), g = (q >> 3) & 3;  // q = bnd_coord(st, g)
  pivB[((long)(i >> 4) * (kp >> 2) + st) * 64 + g * 16 + (i & 15)] = v;
  if (pivA) {
    const int c = 4 * ((i & 15) & 3) + ((i & 15) >> 2);
    pivA[((long)(i >> 4) * (kp >> 2) + st) * 64 + g * 16 + c] = (double)v;
  }
}

// Bucket (nearest pivot) of every target, then every source: the pivot with the
// least |p|^2 - 2 x.p in FP32 (|x - p|^2 less the point's own |x|^2), ties to
// the lowest pivot.  One wave per 16 points at a time (targets, then sources,
// in call order): their dot products with the 128 pivots on the FP32 matrix
// cores (v_mfma_f32_16x16x4f32, the points as the A operand, the pivots as B
// from pivB, eight 16-pivot tiles), |p|^2 from the same fragments.  Up to 64
// coordinates a wave keeps the pivot fragments and norms in registers (128
// VGPRs) and takes blocks of 16 points in a grid-stride loop; beyond, it takes
// one block and reads the fragments as it goes.  The buckets only decide which
// points share a block.
template <int KP>
__global__ __launch_bounds__(256) void mix_bucket_kernel(const Problem p,
                                                         const float *__restrict__ pivB,
                                                         unsigned *key_a, unsigned *key_b) {
  constexpr int NS = KP / 4;
  constexpr int NT = kPivots / 16;
  constexpr bool RES = KP <= 64;
  const int lane = threadIdx.x & 63, p16 = lane & 15, g = lane >> 4;
  const long tot = (long)p.m + p.n;
  const long nblk = (tot + 15) / 16;
  const long nw = (long)gridDim.x * 4;
  long blk = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (blk >= nblk) return;  // wave-uniform
  float pb[RES ? NT : 1][RES ? NS : 1], nn[RES ? NT : 1];
  if constexpr (RES) {
#pragma unroll
    for (int tt = 0; tt < NT; ++tt) {
      float n2 = 0.0f;
#pragma unroll
      for (int st = 0; st < NS; ++st) {
        pb[tt][st] = pivB[((long)tt * NS + st) * 64 + lane];
        n2 = fmaf(pb[tt][st], pb[tt][st], n2);
      }
      n2 += __shfl_xor(n2, 16);
      n2 += __shfl_xor(n2, 32);
      nn[tt] = n2;  // |p|^2 of pivot 16 tt + p16
    }
  }
  // the A operand: point blk 16 + p16 supplies coordinate bnd_coord(st, g) at
  // step st, runs of 8 coordinates loaded as 32-byte vectors where the rows
  // are 32-byte aligned
  const bool vec_a = rows_vec_aligned(p.XA, p.k), vec_b = rows_vec_aligned(p.XB, p.k);
  auto load = [&](const long blk, float (&a)[NS]) {
    const long t = blk * 16 + p16;
    const double *x = nullptr;
    bool vec = false;
    if (t < p.m) {
      x = p.XA + (long)p.amap[t] * p.k;
      vec = vec_a;
    } else if (t < tot) {
      x = p.XB + (long)p.bmap[t - p.m] * p.k;
      vec = vec_b;
    }
#pragma unroll
    for (int cc = 0; cc < KP / 32; ++cc) {
      double v[8];
      load_run8(x, p.k, 32 * cc + 8 * g, vec, v);
#pragma unroll
      for (int u = 0; u < 8; ++u) a[8 * cc + u] = (float)v[u];
    }
  };
  auto block = [&](const long blk, const float (&a)[NS]) {
    float best[4];
    int bi[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      best[v] = INFINITY;
      bi[v] = 0;
    }
#pragma unroll
    for (int tt = 0; tt < NT; ++tt) {
      f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
      float n2 = 0.0f;
      if constexpr (RES) {
#pragma unroll
        for (int st = 0; st < NS; ++st)
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[st], pb[tt][st], acc, 0, 0, 0);
        n2 = nn[tt];
      } else {
        const float *pv = pivB + (long)tt * NS * 64 + lane;
#pragma unroll
        for (int st = 0; st < NS; ++st) {
          const float b = pv[st * 64];
          n2 = fmaf(b, b, n2);
          acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[st], b, acc, 0, 0, 0);
        }
        n2 += __shfl_xor(n2, 16);
        n2 += __shfl_xor(n2, 32);  // |p|^2 of pivot 16 tt + p16
      }
      // acc[v]: point 4 g + v, pivot 16 tt + p16
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float d = fmaf(-2.0f, acc[v], n2);
        if (d < best[v]) {
          best[v] = d;
          bi[v] = 16 * tt + p16;
        }
      }
    }
    // the least over the 16 lanes of the point's lane group, ties to the lower pivot
#pragma unroll
    for (int o = 1; o < 16; o <<= 1)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float ob = __shfl_xor(best[v], o);
        const int oi = __shfl_xor(bi[v], o);
        if (ob < best[v] || (ob == best[v] && oi < bi[v])) {
          best[v] = ob;
          bi[v] = oi;
        }
      }
    if (p16 == 0) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long tp = blk * 16 + 4 * g + v;
        if (tp < p.m) key_a[tp] = (unsigned)bi[v];
        else if (tp < tot) key_b[tp - p.m] = (unsigned)bi[v];
      }
    }
  };
  float a[NS];
  if constexpr (RES) {
    for (; blk < nblk; blk += nw) {
      load(blk, a);
      block(blk, a);
    }
  } else {
    load(blk, a);
    block(blk, a);  // one block per wave: the grid covers them all
  }
}

// ---- pruning bounds (skip tier)
//
// For any two points p_A, p_B, a target a and a source b,
//   |a - b|^2 >= (e_B(a) - e_A(a)) + (e_A(b) - e_B(b)) - |p_A - p_B|^2,
// with e_i(x) = |x - p_i|^2: write a - b = D + (a - p_A) - (b - p_B) with
// D = p_A - p_B and drop the square |(a - p_A) - (b - p_B)|^2 >= 0 from the
// expansion.  With the pivots p_i and each point's bucket as A (targets) or B
// (sources), the per-block minima
//   V[i][blk] = min over the block's points x of e_i(x) - e_own(x)
// bound every pair of a target block T (bucket A) and a source block S (bucket
// B) from below by Vt[B][T] + Vs[A][S] - pp[A][B] (blocks that span several
// buckets take the least value over their ranges).  Across clusters the two
// differences are nearly |D|^2 each, so whole runs of far source blocks are
// proven negligible before any of their operands is loaded.

// pp[i][j] = |p_i - p_j|^2 in FP64 from the FP32 pivots, rounded up to FP32,
// and pn[i] = |p_i|^2 (nearest FP32); with ppd, also both in FP64 (ppd[i][j],
// pn64[i]) for the centred near tier's tables.  Workgroup i, thread j.
__global__ void mix_pivot_gram_kernel(int P, int kp, const float *piv, float *pp, float *pn,
                                      double *ppd, double *pn64) {
  const int i = blockIdx.x, j = threadIdx.x;
  if (j >= P) return;
  double t = 0.0;
  for (int q = 0; q < kp; ++q) {
    const double d = (double)piv[i * kp + q] - (double)piv[j * kp + q];  // exact
    t += d * d;
  }
  if (ppd) ppd[i * P + j] = t;
  t *= 1.0 + 0x1p-40;  // covers the FP64 roundings of the sum
  float f = (float)t;
  if ((double)f < t) f = __uint_as_float(__float_as_uint(f) + 1u);  // next float up (f >= 0)
  pp[i * P + j] = f;
  if (j == i) {
    double n2 = 0.0;
    for (int q = 0; q < kp; ++q) n2 += (double)piv[i * kp + q] * (double)piv[i * kp + q];
    pn[i] = (float)n2;
    if (pn64) pn64[i] = n2;
  }
}

// Outputs of mix_bounds_kernel.  Vt, Vs, rt, rs: the pruning bounds above.
// With Ex non-null, also the operands of the mixed kernel's centred near tier:
//   Ex[i][t] = xs e_i(a_t)                         (pivot i, sorted target t)
//   Qx[i][j] = xs (e_i(b_j) - |p_i - p_own(j)|^2)  (pivot i, sorted source j)
//            with e_i(x) = N(x) - 2 x.p_i + |p_i|^2 and N the caller's squared
//            norm (XA2 / XB2), i.e. |x - p_i|^2 when the norms are exact;
//   Ac, Bc   the coordinates centred on the point's own pivot, x - p_own,
//            each split from the FP64 difference into three BF16 parts
//            (h + m + l within 2^-24 of it), in the lane order of the
//            v_mfma_f32_16x16x32_bf16 operands: block b, chunk cc (32
//            coordinates), part, lane 16 g + p holding coordinates
//            32 cc + 8 g .. 32 cc + 8 g + 7 of point 16 b + p (targets and
//            sources alike);
//   na, nb   |x - p_own| rounded up, per sorted target / source.
// xs = scal log2(e) (FP64); the tables are FP64 values rounded once to FP32.
// For target t in bucket A and source j in bucket B, with a = p_A + alpha and
// b = p_B + beta, the reference's d^2 = N(a) + N(b) - 2 a.b is
//   e_B(a) + (N(b) - 2 p_A.beta - |p_B|^2) - 2 alpha.beta
// and N(b) - 2 p_A.beta - |p_B|^2 = e_A(b) - |p_A - p_B|^2, so the pair's
// exponent (log2 units) is Ex[B][t] + Qx[A][j] - 2 xs alpha.beta, where only
// the last term, a product of two small centred vectors, is left to the kernel.
struct MixTabs {
  float *Vt, *Vs, *Vt1, *Vs1;
  unsigned *rt, *rs;
  float *Ex, *Qx;
  uint4 *Ac, *Bc;
  float *na, *nb;
  long mp, np;  // row strides of Ex (padded targets) and Qx (padded sources)
};

// One wave per 16-point block of the sorted targets (blocks [0, ntb)) or
// sources ([ntb, ntb + nsb)): the dot products p_i . x of its 16 points with the
// 128 pivots on the FP64 matrix cores (v_mfma_f64_16x16x4f64, pivots as the A
// operand from pivA, the points as B; eight 16-pivot tiles), then
//   V[blk][i] = min_x (pn_i - pn_own - 2 (x.p_i - x.p_own)) - mu(x),
// the minimum of e_i(x) - e_own(x) less a margin mu(x) far above the FP64
// roundings of the dot products and norms:
//   mu(x) = 2^-36 (|x| + max |p_i|)^2.
// Targets go to Vt[T][i], sources to Vs[i][S].  Points past m / n do not count
// (+inf); rng[blk] = first | last << 16 bucket of the block's points.  With
// o.Ex, also the centred tier's tables (MixTabs), zero for points past m / n.
template <int KP>
__global__ __launch_bounds__(256, KP <= 64 ? 4 : 2) void mix_bounds_kernel(
    const Problem p, int ntb, int nsb, const double *__restrict__ pivA,
    const float *__restrict__ piv, const float *__restrict__ pn, const double *__restrict__ pn64,
    const double *__restrict__ ppd, const unsigned *__restrict__ keys_a,
    const unsigned *__restrict__ keys_b, const int *__restrict__ prow,
    const int *__restrict__ pcol, double xs, MixTabs o) {
  constexpr int NS = KP / 4;  // MFMA steps of 4 coordinates
  constexpr int NT = kPivots / 16;
  const int lane = threadIdx.x & 63;
  const int blk = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (blk >= ntb + nsb) return;  // wave-uniform
  const bool tgt = blk < ntb;
  const int b0 = tgt ? blk : blk - ntb;
  const int cnt = tgt ? p.m : p.n;
  const unsigned *keys = tgt ? keys_a : keys_b;
  const int p16 = lane & 15, g = lane >> 4;
  // the B operand: point p16 supplies coordinate bnd_coord(st, g) at step st
  const int pos = b0 * 16 + p16;
  const bool valid = pos < cnt;
  const double *x = nullptr;
  double nrm = 0.0;  // the caller's squared norm of the point (XA2 / XB2)
  if (valid) {
    const int xi = tgt ? p.amap[prow[pos]] : p.bmap[pcol[pos]];
    x = (tgt ? p.XA : p.XB) + (long)xi * p.k;
    nrm = (tgt ? p.XA2 : p.XB2)[xi];
  }
  const int own = valid ? (int)keys[pos] : 0;
  // the block's first bucket: its points there make part 0 of the bounds, the
  // others part 1
  const int fbk = b0 * 16 < cnt ? (int)keys[b0 * 16] : 0;
  const bool part0 = own == fbk;
  double a[NS];
  // a lane's coordinates come in runs of 8 (bnd_coord): 32-byte loads where
  // the rows are 32-byte aligned (k a multiple of 4, table aligned)
  const bool vec = rows_vec_aligned(tgt ? p.XA : p.XB, p.k);  // wave-uniform
#pragma unroll
  for (int cc = 0; cc < KP / 32; ++cc) {
    double v[8];
    load_run8(valid ? x : nullptr, p.k, 32 * cc + 8 * g, vec, v);
#pragma unroll
    for (int t = 0; t < 8; ++t) a[8 * cc + t] = v[t];
  }
  double n2 = 0.0;
#pragma unroll
  for (int st = 0; st < NS; ++st) n2 = fma(a[st], a[st], n2);
  n2 += __shfl_xor(n2, 16);
  n2 += __shfl_xor(n2, 32);  // |x|^2 of point p16 on its four lanes
  // x.p_own in FP64 from the point's own coordinates (lanes of the point's four
  // coordinate groups summed)
  double dsel = 0.0;
#pragma unroll
  for (int st = 0; st < NS; ++st)
    dsel = fma(a[st], (double)piv[own * KP + bnd_coord(st, g)], dsel);
  dsel += __shfl_xor(dsel, 16);
  dsel += __shfl_xor(dsel, 32);
  float pmx = fmaxf(pn[lane], pn[64 + lane]);
#pragma unroll
  for (int o2 = 1; o2 < 64; o2 <<= 1) pmx = fmaxf(pmx, __shfl_xor(pmx, o2));
  const float pm = sqrtf(pmx) * 1.0001f;  // >= max |p_i| (pn rounded to nearest)
  const float nx = sqrtf((float)n2) * 1.0001f;
  const float mu = 0x1p-36f * (nx + pm) * (nx + pm);
  const double pno = pn64[own];
  const bool tabs = o.Ex != nullptr;
  float *etab = tgt ? o.Ex : o.Qx;
  const long estride = tgt ? o.mp : o.np;
  // the pivot tiles one per pass (few live accumulators: four waves per SIMD
  // up to k = 64; two passes held two waves, four passes three: 0.30 -> 0.25
  // -> 0.23 ms at mixture_k64)
  constexpr int NH = NT / 8;
#pragma unroll 1
  for (int h = 0; h < 8; ++h) {
    d4 acc[NH];
#pragma unroll
    for (int t = 0; t < NH; ++t) {
      acc[t] = d4{0.0, 0.0, 0.0, 0.0};
      const double *pv = pivA + (long)(h * NH + t) * NS * 64 + lane;
#pragma unroll
      for (int st = 0; st < NS; ++st)
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(pv[st * 64], a[st], acc[t], 0, 0, 0);
    }
    // acc[t][v]: pivot 16 (h NH + t) + 4 g + v, point p16
    float best[NH][4], best1[NH][4];
#pragma unroll
    for (int t = 0; t < NH; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * (h * NH + t) + 4 * g + v;
        const double pni = pn64[i];
        const double diff = (pni - pno) - 2.0 * (acc[t][v] - dsel);  // e_i(x) - e_own(x)
        const float bv = valid ? (float)diff - mu : INFINITY;
        best[t][v] = part0 ? bv : INFINITY;
        best1[t][v] = part0 ? INFINITY : bv;
        if (tabs) {
          // e_i(x) = |x - p_i|^2 with the caller's norm in place of |x|^2: the
          // tables then carry the reference's d^2 = XA2 + XB2 - 2 a.b exactly
          // (norms that are not the points' own, e.g. rounded ones, included)
          double e = nrm + pni - 2.0 * acc[t][v];
          if (!tgt) e -= ppd[i * kPivots + own];
          etab[(long)i * estride + pos] = valid ? (float)(xs * e) : 0.0f;
        }
      }
    // block minima over the 16 points, per part
#pragma unroll
    for (int t = 0; t < NH; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        float b = best[t][v], c = best1[t][v];
        b = fminf(b, __shfl_xor(b, 1));
        c = fminf(c, __shfl_xor(c, 1));
        b = fminf(b, __shfl_xor(b, 2));
        c = fminf(c, __shfl_xor(c, 2));
        b = fminf(b, __shfl_xor(b, 4));
        c = fminf(c, __shfl_xor(c, 4));
        b = fminf(b, __shfl_xor(b, 8));
        c = fminf(c, __shfl_xor(c, 8));
        best[t][v] = b;
        best1[t][v] = c;
      }
    if (p16 == 0) {
#pragma unroll
      for (int t = 0; t < NH; ++t) {
        const int i0 = 16 * (h * NH + t) + 4 * g;
        if (tgt) {
          *reinterpret_cast<float4 *>(&o.Vt[(long)b0 * kPivots + i0]) =
              make_float4(best[t][0], best[t][1], best[t][2], best[t][3]);
          *reinterpret_cast<float4 *>(&o.Vt1[(long)b0 * kPivots + i0]) =
              make_float4(best1[t][0], best1[t][1], best1[t][2], best1[t][3]);
        } else {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            o.Vs[(long)(i0 + v) * nsb + b0] = best[t][v];
            o.Vs1[(long)(i0 + v) * nsb + b0] = best1[t][v];
          }
        }
      }
    }
  }
  if (lane == 0) {
    const int last = min(cnt - b0 * 16, 16) - 1;
    const unsigned f = last >= 0 ? keys[b0 * 16] : 0u, l = last >= 0 ? keys[b0 * 16 + last] : 0u;
    (tgt ? o.rt : o.rs)[b0] = f | (l << 16);
  }
  if (tabs) {
    // the centred coordinates x - p_own (exact FP64 differences, rounded once)
    double c2 = 0.0;
#pragma unroll
    for (int st = 0; st < NS; ++st) {
      const double d = valid ? a[st] - (double)piv[own * KP + bnd_coord(st, g)] : 0.0;
      c2 = fma(d, d, c2);
    }
    c2 += __shfl_xor(c2, 16);
    c2 += __shfl_xor(c2, 32);
    // the same differences, split into three BF16 parts (h + m + l within
    // 2^-24 of each), in the lane order of the v_mfma_f32_16x16x32_bf16
    // operands: lane 16 g + p holds coordinates 32 cc + 8 g .. + 7 of point p
    uint4 *dst = (tgt ? o.Ac : o.Bc) + (long)b0 * (KP / 32) * 3 * 64 + lane;
#pragma unroll
    for (int cc = 0; cc < KP / 32; ++cc) {
      bf8 ph, pm, pl;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int q = 32 * cc + 8 * g + t;
        // a[8 cc + t] holds coordinate q (bnd_coord)
        const double d = (valid && q < p.k) ? a[8 * cc + t] - (double)piv[own * KP + q] : 0.0;
        const __bf16 h = (__bf16)(float)d;
        const double r1 = d - (double)(float)h;
        const __bf16 mm = (__bf16)(float)r1;
        const double r2 = r1 - (double)(float)mm;
        ph[t] = h;
        pm[t] = mm;
        pl[t] = (__bf16)(float)r2;
      }
      dst[(cc * 3 + 0) * 64] = *reinterpret_cast<uint4 *>(&ph);
      dst[(cc * 3 + 1) * 64] = *reinterpret_cast<uint4 *>(&pm);
      dst[(cc * 3 + 2) * 64] = *reinterpret_cast<uint4 *>(&pl);
    }
    if (g == 0) (tgt ? o.na : o.nb)[pos] = (float)sqrt(c2) * (1.0f + 0x1p-20f);
  }
}

// One wave per (row group g of 32 sorted targets, source slice): for each
// group gi of 64 source blocks of the slice, bit b of
// need[(g nsplit + split) ngw + gi] is set unless the bound proves every pair of
// source block 64 gi + b negligible for both 16-row blocks of the group:
//   L = min_B Vt[T][B] + min_A Vs[A][S] - max_{A,B} pp[A][B]  > thr_d2
// over the blocks' bucket ranges, less the FP32 roundings of that sum.  Blocks
// spanning more than four buckets, and row blocks spanning more than four, are
// kept.  The row blocks' Vt rows and the pp rows of their buckets stay in
// registers (lane l holds entries l and 64 + l) and are read with lane
// shuffles, so a group's only memory loads are the blocks' bucket ranges and
// their Vs entries, independent of each other.
// The dispatch order of each slice's row blocks: by descending needed-block
// count (work[split][rb], written by mix_need_kernel), equal counts in index
// order, so the heaviest start first.  Up to kRankMaxRB row blocks per slice
// each row block's position is the rank of its key c = count * 2^13 +
// (8191 - rb) among its slice's keys (distinct keys, so the ranks are a
// permutation and the order is that of a stable sort).  One workgroup per 64
// row blocks of a slice, the slice's keys in LDS, each of its kRankWaves waves
// counting over its share of them.
constexpr int kRankMaxRB = 8192;
constexpr int kRankWaves = 16;
__global__ __launch_bounds__(64 * kRankWaves) void mix_rank_kernel(const unsigned *work,
                                                                   int row_blocks, int *order) {
  __shared__ uint4 s_c[kRankMaxRB / 4];
  __shared__ int s_r[kRankWaves][64];
  unsigned *sc = (unsigned *)s_c;
  const int split = blockIdx.y;
  const unsigned *wk = work + (long)split * row_blocks;
  const int nq = (row_blocks + 3) / 4;
  for (int j = threadIdx.x; j < nq * 4; j += 64 * kRankWaves)
    sc[j] = j < row_blocks ? (min(wk[j], 0x7FFFFu) << 13) | (unsigned)(kRankMaxRB - 1 - j) : 0u;
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int rb = blockIdx.x * 64 + lane;
  const unsigned mine = rb < row_blocks ? sc[rb] : 0xFFFFFFFFu;
  const int q0 = wv * nq / kRankWaves, q1 = (wv + 1) * nq / kRankWaves;
  int r0 = 0, r1 = 0, r2 = 0, r3 = 0;  // padding keys (0) never exceed mine
#pragma unroll 4
  for (int q = q0; q < q1; ++q) {
    const uint4 c = s_c[q];
    r0 += c.x > mine;
    r1 += c.y > mine;
    r2 += c.z > mine;
    r3 += c.w > mine;
  }
  s_r[wv][lane] = (r0 + r1) + (r2 + r3);
  __syncthreads();
  if (wv == 0 && rb < row_blocks) {
    int r = 0;
#pragma unroll
    for (int v = 0; v < kRankWaves; ++v) r += s_r[v][lane];
    order[(long)split * row_blocks + r] = rb;
  }
}

// More row blocks per slice: radix-sort keys in place of the counts -- slice in
// the high bits, then the complement of the count, so an ascending sort puts
// every slice's heaviest row blocks first -- and the row blocks as values.
__global__ __launch_bounds__(256) void mix_order_kernel(unsigned *work_key, int nsplit,
                                                        int row_blocks, int *val) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= nsplit * row_blocks) return;
  const int split = t / row_blocks;
  const unsigned work = work_key[t];
  work_key[t] = ((unsigned)split << 20) | (0xFFFFFu - (work < 0xFFFFFu ? work : 0xFFFFFu));
  val[t] = t - split * row_blocks;
}

__global__ __launch_bounds__(256) void mix_need_kernel(const MixBounds b, int ngrp, int nsplit,
                                                       int cps, int ngw, int n,
                                                       unsigned long long *need, unsigned *work) {
  const int lane = threadIdx.x & 63;
  const long w = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= (long)ngrp * nsplit) return;  // wave-uniform
  const int split = (int)(w % nsplit), g = (int)(w / nsplit);
  const int j0 = split * cps, j1 = min(n, j0 + cps);
  const int nbs = j0 < j1 ? (j1 - j0 + 15) >> 4 : 0;
  unsigned long long *out = need + w * ngw;
  // Each block is taken as two parts: its points in its first bucket (bucket
  // range [f, f], bounds V) and the others ([f + 1, l], bounds V1).  For parts
  // with bucket ranges RA (targets) and RB (sources), every pair obeys
  //   d^2 >= min_{B in RB} V[T][B] + min_{A in RA} V[S][A] - max_{A, B} pp[A][B]
  // (the pivot-hyperplane bound of each pair's own pivots); a block pair is
  // kept when the bound of any of its part pairs fails to prove the cutoff.
  // The two row blocks' rows of V (with the pp row of their first bucket) and
  // V1, in this wave's LDS; entry kPivots repeats entry kPivots - 1, so entry
  // b0 + 1 is the first bucket of a source block's part 1 (b0 + 1, or the last
  // bucket when b0 is the last).  One wave writes and reads its own rows: its
  // LDS operations complete in order.
  __shared__ float2 s_tq[4][2][kPivots + 2];
  __shared__ float s_t1[4][2][kPivots + 2];
  const int wv = threadIdx.x >> 6;
  int a0[2], a1[2];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) {
    const int T = 2 * g + rb;
    const unsigned rtv = b.rt[T];
    a0[rb] = (int)(rtv & 0xFFFFu);
    a1[rb] = (int)(rtv >> 16);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int B = h * 64 + lane;
      const float2 v = make_float2(b.Vt[(long)T * kPivots + B], b.pp[a0[rb] * kPivots + B]);
      const float v1 = b.Vt1[(long)T * kPivots + B];
      s_tq[wv][rb][B] = v;
      s_t1[wv][rb][B] = v1;
      if (B == kPivots - 1) {
        s_tq[wv][rb][kPivots] = v;
        s_t1[wv][rb][kPivots] = v1;
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  unsigned cnt = 0;
  for (int gi = 0; gi < ngw; ++gi) {
    const int sb = gi * 64 + lane;
    const bool in = sb < nbs;
    const int S = min((j0 >> 4) + (in ? sb : 0), b.nsb - 1);  // in range for an empty slice
    const unsigned rsv = b.rs[S];
    const int b0 = (int)(rsv & 0xFFFFu), b1 = (int)(rsv >> 16);
    // the source block's two bound columns at the row blocks' first buckets
    // (V1 is +inf for a block of one bucket), every load issued before the
    // first use
    float vs0[2], vs1[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      vs0[rb] = b.Vs[(long)a0[rb] * b.nsb + S];
      vs1[rb] = b.Vs1[(long)a0[rb] * b.nsb + S];
    }
    // one part pair: false where its bound proves every pair negligible
    auto part_keeps = [&](float tmin, float smin, float ppmax) {
      if (tmin == INFINITY || smin == INFINITY) return false;  // a part with no points
      const float L = (tmin + smin) - ppmax;
      const float err = 0x1p-22f * (fabsf(tmin) + fabsf(smin) + ppmax);
      return !(L - err > b.thr_d2);  // NaN keeps the block
    };
    bool nd = false;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      bool keep = (b1 - b0 >= 4) | (a1[rb] - a0[rb] >= 4);
      // the row block's first bucket against the source's first bucket, and
      // against its other buckets
      const float2 q0 = s_tq[wv][rb][b0], q1 = s_tq[wv][rb][b0 + 1];
      keep = keep | part_keeps(q0.x, vs0[rb], q0.y);
      // (past b0 + 3 the block is kept anyway)
      float tmin = q1.x, ppmax = q1.y;
      const int bl = min(b1, b0 + 3);
      for (int B = b0 + 2; B <= bl; ++B) {  // a source block over three buckets or more: rare
        const float2 q = s_tq[wv][rb][B];
        tmin = fminf(tmin, q.x);
        ppmax = fmaxf(ppmax, q.y);
      }
      keep = keep | part_keeps(tmin, vs1[rb], ppmax);
      if (a1[rb] > a0[rb] && a1[rb] - a0[rb] < 4) {  // wave-uniform: a row block over buckets (rare)
        // the row block's other buckets, with the bounds of every source bucket
        for (int pb = 0; pb < 2; ++pb) {
          if (pb == 1 && b1 == b0) break;  // a source block of one bucket
          const float *vs = pb ? b.Vs1 : b.Vs;
          const int blo = pb ? b0 + 1 : b0, bhi = pb ? bl : b0;
          float tm = s_t1[wv][rb][blo], smin = INFINITY, ppm = 0.0f;
          for (int B = blo + 1; B <= bhi; ++B) tm = fminf(tm, s_t1[wv][rb][B]);
          for (int A = a0[rb] + 1; A <= a1[rb]; ++A) {
            smin = fminf(smin, vs[(long)A * b.nsb + S]);
            for (int B = blo; B <= bhi; ++B) ppm = fmaxf(ppm, b.pp[A * kPivots + B]);
          }
          keep = keep | part_keeps(tm, smin, ppm);
        }
      }
      nd = nd | keep;
    }
    const unsigned long long msk = __ballot(in && nd);
    if (lane == 0) out[gi] = msk;
    cnt += (unsigned)__popcll(msk);
  }
  // the row group's needed-block count in this slice (one wave per workgroup
  // in the skip-tier kernel: the group is its row block)
  if (lane == 0) work[(long)split * ngrp + g] = cnt;
}

// ---- stable counting sort of the bucket keys (keys < kPivots = 2^kKeyBits),
// the targets' and the sources' in the same three launches: per tile of
// kSortTile keys a histogram, per segment the tiles' first positions of every
// key, and per tile one wave placing its keys in order.
constexpr int kSortTile = 1024;
constexpr int kKeyBits = 7;
static_assert(kPivots <= (1 << kKeyBits), "bucket keys take kKeyBits bits");

__global__ __launch_bounds__(256) void mix_hist_kernel(const unsigned *__restrict__ ka, long m,
                                                       const unsigned *__restrict__ kb, long n,
                                                       int tiles_a, unsigned *hist) {
  __shared__ unsigned h[kPivots];
  const int tile = blockIdx.x;
  const bool seg_a = tile < tiles_a;
  const unsigned *k = seg_a ? ka : kb;
  const long cnt = seg_a ? m : n;
  const long t0 = (long)(seg_a ? tile : tile - tiles_a) * kSortTile;
  for (int i = threadIdx.x; i < kPivots; i += 256) h[i] = 0u;
  __syncthreads();
  for (int i = threadIdx.x; i < kSortTile; i += 256)
    if (t0 + i < cnt) atomicAdd(&h[k[t0 + i] & (kPivots - 1)], 1u);
  __syncthreads();
  for (int i = threadIdx.x; i < kPivots; i += 256) hist[(long)tile * kPivots + i] = h[i];
}

// hist[tile][key] becomes the position of the tile's first key `key` in its
// segment's sorted order: the keys below it in the segment, then the same key
// in earlier tiles.  Workgroup 0: targets, 1: sources; thread (part, key) sums
// and then rewrites an eighth of the segment's tiles for its key.
__global__ __launch_bounds__(1024) void mix_scan_kernel(unsigned *hist, int tiles_a,
                                                        int tiles_b) {
  constexpr int kParts = 1024 / kPivots;
  __shared__ unsigned part[kParts][kPivots];
  __shared__ unsigned start[kPivots];
  const int t0 = blockIdx.x ? tiles_a : 0, t1 = blockIdx.x ? tiles_a + tiles_b : tiles_a;
  const int per = (t1 - t0 + kParts - 1) / kParts;
  const int key = threadIdx.x % kPivots, pp = threadIdx.x / kPivots;
  const int a = t0 + pp * per, b = min(t1, a + per);
  unsigned sum = 0u;
  for (int t = a; t < b; ++t) sum += hist[(long)t * kPivots + key];
  part[pp][key] = sum;
  __syncthreads();
  if (pp == 0) {  // the key's total, then (below) the keys before it
    unsigned tot = 0u;
    for (int q = 0; q < kParts; ++q) tot += part[q][key];
    start[key] = tot;
  }
  __syncthreads();
  unsigned run = 0u;
  for (int i = 0; i < key; ++i) run += start[i];
  for (int q = 0; q < pp; ++q) run += part[q][key];
  __syncthreads();  // every thread has read start[] before the tiles change
  for (int t = a; t < b; ++t) {
    const unsigned c = hist[(long)t * kPivots + key];
    hist[(long)t * kPivots + key] = run;
    run += c;
  }
}

// One wave per tile, 64 keys at a time in order: a key's position is its
// tile's running position for that key plus its rank among the equal keys of
// the round (from kKeyBits ballots), so equal keys keep their order.  Writes
// the sorted keys and the call index of each.
__global__ __launch_bounds__(64) void mix_place_kernel(const unsigned *__restrict__ ka, long m,
                                                       const unsigned *__restrict__ kb, long n,
                                                       int tiles_a, const unsigned *__restrict__ off,
                                                       unsigned *outa, int *vala, unsigned *outb,
                                                       int *valb) {
  __shared__ unsigned base[kPivots];
  const int tile = blockIdx.x, lane = threadIdx.x;
  const bool seg_a = tile < tiles_a;
  const unsigned *k = seg_a ? ka : kb;
  const long cnt = seg_a ? m : n;
  unsigned *out = seg_a ? outa : outb;
  int *val = seg_a ? vala : valb;
  const long t0 = (long)(seg_a ? tile : tile - tiles_a) * kSortTile;
  for (int i = lane; i < kPivots; i += 64) base[i] = off[(long)tile * kPivots + i];
  __syncthreads();
  const unsigned long long before = (1ull << lane) - 1ull;
  for (int r = 0; r < kSortTile / 64; ++r) {
    const long e = t0 + r * 64 + lane;
    const bool v = e < cnt;
    const unsigned key = v ? (k[e] & (kPivots - 1)) : 0u;
    unsigned long long same = __ballot(v);
#pragma unroll
    for (int bit = 0; bit < kKeyBits; ++bit) {
      const bool set = (key >> bit) & 1u;
      const unsigned long long bb = __ballot(set);
      same &= set ? bb : ~bb;
    }
    if (v) {
      const unsigned pos = base[key] + (unsigned)__popcll(same & before);
      out[pos] = key;
      val[pos] = (int)e;
    }
    // the last lane of each group of equal keys advances that key's position
    // (one wave: its LDS operations complete in order)
    if (v && (same >> lane) == 1ull) base[key] += (unsigned)__popcll(same);
  }
}

__global__ void mix_identity2_kernel(int *a, long m, int *b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) a[i] = (int)i;
  if (i < n) b[i] = (int)i;
}

// Writes point `pt` (coordinates x, or zeros when x is null) into the blocked
// BF16 / FP32 layout, and with Bd non-null into the FP64 source layout too.
template <int PREC>
__device__ __forceinline__ void pack_chunk(const double *__restrict__ x, int k, int nc, long pt,
                                           int c, int g, bool vec, uint4 *pack) {
  uint4 *o0 = pack + (pt >> 4) * nc * 128 + (c * 2) * 64 + g * 16 + (int)(pt & 15);
  uint4 *o1 = o0 + 64;
  double v[8];
  load_run8(x, k, c * 32 + 8 * g, vec, v);
  if constexpr (PREC == KS_PREC_MIXED_BF16X3) {
    bf8 hi, lo;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const __bf16 h = (__bf16)(float)v[t];
      hi[t] = h;
      lo[t] = (__bf16)(float)(v[t] - (double)(float)h);
    }
    *o0 = *reinterpret_cast<uint4 *>(&hi);
    *o1 = *reinterpret_cast<uint4 *>(&lo);
  } else {
    float f[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) f[t] = (float)v[t];
    *o0 = *reinterpret_cast<uint4 *>(&f[0]);
    *o1 = *reinterpret_cast<uint4 *>(&f[4]);
  }
}

// The FP64 source operand's 16-coordinate chunks 2c and 2c + 1, lane group g:
// coordinates 16 c2 + 4 g .. + 3 of point pt.
__device__ __forceinline__ void pack_f64_chunk(const double *__restrict__ x, int k, int nc, long pt,
                                               int c, int g, double4 *Bd) {
  double4 *dblk = Bd + (pt >> 4) * (2 * nc) * 64 + g * 16 + (int)(pt & 15);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int c2 = 2 * c + h;
    double v[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int idx = c2 * 16 + 4 * g + t;
      v[t] = (x && idx < k) ? x[idx] : 0.0;
    }
    dblk[c2 * 64] = make_double4(v[0], v[1], v[2], v[3]);
  }
}

// The FP64 targets in sorted order, rows of kp = 32 nc doubles (zero past k and
// for padding rows): one thread per pair of coordinates, so that loads of a
// row and stores of the copy are contiguous across lanes.
__global__ void mix_rows_kernel(const Problem p, int kp, long mp, const int *__restrict__ prow,
                                double2 *Ad) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int h = kp / 2;
  if (e >= mp * h) return;
  const long t = e / h;
  const int q = 2 * (int)(e - t * h);
  double2 v = make_double2(0.0, 0.0);
  if (t < p.m) {
    const double *x = p.XA + (long)p.amap[prow[t]] * p.k;
    if (q < p.k) v.x = x[q];
    if (q + 1 < p.k) v.y = x[q + 1];
  }
  Ad[e] = v;
}

// One thread per point of the padded target range [0, mp) then source range [0, np),
// both in sorted order.
template <int PREC, int TYPE>
__global__ void mixed_pack_kernel(const Problem p, int nc, float xs, long mp, long np, MixPack pk) {
  // one thread per (point, 32-coordinate chunk c, lane group g): eight threads
  // read a row of k = 64 together; the (0, 0) thread writes the point's factors
  const int per = 4 * nc;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (mp + np) * per) return;
  const long u = e / per;
  const int cg = (int)(e - u * per), c = cg >> 2, g = cg & 3;
  if (u < mp) {
    const long t = u;
    if (t < p.m) {
      const int ai = p.amap[pk.prow[t]];
      pack_chunk<PREC>(p.XA + (long)ai * p.k, p.k, nc, t, c, g, rows_vec_aligned(p.XA, p.k), pk.Ap);
      if (cg == 0) {
        const double aa = p.XA2[ai];
        const double hi = (TYPE == VAR_BANDWIDTH) ? p.hi[ai] : 0.0;
        pk.ra[t] = make_float2((float)(-0.5 * aa), (float)(hi * -0.5 * (double)kLog2e));
        pk.rd[t] = make_double2(aa, -0.5 * hi);
      }
    } else {  // padding target: its estimated exponent is -inf, so it is never near
      pack_chunk<PREC>(nullptr, p.k, nc, t, c, g, false, pk.Ap);
      if (cg == 0) {
        // Gaussian x2 = -2 xs acc + ...: acc = -inf gives -inf for xs < 0 (a
        // decaying kernel); for xs > 0 the row starts at +inf instead
        const float pad = (TYPE == GAUSSIAN && xs > 0.0f) ? INFINITY : -INFINITY;
        pk.ra[t] = make_float2(pad, -1.0f);
        pk.rd[t] = make_double2(0.0, 0.0);
      }
    }
  } else {
    const long j = u - mp;
    if (j < p.n) {
      const int jj = pk.pcol[j];
      const int bj = p.bmap[jj];
      const double *x = p.XB + (long)bj * p.k;
      pack_chunk<PREC>(x, p.k, nc, j, c, g, rows_vec_aligned(p.XB, p.k), pk.Bp);
      pack_f64_chunk(x, p.k, nc, j, c, g, pk.Bd);
      if (cg == 0) {
        const double bb = p.XB2[bj];
        const double w = p.w[p.wmap[jj]];
        const double hj = (TYPE == VAR_BANDWIDTH) ? p.hj[bj] : 0.0;
        const float nb = (float)sqrt(bb);
        // Gaussian: {xs |b|^2, |b - p_own| (centred near tier), w, |b|}
        if (TYPE == GAUSSIAN)
          pk.cb[j] = make_float4((float)(bb * (double)xs), pk.nb ? pk.nb[j] : 0.0f, (float)w, nb);
        else pk.cb[j] = make_float4((float)bb, (float)hj, (float)w, nb);
        pk.cs[j] = make_float2((float)(bb * (double)xs), nb);
        pk.cd[j] = make_double4(bb, w, hj, 0.0);
      }
    } else {  // padding source: exponent -inf, weight 0
      pack_chunk<PREC>(nullptr, p.k, nc, j, c, g, false, pk.Bp);
      pack_f64_chunk(nullptr, p.k, nc, j, c, g, pk.Bd);
      if (cg == 0) {
        pk.cb[j] = (TYPE == GAUSSIAN) ? make_float4(-INFINITY, 0.0f, 0.0f, 0.0f)
                                      : make_float4(INFINITY, 1.0f, 0.0f, 0.0f);
        pk.cs[j] = make_float2(-INFINITY, 0.0f);
        pk.cd[j] = make_double4(0.0, 0.0, 1.0, 0.0);
      }
    }
  }
}

// ------------------------------------------------------------------ main kernel

template <int PREC, int NC>
struct Frag {
  uint4 v[NC][2];  // per chunk: 32 B (bf16: hi | lo, fp32: floats 0-3 | 4-7)
};

// The fragment of one 16-point block: each load is 64 lanes x 16 B contiguous.
template <int PREC, int NC>
__device__ __forceinline__ void load_frag_packed(const uint4 *__restrict__ blk, int lane,
                                                 Frag<PREC, NC> &f) {
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    f.v[c][0] = blk[(c * 2) * 64 + lane];
    f.v[c][1] = blk[(c * 2 + 1) * 64 + lane];
  }
}

template <int PREC, int NC>
__device__ __forceinline__ f4 cross_block(const Frag<PREC, NC> &a, const Frag<PREC, NC> &b, f4 acc) {
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if constexpr (PREC == KS_PREC_MIXED_BF16X3) {
      bf8 ah = *reinterpret_cast<const bf8 *>(&a.v[c][0]);
      bf8 al = *reinterpret_cast<const bf8 *>(&a.v[c][1]);
      bf8 bh = *reinterpret_cast<const bf8 *>(&b.v[c][0]);
      bf8 bl = *reinterpret_cast<const bf8 *>(&b.v[c][1]);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
    } else {
      const float *af = reinterpret_cast<const float *>(&a.v[c][0]);
      const float *bf = reinterpret_cast<const float *>(&b.v[c][0]);
#pragma unroll
      for (int t = 0; t < 8; ++t) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(af[t], bf[t], acc, 0, 0, 0);
    }
  }
  return acc;
}

// Row of the 16 x 16 result block held in accumulator slot v by lane group g.
// Measured on gfx950: v_mfma_f32_16x16x32_bf16 and v_mfma_f32_16x16x4f32 put
// rows 4g + v of column lane%16 in lane g*16 + col; v_mfma_f64_16x16x4f64 puts
// A-operand row g + 4v there instead, so the FP64 A operand of lane p carries
// target row 4 (p % 4) + p / 4, which lands FP64 results on the same rows.
__device__ __forceinline__ int out_row(int g, int v) { return 4 * g + v; }
__device__ __forceinline__ int f64_a_row(int p16) { return 4 * (p16 & 3) + (p16 >> 2); }

constexpr int kFlush = 8;  // far-pair FP32 partials go to FP64 every 8 source blocks


template <int PREC, int TYPE, int NC>
__global__ __launch_bounds__(64 * mix_waves(PREC, TYPE), NC <= 2 ? 2 : 1) void mixed_kernel(FusedArgs A, MixPack pk,
                                                                  float xs) {
  constexpr int RB = kMixRB;
  constexpr int KC = 2 * NC;  // 16-coordinate chunks of the FP64 operand
  const Problem &p = A.p;
  const int m = p.m;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int p16 = lane & 15;
  const int g = lane >> 4;
  // XCD-aware order: consecutive workgroups go to the XCDs in turn, so XCD x
  // runs workgroups x, x + 8, ...; give it the source slices x, x + 8, ... and
  // let it sweep every row block of one slice before the next, so that the
  // workgroups sharing its L2 read the same slice (nsplit is a multiple of 8).
  // Within a slice the row blocks go heaviest first (pk.order, by their
  // needed-block counts): the few row blocks whose clusters lie in the slice
  // hold most of its work, and started last they would set the launch's end.
  constexpr int WAVES = mix_waves(PREC, TYPE);
  constexpr int RPW = mix_rows_per_wg(PREC, TYPE);
  const int row_blocks = (m + RPW - 1) / RPW;
  const int xcd = (int)(blockIdx.x % kXcds), loc = (int)(blockIdx.x / kXcds);
  const int split = xcd + kXcds * (loc / row_blocks);
  const int rblk = pk.order ? pk.order[split * row_blocks + loc % row_blocks] : loc % row_blocks;
  constexpr int WROWS = RB * 16;
  const int wg_rows = WAVES * WROWS;
  const int row0 = rblk * wg_rows + wave * WROWS;  // sorted target position

  // The skip-tier kernel (BF16x3 Gaussian) keeps the wave's targets centred
  // on their pivots in FP32 (the A operand of its centred near tier, in the
  // MixTabs lane order); the other kernels keep them in FP64 (the A operand of
  // their near blocks), row stride 32 NC + 2 doubles so the 16 rows of one
  // FP64 fragment load fall on different banks; zero past k and for padding
  // rows.
  constexpr bool SK = kSkipTier<PREC, TYPE>;
  constexpr int KS = 32 * NC + 2;
  __shared__ __align__(16) double s_rows[WAVES][SK ? 2 : WROWS * KS];
  __shared__ __align__(16) uint4 s_ac[WAVES][SK ? RB * NC * 3 * 64 : 1];
  __shared__ double s_tot[WAVES][WROWS];
  // FP64 row factors {|a|^2, -hi/2}, read by near blocks only
  __shared__ double2 s_rd[WAVES][WROWS];
  double *arows = s_rows[wave];
  if constexpr (SK) {
    if (pk.Ex) {
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int c = 0; c < NC * 3; ++c)
          s_ac[wave][(rb * NC * 3 + c) * 64 + lane] =
              pk.Ac[((long)((row0 >> 4) + rb) * NC * 3 + c) * 64 + lane];
    }
  } else {
    // the wave's rows are contiguous in the sorted FP64 copy: every load is
    // independent and issued before the first store (a gather through prow
    // and amap made each iteration wait for three dependent round trips)
    constexpr int PER = WROWS * 32 * NC / 128;  // 16-byte loads per lane
    const double2 *src = reinterpret_cast<const double2 *>(pk.Ad + (long)row0 * 32 * NC);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = 2 * (i * 64 + lane);
      const int rl = e / (32 * NC), q = e - rl * (32 * NC);
      *reinterpret_cast<double2 *>(&arows[rl * KS + q]) = src[i * 64 + lane];
    }
  }

  // BF16 / FP32 accumulators start at -|a|^2/2 (the packed row factor), so the
  // block result is c - |a|^2/2 and the exponent is one FMA away: Gaussian
  // x2 = -2 xs acc + xs |b|^2, variable bandwidth d^2 = -2 acc + |b|^2.
  Frag<PREC, NC> af[RB];
  f4 cinit[RB];
  float ra_h[RB][4], facc[RB][4];
  double racc[RB][4];
  float mA = 0.0f;  // largest |a| of the lane's rows
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    load_frag_packed<PREC, NC>(pk.Ap + (long)((row0 >> 4) + rb) * NC * 128, lane, af[rb]);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int t = row0 + rb * 16 + out_row(g, v);
      const float2 f = pk.ra[t];
      const double2 d = pk.rd[t];
      cinit[rb][v] = f.x;
      ra_h[rb][v] = f.y;
      if (p16 == 0) s_rd[wave][rb * 16 + out_row(g, v)] = d;
      mA = fmaxf(mA, (float)sqrt(d.x));
      racc[rb][v] = 0.0;
      facc[rb][v] = 0.0f;
    }
  }
  __syncthreads();  // the staged rows are visible to every lane of the wave
  const float m2xs = -2.0f * xs;

  // Skip-test margin (Gaussian BF16x3).  After the first product,
  // acc = -|a|^2/2 + ah.bh in FP32.  With |ah - a| <= 2^-8 |a| per coordinate
  // (and the same for b), |ah.bh - a.b| <= (2^-7 + 2^-16) |a||b|; the FP32 sums
  // of kp + 1 terms add at most (kp + 3) 2^-24 (|a|^2/2 + |a||b|).  With A the
  // largest |a| of the wave's rows and B = |b| of the lane's source, the
  // exponent x2 = -2 xs acc + xs |b|^2 is off by at most
  //   |2 xs| (2^-7 1.01 A B + (kp + 3) 2^-23 (A^2/2 + A B)) + 2^-23 |xs| B^2
  // (doubled rounding terms cover the FP32 norms and the final FMA); the test
  // also shrinks the exponent by 2^-18 of itself for the FP32 roundings that
  // scale with it (xs, the FMA).
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) mA = fmaxf(mA, __shfl_xor(mA, o));
  mA *= 1.0001f;  // the FP32 rounding of sqrt
  constexpr float kAcc = (32.0f * NC + 3.0f) * 0x1p-23f;
  const float mk1 = fabsf(2.0f * xs) * (0x1p-7f * 1.01f * mA + kAcc * mA);  // times B
  const float mk0 = fabsf(2.0f * xs) * (kAcc * mA * mA * 0.5f);           // constant
  const float mk2 = fabsf(xs) * 0x1p-23f;                                  // times B^2
  constexpr float kShrink = 1.0f - 0x1p-18f;
  const float thr = -kFarT * kLog2e;
  const KerParams P = p.P;
  const int j0 = split * A.cols_per_split;
  const int j1 = min(p.n, j0 + A.cols_per_split);

  // FP64 source fragment of a block being (or about to be) recomputed exactly
  struct F64Frag {
    double bd[KC][4];
    double4 cdv;
    int block;  // the source block it holds (or has in flight), -1 for none
  };
  auto fetch64 = [&](F64Frag &f, int jb) {
    const double4 *src = pk.Bd + (long)(jb >> 4) * KC * 64 + lane;
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      const double4 q = src[c * 64];
      f.bd[c][0] = q.x;
      f.bd[c][1] = q.y;
      f.bd[c][2] = q.z;
      f.bd[c][3] = q.w;
    }
    f.cdv = pk.cd[jb + p16];
    f.block = jb;
  };

  // ---- the centred near tier (skip-tier kernel with pk.Ex)
  // A block pair (T in bucket A, S in bucket B, one bucket each) takes its
  // exponents as x2 = fl(Ex[B][t] + Qx[A][j]) + m2xs c (one FMA), with c the
  // cross term alpha.beta of the centred coordinates on the BF16 matrix cores:
  // with alpha and beta in three BF16 parts each, the products h.h in one FP32
  // accumulator (2 v_mfma_f32_16x16x32_bf16 at k = 64) and h.m, m.h, h.l, m.m,
  // l.h in another (10), so the large terms' sums see no small ones; then one
  // v_exp_f32 per pair.  With u = 2^-24, the exponent's error (log2 units) is
  // at most
  //   u (2 (|Ex| + |Qx|) + (kp + 10) |m2xs| A' |beta|),
  // A' the largest |alpha| of the lane's four rows: the roundings of Ex and Qx
  // (half an ulp each, FP64 values), of their sum and of the FMA (|x2| <=
  // |Ex| + |Qx| + |m2xs c|), of m2xs and of c0 + c1, and the cross term's: the
  // splits (2u), the dropped products m.l, l.m, l.l (2u), the FP32 sums of the
  // kp exact products h.h (kp u) and of the 5 kp small ones (together at most
  // 2^-7 A' |beta|: 5 kp 2^-7 u <= 5u for kp <= 128), with
  // |c| <= A' |beta| (1 + (kp + 10) u).  The bound is checked before the
  // products; a block where any lane's bound exceeds 2^-15 -- or whose target
  // or source block spans buckets -- takes the split-BF16 estimate and the FP64
  // near tier instead, so every kernel value of this tier is within
  // ln2 2^-15 + 2^-22 (v_exp_f32) relative.
  struct C32Frag {
    uint4 b[NC * 3];  // the three BF16 parts of each 32-coordinate chunk
    float4 cb;    // the lane's source record {xs |b|^2, |b - p_own|, w, |b|}
    float q[RB];  // Qx[A][j] of the lane's source for each row block's bucket A
    int block;    // the source block it holds (or has in flight), -1 for none
    int bucket;   // its single bucket, -1 if it spans several
  };
  int abk[RB];      // each row block's single bucket, -1 if it spans several
  float aA[RB];  // largest |alpha| of the lane's four rows in each row block
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) aA[rb] = 0.0f;
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) abk[rb] = -1;
  // Ex[B][rows of each row block] for the last source bucket B seen (runs of
  // source blocks share a bucket once the points are sorted), and the largest
  // |Ex| of the lane's rows
  float4 ecache[RB];
  float emax[RB];
  int ebucket = -1;
  auto fetch32 = [&](C32Frag &f, int jb) {
    const uint4 *src = pk.Bc + (long)(jb >> 4) * NC * 3 * 64 + lane;
#pragma unroll
    for (int c = 0; c < NC * 3; ++c) f.b[c] = src[c * 64];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
      f.q[rb] = abk[rb] >= 0 ? pk.Qx[(long)abk[rb] * pk.np + jb + p16] : 0.0f;
    f.cb = pk.cb[jb + p16];
    const unsigned r = pk.rs[jb >> 4];
    f.bucket = (r & 0xFFFFu) == (r >> 16) ? (int)(r & 0xFFFFu) : -1;
    f.block = jb;
    if (f.bucket >= 0 && f.bucket != ebucket) {
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
        const float4 e = *reinterpret_cast<const float4 *>(pk.Ex + (long)f.bucket * pk.mp + row0 +
                                                           rb * 16 + 4 * g);
        ecache[rb] = e;
        emax[rb] = fmaxf(fmaxf(fabsf(e.x), fabsf(e.y)), fmaxf(fabsf(e.z), fabsf(e.w)));
      }
      ebucket = f.bucket;
    }
  };

  if constexpr (SK) {
    if (pk.Ex) {
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
        const unsigned r = pk.rt[(row0 >> 4) + rb];
        abk[rb] = (r & 0xFFFFu) == (r >> 16) ? (int)(r & 0xFFFFu) : -1;
      }
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
        const float4 na4 = *reinterpret_cast<const float4 *>(pk.na + row0 + rb * 16 + 4 * g);
        aA[rb] = fmaxf(fmaxf(na4.x, na4.y), fmaxf(na4.z, na4.w));
      }
    }
  }
  const float kCross = (32.0f * NC + 10.0f) * fabsf(m2xs);  // times |alpha| |beta|
  constexpr float kBnd = 0x1p-24f * 1.01f;
  // Row block rb against the source block of cf, every pair (near or far):
  // false (nothing added) when the bound hands the block back.
  auto near32 = [&](int rb, const C32Frag &cf) -> bool {
    const float4 &cb = cf.cb;
    const float q = cf.q[rb];
    const float bnd = kBnd * (2.0f * (emax[rb] + fabsf(q)) + kCross * aA[rb] * cb.y);
    if (__ballot(bnd > 0x1p-15f) != 0ull) return false;  // wave-uniform
    // alpha.beta from the BF16 parts: h.h in one accumulator, the five next
    // products (h.m, m.h, h.l, m.m, l.h; at most 2^-7 |alpha||beta| together)
    // in another, so the large terms' FP32 sums see no small ones
    f4 c0 = {0.0f, 0.0f, 0.0f, 0.0f}, c1 = {0.0f, 0.0f, 0.0f, 0.0f};
    const uint4 *sa = &s_ac[wave][rb * NC * 3 * 64 + lane];
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) {
      const uint4 ah = sa[(cc * 3) * 64], am = sa[(cc * 3 + 1) * 64], al = sa[(cc * 3 + 2) * 64];
      const uint4 &bh = cf.b[cc * 3], &bm = cf.b[cc * 3 + 1], &bl = cf.b[cc * 3 + 2];
      auto bfv = [](const uint4 &u) { return *reinterpret_cast<const bf8 *>(&u); };
      c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfv(ah), bfv(bh), c0, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfv(ah), bfv(bm), c1, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfv(am), bfv(bh), c1, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfv(ah), bfv(bl), c1, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfv(am), bfv(bm), c1, 0, 0, 0);
      c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfv(al), bfv(bh), c1, 0, 0, 0);
    }
    const float4 e = ecache[rb];  // cf.bucket == ebucket
    const float ev[4] = {e.x, e.y, e.z, e.w};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const float x2 = fmaf(m2xs, c0[v] + c1[v], ev[v] + q);
      // d^2 clamped at 0
      facc[rb][v] = fmaf(__builtin_amdgcn_exp2f(fminf(x2, 0.0f)), cb.z, facc[rb][v]);
    }
    return true;
  };

  // source blocks start on multiples of 16 and a slice ends on one or at n, so
  // a block's columns past j1 are padding of the packed arrays.  The skip tier
  // loads the hi halves ahead of time and the lo halves (fetch_lo) only for
  // blocks that pass its test.
  auto fetch = [&](int jb, Frag<PREC, NC> &f, float4 &c4) {
    const uint4 *blk = pk.Bp + (long)(jb >> 4) * NC * 128 + lane;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      f.v[c][0] = blk[(c * 2) * 64];
      f.v[c][1] = blk[(c * 2 + 1) * 64];
    }
    c4 = pk.cb[jb + p16];
  };
  uint4 blo[NC];     // lo halves of the source block lo_block (BF16x3)
  float4 bcb;        // and its sources' records {xs |b|^2, 0, w, |b|}
  int lo_block = -1;
  auto fetch_lo = [&](int jb) {
    const uint4 *blk = pk.Bp + (long)(jb >> 4) * NC * 128 + lane;
#pragma unroll
    for (int c = 0; c < NC; ++c) blo[c] = blk[(c * 2 + 1) * 64];
    bcb = pk.cb[jb + p16];
    lo_block = jb;
  };

  int nblk = 0;  // source blocks since the last flush of the FP32 partials
  // The skip-tier kernel's FP64 fallback (blocks the centred tier hands back):
  // two 16-coordinate chunks of the FP64 source fragment in registers at a
  // time (their loads in flight together; one at a time waited on four round
  // trips per block).
  auto near64_lean = [&](int rb, int jb) {
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    const double *ar = pk.Ad + (long)(row0 + rb * 16 + f64_a_row(p16)) * (32 * NC) + 4 * g;
    const double4 *src = pk.Bd + (long)(jb >> 4) * KC * 64 + lane;
#pragma unroll 2
    for (int cc = 0; cc < KC; ++cc) {
      const double4 bq = src[cc * 64];
      const double2 a01 = *reinterpret_cast<const double2 *>(ar + cc * 16);
      const double2 a23 = *reinterpret_cast<const double2 *>(ar + cc * 16 + 2);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a01.x, bq.x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a01.y, bq.y, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a23.x, bq.z, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a23.y, bq.w, acc, 0, 0, 0);
    }
    const double4 cdv = pk.cd[jb + p16];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double2 rd = s_rd[wave][rb * 16 + out_row(g, v)];
      const double K = kval_f64<TYPE>(acc[v], rd.x, cdv.x, rd.y, cdv.z, P);
      racc[rb][v] = fma(K, cdv.y, racc[rb][v]);
    }
  };
  // Near block, exactly: FP64 cross terms on the matrix cores, FP64 kernel
  // values (the skip-tier kernel reads its FP64 rows from the sorted copy).
  auto near64 = [&](int rb, const F64Frag &f64) {
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    const double *ar = SK ? pk.Ad + (long)(row0 + rb * 16 + f64_a_row(p16)) * (32 * NC) + 4 * g
                          : arows + (rb * 16 + f64_a_row(p16)) * KS + 4 * g;
#pragma unroll
    for (int cc = 0; cc < KC; ++cc) {
      const double2 a01 = *reinterpret_cast<const double2 *>(ar + cc * 16);
      const double2 a23 = *reinterpret_cast<const double2 *>(ar + cc * 16 + 2);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a01.x, f64.bd[cc][0], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a01.y, f64.bd[cc][1], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a23.x, f64.bd[cc][2], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a23.y, f64.bd[cc][3], acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double2 rd = s_rd[wave][rb * 16 + out_row(g, v)];
      const double K = kval_f64<TYPE>(acc[v], rd.x, f64.cdv.x, rd.y, f64.cdv.z, P);
      racc[rb][v] = fma(K, f64.cdv.y, racc[rb][v]);
    }
  };
  // Far / near evaluation of row block rb against source block jb, given its
  // full FP32 cross-term estimate c (= a.b - |a|^2/2).  The skip-tier kernel
  // keeps the centred fragment of a block across its row blocks (c32) and
  // reads FP64 only where the centred tier hands a block back; the others keep
  // the FP64 fragment across row blocks and source blocks (f64).
  auto evaluate = [&](int rb, int jb, const f4 &c, const float4 &cb, F64Frag &f64, C32Frag &c32,
                      bool &used64) {
    float x2[4];
    bool nr = false;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if constexpr (TYPE == GAUSSIAN) {
        x2[v] = fmaf(m2xs, c[v], cb.x);
      } else {
        x2[v] = fmaxf(fmaf(-2.0f, c[v], cb.x), 0.0f) * (ra_h[rb][v] * cb.y);
      }
      nr = nr || (x2[v] > thr);
    }
    if (__ballot(nr) == 0ull) {
      // far block: FP32 kernel values from the estimate
#pragma unroll
      for (int v = 0; v < 4; ++v) facc[rb][v] = fmaf(__builtin_amdgcn_exp2f(x2[v]), cb.z, facc[rb][v]);
    } else if constexpr (SK) {
      // near block the centred tier handed back (or had no tables for): FP64
      near64_lean(rb, jb);
    } else {
      if (f64.block != jb) fetch64(f64, jb);
      used64 = true;
      near64(rb, f64);
    }
  };
  // FP32 far partials go to the FP64 row sums every kFlush source blocks
  auto tick = [&](int nb) {
    nblk += nb;
    if (nblk >= kFlush) {
      nblk = 0;
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          racc[rb][v] += (double)facc[rb][v];
          facc[rb][v] = 0.0f;
        }
    }
  };
  const int jlast = j0 + ((j1 - j0 - 1) >> 4) * 16;  // start of the slice's last block

  if constexpr (kSkipTier<PREC, TYPE>) {
    // ---- which source blocks of the slice can matter (pruning bounds)
    // Bit b of mask word gi marks source block 64 gi + b of the slice as one
    // that may hold a pair with K >= 2^-50 for the wave's rows
    // (mix_need_kernel); every other block is proven negligible and never
    // loaded.  Lane gi holds word gi; without masks every block is kept.
    const int nbs = j0 < j1 ? (j1 - j0 + 15) >> 4 : 0;  // source blocks of the slice
    const int ngroups = (nbs + 63) >> 6;
    unsigned long long myword = 0ull;
    if (lane < ngroups) {
      if (pk.need) {
        const long g = row0 >> 5;  // the wave's row group
        myword = pk.need[(g * A.nsplit + split) * pk.ngw + lane];
      } else {
        const int left = nbs - 64 * lane;
        myword = left >= 64 ? ~0ull : ((1ull << left) - 1ull);
      }
    }
    // the needed blocks in order, as a wave-uniform generator
    auto read_mask = [&](int gi) {
      const unsigned lo = __builtin_amdgcn_readlane((unsigned)myword, gi);
      const unsigned hi = __builtin_amdgcn_readlane((unsigned)(myword >> 32), gi);
      return ((unsigned long long)hi << 32) | lo;
    };
    int gcur = 0;
    unsigned long long mcur = ngroups > 0 ? read_mask(0) : 0ull;
    auto next_block = [&]() -> int {
      while (mcur == 0ull) {
        if (++gcur >= ngroups) return -1;
        mcur = read_mask(gcur);
      }
      const int b = __builtin_ctzll(mcur);
      mcur &= mcur - 1ull;
      return j0 + ((gcur << 6) + b) * 16;
    };

    // Two source blocks x RB row blocks per step: their 2 RB first products
    // are independent and issue back to back, their skip tests are four
    // instructions each, and one ballot lets the usual all-skipped step
    // through.  The hi halves of the next two needed blocks load meanwhile.
    // a step loads the hi halves and the skip test's two factors per source
    // (8 bytes, not the 16-byte record with the weight, which only blocks that
    // pass need: it comes with their lo halves)
    struct HB {
      uint4 h[NC];
      float2 cs;
      int jb;  // the source block held, -1 past the last needed one
    };
    auto fetch_h = [&](HB &b) {
      b.jb = next_block();
      if (b.jb < 0) return;  // wave-uniform
      const uint4 *blk = pk.Bp + (long)(b.jb >> 4) * NC * 128 + lane;
#pragma unroll
      for (int c = 0; c < NC; ++c) b.h[c] = blk[(c * 2) * 64];
      b.cs = pk.cs[b.jb + p16];
    };
    auto est = [&](int rb, const HB &b) {
      f4 e = cinit[rb];
#pragma unroll
      for (int cc = 0; cc < NC; ++cc)
        e = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf8 *>(&af[rb].v[cc][0]),
                                                    *reinterpret_cast<const bf8 *>(&b.h[cc]), e, 0,
                                                    0, 0);
      return e;
    };
    // For a decaying kernel (xs < 0) the largest exponent of a lane's four
    // pairs belongs to its largest accumulator: x2 = -2 xs acc + xs |b|^2.
    const bool skip_on = xs < 0.0f;
    const float m2s = m2xs * kShrink;
    auto tval = [&](const HB &b) {
      return fmaf(b.cs.x, kShrink, fmaf(fmaf(mk2, b.cs.y, mk1), b.cs.y, mk0));
    };
    auto hot = [&](const f4 &e, float t) {
      const float mx = fmaxf(fmaxf(e[0], e[1]), fmaxf(e[2], e[3]));
      return !skip_on || fmaf(m2s, mx, t) > kSkipT2;
    };
    // the row blocks of one source block that passed: the two correction
    // products complete the split-BF16 cross term, then far / near
    // centred fragment of the current block, or of the next needed one: after a
    // block that took the centred tier, the next needed block's loads start
    // (passing blocks come in runs once the points are sorted)
    C32Frag c32;
    c32.block = -1;
    auto finish = [&](int jb, int jnext, const HB &b, const f4 *e, const bool *h) {
      bool used64 = false, usedlo = false, used32 = false;
      F64Frag f64;  // unused in the skip-tier kernel (its FP64 fallback loads its own)
      f64.block = -1;
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
        if (!h[rb]) continue;  // wave-uniform
        // every pair of the block in the centred FP32 tier where it applies
        if (pk.Ex && abk[rb] >= 0) {
          if (c32.block != jb) fetch32(c32, jb);
          if (c32.bucket >= 0 && near32(rb, c32)) {
            used32 = true;
            continue;
          }
        }
        // else the split-BF16 estimate, then far (FP32) or near (FP64)
        if (lo_block != jb) fetch_lo(jb);
        usedlo = true;
        f4 c = e[rb];
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
          const bf8 ah = *reinterpret_cast<const bf8 *>(&af[rb].v[cc][0]);
          const bf8 al = *reinterpret_cast<const bf8 *>(&af[rb].v[cc][1]);
          const bf8 bh = *reinterpret_cast<const bf8 *>(&b.h[cc]);
          const bf8 bl = *reinterpret_cast<const bf8 *>(&blo[cc]);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
        }
        evaluate(rb, jb, c, bcb, f64, c32, used64);
      }
      // blocks that pass come in runs once the points are bucketed: after one,
      // start the next needed block's loads
      if (jnext >= 0 && jnext != jb) {
        if (used32) fetch32(c32, jnext);
        if (usedlo) fetch_lo(jnext);
      }
    };
    // one step over the pair (b0, b1); false once the needed blocks are done
    auto step = [&](const HB &b0, const HB &b1, int jafter) {
      if (b0.jb < 0) return false;
      const bool two = b1.jb >= 0;
      f4 e0[RB], e1[RB];
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) e0[rb] = est(rb, b0);
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) e1[rb] = est(rb, b1);
      const float t0 = tval(b0), t1 = tval(b1);
      bool h0[RB], h1[RB], any = false;
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
        h0[rb] = hot(e0[rb], t0);
        h1[rb] = two && hot(e1[rb], t1);
        any = any || h0[rb] || h1[rb];
      }
      if (__ballot(any) != 0ull) {
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
          h0[rb] = __ballot(h0[rb]) != 0ull;
          h1[rb] = __ballot(h1[rb]) != 0ull;
        }
        finish(b0.jb, two ? b1.jb : jafter, b0, e0, h0);
        if (two) finish(b1.jb, jafter, b1, e1, h1);
      }
      tick(two ? 2 : 1);
      return two;
    };
    // two pairs of buffers in turn, one pair ahead; a buffer past the last
    // needed block holds jb = -1 and is neither loaded nor evaluated
    HB pa0, pa1, pb0, pb1;
    auto fetch_pair = [&](HB &x0, HB &x1) {
      fetch_h(x0);
      fetch_h(x1);
    };
    fetch_pair(pa0, pa1);
    for (;;) {
      fetch_pair(pb0, pb1);
      if (!step(pa0, pa1, pb0.jb)) break;
      fetch_pair(pa0, pa1);
      if (!step(pb0, pb1, pa0.jb)) break;
    }
  } else {
    F64Frag f64p;  // persists across blocks: prefetched after a near block
    f64p.block = -1;
    C32Frag c32u;  // unused: the centred tier belongs to the skip-tier kernel
    c32u.block = -1;
    // One source block against the wave's RB row blocks.
    auto block = [&](int jb, int jnext, const Frag<PREC, NC> &bf, const float4 &cb) {
      bool used64 = false;
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
        evaluate(rb, jb, cross_block<PREC, NC>(af[rb], bf, cinit[rb]), cb, f64p, c32u, used64);
      // near blocks come in runs once the points are bucketed: after one,
      // start the next block's FP64 fragment loads
      if (used64 && jnext != jb) fetch64(f64p, jnext);
      tick(1);
    };
    // two fragment buffers in turn: the next block's loads are in flight while
    // this block computes, with no register copies between them (the last
    // block's "next" is itself again, an in-bounds re-read)
    Frag<PREC, NC> b0, b1;
    float4 c0, c1;
    if (j0 < j1) {
      fetch(j0, b0, c0);
      for (int jb = j0;; jb += 32) {
        const int jn1 = min(jb + 16, jlast), jn2 = min(jb + 32, jlast);
        fetch(jn1, b1, c1);
        block(jb, jn1, b0, c0);
        if (jb + 16 >= j1) break;
        fetch(jn2, b0, c0);
        block(jb + 16, jn2, b1, c1);
        if (jb + 32 >= j1) break;
      }
    }
  }

  // ---- row totals: partials summed over the 16 column lanes, one writer lane per row
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double s = row_sum16(racc[rb][v] + (double)facc[rb][v]);
      if (p16 == 0) s_tot[wave][rb * 16 + out_row(g, v)] = s;
    }
  __syncthreads();
  const bool writer = lane < WROWS;
  const int myrow = row0 + (writer ? lane : 0);
  const double mine = writer ? s_tot[wave][lane] : 0.0;
  if (A.nsplit == 1) {
    if (writer && myrow < m) p.u[p.umap[pk.prow[myrow]]] += mine;
    return;
  }
  // the slices' partial sums are added in slice order by mix_reduce_kernel,
  // launched next: a fence and a counter here would make every workgroup
  // write back and invalidate its XCD's L2 (device-scope release / acquire on
  // gfx950), evicting the source slices the other workgroups are reading
  if (writer && myrow < m) A.partial[(long)split * m + myrow] = mine;
}

// ------------------------------------------------------------------ launch

template <int PREC, int TYPE, int NC>
static hipError_t mixed_go(const FusedArgs &A, const MixPack &pk, float xs, int row_blocks,
                           hipStream_t s) {
  dim3 grid((unsigned)row_blocks * (unsigned)A.nsplit);
  hipLaunchKernelGGL((mixed_kernel<PREC, TYPE, NC>), grid, dim3(64 * mix_waves(PREC, TYPE)), 0, s,
                     A, pk, xs);
  if (A.nsplit > 1)
    return launch_split_reduce(A.partial, A.nsplit, A.p.m, A.p.u, A.p.umap, pk.prow, s);
  return hipGetLastError();
}

template <int PREC, int TYPE>
static hipError_t mixed_nc(const FusedArgs &A, const MixPack &pk, float xs, int nc,
                           int row_blocks, hipStream_t s) {
  switch (nc) {
    case 1: return mixed_go<PREC, TYPE, 1>(A, pk, xs, row_blocks, s);
    case 2: return mixed_go<PREC, TYPE, 2>(A, pk, xs, row_blocks, s);
    case 3: return mixed_go<PREC, TYPE, 3>(A, pk, xs, row_blocks, s);
    default: return mixed_go<PREC, TYPE, 4>(A, pk, xs, row_blocks, s);
  }
}

template <int PREC>
static hipError_t mixed_launch(const Problem &p, const FusedPlan &plan, double *partial,
                               const MixPack &pk, int nc, hipStream_t s) {
  FusedArgs A;
  A.p = p;
  A.cols_per_split = plan.cols_per_split;
  A.nsplit = plan.nsplit;
  A.partial = partial;
  A.vec2 = 0;
  const float xs = (float)(p.P.scal * (double)kLog2e);
  const long mp = (long)plan.row_blocks * mix_rows_per_wg(PREC, p.type),
             np = ((long)p.n + 15) / 16 * 16;
  dim3 pg((unsigned)(((mp + np) * 4 * nc + 255) / 256));
  const int kp = 32 * nc;
  hipLaunchKernelGGL(mix_rows_kernel, dim3((unsigned)((mp * (kp / 2) + 255) / 256)), dim3(256), 0,
                     s, p, kp, mp, pk.prow, reinterpret_cast<double2 *>(pk.Ad));
  if (p.type == GAUSSIAN) {
    hipLaunchKernelGGL((mixed_pack_kernel<PREC, GAUSSIAN>), pg, dim3(256), 0, s, p, nc, xs, mp, np,
                       pk);
    return mixed_nc<PREC, GAUSSIAN>(A, pk, xs, nc, plan.row_blocks, s);
  }
  hipLaunchKernelGGL((mixed_pack_kernel<PREC, VAR_BANDWIDTH>), pg, dim3(256), 0, s, p, nc, xs, mp,
                     np, pk);
  return mixed_nc<PREC, VAR_BANDWIDTH>(A, pk, xs, nc, plan.row_blocks, s);
}

template <int KP>
static void launch_bounds(const Problem &p, int ntb, int nsb, const double *pivA, const float *piv,
                          const float *pn, const double *pn64, const double *ppd,
                          const unsigned *ka, const unsigned *kb, const int *prow,
                          const int *pcol, double xs, const MixTabs &o, hipStream_t s) {
  const unsigned wgs = (unsigned)(((long)ntb + nsb + 3) / 4);
  hipLaunchKernelGGL((mix_bounds_kernel<KP>), dim3(wgs), dim3(256), 0, s, p, ntb, nsb, pivA, piv,
                     pn, pn64, ppd, ka, kb, prow, pcol, xs, o);
}

template <int KP>
static void launch_bucket(const Problem &p, const float *pivB, unsigned *ka, unsigned *kb,
                          hipStream_t s) {
  const long blocks = ((long)p.m + p.n + 15) / 16;
  // with the pivots resident (k <= 64), two workgroups per CU in a grid-stride loop
  long wgs = (blocks + 3) / 4;
  if (KP <= 64 && wgs > 2 * kCUs) wgs = 2 * kCUs;
  hipLaunchKernelGGL((mix_bucket_kernel<KP>), dim3((unsigned)wgs), dim3(256), 0, s, p, pivB, ka,
                     kb);
}

int enqueue_mixed(const Problem &p, int precision, hipStream_t s) {
  if (p.type != GAUSSIAN && p.type != VAR_BANDWIDTH)
    return raise_error(KS_ERR_INVALID_ARG, false,
                       "dgsks_device(): mixed precision supports the Gaussian and "
                       "variable-bandwidth kernels only\n");
  if (precision != KS_PREC_MIXED_BF16X3 && precision != KS_PREC_MIXED_FP32)
    return raise_error(KS_ERR_INVALID_ARG, false, "dgsks_device(): unknown precision\n");
  const int nc = (p.k + 31) / 32;
  if (nc > 4)
    return raise_error(KS_ERR_INVALID_ARG, false,
                       "dgsks_device(): mixed precision supports k <= 128\n");
  FusedPlan plan = plan_split(p.m, p.n, mix_rows_per_wg(precision, p.type), 2048, 128);
  // Source slices: a multiple of the 8 XCDs (the kernel gives XCD x the slices
  // x, x + 8, ...), each with at most kSliceHiBytes of packed hi halves so that
  // one slice stays in an XCD's L2, and at least as many as plan_split wants
  // for parallelism.  Slices start on 16-point blocks of the packed layout;
  // trailing ones may be empty.
  {
    const long max_src = kSliceHiBytes / (64l * nc);
    long ns = kXcds * ((p.n + kXcds * max_src - 1) / (kXcds * max_src));
    const long want = (plan.nsplit + kXcds - 1) / kXcds * kXcds;
    if (ns < want) ns = want;
    plan.nsplit = (int)ns;
    plan.cols_per_split = (int)(((p.n + ns - 1) / ns + 15) / 16 * 16);
  }
  const bool bucket = p.m >= kSortMin && p.n >= kSortMin;
  const int P = bucket ? kPivots : 0;
  const int kp = 32 * nc;

  // scratch: partial sums | packed operands and factors (rows padded to whole
  // workgroups, sources to 16-point blocks) | orders | bucket keys | sort space
  const size_t mp = (size_t)plan.row_blocks * mix_rows_per_wg(precision, p.type),
               np = ((size_t)p.n + 15) / 16 * 16;
  const size_t mx = (size_t)(p.m > p.n ? p.m : p.n);
  // sort space for the dispatch order (radix sort of nsplit row_blocks 32-bit keys)
  size_t sort_bytes = 0;
  if (bucket && plan.row_blocks > kRankMaxRB)
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const unsigned *)nullptr,
                                             (unsigned *)nullptr, (const int *)nullptr,
                                             (int *)nullptr, plan.nsplit * plan.row_blocks, 0,
                                             32, s);
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t off = 0;
  const size_t o_Ap = off; off += al(mp * nc * 128);
  const size_t o_Ad = off; off += al(mp * nc * 32 * sizeof(double));
  const size_t o_Bp = off; off += al(np * nc * 128);
  const size_t o_Bd = off; off += al(np * nc * 256);
  const size_t o_ra = off; off += al(mp * sizeof(float2));
  const size_t o_cb = off; off += al(np * sizeof(float4));
  const size_t o_cs = off; off += al(np * sizeof(float2));
  const size_t o_rd = off; off += al(mp * sizeof(double2));
  const size_t o_cd = off; off += al(np * sizeof(double4));
  const size_t o_prow = off; off += al((size_t)p.m * sizeof(int));
  const size_t o_pcol = off; off += al((size_t)p.n * sizeof(int));

  const size_t o_ka = off; off += bucket ? al((size_t)p.m * sizeof(unsigned)) : 0;
  const size_t o_kb = off; off += bucket ? al((size_t)p.n * sizeof(unsigned)) : 0;
  const size_t o_kouta = off; off += bucket ? al((size_t)p.m * sizeof(unsigned)) : 0;
  const size_t o_koutb = off; off += bucket ? al((size_t)p.n * sizeof(unsigned)) : 0;
  const size_t o_piv = off; off += bucket ? al((size_t)P * kp * sizeof(float)) : 0;
  const size_t o_pivf = off; off += bucket ? al((size_t)P * kp * sizeof(float)) : 0;
  // pruning bounds of the skip tier (bucketed BF16x3 Gaussian with a decaying kernel)
  const bool bounds = bucket && precision == KS_PREC_MIXED_BF16X3 && p.type == GAUSSIAN &&
                      p.P.scal < 0.0;
  const size_t ntb = mp / 16, nsb = np / 16;
  const size_t ngrp = mp / 32;  // row groups of one wave
  const size_t ngw = ((size_t)plan.cols_per_split / 16 + 63) / 64;
  const size_t o_pivb = off; off += bounds ? al((size_t)P * kp * sizeof(double)) : 0;
  const size_t o_pp = off; off += bounds ? al((size_t)P * P * sizeof(float)) : 0;
  const size_t o_pn = off; off += bounds ? al((size_t)P * sizeof(float)) : 0;
  const size_t o_ppd = off; off += bounds ? al((size_t)P * P * sizeof(double)) : 0;
  const size_t o_pn64 = off; off += bounds ? al((size_t)P * sizeof(double)) : 0;
  // the centred near tier's tables and operands (MixTabs)
  const size_t o_ex = off; off += bounds ? al((size_t)P * mp * sizeof(float)) : 0;
  const size_t o_qx = off; off += bounds ? al((size_t)P * np * sizeof(float)) : 0;
  const size_t o_ac = off; off += bounds ? al(mp * kp * 3 * sizeof(__bf16)) : 0;
  const size_t o_bc = off; off += bounds ? al(np * kp * 3 * sizeof(__bf16)) : 0;
  const size_t o_na = off; off += bounds ? al(mp * sizeof(float)) : 0;
  const size_t o_nb = off; off += bounds ? al(np * sizeof(float)) : 0;
  const size_t o_vt = off; off += bounds ? al((size_t)P * ntb * sizeof(float)) : 0;
  const size_t o_vs = off; off += bounds ? al((size_t)P * nsb * sizeof(float)) : 0;
  const size_t o_vt1 = off; off += bounds ? al((size_t)P * ntb * sizeof(float)) : 0;
  const size_t o_vs1 = off; off += bounds ? al((size_t)P * nsb * sizeof(float)) : 0;
  const size_t o_rt = off; off += bounds ? al(ntb * sizeof(unsigned)) : 0;
  const size_t o_rs = off; off += bounds ? al(nsb * sizeof(unsigned)) : 0;
  const size_t o_need = off;
  off += bounds ? al(ngrp * plan.nsplit * ngw * sizeof(unsigned long long)) : 0;
  // dispatch order of the row blocks of each slice: keys and values, twice
  const size_t nord = bounds ? (size_t)plan.nsplit * plan.row_blocks : 0;
  const size_t o_ord = off; off += bounds ? al(4 * nord * sizeof(int)) : 0;
  const int tiles_a = (int)((p.m + kSortTile - 1) / kSortTile);
  const int tiles_b = (int)((p.n + kSortTile - 1) / kSortTile);
  const size_t o_hist = off; off += bucket ? al((size_t)(tiles_a + tiles_b) * kPivots * 4) : 0;
  const size_t o_sort = off; off += al(sort_bytes);
  const size_t npart = (size_t)(plan.nsplit > 1 ? plan.nsplit : 0) * p.m;
  double *partial = nullptr;
  ScratchLease lease(s);
  int rc = lease.partials(npart + (off + 255) / 8 + 32, &partial);
  if (rc != KS_OK) return rc;
  char *base = reinterpret_cast<char *>(partial + npart);
  base = (char *)(((uintptr_t)base + 255) & ~(uintptr_t)255);
  MixPack pk;
  pk.Ap = (uint4 *)(base + o_Ap);
  pk.Ad = (double *)(base + o_Ad);
  pk.Bp = (uint4 *)(base + o_Bp);
  pk.Bd = (double4 *)(base + o_Bd);
  pk.ra = (float2 *)(base + o_ra);
  pk.cb = (float4 *)(base + o_cb);
  pk.cs = (float2 *)(base + o_cs);
  pk.rd = (double2 *)(base + o_rd);
  pk.cd = (double4 *)(base + o_cd);
  int *prow = (int *)(base + o_prow);
  int *pcol = (int *)(base + o_pcol);
  pk.prow = prow;
  pk.pcol = pcol;
  pk.need = nullptr;
  pk.ngw = (int)ngw;
  pk.Ex = pk.Qx = nullptr;
  pk.Ac = pk.Bc = nullptr;
  pk.na = pk.nb = nullptr;
  pk.rt = pk.rs = nullptr;
  pk.mp = (long)mp;
  pk.np = (long)np;
  pk.order = nullptr;

  // ---- target and source orders
  if (!bucket) {
    hipLaunchKernelGGL(mix_identity2_kernel, dim3((unsigned)((mx + 255) / 256)), dim3(256), 0, s,
                       prow, (long)p.m, pcol, (long)p.n);
  } else {
    unsigned *ka = (unsigned *)(base + o_ka), *kb = (unsigned *)(base + o_kb);
    unsigned *kouta = (unsigned *)(base + o_kouta), *koutb = (unsigned *)(base + o_koutb);
    float *piv = (float *)(base + o_piv);
    double *pivA = bounds ? (double *)(base + o_pivb) : nullptr;
    float *pivB = (float *)(base + o_pivf);
    hipLaunchKernelGGL(mix_pivot_kernel, dim3((unsigned)((P * kp + 255) / 256)), dim3(256), 0, s, p,
                       P, kp, piv, pivB, pivA);
    switch (nc) {
      case 1: launch_bucket<32>(p, pivB, ka, kb, s); break;
      case 2: launch_bucket<64>(p, pivB, ka, kb, s); break;
      case 3: launch_bucket<96>(p, pivB, ka, kb, s); break;
      default: launch_bucket<128>(p, pivB, ka, kb, s); break;
    }
    // targets and sources ordered by bucket, stably (a counting sort: keys < 128)
    unsigned *hist = (unsigned *)(base + o_hist);
    hipLaunchKernelGGL(mix_hist_kernel, dim3((unsigned)(tiles_a + tiles_b)), dim3(256), 0, s,
                       (const unsigned *)ka, (long)p.m, (const unsigned *)kb, (long)p.n, tiles_a,
                       hist);
    hipLaunchKernelGGL(mix_scan_kernel, dim3(2), dim3(1024), 0, s, hist, tiles_a, tiles_b);
    hipLaunchKernelGGL(mix_place_kernel, dim3((unsigned)(tiles_a + tiles_b)), dim3(64), 0, s,
                       (const unsigned *)ka, (long)p.m, (const unsigned *)kb, (long)p.n, tiles_a,
                       (const unsigned *)hist, kouta, prow, koutb, pcol);
    if (bounds) {
      if (P != kPivots) return KS_ERR_INVALID_ARG;  // the bound kernels take all kPivots tiles
      float *pp = (float *)(base + o_pp), *pn = (float *)(base + o_pn);
      double *ppd = (double *)(base + o_ppd), *pn64 = (double *)(base + o_pn64);
      MixTabs tb;
      tb.Vt = (float *)(base + o_vt);
      tb.Vs = (float *)(base + o_vs);
      tb.Vt1 = (float *)(base + o_vt1);
      tb.Vs1 = (float *)(base + o_vs1);
      tb.rt = (unsigned *)(base + o_rt);
      tb.rs = (unsigned *)(base + o_rs);
      tb.Ex = (float *)(base + o_ex);
      tb.Qx = (float *)(base + o_qx);
      tb.Ac = (uint4 *)(base + o_ac);
      tb.Bc = (uint4 *)(base + o_bc);
      tb.na = (float *)(base + o_na);
      tb.nb = (float *)(base + o_nb);
      tb.mp = (long)mp;
      tb.np = (long)np;
      float *Vt = tb.Vt, *Vs = tb.Vs;
      unsigned *rt = tb.rt, *rs = tb.rs;
      hipLaunchKernelGGL(mix_pivot_gram_kernel, dim3((unsigned)P), dim3((unsigned)P), 0, s, P, kp,
                         (const float *)piv, pp, pn, ppd, pn64);
      // exponents in log2 units: xs = scal log2(e) in FP64 for the tables
      const double xs64 = p.P.scal * 1.4426950408889634;
      switch (nc) {
        case 1: launch_bounds<32>(p, (int)ntb, (int)nsb, pivA, piv, pn, pn64, ppd, kouta, koutb, prow, pcol, xs64, tb, s); break;
        case 2: launch_bounds<64>(p, (int)ntb, (int)nsb, pivA, piv, pn, pn64, ppd, kouta, koutb, prow, pcol, xs64, tb, s); break;
        case 3: launch_bounds<96>(p, (int)ntb, (int)nsb, pivA, piv, pn, pn64, ppd, kouta, koutb, prow, pcol, xs64, tb, s); break;
        default: launch_bounds<128>(p, (int)ntb, (int)nsb, pivA, piv, pn, pn64, ppd, kouta, koutb, prow, pcol, xs64, tb, s); break;
      }
      pk.Ex = tb.Ex;
      pk.Qx = tb.Qx;
      pk.Ac = tb.Ac;
      pk.Bc = tb.Bc;
      pk.na = tb.na;
      pk.nb = tb.nb;
      pk.rt = tb.rt;
      pk.rs = tb.rs;
      MixBounds mb;
      mb.Vt = Vt;
      mb.Vs = Vs;
      mb.Vt1 = tb.Vt1;
      mb.Vs1 = tb.Vs1;
      mb.pp = pp;
      mb.rt = rt;
      mb.rs = rs;
      mb.ntb = (int)ntb;
      mb.nsb = (int)nsb;
      // a block pair whose d^2 lower bound exceeds -kSkipT2 / |xs| has every
      // exponent xs d^2 below kSkipT2 (K < 2^-50); the factor covers xs =
      // scal log2(e) rounded to FP32
      const float xs = (float)(p.P.scal * (double)kLog2e);
      mb.thr_d2 = (float)(-(double)kSkipT2 / fabs((double)xs) * (1.0 + 0x1p-20));
      unsigned long long *need = (unsigned long long *)(base + o_need);
      const long nwaves = (long)ngrp * plan.nsplit;
      // needed-block counts per slice and row block (row groups are row blocks:
      // the skip-tier kernel runs one wave per workgroup)
      static_assert(mix_waves(KS_PREC_MIXED_BF16X3, GAUSSIAN) == 1, "row group != row block");
      unsigned *ok0 = (unsigned *)(base + o_ord), *ok1 = ok0 + nord;
      int *ov0 = (int *)(ok1 + nord), *ov1 = ov0 + nord;
      hipLaunchKernelGGL(mix_need_kernel, dim3((unsigned)((nwaves + 3) / 4)), dim3(256), 0, s, mb,
                         (int)ngrp, plan.nsplit, plan.cols_per_split, (int)ngw, p.n, need, ok0);
      pk.need = need;
      // each slice's row blocks by descending needed-block count (stable: equal
      // counts keep their order), so the heaviest start first
      if (plan.row_blocks <= kRankMaxRB) {
        hipLaunchKernelGGL(mix_rank_kernel,
                           dim3((unsigned)((plan.row_blocks + 63) / 64), (unsigned)plan.nsplit),
                           dim3(64 * kRankWaves), 0, s, ok0, plan.row_blocks, ov1);
      } else {
        hipLaunchKernelGGL(mix_order_kernel, dim3((unsigned)((nord + 255) / 256)), dim3(256), 0,
                           s, ok0, plan.nsplit, plan.row_blocks, ov0);
        size_t obytes = sort_bytes;
        if (hipcub::DeviceRadixSort::SortPairs(base + o_sort, obytes, ok0, ok1, ov0, ov1,
                                               (int)nord, 0, 32, s) != hipSuccess)
          return KS_ERR_DEVICE;
      }
      pk.order = ov1;
    }
  }
  if (hipGetLastError() != hipSuccess) return KS_ERR_DEVICE;

  hipError_t e = (precision == KS_PREC_MIXED_BF16X3)
                     ? mixed_launch<KS_PREC_MIXED_BF16X3>(p, plan, partial, pk, nc, s)
                     : mixed_launch<KS_PREC_MIXED_FP32>(p, plan, partial, pk, nc, s);
  return e == hipSuccess ? KS_OK : KS_ERR_DEVICE;
}

}  // namespace gsks
