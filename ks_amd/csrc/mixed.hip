// Mixed-precision fused summation (Gaussian and variable-bandwidth Gaussian):
//   u[umap[i]] += sum_j exp(x_ij) w[wmap[j]],  x = scal*d^2  or  -0.5 d^2 hi hj.
//
// Three changes against the FP64 drop-in, each where it is safe:
//   1. the cross term a.b runs on the BF16 matrix cores as a split-BF16
//      three-product, a ~= ah + al and b ~= bh + bl (bf16 each),
//      a.b ~= ah.bh + ah.bl + al.bh, accumulated in FP32 by
//      v_mfma_f32_16x16x32_bf16 -- 3 instructions per 32 coordinates of a
//      16 x 16 pair block (KS_PREC_MIXED_BF16X3);  or on the FP32 matrix cores,
//      v_mfma_f32_16x16x4f32, 8 instructions per 32 coordinates
//      (KS_PREC_MIXED_FP32);
//   2. that estimate classifies every pair: a pair whose estimated exponent
//      lies below -kFarT (K < ~1e-7) is "far" and its kernel value is taken in
//      FP32 (v_exp_f32) from the estimate; its error is bounded by the
//      estimate's error times K, far below 1e-6 of u;
//   3. a "near" pair -- the pairs that carry u, and the ones where the
//      |a|^2 + |b|^2 - 2 a.b expansion cancels -- is pushed to a per-wave LDS
//      queue and recomputed in FP64 from the coordinates (FP64 FMA dot
//      product, FP64 norms, FP64 exp) 64 pairs at a time, one pair per lane.
// u accumulates in FP64 across source tiles (FP32 only inside one 16 x 16
// block of far pairs).
//
// Packing: one pre-pass converts the target rows (through amap) and the source
// points (through bmap) to the MFMA operand layout once per call -- for BF16x3
// each point is nc chunks of 32 coordinates, each chunk 4 groups x (8 hi + 8 lo
// bf16); for FP32 each point is nc chunks x 4 groups x 8 floats -- plus per-row
// and per-column FP32 scale factors and weights.  The main kernel streams the
// packed sources; A fragments stay in VGPRs.
#include <hip/hip_runtime.h>

#include "gsks_kernels.h"
#include "gsks_launch.h"
#include "host_util.h"

namespace gsks {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr float kFarT = 16.0f;          // far: estimated x < -16 (natural-log units)
constexpr float kLog2e = 1.4426950408889634f;
constexpr int XWG = 256;                // 4 waves
constexpr int kQueue = 512;             // per-wave near-pair ring (power of two)

// ------------------------------------------------------------------ pre-pass

struct MixPack {
  // packed operands, 128 B per 32-coordinate chunk per point (both precisions)
  uint4 *Ap;   // [m][nc][4 groups][32 B]
  uint4 *Bp;   // [n][nc][4 groups][32 B]
  float2 *ra;  // per target row: {scaled |a|^2 (Gaussian) or |a|^2 (VarBW), hi factor}
  float4 *cb;  // per source: {scaled |b|^2 or |b|^2, hj factor, w, 0}
};

template <int PREC>
__device__ __forceinline__ void pack_point(const double *__restrict__ x, int k, int nc, uint4 *out) {
  for (int c = 0; c < nc; ++c) {
    for (int g = 0; g < 4; ++g) {
      if constexpr (PREC == KS_PREC_MIXED_BF16X3) {
        bf8 hi, lo;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          int idx = c * 32 + 8 * g + t;
          double v = idx < k ? x[idx] : 0.0;
          __bf16 h = (__bf16)(float)v;
          hi[t] = h;
          lo[t] = (__bf16)(float)(v - (double)(float)h);
        }
        uint4 *o = out + (c * 4 + g) * 2;
        o[0] = *reinterpret_cast<uint4 *>(&hi);
        o[1] = *reinterpret_cast<uint4 *>(&lo);
      } else {
        float f[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          int idx = c * 32 + 8 * g + t;
          f[t] = idx < k ? (float)x[idx] : 0.0f;
        }
        uint4 *o = out + (c * 4 + g) * 2;
        o[0] = *reinterpret_cast<uint4 *>(&f[0]);
        o[1] = *reinterpret_cast<uint4 *>(&f[4]);
      }
    }
  }
}

template <int PREC, int TYPE>
__global__ void mixed_pack_kernel(const Problem p, int nc, float xs, MixPack pk) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < p.m) {
    int ai = p.amap[t];
    pack_point<PREC>(p.XA + (long)ai * p.k, p.k, nc, pk.Ap + t * nc * 8);
    double aa = p.XA2[ai];
    if (TYPE == GAUSSIAN) pk.ra[t] = make_float2((float)(aa * (double)xs), 0.0f);
    else pk.ra[t] = make_float2((float)aa, (float)(p.hi[ai] * -0.5 * (double)kLog2e));
  } else if (t < (long)p.m + p.n) {
    long j = t - p.m;
    int bj = p.bmap[j];
    pack_point<PREC>(p.XB + (long)bj * p.k, p.k, nc, pk.Bp + j * nc * 8);
    double bb = p.XB2[bj];
    float w = (float)p.w[p.wmap[j]];
    if (TYPE == GAUSSIAN) pk.cb[j] = make_float4((float)(bb * (double)xs), 0.0f, w, 0.0f);
    else pk.cb[j] = make_float4((float)bb, (float)p.hj[bj], w, 0.0f);
  }
}

// ------------------------------------------------------------------ main kernel

template <int PREC, int NC>
struct Frag {
  uint4 v[NC][2];  // per chunk: 32 B (bf16: hi | lo, fp32: floats 0-3 | 4-7)
};

template <int PREC, int NC>
__device__ __forceinline__ void load_frag_packed(const uint4 *__restrict__ base, int g,
                                                 Frag<PREC, NC> &f) {
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    f.v[c][0] = base[(c * 4 + g) * 2];
    f.v[c][1] = base[(c * 4 + g) * 2 + 1];
  }
}

template <int PREC, int NC>
__device__ __forceinline__ f4 cross_block(const Frag<PREC, NC> &a, const Frag<PREC, NC> &b) {
  f4 acc = {0.f, 0.f, 0.f, 0.f};
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

// FP64 kernel value times weight of one near pair from its exact rank-k sum c.
template <int TYPE>
__device__ __forceinline__ double near_value(const Problem &p, int row, int col, double c) {
  const int ai = p.amap[row], bj = p.bmap[col];
  double d2 = fmax(fma(-2.0, c, p.XA2[ai]) + p.XB2[bj], 0.0);
  double x = (TYPE == GAUSSIAN) ? d2 * p.P.scal : (d2 * (-0.5 * p.hi[ai])) * p.hj[bj];
  return exp(x) * p.w[p.wmap[col]];
}

// Row of the 16 x 16 result block held in accumulator slot v by lane group g.
// Measured on gfx950: v_mfma_f32_16x16x32_bf16 and v_mfma_f32_16x16x4f32 put
// rows 4g + v of column lane%16 in lane g*16 + col (the FP64 16x16x4 shape
// uses g + 4v instead).
__device__ __forceinline__ int out_row(int g, int v) { return 4 * g + v; }

constexpr int kFlush = 8;  // far-pair FP32 partials go to FP64 every 8 source blocks

template <int PREC, int TYPE, int NC, int RB>
__global__ __launch_bounds__(XWG) void mixed_kernel(FusedArgs A, MixPack pk, float xs) {
  const Problem &p = A.p;
  const int m = p.m;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int p16 = lane & 15;
  const int g = lane >> 4;
  const int rblk = blockIdx.x;
  const int split = blockIdx.y;
  constexpr int WROWS = RB * 16;
  const int wg_rows = 4 * WROWS;
  const int row0 = rblk * wg_rows + wave * WROWS;

  __shared__ int s_queue[4][kQueue];
  __shared__ double s_near[4][WROWS];
  // the wave's target rows in FP64 for the near-pair recompute: each is reused
  // by every near pair of its row, so only the source side is gathered later
  // (row stride 32*NC + 2 doubles keeps the 16 lanes' 32-byte reads apart)
  constexpr int KS = 32 * NC + 2;
  __shared__ double s_rows[4][WROWS * KS];
  int *queue = s_queue[wave];
  double *nacc = s_near[wave];
  double *arows = s_rows[wave];
  for (int q = lane; q < WROWS; q += 64) nacc[q] = 0.0;
  for (int e = lane; e < WROWS * 32 * NC; e += 64) {
    const int rl = e / (32 * NC), q = e - rl * (32 * NC);
    const int r = row0 + rl;
    arows[rl * KS + q] = q < p.k ? p.XA[(long)p.amap[r < m ? r : m - 1] * p.k + q] : 0.0;
  }

  Frag<PREC, NC> af[RB];
  float ra_s[RB][4], ra_h[RB][4], facc[RB][4];
  double racc[RB][4];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    int r = row0 + rb * 16 + p16;
    load_frag_packed<PREC, NC>(pk.Ap + (long)(r < m ? r : m - 1) * NC * 8, g, af[rb]);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      int ro = row0 + rb * 16 + out_row(g, v);
      float2 t = pk.ra[ro < m ? ro : m - 1];
      // a row past m can never be near: its estimated exponent is -inf
      if (ro >= m) t.x = (TYPE == GAUSSIAN) ? -INFINITY : INFINITY;
      ra_s[rb][v] = t.x;
      ra_h[rb][v] = t.y;
      racc[rb][v] = 0.0;
      facc[rb][v] = 0.0f;
    }
  }
  const int j0 = split * A.cols_per_split;
  const int j1 = min(p.n, j0 + A.cols_per_split);
  unsigned head = 0, tail = 0;  // wave-uniform ring indices

  Frag<PREC, NC> bf, bn;
  float4 cb = make_float4(0.f, 0.f, 0.f, 0.f), cbn = cb;
  auto fetch = [&](int jb, Frag<PREC, NC> &f, float4 &c4) {
    int col = jb + p16;
    bool ok = col < j1;
    load_frag_packed<PREC, NC>(pk.Bp + (long)(ok ? col : j1 - 1) * NC * 8, g, f);
    c4 = pk.cb[ok ? col : j1 - 1];
    if (!ok) {  // a column past the slice: exponent -inf, weight 0
      c4.x = (TYPE == GAUSSIAN) ? -INFINITY : INFINITY;
      c4.z = 0.0f;
    }
  };
  // Near pairs, up to 64 per call: 16 rounds of 4 pairs x 16 lanes.  In round r
  // the 16 lanes of group s sum pair 4r + s, each over 2*NC consecutive
  // coordinates (contiguous 16-byte loads of both rows), reduce across the 16
  // lanes, and lane 16s + r keeps the sum; afterwards every lane holds one
  // pair's exact rank-k sum and evaluates its kernel value in FP64.
  auto drain = [&](unsigned count) {
    const int grp = lane >> 4, sub = lane & 15;
    constexpr int CPL = 2 * NC;  // coordinates per lane (k <= 32*NC)
    double mine_c = 0.0;
    int mine_e = -1;
    // 4 rounds at a time: all their loads are issued before the first FMA
#pragma unroll 1
    for (int r0 = 0; r0 < 16; r0 += 4) {
      if (4 * r0 >= (int)count) break;
      double2 av[4][CPL / 2], bv[4][CPL / 2];
      int ev[4];
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int idx = 4 * (r0 + rr) + grp;
        const int e = idx < (int)count ? queue[(head + idx) & (kQueue - 1)] : -1;
        ev[rr] = e;
        const int rl = (e >= 0 ? e : 0) & 31, col = e >= 0 ? (e >> 5) : j0;
        const double *b = p.XB + (long)p.bmap[col] * p.k + sub * CPL;
        const double *a = arows + rl * KS + sub * CPL;  // zero-padded past k
#pragma unroll
        for (int t = 0; t < CPL / 2; ++t) {
          const bool in = (e >= 0) && (sub * CPL + 2 * t < p.k);
          av[rr][t] = *reinterpret_cast<const double2 *>(a + 2 * t);
          if (A.vec2) {
            bv[rr][t] = in ? reinterpret_cast<const double2 *>(b)[t] : make_double2(0.0, 0.0);
          } else {
            const bool in1 = in && (sub * CPL + 2 * t + 1 < p.k);
            bv[rr][t] = make_double2(in ? b[2 * t] : 0.0, in1 ? b[2 * t + 1] : 0.0);
          }
        }
      }
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        double c = 0.0;
#pragma unroll
        for (int t = 0; t < CPL / 2; ++t) {
          c = fma(av[rr][t].x, bv[rr][t].x, c);
          c = fma(av[rr][t].y, bv[rr][t].y, c);
        }
        c += __shfl_xor(c, 8, 16);
        c += __shfl_xor(c, 4, 16);
        c += __shfl_xor(c, 2, 16);
        c += __shfl_xor(c, 1, 16);
        if (sub == r0 + rr) {
          mine_c = c;
          mine_e = ev[rr];
        }
      }
    }
    if (mine_e >= 0) {
      const int rl = mine_e & 31, col = mine_e >> 5;
      atomicAdd(&nacc[rl], near_value<TYPE>(p, row0 + rl, col, mine_c));
    }
    head += count;
  };

  if (j0 < j1) fetch(j0, bf, cb);
  int nblk = 0;
  for (int jb = j0; jb < j1; jb += 16) {
    const bool more = jb + 16 < j1;
    if (more) fetch(jb + 16, bn, cbn);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      f4 c = cross_block<PREC, NC>(af[rb], bf);
      unsigned long long nmask[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        float x2;  // estimated exponent, log2 units (-inf for padding)
        if constexpr (TYPE == GAUSSIAN) {
          x2 = fmaf(-2.0f * xs, c[v], ra_s[rb][v] + cb.x);
        } else {
          float d2 = fmaxf(fmaf(-2.0f, c[v], ra_s[rb][v] + cb.x), 0.0f);
          x2 = d2 * (ra_h[rb][v] * cb.y);
        }
        const bool nr = x2 > -kFarT * kLog2e;
        const float K = __builtin_amdgcn_exp2f(x2);
        facc[rb][v] = fmaf(nr ? 0.0f : K, cb.z, facc[rb][v]);
        nmask[v] = __ballot(nr);
      }
      // queue this row block's near pairs (row-local index in the low 5 bits);
      // the ring holds < 64 + 4*64 entries at any time
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const unsigned long long mask = nmask[v];
        if (mask) {
          if ((mask >> lane) & 1ull) {
            const unsigned off = __popcll(mask & ((1ull << lane) - 1ull));
            queue[(tail + off) & (kQueue - 1)] = ((jb + p16) << 5) | (rb * 16 + out_row(g, v));
          }
          tail += (unsigned)__popcll(mask);
        }
      }
      while (tail - head >= 64) drain(64);
    }
    if (++nblk == kFlush) {
      nblk = 0;
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          racc[rb][v] += (double)facc[rb][v];
          facc[rb][v] = 0.0f;
        }
    }
    if (more) {
      bf = bn;
      cb = cbn;
    }
  }
  while (tail != head) drain(tail - head < 64 ? tail - head : 64);

  // ---- row totals: far partials summed over the 16 column lanes + near sums
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double s = racc[rb][v] + (double)facc[rb][v];
      s += __shfl_xor(s, 8, 16);
      s += __shfl_xor(s, 4, 16);
      s += __shfl_xor(s, 2, 16);
      s += __shfl_xor(s, 1, 16);
      racc[rb][v] = s;
    }
  const bool writer = p16 < 4;
  double mine[RB];
  int myrow[RB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    double x = racc[rb][0];
    x = (p16 == 1) ? racc[rb][1] : x;
    x = (p16 == 2) ? racc[rb][2] : x;
    x = (p16 == 3) ? racc[rb][3] : x;
    const int rl = rb * 16 + out_row(g, p16 & 3);
    mine[rb] = x + (writer ? nacc[rl] : 0.0);
    myrow[rb] = row0 + rl;
  }
  if (A.nsplit == 1) {
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
      if (writer && myrow[rb] < m) p.u[p.umap[myrow[rb]]] += mine[rb];
    return;
  }
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
    if (writer && myrow[rb] < m) A.partial[(long)split * m + myrow[rb]] = mine[rb];
  __threadfence();
  __syncthreads();
  __shared__ int s_last;
  if (threadIdx.x == 0) {
    unsigned prev = atomicAdd(&A.counters[rblk], 1u);
    s_last = (prev == (unsigned)(A.nsplit - 1));
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  for (int q = threadIdx.x; q < wg_rows; q += XWG) {
    int r = rblk * wg_rows + q;
    if (r < m) {
      double s = 0.0;
      for (int sp = 0; sp < A.nsplit; ++sp) s += A.partial[(long)sp * m + r];
      p.u[p.umap[r]] += s;
    }
  }
  if (threadIdx.x == 0) A.counters[rblk] = 0u;
}

// ------------------------------------------------------------------ launch

constexpr int kMixRB = 2;
constexpr int kMixRowsPerWG = 4 * kMixRB * 16;

template <int PREC, int TYPE, int NC>
static hipError_t mixed_go(const FusedArgs &A, const MixPack &pk, float xs, int row_blocks,
                           hipStream_t s) {
  dim3 grid(row_blocks, A.nsplit);
  hipLaunchKernelGGL((mixed_kernel<PREC, TYPE, NC, kMixRB>), grid, dim3(XWG), 0, s, A, pk, xs);
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
                               unsigned *counters, const MixPack &pk, int nc, hipStream_t s) {
  FusedArgs A;
  A.p = p;
  A.cols_per_split = plan.cols_per_split;
  A.nsplit = plan.nsplit;
  A.partial = partial;
  A.counters = counters;
  A.vec2 = (p.k % 2 == 0) && ((uintptr_t)p.XA % 16 == 0) && ((uintptr_t)p.XB % 16 == 0);
  const float xs = (float)(p.P.scal * (double)kLog2e);
  long tot = (long)p.m + p.n;
  dim3 pg((unsigned)((tot + 255) / 256));
  if (p.type == GAUSSIAN) {
    hipLaunchKernelGGL((mixed_pack_kernel<PREC, GAUSSIAN>), pg, dim3(256), 0, s, p, nc, xs, pk);
    return mixed_nc<PREC, GAUSSIAN>(A, pk, xs, nc, plan.row_blocks, s);
  }
  hipLaunchKernelGGL((mixed_pack_kernel<PREC, VAR_BANDWIDTH>), pg, dim3(256), 0, s, p, nc, xs, pk);
  return mixed_nc<PREC, VAR_BANDWIDTH>(A, pk, xs, nc, plan.row_blocks, s);
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
  FusedPlan plan = plan_split(p.m, p.n, kMixRowsPerWG, 2048, 128);
  // scratch: partial sums | counters | packed operands | per-row / per-column factors
  size_t packA = (size_t)p.m * nc * 128, packB = (size_t)p.n * nc * 128;
  size_t ra = (size_t)p.m * sizeof(float2), cbb = (size_t)p.n * sizeof(float4);
  size_t extra = packA + packB + ra + cbb + 4 * 256;
  double *partial = nullptr;
  unsigned *counters = nullptr;
  size_t pelems = (size_t)(plan.nsplit > 1 ? plan.nsplit : 0) * p.m + (extra + 7) / 8;
  int rc = get_scratch(s, pelems, (size_t)plan.row_blocks, &partial, &counters);
  if (rc != KS_OK) return rc;
  char *base = reinterpret_cast<char *>(partial + (size_t)(plan.nsplit > 1 ? plan.nsplit : 0) * p.m);
  auto align = [](char *q) { return (char *)(((uintptr_t)q + 255) & ~(uintptr_t)255); };
  MixPack pk;
  char *q = align(base);
  pk.Ap = (uint4 *)q;
  q = align(q + packA);
  pk.Bp = (uint4 *)q;
  q = align(q + packB);
  pk.ra = (float2 *)q;
  q = align(q + ra);
  pk.cb = (float4 *)q;
  hipError_t e = (precision == KS_PREC_MIXED_BF16X3)
                     ? mixed_launch<KS_PREC_MIXED_BF16X3>(p, plan, partial, counters, pk, nc, s)
                     : mixed_launch<KS_PREC_MIXED_FP32>(p, plan, partial, counters, pk, nc, s);
  return e == hipSuccess ? KS_OK : KS_ERR_DEVICE;
}

}  // namespace gsks
