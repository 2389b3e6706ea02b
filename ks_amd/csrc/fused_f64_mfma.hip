// FP64 fused kernel summation with the cross term on the matrix cores
// (v_mfma_f64_16x16x4_f64), for 9 <= k <= 64.
//
// On MI355X the FP64 matrix and FP64 vector peaks are equal, but they are
// separate pipes: here the matrix core forms a.b for a 16 x 16 block of pairs
// while the VALU evaluates the kernel function of the previous block, so the
// k-long cross term no longer competes with exp() for VALU issue slots.
//
// Mapping:
//   * a wave owns RB blocks of 16 target rows; their coordinates are loaded once
//     into the MFMA A-operand layout and stay in VGPRs;
//   * the wave streams 16-point source blocks of its slice; each lane loads its
//     B fragment straight from HBM/L2 through bmap (lane l reads point l%16,
//     coordinates 4*(l/16) .. +3 of each 16-coordinate chunk -- one contiguous
//     32-byte piece per chunk; A uses the same coordinate order, so the MFMA k
//     order is a permutation of 0..k-1 shared by both operands);
//   * the 16 x 16 result block (lane l holds rows l/16 + 4v, v = 0..3, of
//     column l%16 -- the layout measured on gfx950) is turned into d^2 = |a|^2 - 2c + |b|^2, K(d^2) and K*w in
//     registers; each lane keeps 4*RB row partials across the whole slice and
//     the 16 lanes of a row group are summed once at the end;
//   * the split-source reduction is the one of fused_f64.hip (partials + the
//     last workgroup of a row block adds them in split order).
// Laplace is not routed here: its d^2 < 1e-15 test needs the reference's exact
// rank-k sum (see fused_f64.hip).
#include <hip/hip_runtime.h>

#include "gsks_kernels.h"
#include "gsks_launch.h"

namespace gsks {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int MWG = 256;  // 4 waves

// Operand fragment of one point: coordinates c*16 + 4g .. +3 for each 16-chunk c.
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

template <int TYPE, int KC, int RB, bool EXACT>
__global__ __launch_bounds__(MWG) void fused_f64_mfma_kernel(FusedArgs A) {
  const Problem &p = A.p;
  const int m = p.m, k = p.k;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int p16 = lane & 15;
  const int g = lane >> 4;
  const int rblk = blockIdx.x;
  const int split = blockIdx.y;
  const int wg_rows = 4 * RB * 16;
  const int row0 = rblk * wg_rows + wave * (RB * 16);

  // ---- A fragments and the norms / bandwidths of the rows this lane's results cover
  double af[RB][KC][4];
  double aa[RB][4], hi[RB][4], racc[RB][4];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    int r = row0 + rb * 16 + p16;
    int ai = p.amap[r < m ? r : m - 1];
    load_frag<KC, EXACT>(p.XA, (long)ai * k, k, g, af[rb]);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      int ro = row0 + rb * 16 + g + 4 * v;
      int ao = p.amap[ro < m ? ro : m - 1];
      aa[rb][v] = p.XA2[ao];
      hi[rb][v] = (TYPE == VAR_BANDWIDTH) ? -0.5 * p.hi[ao] : 0.0;
      racc[rb][v] = 0.0;
    }
  }

  const int j0 = split * A.cols_per_split;
  const int j1 = min(p.n, j0 + A.cols_per_split);
  const KerParams P = p.P;

  // ---- stream 16-point source blocks; two register buffers used in turn
  double bf0[KC][4], bf1[KC][4];
  double bb0 = 0.0, wj0 = 0.0, hj0 = 0.0, bb1 = 0.0, wj1 = 0.0, hj1 = 0.0;
  auto fetch = [&](int jb, double (&f)[KC][4], double &b2, double &ww, double &hh) {
    int col = jb + p16;
    bool ok = col < j1;
    int bj = p.bmap[ok ? col : j1 - 1];
    load_frag<KC, EXACT>(p.XB, (long)bj * k, k, g, f);
    b2 = p.XB2[bj];
    ww = ok ? p.w[p.wmap[col]] : 0.0;
    hh = (TYPE == VAR_BANDWIDTH) ? p.hj[bj] : 0.0;
  };
  auto block = [&](const double (&f)[KC][4], double bb, double wj, double hj) {
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int c = 0; c < KC; ++c) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(af[rb][c][t], f[c][t], acc, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double K = kval_f64<TYPE>(acc[v], aa[rb][v], bb, hi[rb][v], hj, P);
        racc[rb][v] = fma(K, wj, racc[rb][v]);
      }
    }
  };
  if (j0 < j1) fetch(j0, bf0, bb0, wj0, hj0);
  for (int jb = j0; jb < j1; jb += 32) {
    if (jb + 16 < j1) fetch(jb + 16, bf1, bb1, wj1, hj1);
    block(bf0, bb0, wj0, hj0);
    if (jb + 16 >= j1) break;
    if (jb + 32 < j1) fetch(jb + 32, bf0, bb0, wj0, hj0);
    block(bf1, bb1, wj1, hj1);
  }

  // ---- sum the 16 lanes (source columns) of each row group
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double s = racc[rb][v];
      s += __shfl_xor(s, 8, 16);
      s += __shfl_xor(s, 4, 16);
      s += __shfl_xor(s, 2, 16);
      s += __shfl_xor(s, 1, 16);
      racc[rb][v] = s;
    }
  // lane p16 < 4 of group g writes row g + 4 p16 of each block
  double mine[RB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    double x = racc[rb][0];
    x = (p16 == 1) ? racc[rb][1] : x;
    x = (p16 == 2) ? racc[rb][2] : x;
    x = (p16 == 3) ? racc[rb][3] : x;
    mine[rb] = x;
  }
  const bool writer = p16 < 4;
  if (A.nsplit == 1) {
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      int r = row0 + rb * 16 + g + 4 * p16;
      if (writer && r < m) p.u[p.umap[r]] += mine[rb];
    }
    return;
  }
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    int r = row0 + rb * 16 + g + 4 * p16;
    if (writer && r < m) A.partial[(long)split * m + r] = mine[rb];
  }
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
  for (int q = threadIdx.x; q < wg_rows; q += MWG) {
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

constexpr int kMfmaRB = 2;

bool mfma_f64_eligible(int type, int k) { return type != LAPLACE && k >= 9 && k <= 64; }

int mfma_rows_per_wg() { return 4 * kMfmaRB * 16; }

template <int TYPE, int KC>
static hipError_t go(const FusedArgs &A, int row_blocks, hipStream_t s) {
  dim3 grid(row_blocks, A.nsplit);
  const bool exact = A.p.k == 16 * KC && (uintptr_t)A.p.XA % 16 == 0 &&
                     (uintptr_t)A.p.XB % 16 == 0;
  if (exact)
    hipLaunchKernelGGL((fused_f64_mfma_kernel<TYPE, KC, kMfmaRB, true>), grid, dim3(MWG), 0, s, A);
  else
    hipLaunchKernelGGL((fused_f64_mfma_kernel<TYPE, KC, kMfmaRB, false>), grid, dim3(MWG), 0, s, A);
  return hipGetLastError();
}

template <int TYPE>
static hipError_t go_type(const FusedArgs &A, int row_blocks, hipStream_t s) {
  switch ((A.p.k + 15) / 16) {
    case 1: return go<TYPE, 1>(A, row_blocks, s);
    case 2: return go<TYPE, 2>(A, row_blocks, s);
    case 3: return go<TYPE, 3>(A, row_blocks, s);
    default: return go<TYPE, 4>(A, row_blocks, s);
  }
}

hipError_t launch_fused_f64_mfma(const Problem &p, const FusedPlan &plan, double *partial,
                                 unsigned *counters, hipStream_t s) {
  FusedArgs A;
  A.p = p;
  A.cols_per_split = plan.cols_per_split;
  A.nsplit = plan.nsplit;
  A.partial = partial;
  A.counters = counters;
  A.vec2 = 0;
  switch (p.type) {
    case GAUSSIAN: return go_type<GAUSSIAN>(A, plan.row_blocks, s);
    case POLYNOMIAL: return go_type<POLYNOMIAL>(A, plan.row_blocks, s);
    case VAR_BANDWIDTH: return go_type<VAR_BANDWIDTH>(A, plan.row_blocks, s);
    case TANH: return go_type<TANH>(A, plan.row_blocks, s);
    case QUARTIC: return go_type<QUARTIC>(A, plan.row_blocks, s);
    case MULTIQUADRATIC: return go_type<MULTIQUADRATIC>(A, plan.row_blocks, s);
    case EPANECHNIKOV: return go_type<EPANECHNIKOV>(A, plan.row_blocks, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace gsks
