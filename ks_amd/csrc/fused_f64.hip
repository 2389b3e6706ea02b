// FP64 fused kernel summation for gfx950:
//   u[umap[i]] += sum_j K(XA[amap[i]], XB[bmap[j]]) * w[wmap[j]]
// in one launch, without forming the m x n kernel matrix.
//
// Mapping (replaces the BLIS loop nest and the x86 micro-kernels of
// frame/dgsks.c:367-776 / micro_kernel/x86_64/sandybridge):
//   * a lane owns R target rows; their coordinates stay in VGPRs for the
//     whole launch (gathered once through amap);
//   * a workgroup (4 waves, 256*R rows) walks one slice of the source points;
//     every lane of a wave needs the same source point at the same time, so
//     the source coordinates, |b|^2, w and the index maps are read with
//     wave-uniform scalar loads into SGPRs and feed the FP64 FMAs directly
//     as the scalar operand -- no LDS traffic and no barriers in the loop;
//   * the GEMV partial u_i stays in a VGPR across the slice;
//   * when the source points are split over several workgroups (to fill 256
//     CUs), each writes its partial sums to scratch and the last workgroup to
//     finish a row block (device-scope counter) adds the partials in split
//     order -- a deterministic reduction, no floating-point atomics -- and
//     resets the counter, so the launch leaves scratch ready for the next one.
#include <hip/hip_runtime.h>

#include "gsks_kernels.h"
#include "gsks_launch.h"

namespace gsks {

constexpr int WG = 256;

template <int TYPE, int K, int R>
__global__ __launch_bounds__(WG) void fused_f64_kernel(FusedArgs A) {
  const Problem &p = A.p;
  const int m = p.m;
  __shared__ double2 etab[64];
  if constexpr (uses_exp<TYPE>) {
    load_exp_table(etab);
    __syncthreads();
  }
  const int kdim = (K > 0) ? K : p.k;
  const int tid = threadIdx.x;
  const int rblk = blockIdx.x;
  const int split = blockIdx.y;
  const int row0 = rblk * (WG * R);

  // ---- target rows: gathered once, kept in registers
  constexpr int KR = (K > 0) ? K : 1;
  double a[R][KR];
  double aa[R], hi[R], acc[R];
  const double *arow[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    int i = row0 + q * WG + tid;
    int ii = i < m ? i : m - 1;
    int ai = p.amap[ii];
    arow[q] = p.XA + (long)ai * kdim;
    aa[q] = p.XA2[ai];
    hi[q] = (TYPE == VAR_BANDWIDTH) ? -0.5 * p.hi[ai] : 0.0;
    acc[q] = 0.0;
    if constexpr (K > 0) {
#pragma unroll
      for (int pp = 0; pp < K; ++pp) a[q][pp] = arow[q][pp];
    }
  }

  // ---- source slice
  const int j0 = split * A.cols_per_split;
  const int j1 = min(p.n, j0 + A.cols_per_split);
  cptr<int> bmap = to_const(p.bmap);
  cptr<int> wmap = to_const(p.wmap);
  cptr<double> XB = to_const(p.XB);
  cptr<double> XB2 = to_const(p.XB2);
  cptr<double> W = to_const(p.w);
  cptr<double> HJ = to_const(p.hj);
  const KerParams P = p.P;

#pragma unroll 2
  for (int j = j0; j < j1; ++j) {
    const int bj = bmap[j];
    const double wj = W[wmap[j]];
    const double bb = XB2[bj];
    const double hj = (TYPE == VAR_BANDWIDTH) ? HJ[bj] : 0.0;
    cptr<double> b = XB + (long)bj * kdim;
    if constexpr (K > 0) {
      double bv[K];
#pragma unroll
      for (int pp = 0; pp < K; ++pp) bv[pp] = b[pp];
#pragma unroll
      for (int q = 0; q < R; ++q) {
        double c;
        if constexpr (TYPE == LAPLACE) {
          c = dot_exact<K>(a[q], b);
        } else {
          c = a[q][0] * bv[0];
#pragma unroll
          for (int pp = 1; pp < K; ++pp) c = fma(a[q][pp], bv[pp], c);
        }
        acc[q] = fma(kval_f64<TYPE, uses_exp<TYPE>>(c, aa[q], bb, hi[q], hj, P, etab), wj, acc[q]);
      }
    } else {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        double c;
        if constexpr (TYPE == LAPLACE) {
          c = dot_exact_rt(arow[q], b, kdim);
        } else {
          c = arow[q][0] * b[0];
          for (int pp = 1; pp < kdim; ++pp) c = fma(arow[q][pp], b[pp], c);
        }
        acc[q] = fma(kval_f64<TYPE, uses_exp<TYPE>>(c, aa[q], bb, hi[q], hj, P, etab), wj, acc[q]);
      }
    }
  }

  // ---- results
  if (A.nsplit == 1) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      int i = row0 + q * WG + tid;
      if (i < m) p.u[p.umap[i]] += acc[q];
    }
    return;
  }
  // the splits' partials are added in split order by launch_split_reduce's
  // kernel, launched next (no device-scope fence here: on gfx950 it would
  // write back and invalidate this XCD's L2)
#pragma unroll
  for (int q = 0; q < R; ++q) {
    int i = row0 + q * WG + tid;
    if (i < m) A.partial[(long)split * m + i] = acc[q];
  }
}

// ------------------------------------------------------------------ launch

template <int TYPE, int K, int R>
static hipError_t launch_one(const FusedArgs &A, int row_blocks, hipStream_t s) {
  dim3 grid(row_blocks, A.nsplit);
  hipLaunchKernelGGL((fused_f64_kernel<TYPE, K, R>), grid, dim3(WG), 0, s, A);
  return hipGetLastError();
}

template <int TYPE>
static hipError_t launch_type(FusedArgs A, const FusedPlan &plan, hipStream_t s) {
  switch (plan.kspec) {
#define KS_CASE(KK, RR) \
  case KK: return launch_one<TYPE, KK, RR>(A, plan.row_blocks, s);
    KS_CASE(2, 4)
    KS_CASE(3, 4)
    KS_CASE(4, 4)
    KS_CASE(8, 4)
    KS_CASE(16, 2)
    KS_CASE(32, 1)
    KS_CASE(64, 1)
#undef KS_CASE
    default: return launch_one<TYPE, 0, 1>(A, plan.row_blocks, s);
  }
}

// rows handled by one workgroup for a given k
int fused_rows_per_lane(int k) {
  switch (k) {
    case 2: case 3: case 4: case 8: return 4;
    case 16: return 2;
    default: return 1;
  }
}

int fused_kspec(int k) {
  switch (k) {
    case 2: case 3: case 4: case 8: case 16: case 32: case 64: return k;
    default: return 0;
  }
}

FusedPlan plan_split(int m, int n, int rows_per_wg, int target_wgs, int min_cols) {
  FusedPlan pl;
  pl.kspec = 0;
  pl.rows_per_wg = rows_per_wg;
  pl.row_blocks = (m + rows_per_wg - 1) / rows_per_wg;
  long want = (target_wgs + pl.row_blocks - 1) / pl.row_blocks;
  long maxs = (n + min_cols - 1) / min_cols;
  if (want > maxs) want = maxs;
  if (want < 1) want = 1;
  pl.cols_per_split = (int)((n + want - 1) / want);
  pl.nsplit = (n + pl.cols_per_split - 1) / pl.cols_per_split;
  return pl;
}

FusedPlan plan_fused(int m, int n, int k, int target_wgs, int min_cols) {
  int ks = fused_kspec(k);
  FusedPlan pl = plan_split(m, n, WG * fused_rows_per_lane(ks), target_wgs, min_cols);
  pl.kspec = ks;
  return pl;
}

hipError_t launch_fused_f64(const Problem &p, const FusedPlan &plan, double *partial,
                            hipStream_t s) {
  FusedArgs A;
  A.p = p;
  A.cols_per_split = plan.cols_per_split;
  A.nsplit = plan.nsplit;
  A.partial = partial;
  A.vec2 = 0;
  switch (p.type) {
    case GAUSSIAN: return launch_type<GAUSSIAN>(A, plan, s);
    case POLYNOMIAL:
      switch (poly_type(p.P)) {  // the power fixed at compile time where it can be
        case POLY_SQUARE_T: return launch_type<POLY_SQUARE_T>(A, plan, s);
        case POLY_CUBE_T: return launch_type<POLY_CUBE_T>(A, plan, s);
        case POLY_FOURTH_T: return launch_type<POLY_FOURTH_T>(A, plan, s);
        default: return launch_type<POLYNOMIAL>(A, plan, s);
      }
    case LAPLACE: return launch_type<LAPLACE>(A, plan, s);
    case VAR_BANDWIDTH: return launch_type<VAR_BANDWIDTH>(A, plan, s);
    case TANH: return launch_type<TANH>(A, plan, s);
    case QUARTIC: return launch_type<QUARTIC>(A, plan, s);
    case MULTIQUADRATIC: return launch_type<MULTIQUADRATIC>(A, plan, s);
    case EPANECHNIKOV: return launch_type<EPANECHNIKOV>(A, plan, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace gsks
