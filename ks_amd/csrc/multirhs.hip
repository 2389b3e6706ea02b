// Multi-right-hand-side fused summation and the Gaussian gradient built on it.
//
//   U[umap[i]][r] += sum_j K(XA[amap[i]], XB[bmap[j]]) * W[wmap[j]][r],  r < nrhs
//
// This is the reference's KS_RHS > 1 layout (u and w row-major [npoints][rhs],
// include/ks.h:46; rhs loops in micro_kernel/x86_64/sandybridge/
// ks_kernel_summation_int_d8x4.h): the GEMV becomes a small GEMM epilogue.  The
// kernel matrix is recomputed tile by tile, never stored.  Mapping as in
// fused_f64_kernel: a lane owns one target row (coordinates in VGPRs), the
// workgroup streams one source slice with wave-uniform scalar loads, and each
// K value feeds nrhs FMAs against the source's weight row (also scalar loads);
// split slices are reduced deterministically by the last workgroup.
//
// The gradient of L = g . u for u = K w, K_ij = exp(-|a_i - b_j|^2 / (2 h^2)),
// uses two such passes (identity maps), k + 1 right-hand sides each:
//   rows = a, sources = b, RHS [w b_1..w b_k, w] -> [V_i, u_i]
//   rows = b, sources = a, RHS [g a_1..g a_k, g] -> [Q_j, dw_j]
// and then, per point,
//   dXA_i = -(g_i / h^2) (u_i a_i - V_i)
//   dXB_j =  (w_j / h^2) (Q_j - dw_j b_j)
//   dh    = (1 / h^3) (sum_i g_i (|a_i|^2 u_i - 2 a_i . V_i) + sum_j w_j |b_j|^2 dw_j),
// the last sum being sum_i g_i sum_j K_ij w_j |b_j|^2 regrouped by source, so
// neither pass needs a w |b|^2 column (at k = 16 both have 17 columns: one
// 16-wide matrix-core tile and one column on the VALU).
#include <hip/hip_runtime.h>

#include "gsks_kernels.h"
#include "gsks_launch.h"
#include "host_util.h"

namespace gsks {

struct MrhsArgs {
  Problem p;  // p.u / p.w are the [npoints][nrhs] matrices U / W
  int nrhs;
  int cols_per_split, nsplit;
  double *partial;     // [nsplit][m][nrhs]
};

constexpr int MR_WG = 256;

template <int TYPE, int K, int RMAX>
__global__ __launch_bounds__(MR_WG) void multirhs_kernel(MrhsArgs A) {
  const Problem &p = A.p;
  const int m = p.m, nr = A.nrhs;
  __shared__ double2 etab[64];
  if constexpr (uses_exp<TYPE>) {
    load_exp_table(etab);
    __syncthreads();
  }
  const int kdim = K > 0 ? K : p.k;
  const int tid = threadIdx.x;
  const int i = blockIdx.x * MR_WG + tid;
  const int split = blockIdx.y;
  const int ai = p.amap[i < m ? i : m - 1];
  const double *arow = p.XA + (long)ai * kdim;
  constexpr int KR = K > 0 ? K : 1;
  double a[KR];
  if constexpr (K > 0) {
#pragma unroll
    for (int q = 0; q < K; ++q) a[q] = arow[q];
  }
  const double aa = p.XA2[ai];
  const double hi = (TYPE == VAR_BANDWIDTH) ? -0.5 * p.hi[ai] : 0.0;
  double acc[RMAX];
#pragma unroll
  for (int r = 0; r < RMAX; ++r) acc[r] = 0.0;

  const int j0 = split * A.cols_per_split;
  const int j1 = min(p.n, j0 + A.cols_per_split);
  cptr<int> bmap = to_const(p.bmap);
  cptr<int> wmap = to_const(p.wmap);
  cptr<double> XB = to_const(p.XB);
  cptr<double> XB2 = to_const(p.XB2);
  cptr<double> W = to_const(p.w);
  cptr<double> HJ = to_const(p.hj);
  const KerParams P = p.P;
  for (int j = j0; j < j1; ++j) {
    const int bj = bmap[j];
    const double bb = XB2[bj];
    const double hj = (TYPE == VAR_BANDWIDTH) ? HJ[bj] : 0.0;
    cptr<double> b = XB + (long)bj * kdim;
    double c;
    if constexpr (TYPE == LAPLACE) {
#pragma clang fp contract(off)
      c = (K > 0 ? a[0] : arow[0]) * b[0];
      for (int q = 1; q < kdim; ++q) c = c + (K > 0 ? a[q] : arow[q]) * b[q];
    } else if constexpr (K > 0) {
      c = a[0] * b[0];
#pragma unroll
      for (int q = 1; q < K; ++q) c = fma(a[q], b[q], c);
    } else {
      c = arow[0] * b[0];
      for (int q = 1; q < kdim; ++q) c = fma(arow[q], b[q], c);
    }
    const double Kv = kval_f64<TYPE, uses_exp<TYPE>>(c, aa, bb, hi, hj, P, etab);
    cptr<double> wr = W + (long)wmap[j] * nr;
#pragma unroll
    for (int r = 0; r < RMAX; ++r)
      if (r < nr) acc[r] = fma(Kv, wr[r], acc[r]);
  }

  if (A.nsplit == 1) {
    if (i < m) {
      double *ur = p.u + (long)p.umap[i] * nr;
#pragma unroll
      for (int r = 0; r < RMAX; ++r)
        if (r < nr) ur[r] += acc[r];
    }
    return;
  }
  if (i < m) {
    double *pr = A.partial + ((long)split * m + i) * nr;
#pragma unroll
    for (int r = 0; r < RMAX; ++r)
      if (r < nr) pr[r] = acc[r];
  }
  // the splits' partials are added in split order by a reduction kernel
  // launched next (no device-scope fence here: on gfx950 it would write back
  // and invalidate this XCD's L2)
}

// Matrix-core form (9 <= k <= 64, every type but Laplace, nrhs <= 16 NT): the
// cross term is formed with the sources as the A operand and the targets as the
// B operand of v_mfma_f64_16x16x4f64, so lane (p16, g) ends up holding
// K(target p16, source 4v + g) for v = 0..3 -- exactly the A operand of the
// next product, chunk v of U[targets][rhs] += K[targets][sources] W[sources][rhs].
// Per 16 x 16 block that is 4 KC cross-term MFMAs, the 4 kernel values per lane
// on the VALU, and 4 NT MFMAs for the right-hand sides; nothing is transposed.
// The last NX right-hand sides (nrhs = 16 NT + NX, NX <= 3) are summed on the
// VALU instead of a further 16-wide tile: 4 NX FMAs per lane per block, and the
// 4 lane groups (sources 4v + g) are added by two lane shuffles at the end.
// A wave owns RB groups of 16 targets; splits are reduced as in the VALU form.
constexpr int MRM_RB = 2;
constexpr int MRM_ROWS = 4 * MRM_RB * 16;  // targets per workgroup

template <int TYPE, int KC, int NT, bool EXACT, int NX>
__global__ __launch_bounds__(MR_WG, (NX >= 3 && KC == 1) ? 2 : 1) void multirhs_mfma_kernel(MrhsArgs A) {
  typedef double d4 __attribute__((ext_vector_type(4)));
  constexpr int RB = MRM_RB;
  const Problem &p = A.p;
  const int m = p.m, k = p.k, nr = A.nrhs;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int p16 = lane & 15, g = lane >> 4;
  const int rblk = blockIdx.x, split = blockIdx.y;
  const int row0 = rblk * MRM_ROWS + wave * (RB * 16);
  __shared__ double2 etab[64];
  if constexpr (uses_exp<TYPE>) {
    load_exp_table(etab);
    __syncthreads();
  }

  // targets (B operand): lane (p16, g) -> target row0 + 16 rb + p16
  double tf[RB][KC][4], aa[RB], hi[RB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    const int r = row0 + 16 * rb + p16;
    const int ai = p.amap[r < m ? r : m - 1];
    load_frag<KC, EXACT>(p.XA, (long)ai * k, k, g, tf[rb]);
    aa[rb] = p.XA2[ai];
    hi[rb] = (TYPE == VAR_BANDWIDTH) ? -0.5 * p.hi[ai] : 0.0;
  }
  d4 U[RB][NT];  // lane (p16, g): targets 16 rb + g + 4v' x rhs 16 t + p16
  double ux[RB][NX > 0 ? NX : 1];  // right-hand sides 16 NT + x, lane (p16, g)'s sources of target p16
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
    for (int x = 0; x < NX; ++x) ux[rb][x] = 0.0;
#pragma unroll
    for (int t = 0; t < NT; ++t) U[rb][t] = d4{0.0, 0.0, 0.0, 0.0};
  }

  const int j0 = split * A.cols_per_split;
  const int j1 = min(p.n, j0 + A.cols_per_split);
  const KerParams P = p.P;
  // one 16-source block: the A-operand fragment (lane p16 = source), and for
  // the sources 4v + g of the lane's results their norms, bandwidths and
  // weight rows (the B operand of the right-hand-side product)
  struct SB {
    double f[KC][4];
    double bb[4], hj[4];
    double wv[4][NT];
    double wx[4][NX > 0 ? NX : 1];  // weights 16 NT + x of the sources 4v + g
  };
  // a block's map entries load two blocks ahead of the data they address, so
  // no block waits on two dependent round trips
  struct SI {
    int fb;     // the lane's fragment source (row of XB)
    int bj[4];  // sources 4v + g
    int wi[4];  // their weight rows (-1: past the slice)
  };
  auto fetch_idx = [&](int jb, SI &x) {
    const int col = jb + p16;
    x.fb = p.bmap[col < j1 ? col : j1 - 1];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int src = jb + 4 * v + g;
      const bool ok = src < j1;
      x.bj[v] = p.bmap[ok ? src : j1 - 1];
      x.wi[v] = ok ? p.wmap[src] : -1;
    }
  };
  auto fetch = [&](const SI &x, SB &b) {
    load_frag<KC, EXACT>(p.XB, (long)x.fb * k, k, g, b.f);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int bj = x.bj[v];
      b.bb[v] = p.XB2[bj];
      b.hj[v] = (TYPE == VAR_BANDWIDTH) ? p.hj[bj] : 0.0;
      const long wrow = x.wi[v] >= 0 ? (long)x.wi[v] * nr : -1;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int r = 16 * t + p16;
        b.wv[v][t] = (wrow >= 0 && r < nr) ? p.w[wrow + r] : 0.0;
      }
#pragma unroll
      for (int x = 0; x < NX; ++x) b.wx[v][x] = wrow >= 0 ? p.w[wrow + 16 * NT + x] : 0.0;
    }
  };
  auto block = [&](const SB &b) {
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      d4 c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int cc = 0; cc < KC; ++cc)
#pragma unroll
        for (int t = 0; t < 4; ++t)
          c = __builtin_amdgcn_mfma_f64_16x16x4f64(b.f[cc][t], tf[rb][cc][t], c, 0, 0, 0);
      double Kv[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) Kv[v] = kval_f64<TYPE, uses_exp<TYPE>>(c[v], aa[rb], b.bb[v], hi[rb], b.hj[v], P, etab);
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          U[rb][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(Kv[v], b.wv[v][t], U[rb][t], 0, 0, 0);
#pragma unroll
      for (int x = 0; x < NX; ++x)
#pragma unroll
        for (int v = 0; v < 4; ++v) ux[rb][x] = fma(Kv[v], b.wx[v][x], ux[rb][x]);
    }
  };
  // b0 / b1: data of blocks jb / jb + 16; i0 / i1: map entries of the block
  // whose data loads next into b0 / b1
  SB b0, b1;
  SI i0, i1;
  if (j0 < j1) {
    fetch_idx(j0, i0);
    fetch(i0, b0);
    if (j0 + 16 < j1) fetch_idx(j0 + 16, i1);
  }
  for (int jb = j0; jb < j1; jb += 32) {
    if (jb + 16 < j1) {
      fetch(i1, b1);
      if (jb + 32 < j1) fetch_idx(jb + 32, i0);
    }
    block(b0);
    if (jb + 16 >= j1) break;
    if (jb + 32 < j1) {
      fetch(i0, b0);
      if (jb + 48 < j1) fetch_idx(jb + 48, i1);
    }
    block(b1);
  }

  // ---- results: lane (p16, g) holds target 16 rb + g + 4v, rhs 16 t + p16;
  // the extra right-hand side sums its 4 lane groups into lanes g = 0
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int x = 0; x < NX; ++x) {
      ux[rb][x] += __shfl_xor(ux[rb][x], 16);
      ux[rb][x] += __shfl_xor(ux[rb][x], 32);
    }
  if (A.nsplit == 1) {
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = row0 + 16 * rb + g + 4 * v, r = 16 * t + p16;
          if (i < m && r < nr) p.u[(long)p.umap[i] * nr + r] += U[rb][t][v];
        }
      if (NX > 0) {
        const int i = row0 + 16 * rb + p16;
        if (g == 0 && i < m) {
#pragma unroll
          for (int x = 0; x < NX; ++x) p.u[(long)p.umap[i] * nr + 16 * NT + x] += ux[rb][x];
        }
      }
    }
    return;
  }
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = row0 + 16 * rb + g + 4 * v, r = 16 * t + p16;
        if (i < m && r < nr) A.partial[((long)split * m + i) * nr + r] = U[rb][t][v];
      }
    if (NX > 0) {
      const int i = row0 + 16 * rb + p16;
      if (g == 0 && i < m) {
#pragma unroll
        for (int x = 0; x < NX; ++x) A.partial[((long)split * m + i) * nr + 16 * NT + x] = ux[rb][x];
      }
    }
  }
  // the splits' partials are added in split order by a reduction kernel
  // launched next (no device-scope fence here: on gfx950 it would write back
  // and invalidate this XCD's L2)
}

template <int TYPE, int KC, int NT, int NX>
static hipError_t mrm_go(const MrhsArgs &A, int rblocks, hipStream_t s) {
  dim3 grid(rblocks, A.nsplit);
  const bool exact = A.p.k == 16 * KC && (uintptr_t)A.p.XA % 16 == 0 &&
                     (uintptr_t)A.p.XB % 16 == 0;
  if (exact)
    hipLaunchKernelGGL((multirhs_mfma_kernel<TYPE, KC, NT, true, NX>), grid, dim3(MR_WG), 0, s, A);
  else
    hipLaunchKernelGGL((multirhs_mfma_kernel<TYPE, KC, NT, false, NX>), grid, dim3(MR_WG), 0, s, A);
  return hipGetLastError();
}

template <int TYPE, int NT, int NX>
static hipError_t mrm_kc(const MrhsArgs &A, int rblocks, hipStream_t s) {
  switch ((A.p.k + 15) / 16) {
    case 1: return mrm_go<TYPE, 1, NT, NX>(A, rblocks, s);
    case 2: return mrm_go<TYPE, 2, NT, NX>(A, rblocks, s);
    case 3: return mrm_go<TYPE, 3, NT, NX>(A, rblocks, s);
    default: return mrm_go<TYPE, 4, NT, NX>(A, rblocks, s);
  }
}

// 16-wide tiles on the matrix cores; a 17th right-hand side on the VALU, and
// for the Gaussian kernels (whose gradients take 17 to 19) up to three
template <int TYPE>
static hipError_t mrm_nt(const MrhsArgs &A, int rblocks, hipStream_t s) {
  if (A.nrhs <= 16) return mrm_kc<TYPE, 1, 0>(A, rblocks, s);
  if (A.nrhs == 17) return mrm_kc<TYPE, 1, 1>(A, rblocks, s);
  if constexpr (TYPE == GAUSSIAN || TYPE == VAR_BANDWIDTH) {
    if (A.nrhs == 18) return mrm_kc<TYPE, 1, 2>(A, rblocks, s);
    if (A.nrhs == 19) return mrm_kc<TYPE, 1, 3>(A, rblocks, s);
  }
  return mrm_kc<TYPE, 2, 0>(A, rblocks, s);
}

template <int TYPE, int RMAX>
static hipError_t mr_k(const MrhsArgs &A, int rblocks, hipStream_t s) {
  dim3 grid(rblocks, A.nsplit);
  switch (A.p.k) {
    case 3: hipLaunchKernelGGL((multirhs_kernel<TYPE, 3, RMAX>), grid, dim3(MR_WG), 0, s, A); break;
    case 16: hipLaunchKernelGGL((multirhs_kernel<TYPE, 16, RMAX>), grid, dim3(MR_WG), 0, s, A); break;
    default: hipLaunchKernelGGL((multirhs_kernel<TYPE, 0, RMAX>), grid, dim3(MR_WG), 0, s, A); break;
  }
  return hipGetLastError();
}

template <int TYPE>
static hipError_t mr_r(const MrhsArgs &A, int rblocks, hipStream_t s) {
  if (A.nrhs <= 8) return mr_k<TYPE, 8>(A, rblocks, s);
  if (A.nrhs <= 24) return mr_k<TYPE, 24>(A, rblocks, s);
  return mr_k<TYPE, 72>(A, rblocks, s);
}

int enqueue_multirhs(const Problem &p, int nrhs, hipStream_t s) {
  if (nrhs < 1 || nrhs > 72) return KS_ERR_INVALID_ARG;
  // the matrix-core form for 9 <= k <= 64 (not Laplace) and up to 32 right-hand sides
  const bool mfma = mfma_f64_eligible(p.type, p.k) && nrhs <= 32;
  FusedPlan plan = mfma ? plan_split(p.m, p.n, MRM_ROWS, 2048, 1024)
                        : plan_split(p.m, p.n, MR_WG, 2048, 128);
  double *partial = nullptr;
  ScratchLease lease(s);
  if (plan.nsplit > 1) {
    int rc = lease.partials((size_t)plan.nsplit * p.m * nrhs, &partial);
    if (rc != KS_OK) return rc;
  }
  MrhsArgs A;
  A.p = p;
  A.nrhs = nrhs;
  A.cols_per_split = plan.cols_per_split;
  A.nsplit = plan.nsplit;
  A.partial = partial;
  hipError_t e;
  if (mfma) {
    switch (p.type) {
      case GAUSSIAN: e = mrm_nt<GAUSSIAN>(A, plan.row_blocks, s); break;
      case POLYNOMIAL: e = mrm_nt<POLYNOMIAL>(A, plan.row_blocks, s); break;
      case VAR_BANDWIDTH: e = mrm_nt<VAR_BANDWIDTH>(A, plan.row_blocks, s); break;
      case TANH: e = mrm_nt<TANH>(A, plan.row_blocks, s); break;
      case QUARTIC: e = mrm_nt<QUARTIC>(A, plan.row_blocks, s); break;
      case MULTIQUADRATIC: e = mrm_nt<MULTIQUADRATIC>(A, plan.row_blocks, s); break;
      default: e = mrm_nt<EPANECHNIKOV>(A, plan.row_blocks, s); break;
    }
    if (e == hipSuccess && plan.nsplit > 1)
      e = launch_split_reduce_rhs(partial, plan.nsplit, p.m, nrhs, p.u, p.umap, s);
    return e == hipSuccess ? KS_OK : KS_ERR_DEVICE;
  }
  switch (p.type) {
    case GAUSSIAN: e = mr_r<GAUSSIAN>(A, plan.row_blocks, s); break;
    case POLYNOMIAL: e = mr_r<POLYNOMIAL>(A, plan.row_blocks, s); break;
    case LAPLACE: e = mr_r<LAPLACE>(A, plan.row_blocks, s); break;
    case VAR_BANDWIDTH: e = mr_r<VAR_BANDWIDTH>(A, plan.row_blocks, s); break;
    case TANH: e = mr_r<TANH>(A, plan.row_blocks, s); break;
    case QUARTIC: e = mr_r<QUARTIC>(A, plan.row_blocks, s); break;
    case MULTIQUADRATIC: e = mr_r<MULTIQUADRATIC>(A, plan.row_blocks, s); break;
    default: e = mr_r<EPANECHNIKOV>(A, plan.row_blocks, s); break;
  }
  if (e == hipSuccess && plan.nsplit > 1)
    e = launch_split_reduce_rhs(partial, plan.nsplit, p.m, nrhs, p.u, p.umap, s);
  return e == hipSuccess ? KS_OK : KS_ERR_DEVICE;
}

// ------------------------------------------------------------------ Gaussian gradient

// RHS of one pass: [s x_1 .. s x_k, s] for the points x (scale s = w or g)
__global__ void grad_rhs(const double *X, const double *s, int n, int k, double *W) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int R = k + 1;
  const double sj = s[j];
  for (int q = 0; q < k; ++q) W[j * R + q] = sj * X[j * k + q];
  W[j * R + k] = sj;
}

// dXA, and per-row dh terms (summed afterwards in a fixed order); either output
// may be NULL
__global__ void grad_fin_a(const double *XA, const double *XA2, const double *g, const double *U1,
                           int m, int k, double h, double *dXA, double *dh_rows) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int R = k + 1;
  const double *u = U1 + i * R;
  const double ui = u[k];
  const double f = -g[i] / (h * h);
  double av = 0.0;
  for (int q = 0; q < k; ++q) {
    const double a = XA[i * k + q], V = u[q];
    if (dXA) dXA[i * k + q] = f * (ui * a - V);
    av = fma(a, V, av);
  }
  if (dh_rows) dh_rows[i] = g[i] * (XA2[i] * ui - 2.0 * av);
}

// dXB, dw, and per-source dh terms w_j |b_j|^2 dw_j; any output may be NULL
__global__ void grad_fin_b(const double *XB, const double *XB2, const double *w, const double *U2,
                           int n, int k, double h, double *dXB, double *dw, double *dh_rows) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int R = k + 1;
  const double *u = U2 + j * R;
  const double f = w[j] / (h * h), dwj = u[k];
  if (dw) dw[j] = dwj;
  if (dXB)
    for (int q = 0; q < k; ++q) dXB[j * k + q] = f * (u[q] - dwj * XB[j * k + q]);
  if (dh_rows) dh_rows[j] = w[j] * XB2[j] * dwj;
}

// deterministic sum of the per-point dh terms: one workgroup of 1024, each lane
// with 4 interleaved accumulators (independent loads in flight), fixed tree
constexpr int GF_WG = 1024;
__global__ __launch_bounds__(GF_WG) void grad_fin_h(const double *dh_rows, int m, double h,
                                                    double *dh) {
  __shared__ double s[GF_WG];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  int i = threadIdx.x;
  for (; i + 3 * GF_WG < m; i += 4 * GF_WG)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] += dh_rows[i + q * GF_WG];
  for (int q = 0; i < m; i += GF_WG, ++q) acc[q] += dh_rows[i];
  s[threadIdx.x] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  __syncthreads();
  for (int w = GF_WG / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *dh = s[0] / (h * h * h);
}

__global__ void zero_f64(double *v, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = 0.0;
}

__global__ void iota_i32(int *v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

}  // namespace gsks

using namespace gsks;

extern "C" int dgsks_multirhs_device(const ks_t *kernel, int m, int n, int k, int nrhs,
                                     double *U, const int *umap, const double *XA,
                                     const double *XA2, const int *amap, const double *XB,
                                     const double *XB2, const int *bmap, const double *W,
                                     const int *wmap, void *stream) {
  if (!kernel || m < 0 || n < 0 || k < 0 || nrhs < 1 || nrhs > 72)
    return raise_error(KS_ERR_INVALID_ARG, false, "dgsks_multirhs_device(): invalid argument\n");
  if (m == 0 || n == 0 || k == 0) return KS_OK;
  if (ensure_device() != KS_OK)
    return raise_error(KS_ERR_DEVICE, false, "dgsks_multirhs_device(): no GPU available\n");
  Problem p;
  int rc = make_params(kernel, k, p.P, true);
  if (rc != KS_OK && rc != KS_ERR_LAPLACE_K) return rc;
  const void *arrays[] = {U, umap, XA, XA2, amap, XB, XB2, bmap, W, wmap};
  for (const void *a : arrays)
    if (!a || !device_ptr_cached(a))
      return raise_error(KS_ERR_MIXED_POINTERS, false,
                         "dgsks_multirhs_device(): every array must be device memory\n");
  p.m = m; p.n = n; p.k = k; p.type = ker_type(kernel);
  p.u = U; p.umap = umap; p.XA = XA; p.XA2 = XA2; p.amap = amap;
  p.XB = XB; p.XB2 = XB2; p.bmap = bmap; p.w = W; p.wmap = wmap;
  p.hi = kernel->hi; p.hj = kernel->hj;
  if (p.type == KS_GAUSSIAN_VAR_BANDWIDTH && (!p.hi || !p.hj)) return KS_ERR_NO_BANDWIDTH;
  rc = enqueue_multirhs(p, nrhs, (hipStream_t)stream);
  return rc == KS_OK ? KS_OK : raise_error(rc, false, "dgsks_multirhs_device(): %s\n", err_string(rc));
}

namespace gsks {
namespace {

// Gradient state: [m][k + 1] rows (V_i = sum_j K_ij w_j b_j, then
// u_i = sum_j K_ij w_j, in the frame the precision works in), then the k
// coordinates of that frame's origin (zeros for FP64, the mean of both point
// sets for the mixed form).
size_t grad_state_elems(int m, int k) { return (size_t)m * (k + 1) + k; }

int grad_check(const char *name, int m, int n, int k, double h, int precision, bool &mixed) {
  if (m <= 0 || n <= 0 || k <= 0 || k > 70 || !(h > 0.0))
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): invalid argument\n", name);
  mixed = precision == KS_PREC_MIXED_FP32;
  if (!mixed && precision != KS_PREC_FP64)
    return raise_error(KS_ERR_INVALID_ARG, false,
                       "%s(): precision must be KS_PREC_FP64 or KS_PREC_MIXED_FP32\n", name);
  if (mixed && k + 1 > kMixedMrhsMaxRhs)
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): the mixed gradient takes k <= %d\n",
                       name, kMixedMrhsMaxRhs - 1);
  if (ensure_device() != KS_OK)
    return raise_error(KS_ERR_DEVICE, false, "%s(): no GPU available\n", name);
  return KS_OK;
}

__global__ void grad_take_u(const double *state, int m, int k, double *u) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) u[i] = state[i * (k + 1) + k];
}

// The coordinates both passes work in and their squared norms.  FP64: the
// caller's points.  Mixed: both sets translated by `origin` (written here from
// the mean when `centre_out`, else read), so the FP32 cross terms see the
// smallest coordinates the data allows; the gradient does not change.
struct GradFrame {
  const double *A = nullptr, *B = nullptr;
  double *A2 = nullptr, *B2 = nullptr, *Ac = nullptr, *Bc = nullptr, *zc = nullptr;
  double *cwork = nullptr;
  int *ident = nullptr;

  void layout(Carve &c, int m, int n, int k, bool mixed, bool centre_out) {
    A2 = c.take<double>(m);
    B2 = c.take<double>(n);
    ident = c.take<int>(m > n ? m : n);
    if (mixed) {
      Ac = c.take<double>((size_t)m * k);
      Bc = c.take<double>((size_t)n * k);
      zc = c.take<double>(64);
      if (centre_out) cwork = c.take<double>(centre_work_elems((long)m + n));
    }
  }

  int setup(const double *XA, int m, const double *XB, int n, int k, bool mixed, double *origin,
            bool centre_out, hipStream_t s) {
    const int mx = m > n ? m : n;
    int rc = KS_OK;
    auto chain = [&](int r) {
      if (rc == KS_OK) rc = r;
    };
    hipLaunchKernelGGL(iota_i32, dim3((mx + 255) / 256), dim3(256), 0, s, ident, mx);
    A = XA;
    B = XB;
    if (mixed) {
      hipLaunchKernelGGL(zero_f64, dim3(1), dim3(64), 0, s, zc, 64l);
      if (centre_out) chain(enqueue_centre(XA, m, XB, n, k, ident, cwork, origin, s));
      chain(enqueue_shift(XA, m, k, origin, Ac, s));
      chain(enqueue_shift(XB, n, k, origin, Bc, s));
      A = Ac;
      B = Bc;
    } else if (centre_out) {
      hipLaunchKernelGGL(zero_f64, dim3((k + 63) / 64), dim3(64), 0, s, origin, (long)k);
    }
    chain(ks_sqnorm_device(A, m, k, A2, s));
    chain(ks_sqnorm_device(B, n, k, B2, s));
    return rc;
  }

  // One summation of R = k + 1 right-hand sides from sources S to targets T.
  int pass(bool mixed, const double *T, const double *T2, int mt, const double *S,
           const double *S2, int ns, int k, double h, const double *W, double *U,
           hipStream_t s) const {
    const int R = k + 1;
    hipLaunchKernelGGL(zero_f64, dim3((unsigned)(((long)mt * R + 255) / 256)), dim3(256), 0, s,
                       U, (long)mt * R);
    Problem p;
    p.m = mt; p.n = ns; p.k = k; p.type = GAUSSIAN;
    p.P = KerParams{-1.0 / (2.0 * h * h), 0.0, 0.0, 0, POW_GENERAL};
    p.u = U; p.umap = ident; p.XA = T; p.XA2 = T2; p.amap = ident;
    p.XB = S; p.XB2 = S2; p.bmap = ident; p.w = W; p.wmap = ident;
    p.hi = p.hj = nullptr;
    return mixed ? enqueue_multirhs_mixed(p, R, zc, s) : enqueue_multirhs(p, R, s);
  }
};

// Pass 1: state = K [w b, w] (and its origin); u = the last column.
int grad_forward(int m, int n, int k, const double *XA, const double *XB, const double *w,
                 double h, double *u, double *state, bool mixed, hipStream_t s) {
  GradFrame F;
  double *W1 = nullptr;
  char *buf = nullptr;
  int rc = carve_alloc([&](Carve &c) {
    F.layout(c, m, n, k, mixed, true);
    W1 = c.take<double>((size_t)n * (k + 1));
  }, &buf, s);
  if (rc != KS_OK) return rc;
  double *origin = state + (size_t)m * (k + 1);
  rc = F.setup(XA, m, XB, n, k, mixed, origin, true, s);
  hipLaunchKernelGGL(grad_rhs, dim3((n + 255) / 256), dim3(256), 0, s, F.B, w, n, k, W1);
  if (rc == KS_OK) rc = F.pass(mixed, F.A, F.A2, m, F.B, F.B2, n, k, h, W1, state, s);
  if (rc == KS_OK && u)
    hipLaunchKernelGGL(grad_take_u, dim3((m + 255) / 256), dim3(256), 0, s, state, m, k, u);
  if (rc == KS_OK && hipGetLastError() != hipSuccess) rc = KS_ERR_DEVICE;
  (void)hipFreeAsync(buf, s);
  return rc;
}

// Pass 2: [Q, dw] = K^T [g a, g], then the finishing kernels.  Without a state
// only dXB and dw can be formed (dXA and dh need pass 1).
int grad_backward(int m, int n, int k, const double *XA, const double *XB, const double *w,
                  const double *g, double h, const double *state, double *dXA, double *dXB,
                  double *dw, double *dh, bool mixed, hipStream_t s) {
  const bool need_a = dXA || dh;
  if (need_a && !state) return KS_ERR_INVALID_ARG;
  if (!need_a && !dXB && !dw) return KS_OK;
  const int R = k + 1;
  GradFrame F;
  double *W2 = nullptr, *U2 = nullptr, *hrows = nullptr, *cen = nullptr;
  char *buf = nullptr;
  const bool centre_out = mixed && !state;  // no stored origin: take the mean again
  int rc = carve_alloc([&](Carve &c) {
    F.layout(c, m, n, k, mixed, centre_out);
    W2 = c.take<double>((size_t)m * R);
    U2 = c.take<double>((size_t)n * R);
    if (dh) hrows = c.take<double>((size_t)m + n);
    if (centre_out) cen = c.take<double>(64);
  }, &buf, s);
  if (rc != KS_OK) return rc;
  double *origin = state ? const_cast<double *>(state) + (size_t)m * R : cen;
  rc = F.setup(XA, m, XB, n, k, mixed, origin, centre_out, s);
  hipLaunchKernelGGL(grad_rhs, dim3((m + 255) / 256), dim3(256), 0, s, F.A, g, m, k, W2);
  if (rc == KS_OK) rc = F.pass(mixed, F.B, F.B2, n, F.A, F.A2, m, k, h, W2, U2, s);
  if (rc == KS_OK) {
    if (need_a)
      hipLaunchKernelGGL(grad_fin_a, dim3((m + 255) / 256), dim3(256), 0, s, F.A, F.A2, g, state,
                         m, k, h, dXA, hrows);
    hipLaunchKernelGGL(grad_fin_b, dim3((n + 255) / 256), dim3(256), 0, s, F.B, F.B2, w, U2, n, k,
                       h, dXB, dw, dh ? hrows + m : nullptr);
    if (dh) hipLaunchKernelGGL(grad_fin_h, dim3(1), dim3(GF_WG), 0, s, hrows, m + n, h, dh);
    if (hipGetLastError() != hipSuccess) rc = KS_ERR_DEVICE;
  }
  (void)hipFreeAsync(buf, s);
  return rc;
}

}  // namespace
}  // namespace gsks

// Gradient of L = g . (K w) for the Gaussian kernel with bandwidth h and identity
// maps; all arrays device memory, XA [m][k], XB [n][k].  Outputs dXA [m][k],
// dXB [n][k], dw [n], dh [1] (any output may be NULL).  Temporary buffers come from
// the stream-ordered pool.  KS_PREC_MIXED_FP32 (k <= 32) translates both point
// sets by their mean (the gradient does not change) and runs the two passes
// through the mixed multi-right-hand-side kernel; the right-hand sides and the
// finishing kernels then work on the centred FP64 coordinates, so u a - V and
// Q - dw b are formed from the smallest coordinates the data allows.
extern "C" int ks_gauss_grad_device(int m, int n, int k, const double *XA, const double *XB,
                                    const double *w, const double *g, double h, double *dXA,
                                    double *dXB, double *dw, double *dh, int precision,
                                    void *stream) {
  const char *name = "ks_gauss_grad_device";
  bool mixed = false;
  int rc = grad_check(name, m, n, k, h, precision, mixed);
  if (rc != KS_OK) return rc;
  if (!XA || !XB || !w || !g)
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): invalid argument\n", name);
  hipStream_t s = (hipStream_t)stream;
  double *state = nullptr;
  if (dXA || dh) {
    if (hipMallocAsync((void **)&state, sizeof(double) * grad_state_elems(m, k), s) !=
        hipSuccess) {
      (void)hipGetLastError();
      return raise_error(KS_ERR_ALLOC, false, "%s(): allocation failed\n", name);
    }
    rc = grad_forward(m, n, k, XA, XB, w, h, nullptr, state, mixed, s);
  }
  if (rc == KS_OK) rc = grad_backward(m, n, k, XA, XB, w, g, h, state, dXA, dXB, dw, dh, mixed, s);
  if (state) (void)hipFreeAsync(state, s);
  if (rc != KS_OK) return raise_error(rc, false, "%s(): %s\n", name, err_string(rc));
  return KS_OK;
}

// The first half of ks_gauss_grad_device for a training step: u = K w (u may be
// NULL) and the gradient state (m (k + 1) + k doubles) that
// ks_gauss_grad_backward_device reads, from one summation of k + 1 right-hand
// sides, so the step costs two summations in all.
extern "C" int ks_gauss_grad_forward_device(int m, int n, int k, const double *XA,
                                            const double *XB, const double *w, double h,
                                            double *u, double *state, int precision,
                                            void *stream) {
  const char *name = "ks_gauss_grad_forward_device";
  bool mixed = false;
  int rc = grad_check(name, m, n, k, h, precision, mixed);
  if (rc != KS_OK) return rc;
  if (!XA || !XB || !w || !state)
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): invalid argument\n", name);
  rc = grad_forward(m, n, k, XA, XB, w, h, u, state, mixed, (hipStream_t)stream);
  if (rc != KS_OK) return raise_error(rc, false, "%s(): %s\n", name, err_string(rc));
  return KS_OK;
}

// The second half: the gradient of L = g . u from the state the forward half
// wrote with the same points, weights, h and precision.  state may be NULL
// when dXA and dh are (dXB and dw need the second summation only).
extern "C" int ks_gauss_grad_backward_device(int m, int n, int k, const double *XA,
                                             const double *XB, const double *w, const double *g,
                                             double h, const double *state, double *dXA,
                                             double *dXB, double *dw, double *dh, int precision,
                                             void *stream) {
  const char *name = "ks_gauss_grad_backward_device";
  bool mixed = false;
  int rc = grad_check(name, m, n, k, h, precision, mixed);
  if (rc != KS_OK) return rc;
  if (!XA || !XB || !w || !g || (!state && (dXA || dh)))
    return raise_error(KS_ERR_INVALID_ARG, false,
                       "%s(): invalid argument (dXA and dh need the forward state)\n", name);
  rc = grad_backward(m, n, k, XA, XB, w, g, h, state, dXA, dXB, dw, dh, mixed,
                     (hipStream_t)stream);
  if (rc != KS_OK) return raise_error(rc, false, "%s(): %s\n", name, err_string(rc));
  return KS_OK;
}
