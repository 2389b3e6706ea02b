// Symmetric FP64 fused summation.  When the sources are the targets -- m == n,
// amap == bmap, XA == XB, XA2 == XB2 and a symmetric kernel (for the
// variable-bandwidth Gaussian, hi == hj) -- K(a_i, b_j) = K(a_j, b_i), so every
// pair of different row blocks is evaluated once for both of its terms:
//   u[umap[i]] += K_ij w[wmap[j]]   and   u[umap[j]] += K_ij w[wmap[i]].
// (The reference evaluates both orders, frame/dgsks.c:367-776; its
// omp_dgsks_list_symmetric only shares the point table,
// frame/omp_dgsks_list.cpp:114-166.)  umap and wmap may be anything.
//
// Mapping (row blocks of B = 256 R targets, lane = R rows as in
// fused_f64_kernel, sources as wave-uniform scalar loads):
//   * row block I walks only sources j >= I*B: the sources of its own block
//     ("diagonal", every pair, row sums only) and those of later blocks
//     ("upper", row sums plus column sums);
//   * a column sum sum_i K_ij w_i over the workgroup's rows is reduced within
//     each 16-lane row by DPP per source; lane 16r + (t % 16) of row r keeps
//     the row's share of source t, and the four rows are added with two lane
//     swaps once per 16 sources; the four waves' values meet in LDS once per
//     64 sources, and one value per (row block, source) goes to colp[I][j];
//   * row sums use split-source partials, as fused_f64_kernel does (splits
//     before a block's diagonal hold no work for it);
//   * sym_colsum_kernel then adds sum_{I < J} colp[I][j] (J = j's block) to
//     u[umap[j]] in block order.  Both steps are deterministic.
// Per pair this costs about half the evaluations of the full sum plus one
// FMA and 1/64 of a wave reduction per lane.
#include <hip/hip_runtime.h>

#include "gsks_kernels.h"
#include "gsks_launch.h"

namespace gsks {

constexpr int SWG = 256;

struct SymArgs {
  Problem p;
  int cols_per_split;  // a multiple of the row block
  int nsplit;
  double *partial;     // [nsplit][m] row partials
  double *colp;        // [row blocks][m] column partials of the upper sources
};

template <int TYPE, int K, int R>
__global__ __launch_bounds__(SWG) void fused_f64_sym_kernel(SymArgs A) {
  const Problem &p = A.p;
  const int m = p.m;  // == n
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int rblk = blockIdx.x, split = blockIdx.y;
  constexpr int B = SWG * R;
  const int row0 = rblk * B;
  const int first = row0 / A.cols_per_split;  // the split holding this block's diagonal
  if (split < first) return;                  // earlier sources are other blocks' upper part
  __shared__ double2 etab[64];
  if constexpr (uses_exp<TYPE>) {
    load_exp_table(etab);
    __syncthreads();
  }

  // ---- target rows in registers, with their own weights for the column sums
  double a[R][K], aa[R], hi[R], wr[R], acc[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = row0 + q * SWG + tid;
    const int ii = i < m ? i : m - 1;
    const int ai = p.amap[ii];
    const double *arow = p.XA + (long)ai * K;
#pragma unroll
    for (int pp = 0; pp < K; ++pp) a[q][pp] = arow[pp];
    aa[q] = p.XA2[ai];
    hi[q] = (TYPE == VAR_BANDWIDTH) ? -0.5 * p.hi[ai] : 0.0;
    wr[q] = i < m ? p.w[p.wmap[ii]] : 0.0;  // padding rows add nothing to a column
    acc[q] = 0.0;
  }
  const int c0 = split * A.cols_per_split;
  const int c1 = min(m, c0 + A.cols_per_split);
  cptr<int> bmap = to_const(p.bmap);
  cptr<int> wmap = to_const(p.wmap);
  cptr<double> XB = to_const(p.XB);
  cptr<double> XB2 = to_const(p.XB2);
  cptr<double> W = to_const(p.w);
  cptr<double> HJ = to_const(p.hj);
  const KerParams P = p.P;

  // K(a_q, b_j) for the lane's R rows; the row sums take w_j
  auto pair_values = [&](int j, double (&kv)[R]) {
    const int bj = bmap[j];
    const double wj = W[wmap[j]];
    const double bb = XB2[bj];
    const double hj = (TYPE == VAR_BANDWIDTH) ? HJ[bj] : 0.0;
    cptr<double> b = XB + (long)bj * K;
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
      kv[q] = kval_f64<TYPE, uses_exp<TYPE>>(c, aa[q], bb, hi[q], hj, P, etab);
      acc[q] = fma(kv[q], wj, acc[q]);
    }
  };

  // ---- diagonal sources: every pair, row sums only
  const int dlo = max(c0, row0), dhi = min(c1, row0 + B);
#pragma unroll 2
  for (int j = dlo; j < dhi; ++j) {
    double kv[R];
    pair_values(j, kv);
  }

  // ---- upper sources: row sums and column sums, 64 sources at a time
  __shared__ double s_col[4][64];
  const int ulo = max(c0, row0 + B);
  for (int jc = ulo; jc < c1; jc += 64) {
    const int nj = min(64, c1 - jc);
    // part[g]: this lane's DPP row's share of source jc + 16 g + (lane % 16)
    double part[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int tn = min(16, nj - 16 * g);
#pragma unroll 2
      for (int t = 0; t < tn; ++t) {
        double kv[R];
        pair_values(jc + 16 * g + t, kv);
        double cs = 0.0;
#pragma unroll
        for (int q = 0; q < R; ++q) cs = fma(kv[q], wr[q], cs);
        cs = row_sum16(cs);
        part[g] = ((lane & 15) == t) ? cs : part[g];
      }
    }
    // add the four DPP rows: lane l then holds the wave's sum of source jc + 16 g + l % 16
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      part[g] += __shfl_xor(part[g], 16);
      part[g] += __shfl_xor(part[g], 32);
    }
    if (lane < 16) {
#pragma unroll
      for (int g = 0; g < 4; ++g) s_col[wave][16 * g + lane] = part[g];
    }
    __syncthreads();
    if (tid < nj)
      A.colp[(long)rblk * m + jc + tid] = (s_col[0][tid] + s_col[1][tid]) + (s_col[2][tid] + s_col[3][tid]);
    __syncthreads();
  }

  // ---- row results: the splits from `first` on, added in split order
  if (A.nsplit == 1) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = row0 + q * SWG + tid;
      if (i < m) p.u[p.umap[i]] += acc[q];
    }
    return;
  }
  // the splits' partials (from `first` on) are added in split order by
  // sym_colsum_kernel, launched next, before the column partials (no
  // device-scope fence here: on gfx950 it would write back and invalidate this
  // XCD's L2)
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = row0 + q * SWG + tid;
    if (i < m) A.partial[(long)split * m + i] = acc[q];
  }
}

// The last step of a symmetric sum, one thread per target j of row block J:
// with split partials (nsplit > 1), u[umap[j]] += the row partials of the
// splits from J's first on, in split order; then u[umap[j]] += the column
// partials of the row blocks before J, in order.
__global__ void sym_colsum_kernel(const Problem p, int B, const double *__restrict__ partial,
                                  int nsplit, int cols_per_split,
                                  const double *__restrict__ colp) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= p.m) return;
  const int J = (int)(j / B);
  if (!partial && J == 0) return;
  double x = p.u[p.umap[j]];
  if (partial) {
    double s = 0.0;
    for (int sp = (J * B) / cols_per_split; sp < nsplit; ++sp) s += partial[(long)sp * p.m + j];
    x += s;
  }
  if (J > 0) {
    double s = 0.0;
    for (int I = 0; I < J; ++I) s += colp[(long)I * p.m + j];
    x += s;
  }
  p.u[p.umap[j]] = x;
}

// ------------------------------------------------------------------ launch

bool sym_eligible(const Problem &p) {
  const int ks = fused_kspec(p.k);
  if (ks != 2 && ks != 3 && ks != 4 && ks != 8) return false;
  if (p.m != p.n || p.m < 2 * SWG * sym_rows_per_lane(p.k)) return false;
  if (p.amap != p.bmap || p.XA != p.XB || p.XA2 != p.XB2) return false;
  if (p.type == VAR_BANDWIDTH && p.hi != p.hj) return false;
  return true;
}

SymPlan plan_sym(int m, int k, int target_wgs) {
  return plan_sym_rows(m, target_wgs, SWG * sym_rows_per_lane(k));
}

SymPlan plan_sym_rows(int m, int target_wgs, int rows_per_wg) {
  SymPlan pl;
  pl.rows_per_wg = rows_per_wg;
  pl.row_blocks = (m + pl.rows_per_wg - 1) / pl.rows_per_wg;
  // about target_wgs working workgroups: row block I has nsplit - I*B/S splits
  long nsplit = (2L * target_wgs + pl.row_blocks - 1) / pl.row_blocks;
  if (nsplit > pl.row_blocks) nsplit = pl.row_blocks;
  if (nsplit < 1) nsplit = 1;
  long per = (m + nsplit - 1) / nsplit;
  per = (per + pl.rows_per_wg - 1) / pl.rows_per_wg * pl.rows_per_wg;
  pl.cols_per_split = (int)per;
  pl.nsplit = (int)((m + per - 1) / per);
  return pl;
}

template <int TYPE, int K>
static hipError_t sym_go(const SymArgs &A, const SymPlan &pl, hipStream_t s) {
  constexpr int R = sym_rows_per_lane(K);
  hipLaunchKernelGGL((fused_f64_sym_kernel<TYPE, K, R>), dim3(pl.row_blocks, pl.nsplit), dim3(SWG),
                     0, s, A);
  return hipGetLastError();
}

template <int TYPE>
static hipError_t sym_k(const SymArgs &A, const SymPlan &pl, hipStream_t s) {
  switch (A.p.k) {
    case 2: return sym_go<TYPE, 2>(A, pl, s);
    case 3: return sym_go<TYPE, 3>(A, pl, s);
    case 4: return sym_go<TYPE, 4>(A, pl, s);
    default: return sym_go<TYPE, 8>(A, pl, s);
  }
}

// One thread per point, in map order: rows X[map[i]], norms, weights
// w[wmap[i]], bandwidths and an identity map.  The vector kernels read every
// source row and its factors with wave-uniform scalar loads; through a
// permutation map each is a different line that the scalar cache mostly
// misses, from the compact tables they are consecutive.
__global__ void gather_points_kernel(int count, int k, const double *__restrict__ X,
                                     const double *__restrict__ X2, const int *__restrict__ map,
                                     const double *__restrict__ w, const int *__restrict__ wmap,
                                     const double *__restrict__ h, double *__restrict__ Xc,
                                     double *__restrict__ X2c, double *__restrict__ wc,
                                     double *__restrict__ hc, int *__restrict__ ids) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int a = map[i];
  const double *x = X + (long)a * k;
  for (int q = 0; q < k; ++q) Xc[(long)i * k + q] = x[q];
  X2c[i] = X2[a];
  wc[i] = w[wmap[i]];
  if (hc) hc[i] = h[a];
  ids[i] = i;
}

static size_t up32(size_t e) { return (e + 31) / 32 * 32; }

size_t gather_table_doubles(long count, int k, int type) {
  const size_t c = (size_t)count;
  return up32(c * k) + 2 * up32(c) + (type == VAR_BANDWIDTH ? up32(c) : 0) + up32((c + 1) / 2);
}

// The tables of `count` points (rows X[map], X2[map], w[wmap], h[map] and the
// identity map) laid out in `tab`, filled by one launch.
static hipError_t gather_points(int count, int k, int type, const double *X, const double *X2,
                                const int *map, const double *w, const int *wmap,
                                const double *h, double *tab, const double **Xc,
                                const double **X2c, const double **wc, const double **hc,
                                const int **ids, hipStream_t s) {
  const size_t c = (size_t)count;
  size_t off = 0;
  double *x = tab + off;
  off += up32(c * k);
  double *x2 = tab + off;
  off += up32(c);
  double *ww = tab + off;
  off += up32(c);
  double *hh = nullptr;
  if (type == VAR_BANDWIDTH) {
    hh = tab + off;
    off += up32(c);
  }
  int *id = reinterpret_cast<int *>(tab + off);
  hipLaunchKernelGGL(gather_points_kernel, dim3((unsigned)((c + 255) / 256)), dim3(256), 0, s,
                     count, k, X, X2, map, w, wmap, h, x, x2, ww, hh, id);
  *Xc = x;
  *X2c = x2;
  *wc = ww;
  *hc = hh;
  *ids = id;
  return hipGetLastError();
}

size_t sym_table_doubles(const Problem &p) { return gather_table_doubles(p.m, p.k, p.type); }

hipError_t launch_sym_gather(const Problem &p, double *tab, Problem *q, hipStream_t s) {
  const double *Xc, *X2c, *wc, *hc;
  const int *ids;
  hipError_t e = gather_points(p.m, p.k, p.type, p.XA, p.XA2, p.amap, p.w, p.wmap, p.hi, tab, &Xc,
                               &X2c, &wc, &hc, &ids, s);
  *q = p;
  q->XA = q->XB = Xc;
  q->XA2 = q->XB2 = X2c;
  q->amap = q->bmap = q->wmap = ids;
  q->w = wc;
  if (hc) q->hi = q->hj = hc;
  return e;
}

hipError_t launch_source_gather(const Problem &p, double *tab, Problem *q, hipStream_t s) {
  const double *Xc, *X2c, *wc, *hc;
  const int *ids;
  hipError_t e = gather_points(p.n, p.k, p.type, p.XB, p.XB2, p.bmap, p.w, p.wmap, p.hj, tab, &Xc,
                               &X2c, &wc, &hc, &ids, s);
  *q = p;
  q->XB = Xc;
  q->XB2 = X2c;
  q->bmap = q->wmap = ids;
  q->w = wc;
  if (hc) q->hj = hc;
  return e;
}

hipError_t launch_fused_f64_sym(const Problem &p, const SymPlan &pl, double *partial,
                                double *colp, hipStream_t s) {
  SymArgs A;
  A.p = p;
  A.cols_per_split = pl.cols_per_split;
  A.nsplit = pl.nsplit;
  A.partial = partial;
  A.colp = colp;
  hipError_t e;
  switch (p.type) {
    case GAUSSIAN: e = sym_k<GAUSSIAN>(A, pl, s); break;
    case POLYNOMIAL:
      switch (poly_type(p.P)) {  // the power fixed at compile time where it can be
        case POLY_SQUARE_T: e = sym_k<POLY_SQUARE_T>(A, pl, s); break;
        case POLY_CUBE_T: e = sym_k<POLY_CUBE_T>(A, pl, s); break;
        case POLY_FOURTH_T: e = sym_k<POLY_FOURTH_T>(A, pl, s); break;
        default: e = sym_k<POLYNOMIAL>(A, pl, s); break;
      }
      break;
    case LAPLACE: e = sym_k<LAPLACE>(A, pl, s); break;
    case VAR_BANDWIDTH: e = sym_k<VAR_BANDWIDTH>(A, pl, s); break;
    case TANH: e = sym_k<TANH>(A, pl, s); break;
    case QUARTIC: e = sym_k<QUARTIC>(A, pl, s); break;
    case MULTIQUADRATIC: e = sym_k<MULTIQUADRATIC>(A, pl, s); break;
    case EPANECHNIKOV: e = sym_k<EPANECHNIKOV>(A, pl, s); break;
    default: return hipErrorInvalidValue;
  }
  if (e != hipSuccess) return e;
  return launch_sym_colsum(p, pl, partial, colp, s);
}

hipError_t launch_sym_colsum(const Problem &p, const SymPlan &pl, const double *partial,
                             const double *colp, hipStream_t s) {
  hipLaunchKernelGGL(sym_colsum_kernel, dim3((unsigned)((p.m + 255) / 256)), dim3(256), 0, s, p,
                     pl.rows_per_wg, pl.nsplit > 1 ? partial : nullptr, pl.nsplit,
                     pl.cols_per_split, colp);
  return hipGetLastError();
}

}  // namespace gsks
