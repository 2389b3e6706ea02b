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
// Laplace: the reference drops pairs with d^2 < 1e-15 (frame/dgsks_ref.c:343-345),
// so a near pair's d^2 must be the reference's, bit for bit.  The matrix-core
// and the reference's sequential sums differ by at most
// (2 (k + 2) + 4) 2^-53 (|a|^2 + |b|^2); a pair whose d^2 could lie on the other
// side of the threshold, or whose power d^(2-k) could turn that bound into more
// than 1e-11 relative, is recomputed with the reference's sequential non-FMA
// sum (then d^2, the drop and K are the reference's).  On spread points only
// coincident pairs take that path.  The symmetric and multi-right-hand-side
// matrix-core kernels keep Laplace on the vector units.
#include <hip/hip_runtime.h>

#include "gsks_kernels.h"
#include "gsks_launch.h"

namespace gsks {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int MWG = 256;  // 4 waves

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

  // ---- stream 16-point source blocks; two register buffers used in turn.  A
  // block's map entries (bmap, wmap) load two blocks ahead and the data they
  // address one block ahead, so no block waits on two dependent round trips.
  double bf0[KC][4], bf1[KC][4];
  double bb0 = 0.0, wj0 = 0.0, hj0 = 0.0, bb1 = 0.0, wj1 = 0.0, hj1 = 0.0;
  int ib0 = 0, iw0 = -1, ib1 = 0, iw1 = -1;
  auto fetch_idx = [&](int jb, int &ib, int &iw) {
    const int col = jb + p16;
    const bool ok = col < j1;
    ib = p.bmap[ok ? col : j1 - 1];
    iw = ok ? p.wmap[col] : -1;
  };
  auto fetch = [&](int ib, int iw, double (&f)[KC][4], double &b2, double &ww, double &hh) {
    load_frag<KC, EXACT>(p.XB, (long)ib * k, k, g, f);
    b2 = p.XB2[ib];
    ww = iw >= 0 ? p.w[iw] : 0.0;
    hh = (TYPE == VAR_BANDWIDTH) ? p.hj[ib] : 0.0;
  };
  // Laplace: exact-recompute threshold on d^2, as a multiple of |a|^2 + |b|^2
  const double lap_b = (2.0 * (k + 2) + 4.0) * 0x1p-53;
  const double lap_t1 = lap_b * (1.0 + fmax(0.5 * k - 1.0, 1.0) * 1e11) + lap_b;
  auto block = [&](const double (&f)[KC][4], double bb, double wj, double hj, int ib) {
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int c = 0; c < KC; ++c) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(af[rb][c][t], f[c][t], acc, 0, 0, 0);
      }
      if constexpr (TYPE == LAPLACE) {
        double d2[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          d2[v] = fma(-2.0, acc[v], aa[rb][v]) + bb;  // the assembly of kval_f64
          if (d2[v] < 1e-15 + lap_t1 * (aa[rb][v] + bb)) {
            const int ro = row0 + rb * 16 + g + 4 * v;
            const int ao = p.amap[ro < m ? ro : m - 1];
            const double c = dot_exact_n(p.XA + (long)ao * k, p.XB + (long)ib * k, k);
            d2[v] = fma(-2.0, c, aa[rb][v]) + bb;
          }
        }
        laplace_n_f64<4>(d2, P);  // the 4 kernel values, one power ladder
#pragma unroll
        for (int v = 0; v < 4; ++v) racc[rb][v] = fma(d2[v], wj, racc[rb][v]);
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          double K = kval_f64<TYPE>(acc[v], aa[rb][v], bb, hi[rb][v], hj, P);
          racc[rb][v] = fma(K, wj, racc[rb][v]);
        }
      }
    }
  };
  int sb0 = 0, sb1 = 0;  // the source rows of the blocks held in bf0 / bf1
  if (j0 < j1) {
    fetch_idx(j0, ib0, iw0);
    if (j0 + 16 < j1) fetch_idx(j0 + 16, ib1, iw1);
    fetch(ib0, iw0, bf0, bb0, wj0, hj0);
    sb0 = ib0;
  }
  for (int jb = j0; jb < j1; jb += 32) {
    // here bf0 holds block jb and (ib1, iw1) the map entries of block jb + 16
    if (jb + 16 < j1) {
      fetch(ib1, iw1, bf1, bb1, wj1, hj1);
      sb1 = ib1;
    }
    if (jb + 32 < j1) fetch_idx(jb + 32, ib0, iw0);
    block(bf0, bb0, wj0, hj0, sb0);
    if (jb + 16 >= j1) break;
    if (jb + 32 < j1) {
      fetch(ib0, iw0, bf0, bb0, wj0, hj0);
      sb0 = ib0;
    }
    if (jb + 48 < j1) fetch_idx(jb + 48, ib1, iw1);
    block(bf1, bb1, wj1, hj1, sb1);
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
  // the splits' partials are added in split order by launch_split_reduce's
  // kernel, launched next (no device-scope fence here: on gfx950 it would
  // write back and invalidate this XCD's L2)
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    int r = row0 + rb * 16 + g + 4 * p16;
    if (writer && r < m) A.partial[(long)split * m + r] = mine[rb];
  }
}

// ------------------------------------------------------------------ symmetric sums

// Sources are the targets (one map, one table, one norm vector, hi == hj):
// K_ij = K_ji, so each pair of different row blocks is evaluated once for both
// of its terms, as in fused_f64_sym.hip, with the cross terms on the matrix
// cores.  A workgroup of 8 waves owns a row block of 256 targets (RB 16-row
// blocks per wave) and walks the sources j >= its first row: its own block
// ("diagonal", every pair, row sums only), then the later blocks ("upper", row
// sums and column sums).  A lane's column sum sum_v K[v] w_row[v] covers its
// 4 RB rows of its column; two lane exchanges give the wave's 32 rows, and the
// 8 waves meet in LDS once per 64 sources, in wave order, giving one value per
// (row block, source) in colp[I][j].  Row sums use the split partials of the
// general kernel (splits before a block's diagonal hold no work for it), and
// sym_colsum_kernel adds sum_{I < J} colp[I][j] in block order: deterministic.
constexpr int SMWG = 256;  // 4 waves

struct SymMArgs {
  Problem p;
  int cols_per_split;  // a multiple of the row block
  int nsplit;
  double *partial;     // [nsplit][m] row partials
  unsigned *counters;  // [row blocks], zero between launches
  double *colp;        // [row blocks][m] column partials of the upper sources
};

template <int TYPE, int KC, int RB, bool EXACT>
__global__ __launch_bounds__(SMWG) void fused_f64_sym_mfma_kernel(SymMArgs A) {
  const Problem &p = A.p;
  const int m = p.m, k = p.k;  // m == n
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int p16 = lane & 15, g = lane >> 4;
  const int rblk = blockIdx.x, split = blockIdx.y;
  constexpr int B = (SMWG / 64) * RB * 16;
  const int row0 = rblk * B;
  const int first = row0 / A.cols_per_split;  // the split holding this block's diagonal
  if (split < first) return;                  // earlier sources are other blocks' upper part
  const int wrow0 = row0 + wave * (RB * 16);

  // ---- A fragments; norms, bandwidths and weights of the rows of this lane's results
  double af[RB][KC][4];
  double aa[RB][4], hi[RB][4], wr[RB][4], racc[RB][4];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    const int r = wrow0 + rb * 16 + p16;
    const int ai = p.amap[r < m ? r : m - 1];
    load_frag<KC, EXACT>(p.XA, (long)ai * k, k, g, af[rb]);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int ro = wrow0 + rb * 16 + g + 4 * v;
      const int ro_ = ro < m ? ro : m - 1;
      const int ao = p.amap[ro_];
      aa[rb][v] = p.XA2[ao];
      hi[rb][v] = (TYPE == VAR_BANDWIDTH) ? -0.5 * p.hi[ao] : 0.0;
      wr[rb][v] = ro < m ? p.w[p.wmap[ro_]] : 0.0;  // padding rows add nothing to a column
      racc[rb][v] = 0.0;
    }
  }
  const int c0 = split * A.cols_per_split;
  const int c1 = min(m, c0 + A.cols_per_split);
  const KerParams P = p.P;

  // map entries two blocks ahead, the data they address one block ahead (as
  // in the general kernel)
  double bf0[KC][4], bf1[KC][4];
  double bb0 = 0.0, wj0 = 0.0, hj0 = 0.0, bb1 = 0.0, wj1 = 0.0, hj1 = 0.0;
  int ib0 = 0, iw0 = -1, ib1 = 0, iw1 = -1;
  auto fetch_idx = [&](int jb, int &ib, int &iw) {
    const int col = jb + p16;
    const bool ok = col < c1;
    ib = p.bmap[ok ? col : c1 - 1];
    iw = ok ? p.wmap[col] : -1;
  };
  auto fetch = [&](int ib, int iw, double (&f)[KC][4], double &b2, double &ww, double &hh) {
    load_frag<KC, EXACT>(p.XB, (long)ib * k, k, g, f);
    b2 = p.XB2[ib];
    ww = iw >= 0 ? p.w[iw] : 0.0;
    hh = (TYPE == VAR_BANDWIDTH) ? p.hj[ib] : 0.0;
  };
  // one 16-source block: row sums, and (upper) the lane's column sum
  auto block = [&](const double (&f)[KC][4], double bb, double wj, double hj, bool upper) {
    double cs = 0.0;
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
        const double K = kval_f64<TYPE>(acc[v], aa[rb][v], bb, hi[rb][v], hj, P);
        racc[rb][v] = fma(K, wj, racc[rb][v]);
        if (upper) cs = fma(K, wr[rb][v], cs);
      }
    }
    return cs;
  };

  // ---- diagonal sources: every pair, row sums only
  {
    const int lo = max(c0, row0), hi_ = min(c1, row0 + B);
    if (lo < hi_) {
      fetch_idx(lo, ib0, iw0);
      if (lo + 16 < hi_) fetch_idx(lo + 16, ib1, iw1);
      fetch(ib0, iw0, bf0, bb0, wj0, hj0);
      for (int jb = lo; jb < hi_; jb += 32) {
        if (jb + 16 < hi_) fetch(ib1, iw1, bf1, bb1, wj1, hj1);
        if (jb + 32 < hi_) fetch_idx(jb + 32, ib0, iw0);
        (void)block(bf0, bb0, wj0, hj0, false);
        if (jb + 16 >= hi_) break;
        if (jb + 32 < hi_) fetch(ib0, iw0, bf0, bb0, wj0, hj0);
        if (jb + 48 < hi_) fetch_idx(jb + 48, ib1, iw1);
        (void)block(bf1, bb1, wj1, hj1, false);
      }
    }
  }

  // ---- upper sources: row sums and column sums, combined every 64 sources
  __shared__ double s_col[SMWG / 64][64];
  const int ulo = max(c0, row0 + B);  // a multiple of 64 (of B)
  auto upper_block = [&](int jb, const double (&f)[KC][4], double bb, double wj, double hj) {
    double cs = block(f, bb, wj, hj, true);
    cs += __shfl_xor(cs, 16);
    cs += __shfl_xor(cs, 32);
    const int slot = (jb - ulo) & 63;  // 16 q for the q-th block of the chunk
    if (lane < 16) s_col[wave][slot + lane] = cs;
    if (slot == 48 || jb + 16 >= c1) {
      __syncthreads();
      const int jc = jb - slot;
      if (tid < 64 && jc + tid < c1) {
        double t = 0.0;
#pragma unroll
        for (int w8 = 0; w8 < SMWG / 64; ++w8) t += s_col[w8][tid];
        A.colp[(long)rblk * m + jc + tid] = t;
      }
      __syncthreads();
    }
  };
  if (ulo < c1) {
    fetch_idx(ulo, ib0, iw0);
    if (ulo + 16 < c1) fetch_idx(ulo + 16, ib1, iw1);
    fetch(ib0, iw0, bf0, bb0, wj0, hj0);
    for (int jb = ulo; jb < c1; jb += 32) {
      if (jb + 16 < c1) fetch(ib1, iw1, bf1, bb1, wj1, hj1);
      if (jb + 32 < c1) fetch_idx(jb + 32, ib0, iw0);
      upper_block(jb, bf0, bb0, wj0, hj0);
      if (jb + 16 >= c1) break;
      if (jb + 32 < c1) fetch(ib0, iw0, bf0, bb0, wj0, hj0);
      if (jb + 48 < c1) fetch_idx(jb + 48, ib1, iw1);
      upper_block(jb + 16, bf1, bb1, wj1, hj1);
    }
  }

  // ---- row results: the 16 column lanes of each row group, then the splits
  // from `first` on in split order
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
      const int r = wrow0 + rb * 16 + g + 4 * p16;
      if (writer && r < m) p.u[p.umap[r]] += mine[rb];
    }
    return;
  }
  // the splits' partials (from `first` on) are added in split order by
  // sym_colsum_kernel, launched next, before the column partials (no
  // device-scope fence here: on gfx950 it would write back and invalidate this
  // XCD's L2)
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    const int r = wrow0 + rb * 16 + g + 4 * p16;
    if (writer && r < m) A.partial[(long)split * m + r] = mine[rb];
  }
}

// ------------------------------------------------------------------ launch

constexpr int kMfmaRB = 2;

bool mfma_f64_eligible(int type, int k) { return type != LAPLACE && k >= 9 && k <= 64; }

bool fused_mfma_eligible(int type, int k) { return k >= 9 && k <= 64; }

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
    case POLYNOMIAL:
      switch (poly_type(p.P)) {  // the power fixed at compile time where it can be
        case POLY_SQUARE_T: return go_type<POLY_SQUARE_T>(A, plan.row_blocks, s);
        case POLY_CUBE_T: return go_type<POLY_CUBE_T>(A, plan.row_blocks, s);
        case POLY_FOURTH_T: return go_type<POLY_FOURTH_T>(A, plan.row_blocks, s);
        default: return go_type<POLYNOMIAL>(A, plan.row_blocks, s);
      }
    case VAR_BANDWIDTH: return go_type<VAR_BANDWIDTH>(A, plan.row_blocks, s);
    case TANH: return go_type<TANH>(A, plan.row_blocks, s);
    case QUARTIC: return go_type<QUARTIC>(A, plan.row_blocks, s);
    case MULTIQUADRATIC: return go_type<MULTIQUADRATIC>(A, plan.row_blocks, s);
    case EPANECHNIKOV: return go_type<EPANECHNIKOV>(A, plan.row_blocks, s);
    case LAPLACE: return go_type<LAPLACE>(A, plan.row_blocks, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace gsks

namespace gsks {

bool sym_mfma_eligible(const Problem &p) {
  if (!mfma_f64_eligible(p.type, p.k)) return false;
  if (p.m != p.n || p.m < 2 * sym_mfma_rows()) return false;
  if (p.amap != p.bmap || p.XA != p.XB || p.XA2 != p.XB2) return false;
  if (p.type == VAR_BANDWIDTH && p.hi != p.hj) return false;
  return true;
}

int sym_mfma_rows() { return (SMWG / 64) * kMfmaRB * 16; }

template <int TYPE, int KC>
static hipError_t sym_mfma_go(const SymMArgs &A, const SymPlan &pl, hipStream_t s) {
  dim3 grid(pl.row_blocks, pl.nsplit);
  const bool exact = A.p.k == 16 * KC && (uintptr_t)A.p.XA % 16 == 0;
  if (exact)
    hipLaunchKernelGGL((fused_f64_sym_mfma_kernel<TYPE, KC, kMfmaRB, true>), grid, dim3(SMWG), 0, s, A);
  else
    hipLaunchKernelGGL((fused_f64_sym_mfma_kernel<TYPE, KC, kMfmaRB, false>), grid, dim3(SMWG), 0, s,
                       A);
  return hipGetLastError();
}

template <int TYPE>
static hipError_t sym_mfma_type(const SymMArgs &A, const SymPlan &pl, hipStream_t s) {
  switch ((A.p.k + 15) / 16) {
    case 1: return sym_mfma_go<TYPE, 1>(A, pl, s);
    case 2: return sym_mfma_go<TYPE, 2>(A, pl, s);
    case 3: return sym_mfma_go<TYPE, 3>(A, pl, s);
    default: return sym_mfma_go<TYPE, 4>(A, pl, s);
  }
}

hipError_t launch_fused_f64_sym_mfma(const Problem &p, const SymPlan &pl, double *partial,
                                     unsigned *counters, double *colp, hipStream_t s) {
  SymMArgs A;
  A.p = p;
  A.cols_per_split = pl.cols_per_split;
  A.nsplit = pl.nsplit;
  A.partial = partial;
  A.counters = counters;
  A.colp = colp;
  hipError_t e;
  switch (p.type) {
    case GAUSSIAN: e = sym_mfma_type<GAUSSIAN>(A, pl, s); break;
    case POLYNOMIAL: e = sym_mfma_type<POLYNOMIAL>(A, pl, s); break;
    case VAR_BANDWIDTH: e = sym_mfma_type<VAR_BANDWIDTH>(A, pl, s); break;
    case TANH: e = sym_mfma_type<TANH>(A, pl, s); break;
    case QUARTIC: e = sym_mfma_type<QUARTIC>(A, pl, s); break;
    case MULTIQUADRATIC: e = sym_mfma_type<MULTIQUADRATIC>(A, pl, s); break;
    case EPANECHNIKOV: e = sym_mfma_type<EPANECHNIKOV>(A, pl, s); break;
    default: return hipErrorInvalidValue;
  }
  if (e != hipSuccess) return e;
  return launch_sym_colsum(p, pl, partial, colp, s);
}

}  // namespace gsks
