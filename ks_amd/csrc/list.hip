// Batched summation (dgsks_list): thousands of independent summations of
// uneven m, n over shared point pools in ONE persistent launch.
//
// Work unit = tile: 64 target rows of one task x one slice of that task's
// sources.  The host lays the tiles out once per batch (a plan), largest
// first; every wave of the persistent grid pulls the next tile index from a
// device-scope counter (atomicAdd) and runs it, so uneven tasks balance
// across all 1024 SIMDs without host involvement.  The slice length and the
// tile kind are the autotuned tile shape of the task's (kernel type, k).  Per tile:
//   * k = 3 or 8 (and k outside 9..64): lane = target row, the row in VGPRs,
//     sources as wave-uniform scalar loads (as fused_f64_kernel);
//   * 9 <= k <= 64, row tile: lane = target row, the row in VGPRs padded to
//     16 / 32 / 64 coordinates, sources staged through LDS (Laplace keeps the
//     reference's sequential rank-k sum here);
//   * 9 <= k <= 64, matrix-core tile: 4 blocks of 16 rows on the FP64 matrix
//     cores (as fused_f64_mfma_kernel).  For Laplace, whose d^2 < 1e-15 drop
//     makes the result discontinuous, a per-pair bound on the difference from
//     the reference's sequential sum sends every pair that could be dropped
//     differently, or whose power d^(2-k) could lose more than 1e-11, through
//     that exact sum; on spread points only coincident pairs take it.
// Each tile writes 64 row partials; the last tile of a (task, row block) to
// finish (device-scope counter) adds its partials in slice order into
// R[a_off[t] + i].  The last wave to leave resets the tile counter, so the
// launch leaves its scratch ready for the next one (graph replay friendly).
//
// Merging R into u: tasks may share target points, and the reference adds
// every task's contribution (frame/omp_dgsks_list.cpp:407-414).  To keep the
// result independent of scheduling, the target ids umap are stably
// radix-sorted (hipCUB) together with their positions, and one thread per
// distinct target adds its R entries in position order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "gsks_kernels.h"
#include "host_util.h"
#include "list_launch.h"

namespace gsks {

typedef double d4 __attribute__((ext_vector_type(4)));

struct DevTask {
  int type;
  int m, n;
  int bc;    // source slice per tile (the tuned tile shape)
  int tile;  // launch kind: 0 lane-per-row tile, 1 FP64 matrix-core tile (9 <= k <= 64),
             // 2 the same for Laplace (a launch of its own), 3 symmetric row tile
  int sym;   // symmetric: the sources are the targets (same list, same point table)
  int srb;   // symmetric: where its row blocks' indices start in ListArgs::symrb
  long a0, b0;
  KerParams P;
  const double *hi, *hj;
};

struct RowBlock {
  int task;
  int row0;       // first row of the block within the task
  int nsplit;     // source slices of this block
  int tile_base;  // index of its first tile
};

struct ListArgs {
  int tile0, ntiles, k;  // this launch's tiles: [tile0, tile0 + ntiles)
  const int2 *tiles;  // {row block, slice}
  const RowBlock *rbs;
  const DevTask *tasks;
  const int *amap, *bmap, *wmap;
  const double *XA, *XA2, *XB, *XB2, *w;
  double *partial;     // [ntiles][64]
  unsigned *head;      // tile counter (reset by list_rowsum_kernel after each run)
  double *R;           // [total_a + 64] per-(task, row) totals, then 64 sink entries
  long rsink;          // total_a: where the lanes past a row block's end write
  // symmetric tasks: the column partials of each tile's later sources
  // (colp + coff[tile], one per source from max(c0, row0 + 64) to c1), and each
  // task's row blocks in row order (symrb[srb + row0 / 64] = row-block index)
  double *colp;
  const long *coff;
  const int *symrb;
};

// ------------------------------------------------------------------ per-tile work

template <int TYPE>
__device__ __forceinline__ double dot_ref_order(const double *a, cptr<double> b, int k) {
#pragma clang fp contract(off)
  double c = a[0] * b[0];
  for (int p = 1; p < k; ++p) c = c + a[p] * b[p];
  return c;
}

// lane = row; sources [c0, c1) of task T through wave-uniform scalar loads
template <int TYPE, int K>
__device__ double tile_valu(const ListArgs &L, const DevTask &T, int pos, int c0, int c1,
                           const double2 *etab) {
  const int kdim = K > 0 ? K : L.k;
  const int ai = L.amap[T.a0 + (pos < T.m ? pos : T.m - 1)];
  const double *arow = L.XA + (long)ai * kdim;
  constexpr int KR = K > 0 ? K : 1;
  double a[KR];
  if constexpr (K > 0) {
#pragma unroll
    for (int p = 0; p < K; ++p) a[p] = arow[p];
  }
  const double aa = L.XA2[ai];
  const double hi = (TYPE == VAR_BANDWIDTH) ? -0.5 * T.hi[ai] : 0.0;
  cptr<int> bmap = to_const(L.bmap + T.b0);
  cptr<int> wmap = to_const(L.wmap + T.b0);
  cptr<double> XB = to_const(L.XB);
  cptr<double> XB2 = to_const(L.XB2);
  cptr<double> W = to_const(L.w);
  cptr<double> HJ = to_const(T.hj);
  const KerParams P = T.P;
  double acc = 0.0;
  for (int j = c0; j < c1; ++j) {
    const int bj = bmap[j];
    const double wj = W[wmap[j]];
    const double bb = XB2[bj];
    const double hj = (TYPE == VAR_BANDWIDTH) ? HJ[bj] : 0.0;
    cptr<double> b = XB + (long)bj * kdim;
    double c;
    if constexpr (TYPE == LAPLACE) {
      if constexpr (K > 0) {
#pragma clang fp contract(off)
        c = a[0] * b[0];
#pragma unroll
        for (int p = 1; p < K; ++p) c = c + a[p] * b[p];
      } else {
        c = dot_ref_order<TYPE>(arow, b, kdim);
      }
    } else if constexpr (K > 0) {
      c = a[0] * b[0];
#pragma unroll
      for (int p = 1; p < K; ++p) c = fma(a[p], b[p], c);
    } else {
      c = arow[0] * b[0];
      for (int p = 1; p < kdim; ++p) c = fma(arow[p], b[p], c);
    }
    acc = fma(kval_f64<TYPE, uses_exp<TYPE>>(c, aa, bb, hi, hj, P, etab), wj, acc);
  }
  return acc;
}

// Sources staged per wave in LDS: a chunk of stage_rows(K) source rows of K
// padded coordinates (at most kStageDoubles doubles), then 3 doubles per source
// of the chunk (|b|^2, w, hj).
constexpr int kStageDoubles = 1024;
constexpr int kStageSide = 3 * 64;
constexpr int stage_rows(int K) { return kStageDoubles / K < 64 ? kStageDoubles / K : 64; }
constexpr int stage_doubles(int K) { return (stage_rows(K) * K + 63) / 64 * 64; }

// lane = row, k <= K (K in {3, 8, 16, 32, 64}, coordinates past k are zero):
// the row stays in VGPRs; the sources go through the wave's LDS slot in chunks
// of stage_rows(K) rows -- loaded with vector loads that are all in flight at
// once (consecutive lanes read consecutive doubles of a row) and read back as
// wave-wide broadcasts -- so every source row is fetched once per tile and no
// load waits on the one before it.  Zero padding keeps the Laplace rank-k sum
// exact: the extra terms add +0.
template <int TYPE, int K>
__device__ double tile_valu_lds(const ListArgs &L, const DevTask &T, int pos, int c0, int c1,
                                double *stage, const double2 *etab) {
  constexpr int CH = stage_rows(K);
  constexpr int NV = stage_doubles(K) / 64;  // staged doubles per lane
  const int k = L.k;
  const int lane = threadIdx.x & 63;
  const int ai = L.amap[T.a0 + (pos < T.m ? pos : T.m - 1)];
  const double *arow = L.XA + (long)ai * k;
  double a[K];
#pragma unroll
  for (int p = 0; p < K; ++p) a[p] = p < k ? arow[p] : 0.0;
  const double aa = L.XA2[ai];
  const double hi = (TYPE == VAR_BANDWIDTH) ? -0.5 * T.hi[ai] : 0.0;
  const KerParams P = T.P;
  double *side = stage + stage_doubles(K);  // [CH][3]: |b|^2, w, hj
  double acc = 0.0;
  for (int j0 = c0; j0 < c1; j0 += CH) {
    const int nch = min(CH, c1 - j0);
    double v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int e = lane + 64 * i, r = e / K, p = e % K;
      const int rr = r < nch ? r : nch - 1;
      const double x = L.XB[(long)L.bmap[T.b0 + j0 + rr] * k + (p < k ? p : k - 1)];
      v[i] = (r < nch && p < k) ? x : 0.0;
    }
    // lane r < CH: the per-source factors of source j0 + r
    double sb = 0.0, sw = 0.0, sh = 0.0;
    if (lane < nch) {
      const int bj = L.bmap[T.b0 + j0 + lane];
      sb = L.XB2[bj];
      sw = L.w[L.wmap[T.b0 + j0 + lane]];
      sh = (TYPE == VAR_BANDWIDTH) ? T.hj[bj] : 0.0;
    }
    __builtin_amdgcn_wave_barrier();  // the previous chunk's broadcasts are done
#pragma unroll
    for (int i = 0; i < NV; ++i) stage[lane + 64 * i] = v[i];
    if (lane < CH) {
      side[3 * lane] = sb;
      side[3 * lane + 1] = sw;
      side[3 * lane + 2] = sh;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int r = 0;
    if constexpr (TYPE == LAPLACE && K >= 8) {
      // four sources at a time: their kernel values share one power ladder
      // (k = 3 has none: d^-1), and acc still takes them in source order
#pragma unroll 1
      for (; r + 4 <= nch; r += 4) {
        double d2[4], wq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double *b = stage + (r + q) * K;
          double c;
          {
#pragma clang fp contract(off)
            c = a[0] * b[0];
#pragma unroll
            for (int p = 1; p < K; ++p) c = c + a[p] * b[p];
          }
          d2[q] = fma(-2.0, c, aa) + side[3 * (r + q)];  // the assembly of kval_f64
          wq[q] = side[3 * (r + q) + 1];
        }
        laplace_n_f64<4>(d2, P);
#pragma unroll
        for (int q = 0; q < 4; ++q) acc = fma(d2[q], wq[q], acc);
      }
    }
    // two sources per iteration: their kernel values are independent chains
    // that interleave (acc still takes them in source order); more spills at
    // 4 waves per SIMD, but list_pool k = 3 / 8 12.8 -> 11.7 / 19.4 -> 18.7 ms
#pragma unroll 2
    for (; r < nch; ++r) {
      const double bb = side[3 * r], wj = side[3 * r + 1], hj = side[3 * r + 2];
      const double *b = stage + r * K;
      double c;
      if constexpr (TYPE == LAPLACE) {
#pragma clang fp contract(off)
        c = a[0] * b[0];
#pragma unroll
        for (int p = 1; p < K; ++p) c = c + a[p] * b[p];
      } else {
        c = a[0] * b[0];
#pragma unroll
        for (int p = 1; p < K; ++p) c = fma(a[p], b[p], c);
      }
      acc = fma(kval_f64<TYPE, uses_exp<TYPE>>(c, aa, bb, hi, hj, P, etab), wj, acc);
    }
  }
  __builtin_amdgcn_wave_barrier();
  return acc;
}

// Symmetric row tile (lane = row, as tile_valu_lds) of a task whose sources
// are its targets: rows row0 .. row0 + 63 against sources [c0, c1), c0 >= row0,
// a slice of the upper triangle.  Sources of the rows' own block [row0,
// row0 + 64) give row sums only -- every pair of the diagonal block, both
// orders.  Each later source j gives its row sums and the column sum
// sum_i K(a_i, b_j) w_i over the tile's rows (w_i: row i's weight as a source),
// the term of target j from sources row0 .. row0 + 63, which K(a, b) = K(b, a)
// makes the same pairs' values: it goes to colp[j - max(c0, row0 + 64)].  The
// column sums are reduced within each 16-lane DPP row per source (lane 16q + c
// keeping row q's share of source c of a 16-source group), then over the four
// rows with two lane swaps per group: a fixed order.
template <int TYPE, int K>
__device__ double tile_valu_lds_sym(const ListArgs &L, const DevTask &T, int row0, int c0,
                                    int c1, double *stage, double *colp, const double2 *etab) {
  constexpr int CH = stage_rows(K);  // 64, 32 or 16: whole 16-source groups
  constexpr int NV = stage_doubles(K) / 64;
  const int k = L.k;
  const int lane = threadIdx.x & 63;
  const int pos = row0 + lane;
  const bool rv = pos < T.m;
  const int ai = L.amap[T.a0 + (rv ? pos : T.m - 1)];
  const double *arow = L.XA + (long)ai * k;
  double a[K];
#pragma unroll
  for (int p = 0; p < K; ++p) a[p] = p < k ? arow[p] : 0.0;
  const double aa = L.XA2[ai];
  const double hi = (TYPE == VAR_BANDWIDTH) ? -0.5 * T.hi[ai] : 0.0;
  // the row's weight as a source (its position in the source list is its own);
  // padding rows add nothing to a column
  const double wi = rv ? L.w[L.wmap[T.b0 + pos]] : 0.0;
  const KerParams P = T.P;
  double *side = stage + stage_doubles(K);  // [CH][3]: |b|^2, w, hj
  const int cs = max(c0, row0 + 64);        // the first later source of the tile
  double acc = 0.0, part = 0.0;
  for (int j0 = c0; j0 < c1; j0 += CH) {
    const int nch = min(CH, c1 - j0);
    double v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int e = lane + 64 * i, r = e / K, p = e % K;
      const int rr = r < nch ? r : nch - 1;
      const double x = L.XB[(long)L.bmap[T.b0 + j0 + rr] * k + (p < k ? p : k - 1)];
      v[i] = (r < nch && p < k) ? x : 0.0;
    }
    double sb = 0.0, sw = 0.0, sh = 0.0;
    if (lane < nch) {
      const int bj = L.bmap[T.b0 + j0 + lane];
      sb = L.XB2[bj];
      sw = L.w[L.wmap[T.b0 + j0 + lane]];
      sh = (TYPE == VAR_BANDWIDTH) ? T.hj[bj] : 0.0;
    }
    __builtin_amdgcn_wave_barrier();  // the previous chunk's broadcasts are done
#pragma unroll
    for (int i = 0; i < NV; ++i) stage[lane + 64 * i] = v[i];
    if (lane < CH) {
      side[3 * lane] = sb;
      side[3 * lane + 1] = sw;
      side[3 * lane + 2] = sh;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // chunks start on multiples of CH from c0 (row0 or row0 + s bc, multiples of
    // 64 apart), so a chunk lies wholly before or wholly after row0 + 64
    const bool later = j0 >= row0 + 64;
#pragma unroll 1
    for (int r = 0; r < nch; ++r) {
      const double bb = side[3 * r], wj = side[3 * r + 1], hj = side[3 * r + 2];
      const double *b = stage + r * K;
      double c;
      if constexpr (TYPE == LAPLACE) {
#pragma clang fp contract(off)
        c = a[0] * b[0];
#pragma unroll
        for (int p = 1; p < K; ++p) c = c + a[p] * b[p];
      } else {
        c = a[0] * b[0];
#pragma unroll
        for (int p = 1; p < K; ++p) c = fma(a[p], b[p], c);
      }
      const double kv = kval_f64<TYPE, uses_exp<TYPE>>(c, aa, bb, hi, hj, P, etab);
      acc = fma(kv, wj, acc);
      if (later) {  // wave-uniform
        const double rs = row_sum16(kv * wi);
        part = (lane & 15) == (r & 15) ? rs : part;
        if ((r & 15) == 15 || r == nch - 1) {
          double tot = part + __shfl_xor(part, 16);
          tot += __shfl_xor(tot, 32);
          const int jj = j0 + (r & ~15) + (lane & 15);
          if (lane < 16 && jj <= j0 + r) colp[jj - cs] = tot;
          part = 0.0;
        }
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  return acc;
}

template <int K>
__device__ double sym_tile(const ListArgs &L, const DevTask &T, int row0, int c0, int c1,
                           double *stage, double *colp, const double2 *etab) {
  switch (T.type) {
    case GAUSSIAN: return tile_valu_lds_sym<GAUSSIAN, K>(L, T, row0, c0, c1, stage, colp, etab);
    case POLYNOMIAL:
      switch (poly_type(T.P)) {
        case POLY_SQUARE_T:
          return tile_valu_lds_sym<POLY_SQUARE_T, K>(L, T, row0, c0, c1, stage, colp, etab);
        case POLY_CUBE_T:
          return tile_valu_lds_sym<POLY_CUBE_T, K>(L, T, row0, c0, c1, stage, colp, etab);
        case POLY_FOURTH_T:
          return tile_valu_lds_sym<POLY_FOURTH_T, K>(L, T, row0, c0, c1, stage, colp, etab);
        default: return tile_valu_lds_sym<POLYNOMIAL, K>(L, T, row0, c0, c1, stage, colp, etab);
      }
    case LAPLACE: return tile_valu_lds_sym<LAPLACE, K>(L, T, row0, c0, c1, stage, colp, etab);
    case VAR_BANDWIDTH:
      return tile_valu_lds_sym<VAR_BANDWIDTH, K>(L, T, row0, c0, c1, stage, colp, etab);
    case TANH: return tile_valu_lds_sym<TANH, K>(L, T, row0, c0, c1, stage, colp, etab);
    case QUARTIC: return tile_valu_lds_sym<QUARTIC, K>(L, T, row0, c0, c1, stage, colp, etab);
    case MULTIQUADRATIC:
      return tile_valu_lds_sym<MULTIQUADRATIC, K>(L, T, row0, c0, c1, stage, colp, etab);
    default: return tile_valu_lds_sym<EPANECHNIKOV, K>(L, T, row0, c0, c1, stage, colp, etab);
  }
}

// 64 rows as 4 blocks of 16 on the FP64 matrix cores, in two passes of 2 row
// blocks (the A fragments, norms and accumulators of 64 rows at once would not
// fit in VGPRs beside the source fragments; the second pass re-reads the
// slice's sources from L2); returns the sum of the lane's row (lane = row)
// through the wave's LDS slot `red` (64 doubles).
template <int TYPE, int KC>
__device__ double tile_mfma(const ListArgs &L, const DevTask &T, int row0, int c0, int c1,
                            double *red, const double2 *etab) {
  constexpr int RBP = 2;
  const int k = L.k;
  const bool exact = (k == 16 * KC);  // whole chunks: two 16-byte loads per chunk
  const int lane = threadIdx.x & 63, p16 = lane & 15, g = lane >> 4;
  // coordinates c*16 + 4g .. +3 of point row `base`
  auto frag = [&](const double *X, long base, double (&f)[KC][4]) {
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      if (exact) {
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
          f[c][t] = idx < k ? X[base + idx] : 0.0;
        }
      }
    }
  };
  const KerParams P = T.P;
  const int jlast = c0 + ((c1 - c0 - 1) >> 4) * 16;
  // Laplace: |d2 (matrix cores) - d2 (reference's sum)| <= (2 (k + 2) + 4) 2^-53 (|a|^2 + |b|^2)
  // (both rank-k sums within (k + 1) 2^-53 |a||b| of a.b, plus the assembly);
  // pairs with d2 below the drop threshold plus that bound, or below the bound
  // times 1 + (k/2 - 1) 1e11 (a relative error above 1e-11 after the power
  // d^(2-k)), are recomputed exactly
  const double lap_b = (2.0 * (k + 2) + 4.0) * 0x1p-53;
  const double lap_t0 = 1e-15;
  const double lap_t1 = lap_b * (1.0 + fmax(0.5 * k - 1.0, 1.0) * 1e11) + lap_b;
#pragma unroll 1
  for (int pass = 0; pass < 4 / RBP; ++pass) {
    const int rbase = row0 + pass * RBP * 16;
    double af[RBP][KC][4], aa[RBP][4], hi[RBP][4], racc[RBP][4];
#pragma unroll
    for (int rb = 0; rb < RBP; ++rb) {
      int r = rbase + rb * 16 + p16;
      int ai = L.amap[T.a0 + (r < T.m ? r : T.m - 1)];
      frag(L.XA, (long)ai * k, af[rb]);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int ro = rbase + rb * 16 + g + 4 * v;
        int ao = L.amap[T.a0 + (ro < T.m ? ro : T.m - 1)];
        aa[rb][v] = L.XA2[ao];
        hi[rb][v] = (TYPE == VAR_BANDWIDTH) ? -0.5 * T.hi[ao] : 0.0;
        racc[rb][v] = 0.0;
      }
    }
    // two source-fragment buffers in turn: a block's map entries load two
    // blocks ahead and the data they address one block ahead, so no block
    // waits on two dependent round trips (blocks past the slice re-read its
    // last block, in bounds, and are not evaluated)
    double bf0[KC][4], bf1[KC][4];
    double bb0, wj0, hj0, bb1, wj1, hj1;
    int ib0, iw0, ib1, iw1;
    auto fetch_idx = [&](int jb, int &ib, int &iw) {
      const int col = jb + p16;
      const bool ok = col < c1;
      ib = L.bmap[T.b0 + (ok ? col : c1 - 1)];
      iw = ok ? L.wmap[T.b0 + col] : -1;
    };
    auto fetch = [&](int ib, int iw, double (&f)[KC][4], double &b2, double &ww, double &hh) {
      frag(L.XB, (long)ib * k, f);
      b2 = L.XB2[ib];
      ww = iw >= 0 ? L.w[iw] : 0.0;
      hh = (TYPE == VAR_BANDWIDTH) ? T.hj[ib] : 0.0;
    };
    auto block = [&](const double (&f)[KC][4], double bb, double wj, double hj, int ib) {
#pragma unroll
      for (int rb = 0; rb < RBP; ++rb) {
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < KC; ++c)
#pragma unroll
          for (int t = 0; t < 4; ++t)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(af[rb][c][t], f[c][t], acc, 0, 0, 0);
        if constexpr (TYPE == LAPLACE) {
          double d2[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            d2[v] = fma(-2.0, acc[v], aa[rb][v]) + bb;  // the assembly of kval_f64
            // a pair whose d^2 could fall on the other side of the drop
            // threshold, or lose more than 1e-11 through the power, takes the
            // reference's exact sequential sum (and so its d^2, bit for bit)
            if (d2[v] < lap_t0 + lap_t1 * (aa[rb][v] + bb)) {
              const int ro = rbase + rb * 16 + g + 4 * v;
              const int ao = L.amap[T.a0 + (ro < T.m ? ro : T.m - 1)];
              const double c =
                  dot_ref_order<LAPLACE>(L.XA + (long)ao * k, to_const(L.XB + (long)ib * k), k);
              d2[v] = fma(-2.0, c, aa[rb][v]) + bb;
            }
          }
          laplace_n_f64<4>(d2, P);  // the 4 kernel values, one power ladder
#pragma unroll
          for (int v = 0; v < 4; ++v) racc[rb][v] = fma(d2[v], wj, racc[rb][v]);
        } else {
#pragma unroll
          for (int v = 0; v < 4; ++v)
            racc[rb][v] =
                fma(kval_f64<TYPE, uses_exp<TYPE>>(acc[v], aa[rb][v], bb, hi[rb][v], hj, P, etab), wj, racc[rb][v]);
        }
      }
    };
    fetch_idx(c0, ib0, iw0);
    fetch_idx(min(c0 + 16, jlast), ib1, iw1);
    fetch(ib0, iw0, bf0, bb0, wj0, hj0);
    int sb0 = ib0, sb1;  // the source rows of the blocks held in bf0 / bf1
    for (int jb = c0;; jb += 32) {
      // here bf0 holds block jb and (ib1, iw1) the map entries of block jb + 16
      fetch(ib1, iw1, bf1, bb1, wj1, hj1);
      sb1 = ib1;
      fetch_idx(min(jb + 32, jlast), ib0, iw0);
      block(bf0, bb0, wj0, hj0, sb0);
      if (jb + 16 >= c1) break;
      fetch(ib0, iw0, bf0, bb0, wj0, hj0);
      sb0 = ib0;
      fetch_idx(min(jb + 48, jlast), ib1, iw1);
      block(bf1, bb1, wj1, hj1, sb1);
      if (jb + 32 >= c1) break;
    }
#pragma unroll
    for (int rb = 0; rb < RBP; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double s = racc[rb][v];
        s += __shfl_xor(s, 8, 16);
        s += __shfl_xor(s, 4, 16);
        s += __shfl_xor(s, 2, 16);
        s += __shfl_xor(s, 1, 16);
        if (p16 == v) red[(pass * RBP + rb) * 16 + g + 4 * v] = s;
      }
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  double mine = red[lane];
  __builtin_amdgcn_wave_barrier();
  return mine;
}

// MF = 1, 2: matrix-core tiles (padded K = 16 -> one 16-coordinate chunk, 32 ->
// two, 64 -> three or four); MF = 0: row tiles.
template <int TYPE, int K, int MF>
__device__ double run_tile(const ListArgs &L, const DevTask &T, int row0, int c0, int c1,
                           double *red, double *stage, const double2 *etab) {
  if constexpr (MF >= 1) {
    if constexpr (K == 16) {
      return tile_mfma<TYPE, 1>(L, T, row0, c0, c1, red, etab);
    } else if constexpr (K == 32) {
      return tile_mfma<TYPE, 2>(L, T, row0, c0, c1, red, etab);
    } else {
      if (L.k <= 48) return tile_mfma<TYPE, 3>(L, T, row0, c0, c1, red, etab);
      return tile_mfma<TYPE, 4>(L, T, row0, c0, c1, red, etab);
    }
  } else if constexpr (K > 0) {
    return tile_valu_lds<TYPE, K>(L, T, row0 + (threadIdx.x & 63), c0, c1, stage, etab);
  } else {
    return tile_valu<TYPE, K>(L, T, row0 + (threadIdx.x & 63), c0, c1, etab);
  }
}

// One persistent launch over one tile kind (MF): the row tiles, the matrix-core
// tiles, the Laplace matrix-core tiles (whose exact-recompute path needs more
// registers) and the symmetric row tiles of a batch run as four launches, so
// that each kernel's register allocation is that of its own tiles.
// Row tiles at k <= 8 wait on their gathered source chunks: 4 waves per SIMD
// (128 VGPRs, a few spilled at k = 8) hide more of it than 3 without spills
// (list_pool k = 8: 27.2 -> 24.9 ms); 5 at k = 3 (96 VGPRs, 33 spilled) lost
// (17.1 -> 17.6 ms).
template <int K, int MF>
__global__ __launch_bounds__(256, ((MF == 0 || MF == 3) && K > 0 && K <= 8) ? 4 : 1) void list_kernel(
    ListArgs L) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  __shared__ double s_red[4][64];
  double *red = s_red[wave];
  constexpr int NST = (K > 0 && (MF == 0 || MF == 3)) ? stage_doubles(K > 0 ? K : 1) + kStageSide : 1;
  __shared__ double s_stage[4][NST];
  double *stage = s_stage[wave];
  __shared__ double2 etab[64];
  load_exp_table(etab);
  __syncthreads();
  // The loop body has no lane-dependent branch outside the tile itself and no
  // `continue`: its wave-wide values (the tile index, the split count) come
  // from lane 0 through readfirstlane, which is only right while every lane
  // runs every iteration.  With lane-0-only branches and early `continue`s
  // here, the compiled loop let lanes 1..63 start an iteration without lane 0
  // after a row block's last tile (tiles rerun, counts read from an inactive
  // lane: wrong sums that depended on which wave took which tile).
  while (true) {
    // one atomic per wave: lane 0 adds 1, the other lanes 0
    unsigned t = (unsigned)__builtin_amdgcn_readfirstlane(
        (int)atomicAdd(L.head, lane == 0 ? 1u : 0u));
    if (t >= (unsigned)L.ntiles) break;
    t += (unsigned)L.tile0;
    const int2 tl = L.tiles[t];
    const RowBlock rb = L.rbs[tl.x];
    const DevTask T = L.tasks[rb.task];
    // a symmetric task's slices start at the row block (the upper triangle)
    const int c0 = (MF == 3 ? rb.row0 : 0) + tl.y * T.bc;
    const int c1 = min(T.n, c0 + T.bc);
    double s;
    if constexpr (MF == 3) {
      s = sym_tile<K>(L, T, rb.row0, c0, c1, stage, L.colp + L.coff[t], etab);
    } else if constexpr (MF == 2) {
      s = run_tile<LAPLACE, K, MF>(L, T, rb.row0, c0, c1, red, stage, etab);
    } else switch (T.type) {
      case GAUSSIAN: s = run_tile<GAUSSIAN, K, MF>(L, T, rb.row0, c0, c1, red, stage, etab); break;
      case POLYNOMIAL:
        if constexpr (MF == 0) {  // row tiles: the power fixed at compile time where it can be
          switch (poly_type(T.P)) {
            case POLY_SQUARE_T:
              s = run_tile<POLY_SQUARE_T, K, MF>(L, T, rb.row0, c0, c1, red, stage, etab);
              break;
            case POLY_CUBE_T:
              s = run_tile<POLY_CUBE_T, K, MF>(L, T, rb.row0, c0, c1, red, stage, etab);
              break;
            case POLY_FOURTH_T:
              s = run_tile<POLY_FOURTH_T, K, MF>(L, T, rb.row0, c0, c1, red, stage, etab);
              break;
            default: s = run_tile<POLYNOMIAL, K, MF>(L, T, rb.row0, c0, c1, red, stage, etab); break;
          }
        } else {
          s = run_tile<POLYNOMIAL, K, MF>(L, T, rb.row0, c0, c1, red, stage, etab);
        }
        break;
      case LAPLACE:  // row tiles only: Laplace matrix-core tiles are the MF = 2 launch
        s = MF == 0 ? run_tile<LAPLACE, K, 0>(L, T, rb.row0, c0, c1, red, stage, etab) : 0.0;
        break;
      case VAR_BANDWIDTH: s = run_tile<VAR_BANDWIDTH, K, MF>(L, T, rb.row0, c0, c1, red, stage, etab); break;
      case TANH: s = run_tile<TANH, K, MF>(L, T, rb.row0, c0, c1, red, stage, etab); break;
      case QUARTIC: s = run_tile<QUARTIC, K, MF>(L, T, rb.row0, c0, c1, red, stage, etab); break;
      case MULTIQUADRATIC: s = run_tile<MULTIQUADRATIC, K, MF>(L, T, rb.row0, c0, c1, red, stage, etab); break;
      default: s = run_tile<EPANECHNIKOV, K, MF>(L, T, rb.row0, c0, c1, red, stage, etab); break;
    }
    // A row block of one slice writes its totals; a split one writes the
    // tile's partials, which list_rowsum_kernel adds in slice order after the
    // launch (a device-scope fence and counter here would write back and
    // invalidate this XCD's L2 on gfx950, once per tile).  Lanes past the
    // block's last row write the sink entries R[rsink + lane].
    if (MF != 3 && rb.nsplit == 1) {  // wave-uniform; symmetric rows also take column sums
      const int pos = rb.row0 + lane;
      L.R[pos < T.m ? T.a0 + pos : L.rsink + lane] = s;
    } else {
      L.partial[(long)t * 64 + lane] = s;
    }
  }
}

// After a batch's persistent launches: R of every split row block is the sum
// of its tiles' partials in slice order (one wave per row block), and the
// launches' tile counters are reset for the next run.  A row r of a symmetric
// task (row block J) also takes, before its own slices, the column partials of
// the row blocks I < J for source r (from the slice of I that holds r), in
// order of I.
__global__ __launch_bounds__(256) void list_rowsum_kernel(const RowBlock *__restrict__ rbs, int nrb,
                                                          const DevTask *__restrict__ tasks,
                                                          const double *__restrict__ partial,
                                                          double *R, unsigned *heads,
                                                          const double *__restrict__ colp,
                                                          const long *__restrict__ coff,
                                                          const int *__restrict__ symrb) {
  const int lane = threadIdx.x & 63;
  const int b = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (blockIdx.x == 0 && threadIdx.x < 8) heads[threadIdx.x] = 0u;
  if (b >= nrb) return;
  const RowBlock rb = rbs[b];
  const DevTask T = tasks[rb.task];
  if (rb.nsplit == 1 && !T.sym) return;
  const int pos = rb.row0 + lane;
  if (pos >= T.m) return;
  double tot = 0.0;
  if (T.sym) {
    const int J = rb.row0 >> 6;
    for (int I = 0; I < J; ++I) {
      const int sI = (pos - 64 * I) / T.bc;  // the slice of block I holding source pos
      const long tile = (long)rbs[symrb[T.srb + I]].tile_base + sI;
      const int cs = sI == 0 ? 64 * I + 64 : 64 * I + sI * T.bc;
      tot += colp[coff[tile] + (pos - cs)];
    }
  }
  for (int sp = 0; sp < rb.nsplit; ++sp) tot += partial[(long)(rb.tile_base + sp) * 64 + lane];
  R[T.a0 + pos] = tot;
}

// Whether each candidate task's target list equals its source list: one
// workgroup per candidate, differs[c] = 1 on the first mismatch.
struct SymCand {
  long a0, b0;
  int m, pad;
};
__global__ __launch_bounds__(256) void list_symcheck_kernel(const SymCand *__restrict__ cand,
                                                            const int *__restrict__ amap,
                                                            const int *__restrict__ bmap,
                                                            int *differs) {
  const SymCand c = cand[blockIdx.x];
  int d = 0;
  for (int i = threadIdx.x; i < c.m; i += 256) d |= amap[c.a0 + i] != bmap[c.b0 + i];
  d = __syncthreads_or(d);
  if (threadIdx.x == 0) differs[blockIdx.x] = d;
}

// The compact source pool of a plan run: position i (task offset b0 + j) gets
// row XB[bmap[i]], its norm and the weight w[wmap[i]].  The tiles then read a
// chunk of a task's sources as consecutive rows instead of one line per source.
// K is the row length (k < 8), so a thread's loads are all in flight before its
// stores.
template <int K>
__global__ void list_gather_kernel(long n, const int *__restrict__ bmap,
                                   const int *__restrict__ wmap, const double *__restrict__ XB,
                                   const double *__restrict__ XB2, const double *__restrict__ w,
                                   double *__restrict__ cX, double *__restrict__ cX2,
                                   double *__restrict__ cw) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int bj = bmap[i], wi = wmap[i];
  const double *x = XB + (long)bj * K;
  double v[K];
#pragma unroll
  for (int q = 0; q < K; ++q) v[q] = x[q];
  const double b2 = XB2[bj], ww = w[wi];
#pragma unroll
  for (int q = 0; q < K; ++q) cX[i * K + q] = v[q];
  cX2[i] = b2;
  cw[i] = ww;
}


__global__ void iota_kernel(int *v, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (int)i;
}

// one thread per sorted position; the first position of each target adds the
// target's entries in position order
__global__ void merge_kernel(const int *keys, const int *pos, const double *R, long n, double *u) {
  long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int key = keys[s];
  if (s > 0 && keys[s - 1] == key) return;
  double acc = 0.0;
  for (long q = s; q < n && keys[q] == key; ++q) acc += R[pos[q]];
  u[key] += acc;
}

// ------------------------------------------------------------------ host side

template <int K, int MF>
static void launch_list(const ListArgs &A, int grid, hipStream_t s) {
  hipLaunchKernelGGL((list_kernel<K, MF>), dim3(grid), dim3(256), 0, s, A);
}

bool list_mfma_eligible(int type, int k) { return k >= 9 && k <= 64; }

// Device state of one batch layout: everything a run needs except the point
// data, so a run is kernel launches only.
struct ListPlan {
  int k = 0, precision = KS_PREC_FP64;
  long total_a = 0, ntiles = 0;
  char *buf = nullptr;
  // layout part (tiles, row blocks, tasks, partials, counters, R) of the row-tile
  // launch [0], the matrix-core launch [1] and the Laplace matrix-core launch [2]
  ListArgs A[4]{};
  int grid[4] = {0, 0, 0, 0};
  int nrb = 0;                                // row blocks of the persistent launches
  const int *keys = nullptr, *pos = nullptr;  // umap sorted, with positions
  std::vector<TaskDesc> mixed;                // tasks the mixed kernels run
  // k < 8 without variable-bandwidth tasks: every run first copies the
  // sources into a compact pool (rows, norms, weights in bmap order) that the
  // tiles read through the identity map cid.  A row shorter than a 64-byte line
  // then stops costing a line of its own per chunk (list_pool k = 3: 11.7 ->
  // 10.7 ms); at k = 8 a row is a whole line already, and the copy only adds
  // (18.7 -> 19.6 ms).
  bool compact = false;
  long total_b = 0;
  double *cX = nullptr, *cX2 = nullptr, *cw = nullptr;
  int *cid = nullptr;
  const int *umap = nullptr, *amap = nullptr, *bmap = nullptr, *wmap = nullptr;
  // symmetric tasks (list_plan_build's sym_mode): how many, and the same batch
  // laid out without symmetry for runs whose target and source tables differ
  int nsym = 0;
  ListPlan *alt = nullptr;
};

// The tasks whose target list is their source list, read on the device:
// candidates are tasks of m == n >= 128 (two row blocks or more) and, for the
// variable bandwidth, hi == hj.  Returns the flags (1 = symmetric).
static int find_symmetric(const std::vector<TaskDesc> &tasks, const int *amap, const int *bmap,
                          hipStream_t s, std::vector<char> &sym) {
  sym.assign(tasks.size(), 0);
  std::vector<SymCand> cand;
  std::vector<int> which;
  for (size_t t = 0; t < tasks.size(); ++t) {
    const TaskDesc &td = tasks[t];
    if (td.m != td.n || td.m < 128) continue;
    if (td.type == KS_GAUSSIAN_VAR_BANDWIDTH && td.hi != td.hj) continue;
    cand.push_back(SymCand{td.a0, td.b0, td.m, 0});
    which.push_back((int)t);
  }
  if (cand.empty()) return KS_OK;
  SymCand *dc = nullptr;
  int *dd = nullptr;
  const size_t cb = cand.size() * sizeof(SymCand), db = cand.size() * sizeof(int);
  if (hipMalloc((void **)&dc, cb + db) != hipSuccess) {
    (void)hipGetLastError();
    return KS_ERR_ALLOC;
  }
  dd = reinterpret_cast<int *>(reinterpret_cast<char *>(dc) + cb);
  std::vector<int> differs(cand.size(), 1);
  int rc = KS_OK;
  if (hipMemcpyAsync(dc, cand.data(), cb, hipMemcpyHostToDevice, s) != hipSuccess) rc = KS_ERR_DEVICE;
  if (rc == KS_OK) {
    hipLaunchKernelGGL(list_symcheck_kernel, dim3((unsigned)cand.size()), dim3(256), 0, s, dc, amap,
                       bmap, dd);
    if (hipMemcpyAsync(differs.data(), dd, db, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      rc = KS_ERR_DEVICE;
  }
  (void)hipFree(dc);
  if (rc != KS_OK) return rc;
  for (size_t c = 0; c < cand.size(); ++c) sym[which[c]] = differs[c] == 0;
  return KS_OK;
}

int list_plan_build(ListPlan **out, const std::vector<TaskDesc> &tasks, int k,
                    const int *umap, const int *amap, const int *bmap, const int *wmap,
                    long total_a, int precision, const TileChoice *force, hipStream_t s,
                    int sym_mode) {
  *out = nullptr;
  // symmetric tasks take row tiles with LDS staging, at k in {3, 8}: there they
  // beat the full layout (list_sym k = 3: 19.4 -> 17.4 ms, k = 8: 36.2 -> 26.6
  // ms); at k >= 9 the full layout's matrix-core tiles win (k = 32: 61 ms
  // against 279 ms for symmetric row tiles)
  std::vector<char> symf(tasks.size(), 0);
  if (sym_mode != 0 && (k == 3 || k == 8)) {
    const int rc = find_symmetric(tasks, amap, bmap, s, symf);
    if (rc != KS_OK) return rc;
  }
  ListPlan *pl = new ListPlan();
  pl->k = k;
  pl->precision = precision;
  pl->total_a = total_a;
  pl->umap = umap;
  pl->amap = amap;
  pl->bmap = bmap;
  pl->wmap = wmap;
  // ---- which tasks the persistent FP64 launch runs, with their tile shapes
  std::vector<DevTask> dt(tasks.size());
  std::vector<RowBlock> rbs;
  bool varbw = false;
  for (const TaskDesc &td : tasks) {
    pl->total_b = std::max(pl->total_b, td.b0 + td.n);
    varbw = varbw || td.type == KS_GAUSSIAN_VAR_BANDWIDTH;
  }
  pl->compact = k < 8 && !varbw && pl->total_b > 0;
  for (size_t t = 0; t < tasks.size(); ++t) {
    const TaskDesc &td = tasks[t];
    dt[t] = DevTask{td.type, td.m, td.n, 1024, 0, 0, 0, td.a0, td.b0, td.P, td.hi, td.hj};
    if (td.m == 0 || td.n == 0) continue;
    if (precision != KS_PREC_FP64 && (td.type == KS_GAUSSIAN || td.type == KS_GAUSSIAN_VAR_BANDWIDTH) &&
        k >= kMixedListMinK && td.m >= kMixedListMinSide && td.n >= kMixedListMinSide) {
      pl->mixed.push_back(td);
      continue;
    }
    const TileChoice tc = force ? *force : list_tile_choice(td.type, k, KS_PREC_FP64);
    dt[t].bc = tc.slice;
    if (symf[t]) {
      // symmetric: row tiles over the upper triangle, each row block's slices
      // starting at its own rows (slices are multiples of 64 sources)
      dt[t].sym = 1;
      dt[t].tile = 3;
      pl->nsym += 1;
      for (int r0 = 0; r0 < td.m; r0 += 64)
        rbs.push_back(RowBlock{(int)t, r0, (td.n - r0 + dt[t].bc - 1) / dt[t].bc, 0});
      continue;
    }
    dt[t].tile = (tc.tile == 1 && list_mfma_eligible(td.type, k))
                     ? (td.type == KS_LAPLACE ? 2 : 1)
                     : 0;
    const int nsplit = (td.n + dt[t].bc - 1) / dt[t].bc;
    for (int r0 = 0; r0 < td.m; r0 += 64) rbs.push_back(RowBlock{(int)t, r0, nsplit, 0});
  }
  // row blocks of row tiles first, then those of matrix-core tiles, then those of
  // Laplace matrix-core tiles, then symmetric row tiles; within each kind by
  // work (columns per tile x rows) descending; then lay tiles out
  std::vector<int> order(rbs.size());
  std::iota(order.begin(), order.end(), 0);
  auto work = [&](const RowBlock &r) {
    const DevTask &T = dt[r.task];
    long rows = std::min(64, T.m - r.row0);
    long cols = std::min(T.bc, T.sym ? T.n - r.row0 : T.n);
    return rows * cols;
  };
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    const int kx = dt[rbs[x].task].tile, ky = dt[rbs[y].task].tile;
    if (kx != ky) return kx < ky;
    return work(rbs[x]) > work(rbs[y]);
  });
  std::vector<RowBlock> rbs_sorted(rbs.size());
  for (size_t i = 0; i < order.size(); ++i) rbs_sorted[i] = rbs[order[i]];
  rbs.swap(rbs_sorted);
  std::vector<int2> tiles;
  long nkind[4] = {0, 0, 0, 0};
  for (size_t i = 0; i < rbs.size(); ++i) {
    rbs[i].tile_base = (int)tiles.size();
    for (int sp = 0; sp < rbs[i].nsplit; ++sp) tiles.push_back(make_int2((int)i, sp));
    nkind[dt[rbs[i].task].tile] += rbs[i].nsplit;
  }
  // symmetric tasks: each task's row blocks in row order, and where each tile's
  // column partials go (its sources from max(c0, row0 + 64) to c1)
  std::vector<int> symrb;
  std::vector<long> coff;
  long ncolp = 0;
  if (pl->nsym > 0) {
    for (size_t t = 0; t < dt.size(); ++t)
      if (dt[t].sym) {
        dt[t].srb = (int)symrb.size();
        symrb.resize(symrb.size() + (dt[t].m + 63) / 64, -1);
      }
    coff.assign(tiles.size(), 0);
    for (size_t i = 0; i < rbs.size(); ++i) {
      const DevTask &T = dt[rbs[i].task];
      if (!T.sym) continue;
      symrb[T.srb + rbs[i].row0 / 64] = (int)i;
      for (int sp = 0; sp < rbs[i].nsplit; ++sp) {
        const int c0 = rbs[i].row0 + sp * T.bc, c1 = std::min(T.n, c0 + T.bc);
        const int cs = std::max(c0, rbs[i].row0 + 64);
        coff[rbs[i].tile_base + sp] = ncolp;
        ncolp += std::max(0, c1 - cs);
      }
    }
  }
  pl->ntiles = (long)tiles.size();
  pl->nrb = (int)rbs.size();
  if (pl->ntiles >= (1l << 31)) {
    delete pl;
    return KS_ERR_INVALID_ARG;
  }
  if (pl->ntiles == 0) {
    *out = pl;
    return KS_OK;
  }

  // ---- device buffer: descriptors, partials, counters, R, merge order, sort space
  size_t sort_bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const int *)nullptr,
                                           (int *)nullptr, (const int *)nullptr, (int *)nullptr,
                                           (int)total_a, 0, 32, s);
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t off_tiles = 0;
  size_t off_rbs = off_tiles + al(pl->ntiles * sizeof(int2));
  size_t off_tasks = off_rbs + al(rbs.size() * sizeof(RowBlock));
  size_t off_part = off_tasks + al(dt.size() * sizeof(DevTask));
  size_t off_head = off_part + al((size_t)pl->ntiles * 64 * sizeof(double));
  size_t off_R = off_head + 256;  // the four launches' tile counters (two words each)
  size_t off_iota = off_R + al((size_t)(total_a + 64) * sizeof(double));  // + 64 sink entries
  size_t off_kout = off_iota + al((size_t)total_a * sizeof(int));
  size_t off_vout = off_kout + al((size_t)total_a * sizeof(int));
  size_t off_sort = off_vout + al((size_t)total_a * sizeof(int));
  const size_t cb = pl->compact ? (size_t)pl->total_b : 0;
  size_t off_cx = off_sort + al(sort_bytes);
  size_t off_cx2 = off_cx + al(cb * k * sizeof(double));
  size_t off_cw = off_cx2 + al(cb * sizeof(double));
  size_t off_cid = off_cw + al(cb * sizeof(double));
  size_t off_colp = off_cid + al(cb * sizeof(int));
  size_t off_coff = off_colp + al((size_t)ncolp * sizeof(double));
  size_t off_symrb = off_coff + al(coff.size() * sizeof(long));
  size_t bytes = off_symrb + al(symrb.size() * sizeof(int));
  char *buf = nullptr;
  if (hipMalloc((void **)&buf, bytes) != hipSuccess) {
    (void)hipGetLastError();
    delete pl;
    return KS_ERR_ALLOC;
  }
  pl->buf = buf;
  int rc = KS_OK;
  // the tile counters and R start at zero (list_rowsum_kernel resets the
  // counters after every run, and R rows of tasks without sources are never
  // written)
  if (hipMemcpyAsync(buf + off_tiles, tiles.data(), pl->ntiles * sizeof(int2), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(buf + off_rbs, rbs.data(), rbs.size() * sizeof(RowBlock), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(buf + off_tasks, dt.data(), dt.size() * sizeof(DevTask), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemsetAsync(buf + off_head, 0, off_iota - off_head, s) != hipSuccess)
    rc = KS_ERR_DEVICE;
  if (rc == KS_OK && pl->nsym > 0 &&
      (hipMemcpyAsync(buf + off_coff, coff.data(), coff.size() * sizeof(long), hipMemcpyHostToDevice, s) != hipSuccess ||
       hipMemcpyAsync(buf + off_symrb, symrb.data(), symrb.size() * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess))
    rc = KS_ERR_DEVICE;
  for (int kind = 0; kind < 4; ++kind) {
    ListArgs &A = pl->A[kind];
    long t0 = 0;
    for (int q = 0; q < kind; ++q) t0 += nkind[q];
    A.tile0 = (int)t0;
    A.ntiles = (int)nkind[kind];
    A.k = k;
    A.tiles = (const int2 *)(buf + off_tiles);
    A.rbs = (const RowBlock *)(buf + off_rbs);
    A.tasks = (const DevTask *)(buf + off_tasks);
    A.partial = (double *)(buf + off_part);
    A.head = (unsigned *)(buf + off_head) + 2 * kind;
    A.R = (double *)(buf + off_R);
    A.rsink = total_a;
    A.colp = pl->nsym > 0 ? (double *)(buf + off_colp) : nullptr;
    A.coff = pl->nsym > 0 ? (const long *)(buf + off_coff) : nullptr;
    A.symrb = pl->nsym > 0 ? (const int *)(buf + off_symrb) : nullptr;
  }
  // merge order: target ids stably sorted with their positions
  int *iota = (int *)(buf + off_iota);
  int *kout = (int *)(buf + off_kout);
  int *vout = (int *)(buf + off_vout);
  if (pl->compact) {
    pl->cX = (double *)(buf + off_cx);
    pl->cX2 = (double *)(buf + off_cx2);
    pl->cw = (double *)(buf + off_cw);
    pl->cid = (int *)(buf + off_cid);
    if (rc == KS_OK)
      hipLaunchKernelGGL(iota_kernel, dim3((unsigned)((pl->total_b + 255) / 256)), dim3(256), 0, s,
                         pl->cid, pl->total_b);
  }
  if (rc == KS_OK) {
    const long nb = (total_a + 255) / 256;
    hipLaunchKernelGGL(iota_kernel, dim3((unsigned)nb), dim3(256), 0, s, iota, total_a);
    size_t tb = sort_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(buf + off_sort, tb, umap, kout, (const int *)iota, vout,
                                           (int)total_a, 0, 32, s) != hipSuccess)
      rc = KS_ERR_DEVICE;
  }
  if (rc == KS_OK && hipStreamSynchronize(s) != hipSuccess) rc = KS_ERR_DEVICE;
  pl->keys = kout;
  pl->pos = vout;
  // persistent grid: enough 256-thread workgroups to hold ~4 waves per SIMD
  int dev = 0, ncu = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  for (int kind = 0; kind < 4; ++kind) {
    long grid = std::min<long>((long)ncu * 4, (nkind[kind] + 3) / 4);
    pl->grid[kind] = (int)std::max<long>(grid, 1);
  }
  // sym_mode 1: the target and source tables of a run are not known yet; the
  // same batch without symmetry serves runs where they differ
  if (rc == KS_OK && sym_mode == 1 && pl->nsym > 0)
    rc = list_plan_build(&pl->alt, tasks, k, umap, amap, bmap, wmap, total_a, precision, force, s, 0);
  if (rc != KS_OK) {
    list_plan_free(pl);
    return rc;
  }
  *out = pl;
  return KS_OK;
}

template <int K>
static void launch_list_gather(const ListPlan &pl, const ListArrays &L, hipStream_t s) {
  hipLaunchKernelGGL((list_gather_kernel<K>), dim3((unsigned)((pl.total_b + 255) / 256)), dim3(256),
                     0, s, pl.total_b, pl.bmap, pl.wmap, L.XB, L.XB2, L.w, pl.cX, pl.cX2, pl.cw);
}

int list_plan_exec(const ListPlan &pl, const ListArrays &L, hipStream_t s) {
  // a symmetric task's pairs are evaluated once only if the sources are the
  // targets in the same table
  if (pl.alt && !(L.XA == L.XB && L.XA2 == L.XB2)) return list_plan_exec(*pl.alt, L, s);
  if (pl.ntiles > 0) {
    const int *bm = pl.bmap, *wm = pl.wmap;
    const double *xb = L.XB, *xb2 = L.XB2, *ww = L.w;
    if (pl.compact) {
      switch (pl.k) {
        case 1: launch_list_gather<1>(pl, L, s); break;
        case 2: launch_list_gather<2>(pl, L, s); break;
        case 3: launch_list_gather<3>(pl, L, s); break;
        case 4: launch_list_gather<4>(pl, L, s); break;
        case 5: launch_list_gather<5>(pl, L, s); break;
        case 6: launch_list_gather<6>(pl, L, s); break;
        default: launch_list_gather<7>(pl, L, s); break;
      }
      bm = wm = pl.cid;
      xb = pl.cX;
      xb2 = pl.cX2;
      ww = pl.cw;
    }
    for (int kind = 0; kind < 4; ++kind) {
      ListArgs A = pl.A[kind];
      if (A.ntiles == 0) continue;
      A.amap = pl.amap;
      A.bmap = bm;
      A.wmap = wm;
      A.XA = L.XA;
      A.XA2 = L.XA2;
      A.XB = xb;
      A.XB2 = xb2;
      A.w = ww;
      // k = 3 and 8: rows and sources in registers / scalar loads; 9 <= k <= 64:
      // padded to 16, 32 or 64 (row tiles stage the sources in LDS); other k:
      // generic row tiles.  Matrix-core tiles only exist for 9 <= k <= 64.
      const int kk = pl.k, g = pl.grid[kind];
      if (kind == 1) {
        if (kk <= 16) launch_list<16, 1>(A, g, s);
        else if (kk <= 32) launch_list<32, 1>(A, g, s);
        else launch_list<64, 1>(A, g, s);
      } else if (kind == 2) {
        if (kk <= 16) launch_list<16, 2>(A, g, s);
        else if (kk <= 32) launch_list<32, 2>(A, g, s);
        else launch_list<64, 2>(A, g, s);
      } else if (kind == 3) {  // symmetric row tiles: k in {3, 8}
        if (kk == 3) launch_list<3, 3>(A, g, s);
        else launch_list<8, 3>(A, g, s);
      } else if (kk == 3) launch_list<3, 0>(A, g, s);
      else if (kk == 8) launch_list<8, 0>(A, g, s);
      else if (kk >= 9 && kk <= 16) launch_list<16, 0>(A, g, s);
      else if (kk >= 17 && kk <= 32) launch_list<32, 0>(A, g, s);
      else if (kk >= 33 && kk <= 64) launch_list<64, 0>(A, g, s);
      else launch_list<0, 0>(A, g, s);
    }
    const double *R = pl.A[0].R;
    hipLaunchKernelGGL(list_rowsum_kernel, dim3((unsigned)((pl.nrb + 3) / 4)), dim3(256), 0, s,
                       pl.A[0].rbs, pl.nrb, pl.A[0].tasks, pl.A[0].partial, pl.A[0].R, pl.A[0].head,
                       pl.A[0].colp, pl.A[0].coff, pl.A[0].symrb);
    const long nb = (pl.total_a + 255) / 256;
    hipLaunchKernelGGL(merge_kernel, dim3((unsigned)nb), dim3(256), 0, s, pl.keys, pl.pos, R,
                       pl.total_a, L.u);
    if (hipGetLastError() != hipSuccess) return KS_ERR_DEVICE;
  }
  for (const TaskDesc &t : pl.mixed) {
    Problem p;
    p.m = t.m; p.n = t.n; p.k = pl.k; p.type = t.type; p.P = t.P;
    p.u = L.u; p.umap = pl.umap + t.a0; p.XA = L.XA; p.XA2 = L.XA2; p.amap = pl.amap + t.a0;
    p.XB = L.XB; p.XB2 = L.XB2; p.bmap = pl.bmap + t.b0; p.w = L.w; p.wmap = pl.wmap + t.b0;
    p.hi = t.hi; p.hj = t.hj;
    int rc = enqueue_mixed(p, pl.precision, s);
    if (rc != KS_OK) return rc;
  }
  return KS_OK;
}

void list_plan_free(ListPlan *pl) {
  if (!pl) return;
  list_plan_free(pl->alt);
  if (pl->buf) (void)hipFree(pl->buf);
  delete pl;
}

long list_plan_tiles(const ListPlan *pl) { return pl ? pl->ntiles : 0; }
int list_plan_symmetric_tasks(const ListPlan *pl) { return pl ? pl->nsym : 0; }

}  // namespace gsks
