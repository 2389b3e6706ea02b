// FP64 fused kernel summation with the cross term on the matrix cores
// (v_mfma_f64_16x16x4_f64), for 9 <= k <= 64.
//
// Mapping:
//   * a workgroup of 4 waves owns 4 RB blocks of 16 target rows; each wave
//     loads the coordinates of its RB blocks once, in the MFMA A-operand
//     layout, and keeps them in VGPRs;
//   * the workgroup walks its source slice in stages of S sources.  Its 256
//     lanes stage the sources' coordinates (zero-padded to 16 KC), norms,
//     weights, bandwidths and rows into LDS once (double-buffered), and the four
//     waves read their B fragments from there instead of each wave fetching
//     the same fragments from L2 (lane l reads point l%16, coordinates
//     4*(l/16) .. +3 of each 16-coordinate chunk; A uses the same coordinate
//     order, so the MFMA k order is a permutation of 0..k-1 shared by both
//     operands).  Map entries load two stages ahead and coordinates one stage
//     ahead, so no stage waits on a dependent pair of round trips;
//   * the 16 x 16 result block (lane l holds rows l/16 + 4v, v = 0..3, of
//     column l%16 -- the layout measured on gfx950) is turned into
//     d^2 = |a|^2 - 2c + |b|^2, K(d^2) and K*w in registers; each lane keeps
//     4*RB row partials across the whole slice, and the 16 lanes of a row
//     group are summed once at the end;
//   * within a stage the products of row block rb for block b + 1 are issued
//     as soon as its kernel values for block b are formed (one accumulator per
//     row block): the FP64 matrix and vector instructions share the SIMD's
//     FP64 issue, so the gain is in latency hidden, not in overlap;
//   * Gaussian kernel values take exp_tab (a table-driven exp, about 15
//     vector instructions instead of about 22);
//   * split sources leave partial sums that launch_split_reduce adds in split
//     order.
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

template <int KC>
struct Stage {
  static constexpr int NB = KC <= 2 ? 4 : 2;  // 16-source blocks per stage
  static constexpr int S = 16 * NB;           // sources per stage
  static constexpr int KP = 16 * KC + 2;      // LDS row of one source (doubles, padded)
  static constexpr int TPP = MWG / S;         // lanes per staged source
  static constexpr int CPT = 16 * KC / TPP;   // coordinates per lane (even)
};

// k <= 16 (one chunk) fits 3 waves per SIMD in 168 registers without spills,
// the variable-bandwidth form aside; larger chunks run at 2.
template <int TYPE, int KC, int RB, bool EXACT>
__global__ __launch_bounds__(MWG, (KC == 1 && TYPE != VAR_BANDWIDTH ? 3 : 1)) void fused_f64_mfma_kernel(FusedArgs A) {
  using St = Stage<KC>;
  constexpr int S = St::S, KP = St::KP, NB = St::NB, CPT = St::CPT;
  __shared__ __align__(16) double sx[2][S * KP];
  __shared__ double sb[2][S], sw[2][S], sh[2][S];
  __shared__ int si[2][S];
  __shared__ double2 etab[64];
  load_exp_table(etab);

  const Problem &p = A.p;
  const int m = p.m, k = p.k;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int p16 = lane & 15;
  const int g = lane >> 4;
  const int row0 = blockIdx.x * (4 * RB * 16) + wave * (RB * 16);

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

  const int j0 = blockIdx.y * A.cols_per_split;
  const int j1 = min(p.n, j0 + A.cols_per_split);
  const int nst = (j1 - j0 + S - 1) / S;
  const KerParams P = p.P;

  // this lane's part of a stage: source `pt`, coordinates [q0, q0 + CPT)
  const int pt = threadIdx.x / St::TPP;
  const int q0 = (threadIdx.x % St::TPP) * CPT;
  const bool head = q0 == 0;  // also loads the source's norm, weight, bandwidth
  auto fetch_idx = [&](int st, int &ib, int &iw) {
    const int col = j0 + st * S + pt;
    const bool ok = col < j1;
    ib = p.bmap[ok ? col : j1 - 1];
    iw = (ok && head) ? p.wmap[col] : -1;
  };
  double cx[CPT], cb = 0.0, cw = 0.0, ch = 0.0;
  int cib = 0;
  auto fetch = [&](int ib, int iw) {
    const double *src = p.XB + (long)ib * k + q0;
    if constexpr (EXACT) {
      const double2 *q = reinterpret_cast<const double2 *>(src);
#pragma unroll
      for (int i = 0; i < CPT / 2; ++i) {
        const double2 x = q[i];
        cx[2 * i] = x.x;
        cx[2 * i + 1] = x.y;
      }
    } else {
#pragma unroll
      for (int i = 0; i < CPT; ++i) cx[i] = (q0 + i < k) ? src[i] : 0.0;
    }
    cib = ib;
    if (head) {
      cb = p.XB2[ib];
      cw = iw >= 0 ? p.w[iw] : 0.0;
      ch = (TYPE == VAR_BANDWIDTH) ? p.hj[ib] : 0.0;
    }
  };
  auto stash = [&](int buf) {
    double2 *d = reinterpret_cast<double2 *>(&sx[buf][pt * KP + q0]);
#pragma unroll
    for (int i = 0; i < CPT / 2; ++i) d[i] = make_double2(cx[2 * i], cx[2 * i + 1]);
    if (head) {
      sb[buf][pt] = cb;
      sw[buf][pt] = cw;
      sh[buf][pt] = ch;
      si[buf][pt] = cib;
    }
  };

  const double lap_b = (2.0 * (k + 2) + 4.0) * 0x1p-53;
  const double lap_t1 = lap_b * (1.0 + fmax(0.5 * k - 1.0, 1.0) * 1e11) + lap_b;
  auto frag = [&](int buf, int b, double (&f)[KC][4]) {
    const double *row = &sx[buf][(b * 16 + p16) * KP + 4 * g];
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      const double2 x0 = *reinterpret_cast<const double2 *>(row + 16 * c);
      const double2 x1 = *reinterpret_cast<const double2 *>(row + 16 * c + 2);
      f[c][0] = x0.x;
      f[c][1] = x0.y;
      f[c][2] = x1.x;
      f[c][3] = x1.y;
    }
  };
  auto cross = [&](const double (&f)[KC][4], int rb) {
    d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < KC; ++c)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(af[rb][c][t], f[c][t], acc, 0, 0, 0);
    return acc;
  };
  auto values = [&](int buf, int b, int rb, const d4 &acc) {
    const int col = b * 16 + p16;
    const double bb = sb[buf][col], wj = sw[buf][col];
    const double hj = (TYPE == VAR_BANDWIDTH) ? sh[buf][col] : 0.0;
    if constexpr (TYPE == LAPLACE) {
      double d2[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        d2[v] = fma(-2.0, acc[v], aa[rb][v]) + bb;  // the assembly of kval_f64
        if (d2[v] < 1e-15 + lap_t1 * (aa[rb][v] + bb)) {
          const int ro = row0 + rb * 16 + g + 4 * v;
          const int ao = p.amap[ro < m ? ro : m - 1];
          const double c = dot_exact_n(p.XA + (long)ao * k, p.XB + (long)si[buf][col] * k, k);
          d2[v] = fma(-2.0, c, aa[rb][v]) + bb;
        }
      }
      laplace_n_f64<4>(d2, P);
#pragma unroll
      for (int v = 0; v < 4; ++v) racc[rb][v] = fma(d2[v], wj, racc[rb][v]);
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double K = kval_f64<TYPE, true>(acc[v], aa[rb][v], bb, hi[rb][v], hj, P, etab);
        racc[rb][v] = fma(K, wj, racc[rb][v]);
      }
    }
  };

  int ib1 = 0, iw1 = -1, ib2 = 0, iw2 = -1;
  if (nst > 0) {
    int ib0, iw0;
    fetch_idx(0, ib0, iw0);
    if (nst > 1) fetch_idx(1, ib1, iw1);
    fetch(ib0, iw0);
    stash(0);
  }
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const int buf = st & 1;
    if (st + 1 < nst) fetch(ib1, iw1);
    if (st + 2 < nst) fetch_idx(st + 2, ib2, iw2);
    // a half-block pipeline: row block rb's products for block b + 1 are
    // issued as soon as its kernel values for block b are formed, so the
    // matrix core works on one row block while the VALU works on the other,
    // with one accumulator per row block
    double f[KC][4], fn[KC][4];
    d4 acc[RB];
    frag(buf, 0, f);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) acc[rb] = cross(f, rb);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      if (b + 1 < NB) frag(buf, b + 1, fn);
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
        values(buf, b, rb, acc[rb]);
        if (b + 1 < NB) {
          // empty asm statements that take the row sums and the next fragment
          // as operands: the products for block b + 1 cannot move above the
          // kernel values of block b (left to itself the compiler issues every
          // product of the stage first and runs out of registers)
          asm volatile("" : "+v"(racc[rb][0]), "+v"(racc[rb][1]), "+v"(racc[rb][2]),
                       "+v"(racc[rb][3]));
#pragma unroll
          for (int c = 0; c < KC; ++c)
            asm volatile("" : "+v"(fn[c][0]), "+v"(fn[c][1]), "+v"(fn[c][2]), "+v"(fn[c][3]));
          acc[rb] = cross(fn, rb);
        }
      }
      if (b + 1 < NB)
#pragma unroll
        for (int c = 0; c < KC; ++c)
#pragma unroll
          for (int t = 0; t < 4; ++t) f[c][t] = fn[c][t];
    }
    if (st + 1 < nst) stash(buf ^ 1);
    __syncthreads();
    ib1 = ib2;
    iw1 = iw2;
  }

  // Each 16-lane group holds 4 row sums per row block, spread over its
  // lanes.  A transposing reduction leaves lane p16 with the whole sum of
  // row v = p16 >> 2 (halving the values held at each of the first two
  // steps), and lane 4 v of the group writes row g + 4 v.  (A select of
  // racc[rb][p16] instead would be a lane-indexed read, which puts racc in
  // scratch memory.)
  const bool h8 = p16 & 8, h4 = p16 & 4;
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    const double s0 = (h8 ? racc[rb][2] : racc[rb][0]) +
                      __shfl_xor(h8 ? racc[rb][0] : racc[rb][2], 8, 16);
    const double s1 = (h8 ? racc[rb][3] : racc[rb][1]) +
                      __shfl_xor(h8 ? racc[rb][1] : racc[rb][3], 8, 16);
    double t = (h4 ? s1 : s0) + __shfl_xor(h4 ? s0 : s1, 4, 16);
    t += __shfl_xor(t, 2, 16);
    t += __shfl_xor(t, 1, 16);
    const int r = row0 + rb * 16 + g + 4 * (p16 >> 2);
    if ((p16 & 3) != 0 || r >= m) continue;
    if (A.nsplit == 1)
      p.u[p.umap[r]] += t;
    else
      A.partial[(long)blockIdx.y * m + r] = t;
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
  __shared__ double2 etab[64];
  if constexpr (uses_exp<TYPE>) {
    load_exp_table(etab);
    __syncthreads();
  }

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
        const double K = kval_f64<TYPE, uses_exp<TYPE>>(acc[v], aa[rb][v], bb, hi[rb][v], hj, P, etab);
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
                                 hipStream_t s) {
  FusedArgs A;
  A.p = p;
  A.cols_per_split = plan.cols_per_split;
  A.nsplit = plan.nsplit;
  A.partial = partial;
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
                                     double *colp, hipStream_t s) {
  SymMArgs A;
  A.p = p;
  A.cols_per_split = pl.cols_per_split;
  A.nsplit = pl.nsplit;
  A.partial = partial;
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
