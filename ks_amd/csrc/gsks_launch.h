// Host-visible launch interface of the HIP kernels (no torch types, no host
// state): the C-ABI layer in api.cpp plans and enqueues through these.
#pragma once

#include <hip/hip_runtime.h>

#include "gsks_kernels.h"

namespace gsks {

struct FusedArgs {
  Problem p;
  int cols_per_split;
  int nsplit;
  double *partial;     // [nsplit][m] partial sums (nsplit > 1)
  int vec2;            // point rows are 16-byte aligned (k even, aligned tables)
};

struct FusedPlan {
  int kspec;           // specialised k, 0 = generic
  int rows_per_wg;
  int row_blocks;
  int cols_per_split;
  int nsplit;
};

int fused_rows_per_lane(int k);
int fused_kspec(int k);
// Split the sources of an m x n problem so that about target_wgs workgroups of
// rows_per_wg target rows each run, with at least min_cols sources per split.
FusedPlan plan_split(int m, int n, int rows_per_wg, int target_wgs, int min_cols);
FusedPlan plan_fused(int m, int n, int k, int target_wgs, int min_cols);
hipError_t launch_fused_f64(const Problem &p, const FusedPlan &plan, double *partial,
                            hipStream_t s);

// Symmetric sums (fused_f64_sym.hip): sources == targets, each pair of
// different row blocks evaluated once for both of its terms.
struct SymPlan {
  int rows_per_wg;
  int row_blocks;
  int cols_per_split;  // a multiple of rows_per_wg
  int nsplit;
};
bool sym_eligible(const Problem &p);
// Rows per lane of the symmetric vector kernel at dimension k: six at k <= 4
// (more independent kernel-value chains per source and fewer source loads per
// pair; N-body k = 3: Gaussian 38.5 -> 35.5 ms, Laplace 38.0 -> 35.4 ms; eight
// rows lost occupancy: 38.3 and 42.3 ms), four at k = 8, whose rows take twice
// the registers.  The plan and the kernel template both take it from here.
constexpr int sym_rows_per_lane(int k) { return k <= 4 ? 6 : 4; }
SymPlan plan_sym(int m, int k, int target_wgs);
SymPlan plan_sym_rows(int m, int target_wgs, int rows_per_wg);
// The last step of a symmetric sum: u[umap[j]] += the split partials of j's row
// block (nsplit > 1), then += sum_{I < J} colp[I][j] over row blocks of
// pl.rows_per_wg targets (J = j's block)
hipError_t launch_sym_colsum(const Problem &p, const SymPlan &pl, const double *partial,
                             const double *colp, hipStream_t s);
// partial: [nsplit][m]; colp: [row_blocks][m]
hipError_t launch_fused_f64_sym(const Problem &p, const SymPlan &pl, double *partial,
                                double *colp, hipStream_t s);
// Compact tables of a symmetric problem (the mapped rows, norms, weights and
// bandwidths in map order, and an identity map) in `tab` (sym_table_doubles(p)
// doubles, 256-byte aligned); *q is p over those tables, writing through p's umap.
size_t sym_table_doubles(const Problem &p);
hipError_t launch_sym_gather(const Problem &p, double *tab, Problem *q, hipStream_t s);
// The same for the sources of a general problem (n points through bmap / wmap);
// *q reads its sources from the tables, its targets as p does.
size_t gather_table_doubles(long count, int k, int type);
hipError_t launch_source_gather(const Problem &p, double *tab, Problem *q, hipStream_t s);

// u[umap[r]] += sum over sp < nsplit of partial[sp][i] (added in split order),
// r = prow ? prow[i] : i, for every i < m: the split-source reduction of the
// fused kernels, in a launch of its own.  The launch boundary orders the
// partials' stores before these loads, so the kernels that write them need no
// device-scope fences (which write back and invalidate the XCD's L2 on gfx950).
hipError_t launch_split_reduce(const double *partial, int nsplit, int m, double *u,
                               const int *umap, const int *prow, hipStream_t s);
// The same for nrhs right-hand sides: partial [nsplit][m][nrhs],
// u[umap[i] nrhs + r] += sum over sp of partial[sp][i][r].
hipError_t launch_split_reduce_rhs(const double *partial, int nsplit, int m, int nrhs, double *u,
                                   const int *umap, hipStream_t s);

// FP64 cross term on the matrix cores (fused_f64_mfma.hip)
bool mfma_f64_eligible(int type, int k);
// The single-summation matrix-core kernel: every type, Laplace included.
bool fused_mfma_eligible(int type, int k);
int mfma_rows_per_wg();
// ... and its symmetric form (row blocks of sym_mfma_rows() targets)
bool sym_mfma_eligible(const Problem &p);
int sym_mfma_rows();
hipError_t launch_fused_f64_sym_mfma(const Problem &p, const SymPlan &pl, double *partial,
                                     double *colp, hipStream_t s);
hipError_t launch_fused_f64_mfma(const Problem &p, const FusedPlan &plan, double *partial,
                                 hipStream_t s);

}  // namespace gsks
