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
  unsigned *counters;  // [row_blocks], zero between launches
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
FusedPlan plan_fused(int m, int n, int k, int target_wgs, int min_cols);
hipError_t launch_fused_f64(const Problem &p, const FusedPlan &plan, double *partial,
                            unsigned *counters, hipStream_t s);

}  // namespace gsks
