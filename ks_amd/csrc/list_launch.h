// Interface between the list host code and the list kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "gsks_kernels.h"

namespace gsks {

struct TaskDesc {
  int type;
  KerParams P;
  long a0, b0;  // offsets into amap/umap and bmap/wmap
  int m, n;
  const double *hi, *hj;
};

struct ListArrays {
  double *u = nullptr;
  const int *umap = nullptr, *amap = nullptr, *bmap = nullptr, *wmap = nullptr;
  const double *XA = nullptr, *XA2 = nullptr, *XB = nullptr, *XB2 = nullptr, *w = nullptr;
};

// Tile shape of a task: sources per tile, and the tile kind (1: 64 rows as four
// 16 x 16 blocks on the FP64 matrix cores, for 9 <= k <= 64 except Laplace;
// 0: one row per lane on the vector units).
struct TileChoice {
  int slice;
  int tile;
};
bool list_mfma_eligible(int type, int k);
// The shape for tasks of one kernel type at one k and precision: tuned once per
// process (list_tune.cpp) and cached; KS_AUTOTUNE=0 keeps the default shape.
TileChoice list_tile_choice(int type, int k, int precision);
TileChoice list_default_tile(int type, int k);

// In a mixed-precision batch the Gaussian tasks with k >= kMixedListMinK and
// m, n >= kMixedListMinSide run through the mixed kernels (one pre-pass and
// launch sequence per task); every other task runs in the persistent FP64
// launch.  The sides are those at which the mixed form buckets its points and
// beats the FP64 kernels (the single-precision front end uses the same rule).
// Routing tasks of 2^24 pairs and more, as once, made list_pool k = 32 slower
// than FP64 (41 ms against 27.7 ms): below these sides one persistent launch
// beats a launch sequence per task.
constexpr int kMixedListMinSide = 16384;
constexpr int kMixedListMinK = 16;

// A batch layout on the device (list.hip).  Building it does the host work
// (tile layout, descriptors, the merge order of umap) and synchronises; running
// it only enqueues kernels, so a run can be captured into a HIP graph.  The
// maps are device arrays that must outlive the plan.
// sym_mode: 0 evaluates every task in full; 2 evaluates each pair of a task
// whose target list is its source list once (the caller's target and source
// tables are one: XA == XB, XA2 == XB2); 1 does so when a run's tables are one
// and keeps the full layout beside it for runs where they are not.
struct ListPlan;
int list_plan_build(ListPlan **out, const std::vector<TaskDesc> &tasks, int k, const int *umap,
                    const int *amap, const int *bmap, const int *wmap, long total_a,
                    int precision, const TileChoice *force, hipStream_t s, int sym_mode);
int list_plan_exec(const ListPlan &pl, const ListArrays &L, hipStream_t s);
void list_plan_free(ListPlan *pl);
long list_plan_tiles(const ListPlan *pl);
int list_plan_symmetric_tasks(const ListPlan *pl);

int list_run(int ntask, const ks_t *kernels, int k, double *u, long nu, const long *a_off,
             const int *umap, const int *amap, const long *b_off, const int *bmap,
             const int *wmap, const double *XA, const double *XA2, long nxa_hint,
             const double *XB, const double *XB2, long nxb_hint, const double *w, int precision,
             const char *name);

}  // namespace gsks
