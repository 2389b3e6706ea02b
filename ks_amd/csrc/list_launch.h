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

// Enqueue every task on stream s; all ListArrays pointers are device pointers.
int list_enqueue(const std::vector<TaskDesc> &tasks, int k, const ListArrays &L, int precision,
                 hipStream_t s);

int list_run(int ntask, const ks_t *kernels, int k, double *u, long nu, const long *a_off,
             const int *umap, const int *amap, const long *b_off, const int *bmap,
             const int *wmap, const double *XA, const double *XA2, long nxa_hint,
             const double *XB, const double *XB2, long nxb_hint, const double *w, int precision,
             const char *name);

}  // namespace gsks
