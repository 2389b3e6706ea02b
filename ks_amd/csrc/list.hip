// Batched summation: every task of a list on one stream.
#include <hip/hip_runtime.h>

#include "host_util.h"
#include "list_launch.h"

namespace gsks {

int list_enqueue(const std::vector<TaskDesc> &tasks, int k, const ListArrays &L, int precision,
                 hipStream_t s) {
  for (const TaskDesc &t : tasks) {
    if (t.m == 0 || t.n == 0) continue;
    Problem p;
    p.m = t.m;
    p.n = t.n;
    p.k = k;
    p.type = t.type;
    p.P = t.P;
    p.u = L.u;
    p.umap = L.umap + t.a0;
    p.XA = L.XA;
    p.XA2 = L.XA2;
    p.amap = L.amap + t.a0;
    p.XB = L.XB;
    p.XB2 = L.XB2;
    p.bmap = L.bmap + t.b0;
    p.w = L.w;
    p.wmap = L.wmap + t.b0;
    p.hi = t.hi;
    p.hj = t.hj;
    int rc = precision == KS_PREC_FP64 ? enqueue_fused(p, s) : enqueue_mixed(p, precision, s);
    if (rc != KS_OK) return rc;
  }
  return KS_OK;
}

}  // namespace gsks
