// dgsks_list(): many independent summations over shared point pools in one call
// (the batched path of frame/omp_dgsks_list.cpp:297-420, with one kernel
// descriptor per task).  Host arrays are staged through HBM once for the whole
// batch; device arrays are used in place.  The tasks run on the GPU through a
// batch plan (list.hip), built and freed within the call; ks_list_plan_* keeps
// a plan across runs for callers that replay the same batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "host_util.h"
#include "list_launch.h"

namespace gsks {

static long max_index_range(const int *map, long len) {
  long mx = -1;
  for (long i = 0; i < len; ++i) mx = std::max(mx, (long)map[i]);
  return mx;
}

int list_run(int ntask, const ks_t *kernels, int k, double *u, long nu, const long *a_off,
             const int *umap, const int *amap, const long *b_off, const int *bmap,
             const int *wmap, const double *XA, const double *XA2, long nxa_hint,
             const double *XB, const double *XB2, long nxb_hint, const double *w, int precision,
             const char *name) {
  if (ntask < 0 || k < 0 || nu < 0 || (ntask > 0 && (!kernels || !a_off || !b_off)))
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): invalid argument\n", name);
  if (ntask == 0 || k == 0) return KS_OK;
  for (int t = 0; t < ntask; ++t) {
    if (a_off[t + 1] < a_off[t] || b_off[t + 1] < b_off[t] || a_off[t] < 0 || b_off[t] < 0)
      return raise_error(KS_ERR_LIST, false, "%s(): offsets must be non-decreasing\n", name);
  }
  if (a_off[0] != 0)
    return raise_error(KS_ERR_LIST, false, "%s(): a_off[0] must be 0\n", name);
  const long total_a = a_off[ntask], total_b = b_off[ntask];
  if (total_a == 0 || total_b == 0) return KS_OK;
  if (!u || !umap || !amap || !bmap || !wmap || !XA || !XB || !w)
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): NULL array\n", name);
  if (ensure_device() != KS_OK)
    return raise_error(KS_ERR_DEVICE, true, "%s(): no GPU available\n", name);
  std::vector<TaskDesc> tasks(ntask);
  for (int t = 0; t < ntask; ++t) {
    KerParams P;
    int rc = make_params(&kernels[t], k, P, true);
    if (rc != KS_OK && rc != KS_ERR_LAPLACE_K) return rc;
    tasks[t].type = ker_type(&kernels[t]);
    tasks[t].P = P;
    tasks[t].a0 = a_off[t];
    tasks[t].m = (int)(a_off[t + 1] - a_off[t]);
    tasks[t].b0 = b_off[t];
    tasks[t].n = (int)(b_off[t + 1] - b_off[t]);
    tasks[t].hi = kernels[t].hi;
    tasks[t].hj = kernels[t].hj;
  }
  hipStream_t s = thread_stream();
  if (!s) return raise_error(KS_ERR_DEVICE, true, "%s(): stream creation failed\n", name);

  std::vector<const void *> ptrs = {u, umap, amap, bmap, wmap, XA, XB, w};
  if (XA2) ptrs.push_back(XA2);
  if (XB2) ptrs.push_back(XB2);
  for (auto &td : tasks)
    if (td.type == KS_GAUSSIAN_VAR_BANDWIDTH) {
      ptrs.push_back(td.hi);
      ptrs.push_back(td.hj);
    }
  int ndev = 0;
  for (const void *q : ptrs) ndev += is_device_ptr(q) ? 1 : 0;
  const bool on_device = ndev == (int)ptrs.size();
  if (!on_device && ndev != 0)
    return raise_error(KS_ERR_MIXED_POINTERS, false, "%s(): %s\n", name,
                       err_string(KS_ERR_MIXED_POINTERS));

  ListArrays L;
  std::vector<void *> owned;
  auto dalloc = [&](size_t bytes) -> void * {
    void *d = nullptr;
    if (hipMallocAsync(&d, bytes ? bytes : 8, s) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    owned.push_back(d);
    return d;
  };
  auto release = [&]() {
    for (void *d : owned) (void)hipFreeAsync(d, s);
    owned.clear();
  };
  int rc = KS_OK;
  long nxa = nxa_hint, nxb = nxb_hint;
  if (on_device) {
    L.u = u;
    L.umap = umap;
    L.amap = amap;
    L.bmap = bmap;
    L.wmap = wmap;
    L.XA = XA;
    L.XB = XB;
    L.w = w;
    if (!XA2 || !XB2) {
      if (nxa <= 0 || nxb <= 0) {
        return raise_error(KS_ERR_INVALID_ARG, false,
                           "%s(): device pools without norms need their sizes\n", name);
      }
    }
  } else {
    for (long i = 0; i < total_a; ++i)
      if (amap[i] < 0 || umap[i] < 0 || umap[i] >= nu)
        return raise_error(KS_ERR_INVALID_ARG, false, "%s(): target index out of range\n", name);
    for (long j = 0; j < total_b; ++j)
      if (bmap[j] < 0 || wmap[j] < 0)
        return raise_error(KS_ERR_INVALID_ARG, false, "%s(): negative source index\n", name);
    nxa = std::max(nxa, max_index_range(amap, total_a) + 1);
    nxb = std::max(nxb, max_index_range(bmap, total_b) + 1);
    const long nw = max_index_range(wmap, total_b) + 1;
    const bool same_x = XA == XB && XA2 == XB2;
    if (same_x) nxa = nxb = std::max(nxa, nxb);
    auto up = [&](const void *src, size_t bytes) -> void * {
      void *d = dalloc(bytes);
      if (!d) {
        rc = KS_ERR_ALLOC;
        return nullptr;
      }
      if (bytes && hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, s) != hipSuccess)
        rc = KS_ERR_DEVICE;
      return d;
    };
    L.u = (double *)up(u, (size_t)nu * 8);
    L.umap = (const int *)up(umap, (size_t)total_a * 4);
    L.amap = (const int *)up(amap, (size_t)total_a * 4);
    L.bmap = (const int *)up(bmap, (size_t)total_b * 4);
    L.wmap = (const int *)up(wmap, (size_t)total_b * 4);
    L.XA = (const double *)up(XA, (size_t)nxa * k * 8);
    L.XB = same_x ? L.XA : (const double *)up(XB, (size_t)nxb * k * 8);
    L.w = (const double *)up(w, (size_t)nw * 8);
    if (XA2) L.XA2 = (const double *)up(XA2, (size_t)nxa * 8);
    if (XB2) L.XB2 = same_x ? L.XA2 : (const double *)up(XB2, (size_t)nxb * 8);
    // bandwidth vectors: hi is indexed by target ids (< nxa), hj by source ids
    // (< nxb); a vector used in both roles covers both ranges
    std::map<const double *, long> hext;
    for (auto &td : tasks) {
      if (td.type != KS_GAUSSIAN_VAR_BANDWIDTH) continue;
      hext[td.hi] = std::max(hext[td.hi], nxa);
      hext[td.hj] = std::max(hext[td.hj], nxb);
    }
    std::map<const double *, const double *> hmap;
    for (auto &e : hext) hmap[e.first] = (const double *)up(e.first, (size_t)e.second * 8);
    for (auto &td : tasks) {
      if (td.type != KS_GAUSSIAN_VAR_BANDWIDTH) continue;
      td.hi = hmap[td.hi];
      td.hj = hmap[td.hj];
    }
    if (rc != KS_OK) {
      release();
      (void)hipStreamSynchronize(s);
      return raise_error(rc, false, "%s(): staging failed\n", name);
    }
  }
  if (on_device) {
    L.XA2 = XA2;
    L.XB2 = XB2;
  }
  if (!L.XA2) {
    double *d = (double *)dalloc((size_t)nxa * 8);
    if (!d) rc = KS_ERR_ALLOC;
    else if (ks_sqnorm_device(L.XA, nxa, k, d, s) != KS_OK) rc = KS_ERR_DEVICE;
    L.XA2 = d;
  }
  if (!L.XB2) {
    if (L.XB == L.XA && nxb <= nxa) {
      L.XB2 = L.XA2;
    } else {
      double *d = (double *)dalloc((size_t)nxb * 8);
      if (!d) rc = KS_ERR_ALLOC;
      else if (ks_sqnorm_device(L.XB, nxb, k, d, s) != KS_OK) rc = KS_ERR_DEVICE;
      L.XB2 = d;
    }
  }
  if (rc == KS_OK) {
    ListPlan *pl = nullptr;
    // one table for targets and sources: tasks whose lists are one set evaluate
    // each pair once
    const int sym_mode = (L.XA == L.XB && L.XA2 == L.XB2) ? 2 : 0;
    rc = list_plan_build(&pl, tasks, k, L.umap, L.amap, L.bmap, L.wmap, total_a, precision,
                         nullptr, s, sym_mode);
    if (rc == KS_OK) rc = list_plan_exec(*pl, L, s);
    if (pl) {
      (void)hipStreamSynchronize(s);
      list_plan_free(pl);
    }
  }
  if (rc == KS_OK && !on_device &&
      hipMemcpyAsync(u, L.u, (size_t)nu * 8, hipMemcpyDeviceToHost, s) != hipSuccess)
    rc = KS_ERR_DEVICE;
  release();
  if (hipStreamSynchronize(s) != hipSuccess && rc == KS_OK) rc = KS_ERR_DEVICE;
  if (rc != KS_OK) return raise_error(rc, rc == KS_ERR_DEVICE, "%s(): %s\n", name, err_string(rc));
  return KS_OK;
}

}  // namespace gsks

extern "C" int dgsks_list(int ntask, ks_t *kernels, int k, double *u, long nu,
                          const long *a_off, const int *umap, const int *amap,
                          const long *b_off, const int *bmap, const int *wmap,
                          const double *XA, const double *XA2, const double *XB,
                          const double *XB2, const double *w, int precision) {
  using namespace gsks;
  if (!XA2 || !XB2)
    return raise_error(KS_ERR_INVALID_ARG, false, "dgsks_list(): XA2 and XB2 are required\n");
  int rc = list_run(ntask, kernels, k, u, nu, a_off, umap, amap, b_off, bmap, wmap, XA, XA2, 0,
                    XB, XB2, 0, w, precision, "dgsks_list");
  if (rc == KS_OK)
    for (int t = 0; t < ntask; ++t)  // the Laplace descriptor side effect, per task
      if (ker_type(&kernels[t]) == KS_LAPLACE) laplace_params(k, &kernels[t].powe, &kernels[t].scal);
  return rc;
}

// ---- reusable batch plans (device-resident batches, replayable runs)

struct ks_list_plan_s {
  gsks::ListPlan *pl = nullptr;
  int k = 0;
};

extern "C" int ks_list_plan_create(ks_list_plan **plan, int ntask, const ks_t *kernels, int k,
                                   const long *a_off, const int *umap, const int *amap,
                                   const long *b_off, const int *bmap, const int *wmap,
                                   int precision) {
  using namespace gsks;
  const char *name = "ks_list_plan_create";
  if (!plan) return raise_error(KS_ERR_INVALID_ARG, false, "%s(): NULL plan\n", name);
  *plan = nullptr;
  if (ntask < 0 || k <= 0 || (ntask > 0 && (!kernels || !a_off || !b_off)))
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): invalid argument\n", name);
  if (precision != KS_PREC_FP64 && precision != KS_PREC_MIXED_FP32 &&
      precision != KS_PREC_MIXED_BF16X3)
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): unknown precision\n", name);
  for (int t = 0; t < ntask; ++t)
    if (a_off[t + 1] < a_off[t] || b_off[t + 1] < b_off[t] || a_off[t] < 0 || b_off[t] < 0)
      return raise_error(KS_ERR_LIST, false, "%s(): offsets must be non-decreasing\n", name);
  if (ntask > 0 && a_off[0] != 0)
    return raise_error(KS_ERR_LIST, false, "%s(): a_off[0] must be 0\n", name);
  if (ensure_device() != KS_OK)
    return raise_error(KS_ERR_DEVICE, true, "%s(): no GPU available\n", name);
  const long total_a = ntask > 0 ? a_off[ntask] : 0;
  std::vector<TaskDesc> tasks(ntask);
  for (int t = 0; t < ntask; ++t) {
    KerParams P;
    int rc = make_params(&kernels[t], k, P, true);
    if (rc != KS_OK && rc != KS_ERR_LAPLACE_K) return rc;
    tasks[t].type = ker_type(&kernels[t]);
    tasks[t].P = P;
    tasks[t].a0 = a_off[t];
    tasks[t].m = (int)(a_off[t + 1] - a_off[t]);
    tasks[t].b0 = b_off[t];
    tasks[t].n = (int)(b_off[t + 1] - b_off[t]);
    tasks[t].hi = kernels[t].hi;
    tasks[t].hj = kernels[t].hj;
    if (ker_type(&kernels[t]) == KS_GAUSSIAN_VAR_BANDWIDTH &&
        !(is_device_ptr(kernels[t].hi) && is_device_ptr(kernels[t].hj)))
      return raise_error(KS_ERR_INVALID_ARG, false,
                         "%s(): the bandwidth vectors of a plan must be device arrays\n", name);
  }
  if (total_a > 0 && !(is_device_ptr(umap) && is_device_ptr(amap) && is_device_ptr(bmap) &&
                       is_device_ptr(wmap)))
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): the maps of a plan must be device arrays\n",
                       name);
  hipStream_t s = thread_stream();
  if (!s) return raise_error(KS_ERR_DEVICE, true, "%s(): stream creation failed\n", name);
  ks_list_plan *h = new ks_list_plan();
  h->k = k;
  int rc = total_a > 0 ? list_plan_build(&h->pl, tasks, k, umap, amap, bmap, wmap, total_a,
                                         precision, nullptr, s, 1)
                       : KS_OK;
  if (rc != KS_OK) {
    delete h;
    return raise_error(rc, rc == KS_ERR_DEVICE, "%s(): %s\n", name, err_string(rc));
  }
  *plan = h;
  return KS_OK;
}

extern "C" int ks_list_plan_execute(const ks_list_plan *plan, double *u, const double *XA,
                                    const double *XA2, const double *XB, const double *XB2,
                                    const double *w, void *stream) {
  using namespace gsks;
  const char *name = "ks_list_plan_execute";
  if (!plan) return raise_error(KS_ERR_INVALID_ARG, false, "%s(): NULL plan\n", name);
  if (!plan->pl) return KS_OK;
  if (!u || !XA || !XA2 || !XB || !XB2 || !w)
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): NULL array\n", name);
  if (!(device_ptr_cached(u) && device_ptr_cached(XA) && device_ptr_cached(XA2) &&
        device_ptr_cached(XB) && device_ptr_cached(XB2) && device_ptr_cached(w)))
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): arrays must be device arrays\n", name);
  ListArrays L;
  L.u = u;
  L.XA = XA;
  L.XA2 = XA2;
  L.XB = XB;
  L.XB2 = XB2;
  L.w = w;
  int rc = list_plan_exec(*plan->pl, L, (hipStream_t)stream);
  if (rc != KS_OK) return raise_error(rc, rc == KS_ERR_DEVICE, "%s(): %s\n", name, err_string(rc));
  return KS_OK;
}

extern "C" int ks_list_plan_symmetric_tasks(const ks_list_plan *plan) {
  return plan ? gsks::list_plan_symmetric_tasks(plan->pl) : 0;
}

extern "C" void ks_list_plan_destroy(ks_list_plan *plan) {
  if (!plan) return;
  gsks::list_plan_free(plan->pl);
  delete plan;
}
