// Host-pointer staging for the drop-in entry points dgsks()/dgsks_ref().
//
// The reference API passes no array extents, so the extents of XA/XB/w/hi/hj
// are taken from the index maps (max index + 1).  u is gathered through umap
// into a packed vector, updated on the device, and scattered back in row order,
// which reproduces the reference's gather/scatter (frame/dgsks_ref.c:192-205,
// 454-459): a duplicate umap entry keeps the value of its last row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "host_util.h"

namespace gsks {

bool is_device_ptr(const void *p) {
  if (!p) return false;
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

static int max_index(const int *map, int len) {
  int mx = -1;
  for (int i = 0; i < len; ++i) mx = std::max(mx, map[i]);
  return mx;
}

static bool any_negative(const int *map, int len) {
  for (int i = 0; i < len; ++i)
    if (map[i] < 0) return true;
  return false;
}

namespace {
struct DevArena {
  char *base = nullptr;
  size_t off = 0, cap = 0;
  void *take(size_t bytes) {
    size_t a = (off + 255) & ~(size_t)255;
    off = a + bytes;
    return base + a;
  }
};
}  // namespace

static int finish(int rc, const char *what) {
  if (rc == KS_OK) return KS_OK;
  return raise_error(rc, rc == KS_ERR_DEVICE || rc == KS_ERR_ALLOC, "%s(): %s\n", what,
                     err_string(rc));
}

void run_staged(const ks_t *kernel, int m, int n, int k, double *u, const int *umap,
                const double *XA, const double *XA2, const int *amap, const double *XB,
                const double *XB2, const int *bmap, const double *w, const int *wmap,
                bool reference_algorithm) {
  const char *name = reference_algorithm ? "dgsks_ref" : "dgsks";
  if (ensure_device() != KS_OK) {
    finish(KS_ERR_DEVICE, name);
    return;
  }
  if (!u || !umap || !XA || !XA2 || !amap || !XB || !XB2 || !bmap || !w || !wmap) {
    raise_error(KS_ERR_INVALID_ARG, false, "%s(): NULL array\n", name);
    return;
  }
  Problem p;
  int rc = make_params(kernel, k, p.P, false);
  if (rc != KS_OK && rc != KS_ERR_LAPLACE_K) return;
  p.m = m;
  p.n = n;
  p.k = k;
  p.type = ker_type(kernel);
  const bool varbw = ker_type(kernel) == KS_GAUSSIAN_VAR_BANDWIDTH;
  hipStream_t s = thread_stream();
  if (!s) {
    finish(KS_ERR_DEVICE, name);
    return;
  }

  const void *ptrs[] = {u, umap, XA, XA2, amap, XB, XB2, bmap, w, wmap,
                        varbw ? kernel->hi : nullptr, varbw ? kernel->hj : nullptr};
  int ndev = 0, nptr = 0;
  for (const void *q : ptrs) {
    if (!q) continue;
    ++nptr;
    ndev += is_device_ptr(q) ? 1 : 0;
  }
  if (ndev == nptr) {  // all device: in place, no copies
    p.u = u;
    p.umap = umap;
    p.XA = XA;
    p.XA2 = XA2;
    p.amap = amap;
    p.XB = XB;
    p.XB2 = XB2;
    p.bmap = bmap;
    p.w = w;
    p.wmap = wmap;
    p.hi = kernel->hi;
    p.hj = kernel->hj;
    rc = reference_algorithm ? enqueue_reference(p, s) : enqueue_fused(p, s);
    if (rc == KS_OK) {
      hipError_t e = hipStreamSynchronize(s);
      if (e != hipSuccess) rc = hip_failure(e, "hipStreamSynchronize");
    }
    finish(rc, name);
    return;
  }
  if (ndev != 0) {
    finish(KS_ERR_MIXED_POINTERS, name);
    return;
  }
  if (any_negative(amap, m) || any_negative(umap, m) || any_negative(bmap, n) ||
      any_negative(wmap, n)) {
    raise_error(KS_ERR_INVALID_ARG, false, "%s(): negative index in a map\n", name);
    return;
  }

  // ---- all host: stage through HBM
  const size_t nxa = (size_t)max_index(amap, m) + 1;
  const size_t nxb = (size_t)max_index(bmap, n) + 1;
  const size_t nw = (size_t)max_index(wmap, n) + 1;
  const bool same_x = (XA == XB) && (XA2 == XB2);
  const bool same_h = varbw && kernel->hi == kernel->hj;
  const size_t nx_shared = std::max(nxa, nxb);

  std::vector<double> upack(m);
  for (int i = 0; i < m; ++i) upack[i] = u[umap[i]];
  std::vector<int> ident(m);
  for (int i = 0; i < m; ++i) ident[i] = i;

  size_t bytes = 0;
  auto need = [&](size_t b) { bytes += ((b + 255) & ~(size_t)255) + 256; };
  if (same_x) {
    need(nx_shared * k * 8);
    need(nx_shared * 8);
  } else {
    need(nxa * k * 8);
    need(nxa * 8);
    need(nxb * k * 8);
    need(nxb * 8);
  }
  need(nw * 8);
  need((size_t)m * 4 * 2);
  need((size_t)n * 4 * 2);
  need((size_t)m * 8);
  if (varbw) {
    need((same_h ? nx_shared : nxa) * 8);
    if (!same_h) need(nxb * 8);
  }
  DevArena ar;
  if (hipMallocAsync((void **)&ar.base, bytes, s) != hipSuccess) {
    (void)hipGetLastError();
    finish(KS_ERR_ALLOC, name);
    return;
  }
  ar.cap = bytes;
  auto up = [&](const void *src, size_t nbytes) -> void * {
    void *d = ar.take(nbytes);
    if (nbytes) {
      hipError_t e = hipMemcpyAsync(d, src, nbytes, hipMemcpyHostToDevice, s);
      if (e != hipSuccess) rc = hip_failure(e, "hipMemcpyAsync H2D");
    }
    return d;
  };
  rc = KS_OK;
  if (same_x) {
    size_t rows = nx_shared;
    p.XA = p.XB = (const double *)up(XA, rows * k * 8);
    p.XA2 = p.XB2 = (const double *)up(XA2, rows * 8);
  } else {
    p.XA = (const double *)up(XA, nxa * k * 8);
    p.XA2 = (const double *)up(XA2, nxa * 8);
    p.XB = (const double *)up(XB, nxb * k * 8);
    p.XB2 = (const double *)up(XB2, nxb * 8);
  }
  p.w = (const double *)up(w, nw * 8);
  p.amap = (const int *)up(amap, (size_t)m * 4);
  p.umap = (const int *)up(ident.data(), (size_t)m * 4);
  // one copy of a map that serves as both amap and bmap keeps the sum
  // recognisably symmetric on the device
  p.bmap = (bmap == amap && n == m) ? p.amap : (const int *)up(bmap, (size_t)n * 4);
  p.wmap = (const int *)up(wmap, (size_t)n * 4);
  double *du = (double *)up(upack.data(), (size_t)m * 8);
  p.u = du;
  p.hi = p.hj = nullptr;
  if (varbw) {
    if (same_h) {
      p.hi = p.hj = (const double *)up(kernel->hi, nx_shared * 8);
    } else {
      p.hi = (const double *)up(kernel->hi, nxa * 8);
      p.hj = (const double *)up(kernel->hj, nxb * 8);
    }
  }
  if (rc == KS_OK) rc = reference_algorithm ? enqueue_reference(p, s) : enqueue_fused(p, s);
  if (rc == KS_OK) {
    hipError_t e = hipMemcpyAsync(upack.data(), du, (size_t)m * 8, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) rc = hip_failure(e, "hipMemcpyAsync D2H");
  }
  (void)hipFreeAsync(ar.base, s);
  {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess && rc == KS_OK) rc = hip_failure(e, "hipStreamSynchronize");
  }
  if (rc == KS_OK)
    for (int i = 0; i < m; ++i) u[umap[i]] = upack[i];
  finish(rc, name);
}

void run_staged_single(const ks_t *kernel, int m, int n, int k, float *u, const int *umap,
                       const float *XA, const float *XA2, const int *amap, const float *XB,
                       const float *XB2, const int *bmap, const float *w, const int *wmap) {
  const char *name = "sgsks";
  if (ensure_device() != KS_OK) {
    finish(KS_ERR_DEVICE, name);
    return;
  }
  if (!u || !umap || !XA || !amap || !XB || !bmap || !w || !wmap) {
    raise_error(KS_ERR_INVALID_ARG, false, "%s(): NULL array\n", name);
    return;
  }
  SingleProblem p;
  int rc = make_params(kernel, k, p.P, false);
  if (rc != KS_OK && rc != KS_ERR_LAPLACE_K) return;
  p.m = m;
  p.n = n;
  p.k = k;
  p.type = ker_type(kernel);
  const bool varbw = ker_type(kernel) == KS_GAUSSIAN_VAR_BANDWIDTH;
  const int precision = single_default_precision(p.type, m, n, k);
  hipStream_t s = thread_stream();
  if (!s) {
    finish(KS_ERR_DEVICE, name);
    return;
  }
  const void *ptrs[] = {u, umap, XA, XA2, amap, XB, XB2, bmap, w, wmap,
                        varbw ? kernel->hi : nullptr, varbw ? kernel->hj : nullptr};
  int ndev = 0, nptr = 0;
  for (const void *q : ptrs) {
    if (!q) continue;
    ++nptr;
    ndev += is_device_ptr(q) ? 1 : 0;
  }
  if (ndev == nptr) {  // all device: in place
    p.u = u;
    p.umap = umap;
    p.XA = XA;
    p.XA2 = XA2;
    p.amap = amap;
    p.XB = XB;
    p.XB2 = XB2;
    p.bmap = bmap;
    p.w = w;
    p.wmap = wmap;
    p.hi = varbw ? kernel->hi : nullptr;
    p.hj = varbw ? kernel->hj : nullptr;
    rc = enqueue_single(p, precision, s);
    if (rc == KS_OK) {
      hipError_t e = hipStreamSynchronize(s);
      if (e != hipSuccess) rc = hip_failure(e, "hipStreamSynchronize");
    }
    finish(rc, name);
    return;
  }
  if (ndev != 0) {
    finish(KS_ERR_MIXED_POINTERS, name);
    return;
  }
  if (varbw && (!kernel->hi || !kernel->hj)) {
    finish(KS_ERR_NO_BANDWIDTH, name);
    return;
  }
  if (any_negative(amap, m) || any_negative(umap, m) || any_negative(bmap, n) ||
      any_negative(wmap, n)) {
    raise_error(KS_ERR_INVALID_ARG, false, "%s(): negative index in a map\n", name);
    return;
  }

  // ---- all host: stage through HBM (u gathered through umap, scattered back in
  // row order afterwards, as for dgsks())
  const size_t nxa = (size_t)max_index(amap, m) + 1;
  const size_t nxb = (size_t)max_index(bmap, n) + 1;
  const size_t nw = (size_t)max_index(wmap, n) + 1;
  const bool same_x = (XA == XB) && (XA2 == XB2);
  const bool same_h = varbw && kernel->hi == kernel->hj;
  const size_t nx_shared = std::max(nxa, nxb);
  std::vector<float> upack(m);
  for (int i = 0; i < m; ++i) upack[i] = u[umap[i]];
  std::vector<int> ident(m);
  for (int i = 0; i < m; ++i) ident[i] = i;

  size_t bytes = 0;
  auto need = [&](size_t b) { bytes += ((b + 255) & ~(size_t)255) + 256; };
  if (same_x) {
    need(nx_shared * k * 4);
    if (XA2) need(nx_shared * 4);
  } else {
    need(nxa * k * 4);
    if (XA2) need(nxa * 4);
    need(nxb * k * 4);
    if (XB2) need(nxb * 4);
  }
  need(nw * 4);
  need((size_t)m * 4 * 3);
  need((size_t)n * 4 * 2);
  if (varbw) {
    need((same_h ? nx_shared : nxa) * 8);
    if (!same_h) need(nxb * 8);
  }
  DevArena ar;
  if (hipMallocAsync((void **)&ar.base, bytes, s) != hipSuccess) {
    (void)hipGetLastError();
    finish(KS_ERR_ALLOC, name);
    return;
  }
  ar.cap = bytes;
  rc = KS_OK;
  auto up = [&](const void *src, size_t nbytes) -> void * {
    if (!src) return nullptr;
    void *d = ar.take(nbytes);
    if (nbytes) {
      hipError_t e = hipMemcpyAsync(d, src, nbytes, hipMemcpyHostToDevice, s);
      if (e != hipSuccess) rc = hip_failure(e, "hipMemcpyAsync H2D");
    }
    return d;
  };
  if (same_x) {
    p.XA = p.XB = (const float *)up(XA, nx_shared * k * 4);
    p.XA2 = p.XB2 = (const float *)up(XA2, nx_shared * 4);
  } else {
    p.XA = (const float *)up(XA, nxa * k * 4);
    p.XA2 = (const float *)up(XA2, nxa * 4);
    p.XB = (const float *)up(XB, nxb * k * 4);
    p.XB2 = (const float *)up(XB2, nxb * 4);
  }
  p.w = (const float *)up(w, nw * 4);
  p.amap = (const int *)up(amap, (size_t)m * 4);
  p.umap = (const int *)up(ident.data(), (size_t)m * 4);
  p.bmap = (bmap == amap && n == m) ? p.amap : (const int *)up(bmap, (size_t)n * 4);
  p.wmap = (const int *)up(wmap, (size_t)n * 4);
  float *du = (float *)up(upack.data(), (size_t)m * 4);
  p.u = du;
  p.hi = p.hj = nullptr;
  if (varbw) {
    if (same_h) {
      p.hi = p.hj = (const double *)up(kernel->hi, nx_shared * 8);
    } else {
      p.hi = (const double *)up(kernel->hi, nxa * 8);
      p.hj = (const double *)up(kernel->hj, nxb * 8);
    }
  }
  if (rc == KS_OK) rc = enqueue_single(p, precision, s);
  if (rc == KS_OK) {
    hipError_t e = hipMemcpyAsync(upack.data(), du, (size_t)m * 4, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) rc = hip_failure(e, "hipMemcpyAsync D2H");
  }
  (void)hipFreeAsync(ar.base, s);
  {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess && rc == KS_OK) rc = hip_failure(e, "hipStreamSynchronize");
  }
  if (rc == KS_OK)
    for (int i = 0; i < m; ++i) u[umap[i]] = upack[i];
  finish(rc, name);
}

}  // namespace gsks
