// Single-precision front end (sgsks / sgsks_device): float points, norms,
// weights and potentials in and out, the summation itself run by the FP64 or
// mixed-precision kernels.  The reference declares a single-precision
// micro-kernel for every kernel type and leaves each unimplemented
// (micro_kernel/x86_64/haswell/gaussian_int_d8x6.c:38-54 and its siblings only
// print "not yet implemented"); the arguments here are dgsks()'s
// (frame/dgsks.c:373-390) with float arrays.
//
// Work on the stream, in order:
//  1. gather: the m target rows and n source rows that amap / bmap select are
//     widened into compact FP64 tables (exact: every float is a double), with
//     their squared norms (given norms widened, NULL norms computed in FP64
//     from the float coordinates, left to right), the weights w[wmap[j]] and,
//     for the variable-bandwidth kernel, hi[amap[i]] / hj[bmap[j]]; an FP64 row
//     vector is zeroed and an identity map written;
//  2. the summation over the compact tables with identity maps into that row
//     vector: the FP64 kernels (the symmetric one when targets are the
//     sources) or the mixed-precision kernels;
//  3. u[umap[i]] = float(double(u[umap[i]]) + row sum i): one rounding per
//     target.
// Coordinates are widened by a flattened (row, coordinate) launch, so loads of
// one row and stores of the compact table are contiguous across lanes.
#include <hip/hip_runtime.h>

#include "gsks_launch.h"
#include "host_util.h"

#pragma clang fp contract(off)

namespace gsks {

namespace {

// One side (targets or sources) of the gather.
struct Side {
  const float *X, *X2;  // X2 may be NULL: norms computed from X
  const int *map;
  const double *h;      // variable bandwidths (NULL otherwise)
  double *Xc, *X2c, *hc;
  int rows;
};

__global__ void single_widen_kernel(const float *__restrict__ X, const int *__restrict__ map,
                                    long total, int k, double *__restrict__ Xc) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const long i = t / k;
  const int p = (int)(t - i * k);
  Xc[t] = (double)X[(long)map[i] * k + p];
}

// Per-row part: norms, bandwidths, weights, the zeroed row sums and the identity map.
__global__ void single_rows_kernel(Side a, Side b, int k, const float *__restrict__ w,
                                   const int *__restrict__ wmap, double *__restrict__ wc,
                                   double *__restrict__ uc, int *__restrict__ ident, int nid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nid) ident[i] = i;
  for (int side = 0; side < 2; ++side) {
    const Side &S = side ? b : a;
    if (!S.X2c || i >= S.rows) continue;  // b.X2c NULL: sources share the target tables
    const int src = S.map[i];
    double t;
    if (S.X2) {
      t = (double)S.X2[src];
    } else {
      const double *x = S.Xc + (long)i * k;
      t = 0.0;
      for (int p = 0; p < k; ++p) t += x[p] * x[p];
    }
    S.X2c[i] = t;
    if (S.h) S.hc[i] = S.h[src];
  }
  if (i < b.rows) wc[i] = (double)w[wmap[i]];
  if (i < a.rows) uc[i] = 0.0;
}

__global__ void single_scatter_kernel(float *__restrict__ u, const int *__restrict__ umap,
                                      const double *__restrict__ uc, int m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int t = umap[i];
  u[t] = (float)((double)u[t] + uc[i]);
}

inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

int single_default_precision(int type, int m, int n, int k) {
  // sgsks() sums in FP64: the mixed kernels are accurate to 1e-6 only where near
  // pairs carry the sums (a row made only of far pairs is accurate to about
  // 1e-5, and the skip tier drops pairs with K < 2^-50 that a float could
  // still hold), which depends on the data.  Callers that know their data ask
  // for mixed precision through sgsks_device.
  (void)type;
  (void)m;
  (void)n;
  (void)k;
  return KS_PREC_FP64;
}

int enqueue_single(const SingleProblem &sp, int precision, hipStream_t s) {
  const int m = sp.m, n = sp.n, k = sp.k;
  const bool varbw = sp.type == VAR_BANDWIDTH;
  // targets are the sources: one compact table serves both sides, so the
  // summation below sees one map and one table (the symmetric FP64 path)
  const bool shared = m == n && sp.amap == sp.bmap && sp.XA == sp.XB && sp.XA2 == sp.XB2 &&
                      (!varbw || sp.hi == sp.hj);
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t off = 0;
  const size_t o_XA = off; off += al((size_t)m * k * 8);
  const size_t o_XA2 = off; off += al((size_t)m * 8);
  const size_t o_hi = off; off += varbw ? al((size_t)m * 8) : 0;
  const size_t o_XB = off; off += shared ? 0 : al((size_t)n * k * 8);
  const size_t o_XB2 = off; off += shared ? 0 : al((size_t)n * 8);
  const size_t o_hj = off; off += (varbw && !shared) ? al((size_t)n * 8) : 0;
  const size_t o_w = off; off += al((size_t)n * 8);
  const size_t o_u = off; off += al((size_t)m * 8);
  const int nid = m > n ? m : n;
  const size_t o_id = off; off += al((size_t)nid * 4);
  char *base = nullptr;
  ScratchLease lease(s);
  int rc = lease.aux(off, &base);
  if (rc != KS_OK) return rc;

  Side A{sp.XA, sp.XA2, sp.amap, varbw ? sp.hi : nullptr, (double *)(base + o_XA),
         (double *)(base + o_XA2), varbw ? (double *)(base + o_hi) : nullptr, m};
  Side B{sp.XB, sp.XB2, sp.bmap, varbw ? sp.hj : nullptr, nullptr, nullptr, nullptr, n};
  if (shared) {
    B.Xc = A.Xc;
    B.hc = A.hc;
  } else {
    B.Xc = (double *)(base + o_XB);
    B.X2c = (double *)(base + o_XB2);
    B.hc = varbw ? (double *)(base + o_hj) : nullptr;
  }
  double *wc = (double *)(base + o_w), *uc = (double *)(base + o_u);
  int *ident = (int *)(base + o_id);

  const int bs = 256;
  hipLaunchKernelGGL(single_widen_kernel, dim3(blocks_for((long)m * k, bs)), dim3(bs), 0, s,
                     sp.XA, sp.amap, (long)m * k, k, A.Xc);
  if (!shared)
    hipLaunchKernelGGL(single_widen_kernel, dim3(blocks_for((long)n * k, bs)), dim3(bs), 0, s,
                       sp.XB, sp.bmap, (long)n * k, k, B.Xc);
  hipLaunchKernelGGL(single_rows_kernel, dim3(blocks_for(nid, bs)), dim3(bs), 0, s, A, B, k, sp.w,
                     sp.wmap, wc, uc, ident, nid);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_failure(e, "sgsks gather launch");

  Problem q;
  q.m = m;
  q.n = n;
  q.k = k;
  q.type = sp.type;
  q.P = sp.P;
  q.u = uc;
  q.umap = ident;
  q.XA = A.Xc;
  q.XA2 = A.X2c;
  q.amap = ident;
  q.XB = B.Xc;
  q.XB2 = shared ? A.X2c : B.X2c;
  q.bmap = ident;
  q.w = wc;
  q.wmap = ident;
  q.hi = A.hc;
  q.hj = B.hc;
  rc = precision == KS_PREC_FP64 ? enqueue_fused(q, s) : enqueue_mixed(q, precision, s);
  if (rc != KS_OK) return rc;
  hipLaunchKernelGGL(single_scatter_kernel, dim3(blocks_for(m, bs)), dim3(bs), 0, s, sp.u,
                     sp.umap, uc, m);
  e = hipGetLastError();
  return e == hipSuccess ? KS_OK : hip_failure(e, "sgsks scatter launch");
}

}  // namespace gsks
