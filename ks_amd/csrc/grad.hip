// Gradient of the mapped Gaussian and variable-bandwidth summations
// (ks_grad_device).  No counterpart in the reference, whose summations
// (frame/dgsks.c:367-776, frame/dgsks_ref.c:94-535) have no gradient; the
// kernel values follow frame/dgsks_ref.c:254-288:
//   Gaussian            K_ij = exp(scal d2_ij)
//   variable bandwidth  K_ij = exp(-0.5 hi[amap[i]] hj[bmap[j]] d2_ij)
// with d2_ij = |a_i - b_j|^2, a_i = XA[amap[i]], b_j = XB[bmap[j]].
//
// For L = sum_x g[x] du[x], du[umap[i]] += sum_j K_ij w[wmap[j]], write
// g_i = g[umap[i]], w_j = w[wmap[j]], and c_ij = scal or -0.5 h_i h_j.  Then
//   dL/da_i  = 2 g_i sum_j c_ij K_ij w_j (a_i - b_j)
//   dL/db_j  = 2 w_j sum_i c_ij K_ij g_i (b_j - a_i)
//   dL/dw_j  = sum_i g_i K_ij
//   dL/dscal = sum_ij g_i w_j K_ij d2_ij
//   dL/dh_i  = -0.5 g_i sum_j h_j w_j K_ij d2_ij,  dL/dh_j likewise.
// Each sum is a multi-right-hand-side summation whose right-hand sides are
// the weights times the points, their norms and 1, so the kernel matrix is
// recomputed twice and never stored:
//   pass 1 (targets a, sources b): [V, s, T]_i = sum_j K_ij c_j [b_j, 1, |b_j|^2],
//          c_j = w_j (h_j w_j for the variable bandwidth)
//   pass 2 (targets b, sources a): [Q, r, S, dw]_j = sum_i K_ij e_i [a_i, 1, |a_i|^2, g_i / e_i],
//          e_i = g_i (h_i g_i)
// (the Gaussian needs no norm columns: its dscal is assembled from u, V, r and
// the norms).  Both passes work on the mapped points gathered into compact
// tables and translated by the mean of both sets, which leaves every d2 and
// every gradient unchanged and keeps the differences a s - V small.  The
// per-row results are added into the caller's tables (repeated map entries
// summed in row order, after a stable sort of the map) by one thread per
// (entry, column), so the result is deterministic.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "gsks_kernels.h"
#include "gsks_launch.h"
#include "host_util.h"

namespace gsks {
namespace {

constexpr int GW = 256;

unsigned blocks(long n) { return (unsigned)((n + GW - 1) / GW); }

__global__ void kg_iota(int *v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

__global__ void kg_zero(double *v, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = 0.0;
}

// out[i] = X[map[i]] row by row (X [.][k]); also v[i] = src[vmap[i]] for up to
// two per-point vectors (NULL: skipped)
__global__ void kg_gather(const double *__restrict__ X, const int *__restrict__ map, int rows,
                          int k, double *__restrict__ out, const double *src1,
                          const int *vmap1, double *v1, const double *src2, const int *vmap2,
                          double *v2) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)rows * k) return;
  const long i = t / k;
  const int q = (int)(t - i * k);
  out[t] = X[(long)map[i] * k + q];
  if (q == 0) {
    if (v1) v1[i] = src1[vmap1[i]];
    if (v2) v2[i] = src2[vmap2[i]];
  }
}

// Right-hand sides of one pass over the points X (rows of k): with
// c = s[i] (times h[i] when h is given), W[i] = [c x_i, c, c |x_i|^2, s[i]],
// the last two columns only while R asks for them (R = k + 1 ... k + 3).
__global__ void kg_rhs(const double *__restrict__ X, const double *__restrict__ X2,
                       const double *__restrict__ s, const double *__restrict__ h, int rows,
                       int k, int R, double *__restrict__ W) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const double c = h ? s[i] * h[i] : s[i];
  for (int q = 0; q < k; ++q) W[i * R + q] = c * X[i * k + q];
  W[i * R + k] = c;
  if (R > k + 1) W[i * R + k + 1] = c * X2[i];
  if (R > k + 2) W[i * R + k + 2] = s[i];
}

// Target rows from U1 = [V, s, T]: GA = dL/da (rows of k), Gh = dL/dh_i
// (variable bandwidth) or terms = the rows' share of dL/dscal (Gaussian).
// Outputs may be NULL.
__global__ void kg_fin_a(const double *__restrict__ A, const double *__restrict__ A2,
                         const double *__restrict__ gv, const double *__restrict__ hA,
                         const double *__restrict__ U1, int m, int k, int R, double scal,
                         double *GA, double *Gh, double *terms) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double *u = U1 + i * R;
  const double s = u[k];
  const double f = hA ? -hA[i] * gv[i] : 2.0 * scal * gv[i];
  double av = 0.0;
  for (int q = 0; q < k; ++q) {
    const double a = A[i * k + q], V = u[q];
    if (GA) GA[i * k + q] = f * (a * s - V);
    av = fma(a, V, av);
  }
  const double quad = A2[i] * s - 2.0 * av;  // sum_j K c_j |a_i|^2 - 2 a_i . b_j
  if (hA) {
    if (Gh) Gh[i] = -0.5 * gv[i] * (quad + u[k + 1]);
  } else if (terms) {
    terms[i] = gv[i] * quad;
  }
}

// Source rows from U2 = [Q, r, S, dw]: GB = dL/db, Gw = dL/dw, Gh = dL/dh_j
// (variable bandwidth) or terms (Gaussian, the sources' share of dL/dscal).
__global__ void kg_fin_b(const double *__restrict__ B, const double *__restrict__ B2,
                         const double *__restrict__ wv, const double *__restrict__ hB,
                         const double *__restrict__ U2, int n, int k, int R, double scal,
                         double *GB, double *Gw, double *Gh, double *terms) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double *u = U2 + j * R;
  const double r = u[k];
  const double f = hB ? -hB[j] * wv[j] : 2.0 * scal * wv[j];
  double bq = 0.0;
  for (int q = 0; q < k; ++q) {
    const double b = B[j * k + q], Q = u[q];
    if (GB) GB[j * k + q] = f * (b * r - Q);
    bq = fma(b, Q, bq);
  }
  if (hB) {
    if (Gw) Gw[j] = u[k + 2];
    if (Gh) Gh[j] = -0.5 * wv[j] * (B2[j] * r - 2.0 * bq + u[k + 1]);
  } else {
    if (Gw) Gw[j] = r;
    if (terms) terms[j] = wv[j] * B2[j] * r;
  }
}

// *out += sum of t[0 .. n): one workgroup, four interleaved accumulators per
// lane, then a fixed tree, so the sum is the same on every call
constexpr int SW = 1024;
__global__ __launch_bounds__(SW) void kg_sum(const double *__restrict__ t, long n, double *out) {
  __shared__ double sh[SW];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  long i = threadIdx.x;
  for (; i + 3 * SW < n; i += 4 * SW)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] += t[i + q * SW];
  for (int q = 0; i < n; i += SW, ++q) acc[q & 3] += t[i];
  sh[threadIdx.x] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
  __syncthreads();
  for (int w = SW / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out += sh[0];
}

// table[key * width + q] += sum of src[perm[t] * width + q] over the run of
// equal keys that starts at sorted position t (keys sorted stably, so each run
// adds its rows in row order); one thread per (position, column), the run's
// first position doing the work.
__global__ void kg_scatter_add(const unsigned *__restrict__ key, const int *__restrict__ perm,
                               int rows, int width, const double *__restrict__ src,
                               double *table) {
  const long t0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t0 >= (long)rows * width) return;
  const int t = (int)(t0 / width);
  const int q = (int)(t0 - (long)t * width);
  const unsigned kk = key[t];
  if (t > 0 && key[t - 1] == kk) return;
  double acc = 0.0;
  for (int u = t; u < rows && key[u] == kk; ++u) acc += src[(long)perm[u] * width + q];
  table[(long)kk * width + q] += acc;
}

struct GradWork {
  int *ident = nullptr, *perm = nullptr;
  unsigned *skey = nullptr;
  void *sort_tmp = nullptr;
  size_t sort_bytes = 0;
  double *A = nullptr, *B = nullptr, *A2 = nullptr, *B2 = nullptr, *centre = nullptr,
         *cwork = nullptr, *zc = nullptr;
  double *gv = nullptr, *wv = nullptr, *hA = nullptr, *hB = nullptr;
  double *W1 = nullptr, *U1 = nullptr, *W2 = nullptr, *U2 = nullptr;
  double *GA = nullptr, *GhA = nullptr, *GB = nullptr, *GhB = nullptr, *Gw = nullptr,
         *terms = nullptr;
};

struct GradCall {
  int m, n, k, type;
  double scal;
  bool var, mixed, pass1, pass2;
  int R1, R2;
};

// One pass: U = K [targets T, sources S] W with R right-hand sides, identity maps
// into the compact tables.  hT / hS are the targets' / sources' bandwidths.
int kg_pass(const GradCall &c, const GradWork &g, const double *T, const double *T2, int mt,
            const double *hT, const double *S, const double *S2, int ns, const double *hS,
            const double *W, double *U, int R, hipStream_t s) {
  hipLaunchKernelGGL(kg_zero, dim3(blocks((long)mt * R)), dim3(GW), 0, s, U, (long)mt * R);
  Problem p;
  p.m = mt; p.n = ns; p.k = c.k; p.type = c.type;
  p.P = KerParams{c.scal, 0.0, 0.0, 0, POW_GENERAL};
  p.u = U; p.umap = g.ident; p.XA = T; p.XA2 = T2; p.amap = g.ident;
  p.XB = S; p.XB2 = S2; p.bmap = g.ident; p.w = W; p.wmap = g.ident;
  p.hi = hT; p.hj = hS;
  return c.mixed ? enqueue_multirhs_mixed(p, R, g.zc, s) : enqueue_multirhs(p, R, s);
}

// Sorts map's entries (stably: equal entries keep row order) into g.skey with
// their rows in g.perm, for the scatters that follow.
int kg_sort(const GradWork &g, const int *map, int rows, hipStream_t s) {
  size_t bytes = g.sort_bytes;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(
      g.sort_tmp, bytes, reinterpret_cast<const unsigned *>(map), g.skey, g.ident, g.perm, rows,
      0, 31, s);
  return e == hipSuccess ? KS_OK : hip_failure(e, "gradient map sort");
}

// Adds rows [rows][width] of src into table at the entries of the map last sorted.
void kg_scatter(const GradWork &g, int rows, int width, const double *src, double *table,
                hipStream_t s) {
  if (table)
    hipLaunchKernelGGL(kg_scatter_add, dim3(blocks((long)rows * width)), dim3(GW), 0, s, g.skey,
                       g.perm, rows, width, src, table);
}

}  // namespace
}  // namespace gsks

using namespace gsks;

extern "C" int ks_grad_device(const ks_t *kernel, int m, int n, int k, const double *g,
                              const int *umap, const double *XA, const int *amap,
                              const double *XB, const int *bmap, const double *w,
                              const int *wmap, double *dXA, double *dXB, double *dw,
                              double *dscal, double *dhi, double *dhj, int precision,
                              void *stream) {
  const char *name = "ks_grad_device";
  if (!kernel || m < 0 || n < 0 || k < 0)
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): invalid argument\n", name);
  GradCall c;
  c.m = m; c.n = n; c.k = k;
  c.type = ker_type(kernel);
  c.var = c.type == KS_GAUSSIAN_VAR_BANDWIDTH;
  if (c.type != KS_GAUSSIAN && !c.var)
    return raise_error(KS_ERR_INVALID_ARG, false,
                       "%s(): the Gaussian and variable-bandwidth kernels only\n", name);
  if ((c.var && dscal) || (!c.var && (dhi || dhj)))
    return raise_error(KS_ERR_INVALID_ARG, false,
                       "%s(): dscal is the Gaussian's output, dhi / dhj the variable "
                       "bandwidth's\n", name);
  c.mixed = precision == KS_PREC_MIXED_FP32;
  if (!c.mixed && precision != KS_PREC_FP64)
    return raise_error(KS_ERR_INVALID_ARG, false,
                       "%s(): precision must be KS_PREC_FP64 or KS_PREC_MIXED_FP32\n", name);
  // right-hand sides per pass
  c.R1 = c.var ? k + 2 : k + 1;
  c.R2 = c.var ? k + 3 : k + 1;
  if (c.mixed && (c.var || c.R2 > kMixedMrhsMaxRhs || k > kMixedMrhsMaxK))
    return raise_error(KS_ERR_INVALID_ARG, false,
                       "%s(): the mixed gradient takes the Gaussian kernel and k <= %d\n", name,
                       kMixedMrhsMaxRhs - 1);
  if (c.R2 > 72)
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): k <= %d\n", name, c.var ? 69 : 71);
  if (c.var && (!kernel->hi || !kernel->hj))
    return raise_error(KS_ERR_NO_BANDWIDTH, false, "%s(): hi and hj are required\n", name);
  if (m == 0 || n == 0 || k == 0) return KS_OK;
  if (!g || !umap || !XA || !amap || !XB || !bmap || !w || !wmap)
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): invalid argument\n", name);
  if (ensure_device() != KS_OK)
    return raise_error(KS_ERR_DEVICE, false, "%s(): no GPU available\n", name);
  const void *arrays[] = {g,   umap, XA,  amap,  XB,  bmap, w,   wmap,
                          dXA, dXB,  dw,  dscal, dhi, dhj,  c.var ? kernel->hi : nullptr,
                          c.var ? kernel->hj : nullptr};
  for (const void *a : arrays)
    if (a && !device_ptr_cached(a))
      return raise_error(KS_ERR_MIXED_POINTERS, false, "%s(): every array must be device memory\n",
                         name);
  if (!dXA && !dXB && !dw && !dscal && !dhi && !dhj) return KS_OK;
  c.scal = c.var ? 0.0 : kernel->scal;
  c.pass1 = dXA || dhi || dscal;
  c.pass2 = dXB || dw || dhj || dscal;
  hipStream_t s = (hipStream_t)stream;
  const int mx = m > n ? m : n;

  GradWork W;
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, W.sort_bytes, (const unsigned *)nullptr,
                                         (unsigned *)nullptr, (const int *)nullptr,
                                         (int *)nullptr, mx, 0, 31, s) != hipSuccess)
    return raise_error(KS_ERR_DEVICE, false, "%s(): sort set-up failed\n", name);
  char *buf = nullptr;
  int rc = carve_alloc([&](Carve &cv) {
    W.ident = cv.take<int>(mx);
    W.perm = cv.take<int>(mx);
    W.skey = cv.take<unsigned>(mx);
    W.sort_tmp = cv.take<char>(W.sort_bytes);
    W.A = cv.take<double>((size_t)m * k);
    W.B = cv.take<double>((size_t)n * k);
    W.A2 = cv.take<double>(m);
    W.B2 = cv.take<double>(n);
    W.centre = cv.take<double>(k);
    W.cwork = cv.take<double>(centre_work_elems((long)m + n));
    W.zc = cv.take<double>(k);
    W.gv = cv.take<double>(m);
    W.wv = cv.take<double>(n);
    if (c.var) {
      W.hA = cv.take<double>(m);
      W.hB = cv.take<double>(n);
    }
    if (c.pass1) {
      W.W1 = cv.take<double>((size_t)n * c.R1);
      W.U1 = cv.take<double>((size_t)m * c.R1);
      W.GA = cv.take<double>((size_t)m * k);
      W.GhA = cv.take<double>(m);
    }
    if (c.pass2) {
      W.W2 = cv.take<double>((size_t)m * c.R2);
      W.U2 = cv.take<double>((size_t)n * c.R2);
      W.GB = cv.take<double>((size_t)n * k);
      W.GhB = cv.take<double>(n);
      W.Gw = cv.take<double>(n);
    }
    W.terms = cv.take<double>((size_t)m + n);
  }, &buf, s);
  if (rc != KS_OK) return raise_error(rc, false, "%s(): allocation failed\n", name);

  auto chain = [&](int r) {
    if (rc == KS_OK) rc = r;
  };
  hipLaunchKernelGGL(kg_iota, dim3(blocks(mx)), dim3(GW), 0, s, W.ident, mx);
  hipLaunchKernelGGL(kg_zero, dim3(1), dim3(GW), 0, s, W.zc, (long)k);
  // compact tables: points, g[umap], w[wmap], bandwidths
  hipLaunchKernelGGL(kg_gather, dim3(blocks((long)m * k)), dim3(GW), 0, s, XA, amap, m, k, W.A,
                     g, umap, W.gv, c.var ? kernel->hi : nullptr, amap, W.hA);
  hipLaunchKernelGGL(kg_gather, dim3(blocks((long)n * k)), dim3(GW), 0, s, XB, bmap, n, k, W.B,
                     w, wmap, W.wv, c.var ? kernel->hj : nullptr, bmap, W.hB);
  // translate both sets by their mean, in place; norms of the translated points
  chain(enqueue_centre(W.A, m, W.B, n, k, W.ident, W.cwork, W.centre, s));
  chain(enqueue_shift(W.A, m, k, W.centre, W.A, s));
  chain(enqueue_shift(W.B, n, k, W.centre, W.B, s));
  chain(ks_sqnorm_device(W.A, m, k, W.A2, s));
  chain(ks_sqnorm_device(W.B, n, k, W.B2, s));
  if (rc == KS_OK && c.pass1) {
    hipLaunchKernelGGL(kg_rhs, dim3(blocks(n)), dim3(GW), 0, s, W.B, W.B2, W.wv, W.hB, n, k,
                       c.R1, W.W1);
    chain(kg_pass(c, W, W.A, W.A2, m, W.hA, W.B, W.B2, n, W.hB, W.W1, W.U1, c.R1, s));
    if (rc == KS_OK)
      hipLaunchKernelGGL(kg_fin_a, dim3(blocks(m)), dim3(GW), 0, s, W.A, W.A2, W.gv, W.hA, W.U1,
                         m, k, c.R1, c.scal, W.GA, W.GhA, W.terms);
  }
  if (rc == KS_OK && c.pass2) {
    hipLaunchKernelGGL(kg_rhs, dim3(blocks(m)), dim3(GW), 0, s, W.A, W.A2, W.gv, W.hA, m, k,
                       c.R2, W.W2);
    chain(kg_pass(c, W, W.B, W.B2, n, W.hB, W.A, W.A2, m, W.hA, W.W2, W.U2, c.R2, s));
    if (rc == KS_OK)
      hipLaunchKernelGGL(kg_fin_b, dim3(blocks(n)), dim3(GW), 0, s, W.B, W.B2, W.wv, W.hB, W.U2,
                         n, k, c.R2, c.scal, W.GB, W.Gw, W.GhB, W.terms + m);
  }
  if (rc == KS_OK && dscal)
    hipLaunchKernelGGL(kg_sum, dim3(1), dim3(SW), 0, s, W.terms, (long)m + n, dscal);
  // add the rows into the caller's tables
  if (rc == KS_OK && (dXA || dhi)) {
    chain(kg_sort(W, amap, m, s));
    kg_scatter(W, m, k, W.GA, dXA, s);
    kg_scatter(W, m, 1, W.GhA, dhi, s);
  }
  if (rc == KS_OK && (dXB || dhj || (dw && wmap == bmap))) {
    chain(kg_sort(W, bmap, n, s));
    kg_scatter(W, n, k, W.GB, dXB, s);
    kg_scatter(W, n, 1, W.GhB, dhj, s);
    if (wmap == bmap) kg_scatter(W, n, 1, W.Gw, dw, s);
  }
  if (rc == KS_OK && dw && wmap != bmap) {
    chain(kg_sort(W, wmap, n, s));
    kg_scatter(W, n, 1, W.Gw, dw, s);
  }
  if (rc == KS_OK && hipGetLastError() != hipSuccess) rc = KS_ERR_DEVICE;
  (void)hipFreeAsync(buf, s);
  if (rc != KS_OK) return raise_error(rc, false, "%s(): %s\n", name, err_string(rc));
  return KS_OK;
}
