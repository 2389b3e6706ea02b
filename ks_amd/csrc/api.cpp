// C-ABI of libgsks.so: the reference's dgsks()/dgsks_ref()/wrappers/ks_malloc_aligned
// (include/ks.h:73-111, frame/dgsks.c:783-826, frame/dgsks_ref.c:537-580) plus the
// stream-ordered device entry points.  Host-side only: validation, the Laplace
// parameter side effect, staging of host arrays through HBM, scratch management
// and kernel launches.  Every summation runs on the GPU; there is no CPU path.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <optional>
#include <utility>
#include <vector>

#include "../../include/ks.h"
#include "gsks_launch.h"
#include "host_util.h"

using namespace gsks;

namespace gsks {

// ------------------------------------------------------------------ errors

static std::atomic<int> g_errmode{KS_ERRMODE_EXIT};
static thread_local int t_last_error = KS_OK;
static thread_local char t_detail[256] = "";

// Remember which HIP call failed and why (read back by ks_error_detail()).
int hip_failure(hipError_t e, const char *where) {
  snprintf(t_detail, sizeof(t_detail), "%s: %s", where, hipGetErrorString(e));
  return KS_ERR_DEVICE;
}

const char *err_string(int code) {
  switch (code) {
    case KS_OK: return "ok";
    case KS_ERR_INVALID_TYPE: return "illegal kernel type";
    case KS_ERR_INVALID_ARG: return "invalid argument";
    case KS_ERR_NO_BANDWIDTH: return "bandwidth vector has not been initialized";
    case KS_ERR_LAPLACE_K: return "laplace kernel only supports k > 2";
    case KS_ERR_DEVICE: return "GPU unavailable or HIP call failed";
    case KS_ERR_ALLOC: return "allocation failed";
    case KS_ERR_MIXED_POINTERS: return "inconsistent host/device pointers";
    case KS_ERR_LIST: return "inconsistent list description";
    default: return "unknown error";
  }
}

// Record an error.  In KS_ERRMODE_EXIT a fatal error prints the reference's style
// of message and exits(1) (frame/dgsks.c:465-467); warnings only print.
int raise_error(int code, bool fatal, const char *fmt, ...) {
  t_last_error = code;
  if (g_errmode.load() == KS_ERRMODE_EXIT) {
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
    fflush(stdout);
    if (fatal) exit(1);
  }
  return code;
}

void set_last_error(int code) { t_last_error = code; }

// ------------------------------------------------------------------ device context

struct Scratch {
  std::recursive_mutex mu;  // held by a ScratchLease
  double *partial = nullptr;
  size_t partial_elems = 0;
  char *aux = nullptr;
  size_t aux_bytes = 0;
  bool captured = false;        // a stream capture has handed out these buffers
  std::vector<void *> retired;  // replaced buffers a captured graph may still use
};
static std::mutex g_map_mu;  // guards the map only (its nodes never move)
static std::map<std::pair<int, hipStream_t>, Scratch> g_scratch;

int ensure_device() {
  static int ndev = -1;
  static std::once_flag once;
  std::call_once(once, [] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    (void)hipGetLastError();
    ndev = n;
  });
  return ndev > 0 ? KS_OK : KS_ERR_DEVICE;
}

hipStream_t thread_stream() {
  static thread_local hipStream_t s = nullptr;
  static thread_local int dev = -1;
  int d = 0;
  (void)hipGetDevice(&d);
  if (!s || dev != d) {
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) {
      hip_failure(e, "hipStreamCreateWithFlags");
      s = nullptr;
    }
    dev = d;
  }
  return s;
}

ScratchLease::ScratchLease(hipStream_t s) : s_(s) {
  int d = 0;
  (void)hipGetDevice(&d);
  {
    std::lock_guard<std::mutex> lk(g_map_mu);
    sc_ = &g_scratch[{d, s}];
  }
  sc_->mu.lock();
}

ScratchLease::~ScratchLease() { sc_->mu.unlock(); }

// Whether stream s is being captured into a graph.
static bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return st != hipStreamCaptureStatusNone;
}

// Before buffers are handed out: refuse to grow under capture, and remember
// that a capture has seen this scratch.
static int scratch_prepare(Scratch &sc, hipStream_t s, bool grow, const char *what) {
  if (!stream_capturing(s)) return KS_OK;
  if (grow) {
    snprintf(t_detail, sizeof(t_detail),
             "%s: scratch must grow while the stream is being captured; run the call once on "
             "this stream before capturing it", what);
    return KS_ERR_ALLOC;
  }
  sc.captured = true;
  return KS_OK;
}

// Drop a buffer that is being replaced: freed after the stream drains, or kept
// for the life of the process when a captured graph may still use it.
static int scratch_drop(Scratch &sc, hipStream_t s, void *p) {
  if (!p) return KS_OK;
  if (sc.captured) {
    sc.retired.push_back(p);
    return KS_OK;
  }
  hipError_t se = hipStreamSynchronize(s);
  if (se != hipSuccess) return hip_failure(se, "scratch: hipStreamSynchronize");
  (void)hipFree(p);
  return KS_OK;
}

int ScratchLease::partials(size_t partial_elems, double **partial) {
  Scratch &sc = *sc_;
  const bool grow_p = sc.partial_elems < partial_elems;
  int rc = scratch_prepare(sc, s_, grow_p, "scratch partials");
  if (rc != KS_OK) return rc;
  if (grow_p) {
    rc = scratch_drop(sc, s_, sc.partial);
    if (rc != KS_OK) return rc;
    sc.partial = nullptr;
    sc.partial_elems = 0;
    size_t want = partial_elems + partial_elems / 4;
    if (hipMalloc(&sc.partial, want * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      sc.partial = nullptr;
      return KS_ERR_ALLOC;
    }
    sc.partial_elems = want;
  }
  *partial = sc.partial;
  return KS_OK;
}

int ScratchLease::aux(size_t bytes, char **base) {
  Scratch &sc = *sc_;
  const bool grow = sc.aux_bytes < bytes;
  int rc = scratch_prepare(sc, s_, grow, "scratch arena");
  if (rc != KS_OK) return rc;
  if (grow) {
    rc = scratch_drop(sc, s_, sc.aux);
    if (rc != KS_OK) return rc;
    sc.aux = nullptr;
    sc.aux_bytes = 0;
    size_t want = bytes + bytes / 4 + 256;
    if (hipMalloc(&sc.aux, want) != hipSuccess) {
      (void)hipGetLastError();
      sc.aux = nullptr;
      return KS_ERR_ALLOC;
    }
    sc.aux_bytes = want;
  }
  *base = sc.aux;
  return KS_OK;
}

// Device-pointer check with a small per-thread cache of recently verified buffers.
bool device_ptr_cached(const void *p) {
  static thread_local const void *cache[64];
  static thread_local int next = 0;
  for (const void *c : cache)
    if (c == p) return true;
  if (!is_device_ptr(p)) return false;
  cache[next] = p;
  next = (next + 1) & 63;
  return true;
}

// ------------------------------------------------------------------ kernel parameters

// Polynomial power evaluation mode: the reference's powe == 2 / powe == 4 fast
// paths (frame/dgsks_ref.c:297-319), small integer powers by repeated squaring,
// anything else through pow().
static void poly_mode(double powe, KerParams &P) {
  P.powe = powe;
  P.ipow = 0;
  if (powe == 2.0) {
    P.pmode = POW_SQUARE;
  } else if (powe == 4.0) {
    P.pmode = POW_FOURTH;
  } else if (powe == std::floor(powe) && std::fabs(powe) <= 64.0) {
    P.pmode = POW_INT;
    P.ipow = (int)powe;
  } else {
    P.pmode = POW_GENERAL;
  }
}

void laplace_params(int k, double *powe, double *scal) {
  // frame/dgsks.c:445-447
  *powe = 0.5 * (2.0 - (double)k);
  *scal = std::tgamma(0.5 * k + 1.0) / ((double)k * (double)(k - 2) * std::pow(M_PI, 0.5 * k));
}

int make_params(const ks_t *ker, int k, KerParams &P, bool warn) {
  P = KerParams{ker->scal, ker->cons, ker->powe, 0, POW_GENERAL};
  switch (ker_type(ker)) {
    case KS_GAUSSIAN:
    case KS_TANH:
    case KS_QUARTIC:
    case KS_MULTIQUADRATIC:
    case KS_EPANECHNIKOV:
      return KS_OK;
    case KS_GAUSSIAN_VAR_BANDWIDTH:
      if (!ker->hi || !ker->hj)
        return raise_error(KS_ERR_NO_BANDWIDTH, false,
                           "Error dgsks(): bandwidth vector has been initialized yet.\n");
      return KS_OK;
    case KS_POLYNOMIAL:
      poly_mode(ker->powe, P);
      return KS_OK;
    case KS_LAPLACE:
      if (k < 3 && warn) raise_error(KS_ERR_LAPLACE_K, false,
                                     "Error dgsks(): laplace kernel only supports k > 2.\n");
      laplace_params(k, &P.powe, &P.scal);
      P.ipow = k - 2;  // scal * rsqrt(d2)^(k-2)
      P.pmode = POW_INT;
      return KS_OK;
    default:
      return raise_error(KS_ERR_INVALID_TYPE, true, "Error dgsks(): illegal kernel type\n");
  }
}

// ------------------------------------------------------------------ launch helpers

constexpr int kTargetWGs = 2048;  // ~8 workgroups (32 waves) per CU on 256 CUs
constexpr int kMinCols = 128;     // keep each workgroup's source slice >= 128 points
// The matrix-core kernel stages its sources through LDS 64 (k > 32: 32) at a
// time; a workgroup needs about 16 stages to amortise its set-up (A fragments,
// norms, the first stage's two dependent loads).  Measured at m = n = 8192,
// k = 16: 0.1001 ms with 1024-source slices (512 workgroups), 0.1002 ms with
// 512; before the kernel staged its sources, 0.205 ms with 256 and 0.155 ms
// with 1024.
constexpr int kMinColsMfma = 1024;
// the vector kernel gathers its sources into compact tables from this many targets
constexpr int kGatherMinRows = 1024;
// symmetric sums keep one column partial per (row block, source): at most 4 GiB
constexpr size_t kSymMaxColPartials = (size_t)1 << 29;

size_t fixed_scratch_need(size_t partial_elems) { return partial_elems * sizeof(double); }

// Hands out enqueue_fused's partial sums and tables: from the stream's scratch
// (held by a lease until the caller's launches are enqueued), or from a
// caller-owned fixed buffer.
class FusedScratch {
 public:
  FusedScratch(hipStream_t s, const FixedScratch *fixed) : fixed_(fixed) {
    if (!fixed_) lease_.emplace(s);
  }
  bool fits(size_t partial_elems) const {
    return !fixed_ || fixed_scratch_need(partial_elems) <= fixed_->bytes;
  }
  int get(size_t partial_elems, double **partial) {
    if (lease_) return lease_->partials(partial_elems, partial);
    if (!fits(partial_elems)) return KS_ERR_INVALID_ARG;
    *partial = reinterpret_cast<double *>(fixed_->base);
    return KS_OK;
  }

 private:
  const FixedScratch *fixed_;
  std::optional<ScratchLease> lease_;
};

int enqueue_fused(const Problem &p, hipStream_t s, const FixedScratch *fixed) {
  FusedScratch scratch(s, fixed);
  // sources are the targets: each pair of different row blocks evaluated once,
  // on the vector units (k <= 8) or the matrix cores (9 <= k <= 64, not Laplace)
  const bool sym_valu = sym_eligible(p), sym_mfma = !sym_valu && sym_mfma_eligible(p);
  if (sym_valu || sym_mfma) {
    const SymPlan pl = sym_valu ? plan_sym(p.m, p.k, kTargetWGs)
                                : plan_sym_rows(p.m, kTargetWGs, sym_mfma_rows());
    const size_t ncol = (size_t)pl.row_blocks * p.m;
    // the vector form runs over compact tables gathered beside its partials
    // (N-body k = 3: the source loads stop being a gather through amap)
    const size_t tab = sym_valu ? sym_table_doubles(p) + 32 : 0;
    const size_t need = (size_t)pl.nsplit * p.m + ncol + tab;
    // a fixed buffer too small for the column partials takes the general path
    if (ncol <= kSymMaxColPartials && scratch.fits(need)) {
      double *partial = nullptr;
      int rc = scratch.get(need, &partial);
      if (rc != KS_OK) return rc;
      double *colp = partial + (size_t)pl.nsplit * p.m;
      Problem q = p;
      if (sym_valu) {
        double *t = colp + ncol;
        t = reinterpret_cast<double *>((reinterpret_cast<uintptr_t>(t) + 255) & ~(uintptr_t)255);
        hipError_t ge = launch_sym_gather(p, t, &q, s);
        if (ge != hipSuccess) return hip_failure(ge, "symmetric gather launch");
      }
      hipError_t e = sym_valu ? launch_fused_f64_sym(q, pl, partial, colp, s)
                              : launch_fused_f64_sym_mfma(p, pl, partial, colp, s);
      return e == hipSuccess ? KS_OK : hip_failure(e, "symmetric launch");
    }
  }
  const bool mfma = fused_mfma_eligible(p.type, p.k);
  FusedPlan plan = mfma ? plan_split(p.m, p.n, mfma_rows_per_wg(), kTargetWGs, kMinColsMfma)
                        : plan_fused(p.m, p.n, p.k, kTargetWGs, kMinCols);
  // the vector kernel reads its sources from compact tables when there are
  // enough targets to pay for the copy
  const bool gather = !mfma && p.m >= kGatherMinRows;
  const size_t npart = plan.nsplit > 1 ? (size_t)plan.nsplit * p.m : 0;
  const size_t tab = gather ? gather_table_doubles(p.n, p.k, p.type) + 32 : 0;
  double *partial = nullptr;
  if (npart + tab > 0) {
    int rc = scratch.get(npart + tab, &partial);
    if (rc != KS_OK) return rc;
  }
  Problem q = p;
  if (gather) {
    double *t = partial + npart;
    t = reinterpret_cast<double *>((reinterpret_cast<uintptr_t>(t) + 255) & ~(uintptr_t)255);
    hipError_t ge = launch_source_gather(p, t, &q, s);
    if (ge != hipSuccess) return hip_failure(ge, "source gather launch");
  }
  hipError_t e = mfma ? launch_fused_f64_mfma(p, plan, partial, s)
                      : launch_fused_f64(q, plan, partial, s);
  if (e != hipSuccess)
    return hip_failure(e, mfma ? "fused_f64_mfma launch" : "fused_f64 launch");
  // the source splits' partial sums, added in split order
  if (plan.nsplit > 1) e = launch_split_reduce(partial, plan.nsplit, p.m, p.u, p.umap, nullptr, s);
  return e == hipSuccess ? KS_OK : hip_failure(e, "split reduction launch");
}

int device_problem(const char *name, const ks_t *kernel, int m, int n, int k, double *u,
                   const int *umap, const double *XA, const double *XA2, const int *amap,
                   const double *XB, const double *XB2, const int *bmap, const double *w,
                   const int *wmap, Problem &p, bool &empty) {
  empty = false;
  if (!kernel || m < 0 || n < 0 || k < 0)
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): invalid argument\n", name);
  if (m == 0 || n == 0 || k == 0) {
    empty = true;
    return KS_OK;
  }
  if (!u || !umap || !XA || !XA2 || !amap || !XB || !XB2 || !bmap || !w || !wmap)
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): NULL array\n", name);
  int rc = ensure_device();
  if (rc != KS_OK) return raise_error(rc, false, "%s(): no GPU available\n", name);
  rc = make_params(kernel, k, p.P, true);
  if (rc != KS_OK && rc != KS_ERR_LAPLACE_K) return rc;
  p.m = m;
  p.n = n;
  p.k = k;
  p.type = ker_type(kernel);
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
  // A host pointer handed to a kernel would fault the GPU: check every array once
  // (results are cached per thread for repeated calls with the same buffers).
  const void *arrays[] = {u, umap, XA, XA2, amap, XB, XB2, bmap, w, wmap,
                          p.type == KS_GAUSSIAN_VAR_BANDWIDTH ? p.hi : nullptr,
                          p.type == KS_GAUSSIAN_VAR_BANDWIDTH ? p.hj : nullptr};
  for (const void *a : arrays)
    if (a && !device_ptr_cached(a))
      return raise_error(KS_ERR_MIXED_POINTERS, false, "%s(): every array must be device memory\n",
                         name);
  if (p.type == KS_GAUSSIAN_VAR_BANDWIDTH && (!p.hi || !p.hj))
    return raise_error(KS_ERR_NO_BANDWIDTH, false, "%s(): %s\n", name,
                       err_string(KS_ERR_NO_BANDWIDTH));
  return KS_OK;
}

}  // namespace gsks

// ================================================================== C ABI

extern "C" {

int ks_last_error(void) {
  int e = t_last_error;
  t_last_error = KS_OK;
  return e;
}

const char *ks_error_string(int code) { return err_string(code); }

const char *ks_error_detail(void) { return t_detail; }

void ks_set_error_mode(int mode) {
  g_errmode.store(mode == KS_ERRMODE_RETURN ? KS_ERRMODE_RETURN : KS_ERRMODE_EXIT);
}

int ks_get_error_mode(void) { return g_errmode.load(); }

double *ks_malloc_aligned(int m, int n, int size) {
  // frame/ks_util.c:34-55: posix_memalign, exit(1) on failure
  void *p = nullptr;
  size_t bytes = (size_t)(m > 0 ? m : 0) * (size_t)(n > 0 ? n : 0) * (size_t)(size > 0 ? size : 0);
  if (bytes == 0) bytes = 64;
  if (posix_memalign(&p, 64, bytes) != 0 || !p) {
    raise_error(KS_ERR_ALLOC, true, "ks_malloc_aligned(): posix_memalign() failures\n");
    return nullptr;
  }
  return (double *)p;
}

int dgsks_device(const ks_t *kernel, int m, int n, int k, double *u, const int *umap,
                 const double *XA, const double *XA2, const int *amap, const double *XB,
                 const double *XB2, const int *bmap, const double *w, const int *wmap,
                 int precision, void *stream) {
  Problem p;
  bool empty = false;
  int rc = device_problem("dgsks_device", kernel, m, n, k, u, umap, XA, XA2, amap, XB, XB2, bmap,
                          w, wmap, p, empty);
  if (rc != KS_OK || empty) return rc;
  hipStream_t s = (hipStream_t)stream;  // NULL: the HIP null stream (torch's default stream)
  int r;
  if (precision == KS_PREC_FP64) {
    r = enqueue_fused(p, s);
  } else {
    r = enqueue_mixed(p, precision, s);
  }
  if (r != KS_OK) return raise_error(r, false, "dgsks_device(): launch failed (%s)\n", err_string(r));
  return KS_OK;
}

// Checks shared by sgsks_device() and sgsks() (the host path checks its arrays itself).
static int single_precision_ok(const char *name, int type, int k, int precision) {
  if (precision == KS_PREC_FP64) return KS_OK;
  if (precision != KS_PREC_MIXED_FP32 && precision != KS_PREC_MIXED_BF16X3)
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): unknown precision\n", name);
  if (type != KS_GAUSSIAN && type != KS_GAUSSIAN_VAR_BANDWIDTH)
    return raise_error(KS_ERR_INVALID_ARG, false,
                       "%s(): mixed precision supports the Gaussian and variable-bandwidth "
                       "kernels only\n", name);
  if (k > 128)
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): mixed precision supports k <= 128\n",
                       name);
  return KS_OK;
}

int sgsks_device(const ks_t *kernel, int m, int n, int k, float *u, const int *umap,
                 const float *XA, const float *XA2, const int *amap, const float *XB,
                 const float *XB2, const int *bmap, const float *w, const int *wmap,
                 int precision, void *stream) {
  if (!kernel || m < 0 || n < 0 || k < 0)
    return raise_error(KS_ERR_INVALID_ARG, false, "sgsks_device(): invalid argument\n");
  if (m == 0 || n == 0 || k == 0) return KS_OK;
  if (!u || !umap || !XA || !amap || !XB || !bmap || !w || !wmap)
    return raise_error(KS_ERR_INVALID_ARG, false, "sgsks_device(): NULL array\n");
  int rc = ensure_device();
  if (rc != KS_OK) return raise_error(rc, false, "sgsks_device(): no GPU available\n");
  SingleProblem p;
  rc = make_params(kernel, k, p.P, true);
  if (rc != KS_OK && rc != KS_ERR_LAPLACE_K) return rc;
  rc = single_precision_ok("sgsks_device", ker_type(kernel), k, precision);
  if (rc != KS_OK) return rc;
  const bool varbw = ker_type(kernel) == KS_GAUSSIAN_VAR_BANDWIDTH;
  p.m = m;
  p.n = n;
  p.k = k;
  p.type = ker_type(kernel);
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
  const void *arrays[] = {u, umap, XA, XA2, amap, XB, XB2, bmap, w, wmap, p.hi, p.hj};
  for (const void *a : arrays)
    if (a && !device_ptr_cached(a))
      return raise_error(KS_ERR_MIXED_POINTERS, false,
                         "sgsks_device(): every array must be device memory\n");
  const int r = enqueue_single(p, precision, (hipStream_t)stream);
  if (r != KS_OK) return raise_error(r, false, "sgsks_device(): launch failed (%s)\n", err_string(r));
  return KS_OK;
}

void sgsks(ks_t *kernel, int m, int n, int k, float *u, int *umap, float *XA, float *XA2,
           int *amap, float *XB, float *XB2, int *bmap, float *w, int *wmap) {
  set_last_error(KS_OK);
  // the early return and the Laplace side effect of dgsks(), frame/dgsks.c:394-397,445-447
  if (m == 0 || n == 0 || k == 0) {
    if (g_errmode.load() == KS_ERRMODE_EXIT) printf("sgsks(): early return\n");
    return;
  }
  if (!kernel || m < 0 || n < 0 || k < 0) {
    raise_error(KS_ERR_INVALID_ARG, false, "sgsks(): invalid argument\n");
    return;
  }
  KerParams P;
  int rc = make_params(kernel, k, P, true);
  if (rc != KS_OK && rc != KS_ERR_LAPLACE_K) return;
  if (ker_type(kernel) == KS_LAPLACE) {
    kernel->powe = P.powe;
    kernel->scal = P.scal;
  }
  run_staged_single(kernel, m, n, k, u, umap, XA, XA2, amap, XB, XB2, bmap, w, wmap);
}

void sgsks_wrapper(int m, int n, int k, float *u, int *umap, float *XA, float *XA2, int *alpha,
                   float *XB, float *XB2, int *beta, float *w, int *omega, int type, double scal,
                   double cons, double powe, double *h) {
  ks_t kernel;
  memset(&kernel, 0, sizeof(kernel));
  set_ker_type(&kernel, type);
  kernel.scal = scal;
  kernel.cons = cons;
  kernel.powe = powe;
  kernel.hi = h;
  kernel.hj = h;
  sgsks(&kernel, m, n, k, u, umap, XA, XA2, alpha, XB, XB2, beta, w, omega);
}

void dgsks(ks_t *kernel, int m, int n, int k, double *u, int *umap, double *XA, double *XA2,
           int *amap, double *XB, double *XB2, int *bmap, double *w, int *wmap) {
  set_last_error(KS_OK);
  // frame/dgsks.c:394-397
  if (m == 0 || n == 0 || k == 0) {
    if (g_errmode.load() == KS_ERRMODE_EXIT) printf("dgsks(): early return\n");
    return;
  }
  if (!kernel || m < 0 || n < 0 || k < 0) {
    raise_error(KS_ERR_INVALID_ARG, false, "dgsks(): invalid argument\n");
    return;
  }
  KerParams P;
  int rc = make_params(kernel, k, P, true);
  if (rc != KS_OK && rc != KS_ERR_LAPLACE_K) return;
  if (ker_type(kernel) == KS_LAPLACE) {  // the reference's side effect, frame/dgsks.c:445-447
    kernel->powe = P.powe;
    kernel->scal = P.scal;
  }
  run_staged(kernel, m, n, k, u, umap, XA, XA2, amap, XB, XB2, bmap, w, wmap,
             /*reference_algorithm=*/false);
}

void dgsks_ref(ks_t *kernel, int m, int n, int k, double *u, int *umap, double *XA,
               double *XA2, int *amap, double *XB, double *XB2, int *bmap, double *w,
               int *wmap) {
  set_last_error(KS_OK);
  if (!kernel || m <= 0 || n <= 0 || k <= 0) return;  // frame/dgsks_ref.c has no early-return print
  KerParams P;
  int rc = make_params(kernel, k, P, true);
  if (rc != KS_OK && rc != KS_ERR_LAPLACE_K) return;
  if (ker_type(kernel) == KS_LAPLACE) {  // frame/dgsks_ref.c:146-148
    kernel->powe = P.powe;
    kernel->scal = P.scal;
  }
  run_staged(kernel, m, n, k, u, umap, XA, XA2, amap, XB, XB2, bmap, w, wmap,
             /*reference_algorithm=*/true);
}

// frame/dgsks.c:783-826 / frame/dgsks_ref.c:537-580.  The descriptor is zero-initialised
// and cons is set (the reference leaves kernel.cons uninitialised).
void dgsks_wrapper(int m, int n, int k, double *u, int *umap, double *XA, double *XA2,
                   int *alpha, double *XB, double *XB2, int *beta, double *w, int *omega,
                   int type, double scal, double cons, double powe, double *h) {
  ks_t kernel;
  memset(&kernel, 0, sizeof(kernel));
  set_ker_type(&kernel, type);
  kernel.scal = scal;
  kernel.cons = cons;
  kernel.powe = powe;
  kernel.hi = h;
  kernel.hj = h;
  dgsks(&kernel, m, n, k, u, umap, XA, XA2, alpha, XB, XB2, beta, w, omega);
}

void dgsks_ref_wrapper(int m, int n, int k, double *u, int *umap, double *XA, double *XA2,
                       int *alpha, double *XB, double *XB2, int *beta, double *w, int *omega,
                       int type, double scal, double cons, double powe, double *h) {
  ks_t kernel;
  memset(&kernel, 0, sizeof(kernel));
  set_ker_type(&kernel, type);
  kernel.scal = scal;
  kernel.cons = cons;
  kernel.powe = powe;
  kernel.hi = h;
  kernel.hj = h;
  dgsks_ref(&kernel, m, n, k, u, umap, XA, XA2, alpha, XB, XB2, beta, w, omega);
}

}  // extern "C"
