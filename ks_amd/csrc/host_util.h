// Host-side helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/ks.h"
#include "gsks_kernels.h"

namespace gsks {

// The kernel type of a descriptor as the int the caller stored.  ks_type has
// no fixed underlying type in C++, so loading a value outside 0..7 through the
// enum is undefined behaviour; callers may store any int there (an illegal
// type is an error reported at run time, as the reference reports it).
static_assert(sizeof(ks_type) == sizeof(int), "ks_type is stored as an int");
inline int ker_type(const ks_t *k) {
  int t;
  std::memcpy(&t, &k->type, sizeof(t));
  return t;
}
inline void set_ker_type(ks_t *k, int t) { std::memcpy(&k->type, &t, sizeof(t)); }

// Carves one stream-ordered allocation (hipMallocAsync on the call's stream,
// freed with hipFreeAsync after the last launch that reads it) into 256-byte
// aligned arrays.  A layout function runs twice: on a null base to measure,
// then on the allocation.
struct Carve {
  char *base = nullptr;
  size_t used = 0;
  template <class T>
  T *take(size_t count) {
    T *r = base ? reinterpret_cast<T *>(base + used) : nullptr;
    used += (sizeof(T) * count + 255) & ~(size_t)255;
    return r;
  }
};

template <class Layout>
int carve_alloc(Layout &&layout, char **buf, hipStream_t s) {
  Carve c;
  layout(c);
  if (hipMallocAsync((void **)buf, c.used, s) != hipSuccess) {
    (void)hipGetLastError();
    *buf = nullptr;
    return KS_ERR_ALLOC;
  }
  c.base = *buf;
  c.used = 0;
  layout(c);
  return KS_OK;
}

const char *err_string(int code);
int raise_error(int code, bool fatal, const char *fmt, ...);
void set_last_error(int code);
int hip_failure(hipError_t e, const char *where);
int ensure_device();
hipStream_t thread_stream();

struct Scratch;
// The grow-only scratch of one (device, stream), held from construction to
// destruction.  An enqueue function declares the lease before it asks for
// buffers and keeps it until its last launch is enqueued, so another host thread
// on the same stream cannot grow (and free) a buffer between its hand-out and
// the launch that reads it (the lock is recursive: nested enqueues on one
// thread share it).
//  * partials(): split-source partial sums and tables (FP64).
//  * aux(): a byte arena, separate from the partials, for front ends that
//    re-lay out their inputs before a summation.
// Growing synchronises the stream and frees the old buffer -- unless a stream
// capture has handed out buffers of this (device, stream): a captured graph
// keeps their addresses and may be replayed at any time, so they are retired
// instead and live until the process ends.  Growing during a capture is an
// error (KS_ERR_ALLOC, with ks_error_detail() saying so): run the call once on
// the stream before capturing it.
class ScratchLease {
 public:
  explicit ScratchLease(hipStream_t s);
  ~ScratchLease();
  ScratchLease(const ScratchLease &) = delete;
  ScratchLease &operator=(const ScratchLease &) = delete;
  int partials(size_t partial_elems, double **partial);
  int aux(size_t bytes, char **base);

 private:
  hipStream_t s_;
  Scratch *sc_;
};
void laplace_params(int k, double *powe, double *scal);
int make_params(const ks_t *ker, int k, KerParams &P, bool warn);

// A device buffer owned by the caller of enqueue_fused, used instead of the
// stream's grow-only scratch for the partial sums and tables.  A summation
// whose scratch does not fit is refused with KS_ERR_INVALID_ARG before
// anything is enqueued.
struct FixedScratch {
  char *base;
  size_t bytes;
};
// Bytes of a FixedScratch that partial_elems doubles take.
size_t fixed_scratch_need(size_t partial_elems);

// The checks of the stream-ordered entry points (sizes, NULL arrays, a GPU,
// kernel parameters, device pointers, bandwidths), raising the error they find;
// fills p, or sets empty when m, n or k is zero (nothing to do).
int device_problem(const char *name, const ks_t *kernel, int m, int n, int k, double *u,
                   const int *umap, const double *XA, const double *XA2, const int *amap,
                   const double *XB, const double *XB2, const int *bmap, const double *w,
                   const int *wmap, Problem &p, bool &empty);

// Fused FP64 summation of one problem (all device pointers) on stream s, with
// its scratch from the stream's scratch, or from `fixed` when it is given.
int enqueue_fused(const Problem &p, hipStream_t s, const FixedScratch *fixed = nullptr);
// The reference algorithm (unfused: cross-term panel, elementwise kernel, GEMV).
int enqueue_reference(const Problem &p, hipStream_t s);
// Multi-right-hand-side summation: p.u / p.w are [npoints][nrhs] matrices.
int enqueue_multirhs(const Problem &p, int nrhs, hipStream_t s);
// Mixed-precision summation (KS_PREC_MIXED_*).
int enqueue_mixed(const Problem &p, int precision, hipStream_t s);
// Mixed-precision multi-right-hand-side Gaussian summation (FP32 matrix cores on
// centred points where a per-block bound allows, FP64 elsewhere).  `centre`
// (k doubles, device) translates the points for the FP32 tier; NULL computes
// the mean of the mapped targets and sources.
constexpr int kMixedMrhsMaxK = 64;
constexpr int kMixedMrhsMaxRhs = 33;
int enqueue_multirhs_mixed(const Problem &p, int nrhs, const double *centre, hipStream_t s);
// centre[q] = mean over the m rows of XA and the n rows of XB (identity map ident),
// summed in a fixed order through `work` (centre_work_elems(m + n) doubles).
size_t centre_work_elems(long rows);
int enqueue_centre(const double *XA, int m, const double *XB, int n, int k, const int *ident,
                   double *work, double *centre, hipStream_t s);
// Xc = X - centre, row by row (X [rows][k]).
int enqueue_shift(const double *X, int rows, int k, const double *centre, double *Xc,
                  hipStream_t s);

// Single-precision problem (sgsks / sgsks_device): float arrays, FP64 bandwidths.
struct SingleProblem {
  int m, n, k;
  int type;
  KerParams P;
  float *u;
  const int *umap;
  const float *XA, *XA2;  // XA2 / XB2 may be NULL (norms computed in FP64)
  const int *amap;
  const float *XB, *XB2;
  const int *bmap;
  const float *w;
  const int *wmap;
  const double *hi, *hj;
};
// Widen the selected rows to FP64, sum with the FP64 (precision KS_PREC_FP64)
// or mixed kernels, and add the row sums into the float u, on stream s.
int enqueue_single(const SingleProblem &p, int precision, hipStream_t s);
// The precision sgsks() uses: FP64 (mixed precision only on request, through
// sgsks_device).
int single_default_precision(int type, int m, int n, int k);

// Host entry behind dgsks()/dgsks_ref(): all-device pointers run in place,
// all-host pointers are staged through HBM, then the call synchronises.
void run_staged(const ks_t *kernel, int m, int n, int k, double *u, const int *umap,
                const double *XA, const double *XA2, const int *amap, const double *XB,
                const double *XB2, const int *bmap, const double *w, const int *wmap,
                bool reference_algorithm);

// Host entry behind sgsks(): as run_staged, with float arrays.
void run_staged_single(const ks_t *kernel, int m, int n, int k, float *u, const int *umap,
                       const float *XA, const float *XA2, const int *amap, const float *XB,
                       const float *XB2, const int *bmap, const float *w, const int *wmap);

// True when p is device (or managed) memory that kernels can dereference.
bool is_device_ptr(const void *p);
bool device_ptr_cached(const void *p);

}  // namespace gsks
