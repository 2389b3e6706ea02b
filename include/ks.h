/*
 * ks.h -- public C API of the MI355X fused kernel summation library
 * (ks_amd/libgsks.so).
 *
 * The first half is ABI-identical to the reference header (include/ks.h:46-111
 * of ChenhanYu/ks): the same KS_RHS / KS_NUM_THREAD constants, the same
 * ks_type enumerator order, the same struct kernel_s field order and size
 * (56 bytes on LP64), and the same dgsks / dgsks_ref / ks_malloc_aligned
 * prototypes.  A program written against the reference compiles and links
 * against this header and library unchanged.
 *
 * Semantics kept from the reference:
 *   u[umap[i]] += sum_j K(XA[amap[i]], XB[bmap[j]]) * w[wmap[j]]
 *   - XA / XB are row-major [npoints][k] (k leading), XA2 / XB2 the caller's
 *     squared row norms, all maps are int32 global point ids;
 *   - u is accumulated, never overwritten;
 *   - KS_LAPLACE writes kernel->powe = (2-k)/2 and
 *     kernel->scal = Gamma(k/2+1) / (k (k-2) pi^(k/2))  (frame/dgsks.c:445-447);
 *   - KS_GAUSSIAN_VAR_BANDWIDTH reads kernel->hi[amap[i]] and kernel->hj[bmap[j]]
 *     (global ids) and evaluates exp(-0.5 d^2 hi hj);
 *   - KS_MULTIQUADRATIC is d^2 + cons (no square root);
 *   - m, n or k == 0 returns early.
 * Pointers may be host memory (the library stages them through HBM) or
 * device memory (used in place, no copies).
 *
 * Everything after the "extensions" marker has no counterpart in the reference.
 */
#ifndef __KS_H__
#define __KS_H__

#include <stdio.h>
#include <stdlib.h>
#include <math.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KS_RHS 1          /* reference include/ks.h:46 */
#define KS_NUM_THREAD 68  /* reference include/ks.h:47 (kept for source compatibility) */

/* reference include/ks.h:49-58 -- the enumerator order is ABI */
typedef enum {
  KS_GAUSSIAN,
  KS_POLYNOMIAL,
  KS_LAPLACE,
  KS_GAUSSIAN_VAR_BANDWIDTH,
  KS_TANH,
  KS_QUARTIC,
  KS_MULTIQUADRATIC,
  KS_EPANECHNIKOV
} ks_type;

/* reference include/ks.h:60-69 */
struct kernel_s {
  ks_type type;
  double powe;
  double scal;
  double cons;
  double *hi; /* variable-bandwidth factors, indexed by global target id */
  double *hj; /* variable-bandwidth factors, indexed by global source id */
  double *h;  /* unused by dgsks (as in the reference) */
};

typedef struct kernel_s ks_t;

/* Fused FP64 kernel summation on the GPU.  Replaces reference include/ks.h:73-88
 * (frame/dgsks.c:367-776).  One fused HIP kernel per call; never forms the
 * m x n kernel matrix.  Returns when u holds the result. */
void dgsks(
    ks_t   *kernel,
    int    m,
    int    n,
    int    k,
    double *u,
    int    *umap,
    double *XA,
    double *XA2,
    int    *amap,
    double *XB,
    double *XB2,
    int    *bmap,
    double *w,
    int    *wmap
    );

/* The reference algorithm on the GPU: gather, FP64 GEMM of the cross term,
 * elementwise kernel function, GEMV, scatter -- unfused, in column panels.
 * Replaces reference include/ks.h:90-105 (frame/dgsks_ref.c:94-535). */
void dgsks_ref(
    ks_t   *kernel,
    int    m,
    int    n,
    int    k,
    double *u,
    int    *umap,
    double *XA,
    double *XA2,
    int    *amap,
    double *XB,
    double *XB2,
    int    *bmap,
    double *w,
    int    *wmap
    );

/* 64-byte aligned host allocation of m*n elements of `size` bytes.  Replaces
 * reference include/ks.h:107-111 (frame/ks_util.c:34-55). */
double *ks_malloc_aligned(
    int    m,
    int    n,
    int    size
    );

/* Flat ctypes entry points: replace frame/dgsks.c:783-826 and
 * frame/dgsks_ref.c:537-580 (called by python/gsks.py:85,179).  The descriptor
 * is built with hi = hj = h and, unlike the reference, with cons set. */
void dgsks_wrapper(int m, int n, int k, double *u, int *umap, double *XA, double *XA2,
                   int *alpha, double *XB, double *XB2, int *beta, double *w, int *omega,
                   int type, double scal, double cons, double powe, double *h);
void dgsks_ref_wrapper(int m, int n, int k, double *u, int *umap, double *XA, double *XA2,
                       int *alpha, double *XB, double *XB2, int *beta, double *w, int *omega,
                       int type, double scal, double cons, double powe, double *h);

/* ======================= extensions (no reference counterpart) ============ */

/* Error reporting.  The reference prints and calls exit(1) on an illegal kernel
 * type (frame/dgsks.c:465-467); that stays the default (KS_ERRMODE_EXIT).  With
 * KS_ERRMODE_RETURN the call prints nothing, leaves u untouched and records the
 * code, which ks_last_error() returns (per thread) and then clears.  The void
 * entry points (dgsks, dgsks_ref, sgsks and their wrappers) clear it when they
 * start, so after one of them it describes that call. */
enum {
  KS_OK = 0,
  KS_ERR_INVALID_TYPE = 1,   /* kernel->type outside ks_type */
  KS_ERR_INVALID_ARG = 2,    /* negative size, NULL array, k too large */
  KS_ERR_NO_BANDWIDTH = 3,   /* VAR_BANDWIDTH with hi or hj NULL */
  KS_ERR_LAPLACE_K = 4,      /* LAPLACE with k < 3 (reference prints and continues) */
  KS_ERR_DEVICE = 5,         /* no usable GPU, or a HIP call failed */
  KS_ERR_ALLOC = 6,          /* device or pinned allocation failed */
  KS_ERR_MIXED_POINTERS = 7, /* some arrays on the device, others on the host */
  KS_ERR_LIST = 8            /* inconsistent list description */
};
enum { KS_ERRMODE_EXIT = 0, KS_ERRMODE_RETURN = 1 };

int ks_last_error(void);
const char *ks_error_string(int code);
/* Which HIP call failed last on this thread, with HIP's message. */
const char *ks_error_detail(void);
void ks_set_error_mode(int mode);
int ks_get_error_mode(void);

/* Precision of a summation. */
enum {
  KS_PREC_FP64 = 0,       /* everything FP64 (the drop-in) */
  KS_PREC_MIXED_FP32 = 1, /* FP32 cross term, FP32 K(.), FP64 norms/near pairs/u */
  KS_PREC_MIXED_BF16X3 = 2 /* split-BF16 (hi/lo, 3 products) cross term on MFMA */
};

/* Stream-ordered device entry point: every array is a device pointer, the work
 * is enqueued on `stream` (a hipStream_t; NULL = the HIP null stream), the call
 * does not synchronise and does not allocate once its scratch has grown to the
 * problem size (the first call of a given size on a stream).  Returns a KS_* code.  kernel is read, not written (the
 * Laplace parameters are derived internally).  hi / hj must be device
 * pointers for VAR_BANDWIDTH. */
int dgsks_device(const ks_t *kernel, int m, int n, int k,
                 double *u, const int *umap,
                 const double *XA, const double *XA2, const int *amap,
                 const double *XB, const double *XB2, const int *bmap,
                 const double *w, const int *wmap, int precision, void *stream);

/* dgsks_device without allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The arguments and
 * results are dgsks_device's with precision KS_PREC_FP64 (any other precision
 * returns KS_ERR_INVALID_ARG); the same FP64 kernels run.  Their scratch
 * (split partial sums, counters, compact source tables) lives in a static
 * 64 MiB device array that every GPU the library runs on holds until the
 * process ends; a summation that needs more is refused with
 * KS_ERR_INVALID_ARG before anything is enqueued.  The call allocates nothing
 * and does not synchronise.  All calls on one GPU share that array, which keeps
 * the last call's contents: calls that overlap on one GPU (other streams or
 * threads) race and must be ordered by the caller. */
int dgsks_device_noalloc(const ks_t *kernel, int m, int n, int k,
                         double *u, const int *umap,
                         const double *XA, const double *XA2, const int *amap,
                         const double *XB, const double *XB2, const int *bmap,
                         const double *w, const int *wmap, int precision, void *stream);

/* Single precision.  The reference declares a single-precision micro-kernel
 * for every kernel type and implements none of them
 * (micro_kernel/x86_64/haswell/gaussian_int_d8x6.c:38-54: "not yet
 * implemented"); these take dgsks()'s arguments (frame/dgsks.c:373-390) as
 * float arrays.  The rows the maps select are widened to FP64 (exactly), the
 * summation runs in `precision` (mixed precision: Gaussian kernels, k <= 128),
 * and u[umap[i]] receives float(u[umap[i]] + row sum i), one rounding.  XA2 /
 * XB2 may be NULL: the squared norms are then computed in FP64 from the float
 * coordinates.  hi / hj (variable bandwidth) stay double.  sgsks_device takes
 * device pointers and is enqueued on `stream` (NULL = the HIP null stream); a
 * umap entry may not repeat there.  sgsks takes all-host or all-device
 * pointers, synchronises and sums in KS_PREC_FP64 (the mixed forms' accuracy
 * depends on the data, so only sgsks_device offers them); a repeated umap
 * entry keeps the value of its last row, as in dgsks. */
int sgsks_device(const ks_t *kernel, int m, int n, int k,
                 float *u, const int *umap,
                 const float *XA, const float *XA2, const int *amap,
                 const float *XB, const float *XB2, const int *bmap,
                 const float *w, const int *wmap, int precision, void *stream);
void sgsks(ks_t *kernel, int m, int n, int k, float *u, int *umap, float *XA, float *XA2,
           int *amap, float *XB, float *XB2, int *bmap, float *w, int *wmap);
/* sgsks with the descriptor built from its fields (as dgsks_wrapper). */
void sgsks_wrapper(int m, int n, int k, float *u, int *umap, float *XA, float *XA2,
                   int *alpha, float *XB, float *XB2, int *beta, float *w, int *omega,
                   int type, double scal, double cons, double powe, double *h);

/* Squared row norms X2[i] = sum_p X[i*k+p]^2 on the device (the norm
 * precompute of frame/omp_dgsks_list.cpp:75-81), enqueued on `stream`
 * (NULL = the HIP null stream). */
int ks_sqnorm_device(const double *X, long nx, int k, double *X2, void *stream);

/* Several right-hand sides at once (the reference's KS_RHS > 1 layout,
 * include/ks.h:46): U[umap[i]*nrhs + r] += sum_j K(a_i, b_j) W[wmap[j]*nrhs + r]
 * for r < nrhs <= 72.  Device pointers, enqueued on `stream`. */
int dgsks_multirhs_device(const ks_t *kernel, int m, int n, int k, int nrhs,
                          double *U, const int *umap,
                          const double *XA, const double *XA2, const int *amap,
                          const double *XB, const double *XB2, const int *bmap,
                          const double *W, const int *wmap, void *stream);

/* Mixed-precision form of dgsks_multirhs_device for the Gaussian kernel
 * (k <= 64, nrhs <= 33; no counterpart in the reference, whose summations are
 * FP64 only).  Every 16 x 16 pair block whose exponent error, bounded from the
 * norms of the points translated by their mean, is at most 2^-15 is evaluated
 * with FP32 cross terms and kernel values on the matrix cores (FP64 norms,
 * distance assembly and accumulation); every other block exactly as the FP64
 * form.  Device pointers, enqueued on `stream`. */
int dgsks_multirhs_mixed_device(const ks_t *kernel, int m, int n, int k, int nrhs,
                                double *U, const int *umap,
                                const double *XA, const double *XA2, const int *amap,
                                const double *XB, const double *XB2, const int *bmap,
                                const double *W, const int *wmap, void *stream);

/* Gradient of L = sum_i g[i] u[i] for the Gaussian summation u = K w,
 * K_ij = exp(-|a_i - b_j|^2 / (2 h^2)), identity maps, XA [m][k], XB [n][k]:
 * dXA [m][k], dXB [n][k], dw [n] (the transposed summation) and dh [1]
 * (outputs may be NULL).  The kernel matrix is recomputed, never stored.
 * precision: KS_PREC_FP64 (k <= 70) or KS_PREC_MIXED_FP32 (k <= 32: the two
 * summations through dgsks_multirhs_mixed_device's kernel on centred points).
 * Device pointers, enqueued on `stream`. */
int ks_gauss_grad_device(int m, int n, int k, const double *XA, const double *XB,
                         const double *w, const double *g, double h,
                         double *dXA, double *dXB, double *dw, double *dh, int precision,
                         void *stream);

/* The two halves of ks_gauss_grad_device, for a training step (no counterpart
 * in the reference, whose summations have no gradient).  The forward half
 * computes u = K w (u may be NULL) and writes the gradient state -- V = K [w b]
 * and u as [m][k + 1] rows, then the k coordinates of the frame the precision
 * works in -- into `state` (m (k + 1) + k doubles) from one summation of k + 1
 * right-hand sides.  The backward half takes g and the same points, weights,
 * h and precision, runs the second summation and forms dXA, dXB, dw and dh
 * (any may be NULL); state may be NULL when dXA and dh are.  A step is then
 * two summations instead of three.  Device pointers, enqueued on `stream`. */
int ks_gauss_grad_forward_device(int m, int n, int k, const double *XA, const double *XB,
                                 const double *w, double h, double *u, double *state,
                                 int precision, void *stream);
int ks_gauss_grad_backward_device(int m, int n, int k, const double *XA, const double *XB,
                                  const double *w, const double *g, double h,
                                  const double *state, double *dXA, double *dXB,
                                  double *dw, double *dh, int precision, void *stream);

/* Gradient of the mapped summation (no counterpart in the reference, whose
 * summations have no gradient).  For L = sum_x g[x] du[x] with
 *   du[umap[i]] += sum_j K(XA[amap[i]], XB[bmap[j]]) w[wmap[j]],
 * the kernel a KS_GAUSSIAN (K = exp(scal d2)) or KS_GAUSSIAN_VAR_BANDWIDTH
 * (K = exp(-0.5 hi[amap[i]] hj[bmap[j]] d2)) descriptor gives, as
 * frame/dgsks_ref.c:254-288 evaluates it, with the norms in d2 those of the
 * points themselves, ADD the gradient into the tables at the mapped entries:
 *   dXA [.][k] at amap[i], dXB [.][k] at bmap[j], dw at wmap[j],
 *   *dscal (Gaussian only), dhi at amap[i] and dhj at bmap[j] (variable
 *   bandwidth only),
 * a repeated map entry receiving the sum of its rows' terms, in row order (the
 * result is the same on every call).  Any output may be NULL; outputs may be
 * one table (dXA == dXB when XA == XB, dhi == dhj when hi == hj): the additions
 * run one after the other.  Two multi-right-hand-side summations recompute the
 * kernel matrix (k + 1 right-hand sides each for the Gaussian, k + 2 and k + 3
 * for the variable bandwidth), on the mapped points translated by their mean.
 * precision: KS_PREC_FP64 (Gaussian k <= 71, variable bandwidth k <= 69) or
 * KS_PREC_MIXED_FP32 (Gaussian, k <= 32: dgsks_multirhs_mixed_device's
 * kernel).  Device pointers (map entries below 2^31), enqueued on `stream`;
 * temporary buffers from the stream-ordered pool. */
int ks_grad_device(const ks_t *kernel, int m, int n, int k,
                   const double *g, const int *umap,
                   const double *XA, const int *amap,
                   const double *XB, const int *bmap,
                   const double *w, const int *wmap,
                   double *dXA, double *dXB, double *dw,
                   double *dscal, double *dhi, double *dhj, int precision, void *stream);

/* Batched summation: ntask independent summations over shared point pools,
 * one persistent launch.  Task t owns the ranges [a_off[t], a_off[t+1]) of
 * amap/umap and [b_off[t], b_off[t+1]) of bmap/wmap and uses kernels[t]
 * (one descriptor per task; the reference's omp_dgsks_list takes a single
 * descriptor, frame/omp_dgsks_list.cpp:297-420).  u[x] accumulates the sum of
 * every task's contribution to point x, so tasks may share target points.
 * Host pointers (staged) or device pointers.  kernels and the offset arrays
 * are always host memory. */
int dgsks_list(int ntask, ks_t *kernels, int k, double *u, long nu,
               const long *a_off, const int *umap, const int *amap,
               const long *b_off, const int *bmap, const int *wmap,
               const double *XA, const double *XA2,
               const double *XB, const double *XB2, const double *w,
               int precision);

/* Reusable batch plans.  No counterpart in the reference (its batches are
 * re-laid out on every omp_dgsks_list call, frame/omp_dgsks_list.cpp:345-359):
 * for callers that run the same batch structure many times, and for HIP graph
 * capture.  ks_list_plan_create does all host work once -- tile layout with
 * the autotuned tile shapes, device descriptors, the merge order of umap --
 * and synchronises; it must not run while a stream of the process is being
 * captured.  Arguments as for dgsks_list, except that umap/amap/bmap/wmap (and
 * the hi/hj vectors of variable-bandwidth descriptors) are DEVICE arrays that
 * must stay valid and unchanged while the plan lives; the Laplace descriptor
 * side effect is not applied (kernels is read only).
 * ks_list_plan_execute enqueues one run on `stream` (NULL: the HIP null
 * stream): u[umap[i]] += task sums, with device XA/XA2/XB/XB2/w of the sizes
 * the maps index.  It only launches kernels, so it can be captured into a HIP
 * graph and replayed (FP64 plans; a mixed-precision plan with tasks routed to
 * the mixed kernels needs one eager run on the same stream first, which sizes
 * its scratch).  Runs of one plan must not overlap. */
typedef struct ks_list_plan_s ks_list_plan;
int ks_list_plan_create(ks_list_plan **plan, int ntask, const ks_t *kernels, int k,
                        const long *a_off, const int *umap, const int *amap,
                        const long *b_off, const int *bmap, const int *wmap,
                        int precision);
int ks_list_plan_execute(const ks_list_plan *plan, double *u, const double *XA,
                         const double *XA2, const double *XB, const double *XB2,
                         const double *w, void *stream);
void ks_list_plan_destroy(ks_list_plan *plan);

/* Symmetric tasks of a plan.  No counterpart in the reference (its
 * omp_dgsks_list_symmetric shares the point table but evaluates every pair,
 * frame/omp_dgsks_list.cpp:114-166).  A task whose target list is its source
 * list (m == n >= 128, amap[a_off[t] + i] == bmap[b_off[t] + i] for every i;
 * for the variable bandwidth also hi == hj), at k = 3 or 8 (where the vector
 * units evaluate the pairs; at larger k the matrix-core layout of every pair
 * is faster), evaluates each pair of its different 64-row blocks once for
 * both terms in runs whose target and source tables are one (XA == XB and
 * XA2 == XB2); umap and wmap may be anything.  Other runs use the full layout
 * the plan keeps beside it.  dgsks_list does the same when its XA == XB and
 * XA2 == XB2.  Returns how many tasks of the plan are symmetric. */
int ks_list_plan_symmetric_tasks(const ks_list_plan *plan);

/* Tile shape of the batched launch for tasks of kernel `type` at dimension k
 * and `precision`: *slice sources per 64-row tile, *tile = 1 for FP64
 * matrix-core tiles, 0 for one row per lane.  Tuned on first use (timing every
 * candidate on a synthetic batch of that type) and cached for the process and,
 * when the environment variable KS_TUNE_CACHE names a file, across processes;
 * KS_AUTOTUNE=0 returns the default shape without tuning.  No counterpart in
 * the reference (its blocking is fixed at compile time,
 * micro_kernel/x86_64/sandybridge/gsks_config.h:3-5). */
int ks_list_autotune(int type, int k, int precision, int *slice, int *tile);

#ifdef __cplusplus
}
#endif

#endif /* __KS_H__ */
