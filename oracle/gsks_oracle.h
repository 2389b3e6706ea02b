/*
 * CPU oracle for the fused kernel summation  u[umap[i]] += sum_j K(a_i, b_j) w[wmap[j]].
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the CPU baseline; the product library (ks_amd/libgsks.so) never calls it.
 *
 * The arithmetic is a plain C restatement of the reference oracle dgsks_ref
 * (frame/dgsks_ref.c:94-535): per pair the same operation order (gather,
 * sequential rank-k sum, scale by -2 or +1, + |a|^2, + |b|^2, kernel function,
 * then a sequential GEMV over j starting from the current u), compiled without
 * FMA contraction so that it reproduces the reference's naive-loop build.  It
 * never materialises the m x n matrix, and it can evaluate any subset of the
 * target rows (row-subsampled checks of the large shapes).
 *
 * Pinning (what checks this restatement): the compiled reference
 * (oracle/_ref, built from the reference sources by oracle/Makefile.ref), the
 * golden fixtures under tests/golden/ that were written from it, an
 * independent numpy restatement, mpmath brute force on tiny inputs, and closed
 * forms (multiquadric, integer polynomial, scal = 0) -- see tests/test_oracle.py.
 * The gradients (orc_gauss_grad, orc_ks_grad; the reference has none) are
 * pinned by central differences of this oracle's own forward, with repeated
 * map entries and one table on both sides, and orc_ks_grad with identity maps
 * against orc_gauss_grad (tests/test_oracle.py).
 */
#ifndef GSKS_ORACLE_H
#define GSKS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel type codes: the order of ks_type in include/ks.h:49-58 of the reference. */
enum {
  ORC_GAUSSIAN = 0,
  ORC_POLYNOMIAL = 1,
  ORC_LAPLACE = 2,
  ORC_GAUSSIAN_VAR_BANDWIDTH = 3,
  ORC_TANH = 4,
  ORC_QUARTIC = 5,
  ORC_MULTIQUADRATIC = 6,
  ORC_EPANECHNIKOV = 7
};

typedef struct {
  int type;
  double powe;
  double scal;
  double cons;
  const double *hi; /* indexed by global point id (VAR_BANDWIDTH only) */
  const double *hj;
  /* 0: dgsks_ref's semantics (no clamp).  1: the fused micro-kernels' --
   * d^2 = max(d^2, 0) before the Gaussian, variable-bandwidth, quartic,
   * multiquadric and Epanechnikov kernels
   * (micro_kernel/x86_64/haswell/sq2nrm_int_d8x6.h:64-78), which is what
   * dgsks() computes; Laplace is unaffected (d^2 < 1e-15 is replaced either
   * way). */
  int clamp;
} orc_kernel;

/* Laplace parameters the reference writes into ks_t (frame/dgsks_ref.c:146-148). */
void orc_laplace_params(int k, double *powe, double *scal);

/* X2[i] = sum_p X[i*k+p]^2, left to right (frame/omp_dgsks_list.cpp:75-81). */
void orc_sqnorm(const double *X, long nx, int k, double *X2);

/* K value of one pair, given the rank-k sum c = a.b (frame/dgsks_ref.c:253-424).
 * hi_i / hj_j are only read for VAR_BANDWIDTH. */
double orc_pair(const orc_kernel *ker, double c, double aa, double bb,
                double hi_i, double hj_j);

/* out[r] = u[umap[rows[r]]] + sum_{j=0..n-1} K(rows[r], j) * w[wmap[j]], with the
 * sum taken left to right exactly as the reference GEMV (frame/dgsks_ref.c:437-444).
 * u is read, not written.  The kernel descriptor is used as given (the caller
 * applies orc_laplace_params first).  Rows are independent; nthreads > 1 splits
 * them over OpenMP threads without changing any result. */
void orc_dgsks_rows(const orc_kernel *ker, int m, int n, int k,
                    const double *u, const int *umap,
                    const double *XA, const double *XA2, const int *amap,
                    const double *XB, const double *XB2, const int *bmap,
                    const double *w, const int *wmap,
                    int nrows, const int *rows, double *out, int nthreads);

/* Whole call with the reference's semantics (frame/dgsks_ref.c:94-535): sets the
 * Laplace parameters in *ker, gathers u[umap], accumulates, and scatters back in
 * index order (a duplicate umap entry keeps the last row's value). */
void orc_dgsks_ref(orc_kernel *ker, int m, int n, int k,
                   double *u, const int *umap,
                   const double *XA, const double *XA2, const int *amap,
                   const double *XB, const double *XB2, const int *bmap,
                   const double *w, const int *wmap, int nthreads);

/* Batched summation with the semantics of omp_dgsks_list
 * (frame/omp_dgsks_list.cpp:297-420), generalised to one kernel descriptor per
 * task: u[x] += sum over tasks t of the task's contributions to row x.  Task t
 * uses the index ranges [off_a[t], off_a[t+1]) of amap/umap and
 * [off_b[t], off_b[t+1]) of bmap/wmap.  Laplace parameters are applied to a
 * copy of each task's descriptor.  As in the reference, tasks are dealt to
 * nthreads workers by least load (deterministically), each worker sums into a
 * u_local of nu doubles, and the u_local are merged into u in worker order. */
void orc_dgsks_list(int ntask, const orc_kernel *kers, int k, double *u, long nu,
                    const long *off_a, const int *umap, const int *amap,
                    const long *off_b, const int *bmap, const int *wmap,
                    const double *XA, const double *XA2,
                    const double *XB, const double *XB2,
                    const double *w, int nthreads);

/* FP64 reference gradient of L = sum_i g[i] * u_i for the Gaussian summation
 * u_i = sum_j K_ij w_j with identity maps, K_ij = exp(-d2_ij / (2 h^2)) and
 * d2 assembled as |a|^2 + |b|^2 - 2 a.b exactly as the forward.  Output rows
 * can be restricted to subsets (rows_a for dXA, rows_b for dXB and dw;
 * NULL = all, outputs are indexed by position in the subset):
 *   dw[j]    = sum_i g_i K_ij
 *   dXA[i,:] = sum_j g_i w_j K_ij * (-(a_i - b_j) / h^2)
 *   dXB[j,:] = sum_i g_i w_j K_ij * ( (a_i - b_j) / h^2)
 *   dh       = sum_ij g_i w_j K_ij * d2_ij / h^3      (skipped when dh == NULL) */
void orc_gauss_grad(int m, int n, int k, const double *XA, const double *XB,
                    const double *w, const double *g, double h,
                    int nra, const int *rows_a, double *dXA,
                    int nrb, const int *rows_b, double *dXB, double *dw,
                    double *dh, int nthreads);

/* Gradient of L = sum_x g[x] du[x] for the mapped summation
 *   du[umap[i]] += sum_j K(XA[amap[i]], XB[bmap[j]]) w[wmap[j]]
 * (frame/dgsks_ref.c:94-535's loops) with the Gaussian kernel (ORC_GAUSSIAN,
 * K = exp(scal d2), :254-270) or the variable-bandwidth Gaussian
 * (ORC_GAUSSIAN_VAR_BANDWIDTH, K = exp(-0.5 hi[amap[i]] hj[bmap[j]] d2),
 * :271-288), where the norms in d2 are those of the points themselves.  The
 * reference has no gradient; this is the chain rule applied pair by pair:
 * with c = scal (Gaussian) or -0.5 hi hj (variable bandwidth),
 *   dK/da = 2 c K (a - b),  dK/db = -2 c K (a - b),  dK/dscal = K d2,
 *   dK/dhi = -0.5 hj d2 K,  dK/dhj = -0.5 hi d2 K,  d2 = sum_p (a_p - b_p)^2,
 * K itself from orc_pair.  Every output is a table that ACCUMULATES (+=) at
 * the mapped entries -- dXA [.][k] at amap[i], dXB at bmap[j], dw at wmap[j],
 * dhi at amap[i], dhj at bmap[j], *dscal once -- a repeated map entry adding
 * each of its rows' terms, in row order.  Any output may be NULL; dscal is
 * for the Gaussian, dhi / dhj for the variable bandwidth. */
void orc_ks_grad(const orc_kernel *ker, int m, int n, int k, const double *g, const int *umap,
                 const double *XA, const int *amap, const double *XB, const int *bmap,
                 const double *w, const int *wmap, double *dXA, double *dXB, double *dw,
                 double *dscal, double *dhi, double *dhj, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
