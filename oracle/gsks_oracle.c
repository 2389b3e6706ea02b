/*
 * CPU oracle: a plain C restatement of the reference oracle dgsks_ref
 * (frame/dgsks_ref.c:94-535).  TEST INFRASTRUCTURE ONLY -- the header
 * gsks_oracle.h states who may use it and what pins it.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp -shared -fPIC (oracle/Makefile).
 * -ffp-contract=off matters: the reference's naive loops compile to separate
 * multiply and add (no FMA under -mavx), and the restatement keeps that.
 */
#include "gsks_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* frame/dgsks_ref.c:141-148: powe = (2 - k) / 2, scal = Gamma(k/2 + 1) / (k (k - 2) pi^(k/2)). */
void orc_laplace_params(int k, double *powe, double *scal) {
  *powe = 0.5 * (2.0 - (double)k);
  *scal = tgamma(0.5 * k + 1.0) / ((double)k * (double)(k - 2) * pow(M_PI, 0.5 * k));
}

/* frame/omp_dgsks_list.cpp:75-81 (and test/test_dgsks.c:179-185). */
void orc_sqnorm(const double *X, long nx, int k, double *X2) {
  for (long i = 0; i < nx; i++) {
    double t = 0.0;
    for (int p = 0; p < k; p++) t += X[i * k + p] * X[i * k + p];
    X2[i] = t;
  }
}

/* rank_k_scale per type: frame/dgsks_ref.c:121-173. */
static double rank_k_scale(int type) {
  return (type == ORC_POLYNOMIAL || type == ORC_TANH) ? 1.0 : -2.0;
}

/* The elementwise kernel block of frame/dgsks_ref.c:253-424, one pair at a time,
 * in the reference's operation order.  c is the plain rank-k sum a.b. */
double orc_pair(const orc_kernel *ker, double c, double aa, double bb, double hi_i,
                double hj_j) {
  double C = c * rank_k_scale(ker->type); /* :236-243 */
  switch (ker->type) {
    case ORC_GAUSSIAN: /* :254-270 */
      C += aa;
      C += bb;
      if (ker->clamp && C < 0.0) C = 0.0;
      C *= ker->scal;
      return exp(C);
    case ORC_GAUSSIAN_VAR_BANDWIDTH: /* :271-288 */
      C += aa;
      C += bb;
      if (ker->clamp && C < 0.0) C = 0.0;
      C *= -0.5;
      C *= hi_i;
      C *= hj_j;
      return exp(C);
    case ORC_POLYNOMIAL: /* :296-336 */
      C *= ker->scal;
      C += ker->cons;
      if (ker->powe == 2.0) return C * C;
      if (ker->powe == 4.0) {
        C = C * C;
        return C * C;
      }
      return pow(C, ker->powe);
    case ORC_LAPLACE: /* :337-361 */
      C += aa;
      C += bb;
      if (C < 1E-15) C = 1.79E+308;
      C = pow(C, ker->powe);
      return C * ker->scal;
    case ORC_TANH: /* :362-378 */
      C *= ker->scal;
      C += ker->cons;
      return tanh(C);
    case ORC_QUARTIC: /* :379-394 */
      C += aa;
      C += bb;
      if (ker->clamp && C < 0.0) C = 0.0;
      if (C < 1.0) {
        C = 1.0 - C;
        return (15.0 / 16.0) * C * C;
      }
      return 0.0;
    case ORC_MULTIQUADRATIC: /* :395-404 */
      C += aa;
      C += bb;
      if (ker->clamp && C < 0.0) C = 0.0;
      return C + ker->cons;
    case ORC_EPANECHNIKOV: /* :405-420 */
      C += aa;
      C += bb;
      if (ker->clamp && C < 0.0) C = 0.0;
      if (C < 1.0) return (3.0 / 4.0) * (1.0 - C);
      return 0.0;
    default:
      return NAN;
  }
}

#define ORC_JB 16 /* columns per block: independent sums, each in reference order */

static void row_block(const orc_kernel *ker, int n, int k, const double *a, double aa,
                      double hi_i, const double *XB, const double *XB2, const int *bmap,
                      const double *w, const int *wmap, double *acc, double *bt) {
  for (int j0 = 0; j0 < n; j0 += ORC_JB) {
    int nb = n - j0 < ORC_JB ? n - j0 : ORC_JB;
    /* transpose the block so the rank-k loop vectorises across columns;
       each column's sum is still taken over p in order */
    for (int jj = 0; jj < nb; jj++) {
      const double *b = XB + (long)bmap[j0 + jj] * k;
      for (int p = 0; p < k; p++) bt[p * ORC_JB + jj] = b[p];
    }
    double c[ORC_JB];
    for (int jj = 0; jj < ORC_JB; jj++) c[jj] = 0.0;
    for (int p = 0; p < k; p++) {
      double ap = a[p];
      for (int jj = 0; jj < ORC_JB; jj++) c[jj] += ap * bt[p * ORC_JB + jj];
    }
    double s = *acc;
    for (int jj = 0; jj < nb; jj++) {
      int bj = bmap[j0 + jj];
      double hj = ker->hj ? ker->hj[bj] : 0.0;
      double K = orc_pair(ker, c[jj], aa, XB2[bj], hi_i, hj);
      s += K * w[wmap[j0 + jj]]; /* frame/dgsks_ref.c:442 */
    }
    *acc = s;
  }
}

void orc_dgsks_rows(const orc_kernel *ker, int m, int n, int k, const double *u,
                    const int *umap, const double *XA, const double *XA2,
                    const int *amap, const double *XB, const double *XB2,
                    const int *bmap, const double *w, const int *wmap, int nrows,
                    const int *rows, double *out, int nthreads) {
  (void)m;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
  {
    double *bt = (double *)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1) * ORC_JB);
#pragma omp for schedule(dynamic, 4)
    for (int r = 0; r < nrows; r++) {
      int i = rows[r];
      int ai = amap[i];
      double acc = u[umap[i]];
      double hi = ker->hi ? ker->hi[ai] : 0.0;
      row_block(ker, n, k, XA + (long)ai * k, XA2[ai], hi, XB, XB2, bmap, w, wmap, &acc,
                bt);
      out[r] = acc;
    }
    free(bt);
  }
}

void orc_dgsks_ref(orc_kernel *ker, int m, int n, int k, double *u, const int *umap,
                   const double *XA, const double *XA2, const int *amap,
                   const double *XB, const double *XB2, const int *bmap,
                   const double *w, const int *wmap, int nthreads) {
  if (m <= 0 || n <= 0 || k <= 0) return;
  if (ker->type == ORC_LAPLACE) orc_laplace_params(k, &ker->powe, &ker->scal);
  int *rows = (int *)malloc(sizeof(int) * m);
  double *us = (double *)malloc(sizeof(double) * m);
  for (int i = 0; i < m; i++) rows[i] = i;
  orc_dgsks_rows(ker, m, n, k, u, umap, XA, XA2, amap, XB, XB2, bmap, w, wmap, m, rows,
                 us, nthreads);
  for (int i = 0; i < m; i++) u[umap[i]] = us[i]; /* frame/dgsks_ref.c:454-459 */
  free(rows);
  free(us);
}

void orc_dgsks_list(int ntask, const orc_kernel *kers, int k, double *u, long nu,
                    const long *off_a, const int *umap, const int *amap,
                    const long *off_b, const int *bmap, const int *wmap,
                    const double *XA, const double *XA2, const double *XB,
                    const double *XB2, const double *w, int nthreads) {
  /* frame/omp_dgsks_list.cpp:330-414: every task is one dgsks_ref into a zeroed
     buffer, and the buffer is then added into u.  Here each task's row sums
     (frame/dgsks_ref.c:437-444, each from 0) are computed in parallel over
     chunks of rows of all tasks, then scattered through umap task by task, in
     order, and the buffer is added into u: the same sums, in a fixed order, so
     the result does not depend on the thread count or timing. */
  if (nthreads < 1) nthreads = 1;
  const long total_a = ntask > 0 ? off_a[ntask] : 0;
  double *R = (double *)malloc(sizeof(double) * (size_t)(total_a > 0 ? total_a : 1));
  enum { CH = 64 };
  long nchunk = 0;
  for (int t = 0; t < ntask; t++) nchunk += (off_a[t + 1] - off_a[t] + CH - 1) / CH;
  int *ct = (int *)malloc(sizeof(int) * (size_t)(nchunk > 0 ? nchunk : 1));
  long *c0 = (long *)malloc(sizeof(long) * (size_t)(nchunk > 0 ? nchunk : 1));
  long c = 0;
  for (int t = 0; t < ntask; t++)
    for (long r = off_a[t]; r < off_a[t + 1]; r += CH) {
      ct[c] = t;
      c0[c++] = r;
    }
#pragma omp parallel num_threads(nthreads)
  {
    double *bt = (double *)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1) * ORC_JB);
#pragma omp for schedule(dynamic, 1)
    for (long q = 0; q < nchunk; q++) {
      const int t = ct[q];
      const long b0 = off_b[t];
      const int n = (int)(off_b[t + 1] - b0);
      const long r1 = c0[q] + CH < off_a[t + 1] ? c0[q] + CH : off_a[t + 1];
      orc_kernel ker = kers[t];
      if (ker.type == ORC_LAPLACE) orc_laplace_params(k, &ker.powe, &ker.scal);
      for (long r = c0[q]; r < r1; r++) {
        const int ai = amap[r];
        double acc = 0.0;
        if (n > 0) {
          double hi = ker.hi ? ker.hi[ai] : 0.0;
          row_block(&ker, n, k, XA + (long)ai * k, XA2[ai], hi, XB, XB2, bmap + b0, w,
                    wmap + b0, &acc, bt);
        }
        R[r] = acc;
      }
    }
    free(bt);
  }
  double *buf = (double *)calloc((size_t)nu, sizeof(double));
  for (long r = 0; r < total_a; r++) buf[umap[r]] += R[r];
  for (long x = 0; x < nu; x++) u[x] += buf[x];
  free(buf);
  free(R);
  free(ct);
  free(c0);
}

void orc_gauss_grad(int m, int n, int k, const double *XA, const double *XB,
                    const double *w, const double *g, double h, int nra,
                    const int *rows_a, double *dXA, int nrb, const int *rows_b,
                    double *dXB, double *dw, double *dh, int nthreads) {
  double scal = -1.0 / (2.0 * h * h);
  double *XA2 = (double *)malloc(sizeof(double) * (size_t)m);
  double *XB2 = (double *)malloc(sizeof(double) * (size_t)n);
  orc_sqnorm(XA, m, k, XA2);
  orc_sqnorm(XB, n, k, XB2);
  if (nthreads < 1) nthreads = 1;
  orc_kernel ker = {ORC_GAUSSIAN, 0.0, scal, 0.0, NULL, NULL};
  int nA = rows_a ? nra : m, nB = rows_b ? nrb : n;
  /* dXA for the requested rows */
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
  for (int q = 0; q < nA; q++) {
    int r = rows_a ? rows_a[q] : q;
    const double *a = XA + (long)r * k;
    double *grad = dXA + (long)q * k;
    for (int p = 0; p < k; p++) grad[p] = 0.0;
    for (int j = 0; j < n; j++) {
      const double *b = XB + (long)j * k;
      double c = 0.0;
      for (int p = 0; p < k; p++) c += a[p] * b[p];
      double Kij = orc_pair(&ker, c, XA2[r], XB2[j], 0.0, 0.0);
      double f = -(g[r] * w[j] * Kij) / (h * h);
      for (int p = 0; p < k; p++) grad[p] += f * (a[p] - b[p]);
    }
  }
  /* dw and dXB for the requested columns */
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
  for (int q = 0; q < nB; q++) {
    int j = rows_b ? rows_b[q] : q;
    const double *b = XB + (long)j * k;
    double *grad = dXB + (long)q * k;
    for (int p = 0; p < k; p++) grad[p] = 0.0;
    double s = 0.0;
    for (int i = 0; i < m; i++) {
      const double *a = XA + (long)i * k;
      double c = 0.0;
      for (int p = 0; p < k; p++) c += a[p] * b[p];
      double Kij = orc_pair(&ker, c, XA2[i], XB2[j], 0.0, 0.0);
      s += g[i] * Kij;
      double f = g[i] * w[j] * Kij / (h * h);
      for (int p = 0; p < k; p++) grad[p] += f * (a[p] - b[p]);
    }
    dw[q] = s;
  }
  /* dL/dh over all pairs (optional: dh == NULL skips the O(mn) pass) */
  if (dh) {
    double dh_sum = 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4) reduction(+ : dh_sum)
    for (int r = 0; r < m; r++) {
      const double *a = XA + (long)r * k;
      double row = 0.0;
      for (int j = 0; j < n; j++) {
        const double *b = XB + (long)j * k;
        double c = 0.0;
        for (int p = 0; p < k; p++) c += a[p] * b[p];
        double Kij = orc_pair(&ker, c, XA2[r], XB2[j], 0.0, 0.0);
        double d2 = (c * -2.0 + XA2[r]) + XB2[j];
        row += g[r] * w[j] * Kij * d2;
      }
      dh_sum += row;
    }
    *dh = dh_sum / (h * h * h);
  }
  free(XA2);
  free(XB2);
}

/* Per-pair chain rule of orc_ks_grad (declaration in gsks_oracle.h): row
 * terms per target i and column terms per source j in two O(m n k) passes,
 * then a sequential scatter in row order. */
void orc_ks_grad(const orc_kernel *ker, int m, int n, int k, const double *g, const int *umap,
                 const double *XA, const int *amap, const double *XB, const int *bmap,
                 const double *w, const int *wmap, double *dXA, double *dXB, double *dw,
                 double *dscal, double *dhi, double *dhj, int nthreads) {
  const int var = ker->type == ORC_GAUSSIAN_VAR_BANDWIDTH;
  if (nthreads < 1) nthreads = 1;
  double *ra = (double *)calloc((size_t)m * (k + 2) + 1, sizeof(double)); /* dA | dhi | dscal */
  double *cb = (double *)calloc((size_t)n * (k + 2) + 1, sizeof(double)); /* dB | dhj | dw */
  double *aa = (double *)malloc(sizeof(double) * (size_t)m);
  double *bb = (double *)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < m; i++) orc_sqnorm(XA + (long)amap[i] * k, 1, k, aa + i);
  for (int j = 0; j < n; j++) orc_sqnorm(XB + (long)bmap[j] * k, 1, k, bb + j);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
  for (int i = 0; i < m; i++) {
    const double *a = XA + (long)amap[i] * k;
    const double gi = g[umap[i]];
    const double hi = var ? ker->hi[amap[i]] : 0.0;
    double *row = ra + (long)i * (k + 2);
    for (int j = 0; j < n; j++) {
      const double *b = XB + (long)bmap[j] * k;
      const double hj = var ? ker->hj[bmap[j]] : 0.0;
      double c = 0.0, d2 = 0.0;
      for (int p = 0; p < k; p++) {
        c += a[p] * b[p];
        d2 += (a[p] - b[p]) * (a[p] - b[p]);
      }
      const double Kw = orc_pair(ker, c, aa[i], bb[j], hi, hj) * w[wmap[j]] * gi;
      const double cf = var ? -0.5 * hi * hj : ker->scal;
      for (int p = 0; p < k; p++) row[p] += 2.0 * cf * Kw * (a[p] - b[p]);
      row[k] += var ? -0.5 * hj * d2 * Kw : 0.0;
      row[k + 1] += var ? 0.0 : d2 * Kw;
    }
  }
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
  for (int j = 0; j < n; j++) {
    const double *b = XB + (long)bmap[j] * k;
    const double wj = w[wmap[j]];
    const double hj = var ? ker->hj[bmap[j]] : 0.0;
    double *col = cb + (long)j * (k + 2);
    for (int i = 0; i < m; i++) {
      const double *a = XA + (long)amap[i] * k;
      const double hi = var ? ker->hi[amap[i]] : 0.0;
      double c = 0.0, d2 = 0.0;
      for (int p = 0; p < k; p++) {
        c += a[p] * b[p];
        d2 += (a[p] - b[p]) * (a[p] - b[p]);
      }
      const double Kg = orc_pair(ker, c, aa[i], bb[j], hi, hj) * g[umap[i]];
      const double cf = var ? -0.5 * hi * hj : ker->scal;
      for (int p = 0; p < k; p++) col[p] -= 2.0 * cf * Kg * wj * (a[p] - b[p]);
      col[k] += var ? -0.5 * hi * d2 * Kg * wj : 0.0;
      col[k + 1] += Kg;
    }
  }
  double s = 0.0;
  for (int i = 0; i < m; i++) {
    const double *row = ra + (long)i * (k + 2);
    if (dXA)
      for (int p = 0; p < k; p++) dXA[(long)amap[i] * k + p] += row[p];
    if (dhi) dhi[amap[i]] += row[k];
    s += row[k + 1];
  }
  if (dscal) *dscal += s;
  for (int j = 0; j < n; j++) {
    const double *col = cb + (long)j * (k + 2);
    if (dXB)
      for (int p = 0; p < k; p++) dXB[(long)bmap[j] * k + p] += col[p];
    if (dhj) dhj[bmap[j]] += col[k];
    if (dw) dw[wmap[j]] += col[k + 1];
  }
  free(ra);
  free(cb);
  free(aa);
  free(bb);
}
