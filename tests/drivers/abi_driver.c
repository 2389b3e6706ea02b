/*
 * abi_driver.c -- runs the reference-facing C ABI of ks_amd/libgsks.so from a
 * plain C program, as a program written for the reference would: dgsks,
 * dgsks_ref (include/ks.h:73-105), dgsks_wrapper, dgsks_ref_wrapper
 * (frame/dgsks.c:783-826, frame/dgsks_ref.c:537-580) and ks_malloc_aligned
 * (include/ks.h:107-111), on host arrays, for all eight kernel types.
 *
 * Data convention of the reference's own driver (test/test_dgsks.c:144-207,
 * 325-355): one point table used as both XA and XB, coordinates
 * (integer in [0,100)) / 1000, weights (integer in [0,1000)) / 1000, norms
 * summed left to right, the default kernel parameters per type, and for the
 * variable bandwidth hi = hj = -1 / (2 (1 + 0.5 / (1 + exp(-|x|^2)))^2).
 * Two map layouts: the driver's identity maps into a zero u, and random maps
 * (umap a permutation of other targets than amap) into a nonzero u.
 *
 * The checker is the CPU oracle (oracle/gsks_oracle.c, a restatement of
 * frame/dgsks_ref.c); the error metric is the reference's relative L2 norm
 * (test/test_dgsks.c:43-79).  Bars: 1e-12 for dgsks and its wrapper, 1e-13
 * for the reference algorithm (dgsks_ref, unfused, on the GPU).  Also checks
 * the Laplace descriptor side effect (frame/dgsks.c:445-447).
 *
 * usage: abi_driver m n k [seed]   -- exits 1 if any check fails
 * Test infrastructure: built by tests/drivers/Makefile, run by
 * tests/test_gpu_drivers.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gsks_oracle.h"
#include "ks.h"

static uint64_t g_rng;

static uint32_t next_u32(void) {
  /* 64-bit LCG (Knuth's MMIX constants), high half */
  g_rng = g_rng * 6364136223846793005ull + 1442695040888963407ull;
  return (uint32_t)(g_rng >> 32);
}

static int below(int n) { return (int)(next_u32() % (uint32_t)n); }

static void shuffle(int *x, int n) {
  for (int i = n - 1; i > 0; --i) {
    int j = below(i + 1), t = x[i];
    x[i] = x[j];
    x[j] = t;
  }
}

/* relative L2 of x against ref over the entries idx (test/test_dgsks.c:43-79),
   both scaled by max |ref| first: Laplace at large k reaches ~1e230, whose
   squares would overflow */
static double rel_l2(const double *x, const double *ref, int n, const int *idx) {
  double sc = 0.0, num = 0.0, den = 0.0;
  for (int i = 0; i < n; ++i) sc = fmax(sc, fabs(ref[idx ? idx[i] : i]));
  if (sc == 0.0 || !isfinite(sc)) sc = 1.0;
  for (int i = 0; i < n; ++i) {
    int r = idx ? idx[i] : i;
    double d = x[r] / sc - ref[r] / sc;
    num += d * d;
    den += (ref[r] / sc) * (ref[r] / sc);
  }
  return den > 0.0 ? sqrt(num / den) : sqrt(num);
}

struct kcase {
  const char *name;
  ks_type type;
  double scal, cons, powe;
};

/* test/test_dgsks.c:325-355 */
static const struct kcase kCases[] = {
    {"Gaussian", KS_GAUSSIAN, -0.5, 0.0, 0.0},
    {"Polynomial", KS_POLYNOMIAL, 0.1, 0.1, 4.0},
    {"Laplace", KS_LAPLACE, 0.0, 0.0, 0.0},
    {"Var_bandwidth", KS_GAUSSIAN_VAR_BANDWIDTH, 0.0, 0.0, 0.0},
    {"Tanh", KS_TANH, 0.1, 0.1, 0.0},
    {"Quartic", KS_QUARTIC, 0.0, 0.0, 0.0},
    {"Multiquadratic", KS_MULTIQUADRATIC, 0.0, 1.0, 0.0},
    {"Epanechnikov", KS_EPANECHNIKOV, 0.0, 0.0, 0.0},
};

static int g_failures = 0;

static void report(const char *entry, const char *kname, const char *maps, int m, int n, int k,
                   double err, double bar) {
  int ok = err <= bar;
  printf("%s %-17s %-13s %-8s m=%d n=%d k=%d rel L2 = %.3e (bar %.0e)\n", ok ? "PASS" : "FAIL",
         entry, kname, maps, m, n, k, err, bar);
  fflush(stdout);
  if (!ok) ++g_failures;
}

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s m n k [seed]\n", argv[0]);
    return 2;
  }
  const int m = atoi(argv[1]), n = atoi(argv[2]), k = atoi(argv[3]);
  g_rng = argc > 4 ? strtoull(argv[4], NULL, 10) : 1;
  const int nx = (m > n ? m : n) + 1000; /* a pool larger than either side */

  /* host arrays through the library's own aligned allocator */
  double *X = ks_malloc_aligned(k, nx, sizeof(double));
  double *X2 = ks_malloc_aligned(1, nx, sizeof(double));
  double *w = ks_malloc_aligned(1, nx, sizeof(double));
  double *hv = ks_malloc_aligned(1, nx, sizeof(double));
  double *u0 = ks_malloc_aligned(1, nx, sizeof(double));
  double *u = ks_malloc_aligned(1, nx, sizeof(double));
  double *ref = ks_malloc_aligned(1, nx, sizeof(double));
  int *amap = (int *)malloc(sizeof(int) * m), *umap = (int *)malloc(sizeof(int) * m);
  int *bmap = (int *)malloc(sizeof(int) * n), *wmap = (int *)malloc(sizeof(int) * n);
  int *perm = (int *)malloc(sizeof(int) * nx);
  char *written = (char *)malloc(nx);
  if (((uintptr_t)X & 63) || ((uintptr_t)u & 63)) {
    printf("FAIL ks_malloc_aligned: not 64-byte aligned\n");
    return 1;
  }
  for (long i = 0; i < (long)nx * k; ++i) X[i] = (double)below(100) / 1000.0;
  for (int i = 0; i < nx; ++i) {
    double t = 0.0;
    for (int p = 0; p < k; ++p) t += X[(long)i * k + p] * X[(long)i * k + p];
    X2[i] = t;
    w[i] = (double)below(1000) / 1000.0;
    double h = 1.0 + 0.5 / (1.0 + exp(-1.0 * X2[i]));
    hv[i] = -1.0 / (2.0 * h * h);
  }

  for (int layout = 0; layout < 2; ++layout) {
    const char *maps = layout == 0 ? "identity" : "random";
    if (layout == 0) {
      for (int i = 0; i < m; ++i) amap[i] = umap[i] = i;
      for (int j = 0; j < n; ++j) bmap[j] = wmap[j] = j;
    } else {
      for (int i = 0; i < nx; ++i) perm[i] = i;
      shuffle(perm, nx);
      for (int i = 0; i < m; ++i) amap[i] = perm[i];
      shuffle(perm, nx);
      for (int i = 0; i < m; ++i) umap[i] = perm[i];
      shuffle(perm, nx);
      for (int j = 0; j < n; ++j) bmap[j] = perm[j];
      shuffle(perm, nx);
      for (int j = 0; j < n; ++j) wmap[j] = perm[j];
    }
    for (int i = 0; i < nx; ++i) u0[i] = layout == 0 ? 0.0 : (double)below(1000) / 1000.0 - 0.5;
    memset(written, 0, nx);
    for (int i = 0; i < m; ++i) written[umap[i]] = 1;

    for (size_t c = 0; c < sizeof(kCases) / sizeof(kCases[0]); ++c) {
      const struct kcase *kc = &kCases[c];
      if (kc->type == KS_LAPLACE && k < 3) continue; /* the reference only warns there */
      const int varbw = kc->type == KS_GAUSSIAN_VAR_BANDWIDTH;
      /* expected: the oracle's dgsks_ref on a copy of u0 */
      orc_kernel ok;
      memset(&ok, 0, sizeof(ok));
      ok.type = (int)kc->type;
      ok.scal = kc->scal;
      ok.cons = kc->cons;
      ok.powe = kc->powe;
      ok.hi = varbw ? hv : NULL;
      ok.hj = varbw ? hv : NULL;
      memcpy(ref, u0, sizeof(double) * nx);
      orc_dgsks_ref(&ok, m, n, k, ref, umap, X, X2, amap, X, X2, bmap, w, wmap, 8);

      for (int entry = 0; entry < 4; ++entry) {
        static const char *names[] = {"dgsks", "dgsks_ref", "dgsks_wrapper", "dgsks_ref_wrapper"};
        memcpy(u, u0, sizeof(double) * nx);
        ks_t kernel;
        memset(&kernel, 0, sizeof(kernel));
        kernel.type = kc->type;
        kernel.scal = kc->scal;
        kernel.cons = kc->cons;
        kernel.powe = kc->powe;
        kernel.hi = varbw ? hv : NULL;
        kernel.hj = varbw ? hv : NULL;
        switch (entry) {
          case 0:
            dgsks(&kernel, m, n, k, u, umap, X, X2, amap, X, X2, bmap, w, wmap);
            break;
          case 1:
            dgsks_ref(&kernel, m, n, k, u, umap, X, X2, amap, X, X2, bmap, w, wmap);
            break;
          case 2:
            dgsks_wrapper(m, n, k, u, umap, X, X2, amap, X, X2, bmap, w, wmap, (int)kc->type,
                          kc->scal, kc->cons, kc->powe, varbw ? hv : NULL);
            break;
          default:
            dgsks_ref_wrapper(m, n, k, u, umap, X, X2, amap, X, X2, bmap, w, wmap,
                              (int)kc->type, kc->scal, kc->cons, kc->powe, varbw ? hv : NULL);
            break;
        }
        int code = ks_last_error();
        if (code != KS_OK) {
          printf("FAIL %s %s: ks_last_error() = %d (%s)\n", names[entry], kc->name, code,
                 ks_error_string(code));
          ++g_failures;
          continue;
        }
        /* targets: the rows the call wrote (umap), then the whole of u (untouched
           entries must keep their values) */
        const double bar = (entry & 1) ? 1e-13 : 1e-12;
        report(names[entry], kc->name, maps, m, n, k, rel_l2(u, ref, m, umap), bar);
        int changed = 0;
        for (int i = 0; i < nx; ++i) changed += !written[i] && u[i] != u0[i];
        if (changed) {
          printf("FAIL %s %s %s: %d entries outside umap changed\n", names[entry], kc->name,
                 maps, changed);
          ++g_failures;
        }
        if (kc->type == KS_LAPLACE && entry < 2) {
          double pw, sc;
          orc_laplace_params(k, &pw, &sc);
          if (kernel.powe != pw || fabs(kernel.scal - sc) > 1e-15 * fabs(sc)) {
            printf("FAIL %s Laplace: descriptor not updated (powe %g scal %g, want %g %g)\n",
                   names[entry], kernel.powe, kernel.scal, pw, sc);
            ++g_failures;
          }
        }
      }
    }
  }
  printf("%s: %d failure(s)\n", g_failures ? "FAILED" : "OK", g_failures);
  free(X);
  free(X2);
  free(w);
  free(hv);
  free(u0);
  free(u);
  free(ref);
  free(amap);
  free(umap);
  free(bmap);
  free(wmap);
  free(perm);
  free(written);
  return g_failures ? 1 : 0;
}
