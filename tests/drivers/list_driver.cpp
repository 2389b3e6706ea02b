// list_driver.cpp -- runs the reference's C++ batched interface of
// ks_amd/libgsks.so from a C++ program, as a program written for the reference
// would: all five overloads of include/omp_dgsks_list.hpp:48-115
// (frame/omp_dgsks_list.cpp:55-420), including the separated-u symmetric one
// that the reference's own driver calls (test/test_dgsks_list.cpp:193).
//
// Tasks are built the way test/test_dgsks_list.cpp:75-150 builds them: each
// task's target and source lists are prefixes of random permutations of the
// point pool, of 512 + (random below `range`) points each, over coordinates
// (integer in [0,100)) / 1000.  Weights are drawn in [0.5, 1) and the weight
// lists are other permutations than the source lists, so a gather through the
// wrong map shows; the separated-u variants scatter through yet another
// permutation.  u starts nonzero (the calls accumulate).
//
// The checker is the CPU oracle's batched restatement (orc_dgsks_list in
// oracle/gsks_oracle.c, per task the reference's dgsks_ref, merged as
// frame/omp_dgsks_list.cpp:407-414 merges u_local).  Bars: relative L2 <= 1e-12
// (test/test_dgsks.c:43-79) and the reference driver's per-element check,
// |u - u_ref| <= 1e-10 |u_ref| (test/test_dgsks_list.cpp:241-247); every value
// is positive here, so the per-element check is well posed.  Also checks the
// Laplace descriptor side effect.
//
// usage: list_driver k ntask [seed]   -- exits 1 if any check fails
// Test infrastructure: built by tests/drivers/Makefile, run by
// tests/test_gpu_drivers.py.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gsks_oracle.h"
#include "omp_dgsks_list.hpp"

namespace {

using VV = std::vector<std::vector<int>>;

uint64_t g_rng = 1;

uint32_t next_u32() {
  g_rng = g_rng * 6364136223846793005ull + 1442695040888963407ull;
  return (uint32_t)(g_rng >> 32);
}

int below(int n) { return (int)(next_u32() % (uint32_t)n); }

std::vector<int> permutation(int n) {
  std::vector<int> p(n);
  for (int i = 0; i < n; ++i) p[i] = i;
  for (int i = n - 1; i > 0; --i) std::swap(p[i], p[below(i + 1)]);
  return p;
}

std::vector<int> prefix(int n, int len) {
  std::vector<int> p = permutation(n);
  p.resize(len);
  return p;
}

void flatten(const VV &lists, std::vector<long> &off, std::vector<int> &flat) {
  off.assign(lists.size() + 1, 0);
  flat.clear();
  for (size_t t = 0; t < lists.size(); ++t) {
    flat.insert(flat.end(), lists[t].begin(), lists[t].end());
    off[t + 1] = (long)flat.size();
  }
}

int g_failures = 0;

void check(const char *entry, const char *kname, int k, const std::vector<double> &u,
           const std::vector<double> &ref) {
  double num = 0.0, den = 0.0;
  long bad = 0;
  for (size_t i = 0; i < u.size(); ++i) {
    double d = u[i] - ref[i];
    num += d * d;
    den += ref[i] * ref[i];
    if (std::fabs(d) > 1e-10 * std::fabs(ref[i])) ++bad;
  }
  const double err = den > 0.0 ? std::sqrt(num / den) : std::sqrt(num);
  const bool ok = err <= 1e-12 && bad == 0;
  printf("%s %-42s %-10s k=%d rel L2 = %.3e, entries off by > 1e-10: %ld\n", ok ? "PASS" : "FAIL",
         entry, kname, k, err, bad);
  fflush(stdout);
  if (!ok) ++g_failures;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s k ntask [seed]\n", argv[0]);
    return 2;
  }
  const int k = atoi(argv[1]), ntask = atoi(argv[2]);
  g_rng = argc > 3 ? strtoull(argv[3], nullptr, 10) : 1;
  const int nxa = 4096 * 5, nxb = 12000, range = 1024;

  std::vector<double> XA((size_t)nxa * k), XB((size_t)nxb * k), XA2(nxa), XB2(nxb);
  for (double &x : XA) x = (double)below(100) / 1000.0;
  for (double &x : XB) x = (double)below(100) / 1000.0;
  orc_sqnorm(XA.data(), nxa, k, XA2.data());
  orc_sqnorm(XB.data(), nxb, k, XB2.data());
  const int nw = nxa > nxb ? nxa : nxb;
  std::vector<double> w(nw), u0(nxa);
  for (double &x : w) x = 0.5 + (double)below(1000) / 2000.0;
  for (double &x : u0) x = 1.0 + (double)below(1000) / 1000.0;

  // task lists: alist over XA, blist over XA (symmetric) or XB (unsymmetric),
  // wlist another permutation prefix of the weights, ulist another of u
  VV alist(ntask), ulist(ntask), blist_a(ntask), blist_b(ntask), wlist(ntask);
  for (int t = 0; t < ntask; ++t) {
    const int na = 512 + below(range), nb = 512 + below(range);
    alist[t] = prefix(nxa, na);
    ulist[t] = prefix(nxa, na);
    blist_a[t] = prefix(nxa, nb);
    blist_b[t] = prefix(nxb, nb);
    wlist[t] = prefix(nw, nb);
  }

  struct KCase {
    const char *name;
    ks_type type;
    double scal, cons, powe;
  };
  const KCase cases[] = {{"Gaussian", KS_GAUSSIAN, -0.5, 0.0, 0.0},  // test/test_dgsks_list.cpp:176
                         {"Laplace", KS_LAPLACE, 0.0, 0.0, 0.0},
                         {"Polynomial", KS_POLYNOMIAL, 1.0 / k, 1.0, 2.0}};
  for (const KCase &kc : cases) {
    if (kc.type == KS_LAPLACE && k < 3) continue;
    for (int variant = 0; variant < 5; ++variant) {
      static const char *names[] = {
          "omp_dgsks_list_unsymmetric", "omp_dgsks_list_symmetric",
          "omp_dgsks_list_separated_u_unsymmetric", "omp_dgsks_list_separated_u_symmetric",
          "omp_dgsks_list"};
      const bool sym = variant == 1 || variant == 3;
      const bool sep_u = variant >= 2;
      VV &blist = sym ? blist_a : blist_b;
      VV &ul = sep_u ? ulist : alist;
      double *xb = sym ? XA.data() : XB.data();
      double *xb2 = sym ? XA2.data() : XB2.data();
      ks_t kernel;
      std::memset(&kernel, 0, sizeof(kernel));
      kernel.type = kc.type;
      kernel.scal = kc.scal;
      kernel.cons = kc.cons;
      kernel.powe = kc.powe;
      std::vector<double> u = u0;
      switch (variant) {
        case 0:
          omp_dgsks_list_unsymmetric(&kernel, k, u, nxa, XA.data(), alist, nxb, XB.data(), blist,
                                     w.data(), wlist);
          break;
        case 1:
          omp_dgsks_list_symmetric(&kernel, k, u, XA.data(), nxa, alist, blist, w.data(), wlist);
          break;
        case 2:
          omp_dgsks_list_separated_u_unsymmetric(&kernel, k, u, ulist, nxa, XA.data(), alist, nxb,
                                                 XB.data(), blist, w.data(), wlist);
          break;
        case 3:
          omp_dgsks_list_separated_u_symmetric(&kernel, k, u, ulist, XA.data(), nxa, alist, blist,
                                               w.data(), wlist);
          break;
        default:
          omp_dgsks_list(&kernel, k, u, ulist, XA.data(), XA2.data(), alist, xb, xb2, blist,
                         w.data(), wlist);
          break;
      }
      int code = ks_last_error();
      if (code != KS_OK) {
        printf("FAIL %s %s: ks_last_error() = %d (%s)\n", names[variant], kc.name, code,
               ks_error_string(code));
        ++g_failures;
        continue;
      }
      // the oracle over the flattened batch
      std::vector<long> aoff, boff, tmp;
      std::vector<int> amap, umap, bmap, wmap;
      flatten(alist, aoff, amap);
      flatten(ul, tmp, umap);
      flatten(blist, boff, bmap);
      flatten(wlist, tmp, wmap);
      orc_kernel ok;
      std::memset(&ok, 0, sizeof(ok));
      ok.type = (int)kc.type;
      ok.scal = kc.scal;
      ok.cons = kc.cons;
      ok.powe = kc.powe;
      std::vector<orc_kernel> oks(ntask, ok);
      std::vector<double> ref = u0;
      orc_dgsks_list(ntask, oks.data(), k, ref.data(), (long)ref.size(), aoff.data(), umap.data(),
                     amap.data(), boff.data(), bmap.data(), wmap.data(), XA.data(), XA2.data(),
                     xb, xb2, w.data(), 8);
      check(names[variant], kc.name, k, u, ref);
      if (kc.type == KS_LAPLACE) {
        double pw, sc;
        orc_laplace_params(k, &pw, &sc);
        if (kernel.powe != pw || std::fabs(kernel.scal - sc) > 1e-15 * std::fabs(sc)) {
          printf("FAIL %s Laplace: descriptor not updated\n", names[variant]);
          ++g_failures;
        }
      }
    }
  }
  printf("%s: %d failure(s)\n", g_failures ? "FAILED" : "OK", g_failures);
  return g_failures ? 1 : 0;
}
