// C++ batched interface of include/omp_dgsks_list.hpp (reference
// frame/omp_dgsks_list.cpp:55-420): each overload flattens its lists into
// offset arrays and runs the whole batch through list_run (one GPU pass).
#include "../../include/omp_dgsks_list.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "host_util.h"
#include "list_launch.h"

namespace {

using VV = std::vector<std::vector<int>>;

void flatten(const VV &lists, std::vector<long> &off, std::vector<int> &flat) {
  off.assign(lists.size() + 1, 0);
  for (size_t t = 0; t < lists.size(); ++t) off[t + 1] = off[t] + (long)lists[t].size();
  flat.resize(off.back());
  for (size_t t = 0; t < lists.size(); ++t)
    for (size_t i = 0; i < lists[t].size(); ++i) flat[off[t] + i] = lists[t][i];
}

// Common body of the five overloads.  XA2 / XB2 may be NULL: the norms are then
// computed on the GPU (the reference computes them on the host,
// frame/omp_dgsks_list.cpp:75-92).
void run(ks_t *kernel, int k, std::vector<double> &u, VV &ulist, double *XA, double *XA2,
         long nxa, VV &alist, double *XB, double *XB2, long nxb, VV &blist, double *w,
         VV &wlist) {
  // frame/omp_dgsks_list.cpp:315-324
  if (alist.size() == 0 || blist.size() == 0) return;
  if (alist.size() != blist.size() || blist.size() != wlist.size()) {
    printf("omp_dgsks_list(): alist, blist and wlist must have the same sizes.\n");
    exit(1);
  }
  if (ulist.size() != alist.size()) {
    printf("omp_dgsks_list(): ulist and alist must have the same sizes.\n");
    exit(1);
  }
  for (size_t t = 0; t < alist.size(); ++t) {
    if (ulist[t].size() != alist[t].size() || wlist[t].size() != blist[t].size()) {
      printf("omp_dgsks_list(): task %zu: map sizes differ.\n", t);
      exit(1);
    }
  }
  const int ntask = (int)alist.size();
  std::vector<long> aoff, boff, tmp;
  std::vector<int> amap, umap, bmap, wmap;
  flatten(alist, aoff, amap);
  flatten(ulist, tmp, umap);
  flatten(blist, boff, bmap);
  flatten(wlist, tmp, wmap);
  std::vector<ks_t> kernels(ntask, *kernel);
  int rc = gsks::list_run(ntask, kernels.data(), k, u.data(), (long)u.size(), aoff.data(),
                          umap.data(), amap.data(), boff.data(), bmap.data(), wmap.data(), XA,
                          XA2, nxa, XB, XB2, nxb, w, KS_PREC_FP64, "omp_dgsks_list");
  if (rc == KS_OK && gsks::ker_type(kernel) == KS_LAPLACE)
    gsks::laplace_params(k, &kernel->powe, &kernel->scal);
}

}  // namespace

void omp_dgsks_list_unsymmetric(ks_t *kernel, int k, std::vector<double> &u, int nxa,
                                double *XA, VV &alist, int nxb, double *XB, VV &blist,
                                double *w, VV &wlist) {
  run(kernel, k, u, alist, XA, nullptr, nxa, alist, XB, nullptr, nxb, blist, w, wlist);
}

void omp_dgsks_list_symmetric(ks_t *kernel, int k, std::vector<double> &u, double *XA, int nxa,
                              VV &alist, VV &blist, double *w, VV &wlist) {
  run(kernel, k, u, alist, XA, nullptr, nxa, alist, XA, nullptr, nxa, blist, w, wlist);
}

void omp_dgsks_list_separated_u_unsymmetric(ks_t *kernel, int k, std::vector<double> &u,
                                            VV &ulist, int nxa, double *XA, VV &alist, int nxb,
                                            double *XB, VV &blist, double *w, VV &wlist) {
  run(kernel, k, u, ulist, XA, nullptr, nxa, alist, XB, nullptr, nxb, blist, w, wlist);
}

void omp_dgsks_list_separated_u_symmetric(ks_t *kernel, int k, std::vector<double> &u,
                                          VV &ulist, double *XA, int nxa, VV &alist, VV &blist,
                                          double *w, VV &wlist) {
  run(kernel, k, u, ulist, XA, nullptr, nxa, alist, XA, nullptr, nxa, blist, w, wlist);
}

void omp_dgsks_list(ks_t *kernel, int k, std::vector<double> &u, VV &ulist, double *XA,
                    double *XA2, VV &alist, double *XB, double *XB2, VV &blist, double *w,
                    VV &wlist) {
  run(kernel, k, u, ulist, XA, XA2, 0, alist, XB, XB2, 0, blist, w, wlist);
}
