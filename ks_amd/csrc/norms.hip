// Squared row norms on the device: X2[i] = sum_p X[i*k+p]^2, summed left to
// right like the reference's norm precompute (frame/omp_dgsks_list.cpp:75-81).
#include <hip/hip_runtime.h>

#include "gsks_launch.h"
#include "host_util.h"

#pragma clang fp contract(off)

namespace gsks {

__global__ void sqnorm_kernel(const double *X, long nx, int k, double *X2) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  const double *x = X + i * k;
  double t = 0.0;
  for (int p = 0; p < k; ++p) t += x[p] * x[p];
  X2[i] = t;
}

__global__ void split_reduce_kernel(const double *__restrict__ partial, int nsplit, int m,
                                    double *u, const int *__restrict__ umap,
                                    const int *__restrict__ prow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double acc = 0.0;
  for (int sp = 0; sp < nsplit; ++sp) acc += partial[(long)sp * m + i];
  u[umap[prow ? prow[i] : i]] += acc;
}

hipError_t launch_split_reduce(const double *partial, int nsplit, int m, double *u,
                               const int *umap, const int *prow, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(split_reduce_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s,
                     partial, nsplit, m, u, umap, prow);
  return hipGetLastError();
}

__global__ void split_reduce_rhs_kernel(const double *__restrict__ partial, int nsplit, int m,
                                        int nrhs, double *u, const int *__restrict__ umap) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * nrhs) return;
  const long i = e / nrhs;
  const int r = (int)(e - i * nrhs);
  double acc = 0.0;
  for (int sp = 0; sp < nsplit; ++sp) acc += partial[(long)sp * m * nrhs + e];
  u[(long)umap[i] * nrhs + r] += acc;
}

hipError_t launch_split_reduce_rhs(const double *partial, int nsplit, int m, int nrhs, double *u,
                                   const int *umap, hipStream_t s) {
  const long tot = (long)m * nrhs;
  if (tot <= 0) return hipSuccess;
  hipLaunchKernelGGL(split_reduce_rhs_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                     partial, nsplit, m, nrhs, u, umap);
  return hipGetLastError();
}

}  // namespace gsks

extern "C" int ks_sqnorm_device(const double *X, long nx, int k, double *X2, void *stream) {
  using namespace gsks;
  if (nx < 0 || k < 0 || (nx > 0 && (!X || !X2)))
    return raise_error(KS_ERR_INVALID_ARG, false, "ks_sqnorm_device(): invalid argument\n");
  if (nx == 0) return KS_OK;
  if (ensure_device() != KS_OK)
    return raise_error(KS_ERR_DEVICE, false, "ks_sqnorm_device(): no GPU available\n");
  hipStream_t s = (hipStream_t)stream;
  long blocks = (nx + 255) / 256;
  hipLaunchKernelGGL(sqnorm_kernel, dim3((unsigned)blocks), dim3(256), 0, s, X, nx, k, X2);
  return hipGetLastError() == hipSuccess ? KS_OK : KS_ERR_DEVICE;
}
