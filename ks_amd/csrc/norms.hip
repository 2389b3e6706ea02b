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

__global__ void zero_u32_kernel(unsigned *x, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0u;
}

hipError_t launch_zero_u32(unsigned *x, long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n);
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
