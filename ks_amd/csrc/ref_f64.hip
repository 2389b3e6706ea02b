// dgsks_ref() on the GPU: the reference algorithm of frame/dgsks_ref.c:94-535,
// step for step -- gather u, the cross-term matrix C = s * A^T B for a panel of
// source columns (s = -2 or +1, each entry a sequential sum over p without FMA,
// as the reference's naive loops), the elementwise kernel function in the
// reference's operation order (no clamps), the GEMV u += C w taken over the
// columns in order, scatter.  Unlike dgsks() it materialises the kernel matrix,
// one panel of at most kPanelBytes at a time.  It is the GPU's independent
// cross-check of the fused path, not a fast path.
#include <hip/hip_runtime.h>

#include "gsks_kernels.h"
#include "host_util.h"

#pragma clang fp contract(off)

namespace gsks {

constexpr size_t kPanelBytes = 256ull << 20;

template <int TYPE>
__device__ double kval_ref(double c, double aa, double bb, double hi, double hj,
                           const KerParams &P) {
  const double scale = (TYPE == POLYNOMIAL || TYPE == TANH) ? 1.0 : -2.0;
  double C = c * scale;
  if constexpr (TYPE == GAUSSIAN) {
    C += aa; C += bb; C *= P.scal;
    return exp(C);
  } else if constexpr (TYPE == VAR_BANDWIDTH) {
    C += aa; C += bb; C *= -0.5; C *= hi; C *= hj;
    return exp(C);
  } else if constexpr (TYPE == POLYNOMIAL) {
    C *= P.scal; C += P.cons;
    if (P.powe == 2.0) return C * C;
    if (P.powe == 4.0) { C = C * C; return C * C; }
    return pow(C, P.powe);
  } else if constexpr (TYPE == LAPLACE) {
    C += aa; C += bb;
    if (C < 1E-15) C = 1.79E+308;
    C = pow(C, P.powe);
    return C * P.scal;
  } else if constexpr (TYPE == TANH) {
    C *= P.scal; C += P.cons;
    return tanh(C);
  } else if constexpr (TYPE == QUARTIC) {
    C += aa; C += bb;
    if (C < 1.0) { C = 1.0 - C; return (15.0 / 16.0) * C * C; }
    return 0.0;
  } else if constexpr (TYPE == MULTIQUADRATIC) {
    C += aa; C += bb;
    return C + P.cons;
  } else {
    C += aa; C += bb;
    if (C < 1.0) return (3.0 / 4.0) * (1.0 - C);
    return 0.0;
  }
}

__global__ void ref_gather_u(const Problem p, double *us) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < p.m) us[i] = p.u[p.umap[i]];
}

template <int TYPE>
__global__ void ref_panel(const Problem p, int j0, int nb, double *Cs) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int jj = blockIdx.y;
  if (i >= p.m || jj >= nb) return;
  int ai = p.amap[i], bj = p.bmap[j0 + jj];
  const double *a = p.XA + (long)ai * p.k;
  const double *b = p.XB + (long)bj * p.k;
  double c = 0.0;
  for (int q = 0; q < p.k; ++q) c += a[q] * b[q];
  double hi = (TYPE == VAR_BANDWIDTH) ? p.hi[ai] : 0.0;
  double hj = (TYPE == VAR_BANDWIDTH) ? p.hj[bj] : 0.0;
  Cs[(long)jj * p.m + i] = kval_ref<TYPE>(c, p.XA2[ai], p.XB2[bj], hi, hj, p.P);
}

__global__ void ref_gemv(const Problem p, int j0, int nb, const double *Cs, double *us) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.m) return;
  double s = us[i];
  for (int jj = 0; jj < nb; ++jj) s += Cs[(long)jj * p.m + i] * p.w[p.wmap[j0 + jj]];
  us[i] = s;
}

__global__ void ref_scatter_u(const Problem p, const double *us) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < p.m) p.u[p.umap[i]] = us[i];
}

template <int TYPE>
static void panel_launch(const Problem &p, int j0, int nb, double *Cs, hipStream_t s) {
  dim3 grid((p.m + 255) / 256, nb);
  hipLaunchKernelGGL(ref_panel<TYPE>, grid, dim3(256), 0, s, p, j0, nb, Cs);
}

int enqueue_reference(const Problem &p, hipStream_t s) {
  long per_col = (long)p.m * 8;
  long nb = (long)(kPanelBytes / (size_t)per_col);
  if (nb < 1) nb = 1;
  if (nb > p.n) nb = p.n;
  if (nb > 65535) nb = 65535;
  double *buf = nullptr;
  size_t bytes = (size_t)p.m * 8 + (size_t)nb * p.m * 8 + 256;
  if (hipMallocAsync((void **)&buf, bytes, s) != hipSuccess) {
    (void)hipGetLastError();
    return KS_ERR_ALLOC;
  }
  double *us = buf;
  double *Cs = buf + (((size_t)p.m + 31) & ~(size_t)31);
  dim3 g1((p.m + 255) / 256);
  hipLaunchKernelGGL(ref_gather_u, g1, dim3(256), 0, s, p, us);
  for (int j0 = 0; j0 < p.n; j0 += (int)nb) {
    int cnt = (int)((p.n - j0) < nb ? (p.n - j0) : nb);
    switch (p.type) {
      case GAUSSIAN: panel_launch<GAUSSIAN>(p, j0, cnt, Cs, s); break;
      case POLYNOMIAL: panel_launch<POLYNOMIAL>(p, j0, cnt, Cs, s); break;
      case LAPLACE: panel_launch<LAPLACE>(p, j0, cnt, Cs, s); break;
      case VAR_BANDWIDTH: panel_launch<VAR_BANDWIDTH>(p, j0, cnt, Cs, s); break;
      case TANH: panel_launch<TANH>(p, j0, cnt, Cs, s); break;
      case QUARTIC: panel_launch<QUARTIC>(p, j0, cnt, Cs, s); break;
      case MULTIQUADRATIC: panel_launch<MULTIQUADRATIC>(p, j0, cnt, Cs, s); break;
      default: panel_launch<EPANECHNIKOV>(p, j0, cnt, Cs, s); break;
    }
    hipLaunchKernelGGL(ref_gemv, g1, dim3(256), 0, s, p, j0, cnt, (const double *)Cs, us);
  }
  hipLaunchKernelGGL(ref_scatter_u, g1, dim3(256), 0, s, p, (const double *)us);
  hipError_t e = hipGetLastError();
  (void)hipFreeAsync(buf, s);
  return e == hipSuccess ? KS_OK : KS_ERR_DEVICE;
}

}  // namespace gsks
