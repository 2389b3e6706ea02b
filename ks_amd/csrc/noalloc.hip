// dgsks_device_noalloc(): the FP64 summation of dgsks_device with its scratch
// (split partial sums, column partials, compact source tables) in a
// static device array instead of memory the library allocates, for callers
// that cannot allocate.  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "host_util.h"

namespace gsks {
// 64 MiB per GPU.  What the FP64 paths take from it: the matrix-core kernels
// (9 <= k <= 64) at most about 2 MiB of split partial sums at any size (about
// 2048 workgroups of 128 targets); the vector kernels up to 16 MiB of partial
// sums plus, from 1024 targets on, compact source tables of about n (k + 2.5)
// doubles (m = n = 262144 at k = 3: 30 MiB in all); symmetric sums their column
// partials [row blocks][m] when those fit, the general path otherwise.
// Larger summations are refused (dgsks_device serves them).
constexpr size_t kNoAllocBytes = (size_t)64 << 20;
}  // namespace gsks

// The static scratch: one copy on every GPU the library runs on, held until the
// process ends.  Host code never passes this name to a kernel (on the host it
// is only a stand-in for the symbol); it uses the address that
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[gsks::kNoAllocBytes / sizeof(double)];

namespace gsks {
namespace {

std::mutex g_mu;
std::map<int, char *> g_scratch_ptr;  // device id -> that device's copy of scratch

// The address of the current device's scratch, looked up once per device with
// that device current.
int scratch_address(char **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_failure(e, "dgsks_device_noalloc: hipGetDevice");
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_failure(e, "dgsks_device_noalloc: hipGetSymbolAddress");
    if (!ptr) return hip_failure(hipErrorInvalidSymbol, "dgsks_device_noalloc: hipGetSymbolAddress");
    it = g_scratch_ptr.emplace(dev, static_cast<char *>(ptr)).first;
  }
  *out = it->second;
  return KS_OK;
}

}  // namespace
}  // namespace gsks

extern "C" int dgsks_device_noalloc(const ks_t *kernel, int m, int n, int k, double *u,
                                    const int *umap, const double *XA, const double *XA2,
                                    const int *amap, const double *XB, const double *XB2,
                                    const int *bmap, const double *w, const int *wmap,
                                    int precision, void *stream) {
  using namespace gsks;
  const char *name = "dgsks_device_noalloc";
  Problem p;
  bool empty = false;
  int rc = device_problem(name, kernel, m, n, k, u, umap, XA, XA2, amap, XB, XB2, bmap, w, wmap,
                          p, empty);
  if (rc != KS_OK || empty) return rc;
  if (precision != KS_PREC_FP64)
    return raise_error(KS_ERR_INVALID_ARG, false, "%s(): only KS_PREC_FP64 is supported\n", name);
  char *base = nullptr;
  rc = scratch_address(&base);
  if (rc != KS_OK) return raise_error(rc, false, "%s(): %s\n", name, err_string(rc));
  const FixedScratch fs{base, kNoAllocBytes};
  rc = enqueue_fused(p, (hipStream_t)stream, &fs);
  if (rc == KS_ERR_INVALID_ARG)
    return raise_error(rc, false,
                       "%s(): the summation needs more scratch than the %zu MiB static buffer "
                       "(use dgsks_device)\n", name, kNoAllocBytes >> 20);
  if (rc != KS_OK) return raise_error(rc, false, "%s(): launch failed (%s)\n", name, err_string(rc));
  return KS_OK;
}
