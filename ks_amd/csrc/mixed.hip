// Mixed-precision summation (placeholder until the MFMA kernels land).
#include <hip/hip_runtime.h>

#include "host_util.h"

namespace gsks {

int enqueue_mixed(const Problem &p, int precision, hipStream_t s) {
  (void)p;
  (void)precision;
  (void)s;
  return KS_ERR_INVALID_ARG;
}

}  // namespace gsks
