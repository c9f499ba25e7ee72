// mfma_kernels.hip -- k >= 5 gate passes on MFMA (split-precision inputs,
// FP32 accumulation).  Filled in by the precision path.

#include <hip/hip_runtime.h>

#include "internal.h"

namespace qhip {

int LaunchGateMfma(float2* s, unsigned n, const unsigned* qs, unsigned k,
                   const double* m64, int method) {
  (void)s; (void)n; (void)qs; (void)k; (void)m64; (void)method;
  return QSIM_HIP_ERR_UNSUPPORTED;
}

}  // namespace qhip
