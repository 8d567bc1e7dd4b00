// Mixed-precision BSE (FP32 factorisation + FP64 iterative refinement).
#include "ldga_internal.h"

int launch_bse_mixed(ldga_ctx *ctx, int, int, int, int, int, const double2 *, const double2 *, double, double, double,
                     double *, double2 *, unsigned long long *, hipStream_t) {
  return ldga_fail(ctx, LDGA_ERR_ARG, "mixed-precision BSE is not available yet");
}
