// The allocation-free form of the Bethe-Salpeter entry point.  Its scratch --
// the 16 report counters and the list of matrices whose mixed-precision
// refinement failed -- lives in a static device array of fixed size instead of
// context buffers, so a call makes no device allocation (for callers under HIP
// graph capture, in tight serving loops, or under harnesses that forbid
// hipMalloc).  Every GPU the library runs on holds its own copy of the array
// until the process ends; all calls on one GPU share it, so calls that overlap
// on one GPU race.
#include <algorithm>
#include <cstring>

#include "ldga_internal.h"

namespace {
constexpr size_t kNoallocCounterBytes = 256;           // 16 x u64 counters, padded to 256 bytes
constexpr size_t kNoallocMaxMatrices = size_t(1) << 20;  // (q, w) matrices per call
constexpr size_t kNoallocBytes = kNoallocCounterBytes + kNoallocMaxMatrices * sizeof(int);
}  // namespace

// the counters at offset 0, the re-solve list at offset 256
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

int ldga_noalloc_setup(ldga_ctx *ctx) {
  // the caller (ldga_create) has made ctx->device current; each device holds
  // its own copy of the array at its own address
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return ldga_fail(ctx, LDGA_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
  ctx->noalloc_scratch = ptr;
  return LDGA_OK;
}

extern "C" int ldga_chi_gamma_noalloc_dev(ldga_ctx *ctx, int s, int precision, const ldga_freq *fr, double U,
                                          double beta, const void *Gamma, const void *chi0, void *chi, void *gamma,
                                          ldga_bse_report *report, void *stream) {
  if (!ctx || !Gamma || !chi0 || !chi || !gamma) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  if (!ctx->has_kgrid) return ldga_fail(ctx, LDGA_ERR_STATE, "ldga_set_kgrid has not been called");
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  if (precision != LDGA_PREC_FP64 && precision != LDGA_PREC_MIXED)
    return ldga_fail(ctx, LDGA_ERR_ARG, "unknown precision mode %d", precision);
  const int N = 2 * fr->n_inu;
  const size_t nmat = (size_t)ctx->Nq * (2 * (size_t)fr->n_iw + 1);
  // the register-resident kernels only: above Nnu = 128 the FP64 solve needs a workspace
  if (N > 128)
    return ldga_fail(ctx, LDGA_ERR_SHAPE, "the allocation-free BSE supports Nnu <= 128 (got %d)", N);
  if (nmat > kNoallocMaxMatrices)
    return ldga_fail(ctx, LDGA_ERR_SHAPE, "the allocation-free BSE supports at most %zu (q, w) matrices (got %zu)",
                     kNoallocMaxMatrices, nmat);
  if (!ctx->noalloc_scratch) return ldga_fail(ctx, LDGA_ERR_STATE, "no scratch address for this device");
  hipSetDevice(ctx->device);
  unsigned char *base = static_cast<unsigned char *>(ctx->noalloc_scratch);
  unsigned long long *cnt = reinterpret_cast<unsigned long long *>(base);
  int *redo = reinterpret_cast<int *>(base + kNoallocCounterBytes);
  return launch_bse_with(ctx, s, precision, fr, U, beta, (const double2 *)Gamma, (const double2 *)chi0, (double *)chi,
                         (double2 *)gamma, report, ldga_stream(ctx, stream), cnt, redo);
}
