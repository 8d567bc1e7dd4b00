// Batched Bethe-Salpeter solve (calc_χγ, src/BSETools.jl:192-248): dispatch.
//
// The reference inverts A = Gamma_r(w) + diag(1/chi0(q,.,w)) for every (q,w)
// with zgetrf+zgetri (src/LapackWrapper.jl:64-72) and then only uses the
// column sums of A^-1:
//     chi       = Re sum_{nu nu'} A^-1 / beta^2
//     gamma[nu'] = sum_nu A^-1[nu,nu'] / (chi0[q,nu',w] (1 + s U chi))
// i.e. the row vector y^T = 1^T A^-1, the solution of A^T y = 1.  The kernels:
//   * FP64: bse_blk.hip (register-resident blocked Gauss-Jordan with column
//     pivoting over the whole row, Nnu <= 128), bse_gmem.hip above;
//   * mixed: bse_bgj.hip (FP32 16-row blocked Gauss-Jordan on MFMA and FP64
//     refinement, Nnu <= 160); its refinement failures are re-solved in FP64.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "ldga_internal.h"


int launch_bse_mixed(ldga_ctx *ctx, int N, int Nq, int Nw, int Nc, int shell, const double2 *Gam, const double2 *chi0,
                     double s, double U, double beta, double *chi, double2 *gamma, unsigned long long *counters,
                     int *redo, hipStream_t st);
int launch_bse_blk_f64(ldga_ctx *ctx, int N, int Nq, int Nw, int Nc, int shell, const double2 *Gam,
                       const double2 *chi0, double s, double U, double beta, double *chi, double2 *gamma,
                       unsigned long long *cnt, const int *list, const unsigned long long *count, int max_items,
                       hipStream_t st);
int launch_bse_gmem_f64(ldga_ctx *ctx, int N, int Nq, int Nw, int Nc, int shell, const double2 *Gam,
                        const double2 *chi0, double s, double U, double beta, double *chi, double2 *gamma,
                        unsigned long long *cnt, const int *only, int n_only, const unsigned long long *n_dev,
                        hipStream_t st, double2 *inv);

// FP64 over every matrix (list == NULL) or over the list[0 .. *count - 1] of
// at most max_items matrices, count read on the device: the blocked MFMA
// kernel up to Nnu = 128, the global-memory kernel above.
static int launch_fp64(ldga_ctx *ctx, int N, int Nq, int Nw, int Nc, int shell, const double2 *Gam,
                       const double2 *chi0, double s, double U, double beta, double *chi, double2 *gamma,
                       unsigned long long *cnt, const int *list, const unsigned long long *count, int max_items,
                       hipStream_t st) {
  if (N <= 128)
    return launch_bse_blk_f64(ctx, N, Nq, Nw, Nc, shell, Gam, chi0, s, U, beta, chi, gamma, cnt, list, count,
                              max_items, st);
  return launch_bse_gmem_f64(ctx, N, Nq, Nw, Nc, shell, Gam, chi0, s, U, beta, chi, gamma, cnt, list,
                             list ? max_items : 0, count, st, nullptr);
}

// Clears the report counters with a kernel rather than hipMemsetAsync, so that
// a caller may capture the FP64 path (no host synchronisation without a
// report) in a HIP graph.
namespace {
__global__ void bse_clear_counters(unsigned long long *cnt, int n) {
  if ((int)threadIdx.x < n) cnt[threadIdx.x] = 0ull;
}
}  // namespace

// calc_gen_χ (BSETools.jl:259-275): the full inverses of Gamma[:,:,w] +
// diag(1/chi0[q, shell+1:shell+Nnu, w]) by the global-memory FP64 kernel into
// inv (Nnu x Nnu x Nq x Nw).  Synchronises the stream to read the zero-pivot
// count (LDGA_ERR_ZERO_PIVOT: those inverses hold NaN/Inf, as inv would).
int launch_gen_chi(ldga_ctx *ctx, const ldga_freq *fr, const double2 *Gam, const double2 *chi0, double2 *inv,
                   hipStream_t st) {
  const int N = 2 * fr->n_inu, Nw = 2 * fr->n_iw + 1, Nq = ctx->Nq, Nc = N + 2 * fr->n_inu_shell;
  unsigned long long *cnt = (unsigned long long *)ldga_scratch(ctx, "gen_chi_counters", 8 * 8, st);
  if (!cnt) return LDGA_ERR_HIP;
  bse_clear_counters<<<1, 8, 0, st>>>(cnt, 8);
  LDGA_HIP_TRY(ctx, hipGetLastError());
  int rc = launch_bse_gmem_f64(ctx, N, Nq, Nw, Nc, fr->n_inu_shell, Gam, chi0, 1.0, 0.0, 1.0, nullptr, nullptr, cnt,
                               nullptr, 0, nullptr, st, inv);
  if (rc) return rc;
  unsigned long long zeros = 0;
  LDGA_HIP_TRY(ctx, hipMemcpyAsync(&zeros, cnt, 8, hipMemcpyDeviceToHost, st));
  LDGA_HIP_TRY(ctx, hipStreamSynchronize(st));
  if (zeros) return ldga_fail(ctx, LDGA_ERR_ZERO_PIVOT, "%llu matrices with an exactly zero pivot", zeros);
  return LDGA_OK;
}

int launch_bse(ldga_ctx *ctx, int s, int precision, const ldga_freq *fr, double U, double beta, const double2 *Gam,
               const double2 *chi0, double *chi, double2 *gamma, ldga_bse_report *rep, hipStream_t st) {
  const int Nw = 2 * fr->n_iw + 1, Nq = ctx->Nq;
  // counters: [0] zero pivots, [1] refinement steps max, [2] unconverged refinements,
  // [3] largest final refinement ratio (bits), [5] refinement steps summed over the
  // matrices; [8..15]: the same for the FP64 re-solve
  unsigned long long *cnt = (unsigned long long *)ldga_scratch(ctx, "bse_counters", 16 * 8, st);
  if (!cnt) return LDGA_ERR_HIP;
  // mixed mode: the refinement's failures are listed (room for every matrix)
  // and re-solved in FP64
  int *redo = nullptr;
  if (precision == LDGA_PREC_MIXED && !(redo = (int *)ldga_scratch(ctx, "bse_redo", (size_t)Nq * Nw * sizeof(int), st)))
    return LDGA_ERR_HIP;
  return launch_bse_with(ctx, s, precision, fr, U, beta, Gam, chi0, chi, gamma, rep, st, cnt, redo);
}

// The BSE with caller-provided scratch: cnt (16 counters) and, in mixed mode,
// redo (room for Nq Nw ints).
int launch_bse_with(ldga_ctx *ctx, int s, int precision, const ldga_freq *fr, double U, double beta,
                    const double2 *Gam, const double2 *chi0, double *chi, double2 *gamma, ldga_bse_report *rep,
                    hipStream_t st, unsigned long long *cnt, int *redo) {
  if (s != 1 && s != -1) return ldga_fail(ctx, LDGA_ERR_ARG, "channel sign s must be +1 (m) or -1 (d)");
  const int N = 2 * fr->n_inu, Nw = 2 * fr->n_iw + 1, Nq = ctx->Nq, Nc = N + 2 * fr->n_inu_shell;
  const bool resolve = precision == LDGA_PREC_MIXED;
  bse_clear_counters<<<1, 16, 0, st>>>(cnt, 16);
  LDGA_HIP_TRY(ctx, hipGetLastError());
  int rc;
  if (precision == LDGA_PREC_FP64) {
    rc = launch_fp64(ctx, N, Nq, Nw, Nc, fr->n_inu_shell, Gam, chi0, s, U, beta, chi, gamma, cnt, nullptr, nullptr, 0,
                     st);
  } else if (precision == LDGA_PREC_MIXED) {
    rc = launch_bse_mixed(ctx, N, Nq, Nw, Nc, fr->n_inu_shell, Gam, chi0, s, U, beta, chi, gamma, cnt, redo, st);
    // the refinement failures listed by the mixed kernel (their number in
    // cnt[2], read on the device) are re-solved in FP64, overwriting their
    // chi and gamma; no host synchronisation
    if (!rc)
      rc = launch_fp64(ctx, N, Nq, Nw, Nc, fr->n_inu_shell, Gam, chi0, s, U, beta, chi, gamma, cnt + 8, redo, cnt + 2,
                       Nq * Nw, st);
  } else {
    return ldga_fail(ctx, LDGA_ERR_ARG, "unknown precision mode %d", precision);
  }
  if (rc) return rc;
  if (!rep) return LDGA_OK;  // asynchronous: zero pivots show only as NaN outputs
  unsigned long long h[6] = {0, 0, 0, 0, 0, 0}, h2[6] = {0, 0, 0, 0, 0, 0};
  LDGA_HIP_TRY(ctx, hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, st));
  if (resolve) LDGA_HIP_TRY(ctx, hipMemcpyAsync(h2, cnt + 8, sizeof h2, hipMemcpyDeviceToHost, st));
  LDGA_HIP_TRY(ctx, hipStreamSynchronize(st));
  const unsigned long long nzero = h[0] + h2[0];
  rep->n_matrices = (int64_t)Nq * Nw;
  rep->n_zero_pivot = (int64_t)nzero;
  rep->ir_steps_max = (int32_t)h[1];
  rep->n_fallback = (int32_t)h[2];
  double r;
  memcpy(&r, &h[3], 8);
  rep->max_rel_resid = r;
  rep->ir_steps_mean = (double)h[5] / (double)((int64_t)Nq * Nw);
  rep->n_fp64_resolved = resolve ? (int64_t)h[2] : 0;
  if (nzero) return ldga_fail(ctx, LDGA_ERR_ZERO_PIVOT, "%llu BSE matrices had an exactly zero pivot", nzero);
  return LDGA_OK;
}
