// Batched Bethe-Salpeter solve (calc_χγ, src/BSETools.jl:192-248).
//
// The reference inverts A = Gamma_r(w) + diag(1/chi0(q,.,w)) for every (q,w)
// with zgetrf+zgetri (src/LapackWrapper.jl:64-72) and then only uses the
// column sums of A^-1:
//     chi       = Re sum_{nu nu'} A^-1 / beta^2
//     gamma[nu'] = sum_nu A^-1[nu,nu'] / (chi0[q,nu',w] (1 + s U chi))
// i.e. the row vector y^T = 1^T A^-1, the solution of A^T y = 1.  Each
// workgroup solves one (q,w) system by Gauss-Jordan elimination with partial
// pivoting on [A^T | 1] held entirely in registers:
//   * thread (tr, tc) owns rows tr + 16a and columns tc + TC b of the padded
//     system (the right-hand side is column N);
//   * column k is owned by 16 lanes of one wave: the pivot search is a
//     16-lane butterfly of (|re|+|im|, row) -- LAPACK's izamax criterion --
//     and those lanes publish the multipliers through LDS (one barrier per
//     column);
//   * the pivot row is exchanged wave-locally through LDS without a
//     workgroup barrier, because all owners of a column block sit in the
//     same wave;
//   * rows are never swapped: pivoted rows are masked out of later pivot
//     searches and every other row is eliminated, so after N steps row p_k
//     holds piv_k * y_k = b_{p_k}.
// Partial pivoting picks the same pivot sequence as zgetrf (up to ties).
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "ldga_internal.h"

namespace {

// Element b (wave-uniform, not a compile-time constant) of a register row, as
// a chain of selects.  The empty asm keeps the compiler from folding the chain
// back into a dynamically indexed array, which would move the whole register
// tile to scratch memory.
template <int CPT>
__device__ __forceinline__ double2 sel(const double (&re)[CPT], const double (&im)[CPT], int b) {
  double vr = re[0], vi = im[0];
#pragma unroll
  for (int i = 1; i < CPT; ++i) {
    const bool m = (i == b);
    vr = m ? re[i] : vr;
    vi = m ? im[i] : vi;
    __asm__ volatile("" : "+v"(vr), "+v"(vi));
  }
  return cz(vr, vi);
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wave's LDS stores have landed
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DPP lane exchange inside a 16-lane row (double split into two dwords).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}

// keep (c, i, v) of the better candidate: larger |.|, then smaller row index
template <int CTRL>
__device__ __forceinline__ void pivot_combine(double &bc, int &bi, double &bx, double &by) {
  const double oc = dpp_f64<CTRL>(bc), ox = dpp_f64<CTRL>(bx), oy = dpp_f64<CTRL>(by);
  const int oi = dpp_i32<CTRL>(bi);
  const bool take = (oc > bc) || (oc == bc && oi < bi);
  bc = take ? oc : bc;
  bi = take ? oi : bi;
  bx = take ? ox : bx;
  by = take ? oy : by;
}

// Pivot of column k among the unpivoted rows, run by the 16 lanes (one DPP row)
// that own column k; writes the multipliers l_i = B[i][k] / piv (0 for the
// pivot row) for every row and records (p_k, piv_k).
template <int RPT>
__device__ __forceinline__ void bse_pivot(int k, int tr, unsigned pmask, const double2 (&col)[RPT], double2 *lrow,
                                          int *piv_row, double2 *piv_val, unsigned long long *nzero) {
  double bc = -1.0, bx = 0.0, by = 0.0;
  int bi = 0x7fffffff;
#pragma unroll
  for (int a = 0; a < RPT; ++a) {
    double c = fabs(col[a].x) + fabs(col[a].y);
    if (!(c >= 0.0)) c = 0.0;  // NaN input: still pick a row, NaN propagates
    if (!((pmask >> a) & 1u) && c > bc) {
      bc = c;
      bi = tr + 16 * a;
      bx = col[a].x;
      by = col[a].y;
    }
  }
  // butterfly over the 16-lane row: quad xor 1, quad xor 2, half-row mirror, row mirror
  pivot_combine<0xB1>(bc, bi, bx, by);
  pivot_combine<0x4E>(bc, bi, bx, by);
  pivot_combine<0x141>(bc, bi, bx, by);
  pivot_combine<0x140>(bc, bi, bx, by);
  const double2 inv = crecip(cz(bx, by));
#pragma unroll
  for (int a = 0; a < RPT; ++a) {
    const int i = tr + 16 * a;
    lrow[i] = (i == bi) ? cz(0.0, 0.0) : cmul(col[a], inv);
  }
  if (tr == 0) {
    piv_row[k] = bi;
    piv_val[k] = cz(bx, by);
    if (bc == 0.0) atomicAdd(nzero, 1ull);
  }
}

// B[a][b] -= l[a] * u for every row of the thread (one column slot b)
template <int RPT, int CPT>
__device__ __forceinline__ void bse_update_slot(int b, double (&Br)[RPT][CPT], double (&Bi)[RPT][CPT],
                                                const double2 (&l)[RPT], double2 u) {
#pragma unroll
  for (int a = 0; a < RPT; ++a) {
    double xr = Br[a][b], xi = Bi[a][b];
    xr = fma(-l[a].x, u.x, xr);
    xr = fma(l[a].y, u.y, xr);
    xi = fma(-l[a].x, u.y, xi);
    xi = fma(-l[a].y, u.x, xi);
    Br[a][b] = xr;
    Bi[a][b] = xi;
  }
}

template <int RPT, int CPT, int TC>
__global__ __launch_bounds__(16 * TC) void bse_gj_f64_kernel(int N, int Nq, int Nc, int shell,
                                                              const double2 *__restrict__ Gam,
                                                              const double2 *__restrict__ chi0, double s, double U,
                                                              double beta, double *__restrict__ chi_out,
                                                              double2 *__restrict__ gamma_out,
                                                              unsigned long long *__restrict__ nzero) {
  constexpr int NT = 16 * TC, NW = NT / 64, ROWS = 16 * RPT;
  __shared__ double2 lbuf[2][ROWS];
  __shared__ double2 prow[NW][4][CPT];
  __shared__ int piv_row[ROWS];
  __shared__ double2 piv_val[ROWS];
  __shared__ double2 bvec[ROWS];
  __shared__ double red[NW];

  const int t = threadIdx.x, tr = t & 15, tc = t >> 4, wave = t >> 6, tcl = (t >> 4) & 3;
  const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int q = lb % Nq, w = lb / Nq;
  const double2 *G = Gam + (size_t)w * N * N;

  // B = A^T: B[i][j] = A[j][i] = Gamma[j, i, w] (+ 1/chi0 on the diagonal); column N = 1
  // real and imaginary parts in separate plain arrays so that they stay in VGPRs
  double Br[RPT][CPT], Bi[RPT][CPT];
  unsigned pmask = 0;
#pragma unroll
  for (int a = 0; a < RPT; ++a) {
    const int i = tr + 16 * a;
    if (i >= N) pmask |= 1u << a;
#pragma unroll
    for (int b = 0; b < CPT; ++b) {
      const int j = tc + TC * b;
      double2 v = cz(0.0, 0.0);
      if (i < N) {
        if (j < N)
          v = G[j + (size_t)N * i];
        else if (j == N)
          v = cz(1.0, 0.0);
        if (i == j) v = cadd(v, crecip(chi0[q + (size_t)Nq * (shell + i + (size_t)Nc * w)]));
      }
      Br[a][b] = v.x;
      Bi[a][b] = v.y;
    }
  }

  if (tc == 0) {
    double2 col[RPT];
#pragma unroll
    for (int a = 0; a < RPT; ++a) col[a] = cz(Br[a][0], Bi[a][0]);
    bse_pivot<RPT>(0, tr, pmask, col, lbuf[0], piv_row, piv_val, nzero);
  }

  for (int k = 0; k < N; ++k) {
    __syncthreads();  // multipliers and pivot of column k are published
    // ---- eliminate column k from every other row -------------------------
    const int p = piv_row[k];
    const int ap = p >> 4, trp = p & 15;
    if (tr == trp) {
      pmask |= 1u << ap;
      // ap is wave-uniform: a scalar branch picks the register row, no selects
#pragma unroll
      for (int a = 0; a < RPT; ++a) {
        if (a == ap) {
#pragma unroll
          for (int b = 0; b < CPT; ++b) prow[wave][tcl][b] = cz(Br[a][b], Bi[a][b]);
        }
      }
    }
    wave_lds_sync();
    double2 l[RPT];
#pragma unroll
    for (int a = 0; a < RPT; ++a) l[a] = lbuf[k & 1][tr + 16 * a];
    // Column slots whose every column (in this wave) is <= k are finished: their
    // entries are never read again, so the wave skips them.
    const int wmax = 4 * wave + 3;
    const int kn = k + 1, bn = kn / TC;
    const bool pivot_wave = kn < N && ((kn % TC) >> 2) == wave;
    if (pivot_wave) {
      // look-ahead: column k+1 first, then its pivot, then the remaining slots
      double2 col[RPT];
#pragma unroll
      for (int b = 0; b < CPT; ++b) {
        if (b == bn) {
          bse_update_slot<RPT, CPT>(b, Br, Bi, l, prow[wave][tcl][b]);
#pragma unroll
          for (int a = 0; a < RPT; ++a) col[a] = cz(Br[a][b], Bi[a][b]);
        }
      }
      if (tc == kn % TC) bse_pivot<RPT>(kn, tr, pmask, col, lbuf[kn & 1], piv_row, piv_val, nzero);
#pragma unroll
      for (int b = 0; b < CPT; ++b)
        if (b != bn && wmax + TC * b > k) bse_update_slot<RPT, CPT>(b, Br, Bi, l, prow[wave][tcl][b]);
    } else {
#pragma unroll
      for (int b = 0; b < CPT; ++b)
        if (wmax + TC * b > k) bse_update_slot<RPT, CPT>(b, Br, Bi, l, prow[wave][tcl][b]);
    }
  }

  // ---- y_k = b_{p_k} / piv_k, chi and gamma -------------------------------
  if (tc == N % TC) {
#pragma unroll
    for (int a = 0; a < RPT; ++a) bvec[tr + 16 * a] = sel<CPT>(Br[a], Bi[a], N / TC);
  }
  __syncthreads();
  double part = 0.0;
  for (int k = t; k < N; k += NT) {
    const double2 y = cdiv(bvec[piv_row[k]], piv_val[k]);
    lbuf[0][k] = y;
    part += y.x;
  }
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
  if ((t & 63) == 0) red[wave] = part;
  __syncthreads();
  double tot = 0.0;
#pragma unroll
  for (int i = 0; i < NW; ++i) tot += red[i];
  const double chi = tot / (beta * beta);
  const double fac = 1.0 + s * U * chi;
  for (int k = t; k < N; k += NT) {
    const double2 c0 = chi0[q + (size_t)Nq * (k + (size_t)Nc * w)];  // no shell offset (BSETools.jl:236)
    gamma_out[q + (size_t)Nq * (k + (size_t)N * w)] = cdiv(lbuf[0][k], cscale(c0, fac));
  }
  if (t == 0) chi_out[q + (size_t)Nq * w] = chi;
}

template <int RPT, int CPT, int TC>
int run_gj(ldga_ctx *ctx, int N, int Nq, int Nw, int Nc, int shell, const double2 *Gam, const double2 *chi0, double s,
           double U, double beta, double *chi, double2 *gamma, unsigned long long *nzero, hipStream_t st) {
  if (N > 16 * RPT || N + 1 > TC * CPT) return ldga_fail(ctx, LDGA_ERR_SHAPE, "BSE tile too small for N=%d", N);
  bse_gj_f64_kernel<RPT, CPT, TC><<<dim3((unsigned)Nq * Nw), dim3(16 * TC), 0, st>>>(N, Nq, Nc, shell, Gam, chi0, s, U,
                                                                                   beta, chi, gamma, nzero);
  LDGA_HIP_TRY(ctx, hipGetLastError());
  return LDGA_OK;
}

}  // namespace

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

// The register-tile Gauss-Jordan kernel above (Nnu <= 127) is kept as the
// second FP64 implementation; LDGA_BSE_F64=gj selects it instead of the
// blocked MFMA kernel (bse_blk.hip).
static bool use_register_gj() {
  const char *e = getenv("LDGA_BSE_F64");
  return e && strcmp(e, "gj") == 0;
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
  return launch_bse_with(ctx, s, precision, fr, U, beta, Gam, chi0, chi, gamma, rep, st, cnt, redo, use_register_gj());
}

// The BSE with caller-provided scratch: cnt (16 counters) and, in mixed mode,
// redo (room for Nq Nw ints).  register_gj selects the register-tile FP64
// kernel instead of the blocked one.
int launch_bse_with(ldga_ctx *ctx, int s, int precision, const ldga_freq *fr, double U, double beta,
                    const double2 *Gam, const double2 *chi0, double *chi, double2 *gamma, ldga_bse_report *rep,
                    hipStream_t st, unsigned long long *cnt, int *redo, bool register_gj) {
  if (s != 1 && s != -1) return ldga_fail(ctx, LDGA_ERR_ARG, "channel sign s must be +1 (m) or -1 (d)");
  const int N = 2 * fr->n_inu, Nw = 2 * fr->n_iw + 1, Nq = ctx->Nq, Nc = N + 2 * fr->n_inu_shell;
  const bool resolve = precision == LDGA_PREC_MIXED;
  bse_clear_counters<<<1, 16, 0, st>>>(cnt, 16);
  LDGA_HIP_TRY(ctx, hipGetLastError());
  int rc;
  if (precision == LDGA_PREC_FP64 && !register_gj) {
    rc = launch_fp64(ctx, N, Nq, Nw, Nc, fr->n_inu_shell, Gam, chi0, s, U, beta, chi, gamma, cnt, nullptr, nullptr, 0,
                     st);
  } else if (precision == LDGA_PREC_FP64) {
    if (N <= 47)
      rc = run_gj<3, 3, 16>(ctx, N, Nq, Nw, Nc, fr->n_inu_shell, Gam, chi0, s, U, beta, chi, gamma, cnt, st);
    else if (N <= 79)
      rc = run_gj<5, 3, 32>(ctx, N, Nq, Nw, Nc, fr->n_inu_shell, Gam, chi0, s, U, beta, chi, gamma, cnt, st);
    else if (N <= 95)
      rc = run_gj<6, 3, 32>(ctx, N, Nq, Nw, Nc, fr->n_inu_shell, Gam, chi0, s, U, beta, chi, gamma, cnt, st);
    else if (N <= 127)
      rc = run_gj<8, 4, 32>(ctx, N, Nq, Nw, Nc, fr->n_inu_shell, Gam, chi0, s, U, beta, chi, gamma, cnt, st);
    else
      return ldga_fail(ctx, LDGA_ERR_SHAPE, "the register Gauss-Jordan kernel (LDGA_BSE_F64=gj) supports Nnu <= 127 (got %d)", N);
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
