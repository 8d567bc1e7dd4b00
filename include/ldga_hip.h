/*
 * ldga_hip.h -- C ABI of libldga_hip.so, the MI355X (gfx950) ladder hot path
 * of LadderDGA.jl v0.15.1: bubble, Bethe-Salpeter solve, EoM self-energy and
 * the lambda-correction sums.
 *
 * Conventions (all entry points):
 *   - Arrays are column-major with Julia's index order, exactly as the
 *     reference stores them; "complex" arrays are interleaved (re, im)
 *     float64 pairs (Julia ComplexF64).
 *   - k arrays run over the irreducible wedge given to ldga_set_kgrid; the G
 *     tables run over the full Ns^D grid (first axis fastest), then over
 *     fft_range (origin freq->fft_first).
 *   - Functions without the _dev suffix take HOST pointers (Julia arrays):
 *     they upload, compute, download and synchronise.  The _dev variants take
 *     DEVICE pointers and a hipStream_t (passed as void*, NULL = the
 *     context's own stream); they are asynchronous.
 *   - Return value: LDGA_OK (0) or a negative LDGA_ERR_* code; the message is
 *     available from ldga_last_error().  The library never frees caller
 *     memory.  A context must be used from one host thread at a time.
 */
#ifndef LDGA_HIP_H
#define LDGA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LDGA_OK 0
#define LDGA_ERR_SHAPE (-1)      /* sizes inconsistent or above a kernel limit   */
#define LDGA_ERR_HIP (-2)        /* a HIP runtime call failed                    */
#define LDGA_ERR_ZERO_PIVOT (-3) /* BSE matrix exactly singular; outputs hold NaN */
#define LDGA_ERR_ARG (-4)        /* invalid argument (NULL pointer, bad flag)    */
#define LDGA_ERR_STATE (-5)      /* call order (e.g. no k grid set)              */

/* BSE precision modes */
#define LDGA_PREC_FP64 0  /* FP64 Gauss-Jordan elimination                     */
#define LDGA_PREC_MIXED 1 /* FP32 factorisation + FP64 iterative refinement    */

/* EoM kernel selectors: src/ladderDGATools.jl:260-271 (eom, eom_χ_m, ..., eom_sp_01, eom_sp_02, eom_sp, eom_ch,
 * eom_rest; the real-χ methods) */
#define LDGA_EOM_FULL 0
#define LDGA_EOM_CHI_M 1
#define LDGA_EOM_GAMMA_M 2
#define LDGA_EOM_CHI_D 3
#define LDGA_EOM_GAMMA_D 4
#define LDGA_EOM_REST_01 5
#define LDGA_EOM_REST 6
#define LDGA_EOM_SP_01 7
#define LDGA_EOM_SP_02 8
#define LDGA_EOM_SP 9
#define LDGA_EOM_CH 10

typedef struct ldga_ctx ldga_ctx;

/* SimulationParameters (src/Config.jl:66-75) as far as the path needs it. */
typedef struct {
  int32_t n_iw;        /* sP.n_iω                                        */
  int32_t n_inu;       /* sP.n_iν                                        */
  int32_t n_inu_shell; /* sP.n_iν_shell                                  */
  int32_t shift;       /* sP.shift (0/1)                                 */
  int32_t fft_first;   /* first(sP.fft_range)                            */
  int32_t n_fft;       /* length(sP.fft_range)                           */
} ldga_freq;

/* ModelParameters (src/Config.jl:39-47) fields used by the path, and the
 * 1/ω² tail coefficients of the two susceptibilities the path sums
 * (chi_m.tail_c[3] and chi_d.tail_c[3], src/DataTypes.jl:177-194): the c2 of
 * sum_kω(χm) in the EoM tail and of sum_kω(χr) in the λdm residual
 * (DataTypes.jl:215-249).  calc_χγ sets both to Ekin_DMFT
 * (BSETools.jl:241-243); update_tail! (DataTypes.jl:315-345) changes them. */
typedef struct {
  double U, mu, beta, n, Ekin_DMFT;
  double chi_m_c2, chi_d_c2;
} ldga_model;

/* Per-call BSE report. */
typedef struct {
  int64_t n_matrices;    /* matrices factorised                              */
  int64_t n_zero_pivot;  /* matrices with an exactly zero pivot              */
  int32_t ir_steps_max;  /* mixed mode: largest refinement step count used   */
  int32_t n_fallback;    /* mixed mode: matrices whose refinement ended above 2^-30
                            (or met a zero FP32 pivot)                       */
  double max_rel_resid;  /* mixed mode: largest final correction |dy|/|y|    */
  double ir_steps_mean;  /* mixed mode: mean refinement step count           */
  int64_t n_fp64_resolved; /* mixed mode: of n_fallback, matrices re-solved in
                            FP64 (all of them; their chi and gamma are the FP64
                            results)                                         */
} ldga_bse_report;

/* ---- context --------------------------------------------------------- */
int ldga_version(void);
/* New context on HIP device `device` (own non-blocking stream, default priority). */
int ldga_create(int device, ldga_ctx **out);
int ldga_destroy(ldga_ctx *ctx);
const char *ldga_last_error(const ldga_ctx *ctx);
/* Synchronise the context's stream (host-array calls already do). */
int ldga_sync(ldga_ctx *ctx);

/* Device memory for callers that keep the path's arrays resident (the _dev
 * entry points): plain hipMalloc / hipFree / synchronous copies on the
 * context's device. */
int ldga_malloc(ldga_ctx *ctx, uint64_t bytes, void **dptr);
int ldga_free(ldga_ctx *ctx, void *dptr);
int ldga_memcpy_h2d(ldga_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int ldga_memcpy_d2h(ldga_ctx *ctx, void *dst, const void *src, uint64_t bytes);

/* ---- k grid -----------------------------------------------------------
 * Replaces the KGrid of Dispersions.jl used at src/BSETools.jl:165,
 * src/ladderDGATools.jl:151-157 and DataTypes.jl:240-246 (kintegrate).
 * D in {2,3}; Ns even, Ns <= 64 (2D) / 32 (3D).  wedge[d + D*q] are the
 * integer grid coordinates (0..Ns/2, descending over d) of wedge point q;
 * kMult is derived from them.  eps[q] is the dispersion on the wedge. */
int ldga_set_kgrid(ldga_ctx *ctx, int D, int Ns, int Nq, const int32_t *wedge, const double *eps);
/* The library's own wedge for (D, Ns): fills wedge (D*Nq ints) and kmult (Nq)
 * when non-NULL and returns Nq (> 0) or an error code. */
int ldga_default_wedge(int D, int Ns, int32_t *wedge, double *kmult);

/* ---- bubble: calc_bubble (src/BSETools.jl:157-172) ---------------------
 * chi0[q, nu, w] (Nq x 2(n_inu+shell) x (2 n_iw+1)) = -beta conv_k(G(nu), G(nu+w)).
 * G_fft = fft(G_k), G_rfft = fft(reverse(G_k)) (src/helpers.jl:164-183). */
int ldga_bubble(ldga_ctx *ctx, const ldga_freq *fr, double beta, const double *G_fft, const double *G_rfft,
                double *chi0);
int ldga_bubble_dev(ldga_ctx *ctx, const ldga_freq *fr, double beta, const void *G_fft, const void *G_rfft,
                    void *chi0, void *stream);

/* ---- G tables: G_fft / G_rfft (src/helpers.jl:153-183) ------------------
 * From G on the wedge over fft_range (Nq x n_fft, column-major, e.g. the
 * G_ladder of ldga_lambdadm_residual), the full-grid tables
 * G_fft = fft(expandKArr(G)) and G_rfft = fft(reverse(expandKArr(G)))
 * (Ns^D x n_fft each), the inputs of ldga_bubble and ldga_sigma. */
int ldga_g_tables(ldga_ctx *ctx, const ldga_freq *fr, const double *G_wedge, double *G_fft, double *G_rfft);
int ldga_g_tables_dev(ldga_ctx *ctx, const ldga_freq *fr, const void *G_wedge, void *G_fft, void *G_rfft,
                      void *stream);

/* ---- Bethe-Salpeter: calc_chi_gamma (src/BSETools.jl:192-248) -----------
 * Replaces the per-(q,w) LapackWrapper.inv! (src/LapackWrapper.jl:64-72,
 * zgetrf_+zgetri_) and the plain sums: chi = Re sum(A^-1)/beta^2,
 * gamma[nu'] = sum_nu A^-1[nu,nu'] / (chi0[q,nu',w](1 + s U chi)), with
 * A = Gamma[:,:,w] + diag(1/chi0[q, shell+1:shell+Nnu, w]).
 * s = +1 for :m, -1 for :d.  Gamma is (Nnu x Nnu x Nw), chi (Nq x Nw) float64,
 * gamma (Nq x Nnu x Nw).  report may be NULL. */
int ldga_chi_gamma(ldga_ctx *ctx, int s, int precision, const ldga_freq *fr, double U, double beta,
                   const double *Gamma, const double *chi0, double *chi, double *gamma, ldga_bse_report *report);
int ldga_chi_gamma_dev(ldga_ctx *ctx, int s, int precision, const ldga_freq *fr, double U, double beta,
                       const void *Gamma, const void *chi0, void *chi, void *gamma, ldga_bse_report *report,
                       void *stream);

/* ---- allocation-free BSE ---------------------------------------------------
 * No counterpart in the reference.  ldga_chi_gamma_dev without any device
 * allocation, for callers that cannot allocate: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  Same kernels and
 * results as ldga_chi_gamma_dev; its scratch (report counters, the mixed re-solve list)
 * is a static device array of 4 MiB + 256 bytes per GPU, held until the process
 * ends.  Limits: Nnu <= 128 and Nq (2 n_iw + 1) <= 1048576 matrices, else
 * LDGA_ERR_SHAPE.  All calls on one GPU share the array: calls that overlap on
 * one GPU race.  Without a report the call is asynchronous and makes no host
 * synchronisation. */
int ldga_chi_gamma_noalloc_dev(ldga_ctx *ctx, int s, int precision, const ldga_freq *fr, double U, double beta,
                               const void *Gamma, const void *chi0, void *chi, void *gamma, ldga_bse_report *report,
                               void *stream);

/* ---- generalized susceptibility: calc_gen_χ (src/BSETools.jl:259-275) --
 * inv[:,:,q,w] = (Gamma[:,:,w] + diag(1/chi0[q, shell+1:shell+Nnu, w]))^-1,
 * inv (Nnu x Nnu x Nq x Nw) complex, in FP64 by Gauss-Jordan with column
 * pivoting in global memory (meant for analysis, not the hot path).
 * Synchronises the stream; LDGA_ERR_ZERO_PIVOT when a matrix is exactly
 * singular (its inverse then holds NaN/Inf). */
int ldga_gen_chi(ldga_ctx *ctx, const ldga_freq *fr, const double *Gamma, const double *chi0, double *inv);
int ldga_gen_chi_dev(ldga_ctx *ctx, const ldga_freq *fr, const void *Gamma, const void *chi0, void *inv, void *stream);

/* ---- local full vertex: F_from_χ (src/BSETools.jl:40-53) ---------------
 * F[nu,nu',w] = -(chi[nu,nu',w] + diag_term delta_{nu nu'} beta G(nu) G(nu+w))
 *               / (G(nu) G(nu+w) G(nu') G(nu'+w)),
 * chi and F (Nnu x Nnu x Nw) complex, G (n_fft) the local G over fft_range
 * (e.g. gImp).  Needs no k grid.  LDGA_ERR_SHAPE when fft_range does not
 * cover every nu and nu+w the formula reads. */
int ldga_F_from_chi(ldga_ctx *ctx, const ldga_freq *fr, double beta, int diag_term, const double *chi,
                    const double *G, double *F);
int ldga_F_from_chi_dev(ldga_ctx *ctx, const ldga_freq *fr, double beta, int diag_term, const void *chi,
                        const void *G, void *F, void *stream);

/* ---- lambda_0: calc_λ0 plain branch (src/BSETools.jl:128-141) ----------
 * lambda0[q,nu',w] = chi0[q, shell+nu', w]/beta^2 * sum_nu F[nu,nu',w]. */
int ldga_lambda0(ldga_ctx *ctx, const ldga_freq *fr, double beta, const double *F, const double *chi0,
                 double *lambda0);
int ldga_lambda0_dev(ldga_ctx *ctx, const ldga_freq *fr, double beta, const void *F, const void *chi0,
                     void *lambda0, void *stream);

/* ---- bosonic sums: sum_kω (src/DataTypes.jl:215-249) with chi_lambda ----
 * (LambdaCorrection/helpers.jl:18,61).  For chi (Nq x Nw) over the usable
 * range w_lo..w_hi (0-based, inclusive):
 *   out[0] = sum_kω(chi_λ(chi, lam))
 *   out[1] = sum_kω(dchi_λ(chi, lam)) exactly as the reference's df_c1
 *            (transform = dchi_λ, tail terms included; conditions_new.jl:153)
 *   out[2] = d out[0] / d lam  (the true derivative, no tail terms)
 * with tail coefficient c2 (chi.tail_c[3]); lam = 0 means no transform.
 * out_dev of the _dev form is a device array of 3 doubles. */
int ldga_sum_kw(ldga_ctx *ctx, const double *chi, int Nw, int w_lo, int w_hi, double c2, double beta, double lam,
                double out[3]);
int ldga_sum_kw_dev(ldga_ctx *ctx, const void *chi, int Nw, int w_lo, int w_hi, double c2, double beta,
                    double lam, void *out_dev, void *stream);

/* ---- sum_ω / sum_ωk (src/DataTypes.jl:257-308) ---------------------------
 * Per q, over the usable slice w_lo..w_hi (0-based, inclusive) of chi
 * (Nq x Nw), as the reference's sum_ω! / sum_ω:
 *   res[q] = (sum_{n=0}^{L-1} [chi_λ(q, w_lo + n) - t(n)] - c2 beta^2/12) / beta,
 * L = w_hi - w_lo + 1, t(n) = Re(c2 / (i w_m)^2) with m = n - n_iw and t = 0
 * at m = 0: slice element n is paired with the n-th frequency of the whole
 * grid, exactly as DataTypes.jl:293-305 indexes ωn_arr (on the full range
 * this is each element's own frequency).  chi_λ = chi/(lam chi + 1), lam = 0
 * meaning no map (sum_ωk's χ_λ!, DataTypes.jl:257-264).  out_q (Nq doubles,
 * nullable) receives res; out_k (1 double, nullable; needs the k grid, whose
 * Nq must match) receives kintegrate(res) = sum_ωk. */
int ldga_sum_w(ldga_ctx *ctx, const double *chi, int Nq, int Nw, int w_lo, int w_hi, double c2, double beta,
               double lam, double *out_q, double *out_k);
int ldga_sum_w_dev(ldga_ctx *ctx, const void *chi, int Nq, int Nw, int w_lo, int w_hi, double c2, double beta,
                   double lam, void *out_q, void *out_k, void *stream);

/* ---- EoM self-energy: calc_Σ (src/ladderDGATools_singleCore.jl:114-136) -
 * Sigma[k, nu] (Nq x numax, nu = 0..numax-1) =
 *   (1/beta) sum_w conv_q(K(q,nu,w), G(k+q, nu+w)) + tail/(i nu_n) + n U/2,
 * K = eom(U, gamma_m, gamma_d, chi_m^λm, chi_d^λd, lambda0)
 * (src/ladderDGATools.jl:260), tail = -U (sum_kω(chi_m^λm) - chiloc_m_sum)
 * (tail coefficient mp->chi_m_c2) when tc != 0.  eom_kind selects another of the reference's EoM functions
 * (LDGA_EOM_*: the parts of calc_Σ_parts, and the eomf argument of calc_Σ_ω!,
 * src/ladderDGATools_singleCore.jl:49-69); the Hartree term is added to FULL
 * and REST only, the tail only when tc != 0 and eom_kind == FULL.  chi_m/chi_d (Nq x Nw) are
 * read-only: the lambda map is applied on the fly and never stored. */
int ldga_sigma(ldga_ctx *ctx, const ldga_freq *fr, const ldga_model *mp, int numax, double lam_m, double lam_d,
               int tc, int eom_kind, double chiloc_m_sum, const double *chi_m, const double *gamma_m,
               const double *chi_d, const double *gamma_d, const double *lambda0, const double *G_rfft,
               double *sigma);
int ldga_sigma_dev(ldga_ctx *ctx, const ldga_freq *fr, const ldga_model *mp, int numax, double lam_m, double lam_d,
                   int tc, int eom_kind, double chiloc_m_sum, const void *chi_m, const void *gamma_m,
                   const void *chi_d, const void *gamma_d, const void *lambda0, const void *G_rfft, void *sigma,
                   void *stream);


/* ---- Matsubara shell sum: G_shell_sum (src/GFTools.jl:281-283) ----------
 * (1/beta) sum_{n in shell} 1/(i nu_n)^2 + beta/4 for indices beyond N-1,
 * = beta * trigamma(N + 1/2) / (2 pi^2).  Host-only, no context needed. */
double ldga_G_shell_sum(int N, double beta);

/* ---- lambda_dm residual: residual_vals / residual_f of λdm_correction ------
 * (LambdaCorrection/conditions_new.jl:250-287).  For (lam_m, lam_d):
 * Sigma = calc_Σ(...; λm, λd, tc) over nu = 0..numax-1, mu refit so that
 * filling_pos(G_ladder[:, 0:numax-1]) = n (GFTools.jl:162-186; solved to
 * |f| < 1e-14 when fit_mu != 0, else mu = mu0; a fit that ends with
 * |f| > 1e-8, the reference's find_zero tolerance, gives mu = NaN and NaN
 * results, as G_from_Σladder! returns NaN), E_kin/E_pot1 from calc_E
 * (thermodynamics.jl:142-168), the chi sums sum_kω(chi_r^λr) (tail
 * coefficients mp->chi_m_c2, mp->chi_d_c2) and
 *   res_pp   = (Σχd + Σχm)/2 - n/2 (1 - n/2)
 *   res_epot = E_pot1 - [(U/2)(Σχd - Σχm) + U n^2/4].
 * want_jac != 0 also returns d(res_pp, res_epot)/d(lam_m, lam_d) from the
 * hand-written adjoint (implicit differentiation through the mu fit and a
 * transposed EoM convolution); it has no counterpart in the reference, which
 * uses a forward finite-difference Jacobian (LambdaCorrection/helpers.jl:177-180).
 * Sigma_loc (nu = 0..last(fft_range), complex) is only read when G_ladder is
 * requested.  Outputs sigma (Nq x numax) and G_ladder (Nq x n_fft) are optional. */
typedef struct {
  double n, mu, Ekin, Epot1, Epot2, lhs_c1, res_pp, res_epot;
  double jac[4]; /* row-major: [dres_pp/dlam_m, dres_pp/dlam_d, dres_epot/dlam_m, dres_epot/dlam_d] */
  int32_t mu_iters;
  int32_t mu_converged; /* 1 when |f(mu)| <= 1e-8 (or no fit was asked for) */
  double n_full;        /* filling(G_ladder, kG, U, mu, beta) over the whole fft_range (GFTools.jl:234-241),
                           the trace's n (conditions_new.jl:511); NaN when Sigma_loc is not given */
} ldga_lambdadm_out;

int ldga_lambdadm_residual(ldga_ctx *ctx, const ldga_freq *fr, const ldga_model *mp, int numax, double lam_m,
                           double lam_d, int tc, double chiloc_m_sum, int fit_mu, double mu0, const double *chi_m,
                           const double *gamma_m, const double *chi_d, const double *gamma_d, const double *lambda0,
                           const double *G_rfft, const double *Sigma_loc, int want_jac, ldga_lambdadm_out *out,
                           double *sigma, double *G_ladder);
/* Device-pointer form: inputs resident on the device; synchronises the stream
 * to return `out` on the host; sigma / G_ladder are device pointers (nullable). */
int ldga_lambdadm_residual_dev(ldga_ctx *ctx, const ldga_freq *fr, const ldga_model *mp, int numax, double lam_m,
                               double lam_d, int tc, double chiloc_m_sum, int fit_mu, double mu0, const void *chi_m,
                               const void *gamma_m, const void *chi_d, const void *gamma_d, const void *lambda0,
                               const void *G_rfft, const void *Sigma_loc, int want_jac, ldga_lambdadm_out *out,
                               void *sigma, void *G_ladder, void *stream);

/* ---- G_from_Σladder! and calc_E on a given Sigma_ladder ------------------
 * (GFTools.jl:162-186, thermodynamics.jl:142-168): the step of the
 * self-consistency loop after the EoM (conditions_new.jl:484-488), device
 * pointers.  sigma (Nq x numax) is in/out: with sigma_prev != NULL and
 * mixing != 0 it is first replaced by (1 - mixing) sigma + mixing sigma_prev
 * (conditions_new.jl:484).  Then mu is refit (fit_mu != 0) so that
 * filling_pos(G_ladder[:, 0:numax-1]) = n, G_ladder (Nq x n_fft, nullable) is
 * written, and `out` carries n, mu, E_kin, E_pot1 and, from chi_m / chi_d at
 * (lam_m, lam_d), lhs_c1, E_pot2 and the two residuals (jac is not filled).
 * chi_m and chi_d may both be NULL (G_from_Σladder alone): lhs_c1, E_pot2 and
 * the residuals are then NaN.  Sigma_loc is read only for G_ladder.
 * Synchronises the stream. */
int ldga_g_from_sigma_ladder_dev(ldga_ctx *ctx, const ldga_freq *fr, const ldga_model *mp, int numax, double lam_m,
                                 double lam_d, int fit_mu, double mu0, const void *chi_m, const void *chi_d,
                                 void *sigma, const void *sigma_prev, double mixing, const void *Sigma_loc,
                                 ldga_lambdadm_out *out, void *G_ladder, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* LDGA_HIP_H */
