// The λdm residual of LambdaCorrection/conditions_new.jl:250-287 on the GPU:
//   EoM Σ(λm, λd)  ->  μ refit (device secant)  ->  calc_E, filling_pos, Σχ sums
// and, on request, its Jacobian by implicit differentiation plus the EoM
// adjoint (launch_eom_adjoint).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ldga_internal.h"

namespace {

constexpr int kFT = 1024;  // one workgroup does all scalar reductions

// sum over the workgroup (all threads must call)
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*red)[kFT / 64]) {
#pragma unroll
  for (int i = 0; i < NV; ++i)
    for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) red[i][wv] = v[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
    for (int j = 0; j < kFT / 64; ++j) s += red[i][j];
    v[i] = s;
  }
}

struct FinalArgs {
  const double2 *sigma;  // Nq x numax
  const double *chim, *chid;
  int Nq, numax, Nw, n_iw;
  double Nk, beta, U, n, mu0, shell, lam_m, lam_d, c2m, c2d, tol;
  int fit_mu, want_jac;
  double *res;   // 16 doubles: n mu Ekin Epot1 Epot2 lhs res_pp res_epot | Sm' Sd' - Fmu Emu it conv n_full
  double2 *W;    // Nq x numax adjoint weight (want_jac)
  double2 *Wk;   // numax: sum_k kMult W(k, nu)
};

// G(k, nu) = 1 / (i nu_n + mu - eps_k - Sigma(k, nu))
__device__ __forceinline__ double2 green(double nu_n, double mu, double eps, double2 sig) {
  return crecip(cz(mu - eps - sig.x, nu_n - sig.y));
}

__global__ __launch_bounds__(kFT) void lambdadm_final_kernel(OctTables T, FinalArgs A) {
  __shared__ double red[8][kFT / 64];
  const int tid = threadIdx.x;
  const double beta = A.beta, U = A.U, n = A.n, Nk = A.Nk;
  const double pib = M_PI / beta;
  // ---- χ sums: sum_kω(χ^λ) over all ω and their λ derivatives (DataTypes.jl:230-249)
  double v4[4] = {0.0, 0.0, 0.0, 0.0};
  const bool have_chi = A.chim != nullptr && A.chid != nullptr;  // NULL: G and energies only
  const size_t tot = have_chi ? (size_t)A.Nq * A.Nw : 0;
  for (size_t i = tid; i < tot; i += kFT) {
    const int q = (int)(i % A.Nq);
    const double km = T.kmult[q];
    double xm = A.chim[i], xd = A.chid[i];
    if (A.lam_m != 0.0) xm = xm / (A.lam_m * xm + 1.0);
    if (A.lam_d != 0.0) xd = xd / (A.lam_d * xd + 1.0);
    v4[0] = fma(km, xm, v4[0]);
    v4[1] = fma(km, -xm * xm, v4[1]);
    v4[2] = fma(km, xd, v4[2]);
    v4[3] = fma(km, -xd * xd, v4[3]);
  }
  block_sum<4>(v4, red);
  // the 1/(i w_m)^2 tail over all w per unit c2 (DataTypes.jl:223-225); each sum has its own c2
  double tail1 = 0.0;
  for (int w = 0; w < A.Nw; ++w) {
    const int m = w - A.n_iw;
    if (m != 0) tail1 += -beta * beta / (4.0 * M_PI * M_PI * (double)m * (double)m);
  }
  const double Sm = (v4[0] / Nk - A.c2m * tail1 - beta * beta / 12.0 * A.c2m) / beta, dSm = v4[1] / Nk / beta;
  const double Sd = (v4[2] / Nk - A.c2d * tail1 - beta * beta / 12.0 * A.c2d) / beta, dSd = v4[3] / Nk / beta;

  // ---- filling_pos(G[:, 0:numax-1]) (GFTools.jl:255-273) as a function of mu
  const size_t ne = (size_t)A.Nq * A.numax;
  auto sumG = [&](double mu) -> double {
    double v[1] = {0.0};
    for (size_t i = tid; i < ne; i += kFT) {
      const int q = (int)(i % A.Nq), nu = (int)(i / A.Nq);
      const double2 g = green((2.0 * nu + 1.0) * pib, mu, T.eps[q], A.sigma[i]);
      v[0] = fma(T.kmult[q], g.x, v[0]);
    }
    block_sum<1>(v, red);
    return v[0] / Nk;
  };
  auto filling = [&](double mu, double reSG) {
    return 2.0 * (2.0 * reSG / beta + 0.5 + mu * A.shell) / (1.0 + U * A.shell);
  };
  // secant from mu0 (the reference calls Roots.find_zero, GFTools.jl:172)
  double mu = A.mu0;
  int iters = 0, conv = 1;
  if (A.fit_mu) {
    double x0 = A.mu0, x1 = (x0 == 0.0) ? 1e-4 : x0 * (1.0 + 1e-4);
    double f0 = filling(x0, sumG(x0)) - n, f1 = filling(x1, sumG(x1)) - n;
    double fmu = f1;  // f at the returned mu
    conv = 0;
    for (iters = 0; iters < 100; ++iters) {
      if (fabs(f1) < A.tol) {
        conv = 1;
        break;
      }
      if (f1 == f0) break;
      const double x2 = x1 - f1 * (x1 - x0) / (f1 - f0);
      x0 = x1;
      f0 = f1;
      x1 = x2;
      f1 = filling(x1, sumG(x1)) - n;
    }
    mu = x1;
    fmu = f1;
    if (!conv) {
      // fallback, as a bracketing solver would (the filling grows with mu): step
      // out from mu0 until the sign changes, then Illinois false position
      double a = A.mu0, fa = filling(a, sumG(a)) - n;
      const double dir = fa < 0.0 ? 1.0 : -1.0;
      double h = 1.0, b = a + dir * h, fb = filling(b, sumG(b)) - n;
      for (int k = 0; k < 60 && !(fa * fb <= 0.0); ++k) {
        a = b;
        fa = fb;
        h *= 2.0;
        b = a + dir * h;
        fb = filling(b, sumG(b)) - n;
      }
      if (fa * fb <= 0.0) {
        int side = 0;
        for (int k = 0; k < 200; ++k, ++iters) {
          const double c = (a * fb - b * fa) / (fb - fa);
          const double fc = filling(c, sumG(c)) - n;
          mu = c;
          fmu = fc;
          if (fabs(fc) < A.tol || fabs(b - a) <= 4e-16 * fmax(fabs(a), fabs(b))) {
            conv = 1;
            break;
          }
          if (fc * fb > 0.0) {
            b = c;
            fb = fc;
            if (side == -1) fa *= 0.5;
            side = -1;
          } else {
            a = c;
            fa = fc;
            if (side == 1) fb *= 0.5;
            side = 1;
          }
        }
      }
    }
    // a fit that did not reach the reference's tolerance (find_zero atol = 1e-8,
    // GFTools.jl:171-176) fails: mu = NaN, and every result below is NaN
    conv = fabs(fmu) <= 1e-8;
    if (!conv) mu = NAN;
  }
  // ---- filling, calc_E (thermodynamics.jl:142-168), and for the Jacobian
  //      d(filling)/dmu and dE_pot/dmu
  const double Sh = n * U / 2.0;
  double v6[6] = {0, 0, 0, 0, 0, 0};  // Re sG, Ekin, Epot, Re sum G^2 (dF/dmu), dEpot/dmu
  for (size_t i = tid; i < ne; i += kFT) {
    const int q = (int)(i % A.Nq), nu = (int)(i / A.Nq);
    const double e = T.eps[q], km = T.kmult[q];
    const double nun = (2.0 * nu + 1.0) * pib;
    const double inu2 = -nun * nun;  // (i nu_n)^2
    const double2 sig = A.sigma[i];
    const double2 g = green(nun, mu, e, sig);
    const double Ekt = e + Sh - mu;
    const double Ept = U * U * 0.5 * n * (1.0 - 0.5 * n) + Sh * (e + Sh - mu);
    const double2 gs = cmul(g, sig);
    const double2 g2 = cmul(g, g);
    v6[0] = fma(km, g.x, v6[0]);
    v6[1] = fma(km, 4.0 * e * (g.x - Ekt / inu2), v6[1]);
    v6[2] = fma(km, 2.0 * (gs.x - Ept / inu2), v6[2]);
    v6[3] = fma(km, -g2.x, v6[3]);
    v6[4] = fma(km, 2.0 * (-cmul(g2, sig).x + Sh / inu2), v6[4]);
    if (nu == 0) {  // the per-k tail terms, once per k
      v6[1] = fma(km, (beta / 2.0) * e * (1.0 - beta * Ekt), v6[1]);
      v6[2] = fma(km, (beta / 2.0) * Sh - (beta * beta / 4.0) * Ept, v6[2]);
      v6[4] = fma(km, beta * beta * Sh / 4.0, v6[4]);
    }
  }
  block_sum<6>(v6, red);
  const double nval = filling(mu, v6[0] / Nk);
  const double Ekin = v6[1] / Nk / beta, Epot1 = v6[2] / Nk / beta;
  const double lhs = have_chi ? 0.5 * (Sd + Sm) : NAN;
  const double Epot2 = have_chi ? 0.5 * U * (Sd - Sm) + U * (n / 2.0) * (n / 2.0) : NAN;
  if (tid == 0) {
    double *r = A.res;
    r[0] = nval;
    r[1] = mu;
    r[2] = Ekin;
    r[3] = Epot1;
    r[4] = Epot2;
    r[5] = lhs;
    r[6] = lhs - n / 2.0 * (1.0 - n / 2.0);
    r[7] = Epot1 - Epot2;
    r[8] = dSm;
    r[9] = dSd;
    r[13] = (double)iters;
    r[14] = (double)conv;
    r[15] = NAN;  // the full filling, written by nfull_kernel when G_ladder is formed
  }
  if (!A.want_jac) return;
  // ---- adjoint weight W = dR_epot/dSigma (full-grid function, wedge values)
  const double Fmu = 2.0 * (2.0 * (v6[3] / Nk) / beta + A.shell) / (1.0 + U * A.shell);
  const double Emu = v6[4] / Nk / beta;
  const double cmu = -(Emu / Fmu) * 4.0 / (beta * (1.0 + U * A.shell));
  for (size_t i = tid; i < ne; i += kFT) {
    const int q = (int)(i % A.Nq), nu = (int)(i / A.Nq);
    const double2 sig = A.sigma[i];
    const double2 g = green((2.0 * nu + 1.0) * pib, mu, T.eps[q], sig);
    const double2 g2 = cmul(g, g);
    const double2 wv = cadd(cscale(cadd(g, cmul(sig, g2)), 2.0 / beta), cscale(g2, cmu));
    A.W[i] = cscale(wv, 1.0 / Nk);
  }
  // sum_k kMult W(k, nu), recomputed per nu (no read-back of W across threads)
  for (int nu = tid; nu < A.numax; nu += kFT) {
    double2 s = cz(0.0, 0.0);
    for (int q = 0; q < A.Nq; ++q) {
      const double2 sig = A.sigma[q + (size_t)A.Nq * nu];
      const double2 g = green((2.0 * nu + 1.0) * pib, mu, T.eps[q], sig);
      const double2 g2 = cmul(g, g);
      const double2 wv = cadd(cscale(cadd(g, cmul(sig, g2)), 2.0 / beta), cscale(g2, cmu));
      s = cadd(s, cscale(wv, T.kmult[q] / Nk));
    }
    A.Wk[nu] = s;
  }
  if (tid == 0) {
    A.res[11] = Fmu;
    A.res[12] = Emu;
  }
}

// sigma <- (1 - m) sigma + m sigma_prev (conditions_new.jl:484)
__global__ void mix_kernel(double2 *__restrict__ sigma, const double2 *__restrict__ prev, double m, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) sigma[i] = cadd(cscale(sigma[i], 1.0 - m), cscale(prev[i], m));
}

// G_ladder over fft_range: Sigma_ladder below numax, Sigma_loc above, negative
// frequencies by conjugation (GFTools.jl:162-186).  Each workgroup also leaves
// sum kMult Re G over its elements in part[blockIdx.x], the pieces of the full
// filling (GFTools.jl:234-241) that nfull_kernel adds up.
constexpr int kGL = 256;
__global__ __launch_bounds__(kGL) void gladder_kernel(OctTables T, int numax, int fft_first, int n_fft, double beta,
                                                      const double *res, const double2 *__restrict__ sigma,
                                                      const double2 *__restrict__ sigloc, double2 *__restrict__ G,
                                                      double *__restrict__ part) {
  __shared__ double red[kGL / 64];
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double re = 0.0;
  if (i < (size_t)T.Nq * n_fft) {
    const int q = (int)(i % T.Nq), col = (int)(i / T.Nq);
    const int nun = fft_first + col;
    const int m = nun >= 0 ? nun : -nun - 1;
    const double2 sig = (m < numax) ? sigma[q + (size_t)T.Nq * m] : sigloc[m];
    double2 g = green((2.0 * m + 1.0) * M_PI / beta, res[1], T.eps[q], sig);
    if (nun < 0) g.y = -g.y;
    G[i] = g;
    re = T.kmult[q] * g.x;
  }
  for (int off = 32; off > 0; off >>= 1) re += __shfl_xor(re, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = re;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// filling(G_ladder, kG, U, mu, beta) = 2 (Re sum_nu kintegrate(G)/beta + 1/2 + mu shell)/(1 + U shell),
// shell = G_shell_sum(floor(n_fft / 2), beta)  (GFTools.jl:234-241), into res[15]
__global__ __launch_bounds__(kGL) void nfull_kernel(const double *__restrict__ part, int nparts, double Nk,
                                                    double beta, double U, double shell, double *res) {
  __shared__ double red[kGL / 64];
  double s = 0.0;
  for (int k = threadIdx.x; k < nparts; k += kGL) s += part[k];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double reSG = (red[0] + red[1] + red[2] + red[3]) / Nk;
    res[15] = 2.0 * (reSG / beta + 0.5 + res[1] * shell) / (1.0 + U * shell);
  }
}

// trigamma(x) for x > 0: recurrence up to x >= 20, then the asymptotic series
double trigamma(double x) {
  double acc = 0.0;
  while (x < 20.0) {
    acc += 1.0 / (x * x);
    x += 1.0;
  }
  const double x2 = 1.0 / (x * x);
  // 1/x + 1/(2x^2) + sum_k B_2k / x^(2k+1): B2=1/6, B4=-1/30, B6=1/42, B8=-1/30, B10=5/66
  const double series = 1.0 / x + 0.5 * x2 +
                        (1.0 / x) * x2 * (1.0 / 6.0 + x2 * (-1.0 / 30.0 + x2 * (1.0 / 42.0 + x2 * (-1.0 / 30.0 + x2 * (5.0 / 66.0)))));
  return acc + series;
}

}  // namespace

extern "C" double ldga_G_shell_sum(int N, double beta) {
  // (psi1(N+1/2) - psi1(1/2-N)) beta/(4 pi^2) + beta/4 with the reflection
  // psi1(1/2-N) = pi^2 - psi1(N+1/2)  ->  beta psi1(N+1/2) / (2 pi^2)
  return beta * trigamma((double)N + 0.5) / (2.0 * M_PI * M_PI);
}

static int lambdadm_dev(ldga_ctx *ctx, const ldga_freq *fr, const ldga_model *mp, int numax, double lam_m,
                        double lam_d, int tc, double chiloc_m_sum, int fit_mu, double mu0, const double *chi_m,
                        const double2 *gamma_m, const double *chi_d, const double2 *gamma_d, const double2 *lambda0,
                        const double2 *Grfft, const double2 *sigloc, int want_jac, ldga_lambdadm_out *out,
                        double2 *sigma_out, double2 *G_out, hipStream_t st, bool eom = true,
                        const double2 *sigma_prev = nullptr, double mixing = 0.0) {
  // eom = false: sigma_out holds a given Sigma_ladder (optionally mixed with
  // sigma_prev first) and only the mu fit, the energies and G_ladder follow
  if (!out) return ldga_fail(ctx, LDGA_ERR_ARG, "out is NULL");
  if (!eom && (!sigma_out || want_jac)) return ldga_fail(ctx, LDGA_ERR_ARG, "a given Sigma needs sigma and no Jacobian");
  if (!ctx->has_kgrid) return ldga_fail(ctx, LDGA_ERR_STATE, "ldga_set_kgrid has not been called");
  if (G_out && !sigloc) return ldga_fail(ctx, LDGA_ERR_ARG, "G_ladder needs Sigma_loc");
  const int Nq = ctx->Nq, Nw = 2 * fr->n_iw + 1;
  memset(out, 0, sizeof *out);
  if (lam_m > 1e5 || lam_d > 1e5) {  // conditions_new.jl:251
    const double nan = std::nan("");
    out->n = out->mu = out->Ekin = out->Epot1 = out->Epot2 = out->lhs_c1 = out->res_pp = out->res_epot = nan;
    out->n_full = nan;
    for (double &j : out->jac) j = nan;
    return LDGA_OK;
  }
  double2 *sigma = sigma_out ? sigma_out : (double2 *)ldga_buf(ctx, "ldm_sigma", (size_t)Nq * numax * 16);
  double *res = (double *)ldga_buf(ctx, "ldm_res", 16 * 8 + 2 * 16);
  if (!sigma || !res) return LDGA_ERR_HIP;
  int rc = LDGA_OK;
  if (eom) {
    rc = launch_sigma(ctx, fr, mp, numax, lam_m, lam_d, tc, LDGA_EOM_FULL, chiloc_m_sum, chi_m, gamma_m, chi_d,
                      gamma_d, lambda0, Grfft, sigma, st);
    if (rc) return rc;
  } else if (sigma_prev && mixing != 0.0) {
    const size_t tot = (size_t)Nq * numax;
    mix_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(sigma, sigma_prev, mixing, tot);
    LDGA_HIP_TRY(ctx, hipGetLastError());
  }
  FinalArgs A{};
  A.sigma = sigma;
  A.chim = chi_m;
  A.chid = chi_d;
  A.Nq = Nq;
  A.numax = numax;
  A.Nw = Nw;
  A.n_iw = fr->n_iw;
  A.Nk = ctx->oct.Nk;
  A.beta = mp->beta;
  A.U = mp->U;
  A.n = mp->n;
  A.mu0 = mu0;
  A.shell = ldga_G_shell_sum(numax + 1, mp->beta);
  A.lam_m = lam_m;
  A.lam_d = lam_d;
  A.c2m = mp->chi_m_c2;
  A.c2d = mp->chi_d_c2;
  A.tol = 1e-14;
  A.fit_mu = fit_mu;
  A.want_jac = want_jac;
  A.res = res;
  double2 *adj = (double2 *)(res + 16);
  if (want_jac) {
    A.W = (double2 *)ldga_buf(ctx, "ldm_W", (size_t)Nq * numax * 16);
    A.Wk = (double2 *)ldga_buf(ctx, "ldm_Wk", (size_t)numax * 16);
    if (!A.W || !A.Wk) return LDGA_ERR_HIP;
  }
  lambdadm_final_kernel<<<1, kFT, 0, st>>>(ctx->oct, A);
  LDGA_HIP_TRY(ctx, hipGetLastError());
  if (want_jac) {
    rc = launch_eom_adjoint(ctx, fr, mp, numax, lam_m, lam_d, chi_m, gamma_m, chi_d, gamma_d, Grfft, A.W, adj, st);
    if (rc) return rc;
  }
  if (G_out) {
    const size_t tot = (size_t)Nq * fr->n_fft;
    const unsigned nb = (unsigned)((tot + kGL - 1) / kGL);
    double *part = (double *)ldga_buf(ctx, "ldm_gl_part", (size_t)nb * 8);
    if (!part) return LDGA_ERR_HIP;
    gladder_kernel<<<nb, kGL, 0, st>>>(ctx->oct, numax, fr->fft_first, fr->n_fft, mp->beta, res, sigma, sigloc, G_out,
                                       part);
    LDGA_HIP_TRY(ctx, hipGetLastError());
    nfull_kernel<<<1, kGL, 0, st>>>(part, (int)nb, ctx->oct.Nk, mp->beta, mp->U,
                                    ldga_G_shell_sum(fr->n_fft / 2, mp->beta), res);
    LDGA_HIP_TRY(ctx, hipGetLastError());
  }
  double h[20];
  LDGA_HIP_TRY(ctx, hipMemcpyAsync(h, res, sizeof h, hipMemcpyDeviceToHost, st));
  std::vector<double2> wk;
  if (want_jac && tc) {
    wk.resize(numax);
    LDGA_HIP_TRY(ctx, hipMemcpyAsync(wk.data(), A.Wk, numax * 16, hipMemcpyDeviceToHost, st));
  }
  LDGA_HIP_TRY(ctx, hipStreamSynchronize(st));
  out->n = h[0];
  out->mu = h[1];
  out->Ekin = h[2];
  out->Epot1 = h[3];
  out->Epot2 = h[4];
  out->lhs_c1 = h[5];
  out->res_pp = h[6];
  out->res_epot = h[7];
  out->mu_iters = (int32_t)h[13];
  out->mu_converged = (int32_t)h[14];
  out->n_full = h[15];
  if (want_jac) {
    const double dSm = h[8], dSd = h[9], U = mp->U, beta = mp->beta;
    // d res_pp / d lam = (1/2) dΣχ/dlam
    out->jac[0] = 0.5 * dSm;
    out->jac[1] = 0.5 * dSd;
    // tail of Σ: d/dlam_m [-U (Σχm - chiloc)] / (i nu_n), contracted with W
    double tail_m = 0.0;
    if (tc)
      for (int nu = 0; nu < numax; ++nu) {
        const double nun = (2.0 * nu + 1.0) * M_PI / beta;  // 1/(i nu) = -i/nu
        const double2 d = cz(0.0, -(-U * dSm) / nun);
        tail_m += wk[nu].x * d.x - wk[nu].y * d.y;
      }
    const double Am = h[16] / beta, Ad = h[18] / beta;  // Re of the adjoint contractions
    // E_pot2 = (U/2)(Σχd - Σχm) + const
    out->jac[2] = Am + tail_m + 0.5 * U * dSm;
    out->jac[3] = Ad - 0.5 * U * dSd;
  }
  return LDGA_OK;
}

extern "C" {

int ldga_lambdadm_residual_dev(ldga_ctx *ctx, const ldga_freq *fr, const ldga_model *mp, int numax, double lam_m,
                               double lam_d, int tc, double chiloc_m_sum, int fit_mu, double mu0, const void *chi_m,
                               const void *gamma_m, const void *chi_d, const void *gamma_d, const void *lambda0,
                               const void *G_rfft, const void *Sigma_loc, int want_jac, ldga_lambdadm_out *out,
                               void *sigma, void *G_ladder, void *stream) {
  if (!ctx || !mp || !chi_m || !gamma_m || !chi_d || !gamma_d || !lambda0 || !G_rfft)
    return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  return lambdadm_dev(ctx, fr, mp, numax, lam_m, lam_d, tc, chiloc_m_sum, fit_mu, mu0, (const double *)chi_m,
                      (const double2 *)gamma_m, (const double *)chi_d, (const double2 *)gamma_d,
                      (const double2 *)lambda0, (const double2 *)G_rfft, (const double2 *)Sigma_loc, want_jac, out,
                      (double2 *)sigma, (double2 *)G_ladder, ldga_stream(ctx, stream));
}

int ldga_lambdadm_residual(ldga_ctx *ctx, const ldga_freq *fr, const ldga_model *mp, int numax, double lam_m,
                           double lam_d, int tc, double chiloc_m_sum, int fit_mu, double mu0, const double *chi_m,
                           const double *gamma_m, const double *chi_d, const double *gamma_d, const double *lambda0,
                           const double *G_rfft, const double *Sigma_loc, int want_jac, ldga_lambdadm_out *out,
                           double *sigma, double *G_ladder) {
  if (!ctx || !mp || !chi_m || !gamma_m || !chi_d || !gamma_d || !lambda0 || !G_rfft)
    return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  if (!ctx->has_kgrid) return ldga_fail(ctx, LDGA_ERR_STATE, "ldga_set_kgrid has not been called");
  hipSetDevice(ctx->device);
  const size_t Nq = ctx->Nq, Nnu = 2 * (size_t)fr->n_inu, Nw = 2 * (size_t)fr->n_iw + 1;
  const size_t xb = Nq * Nw * 8, yb = Nq * Nnu * Nw * 16, gb = (size_t)ctx->oct.Nk * fr->n_fft * 16;
  const size_t lb = (size_t)(fr->fft_first + fr->n_fft) * 16;  // Sigma_loc over 0..last(fft_range)
  struct In {
    const char *role;
    const void *src;
    size_t bytes;
    void *dev;
  } in[] = {{"chi_m", chi_m, xb, nullptr},     {"chi_d", chi_d, xb, nullptr},     {"gamma_m", gamma_m, yb, nullptr},
            {"gamma_d", gamma_d, yb, nullptr}, {"lambda0", lambda0, yb, nullptr}, {"G_rfft", G_rfft, gb, nullptr},
            {"sigma_loc", Sigma_loc, lb, nullptr}};
  for (auto &x : in) {
    if (!x.src) continue;
    if (!(x.dev = ldga_buf(ctx, x.role, x.bytes))) return LDGA_ERR_HIP;
    LDGA_HIP_TRY(ctx, hipMemcpyAsync(x.dev, x.src, x.bytes, hipMemcpyHostToDevice, ctx->stream));
  }
  void *so = ldga_buf(ctx, "sigma", Nq * (size_t)std::max(numax, 1) * 16);
  void *go = G_ladder ? ldga_buf(ctx, "G_ladder", Nq * (size_t)fr->n_fft * 16) : nullptr;
  if (!so || (G_ladder && !go)) return LDGA_ERR_HIP;
  rc = lambdadm_dev(ctx, fr, mp, numax, lam_m, lam_d, tc, chiloc_m_sum, fit_mu, mu0, (const double *)in[0].dev,
                    (const double2 *)in[2].dev, (const double *)in[1].dev, (const double2 *)in[3].dev,
                    (const double2 *)in[4].dev, (const double2 *)in[5].dev, (const double2 *)in[6].dev, want_jac, out,
                    (double2 *)so, (double2 *)go, ctx->stream);
  if (rc) return rc;
  if (sigma) LDGA_HIP_TRY(ctx, hipMemcpy(sigma, so, Nq * (size_t)numax * 16, hipMemcpyDeviceToHost));
  if (G_ladder) LDGA_HIP_TRY(ctx, hipMemcpy(G_ladder, go, Nq * (size_t)fr->n_fft * 16, hipMemcpyDeviceToHost));
  return LDGA_OK;
}

int ldga_g_from_sigma_ladder_dev(ldga_ctx *ctx, const ldga_freq *fr, const ldga_model *mp, int numax, double lam_m,
                                 double lam_d, int fit_mu, double mu0, const void *chi_m, const void *chi_d,
                                 void *sigma, const void *sigma_prev, double mixing, const void *Sigma_loc,
                                 ldga_lambdadm_out *out, void *G_ladder, void *stream) {
  if (!ctx || !mp || !sigma) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  if (!chi_m != !chi_d) return ldga_fail(ctx, LDGA_ERR_ARG, "chi_m and chi_d are both given or both NULL");
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  return lambdadm_dev(ctx, fr, mp, numax, lam_m, lam_d, 0, 0.0, fit_mu, mu0, (const double *)chi_m, nullptr,
                      (const double *)chi_d, nullptr, nullptr, nullptr, (const double2 *)Sigma_loc, 0, out,
                      (double2 *)sigma, (double2 *)G_ladder, ldga_stream(ctx, stream), false,
                      (const double2 *)sigma_prev, mixing);
}

}  // extern "C"
