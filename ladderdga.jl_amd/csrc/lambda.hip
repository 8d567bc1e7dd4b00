// The λdm residual of LambdaCorrection/conditions_new.jl:250-287 on the GPU:
//   EoM Σ(λm, λd)  ->  μ refit (device secant)  ->  calc_E, filling_pos, Σχ sums
// and, on request, its Jacobian by implicit differentiation plus the EoM
// adjoint (launch_eom_adjoint).
//
// Every sum runs across the device: a reduction kernel leaves one partial per
// workgroup, and the last workgroup to finish (an atomic ticket, no waiting)
// adds the partials in a fixed order and finalises.  The μ secant is a short
// sequence of such launches, one filling evaluation each, with its state in
// device memory; a single-workgroup kernel continues it (and falls back to
// bracketing) only when those launches did not converge.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ldga_internal.h"
#include "ldga_reduce.h"

namespace {

using ldga_red::block_sum;
using ldga_red::device_sum;

constexpr int kFT = 1024;     // the single-workgroup fallback of the μ fit
constexpr int kRB = 256;      // threads per reduction workgroup
constexpr int kRGrid = 64;    // reduction workgroups (at most)
constexpr int kSecantSteps = 4;  // secant evaluations launched before the fallback kernel

// state of one residual evaluation, in doubles after the 20 reported values
enum : int {
  S_X0 = 24, S_X1, S_F0, S_F1, S_IT, S_DONE, S_STALL, S_MU, S_FMU, S_SM, S_DSM, S_SD, S_DSD, S_CMU,
  S_END
};
enum : int { C_MU0 = 0, C_STEP, C_EN, C_NCNT };  // ticket counters

struct FinalArgs {
  const double2 *sigma;  // Nq x numax
  const double *chim, *chid;
  int Nq, numax, Nw, n_iw;
  double Nk, beta, U, n, mu0, shell, lam_m, lam_d, c2m, c2d, tol;
  int fit_mu, want_jac;
  double *res;      // 20 reported doubles (n mu Ekin Epot1 Epot2 lhs res_pp res_epot | Sm' Sd' - Fmu Emu it conv
                    // n_full | adjoint contractions), then the state S_*
  double *part;     // partial sums: kRGrid x 8
  unsigned *ticket; // C_NCNT ticket counters
  double2 *W;       // Nq x numax adjoint weight (want_jac)
  double2 *Wk;      // numax: sum_k kMult W(k, nu)
};

// G(k, nu) = 1 / (i nu_n + mu - eps_k - Sigma(k, nu))
__device__ __forceinline__ double2 green(double nu_n, double mu, double eps, double2 sig) {
  return crecip(cz(mu - eps - sig.x, nu_n - sig.y));
}

// filling_pos with the shell correction (GFTools.jl:255-268) from Re sum_k G / Nk
__device__ __forceinline__ double filling_of(const FinalArgs &A, double mu, double reSG) {
  return 2.0 * (2.0 * reSG / A.beta + 0.5 + mu * A.shell) / (1.0 + A.U * A.shell);
}

// χ sums over all ω with their λ derivatives (DataTypes.jl:230-249), and the
// first two filling evaluations of the secant, f(mu0) and f(mu0 (1 + 1e-4))
__global__ __launch_bounds__(kRB) void ldm_start_kernel(OctTables T, FinalArgs A) {
  const double pib = M_PI / A.beta;
  const double x0 = A.mu0, x1 = (x0 == 0.0) ? 1e-4 : x0 * (1.0 + 1e-4);
  double v[6] = {0, 0, 0, 0, 0, 0};
  const int stride = gridDim.x * blockDim.x, t0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (A.chim && A.chid) {
    const int tot = A.Nq * A.Nw;
    for (int i = t0; i < tot; i += stride) {
      const double km = T.kmult[i % A.Nq];
      double xm = A.chim[i], xd = A.chid[i];
      if (A.lam_m != 0.0) xm = xm / (A.lam_m * xm + 1.0);
      if (A.lam_d != 0.0) xd = xd / (A.lam_d * xd + 1.0);
      v[0] = fma(km, xm, v[0]);
      v[1] = fma(km, -xm * xm, v[1]);
      v[2] = fma(km, xd, v[2]);
      v[3] = fma(km, -xd * xd, v[3]);
    }
  }
  if (A.fit_mu) {
    const int ne = A.Nq * A.numax;
    for (int i = t0; i < ne; i += stride) {
      const int q = (int)(i % A.Nq), nu = (int)(i / A.Nq);
      const double nun = (2.0 * nu + 1.0) * pib, km = T.kmult[q], e = T.eps[q];
      const double2 sig = A.sigma[i];
      v[4] = fma(km, green(nun, x0, e, sig).x, v[4]);
      v[5] = fma(km, green(nun, x1, e, sig).x, v[5]);
    }
  }
  double tot[6];
  if (!device_sum<6>(v, A.part, A.ticket + C_MU0, tot) || threadIdx.x != 0) return;
  double *r = A.res;
  // the 1/(i w_m)^2 tail over all w per unit c2 (DataTypes.jl:223-225); each sum has its own c2
  const double beta = A.beta;
  double tail1 = 0.0;
  for (int w = 0; w < A.Nw; ++w) {
    const int m = w - A.n_iw;
    if (m != 0) tail1 += -beta * beta / (4.0 * M_PI * M_PI * (double)m * (double)m);
  }
  r[S_SM] = (tot[0] / A.Nk - A.c2m * tail1 - beta * beta / 12.0 * A.c2m) / beta;
  r[S_DSM] = tot[1] / A.Nk / beta;
  r[S_SD] = (tot[2] / A.Nk - A.c2d * tail1 - beta * beta / 12.0 * A.c2d) / beta;
  r[S_DSD] = tot[3] / A.Nk / beta;
  r[S_IT] = 0.0;
  r[S_STALL] = 0.0;
  if (!A.fit_mu) {
    r[S_MU] = A.mu0;
    r[S_FMU] = 0.0;
    r[S_DONE] = 1.0;
    return;
  }
  const double f0 = filling_of(A, x0, tot[4] / A.Nk) - A.n, f1 = filling_of(A, x1, tot[5] / A.Nk) - A.n;
  r[S_X0] = x0;
  r[S_X1] = x1;
  r[S_F0] = f0;
  r[S_F1] = f1;
  r[S_MU] = x1;
  r[S_FMU] = f1;
  // the checks at the top of the secant loop (GFTools.jl:172 via Roots.find_zero)
  r[S_DONE] = fabs(f1) < A.tol ? 1.0 : 0.0;
  r[S_STALL] = (r[S_DONE] == 0.0 && f1 == f0) ? 1.0 : 0.0;
}

// one secant step: x2 = x1 - f1 (x1 - x0) / (f1 - f0), f(x2) across the device
__global__ __launch_bounds__(kRB) void ldm_secant_kernel(OctTables T, FinalArgs A) {
  double *r = A.res;
  if (r[S_DONE] != 0.0 || r[S_STALL] != 0.0) return;  // uniform over the grid
  const double x0 = r[S_X0], x1 = r[S_X1], f0 = r[S_F0], f1 = r[S_F1];
  const double x2 = x1 - f1 * (x1 - x0) / (f1 - f0);
  const double pib = M_PI / A.beta;
  double v[1] = {0.0};
  const int ne = A.Nq * A.numax, stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += stride) {
    const int q = (int)(i % A.Nq), nu = (int)(i / A.Nq);
    v[0] = fma(T.kmult[q], green((2.0 * nu + 1.0) * pib, x2, T.eps[q], A.sigma[i]).x, v[0]);
  }
  double tot[1];
  if (!device_sum<1>(v, A.part, A.ticket + C_STEP, tot) || threadIdx.x != 0) return;
  const double f2 = filling_of(A, x2, tot[0] / A.Nk) - A.n;
  r[S_X0] = x1;
  r[S_F0] = f1;
  r[S_X1] = x2;
  r[S_F1] = f2;
  r[S_MU] = x2;
  r[S_FMU] = f2;
  r[S_IT] += 1.0;
  r[S_DONE] = fabs(f2) < A.tol ? 1.0 : 0.0;
  r[S_STALL] = (r[S_DONE] == 0.0 && (f2 == f1 || r[S_IT] >= 100.0)) ? 1.0 : 0.0;
}

// One workgroup: the rest of the secant (to 100 steps), then, as a bracketing
// solver would, stepping out from mu0 until the filling changes sign and
// Illinois false position.  Returns at once when the secant has converged.
__global__ __launch_bounds__(kFT) void ldm_mu_fallback_kernel(OctTables T, FinalArgs A) {
  __shared__ double red[1][ldga_red::kMaxWaves];
  double *r = A.res;
  if (r[S_DONE] != 0.0) return;
  const int tid = threadIdx.x;
  const double pib = M_PI / A.beta, n = A.n;
  const int ne = A.Nq * A.numax;
  auto sumG = [&](double mu) -> double {
    double v[1] = {0.0};
    for (int i = tid; i < ne; i += kFT) {
      const int q = (int)(i % A.Nq), nu = (int)(i / A.Nq);
      v[0] = fma(T.kmult[q], green((2.0 * nu + 1.0) * pib, mu, T.eps[q], A.sigma[i]).x, v[0]);
    }
    block_sum<1>(v, red);
    return v[0] / A.Nk;
  };
  double x0 = r[S_X0], x1 = r[S_X1], f0 = r[S_F0], f1 = r[S_F1];
  int iters = (int)r[S_IT];
  bool conv = false;
  for (; iters < 100; ++iters) {
    if (fabs(f1) < A.tol) {
      conv = true;
      break;
    }
    if (f1 == f0) break;
    const double x2 = x1 - f1 * (x1 - x0) / (f1 - f0);
    x0 = x1;
    f0 = f1;
    x1 = x2;
    f1 = filling_of(A, x1, sumG(x1)) - n;
  }
  double mu = x1, fmu = f1;
  if (!conv) {
    double a = A.mu0, fa = filling_of(A, a, sumG(a)) - n;
    const double dir = fa < 0.0 ? 1.0 : -1.0;
    double h = 1.0, b = a + dir * h, fb = filling_of(A, b, sumG(b)) - n;
    for (int k = 0; k < 60 && !(fa * fb <= 0.0); ++k) {
      a = b;
      fa = fb;
      h *= 2.0;
      b = a + dir * h;
      fb = filling_of(A, b, sumG(b)) - n;
    }
    if (fa * fb <= 0.0) {
      int side = 0;
      for (int k = 0; k < 200; ++k, ++iters) {
        const double c = (a * fb - b * fa) / (fb - fa);
        const double fc = filling_of(A, c, sumG(c)) - n;
        mu = c;
        fmu = fc;
        if (fabs(fc) < A.tol || fabs(b - a) <= 4e-16 * fmax(fabs(a), fabs(b))) break;
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
  __syncthreads();  // every thread has read the state before it is written
  if (tid == 0) {
    r[S_MU] = mu;
    r[S_FMU] = fmu;
    r[S_IT] = (double)iters;
  }
}

// filling, calc_E (thermodynamics.jl:142-168) and, for the Jacobian,
// d(filling)/dmu and dE_pot/dmu; the last workgroup writes the report.  A fit
// that did not reach the reference's tolerance (find_zero atol = 1e-8,
// GFTools.jl:171-176) fails: mu = NaN, and every result is NaN.
__global__ __launch_bounds__(kRB) void ldm_energies_kernel(OctTables T, FinalArgs A) {
  double *r = A.res;
  const bool conv = !A.fit_mu || fabs(r[S_FMU]) <= 1e-8;
  const double mu = conv ? r[S_MU] : NAN;
  const double beta = A.beta, U = A.U, n = A.n, pib = M_PI / beta;
  const double Sh = n * U / 2.0;
  double v[5] = {0, 0, 0, 0, 0};  // Re sG, Ekin, Epot, Re sum -G^2 (dF/dmu), dEpot/dmu
  const int ne = A.Nq * A.numax, stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += stride) {
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
    v[0] = fma(km, g.x, v[0]);
    v[1] = fma(km, 4.0 * e * (g.x - Ekt / inu2), v[1]);
    v[2] = fma(km, 2.0 * (gs.x - Ept / inu2), v[2]);
    v[3] = fma(km, -g2.x, v[3]);
    v[4] = fma(km, 2.0 * (-cmul(g2, sig).x + Sh / inu2), v[4]);
    if (nu == 0) {  // the per-k tail terms, once per k
      v[1] = fma(km, (beta / 2.0) * e * (1.0 - beta * Ekt), v[1]);
      v[2] = fma(km, (beta / 2.0) * Sh - (beta * beta / 4.0) * Ept, v[2]);
      v[4] = fma(km, beta * beta * Sh / 4.0, v[4]);
    }
  }
  double tot[5];
  if (!device_sum<5>(v, A.part, A.ticket + C_EN, tot) || threadIdx.x != 0) return;
  const double Nk = A.Nk;
  const bool have_chi = A.chim != nullptr && A.chid != nullptr;
  const double Sm = r[S_SM], Sd = r[S_SD];
  const double nval = filling_of(A, mu, tot[0] / Nk);
  const double Ekin = tot[1] / Nk / beta, Epot1 = tot[2] / Nk / beta;
  const double lhs = have_chi ? 0.5 * (Sd + Sm) : NAN;
  const double Epot2 = have_chi ? 0.5 * U * (Sd - Sm) + U * (n / 2.0) * (n / 2.0) : NAN;
  r[0] = nval;
  r[1] = mu;
  r[2] = Ekin;
  r[3] = Epot1;
  r[4] = Epot2;
  r[5] = lhs;
  r[6] = lhs - n / 2.0 * (1.0 - n / 2.0);
  r[7] = Epot1 - Epot2;
  r[8] = r[S_DSM];
  r[9] = r[S_DSD];
  r[13] = r[S_IT];
  r[14] = conv ? 1.0 : 0.0;
  r[15] = NAN;  // the full filling, written by nfull_kernel when G_ladder is formed
  const double Fmu = 2.0 * (2.0 * (tot[3] / Nk) / beta + A.shell) / (1.0 + U * A.shell);
  const double Emu = tot[4] / Nk / beta;
  r[11] = Fmu;
  r[12] = Emu;
  r[S_CMU] = -(Emu / Fmu) * 4.0 / (beta * (1.0 + U * A.shell));
}

// adjoint weight W = dR_epot/dSigma (a full-grid function, wedge values), elementwise
__global__ __launch_bounds__(kRB) void ldm_adjw_kernel(OctTables T, FinalArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.Nq * A.numax) return;
  const double mu = A.res[1], cmu = A.res[S_CMU], beta = A.beta;
  const int q = (int)(i % A.Nq), nu = (int)(i / A.Nq);
  const double2 sig = A.sigma[i];
  const double2 g = green((2.0 * nu + 1.0) * M_PI / beta, mu, T.eps[q], sig);
  const double2 g2 = cmul(g, g);
  const double2 wv = cadd(cscale(cadd(g, cmul(sig, g2)), 2.0 / beta), cscale(g2, cmu));
  A.W[i] = cscale(wv, 1.0 / A.Nk);
}

// Wk[nu] = sum_k kMult W(k, nu): one workgroup per nu
__global__ __launch_bounds__(kRB) void ldm_wk_kernel(OctTables T, FinalArgs A) {
  __shared__ double red[2][ldga_red::kMaxWaves];
  const int nu = blockIdx.x;
  double v[2] = {0.0, 0.0};
  for (int q = threadIdx.x; q < A.Nq; q += blockDim.x) {
    const double2 w = A.W[q + (size_t)A.Nq * nu];
    v[0] = fma(T.kmult[q], w.x, v[0]);
    v[1] = fma(T.kmult[q], w.y, v[1]);
  }
  block_sum<2>(v, red);
  if (threadIdx.x == 0) A.Wk[nu] = cz(v[0], v[1]);
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
  double2 *sigma = sigma_out ? sigma_out : (double2 *)ldga_scratch(ctx, "ldm_sigma", (size_t)Nq * numax * 16, st);
  double *res = (double *)ldga_scratch(ctx, "ldm_res", S_END * 8, st);
  double *part = (double *)ldga_scratch(ctx, "ldm_part", (size_t)kRGrid * 8 * 8, st);
  unsigned *ticket = ldga_tickets(ctx, "ldm_ticket", C_NCNT, st);
  if (!sigma || !res || !part || !ticket) return LDGA_ERR_HIP;
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
  A.part = part;
  A.ticket = ticket;
  double2 *adj = (double2 *)(res + 16);
  if (want_jac) {
    A.W = (double2 *)ldga_scratch(ctx, "ldm_W", (size_t)Nq * numax * 16, st);
    A.Wk = (double2 *)ldga_scratch(ctx, "ldm_Wk", (size_t)numax * 16, st);
    if (!A.W || !A.Wk) return LDGA_ERR_HIP;
  }
  // the reductions: enough workgroups for one element per thread, at most kRGrid
  const size_t ne = (size_t)Nq * numax, nchi = (size_t)Nq * Nw;
  const unsigned g_fit = (unsigned)std::min<size_t>(kRGrid, (std::max(ne, nchi) + kRB - 1) / kRB);
  const unsigned g_ne = (unsigned)std::min<size_t>(kRGrid, (ne + kRB - 1) / kRB);
  ldm_start_kernel<<<g_fit, kRB, 0, st>>>(ctx->oct, A);
  if (fit_mu) {
    for (int k = 0; k < kSecantSteps; ++k) ldm_secant_kernel<<<g_ne, kRB, 0, st>>>(ctx->oct, A);
    ldm_mu_fallback_kernel<<<1, kFT, 0, st>>>(ctx->oct, A);
  }
  ldm_energies_kernel<<<g_ne, kRB, 0, st>>>(ctx->oct, A);
  if (want_jac) {
    ldm_adjw_kernel<<<(unsigned)((ne + kRB - 1) / kRB), kRB, 0, st>>>(ctx->oct, A);
    ldm_wk_kernel<<<(unsigned)numax, kRB, 0, st>>>(ctx->oct, A);
  }
  LDGA_HIP_TRY(ctx, hipGetLastError());
  if (want_jac) {
    rc = launch_eom_adjoint(ctx, fr, mp, numax, lam_m, lam_d, chi_m, gamma_m, chi_d, gamma_d, Grfft, A.W, adj, st);
    if (rc) return rc;
  }
  if (G_out) {
    const size_t tot = (size_t)Nq * fr->n_fft;
    const unsigned nb = (unsigned)((tot + kGL - 1) / kGL);
    double *part = (double *)ldga_scratch(ctx, "ldm_gl_part", (size_t)nb * 8, st);
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
