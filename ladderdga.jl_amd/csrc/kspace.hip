// k-space convolutions of the ladder path on the real-space octant:
//   * bubble  chi0(q,nu,w)  = -beta conv_k(G(nu), G(nu+w))      (BSETools.jl:157-172)
//   * EoM     Sigma(k,nu)   = 1/beta sum_w conv_q(K, G(nu+w))   (ladderDGATools_singleCore.jl:49-81)
//   * lambda0 (plain)                                           (BSETools.jl:128-141)
//   * sum_kw with the chi_lambda map                            (DataTypes.jl:215-249)
//
// conv(a,b)[q] = (1/Nk) sum_k a(k) b(k+q).  With A = fft(a) and B = fft(b)
// (B(-R) read from the reversed-table fft(reverse(b)) times exp(-2 pi i sum R/Ns))
// conv[q] = 1/Nk^2 sum_R A(R) B(-R) exp(-i q.R).  All functions of the path
// are even in every axis, so the sum folds onto the octant R in [0,Ns/2]^D
// with multiplicity w(R) and a separable cosine kernel: one small dense
// "mode product" per axis, evaluated in LDS; the wedge points are then read
// from the transformed octant.
#include <algorithm>

#include "ldga_internal.h"
#include "ldga_reduce.h"

namespace {

constexpr int kNT = 256;  // threads per octant workgroup

// In-place mode product along axis `ax` over the whole octant held in LDS:
//   buf[o] <- sum_r cos(2 pi q_ax r / Ns) * buf[o with coordinate ax = r].
// The octant splits into lines along `ax` (the points sharing the other
// coordinates).  Outputs q and Ns/2 - q share their products, since
// cos(2 pi (Ns/2 - q) r / Ns) = (-1)^r cos(2 pi q r / Ns): with E_q and O_q the
// sums over even and odd r, y[q] = E_q + O_q and y[Ns/2 - q] = E_q - O_q, for
// q = 0 .. Ns/4.  Thread (line, q-block) reads the n values of its line once
// and forms kQ such pairs; the cosines come from the Chebyshev recurrence
// c(r+1) = 2 cos(theta) c(r) - c(r-1), seeded with the exact cos(theta) = cs[q]
// (error below 2e-14 for Ns <= 64).  Lines go in rounds of kNT / nqb lines; a
// round writes only its own lines, so one barrier per round separates its
// reads from its writes.
constexpr int kQ = 4;
__device__ __forceinline__ void oct_pass(double2 *buf, const double *cs, int ax, const OctTables &T) {
  const int n = T.n, half = T.Ns / 2, stride = T.stride[ax];
  const int np = T.Ns / 4 + 1;  // q = 0 .. Ns/4 with their partners Ns/2 - q cover 0 .. Ns/2
  const int nqb = (np + kQ - 1) / kQ, nlines = T.NO / n, per_round = kNT / nqb;
  const int li = threadIdx.x / nqb, q0 = (threadIdx.x - li * nqb) * kQ;
  double tw[kQ];
#pragma unroll
  for (int j = 0; j < kQ; ++j) tw[j] = (q0 + j < np) ? 2.0 * cs[q0 + j] : 0.0;
  for (int l0 = 0; l0 < nlines; l0 += per_round) {
    const int line = l0 + li;
    const bool active = li < per_round && line < nlines;
    // octant index of the line's point with coordinate 0 on axis ax
    const int base = (line % stride) + (line / stride) * stride * n;
    double2 E[kQ], O[kQ];
#pragma unroll
    for (int j = 0; j < kQ; ++j) E[j] = O[j] = cz(0.0, 0.0);
    if (active) {
      double c[kQ], cm[kQ];
#pragma unroll
      for (int j = 0; j < kQ; ++j) {
        c[j] = 1.0;            // cos(0)
        cm[j] = 0.5 * tw[j];   // cos(-theta)
      }
      for (int r = 0; r < n; r += 2) {
        const double2 xe = buf[base + r * stride];
        const double2 xo = (r + 1 < n) ? buf[base + (r + 1) * stride] : cz(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < kQ; ++j) {
          E[j].x = fma(c[j], xe.x, E[j].x);
          E[j].y = fma(c[j], xe.y, E[j].y);
          const double c1 = fma(tw[j], c[j], -cm[j]);
          O[j].x = fma(c1, xo.x, O[j].x);
          O[j].y = fma(c1, xo.y, O[j].y);
          cm[j] = c1;
          c[j] = fma(tw[j], c1, -c[j]);
        }
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int j = 0; j < kQ; ++j) {
        const int q = q0 + j;
        if (q < np) {
          buf[base + q * stride] = cadd(E[j], O[j]);
          if (half - q != q) buf[base + (half - q) * stride] = csub(E[j], O[j]);
        }
      }
    }
  }
  __syncthreads();
}

__device__ __forceinline__ int trunc_half(int x) { return x / 2; }  // C division truncates toward 0 like Julia trunc

// Wedge points per thread (q = threadIdx.x + k kNT): Nq <= 969 for the grids
// ldga_set_kgrid accepts (2D Ns <= 64: 561, 3D Ns <= 32: 969).
constexpr int kMaxQ = 4;

// This thread's octant points o = threadIdx.x + p kNT, p < MAXPT: their index
// in the full grid and octphase * octw, loaded once with every load in flight
// together (the points past NO get index 0 and weight 0).
template <int MAXPT>
struct OctPts {
  int f[MAXPT];
  double2 pw[MAXPT];
  __device__ __forceinline__ void load(const OctTables &T) {
#pragma unroll
    for (int p = 0; p < MAXPT; ++p) {
      const int o = threadIdx.x + p * kNT;
      const bool v = o < T.NO;
      f[p] = v ? T.octfull[o] : 0;
      pw[p] = v ? T.octpw[o] : cz(0.0, 0.0);
    }
  }
};

// the octant index of this thread's wedge points (index 0 past Nq), for the
// final reads of the transformed octant
__device__ __forceinline__ void load_w2o(const OctTables &T, int (&w2o)[kMaxQ]) {
#pragma unroll
  for (int k = 0; k < kMaxQ; ++k) {
    const int q = threadIdx.x + k * kNT;
    w2o[k] = q < T.Nq ? T.wedge2oct[q] : 0;
  }
}

// ---------------------------------------------------------------------------
// bubble: one workgroup per (nu, w) pair (the octant transform on all D axes, then
// the wedge points read from it)
// ---------------------------------------------------------------------------
template <int MAXPT>
__global__ __launch_bounds__(kNT) void bubble_kernel(OctTables T, int Nnu, int n_inu_tot, int n_iw, int shift,
                                                     int fft_first, const double2 *__restrict__ Gfft,
                                                     const double2 *__restrict__ Grfft, double2 *__restrict__ chi0,
                                                     double scale) {
  extern __shared__ double2 smem[];
  double2 *buf = smem;
  double *cs = reinterpret_cast<double *>(smem + T.NO);
  const int pair = blockIdx.x;
  const int nui = pair % Nnu, wi = pair / Nnu;
  const int wn = wi - n_iw;
  const int nun = -n_inu_tot + nui - trunc_half(shift * wn);
  const size_t fs = (size_t)T.Nk;
  const double2 *a = Gfft + (size_t)(nun - fft_first) * fs;
  const double2 *b = Grfft + (size_t)(nun + wn - fft_first) * fs;
  for (int m = threadIdx.x; m < T.Ns; m += kNT) cs[m] = T.cosT[m];
  int w2o[kMaxQ];
  load_w2o(T, w2o);  // used after the passes: their latency hides the loads
  OctPts<MAXPT> P;
  P.load(T);
  // the products in batches of up to 10 points, every load of a batch in flight together
  constexpr int B = MAXPT < 10 ? MAXPT : 10;
#pragma unroll
  for (int p0 = 0; p0 < MAXPT; p0 += B) {
    double2 va[B], vb[B];
#pragma unroll
    for (int k = 0; k < B; ++k) {
      if (p0 + k < MAXPT) {
        va[k] = a[P.f[p0 + k]];
        vb[k] = b[P.f[p0 + k]];
      }
    }
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const int o = threadIdx.x + (p0 + k) * kNT;
      if (p0 + k < MAXPT && o < T.NO) buf[o] = cmul(cmul(va[k], vb[k]), P.pw[p0 + k]);
    }
  }
  __syncthreads();
  for (int ax = 0; ax < T.D; ++ax) oct_pass(buf, cs, ax, T);
  double2 *out = chi0 + (size_t)T.Nq * (nui + (size_t)Nnu * wi);
#pragma unroll
  for (int k = 0; k < kMaxQ; ++k) {
    const int q = threadIdx.x + k * kNT;
    if (q < T.Nq) out[q] = cscale(buf[w2o[k]], scale);
  }
}

// ---------------------------------------------------------------------------
// EoM: partial sums over an omega chunk, then the inverse transform per nu
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 eom_value(int kind, double U, double2 gm, double2 gd, double cm, double cd,
                                             double2 l0) {
  // src/ladderDGATools.jl:260-266
  switch (kind) {
    case LDGA_EOM_CHI_M:
      return cscale(gm, U * 1.5 * U * cm);
    case LDGA_EOM_GAMMA_M:
      return cscale(gm, U * 1.5);
    case LDGA_EOM_CHI_D:
      return cscale(gd, -U * 0.5 * (-U * cd));
    case LDGA_EOM_GAMMA_D:
      return cscale(gd, -U * 0.5);
    case LDGA_EOM_REST_01:
      return cz(-U, 0.0);
    case LDGA_EOM_REST:
      return cscale(l0, U);
    default: {
      const double am = 1.5 * (1.0 + U * cm), ad = 0.5 * (1.0 - U * cd);
      return cscale(cz(gm.x * am - gd.x * ad - 1.5 + 0.5 + l0.x, gm.y * am - gd.y * ad + l0.y), U);
    }
  }
}

template <int MAXPT>
__global__ __launch_bounds__(kNT) void eom_accum_kernel(OctTables T, int Nw, int Nnu, int n_inu, int n_iw, int shift,
                                                        int fft_first, int numax, int wchunk, int kind, double U,
                                                        double lam_m, double lam_d, const double *__restrict__ chim,
                                                        const double *__restrict__ chid,
                                                        const double2 *__restrict__ gm,
                                                        const double2 *__restrict__ gd,
                                                        const double2 *__restrict__ l0,
                                                        const double2 *__restrict__ Grfft,
                                                        double2 *__restrict__ partial) {
  extern __shared__ double2 smem[];
  double2 *buf = smem;
  double2 *Kw = smem + T.NO;  // K on the wedge points
  double *cs = reinterpret_cast<double *>(Kw + T.Nq);
  const int nuii = blockIdx.x % numax;
  const int chunk = blockIdx.x / numax;
  for (int m = threadIdx.x; m < T.Ns; m += kNT) cs[m] = T.cosT[m];
  double2 S[MAXPT];
#pragma unroll
  for (int p = 0; p < MAXPT; ++p) S[p] = cz(0.0, 0.0);
  // this thread's octant points: full-grid index, phase x weight, wedge index and
  // weight, loaded once for the whole omega chunk
  OctPts<MAXPT> P;
  P.load(T);
  int o2w[MAXPT];
  double ow[MAXPT];
#pragma unroll
  for (int p = 0; p < MAXPT; ++p) {
    const int o = threadIdx.x + p * kNT;
    o2w[p] = o < T.NO ? T.oct2wedge[o] : 0;
    ow[p] = o < T.NO ? T.octw[o] : 0.0;
  }
  const int w0 = chunk * wchunk, w1 = min(Nw, w0 + wchunk);
  const size_t fs = (size_t)T.Nk;
  for (int wi = w0; wi < w1; ++wi) {
    const int wn = wi - n_iw;
    // nu-window of gamma at this w: nu0 index (1-based) = n_inu + shift*trunc(wn/2) + 1  (helpers.jl:58)
    const int nui = n_inu + shift * trunc_half(wn) + nuii;  // 0-based index of nu_n = nuii
    if (nui > 2 * n_inu - 1) continue;                     // nu list ends at 2 n_inu (singleCore.jl:59-60)
    // K on the wedge points, read once and coalesced (q is the fastest index of
    // gamma, lambda0 and chi), then expanded to the weighted octant from LDS
    for (int q = threadIdx.x; q < T.Nq; q += kNT) {
      double cm = chim[q + (size_t)T.Nq * wi], cd = chid[q + (size_t)T.Nq * wi];
      if (lam_m != 0.0) cm = cm / (lam_m * cm + 1.0);
      if (lam_d != 0.0) cd = cd / (lam_d * cd + 1.0);
      const size_t idx = q + (size_t)T.Nq * (nui + (size_t)Nnu * wi);
      Kw[q] = eom_value(kind, U, gm[idx], gd[idx], cm, cd, l0[idx]);
    }
    __syncthreads();  // Kw written; the previous iteration is done with buf
#pragma unroll
    for (int p = 0; p < MAXPT; ++p) {
      const int o = threadIdx.x + p * kNT;
      if (o < T.NO) buf[o] = cscale(Kw[o2w[p]], ow[p]);
    }
    __syncthreads();
    // G(nu + w) at this thread's points, loaded before the passes, which hide the latency
    const double2 *g = Grfft + (size_t)(nuii + wn - fft_first) * fs;
    double2 gv[MAXPT];
#pragma unroll
    for (int p = 0; p < MAXPT; ++p) gv[p] = g[P.f[p]];
    for (int ax = 0; ax < T.D; ++ax) oct_pass(buf, cs, ax, T);
#pragma unroll
    for (int p = 0; p < MAXPT; ++p) {
      const int o = threadIdx.x + p * kNT;
      if (o < T.NO) S[p] = cadd(S[p], cmul(buf[o], cmul(gv[p], P.pw[p])));
    }
  }
  double2 *dst = partial + (size_t)T.NO * (nuii + (size_t)numax * chunk);
#pragma unroll
  for (int p = 0; p < MAXPT; ++p) {
    const int o = threadIdx.x + p * kNT;
    if (o < T.NO) dst[o] = S[p];
  }
}

template <int MAXPT>
__global__ __launch_bounds__(kNT) void eom_final_kernel(OctTables T, int numax, int nchunks,
                                                        const double2 *__restrict__ partial, double scale, double U,
                                                        double beta, double hartree, int add_tail,
                                                        const double *__restrict__ tail_sum, double chiloc_m_sum,
                                                        double2 *__restrict__ sigma) {
  extern __shared__ double2 smem[];
  double2 *buf = smem;
  double *cs = reinterpret_cast<double *>(smem + T.NO);
  const int nuii = blockIdx.x;
  for (int m = threadIdx.x; m < T.Ns; m += kNT) cs[m] = T.cosT[m];
  int w2o[kMaxQ];
  load_w2o(T, w2o);
  // the chunk partials of this nu, four chunks in flight per step
  const size_t cstride = (size_t)T.NO * numax;
  const double2 *pp = partial + (size_t)T.NO * nuii;
  for (int o = threadIdx.x; o < T.NO; o += kNT) {
    double2 s = cz(0.0, 0.0);
    int c = 0;
    for (; c + 4 <= nchunks; c += 4) {
      const double2 a0 = pp[cstride * c + o], a1 = pp[cstride * (c + 1) + o];
      const double2 a2 = pp[cstride * (c + 2) + o], a3 = pp[cstride * (c + 3) + o];
      s = cadd(s, cadd(cadd(a0, a1), cadd(a2, a3)));
    }
    for (; c < nchunks; ++c) s = cadd(s, pp[cstride * c + o]);
    buf[o] = s;
  }
  __syncthreads();
  for (int ax = 0; ax < T.D; ++ax) oct_pass(buf, cs, ax, T);
  // tail/(i nu_n) = -i * tail * beta / ((2n+1) pi)   (singleCore.jl:78-79)
  double tail = 0.0;
  if (add_tail) tail = -U * (tail_sum[0] - chiloc_m_sum);
  const double tail_im = -tail * beta / ((2.0 * nuii + 1.0) * M_PI);
  double2 *out = sigma + (size_t)T.Nq * nuii;
#pragma unroll
  for (int k = 0; k < kMaxQ; ++k) {
    const int q = threadIdx.x + k * kNT;
    if (q < T.Nq) {
      const double2 v = cscale(buf[w2o[k]], scale);
      out[q] = cz(v.x + hartree, v.y + tail_im);
    }
  }
}

// ---------------------------------------------------------------------------
// lambda0 (plain sums): fsum[nu',w] = sum_nu F[nu,nu',w]; lambda0 = chi0 fsum / beta^2
// ---------------------------------------------------------------------------
__global__ void fsum_kernel(int Nnu, int Nw, const double2 *__restrict__ F, double2 *__restrict__ fsum) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;  // (nu', w)
  if (col >= Nnu * Nw) return;
  const double2 *c = F + (size_t)Nnu * col;
  double2 s = cz(0.0, 0.0);
  for (int v = 0; v < Nnu; ++v) s = cadd(s, c[v]);
  fsum[col] = s;
}

__global__ void lambda0_kernel(int Nq, int Nnu, int Nw, int shell, double inv_b2, const double2 *__restrict__ fsum,
                               const double2 *__restrict__ chi0, double2 *__restrict__ l0) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)Nq * Nnu * Nw;
  if (i >= total) return;
  const int q = (int)(i % Nq);
  const size_t col = i / Nq;  // nu' + Nnu w
  const int nu = (int)(col % Nnu), w = (int)(col / Nnu);
  const double2 c = chi0[q + (size_t)Nq * (shell + nu + (size_t)(Nnu + 2 * shell) * w)];
  l0[i] = cscale(cmul(fsum[col], c), inv_b2);
}

// ---------------------------------------------------------------------------
// sum_kw: a device-wide FP64 reduction over the usable (q, w) block; the last
// workgroup adds the tail terms and writes the three sums
// ---------------------------------------------------------------------------
constexpr int kSumNT = 1024;  // the single-workgroup k integral below
constexpr int kSumRB = 256, kSumGrid = 64;
__global__ __launch_bounds__(kSumRB) void sum_kw_kernel(const double *__restrict__ chi, const double *__restrict__ kmult,
                                                        int Nq, int w_lo, int w_hi, double Nk, double lam, double c2,
                                                        double beta, int n_iw, double *part, unsigned *ticket,
                                                        double *__restrict__ out) {
  double v[2] = {0.0, 0.0};
  const int nw = w_hi - w_lo + 1, total = Nq * nw, stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    const int q = i % Nq;
    const double x = chi[i + Nq * w_lo];  // chi[q + Nq w], w = w_lo + i / Nq
    const double f = (lam != 0.0) ? x / (lam * x + 1.0) : x;
    v[0] = fma(kmult[q], f, v[0]);
    v[1] = fma(kmult[q], -f * f, v[1]);
  }
  double tot[2];
  if (!ldga_red::device_sum<2>(v, part, ticket, tot) || threadIdx.x != 0) return;
  // Re(c2/(i w_n)^2) = -c2 beta^2 / (4 pi^2 m^2), zero at m = 0  (DataTypes.jl:223-225)
  double tail = 0.0;
  for (int w = w_lo; w <= w_hi; ++w) {
    const int m = w - n_iw;
    if (m != 0) tail += -c2 * beta * beta / (4.0 * M_PI * M_PI * (double)m * (double)m);
  }
  const double konst = -beta * beta / 12.0 * c2;
  out[0] = (tot[0] / Nk - tail + konst) / beta;
  out[1] = (tot[1] / Nk - tail + konst) / beta;
  out[2] = (tot[1] / Nk) / beta;
}

// ---------------------------------------------------------------------------
// sum_ω per q (DataTypes.jl:285-305): one thread per q walks the usable slice
// w_lo..w_hi, pairing slice element n with the n-th frequency m = n - n_iw of
// the whole grid for the 1/(i w_m)^2 tail, as the reference's loop does;
// res[q] = (sum_n [chi_λ - Re(c2/(i w_m)^2)] - c2 beta^2/12) / beta.
// Threads run over q, the fastest index of chi, so every load is coalesced.
// ---------------------------------------------------------------------------
__global__ void sum_w_kernel(const double *__restrict__ chi, int Nq, int w_lo, int w_hi, int n_iw, double lam,
                             double c2, double beta, double *__restrict__ res) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Nq) return;
  const double tfac = -c2 * beta * beta / (4.0 * M_PI * M_PI);  // Re(c2/(i w_m)^2) = tfac / m^2
  double r = 0.0;
  for (int n = 0; n <= w_hi - w_lo; ++n) {
    double x = chi[q + (size_t)Nq * (w_lo + n)];
    if (lam != 0.0) x = x / (lam * x + 1.0);
    const int m = n - n_iw;
    r += (m != 0) ? x - tfac / ((double)m * (double)m) : x;
  }
  res[q] = (r + c2 * (-beta * beta / 12.0)) / beta;
}

// kintegrate(res) = sum_q kMult[q] res[q] / sum_q kMult[q]: one workgroup (Nq <= a few thousand)
__global__ __launch_bounds__(kSumNT) void kintegrate_kernel(const double *__restrict__ res,
                                                            const double *__restrict__ kmult, int Nq, double Nk,
                                                            double *__restrict__ out) {
  __shared__ double red[kSumNT / 64];
  double s = 0.0;
  for (int q = threadIdx.x; q < Nq; q += kSumNT) s = fma(kmult[q], res[q], s);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0;
    for (int i = 0; i < kSumNT / 64; ++i) a += red[i];
    out[0] = a / Nk;
  }
}

// ---------------------------------------------------------------------------
// Adjoint of the EoM with respect to (lam_m, lam_d).  Given the weight W(k,nu)
// = dR/dSigma(k,nu) (a symmetric function on the full grid, stored on the
// wedge), sum_k W(k) conv(K', G)(k) = sum_q K'(q) conv(W, G)(q), so one
// transposed convolution T = conv(W(nu), G(nu+w)) per (nu, w) contracts with
//   dK/dlam_m = -1.5 U^2 gamma_m (chi_m^lam)^2,  dK/dlam_d = -0.5 U^2 gamma_d (chi_d^lam)^2
// (the derivatives of eom, src/ladderDGATools.jl:260, through chi/(lam chi + 1)).
// Per workgroup: one nu, an omega chunk; partial sums of sum_q kMult K' T.
// ---------------------------------------------------------------------------
template <int MAXPT>
__global__ __launch_bounds__(kNT) void eom_adjoint_kernel(OctTables T, int Nw, int Nnu, int n_inu, int n_iw, int shift,
                                                          int fft_first, int numax, int wchunk, double U, double lam_m,
                                                          double lam_d, const double *__restrict__ chim,
                                                          const double *__restrict__ chid,
                                                          const double2 *__restrict__ gm,
                                                          const double2 *__restrict__ gd,
                                                          const double2 *__restrict__ Grfft,
                                                          const double2 *__restrict__ W, double scale,
                                                          double2 *__restrict__ partial) {
  extern __shared__ double2 smem[];
  double2 *buf = smem;
  double *cs = reinterpret_cast<double *>(smem + T.NO);
  __shared__ double2 red[2][kNT / 64];
  const int nuii = blockIdx.x % numax;
  const int chunk = blockIdx.x / numax;
  for (int m = threadIdx.x; m < T.Ns; m += kNT) cs[m] = T.cosT[m];
  // forward transform of W(., nu) (weights w(q) of the octant fold)
  for (int o = threadIdx.x; o < T.NO; o += kNT)
    buf[o] = cscale(W[T.oct2wedge[o] + (size_t)T.Nq * nuii], T.octw[o]);
  __syncthreads();
  for (int ax = 0; ax < T.D; ++ax) oct_pass(buf, cs, ax, T);
  double2 What[MAXPT];
#pragma unroll
  for (int p = 0; p < MAXPT; ++p) {
    const int o = threadIdx.x + p * kNT;
    What[p] = (o < T.NO) ? buf[o] : cz(0.0, 0.0);
  }
  double2 am = cz(0.0, 0.0), ad = cz(0.0, 0.0);
  const int w0 = chunk * wchunk, w1 = min(Nw, w0 + wchunk);
  const size_t fs = (size_t)T.Nk;
  for (int wi = w0; wi < w1; ++wi) {
    const int wn = wi - n_iw;
    const int nui = n_inu + shift * trunc_half(wn) + nuii;
    if (nui > 2 * n_inu - 1) continue;
    const double2 *g = Grfft + (size_t)(nuii + wn - fft_first) * fs;
    __syncthreads();  // previous users of buf are done
#pragma unroll
    for (int p = 0; p < MAXPT; ++p) {
      const int o = threadIdx.x + p * kNT;
      if (o < T.NO) buf[o] = cscale(cmul(What[p], cmul(g[T.octfull[o]], T.octphase[o])), T.octw[o]);
    }
    __syncthreads();
    for (int ax = 0; ax < T.D; ++ax) oct_pass(buf, cs, ax, T);
    for (int q = threadIdx.x; q < T.Nq; q += kNT) {
      const double2 t = cscale(buf[T.wedge2oct[q]], scale);
      double cm = chim[q + (size_t)T.Nq * wi], cd = chid[q + (size_t)T.Nq * wi];
      if (lam_m != 0.0) cm = cm / (lam_m * cm + 1.0);
      if (lam_d != 0.0) cd = cd / (lam_d * cd + 1.0);
      const size_t idx = q + (size_t)T.Nq * (nui + (size_t)Nnu * wi);
      const double km = T.kmult[q];
      am = cadd(am, cscale(cmul(gm[idx], t), -1.5 * U * U * cm * cm * km));
      ad = cadd(ad, cscale(cmul(gd[idx], t), -0.5 * U * U * cd * cd * km));
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    am.x += __shfl_xor(am.x, off);
    am.y += __shfl_xor(am.y, off);
    ad.x += __shfl_xor(ad.x, off);
    ad.y += __shfl_xor(ad.y, off);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = am;
    red[1][threadIdx.x >> 6] = ad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 sm = cz(0.0, 0.0), sd = cz(0.0, 0.0);
    for (int i = 0; i < kNT / 64; ++i) {
      sm = cadd(sm, red[0][i]);
      sd = cadd(sd, red[1][i]);
    }
    partial[2 * blockIdx.x] = sm;
    partial[2 * blockIdx.x + 1] = sd;
  }
}

__global__ void sum_pairs_kernel(int n, const double2 *__restrict__ partial, double2 *__restrict__ out) {
  __shared__ double2 red[2][4];
  double2 a = cz(0.0, 0.0), b = cz(0.0, 0.0);
  for (int i = threadIdx.x; i < n; i += 256) {
    a = cadd(a, partial[2 * i]);
    b = cadd(b, partial[2 * i + 1]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    a.x += __shfl_xor(a.x, off);
    a.y += __shfl_xor(a.y, off);
    b.x += __shfl_xor(b.x, off);
    b.y += __shfl_xor(b.y, off);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < 4; ++i) {
      red[0][0] = cadd(red[0][0], red[0][i]);
      red[1][0] = cadd(red[1][0], red[1][i]);
    }
    out[0] = red[0][0];
    out[1] = red[1][0];
  }
}

// ---------------------------------------------------------------------------
// G tables (src/helpers.jl:153-183): G_fft = fft(expand(G)) and
// G_rfft = fft(reverse(expand(G))) for every nu of fft_range, from G on the
// wedge.  G is even in every axis, so fft(expand(G)) is the octant cosine
// transform of the weighted octant values, and is itself even in R.
// reverse() shifts every axis by one point (reverse(G)[i] = G[i + 1] for an
// even G), so G_rfft(R) = exp(2 pi i sum_d R_d / Ns) G_fft(R).  One workgroup
// per nu: the octant in LDS, D passes, then both full grids written.
// ---------------------------------------------------------------------------
template <int MAXPT>
__global__ __launch_bounds__(kNT) void gtables_kernel(OctTables T, const double2 *__restrict__ Gw,
                                                      double2 *__restrict__ Gfft, double2 *__restrict__ Grfft) {
  extern __shared__ double2 smem[];
  double2 *buf = smem;
  double *cs = reinterpret_cast<double *>(smem + T.NO);
  const int f = blockIdx.x;  // position in fft_range
  const double2 *g = Gw + (size_t)T.Nq * f;
  for (int m = threadIdx.x; m < T.Ns; m += kNT) cs[m] = T.cosT[m];
  for (int o = threadIdx.x; o < T.NO; o += kNT) buf[o] = cscale(g[T.oct2wedge[o]], T.octw[o]);
  __syncthreads();
  for (int ax = 0; ax < T.D; ++ax) oct_pass(buf, cs, ax, T);
  // the full grid, first axis fastest (column-major like the reference's arrays)
  double2 *a = Gfft + (size_t)T.Nk * f, *b = Grfft + (size_t)T.Nk * f;
  const int Nk = (int)T.Nk;
  for (int i = threadIdx.x; i < Nk; i += kNT) {
    int rem = i, o = 0, sum = 0;
    for (int d = 0; d < T.D; ++d) {
      const int r = rem % T.Ns;
      rem /= T.Ns;
      o += min(r, T.Ns - r) * T.stride[d];
      sum += r;
    }
    const double2 v = buf[o];
    a[i] = v;
    double sn, cn;
    sincospi(2.0 * (double)(sum % T.Ns) / (double)T.Ns, &sn, &cn);
    b[i] = cmul(v, cz(cn, sn));
  }
}

}  // namespace

// number of register slots per thread needed for the octant (NO / 256 rounded up)
static int oct_maxpt(int NO) { return (NO + kNT - 1) / kNT; }

// Launch KERNEL<MAXPT> with the smallest register slot count that covers the
// octant; large octants (3D) need more than the default 64 KiB of LDS.
#define OCT_DISPATCH(NO, SMEM, KERNEL, GRID, STREAM, ...)                                          \
  do {                                                                                             \
    const int mp_ = oct_maxpt(NO);                                                                 \
    auto go_ = [&](auto kern) -> int {                                                             \
      if ((SMEM) > 65536)                                                                          \
        LDGA_HIP_TRY(ctx, hipFuncSetAttribute((const void *)kern,                                  \
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)(SMEM))); \
      kern<<<(GRID), kNT, (SMEM), (STREAM)>>>(__VA_ARGS__);                                        \
      LDGA_HIP_TRY(ctx, hipGetLastError());                                                        \
      return LDGA_OK;                                                                              \
    };                                                                                             \
    int rc_;                                                                                       \
    if (mp_ <= 5)                                                                                  \
      rc_ = go_(KERNEL<5>);                                                                        \
    else if (mp_ <= 9)                                                                             \
      rc_ = go_(KERNEL<9>);                                                                        \
    else if (mp_ <= 20)                                                                            \
      rc_ = go_(KERNEL<20>);                                                                       \
    else                                                                                           \
      rc_ = ldga_fail(ctx, LDGA_ERR_SHAPE, "octant of %d points is too large", NO);                \
    if (rc_) return rc_;                                                                           \
  } while (0)

int launch_bubble(ldga_ctx *ctx, const ldga_freq *fr, double beta, const double2 *Gfft, const double2 *Grfft,
                  double2 *chi0, hipStream_t st) {
  const OctTables &T = ctx->oct;
  const int ntot = fr->n_inu + fr->n_inu_shell;
  const int Nnu = 2 * ntot, Nw = 2 * fr->n_iw + 1;
  // every (nu, nu+w) must lie inside fft_range
  int lo = 1 << 30, hi = -(1 << 30);
  for (int wn = -fr->n_iw; wn <= fr->n_iw; ++wn) {
    const int a = -ntot - (fr->shift * wn) / 2, b = ntot - 1 - (fr->shift * wn) / 2;
    lo = std::min(lo, std::min(a, a + wn));
    hi = std::max(hi, std::max(b, b + wn));
  }
  if (lo < fr->fft_first || hi > fr->fft_first + fr->n_fft - 1)
    return ldga_fail(ctx, LDGA_ERR_SHAPE, "fft_range [%d,%d] does not cover nu, nu+w in [%d,%d]", fr->fft_first,
                     fr->fft_first + fr->n_fft - 1, lo, hi);
  const size_t smem = (size_t)T.NO * 16 + (size_t)T.Ns * 8;
  const double scale = -beta / (T.Nk * T.Nk);
  OCT_DISPATCH(T.NO, smem, bubble_kernel, Nnu * Nw, st, T, Nnu, ntot, fr->n_iw, fr->shift, fr->fft_first, Gfft,
               Grfft, chi0, scale);
  return LDGA_OK;
}

int launch_gtables(ldga_ctx *ctx, const ldga_freq *fr, const double2 *Gw, double2 *Gfft, double2 *Grfft,
                   hipStream_t st) {
  const OctTables &T = ctx->oct;
  const size_t smem = (size_t)T.NO * 16 + (size_t)T.Ns * 8;
  OCT_DISPATCH(T.NO, smem, gtables_kernel, fr->n_fft, st, T, Gw, Gfft, Grfft);
  return LDGA_OK;
}

int launch_lambda0(ldga_ctx *ctx, const ldga_freq *fr, double beta, const double2 *F, const double2 *chi0,
                   double2 *lambda0, hipStream_t st) {
  const int Nnu = 2 * fr->n_inu, Nw = 2 * fr->n_iw + 1, Nq = ctx->Nq;
  double2 *fsum = (double2 *)ldga_scratch(ctx, "fsum", (size_t)Nnu * Nw * 16, st);
  if (!fsum) return LDGA_ERR_HIP;
  fsum_kernel<<<(Nnu * Nw + 255) / 256, 256, 0, st>>>(Nnu, Nw, F, fsum);
  const size_t total = (size_t)Nq * Nnu * Nw;
  lambda0_kernel<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(Nq, Nnu, Nw, fr->n_inu_shell, 1.0 / (beta * beta),
                                                                  fsum, chi0, lambda0);
  LDGA_HIP_TRY(ctx, hipGetLastError());
  return LDGA_OK;
}

// F[nu, nu', w] = -(chi[nu, nu', w] + delta_{nu nu'} beta G(nu) G(nu+w)) / (G(nu) G(nu+w) G(nu') G(nu'+w))
// (F_from_χ, BSETools.jl:40-53); nu_n of index a at w: a - n_inu - trunc(shift w / 2) (helpers.jl:46-50).
// One thread per element, nu fastest, as the column-major arrays.
__global__ void fvertex_kernel(int Nnu, int Nw, int n_iw, int n_inu, int shift, int fft_first, double beta, int diag,
                               const double2 *__restrict__ chi, const double2 *__restrict__ G,
                               double2 *__restrict__ F) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)Nnu * Nnu * Nw) return;
  const int a = (int)(i % Nnu);
  const size_t r = i / Nnu;
  const int b = (int)(r % Nnu), wi = (int)(r / Nnu);
  const int wn = wi - n_iw, off = -n_inu - trunc_half(shift * wn) - fft_first;
  const double2 g12 = cmul(G[a + off], G[a + wn + off]);
  const double2 g34 = cmul(G[b + off], G[b + wn + off]);
  double2 num = chi[i];
  if (diag && a == b) num = cadd(num, cscale(g12, beta));
  F[i] = cscale(cdiv(num, cmul(g12, g34)), -1.0);
}

int launch_F_from_chi(ldga_ctx *ctx, const ldga_freq *fr, double beta, int diag_term, const double2 *chi,
                      const double2 *G, double2 *F, hipStream_t st) {
  const int Nnu = 2 * fr->n_inu, Nw = 2 * fr->n_iw + 1;
  // every G index the kernel forms must lie in fft_range
  int lo = 0, hi = 0;
  for (int wi = 0; wi < Nw; ++wi) {
    const int wn = wi - fr->n_iw, first = -fr->n_inu - (fr->shift * wn) / 2, last = first + Nnu - 1;
    const int l = std::min(first, first + wn), h = std::max(last, last + wn);
    if (wi == 0 || l < lo) lo = l;
    if (wi == 0 || h > hi) hi = h;
  }
  if (lo < fr->fft_first || hi > fr->fft_first + fr->n_fft - 1)
    return ldga_fail(ctx, LDGA_ERR_SHAPE, "G over fft_range [%d,%d] does not cover nu in [%d,%d]", fr->fft_first,
                     fr->fft_first + fr->n_fft - 1, lo, hi);
  const size_t total = (size_t)Nnu * Nnu * Nw;
  fvertex_kernel<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(Nnu, Nw, fr->n_iw, fr->n_inu, fr->shift,
                                                                  fr->fft_first, beta, diag_term, chi, G, F);
  LDGA_HIP_TRY(ctx, hipGetLastError());
  return LDGA_OK;
}

int launch_sum_kw(ldga_ctx *ctx, const double *chi, int Nw, int w_lo, int w_hi, double c2, double beta, double lam,
                  double *out_dev, hipStream_t st) {
  if (Nw < 1 || (Nw & 1) == 0 || w_lo < 0 || w_hi >= Nw || w_lo > w_hi)
    return ldga_fail(ctx, LDGA_ERR_SHAPE, "bad omega range [%d,%d] of %d", w_lo, w_hi, Nw);
  const OctTables &T = ctx->oct;
  double *part = (double *)ldga_scratch(ctx, "sumkw_part", 2 * kSumGrid * 8, st);
  unsigned *ticket = ldga_tickets(ctx, "sumkw_ticket", 1, st);
  if (!part || !ticket) return LDGA_ERR_HIP;
  const int total = T.Nq * (w_hi - w_lo + 1);
  const unsigned grid = (unsigned)std::min(kSumGrid, (total + kSumRB - 1) / kSumRB);
  sum_kw_kernel<<<grid, kSumRB, 0, st>>>(chi, T.kmult, T.Nq, w_lo, w_hi, T.Nk, lam, c2, beta, (Nw - 1) / 2, part,
                                         ticket, out_dev);
  LDGA_HIP_TRY(ctx, hipGetLastError());
  return LDGA_OK;
}

// sum_ω per q into res (Nq doubles) and, when out_k != NULL, its k integral
int launch_sum_w(ldga_ctx *ctx, const double *chi, int Nq, int Nw, int w_lo, int w_hi, double c2, double beta,
                 double lam, double *res, double *out_k, hipStream_t st) {
  if (Nq < 1 || Nw < 1 || (Nw & 1) == 0 || w_lo < 0 || w_hi >= Nw || w_lo > w_hi)
    return ldga_fail(ctx, LDGA_ERR_SHAPE, "bad sum_ω shape: Nq=%d, omega range [%d,%d] of %d", Nq, w_lo, w_hi, Nw);
  if (out_k && (!ctx->has_kgrid || ctx->Nq != Nq))
    return ldga_fail(ctx, LDGA_ERR_STATE, "sum_ωk needs the k grid of the Nq=%d points of chi", Nq);
  sum_w_kernel<<<(Nq + 255) / 256, 256, 0, st>>>(chi, Nq, w_lo, w_hi, (Nw - 1) / 2, lam, c2, beta, res);
  LDGA_HIP_TRY(ctx, hipGetLastError());
  if (out_k) {
    kintegrate_kernel<<<1, kSumNT, 0, st>>>(res, ctx->oct.kmult, Nq, ctx->oct.Nk, out_k);
    LDGA_HIP_TRY(ctx, hipGetLastError());
  }
  return LDGA_OK;
}

// omega chunking of the EoM: as many (nu, chunk) workgroups as the device holds
// at once (CUs x resident eom_accum workgroups per CU), so that they run in one
// round; each chunk covers wchunk consecutive omegas.
int eom_chunks(ldga_ctx *ctx, int Nw, int numax, int *wchunk) {
  const OctTables &T = ctx->oct;
  const int mp = oct_maxpt(T.NO);
  const void *kern = mp <= 5 ? (const void *)eom_accum_kernel<5> : mp <= 9 ? (const void *)eom_accum_kernel<9>
                                                                             : (const void *)eom_accum_kernel<20>;
  const size_t smem = (size_t)T.NO * 16 + (size_t)T.Ns * 8 + (size_t)T.Nq * 16;
  if (ctx->eom_resident_key != (size_t)T.NO * 4096 + T.Nq) {
    int ncu = 0, per_cu = 0;
    if (smem > 65536) hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device);
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kNT, smem);
    ctx->eom_resident = std::max(1, ncu) * std::max(1, per_cu);
    ctx->eom_resident_key = (size_t)T.NO * 4096 + T.Nq;
  }
  int nchunks = std::max(1, std::min(Nw, ctx->eom_resident / std::max(numax, 1)));
  *wchunk = (Nw + nchunks - 1) / nchunks;
  return (Nw + *wchunk - 1) / *wchunk;
}

int launch_sigma(ldga_ctx *ctx, const ldga_freq *fr, const ldga_model *mp, int numax, double lam_m, double lam_d,
                 int tc, int eom_kind, double chiloc_m_sum, const double *chi_m, const double2 *gamma_m,
                 const double *chi_d, const double2 *gamma_d, const double2 *lambda0, const double2 *Grfft,
                 double2 *sigma, hipStream_t st) {
  const OctTables &T = ctx->oct;
  const int Nnu = 2 * fr->n_inu, Nw = 2 * fr->n_iw + 1;
  if (numax < 1 || numax > 2 * fr->n_inu) return ldga_fail(ctx, LDGA_ERR_SHAPE, "numax=%d out of range", numax);
  if (eom_kind < LDGA_EOM_FULL || eom_kind > LDGA_EOM_REST) return ldga_fail(ctx, LDGA_ERR_ARG, "bad eom kind");
  // G(nu + w) must lie inside fft_range for every nu in 0..numax-1
  if (-fr->n_iw < fr->fft_first || numax - 1 + fr->n_iw > fr->fft_first + fr->n_fft - 1)
    return ldga_fail(ctx, LDGA_ERR_SHAPE, "fft_range does not cover nu+w");
  // the gamma window must contain nu_n = 0 at every w (helpers.jl:58)
  if (fr->n_inu - fr->shift * (fr->n_iw / 2) < 0)
    return ldga_fail(ctx, LDGA_ERR_SHAPE, "nu window does not contain nu_n = 0 at every w");
  int wchunk = 1;
  const int nchunks = eom_chunks(ctx, Nw, numax, &wchunk);
  double2 *partial = (double2 *)ldga_scratch(ctx, "eom_partial", (size_t)T.NO * numax * nchunks * 16, st);
  double *tsum = (double *)ldga_scratch(ctx, "eom_tail", 3 * 8, st);
  if (!partial || !tsum) return LDGA_ERR_HIP;
  const bool add_tail = tc && eom_kind == LDGA_EOM_FULL;
  if (add_tail) {
    int rc = launch_sum_kw(ctx, chi_m, Nw, 0, Nw - 1, mp->chi_m_c2, mp->beta, lam_m, tsum, st);
    if (rc) return rc;
  }
  const size_t smem = (size_t)T.NO * 16 + (size_t)T.Ns * 8;
  const size_t smem_acc = smem + (size_t)T.Nq * 16;
  OCT_DISPATCH(T.NO, smem_acc, eom_accum_kernel, numax * nchunks, st, T, Nw, Nnu, fr->n_inu, fr->n_iw, fr->shift,
               fr->fft_first, numax, wchunk, eom_kind, mp->U, lam_m, lam_d, chi_m, chi_d, gamma_m, gamma_d, lambda0,
               Grfft, partial);
  const double hartree = (eom_kind == LDGA_EOM_FULL || eom_kind == LDGA_EOM_REST) ? mp->n * mp->U / 2.0 : 0.0;
  const double scale = 1.0 / (mp->beta * T.Nk * T.Nk);
  OCT_DISPATCH(T.NO, smem, eom_final_kernel, numax, st, T, numax, nchunks, partial, scale, mp->U, mp->beta, hartree,
               add_tail ? 1 : 0, tsum, chiloc_m_sum, sigma);
  return LDGA_OK;
}

// Adjoint contraction sum_{nu,w,q} kMult K'_r T / beta for r = m, d into out_dev[0..1] (complex).
int launch_eom_adjoint(ldga_ctx *ctx, const ldga_freq *fr, const ldga_model *mp, int numax, double lam_m,
                       double lam_d, const double *chi_m, const double2 *gamma_m, const double *chi_d,
                       const double2 *gamma_d, const double2 *Grfft, const double2 *W, double2 *out_dev,
                       hipStream_t st) {
  const OctTables &T = ctx->oct;
  const int Nnu = 2 * fr->n_inu, Nw = 2 * fr->n_iw + 1;
  int wchunk = 1;
  const int nchunks = eom_chunks(ctx, Nw, numax, &wchunk);
  double2 *partial = (double2 *)ldga_scratch(ctx, "adj_partial", (size_t)2 * numax * nchunks * 16, st);
  if (!partial) return LDGA_ERR_HIP;
  const size_t smem = (size_t)T.NO * 16 + (size_t)T.Ns * 8;
  const double scale = 1.0 / (T.Nk * T.Nk);
  OCT_DISPATCH(T.NO, smem, eom_adjoint_kernel, numax * nchunks, st, T, Nw, Nnu, fr->n_inu, fr->n_iw, fr->shift,
               fr->fft_first, numax, wchunk, mp->U, lam_m, lam_d, chi_m, chi_d, gamma_m, gamma_d, Grfft, W, scale,
               partial);
  sum_pairs_kernel<<<1, 256, 0, st>>>(numax * nchunks, partial, out_dev);
  LDGA_HIP_TRY(ctx, hipGetLastError());
  return LDGA_OK;
}
