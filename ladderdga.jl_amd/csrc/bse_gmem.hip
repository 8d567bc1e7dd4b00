// Bethe-Salpeter solve in FP64 with the matrix in global memory (calc_χγ,
// src/BSETools.jl:192-248; replaces the per-(q,w) zgetrf+zgetri of
// src/LapackWrapper.jl:64-72) for the sizes the register-resident kernels do
// not hold (Nnu > 128): the FP64 re-solve of the matrices whose mixed
// refinement failed, and FP64 requests at those sizes.
//
// One workgroup per matrix runs the same Gauss-Jordan elimination with column
// pivoting and implicit pivots as the blocked kernel, one row at a time, on
// M = A^T copied to a workspace slot (row-major, L2-resident for a few
// matrices at a time):
//   p = argmax_{c unpivoted} |Re M[r][c]| + |Im M[r][c]|  (first maximum),
//   M[i][c] -= M[i][p] M[r][c] / piv,  M[i][p] /= piv       (i != r, c != p),
//   M[r][c] = -M[r][c] / piv,          M[r][p] = 1 / piv;
// then y[p_r] = sum_c M[r][c] solves A^T y = 1.  Every step reads and writes
// the whole matrix, so this kernel is meant for few matrices; its cost per
// matrix is about N steps of 2 N^2 * 16 bytes of L2 traffic.
#include <cfloat>

#include "ldga_internal.h"

namespace {

// NTH threads per matrix: 256 when the launch covers every (q, w) (several
// workgroups share a CU), 1024 for a short re-solve list (a few workgroups on
// an otherwise idle device).  The step's update walks the matrix by flat index
// with GM_UNROLL loads in flight per thread.
constexpr int GM_UNROLL = 4;

template <int NTH>
__global__ __launch_bounds__(NTH) void bse_gmem_f64_kernel(int N, int Nq, int Nc, int shell,
                                                                  const double2 *__restrict__ Gam,
                                                                  const double2 *__restrict__ chi0, double s,
                                                                  double U, double beta, double *__restrict__ chi_out,
                                                                  double2 *__restrict__ gamma_out,
                                                                  unsigned long long *__restrict__ counters,
                                                                  const int *__restrict__ only, int n_items,
                                                                  const unsigned long long *__restrict__ n_dev,
                                                                  double2 *__restrict__ ws,
                                                                  double2 *__restrict__ inv_out) {
  extern __shared__ double2 lds[];
  double2 *rs = lds;         // pivot row / piv
  double2 *fcol = rs + N;    // pivot column before the step
  double2 *y = fcol + N;     // solution
  int *pivoted = reinterpret_cast<int *>(y + N);
  int *pcol = pivoted + N;   // pivot column of each row
  __shared__ double red_v[NTH / 64];
  __shared__ int red_c[NTH / 64];
  __shared__ int piv_col;
  __shared__ double red_chi[NTH / 64];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int NN = N * N;
  double2 *M = ws + (size_t)blockIdx.x * N * N;
  // this thread's first (row, column) of the flat row-major walk, and the step
  const int i0 = tid / N, c0 = tid - i0 * N, di = NTH / N, dc = NTH - di * N;
  // a list whose length is known only on the device: at most n_items of it
  if (n_dev) n_items = min(n_items, __builtin_amdgcn_readfirstlane((int)*n_dev));

  for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int lb = only ? only[item] : item;
    const int q = lb % Nq, w = lb / Nq;
    const double2 *G = Gam + (size_t)w * N * N;
    // M[i][c] = Gamma[c, i, w] + delta_ic / chi0[q, shell + i, w]
    for (int i = i0, c = c0; i < N;) {
      double2 v = G[c + (size_t)N * i];
      if (c == i) v = cadd(v, crecip(chi0[q + (size_t)Nq * (shell + i + (size_t)Nc * w)]));
      M[(size_t)N * i + c] = v;
      c += dc;
      i += di;
      if (c >= N) {
        c -= N;
        ++i;
      }
    }
    for (int c = tid; c < N; c += NTH) pivoted[c] = 0;
    int zero = 0;
    __syncthreads();

    for (int r = 0; r < N; ++r) {
      const double2 *Mr = M + (size_t)N * r;
      // argmax over the unpivoted columns: value descending, column ascending;
      // a NaN counts as the largest value (its column is taken, the NaN propagates)
      double bv = -1.0;
      int bc = N;
      for (int c = tid; c < N; c += NTH) {
        if (pivoted[c]) continue;
        const double2 m = Mr[c];
        double v = fabs(m.x) + fabs(m.y);
        if (!(v <= DBL_MAX)) v = INFINITY;
        if (v > bv) {
          bv = v;
          bc = c;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(bv, off);
        const int oc = __shfl_xor(bc, off);
        if (ov > bv || (ov == bv && oc < bc)) {
          bv = ov;
          bc = oc;
        }
      }
      if (lane == 0) {
        red_v[wave] = bv;
        red_c[wave] = bc;
      }
      __syncthreads();
      if (tid == 0) {
        double v = red_v[0];
        int c = red_c[0];
        for (int k = 1; k < NTH / 64; ++k)
          if (red_v[k] > v || (red_v[k] == v && red_c[k] < c)) {
            v = red_v[k];
            c = red_c[k];
          }
        piv_col = c;
        pivoted[c] = 1;
        pcol[r] = c;
      }
      __syncthreads();
      const int p = piv_col;
      const double2 piv = Mr[p];
      zero |= piv.x == 0.0 && piv.y == 0.0;
      const double2 ipiv = crecip(piv);
      for (int c = tid; c < N; c += NTH) rs[c] = cmul(Mr[c], ipiv);
      for (int i = tid; i < N; i += NTH) fcol[i] = M[(size_t)N * i + p];
      __syncthreads();
      for (int k0 = tid; k0 < NN; k0 += GM_UNROLL * NTH) {
        double2 mv[GM_UNROLL];
#pragma unroll
        for (int u = 0; u < GM_UNROLL; ++u) {
          const int k = k0 + u * NTH;
          if (k < NN) mv[u] = M[k];
        }
#pragma unroll
        for (int u = 0; u < GM_UNROLL; ++u) {
          const int k = k0 + u * NTH;
          if (k < NN) {
            const int i = k / N, c = k - i * N;
            double2 m;
            if (i == r)
              m = c == p ? ipiv : cz(-rs[c].x, -rs[c].y);
            else if (c == p)
              m = cmul(fcol[i], ipiv);
            else
              m = csub(mv[u], cmul(fcol[i], rs[c]));
            M[k] = m;
          }
        }
      }
      __syncthreads();
    }

    // inv_out: the full inverse, A^-1[i][p_r] = M[r][p_i] (the matrix is M = A^T), stored
    // column-major per (q, w) as calc_gen_χ returns it (BSETools.jl:259-275)
    if (inv_out) {
      double2 *X = inv_out + (size_t)(q + (size_t)Nq * w) * N * N;
      for (int k = tid; k < N * N; k += NTH) {
        const int i = k % N, r = k / N;
        X[i + (size_t)N * pcol[r]] = M[(size_t)N * r + pcol[i]];
      }
    }
    // y[p_r] = sum_c M[r][c]; chi = Re sum y / beta^2; gamma = y / (chi0 (1 + s U chi))
    for (int r = tid; r < N; r += NTH) {
      double2 acc = cz(0.0, 0.0);
      for (int c = 0; c < N; ++c) acc = cadd(acc, M[(size_t)N * r + c]);
      y[pcol[r]] = acc;
    }
    __syncthreads();
    double part = 0.0;
    for (int k = tid; k < N; k += NTH) part += y[k].x;
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
    if (lane == 0) red_chi[wave] = part;
    __syncthreads();
    double tot = 0.0;
    for (int k = 0; k < NTH / 64; ++k) tot += red_chi[k];
    const double chi = tot / (beta * beta);
    const double fac = 1.0 + s * U * chi;
    for (int k = tid; gamma_out && k < N; k += NTH) {
      const double2 c0v = chi0[q + (size_t)Nq * (k + (size_t)Nc * w)];  // no shell offset (BSETools.jl:236)
      gamma_out[q + (size_t)Nq * (k + (size_t)N * w)] = cdiv(y[k], cscale(c0v, fac));
    }
    if (tid == 0) {
      if (chi_out) chi_out[q + (size_t)Nq * w] = chi;
      if (zero) atomicAdd(&counters[0], 1ull);
    }
    __syncthreads();  // LDS and the workspace slot are reused by the next matrix
  }
}

}  // namespace

// FP64 Gauss-Jordan in global memory: every (q, w), or the matrices (q + Nq w)
// listed in only: n_only of them, or, with n_dev != NULL, the count *n_dev read
// on the device (at most n_only).  Any Nnu up to 1024.  chi and gamma may be
// NULL; inv (NULL or Nnu x Nnu x Nq x Nw) receives the full inverses.
int launch_bse_gmem_f64(ldga_ctx *ctx, int N, int Nq, int Nw, int Nc, int shell, const double2 *Gam,
                        const double2 *chi0, double s, double U, double beta, double *chi, double2 *gamma,
                        unsigned long long *cnt, const int *only, int n_only, const unsigned long long *n_dev,
                        hipStream_t st, double2 *inv) {
  if (N < 1 || N > 1024) return ldga_fail(ctx, LDGA_ERR_SHAPE, "FP64 global-memory BSE supports Nnu <= 1024 (got %d)", N);
  const int n_items = only ? n_only : Nq * Nw;
  if (n_items == 0) return LDGA_OK;
  // workspace: one N x N slot per workgroup, at most 1 GiB
  const size_t slot = (size_t)N * N * sizeof(double2);
  int grid = n_items < 1024 ? n_items : 1024;
  while (grid > 1 && (size_t)grid * slot > (size_t(1) << 30)) grid /= 2;
  double2 *ws = (double2 *)ldga_scratch(ctx, "bse_gmem_ws", (size_t)grid * slot, st);
  if (!ws) return LDGA_ERR_HIP;
  const size_t lds = (size_t)N * (3 * sizeof(double2) + 2 * sizeof(int));
  if (only)  // a re-solve list: few matrices, each on a whole CU
    bse_gmem_f64_kernel<1024><<<dim3((unsigned)grid), dim3(1024), lds, st>>>(N, Nq, Nc, shell, Gam, chi0, s, U,
                                                                             beta, chi, gamma, cnt, only, n_items,
                                                                             n_dev, ws, inv);
  else
    bse_gmem_f64_kernel<256><<<dim3((unsigned)grid), dim3(256), lds, st>>>(N, Nq, Nc, shell, Gam, chi0, s, U,
                                                                           beta, chi, gamma, cnt, only, n_items,
                                                                           n_dev, ws, inv);
  LDGA_HIP_TRY(ctx, hipGetLastError());
  return LDGA_OK;
}
