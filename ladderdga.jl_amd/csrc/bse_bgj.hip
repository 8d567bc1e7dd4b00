// Batched Bethe-Salpeter solve, mixed precision: a 16-row blocked
// Gauss-Jordan inversion on MFMA with FP64 iterative refinement
// (calc_χγ, src/BSETools.jl:192-248; replaces the per-(q,w) zgetrf+zgetri of
// src/LapackWrapper.jl:64-72).
//
// For every (q,w) one workgroup of NT waves inverts M = A^T,
// A = Gamma(w) + diag(1/chi0(q,.,w)), padded to NC = 16 NT with identity rows
// and columns.  Wave w holds rows 16w .. 16w+15 in MFMA accumulators as NT
// tiles of 16x16 (real and imaginary planes; f32 16x16x4 layout: lane
// (g = lane/16, n = lane%16) holds, in element j, row 4g+j of the tile at
// column n).  Block K = the rows and columns 16K .. 16K+15.
//
// Sweep of block K (the in-place block Gauss-Jordan step, P = inv(M[K,K])):
//   M[K,K] <- P,  M[K,J] <- -P M[K,J],  M[I,K] <- M[I,K] P,
//   M[I,J] <- M[I,J] - M[I,K] P M[K,J]        (I, J != K),
// written as  W = (new rows of K) - E_K,  M[I,:] += M[I,K] W  for I != K.
// After the sweeps of all blocks M holds inv(M) in natural order, and the
// solution of A^T y = 1 is y = inv(M) 1 (row sums).
//
// Pivoting is local to the diagonal block: D = M[K,K] is inverted by scalar
// Gauss-Jordan steps with column pivoting among D's own unpivoted columns
// (LAPACK's izamax criterion |re|+|im|, first maximum).  A D that is exactly
// singular (zero pivot), or a refinement that does not converge, sends the
// matrix to the FP64 re-solve with partial pivoting over the whole row; the
// refinement below restores FP64 accuracy for every other matrix.
//
// Step k (owner = wave k, the owner of the block's rows):
//   barrier 1: P = inv(D_k) and the owner's rows M[K,:] are in LDS;
//   phase 1: wave w forms tile w of W: -P M[K, tile w] (four f32 K=4 MFMA
//     chunks, complex 4M scheme), or P - I for w = K, and publishes it split
//     for the bf16 MFMA B operand;
//   barrier 2;
//   phase 2: the owner takes W + E_K as its rows; every other wave applies
//     M[I,:] += M[I,K] W as four rank-4 complex bf16 MFMA updates per tile
//     (A operand: its tile K, transposed through LDS); the owner of step k+1
//     (raised wave priority) then publishes its rows and inverts its diagonal
//     tile, so that its chain (rank-16 update, sixteen pivot steps) is the
//     critical path of the step.
//
// FP32 products on bf16 MFMA: an FP32 value is split into three bf16 parts
// (bse_common.h, split3); seven of the nine part products are taken, so each
// product is below one FP32 multiplication's error.
//
// Refinement: r = 1 - A^T y with Gamma and chi0 in FP64 (Gamma is L2-resident:
// all q of one w share it), y += X r, until |X r|_inf <= 2^-50 |y|_inf, or the
// next correction extrapolated from the contraction rate is <= 2^-42 |y|_inf,
// or the correction stops halving (the FP64 residual's rounding level).
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "ldga_internal.h"

#include "bse_common.h"

namespace {

using namespace ldga_bse;

template <int NT>
struct BgjShared {
  static constexpr int NC = 16 * NT;
  // the step owner's rows as of the previous step, tile column w written by
  // wave w; in phase 2 wave w also uses its tile column as transposition scratch
  float2 rows[16][NC + 1];
  float2 R[16][NC + 1];     // the next owner's rows before the step's update
  u32x4 Wb[4][NT][64];      // W split for the bf16 B operand: [chunk][tile][lane (k row = lane / 16)]
  float2 P[16][17];         // inv(D) of the step's diagonal block, natural order
  double2 y[NC];            // solution (FP64)
  double2 dinv[NC];         // 1/chi0 on the diagonal (FP64)
  float2 res[NC];           // refinement residual (rounded to FP32 for X r)
  double red[NT];
  float redf[2][2][NT];
  int zero_pivot;
  int r_ready;  // the step whose R is published (set by its owner, read by the other waves)
};

// maximum over the 16 lanes of each DPP row, in every lane of the row
__device__ __forceinline__ unsigned row_max16_u32(unsigned k) {
  __asm__("s_nop 1\n\tv_max_u32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(k));
  __asm__("s_nop 1\n\tv_max_u32_dpp %0, %0, %0 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf" : "+v"(k));
  __asm__("s_nop 1\n\tv_max_u32_dpp %0, %0, %0 row_half_mirror row_mask:0xf bank_mask:0xf" : "+v"(k));
  __asm__("s_nop 1\n\tv_max_u32_dpp %0, %0, %0 row_mirror row_mask:0xf bank_mask:0xf" : "+v"(k));
  return k;
}

// the value of v in lane src (any lane of the wave)
__device__ __forceinline__ float bperm(int src, float v) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}

// f(std::integral_constant<int, I>) for I = B .. E-1 (compile-time indices)
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// the value of v in lane L of each 16-lane row, in every lane of the row (DPP row_newbcast)
template <int L>
__device__ __forceinline__ float row_bcast(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x150 + L, 0xf, 0xf, false));
}

// per lane: b where bit lane of the (wave-uniform) mask is set, else a.  A
// select with a constant lane mask, kept a select (the compiler otherwise
// turns a lane-dependent select feeding arithmetic into a divergent branch).
__device__ __forceinline__ float vsel(float a, float b, unsigned long long mask) {
  float r;
  __asm__("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(mask));
  return r;
}

// an FP32 value from its packed split (h + m + l, exact)
__device__ __forceinline__ float unsplit_lo(u32x4 u) { return (bf_lo(u[0]) + bf_lo(u[1])) + bf_lo(u[2]); }
__device__ __forceinline__ float unsplit_hi(u32x4 u) { return (bf_hi(u[0]) + bf_hi(u[1])) + bf_hi(u[2]); }

// four waves per SIMD (<= 128 VGPRs) up to eight tiles: two workgroups share a
// CU and overlap their serial chains
template <int NT>
constexpr int bgj_min_waves() { return NT <= 8 ? 4 : 1; }

// tile tq (wave-uniform runtime index) of an accumulator plane, picked by
// scalar branches (no runtime-indexed access to the register array)
template <typename P>
__device__ __forceinline__ mfma_traits<float>::v4 tile_of(const P &pl, int tq) {
  tq = __builtin_amdgcn_readfirstlane(tq);
  mfma_traits<float>::v4 r = pl.tile(0);
#pragma unroll
  for (int t = 1; t < P::ntiles; ++t)
    if (tq == t) {
      r = pl.tile(t);
      __asm__ volatile("" : "+v"(r));
    }
  return r;
}

// One workgroup solves matrix lb = q + Nq w.
template <int NT>
__device__ __forceinline__ void bgj_solve(BgjShared<NT> &sh, const int lb, const int N, const int Nq, const int Nc,
                                          const int shell, const double2 *__restrict__ Gam,
                                          const double2 *__restrict__ chi0, const double s, const double U,
                                          const double beta, const int max_ir, double *__restrict__ chi_out,
                                          double2 *__restrict__ gamma_out, unsigned long long *__restrict__ counters,
                                          int *__restrict__ redo) {
  typedef mfma_traits<float> MT;
  typedef MT::v4 v4;
  constexpr int NC = 16 * NT;

  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6), g = lane >> 4,
            n = lane & 15;
  const int q = lb % Nq, w = lb / Nq;
  const double2 *G = Gam + (size_t)w * N * N;
  const int NB = (N + 15) >> 4;  // blocks holding matrix rows; the rest are identity blocks

  for (int i = tid; i < NC; i += 64 * NT) {
    double2 d = cz(0.0, 0.0);
    if (i < N) d = crecip(chi0[q + (size_t)Nq * (shell + i + (size_t)Nc * w)]);
    sh.dinv[i] = d;
  }
  if (tid == 0) {
    sh.zero_pivot = 0;
    sh.r_ready = -1;
  }
  __syncthreads();

  // ---- M = A^T into the accumulators: M[i][c] = Gamma[c, i, w] + delta_ic / chi0[i]
  // (identity on the padding).  Row i = 16 wave + 4g + j, column c = 16t + n.
  // Rows and columns past N - 1 are clamped for the load and zeroed after it.
  // Block 0's rows are needed in registers only for its diagonal tile (the
  // owner inverts it before step 0, and step 0 replaces all of the owner's
  // rows); the other tiles of those rows go to sh.rows, the operand of step
  // 0, each loaded by the wave of its tile column.  So with NT > 1 wave 0
  // loads one tile and starts inverting while the other waves load.
  const bool lead = NT > 1 && wave == 0;
  AccPlane<float, NT> Vr, Vi;
  {
    constexpr int LG = 2;
    const char *Gb = reinterpret_cast<const char *>(G);
    const int rb = 16 * wave + 4 * g;
    unsigned rofs[4];
    float dgx[4], dgy[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = rb + j;
      rofs[j] = 16u * (unsigned)N * (unsigned)min(i, N - 1);
      const double2 d = i < N ? sh.dinv[i] : cz(1.0, 0.0);
      const bool here = n == 4 * g + j;
      dgx[j] = here ? (float)d.x : 0.0f;
      dgy[j] = here ? (float)d.y : 0.0f;
    }
#pragma unroll
    for (int t0 = 0; t0 < NT; t0 += LG) {
      int nl = n;
      __asm__ volatile("" : "+v"(nl));
      double2 v[LG][4];
#pragma unroll
      for (int tt = 0; tt < LG; ++tt) {
        if (t0 + tt < NT && !(lead && t0 + tt > 0)) {
          const unsigned cofs = 16u * (unsigned)min(16 * (t0 + tt) + nl, N - 1);
#pragma unroll
          for (int j = 0; j < 4; ++j) v[tt][j] = *reinterpret_cast<const double2 *>(Gb + (cofs + rofs[j]));
        }
      }
#pragma unroll
      for (int tt = 0; tt < LG; ++tt) {
        const int t = t0 + tt;
        if (t < NT) {
          const bool cok = 16 * t + n < N && !(lead && t > 0);
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const bool ok = cok && rb + j < N;
            float x = ok ? (float)v[tt][j].x : 0.0f, y = ok ? (float)v[tt][j].y : 0.0f;
            if (t == wave) {
              x += dgx[j];
              y += dgy[j];
            }
            Vr.set(t, j, x);
            Vi.set(t, j, y);
          }
          __asm__ volatile("" : "+v"(Vr.regs(t)), "+v"(Vi.regs(t)));
        }
      }
      __asm__ volatile("" ::: "memory");
    }
    if (NT > 1 && wave > 0) {
      // tile column `wave` of block 0's rows (no diagonal there) into sh.rows
      const int c = 16 * wave + n;
      const unsigned cofs = 16u * (unsigned)min(c, N - 1);
      double2 b[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = *reinterpret_cast<const double2 *>(Gb + (cofs + 16u * (unsigned)N * (unsigned)min(4 * g + j, N - 1)));
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bool ok = c < N && 4 * g + j < N;
        sh.rows[4 * g + j][c] = make_float2(ok ? (float)b[j].x : 0.0f, ok ? (float)b[j].y : 0.0f);
      }
    }
  }

  int zero_seen = 0;
  int Lg = g, Ln = n, Ll = lane;

  // this wave's rows -> sh.rows (the operand of the next step's phase 1)
  // this wave's tile kt, transposed into its own tile column of sh.rows
  // (A operand: row n, column 4c + g read back as sh.rows[n][16 wave + 4c + g])
  auto transpose_tile = [&](int kt) {
    const v4 ar = Vr.select(kt), ai = Vi.select(kt);
#pragma unroll
    for (int j = 0; j < 4; ++j) sh.rows[4 * Lg + j][16 * wave + Ln] = make_float2(ar[j], ai[j]);
    wave_sync_lds();
  };
  auto publish_rows = [&]() {
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) sh.rows[4 * Lg + j][16 * t + Ln] = make_float2(Vr.get(t, j), Vi.get(t, j));
  };

  // Invert D = (tr, ti), a tile of this wave, in place by Gauss-Jordan steps on its
  // diagonal (step st pivots on D[st][st], iv = 1/D[st][st]):
  //   row st   <- (D[st][:] with D[st][st] := 1) * iv
  //   row i    <- (D[i][:] with D[i][st] := 0) - D[i][st] * (new row st)    (i != st)
  // and write inv(D) to sh.P.  Every lane exchange has a fixed source: the
  // pivot by readlane, the pivot row from lane group st/4 (a lane permute
  // issued at the start of the step), the pivot column from lane st of each
  // row (DPP).  In the last block the rows
  // past N - 1 are identity rows with zeros in the matrix columns: their
  // steps have piv = 1 and change nothing (exact products by 0 and 1).
  auto invert_diag = [&](const v4 tr, const v4 ti) {
    float dr[4], di[4], dmin = 1.0f;  // dmin: the smallest |piv|^2 (a zero pivot is flagged once, after the steps)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      dr[j] = tr[j];
      di[j] = ti[j];
    }
    static_for<0, 16>([&](auto ST) {
      constexpr int st = decltype(ST)::value, gs = st >> 2, js = st & 3;
      // lanes holding column st (one per row) and lane group gs (row st)
      constexpr unsigned long long atp = 0x0001000100010001ull << st, own = 0xffffull << (16 * gs);
      const float rx = bperm(16 * gs + Ln, dr[js]), ry = bperm(16 * gs + Ln, di[js]);  // pivot row, column n
      const float px = rdlane(dr[js], 16 * gs + st), py = rdlane(di[js], 16 * gs + st);
      float cx[4], cy[4];  // pivot column, rows 4g + j
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        cx[j] = row_bcast<st>(dr[j]);
        cy[j] = row_bcast<st>(di[j]);
      }
      // 1 / piv = conj(piv) / |piv|^2, reciprocal refined by a Newton step
      const float dd = px * px + py * py;
      dmin = fminf(dmin, dd);
      float rd = fast_rcp(dd);
      rd = fmaf(rd, fmaf(-dd, rd, 1.0f), rd);
      const float ivx = px * rd, ivy = -py * rd;
      const float ax = vsel(rx, 1.0f, atp), ay = vsel(ry, 0.0f, atp);
      const float nrx = ax * ivx - ay * ivy, nry = ax * ivy + ay * ivx;  // new row st at column n
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        float x = vsel(dr[j], 0.0f, atp), y = vsel(di[j], 0.0f, atp);
        x = fmaf(-cx[j], nrx, fmaf(cy[j], nry, x));
        y = fmaf(-cx[j], nry, fmaf(-cy[j], nrx, y));
        if (j == js) {
          x = vsel(x, nrx, own);
          y = vsel(y, nry, own);
        }
        dr[j] = x;
        di[j] = y;
      }
    });
    zero_seen |= dmin == 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) sh.P[4 * Lg + j][Ln] = make_float2(dr[j], di[j]);
  };
  // k = -1: only the owner of block 0 inverts its block (and, with NT = 1,
  // publishes its rows)
  for (int k = -1; k < NB; ++k) {
    __asm__ volatile("" : "+v"(Lg), "+v"(Ln), "+v"(Ll));
    const bool next_owner = wave == k + 1 && k + 1 < NB;
    if (k >= 0) {
      __syncthreads();  // P and the owner's rows of block k published
      // ---- phase 1: tile `wave` of W
      {
        v4 wr = (v4)0.0f, wi = (v4)0.0f;
        if (wave == k) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const float2 pv = sh.P[4 * Lg + j][Ln];
            wr[j] = pv.x - ((4 * g + j) == n ? 1.0f : 0.0f);
            wi[j] = pv.y;
          }
        } else {
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const float2 a = sh.P[Ln][4 * c + Lg];
            const float2 b = sh.rows[4 * c + Lg][16 * wave + Ln];
            wr = MT::mma(-a.x, b.x, wr);
            wr = MT::mma(a.y, b.y, wr);
            wi = MT::mma(-a.x, b.y, wi);
            wi = MT::mma(-a.y, b.x, wi);
          }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) sh.Wb[Lg][wave][16 * j + Ln] = bf_pack_b(wr[j], wi[j]);
      }
      __syncthreads();  // W published
      // ---- phase 2
      if (!next_owner) {
        if (wave == k) {
          // the owner's rows become W + E_K
#pragma unroll
          for (int t = 0; t < NT; ++t) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const u32x4 u = sh.Wb[Lg][t][16 * j + Ln];
              float x = unsplit_lo(u);
              if (t == k && 4 * g + j == n) x += 1.0f;
              Vr.set(t, j, x);
              Vi.set(t, j, unsplit_hi(u));
            }
          }
        } else {
          // M[I,:] += M[I,K] W; A operand: this wave's tile k (before the update), transposed
          transpose_tile(k);
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const float2 a = sh.rows[Ln][16 * wave + 4 * c + Lg];
            const BfA o = bf_prep_a(a.x, a.y);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
              v4 cr = Vr.tile(t), ci = Vi.tile(t);
              bf_cmma(o, sh.Wb[c][t][Ll], cr, ci);
              Vr.set_tile(t, cr);
              Vi.set_tile(t, ci);
            }
          }
        }
        if (k + 1 < NB) {
          // tile column `wave` of the next owner's rows, up to date:
          //   rows[:, wave] = R[:, wave] + R[:, k] W[:, wave]
          // R is published at the start of the next owner's phase 2; the wait
          // is bounded, and a matrix whose wait ran out is re-solved in FP64
          int spins = 0;
          while (*(volatile int *)&sh.r_ready != k + 1 && ++spins < (1 << 22)) __builtin_amdgcn_s_sleep(1);
          if (spins >= (1 << 22)) zero_seen = 1;
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
          v4 cr, ci;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const float2 v = sh.R[4 * Lg + j][16 * wave + Ln];
            cr[j] = v.x;
            ci[j] = v.y;
          }
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const float2 a = sh.R[Ln][16 * k + 4 * c + Lg];
            bf_cmma(bf_prep_a(a.x, a.y), sh.Wb[c][wave][Ll], cr, ci);
          }
#pragma unroll
          for (int j = 0; j < 4; ++j) sh.rows[4 * Lg + j][16 * wave + Ln] = make_float2(cr[j], ci[j]);
        }
      }
    }
    if (next_owner) {
      // the owner of block k+1 (k = -1: block 0, its rows as loaded): publish
      // R (its rows before this step's update; the other waves bring them up
      // to date, one tile column each), bring only its diagonal tile up to
      // date, and invert it
      __builtin_amdgcn_s_setprio(3);
      v4 dtr = Vr.select(k + 1), dti = Vi.select(k + 1);
      if (k >= 0) {
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
          for (int j = 0; j < 4; ++j) sh.R[4 * Lg + j][16 * t + Ln] = make_float2(Vr.get(t, j), Vi.get(t, j));
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) *(volatile int *)&sh.r_ready = k + 1;
        transpose_tile(k);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const float2 a = sh.rows[Ln][16 * wave + 4 * c + Lg];
          bf_cmma(bf_prep_a(a.x, a.y), sh.Wb[c][k + 1][Ll], dtr, dti);
        }
      } else if (NT == 1) {
        publish_rows();  // (with NT > 1 the other waves loaded block 0's rows into sh.rows)
      }
      invert_diag(dtr, dti);
      __builtin_amdgcn_s_setprio(0);
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (zero_seen && lane == 0) sh.zero_pivot = 1;

  // ---- y = X 1 (row sums) ------------------------------------------------------
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    float sx = 0.0f, sy = 0.0f;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      sx += Vr.get(t, j);
      sy += Vi.get(t, j);
    }
    sx = row_sum16(sx);
    sy = row_sum16(sy);
    if (n == 0) sh.y[16 * wave + 4 * g + j] = cz((double)sx, (double)sy);
  }
  __syncthreads();

  // ---- FP64 iterative refinement with the FP32 inverse ----------------------------
  int it = 0;
  double last_ratio = 0.0;
  bool converged_stop = false;
  for (; it < max_ir; ++it) {
    // r_i = 1 - sum_j Gamma[j, i] y_j - y_i / chi0_i; wave w: rows 4k + g for k = w, w + NT, ...
    for (int ib = 4 * wave; ib < N; ib += 4 * NT) {
      const int i = ib + g;
      double2 acc = cz(0.0, 0.0);
      if (i < N) {
        const double2 *col = G + (size_t)N * i;
        double2 gv[NT];
        if (N > 16 * (NT - 1)) {
#pragma unroll
          for (int k = 0; k < NT - 1; ++k) gv[k] = col[n + 16 * k];
          gv[NT - 1] = (n + 16 * (NT - 1) < N) ? col[n + 16 * (NT - 1)] : cz(0.0, 0.0);
        } else {
#pragma unroll
          for (int k = 0; k < NT; ++k) gv[k] = (n + 16 * k < N) ? col[n + 16 * k] : cz(0.0, 0.0);
        }
#pragma unroll
        for (int k = 0; k < NT; ++k) {
          const double2 yk = sh.y[n + 16 * k];
          acc.x = fma(gv[k].x, yk.x, fma(-gv[k].y, yk.y, acc.x));
          acc.y = fma(gv[k].x, yk.y, fma(gv[k].y, yk.x, acc.y));
        }
      }
      acc.x = row_sum16(acc.x);
      acc.y = row_sum16(acc.y);
      if (n == 0 && i < N) {
        const double2 yi = sh.y[i];
        const double2 d = cmul(yi, sh.dinv[i]);
        sh.res[i] = make_float2((float)(1.0 - acc.x - d.x), (float)(-acc.y - d.y));
      }
    }
    if (tid < NC - N) sh.res[N + tid] = make_float2(0.0f, 0.0f);
    __syncthreads();
    // dy = X r (FP32 products)
    float dmax = 0.0f, ymax = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float sx = 0.0f, sy = 0.0f;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const float2 rv = sh.res[16 * t + n];
        sx = fmaf(Vr.get(t, j), rv.x, fmaf(-Vi.get(t, j), rv.y, sx));
        sy = fmaf(Vr.get(t, j), rv.y, fmaf(Vi.get(t, j), rv.x, sy));
      }
      sx = row_sum16(sx);
      sy = row_sum16(sy);
      const int row = 16 * wave + 4 * g + j;
      if (n == 0 && row < N) {
        const double2 yo = sh.y[row];
        const double2 yn = cz(yo.x + (double)sx, yo.y + (double)sy);
        sh.y[row] = yn;
        dmax = fmaxf(dmax, fabsf(sx) + fabsf(sy));
        ymax = fmaxf(ymax, (float)(fabs(yn.x) + fabs(yn.y)));
      }
    }
    const unsigned dk = wave_max_u32(__float_as_uint(dmax)), yk = wave_max_u32(__float_as_uint(ymax));
    if (lane == 63) {
      sh.redf[it & 1][0][wave] = __uint_as_float(dk);
      sh.redf[it & 1][1][wave] = __uint_as_float(yk);
    }
    __syncthreads();
    float dm = 0.0f, ym = 0.0f;
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      dm = fmaxf(dm, sh.redf[it & 1][0][k]);
      ym = fmaxf(ym, sh.redf[it & 1][1][k]);
    }
    const double ratio = (double)dm / (double)ym;
    const bool converged = ratio <= 0x1p-50 || (it > 0 && ratio * ratio <= 0x1p-42 * last_ratio);
    const bool stagnant = it > 0 && !(ratio < 0.5 * last_ratio);
    const bool stop = converged || stagnant;
    last_ratio = ratio;
    if (stop) {
      converged_stop = converged;
      ++it;
      break;
    }
  }
  if (tid == 0) {
    atomicMax(&counters[1], (unsigned long long)it);
    atomicAdd(&counters[5], (unsigned long long)it);
    if ((!converged_stop && !(last_ratio <= 0x1p-30)) || sh.zero_pivot) {
      const unsigned long long kk = atomicAdd(&counters[2], 1ull);
      if (redo) redo[kk] = lb;
    }
    unsigned long long bits;
    const double lr = last_ratio == last_ratio ? last_ratio : 1.0;
    memcpy(&bits, &lr, 8);
    atomicMax(&counters[3], bits);
  }

  // ---- chi and gamma -------------------------------------------------------------
  double part = 0.0;
  for (int k = tid; k < N; k += 64 * NT) part += sh.y[k].x;
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
  if (lane == 0) sh.red[wave] = part;
  __syncthreads();
  double tot = 0.0;
#pragma unroll
  for (int k = 0; k < NT; ++k) tot += sh.red[k];
  const double chi = tot / (beta * beta);
  const double fac = 1.0 + s * U * chi;
  for (int k = tid; k < N; k += 64 * NT) {
    const double2 c0 = chi0[q + (size_t)Nq * (k + (size_t)Nc * w)];  // no shell offset (BSETools.jl:236)
    gamma_out[q + (size_t)Nq * (k + (size_t)N * w)] = cdiv(sh.y[k], cscale(c0, fac));
  }
  if (tid == 0) {
    chi_out[q + (size_t)Nq * w] = chi;
    if (sh.zero_pivot && !redo) atomicAdd(&counters[0], 1ull);
  }
}

template <int NT>
__global__ __launch_bounds__(64 * NT, (bgj_min_waves<NT>())) void bse_bgj_kernel(
    int N, int Nq, int Nc, int shell, const double2 *__restrict__ Gam, const double2 *__restrict__ chi0, double s,
    double U, double beta, int max_ir, double *__restrict__ chi_out, double2 *__restrict__ gamma_out,
    unsigned long long *__restrict__ counters, int *__restrict__ redo) {
  __shared__ BgjShared<NT> sh;
  bgj_solve<NT>(sh, xcd_logical_block(blockIdx.x, gridDim.x), N, Nq, Nc, shell, Gam, chi0, s, U, beta, max_ir,
                chi_out, gamma_out, counters, redo);
}

template <int NT>
int run_bgj(ldga_ctx *ctx, int N, int Nq, int Nw, int Nc, int shell, const double2 *Gam, const double2 *chi0,
            double s, double U, double beta, int max_ir, double *chi, double2 *gamma, unsigned long long *cnt,
            int *redo, hipStream_t st) {
  if (N > 16 * NT) return ldga_fail(ctx, LDGA_ERR_SHAPE, "BSE tile too small for N=%d", N);
  const unsigned nblocks = (unsigned)Nq * Nw;
  if (nblocks == 0) return LDGA_OK;
  bse_bgj_kernel<NT><<<dim3(nblocks), dim3(64 * NT), 0, st>>>(N, Nq, Nc, shell, Gam, chi0, s, U, beta, max_ir, chi,
                                                             gamma, cnt, redo);
  LDGA_HIP_TRY(ctx, hipGetLastError());
  return LDGA_OK;
}

}  // namespace

// Mixed, block-pivoted: FP32 16-row blocked Gauss-Jordan on MFMA + FP64
// refinement, Nnu <= 160.  With redo != NULL (room for Nq Nw ints) the matrices
// whose refinement failed (or whose diagonal block met a zero pivot) are listed
// there; counters[2] holds their number.
int launch_bse_mixed(ldga_ctx *ctx, int N, int Nq, int Nw, int Nc, int shell, const double2 *Gam,
                     const double2 *chi0, double s, double U, double beta, double *chi, double2 *gamma,
                     unsigned long long *cnt, int *redo, hipStream_t st) {
  const int nt = (N + 15) / 16, ir = 6;
#define LDGA_BGJ(NT_) return run_bgj<NT_>(ctx, N, Nq, Nw, Nc, shell, Gam, chi0, s, U, beta, ir, chi, gamma, cnt, redo, st)
  switch (nt) {
    case 1:
    case 2:
    case 3: LDGA_BGJ(3);
    case 4: LDGA_BGJ(4);
    case 5: LDGA_BGJ(5);
    case 6: LDGA_BGJ(6);
    case 7: LDGA_BGJ(7);
    case 8: LDGA_BGJ(8);
    case 9: LDGA_BGJ(9);
    case 10: LDGA_BGJ(10);
    default: return ldga_fail(ctx, LDGA_ERR_SHAPE, "mixed BSE supports Nnu <= 160 (got %d)", N);
  }
#undef LDGA_BGJ
}
