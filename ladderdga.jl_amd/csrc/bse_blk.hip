// Batched Bethe-Salpeter solve on MFMA: blocked Gauss-Jordan inversion
// (calc_χγ, src/BSETools.jl:192-248; replaces the per-(q,w) zgetrf+zgetri of
// src/LapackWrapper.jl:64-72).
//
// For every (q,w) one workgroup inverts M = A^T, A = Gamma(w) + diag(1/chi0(q,.,w)),
// in place, by Gauss-Jordan elimination with column pivoting and implicit
// pivots (no data is ever swapped):
//   step for row r: p = argmax_{c unpivoted} |Re M[r][c]| + |Im M[r][c]|
//   (LAPACK's izamax criterion, first maximum), piv = M[r][p], then
//     M[i][c] -= M[i][p] M[r][c] / piv,  M[i][p] /= piv       (i != r, c != p)
//     M[r][c] = -M[r][c] / piv,          M[r][p] = 1 / piv.
// Afterwards M[r][c] = inv(M)[p_r][row_of(c)], so the solution of A^T y = 1,
// y = inv(M) 1, is y[p_r] = sum_c M[r][c] (a row sum of the stored matrix).
//
// Layout: the padded NC x NC matrix (NC = 16 NT, identity on the padding)
// lives in MFMA accumulators.  Wave w holds 16 rows as NT tiles of 16x16
// (real and imaginary planes); lane (g = lane/16, n = lane%16) holds column
// 16t+n of tile t and, in element j, matrix row 4 (j NT + w) + g (tile row
// trow(g, j)).  The row order is free: the elimination tracks which column
// each row pivots on.
//
// Rows are processed in panels of four: panel (j, w) = element j of every
// lane group of wave w.  The owner wave stages its rows in LDS, loads the
// panel rows one column per lane (columns lane + 64u), runs the four pivot
// steps there (64-lane DPP argmax, readlane broadcasts, no workgroup
// barrier), and publishes W = (panel after) - E_P.  After one barrier every
// wave applies the panel to its own rows as a rank-4 block update on MFMA,
//     M[S,:] += M[S,P] W     (S = rows outside the panel, M[S,P] before),
// the complex product split into four real K=4 products (4M scheme).  The
// panel ownership rotates over the waves.
//
// The FP64 path (f64 16x16x4 MFMA): its pivot sequence is LAPACK's (izamax
// over the whole row, up to ties), so its rounding tracks zgetrf+zgetri.  The
// mixed path is bse_bgj.hip.
#include <algorithm>
#include <climits>
#include <cstring>
#include <type_traits>

#include "ldga_internal.h"

#include "bse_common.h"

namespace {

using namespace ldga_bse;

template <typename T, int NT>
struct BlkShared {
  typedef typename mfma_traits<T>::c2 c2;
  static constexpr int NC = 16 * NT;
  // rows padded so that rows 1024 bytes apart (NC = 128) do not fall on the
  // same LDS banks (the column reads of S and the per-group reads of W)
  static constexpr int SPAD = sizeof(T) == 8 ? 4 : 2, WPAD = sizeof(T) == 8 ? 32 : 16;
  c2 W[3][4][NC + WPAD];   // published panel update (triple-buffered)
  c2 S[3][16][NC + SPAD];  // all sixteen rows of a panel's owner, staged (triple-buffered)
  T Cr[NT][16][4];     // per wave: own rows at the four pivot columns (real)
  T Ci[NT][16][4];     // (imaginary)
  int P[3][4];         // pivot columns of the panel
  unsigned pm[(NC + 31) / 32];  // pivoted columns (bit c % 32 of word c / 32); the skipped rows' too
  int pcol_of_row[NC]; // pivot column of the row at each accumulator position
  double2 y[NC];       // solution (FP64)
  double2 dinv[NC];    // 1/chi0 on the diagonal (FP64)
  double red[NT];
  int zero_pivot;
};

// One workgroup solves matrix lb = q + Nq w (the body of both kernels below).
template <typename T, int NT>
__device__ __forceinline__ void blk_solve(BlkShared<T, NT> &sh, const int lb, int N, int Nq, int Nc, int shell,
                                          const double2 *__restrict__ Gam, const double2 *__restrict__ chi0, double s,
                                          double U, double beta, double *__restrict__ chi_out,
                                          double2 *__restrict__ gamma_out, unsigned long long *__restrict__ counters) {
  typedef mfma_traits<T> MT;
  typedef typename MT::v4 v4;
  typedef typename MT::c2 c2;
  constexpr int NC = 16 * NT, NU = (NC + 63) / 64;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = lane >> 4, n = lane & 15;
  const int q = lb % Nq, w = lb / Nq;
  const double2 *G = Gam + (size_t)w * N * N;

  // Matrix row i is stored at accumulator element j of lane group g of wave wv
  // with i = 4 (j NT + wv) + g, so that panel kp (element kp / NT of wave
  // kp % NT) holds rows 4kp .. 4kp + 3: the padding rows come last and the
  // panels holding only padding rows (identity rows, whose updates are zero)
  // are skipped.
  const int NPe = (N + 3) / 4;  // panels processed (of 4 NT)
  auto pos_of_row = [&](int i) {
    const int kp = i >> 2;
    return 16 * (kp % NT) + MT::trow(i & 3, kp / NT);
  };
  for (int i = tid; i < NC; i += 64 * NT) {
    double2 d = cz(0.0, 0.0);
    if (i < N) d = crecip(chi0[q + (size_t)Nq * (shell + i + (size_t)Nc * w)]);
    sh.dinv[i] = d;
    sh.pcol_of_row[pos_of_row(i)] = i;  // the skipped identity rows keep their own column
  }
  if (tid == 0) sh.zero_pivot = 0;
  // the columns of the skipped padding rows (>= 4 NPe) never pivot
  for (int k = tid; k < (NC + 31) / 32; k += 64 * NT) {
    const int lo = 4 * NPe - 32 * k;
    sh.pm[k] = lo <= 0 ? ~0u : lo >= 32 ? 0u : ~0u << lo;
  }
  __syncthreads();

  // ---- M = A^T into the accumulators: M[i][c] = Gamma[c, i, w] + delta_ic / chi0[i]
  // (identity on the padding).  The loads carry no branches, so the loads of a
  // group of LG tiles are in flight together; groups are issued one after the
  // other to bound the registers they occupy.  Rows and columns past N - 1 are
  // clamped for the load (a valid address) and zeroed after the conversion.
  // Row i's diagonal lies in tile (j NT + wave) / 4, the same for the whole
  // wave, at lane i % 16: it is added there only, under a scalar branch.
  AccPlane<T, NT> Vr, Vi;  // every element is set below
  constexpr int LG = 4;
  const char *Gb = reinterpret_cast<const char *>(G);
  int tdiag[4];
  unsigned rofs[4];  // byte offset of row i (clamped) in Gamma(w)
  T dgx[4], dgy[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int i = 4 * (j * NT + wave) + g;
    rofs[j] = 16u * (unsigned)N * (unsigned)min(i, N - 1);
    tdiag[j] = __builtin_amdgcn_readfirstlane((j * NT + wave) >> 2);
    const double2 d = i < N ? sh.dinv[i] : cz(1.0, 0.0);  // 1/chi0[i], or 1 on the padding
    const bool here = n == (i & 15);
    dgx[j] = here ? (T)d.x : (T)0;
    dgy[j] = here ? (T)d.y : (T)0;
  }
#pragma unroll
  for (int t0 = 0; t0 < NT; t0 += LG) {
    // the lane index is laundered per group so that no address of a later
    // group is formed (and held in registers) early
    int nl = n;
    __asm__ volatile("" : "+v"(nl));
    double2 v[LG][4];
#pragma unroll
    for (int tt = 0; tt < LG; ++tt) {
      if (t0 + tt < NT) {
        const unsigned cofs = 16u * (unsigned)min(16 * (t0 + tt) + nl, N - 1);
#pragma unroll
        for (int j = 0; j < 4; ++j) v[tt][j] = *reinterpret_cast<const double2 *>(Gb + (cofs + rofs[j]));
      }
    }
#pragma unroll
    for (int tt = 0; tt < LG; ++tt) {
      const int t = t0 + tt;
      if (t < NT) {
        const bool cok = 16 * t + n < N;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const bool ok = cok && 4 * (j * NT + wave) + g < N;
          T x = ok ? (T)v[tt][j].x : (T)0, y = ok ? (T)v[tt][j].y : (T)0;
          if (t == tdiag[j]) {
            x += dgx[j];
            y += dgy[j];
            __asm__ volatile("" : "+v"(x), "+v"(y));  // kept a branch, not a select per tile
          }
          Vr.set(t, j, x);
          Vi.set(t, j, y);
        }
        __asm__ volatile("" : "+v"(Vr.regs(t)), "+v"(Vi.regs(t)));  // converted here, not later
      }
    }
    __asm__ volatile("" ::: "memory");
  }
  int zero_seen = 0;

  // ---- panels: a two-stage software pipeline -----------------------------------
  // Panel kp = element kp / NT of the lane groups of wave kp % NT.  Its owner
  // stages all 16 of its rows in LDS (S[kp % 3]) at the end of iteration kp-2,
  // after its MFMA update by W_{kp-2}, and factors the panel one iteration
  // early: in iteration kp-1 it loads the panel rows from S one column per
  // lane, applies W_{kp-1} to them in registers (the rows' values at the pivot
  // columns by readlane), factors them, publishes W_kp and goes straight to
  // the barrier.  Its own tile misses the update by W_{kp-1}: in iteration kp
  // it applies W_{kp-1} and W_kp together, both A operands taken from S, and
  // then writes the panel's final rows.  The serial chain per iteration is
  // thus: barrier, rank-4 update of four staged rows, four pivot steps,
  // publish; the other waves' MFMA updates overlap it.  W, P and S are
  // triple-buffered: W_kp is read up to iteration kp+1 and rewritten by the
  // owner of panel kp+3 after the barrier of iteration kp+2; S[kp % 3] is read
  // up to iteration kp and rewritten at the end of iteration kp+1.
  static_assert(NT >= 2, "the pipeline needs two waves");
  // Per-lane parts of the LDS addresses, re-laundered every iteration: the
  // compiler then forms each address as base + immediate offset instead of
  // keeping (and spilling) one hoisted register per array element.
  int Lg = g, Ln = n, Ll = lane;

  // all sixteen rows of this wave (the owner of panel m) -> staging buffer m % 3 (tile row order)
  auto stage_rows = [&](int m) {
    const int sb = m % 3;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) sh.S[sb][MT::trow(Lg, jj)][16 * t + Ln] = mkc<T>(Vr.get(t, jj), Vi.get(t, jj));
  };

  // this wave's rows at the pivot columns q0..q3 -> Cr/Ci[wave] (the A operand
  // of its update; the lanes holding column q write it from tile q / 16)
  auto gather_cols = [&](int q0, int q1, int q2, int q3) {
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      const int q = st == 0 ? q0 : st == 1 ? q1 : st == 2 ? q2 : q3;
      if (n == (q & 15)) {
        const v4 xr = Vr.select(q >> 4), xi = Vi.select(q >> 4);
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          sh.Cr[wave][MT::trow(Lg, jj)][st] = xr[jj];
          sh.Ci[wave][MT::trow(Lg, jj)][st] = xi[jj];
        }
      }
    }
  };
  // tile t += A W[wb] (a: tile row n, k = g): four K=4 f64 MFMAs (complex 4M scheme)
  auto mma_tile = [&](int t, const c2 &a, int wb) {
    v4 cr = Vr.tile(t), ci = Vi.tile(t);
    const c2 bb = sh.W[wb][Lg][16 * t + Ln];
    cr = MT::mma(a.x, bb.x, cr);
    cr = MT::mma(-a.y, bb.y, cr);
    ci = MT::mma(a.x, bb.y, ci);
    ci = MT::mma(a.y, bb.x, ci);
    Vr.set_tile(t, cr);
    Vi.set_tile(t, ci);
  };
  // the owner of panel kp, after factoring it: its tile += A1 W[w1] + A2 W[w2]
  // for panels kp-1 (pivots p, deferred) and kp (pivots q).  Its rows are as
  // of panel kp-2 (it factored panel kp meanwhile): in its accumulators and
  // staged in S[sb].  A1 = M[:, P], A2 = M[:, Q] + A1 W[w1][:, Q].
  // (The panel's own rows of the tile are overwritten with their final values
  // afterwards.)
  auto update_tile2 = [&](int sb, int p0, int p1, int p2, int p3, int w1, int q0, int q1, int q2, int q3, int w2) {
    const int qg = Lg == 0 ? q0 : Lg == 1 ? q1 : Lg == 2 ? q2 : q3;
    const int pg = Lg == 0 ? p0 : Lg == 1 ? p1 : Lg == 2 ? p2 : p3;
    const c2 a1 = sh.S[sb][Ln][pg];
    c2 a2 = sh.S[sb][Ln][qg], cp[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) cp[r] = sh.S[sb][Ln][r == 0 ? p0 : r == 1 ? p1 : r == 2 ? p2 : p3];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const c2 c = cp[r];
      const c2 wv = sh.W[w1][r][qg];
      a2.x = fma(c.x, wv.x, fma(-c.y, wv.y, a2.x));
      a2.y = fma(c.x, wv.y, fma(c.y, wv.x, a2.y));
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) mma_tile(t, a1, w1);
#pragma unroll
    for (int t = 0; t < NT; ++t) mma_tile(t, a2, w2);
    __builtin_amdgcn_wave_barrier();
  };
  // this wave's tile += tile[:, Q] W[wb]
  auto update_tile = [&](int q0, int q1, int q2, int q3, int wb) {
    gather_cols(q0, q1, q2, q3);
    wave_sync_lds();
    const c2 a = mkc<T>(sh.Cr[wave][Ln][Lg], sh.Ci[wave][Ln][Lg]);
#pragma unroll
    for (int t = 0; t < NT; ++t) mma_tile(t, a, wb);
    __builtin_amdgcn_wave_barrier();
  };

  // owner of panel kp1: load its staged rows one column per lane, apply panel
  // kp1-1 (pivots q0..q3, W[wb]) when pre, factor them and publish W_kp1.
  auto factor = [&](auto PRE, int kp1, int j1, int wb, int q0, int q1, int q2, int q3) {
    constexpr bool pre = decltype(PRE)::value;
    __builtin_amdgcn_s_setprio(3);
    const int sb = kp1 % 3;
    if constexpr (!pre) wave_sync_lds();  // own staging just before (later: a barrier in between)
    // the panel's rows in the staged copy (row r: element j1 of lane group r)
    auto prow = [&](int r) { return MT::trow(r, j1); };
    // the panel rows from the staged copy, one column per lane
    T pr[NU][4], pim[NU][4];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int c = lane + 64 * u;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        c2 v = mkc<T>(0, 0);
        if (c < NC) v = sh.S[sb][prow(r)][Ll + 64 * u];
        pr[u][r] = v.x;
        pim[u][r] = v.y;
      }
    }
    // the pivoted columns among lane + 64u, bit u (and the columns past NC)
    unsigned pmask = 0;
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int c = lane + 64 * u;
      const unsigned bit = c < NC ? (sh.pm[(Ll + 64 * u) >> 5] >> (c & 31)) & 1u : 1u;
      pmask |= bit << u;
    }
    if constexpr (pre) {
      // panel rows += rows[:, Q] W[wb]: W one column per lane; the rows' values
      // at the pivot columns (taken before the row changes) by readlane from
      // the lanes holding those columns
      c2 wv[4][NU];
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        const int c = lane + 64 * u;
#pragma unroll
        for (int k = 0; k < 4; ++k) wv[k][u] = c < NC ? sh.W[wb][k][Ll + 64 * u] : mkc<T>(0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        T cx[4], cy[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int qk = k == 0 ? q0 : k == 1 ? q1 : k == 2 ? q2 : q3;
          const int uk = qk >> 6, lk = qk & 63;
          T xr = pr[0][r], xi = pim[0][r];
#pragma unroll
          for (int u = 1; u < NU; ++u)
            if (uk == u) {
              xr = pr[u][r];
              xi = pim[u][r];
            }
          cx[k] = rdlane(xr, lk);
          cy[k] = rdlane(xi, lk);
        }
#pragma unroll
        for (int u = 0; u < NU; ++u) {
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            pr[u][r] = fma(cx[k], wv[k][u].x, fma(-cy[k], wv[k][u].y, pr[u][r]));
            pim[u][r] = fma(cx[k], wv[k][u].y, fma(cy[k], wv[k][u].x, pim[u][r]));
          }
        }
      }
    }
    int P[4];
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      // argmax over the unpivoted columns of panel row st, as one 32-bit key per
      // lane: the float bits of |re| + |im| (monotone for non-negative values)
      // with the low byte replaced by 255 - column, so ties go to the smaller
      // column; pivoted and padding columns have key 0.  A NaN value has the
      // largest key: its column is picked and the NaN propagates.
      unsigned key = 0;
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        const float v = (float)(fabs(pr[u][st]) + fabs(pim[u][st]));
        const unsigned k = (__float_as_uint(v) & 0xffffff00u) | (unsigned)(255 - (lane + 64 * u));
        if (!((pmask >> u) & 1u)) key = max(key, k);
      }
      key = wave_max_u32(key);
      const int p = 255 - (int)(__builtin_amdgcn_readlane((int)key, 63) & 0xff);
      P[st] = p;
      const int up = p >> 6, lp = p & 63;
      if (lane == lp) pmask |= 1u << up;
      // the pivot column at the four panel rows (row st gives the pivot)
      T cpx[4], cpy[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        T xr = pr[0][r], xi = pim[0][r];
#pragma unroll
        for (int u = 1; u < NU; ++u)
          if (up == u) {
            xr = pr[u][r];
            xi = pim[u][r];
          }
        cpx[r] = rdlane(xr, lp);
        cpy[r] = rdlane(xi, lp);
      }
      const T px = cpx[st], py = cpy[st];
      zero_seen |= (px == (T)0 && py == (T)0);
      // 1 / piv = conj(piv) / |piv|^2, reciprocal refined by Newton steps
      const T dd = px * px + py * py;
      T rd = fast_rcp(dd);
      rd = fma(rd, fma(-dd, rd, (T)1), rd);
      rd = fma(rd, fma(-dd, rd, (T)1), rd);
      const T ivx = px * rd, ivy = -py * rd;
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        const T rsx = pr[u][st] * ivx - pim[u][st] * ivy;
        const T rsy = pr[u][st] * ivy + pim[u][st] * ivx;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (r == st) continue;
          pr[u][r] = fma(-cpx[r], rsx, fma(cpy[r], rsy, pr[u][r]));
          pim[u][r] = fma(-cpx[r], rsy, fma(-cpy[r], rsx, pim[u][r]));
        }
        pr[u][st] = -rsx;
        pim[u][st] = -rsy;
      }
      // pivot column: M[i][p] = M[i][p] / piv, M[r][p] = 1 / piv
      // (as selects: a taken branch costs more than the selects)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const T nx = r == st ? ivx : cpx[r] * ivx - cpy[r] * ivy;
        const T ny = r == st ? ivy : cpx[r] * ivy + cpy[r] * ivx;
#pragma unroll
        for (int u = 0; u < NU; ++u) {
          const bool hit = u == up && lane == lp;
          pr[u][r] = hit ? nx : pr[u][r];
          pim[u][r] = hit ? ny : pim[u][r];
        }
      }
    }
    // publish W = panel - E_P, the pivots and each row's pivot column
    const int bo = kp1 % 3;
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int c = lane + 64 * u;
      if (c < NC) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const T wx = pr[u][r] - (c == P[r] ? (T)1 : (T)0);
          sh.W[bo][r][Ll + 64 * u] = mkc<T>(wx, pim[u][r]);
        }
      }
    }
    if (lane < 4) {
      const int p = lane == 0 ? P[0] : lane == 1 ? P[1] : lane == 2 ? P[2] : P[3];
      const int row = 16 * wave + MT::trow(lane, j1);
      sh.P[bo][lane] = p;
      sh.pcol_of_row[row] = p;
      atomicOr(&sh.pm[p >> 5], 1u << (p & 31));
    }
    __builtin_amdgcn_s_setprio(0);
  };

  if (wave == 0) stage_rows(0);
  if (wave == 1) stage_rows(1);
  if (wave == 0) factor(std::false_type{}, 0, 0, 0, 0, 0, 0, 0);
  for (int kp = 0; kp < NPe; ++kp) {
    const int o = kp % NT, j = kp / NT, bw = kp % 3;
    const int kp1 = kp + 1, o1 = kp1 % NT;
    __asm__ volatile("" : "+v"(Lg), "+v"(Ln), "+v"(Ll));
    __syncthreads();  // W_kp, P_kp published
    int Ps[4];
#pragma unroll
    for (int st = 0; st < 4; ++st) Ps[st] = __builtin_amdgcn_readfirstlane(sh.P[bw][st]);
    if (kp1 < NPe && wave == o1) {
      // the critical path: straight from the barrier to the next barrier; this
      // wave's own MFMA update by W_kp is applied in the next iteration
      factor(std::true_type{}, kp1, kp1 / NT, bw, Ps[0], Ps[1], Ps[2], Ps[3]);
    } else {
      if (wave == o && kp > 0) {  // with the update by W_{kp-1}, deferred while this wave factored panel kp
        const int bp = (kp + 2) % 3;
        update_tile2(kp % 3, __builtin_amdgcn_readfirstlane(sh.P[bp][0]), __builtin_amdgcn_readfirstlane(sh.P[bp][1]),
                     __builtin_amdgcn_readfirstlane(sh.P[bp][2]), __builtin_amdgcn_readfirstlane(sh.P[bp][3]), bp,
                     Ps[0], Ps[1], Ps[2], Ps[3], bw);
      } else {
        update_tile(Ps[0], Ps[1], Ps[2], Ps[3], bw);
      }
      if (wave == o) {  // panel rows: their final values
        const int pg = Ps[0] * (g == 0) + Ps[1] * (g == 1) + Ps[2] * (g == 2) + Ps[3] * (g == 3);
        auto body = [&](auto J) {
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const c2 v = sh.W[bw][Lg][16 * t + Ln];
            Vr.set(t, J, v.x + ((16 * t + n) == pg ? (T)1 : (T)0));
            Vi.set(t, J, v.y);
          }
        };
        if constexpr (decltype(Vr)::bank) body(__builtin_amdgcn_readfirstlane(j));
        else with_j(j, body);
      }
      // the owner of panel kp+2 stages its rows (updated through panel kp)
      if (kp + 2 < NPe && wave == (kp + 2) % NT) stage_rows(kp + 2);
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (zero_seen && lane == 0) sh.zero_pivot = 1;

  // ---- y[p_r] = sum_c M[r][c] ------------------------------------------------
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    T sx = 0, sy = 0;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      sx += Vr.get(t, jj);
      sy += Vi.get(t, jj);
    }
    sx = row_sum16(sx);
    sy = row_sum16(sy);
    if (n == 0) {
      const int row = 16 * wave + MT::trow(g, jj);
      sh.y[sh.pcol_of_row[row]] = cz((double)sx, (double)sy);
    }
  }
  __syncthreads();

  // ---- chi and gamma ---------------------------------------------------------
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
    if (sh.zero_pivot) atomicAdd(&counters[0], 1ull);
  }
}

// every (q, w), one workgroup each, XCD-aware
template <typename T, int NT>
__global__ __launch_bounds__(64 * NT, 1) void bse_blk_kernel(int N, int Nq, int Nc, int shell,
                                                          const double2 *__restrict__ Gam,
                                                          const double2 *__restrict__ chi0, double s, double U,
                                                          double beta, double *__restrict__ chi_out,
                                                          double2 *__restrict__ gamma_out,
                                                          unsigned long long *__restrict__ counters) {
  __shared__ BlkShared<T, NT> sh;
  blk_solve<T, NT>(sh, xcd_logical_block(blockIdx.x, gridDim.x), N, Nq, Nc, shell, Gam, chi0, s, U, beta, chi_out,
                   gamma_out, counters);
}

// over the matrices listed in list[0 .. *count - 1] (q + Nq w each): the
// re-solve of a mixed call's refinement failures.  The count is read on the
// device, so the host never waits for it; each workgroup takes the items
// blockIdx.x, blockIdx.x + gridDim.x, ...
template <typename T, int NT>
__global__ __launch_bounds__(64 * NT, 1) void bse_blk_list_kernel(
    int N, int Nq, int Nc, int shell, const double2 *__restrict__ Gam, const double2 *__restrict__ chi0, double s,
    double U, double beta, double *__restrict__ chi_out, double2 *__restrict__ gamma_out,
    unsigned long long *__restrict__ counters, const int *__restrict__ list,
    const unsigned long long *__restrict__ count) {
  __shared__ BlkShared<T, NT> sh;
  const int nitems = __builtin_amdgcn_readfirstlane((int)*count);
  for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
    blk_solve<T, NT>(sh, list[item], N, Nq, Nc, shell, Gam, chi0, s, U, beta, chi_out, gamma_out, counters);
    __syncthreads();  // the shared state is reused by the next item
  }
}

template <int NT>
int run_blk(ldga_ctx *ctx, int N, int Nq, int Nw, int Nc, int shell, const double2 *Gam, const double2 *chi0,
            double s, double U, double beta, double *chi, double2 *gamma, unsigned long long *cnt, const int *list,
            const unsigned long long *count, int max_items, hipStream_t st) {
  if (N > 16 * NT) return ldga_fail(ctx, LDGA_ERR_SHAPE, "BSE tile too small for N=%d", N);
  if (list) {
    if (max_items <= 0) return LDGA_OK;
    const unsigned grid = (unsigned)std::min(max_items, 1024);
    bse_blk_list_kernel<double, NT><<<dim3(grid), dim3(64 * NT), 0, st>>>(N, Nq, Nc, shell, Gam, chi0, s, U, beta,
                                                                          chi, gamma, cnt, list, count);
    LDGA_HIP_TRY(ctx, hipGetLastError());
    return LDGA_OK;
  }
  const unsigned nblocks = (unsigned)Nq * Nw;
  if (nblocks == 0) return LDGA_OK;
  bse_blk_kernel<double, NT><<<dim3(nblocks), dim3(64 * NT), 0, st>>>(N, Nq, Nc, shell, Gam, chi0, s, U, beta, chi,
                                                                      gamma, cnt);
  LDGA_HIP_TRY(ctx, hipGetLastError());
  return LDGA_OK;
}

}  // namespace

// FP64 blocked kernel: Nnu <= 128.  With list != NULL it solves just the
// matrices (q + Nq w) list[0 .. *count - 1], count read on the device, at most
// max_items of them.
int launch_bse_blk_f64(ldga_ctx *ctx, int N, int Nq, int Nw, int Nc, int shell, const double2 *Gam,
                       const double2 *chi0, double s, double U, double beta, double *chi, double2 *gamma,
                       unsigned long long *cnt, const int *list, const unsigned long long *count, int max_items,
                       hipStream_t st) {
  const int nt = (N + 15) / 16;
#define LDGA_BLK_F64(NT_) \
  return run_blk<NT_>(ctx, N, Nq, Nw, Nc, shell, Gam, chi0, s, U, beta, chi, gamma, cnt, list, count, max_items, st)
  switch (nt) {
    case 1:
    case 2:
    case 3: LDGA_BLK_F64(3);
    case 4: LDGA_BLK_F64(4);
    case 5: LDGA_BLK_F64(5);
    case 6: LDGA_BLK_F64(6);
    case 7: LDGA_BLK_F64(7);
    case 8: LDGA_BLK_F64(8);
    default: return ldga_fail(ctx, LDGA_ERR_SHAPE, "FP64 blocked BSE supports Nnu <= 128 (got %d)", N);
  }
#undef LDGA_BLK_F64
}
