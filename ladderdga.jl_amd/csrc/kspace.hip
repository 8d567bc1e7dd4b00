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
// from the transformed octant.  The functions are also symmetric under
// permutations of the axes (they depend on k through the dispersion of the
// square / simple-cubic lattice), so every product and sum over the octant is
// formed on the wedge points only and written to the permuted copies (the
// orbit) of each wedge point in LDS.
#include <algorithm>
#include <type_traits>

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
constexpr int kQ = 3;
__device__ __forceinline__ void oct_pass_paired(double2 *buf, const double *cs, int ax, const OctTables &T) {
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

// The same mode product for Ns divisible by 4 (h = Ns/2 = 2g), with one more
// split.  Folding the line, u[r] = x[r] + x[h-r] and v[r] = x[r] - x[h-r]
// (r < g; u[g] = x[g]), gives
//   y[2m]   = sum_{r<=g} u[r] cos(pi 2m r / h)        (cos is even about r = g)
//   y[2m+1] = sum_{r<g}  v[r] cos(pi (2m+1) r / h)    (odd about r = g, zero at g)
// and each half keeps the E/O pairing: outputs 2m and h-2m from u, outputs 2m+1
// and h-2m-1 from v.  The lines are half as long: half the multiply-adds of
// oct_pass_paired.  The fold is done in place (u at r, v at h-r) by the
// threads of the line, one barrier before the sums.
__device__ __forceinline__ void oct_pass_fold(double2 *buf, const double *cs, int ax, const OctTables &T) {
  const int n = T.n, h = T.Ns / 2, g = h / 2, stride = T.stride[ax];
  const int ne = g / 2 + 1, nod = (g + 1) / 2;  // pairs from u, pairs from v
  const int nbe = (ne + kQ - 1) / kQ, nqb = nbe + (nod + kQ - 1) / kQ;
  const int nlines = T.NO / n, per_round = kNT / nqb;
  const int li = threadIdx.x / nqb, b = threadIdx.x - li * nqb;
  const bool odd = b >= nbe;
  const int m0 = (odd ? b - nbe : b) * kQ, nm = odd ? nod : ne;
  const int cnt = odd ? g : g + 1;             // points of the folded line
  const int rstep = odd ? -stride : stride;    // v[r] sits at h - r
  double tw[kQ];
#pragma unroll
  for (int j = 0; j < kQ; ++j) {
    const int m = m0 + j;
    tw[j] = m < nm ? 2.0 * cs[odd ? 2 * m + 1 : 2 * m] : 0.0;
  }
  for (int l0 = 0; l0 < nlines; l0 += per_round) {
    const int line = l0 + li;
    const bool active = li < per_round && line < nlines;
    const int base = (line % stride) + (line / stride) * stride * n;
    if (active) {
      for (int r = b; r < g; r += nqb) {
        const double2 x0 = buf[base + r * stride], x1 = buf[base + (h - r) * stride];
        buf[base + r * stride] = cadd(x0, x1);
        buf[base + (h - r) * stride] = csub(x0, x1);
      }
    }
    __syncthreads();
    double2 E[kQ], O[kQ];
#pragma unroll
    for (int j = 0; j < kQ; ++j) E[j] = O[j] = cz(0.0, 0.0);
    if (active) {
      double c[kQ], cm[kQ];
#pragma unroll
      for (int j = 0; j < kQ; ++j) {
        c[j] = 1.0;
        cm[j] = 0.5 * tw[j];
      }
      auto step = [&](double2 xe, double2 xo) {
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
      };
      // nit steps of two points, the same count for both kinds; the reads of
      // the last step stay inside the line (points up to g + 1 <= h from either
      // end) and its points past cnt are zeroed
      const int nit = (g + 2) / 2;
      int pos = base + (odd ? h * stride : 0);
      for (int it = 0; it + 1 < nit; ++it, pos += 2 * rstep) step(buf[pos], buf[pos + rstep]);
      const int r = 2 * (nit - 1);
      double2 xe = buf[pos], xo = buf[pos + rstep];
      if (r >= cnt) xe = cz(0.0, 0.0);
      if (r + 1 >= cnt) xo = cz(0.0, 0.0);
      step(xe, xo);
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int j = 0; j < kQ; ++j) {
        const int m = m0 + j;
        if (m < nm) {
          const int q = odd ? 2 * m + 1 : 2 * m;
          buf[base + q * stride] = cadd(E[j], O[j]);
          if (h - q != q) buf[base + (h - q) * stride] = csub(E[j], O[j]);
        }
      }
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void oct_pass(double2 *buf, const double *cs, int ax, const OctTables &T) {
  if ((T.Ns & 3) == 0)
    oct_pass_fold(buf, cs, ax, T);
  else
    oct_pass_paired(buf, cs, ax, T);
}

// octant index of the point with packed coordinates pc (c0 | c1 << 8 | c2 << 16)
__device__ __forceinline__ int oct_of(int pc, const OctTables &T) {
  return (pc & 0xff) + T.n * (((pc >> 8) & 0xff) + T.n * (pc >> 16));
}

// The passes on the vector ALU (any even Ns).
struct ValuPass {
  __device__ __forceinline__ void init(const OctTables &) {}
  __device__ __forceinline__ void run(double2 *buf, const double *cs, int ax, const OctTables &T) const {
    oct_pass(buf, cs, ax, T);
  }
  // all D passes, for outputs read at wedge points only (oct_out)
  __device__ __forceinline__ void transform(double2 *buf, const double *cs, const OctTables &T) const {
    for (int ax = 0; ax < T.D; ++ax) run(buf, cs, ax, T);
  }
  __device__ __forceinline__ int oct_out(int pc, const OctTables &T) const { return oct_of(pc, T); }
};

// The folded pass with one real column (the real or imaginary part of a line)
// per lane and the whole line in registers, for Ns = NS divisible by 4: with
// the fold u, v of the line,
//   y[2m]   = sum_{r<=g} cos(2 pi 2m r / Ns) u[r]      (pairs m, g - m: E +- O)
//   y[2m+1] = sum_{r<g}  cos(2 pi (2m+1) r / Ns) v[r]  (pairs m, g - 1 - m)
// with E, O the sums over even and odd r.  Every index is a compile-time
// constant: the cosines are wave-uniform (the quarter table cos(2 pi j / Ns),
// j <= Ns/4, held in scalar registers, signs folded into the multiply-adds;
// the zero at j = Ns/4 skipped) and there is no recurrence.  A lane reads and
// writes only its own column, so the pass has one barrier, at its end.
template <int NS>
struct RegPass {
  static constexpr int h = NS / 2, n = h + 1, g = h / 2, nc = NS / 4;
  double cv[nc];  // cos(2 pi j / NS), j = 0 .. NS/4 - 1
  __device__ __forceinline__ void init(const OctTables &T) {
#pragma unroll
    for (int j = 0; j < nc; ++j) cv[j] = T.cosT[j];
  }
  static constexpr int fold_j(int j) {  // |index| into the quarter table, or nc for cos = 0
    j %= NS;
    j = j > NS / 2 ? NS - j : j;
    return j > NS / 4 ? NS / 2 - j : j;
  }
  static constexpr bool neg_j(int j) {
    j %= NS;
    j = j > NS / 2 ? NS - j : j;
    return j > NS / 4;
  }
  // acc + cos(2 pi j / NS) x
  __device__ __forceinline__ double mac(int j, double x, double acc) const {
    const int f = fold_j(j);
    if (f == nc) return acc;  // cos(pi / 2) = 0
    if (f == 0) return neg_j(j) ? acc - x : acc + x;
    return fma(neg_j(j) ? -cv[f] : cv[f], x, acc);
  }
  // the transform of one real column: point r at p[r st]
  __device__ __forceinline__ void column(double *p, int st) const {
    double u[g + 1], v[g];
#pragma unroll
    for (int r = 0; r < g; ++r) {
      const double a = p[r * st], b = p[(h - r) * st];
      u[r] = a + b;
      v[r] = a - b;
    }
    u[g] = p[g * st];
#pragma unroll
    for (int m = 0; 2 * m <= g; ++m) {
      double E = u[0], O = 0.0;
#pragma unroll
      for (int r = 1; r <= g; ++r) {
        if (r % 2 == 0)
          E = mac(2 * m * r, u[r], E);
        else
          O = mac(2 * m * r, u[r], O);
      }
      p[2 * m * st] = E + O;
      if (2 * m != g) p[(h - 2 * m) * st] = E - O;
    }
#pragma unroll
    for (int m = 0; 2 * m < g; ++m) {
      double E = v[0], O = 0.0;
#pragma unroll
      for (int r = 1; r < g; ++r) {
        if (r % 2 == 0)
          E = mac((2 * m + 1) * r, v[r], E);
        else
          O = mac((2 * m + 1) * r, v[r], O);
      }
      p[(2 * m + 1) * st] = E + O;
      if (4 * m + 2 != h) p[(h - 2 * m - 1) * st] = E - O;
    }
  }
  // this thread's index among the columns of a round: the columns of a partial
  // round go to different waves (SIMDs) in each pass and workgroup
  __device__ __forceinline__ int col0(int ax) const {
    constexpr int nw = kNT / 64;
    return (((threadIdx.x >> 6) + blockIdx.x + ax) & (nw - 1)) * 64 + (threadIdx.x & 63);
  }
  __device__ __forceinline__ void run(double2 *buf, const double *, int ax, const OctTables &T) const {
    const int stride = T.stride[ax], nlines = T.NO / n, ncol = 2 * nlines;
    double *bd = reinterpret_cast<double *>(buf);
    for (int col = col0(ax); col < ncol; col += kNT) {
      const int line = col >> 1;
      int base;  // octant index of the line's point 0
      if (ax == 0)
        base = line * n;
      else if (ax == T.D - 1)
        base = line;
      else {  // 3D, axis 1: line = a + n c
        const int c = (int)(((float)line + 0.5f) * (1.0f / n));
        base = line + c * (n * n - n);
      }
      column(bd + 2 * base + (col & 1), 2 * stride);
    }
    __syncthreads();
  }
  // All D passes for outputs read at wedge points (oct_out).  The output is
  // symmetric under permutations of the axes, so the last pass transforms only
  // the lines the wedge reads need: in 3D the lines (x, y) with x >= y (the
  // wedge point (c0, c1, c2) is read at x = c0, y = c1, z = c2); in 2D the
  // lines x < h, the points (h, c1) being read at (c1, h), and the corner
  // (h, h) = sum_y (-1)^y T(h, y) of the first pass's output summed apart.
  __device__ __forceinline__ void transform(double2 *buf, const double *cs, const OctTables &T) const {
    for (int ax = 0; ax + 1 < T.D; ++ax) run(buf, cs, ax, T);
    const int ax = T.D - 1;
    double *bd = reinterpret_cast<double *>(buf);
    if (T.D == 2) {
      const int c = col0(ax);
      if (c < 2 * h) {
        column(bd + 2 * (c >> 1) + (c & 1), 2 * n);
      } else if (c < 2 * h + 2) {  // the corner, one lane per part, in a wave with no column
        const int part = c & 1;
        double s = 0.0;
#pragma unroll
        for (int y = 0; y <= h; ++y) s += (y & 1) ? -bd[2 * (h + n * y) + part] : bd[2 * (h + n * y) + part];
        bd[2 * (h + n * h) + part] = s;
      }
    } else {
      constexpr int nl = n * (n + 1) / 2;
      for (int col = col0(ax); col < 2 * nl; col += kNT) {
        const int l = col >> 1;
        const int x = (int)((sqrtf(8.0f * (float)l + 1.0f) - 1.0f) * 0.5f);
        const int y = l - x * (x + 1) / 2;  // l = x (x + 1) / 2 + y, y <= x
        column(bd + 2 * (x + n * y) + (col & 1), 2 * n * n);
      }
    }
    __syncthreads();
  }
  __device__ __forceinline__ int oct_out(int pc, const OctTables &T) const {
    const int a = pc & 0xff, b = (pc >> 8) & 0xff;
    if (T.D == 2) return a == h ? b + n * h : a + n * b;
    return oct_of(pc, T);
  }
};

// sum of v over the four 16-lane rows of the wave, in every lane (permlane swaps)
__device__ __forceinline__ double row_sum4(double v) {
  const unsigned lo = __double2loint(v), hi = __double2hiint(v);
  const auto l32 = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto h32 = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  const double y = __hiloint2double(h32[0], l32[0]) + __hiloint2double(h32[1], l32[1]);
  const unsigned ylo = __double2loint(y), yhi = __double2hiint(y);
  const auto l16 = __builtin_amdgcn_permlane16_swap(ylo, ylo, false, false);
  const auto h16 = __builtin_amdgcn_permlane16_swap(yhi, yhi, false, false);
  return __hiloint2double(h16[0], l16[0]) + __hiloint2double(h16[1], l16[1]);
}

// The folded pass on the FP64 matrix cores, for Ns = 16 NKB (NKB = 2, 3, 4:
// Ns = 32, 48, 64; g = Ns/4 = 4 NKB).  A wave takes 8 lines (16 real columns: the
// real and imaginary parts) at a time.  With u and v of the fold,
//   y[2m]   = sum_{r<g} C1[m][r] u[r] + (-1)^m x[g]     C1[m][r] = cos(pi 2m r / h)
//   y[2m+1] = sum_{r<g} C2[m][r] v[r]                   C2[m][r] = cos(pi (2m+1) r / h)
//   y[h]    = sum_{r<g} (-1)^r u[r] + (-1)^g x[g]
// for m < g: two 16 x 16 x g products (v_mfma_f64_16x16x4f64, g/4 steps each),
// the cosine tiles held in registers for the whole kernel (A operand: lane l,
// row l % 16, k = l / 16), the fold done as the B operand is read (column l %
// 16, k = l / 16), and y[h] as a lane sum over the four rows.  Accumulator
// element e of lane l is row l / 16 + 4 e, column l % 16.  A wave reads and
// writes only its own lines, so a pass needs one barrier, at its end.
typedef double dbl4 __attribute__((ext_vector_type(4)));
template <int NKB>
struct MfmaPass {
  static constexpr int Ns = 16 * NKB, h = Ns / 2, g = h / 2;
  double ae[NKB], ao[NKB];
  __device__ __forceinline__ void init(const OctTables &T) {
    const int lane = threadIdx.x & 63, i = lane & 15, k = lane >> 4;
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb) {
      const int r = 4 * kb + k;
      const bool v = i < g && r < g;
      ae[kb] = v ? T.cosT[(2 * i * r) % Ns] : 0.0;
      ao[kb] = v ? T.cosT[((2 * i + 1) * r) % Ns] : 0.0;
    }
  }
  __device__ __forceinline__ void run(double2 *buf, const double *, int ax, const OctTables &T) const {
    run_lines(buf, ax, T, T.NO / (h + 1));
  }
  // pass ax over the lines 0 .. nlines-1 (the first nlines of the octant's line order)
  __device__ __forceinline__ void run_lines(double2 *buf, int ax, const OctTables &T, int nlines) const {
    constexpr int n = h + 1;
    const int stride = T.stride[ax];
    const int nct = (nlines + 7) / 8;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 15, k = lane >> 4, part = j & 1;
    double *bd = reinterpret_cast<double *>(buf);
    // the extra tile of a pass goes to a different wave in each pass and workgroup
    const int rot = (blockIdx.x + ax) & (kNT / 64 - 1);
    for (int ct = (wave + kNT / 64 - rot) & (kNT / 64 - 1); ct < nct; ct += kNT / 64) {
      const int line = ct * 8 + (j >> 1);
      const bool lv = line < nlines;
      const int lc = lv ? line : nlines - 1;  // past the last line: read a valid one, write nothing
      int base;                               // octant index of the line's point 0
      if (ax == 0)
        base = lc * n;
      else if (ax == T.D - 1)
        base = lc;
      else {  // 3D, axis 1: line = a + n c
        const int c = (int)(((float)lc + 0.5f) * (1.0f / n));
        base = lc + c * (n * n - n);
      }
      const int step = 8 * stride;  // doubles per 4 points
      const int plo = 2 * (base + k * stride) + part, phi = 2 * (base + (h - k) * stride) + part;
      double x0[NKB], x1[NKB];
#pragma unroll
      for (int kb = 0; kb < NKB; ++kb) {
        x0[kb] = bd[plo + kb * step];
        x1[kb] = bd[phi - kb * step];
      }
      const double xg = bd[2 * (base + g * stride) + part];
      dbl4 de = {0.0, 0.0, 0.0, 0.0}, dd = {0.0, 0.0, 0.0, 0.0};
      double s = 0.0;
#pragma unroll
      for (int kb = 0; kb < NKB; ++kb) {
        const double u = x0[kb] + x1[kb], v = x0[kb] - x1[kb];
        de = __builtin_amdgcn_mfma_f64_16x16x4f64(ae[kb], u, de, 0, 0, 0);
        dd = __builtin_amdgcn_mfma_f64_16x16x4f64(ao[kb], v, dd, 0, 0, 0);
        s += u;
      }
      const double xs = (k & 1) ? -xg : xg;  // (-1)^m x[g] for the rows m = k + 4e
      s = row_sum4((k & 1) ? -s : s);       // (-1)^r = (-1)^k for r = 4 kb + k
      if (lv) {
        const int w = 2 * (base + 2 * k * stride) + part, wstep = 16 * stride;
#pragma unroll
        for (int e = 0; e < NKB; ++e) {  // rows m = k + 4e < g
          bd[w + e * wstep] = de[e] + xs;
          bd[w + 2 * stride + e * wstep] = dd[e];
        }
        if (k == 0) bd[2 * (base + h * stride) + part] = s + ((g & 1) ? -xg : xg);
      }
    }
    __syncthreads();
  }
  // All D passes for outputs read at wedge points (oct_out).  In 2D the last
  // pass skips the line x = h: the points (h, c1) are read at (c1, h) (the
  // output is symmetric), and the corner (h, h) = sum_y (-1)^y T(h, y) of the
  // first pass's output is summed apart by two lanes before the last pass (no
  // tile of the last pass touches the line x = h).
  __device__ __forceinline__ void transform(double2 *buf, const double *cs, const OctTables &T) const {
    if (T.D != 2) {
      for (int ax = 0; ax < T.D; ++ax) run(buf, cs, ax, T);
      return;
    }
    run(buf, cs, 0, T);
    constexpr int n = h + 1;
    if (threadIdx.x < 2) {  // read before the last pass: it leaves the line x = h alone
      double *bd = reinterpret_cast<double *>(buf);
      const int part = threadIdx.x;
      double s = 0.0;
#pragma unroll
      for (int y = 0; y <= h; ++y) s += (y & 1) ? -bd[2 * (h + n * y) + part] : bd[2 * (h + n * y) + part];
      bd[2 * (h + n * h) + part] = s;
    }
    run_lines(buf, 1, T, h);
  }
  __device__ __forceinline__ int oct_out(int pc, const OctTables &T) const {
    const int a = pc & 0xff, b = (pc >> 8) & 0xff;
    if (T.D == 2) return a == h ? b + h * (h + 1) : a + (h + 1) * b;
    return oct_of(pc, T);
  }
};


__device__ __forceinline__ int trunc_half(int x) { return x / 2; }  // C division truncates toward 0 like Julia trunc

// Wedge points per thread (q = threadIdx.x + k kNT, k < NQS): NQS = Nq / kNT
// rounded up, at most 4 (Nq <= 969 for the grids ldga_set_kgrid accepts: 2D
// Ns <= 64: 561, 3D Ns <= 32: 969).  The kernels are instantiated for NQS = 1..4.

// This thread's wedge points: packed coordinates (-1 past Nq) and the octant
// index of the point itself.
template <int NQS>
struct WedgeSlots {
  int pc[NQS];
  __device__ __forceinline__ void load(const OctTables &T) {
#pragma unroll
    for (int k = 0; k < NQS; ++k) {
      const int q = threadIdx.x + k * kNT;
      pc[k] = q < T.Nq ? T.wedgecoord[q] : -1;
    }
  }
  __device__ __forceinline__ bool valid(int k) const { return pc[k] >= 0; }
  __device__ __forceinline__ int oct(int k, const OctTables &T) const {
    return (pc[k] & 0xff) + T.n * (((pc[k] >> 8) & 0xff) + T.n * (pc[k] >> 16));
  }
};

// v stored at every octant point of the orbit of packed coordinates pc (a
// repeated coordinate writes one point twice, with the same value)
__device__ __forceinline__ void orbit_store(double2 *buf, const OctTables &T, int pc, double2 v) {
  const int a = pc & 0xff, b = (pc >> 8) & 0xff, c = pc >> 16, n = T.n;
  if (T.D == 2) {
    buf[a + n * b] = v;
    buf[b + n * a] = v;
  } else {
    const int n2 = n * n;
    buf[a + n * b + n2 * c] = v;
    buf[a + n * c + n2 * b] = v;
    buf[b + n * a + n2 * c] = v;
    buf[b + n * c + n2 * a] = v;
    buf[c + n * a + n2 * b] = v;
    buf[c + n * b + n2 * a] = v;
  }
}

// ---------------------------------------------------------------------------
// The G tables on the wedge: dst[q + Nq f] = src[full index of wedge point q +
// Nk f] for the planes f of fft_range.  The octant kernels gather the wedge
// points of a plane for every (nu, w); on the full grid each such gather reads
// a cache line per point (the wedge is 1/2^D/D! of the grid), on the wedge it
// is one contiguous Nq-vector that stays in L2.
// ---------------------------------------------------------------------------
// (blockIdx.z selects the table: one launch compacts up to two tables)
__global__ void wedge_tables_kernel(OctTables T, int nplanes, const double2 *__restrict__ src0,
                                    double2 *__restrict__ dst0, const double2 *__restrict__ src1,
                                    double2 *__restrict__ dst1) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x, f = blockIdx.y;
  if (q >= T.Nq || f >= nplanes) return;
  const double2 *src = blockIdx.z ? src1 : src0;
  double2 *dst = blockIdx.z ? dst1 : dst0;
  dst[q + (size_t)T.Nq * f] = src[T.wedgefull[q] + (size_t)T.Nk * f];
}

// ---------------------------------------------------------------------------
// bubble: per (nu, w) pair the product of the two G tables on the octant, the
// octant transform on all D axes, then the wedge points read from it.  A
// workgroup takes a contiguous range of pairs (w-major, nu fastest) and gathers
// the next pair's table values while it transforms the current one, so that
// the gather latency hides behind the passes.
// ---------------------------------------------------------------------------
template <int NQS, class PS>
__global__ __launch_bounds__(kNT) void bubble_kernel(OctTables T, int Nnu, int n_inu_tot, int n_iw, int shift,
                                                     int fft_first, int per_wg, const double2 *__restrict__ Gfft,
                                                     const double2 *__restrict__ Grfft, double2 *__restrict__ chi0,
                                                     double scale) {
  extern __shared__ double2 smem[];
  double2 *buf = smem;
  double *cs = reinterpret_cast<double *>(smem + T.NO);
  const int npairs = Nnu * (2 * n_iw + 1);
  const int p0 = xcd_logical_block(blockIdx.x, gridDim.x) * per_wg, p1 = min(npairs, p0 + per_wg);
  if (p0 >= p1) return;
  for (int m = threadIdx.x; m < T.Ns; m += kNT) cs[m] = T.cosT[m];
  PS ps;
  ps.init(T);
  WedgeSlots<NQS> W;
  W.load(T);
  int f[NQS];
  double2 pw[NQS];
#pragma unroll
  for (int k = 0; k < NQS; ++k) {
    const int q = threadIdx.x + k * kNT;
    f[k] = W.valid(k) ? q : 0;
    pw[k] = W.valid(k) ? T.wedgepw[q] : cz(0.0, 0.0);
  }
  const size_t fs = (size_t)T.Nq;  // the G tables on the wedge, one plane per frequency
  double2 va[NQS], vb[NQS];
  auto gather = [&](int P) {
    const int nui = P % Nnu, wn = P / Nnu - n_iw;
    const int nun = -n_inu_tot + nui - trunc_half(shift * wn);
    const double2 *a = Gfft + (size_t)(nun - fft_first) * fs;
    const double2 *b = Grfft + (size_t)(nun + wn - fft_first) * fs;
#pragma unroll
    for (int k = 0; k < NQS; ++k) {
      va[k] = a[f[k]];
      vb[k] = b[f[k]];
    }
  };
  gather(p0);
  for (int P = p0; P < p1; ++P) {
#pragma unroll
    for (int k = 0; k < NQS; ++k)
      if (W.valid(k)) orbit_store(buf, T, W.pc[k], cmul(cmul(va[k], vb[k]), pw[k]));
    if (P + 1 < p1) gather(P + 1);
    __syncthreads();
    ps.transform(buf, cs, T);
    double2 *out = chi0 + (size_t)T.Nq * P;  // chi0(q, nu, w), P = nu + Nnu w
#pragma unroll
    for (int k = 0; k < NQS; ++k) {
      const int q = threadIdx.x + k * kNT;
      if (W.valid(k)) out[q] = cscale(buf[ps.oct_out(W.pc[k], T)], scale);
    }
    __syncthreads();  // the reads of buf are done before the next pair's stores
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
    // src/ladderDGATools.jl:267-270
    case LDGA_EOM_SP_01:
      return cscale(cz(gm.x * 0.5 * (1.0 + U * cm) - 0.5, gm.y * 0.5 * (1.0 + U * cm)), U);
    case LDGA_EOM_SP_02:
      return cscale(cz(gm.x * (1.0 + U * cm) - 1.0, gm.y * (1.0 + U * cm)), U);
    case LDGA_EOM_SP:
      return cscale(cz(gm.x * 1.5 * (1.0 + U * cm) - 1.5, gm.y * 1.5 * (1.0 + U * cm)), U);
    case LDGA_EOM_CH:
      return cscale(cz(gd.x * 0.5 * (1.0 - U * cd) - 0.5, gd.y * 0.5 * (1.0 - U * cd)), -U);
    default: {
      const double am = 1.5 * (1.0 + U * cm), ad = 0.5 * (1.0 - U * cd);
      return cscale(cz(gm.x * am - gd.x * ad - 1.5 + 0.5 + l0.x, gm.y * am - gd.y * ad + l0.y), U);
    }
  }
}

// K(q, nu, w) is formed on this thread's wedge points from gamma_m, gamma_d,
// chi_m^lam_m, chi_d^lam_d and lambda0 (the lambda map applied in registers;
// chi is never written), times the fold weight, and stored on each point's
// orbit; after the passes the product with G_rfft(nu + w) is accumulated on the
// wedge points, and the phase x weight of the back transform applied once at
// the end.  With tpart != NULL every workgroup L also sums kMult[q] chi_m^lam_m
// over its share (elements L kNT + i + j gridDim.x kNT of the Nq x Nw block) into
// tpart[L], for the 1/(i nu) tail of eom_final_kernel: the k-omega sum needs no
// launch or device-wide fence of its own.
template <int NQS, class PS>
__global__ __launch_bounds__(kNT) void eom_accum_kernel(OctTables T, int Nw, int Nnu, int n_inu, int n_iw, int shift,
                                                        int fft_first, int numax, int wchunk, int kind, double U,
                                                        double lam_m, double lam_d, const double *__restrict__ chim,
                                                        const double *__restrict__ chid,
                                                        const double2 *__restrict__ gm,
                                                        const double2 *__restrict__ gd,
                                                        const double2 *__restrict__ l0,
                                                        const double2 *__restrict__ Grfft,
                                                        double2 *__restrict__ partial, double *__restrict__ tpart) {
  extern __shared__ double2 smem[];
  double2 *buf = smem;
  double *cs = reinterpret_cast<double *>(smem + T.NO);
  // neighbouring nu (which gather overlapping G planes) on one XCD
  const int L = xcd_logical_block(blockIdx.x, gridDim.x);
  const int nuii = L % numax;
  const int chunk = L / numax;
  for (int m = threadIdx.x; m < T.Ns; m += kNT) cs[m] = T.cosT[m];
  PS ps;
  ps.init(T);
  WedgeSlots<NQS> W;
  W.load(T);
  int f[NQS];
  double wq[NQS];
  double2 S[NQS];
#pragma unroll
  for (int k = 0; k < NQS; ++k) {
    const int q = threadIdx.x + k * kNT;
    f[k] = W.valid(k) ? q : 0;
    wq[k] = W.valid(k) ? T.wedgew[q] : 0.0;
    S[k] = cz(0.0, 0.0);
  }
  const int w0 = chunk * wchunk, w1 = min(Nw, w0 + wchunk);
  const size_t fs = (size_t)T.Nq;  // the G tables on the wedge, one plane per frequency
  // nu-window of gamma at w: nu0 index (1-based) = n_inu + shift*trunc(wn/2) + 1  (helpers.jl:58); the
  // 0-based index of nu_n = nuii, skipped past the end of the nu list (singleCore.jl:59-60)
  auto nui_of = [&](int wi) { return n_inu + shift * trunc_half(wi - n_iw) + nuii; };
  auto next_w = [&](int wi) {
    while (wi < w1 && nui_of(wi) > 2 * n_inu - 1) ++wi;
    return wi;
  };
  // the inputs of K at this thread's wedge points, loaded one omega ahead: the
  // passes of one omega hide the loads of the next
  double km_[NQS], kd_[NQS];
  double2 kgm[NQS], kgd[NQS], kl0[NQS];
  auto load_k = [&](int wi) {
    const size_t col = (size_t)T.Nq * (nui_of(wi) + (size_t)Nnu * wi);
#pragma unroll
    for (int k = 0; k < NQS; ++k) {
      const int q = W.valid(k) ? threadIdx.x + k * kNT : 0;
      km_[k] = chim[q + (size_t)T.Nq * wi];
      kd_[k] = chid[q + (size_t)T.Nq * wi];
      kgm[k] = gm[q + col];
      kgd[k] = gd[q + col];
      kl0[k] = l0[q + col];
    }
  };
  int wi = next_w(w0);
  if (wi < w1) load_k(wi);
  if (tpart) {  // this workgroup's share of sum_{q,w} kMult[q] chi_m^lam_m (loads in flight with the first K inputs)
    const int total = T.Nq * Nw, step = gridDim.x * kNT;
    double v = 0.0;
    for (int i0 = L * kNT + threadIdx.x; i0 < total; i0 += 2 * step) {
      const int i1 = i0 + step;
      const double x0 = chim[i0], m0 = T.kmult[i0 % T.Nq];
      const double x1 = i1 < total ? chim[i1] : 0.0, m1 = i1 < total ? T.kmult[i1 % T.Nq] : 0.0;
      v = fma(m0, (lam_m != 0.0) ? x0 / (lam_m * x0 + 1.0) : x0, v);
      v = fma(m1, (lam_m != 0.0) ? x1 / (lam_m * x1 + 1.0) : x1, v);
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    double *red = reinterpret_cast<double *>(buf);  // buf is first written after the loop's first barrier
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < kNT / 64; ++w) t += red[w];
      tpart[L] = t;
    }
    __syncthreads();
  }
  while (wi < w1) {
    const int wn = wi - n_iw;
    // G(nu + w) at this thread's points, loaded first: the passes hide the latency
    const double2 *g = Grfft + (size_t)(nuii + wn - fft_first) * fs;
    double2 gv[NQS];
#pragma unroll
    for (int k = 0; k < NQS; ++k) gv[k] = g[f[k]];
    __syncthreads();  // the previous omega's reads of buf are done
#pragma unroll
    for (int k = 0; k < NQS; ++k) {
      if (W.valid(k)) {
        double cm = km_[k], cd = kd_[k];
        if (lam_m != 0.0) cm = cm / (lam_m * cm + 1.0);
        if (lam_d != 0.0) cd = cd / (lam_d * cd + 1.0);
        orbit_store(buf, T, W.pc[k], cscale(eom_value(kind, U, kgm[k], kgd[k], cm, cd, kl0[k]), wq[k]));
      }
    }
    const int wnext = next_w(wi + 1);
    if (wnext < w1) load_k(wnext);
    __syncthreads();
    ps.transform(buf, cs, T);
#pragma unroll
    for (int k = 0; k < NQS; ++k)
      if (W.valid(k)) S[k] = cadd(S[k], cmul(buf[ps.oct_out(W.pc[k], T)], gv[k]));
    wi = wnext;
  }
  double2 *dst = partial + (size_t)T.Nq * (nuii + (size_t)numax * chunk);
#pragma unroll
  for (int k = 0; k < NQS; ++k) {
    const int q = threadIdx.x + k * kNT;
    if (W.valid(k)) dst[q] = cmul(S[k], T.wedgepw[q]);
  }
}

// Sums the chunk partials of one nu on the wedge, stores them on the orbits
// and transforms back; adds the 1/(i nu) tail and the Hartree term.  The tail
// (with ntp > 0) comes from the ntp partial k-omega sums of chi_m^lam_m, added
// in order: sum_kw = (sum / Nk - tail_w + konst) / beta (DataTypes.jl:215-249),
// tail_w the omega sum of the 1/(i w)^2 tail and konst = -beta^2 c2 / 12.
template <int NQS, class PS>
__global__ __launch_bounds__(kNT) void eom_final_kernel(OctTables T, int numax, int nchunks,
                                                        const double2 *__restrict__ partial, double scale, double U,
                                                        double beta, double hartree, const double *__restrict__ tpart,
                                                        int ntp, double tail_w, double konst, double chiloc_m_sum,
                                                        double2 *__restrict__ sigma) {
  extern __shared__ double2 smem[];
  __shared__ double tred[kNT / 64];
  double2 *buf = smem;
  double *cs = reinterpret_cast<double *>(smem + T.NO);
  const int nuii = blockIdx.x;
  for (int m = threadIdx.x; m < T.Ns; m += kNT) cs[m] = T.cosT[m];
  // the tail's partial sums, reduced in a fixed order (thread strides, then
  // the wave tree, then the waves in order after the passes' barriers)
  if (ntp > 0) {
    double t = 0.0;
    for (int b = threadIdx.x; b < ntp; b += kNT) t += tpart[b];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off);
    if ((threadIdx.x & 63) == 0) tred[threadIdx.x >> 6] = t;
  }
  PS ps;
  ps.init(T);
  WedgeSlots<NQS> W;
  W.load(T);
  // the chunk partials of this nu, summed in chunk order: eight chunks of
  // every wedge slot in flight per round (one memory round trip per eight
  // chunks, not per chunk and slot)
  const size_t cstride = (size_t)T.Nq * numax;
  const double2 *pp = partial + (size_t)T.Nq * nuii;
  constexpr int CB = 8;
  int f[NQS];
  double2 s[NQS];
#pragma unroll
  for (int k = 0; k < NQS; ++k) {
    f[k] = W.valid(k) ? threadIdx.x + k * kNT : 0;
    s[k] = cz(0.0, 0.0);
  }
  for (int c0 = 0; c0 < nchunks; c0 += CB) {
    double2 a[NQS][CB];
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int k = 0; k < NQS; ++k) a[k][c] = c0 + c < nchunks ? pp[cstride * (c0 + c) + f[k]] : cz(0.0, 0.0);
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int k = 0; k < NQS; ++k) s[k] = cadd(s[k], a[k][c]);
  }
#pragma unroll
  for (int k = 0; k < NQS; ++k)
    if (W.valid(k)) orbit_store(buf, T, W.pc[k], s[k]);
  __syncthreads();
  ps.transform(buf, cs, T);
  // tail/(i nu_n) = -i * tail * beta / ((2n+1) pi)   (singleCore.jl:78-79)
  double tail = 0.0;
  if (ntp > 0) {
    double t = 0.0;
    for (int w = 0; w < kNT / 64; ++w) t += tred[w];
    tail = -U * ((t / T.Nk - tail_w + konst) / beta - chiloc_m_sum);
  }
  const double tail_im = -tail * beta / ((2.0 * nuii + 1.0) * M_PI);
  double2 *out = sigma + (size_t)T.Nq * nuii;
#pragma unroll
  for (int k = 0; k < NQS; ++k) {
    const int q = threadIdx.x + k * kNT;
    if (W.valid(k)) {
      const double2 v = cscale(buf[ps.oct_out(W.pc[k], T)], scale);
      out[q] = cz(v.x + hartree, v.y + tail_im);
    }
  }
}

// ---------------------------------------------------------------------------
// lambda0 (plain sums): fsum[nu',w] = sum_nu F[nu,nu',w]; lambda0 = chi0 fsum / beta^2
// ---------------------------------------------------------------------------
// One wave per column (nu', w): the lanes read the column's contiguous nu
// entries (coalesced) and reduce them with shuffles.
constexpr int kFsumNT = 256;
__global__ __launch_bounds__(kFsumNT) void fsum_kernel(int Nnu, int Nw, const double2 *__restrict__ F,
                                                       double2 *__restrict__ fsum) {
  const int col = blockIdx.x * (kFsumNT / 64) + (threadIdx.x >> 6);  // (nu', w)
  if (col >= Nnu * Nw) return;
  const int lane = threadIdx.x & 63;
  const double2 *c = F + (size_t)Nnu * col;
  double2 s = cz(0.0, 0.0);
  for (int v = lane; v < Nnu; v += 64) s = cadd(s, c[v]);
  for (int off = 32; off > 0; off >>= 1) {
    s.x += __shfl_xor(s.x, off);
    s.y += __shfl_xor(s.y, off);
  }
  if (lane == 0) fsum[col] = s;
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
// workgroup adds the tail terms (the w sum of the 1/(i w)^2 tail comes from the
// host) and writes the three sums.  Up to kSumGrid workgroups, each thread with
// kSumPer elements whose loads are all issued before the sums.
// ---------------------------------------------------------------------------
constexpr int kSumNT = 1024;  // the single-workgroup k integral below
constexpr int kSumRB = 256, kSumGrid = 256, kSumPer = 4;
__global__ __launch_bounds__(kSumRB) void sum_kw_kernel(const double *__restrict__ chi, const double *__restrict__ kmult,
                                                        int Nq, int w_lo, int w_hi, double Nk, double lam, double c2,
                                                        double beta, double tail, double *part, unsigned *ticket,
                                                        double *__restrict__ out) {
  double v[2] = {0.0, 0.0};
  const int nw = w_hi - w_lo + 1, total = Nq * nw, stride = gridDim.x * blockDim.x;
  for (int i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < total; i0 += kSumPer * stride) {
    double x[kSumPer], km[kSumPer];
#pragma unroll
    for (int u = 0; u < kSumPer; ++u) {
      const int i = i0 + u * stride;
      const bool ok = i < total;
      const int ic = ok ? i : 0;
      x[u] = ok ? chi[ic + Nq * w_lo] : 0.0;  // chi[q + Nq w], w = w_lo + i / Nq
      km[u] = ok ? kmult[ic - (ic / Nq) * Nq] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kSumPer; ++u) {
      const double f = (lam != 0.0) ? x[u] / (lam * x[u] + 1.0) : x[u];
      v[0] = fma(km[u], f, v[0]);
      v[1] = fma(km[u], -f * f, v[1]);
    }
  }
  double tot[2];
  if (!ldga_red::device_sum<2>(v, part, ticket, tot) || threadIdx.x != 0) return;
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
template <int NQS, class PS>
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
  // neighbouring nu (which gather overlapping G planes) on one XCD
  const int L = xcd_logical_block(blockIdx.x, gridDim.x);
  const int nuii = L % numax;
  const int chunk = L / numax;
  for (int m = threadIdx.x; m < T.Ns; m += kNT) cs[m] = T.cosT[m];
  PS ps;
  ps.init(T);
  WedgeSlots<NQS> P;
  P.load(T);
  int f[NQS];
#pragma unroll
  for (int k = 0; k < NQS; ++k) f[k] = P.valid(k) ? threadIdx.x + k * kNT : 0;
  // forward transform of W(., nu) (weights w(q) of the octant fold)
#pragma unroll
  for (int k = 0; k < NQS; ++k) {
    const int q = threadIdx.x + k * kNT;
    if (P.valid(k)) orbit_store(buf, T, P.pc[k], cscale(W[q + (size_t)T.Nq * nuii], T.wedgew[q]));
  }
  __syncthreads();
  ps.transform(buf, cs, T);
  // W^ times the phase x weight of the second transform, on this thread's wedge points
  double2 What[NQS];
#pragma unroll
  for (int k = 0; k < NQS; ++k)
    What[k] = P.valid(k) ? cmul(buf[ps.oct_out(P.pc[k], T)], T.wedgepw[threadIdx.x + k * kNT]) : cz(0.0, 0.0);
  double2 am = cz(0.0, 0.0), ad = cz(0.0, 0.0);
  const int w0 = chunk * wchunk, w1 = min(Nw, w0 + wchunk);
  const size_t fs = (size_t)T.Nq;  // the G tables on the wedge, one plane per frequency
  for (int wi = w0; wi < w1; ++wi) {
    const int wn = wi - n_iw;
    const int nui = n_inu + shift * trunc_half(wn) + nuii;
    if (nui > 2 * n_inu - 1) continue;
    const double2 *g = Grfft + (size_t)(nuii + wn - fft_first) * fs;
    double2 gv[NQS];
#pragma unroll
    for (int k = 0; k < NQS; ++k) gv[k] = g[f[k]];
    __syncthreads();  // previous users of buf are done
#pragma unroll
    for (int k = 0; k < NQS; ++k)
      if (P.valid(k)) orbit_store(buf, T, P.pc[k], cmul(What[k], gv[k]));
    __syncthreads();
    ps.transform(buf, cs, T);
#pragma unroll
    for (int k = 0; k < NQS; ++k) {
      const int q = threadIdx.x + k * kNT;
      if (!P.valid(k)) continue;
      const double2 t = cscale(buf[ps.oct_out(P.pc[k], T)], scale);
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
template <class PS>
__global__ __launch_bounds__(kNT) void gtables_kernel(OctTables T, const double2 *__restrict__ Gw,
                                                      double2 *__restrict__ Gfft, double2 *__restrict__ Grfft) {
  extern __shared__ double2 smem[];
  double2 *buf = smem;
  double *cs = reinterpret_cast<double *>(smem + T.NO);
  const int f = blockIdx.x;  // position in fft_range
  const double2 *g = Gw + (size_t)T.Nq * f;
  for (int m = threadIdx.x; m < T.Ns; m += kNT) cs[m] = T.cosT[m];
  PS ps;
  ps.init(T);
  for (int q = threadIdx.x; q < T.Nq; q += kNT) orbit_store(buf, T, T.wedgecoord[q], cscale(g[q], T.wedgew[q]));
  __syncthreads();
  for (int ax = 0; ax < T.D; ++ax) ps.run(buf, cs, ax, T);
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

// Launch an octant kernel (kNT threads); 3D octants need more than the
// default 64 KiB of LDS.
template <typename... KA, typename... A>
static int oct_launch(ldga_ctx *ctx, void (*kern)(KA...), int grid, size_t smem, hipStream_t st, A... args) {
  if (smem > 65536)
    LDGA_HIP_TRY(ctx, hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  kern<<<grid, kNT, smem, st>>>(args...);
  LDGA_HIP_TRY(ctx, hipGetLastError());
  return LDGA_OK;
}

// workgroups of kern (kNT threads, smem bytes of LDS) resident on the whole
// device at once; queried once per (kernel, smem) and cached in the context
static int resident_wgs(ldga_ctx *ctx, const void *kern, size_t smem) {
  const auto key = std::make_pair(kern, smem);
  const auto it = ctx->resident.find(key);
  if (it != ctx->resident.end()) return it->second;
  int ncu = 0, per_cu = 0;
  if (smem > 65536) (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) ncu = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kNT, smem) != hipSuccess) per_cu = 1;
  const int r = std::max(1, ncu) * std::max(1, per_cu);
  ctx->resident[key] = r;
  return r;
}

// calls f(std::integral_constant<int, NQS>) for the wedge slot count of the k grid
template <typename F>
static int with_nqs(ldga_ctx *ctx, F &&f) {
  switch ((ctx->oct.Nq + kNT - 1) / kNT) {
    case 1:
      return f(std::integral_constant<int, 1>{});
    case 2:
      return f(std::integral_constant<int, 2>{});
    case 3:
      return f(std::integral_constant<int, 3>{});
    case 4:
      return f(std::integral_constant<int, 4>{});
    default:
      return ldga_fail(ctx, LDGA_ERR_SHAPE, "wedge of %d points is too large", ctx->oct.Nq);
  }
}

// calls f(std::integral_constant<int, NQS>, pass) with the pass of the k grid
// (the wedge size of each grid below fixes NQS): the matrix-core pass for the
// 2D grids with 33 and 25 points a line (their 66 and 50 real columns fill a
// wave's 16-column tiles better than one column per lane), the
// register-resident pass for the other common Ns, the general vector pass
// otherwise
template <typename F>
static int with_cfg(ldga_ctx *ctx, F &&f) {
  const int Ns = ctx->oct.Ns, D = ctx->oct.D, nqs = (ctx->oct.Nq + kNT - 1) / kNT;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I4 = std::integral_constant<int, 4>;
  // the wedge of each (D, Ns) below has the slot count named with it
  if (D == 2 && Ns == 64 && nqs == 3) return f(I3{}, MfmaPass<4>{});
  if (D == 2 && Ns == 48 && nqs == 2) return f(I2{}, MfmaPass<3>{});
  if (D == 2 && Ns == 32 && nqs == 1) return f(I1{}, RegPass<32>{});
  if (D == 2 && Ns == 24 && nqs == 1) return f(I1{}, RegPass<24>{});
  if (D == 2 && Ns == 16 && nqs == 1) return f(I1{}, RegPass<16>{});
  if (D == 3 && Ns == 32 && nqs == 4) return f(I4{}, RegPass<32>{});
  if (D == 3 && Ns == 24 && nqs == 2) return f(I2{}, RegPass<24>{});
  if (D == 3 && Ns == 16 && nqs == 1) return f(I1{}, RegPass<16>{});
  return with_nqs(ctx, [&](auto c) { return f(c, ValuPass{}); });
}

// the full-grid G table src (Nk x n_fft) on the wedge, in the context's scratch of the given role
// the full-grid tables src0 (and src1, when given) on the wedge, in one launch;
// returns the compacted src0 (*dst1 receives the compacted src1)
static const double2 *wedge_tables(ldga_ctx *ctx, const ldga_freq *fr, const double2 *src0, const char *role0,
                                   hipStream_t st, const double2 *src1 = nullptr, const char *role1 = nullptr,
                                   const double2 **dst1 = nullptr) {
  const OctTables &T = ctx->oct;
  const size_t bytes = (size_t)T.Nq * fr->n_fft * 16;
  double2 *d0 = (double2 *)ldga_scratch(ctx, role0, bytes, st), *d1 = nullptr;
  if (!d0) return nullptr;
  if (src1 && !(d1 = (double2 *)ldga_scratch(ctx, role1, bytes, st))) return nullptr;
  wedge_tables_kernel<<<dim3((T.Nq + 255) / 256, fr->n_fft, src1 ? 2 : 1), 256, 0, st>>>(T, fr->n_fft, src0, d0,
                                                                                         src1, d1);
  if (hipGetLastError() != hipSuccess) {
    ldga_fail(ctx, LDGA_ERR_HIP, "wedge_tables_kernel launch failed");
    return nullptr;
  }
  if (dst1) *dst1 = d1;
  return d0;
}

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
  const double2 *Grw = nullptr;
  const double2 *Gw = wedge_tables(ctx, fr, Gfft, "bubble_gw", st, Grfft, "bubble_grw", &Grw);
  if (!Gw || !Grw) return LDGA_ERR_HIP;
  // one round of resident workgroups, each over a contiguous range of pairs
  return with_cfg(ctx, [&](auto c, auto ps) {
    auto kern = bubble_kernel<decltype(c)::value, decltype(ps)>;
    const int npairs = Nnu * Nw, res = resident_wgs(ctx, (const void *)kern, smem);
    const int per_wg = (npairs + res - 1) / res;
    return oct_launch(ctx, kern, (npairs + per_wg - 1) / per_wg, smem, st, T, Nnu, ntot, fr->n_iw, fr->shift,
                      fr->fft_first, per_wg, Gw, Grw, chi0, scale);
  });
}

int launch_gtables(ldga_ctx *ctx, const ldga_freq *fr, const double2 *Gw, double2 *Gfft, double2 *Grfft,
                   hipStream_t st) {
  const OctTables &T = ctx->oct;
  const size_t smem = (size_t)T.NO * 16 + (size_t)T.Ns * 8;
  return with_cfg(ctx, [&](auto, auto ps) {
    return oct_launch(ctx, gtables_kernel<decltype(ps)>, fr->n_fft, smem, st, T, Gw, Gfft, Grfft);
  });
}

int launch_lambda0(ldga_ctx *ctx, const ldga_freq *fr, double beta, const double2 *F, const double2 *chi0,
                   double2 *lambda0, hipStream_t st) {
  const int Nnu = 2 * fr->n_inu, Nw = 2 * fr->n_iw + 1, Nq = ctx->Nq;
  double2 *fsum = (double2 *)ldga_scratch(ctx, "fsum", (size_t)Nnu * Nw * 16, st);
  if (!fsum) return LDGA_ERR_HIP;
  fsum_kernel<<<(Nnu * Nw + kFsumNT / 64 - 1) / (kFsumNT / 64), kFsumNT, 0, st>>>(Nnu, Nw, F, fsum);
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
  // Re(c2/(i w_n)^2) = -c2 beta^2 / (4 pi^2 m^2), zero at m = 0, summed over the slice  (DataTypes.jl:223-225)
  const int n_iw = (Nw - 1) / 2;
  double tail = 0.0;
  for (int w = w_lo; w <= w_hi; ++w) {
    const int m = w - n_iw;
    if (m != 0) tail += -c2 * beta * beta / (4.0 * M_PI * M_PI * (double)m * (double)m);
  }
  sum_kw_kernel<<<grid, kSumRB, 0, st>>>(chi, T.kmult, T.Nq, w_lo, w_hi, T.Nk, lam, c2, beta, tail, part, ticket,
                                         out_dev);
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
  const size_t smem = (size_t)T.NO * 16 + (size_t)T.Ns * 8;
  int resident = 1;
  with_cfg(ctx, [&](auto c, auto ps) {
    resident = resident_wgs(ctx, (const void *)eom_accum_kernel<decltype(c)::value, decltype(ps)>, smem);
    return LDGA_OK;
  });
  int nchunks = std::max(1, std::min(Nw, resident / std::max(numax, 1)));
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
  if (eom_kind < LDGA_EOM_FULL || eom_kind > LDGA_EOM_CH) return ldga_fail(ctx, LDGA_ERR_ARG, "bad eom kind");
  // G(nu + w) must lie inside fft_range for every nu in 0..numax-1
  if (-fr->n_iw < fr->fft_first || numax - 1 + fr->n_iw > fr->fft_first + fr->n_fft - 1)
    return ldga_fail(ctx, LDGA_ERR_SHAPE, "fft_range does not cover nu+w");
  // the gamma window must contain nu_n = 0 at every w (helpers.jl:58)
  if (fr->n_inu - fr->shift * (fr->n_iw / 2) < 0)
    return ldga_fail(ctx, LDGA_ERR_SHAPE, "nu window does not contain nu_n = 0 at every w");
  int wchunk = 1;
  const int nchunks = eom_chunks(ctx, Nw, numax, &wchunk);
  const int nwg = numax * nchunks;
  double2 *partial = (double2 *)ldga_scratch(ctx, "eom_partial", (size_t)T.Nq * numax * nchunks * 16, st);
  const bool add_tail = tc && eom_kind == LDGA_EOM_FULL;
  // the accumulate workgroups' partial k-omega sums of chi_m for the tail (sum_kw over the whole omega range)
  double *tpart = add_tail ? (double *)ldga_scratch(ctx, "eom_tail_part", (size_t)nwg * 8, st) : nullptr;
  if (!partial || (add_tail && !tpart)) return LDGA_ERR_HIP;
  // Re(c2/(i w_n)^2) = -c2 beta^2 / (4 pi^2 m^2), zero at m = 0, summed over the omega grid  (DataTypes.jl:223-225)
  const double c2 = mp->chi_m_c2, beta = mp->beta;
  double tail_w = 0.0;
  for (int w = 0; w < Nw; ++w) {
    const int m = w - fr->n_iw;
    if (m != 0) tail_w += -c2 * beta * beta / (4.0 * M_PI * M_PI * (double)m * (double)m);
  }
  const double konst = -beta * beta / 12.0 * c2;
  const size_t smem = (size_t)T.NO * 16 + (size_t)T.Ns * 8;
  const double2 *Grw = wedge_tables(ctx, fr, Grfft, "eom_grw", st);
  if (!Grw) return LDGA_ERR_HIP;
  int rc = with_cfg(ctx, [&](auto c, auto ps) {
    return oct_launch(ctx, eom_accum_kernel<decltype(c)::value, decltype(ps)>, nwg, smem, st, T, Nw, Nnu, fr->n_inu,
                      fr->n_iw, fr->shift, fr->fft_first, numax, wchunk, eom_kind, mp->U, lam_m, lam_d, chi_m, chi_d,
                      gamma_m, gamma_d, lambda0, Grw, partial, tpart);
  });
  if (rc) return rc;
  const double hartree = (eom_kind == LDGA_EOM_FULL || eom_kind == LDGA_EOM_REST) ? mp->n * mp->U / 2.0 : 0.0;
  const double scale = 1.0 / (mp->beta * T.Nk * T.Nk);
  return with_cfg(ctx, [&](auto c, auto ps) {
    return oct_launch(ctx, eom_final_kernel<decltype(c)::value, decltype(ps)>, numax, smem, st, T, numax, nchunks, partial, scale,
                      mp->U, mp->beta, hartree, (const double *)tpart, add_tail ? nwg : 0, tail_w, konst,
                      chiloc_m_sum, sigma);
  });
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
  const double2 *Grw = wedge_tables(ctx, fr, Grfft, "eom_grw", st);
  if (!Grw) return LDGA_ERR_HIP;
  const int rc = with_cfg(ctx, [&](auto c, auto ps) {
    return oct_launch(ctx, eom_adjoint_kernel<decltype(c)::value, decltype(ps)>, numax * nchunks, smem, st, T, Nw, Nnu, fr->n_inu,
                      fr->n_iw, fr->shift, fr->fft_first, numax, wchunk, mp->U, lam_m, lam_d, chi_m, chi_d, gamma_m,
                      gamma_d, Grw, W, scale, partial);
  });
  if (rc) return rc;
  sum_pairs_kernel<<<1, 256, 0, st>>>(numax * nchunks, partial, out_dev);
  LDGA_HIP_TRY(ctx, hipGetLastError());
  return LDGA_OK;
}
