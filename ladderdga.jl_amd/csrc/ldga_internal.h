// Internal declarations shared by the libldga_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/ldga_hip.h"

// ---------------------------------------------------------------------------
// complex double helpers (double2 = (re, im), the ComplexF64 memory layout)
// ---------------------------------------------------------------------------
__host__ __device__ inline double2 cz(double re, double im) { return make_double2(re, im); }
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return cz(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return cz(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b) { return cz(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return cz(a.x * s, a.y * s); }
// 1/a as conj(a)/|a|^2 (a = 0 gives NaN, as a singular pivot should)
__host__ __device__ inline double2 crecip(double2 a) {
  double d = a.x * a.x + a.y * a.y;
  return cz(a.x / d, -a.y / d);
}
__host__ __device__ inline double2 cdiv(double2 a, double2 b) { return cmul(a, crecip(b)); }

// ---------------------------------------------------------------------------
// octant tables: the k-space convolutions run on the first octant
// R in [0, Ns/2]^D of the real-space (FFT) grid.  Every quantity of the path
// is even in k (eps_k = -2t sum cos k_d and the lattice symmetries map each k
// to its wedge point), so its transform is real-even and a cosine sum over the
// octant with weights prod_d w(R_d), w = 1 at R_d in {0, Ns/2} and 2 elsewhere,
// reproduces the full-grid FFT.
// ---------------------------------------------------------------------------
struct OctTables {
  int D, Ns, n, NO, Nq;
  // Per wedge point q (coordinates c0 >= c1 >= c2 in 0..Ns/2).  Every function
  // of the path is symmetric under permutations of the axes, so the octant
  // points of q's orbit (its coordinates permuted) hold q's value.
  const int32_t *wedgecoord;  // Nq: c0 | c1 << 8 | c2 << 16
  const int32_t *wedgefull;   // Nq: linear index of (c0, c1, c2) in the full Ns^D grid
  const double *wedgew;       // Nq: fold weight prod_d w(c_d), w = 1 at 0 and Ns/2, else 2
  const double2 *wedgephase;  // Nq: exp(-2 pi i sum_d c_d / Ns), undoes the reverse() of G_rfft
  const double2 *wedgepw;     // Nq: wedgephase * wedgew
  const double *cosT;        // Ns: cos(2 pi m / Ns)
  const double *kmult;       // Nq: wedge multiplicities
  const double *eps;         // Nq: dispersion on the wedge
  double Nk;                 // Ns^D
  int stride[3];             // octant strides 1, n, n^2
};

// Workgroups are dispatched round-robin over the 8 XCDs of the GPU (block b
// runs on XCD b % 8), each XCD with its own L2.  This maps the k-th block of
// XCD x to a contiguous range of logical indices, so that the blocks one XCD
// runs together work on neighbouring (q, w) points and share L2 lines of the
// q-fastest arrays they read and write.
__device__ __forceinline__ int xcd_logical_block(int b, int nblocks) {
  const int x = b & 7, k = b >> 3, per = nblocks >> 3, rem = nblocks & 7;
  return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  bool fresh = false;  // allocated since its last use as a ticket buffer
};

struct ldga_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // k grid
  bool has_kgrid = false;
  int D = 0, Ns = 0, Nq = 0, n = 0, NO = 0;
  std::vector<int32_t> h_wedge;
  std::vector<double> h_kmult, h_eps;
  OctTables oct{};
  // cached device buffers (grow-only, keyed by role)
  std::map<std::string, DevBuf> bufs;
  // device address of this device's copy of the allocation-free entry
  // point's static scratch array (set by ldga_create)
  void *noalloc_scratch = nullptr;
  // workgroups of a kernel resident on the whole device at once, by (kernel, LDS bytes)
  std::map<std::pair<const void *, size_t>, int> resident;
};

// looks up the static scratch array of the allocation-free entry point on the
// current device (hipGetSymbolAddress); called once per context at creation
int ldga_noalloc_setup(ldga_ctx *ctx);

// error plumbing ------------------------------------------------------------
int ldga_fail(ldga_ctx *ctx, int code, const char *fmt, ...);
#define LDGA_HIP_TRY(ctx, expr)                                                                             \
  do {                                                                                                      \
    hipError_t e_ = (expr);                                                                                 \
    if (e_ != hipSuccess) return ldga_fail((ctx), LDGA_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                           __FILE__, __LINE__);                                             \
  } while (0)

// cached device buffers (returns nullptr on failure and sets ctx->err): ldga_buf
// for the host-array forms' staging, ldga_scratch for the launchers' scratch
// (one per role and stream), ldga_tickets for device_sum's ticket counters
void *ldga_buf(ldga_ctx *ctx, const char *role, size_t bytes);
void *ldga_scratch(ldga_ctx *ctx, const char *role, size_t bytes, hipStream_t st);
unsigned *ldga_tickets(ldga_ctx *ctx, const char *role, int n, hipStream_t st);
hipStream_t ldga_stream(ldga_ctx *ctx, void *stream);
int ldga_check_freq(ldga_ctx *ctx, const ldga_freq *fr);

// kernel launchers (device pointers, asynchronous) ---------------------------
int launch_gtables(ldga_ctx *ctx, const ldga_freq *fr, const double2 *Gw, double2 *Gfft, double2 *Grfft,
                   hipStream_t st);
int launch_bubble(ldga_ctx *ctx, const ldga_freq *fr, double beta, const double2 *Gfft, const double2 *Grfft,
                  double2 *chi0, hipStream_t st);
int launch_bse(ldga_ctx *ctx, int s, int precision, const ldga_freq *fr, double U, double beta, const double2 *Gam,
               const double2 *chi0, double *chi, double2 *gamma, ldga_bse_report *rep, hipStream_t st);
int launch_bse_with(ldga_ctx *ctx, int s, int precision, const ldga_freq *fr, double U, double beta,
                    const double2 *Gam, const double2 *chi0, double *chi, double2 *gamma, ldga_bse_report *rep,
                    hipStream_t st, unsigned long long *cnt, int *redo);
int launch_lambda0(ldga_ctx *ctx, const ldga_freq *fr, double beta, const double2 *F, const double2 *chi0,
                   double2 *lambda0, hipStream_t st);
int launch_sum_kw(ldga_ctx *ctx, const double *chi, int Nw, int w_lo, int w_hi, double c2, double beta, double lam,
                  double *out_dev, hipStream_t st);
int launch_F_from_chi(ldga_ctx *ctx, const ldga_freq *fr, double beta, int diag_term, const double2 *chi,
                      const double2 *G, double2 *F, hipStream_t st);
int launch_gen_chi(ldga_ctx *ctx, const ldga_freq *fr, const double2 *Gam, const double2 *chi0, double2 *inv,
                   hipStream_t st);
int launch_sum_w(ldga_ctx *ctx, const double *chi, int Nq, int Nw, int w_lo, int w_hi, double c2, double beta,
                 double lam, double *res, double *out_k, hipStream_t st);
// EoM partial sums: one (nu, omega-chunk) workgroup each; returns the chunking used
int eom_chunks(ldga_ctx *ctx, int Nw, int numax, int *wchunk);
int launch_sigma(ldga_ctx *ctx, const ldga_freq *fr, const ldga_model *mp, int numax, double lam_m, double lam_d,
                 int tc, int eom_kind, double chiloc_m_sum, const double *chi_m, const double2 *gamma_m,
                 const double *chi_d, const double2 *gamma_d, const double2 *lambda0, const double2 *Grfft,
                 double2 *sigma, hipStream_t st);
int launch_eom_adjoint(ldga_ctx *ctx, const ldga_freq *fr, const ldga_model *mp, int numax, double lam_m,
                       double lam_d, const double *chi_m, const double2 *gamma_m, const double *chi_d,
                       const double2 *gamma_d, const double2 *Grfft, const double2 *W, double2 *out_dev,
                       hipStream_t st);
