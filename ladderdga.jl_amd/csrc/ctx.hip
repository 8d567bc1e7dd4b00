// Context, k-grid tables, buffer cache and the host-array (Julia-facing) entry
// points of libldga_hip.so.  The host-array forms stage through cached device
// buffers and call the same launchers as the _dev forms.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>

#include "ldga_internal.h"

int ldga_fail(ldga_ctx *ctx, int code, const char *fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return code;
}

void *ldga_buf(ldga_ctx *ctx, const char *role, size_t bytes) {
  DevBuf &b = ctx->bufs[role];
  if (b.bytes >= bytes && b.ptr) return b.ptr;
  if (b.ptr) {
    hipStreamSynchronize(ctx->stream);
    hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
  }
  if (hipMalloc(&b.ptr, std::max<size_t>(bytes, 256)) != hipSuccess) {
    b.ptr = nullptr;
    ldga_fail(ctx, LDGA_ERR_HIP, "hipMalloc of %zu bytes for %s failed", bytes, role);
    return nullptr;
  }
  b.bytes = bytes;
  return b.ptr;
}

hipStream_t ldga_stream(ldga_ctx *ctx, void *stream) { return stream ? (hipStream_t)stream : ctx->stream; }

// Scratch of the launchers, one buffer per (role, stream): _dev calls on
// different caller streams never share partial sums, counters or lists.
void *ldga_scratch(ldga_ctx *ctx, const char *role, size_t bytes, hipStream_t st) {
  char key[96];
  snprintf(key, sizeof key, "%s#%p", role, (void *)st);
  DevBuf &b = ctx->bufs[key];
  if (b.bytes >= bytes && b.ptr) return b.ptr;
  if (b.ptr) {
    hipStreamSynchronize(st);
    hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
  }
  if (hipMalloc(&b.ptr, std::max<size_t>(bytes, 256)) != hipSuccess) {
    b.ptr = nullptr;
    ldga_fail(ctx, LDGA_ERR_HIP, "hipMalloc of %zu bytes for %s failed", bytes, role);
    return nullptr;
  }
  b.bytes = bytes;
  b.fresh = true;
  return b.ptr;
}

namespace {
__global__ void zero_u32_kernel(unsigned *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}
}  // namespace

// n ticket counters of device_sum (ldga_reduce.h): zeroed by a kernel when the
// buffer is new; the kernels that use them leave them at zero
unsigned *ldga_tickets(ldga_ctx *ctx, const char *role, int n, hipStream_t st) {
  unsigned *p = (unsigned *)ldga_scratch(ctx, role, (size_t)n * sizeof(unsigned), st);
  if (!p) return nullptr;
  char key[96];
  snprintf(key, sizeof key, "%s#%p", role, (void *)st);
  DevBuf &b = ctx->bufs[key];
  if (b.fresh) {
    zero_u32_kernel<<<(n + 255) / 256, 256, 0, st>>>(p, n);
    if (hipGetLastError() != hipSuccess) {
      ldga_fail(ctx, LDGA_ERR_HIP, "clearing the %s tickets failed", role);
      return nullptr;
    }
    b.fresh = false;
  }
  return p;
}

int ldga_check_freq(ldga_ctx *ctx, const ldga_freq *fr) {
  if (!fr) return ldga_fail(ctx, LDGA_ERR_ARG, "freq is NULL");
  if (fr->n_iw < 0 || fr->n_inu < 1 || fr->n_inu_shell < 0 || fr->n_fft < 1 || (fr->shift != 0 && fr->shift != 1))
    return ldga_fail(ctx, LDGA_ERR_SHAPE, "invalid frequency parameters");
  return LDGA_OK;
}

extern "C" {

int ldga_version(void) { return 10000; }

int ldga_create(int device, ldga_ctx **out) {
  if (!out) return LDGA_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return LDGA_ERR_HIP;
  ldga_ctx *ctx = new ldga_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return LDGA_ERR_HIP;
  }
  if (ldga_noalloc_setup(ctx) != LDGA_OK) {
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return LDGA_ERR_HIP;
  }
  *out = ctx;
  return LDGA_OK;
}

static void free_kgrid(ldga_ctx *ctx) {
  const void *p[] = {ctx->oct.wedgecoord, ctx->oct.wedgefull, ctx->oct.wedgew, ctx->oct.wedgephase,
                     ctx->oct.wedgepw,    ctx->oct.cosT,      ctx->oct.kmult,  ctx->oct.eps};
  for (const void *q : p)
    if (q) hipFree(const_cast<void *>(q));
  ctx->oct = OctTables{};
  ctx->has_kgrid = false;
}

int ldga_destroy(ldga_ctx *ctx) {
  if (!ctx) return LDGA_OK;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  free_kgrid(ctx);
  for (auto &kv : ctx->bufs)
    if (kv.second.ptr) hipFree(kv.second.ptr);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return LDGA_OK;
}

const char *ldga_last_error(const ldga_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int ldga_sync(ldga_ctx *ctx) {
  if (!ctx) return LDGA_ERR_ARG;
  LDGA_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return LDGA_OK;
}

int ldga_malloc(ldga_ctx *ctx, uint64_t bytes, void **dptr) {
  if (!ctx || !dptr) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  hipSetDevice(ctx->device);
  LDGA_HIP_TRY(ctx, hipMalloc(dptr, std::max<uint64_t>(bytes, 1)));
  return LDGA_OK;
}

int ldga_free(ldga_ctx *ctx, void *dptr) {
  if (!ctx) return LDGA_ERR_ARG;
  hipSetDevice(ctx->device);
  LDGA_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  LDGA_HIP_TRY(ctx, hipFree(dptr));
  return LDGA_OK;
}

int ldga_memcpy_h2d(ldga_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!ctx || !dst || !src) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  hipSetDevice(ctx->device);
  LDGA_HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  LDGA_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return LDGA_OK;
}

int ldga_memcpy_d2h(ldga_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!ctx || !dst || !src) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  hipSetDevice(ctx->device);
  LDGA_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  LDGA_HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  LDGA_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return LDGA_OK;
}

// ---------------------------------------------------------------------------
// k grid
// ---------------------------------------------------------------------------
static int canon_key(const int *c, int D, int n) {
  int s[3] = {0, 0, 0};
  for (int d = 0; d < D; ++d) s[d] = c[d];
  std::sort(s, s + D, [](int a, int b) { return a > b; });
  int key = 0;
  for (int d = D - 1; d >= 0; --d) key = key * n + s[d];
  return key;
}

int ldga_default_wedge(int D, int Ns, int32_t *wedge, double *kmult) {
  if ((D != 2 && D != 3) || Ns < 2 || (Ns & 1)) return LDGA_ERR_SHAPE;
  const int n = Ns / 2 + 1;
  int Nq = 0;
  // descending coordinates c0 >= c1 (>= c2), enumerated with c0 slowest
  std::vector<int> pts;
  if (D == 2) {
    for (int a = 0; a < n; ++a)
      for (int b = 0; b <= a; ++b) pts.insert(pts.end(), {a, b});
  } else {
    for (int a = 0; a < n; ++a)
      for (int b = 0; b <= a; ++b)
        for (int c = 0; c <= b; ++c) pts.insert(pts.end(), {a, b, c});
  }
  Nq = (int)pts.size() / D;
  if (wedge)
    for (int i = 0; i < Nq * D; ++i) wedge[i] = pts[i];
  if (kmult) {
    std::map<int, int> idx;
    for (int q = 0; q < Nq; ++q) idx[canon_key(&pts[q * D], D, n)] = q;
    std::fill(kmult, kmult + Nq, 0.0);
    int c[3] = {0, 0, 0};
    const int Nk = D == 2 ? Ns * Ns : Ns * Ns * Ns;
    for (int f = 0; f < Nk; ++f) {
      int r = f;
      for (int d = 0; d < D; ++d) {
        int x = r % Ns;
        r /= Ns;
        c[d] = std::min(x, Ns - x);
      }
      kmult[idx[canon_key(c, D, n)]] += 1.0;
    }
  }
  return Nq;
}

int ldga_set_kgrid(ldga_ctx *ctx, int D, int Ns, int Nq, const int32_t *wedge, const double *eps) {
  if (!ctx || !wedge) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  if ((D != 2 && D != 3) || Ns < 2 || (Ns & 1) || (D == 2 && Ns > 64) || (D == 3 && Ns > 32))
    return ldga_fail(ctx, LDGA_ERR_SHAPE, "k grid D=%d Ns=%d outside the supported range (2D Ns<=64, 3D Ns<=32)", D, Ns);
  const int n = Ns / 2 + 1;
  const int NO = D == 2 ? n * n : n * n * n;
  const int Nexp = ldga_default_wedge(D, Ns, nullptr, nullptr);
  if (Nq != Nexp) return ldga_fail(ctx, LDGA_ERR_SHAPE, "wedge has %d points, expected %d", Nq, Nexp);
  hipSetDevice(ctx->device);
  // wedge lookup by canonical key
  std::map<int, int> idx;
  for (int q = 0; q < Nq; ++q) {
    for (int d = 0; d < D; ++d) {
      int c = wedge[d + D * q];
      if (c < 0 || c >= n) return ldga_fail(ctx, LDGA_ERR_SHAPE, "wedge coordinate %d of point %d outside [0,Ns/2]", c, q);
      if (d > 0 && c > wedge[d - 1 + D * q])
        return ldga_fail(ctx, LDGA_ERR_SHAPE, "wedge point %d coordinates are not descending", q);
    }
    int key = canon_key(&wedge[D * q], D, n);
    if (idx.count(key)) return ldga_fail(ctx, LDGA_ERR_SHAPE, "wedge point %d repeated", q);
    idx[key] = q;
  }
  // every octant point is a permutation of exactly one wedge point
  for (int o = 0; o < NO; ++o) {
    int c[3] = {o % n, (o / n) % n, D == 3 ? o / (n * n) : 0};
    if (!idx.count(canon_key(c, D, n)))
      return ldga_fail(ctx, LDGA_ERR_SHAPE, "octant point %d has no wedge point", o);
  }
  std::vector<int32_t> wcoord(Nq), wfull(Nq);
  std::vector<double> ww(Nq), cosT(Ns), kmult(Nq, 0.0);
  std::vector<double2> wph(Nq), wpw(Nq);
  const int Nk = D == 2 ? Ns * Ns : Ns * Ns * Ns;
  for (int q = 0; q < Nq; ++q) {
    const int *c = &wedge[D * q];
    const int c2 = D == 3 ? c[2] : 0;
    wcoord[q] = c[0] | (c[1] << 8) | (c2 << 16);
    wfull[q] = c[0] + Ns * c[1] + Ns * Ns * c2;
    double w = 1.0;
    int sum = 0;
    for (int d = 0; d < D; ++d) {
      w *= (c[d] == 0 || c[d] == Ns / 2) ? 1.0 : 2.0;
      sum += c[d];
    }
    ww[q] = w;
    const double ang = -2.0 * M_PI * (double)(sum % Ns) / (double)Ns;
    wph[q] = make_double2(std::cos(ang), std::sin(ang));
    wpw[q] = make_double2(w * wph[q].x, w * wph[q].y);
  }
  for (int m = 0; m < Ns; ++m) cosT[m] = std::cos(2.0 * M_PI * (double)m / (double)Ns);
  {
    int c[3];
    for (int f = 0; f < Nk; ++f) {
      int r = f;
      for (int d = 0; d < D; ++d) {
        int x = r % Ns;
        r /= Ns;
        c[d] = std::min(x, Ns - x);
      }
      kmult[idx.at(canon_key(c, D, n))] += 1.0;
    }
  }
  free_kgrid(ctx);
  OctTables &T = ctx->oct;
  auto up = [&](const void *src, size_t bytes, void **dst) -> int {
    LDGA_HIP_TRY(ctx, hipMalloc(dst, bytes));
    LDGA_HIP_TRY(ctx, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return LDGA_OK;
  };
  int rc = 0;
  void *p;
  if ((rc = up(wcoord.data(), Nq * 4, &p))) return rc;
  T.wedgecoord = (int32_t *)p;
  if ((rc = up(wfull.data(), Nq * 4, &p))) return rc;
  T.wedgefull = (int32_t *)p;
  if ((rc = up(ww.data(), Nq * 8, &p))) return rc;
  T.wedgew = (double *)p;
  if ((rc = up(wph.data(), Nq * 16, &p))) return rc;
  T.wedgephase = (double2 *)p;
  if ((rc = up(wpw.data(), Nq * 16, &p))) return rc;
  T.wedgepw = (double2 *)p;
  if ((rc = up(cosT.data(), Ns * 8, &p))) return rc;
  T.cosT = (double *)p;
  if ((rc = up(kmult.data(), Nq * 8, &p))) return rc;
  T.kmult = (double *)p;
  {
    std::vector<double> e(Nq, 0.0);
    if (eps) e.assign(eps, eps + Nq);
    if ((rc = up(e.data(), Nq * 8, &p))) return rc;
    T.eps = (double *)p;
  }
  T.D = D;
  T.Ns = Ns;
  T.n = n;
  T.NO = NO;
  T.Nq = Nq;
  T.Nk = (double)Nk;
  T.stride[0] = 1;
  T.stride[1] = n;
  T.stride[2] = n * n;
  ctx->D = D;
  ctx->Ns = Ns;
  ctx->Nq = Nq;
  ctx->n = n;
  ctx->NO = NO;
  ctx->h_wedge.assign(wedge, wedge + D * Nq);
  ctx->h_kmult = kmult;
  ctx->h_eps.assign(Nq, 0.0);
  if (eps) ctx->h_eps.assign(eps, eps + Nq);
  ctx->has_kgrid = true;
  return LDGA_OK;
}

// ---------------------------------------------------------------------------
// host-array entry points
// ---------------------------------------------------------------------------
#define NEED_KGRID(ctx) \
  if (!(ctx)->has_kgrid) return ldga_fail((ctx), LDGA_ERR_STATE, "ldga_set_kgrid has not been called")

static int h2d(ldga_ctx *ctx, const char *role, const void *src, size_t bytes, void **dst) {
  *dst = ldga_buf(ctx, role, bytes);
  if (!*dst) return LDGA_ERR_HIP;
  LDGA_HIP_TRY(ctx, hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  return LDGA_OK;
}

static int d2h(ldga_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (bytes) LDGA_HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  LDGA_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return LDGA_OK;
}

int ldga_bubble(ldga_ctx *ctx, const ldga_freq *fr, double beta, const double *G_fft, const double *G_rfft,
                double *chi0) {
  if (!ctx || !G_fft || !G_rfft || !chi0) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  NEED_KGRID(ctx);
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  const size_t gbytes = (size_t)ctx->oct.Nk * fr->n_fft * 16;
  const size_t obytes = (size_t)ctx->Nq * 2 * (fr->n_inu + fr->n_inu_shell) * (2 * fr->n_iw + 1) * 16;
  void *a, *b, *o;
  if ((rc = h2d(ctx, "G_fft", G_fft, gbytes, &a))) return rc;
  if ((rc = h2d(ctx, "G_rfft", G_rfft, gbytes, &b))) return rc;
  if (!(o = ldga_buf(ctx, "chi0", obytes))) return LDGA_ERR_HIP;
  if ((rc = launch_bubble(ctx, fr, beta, (const double2 *)a, (const double2 *)b, (double2 *)o, ctx->stream))) return rc;
  return d2h(ctx, chi0, o, obytes);
}

int ldga_bubble_dev(ldga_ctx *ctx, const ldga_freq *fr, double beta, const void *G_fft, const void *G_rfft,
                    void *chi0, void *stream) {
  if (!ctx || !G_fft || !G_rfft || !chi0) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  NEED_KGRID(ctx);
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  return launch_bubble(ctx, fr, beta, (const double2 *)G_fft, (const double2 *)G_rfft, (double2 *)chi0,
                       ldga_stream(ctx, stream));
}

int ldga_g_tables(ldga_ctx *ctx, const ldga_freq *fr, const double *G_wedge, double *G_fft, double *G_rfft) {
  if (!ctx || !G_wedge || !G_fft || !G_rfft) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  NEED_KGRID(ctx);
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  const size_t wbytes = (size_t)ctx->Nq * fr->n_fft * 16, gbytes = (size_t)ctx->oct.Nk * fr->n_fft * 16;
  void *w, *a, *b;
  if ((rc = h2d(ctx, "G_wedge", G_wedge, wbytes, &w))) return rc;
  if (!(a = ldga_buf(ctx, "G_fft", gbytes)) || !(b = ldga_buf(ctx, "G_rfft", gbytes))) return LDGA_ERR_HIP;
  if ((rc = launch_gtables(ctx, fr, (const double2 *)w, (double2 *)a, (double2 *)b, ctx->stream))) return rc;
  if ((rc = d2h(ctx, G_fft, a, gbytes))) return rc;
  return d2h(ctx, G_rfft, b, gbytes);
}

int ldga_g_tables_dev(ldga_ctx *ctx, const ldga_freq *fr, const void *G_wedge, void *G_fft, void *G_rfft,
                      void *stream) {
  if (!ctx || !G_wedge || !G_fft || !G_rfft) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  NEED_KGRID(ctx);
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  return launch_gtables(ctx, fr, (const double2 *)G_wedge, (double2 *)G_fft, (double2 *)G_rfft,
                        ldga_stream(ctx, stream));
}

int ldga_chi_gamma(ldga_ctx *ctx, int s, int precision, const ldga_freq *fr, double U, double beta,
                   const double *Gamma, const double *chi0, double *chi, double *gamma, ldga_bse_report *report) {
  if (!ctx || !Gamma || !chi0 || !chi || !gamma) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  NEED_KGRID(ctx);
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  const size_t Nnu = 2 * (size_t)fr->n_inu, Nw = 2 * (size_t)fr->n_iw + 1, Nq = ctx->Nq;
  const size_t gb = Nnu * Nnu * Nw * 16, cb = Nq * (Nnu + 2 * fr->n_inu_shell) * Nw * 16;
  void *g, *c, *xo, *yo;
  if ((rc = h2d(ctx, "Gamma", Gamma, gb, &g))) return rc;
  if ((rc = h2d(ctx, "chi0", chi0, cb, &c))) return rc;
  if (!(xo = ldga_buf(ctx, "chi", Nq * Nw * 8))) return LDGA_ERR_HIP;
  if (!(yo = ldga_buf(ctx, "gamma", Nq * Nnu * Nw * 16))) return LDGA_ERR_HIP;
  ldga_bse_report local;  // the host form always reads the counters (zero pivots -> status)
  int st = launch_bse(ctx, s, precision, fr, U, beta, (const double2 *)g, (const double2 *)c, (double *)xo,
                      (double2 *)yo, report ? report : &local, ctx->stream);
  if (st && st != LDGA_ERR_ZERO_PIVOT) return st;
  if ((rc = d2h(ctx, chi, xo, Nq * Nw * 8))) return rc;
  if ((rc = d2h(ctx, gamma, yo, Nq * Nnu * Nw * 16))) return rc;
  return st;
}

int ldga_chi_gamma_dev(ldga_ctx *ctx, int s, int precision, const ldga_freq *fr, double U, double beta,
                       const void *Gamma, const void *chi0, void *chi, void *gamma, ldga_bse_report *report,
                       void *stream) {
  if (!ctx || !Gamma || !chi0 || !chi || !gamma) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  NEED_KGRID(ctx);
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  return launch_bse(ctx, s, precision, fr, U, beta, (const double2 *)Gamma, (const double2 *)chi0, (double *)chi,
                    (double2 *)gamma, report, ldga_stream(ctx, stream));
}

int ldga_lambda0(ldga_ctx *ctx, const ldga_freq *fr, double beta, const double *F, const double *chi0,
                 double *lambda0) {
  if (!ctx || !F || !chi0 || !lambda0) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  NEED_KGRID(ctx);
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  const size_t Nnu = 2 * (size_t)fr->n_inu, Nw = 2 * (size_t)fr->n_iw + 1, Nq = ctx->Nq;
  const size_t fb = Nnu * Nnu * Nw * 16, cb = Nq * (Nnu + 2 * fr->n_inu_shell) * Nw * 16, ob = Nq * Nnu * Nw * 16;
  void *f, *c, *o;
  if ((rc = h2d(ctx, "F", F, fb, &f))) return rc;
  if ((rc = h2d(ctx, "chi0", chi0, cb, &c))) return rc;
  if (!(o = ldga_buf(ctx, "lambda0", ob))) return LDGA_ERR_HIP;
  if ((rc = launch_lambda0(ctx, fr, beta, (const double2 *)f, (const double2 *)c, (double2 *)o, ctx->stream)))
    return rc;
  return d2h(ctx, lambda0, o, ob);
}

int ldga_lambda0_dev(ldga_ctx *ctx, const ldga_freq *fr, double beta, const void *F, const void *chi0,
                     void *lambda0, void *stream) {
  if (!ctx || !F || !chi0 || !lambda0) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  NEED_KGRID(ctx);
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  return launch_lambda0(ctx, fr, beta, (const double2 *)F, (const double2 *)chi0, (double2 *)lambda0,
                        ldga_stream(ctx, stream));
}

int ldga_gen_chi(ldga_ctx *ctx, const ldga_freq *fr, const double *Gamma, const double *chi0, double *inv) {
  if (!ctx || !Gamma || !chi0 || !inv) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  NEED_KGRID(ctx);
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  const size_t Nnu = 2 * (size_t)fr->n_inu, Nw = 2 * (size_t)fr->n_iw + 1, Nq = ctx->Nq;
  const size_t gb = Nnu * Nnu * Nw * 16, cb = Nq * (Nnu + 2 * fr->n_inu_shell) * Nw * 16, ob = Nnu * Nnu * Nq * Nw * 16;
  void *g, *c, *o;
  if ((rc = h2d(ctx, "Gamma", Gamma, gb, &g))) return rc;
  if ((rc = h2d(ctx, "chi0", chi0, cb, &c))) return rc;
  // the inverses (Nnu^2 Nq Nw complex: 15.6 GB at 64^2, Nnu = 120) are staged
  // for this call only, not cached in the context
  LDGA_HIP_TRY(ctx, hipMalloc(&o, std::max<size_t>(ob, 1)));
  int st = launch_gen_chi(ctx, fr, (const double2 *)g, (const double2 *)c, (double2 *)o, ctx->stream);
  if (!st || st == LDGA_ERR_ZERO_PIVOT) {
    rc = d2h(ctx, inv, o, ob);
    if (rc) st = rc;
  }
  hipStreamSynchronize(ctx->stream);
  hipFree(o);
  return st;
}

int ldga_gen_chi_dev(ldga_ctx *ctx, const ldga_freq *fr, const void *Gamma, const void *chi0, void *inv, void *stream) {
  if (!ctx || !Gamma || !chi0 || !inv) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  NEED_KGRID(ctx);
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  return launch_gen_chi(ctx, fr, (const double2 *)Gamma, (const double2 *)chi0, (double2 *)inv,
                        ldga_stream(ctx, stream));
}

int ldga_F_from_chi(ldga_ctx *ctx, const ldga_freq *fr, double beta, int diag_term, const double *chi,
                    const double *G, double *F) {
  if (!ctx || !chi || !G || !F) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  const size_t Nnu = 2 * (size_t)fr->n_inu, Nw = 2 * (size_t)fr->n_iw + 1, bytes = Nnu * Nnu * Nw * 16;
  void *c, *g, *f;
  if ((rc = h2d(ctx, "fv_chi", chi, bytes, &c))) return rc;
  if ((rc = h2d(ctx, "fv_G", G, (size_t)fr->n_fft * 16, &g))) return rc;
  if (!(f = ldga_buf(ctx, "fv_F", bytes))) return LDGA_ERR_HIP;
  if ((rc = launch_F_from_chi(ctx, fr, beta, diag_term, (const double2 *)c, (const double2 *)g, (double2 *)f,
                              ctx->stream)))
    return rc;
  return d2h(ctx, F, f, bytes);
}

int ldga_F_from_chi_dev(ldga_ctx *ctx, const ldga_freq *fr, double beta, int diag_term, const void *chi,
                        const void *G, void *F, void *stream) {
  if (!ctx || !chi || !G || !F) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  return launch_F_from_chi(ctx, fr, beta, diag_term, (const double2 *)chi, (const double2 *)G, (double2 *)F,
                           ldga_stream(ctx, stream));
}

int ldga_sum_kw(ldga_ctx *ctx, const double *chi, int Nw, int w_lo, int w_hi, double c2, double beta, double lam,
                double out[3]) {
  if (!ctx || !chi || !out) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  NEED_KGRID(ctx);
  hipSetDevice(ctx->device);
  void *c, *o;
  int rc;
  if ((rc = h2d(ctx, "sum_chi", chi, (size_t)ctx->Nq * Nw * 8, &c))) return rc;
  if (!(o = ldga_buf(ctx, "sum_out", 3 * 8))) return LDGA_ERR_HIP;
  if ((rc = launch_sum_kw(ctx, (const double *)c, Nw, w_lo, w_hi, c2, beta, lam, (double *)o, ctx->stream)))
    return rc;
  return d2h(ctx, out, o, 3 * 8);
}

int ldga_sum_kw_dev(ldga_ctx *ctx, const void *chi, int Nw, int w_lo, int w_hi, double c2, double beta,
                    double lam, void *out_dev, void *stream) {
  if (!ctx || !chi || !out_dev) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  NEED_KGRID(ctx);
  hipSetDevice(ctx->device);
  return launch_sum_kw(ctx, (const double *)chi, Nw, w_lo, w_hi, c2, beta, lam, (double *)out_dev,
                       ldga_stream(ctx, stream));
}

int ldga_sum_w(ldga_ctx *ctx, const double *chi, int Nq, int Nw, int w_lo, int w_hi, double c2, double beta,
               double lam, double *out_q, double *out_k) {
  if (!ctx || !chi) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  if (Nq < 1 || Nw < 1) return ldga_fail(ctx, LDGA_ERR_SHAPE, "bad sum_ω shape Nq=%d Nw=%d", Nq, Nw);
  hipSetDevice(ctx->device);
  void *c, *r, *k = nullptr;
  int rc;
  if ((rc = h2d(ctx, "sum_chi", chi, (size_t)Nq * Nw * 8, &c))) return rc;
  if (!(r = ldga_buf(ctx, "sum_w_res", (size_t)Nq * 8))) return LDGA_ERR_HIP;
  if (out_k && !(k = ldga_buf(ctx, "sum_out", 3 * 8))) return LDGA_ERR_HIP;
  if ((rc = launch_sum_w(ctx, (const double *)c, Nq, Nw, w_lo, w_hi, c2, beta, lam, (double *)r, (double *)k,
                         ctx->stream)))
    return rc;
  if (out_q && (rc = d2h(ctx, out_q, r, (size_t)Nq * 8))) return rc;
  if (out_k && (rc = d2h(ctx, out_k, k, 8))) return rc;
  return d2h(ctx, nullptr, nullptr, 0);
}

int ldga_sum_w_dev(ldga_ctx *ctx, const void *chi, int Nq, int Nw, int w_lo, int w_hi, double c2, double beta,
                   double lam, void *out_q, void *out_k, void *stream) {
  if (!ctx || !chi) return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  if (Nq < 1) return ldga_fail(ctx, LDGA_ERR_SHAPE, "bad sum_ω shape Nq=%d", Nq);
  hipSetDevice(ctx->device);
  double *r = (double *)out_q;
  if (!r && !(r = (double *)ldga_buf(ctx, "sum_w_res", (size_t)Nq * 8))) return LDGA_ERR_HIP;
  return launch_sum_w(ctx, (const double *)chi, Nq, Nw, w_lo, w_hi, c2, beta, lam, r, (double *)out_k,
                      ldga_stream(ctx, stream));
}

int ldga_sigma(ldga_ctx *ctx, const ldga_freq *fr, const ldga_model *mp, int numax, double lam_m, double lam_d,
               int tc, int eom_kind, double chiloc_m_sum, const double *chi_m, const double *gamma_m,
               const double *chi_d, const double *gamma_d, const double *lambda0, const double *G_rfft,
               double *sigma) {
  if (!ctx || !mp || !chi_m || !gamma_m || !chi_d || !gamma_d || !lambda0 || !G_rfft || !sigma)
    return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  NEED_KGRID(ctx);
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  const size_t Nnu = 2 * (size_t)fr->n_inu, Nw = 2 * (size_t)fr->n_iw + 1, Nq = ctx->Nq;
  const size_t xb = Nq * Nw * 8, yb = Nq * Nnu * Nw * 16, gb = (size_t)ctx->oct.Nk * fr->n_fft * 16;
  void *cm, *gm, *cd, *gd, *l0, *gr, *so;
  if ((rc = h2d(ctx, "chi_m", chi_m, xb, &cm))) return rc;
  if ((rc = h2d(ctx, "chi_d", chi_d, xb, &cd))) return rc;
  if ((rc = h2d(ctx, "gamma_m", gamma_m, yb, &gm))) return rc;
  if ((rc = h2d(ctx, "gamma_d", gamma_d, yb, &gd))) return rc;
  if ((rc = h2d(ctx, "lambda0", lambda0, yb, &l0))) return rc;
  if ((rc = h2d(ctx, "G_rfft", G_rfft, gb, &gr))) return rc;
  if (!(so = ldga_buf(ctx, "sigma", Nq * (size_t)std::max(numax, 1) * 16))) return LDGA_ERR_HIP;
  if ((rc = launch_sigma(ctx, fr, mp, numax, lam_m, lam_d, tc, eom_kind, chiloc_m_sum, (const double *)cm,
                         (const double2 *)gm, (const double *)cd, (const double2 *)gd, (const double2 *)l0,
                         (const double2 *)gr, (double2 *)so, ctx->stream)))
    return rc;
  return d2h(ctx, sigma, so, Nq * (size_t)numax * 16);
}

int ldga_sigma_dev(ldga_ctx *ctx, const ldga_freq *fr, const ldga_model *mp, int numax, double lam_m, double lam_d,
                   int tc, int eom_kind, double chiloc_m_sum, const void *chi_m, const void *gamma_m,
                   const void *chi_d, const void *gamma_d, const void *lambda0, const void *G_rfft, void *sigma,
                   void *stream) {
  if (!ctx || !mp || !chi_m || !gamma_m || !chi_d || !gamma_d || !lambda0 || !G_rfft || !sigma)
    return ldga_fail(ctx, LDGA_ERR_ARG, "NULL argument");
  NEED_KGRID(ctx);
  int rc = ldga_check_freq(ctx, fr);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  return launch_sigma(ctx, fr, mp, numax, lam_m, lam_d, tc, eom_kind, chiloc_m_sum, (const double *)chi_m,
                      (const double2 *)gamma_m, (const double *)chi_d, (const double2 *)gamma_d,
                      (const double2 *)lambda0, (const double2 *)G_rfft, (double2 *)sigma, ldga_stream(ctx, stream));
}

}  // extern "C"
