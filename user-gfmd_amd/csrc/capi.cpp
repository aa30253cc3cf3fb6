// C ABI of libgfmd_hip.so (declarations and semantics: include/gfmd_hip.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/gfmd_hip.h"
#include "gfmd_internal.h"

using namespace gfmd;

namespace gfmd {
static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}
}  // namespace gfmd

struct gfmd_solver {
  int nx = 0, ny = 0, nky = 0, d = 0, precision = GFMD_FP64;
  bool fast = false;
  bool has_phi = false;
  hipStream_t stream = nullptr;
  void *spec = nullptr, *spec2 = nullptr, *xbuf = nullptr, *phi = nullptr;
  void *twx = nullptr, *twy = nullptr;
  double *linf = nullptr;
  bool linf_set = false;
  double *partials = nullptr;
  int nparts = 0;
  double *u0 = nullptr, *epot = nullptr;
  int *status = nullptr;
  double *u_io = nullptr, *f_io = nullptr;
  size_t bytes = 0;
  int ncu = 0;
  int row_var = 0, col_var = 0;
};

namespace {

// Autotuned fast-path variants, shared by all handles of one grid shape.
using TuneKey = std::tuple<int, int, int, int>;  // nx, ny, ndof, precision
std::mutex g_tune_mu;
std::map<TuneKey, std::pair<int, int>> g_tuned;

bool is_fp64(const gfmd_solver_t *s) { return s->precision == GFMD_FP64; }
size_t real_size(const gfmd_solver_t *s) { return is_fp64(s) ? 8 : 4; }

int dev_alloc(gfmd_solver_t *s, void **p, size_t n) {
  if (n == 0) n = 8;
  hipError_t e = hipMalloc(p, n);
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
    *p = nullptr;
    return ST_HIP;
  }
  s->bytes += n;
  return ST_OK;
}

// exp(-2 pi i k / n) for k in [0, n), in the handle's real type.
int upload_twiddles(gfmd_solver_t *s, int n, void **dst) {
  int rc = dev_alloc(s, dst, (size_t)n * 2 * real_size(s));
  if (rc) return rc;
  std::vector<double> w(2 * (size_t)n);
  for (int k = 0; k < n; k++) {
    const long double a = -2.0L * 3.141592653589793238462643383279502884L *
                          (long double)k / (long double)n;
    w[2 * k] = (double)cosl(a);
    w[2 * k + 1] = (double)sinl(a);
  }
  if (is_fp64(s)) {
    GFMD_HIP_TRY(hipMemcpy(*dst, w.data(), w.size() * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> wf(w.begin(), w.end());
    GFMD_HIP_TRY(hipMemcpy(*dst, wf.data(), wf.size() * 4, hipMemcpyHostToDevice));
  }
  return ST_OK;
}

void free_all(gfmd_solver_t *s) {
  void *ptrs[] = {s->spec, s->spec2, s->xbuf, s->phi,  s->twx,  s->twy,    s->linf,
                  s->partials, s->u0, s->epot, s->status, s->u_io, s->f_io};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (s->stream) (void)hipStreamDestroy(s->stream);
}

ForceArgs make_args(gfmd_solver_t *s, const void *u, void *f) {
  ForceArgs a;
  a.nx = s->nx;
  a.ny = s->ny;
  a.nky = s->nky;
  a.d = s->d;
  a.u = u;
  a.f = f;
  a.spec = s->spec;
  a.spec2 = s->spec2;
  a.xbuf = s->xbuf;
  a.phi = s->phi;
  a.twx = s->twx;
  a.twy = s->twy;
  a.linf = s->linf_set ? s->linf : nullptr;
  a.partials = s->partials;
  a.nparts = s->nparts;
  a.u0 = s->u0;
  a.epot = s->epot;
  a.ncu = s->ncu;
  a.row_var = s->row_var;
  a.col_var = s->col_var;
  return a;
}

int run_force(gfmd_solver_t *s, const void *u, void *f, double *epot_dev,
              bool io_double) {
  if (!s->has_phi) {
    set_error("no stiffness kernel set (call gfmd_set_kernel or gfmd_set_phi)");
    return ST_STATE;
  }
  ForceArgs a = make_args(s, u, f);
  if (epot_dev) a.epot = epot_dev;
  int rc = s->fast ? launch_force_fast(a, is_fp64(s), io_double, s->stream)
                   : launch_force_generic(a, is_fp64(s), io_double, s->stream);
  if (rc) {
    set_error("force launch failed (status %d)", rc);
    return rc;
  }
  GFMD_HIP_TRY(hipGetLastError());
  return ST_OK;
}

}  // namespace

extern "C" {

const char *gfmd_last_error(void) { return g_err.c_str(); }

const char *gfmd_version(void) { return "gfmd_hip 0.1 (gfx950)"; }

int gfmd_create(int nx, int ny, int ndof, int precision, gfmd_solver_t **out) {
  if (!out) {
    set_error("gfmd_create: out is NULL");
    return ST_ARG;
  }
  *out = nullptr;
  if (nx < 1 || ny < 1 || ndof < 3 || ndof % 3 != 0 || ndof > 24) {
    set_error("gfmd_create: invalid grid %d x %d, ndof %d", nx, ny, ndof);
    return ST_ARG;
  }
  if (precision != GFMD_FP64 && precision != GFMD_FP32) {
    set_error("gfmd_create: unknown precision %d", precision);
    return ST_ARG;
  }
  if ((size_t)nx * ny > (size_t)1 << 28) {
    set_error("gfmd_create: grid too large");
    return ST_ARG;
  }
  gfmd_solver_t *s = new gfmd_solver_t;
  s->nx = nx;
  s->ny = ny;
  s->nky = ny / 2 + 1;
  s->d = ndof;
  s->precision = precision;
  s->fast = force_fast_supported(nx, ny, ndof, precision == GFMD_FP64);
  int rc = ST_OK;
  auto fail = [&](int code) {
    free_all(s);
    delete s;
    return code;
  };
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
    return fail(ST_HIP);
  }
  int dev = 0;
  e = hipGetDevice(&dev);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&s->ncu, hipDeviceAttributeMultiprocessorCount,
                              dev);
  if (e != hipSuccess || s->ncu < 1) {
    set_error("cannot query the compute-unit count: %s", hipGetErrorString(e));
    return fail(ST_HIP);
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(TuneKey(nx, ny, ndof, precision));
    if (it != g_tuned.end()) {
      s->row_var = it->second.first;
      s->col_var = it->second.second;
    }
  }
  const size_t nspec = (size_t)ndof * s->nky * nx * 2 * real_size(s);
  const size_t nphi = (size_t)s->nky * nx * ndof * ndof * real_size(s);
  const size_t nfield = (size_t)ndof * nx * ny * sizeof(double);
  s->nparts = s->fast ? force_fast_nparts(nx, ny, ndof, is_fp64(s))
                      : force_generic_nparts(nx, ny, ndof);
  if ((rc = dev_alloc(s, &s->spec, nspec)) ||
      (!s->fast && (rc = dev_alloc(s, &s->spec2, nspec))) ||
      (s->fast && (rc = dev_alloc(s, &s->xbuf,
                                  (size_t)ndof * nx * 2 * real_size(s)))) ||
      (rc = dev_alloc(s, &s->phi, nphi)) ||
      (rc = dev_alloc(s, (void **)&s->linf, 8 * 8)) ||
      (rc = dev_alloc(s, (void **)&s->partials, (size_t)s->nparts * 8)) ||
      (rc = dev_alloc(s, (void **)&s->u0, 24 * 8)) ||
      (rc = dev_alloc(s, (void **)&s->epot, 8)) ||
      (rc = dev_alloc(s, (void **)&s->status, 8)) ||
      (rc = dev_alloc(s, (void **)&s->u_io, nfield)) ||
      (rc = dev_alloc(s, (void **)&s->f_io, nfield)) ||
      (rc = upload_twiddles(s, nx, &s->twx)) ||
      (rc = upload_twiddles(s, ny, &s->twy)))
    return fail(rc);
  e = hipMemset(s->linf, 0, 64);
  if (e == hipSuccess) e = hipMemset(s->u0, 0, 24 * 8);
  if (e != hipSuccess) {
    set_error("hipMemset failed: %s", hipGetErrorString(e));
    return fail(ST_HIP);
  }
  *out = s;
  return ST_OK;
}

int gfmd_destroy(gfmd_solver_t *s) {
  if (!s) return ST_OK;
  (void)hipStreamSynchronize(s->stream);
  free_all(s);
  delete s;
  return ST_OK;
}

int gfmd_set_kernel(gfmd_solver_t *s, const char *name, const char *args,
                    int normalize) {
  if (!s || !name) {
    set_error("gfmd_set_kernel: NULL argument");
    return ST_ARG;
  }
  KernelDesc k{};
  std::vector<std::string> tok;
  if (args) {
    std::string a(args);
    size_t i = 0;
    while (i < a.size()) {
      while (i < a.size() && isspace((unsigned char)a[i])) i++;
      size_t j = i;
      while (j < a.size() && !isspace((unsigned char)a[j])) j++;
      if (j > i) tok.push_back(a.substr(i, j - i));
      i = j;
    }
  }
  if (!strcmp(name, "sc100")) {
    // SC100StiffnessKernel::SC100StiffnessKernel
    // (sc100_stiffness.cpp:140-160): optional "height <int>".
    k.kind = KIND_SC100;
    k.d = 3;
    k.height = 0;
    size_t i = 0;
    if (i + 1 < tok.size() && tok[i] == "height") {
      char *end = nullptr;
      long h = strtol(tok[i + 1].c_str(), &end, 10);
      if (end == tok[i + 1].c_str() || *end) {
        set_error("SC100StiffnessKernel: Error converting height argument to "
                  "number.");
        return ST_ARG;
      }
      k.height = (int)h;
      i += 2;
    }
    if (i != tok.size()) {
      set_error("gfmd_set_kernel: unexpected argument '%s' for sc100",
                tok[i].c_str());
      return ST_ARG;
    }
  } else {
    set_error("gfmd_set_kernel: kernel '%s' has no device tabulation; "
              "tabulate it on the host and use gfmd_set_phi",
              name);
    return ST_UNSUPPORTED;
  }
  if (k.d != s->d) {
    set_error("gfmd_set_kernel: kernel '%s' has %d degrees of freedom, the "
              "grid was created with %d",
              name, k.d, s->d);
    return ST_ARG;
  }
  GFMD_HIP_TRY(hipMemsetAsync(s->status, 0, sizeof(int), s->stream));
  int rc = launch_phi_generate(k, s->nx, s->ny, normalize, is_fp64(s), s->phi,
                               s->status, s->stream);
  if (rc) {
    set_error("Phi tabulation failed (status %d)", rc);
    return rc;
  }
  GFMD_HIP_TRY(hipGetLastError());
  int st = 0;
  GFMD_HIP_TRY(hipMemcpyAsync(&st, s->status, sizeof(int),
                              hipMemcpyDeviceToHost, s->stream));
  GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  if (st) {
    set_error("Stiffness kernel '%s': singular matrix or out of iterations "
              "while evaluating the continued fraction.",
              name);
    return ST_KERNEL;
  }
  s->has_phi = true;
  return ST_OK;
}

int gfmd_set_phi(gfmd_solver_t *s, const double *phi, int normalized) {
  if (!s || !phi) {
    set_error("gfmd_set_phi: NULL argument");
    return ST_ARG;
  }
  const int nx = s->nx, ny = s->ny, d = s->d, dd = d * d;
  // Check Phi(-q) = conj(Phi(q)) on the host before relying on it.
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++) {
      const int jx = (nx - ix) % nx, jy = (ny - iy) % ny;
      const double *a = phi + ((size_t)ix * ny + iy) * dd * 2;
      const double *b = phi + ((size_t)jx * ny + jy) * dd * 2;
      double na = 0, diff = 0;
      for (int k = 0; k < dd; k++) {
        na += a[2 * k] * a[2 * k] + a[2 * k + 1] * a[2 * k + 1];
        const double dr = a[2 * k] - b[2 * k], di = a[2 * k + 1] + b[2 * k + 1];
        diff += dr * dr + di * di;
      }
      if (diff > 1e-12 * (na > 1e-300 ? na : 1e-300) && diff > 1e-24) {
        set_error("gfmd_set_phi: Phi(-q) != conj(Phi(q)) at grid point "
                  "(%d, %d)",
                  ix, iy);
        return ST_ARG;
      }
    }
  const size_t bytes = (size_t)nx * ny * dd * 2 * sizeof(double);
  double *tmp = nullptr;
  GFMD_HIP_TRY(hipMalloc((void **)&tmp, bytes));
  hipError_t e = hipMemcpyAsync(tmp, phi, bytes, hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) {
    const double scale = normalized ? 1.0 : 1.0 / ((double)nx * (double)ny);
    launch_phi_pack(tmp, nx, ny, d, scale, is_fp64(s), s->phi, s->stream);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) {
    set_error("gfmd_set_phi: %s", hipGetErrorString(e));
    return ST_HIP;
  }
  s->has_phi = true;
  return ST_OK;
}

int gfmd_set_linf(gfmd_solver_t *s, const double *linf) {
  if (!s || !linf) {
    set_error("gfmd_set_linf: NULL argument");
    return ST_ARG;
  }
  GFMD_HIP_TRY(hipMemcpyAsync(s->linf, linf, sizeof(double) * (s->d / 3),
                              hipMemcpyHostToDevice, s->stream));
  GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  s->linf_set = false;
  for (int i = 0; i < s->d / 3; i++)
    if (linf[i] != 0.0) s->linf_set = true;
  return ST_OK;
}

int gfmd_post_force(gfmd_solver_t *s, const double *u, double *f,
                    double *epot) {
  if (!s || !u || !f) {
    set_error("gfmd_post_force: NULL argument");
    return ST_ARG;
  }
  if ((const void *)u == (const void *)f) {
    set_error("gfmd_post_force: input and output buffers must differ");
    return ST_ARG;
  }
  const size_t bytes = (size_t)s->d * s->nx * s->ny * sizeof(double);
  GFMD_HIP_TRY(hipMemcpyAsync(s->u_io, u, bytes, hipMemcpyHostToDevice, s->stream));
  int rc = run_force(s, s->u_io, s->f_io, nullptr, true);
  if (rc) return rc;
  double e = 0;
  GFMD_HIP_TRY(hipMemcpyAsync(f, s->f_io, bytes, hipMemcpyDeviceToHost, s->stream));
  GFMD_HIP_TRY(hipMemcpyAsync(&e, s->epot, sizeof(double), hipMemcpyDeviceToHost,
                              s->stream));
  GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  if (epot) *epot = e;
  if (!std::isfinite(e)) {
    set_error("gfmd_post_force: non-finite elastic energy");
    return ST_NONFINITE;
  }
  return ST_OK;
}

int gfmd_post_force_device(gfmd_solver_t *s, const void *u_dev, void *f_dev,
                           double *epot_dev) {
  if (!s || !u_dev || !f_dev) {
    set_error("gfmd_post_force_device: NULL argument");
    return ST_ARG;
  }
  if (u_dev == f_dev) {
    set_error("gfmd_post_force_device: input and output buffers must differ");
    return ST_ARG;
  }
  return run_force(s, u_dev, f_dev, epot_dev, is_fp64(s));
}

int gfmd_get_u0(gfmd_solver_t *s, double *u0) {
  if (!s || !u0) {
    set_error("gfmd_get_u0: NULL argument");
    return ST_ARG;
  }
  GFMD_HIP_TRY(hipMemcpyAsync(u0, s->u0, sizeof(double) * s->d,
                              hipMemcpyDeviceToHost, s->stream));
  GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  return ST_OK;
}

int gfmd_time_force_device(gfmd_solver_t *s, const void *u_dev, void *f_dev,
                           double *epot_dev, int iters, double *ms) {
  if (!s || !u_dev || !f_dev || !ms || iters < 1) {
    set_error("gfmd_time_force_device: invalid argument");
    return ST_ARG;
  }
  if (!s->fast) {
    set_error("gfmd_time_force_device: only the fast FFT path is timed per "
              "sweep");
    return ST_UNSUPPORTED;
  }
  if (!s->has_phi) {
    set_error("no stiffness kernel set");
    return ST_STATE;
  }
  hipEvent_t ev[4];
  for (int i = 0; i < 4; i++) GFMD_HIP_TRY(hipEventCreate(&ev[i]));
  double acc[4] = {0, 0, 0, 0};
  int rc = ST_OK;
  ForceArgs a = make_args(s, u_dev, f_dev);
  if (epot_dev) a.epot = epot_dev;
  for (int it = 0; it < iters && rc == ST_OK; it++) {
    rc = launch_force_fast(a, is_fp64(s), is_fp64(s), s->stream, ev);
    if (rc) break;
    if (hipEventSynchronize(ev[3]) != hipSuccess) {
      rc = ST_HIP;
      break;
    }
    float t;
    for (int k = 0; k < 3; k++) {
      (void)hipEventElapsedTime(&t, ev[k], ev[k + 1]);
      acc[k] += t;
    }
    (void)hipEventElapsedTime(&t, ev[0], ev[3]);
    acc[3] += t;
  }
  for (int i = 0; i < 4; i++) (void)hipEventDestroy(ev[i]);
  if (rc) {
    set_error("gfmd_time_force_device: launch failed (%d)", rc);
    return rc;
  }
  for (int k = 0; k < 4; k++) ms[k] = acc[k] / iters;
  return ST_OK;
}

int gfmd_dump_stiffness(gfmd_solver_t *s, double *out) {
  if (!s || !out) {
    set_error("gfmd_dump_stiffness: NULL argument");
    return ST_ARG;
  }
  if (!s->has_phi) {
    set_error("no stiffness kernel set");
    return ST_STATE;
  }
  const size_t n = (size_t)s->nky * s->nx * s->d * s->d;
  if (is_fp64(s)) {
    GFMD_HIP_TRY(hipMemcpyAsync(out, s->phi, n * 8, hipMemcpyDeviceToHost,
                                s->stream));
    GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  } else {
    std::vector<float> tmp(n);
    GFMD_HIP_TRY(hipMemcpyAsync(tmp.data(), s->phi, n * 4,
                                hipMemcpyDeviceToHost, s->stream));
    GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
    for (size_t i = 0; i < n; i++) out[i] = tmp[i];
  }
  return ST_OK;
}

int gfmd_set_variant(gfmd_solver_t *s, int row_var, int col_var) {
  if (!s) {
    set_error("gfmd_set_variant: NULL handle");
    return ST_ARG;
  }
  if (!s->fast) {
    set_error("gfmd_set_variant: grid uses the generic path (no variants)");
    return ST_UNSUPPORTED;
  }
  if (!force_fast_variant_ok(s->nx, s->ny, s->d, is_fp64(s), row_var,
                             col_var)) {
    set_error("gfmd_set_variant: variant (%d, %d) not available for this grid",
              row_var, col_var);
    return ST_ARG;
  }
  s->row_var = row_var;
  s->col_var = col_var;
  return ST_OK;
}

int gfmd_autotune(gfmd_solver_t *s, int iters, int *row_var, int *col_var,
                  double *ms_best) {
  if (!s || iters < 1) {
    set_error("gfmd_autotune: invalid argument");
    return ST_ARG;
  }
  if (!s->fast || !s->has_phi) {
    if (row_var) *row_var = s->row_var;
    if (col_var) *col_var = s->col_var;
    if (ms_best) *ms_best = 0.0;
    if (!s->has_phi) {
      set_error("gfmd_autotune: no stiffness kernel set");
      return ST_STATE;
    }
    return ST_OK;  // generic path: nothing to tune
  }
  const size_t nfield = (size_t)s->d * s->nx * s->ny * sizeof(double);
  GFMD_HIP_TRY(hipMemsetAsync(s->u_io, 0, nfield, s->stream));
  hipEvent_t e0, e1;
  GFMD_HIP_TRY(hipEventCreate(&e0));
  GFMD_HIP_TRY(hipEventCreate(&e1));
  double best = 1e30;
  int br = s->row_var, bc = s->col_var;
  int rc = ST_OK;
  for (int rv = 0; rv < N_ROW_VARIANTS && rc == ST_OK; rv++)
    for (int cv = 0; cv < N_COL_VARIANTS && rc == ST_OK; cv++) {
      if (!force_fast_variant_ok(s->nx, s->ny, s->d, is_fp64(s), rv, cv))
        continue;
      ForceArgs a = make_args(s, s->u_io, s->f_io);
      a.row_var = rv;
      a.col_var = cv;
      // warm-up (first launch also queries the occupancy)
      rc = launch_force_fast(a, is_fp64(s), true, s->stream);
      if (rc) break;
      (void)hipEventRecord(e0, s->stream);
      for (int i = 0; i < iters && rc == ST_OK; i++)
        rc = launch_force_fast(a, is_fp64(s), true, s->stream);
      (void)hipEventRecord(e1, s->stream);
      if (hipEventSynchronize(e1) != hipSuccess) rc = ST_HIP;
      if (rc) break;
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      const double per = ms / iters;
      if (per < best) {
        best = per;
        br = rv;
        bc = cv;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {
    set_error("gfmd_autotune: launch failed (%d)", rc);
    return rc;
  }
  s->row_var = br;
  s->col_var = bc;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[TuneKey(s->nx, s->ny, s->d, s->precision)] = {br, bc};
  }
  if (row_var) *row_var = br;
  if (col_var) *col_var = bc;
  if (ms_best) *ms_best = best;
  return ST_OK;
}

void *gfmd_stream(gfmd_solver_t *s) { return s ? (void *)s->stream : nullptr; }

int gfmd_synchronize(gfmd_solver_t *s) {
  if (!s) {
    set_error("gfmd_synchronize: NULL handle");
    return ST_ARG;
  }
  GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  return ST_OK;
}

int gfmd_get_info(gfmd_solver_t *s, int *nx, int *ny, int *ndof,
                  int *precision, int *fast_path) {
  if (!s) {
    set_error("gfmd_get_info: NULL handle");
    return ST_ARG;
  }
  if (nx) *nx = s->nx;
  if (ny) *ny = s->ny;
  if (ndof) *ndof = s->d;
  if (precision) *precision = s->precision;
  if (fast_path) *fast_path = s->fast ? 1 : 0;
  return ST_OK;
}

size_t gfmd_memory_usage(gfmd_solver_t *s) { return s ? s->bytes : 0; }

}  // extern "C"
