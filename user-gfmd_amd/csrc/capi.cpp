// C ABI of libgfmd_hip.so (declarations and semantics: include/gfmd_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cfloat>
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
#include "contact_cg.h"
#include "ft_lattice.h"
#include "fix_glue.h"
#include "relax.h"
#include "wall_map.h"

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

// the same, for the library's other translation units
void gfmd_set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  gfmd::g_err = buf;
}

struct gfmd_solver {
  int nx = 0, ny = 0, nky = 0, d = 0, precision = GFMD_FP64;
  bool fast = false;
  bool has_phi = false;
  bool phi_mirror = false;  // Phi symmetric under x -> -x (checked)
  // ndof 3, fast path: Phi_xy real, Phi_xz and Phi_yz imaginary (checked);
  // the force sweeps then read the compact table phi_c (6 reals per q)
  bool phi_par = false;
  void *phi_c = nullptr;
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
  // relaxation state (allocated on first use): T[d][nx][ny] each
  void *ru = nullptr, *rv = nullptr, *rf = nullptr, *rfel = nullptr;
  void *rwall = nullptr;
  double *rstate = nullptr, *rpart = nullptr, *repot = nullptr;
  int rnblk = 0;
  // contact solver (allocated on first use): compliance tables gz
  // T[nky][nx] and gfull T[nky][d*d][nx], fields T[6][nx*ny]
  void *gz = nullptr, *gfull = nullptr, *pk = nullptr;
  double *pkstate = nullptr, *pkpart = nullptr;
  bool gz_valid = false;
  // one captured contact-solver iteration, replayed; rebuilt when the
  // sweep variants change
  hipGraphExec_t pk_exec = nullptr;
  int pk_rv = -1, pk_cv = -1;
  int *gate = nullptr;  // device flag of batched runs (1: iterate)
  // prec_gradient: the table -(Phi~ + cavg)^-1 T[nky][d*d][nx] for the
  // cavg in prec_c (device) / prec_key (host), valid until Phi changes
  void *prec = nullptr;
  double *prec_c = nullptr;
  std::vector<double> prec_key;
  bool prec_valid = false;
  bool prec_mirror = false;
  // fix gfmd glue: state [FS_COUNT], reduction scratch, frame
  double *fix_st = nullptr, *fix_part = nullptr;
  FixFrame fix_frame{};
  bool fix_ready = false;
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
  void *ptrs[] = {s->phi_c, s->spec,  s->spec2,  s->xbuf,   s->phi,    s->twx,
                  s->twy,   s->linf,   s->partials, s->u0,   s->epot,
                  s->status, s->u_io,  s->f_io,   s->ru,     s->rv,
                  s->rf,    s->rfel,   s->rwall,  s->rstate, s->rpart,
                  s->repot, s->gz,     s->gfull,  s->pk,     s->pkstate,
                  s->pkpart, s->gate,  s->prec,   s->prec_c, s->fix_st,
                  s->fix_part};
  if (s->pk_exec) (void)hipGraphExecDestroy(s->pk_exec);
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
  a.mirror = s->phi_mirror ? 1 : 0;
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

// The force sweeps of the handle's own table read the compact ndof-3 table
// when it exists (fast path only).  Callers that pass another table
// (compliance, preconditioner) keep make_args' full-table defaults.
void use_compact_table(gfmd_solver_t *s, ForceArgs &a) {
  if (s->fast && s->phi_par && s->phi_c && a.d == 3) {
    a.phi = s->phi_c;
    a.par = 1;
  }
}

int run_force(gfmd_solver_t *s, const void *u, void *f, double *epot_dev,
              bool io_double) {
  if (!s->has_phi) {
    set_error("no stiffness kernel set (call gfmd_set_kernel or gfmd_set_phi)");
    return ST_STATE;
  }
  ForceArgs a = make_args(s, u, f);
  use_compact_table(s, a);
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

// ---------------------------------------------------------------------------
// Scratch of the allocation-free force evaluation (gfmd_post_force_noalloc):
// the spectrum, the column-0 exchange, the energy partials, u0 and the
// twiddle tables of one evaluation.  Every device the library runs on holds
// its own copy for the life of the process.  Host code never uses the
// array's name as an address: the device's copy is found once per device
// with hipGetSymbolAddress and kept in g_scratch_ptr.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char
    scratch[GFMD_NOALLOC_SCRATCH_BYTES];

namespace {

// Timing events, destroyed on every return path of the entry point that
// created them.
template <int N>
struct TimingEvents {
  hipEvent_t e[N] = {};
  hipError_t create() {
    for (int i = 0; i < N; i++) {
      const hipError_t r = hipEventCreate(&e[i]);
      if (r != hipSuccess) return r;
    }
    return hipSuccess;
  }
  ~TimingEvents() {
    for (int i = 0; i < N; i++)
      if (e[i]) (void)hipEventDestroy(e[i]);
  }
};

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];
int g_scratch_ncu[kMaxDevices];

// The current device's scratch address and compute-unit count, looked up
// on the first call on that device (with it current).
int noalloc_scratch(void **ptr, int *ncu) {
  int dev = 0;
  GFMD_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("gfmd_post_force_noalloc: device %d beyond the %d supported",
              dev, kMaxDevices);
    return ST_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("gfmd_post_force_noalloc: hipGetSymbolAddress(scratch) "
                "failed on device %d: %s", dev, hipGetErrorString(e));
      return ST_HIP;
    }
    int n = 0;
    GFMD_HIP_TRY(hipDeviceGetAttribute(
        &n, hipDeviceAttributeMultiprocessorCount, dev));
    g_scratch_ncu[dev] = n;
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  *ncu = g_scratch_ncu[dev];
  return ST_OK;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

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

// After every write of the Phi table (ndof 3, fast path): has it the
// structure of the compact table (Phi_xy real, Phi_xz and Phi_yz imaginary,
// exactly)?  Then the compact copy the force sweeps read is written.
static int check_compact(gfmd_solver_t *s) {
  s->phi_par = false;
  if (s->d != 3 || !s->fast) return ST_OK;
  GFMD_HIP_TRY(hipMemsetAsync(s->status, 0, sizeof(int), s->stream));
  launch_phi_par_check(s->phi, s->nx, s->ny, is_fp64(s), s->status,
                       s->stream);
  GFMD_HIP_TRY(hipGetLastError());
  int bad = 1;
  GFMD_HIP_TRY(hipMemcpyAsync(&bad, s->status, sizeof(int),
                              hipMemcpyDeviceToHost, s->stream));
  GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  if (bad) return ST_OK;
  if (!s->phi_c) {
    const int rc = dev_alloc(s, &s->phi_c,
                             (size_t)s->nky * s->nx * 6 * real_size(s));
    if (rc) return rc;
  }
  launch_phi_compact(s->phi, s->phi_c, s->nx, s->ny, is_fp64(s), s->stream);
  GFMD_HIP_TRY(hipGetLastError());
  s->phi_par = true;
  return ST_OK;
}

// After every write of the Phi table: is it symmetric under x -> -x (to
// rounding)?  Then the column sweep reads half of it.  A device-tabulated
// table passes within 1e-12 (FP64) / 1e-6 (FP32) of the trace (its two
// halves come from the same formulas at +qx and -qx); a table the host
// uploaded is mirrored only when its halves agree to a few ulp of the
// trace, so the read never changes a caller's entries by more than the
// rounding of the arithmetic itself.
static int check_mirror(gfmd_solver_t *s, bool host_table = false,
                        bool compact = true) {
  s->phi_mirror = false;
  const double tol = host_table ? (is_fp64(s) ? 4 * DBL_EPSILON
                                              : 4 * (double)FLT_EPSILON)
                                : (is_fp64(s) ? 1e-12 : 1e-6);
  GFMD_HIP_TRY(hipMemsetAsync(s->status, 0, sizeof(int), s->stream));
  launch_phi_mirror_check(s->phi, s->nx, s->ny, s->d, is_fp64(s), tol,
                          s->status, s->stream);
  GFMD_HIP_TRY(hipGetLastError());
  int bad = 1;
  GFMD_HIP_TRY(hipMemcpyAsync(&bad, s->status, sizeof(int),
                              hipMemcpyDeviceToHost, s->stream));
  GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  s->phi_mirror = (bad == 0) && s->nx % 2 == 0;
  if (!compact) {
    s->phi_par = false;
    return ST_OK;
  }
  return check_compact(s);
}

int gfmd_set_kernel(gfmd_solver_t *s, const char *name, const char *args,
                    int normalize) {
  if (!s || !name) {
    set_error("gfmd_set_kernel: NULL argument");
    return ST_ARG;
  }
  KernelDesc k{};
  FTTable ft;
  double linf[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // per-layer Gamma-point force
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
  } else if (!strcmp(name, "isotropic")) {
    // IsotropicStiffnessKernel::IsotropicStiffnessKernel
    // (isotropic_stiffness.cpp:44-86): Poisson number, shear modulus,
    // gamma.  The q = 0 entry uses the domain lengths, which the kernel
    // reads from LAMMPS; here they default to the grid size (lattice
    // constant 1) and are set with "xprd L" / "yprd L" (keywords that the
    // reference's syntax does not have: the fix passes domain->xprd/yprd).
    k.kind = KIND_ISOTROPIC;
    k.d = 3;
    k.height = 0;
    const char *what[3] = {"nu", "myu", "gamma"};
    size_t i = 0;
    char *end = nullptr;
    for (int n = 0; n < 3; n++, i++) {
      if (i >= tok.size() ||
          (k.p[n] = strtod(tok[i].c_str(), &end), end == tok[i].c_str())) {
        set_error("IsotropicStiffnessKernel::IsotropicStiffnessKernel: "
                  "Can't convert %s to float.", what[n]);
        return ST_ARG;
      }
    }
    double xprd = s->nx, yprd = s->ny;
    while (i + 1 < tok.size() && (tok[i] == "xprd" || tok[i] == "yprd")) {
      const double v = strtod(tok[i + 1].c_str(), &end);
      if (end == tok[i + 1].c_str() || !(v > 0.0)) {
        set_error("isotropic: invalid %s '%s'", tok[i].c_str(),
                  tok[i + 1].c_str());
        return ST_ARG;
      }
      (tok[i] == "xprd" ? xprd : yprd) = v;
      i += 2;
    }
    if (i != tok.size()) {
      set_error("gfmd_set_kernel: unexpected argument '%s' for isotropic",
                tok[i].c_str());
      return ST_ARG;
    }
    if (yprd != xprd) {
      set_error("isotropic kernel work with square surfaces only.");
      return ST_ARG;
    }
    k.p[3] = xprd;
  } else if (!strcmp(name, "nonperiodic")) {
    // NonperiodicStiffnessKernel::NonperiodicStiffnessKernel
    // (nonperiodic_stiffness.cpp:44-103): Poisson number, shear modulus.
    // The grid spacing comes from the domain lengths (LAMMPS); here they
    // default to the grid size (lattice constant 1) and are set with
    // "xprd L" / "yprd L" (the fix passes domain->xprd/yprd).
    k.kind = KIND_NONPERIODIC;
    k.d = 3;
    k.height = 1;
    char *end = nullptr;
    const char *what[2] = {"Poisson number", "shear modulus"};
    size_t i = 0;
    for (int n = 0; n < 2; n++, i++) {
      if (i >= tok.size()) {
        set_error("Missing %s.", what[n]);
        return ST_ARG;
      }
      k.p[n] = strtod(tok[i].c_str(), &end);
      if (end == tok[i].c_str()) {
        set_error("Can't convert %s '%s' to floating point.", what[n],
                  tok[i].c_str());
        return ST_ARG;
      }
    }
    double xprd = s->nx, yprd = s->ny;
    while (i + 1 < tok.size() && (tok[i] == "xprd" || tok[i] == "yprd")) {
      const double v = strtod(tok[i + 1].c_str(), &end);
      if (end == tok[i + 1].c_str() || !(v > 0.0)) {
        set_error("nonperiodic: invalid %s '%s'", tok[i].c_str(),
                  tok[i + 1].c_str());
        return ST_ARG;
      }
      (tok[i] == "xprd" ? xprd : yprd) = v;
      i += 2;
    }
    if (i != tok.size()) {
      set_error("gfmd_set_kernel: unexpected argument '%s' for nonperiodic",
                tok[i].c_str());
      return ST_ARG;
    }
    if (s->nx > 4096 || s->ny > 4096) {
      set_error("nonperiodic: grids up to 4096 x 4096");
      return ST_UNSUPPORTED;
    }
    k.p[2] = xprd / s->nx;
    k.p[3] = yprd / s->ny;
  } else if (!strcmp(name, "ft")) {
    // FTStiffnessKernel (ft_stiffness.cpp:47-100): <surface> <lattice
    // constant> <nu> <force constants ...> [height H]; tables on the host,
    // Phi(q) on the device
    std::string err = ft_build(tok, ft);
    if (!err.empty()) {
      set_error("ft: %s", err.c_str());
      return ST_ARG;
    }
    if (ft.nat > 4) {
      set_error("ft: %d atoms per surface cell; the device tabulation "
                "takes up to 4 (ndof 12)", ft.nat);
      return ST_UNSUPPORTED;
    }
    k.kind = KIND_FT;
    k.d = 3 * ft.nat;
    k.height = ft.height;
    k.ft = &ft;
    for (int n = 0; n < ft.nat; n++) linf[n] = ft.linf[n];
  } else if (!strcmp(name, "fcc100")) {
    // FCC100StiffnessKernel::FCC100StiffnessKernel
    // (fcc100_stiffness.cpp:59-242): [x] nu k1 [kappa1] (k2 [kappa2] k3
    // [kappa3] for nu >= 2), then the keywords height, linf (nu values
    // summing to zero), normal (+-1) and mulphase (parsed, unused).
    k.kind = KIND_FCC100;
    k.height = 0;
    k.p[6] = 1.0;
    size_t i = 0;
    if (i >= tok.size()) {
      set_error("FCC100StiffnessKernel::FCC100StiffnessKernel: Expected "
                "number of interacting neighbors.");
      return ST_ARG;
    }
    const bool with_kappa = tok[i].back() == 'x';
    if (with_kappa) i++;
    if (i >= tok.size()) {
      set_error("FCC100StiffnessKernel::FCC100StiffnessKernel: Error "
                "converting number of layers to integer.");
      return ST_ARG;
    }
    char *end = nullptr;
    const long nu = strtol(tok[i].c_str(), &end, 10);
    if (end == tok[i].c_str()) {
      set_error("FCC100StiffnessKernel::FCC100StiffnessKernel: Error "
                "converting number of layers to integer.");
      return ST_ARG;
    }
    i++;
    if (nu < 1) {
      set_error("FCC100StiffnessKernel::FCC100StiffnessKernel: Number of "
                "bounday GF layers must be at least 1.");
      return ST_ARG;
    }
    if (nu > 4) {
      set_error("fcc100: at most 4 layers per cell (the reference's "
                "per-layer matrices cover nu <= 4)");
      return ST_UNSUPPORTED;
    }
    const int nk = nu == 1 ? 1 : 3;
    for (int n = 0; n < nk; n++)
      for (int w = 0; w < (with_kappa ? 2 : 1); w++) {
        if (i >= tok.size()) {
          set_error("FCC100StiffnessKernel::FCC100StiffnessKernel: Spring "
                    "constant expected.");
          return ST_ARG;
        }
        const double v = strtod(tok[i].c_str(), &end);
        if (end == tok[i].c_str()) {
          set_error("FCC100StiffnessKernel::FCC100StiffnessKernel: Error "
                    "converting spring constant to number.");
          return ST_ARG;
        }
        k.p[w ? 3 + n : n] = v;
        i++;
      }
    bool found = true;
    while (i + 1 < tok.size() && found) {
      found = false;
      if (tok[i] == "height") {
        const long h = strtol(tok[i + 1].c_str(), &end, 10);
        if (end == tok[i + 1].c_str()) {
          set_error("FCC100StiffnessKernel::FCC100StiffnessKernel: Error "
                    "converting height argument to number.");
          return ST_ARG;
        }
        k.height = (int)h;
        i += 2;
        found = true;
      } else if (tok[i] == "linf") {
        double sum = 0.0;
        for (int n = 0; n < nu; n++) {
          i++;
          if (i >= tok.size()) {
            set_error("fix gfmd: Stiffness kernel report %ld layers, please "
                      "provide that many *linf* parameters.", nu);
            return ST_ARG;
          }
          linf[n] = atof(tok[i].c_str());
          sum += linf[n];
        }
        if (fabs(sum) > 1e-12) {
          set_error("fix gfmd: Linear force contributions must sum to zero "
                    "but sums to %e.", sum);
          return ST_ARG;
        }
        i++;
        found = true;
      } else if (tok[i] == "normal") {
        const long z = strtol(tok[i + 1].c_str(), &end, 10);
        if (end == tok[i + 1].c_str()) {
          set_error("fix gfmd: Could not convert surface normal (%s) to "
                    "integer.", tok[i + 1].c_str());
          return ST_ARG;
        }
        if (z != 1 && z != -1) {
          set_error("fix gfmd: Surface normal must be either +1 or -1, but "
                    "is %ld.", z);
          return ST_ARG;
        }
        k.p[6] = (double)z;
        i += 2;
        found = true;
      } else if (tok[i] == "mulphase") {
        (void)strtod(tok[i + 1].c_str(), &end);
        if (end == tok[i + 1].c_str()) {
          set_error("fix gfmd: Could not convert mulphase (%s) to double.",
                    tok[i + 1].c_str());
          return ST_ARG;
        }
        i += 2;
        found = true;
      }
    }
    if (i != tok.size()) {
      set_error("gfmd_set_kernel: unexpected argument '%s' for fcc100",
                tok[i].c_str());
      return ST_ARG;
    }
    k.d = 3 * (int)nu;
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
  // the kernel's Gamma-point forces (StiffnessKernel::
  // get_force_at_gamma_point): zero unless given with "linf"
  GFMD_HIP_TRY(hipMemcpyAsync(s->linf, linf, sizeof(double) * (s->d / 3),
                              hipMemcpyHostToDevice, s->stream));
  GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  s->linf_set = false;
  for (int n = 0; n < s->d / 3; n++)
    if (linf[n] != 0.0) s->linf_set = true;
  s->has_phi = true;
  s->gz_valid = false;
  s->prec_valid = false;
  // the open-boundary table (direct transforms, rounding in every entry) has
  // no exact zeros to compact
  return check_mirror(s, false, k.kind != KIND_NONPERIODIC);
}

int gfmd_set_phi(gfmd_solver_t *s, const double *phi, int normalized) {
  if (!s || !phi) {
    set_error("gfmd_set_phi: NULL argument");
    return ST_ARG;
  }
  const int nx = s->nx, ny = s->ny, d = s->d, dd = d * d;
  // Each Phi(q) must be Hermitian, as fill_phi_buffer checks
  // (gfmd_misc.cpp:53-73 with _is_Hermitian, linearalgebra.cpp:81-100):
  // |Phi_ij - conj Phi_ji| <= 1e-6 |tr Phi| (non-finite entries fail).
  // Phi(-q) = conj Phi(q) is not required: the packing applies what the
  // reference's real-part inverse transform applies.
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++) {
      const double *a = phi + ((size_t)ix * ny + iy) * dd * 2;
      double tr_re = 0, tr_im = 0;
      for (int i = 0; i < d; i++) {
        tr_re += a[2 * (i * d + i)];
        tr_im += a[2 * (i * d + i) + 1];
      }
      const double tr = std::hypot(tr_re, tr_im);
      for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) {
          const double dr = a[2 * (i * d + j)] - a[2 * (j * d + i)];
          const double di = a[2 * (i * d + j) + 1] + a[2 * (j * d + i) + 1];
          if (!(std::hypot(dr, di) <= 1e-6 * tr)) {
            const double qx = ix <= nx / 2 ? (double)ix / nx : (double)(ix - nx) / nx;
            const double qy = iy <= ny / 2 ? (double)iy / ny : (double)(iy - ny) / ny;
            set_error("gfmd_set_phi: Kernel is not Hermitian at q = 2 pi "
                      "(%f, %f)!", qx, qy);
            return ST_ARG;
          }
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
  s->gz_valid = false;
  s->prec_valid = false;
  return check_mirror(s, true);
}

int gfmd_ft_cell(const char *args, int *nat, double *cell, double *positions,
                 double *linf, int max_atoms) {
  if (!args || !nat) {
    set_error("gfmd_ft_cell: NULL argument");
    return ST_ARG;
  }
  std::vector<std::string> tok;
  std::string a(args);
  for (size_t i = 0; i < a.size();) {
    while (i < a.size() && isspace((unsigned char)a[i])) i++;
    size_t j = i;
    while (j < a.size() && !isspace((unsigned char)a[j])) j++;
    if (j > i) tok.push_back(a.substr(i, j - i));
    i = j;
  }
  FTTable t;
  std::string err = ft_build(tok, t);
  if (!err.empty()) {
    set_error("ft: %s", err.c_str());
    return ST_ARG;
  }
  *nat = t.nat;
  if (t.nat > max_atoms) return ST_OK;
  if (cell) std::memcpy(cell, t.cell, sizeof t.cell);
  if (positions) std::memcpy(positions, t.pos.data(), t.pos.size() * 8);
  if (linf) std::memcpy(linf, t.linf.data(), t.linf.size() * 8);
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

int gfmd_post_force_noalloc(int nx, int ny, int ndof, int precision,
                            const void *phi_dev, int phi_mirror,
                            const double *linf_dev, const void *u_dev,
                            void *f_dev, double *epot_dev, double *u0_dev,
                            void *stream) {
  if (!phi_dev || !u_dev || !f_dev || !epot_dev) {
    set_error("gfmd_post_force_noalloc: NULL argument");
    return ST_ARG;
  }
  if (u_dev == f_dev) {
    set_error("gfmd_post_force_noalloc: input and output buffers must differ");
    return ST_ARG;
  }
  if (precision != GFMD_FP64 && precision != GFMD_FP32) {
    set_error("gfmd_post_force_noalloc: unknown precision %d", precision);
    return ST_ARG;
  }
  const bool fp64 = precision == GFMD_FP64;
  if (nx < 1 || ny < 1 || ndof < 3 || ndof % 3 != 0 ||
      !force_fast_supported(nx, ny, ndof, fp64)) {
    set_error("gfmd_post_force_noalloc: grid %d x %d, ndof %d not supported "
              "(power-of-two sides 64..4096, ndof 3 or 6)", nx, ny, ndof);
    return ST_UNSUPPORTED;
  }
  const size_t w = fp64 ? 8 : 4;
  const int nparts = force_fast_nparts(nx, ny, ndof, fp64);
  const size_t o_spec = 0;
  const size_t o_xbuf = o_spec + align256((size_t)ndof * (ny / 2) * nx * 2 * w);
  const size_t o_part = o_xbuf + align256((size_t)ndof * nx * 2 * w);
  const size_t o_u0 = o_part + align256((size_t)nparts * 8);
  const size_t o_twx = o_u0 + align256((size_t)ndof * 8);
  const size_t o_twy = o_twx + align256((size_t)nx * 2 * w);
  const size_t total = o_twy + align256((size_t)ny * 2 * w);
  if (total > GFMD_NOALLOC_SCRATCH_BYTES) {
    set_error("gfmd_post_force_noalloc: a %d x %d grid with ndof %d in %s "
              "needs %zu bytes of scratch, more than the %zu the library "
              "holds per device", nx, ny, ndof, fp64 ? "FP64" : "FP32", total,
              (size_t)GFMD_NOALLOC_SCRATCH_BYTES);
    return ST_ARG;
  }
  void *base = nullptr;
  int ncu = 0;
  int rc = noalloc_scratch(&base, &ncu);
  if (rc) return rc;
  char *sc = (char *)base;
  hipStream_t st = (hipStream_t)stream;
  launch_twiddles(sc + o_twx, nx, fp64, st);
  launch_twiddles(sc + o_twy, ny, fp64, st);
  ForceArgs a;
  a.nx = nx;
  a.ny = ny;
  a.nky = ny / 2 + 1;
  a.d = ndof;
  a.u = u_dev;
  a.f = f_dev;
  a.spec = sc + o_spec;
  a.spec2 = nullptr;
  a.xbuf = sc + o_xbuf;
  a.phi = phi_dev;
  a.mirror = phi_mirror ? 1 : 0;
  a.twx = sc + o_twx;
  a.twy = sc + o_twy;
  a.linf = linf_dev;
  a.partials = (double *)(sc + o_part);
  a.nparts = nparts;
  a.u0 = u0_dev ? u0_dev : (double *)(sc + o_u0);
  a.epot = epot_dev;
  a.ncu = ncu;
  a.row_var = 0;
  a.col_var = 0;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(TuneKey(nx, ny, ndof, precision));
    if (it != g_tuned.end()) {
      a.row_var = it->second.first;
      a.col_var = it->second.second;
    }
  }
  rc = launch_force_fast(a, fp64, fp64, st);
  if (rc) {
    set_error("gfmd_post_force_noalloc: force launch failed (status %d)", rc);
    return rc;
  }
  GFMD_HIP_TRY(hipGetLastError());
  return ST_OK;
}

int gfmd_time_force_graph(gfmd_solver_t *s, const void *u_dev, void *f_dev,
                          double *epot_dev, int iters, int per_graph,
                          double *ms_eager, double *ms_graph) {
  if (!s || !u_dev || !f_dev || iters < 1 || per_graph < 1 ||
      iters % per_graph || !ms_eager || !ms_graph || u_dev == f_dev) {
    set_error("gfmd_time_force_graph: invalid argument");
    return ST_ARG;
  }
  if (!s->fast) {
    set_error("gfmd_time_force_graph: fast FFT path only");
    return ST_UNSUPPORTED;
  }
  if (!s->has_phi) {
    set_error("no stiffness kernel set");
    return ST_STATE;
  }
  TimingEvents<2> tev;
  GFMD_HIP_TRY(tev.create());
  hipEvent_t e0 = tev.e[0], e1 = tev.e[1];
  int rc = run_force(s, u_dev, f_dev, epot_dev, is_fp64(s));  // warm-up
  float t = 0.f;
  // eager: `iters` evaluations launched back to back
  if (rc == ST_OK && hipEventRecord(e0, s->stream) != hipSuccess) rc = ST_HIP;
  for (int i = 0; i < iters && rc == ST_OK; i++)
    rc = run_force(s, u_dev, f_dev, epot_dev, is_fp64(s));
  if (rc == ST_OK && (hipEventRecord(e1, s->stream) != hipSuccess ||
                      hipEventSynchronize(e1) != hipSuccess))
    rc = ST_HIP;
  if (rc == ST_OK) {
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms_eager = (double)t / iters;
  }
  // graph: per_graph evaluations captured once (the sweeps' kernels only;
  // the fast path records no memset), replayed iters / per_graph times
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (rc == ST_OK) {
    hipError_t e = hipStreamBeginCapture(s->stream,
                                         hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
      for (int i = 0; i < per_graph && rc == ST_OK; i++)
        rc = run_force(s, u_dev, f_dev, epot_dev, is_fp64(s));
      const hipError_t e2 = hipStreamEndCapture(s->stream, &graph);
      if (rc == ST_OK && e2 != hipSuccess) rc = ST_HIP;
    } else {
      rc = ST_HIP;
    }
    if (rc == ST_OK &&
        hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
      rc = ST_HIP;
  }
  if (rc == ST_OK && hipGraphLaunch(exec, s->stream) != hipSuccess)  // warm-up
    rc = ST_HIP;
  if (rc == ST_OK && hipEventRecord(e0, s->stream) != hipSuccess) rc = ST_HIP;
  for (int i = 0; i < iters / per_graph && rc == ST_OK; i++)
    if (hipGraphLaunch(exec, s->stream) != hipSuccess) rc = ST_HIP;
  if (rc == ST_OK && (hipEventRecord(e1, s->stream) != hipSuccess ||
                      hipEventSynchronize(e1) != hipSuccess))
    rc = ST_HIP;
  if (rc == ST_OK) {
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms_graph = (double)t / iters;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (rc) {
    set_error("gfmd_time_force_graph: launch, capture or replay failed (%d)",
              rc);
    (void)hipGetLastError();
  }
  return rc;
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
  TimingEvents<5> tev;
  GFMD_HIP_TRY(tev.create());
  hipEvent_t *ev = tev.e;
  double acc[4] = {0, 0, 0, 0};
  int rc = ST_OK;
  ForceArgs a = make_args(s, u_dev, f_dev);
  use_compact_table(s, a);
  if (epot_dev) a.epot = epot_dev;
  for (int it = 0; it < iters && rc == ST_OK; it++) {
    rc = launch_force_fast(a, is_fp64(s), is_fp64(s), s->stream, ev);
    if (rc) break;
    if (hipEventSynchronize(ev[4]) != hipSuccess) {
      rc = ST_HIP;
      break;
    }
    float t;
    (void)hipEventElapsedTime(&t, ev[0], ev[1]);
    acc[0] += t;
    (void)hipEventElapsedTime(&t, ev[1], ev[2]);
    acc[1] += t;
    (void)hipEventElapsedTime(&t, ev[3], ev[4]);
    acc[2] += t;
    (void)hipEventElapsedTime(&t, ev[0], ev[4]);
    acc[3] += t;
  }
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

int gfmd_stiffness_at(gfmd_solver_t *s, int count, const int *kx,
                      const int *ky, double *out) {
  if (!s || count < 0 || (count > 0 && (!kx || !ky || !out))) {
    set_error("gfmd_stiffness_at: invalid argument");
    return ST_ARG;
  }
  if (!s->has_phi) {
    set_error("no stiffness kernel set");
    return ST_STATE;
  }
  const int d = s->d, dd = d * d;
  const size_t es = real_size(s);
  std::vector<char> raw((size_t)count * dd * es);
  for (int q = 0; q < count; q++) {
    if (kx[q] < 0 || kx[q] >= s->nx || ky[q] < 0 || ky[q] > s->ny / 2) {
      set_error("gfmd_stiffness_at: wavevector (%d, %d) outside the stored "
                "half spectrum (0 <= kx < %d, 0 <= ky <= %d)", kx[q], ky[q],
                s->nx, s->ny / 2);
      return ST_ARG;
    }
    // entries (ky*dd + e)*nx + kx, e = 0..dd-1: a column of pitch nx
    const char *src = (const char *)s->phi +
                      ((size_t)ky[q] * dd * s->nx + kx[q]) * es;
    GFMD_HIP_TRY(hipMemcpy2DAsync(raw.data() + (size_t)q * dd * es, es, src,
                                  (size_t)s->nx * es, es, dd,
                                  hipMemcpyDeviceToHost, s->stream));
  }
  GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  for (int q = 0; q < count; q++) {
    double pk[24 * 24];  // ndof <= 24
    for (int e = 0; e < dd; e++)
      pk[e] = es == 8 ? ((const double *)raw.data())[(size_t)q * dd + e]
                      : ((const float *)raw.data())[(size_t)q * dd + e];
    double *m = out + (size_t)q * dd * 2;
    for (int i = 0; i < d; i++) {
      m[2 * (i * d + i)] = pk[i];
      m[2 * (i * d + i) + 1] = 0.0;
    }
    int p = d;
    for (int i = 0; i < d; i++)
      for (int j = i + 1; j < d; j++) {
        const double re = pk[p], im = pk[p + 1];
        p += 2;
        m[2 * (i * d + j)] = re;
        m[2 * (i * d + j) + 1] = im;
        m[2 * (j * d + i)] = re;
        m[2 * (j * d + i) + 1] = -im;
      }
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
  TimingEvents<2> tev;
  GFMD_HIP_TRY(tev.create());
  hipEvent_t e0 = tev.e[0], e1 = tev.e[1];
  double best = 1e30;
  int br = s->row_var, bc = s->col_var;
  int rc = ST_OK;
  // three passes over the combinations, each timed at its fastest pass:
  // the first combinations of one pass run while the clocks still ramp up
  double per_combo[N_ROW_VARIANTS][N_COL_VARIANTS];
  bool okv[N_ROW_VARIANTS][N_COL_VARIANTS];
  for (auto &r : per_combo)
    for (double &x : r) x = 1e30;
  for (int pass = 0; pass < 3 && rc == ST_OK; pass++)
    for (int rv = 0; rv < N_ROW_VARIANTS && rc == ST_OK; rv++)
      for (int cv = 0; cv < N_COL_VARIANTS && rc == ST_OK; cv++) {
        okv[rv][cv] =
            force_fast_variant_ok(s->nx, s->ny, s->d, is_fp64(s), rv, cv);
        if (!okv[rv][cv]) continue;
        ForceArgs a = make_args(s, s->u_io, s->f_io);
        use_compact_table(s, a);
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
        double &per = per_combo[rv][cv];
        per = std::min(per, (double)ms / iters);
      }
  if (rc) {
    set_error("gfmd_autotune: launch failed (%d)", rc);
    return rc;
  }
  // The row variant sets the two row sweeps and the column variant the
  // column sweep, so their costs add: with every combination available,
  // each is chosen by its mean over the other's variants (a row choice
  // rests on 4 timings, a column choice on 3, instead of one difference of
  // a few microseconds in a whole evaluation).  Otherwise the fastest
  // combination.
  bool full = true;
  for (int rv = 0; rv < N_ROW_VARIANTS; rv++)
    for (int cv = 0; cv < N_COL_VARIANTS; cv++) full = full && okv[rv][cv];
  if (full) {
    double brm = 1e30, bcm = 1e30;
    for (int rv = 0; rv < N_ROW_VARIANTS; rv++) {
      double m = 0;
      for (int cv = 0; cv < N_COL_VARIANTS; cv++) m += per_combo[rv][cv];
      if (m < brm) brm = m, br = rv;
    }
    for (int cv = 0; cv < N_COL_VARIANTS; cv++) {
      double m = 0;
      for (int rv = 0; rv < N_ROW_VARIANTS; rv++) m += per_combo[rv][cv];
      if (m < bcm) bcm = m, bc = cv;
    }
    best = per_combo[br][bc];
  } else {
    for (int rv = 0; rv < N_ROW_VARIANTS; rv++)
      for (int cv = 0; cv < N_COL_VARIANTS; cv++)
        if (okv[rv][cv] && per_combo[rv][cv] < best) {
          best = per_combo[rv][cv];
          br = rv;
          bc = cv;
        }
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

void gfmd_fire_defaults(gfmd_fire_params_t *p) {
  if (!p) return;
  p->dt = 0.001;
  p->dtmax = 0.001;
  p->incfac = 1.2;
  p->decfac = 0.5;
  p->mix_in = 0.1;
  p->mixdec = 0.99;
  p->xlimit = 0.0;
  p->mass = 1.0;
  p->minsteps = 10;
  p->adapt_dt = 0;
}

static int relax_alloc(gfmd_solver_t *s) {
  if (s->ru) return ST_OK;
  const size_t nf = (size_t)s->d * s->nx * s->ny * real_size(s);
  s->rnblk = relax_nblocks(s->nx, s->ny, s->ncu);
  int rc;
  if ((rc = dev_alloc(s, &s->ru, nf)) || (rc = dev_alloc(s, &s->rv, nf)) ||
      (rc = dev_alloc(s, &s->rf, nf)) || (rc = dev_alloc(s, &s->rfel, nf)) ||
      (rc = dev_alloc(s, &s->rwall,
                      (size_t)s->nx * s->ny * real_size(s))) ||
      (rc = dev_alloc(s, (void **)&s->rstate, S_COUNT * 8)) ||
      (rc = dev_alloc(s, (void **)&s->rpart, (size_t)16 * s->rnblk * 8)) ||
      (rc = dev_alloc(s, (void **)&s->repot, 8)))
    return rc;
  return ST_OK;
}

static int relax_params(gfmd_solver_t *s, const gfmd_contact_t *c,
                        const gfmd_fire_params_t *fp, RelaxDev &p) {
  // p.wall points at the handle's wall buffer: allocate it first, so that
  // no kernel is ever handed an unallocated wall
  const int arc = relax_alloc(s);
  if (arc) return arc;
  p = RelaxDev{};
  p.nx = s->nx;
  p.ny = s->ny;
  p.d = s->d;
  p.dt0 = fp->dt;
  p.dtmax = fp->dtmax;
  p.mix_in = fp->mix_in;
  p.mixdec = fp->mixdec;
  p.incfac = fp->incfac;
  p.decfac = fp->decfac;
  p.xlimit = fp->xlimit;
  p.mass = fp->mass;
  p.minsteps = fp->minsteps;
  p.adapt_dt = fp->adapt_dt;
  if (!(fp->dt > 0) || !(fp->mass > 0) || !(fp->dtmax > 0)) {
    set_error("gfmd_relax: dt, dtmax and mass must be positive");
    return ST_ARG;
  }
  p.contact = c ? c->type : CONTACT_NONE;
  p.ax = p.ay = 1.0;
  p.ncat = 1;
  for (int l = 0; l < 4; l++) p.aoff[l][0] = p.aoff[l][1] = p.aoff[l][2] = 0.0;
  if (c) {
    p.pressure = c->pressure;
    const double a = c->lattice[0] > 0 ? c->lattice[0] : 1.0;
    p.ax = c->cell_xy[0] > 0 ? c->cell_xy[0] : a;
    p.ay = c->cell_xy[1] > 0 ? c->cell_xy[1] : a;
    p.x0 = c->lattice[1];
    p.y0 = c->lattice[2];
    p.z0 = c->lattice[3];
    p.wall_offset = c->wall_offset;
    if (c->contact_atoms < 0 || c->contact_atoms > 4 ||
        c->contact_atoms > s->d / 3) {
      set_error("gfmd_relax: contact_atoms %d (at most 4 and ndof/3)",
                c->contact_atoms);
      return ST_ARG;
    }
    if (c->contact_atoms > 1 && c->type != CONTACT_LJ_SPHERE &&
        c->type != CONTACT_WALL_MAP) {
      set_error("gfmd_relax: several contact atoms per cell need the LJ "
                "sphere or a wall map (a hard wall acts on atom 0)");
      return ST_ARG;
    }
    if (c->contact_atoms > 0) {
      p.ncat = c->contact_atoms;
      for (int l = 0; l < p.ncat; l++)
        for (int k = 0; k < 3; k++) p.aoff[l][k] = c->atom_offset[l][k];
    }
  }
  p.lx = p.ax * s->nx;
  p.ly = p.ay * s->ny;
  if (p.contact == CONTACT_LJ_SPHERE) {
    // FixContactSphere constructor (fix_contact_sphere.cpp:124-140)
    const double R = c->sphere[3], eps = c->sphere[4], sig = c->sphere[5];
    const double cut = c->sphere[6];
    if (!(R > 0) || !(cut > 0)) {
      set_error("gfmd_relax: LJ sphere needs radius > 0 and cutoff > 0");
      return ST_ARG;
    }
    for (int k = 0; k < 3; k++) p.sph_c[k] = c->sphere[k];
    p.sph_R = R;
    p.sph_rpc2 = (R + cut) * (R + cut);
    p.lj1 = 48.0 * eps * pow(sig, 12.0);
    p.lj2 = 24.0 * eps * pow(sig, 6.0);
    p.lj3 = 4.0 * eps * pow(sig, 12.0);
    p.lj4 = 4.0 * eps * pow(sig, 6.0);
    const double r2inv = 1.0 / (cut * cut);
    const double r6inv = r2inv * r2inv * r2inv;
    p.lj_off = r6inv * (p.lj3 * r6inv - p.lj4);
  } else if (p.contact == CONTACT_HARDWALL) {
    if (!c->wall) {
      set_error("gfmd_relax: hard wall needs a wall map");
      return ST_ARG;
    }
  } else if (p.contact == CONTACT_WALL_MAP) {
    const gfmd_wall_map_t *w = c->wall_map;
    if (!w) {
      set_error("gfmd_relax: wall/map contact needs a gfmd_wall_map_t");
      return ST_ARG;
    }
    if (std::fabs(w->xprd - p.lx) > 1e-9 * p.lx ||
        std::fabs(w->yprd - p.ly) > 1e-9 * p.ly) {
      set_error("gfmd_relax: the wall map spans %g x %g, the surface %g x %g",
                w->xprd, w->yprd, p.lx, p.ly);
      return ST_ARG;
    }
    p.wm = w->dev;
  } else if (p.contact != CONTACT_NONE) {
    set_error("gfmd_relax: unknown contact type %d", p.contact);
    return ST_ARG;
  }
  p.wall = s->rwall;
  if (p.contact == CONTACT_HARDWALL && !p.wall) {
    set_error("relaxation: wall buffer missing");
    return ST_STATE;
  }
  return ST_OK;
}

// FIRE from the current device state of `s` (u in s->ru): setup force,
// then steps until |f| <= tol * |f_ext| (|f| <= tol without external force;
// |f| <= tol * |f(setup)| when rel_initial) or max_steps.  st receives the
// final state scalars.
static int fire_loop(gfmd_solver_t *s, const RelaxDev &p, int max_steps,
                     double tol, int check_every, int *steps_out,
                     double *st, bool *conv, bool rel_initial = false,
                     bool keep_state = false, double abs_floor = 0.0) {
  const bool f64 = is_fp64(s);
  RelaxBufs b{s->ru, s->rv, s->rf, s->rfel, s->rstate, s->rpart, s->repot,
              s->rnblk};
  int rc;
  // keep_state continues with the velocities and the FIRE scalars of the
  // previous call (momentum survives a change of the external field)
  if (!keep_state) launch_relax_step(p, b, f64, STEP_INIT, s->stream);
  launch_relax_step(p, b, f64, STEP_CONTACT_RESET, s->stream);
  if ((rc = run_force(s, s->ru, s->rfel, s->repot, f64))) return rc;
  launch_relax_step(p, b, f64, STEP_FORCE, s->stream);
  GFMD_HIP_TRY(hipGetLastError());
  auto fetch = [&]() -> int {
    GFMD_HIP_TRY(hipMemcpyAsync(st, s->rstate, sizeof(double) * S_COUNT,
                                hipMemcpyDeviceToHost, s->stream));
    GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
    return ST_OK;
  };
  double target = -1.0;  // absolute residual target (rel_initial)
  auto check_state = [&](bool &done) -> int {
    if (st[S_INSIDE] > 0) {
      if (p.contact == CONTACT_WALL_MAP)
        set_error("fix contact/map: Atom too close: Implement short ranged "
                  "spline for lj126 potential. (%g atoms)", st[S_INSIDE]);
      else
        set_error("fix contact/sphere: atom inside sphere (%g atoms)",
                  st[S_INSIDE]);
      return ST_CONTACT;
    }
    if (!std::isfinite(st[S_RES]) || !std::isfinite(st[S_ETOT])) {
      set_error("gfmd_relax: non-finite force or energy");
      return ST_NONFINITE;
    }
    if (rel_initial) {
      // the relative drop tol, or the absolute residual abs_floor that
      // already meets the caller's goal, whichever comes first
      if (target < 0) target = std::max(tol * st[S_RES], abs_floor);
      done = st[S_RES] <= target;
    } else {
      const double denom = st[S_FEXT] > 0 ? st[S_FEXT] : 1.0;
      done = st[S_RES] / denom <= tol;
    }
    return ST_OK;
  };
  if ((rc = fetch())) return rc;
  bool done = false;
  if ((rc = check_state(done))) return rc;
  if (rel_initial) done = false;  // always take at least one check period
  int steps = 0;
  while (!done && steps < max_steps) {
    launch_relax_step(p, b, f64, STEP_MIX_MOVE, s->stream);
    if ((rc = run_force(s, s->ru, s->rfel, s->repot, f64))) return rc;
    launch_relax_step(p, b, f64, STEP_FORCE_KICK, s->stream);
    steps++;
    if (steps % check_every == 0 || steps == max_steps) {
      GFMD_HIP_TRY(hipGetLastError());
      if ((rc = fetch())) return rc;
      if ((rc = check_state(done))) {
        *steps_out = steps;
        return rc;
      }
    }
  }
  *steps_out = steps;
  *conv = done;
  return ST_OK;
}

using Clock = std::chrono::steady_clock;

static void fill_result(gfmd_relax_result_t *res, const double *st,
                        int steps, bool conv, Clock::time_point t0,
                        double device_seconds = 0.0) {
  if (!res) return;
  res->device_seconds = device_seconds;
  res->steps = steps;
  res->converged = conv ? 1 : 0;
  res->residual = st[S_RES];
  res->fext_norm = st[S_FEXT];
  res->energy = st[S_ETOT];
  res->contact_energy = st[S_ECON];
  res->ncontact = st[S_NCON];
  res->seconds = std::chrono::duration<double>(
                     std::chrono::steady_clock::now() - t0)
                     .count();
}

int gfmd_relax(gfmd_solver_t *s, const gfmd_contact_t *contact,
               const gfmd_fire_params_t *fire, double *u, int max_steps,
               double tol, int check_every, gfmd_relax_result_t *res) {
  if (!s || !fire || !u || max_steps < 0 || check_every < 1) {
    set_error("gfmd_relax: invalid argument");
    return ST_ARG;
  }
  if (!s->has_phi) {
    set_error("gfmd_relax: no stiffness kernel set");
    return ST_STATE;
  }
  int rc = relax_alloc(s);
  if (rc) return rc;
  RelaxDev p;
  if ((rc = relax_params(s, contact, fire, p))) return rc;
  const bool f64 = is_fp64(s);
  const size_t n = (size_t)s->d * s->nx * s->ny;
  const auto t0 = std::chrono::steady_clock::now();
  // upload u (and the wall map) through the FP64 staging buffers
  GFMD_HIP_TRY(hipMemcpyAsync(s->u_io, u, n * 8, hipMemcpyHostToDevice,
                              s->stream));
  launch_convert(s->u_io, true, s->ru, f64, n, s->stream);
  if (p.contact == CONTACT_HARDWALL) {
    const size_t nw = (size_t)s->nx * s->ny;
    GFMD_HIP_TRY(hipMemcpyAsync(s->f_io, contact->wall, nw * 8,
                                hipMemcpyHostToDevice, s->stream));
    launch_convert(s->f_io, true, s->rwall, f64, nw, s->stream);
  }
  double st[S_COUNT];
  int steps = 0;
  bool conv = false;
  GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  const auto td = Clock::now();
  rc = fire_loop(s, p, max_steps, tol, check_every, &steps, st, &conv);
  const double dev_s =
      std::chrono::duration<double>(Clock::now() - td).count();
  if (rc) {
    fill_result(res, st, steps, false, t0, dev_s);
    return rc;
  }
  launch_convert(s->ru, f64, s->u_io, true, n, s->stream);
  GFMD_HIP_TRY(hipMemcpyAsync(u, s->u_io, n * 8, hipMemcpyDeviceToHost,
                              s->stream));
  GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  fill_result(res, st, steps, conv, t0, dev_s);
  return ST_OK;
}

int gfmd_dynamics(gfmd_solver_t *s, const gfmd_contact_t *contact,
                  const gfmd_dyn_params_t *prm, double *u, double *v,
                  int nsteps, double *history, gfmd_dyn_result_t *res) {
  if (!s || !prm || !u || !v || nsteps < 0 || prm->record_every < 1 ||
      !(prm->dt > 0) || !(prm->mass > 0) || !(prm->damping >= 0)) {
    set_error("gfmd_dynamics: invalid argument");
    return ST_ARG;
  }
  if (!s->has_phi) {
    set_error("gfmd_dynamics: no stiffness kernel set");
    return ST_STATE;
  }
  if (contact && contact->type == GFMD_CONTACT_HARDWALL) {
    set_error("gfmd_dynamics: a hard wall is not smooth; use the LJ sphere");
    return ST_UNSUPPORTED;
  }
  int rc = relax_alloc(s);
  if (rc) return rc;
  gfmd_fire_params_t fp;
  gfmd_fire_defaults(&fp);
  fp.dt = fp.dtmax = prm->dt;
  fp.mass = prm->mass;
  RelaxDev p;
  if ((rc = relax_params(s, contact, &fp, p))) return rc;
  DynDev q{prm->dt, prm->damping, prm->record_every};
  const bool f64 = is_fp64(s);
  const size_t n = (size_t)s->d * s->nx * s->ny;
  const auto t0 = std::chrono::steady_clock::now();
  const size_t nrec = (size_t)nsteps / prm->record_every + 1;
  double *hist = nullptr;
  if (history) GFMD_HIP_TRY(hipMalloc((void **)&hist, nrec * 3 * 8));
  auto cleanup = [&]() {
    if (hist) (void)hipFree(hist);
  };
  GFMD_HIP_TRY(hipMemcpyAsync(s->u_io, u, n * 8, hipMemcpyHostToDevice,
                              s->stream));
  launch_convert(s->u_io, true, s->ru, f64, n, s->stream);
  GFMD_HIP_TRY(hipMemcpyAsync(s->u_io, v, n * 8, hipMemcpyHostToDevice,
                              s->stream));
  launch_convert(s->u_io, true, s->rv, f64, n, s->stream);
  RelaxBufs b{s->ru, s->rv, s->rf, s->rfel, s->rstate, s->rpart, s->repot,
              s->rnblk};
  launch_dyn_init(s->rstate, s->stream);
  launch_relax_step(p, b, f64, STEP_CONTACT_RESET, s->stream);
  // state 0: force, first half kick and drift (eager; also primes the
  // force path's one-time launch set-up before the capture)
  if ((rc = run_force(s, s->ru, s->rfel, s->repot, f64))) {
    cleanup();
    return rc;
  }
  launch_dyn_step(p, q, b, f64, nsteps == 0, hist, s->stream);
  if (nsteps > 1) {
    // states 1 .. nsteps-1: one captured graph per step, replayed
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
      rc = run_force(s, s->ru, s->rfel, s->repot, f64);
      launch_dyn_step(p, q, b, f64, false, hist, s->stream);
      e = hipStreamEndCapture(s->stream, &graph);
    }
    if (e == hipSuccess && rc == ST_OK)
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    for (int it = 1; it < nsteps && e == hipSuccess && rc == ST_OK; it++)
      e = hipGraphLaunch(exec, s->stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (rc == ST_OK && e != hipSuccess) {
      set_error("gfmd_dynamics: graph capture/launch failed: %s",
                hipGetErrorString(e));
      rc = ST_HIP;
    }
    if (rc) {
      (void)hipStreamSynchronize(s->stream);
      cleanup();
      return rc;
    }
  }
  if (nsteps > 0) {
    if ((rc = run_force(s, s->ru, s->rfel, s->repot, f64))) {
      cleanup();
      return rc;
    }
    launch_dyn_step(p, q, b, f64, true, hist, s->stream);
  }
  double st[DS_COUNT];
  hipError_t e = hipMemcpyAsync(st, s->rstate, sizeof(st),
                                hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) {
    launch_convert(s->ru, f64, s->u_io, true, n, s->stream);
    e = hipMemcpyAsync(u, s->u_io, n * 8, hipMemcpyDeviceToHost, s->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) {
    launch_convert(s->rv, f64, s->u_io, true, n, s->stream);
    e = hipMemcpyAsync(v, s->u_io, n * 8, hipMemcpyDeviceToHost, s->stream);
  }
  if (e == hipSuccess && history)
    e = hipMemcpyAsync(history, hist, nrec * 3 * 8, hipMemcpyDeviceToHost,
                       s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  cleanup();
  if (e != hipSuccess) {
    set_error("gfmd_dynamics: %s", hipGetErrorString(e));
    return ST_HIP;
  }
  if (res) {
    res->steps = nsteps;
    res->ekin = st[DS_KE];
    res->epot = st[DS_PE];
    res->ncontact = (int)st[DS_NCON];
    res->seconds = std::chrono::duration<double>(
                       std::chrono::steady_clock::now() - t0)
                       .count();
  }
  if (st[DS_INSIDE] > 0) {
    if (p.contact == CONTACT_WALL_MAP)
      set_error("fix contact/map: Atom too close: Implement short ranged "
                "spline for lj126 potential. (dynamics)");
    else
      set_error("fix contact/sphere: atom inside sphere (dynamics)");
    return ST_CONTACT;
  }
  if (!std::isfinite(st[DS_KE]) || !std::isfinite(st[DS_PE])) {
    set_error("gfmd_dynamics: energy not finite (time step too large?)");
    return ST_NONFINITE;
  }
  return ST_OK;
}

// ---------------------------------------------------------------------------
// Hard-wall contact by constrained conjugate gradients (contact_cg.hip).
// ---------------------------------------------------------------------------

// dst = G * src with a compliance table (d = 1: the scalar G_zz, d = s->d:
// the full G), on the force sweeps in the handle's compute type.
static int run_conv(gfmd_solver_t *s, const void *table, int d,
                    const void *src, void *dst, const int *gate = nullptr) {
  ForceArgs a = make_args(s, src, dst);
  a.gate = gate;
  a.d = d;
  a.phi = table;
  a.linf = nullptr;
  const bool f64 = is_fp64(s);
  if (!force_fast_variant_ok(s->nx, s->ny, d, f64, a.row_var, a.col_var))
    a.row_var = a.col_var = 0;
  int rc = force_fast_variant_ok(s->nx, s->ny, d, f64, a.row_var, a.col_var)
               ? launch_force_fast(a, f64, f64, s->stream)
               : launch_force_generic(a, f64, f64, s->stream);
  if (rc) set_error("compliance convolution failed (status %d)", rc);
  return rc;
}

static PKBufs pk_bufs(gfmd_solver_t *s) {
  const size_t n = (size_t)s->nx * s->ny;
  char *b = (char *)s->pk;
  const size_t w = n * real_size(s);
  return PKBufs{b,          b + w,     b + 2 * w, b + 3 * w,
                b + 4 * w,  b + 5 * w, s->pkstate, s->pkpart,
                s->rnblk,   n,         s->gate};
}

// Buffers and compliance tables; G_zz(q = 0) of the surface z component
// into *g0 (the uniform compliance that turns the load into an approach).
static int pk_prepare(gfmd_solver_t *s, double *g0) {
  int rc = relax_alloc(s);
  if (rc) return rc;
  const size_t n = (size_t)s->nx * s->ny;
  const size_t nh = (size_t)s->nky * s->nx;
  if (!s->pk &&
      ((rc = dev_alloc(s, &s->gz, nh * real_size(s))) ||
       (rc = dev_alloc(s, &s->gfull, nh * s->d * s->d * real_size(s))) ||
       (rc = dev_alloc(s, &s->pk, 6 * n * real_size(s))) ||
       (rc = dev_alloc(s, (void **)&s->pkstate, PK_COUNT * 8)) ||
       (rc = dev_alloc(s, (void **)&s->pkpart, (size_t)8 * s->rnblk * 8)) ||
       (rc = dev_alloc(s, (void **)&s->gate, sizeof(int)))))
    return rc;
  if (!s->gz_valid) {
    GFMD_HIP_TRY(hipMemsetAsync(s->status, 0, sizeof(int), s->stream));
    rc = launch_compliance(s->phi, s->nx, s->ny, s->d, is_fp64(s), 2, s->gz,
                           s->gfull, s->status, s->stream);
    if (rc) {
      set_error("contact solver: no compliance for ndof %d", s->d);
      return rc;
    }
    int st = 0;
    GFMD_HIP_TRY(hipMemcpyAsync(&st, s->status, sizeof(int),
                                hipMemcpyDeviceToHost, s->stream));
    GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
    if (st) {
      set_error("contact solver: Phi(q) is singular (a finite-height "
                "kernel is needed: the load sets the q = 0 displacement)");
      return ST_KERNEL;
    }
    s->gz_valid = true;
  }
  double gz0 = 0.0;
  if (is_fp64(s)) {
    GFMD_HIP_TRY(hipMemcpy(&gz0, s->gz, 8, hipMemcpyDeviceToHost));
  } else {
    float g = 0.f;
    GFMD_HIP_TRY(hipMemcpy(&g, s->gz, 4, hipMemcpyDeviceToHost));
    gz0 = g;
  }
  // gz = -G/N  ->  G_zz(0) = -N gz[0]
  *g0 = -gz0 * (double)s->nx * (double)s->ny;
  return ST_OK;
}

// One iteration on the handle's stream: u = G*p, gap and direction,
// r = G*t, step and projection.
static int pk_iteration(gfmd_solver_t *s) {
  const bool f64 = is_fp64(s);
  PKBufs b = pk_bufs(s);
  int rc = run_conv(s, s->gz, 1, b.p, b.u, s->gate);
  if (!rc) launch_pk(b, f64, PK_GAP, s->stream);
  if (!rc) rc = run_conv(s, s->gz, 1, b.t, b.r, s->gate);
  if (!rc) launch_pk(b, f64, PK_CURV_UPDATE, s->stream);
  return rc;
}

// `iters` iterations as replays of one captured graph.
static int pk_run(gfmd_solver_t *s, int iters) {
  if (iters <= 0) return ST_OK;
  int rc = ST_OK;
  hipError_t e = hipSuccess;
  if (!s->pk_exec || s->pk_rv != s->row_var || s->pk_cv != s->col_var) {
    if (s->pk_exec) (void)hipGraphExecDestroy(s->pk_exec);
    s->pk_exec = nullptr;
    hipGraph_t graph = nullptr;
    e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
      rc = pk_iteration(s);
      e = hipStreamEndCapture(s->stream, &graph);
    }
    if (e == hipSuccess && rc == ST_OK)
      e = hipGraphInstantiate(&s->pk_exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    s->pk_rv = s->row_var;
    s->pk_cv = s->col_var;
  }
  for (int it = 0; it < iters && e == hipSuccess && rc == ST_OK; it++)
    e = hipGraphLaunch(s->pk_exec, s->stream);
  if (rc == ST_OK && e != hipSuccess) {
    set_error("contact solver: graph failed: %s", hipGetErrorString(e));
    rc = ST_HIP;
  }
  return rc;
}

// Rebuild the displacement from the pressures (u = G * traction, pressure
// sites on the wall) and evaluate the projected-force residual as FIRE
// does; st receives the relaxation scalars.
static int pk_state3d(gfmd_solver_t *s, const RelaxDev &p, double *st) {
  const bool f64 = is_fp64(s);
  PKBufs b = pk_bufs(s);
  launch_pk_traction(b, f64, s->d, 2, p.pressure, s->rf, s->stream);
  int rc = run_conv(s, s->gfull, s->d, s->rf, s->ru);
  if (rc) return rc;
  launch_pk_snap(b, f64, 2, s->rwall, p.wall_offset, s->ru, s->stream);
  RelaxBufs rb{s->ru, s->rv, s->rf, s->rfel, s->rstate, s->rpart, s->repot,
               s->rnblk};
  launch_relax_step(p, rb, f64, STEP_INIT, s->stream);
  if ((rc = run_force(s, s->ru, s->rfel, s->repot, f64))) return rc;
  launch_relax_step(p, rb, f64, STEP_FORCE, s->stream);
  GFMD_HIP_TRY(hipMemcpyAsync(st, s->rstate, sizeof(double) * S_COUNT,
                              hipMemcpyDeviceToHost, s->stream));
  GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  return ST_OK;
}

int gfmd_contact_solve(gfmd_solver_t *s64, gfmd_solver_t *s32,
                       const gfmd_contact_t *contact, double *u,
                       const double *p_init, int max_iter, double tol,
                       gfmd_relax_result_t *res, int *iters) {
  if (!s64 || !contact || !u || max_iter < 1 || !(tol > 0)) {
    set_error("gfmd_contact_solve: invalid argument");
    return ST_ARG;
  }
  if (contact->type != GFMD_CONTACT_HARDWALL || !contact->wall) {
    set_error("gfmd_contact_solve: needs a hard wall with a wall map");
    return ST_ARG;
  }
  if (!is_fp64(s64) ||
      (s32 && (is_fp64(s32) || s32->nx != s64->nx || s32->ny != s64->ny ||
               s32->d != s64->d))) {
    set_error("gfmd_contact_solve: needs an FP64 handle (and optionally an "
              "FP32 handle of the same grid)");
    return ST_ARG;
  }
  if (!s64->has_phi || (s32 && !s32->has_phi)) {
    set_error("gfmd_contact_solve: no stiffness kernel set");
    return ST_STATE;
  }
  const auto t0 = std::chrono::steady_clock::now();
  const size_t n = (size_t)s64->nx * s64->ny;
  gfmd_fire_params_t fp;
  gfmd_fire_defaults(&fp);
  int rc;
  int it32 = 0, it64 = 0;
  double st[S_COUNT] = {0};
  auto upload_wall = [&](gfmd_solver_t *s) -> int {
    GFMD_HIP_TRY(hipMemcpyAsync(s->f_io, contact->wall, n * 8,
                                hipMemcpyHostToDevice, s->stream));
    launch_convert(s->f_io, true, s->rwall, is_fp64(s), n, s->stream);
    return ST_OK;
  };
  auto fetch_pk = [&](gfmd_solver_t *s, double *ps) -> int {
    GFMD_HIP_TRY(hipMemcpyAsync(ps, s->pkstate, sizeof(double) * PK_COUNT,
                                hipMemcpyDeviceToHost, s->stream));
    GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
    return ST_OK;
  };
  // the change measure sum|dp| / sum p below which the FP32 stage ends
  // (measured at 2048^2: 1e-4 .. 3e-6 give the same total; below, FP32
  // round-off stalls the stage)
  const double eps32 = 1e-5;
  const int chunk = 5;  // iterations between convergence checks
  double dev_s = 0.0;  // device phases, timed between synchronisations
  bool have_p = false;
  bool done = false;  // the given start already meets tol
  if (s32 && p_init) {
    // FP64 residual of the given start: a start that the FP32 stage cannot
    // improve (relative residual below its floor) goes straight to the FP64
    // stage, one that meets tol is returned as it is
    double g0;
    if ((rc = pk_prepare(s64, &g0)) || (rc = upload_wall(s64))) return rc;
    RelaxDev ps0;
    if ((rc = relax_params(s64, contact, &fp, ps0))) return rc;
    PKBufs b0 = pk_bufs(s64);
    GFMD_HIP_TRY(hipMemcpyAsync(s64->rv, p_init, n * 8, hipMemcpyHostToDevice,
                                s64->stream));
    launch_pk_init(b0, true, s64->rv, contact->pressure, s64->stream);
    launch_pk_h(b0, true, s64->rwall, contact->wall_offset,
                g0 * contact->pressure, s64->stream);
    GFMD_HIP_TRY(hipStreamSynchronize(s64->stream));
    const auto ts = Clock::now();
    if ((rc = pk_state3d(s64, ps0, st))) return rc;
    dev_s += std::chrono::duration<double>(Clock::now() - ts).count();
    const double r0 = st[S_RES] / (st[S_FEXT] > 0 ? st[S_FEXT] : 1.0);
    if (r0 <= tol)
      done = true;
    else if (r0 <= 1e-5)
      have_p = true;
  }
  // --- FP32 stage ---------------------------------------------------------
  if (s32 && !have_p && !done) {
    double g0;
    if ((rc = pk_prepare(s32, &g0)) || (rc = upload_wall(s32))) return rc;
    RelaxDev p32;
    if ((rc = relax_params(s32, contact, &fp, p32))) return rc;
    PKBufs b = pk_bufs(s32);
    const void *pin = nullptr;
    if (p_init) {
      GFMD_HIP_TRY(hipMemcpyAsync(s32->u_io, p_init, n * 8,
                                  hipMemcpyHostToDevice, s32->stream));
      launch_convert(s32->u_io, true, s32->rv, false, n, s32->stream);
      pin = s32->rv;
    }
    launch_pk_init(b, false, pin, contact->pressure, s32->stream);
    launch_pk_h(b, false, s32->rwall, contact->wall_offset,
                g0 * contact->pressure, s32->stream);
    GFMD_HIP_TRY(hipStreamSynchronize(s32->stream));
    const auto td = Clock::now();
    double ps[PK_COUNT];
    double best = 1e300;
    int stall = 0;
    while (it32 < max_iter) {
      if ((rc = pk_run(s32, chunk))) return rc;
      it32 += chunk;
      if ((rc = fetch_pk(s32, ps))) return rc;
      if (!std::isfinite(ps[PK_PSUM])) {
        set_error("gfmd_contact_solve: non-finite pressure (FP32 stage)");
        return ST_NONFINITE;
      }
      const double m = ps[PK_PSUM] > 0 ? ps[PK_DP] / ps[PK_PSUM] : 0.0;
      if (ps[PK_OVL] == 0 && m < eps32) break;
      // FP32 round-off floor: no progress over 5 chunks
      if (m < 0.5 * best) {
        best = m;
        stall = 0;
      } else if (++stall >= 5) {
        break;
      }
    }
    // hand the pressures to the FP64 handle
    rc = pk_prepare(s64, &g0);
    if (rc) return rc;
    PKBufs b64 = pk_bufs(s64);
    launch_convert(b.p, false, s64->rv, true, n, s32->stream);
    GFMD_HIP_TRY(hipStreamSynchronize(s32->stream));
    dev_s += std::chrono::duration<double>(Clock::now() - td).count();
    launch_pk_init(b64, true, s64->rv, contact->pressure, s64->stream);
    have_p = true;
  }
  // --- FP64 stage ---------------------------------------------------------
  double g0;
  if ((rc = pk_prepare(s64, &g0)) || (rc = upload_wall(s64))) return rc;
  RelaxDev p64;
  if ((rc = relax_params(s64, contact, &fp, p64))) return rc;
  PKBufs b64 = pk_bufs(s64);
  if (!have_p && !done) {
    const void *pin = nullptr;
    if (p_init) {
      GFMD_HIP_TRY(hipMemcpyAsync(s64->rv, p_init, n * 8,
                                  hipMemcpyHostToDevice, s64->stream));
      pin = s64->rv;
    }
    launch_pk_init(b64, true, pin, contact->pressure, s64->stream);
  }
  launch_pk_h(b64, true, s64->rwall, contact->wall_offset,
              g0 * contact->pressure, s64->stream);
  GFMD_HIP_TRY(hipStreamSynchronize(s64->stream));
  const auto td64 = Clock::now();
  double ps[PK_COUNT];
  double thresh = tol;  // change measure that triggers a full check
  bool conv = done;
  const int left = max_iter - it32;
  while (!done && it64 < left) {
    const int k = std::min(chunk, left - it64);
    if ((rc = pk_run(s64, k))) return rc;
    it64 += k;
    if ((rc = fetch_pk(s64, ps))) return rc;
    if (!std::isfinite(ps[PK_PSUM])) {
      set_error("gfmd_contact_solve: non-finite pressure");
      return ST_NONFINITE;
    }
    const double m = ps[PK_PSUM] > 0 ? ps[PK_DP] / ps[PK_PSUM] : 0.0;
    if (ps[PK_OVL] == 0 && (m < thresh || it64 >= left)) {
      if ((rc = pk_state3d(s64, p64, st))) return rc;
      const double denom = st[S_FEXT] > 0 ? st[S_FEXT] : 1.0;
      if (st[S_RES] / denom <= tol) {
        conv = true;
        break;
      }
      thresh *= 0.1;
    }
  }
  if (!conv && (rc = pk_state3d(s64, p64, st))) return rc;
  dev_s += std::chrono::duration<double>(Clock::now() - td64).count();
  launch_convert(s64->ru, true, s64->u_io, true, (size_t)s64->d * n,
                 s64->stream);
  GFMD_HIP_TRY(hipMemcpyAsync(u, s64->u_io, (size_t)s64->d * n * 8,
                              hipMemcpyDeviceToHost, s64->stream));
  GFMD_HIP_TRY(hipStreamSynchronize(s64->stream));
  if (!conv) {
    const double denom = st[S_FEXT] > 0 ? st[S_FEXT] : 1.0;
    conv = st[S_RES] / denom <= tol;
  }
  fill_result(res, st, it32 + it64, conv, t0, dev_s);
  if (iters) {
    iters[0] = it32;
    iters[1] = it64;
  }
  return ST_OK;
}

// A graph holding one iteration of every listed handle, each on its own
// stream (independent branches forked from and joined to `origin`).
// (Measured: chaining the large members one after another inside the graph
// instead of running all branches concurrently is slower.)
static int pk_build_batch(const std::vector<gfmd_solver_t *> &hs,
                          hipStream_t origin, hipGraphExec_t *exec) {
  std::vector<hipEvent_t> ev(hs.size() + 1, nullptr);
  int rc = ST_OK;
  hipError_t e = hipSuccess;
  for (auto &x : ev)
    if (e == hipSuccess)
      e = hipEventCreateWithFlags(&x, hipEventDisableTiming);
  hipGraph_t graph = nullptr;
  if (e == hipSuccess)
    e = hipStreamBeginCapture(origin, hipStreamCaptureModeRelaxed);
  if (e == hipSuccess) {
    e = hipEventRecord(ev[0], origin);
    for (size_t i = 0; i < hs.size() && e == hipSuccess && rc == ST_OK; i++) {
      e = hipStreamWaitEvent(hs[i]->stream, ev[0], 0);
      if (e == hipSuccess) rc = pk_iteration(hs[i]);
      if (e == hipSuccess && rc == ST_OK)
        e = hipEventRecord(ev[i + 1], hs[i]->stream);
      if (e == hipSuccess && rc == ST_OK)
        e = hipStreamWaitEvent(origin, ev[i + 1], 0);
    }
    const hipError_t e2 = hipStreamEndCapture(origin, &graph);
    if (e == hipSuccess) e = e2;
  }
  if (e == hipSuccess && rc == ST_OK)
    e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  for (auto x : ev)
    if (x) (void)hipEventDestroy(x);
  if (rc == ST_OK && e != hipSuccess) {
    set_error("contact batch: graph failed: %s", hipGetErrorString(e));
    rc = ST_HIP;
  }
  return rc;
}

int gfmd_contact_solve_batch(int count, gfmd_solver_t *const *s64,
                             gfmd_solver_t *const *s32,
                             const gfmd_contact_t *contacts, double *const *u,
                             int max_iter, double tol,
                             gfmd_relax_result_t *res) {
  if (count < 1 || !s64 || !contacts || !u || max_iter < 1 || !(tol > 0)) {
    set_error("gfmd_contact_solve_batch: invalid argument");
    return ST_ARG;
  }
  for (int i = 0; i < count; i++) {
    const gfmd_solver_t *a = s64[i], *b = s32 ? s32[i] : nullptr;
    if (!a || !u[i] || !is_fp64(a) || !a->has_phi ||
        (s32 && (!b || is_fp64(b) || !b->has_phi || b->nx != a->nx ||
                 b->ny != a->ny || b->d != a->d)) ||
        contacts[i].type != GFMD_CONTACT_HARDWALL || !contacts[i].wall) {
      set_error("gfmd_contact_solve_batch: member %d: needs an FP64 handle "
                "with a kernel (and an FP32 one of the same grid), a hard "
                "wall and an output array", i);
      return ST_ARG;
    }
    for (int j = 0; j < i; j++)
      if (s64[j] == a || (s32 && s32[j] == b)) {
        set_error("gfmd_contact_solve_batch: member %d repeats a handle", i);
        return ST_ARG;
      }
  }
  const auto t0 = Clock::now();
  gfmd_fire_params_t fp;
  gfmd_fire_defaults(&fp);
  struct Member {
    gfmd_solver_t *h64, *h32;
    RelaxDev p64;
    int it32 = 0, it64 = 0, stall = 0;
    double best = 1e300, thresh = 0.0;
    bool fp32_done = false, done = false, conv = false;
    double st[S_COUNT] = {0};
  };
  std::vector<Member> m(count);
  int rc;
  auto setup = [&](gfmd_solver_t *s, const gfmd_contact_t &c,
                   const void *p_dev) -> int {
    double g0;
    int r = pk_prepare(s, &g0);
    if (r) return r;
    const size_t n = (size_t)s->nx * s->ny;
    GFMD_HIP_TRY(hipMemcpyAsync(s->f_io, c.wall, n * 8, hipMemcpyHostToDevice,
                                s->stream));
    launch_convert(s->f_io, true, s->rwall, is_fp64(s), n, s->stream);
    PKBufs b = pk_bufs(s);
    launch_pk_init(b, is_fp64(s), p_dev, c.pressure, s->stream);
    launch_pk_h(b, is_fp64(s), s->rwall, c.wall_offset, g0 * c.pressure,
                s->stream);
    return ST_OK;
  };
  for (int i = 0; i < count; i++) {
    m[i].h64 = s64[i];
    m[i].h32 = s32 ? s32[i] : nullptr;
    m[i].thresh = tol;
    if ((rc = relax_params(m[i].h64, &contacts[i], &fp, m[i].p64))) return rc;
    if ((rc = setup(m[i].h32 ? m[i].h32 : m[i].h64, contacts[i], nullptr)))
      return rc;
    if (!m[i].h32) m[i].fp32_done = true;
  }
  for (int i = 0; i < count; i++) {
    gfmd_solver_t *h = m[i].h32 ? m[i].h32 : m[i].h64;
    GFMD_HIP_TRY(hipStreamSynchronize(h->stream));
  }
  hipStream_t origin = nullptr;
  GFMD_HIP_TRY(hipStreamCreateWithFlags(&origin, hipStreamNonBlocking));
  const auto td = Clock::now();
  const int chunk = 5;  // iterations between convergence checks
  const double eps32 = 1e-5;
  std::vector<double> ps((size_t)count * PK_COUNT);
  auto fetch = [&](bool fp32) -> int {
    for (int i = 0; i < count; i++) {
      gfmd_solver_t *h = fp32 ? m[i].h32 : m[i].h64;
      GFMD_HIP_TRY(hipMemcpyAsync(&ps[(size_t)i * PK_COUNT], h->pkstate,
                                  sizeof(double) * PK_COUNT,
                                  hipMemcpyDeviceToHost, origin));
    }
    GFMD_HIP_TRY(hipStreamSynchronize(origin));
    return ST_OK;
  };
  auto run_stage = [&](bool fp32) -> int {
    std::vector<gfmd_solver_t *> hs;
    for (auto &x : m) hs.push_back(fp32 ? x.h32 : x.h64);
    hipGraphExec_t exec = nullptr;
    int r = pk_build_batch(hs, origin, &exec);
    if (r) return r;
    int done_count = 0, it = 0;
    while (done_count < count && it < max_iter) {
      hipError_t e = hipSuccess;
      for (int k = 0; k < chunk && e == hipSuccess; k++)
        e = hipGraphLaunch(exec, origin);
      if (e != hipSuccess) {
        (void)hipGraphExecDestroy(exec);
        set_error("contact batch: launch failed: %s", hipGetErrorString(e));
        return ST_HIP;
      }
      it += chunk;
      if ((r = fetch(fp32))) break;
      for (int i = 0; i < count && r == ST_OK; i++) {
        Member &x = m[i];
        const double *q = &ps[(size_t)i * PK_COUNT];
        if (!std::isfinite(q[PK_PSUM])) {
          set_error("contact batch: member %d: non-finite pressure", i);
          r = ST_NONFINITE;
          break;
        }
        const double meas = q[PK_PSUM] > 0 ? q[PK_DP] / q[PK_PSUM] : 0.0;
        if (fp32) {
          if (x.fp32_done) continue;
          x.it32 = it;
          bool end = q[PK_OVL] == 0 && meas < eps32;
          if (meas < 0.5 * x.best) {
            x.best = meas;
            x.stall = 0;
          } else if (++x.stall >= 5) {
            end = true;
          }
          if (end) {
            x.fp32_done = true;
            done_count++;
            static const int zero = 0;
            GFMD_HIP_TRY(hipMemcpyAsync(x.h32->gate, &zero, sizeof(int),
                                        hipMemcpyHostToDevice, origin));
          }
        } else {
          if (x.done) continue;
          x.it64 = it;
          if ((q[PK_OVL] == 0 && meas < x.thresh) || x.it32 + it >= max_iter) {
            if ((r = pk_state3d(x.h64, x.p64, x.st))) break;
            const double den = x.st[S_FEXT] > 0 ? x.st[S_FEXT] : 1.0;
            x.conv = x.st[S_RES] / den <= tol;
            if (x.conv || x.it32 + it >= max_iter) {
              // the final state stays in the handle's u buffer (the
              // iterations do not touch it); copied out after the solve
              x.done = true;
              done_count++;
              static const int zero = 0;
              GFMD_HIP_TRY(hipMemcpyAsync(x.h64->gate, &zero, sizeof(int),
                                          hipMemcpyHostToDevice, origin));
            } else {
              x.thresh *= 0.1;
            }
          }
        }
      }
      if (r) break;
      if (fp32 && done_count == count) break;
    }
    (void)hipGraphExecDestroy(exec);
    if (r) return r;
    if (fp32) {
      // hand every member's FP32 pressures to its FP64 handle
      for (int i = 0; i < count; i++) {
        Member &x = m[i];
        const size_t n = (size_t)x.h64->nx * x.h64->ny;
        PKBufs b32 = pk_bufs(x.h32);
        double g0;
        if ((r = pk_prepare(x.h64, &g0))) return r;
        launch_convert(b32.p, false, x.h64->rv, true, n, x.h32->stream);
        GFMD_HIP_TRY(hipStreamSynchronize(x.h32->stream));
        if ((r = setup(x.h64, contacts[i], x.h64->rv))) return r;
      }
      for (int i = 0; i < count; i++)
        GFMD_HIP_TRY(hipStreamSynchronize(m[i].h64->stream));
    }
    return ST_OK;
  };
  rc = ST_OK;
  if (s32) rc = run_stage(true);
  if (!rc) rc = run_stage(false);
  (void)hipStreamSynchronize(origin);
  (void)hipStreamDestroy(origin);
  if (rc) return rc;
  for (int i = 0; i < count; i++) {
    Member &x = m[i];
    if (!x.done) {
      // iteration budget spent before the change test fired
      if ((rc = pk_state3d(x.h64, x.p64, x.st))) return rc;
      const double den = x.st[S_FEXT] > 0 ? x.st[S_FEXT] : 1.0;
      x.conv = x.st[S_RES] / den <= tol;
    }
  }
  const double dev_s = std::chrono::duration<double>(Clock::now() - td).count();
  for (int i = 0; i < count; i++) {
    Member &x = m[i];
    const size_t n = (size_t)x.h64->d * x.h64->nx * x.h64->ny;
    GFMD_HIP_TRY(hipMemcpyAsync(u[i], x.h64->ru, n * 8, hipMemcpyDeviceToHost,
                                x.h64->stream));
    GFMD_HIP_TRY(hipStreamSynchronize(x.h64->stream));
    if (res) fill_result(&res[i], x.st, x.it32 + x.it64, x.conv, t0, dev_s);
  }
  return ST_OK;
}

// ---------------------------------------------------------------------------
// Ensembles: members of one grid sharing the stiffness tables of prototype
// handles, fields contiguous per member, every sweep one launch over all
// members (blockIdx.y = member).
// ---------------------------------------------------------------------------
struct EnsStage {
  gfmd_solver_t *h = nullptr;  // prototype of this precision
  void *pk = nullptr;          // T[6][M][n]: p, t, g, r, u, h
  double *state = nullptr, *part = nullptr;   // [M][PK_COUNT], [M][8][nblk]
  int *gate = nullptr;                        // [M]
  void *spec = nullptr, *xbuf = nullptr;      // C2<T>[M][nky][nx], [M][nx]
  double *partials = nullptr, *u0 = nullptr, *epot = nullptr;
  void *wall = nullptr;                       // T[M][n]
  double *p0m = nullptr, *cm = nullptr;       // [M]
};

}  // extern "C"

struct gfmd_ensemble {
  int M = 0;
  size_t n = 0;
  EnsStage st[2];  // 0: FP64, 1: FP32 (h == nullptr when absent)
  double *uout = nullptr;  // final displacements, FP64 [M][d][n]
  hipStream_t stream = nullptr;
  size_t bytes = 0;
};

extern "C" {

static void ens_free(gfmd_ensemble_t *e) {
  for (auto &g : e->st) {
    void *ptrs[] = {g.pk,   g.state, g.part,  g.gate, g.spec, g.xbuf,
                    g.partials, g.u0, g.epot, g.wall, g.p0m,  g.cm};
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
  }
  if (e->uout) (void)hipFree(e->uout);
  if (e->stream) (void)hipStreamDestroy(e->stream);
}

// Buffers of members [m0, m0 + cnt) of stage k (all members by default).
static PKBufs ens_bufs(gfmd_ensemble_t *e, int k, int m0 = 0, int cnt = -1) {
  EnsStage &g = e->st[k];
  if (cnt < 0) cnt = e->M;
  const size_t w = e->M * e->n * real_size(g.h);
  char *b = (char *)g.pk + m0 * e->n * real_size(g.h);
  PKBufs r{b,
           b + w,
           b + 2 * w,
           b + 3 * w,
           b + 4 * w,
           b + 5 * w,
           g.state + (size_t)m0 * PK_COUNT,
           g.part + (size_t)m0 * 8 * g.h->rnblk,
           g.h->rnblk,
           e->n,
           g.gate + m0};
  r.members = cnt;
  return r;
}

static int ens_conv(gfmd_ensemble_t *e, int k, const void *src, void *dst,
                    int m0, int cnt) {
  EnsStage &g = e->st[k];
  gfmd_solver_t *h = g.h;
  ForceArgs a = make_args(h, src, dst);
  const size_t ms_s = (size_t)h->nky * h->nx;
  a.d = 1;
  a.phi = h->gz;
  a.linf = nullptr;
  a.spec = (char *)g.spec + m0 * ms_s * 2 * real_size(h);
  a.xbuf = (char *)g.xbuf + (size_t)m0 * h->nx * 2 * real_size(h);
  a.partials = g.partials + (size_t)m0 * h->nparts;
  a.u0 = g.u0 + m0;
  a.epot = g.epot + m0;
  a.gate = g.gate + m0;
  a.members = cnt;
  a.ms_u = a.ms_f = e->n;
  a.ms_s = ms_s;
  const bool f64 = is_fp64(h);
  if (!force_fast_variant_ok(h->nx, h->ny, 1, f64, a.row_var, a.col_var))
    a.row_var = a.col_var = 0;
  const int rc = launch_force_fast(a, f64, f64, e->stream);
  if (rc) set_error("ensemble convolution failed (status %d)", rc);
  return rc;
}

// One iteration of all members: every sweep one launch over the members.
// (Measured: splitting the members into cache-sized waves is slower.)
static int ens_iteration(gfmd_ensemble_t *e, int k) {
  const bool f64 = is_fp64(e->st[k].h);
  PKBufs b = ens_bufs(e, k);
  int rc = ens_conv(e, k, b.p, b.u, 0, e->M);
  if (!rc) launch_pk(b, f64, PK_GAP, e->stream);
  if (!rc) rc = ens_conv(e, k, b.t, b.r, 0, e->M);
  if (!rc) launch_pk(b, f64, PK_CURV_UPDATE, e->stream);
  return rc;
}

int gfmd_ensemble_create(gfmd_solver_t *s64, gfmd_solver_t *s32, int members,
                         gfmd_ensemble_t **out) {
  if (!out || !s64 || members < 1) {
    set_error("gfmd_ensemble_create: invalid argument");
    return ST_ARG;
  }
  *out = nullptr;
  if (!is_fp64(s64) || !s64->has_phi ||
      (s32 && (is_fp64(s32) || !s32->has_phi || s32->nx != s64->nx ||
               s32->ny != s64->ny || s32->d != s64->d))) {
    set_error("gfmd_ensemble_create: needs an FP64 handle with a kernel "
              "(and optionally an FP32 one of the same grid)");
    return ST_ARG;
  }
  if (!force_fast_supported(s64->nx, s64->ny, 1, true) ||
      !force_fast_supported(s64->nx, s64->ny, 1, false)) {
    set_error("gfmd_ensemble_create: ensembles need power-of-two grids "
              "from 64 to 4096");
    return ST_UNSUPPORTED;
  }
  gfmd_ensemble_t *e = new gfmd_ensemble_t;
  e->M = members;
  e->n = (size_t)s64->nx * s64->ny;
  auto fail = [&](int rc) {
    ens_free(e);
    delete e;
    return rc;
  };
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) !=
      hipSuccess) {
    set_error("gfmd_ensemble_create: stream creation failed");
    return fail(ST_HIP);
  }
  gfmd_solver_t *hs[2] = {s64, s32};
  for (int k = 0; k < 2; k++) {
    if (!hs[k]) continue;
    gfmd_solver_t *h = hs[k];
    double g0;
    int rc = pk_prepare(h, &g0);
    if (rc) return fail(rc);
    EnsStage &g = e->st[k];
    g.h = h;
    const size_t M = members, w = real_size(h), n = e->n;
    const size_t nspec = (size_t)h->nky * h->nx * 2 * w;
    auto alloc = [&](void **p, size_t bytes) -> int {
      if (hipMalloc(p, bytes ? bytes : 8) != hipSuccess) {
        *p = nullptr;
        set_error("gfmd_ensemble_create: out of device memory (%zu bytes)",
                  bytes);
        return ST_HIP;
      }
      e->bytes += bytes;
      return ST_OK;
    };
    if ((rc = alloc(&g.pk, 6 * M * n * w)) ||
        (rc = alloc((void **)&g.state, M * PK_COUNT * 8)) ||
        (rc = alloc((void **)&g.part, M * 8 * h->rnblk * 8)) ||
        (rc = alloc((void **)&g.gate, M * sizeof(int))) ||
        (rc = alloc(&g.spec, M * nspec)) ||
        (rc = alloc(&g.xbuf, M * h->nx * 2 * w)) ||
        (rc = alloc((void **)&g.partials, M * h->nparts * 8)) ||
        (rc = alloc((void **)&g.u0, M * 8)) ||
        (rc = alloc((void **)&g.epot, M * 8)) ||
        (rc = alloc(&g.wall, M * n * w)) ||
        (rc = alloc((void **)&g.p0m, M * 8)) ||
        (rc = alloc((void **)&g.cm, M * 8)))
      return fail(rc);
  }
  if (hipMalloc((void **)&e->uout, (size_t)members * s64->d * e->n * 8) !=
      hipSuccess) {
    e->uout = nullptr;
    set_error("gfmd_ensemble_create: out of device memory");
    return fail(ST_HIP);
  }
  *out = e;
  return ST_OK;
}

int gfmd_ensemble_destroy(gfmd_ensemble_t *e) {
  if (!e) return ST_OK;
  (void)hipStreamSynchronize(e->stream);
  ens_free(e);
  delete e;
  return ST_OK;
}

int gfmd_ensemble_contact_solve(int count, gfmd_ensemble_t *const *ens,
                                const gfmd_contact_t *contacts,
                                double *const *u, int max_iter, double tol,
                                gfmd_relax_result_t *res) {
  if (count < 1 || !ens || !contacts || !u || max_iter < 1 || !(tol > 0)) {
    set_error("gfmd_ensemble_contact_solve: invalid argument");
    return ST_ARG;
  }
  std::vector<int> first(count + 1, 0);
  for (int k = 0; k < count; k++) {
    if (!ens[k]) {
      set_error("gfmd_ensemble_contact_solve: NULL ensemble %d", k);
      return ST_ARG;
    }
    for (int j = 0; j < k; j++)
      if (ens[j] == ens[k]) {
        set_error("gfmd_ensemble_contact_solve: ensemble %d repeated", k);
        return ST_ARG;
      }
    first[k + 1] = first[k] + ens[k]->M;
  }
  const int total = first[count];
  for (int i = 0; i < total; i++)
    if (contacts[i].type != GFMD_CONTACT_HARDWALL || !contacts[i].wall ||
        !u[i]) {
      set_error("gfmd_ensemble_contact_solve: member %d needs a hard wall "
                "and an output array", i);
      return ST_ARG;
    }
  const bool mixed = ens[0]->st[1].h != nullptr;
  for (int k = 0; k < count; k++)
    if ((ens[k]->st[1].h != nullptr) != mixed) {
      set_error("gfmd_ensemble_contact_solve: all ensembles need an FP32 "
                "stage, or none");
      return ST_ARG;
    }
  const auto t0 = Clock::now();
  gfmd_fire_params_t fp;
  gfmd_fire_defaults(&fp);
  int rc;
  // per-member set-up of stage kk: walls, p0, c = G_zz(0) p0 - offset,
  // then p = p_init (or p0) and h
  auto setup = [&](int k, int kk, const void *p_init) -> int {
    gfmd_ensemble_t *e = ens[k];
    EnsStage &g = e->st[kk];
    double g0;
    int r = pk_prepare(g.h, &g0);
    if (r) return r;
    const int M = e->M;
    std::vector<double> p0(M), cm(M);
    for (int m = 0; m < M; m++) {
      const gfmd_contact_t &c = contacts[first[k] + m];
      p0[m] = c.pressure;
      cm[m] = g0 * c.pressure - c.wall_offset;
      GFMD_HIP_TRY(hipMemcpyAsync(g.h->f_io, c.wall, e->n * 8,
                                  hipMemcpyHostToDevice, e->stream));
      launch_convert(g.h->f_io, true,
                     (char *)g.wall + m * e->n * real_size(g.h),
                     is_fp64(g.h), e->n, e->stream);
      GFMD_HIP_TRY(hipStreamSynchronize(e->stream));
    }
    GFMD_HIP_TRY(hipMemcpyAsync(g.p0m, p0.data(), M * 8,
                                hipMemcpyHostToDevice, e->stream));
    GFMD_HIP_TRY(hipMemcpyAsync(g.cm, cm.data(), M * 8,
                                hipMemcpyHostToDevice, e->stream));
    PKBufs b = ens_bufs(e, kk);
    launch_pk_init(b, is_fp64(g.h), p_init, 0.0, e->stream, g.p0m);
    launch_pk_h(b, is_fp64(g.h), g.wall, 0.0, 0.0, e->stream, g.cm);
    GFMD_HIP_TRY(hipStreamSynchronize(e->stream));
    return ST_OK;
  };
  struct Mem {
    int it32 = 0, it64 = 0, stall = 0;
    double best = 1e300, thresh = 0.0;
    bool fp32_done = false, done = false, conv = false;
    double st[S_COUNT] = {0};
  };
  std::vector<Mem> mem(total);
  for (auto &x : mem) x.thresh = tol;
  for (int k = 0; k < count; k++)
    if ((rc = setup(k, mixed ? 1 : 0, nullptr))) return rc;
  hipStream_t origin = nullptr;
  GFMD_HIP_TRY(hipStreamCreateWithFlags(&origin, hipStreamNonBlocking));
  const auto td = Clock::now();
  const int chunk = 5;  // iterations between convergence checks
  const double eps32 = 1e-5;
  std::vector<double> ps((size_t)total * PK_COUNT);
  static const int zero = 0;
  // final check of one member on its FP64 prototype
  auto check3d = [&](int k, int m) -> int {
    gfmd_ensemble_t *e = ens[k];
    EnsStage &g = e->st[0];
    gfmd_solver_t *h = g.h;
    Mem &x = mem[first[k] + m];
    PKBufs bp = pk_bufs(h);
    PKBufs be = ens_bufs(e, 0);
    GFMD_HIP_TRY(hipMemcpyAsync(bp.p, (char *)be.p + m * e->n * 8, e->n * 8,
                                hipMemcpyDeviceToDevice, h->stream));
    GFMD_HIP_TRY(hipMemcpyAsync(h->rwall, (char *)g.wall + m * e->n * 8,
                                e->n * 8, hipMemcpyDeviceToDevice, h->stream));
    RelaxDev p;
    int r = relax_params(h, &contacts[first[k] + m], &fp, p);
    if (!r) r = pk_state3d(h, p, x.st);
    if (r) return r;
    const double den = x.st[S_FEXT] > 0 ? x.st[S_FEXT] : 1.0;
    x.conv = x.st[S_RES] / den <= tol;
    return ST_OK;
  };
  // keep member m's final state on the device (the prototype's buffer is
  // reused by the next member's test); copied to the host after the solve
  auto keep = [&](int k, int m) -> int {
    gfmd_solver_t *h = ens[k]->st[0].h;
    const size_t nn = (size_t)h->d * ens[k]->n;
    GFMD_HIP_TRY(hipMemcpyAsync(ens[k]->uout + m * nn, h->ru, nn * 8,
                                hipMemcpyDeviceToDevice, h->stream));
    return ST_OK;
  };
  auto run_stage = [&](int kk) -> int {
    // one graph: a branch per ensemble, each one batched iteration
    std::vector<hipEvent_t> ev(count + 1, nullptr);
    hipError_t e = hipSuccess;
    int r = ST_OK;
    for (auto &x : ev)
      if (e == hipSuccess)
        e = hipEventCreateWithFlags(&x, hipEventDisableTiming);
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (e == hipSuccess)
      e = hipStreamBeginCapture(origin, hipStreamCaptureModeRelaxed);
    if (e == hipSuccess) {
      e = hipEventRecord(ev[0], origin);
      for (int k = 0; k < count && e == hipSuccess && r == ST_OK; k++) {
        e = hipStreamWaitEvent(ens[k]->stream, ev[0], 0);
        if (e == hipSuccess) r = ens_iteration(ens[k], kk);
        if (e == hipSuccess && r == ST_OK)
          e = hipEventRecord(ev[k + 1], ens[k]->stream);
        if (e == hipSuccess && r == ST_OK)
          e = hipStreamWaitEvent(origin, ev[k + 1], 0);
      }
      const hipError_t e2 = hipStreamEndCapture(origin, &graph);
      if (e == hipSuccess) e = e2;
    }
    if (e == hipSuccess && r == ST_OK)
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto x : ev)
      if (x) (void)hipEventDestroy(x);
    if (r == ST_OK && e != hipSuccess) {
      set_error("ensemble: graph failed: %s", hipGetErrorString(e));
      r = ST_HIP;
    }
    int done_count = 0, it = 0;
    while (r == ST_OK && done_count < total && it < max_iter) {
      for (int c = 0; c < chunk && e == hipSuccess; c++)
        e = hipGraphLaunch(exec, origin);
      if (e != hipSuccess) {
        set_error("ensemble: launch failed: %s", hipGetErrorString(e));
        r = ST_HIP;
        break;
      }
      it += chunk;
      for (int k = 0; k < count; k++)
        if (hipMemcpyAsync(&ps[(size_t)first[k] * PK_COUNT],
                           ens[k]->st[kk].state,
                           sizeof(double) * PK_COUNT * ens[k]->M,
                           hipMemcpyDeviceToHost, origin) != hipSuccess)
          r = ST_HIP;
      if (r || hipStreamSynchronize(origin) != hipSuccess) {
        set_error("ensemble: state transfer failed");
        r = ST_HIP;
        break;
      }
      for (int k = 0; k < count && r == ST_OK; k++)
        for (int m = 0; m < ens[k]->M && r == ST_OK; m++) {
          const int i = first[k] + m;
          Mem &x = mem[i];
          const double *q = &ps[(size_t)i * PK_COUNT];
          if (!std::isfinite(q[PK_PSUM])) {
            set_error("ensemble: member %d: non-finite pressure", i);
            r = ST_NONFINITE;
            break;
          }
          const double meas = q[PK_PSUM] > 0 ? q[PK_DP] / q[PK_PSUM] : 0.0;
          bool off = false;
          if (kk == 1) {
            if (x.fp32_done) continue;
            x.it32 = it;
            bool end = q[PK_OVL] == 0 && meas < eps32;
            if (meas < 0.5 * x.best) {
              x.best = meas;
              x.stall = 0;
            } else if (++x.stall >= 5) {
              end = true;
            }
            if (end) x.fp32_done = off = true;
          } else {
            if (x.done) continue;
            x.it64 = it;
            if ((q[PK_OVL] == 0 && meas < x.thresh) ||
                x.it32 + it >= max_iter) {
              if ((r = check3d(k, m))) break;
              if (x.conv || x.it32 + it >= max_iter) {
                if ((r = keep(k, m))) break;
                x.done = off = true;
              } else {
                x.thresh *= 0.1;
              }
            }
          }
          if (off) {
            done_count++;
            if (hipMemcpyAsync(ens[k]->st[kk].gate + m, &zero, sizeof(int),
                               hipMemcpyHostToDevice, origin) != hipSuccess)
              r = ST_HIP;
          }
        }
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    (void)hipStreamSynchronize(origin);
    return r;
  };
  rc = ST_OK;
  if (mixed) {
    rc = run_stage(1);
    // FP32 pressures into the FP64 stage, then its set-up
    for (int k = 0; k < count && rc == ST_OK; k++) {
      gfmd_ensemble_t *e = ens[k];
      const size_t mn = (size_t)e->M * e->n;
      launch_convert(e->st[1].pk, false, e->st[0].pk, true, mn, e->stream);
      rc = setup(k, 0, e->st[0].pk);
    }
  }
  if (rc == ST_OK) rc = run_stage(0);
  for (int k = 0; k < count && rc == ST_OK; k++)
    for (int m = 0; m < ens[k]->M && rc == ST_OK; m++)
      if (!mem[first[k] + m].done) {
        // iteration budget spent before the change test fired
        if (!(rc = check3d(k, m))) rc = keep(k, m);
      }
  for (int k = 0; k < count && rc == ST_OK; k++)
    if (hipStreamSynchronize(ens[k]->st[0].h->stream) != hipSuccess)
      rc = ST_HIP;
  const double dev_s = std::chrono::duration<double>(Clock::now() - td).count();
  (void)hipStreamDestroy(origin);
  if (rc) return rc;
  for (int k = 0; k < count; k++) {
    const size_t nn = (size_t)ens[k]->st[0].h->d * ens[k]->n;
    for (int m = 0; m < ens[k]->M; m++) {
      Mem &x = mem[first[k] + m];
      GFMD_HIP_TRY(hipMemcpy(u[first[k] + m], ens[k]->uout + m * nn, nn * 8,
                             hipMemcpyDeviceToHost));
      if (res) fill_result(&res[first[k] + m], x.st, x.it32 + x.it64, x.conv,
                           t0, dev_s);
    }
  }
  return ST_OK;
}

int gfmd_relax_refine(gfmd_solver_t *s64, gfmd_solver_t *s32,
                      const gfmd_contact_t *contact,
                      const gfmd_fire_params_t *fire, double *u,
                      int max_outer, double tol, double inner_tol,
                      int inner_max_steps, int check_every,
                      gfmd_relax_result_t *res, int *outer_iters) {
  if (!s64 || !s32 || !fire || !u || max_outer < 1 || check_every < 1 ||
      inner_max_steps < 1) {
    set_error("gfmd_relax_refine: invalid argument");
    return ST_ARG;
  }
  if (!is_fp64(s64) || is_fp64(s32) || s64->nx != s32->nx ||
      s64->ny != s32->ny || s64->d != s32->d) {
    set_error("gfmd_relax_refine: needs an FP64 and an FP32 handle of the "
              "same grid");
    return ST_ARG;
  }
  if (!s64->has_phi || !s32->has_phi) {
    set_error("gfmd_relax_refine: no stiffness kernel set");
    return ST_STATE;
  }
  int rc;
  if ((rc = relax_alloc(s64)) || (rc = relax_alloc(s32))) return rc;
  RelaxDev p64, p32;
  if ((rc = relax_params(s64, contact, fire, p64)) ||
      (rc = relax_params(s32, contact, fire, p32)))
    return rc;
  // the FP32 problem: correction d with force f_el(d) + r, wall d_z <= g
  // (hard wall: r = f_el(u64) + load, g the gap); for a nonlinear contact
  // r = the FP64 residual at u64 and the contact enters as
  // f_c(u64 + d) - f_c(u64), both evaluated in FP64 from the FP64 base
  p32.pressure = 0.0;
  p32.wall_offset = 0.0;
  p32.fext = s32->u_io;  // r32 (float, fits the double staging buffer)
  const bool wall = p64.contact == CONTACT_HARDWALL;
  const bool nonlinear = p64.contact == CONTACT_LJ_SPHERE ||
                         p64.contact == CONTACT_WALL_MAP;
  // per-site clearances (FP32 handle's wall buffer, unused otherwise) let
  // the inner passes skip atoms that stay out of contact; lateral moves of
  // up to one spline box of the wall map are covered
  double rho = 0.0;
  if (nonlinear) {
    p32.ubase = (const double *)s64->ru;
    p32.fcbase = (const double *)s64->rv;  // unused by the FP64 handle here
    p32.clear = (const float *)s32->rwall;
    if (p64.contact == CONTACT_WALL_MAP)
      rho = 1.0 / std::max(p64.wm.qx, p64.wm.qy);
    p32.clear_rho = rho;
  }
  const size_t n = (size_t)s64->d * s64->nx * s64->ny;
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t st64 = s64->stream;
  GFMD_HIP_TRY(hipMemcpyAsync(s64->u_io, u, n * 8, hipMemcpyHostToDevice,
                              st64));
  launch_convert(s64->u_io, true, s64->ru, true, n, st64);
  if (wall) {
    const size_t nw = (size_t)s64->nx * s64->ny;
    GFMD_HIP_TRY(hipMemcpyAsync(s64->rwall, contact->wall, nw * 8,
                                hipMemcpyHostToDevice, st64));
  }
  RelaxBufs b64{s64->ru, s64->rv, s64->rf, s64->rfel, s64->rstate,
                s64->rpart, s64->repot, s64->rnblk};
  double st[S_COUNT], sti[S_COUNT];
  int total_steps = 0, outer = 0;
  bool conv = false;
  GFMD_HIP_TRY(hipStreamSynchronize(st64));
  const auto td = Clock::now();
  for (;;) {
    // FP64 residual of the current state
    if ((rc = run_force(s64, s64->ru, s64->rfel, s64->repot, true))) return rc;
    launch_relax_step(p64, b64, true, STEP_INIT, st64);
    launch_relax_step(p64, b64, true, STEP_FORCE, st64);
    GFMD_HIP_TRY(hipMemcpyAsync(st, s64->rstate, sizeof(st),
                                hipMemcpyDeviceToHost, st64));
    GFMD_HIP_TRY(hipStreamSynchronize(st64));
    if (!std::isfinite(st[S_RES])) {
      set_error("gfmd_relax_refine: non-finite residual");
      return ST_NONFINITE;
    }
    const double denom = st[S_FEXT] > 0 ? st[S_FEXT] : 1.0;
    conv = st[S_RES] / denom <= tol;
    if (conv || outer >= max_outer) break;
    // FP32 correction problem
    if (nonlinear) {
      // r32 = the total FP64 force of the check (elastic + load + contact)
      launch_refine_base(p64, (const double *)s64->ru, (double *)s64->rv,
                         (float *)s32->rwall, rho, st64);
      launch_convert(s64->rf, true, s32->u_io, false, n, st64);
      GFMD_HIP_TRY(hipMemsetAsync(s32->ru, 0, n * sizeof(float), st64));
    } else {
      launch_refine_setup(p64, (const double *)s64->ru,
                          (const double *)s64->rfel,
                          wall ? (const double *)s64->rwall : nullptr,
                          (float *)s32->u_io, (float *)s32->rwall,
                          (float *)s32->ru, st64);
    }
    GFMD_HIP_TRY(hipStreamSynchronize(st64));
    int steps = 0;
    bool iconv = false;
    // no inner solve needs to go below half the outer goal
    rc = fire_loop(s32, p32, inner_max_steps, inner_tol, check_every, &steps,
                   sti, &iconv, true, outer > 0, 0.5 * tol * denom);
    total_steps += steps;
    if (rc) return rc;
    GFMD_HIP_TRY(hipStreamSynchronize(s32->stream));
    launch_refine_update(p64, (double *)s64->ru, (const float *)s32->ru,
                         wall ? (const double *)s64->rwall : nullptr, st64);
    outer++;
  }
  const double dev_s =
      std::chrono::duration<double>(Clock::now() - td).count();
  launch_convert(s64->ru, true, s64->u_io, true, n, st64);
  GFMD_HIP_TRY(hipMemcpyAsync(u, s64->u_io, n * 8, hipMemcpyDeviceToHost,
                              st64));
  GFMD_HIP_TRY(hipStreamSynchronize(st64));
  fill_result(res, st, total_steps, conv, t0, dev_s);
  if (outer_iters) *outer_iters = outer;
  return ST_OK;
}

int gfmd_rough_surface(int nx, int ny, double hurst, double rms_slope,
                       double lambda_short, double lambda_long, double a,
                       unsigned long long seed, double *out) {
  if (!out || nx < 2 || ny < 2 || !(a > 0) || !(rms_slope >= 0) ||
      !(lambda_short > 0)) {
    set_error("gfmd_rough_surface: invalid argument");
    return ST_ARG;
  }
  gfmd_solver_t *s = nullptr;
  int rc = gfmd_create(nx, ny, 3, GFMD_FP64, &s);
  if (rc) return rc;
  KernelDesc k{};
  k.kind = KIND_FILTER;
  k.d = 3;
  k.p[0] = hurst;
  k.p[1] = 2.0 * M_PI / lambda_short;
  const double L = (nx > ny ? nx : ny) * a;
  k.p[2] = lambda_long > 0 ? 2.0 * M_PI / lambda_long : 2.0 * M_PI / L;
  k.p[3] = a;
  const int nblk = 1024;
  double *part = nullptr;
  hipError_t e = hipMalloc((void **)&part, nblk * sizeof(double));
  if (e != hipSuccess) {
    gfmd_destroy(s);
    set_error("gfmd_rough_surface: %s", hipGetErrorString(e));
    return ST_HIP;
  }
  rc = launch_phi_generate(k, nx, ny, 1, true, s->phi, s->status, s->stream);
  if (!rc) {
    s->has_phi = true;
  s->gz_valid = false;
    launch_white_noise(s->u_io, nx, ny, seed, s->stream);
    rc = run_force(s, s->u_io, s->f_io, nullptr, true);
  }
  const size_t N = (size_t)nx * ny;
  if (!rc) {
    launch_slope_partials(s->f_io, nx, ny, a, part, nblk, s->stream);
    launch_scale(s->f_io, N, part, nblk, rms_slope, s->stream);
    e = hipMemcpyAsync(out, s->f_io, N * sizeof(double), hipMemcpyDeviceToHost,
                       s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
      set_error("gfmd_rough_surface: %s", hipGetErrorString(e));
      rc = ST_HIP;
    }
  }
  (void)hipFree(part);
  gfmd_destroy(s);
  return rc;
}

// ---------------------------------------------------------------------------
// fix wall/map: set-up and the wall force on a list of atoms.
// ---------------------------------------------------------------------------
static void wall_map_free(gfmd_wall_map_t *w) {
  if (!w) return;
  if (w->buf) (void)hipFree(w->buf);
  if (w->work) (void)hipFree(w->work);
  if (w->stream) (void)hipStreamDestroy(w->stream);
  delete w;
}

// Low-pass filter of the map (FixWallMap::filter_map) by one force sweep of
// a map-sized handle with the filter as its stiffness: f = A(q) u.
static int wall_map_filter(double *values, int mx, int my, double lambda,
                           double xprd, double yprd) {
  gfmd_solver_t *t = nullptr;
  int rc = gfmd_create(mx, my, 3, GFMD_FP64, &t);
  if (rc) return rc;
  KernelDesc k{};
  k.kind = KIND_LOWPASS;
  k.d = 3;
  k.p[0] = 1.0 / lambda;
  k.p[1] = xprd;
  k.p[2] = yprd;
  const size_t N = (size_t)mx * my;
  rc = launch_phi_generate(k, mx, my, 1, true, t->phi, t->status, t->stream);
  hipError_t e = hipSuccess;
  if (!rc) {
    t->has_phi = true;
    t->gz_valid = false;
    e = hipMemsetAsync(t->u_io, 0, 3 * N * sizeof(double), t->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(t->u_io, values, N * sizeof(double),
                         hipMemcpyDeviceToDevice, t->stream);
    if (e == hipSuccess) rc = run_force(t, t->u_io, t->f_io, nullptr, true);
    if (e == hipSuccess && !rc)
      e = hipMemcpyAsync(values, t->f_io, N * sizeof(double),
                         hipMemcpyDeviceToDevice, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
  }
  gfmd_destroy(t);
  if (e != hipSuccess) {
    set_error("gfmd_wall_map_create: filter: %s", hipGetErrorString(e));
    return ST_HIP;
  }
  return rc;
}

int gfmd_wall_map_create(const char *args, const double *map, int map_nx,
                         int map_ny, double xprd, double yprd,
                         gfmd_wall_map_t **out) {
  if (!out || !args) {
    set_error("gfmd_wall_map_create: invalid argument");
    return ST_ARG;
  }
  *out = nullptr;
  WallMapSpec spec;
  const std::string msg =
      wall_map_parse(args, map, map_nx, map_ny, xprd, yprd, spec);
  if (!msg.empty()) {
    set_error("%s", msg.c_str());
    return ST_ARG;
  }
  gfmd_wall_map_t *w = new gfmd_wall_map;
  w->dev = spec.dev;
  w->xprd = xprd;
  w->yprd = yprd;
  const bool table = spec.dev.map == WM_TABLE || spec.dev.map == WM_RAW;
  const int mx = table ? spec.dev.mx : 0, my = table ? spec.dev.my : 0;
  const size_t nv = (size_t)mx * my;
  const int ntx = (mx + kTile - 1) / kTile, nty = (my + kTile - 1) / kTile;
  const size_t nt = spec.dev.map == WM_TABLE ? (size_t)ntx * nty : 0;
  auto fail = [&](hipError_t e) {
    set_error("gfmd_wall_map_create: %s", hipGetErrorString(e));
    wall_map_free(w);
    return ST_HIP;
  };
  hipError_t e = hipStreamCreate(&w->stream);
  if (e != hipSuccess) return fail(e);
  // [0, 8): statistics; then the spline's tile bounds; then the heights
  e = hipMalloc((void **)&w->buf, (8 + 2 * nt + nv) * sizeof(double));
  if (e != hipSuccess) return fail(e);
  double *tiles = w->buf + 8;
  double *vals = tiles + 2 * nt;
  if (table)
    e = hipMemcpyAsync(vals, spec.values.data(), nv * sizeof(double),
                       hipMemcpyHostToDevice, w->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
  if (e != hipSuccess) return fail(e);
  w->dev.values = table ? vals : nullptr;
  w->dev.tiles = nullptr;
  w->dev.nty = nty;
  if (table) {
    if (spec.lambda > 0.0) {
      const int rc = wall_map_filter(vals, mx, my, spec.lambda, xprd, yprd);
      if (rc) {
        wall_map_free(w);
        return rc;
      }
    }
    launch_map_displace(vals, mx, my, spec.dev.z0, spec.z0_type, w->buf,
                        w->stream);
    if (nt) {
      launch_map_tile_bounds(w->dev, tiles, w->stream);
      w->dev.tiles = tiles;
    }
    double st[5];
    e = hipMemcpyAsync(st, w->buf, sizeof st, hipMemcpyDeviceToHost,
                       w->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) return fail(e);
    for (int k = 0; k < 4; k++) w->info[k] = st[k];
    w->info[4] = mx;
    w->info[5] = my;
  }
  *out = w;
  return ST_OK;
}

int gfmd_wall_map_destroy(gfmd_wall_map_t *w) {
  if (w) {
    (void)hipStreamSynchronize(w->stream);
    wall_map_free(w);
  }
  return ST_OK;
}

int gfmd_wall_map_info(const gfmd_wall_map_t *w, double *info) {
  if (!w || !info) {
    set_error("gfmd_wall_map_info: invalid argument");
    return ST_ARG;
  }
  for (int k = 0; k < 6; k++) info[k] = w->info[k];
  return ST_OK;
}

int gfmd_wall_map_force(gfmd_wall_map_t *w, int n, double *x, double *f,
                        double *fmap) {
  if (!w || n < 0 || (n > 0 && (!x || !f)) || !fmap) {
    set_error("gfmd_wall_map_force: invalid argument");
    return ST_ARG;
  }
  for (int k = 0; k < 7; k++) fmap[k] = 0.0;
  if (n == 0) return ST_OK;
  const int nblk = std::min((n + 255) / 256, 1024);
  const size_t need = (size_t)6 * n + 8 * (size_t)nblk + 8;
  if (need > w->work_cap) {
    if (w->work) (void)hipFree(w->work);
    w->work = nullptr;
    w->work_cap = 0;
    GFMD_HIP_TRY(hipMalloc((void **)&w->work, need * sizeof(double)));
    w->work_cap = need;
  }
  double *xd = w->work, *fd = xd + 3 * (size_t)n, *part = fd + 3 * (size_t)n;
  double *res = part + 8 * (size_t)nblk;
  const size_t nb = 3 * (size_t)n * sizeof(double);
  GFMD_HIP_TRY(hipMemcpyAsync(xd, x, nb, hipMemcpyHostToDevice, w->stream));
  GFMD_HIP_TRY(hipMemcpyAsync(fd, f, nb, hipMemcpyHostToDevice, w->stream));
  launch_wall_map_force(w->dev, n, xd, fd, part, nblk, res, w->stream);
  GFMD_HIP_TRY(hipGetLastError());
  double r[8];
  GFMD_HIP_TRY(hipMemcpyAsync(x, xd, nb, hipMemcpyDeviceToHost, w->stream));
  GFMD_HIP_TRY(hipMemcpyAsync(f, fd, nb, hipMemcpyDeviceToHost, w->stream));
  GFMD_HIP_TRY(hipMemcpyAsync(r, res, sizeof r, hipMemcpyDeviceToHost,
                              w->stream));
  GFMD_HIP_TRY(hipStreamSynchronize(w->stream));
  for (int k = 0; k < 7; k++) fmap[k] = r[k];
  if (r[7] > 0) {
    set_error("fix contact/map: Atom too close: Implement short ranged spline "
              "for lj126 potential. (%g atoms)", r[7]);
    return ST_CONTACT;
  }
  return ST_OK;
}

// ---------------------------------------------------------------------------
// fix gfmd glue: atoms <-> grid on the device.
// ---------------------------------------------------------------------------
static bool fix_atoms(const gfmd_atoms_t *a, FixAtoms &fa, const char *who) {
  if (!a || a->n < 0 || (a->n > 0 && (!a->gid || !a->x || !a->xeq))) {
    set_error("%s: atoms need gid, x and xeq", who);
    return false;
  }
  fa = FixAtoms{a->n, a->mask, a->groupbit, a->gid, a->x, a->xeq, a->v,
                a->image, a->mass};
  return true;
}

int gfmd_fix_setup(gfmd_solver_t *s, const gfmd_atoms_t *a,
                   const double box[3], const double invcell[9]) {
  FixAtoms fa;
  if (!s || !box || !invcell || !fix_atoms(a, fa, "gfmd_fix_setup")) {
    if (!s || !box || !invcell) set_error("gfmd_fix_setup: invalid argument");
    return ST_ARG;
  }
  if (!(box[0] > 0) || !(box[1] > 0) || !(box[2] > 0)) {
    set_error("gfmd_fix_setup: box lengths must be positive");
    return ST_ARG;
  }
  int rc;
  if (!s->fix_st && (rc = dev_alloc(s, (void **)&s->fix_st, FS_COUNT * 8)))
    return rc;
  if (!s->fix_part &&
      (rc = dev_alloc(s, (void **)&s->fix_part, 5 * 1024 * 8)))
    return rc;
  FixFrame &fr = s->fix_frame;
  fr.nx = s->nx;
  fr.ny = s->ny;
  fr.nl = s->d / 3;
  for (int k = 0; k < 3; k++) fr.box[k] = box[k];
  for (int k = 0; k < 6; k++) fr.invcell[k] = invcell[k];
  GFMD_HIP_TRY(hipMemsetAsync(s->fix_st, 0, FS_COUNT * 8, s->stream));
  launch_fix_com(fa, fr, true, s->fix_part, s->fix_st, s->stream);
  GFMD_HIP_TRY(hipGetLastError());
  s->fix_ready = true;
  return ST_OK;
}

int gfmd_fix_pre_force(gfmd_solver_t *s, const gfmd_atoms_t *a, void *u_dev,
                       void *v_grid_dev) {
  FixAtoms fa;
  if (!s || !u_dev) {
    set_error("gfmd_fix_pre_force: invalid argument");
    return ST_ARG;
  }
  if (!fix_atoms(a, fa, "gfmd_fix_pre_force")) return ST_ARG;
  if (!s->fix_ready) {
    set_error("gfmd_fix_pre_force: call gfmd_fix_setup first");
    return ST_STATE;
  }
  launch_fix_com(fa, s->fix_frame, false, s->fix_part, s->fix_st, s->stream);
  launch_fix_gather(fa, s->fix_frame, is_fp64(s), u_dev, v_grid_dev,
                    s->fix_part, s->fix_st, s->stream);
  GFMD_HIP_TRY(hipGetLastError());
  return ST_OK;
}

int gfmd_fix_post_force(gfmd_solver_t *s, const gfmd_atoms_t *a,
                        const void *f_grid_dev, double *f_atoms_dev,
                        const int fix_center[3]) {
  FixAtoms fa;
  if (!s || !f_grid_dev || !f_atoms_dev) {
    set_error("gfmd_fix_post_force: invalid argument");
    return ST_ARG;
  }
  if (!fix_atoms(a, fa, "gfmd_fix_post_force")) return ST_ARG;
  if (!s->fix_ready) {
    set_error("gfmd_fix_post_force: call gfmd_fix_setup first");
    return ST_STATE;
  }
  const int zero[3] = {0, 0, 0};
  launch_fix_scatter(fa, s->fix_frame, is_fp64(s), f_grid_dev, f_atoms_dev,
                     fix_center ? fix_center : zero, s->fix_part, s->fix_st,
                     s->stream);
  GFMD_HIP_TRY(hipGetLastError());
  return ST_OK;
}

int gfmd_fix_state(gfmd_solver_t *s, gfmd_fix_state_t *out) {
  if (!s || !out) {
    set_error("gfmd_fix_state: invalid argument");
    return ST_ARG;
  }
  if (!s->fix_ready) {
    set_error("gfmd_fix_state: call gfmd_fix_setup first");
    return ST_STATE;
  }
  double st[FS_COUNT];
  GFMD_HIP_TRY(hipMemcpyAsync(st, s->fix_st, sizeof st, hipMemcpyDeviceToHost,
                              s->stream));
  GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  for (int k = 0; k < 3; k++) {
    out->xeqcm[k] = st[FS_XEQCM + k];
    out->xcm[k] = st[FS_XCM + k];
    out->fsum[k] = st[FS_FSUM + k];
  }
  out->xshift = (int)st[FS_XSHIFT];
  out->yshift = (int)st[FS_YSHIFT];
  out->placed = (int)st[FS_PLACED];
  out->unplaced = (int)st[FS_UNPLACED];
  out->unforced = (int)st[FS_UNFORCED];
  out->masstotal = st[FS_MASSTOTAL];
  out->natoms = (int)st[FS_NATOMS];
  return ST_OK;
}

// ---------------------------------------------------------------------------
// Preconditioned gradient: gP = F^-1 [ (Phi~(q) + cavg)^-1 F[g] ].
// ---------------------------------------------------------------------------
static int prec_prepare(gfmd_solver_t *s, const double *cavg,
                        const char *who) {
  if (!s->has_phi) {
    set_error("%s: no stiffness kernel set", who);
    return ST_STATE;
  }
  const int d = s->d;
  std::vector<double> c((size_t)d * d, 0.0);
  if (cavg) c.assign(cavg, cavg + (size_t)d * d);
  for (int i = 0; i < d; i++)
    for (int j = 0; j < i; j++)
      if (c[i * d + j] != c[j * d + i]) {
        set_error("%s: cavg must be symmetric (entries %d,%d and %d,%d "
                  "differ)", who, i, j, j, i);
        return ST_ARG;
      }
  if (s->prec_valid && c == s->prec_key) return ST_OK;
  int rc;
  if (!s->prec &&
      (rc = dev_alloc(s, &s->prec,
                      (size_t)s->nky * d * d * s->nx * real_size(s))))
    return rc;
  if (!s->prec_c &&
      (rc = dev_alloc(s, (void **)&s->prec_c, (size_t)d * d * 8)))
    return rc;
  GFMD_HIP_TRY(hipMemcpyAsync(s->prec_c, c.data(), c.size() * 8,
                              hipMemcpyHostToDevice, s->stream));
  GFMD_HIP_TRY(hipMemsetAsync(s->status, 0, sizeof(int), s->stream));
  rc = launch_compliance(s->phi, s->nx, s->ny, d, is_fp64(s), 2, nullptr,
                         s->prec, s->status, s->stream, s->prec_c);
  if (rc) {
    set_error("%s: no preconditioner kernel for ndof %d", who, d);
    return rc;
  }
  int st = 0;
  GFMD_HIP_TRY(hipMemcpyAsync(&st, s->status, sizeof(int),
                              hipMemcpyDeviceToHost, s->stream));
  GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  if (st) {
    set_error("%s: Phi(q) + cavg is singular", who);
    return ST_KERNEL;
  }
  // the table keeps Phi's x -> -x symmetry when cavg does not couple an x
  // component with a y or z component
  bool mir = s->phi_mirror;
  for (int i = 0; i < d && mir; i++)
    for (int j = 0; j < d; j++)
      if (((i % 3 == 0) != (j % 3 == 0)) && c[i * d + j] != 0.0) mir = false;
  s->prec_mirror = mir;
  s->prec_key = c;
  s->prec_valid = true;
  return ST_OK;
}

static int run_prec(gfmd_solver_t *s, const void *g, void *gp,
                    bool io_double) {
  ForceArgs a = make_args(s, g, gp);
  a.phi = s->prec;
  a.linf = nullptr;
  a.mirror = s->prec_mirror ? 1 : 0;
  const bool f64 = is_fp64(s);
  int rc = s->fast ? launch_force_fast(a, f64, io_double, s->stream)
                   : launch_force_generic(a, f64, io_double, s->stream);
  if (rc) {
    set_error("gfmd_prec_gradient: launch failed (status %d)", rc);
    return rc;
  }
  GFMD_HIP_TRY(hipGetLastError());
  return ST_OK;
}

int gfmd_prec_gradient(gfmd_solver_t *s, const double *cavg, const double *g,
                       double *gp) {
  if (!s || !g || !gp) {
    set_error("gfmd_prec_gradient: invalid argument");
    return ST_ARG;
  }
  if (g == gp) {
    set_error("gfmd_prec_gradient: g and gP must not alias");
    return ST_ARG;
  }
  int rc = prec_prepare(s, cavg, "gfmd_prec_gradient");
  if (rc) return rc;
  const size_t n = (size_t)s->d * s->nx * s->ny;
  GFMD_HIP_TRY(hipMemcpyAsync(s->u_io, g, n * 8, hipMemcpyHostToDevice,
                              s->stream));
  if ((rc = run_prec(s, s->u_io, s->f_io, true))) return rc;
  GFMD_HIP_TRY(hipMemcpyAsync(gp, s->f_io, n * 8, hipMemcpyDeviceToHost,
                              s->stream));
  GFMD_HIP_TRY(hipStreamSynchronize(s->stream));
  return ST_OK;
}

int gfmd_prec_gradient_device(gfmd_solver_t *s, const double *cavg,
                              const void *g_dev, void *gp_dev) {
  if (!s || !g_dev || !gp_dev || g_dev == gp_dev) {
    set_error("gfmd_prec_gradient_device: invalid argument");
    return ST_ARG;
  }
  int rc = prec_prepare(s, cavg, "gfmd_prec_gradient_device");
  if (rc) return rc;
  return run_prec(s, g_dev, gp_dev, is_fp64(s));
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
  if (fast_path)
    *fast_path = s->fast ? 1 : force_mixed_radix(s->nx, s->ny) ? 2 : 0;
  return ST_OK;
}

int gfmd_phi_entries(gfmd_solver_t *s, int *entries) {
  if (!s || !entries) {
    set_error("gfmd_phi_entries: NULL argument");
    return ST_ARG;
  }
  *entries = (s->fast && s->phi_par && s->d == 3) ? 6 : s->d * s->d;
  return ST_OK;
}

int gfmd_phi_mirror(gfmd_solver_t *s, int *mirror) {
  if (!s || !mirror) {
    set_error("gfmd_phi_mirror: NULL argument");
    return ST_ARG;
  }
  *mirror = s->phi_mirror ? 1 : 0;
  return ST_OK;
}

size_t gfmd_memory_usage(gfmd_solver_t *s) { return s ? s->bytes : 0; }

}  // extern "C"
