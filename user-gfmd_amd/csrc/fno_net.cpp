// Learned warm start: the Fourier neural operator's device state, forward,
// backward and Adam step, and its C ABI (gfmd_fno_*).  The architecture and
// the flat parameter layout are fixed in fno.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gfmd_hip.h"
#include "fno.h"
#include "gfmd_internal.h"

using namespace gfmd;
using namespace gfmd::fno;

void gfmd_set_error(const char *fmt, ...);

#define FNO_TRY(x)                                                         \
  do {                                                                     \
    hipError_t e_ = (x);                                                   \
    if (e_ != hipSuccess) {                                                \
      gfmd_set_error("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_),   \
                     __FILE__, __LINE__);                                  \
      return ST_HIP;                                                       \
    }                                                                      \
  } while (0)

struct gfmd_fno {
  int nx = 0, ny = 0, bmax = 0;
  size_t npmax = 0;
  hipStream_t st = nullptr;
  float *prm = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
  int step = 0;
  bf16 *Wb[L] = {};
  bf16 *Dy = nullptr, *Ax = nullptr, *Axi = nullptr, *Ey = nullptr;
  float *a = nullptr, *t = nullptr;           // inputs, targets
  bf16 *vb[L + 1] = {};                       // activations (BF16)
  float *z[L] = {};                           // pre-activations
  float *vL = nullptr, *acc = nullptr, *u = nullptr, *du = nullptr;
  float *dv = nullptr, *dz = nullptr;
  bf16 *dzb = nullptr;
  float *Zx[L] = {};                          // x-transform outputs
  float *sp = nullptr;                        // spectral temporaries
  bf16 *spb = nullptr;
  float *O = nullptr;
  bf16 *Ob = nullptr;
  float *gWpart = nullptr;
  double *part = nullptr, *loss = nullptr;
  std::vector<void *> allocs;
};

namespace {

int alloc(gfmd_fno_t *f, void **p, size_t bytes) {
  if (hipMalloc(p, bytes ? bytes : 8) != hipSuccess) {
    *p = nullptr;
    gfmd_set_error("gfmd_fno: out of device memory (%zu bytes)", bytes);
    return ST_HIP;
  }
  f->allocs.push_back(*p);
  return ST_OK;
}

uint16_t host_bf16(double x) {
  float f = (float)x;
  uint32_t u;
  memcpy(&u, &f, 4);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// DFT matrices of the truncated transforms (BF16): see the layout notes at
// gfmd_fno_create.
int build_dft(gfmd_fno_t *f) {
  const int nx = f->nx, ny = f->ny;
  const double tp = 6.283185307179586476925286766559;
  std::vector<uint16_t> dy((size_t)ROWS_Y * ny), ey((size_t)ny * ROWS_Y);
  for (int r = 0; r < NKY; r++)
    for (int y = 0; y < ny; y++) {
      const double th = tp * r * y / ny;
      dy[(size_t)r * ny + y] = host_bf16(cos(th));
      dy[(size_t)(NKY + r) * ny + y] = host_bf16(-sin(th));
      const double w = (r == 0 ? 1.0 : 2.0) / ((double)nx * ny);
      ey[(size_t)y * ROWS_Y + r] = host_bf16(w * cos(th));
      ey[(size_t)y * ROWS_Y + NKY + r] = host_bf16(-w * sin(th));
    }
  std::vector<uint16_t> ax((size_t)2 * NKX * 2 * nx), axi((size_t)2 * nx * 2 * NKX);
  for (int a = 0; a < NKX; a++) {
    const int kx = a < M ? a : nx - 2 * M + a;
    for (int x = 0; x < nx; x++) {
      const double th = tp * (double)kx * x / nx;
      const double c = cos(th), s = sin(th);
      const size_t K2 = 2 * nx;
      ax[(2 * a) * K2 + 2 * x] = host_bf16(c);
      ax[(2 * a) * K2 + 2 * x + 1] = host_bf16(s);
      ax[(2 * a + 1) * K2 + 2 * x] = host_bf16(-s);
      ax[(2 * a + 1) * K2 + 2 * x + 1] = host_bf16(c);
      const size_t K4 = 2 * NKX;
      axi[(2 * x) * K4 + 2 * a] = host_bf16(c);
      axi[(2 * x) * K4 + 2 * a + 1] = host_bf16(-s);
      axi[(2 * x + 1) * K4 + 2 * a] = host_bf16(s);
      axi[(2 * x + 1) * K4 + 2 * a + 1] = host_bf16(c);
    }
  }
  FNO_TRY(hipMemcpy(f->Dy, dy.data(), dy.size() * 2, hipMemcpyHostToDevice));
  FNO_TRY(hipMemcpy(f->Ey, ey.data(), ey.size() * 2, hipMemcpyHostToDevice));
  FNO_TRY(hipMemcpy(f->Ax, ax.data(), ax.size() * 2, hipMemcpyHostToDevice));
  FNO_TRY(hipMemcpy(f->Axi, axi.data(), axi.size() * 2, hipMemcpyHostToDevice));
  return ST_OK;
}

void gemm(gfmd_fno_t *f, int Mm, int Nn, int Kk, int batch, const bf16 *A,
          long long sA, long long rsA, long long csA, const bf16 *B,
          long long sB, long long rsB, long long csB, float *Cm, long long sC,
          long long rsC, long long csC, int accumulate) {
  Gemm g{Mm, Nn, Kk, batch, A, sA, rsA, csA, B, sB, rsB, csB,
         Cm, sC, rsC, csC, 1.0f, accumulate};
  launch_gemm(g, f->st);
}

void refresh_weights(gfmd_fno_t *f) {
  for (int l = 0; l < L; l++)
    launch_to_bf16(f->prm + off_W(l), f->Wb[l], C * C, f->st);
}

// Forward of `batch` samples whose inputs are in f->a; u in f->u.
void forward(gfmd_fno_t *f, int batch) {
  const int nx = f->nx, ny = f->ny;
  const size_t np = (size_t)batch * nx * ny;
  const int RW = NKY * C;  // 544: row length of the x-transform layouts
  launch_lift(f->a, f->prm, f->acc, f->vb[0], np, f->st);
  for (int l = 0; l < L; l++) {
    float *Zy = f->sp;
    bf16 *Zyb = f->spb;
    // y transform of every (sample, x) row: Zy = Dy (34 x ny) v (ny x C)
    gemm(f, ROWS_Y, C, ny, batch * nx, f->Dy, 0, ny, 1, f->vb[l],
         (long long)ny * C, C, 1, Zy, (long long)ROWS_Y * C, C, 1, 0);
    launch_to_bf16(Zy, Zyb, (size_t)batch * nx * ROWS_Y * C, f->st);
    // x transform: Zx = Ax (64 x 2nx) Zy (2nx x 544) per sample
    gemm(f, 2 * NKX, RW, 2 * nx, batch, f->Ax, 0, 2 * nx, 1, Zyb,
         (long long)2 * nx * RW, RW, 1, f->Zx[l], (long long)2 * NKX * RW, RW,
         1, 0);
    launch_mix(f->Zx[l], f->prm + off_R(l), f->O, batch, f->st);
    launch_to_bf16(f->O, f->Ob, (size_t)batch * 2 * NKX * RW, f->st);
    // inverse x: Yx = Axi (2nx x 64) O (64 x 544)
    float *Yx = f->sp;
    bf16 *Yxb = f->spb;
    gemm(f, 2 * nx, RW, 2 * NKX, batch, f->Axi, 0, 2 * NKX, 1, f->Ob,
         (long long)2 * NKX * RW, RW, 1, Yx, (long long)2 * nx * RW, RW, 1, 0);
    launch_to_bf16(Yx, Yxb, (size_t)batch * 2 * nx * RW, f->st);
    // inverse y: acc = Ey (ny x 34) Yx (34 x C) per (sample, x)
    gemm(f, ny, C, ROWS_Y, batch * nx, f->Ey, 0, ROWS_Y, 1, Yxb,
         (long long)ROWS_Y * C, C, 1, f->acc, (long long)ny * C, C, 1, 0);
    // pointwise channel mixing: acc += v W
    gemm(f, (int)np, C, C, 1, f->vb[l], 0, C, 1, f->Wb[l], 0, C, 1, f->acc,
         0, C, 1, 1);
    launch_bias_act(f->acc, f->prm + off_b(l), f->z[l],
                    l == L - 1 ? f->vL : nullptr, f->vb[l + 1], np, l < L - 1,
                    f->st);
  }
  launch_project(f->vL, f->prm, f->u, np, f->st);
}

// Loss against f->t and the gradients (f->grad) of the last forward.
void backward(gfmd_fno_t *f, int batch) {
  const int nx = f->nx, ny = f->ny;
  const size_t np = (size_t)batch * nx * ny;
  const int RW = NKY * C;
  (void)hipMemsetAsync(f->grad, 0, NPARAM * sizeof(float), f->st);
  launch_loss(f->u, f->t, f->du, f->part, f->loss, np * COUT, f->st);
  launch_project_back(f->vL, f->du, f->prm, f->dv, f->grad, f->part, np,
                      f->st);
  const int S = np >= 512 * 64 ? 512 : 1;  // split-K chunks of the dW sums
  const long long Lc = (long long)(np / S);
  for (int l = L - 1; l >= 0; l--) {
    launch_act_back(f->dv, f->z[l], f->dz, f->dzb, f->grad + off_b(l),
                    f->part, np, l < L - 1, f->st);
    // dW = v^T dz, in S chunks of the sites
    gemm(f, C, C, (int)Lc, S, f->vb[l], Lc * C, 1, C, f->dzb, Lc * C, C, 1,
         f->gWpart, C * C, C, 1, 0);
    if (np % S)  // sites left over by the chunks
      gemm(f, C, C, (int)(np - S * Lc), 1, f->vb[l] + S * Lc * C, 0, 1, C,
           f->dzb + S * Lc * C, 0, C, 1, f->gWpart, 0, C, 1, 1);
    launch_sum_rows(f->gWpart, S, C * C, f->grad + off_W(l), true, f->st);
    // dv = dz W^T (pointwise part)
    gemm(f, (int)np, C, C, 1, f->dzb, 0, C, 1, f->Wb[l], 0, 1, C, f->dv, 0,
         C, 1, 0);
    // spectral part, transposed products in reverse order
    float *dYx = f->sp;
    bf16 *dYxb = f->spb;
    gemm(f, ROWS_Y, C, ny, batch * nx, f->Ey, 0, 1, ROWS_Y, f->dzb,
         (long long)ny * C, C, 1, dYx, (long long)ROWS_Y * C, C, 1, 0);
    launch_to_bf16(dYx, dYxb, (size_t)batch * nx * ROWS_Y * C, f->st);
    gemm(f, 2 * NKX, RW, 2 * nx, batch, f->Axi, 0, 1, 2 * NKX, dYxb,
         (long long)2 * nx * RW, RW, 1, f->O, (long long)2 * NKX * RW, RW, 1,
         0);
    float *dZx = f->sp;
    launch_mix_back(f->Zx[l], f->O, f->prm + off_R(l), dZx,
                    f->grad + off_R(l), batch, f->st);
    launch_to_bf16(dZx, f->Ob, (size_t)batch * 2 * NKX * RW, f->st);
    float *dZy = f->sp;
    bf16 *dZyb = f->spb;
    gemm(f, 2 * nx, RW, 2 * NKX, batch, f->Ax, 0, 1, 2 * nx, f->Ob,
         (long long)2 * NKX * RW, RW, 1, dZy, (long long)2 * nx * RW, RW, 1,
         0);
    launch_to_bf16(dZy, dZyb, (size_t)batch * 2 * nx * RW, f->st);
    gemm(f, ny, C, ROWS_Y, batch * nx, f->Dy, 0, 1, ny, dZyb,
         (long long)ROWS_Y * C, C, 1, f->dv, (long long)ny * C, C, 1, 1);
  }
  launch_lift_back(f->a, f->dv, f->grad, f->part, np, f->st);
}

int check_batch(gfmd_fno_t *f, int batch) {
  if (!f || batch < 1 || batch > f->bmax) {
    gfmd_set_error("gfmd_fno: batch %d outside 1..%d", batch,
                   f ? f->bmax : 0);
    return ST_ARG;
  }
  return ST_OK;
}

}  // namespace

extern "C" {

size_t gfmd_fno_param_count(void) { return NPARAM; }

int gfmd_fno_create(int nx, int ny, int max_batch, gfmd_fno_t **out) {
  if (!out) return ST_ARG;
  *out = nullptr;
  if (nx < 2 * M || ny < 2 * M + 2 || max_batch < 1 ||
      (size_t)nx * ny * max_batch > ((size_t)1 << 30)) {
    gfmd_set_error("gfmd_fno_create: grid %d x %d (batch %d) unsupported: "
                   "nx >= %d, ny >= %d", nx, ny, max_batch, 2 * M, 2 * M + 2);
    return ST_ARG;
  }
  gfmd_fno_t *f = new gfmd_fno_t;
  f->nx = nx;
  f->ny = ny;
  f->bmax = max_batch;
  f->npmax = (size_t)nx * ny * max_batch;
  auto fail = [&](int rc) {
    for (void *p : f->allocs) (void)hipFree(p);
    if (f->st) (void)hipStreamDestroy(f->st);
    delete f;
    return rc;
  };
  if (hipStreamCreateWithFlags(&f->st, hipStreamNonBlocking) != hipSuccess) {
    gfmd_set_error("gfmd_fno_create: stream creation failed");
    return fail(ST_HIP);
  }
  const size_t np = f->npmax, RW = NKY * C;
  const size_t nsp = std::max((size_t)max_batch * 2 * nx * RW,
                              (size_t)max_batch * 2 * NKX * RW);
  int rc = ST_OK;
  auto A = [&](void **p, size_t bytes) {
    if (rc == ST_OK) rc = alloc(f, p, bytes);
  };
  A((void **)&f->prm, NPARAM * 4);
  A((void **)&f->grad, NPARAM * 4);
  A((void **)&f->am, NPARAM * 4);
  A((void **)&f->av, NPARAM * 4);
  for (int l = 0; l < L; l++) A((void **)&f->Wb[l], C * C * 2);
  A((void **)&f->Dy, (size_t)ROWS_Y * ny * 2);
  A((void **)&f->Ey, (size_t)ny * ROWS_Y * 2);
  A((void **)&f->Ax, (size_t)2 * NKX * 2 * nx * 2);
  A((void **)&f->Axi, (size_t)2 * nx * 2 * NKX * 2);
  A((void **)&f->a, np * CIN * 4);
  A((void **)&f->t, np * COUT * 4);
  for (int l = 0; l <= L; l++) A((void **)&f->vb[l], np * C * 2);
  for (int l = 0; l < L; l++) A((void **)&f->z[l], np * C * 4);
  for (int l = 0; l < L; l++)
    A((void **)&f->Zx[l], (size_t)max_batch * 2 * NKX * RW * 4);
  A((void **)&f->vL, np * C * 4);
  A((void **)&f->acc, np * C * 4);
  A((void **)&f->u, np * COUT * 4);
  A((void **)&f->du, np * COUT * 4);
  A((void **)&f->dv, np * C * 4);
  A((void **)&f->dz, np * C * 4);
  A((void **)&f->dzb, np * C * 2);
  A((void **)&f->sp, nsp * 4);
  A((void **)&f->spb, nsp * 2);
  A((void **)&f->O, (size_t)max_batch * 2 * NKX * RW * 4);
  A((void **)&f->Ob, (size_t)max_batch * 2 * NKX * RW * 2);
  A((void **)&f->gWpart, (size_t)512 * C * C * 4);
  A((void **)&f->part, (size_t)1024 * 128 * 8);
  A((void **)&f->loss, 8);
  if (rc) return fail(rc);
  if ((rc = build_dft(f))) return fail(rc);
  (void)hipMemsetAsync(f->prm, 0, NPARAM * 4, f->st);
  (void)hipMemsetAsync(f->am, 0, NPARAM * 4, f->st);
  (void)hipMemsetAsync(f->av, 0, NPARAM * 4, f->st);
  if (hipStreamSynchronize(f->st) != hipSuccess) return fail(ST_HIP);
  *out = f;
  return ST_OK;
}

int gfmd_fno_destroy(gfmd_fno_t *f) {
  if (!f) return ST_OK;
  (void)hipStreamSynchronize(f->st);
  for (void *p : f->allocs) (void)hipFree(p);
  (void)hipStreamDestroy(f->st);
  delete f;
  return ST_OK;
}

int gfmd_fno_set_params(gfmd_fno_t *f, const float *params) {
  if (!f || !params) return ST_ARG;
  FNO_TRY(hipMemcpyAsync(f->prm, params, NPARAM * 4, hipMemcpyHostToDevice,
                         f->st));
  refresh_weights(f);
  (void)hipMemsetAsync(f->am, 0, NPARAM * 4, f->st);
  (void)hipMemsetAsync(f->av, 0, NPARAM * 4, f->st);
  f->step = 0;
  FNO_TRY(hipStreamSynchronize(f->st));
  return ST_OK;
}

int gfmd_fno_get_params(gfmd_fno_t *f, float *params) {
  if (!f || !params) return ST_ARG;
  FNO_TRY(hipMemcpyAsync(params, f->prm, NPARAM * 4, hipMemcpyDeviceToHost,
                         f->st));
  FNO_TRY(hipStreamSynchronize(f->st));
  return ST_OK;
}

int gfmd_fno_forward_device(gfmd_fno_t *f, int batch, const float *in,
                            float *out) {
  int rc = check_batch(f, batch);
  if (rc) return rc;
  const size_t np = (size_t)batch * f->nx * f->ny;
  FNO_TRY(hipMemcpyAsync(f->a, in, np * CIN * 4, hipMemcpyDefault, f->st));
  forward(f, batch);
  FNO_TRY(hipMemcpyAsync(out, f->u, np * COUT * 4, hipMemcpyDefault, f->st));
  FNO_TRY(hipGetLastError());
  return ST_OK;
}

int gfmd_fno_forward(gfmd_fno_t *f, int batch, const float *in, float *out) {
  int rc = gfmd_fno_forward_device(f, batch, in, out);
  if (rc) return rc;
  FNO_TRY(hipStreamSynchronize(f->st));
  return ST_OK;
}

int gfmd_fno_gradients(gfmd_fno_t *f, int batch, const float *in,
                       const float *target, double *loss, float *grad) {
  int rc = check_batch(f, batch);
  if (rc) return rc;
  const size_t np = (size_t)batch * f->nx * f->ny;
  FNO_TRY(hipMemcpyAsync(f->a, in, np * CIN * 4, hipMemcpyDefault, f->st));
  FNO_TRY(hipMemcpyAsync(f->t, target, np * COUT * 4, hipMemcpyDefault,
                         f->st));
  forward(f, batch);
  backward(f, batch);
  FNO_TRY(hipGetLastError());
  if (loss)
    FNO_TRY(hipMemcpyAsync(loss, f->loss, 8, hipMemcpyDeviceToHost, f->st));
  if (grad)
    FNO_TRY(hipMemcpyAsync(grad, f->grad, NPARAM * 4, hipMemcpyDefault,
                           f->st));
  FNO_TRY(hipStreamSynchronize(f->st));
  return ST_OK;
}

int gfmd_fno_train_step(gfmd_fno_t *f, int batch, const float *in,
                        const float *target, float lr, double *loss) {
  int rc = check_batch(f, batch);
  if (rc) return rc;
  if (!(lr > 0)) return ST_ARG;
  const size_t np = (size_t)batch * f->nx * f->ny;
  FNO_TRY(hipMemcpyAsync(f->a, in, np * CIN * 4, hipMemcpyDefault, f->st));
  FNO_TRY(hipMemcpyAsync(f->t, target, np * COUT * 4, hipMemcpyDefault,
                         f->st));
  forward(f, batch);
  backward(f, batch);
  f->step++;
  launch_adam(f->prm, f->grad, f->am, f->av, NPARAM, lr, 0.9f, 0.999f, 1e-8f,
              f->step, f->st);
  refresh_weights(f);
  FNO_TRY(hipGetLastError());
  if (loss) {
    FNO_TRY(hipMemcpyAsync(loss, f->loss, 8, hipMemcpyDeviceToHost, f->st));
    FNO_TRY(hipStreamSynchronize(f->st));
  }
  return ST_OK;
}

int gfmd_fno_synchronize(gfmd_fno_t *f) {
  if (!f) return ST_ARG;
  FNO_TRY(hipStreamSynchronize(f->st));
  return ST_OK;
}

int gfmd_fno_gemm_test(int Mm, int Nn, int Kk, int batch, const uint16_t *A,
                       const long long *sa, const uint16_t *B,
                       const long long *sb, float *Cm, const long long *sc,
                       int accumulate) {
  // device pointers; sa/sb/sc: {batch, row, column} element strides
  Gemm g{Mm,    Nn,    Kk,    batch, A,   sa[0], sa[1], sa[2], B,
         sb[0], sb[1], sb[2], Cm,    sc[0], sc[1], sc[2], 1.0f, accumulate};
  launch_gemm(g, nullptr);
  FNO_TRY(hipGetLastError());
  FNO_TRY(hipDeviceSynchronize());
  return ST_OK;
}

}  // extern "C"
