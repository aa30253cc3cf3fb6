/*
 * gfmd_oracle.c -- CPU FP64 restatement of the reference's elastic-force hot
 * path (Atomistica/user-gfmd).  TEST INFRASTRUCTURE ONLY; see gfmd_oracle.h
 * for what pins each function.  The product library never uses this file.
 */
#include "gfmd_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------------------------------------------------------------------
 * FFT: Stockham autosort, mixed radix (4, 2, 3, 5, then any prime by a
 * direct O(r^2) butterfly), twiddles tabulated per plan from cos/sin of the
 * exact angle.  Plans are cached per (n, sign).
 * ------------------------------------------------------------------------- */

typedef struct {
  int n, sign, npass;
  int radix[64];
  double complex *tw[64]; /* per pass: Ns*r twiddles, [jm*r + k] */
  double complex *work;
} fft_plan_t;

#define PLAN_CACHE 16
static fft_plan_t *plan_cache[PLAN_CACHE];
static int plan_next = 0;

static void plan_free(fft_plan_t *p) {
  if (!p) return;
  for (int i = 0; i < p->npass; i++) free(p->tw[i]);
  free(p->work);
  free(p);
}

static fft_plan_t *plan_get(int n, int sign) {
  for (int i = 0; i < PLAN_CACHE; i++)
    if (plan_cache[i] && plan_cache[i]->n == n && plan_cache[i]->sign == sign)
      return plan_cache[i];
  fft_plan_t *p = (fft_plan_t *)calloc(1, sizeof(fft_plan_t));
  p->n = n;
  p->sign = sign;
  int m = n;
  while (m > 1) {
    int r;
    if (m % 4 == 0)
      r = 4;
    else if (m % 2 == 0)
      r = 2;
    else if (m % 3 == 0)
      r = 3;
    else if (m % 5 == 0)
      r = 5;
    else {
      r = 7;
      while (m % r) r += 2;
    }
    p->radix[p->npass++] = r;
    m /= r;
  }
  int Ns = 1;
  for (int ip = 0; ip < p->npass; ip++) {
    int r = p->radix[ip];
    p->tw[ip] = (double complex *)malloc(sizeof(double complex) * Ns * r);
    for (int jm = 0; jm < Ns; jm++)
      for (int k = 0; k < r; k++) {
        /* exp(sign * 2 pi i * jm * k / (Ns * r)); reduce the integer
           product first so the angle argument stays small and exact. */
        long long num = ((long long)jm * k) % ((long long)Ns * r);
        double a = sign * 2.0 * M_PI * (double)num / (double)(Ns * r);
        p->tw[ip][jm * r + k] = cos(a) + I * sin(a);
      }
    Ns *= r;
  }
  p->work = (double complex *)malloc(sizeof(double complex) * (n > 0 ? n : 1));
  if (plan_cache[plan_next]) plan_free(plan_cache[plan_next]);
  plan_cache[plan_next] = p;
  plan_next = (plan_next + 1) % PLAN_CACHE;
  return p;
}

static void dft_small(int r, int sign, double complex *v) {
  double complex out[64];
  double complex *o = out;
  double complex *heap = NULL;
  if (r > 64) heap = o = (double complex *)malloc(sizeof(double complex) * r);
  for (int k = 0; k < r; k++) {
    double complex s = 0;
    for (int m = 0; m < r; m++) {
      long long num = ((long long)m * k) % r;
      double a = sign * 2.0 * M_PI * (double)num / (double)r;
      s += v[m] * (cos(a) + I * sin(a));
    }
    o[k] = s;
  }
  memcpy(v, o, sizeof(double complex) * r);
  free(heap);
}

static void fft_exec(const fft_plan_t *p, double complex *x,
                     double complex *work);

void orc_fft_c2c(int n, double complex *x, int sign) {
  if (n <= 1) return;
  fft_plan_t *p = plan_get(n, sign);
  fft_exec(p, x, p->work);
}

/* The transform of plan p on x, with the caller's work buffer (n). */
static void fft_exec(const fft_plan_t *p, double complex *x,
                     double complex *work) {
  const int n = p->n, sign = p->sign;
  double complex *in = x, *out = work;
  int Ns = 1;
  for (int ip = 0; ip < p->npass; ip++) {
    int r = p->radix[ip];
    int stride = n / r;
    const double complex *tw = p->tw[ip];
    double complex v[64];
    double complex *vv = v;
    double complex *heap = NULL;
    if (r > 64) heap = vv = (double complex *)malloc(sizeof(double complex) * r);
    for (int j = 0; j < stride; j++) {
      int jm = j % Ns;
      for (int k = 0; k < r; k++) vv[k] = in[j + k * stride] * tw[jm * r + k];
      if (r == 2) {
        double complex a = vv[0], b = vv[1];
        vv[0] = a + b;
        vv[1] = a - b;
      } else if (r == 4) {
        double complex a0 = vv[0], a1 = vv[1], a2 = vv[2], a3 = vv[3];
        double complex si = sign * I;
        double complex t0 = a0 + a2, t1 = a0 - a2;
        double complex t2 = a1 + a3, t3 = si * (a1 - a3);
        vv[0] = t0 + t2;
        vv[1] = t1 + t3;
        vv[2] = t0 - t2;
        vv[3] = t1 - t3;
      } else {
        dft_small(r, sign, vv);
      }
      int base = (j / Ns) * Ns * r + jm;
      for (int k = 0; k < r; k++) out[base + k * Ns] = vv[k];
    }
    free(heap);
    double complex *t = in;
    in = out;
    out = t;
    Ns *= r;
  }
  if (in != x) memcpy(x, in, sizeof(double complex) * n);
}

void orc_fft2d(int nx, int ny, double complex *x, int sign) {
  for (int ix = 0; ix < nx; ix++) orc_fft_c2c(ny, x + (size_t)ix * ny, sign);
  double complex *col =
      (double complex *)malloc(sizeof(double complex) * (nx > 0 ? nx : 1));
  for (int iy = 0; iy < ny; iy++) {
    for (int ix = 0; ix < nx; ix++) col[ix] = x[(size_t)ix * ny + iy];
    orc_fft_c2c(nx, col, sign);
    for (int ix = 0; ix < nx; ix++) x[(size_t)ix * ny + iy] = col[ix];
  }
  free(col);
}

/* ---------------------------------------------------------------------------
 * Small dense complex linear algebra (row-major, m[i*dim+j]).
 * ------------------------------------------------------------------------- */

int orc_invert(int dim, double complex *m) {
  int indxr[64], indxc[64], piv[64];
  if (dim > 64) return -1;
  for (int i = 0; i < dim; i++) piv[i] = 0;
  for (int step = 0; step < dim; step++) {
    double big = -1.0;
    int irow = -1, icol = -1;
    for (int j = 0; j < dim; j++) {
      if (piv[j]) continue;
      for (int k = 0; k < dim; k++) {
        if (piv[k]) continue;
        double a = cabs(m[j * dim + k]);
        if (a > big) {
          big = a;
          irow = j;
          icol = k;
        }
      }
    }
    if (irow < 0 || big == 0.0) return -1;
    piv[icol] = 1;
    if (irow != icol)
      for (int l = 0; l < dim; l++) {
        double complex t = m[irow * dim + l];
        m[irow * dim + l] = m[icol * dim + l];
        m[icol * dim + l] = t;
      }
    indxr[step] = irow;
    indxc[step] = icol;
    double complex pinv = 1.0 / m[icol * dim + icol];
    m[icol * dim + icol] = 1.0;
    for (int l = 0; l < dim; l++) m[icol * dim + l] *= pinv;
    for (int r = 0; r < dim; r++) {
      if (r == icol) continue;
      double complex f = m[r * dim + icol];
      m[r * dim + icol] = 0.0;
      for (int l = 0; l < dim; l++) m[r * dim + l] -= m[icol * dim + l] * f;
    }
  }
  for (int l = dim - 1; l >= 0; l--) {
    if (indxr[l] == indxc[l]) continue;
    for (int k = 0; k < dim; k++) {
      double complex t = m[k * dim + indxr[l]];
      m[k * dim + indxr[l]] = m[k * dim + indxc[l]];
      m[k * dim + indxc[l]] = t;
    }
  }
  return 0;
}

static void matmul(int dim, const double complex *a, const double complex *b,
                   double complex *c) {
  for (int i = 0; i < dim; i++)
    for (int j = 0; j < dim; j++) {
      double complex s = 0;
      for (int k = 0; k < dim; k++) s += a[i * dim + k] * b[k * dim + j];
      c[i * dim + j] = s;
    }
}

/* ---------------------------------------------------------------------------
 * Stiffness kernels.
 * ------------------------------------------------------------------------- */

/* sc100: nearest and next-nearest neighbour central springs, k = 1, a = 1
   (reference sc100_stiffness.cpp:167-227). */
void orc_sc100_layer(double qx, double qy, double complex *U0,
                     double complex *U, double complex *V) {
  double cx = cos(qx), cy = cos(qy), sx = sin(qx), sy = sin(qy);
  for (int i = 0; i < 9; i++) U0[i] = U[i] = V[i] = 0;
  U[0] = 6 - 2 * cx * (1 + cy);
  U[4] = 6 - 2 * cy * (1 + cx);
  U[8] = 6;
  U[1] = U[3] = 2 * sx * sy;
  U0[0] = 5 - 2 * cx * (1 + cy);
  U0[4] = 5 - 2 * cy * (1 + cx);
  U0[8] = 3;
  U0[1] = U0[3] = 2 * sx * sy;
  V[0] = -cx;
  V[4] = -cy;
  V[8] = -1 - cx - cy;
  V[2] = I * sx;
  V[5] = I * sy;
  V[6] = I * sx;
  V[7] = I * sy;
}

/* Reference: greens_function_transfer_matrix_stiffness
   (surface_stiffness.cpp:811-873) and iterate_Gnn (493-548).
   X <- V (U + X)^-1 (-V^+), starting from X = V U^-1 (-V^+); the loop stops
   when X no longer changes (|dX|^2 <= eps^2) or after maxit passes. */
int orc_transfer_stiffness(int dim, int height, const double complex *U0,
                           const double complex *U, const double complex *V,
                           double complex *phi) {
  int dsq = dim * dim;
  double complex Vd[576], X[576], tmp[576], last[576];
  if (dim > 24) return -1;
  if (height == 0) {
    memcpy(phi, U0, sizeof(double complex) * dsq);
    return 0;
  }
  for (int i = 0; i < dim; i++)
    for (int j = 0; j < dim; j++) Vd[i * dim + j] = -conj(V[j * dim + i]);
  memcpy(X, U, sizeof(double complex) * dsq);
  if (orc_invert(dim, X)) return -1;
  matmul(dim, X, Vd, tmp);
  matmul(dim, V, tmp, X);

  double eps2;
  int maxit;
  if (height < 0) {
    eps2 = 1e-16;
    maxit = 100000;
  } else {
    eps2 = 0.0;
    maxit = height - 1;
  }
  memcpy(last, X, sizeof(double complex) * dsq);
  double dnorm = 1.0;
  int it = 0;
  while (dnorm > eps2 && it < maxit) {
    for (int i = 0; i < dsq; i++) X[i] += U[i];
    if (orc_invert(dim, X)) return -1;
    matmul(dim, X, Vd, tmp);
    matmul(dim, V, tmp, X);
    dnorm = 0.0;
    for (int i = 0; i < dsq; i++) {
      double complex dd = X[i] - last[i];
      dnorm += creal(dd) * creal(dd) + cimag(dd) * cimag(dd);
    }
    memcpy(last, X, sizeof(double complex) * dsq);
    it++;
  }
  if (height < 0 && it >= maxit) return -1;
  for (int i = 0; i < dsq; i++) phi[i] = U0[i] + X[i];
  return 0;
}

/* fcc100: nu single-atom layers (1, 2 or 4) with up to three neighbour
   shells (k, kappa), surface normal z = +-1.  Reference:
   FCC100StiffnessKernel::get_per_layer_dynamical_matrices
   (src/stiffness_kernels/fcc100_stiffness.cpp:275-487).  D receives the
   (nu+1)^2 per-layer 3x3 blocks in the reference's order (U0..U_{nu-1}, U,
   V0.., V, W0.., W, ...), xs/ys the layer shifts. */
void orc_fcc100_layer(int nu, const double *k, const double *kap, int z,
                      double qx, double qy, double complex *D, double *xs,
                      double *ys) {
  const double cx = cos(qx), cy = cos(qy), cx2 = cos(qx / 2), cy2 = cos(qy / 2);
  const double sx = sin(qx), sy = sin(qy), sx2 = sin(qx / 2), sy2 = sin(qy / 2);
  const double s2 = sqrt(2.0);
  const int nm = (nu + 1) * (nu + 1);
  for (int i = 0; i < nm * 9; i++) D[i] = 0;
  for (int i = 0; i < nu; i++) xs[i] = ys[i] = 0.0;
#define B(n) (D + 9 * (n))
  double complex *U0 = B(0), *U = B(nu), *V0 = B(nu + 1), *V = B(2 * nu + 1);
  /* nearest neighbours */
  U[0] = k[0] * (4 - 2 * cx) + kap[0] * (8 - 2 * cy);
  U[4] = k[0] * (4 - 2 * cy) + kap[0] * (8 - 2 * cx);
  U[8] = k[0] * 4 + kap[0] * (8 - 2 * cx - 2 * cy);
  U0[0] = U[0] - k[0] * 1 - kap[0] * 3;
  U0[4] = U[4] - k[0] * 1 - kap[0] * 3;
  U0[8] = U[8] - k[0] * 2 - kap[0] * 2;
  V[0] = -k[0] * cx2 * cy2 - kap[0] * 3 * cx2 * cy2;
  V[4] = -k[0] * cx2 * cy2 - kap[0] * 3 * cx2 * cy2;
  V[8] = -k[0] * 2 * cx2 * cy2 - kap[0] * 2 * cx2 * cy2;
  V[1] = V[3] = k[0] * sx2 * sy2 - kap[0] * sx2 * sy2;
  V[2] = V[6] = z * (k[0] * s2 * sx2 * cy2 - kap[0] * s2 * sx2 * cy2) * I;
  V[5] = V[7] = z * (k[0] * s2 * cx2 * sy2 - kap[0] * s2 * cx2 * sy2) * I;
  if (nu >= 2) {
    double complex *U1 = B(1), *V1 = B(nu + 2), *W0 = B(2 * nu + 2),
                   *W1 = B(2 * nu + 3), *W = B(3 * nu + 2);
    double complex U2n[9] = {0};
    U2n[0] = k[1] * (2 - 2 * cx * cy) + kap[1] * (4 - 2 * cx * cy) + k[2] * 8 +
             kap[2] * 16;
    U2n[4] = U2n[0];
    U2n[8] = k[1] * 2 + kap[1] * (4 - 4 * cx * cy) + k[2] * 8 + kap[2] * 16;
    U2n[1] = U2n[3] = k[1] * 2 * sx * sy - kap[1] * 2 * sx * sy;
    for (int i = 0; i < 9; i++) {
      U0[i] += U2n[i];
      U[i] += U2n[i];
    }
    for (int i = 0; i < 9; i++) U1[i] = U[i];
    U0[0] -= kap[1] + k[2] * 4 + kap[2] * 8;
    U0[4] -= kap[1] + k[2] * 4 + kap[2] * 8;
    U0[8] -= k[1] + k[2] * 4 + kap[2] * 8;
    U1[0] -= kap[1] + k[2] * 2 / 3 + kap[2] * 10 / 3;
    U1[4] -= kap[1] + k[2] * 2 / 3 + kap[2] * 10 / 3;
    U1[8] -= k[1] + k[2] * 8 / 3 + kap[2] * 4 / 3;
    V[0] += -k[2] * 1 / 3 * cx2 * cy2 * (-10 + 18 * cx + 2 * cy) -
            kap[2] * 1 / 3 * cx2 * cy2 * (-14 + 6 * cx + 22 * cy);
    V[4] += -k[2] * 1 / 3 * cx2 * cy2 * (-10 + 2 * cx + 18 * cy) -
            kap[2] * 1 / 3 * cx2 * cy2 * (-14 + 22 * cx + 6 * cy);
    V[8] += -k[2] * 2 / 3 * cx2 * cy2 * (-2 + 2 * (cx + cy)) -
            kap[2] * 2 / 3 * cx2 * cy2 * (-10 + 10 * (cx + cy));
    const double v01 = -k[2] * (-2 - 2 * (cx + cy)) * sx2 * sy2 -
                       kap[2] * (2 + 2 * (cx + cy)) * sx2 * sy2;
    V[1] += v01;
    V[3] += v01;
    const double complex v02 =
        z * (-k[2] * 1 / 3 * s2 * cy2 * (-2 - 2 * (3 * cx + cy)) * sx2 -
             kap[2] * 1 / 3 * s2 * cy2 * (2 + 2 * (3 * cx + cy)) * sx2) *
        I;
    const double complex v12 =
        z * (-k[2] * 1 / 3 * s2 * cx2 * (-2 - 2 * (cx + 3 * cy)) * sy2 -
             kap[2] * 1 / 3 * s2 * cx2 * (2 + 2 * (cx + 3 * cy)) * sy2) *
        I;
    V[2] += v02;
    V[6] += v02;
    V[5] += v12;
    V[7] += v12;
    W[0] = -kap[1] - k[2] * 2 / 3 * cx - kap[2] * (4 * cx / 3 + 2 * cy);
    W[4] = -kap[1] - k[2] * 2 / 3 * cy - kap[2] * (2 * cx + 4 * cy / 3);
    W[8] = -k[1] - k[2] * 4 / 3 * (cx + cy) - kap[2] * 2 / 3 * (cx + cy);
    W[2] = W[6] = z * (k[2] * 2 / 3 * s2 * sx - kap[2] * 2 / 3 * s2 * sx) * I;
    W[5] = W[7] = z * (k[2] * 2 / 3 * s2 * sy - kap[2] * 2 / 3 * s2 * sy) * I;
    for (int i = 0; i < 9; i++) {
      V1[i] = V[i];
      W0[i] = W[i];
      W1[i] = W[i];
    }
    if (nu > 2) {
      for (int i = 0; i < 9; i++) {
        B(2)[i] = U[i];
        B(nu + 3)[i] = V[i];
        B(2 * nu + 4)[i] = W[i];
      }
    }
    if (nu > 3) {
      for (int i = 0; i < 9; i++) {
        B(3)[i] = U[i];
        B(nu + 4)[i] = V[i];
        B(2 * nu + 5)[i] = W[i];
      }
    }
  }
  for (int i = 0; i < 9; i++) V0[i] = V[i];
  if (nu == 2 || nu == 4) {
    xs[1] = ys[1] = 0.5;
  }
  if (nu == 4) {
    xs[3] = ys[3] = 0.5;
  }
#undef B
}

/* Cell-level U0, U, V of dimension 3*nu from per-layer blocks with layer
   phase shifts.  Reference: StiffnessKernel::get_dynamical_matrices
   (src/main/surface_stiffness.cpp:127-292). */
void orc_assemble(int nu, const double complex *D, const double *xs,
                  const double *ys, double qx, double qy, double complex *U0,
                  double complex *U, double complex *V) {
  const int dim = 3 * nu, n1 = nu + 1;
  for (int i = 0; i < dim * dim; i++) U0[i] = U[i] = V[i] = 0;
  for (int i = 0; i < nu; i++)
    for (int j = i; j < nu; j++) {
      const double complex *m1 = D + 9 * ((j - i) * n1 + i);
      const double complex *m2 = D + 9 * ((j - i) * n1 + nu);
      if (i == j) {
        for (int k = 0; k < 3; k++)
          for (int l = 0; l < 3; l++) {
            U0[(k + 3 * i) * dim + l + 3 * j] = m1[k * 3 + l];
            U[(k + 3 * i) * dim + l + 3 * j] = m2[k * 3 + l];
          }
      } else {
        const double complex fac =
            cexp(I * ((xs[j] - xs[i]) * qx + (ys[j] - ys[i]) * qy));
        for (int k = 0; k < 3; k++)
          for (int l = 0; l < 3; l++) {
            U0[(k + 3 * i) * dim + l + 3 * j] = fac * m1[k * 3 + l];
            U[(k + 3 * i) * dim + l + 3 * j] = fac * m2[k * 3 + l];
            U0[(k + 3 * j) * dim + l + 3 * i] = conj(fac * m1[l * 3 + k]);
            U[(k + 3 * j) * dim + l + 3 * i] = conj(fac * m2[l * 3 + k]);
          }
      }
    }
  for (int i = 0; i < nu; i++)
    for (int j = 0; j <= i; j++) {
      const double complex *m1 = D + 9 * ((nu + j - i) * n1 + nu);
      const double complex fac =
          cexp(I * ((xs[j] - xs[i]) * qx + (ys[j] - ys[i]) * qy));
      for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++)
          V[(k + 3 * i) * dim + l + 3 * j] = fac * m1[k * 3 + l];
    }
}

/* Isotropic continuum half space: the closed-form surface compliance G(q)
   (Poisson number nu, shear modulus myu), Phi = G^-1; at q = 0 (|qx|, |qy|
   <= 1e-6) the diagonal compliance (alpha + 1/2, alpha + 1/2, 1) *
   xprd/(gamma pi).  Reference: IsotropicStiffnessKernel::
   get_stiffness_matrix (src/stiffness_kernels/isotropic_stiffness.cpp:
   95-167), invariant_ false as set by the base constructor
   (surface_stiffness.cpp:77). */
int orc_isotropic(double nu, double myu, double gamma, double xprd, double qx,
                  double qy, double complex *phi) {
  const double alpha = 1.0 / (2.0 * (1 - nu));
  const double b = sqrt(qx * qx + qy * qy);
  for (int i = 0; i < 9; i++) phi[i] = 0;
  if (fabs(qx) <= 1e-6 && fabs(qy) <= 1e-6) {
    phi[0] = phi[4] = (alpha + 0.5) * xprd / (gamma * M_PI);
    phi[8] = xprd / (gamma * M_PI);
  } else {
    const double c = (2.0 - 1.0 / alpha) / (2.0 * myu * b * b * b);
    const double s = (1.0 - 1.0 / alpha) / (2.0 * myu * b * b);
    phi[0] = 1.0 / (myu * b) - qx * qx * c;
    phi[4] = 1.0 / (myu * b) - qy * qy * c;
    phi[8] = 1.0 / (2.0 * myu * alpha * b);
    phi[1] = phi[3] = -qx * qy * c;
    phi[2] = -I * qx * s;
    phi[5] = -I * qy * s;
    phi[6] = I * qx * s;
    phi[7] = I * qy * s;
  }
  return orc_invert(3, phi);
}

/* ---------------------------------------------------------------------------
 * Li & Berger, J. Elast. 63, 137 (2001): surface displacement of an
 * isotropic half space (shear modulus G, Poisson number nu) at (Bx, By)
 * from uniform tractions (P normal, Qx, Qy) on a triangle, by the corner
 * decomposition.  Sign convention of the reference (li_berger.cpp:19-24):
 * negative P pushes down, negative uz is into the surface.  Reference:
 * tr_uniform_PQ_c / tr_PQ / sq_uniform_PQ (li_berger.cpp:231-304, 453-459)
 * and the plane-geometry predicates (geometry.cpp:42-142).
 * ------------------------------------------------------------------------- */
static double lb_angle(double x1, double y1, double x2, double y2) {
  /* signed angle from (x1, y1) to (x2, y2), clockwise positive */
  const double l = sqrt(x1 * x1 + y1 * y1) * sqrt(x2 * x2 + y2 * y2);
  return atan2((x2 * y1 - x1 * y2) / l, (x1 * x2 + y1 * y2) / l);
}

static int lb_cross(double a1x, double a1y, double a2x, double a2y,
                    double b1x, double b1y, double b2x, double b2y) {
  /* do segments a1-a2 and b1-b2 intersect (touching counts) */
  double c1 = (a1x - a2x) * (b1y - a2y) - (a1y - a2y) * (b1x - a2x);
  double c2 = (a1x - a2x) * (b2y - a2y) - (a1y - a2y) * (b2x - a2x);
  if (c1 * c2 > 0) return 0;
  c1 = (b1x - b2x) * (a1y - b2y) - (b1y - b2y) * (a1x - b2x);
  c2 = (b1x - b2x) * (a2y - b2y) - (b1y - b2y) * (a2x - b2x);
  return !(c1 * c2 > 0);
}

static int lb_inside(double bx, double by, double a1x, double a1y, double a2x,
                     double a2y, double a3x, double a3y) {
  /* strictly inside the triangle a1 a2 a3 */
  const double c1 = (a1x - a2x) * (by - a2y) - (a1y - a2y) * (bx - a2x);
  const double c2 = (a2x - a3x) * (by - a3y) - (a2y - a3y) * (bx - a3x);
  const double c3 = (a3x - a1x) * (by - a1y) - (a3y - a1y) * (bx - a1x);
  return c1 * c2 > 0 && c2 * c3 > 0 && c3 * c1 > 0;
}

static int lb_overlap(double bx, double by, double cx, double cy, double a1x,
                      double a1y, double a2x, double a2y) {
  /* do the triangles (b, a1, a2) and (c, a1, a2) overlap */
  if (lb_inside(bx, by, cx, cy, a1x, a1y, a2x, a2y) ||
      lb_inside(cx, cy, bx, by, a1x, a1y, a2x, a2y))
    return 1;
  return lb_cross(a1x, a1y, bx, by, a2x, a2y, cx, cy) ||
         lb_cross(a2x, a2y, bx, by, a1x, a1y, cx, cy);
}

static double lb_ltan(double phi) { return log(fabs(tan(0.5 * phi))); }

/* load on the triangle (b, a2, a3) whose corner b is the field point */
static void lb_corner(double G, double nu, double P, double Qx, double Qy,
                      double bx, double by, double a2x, double a2y, double a3x,
                      double a3y, double *u) {
  const double area =
      0.5 * fabs((a2x - bx) * (a3y - by) - (a3x - bx) * (a2y - by));
  if (area <= 0.0) return;
  const double l23 = sqrt((a2x - a3x) * (a2x - a3x) + (a2y - a3y) * (a2y - a3y));
  const double n1 = 2 * area / l23;
  const double d = -lb_angle(1.0, 0.0, a3x - a2x, a3y - a2y);
  const double e1 = lb_angle(a2x - a3x, a2y - a3y, bx - a2x, by - a2y);
  const double e2 = lb_angle(a2x - a3x, a2y - a3y, bx - a3x, by - a3y);
  const double sd = sin(d), cd = cos(d), s2d = sin(2 * d), c2d = cos(2 * d);
  /* the integrals of Li & Berger used by a uniform load */
  const double I100 = n1 * (lb_ltan(e1) - lb_ltan(e2));
  const double I201 = -n1 * ((sd * log(fabs(sin(e1))) - e1 * cd) -
                             (sd * log(fabs(sin(e2))) - e2 * cd));
  const double I210 = -n1 * ((cd * log(fabs(sin(e1))) + e1 * sd) -
                             (cd * log(fabs(sin(e2))) + e2 * sd));
#define F311(p) (s2d * lb_ltan(p) + 2 * s2d * cos(p) - 2 * c2d * sin(p))
#define F302(p) (sd * sd * lb_ltan(p) - cos(2 * d - (p)))
#define F320(p) (cd * cd * lb_ltan(p) + cos(2 * d - (p)))
  const double I311 = 0.5 * n1 * (F311(e1) - F311(e2));
  const double I302 = n1 * (F302(e1) - F302(e2));
  const double I320 = n1 * (F320(e1) - F320(e2));
#undef F311
#undef F302
#undef F320
  const double pi = M_PI;
  u[0] += P * (1 - 2 * nu) / (4 * pi * G) * I210;
  u[1] += P * (1 - 2 * nu) / (4 * pi * G) * I201;
  u[2] += P * (1 - nu) / (2 * pi * G) * I100;
  u[0] += Qx * ((1 - nu) / (2 * pi * G) * I100 + nu / (2 * pi * G) * I320);
  u[1] += Qx * nu / (2 * pi * G) * I311;
  u[2] += -Qx * (1 - 2 * nu) / (4 * pi * G) * I210;
  u[0] += Qy * nu / (2 * pi * G) * I311;
  u[1] += Qy * ((1 - nu) / (2 * pi * G) * I100 + nu / (2 * pi * G) * I302);
  u[2] += -Qy * (1 - 2 * nu) / (4 * pi * G) * I201;
}

/* triangle a1 a2 a3 as three corner triangles at b, each with the sign of
   its overlap with the load triangle */
static void lb_triangle(double G, double nu, double P, double Qx, double Qy,
                        const double *A, double bx, double by, double *u) {
  for (int k = 0; k < 3; k++) {
    const int i = k, j = (k + 1) % 3, o = (k + 2) % 3;
    const int s = lb_overlap(bx, by, A[2 * o], A[2 * o + 1], A[2 * i],
                             A[2 * i + 1], A[2 * j], A[2 * j + 1])
                      ? 1
                      : -1;
    lb_corner(G, nu, s * P, s * Qx, s * Qy, bx, by, A[2 * i], A[2 * i + 1],
              A[2 * j], A[2 * j + 1], u);
  }
}

void orc_sq_uniform_pq(double G, double nu, double P, double Qx, double Qy,
                       double a, double b, double bx, double by, double *u) {
  /* the rectangle [-a, a] x [-b, b] as two triangles along its diagonal */
  const double t1[6] = {-a, -b, a, -b, a, b};
  const double t2[6] = {-a, -b, -a, b, a, b};
  lb_triangle(G, nu, P, Qx, Qy, t1, bx, by, u);
  lb_triangle(G, nu, P, Qx, Qy, t2, bx, by, u);
}

int orc_kernel_phi(int kernel, const double *params, double qx, double qy,
                   double complex *phi) {
  if (kernel == ORC_KERNEL_ISOTROPIC)
    return orc_isotropic(params[0], params[1], params[2], params[3], qx, qy,
                         phi);
  if (kernel == ORC_KERNEL_SC100) {
    double complex U0[9], U[9], V[9];
    orc_sc100_layer(qx, qy, U0, U, V);
    return orc_transfer_stiffness(3, (int)params[0], U0, U, V, phi);
  }
  if (kernel == ORC_KERNEL_FCC100) {
    /* params: height, nu, k1, k2, k3, kappa1, kappa2, kappa3, normal */
    const int nu = (int)params[1];
    if (nu < 1 || nu > 4) return -1;
    double complex D[25 * 9], U0[144], U[144], V[144];
    double xs[4], ys[4];
    orc_fcc100_layer(nu, params + 2, params + 5, (int)params[8], qx, qy, D,
                     xs, ys);
    orc_assemble(nu, D, xs, ys, qx, qy, U0, U, V);
    return orc_transfer_stiffness(3 * nu, (int)params[0], U0, U, V, phi);
  }
  return -1;
}

/* Reference: fill_phi_buffer (gfmd_misc.cpp:32-102), single rank. */
int orc_fill_phi(int kernel, const double *params, int d, int nx, int ny,
                 int normalize, double complex *phi) {
  int dsq = d * d;
  double inv = 1.0 / ((double)nx * (double)ny);
  for (int i = 0; i < nx; i++) {
    double qx = (i <= nx / 2) ? 2.0 * M_PI * i / nx : 2.0 * M_PI * (i - nx) / nx;
    for (int j = 0; j < ny; j++) {
      double qy =
          (j <= ny / 2) ? 2.0 * M_PI * j / ny : 2.0 * M_PI * (j - ny) / ny;
      double complex *p = phi + ((size_t)i * ny + j) * dsq;
      if (orc_kernel_phi(kernel, params, qx, qy, p)) return -1;
      /* Hermiticity check with the reference's tolerance (1e-6). */
      for (int a = 0; a < d; a++)
        for (int b = 0; b < d; b++)
          if (cabs(p[a * d + b] - conj(p[b * d + a])) > 1e-6) return -1;
      if (normalize)
        for (int a = 0; a < dsq; a++) p[a] *= inv;
    }
  }
  return 0;
}

/* ---------------------------------------------------------------------------
 * Static force evaluation.  Reference: GFMDSolverStatic::post_force
 * (gfmd_solver_static.cpp:145-249) with GFMDSolverFFT::fft_forward/
 * fft_reverse (gfmd_solver_fft.cpp:96-195): per component a full complex
 * 2D FFT (forward exp(-i q r)), F = Phi~ u(q) and the energy sum over the
 * full grid, Gamma-point linf terms, unnormalised inverse FFT, real part.
 * ------------------------------------------------------------------------- */
double orc_static_force(int nx, int ny, int d, const double complex *phi,
                        const double *linf, const double *u, double *f,
                        double *u0) {
  size_t nxy = (size_t)nx * ny;
  int nu = d / 3;
  double complex *q = (double complex *)malloc(sizeof(double complex) * nxy * d);
  double complex *buf = (double complex *)malloc(sizeof(double complex) * nxy);
  /* fft_forward: q_buffer[idq][idim] (AoS, idim fastest) */
  for (int c = 0; c < d; c++) {
    for (size_t i = 0; i < nxy; i++) buf[i] = u[c * nxy + i];
    orc_fft2d(nx, ny, buf, -1);
    for (size_t i = 0; i < nxy; i++) q[i * d + c] = buf[i];
  }
  if (u0)
    for (int c = 0; c < d; c++) u0[c] = creal(q[c]);

  double epot = 0.0;
  if (linf)
    for (int i = 0; i < nu; i++) epot -= 2 * linf[i] * creal(q[3 * i + 2]);

  double complex F[24];
  for (size_t idq = 0; idq < nxy; idq++) {
    const double complex *p = phi + idq * d * d;
    double complex *qq = q + idq * d;
    for (int i = 0; i < d; i++) {
      double complex s = 0;
      for (int j = 0; j < d; j++) s += p[i * d + j] * qq[j];
      F[i] = s;
    }
    for (int i = 0; i < d; i++) {
      epot += creal(conj(F[i]) * qq[i]);
      qq[i] = -F[i];
    }
  }
  if (linf)
    for (int i = 0; i < nu; i++) q[3 * i + 2] += linf[i];
  epot *= 0.5;

  for (int c = 0; c < d; c++) {
    for (size_t i = 0; i < nxy; i++) buf[i] = q[i * d + c];
    orc_fft2d(nx, ny, buf, +1);
    for (size_t i = 0; i < nxy; i++) f[c * nxy + i] = creal(buf[i]);
  }
  free(q);
  free(buf);
  return epot;
}


/* ---------------------------------------------------------------------------
 * Timing baseline only: orc_static_force with the rows, columns and
 * wavevectors of each step split over OpenMP threads (the same FP64
 * algorithm: per-component full complex 2D transforms, AoS q buffer,
 * full-grid Phi, mat_mul_vec loop).  Plans are built before the parallel
 * regions and only read inside; every thread has its own work buffers.
 * Not used by the tests (the energy sum's order depends on the threads).
 * ------------------------------------------------------------------------- */
static void fft2d_mt(int nx, int ny, double complex *x, const fft_plan_t *py,
                     const fft_plan_t *px) {
#pragma omp parallel
  {
    double complex *w = (double complex *)malloc(
        sizeof(double complex) * (size_t)(nx > ny ? nx : ny) * 2);
    double complex *col = w + (nx > ny ? nx : ny);
#pragma omp for schedule(static)
    for (int ix = 0; ix < nx; ix++)
      if (ny > 1) fft_exec(py, x + (size_t)ix * ny, w);
#pragma omp for schedule(static)
    for (int iy = 0; iy < ny; iy++) {
      for (int ix = 0; ix < nx; ix++) col[ix] = x[(size_t)ix * ny + iy];
      if (nx > 1) fft_exec(px, col, w);
      for (int ix = 0; ix < nx; ix++) x[(size_t)ix * ny + iy] = col[ix];
    }
    free(w);
  }
}

double orc_static_force_mt(int nx, int ny, int d, const double complex *phi,
                           const double *u, double *f, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const size_t nxy = (size_t)nx * ny;
  const fft_plan_t *fy = plan_get(ny, -1), *fx = plan_get(nx, -1);
  const fft_plan_t *by = plan_get(ny, +1), *bx = plan_get(nx, +1);
  double complex *q = (double complex *)malloc(sizeof(double complex) * nxy * d);
  double complex *buf = (double complex *)malloc(sizeof(double complex) * nxy);
  for (int c = 0; c < d; c++) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < nxy; i++) buf[i] = u[c * nxy + i];
    fft2d_mt(nx, ny, buf, fy, fx);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < nxy; i++) q[i * d + c] = buf[i];
  }
  double epot = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : epot)
  for (size_t idq = 0; idq < nxy; idq++) {
    double complex F[24];
    const double complex *p = phi + idq * d * d;
    double complex *qq = q + idq * d;
    for (int i = 0; i < d; i++) {
      double complex s = 0;
      for (int j = 0; j < d; j++) s += p[i * d + j] * qq[j];
      F[i] = s;
    }
    for (int i = 0; i < d; i++) {
      epot += creal(conj(F[i]) * qq[i]);
      qq[i] = -F[i];
    }
  }
  epot *= 0.5;
  for (int c = 0; c < d; c++) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < nxy; i++) buf[i] = q[i * d + c];
    fft2d_mt(nx, ny, buf, by, bx);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < nxy; i++) f[c * nxy + i] = creal(buf[i]);
  }
  free(q);
  free(buf);
  return epot;
}
