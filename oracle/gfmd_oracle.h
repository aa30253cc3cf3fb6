/*
 * gfmd_oracle -- CPU FP64 restatement of the reference's elastic-force hot
 * path.  TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the CPU
 * baseline leg of the benchmark use it as the reference result.  The product library
 * (libgfmd_hip.so) never links, loads or calls it.
 *
 * Complex numbers are C99 `double complex`, i.e. interleaved (re, im) pairs,
 * so arrays are numpy complex128 compatible.
 *
 * Pinning (what each function is checked against in tests/):
 *   orc_fft_c2c / orc_fft2d  : numpy.fft and an O(N^2) brute-force DFT,
 *                              including odd and prime lengths.
 *   orc_sc100_layer, orc_fcc100_layer, orc_assemble,
 *   orc_transfer_stiffness,
 *   orc_fill_phi             : golden Phi(q) vectors written by the
 *                              reference's own C++ core (oracle/_ref), plus
 *                              height=0 => Phi=U0, height=1 closed form,
 *                              Hermiticity and Phi(-q)=conj(Phi(q)).
 *   orc_static_force         : plane-wave closed form, brute-force real-space
 *                              convolution on tiny grids, E = -1/2 sum u.f
 *                              identity, linearity, finite-difference
 *                              f = -dE/du, Gamma-point linf terms.

 * The FIRE state machine and the LJ sphere contact are restated in numpy in
 * oracle/oracle.py (Fire, contact_sphere).
 */
#ifndef GFMD_ORACLE_H
#define GFMD_ORACLE_H

#ifdef __cplusplus
typedef _Complex double orc_complex; /* GNU C++: C99 layout, (re, im) */
extern "C" {
#else
#include <complex.h>
typedef double complex orc_complex;
#endif

/* In-place unnormalised 1D DFT of length n with stride 1:
   X[k] = sum_j x[j] exp(sign * 2 pi i j k / n). Any n >= 1. */
void orc_fft_c2c(int n, orc_complex *x, int sign);

/* In-place unnormalised 2D DFT of an nx*ny grid, y fastest
   (element ix*ny + iy), exponent sign as above. */
void orc_fft2d(int nx, int ny, orc_complex *x, int sign);

/* Complex dim x dim matrix inverse by Gauss-Jordan with full pivoting.
   Returns 0, or -1 when singular. */
int orc_invert(int dim, orc_complex *m);
void orc_sq_uniform_pq(double G, double nu, double P, double Qx, double Qy,
                       double a, double b, double bx, double by, double *u);
int orc_isotropic(double nu, double myu, double gamma, double xprd, double qx,
                  double qy, orc_complex *phi);

/* sc100 per-layer dynamical matrices U0, U, V (3x3, row-major). */
void orc_sc100_layer(double qx, double qy, orc_complex *U0,
                     orc_complex *U, orc_complex *V);

/* Surface stiffness Phi = U0 + V_T from the Green's-function transfer
   matrix with zero-force boundary at depth `height` (height<0: semi-infinite,
   iterate to |dX|^2 < 1e-16; height=0: Phi = U0).  Returns 0 or -1. */
int orc_transfer_stiffness(int dim, int height, const orc_complex *U0,
                           const orc_complex *U, const orc_complex *V,
                           orc_complex *phi);

/* fcc100 per-layer blocks ((nu+1)^2 3x3 blocks, reference order) and
   layer shifts; nu = 1, 2 or 4; k[3], kappa[3] per shell; z = +-1. */
void orc_fcc100_layer(int nu, const double *k, const double *kappa, int z,
                      double qx, double qy, orc_complex *D, double *xs,
                      double *ys);

/* Cell-level U0, U, V (dimension 3*nu) from per-layer blocks. */
void orc_assemble(int nu, const orc_complex *D, const double *xs,
                  const double *ys, double qx, double qy, orc_complex *U0,
                  orc_complex *U, orc_complex *V);

/* Kernel ids for orc_kernel_phi / orc_fill_phi. */
#define ORC_KERNEL_SC100 1
#define ORC_KERNEL_FCC100 2
#define ORC_KERNEL_ISOTROPIC 3 /* params: nu, myu, gamma, xprd */

/* Phi(q) of a kernel.  sc100: params[0] = height (dim 3).  fcc100:
   params = height, nu, k1, k2, k3, kappa1, kappa2, kappa3, normal
   (dim 3 nu). */
int orc_kernel_phi(int kernel, const double *params, double qx, double qy,
                   orc_complex *phi);

/* Full-grid table phi[(ix*ny+iy)*d*d + i*d + j] with the reference's q
   convention and optional 1/(nx*ny) normalisation; returns 0 or -1. */
int orc_fill_phi(int kernel, const double *params, int d, int nx, int ny,
                 int normalize, orc_complex *phi);

/* One static force evaluation: u[idim*nxy + ix*ny + iy] (d components),
   writes f in the same layout and u0[d] = Re u(q=0); returns the energy.
   phi is the normalised full-grid table as above; linf[d/3] may be NULL. */
double orc_static_force(int nx, int ny, int d, const orc_complex *phi,
                        const double *linf, const double *u, double *f,
                        double *u0);
/* Timing baseline: the same evaluation with OpenMP threads (nthreads <= 0:
   the OpenMP default). */
double orc_static_force_mt(int nx, int ny, int d, const orc_complex *phi,
                           const double *u, double *f, int nthreads);

#ifdef __cplusplus
}
#endif

#endif
