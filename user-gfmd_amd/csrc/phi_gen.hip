// Surface stiffness tabulation Phi(q) on the device, one thread per
// wavevector of the half spectrum.  Output layout: T phi[nky][d*d][nx],
// Hermitian-packed (d diagonal reals, then Re/Im of Phi_ij for i < j in
// row-major order), scaled by 1/(nx*ny) when normalising.
//
// Reference semantics: fill_phi_buffer (src/main/gfmd_misc.cpp:32-102) with
// StiffnessKernel::get_stiffness_matrix -> greens_function_transfer_matrix_
// stiffness / iterate_Gnn (src/main/surface_stiffness.cpp:295-398, 493-548,
// 811-873) and the sc100 per-layer matrices
// (src/stiffness_kernels/sc100_stiffness.cpp:167-227).
#include "gfmd_internal.h"

namespace gfmd {

using Z = C2<double>;

__device__ __forceinline__ Z zmul(Z a, Z b) { return cmul(a, b); }
__device__ __forceinline__ double zabs2(Z a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ Z zdiv1(Z a) {  // 1/a
  const double s = 1.0 / zabs2(a);
  return cmake(a.x * s, -a.y * s);
}

// In-place inverse by Gauss-Jordan on [A | I] with partial pivoting; all
// indices are compile-time so the matrices stay in registers.
template <int D>
__device__ bool zinv(Z (&a)[D][D]) {
  Z b[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) b[i][j] = cmake(i == j ? 1.0 : 0.0, 0.0);
#pragma unroll
  for (int k = 0; k < D; k++) {
    int p = k;
    double best = zabs2(a[k][k]);
#pragma unroll
    for (int r = k + 1; r < D; r++) {
      const double v = zabs2(a[r][k]);
      if (v > best) {
        best = v;
        p = r;
      }
    }
    if (!(best > 0.0)) return false;
#pragma unroll
    for (int r = k + 1; r < D; r++) {
      if (p == r) {
#pragma unroll
        for (int j = 0; j < D; j++) {
          Z t = a[k][j];
          a[k][j] = a[r][j];
          a[r][j] = t;
          t = b[k][j];
          b[k][j] = b[r][j];
          b[r][j] = t;
        }
      }
    }
    const Z piv = zdiv1(a[k][k]);
#pragma unroll
    for (int j = 0; j < D; j++) {
      a[k][j] = zmul(a[k][j], piv);
      b[k][j] = zmul(b[k][j], piv);
    }
#pragma unroll
    for (int r = 0; r < D; r++) {
      if (r == k) continue;
      const Z fct = a[r][k];
#pragma unroll
      for (int j = 0; j < D; j++) {
        a[r][j] = a[r][j] - zmul(fct, a[k][j]);
        b[r][j] = b[r][j] - zmul(fct, b[k][j]);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) a[i][j] = b[i][j];
  return true;
}

template <int D>
__device__ __forceinline__ void zmm(const Z (&a)[D][D], const Z (&b)[D][D],
                                    Z (&c)[D][D]) {
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      Z s = cmake(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < D; k++) s = s + zmul(a[i][k], b[k][j]);
      c[i][j] = s;
    }
}

// Phi = U0 + V_T, V_T <- V (U + V_T)^-1 (-V^+) from V_T = V U^-1 (-V^+);
// height > 0: height-1 further passes (stopping early once V_T stops
// changing, as the reference's loop does), height < 0: until |dV_T|^2 <=
// 1e-16, height = 0: Phi = U0.
template <int D>
__device__ bool transfer_stiffness(int height, const Z (&U0)[D][D],
                                   const Z (&U)[D][D], const Z (&V)[D][D],
                                   Z (&phi)[D][D]) {
  if (height == 0) {
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) phi[i][j] = U0[i][j];
    return true;
  }
  Z Vd[D][D], X[D][D], tmp[D][D], VT[D][D];
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) {
      Vd[i][j] = cmake(-V[j][i].x, V[j][i].y);  // -conj(V^T)
      X[i][j] = U[i][j];
    }
  if (!zinv<D>(X)) return false;
  zmm<D>(X, Vd, tmp);
  zmm<D>(V, tmp, VT);
  const double eps2 = height < 0 ? 1e-16 : 0.0;
  const int maxit = height < 0 ? 100000 : height - 1;
  double dnorm = 1.0;
  int it = 0;
  while (dnorm > eps2 && it < maxit) {
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) X[i][j] = VT[i][j] + U[i][j];
    if (!zinv<D>(X)) return false;
    zmm<D>(X, Vd, tmp);
    dnorm = 0.0;
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++) {
        Z s = cmake(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < D; k++) s = s + zmul(V[i][k], tmp[k][j]);
        dnorm += zabs2(s - VT[i][j]);
        VT[i][j] = s;
      }
    it++;
  }
  if (height < 0 && it >= maxit) return false;
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = 0; j < D; j++) phi[i][j] = U0[i][j] + VT[i][j];
  return true;
}

__device__ void sc100_layer(double qx, double qy, Z (&U0)[3][3], Z (&U)[3][3],
                            Z (&V)[3][3]) {
  double sx, cx, sy, cy;
  sincos(qx, &sx, &cx);
  sincos(qy, &sy, &cy);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) U0[i][j] = U[i][j] = V[i][j] = cmake(0.0, 0.0);
  U[0][0].x = 6 - 2 * cx * (1 + cy);
  U[1][1].x = 6 - 2 * cy * (1 + cx);
  U[2][2].x = 6;
  U[0][1].x = U[1][0].x = 2 * sx * sy;
  U0[0][0].x = 5 - 2 * cx * (1 + cy);
  U0[1][1].x = 5 - 2 * cy * (1 + cx);
  U0[2][2].x = 3;
  U0[0][1].x = U0[1][0].x = 2 * sx * sy;
  V[0][0].x = -cx;
  V[1][1].x = -cy;
  V[2][2].x = -1 - cx - cy;
  V[0][2].y = sx;
  V[1][2].y = sy;
  V[2][0].y = sx;
  V[2][1].y = sy;
}

template <int D, typename T>
__device__ __forceinline__ void store_packed(const Z (&phi)[D][D], double scale,
                                             T *__restrict__ out, int nx,
                                             int kx) {
#pragma unroll
  for (int i = 0; i < D; i++) out[(size_t)i * nx + kx] = (T)(phi[i][i].x * scale);
  int p = D;
#pragma unroll
  for (int i = 0; i < D; i++)
#pragma unroll
    for (int j = i + 1; j < D; j++) {
      // Hermitian part: (Phi_ij + conj(Phi_ji)) / 2
      const double re = 0.5 * (phi[i][j].x + phi[j][i].x) * scale;
      const double im = 0.5 * (phi[i][j].y - phi[j][i].y) * scale;
      out[(size_t)p * nx + kx] = (T)re;
      out[(size_t)(p + 1) * nx + kx] = (T)im;
      p += 2;
    }
}

// Reference wavevector convention (gfmd_misc.cpp:42-47).
__device__ __forceinline__ double qcomp(int i, int n) {
  const double tp = 6.283185307179586476925286766559;
  return (i <= n / 2) ? tp * i / n : tp * (i - n) / n;
}

template <typename T>
__global__ void k_phi_sc100(int nx, int ny, int nky, int height, double scale,
                            T *__restrict__ phi, int *__restrict__ status) {
  const size_t n = (size_t)nky * nx;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < n;
       L += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(L % nx);
    const int ky = (int)(L / nx);
    Z U0[3][3], U[3][3], V[3][3], P[3][3];
    sc100_layer(qcomp(kx, nx), qcomp(ky, ny), U0, U, V);
    if (!transfer_stiffness<3>(height, U0, U, V, P)) {
      atomicOr(status, 1);
      continue;
    }
    store_packed<3>(P, scale, phi + (size_t)ky * 9 * nx, nx, kx);
  }
}

// Spectral filter for rough-surface generation: Phi = -A(q) I_3 with
// A(q) = sqrt(C(q)), C(q) = (q/q_l)^(-2(1+H)) for q_l < q <= q_s, C = 1 for
// q <= q_l (roll-off plateau), 0 above q_s and at q = 0; then
// f = -Phi~ u is the filtered field.  p: H, q_s, q_l, lattice constant.
template <typename T>
__global__ void k_phi_filter(int nx, int ny, int nky, double H, double qs,
                             double ql, double a, double scale,
                             T *__restrict__ phi) {
  const size_t n = (size_t)nky * nx;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < n;
       L += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(L % nx);
    const int ky = (int)(L / nx);
    const double qx = qcomp(kx, nx) / a, qy = qcomp(ky, ny) / a;
    const double q = sqrt(qx * qx + qy * qy);
    double A = 0.0;
    if (q > 0.0 && q <= qs) A = q <= ql ? 1.0 : pow(q / ql, -(1.0 + H));
    T *o = phi + (size_t)ky * 9 * nx + kx;
    for (int e = 0; e < 9; e++) o[(size_t)e * nx] = (T)(e < 3 ? -A * scale : 0.0);
  }
}

// full: [nx*ny][d][d] complex128 (reference fill_phi_buffer layout).
template <typename T>
__global__ void k_phi_pack(const double *__restrict__ full, int nx, int ny,
                           int nky, int d, double scale, T *__restrict__ phi) {
  const size_t n = (size_t)nky * nx;
  const int dd = d * d;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < n;
       L += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(L % nx);
    const int ky = (int)(L / nx);
    const double *m = full + ((size_t)kx * ny + ky) * dd * 2;
    T *o = phi + (size_t)ky * dd * nx + kx;
    for (int i = 0; i < d; i++) o[(size_t)i * nx] = (T)(m[2 * (i * d + i)] * scale);
    int p = d;
    for (int i = 0; i < d; i++)
      for (int j = i + 1; j < d; j++) {
        const double re = 0.5 * (m[2 * (i * d + j)] + m[2 * (j * d + i)]);
        const double im = 0.5 * (m[2 * (i * d + j) + 1] - m[2 * (j * d + i) + 1]);
        o[(size_t)p * nx] = (T)(re * scale);
        o[(size_t)(p + 1) * nx] = (T)(im * scale);
        p += 2;
      }
  }
}

static unsigned nblocks(size_t n) {
  size_t b = (n + 255) / 256;
  return (unsigned)(b < 65535 ? b : 65535);
}

int launch_phi_generate(const KernelDesc &k, int nx, int ny, int normalize,
                        bool fp64, void *phi, int *d_status, hipStream_t s) {
  const int nky = ny / 2 + 1;
  const double scale = normalize ? 1.0 / ((double)nx * (double)ny) : 1.0;
  const size_t n = (size_t)nky * nx;
  if (k.kind == KIND_SC100) {
    if (k.d != 3) return ST_ARG;
    if (fp64)
      hipLaunchKernelGGL(k_phi_sc100<double>, dim3(nblocks(n)), dim3(256), 0, s,
                         nx, ny, nky, k.height, scale, (double *)phi, d_status);
    else
      hipLaunchKernelGGL(k_phi_sc100<float>, dim3(nblocks(n)), dim3(256), 0, s,
                         nx, ny, nky, k.height, scale, (float *)phi, d_status);
    return ST_OK;
  }
  if (k.kind == KIND_FILTER) {
    if (k.d != 3) return ST_ARG;
    if (fp64)
      hipLaunchKernelGGL(k_phi_filter<double>, dim3(nblocks(n)), dim3(256), 0,
                         s, nx, ny, nky, k.p[0], k.p[1], k.p[2], k.p[3], scale,
                         (double *)phi);
    else
      hipLaunchKernelGGL(k_phi_filter<float>, dim3(nblocks(n)), dim3(256), 0,
                         s, nx, ny, nky, k.p[0], k.p[1], k.p[2], k.p[3], scale,
                         (float *)phi);
    return ST_OK;
  }
  return ST_UNSUPPORTED;
}

int launch_phi_pack(const double *full, int nx, int ny, int d, double scale,
                    bool fp64, void *phi, hipStream_t s) {
  const int nky = ny / 2 + 1;
  const size_t n = (size_t)nky * nx;
  if (fp64)
    hipLaunchKernelGGL(k_phi_pack<double>, dim3(nblocks(n)), dim3(256), 0, s,
                       full, nx, ny, nky, d, scale, (double *)phi);
  else
    hipLaunchKernelGGL(k_phi_pack<float>, dim3(nblocks(n)), dim3(256), 0, s,
                       full, nx, ny, nky, d, scale, (float *)phi);
  return ST_OK;
}

}  // namespace gfmd
