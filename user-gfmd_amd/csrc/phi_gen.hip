// Surface stiffness tabulation Phi(q) on the device, one thread per
// wavevector of the half spectrum.  Output layout: T phi[nky][d*d][nx],
// Hermitian-packed (d diagonal reals, then Re/Im of Phi_ij for i < j in
// row-major order), scaled by 1/(nx*ny) when normalising.
//
// Reference semantics: fill_phi_buffer (src/main/gfmd_misc.cpp:32-102) with
// StiffnessKernel::get_stiffness_matrix -> greens_function_transfer_matrix_
// stiffness / iterate_Gnn (src/main/surface_stiffness.cpp:295-398, 493-548,
// 811-873) with the sc100 and fcc100 per-layer matrices
// (src/stiffness_kernels/sc100_stiffness.cpp:167-227,
// fcc100_stiffness.cpp:275-487), the closed-form isotropic kernel
// (isotropic_stiffness.cpp:95-167) and the open-boundary `nonperiodic`
// kernel (nonperiodic_stiffness.cpp:116-270).
#include <cmath>
#include <type_traits>
#include <vector>

#include "ft_lattice.h"
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

// fcc100 cell matrices U0, U, V (dimension 3 NU) for NU = 1..4
// single-atom layers per cell.  Per-layer blocks:
// FCC100StiffnessKernel::get_per_layer_dynamical_matrices
// (src/stiffness_kernels/fcc100_stiffness.cpp:275-487); k[3], kap[3]
// spring constants of the three shells, z the surface normal (+-1).
template <int NU>
__device__ void fcc100_cell(double qx, double qy, const double *k,
                            const double *kap, double z,
                            Z (&U0)[3 * NU][3 * NU], Z (&U)[3 * NU][3 * NU],
                            Z (&V)[3 * NU][3 * NU]) {
  double sx, cx, sy, cy, sx2, cx2, sy2, cy2;
  sincos(qx, &sx, &cx);
  sincos(qy, &sy, &cy);
  sincos(0.5 * qx, &sx2, &cx2);
  sincos(0.5 * qy, &sy2, &cy2);
  const double s2 = 1.4142135623730950488016887242097;
  Z bU0[3][3], bU[3][3], bV[3][3], bU1[3][3], bW[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      bU0[i][j] = bU[i][j] = bV[i][j] = bU1[i][j] = bW[i][j] = cmake(0.0, 0.0);
  bU[0][0].x = k[0] * (4 - 2 * cx) + kap[0] * (8 - 2 * cy);
  bU[1][1].x = k[0] * (4 - 2 * cy) + kap[0] * (8 - 2 * cx);
  bU[2][2].x = k[0] * 4 + kap[0] * (8 - 2 * cx - 2 * cy);
  bU0[0][0].x = bU[0][0].x - k[0] * 1 - kap[0] * 3;
  bU0[1][1].x = bU[1][1].x - k[0] * 1 - kap[0] * 3;
  bU0[2][2].x = bU[2][2].x - k[0] * 2 - kap[0] * 2;
  bV[0][0].x = -k[0] * cx2 * cy2 - kap[0] * 3 * cx2 * cy2;
  bV[1][1].x = -k[0] * cx2 * cy2 - kap[0] * 3 * cx2 * cy2;
  bV[2][2].x = -k[0] * 2 * cx2 * cy2 - kap[0] * 2 * cx2 * cy2;
  bV[0][1].x = bV[1][0].x = k[0] * sx2 * sy2 - kap[0] * sx2 * sy2;
  bV[0][2].y = bV[2][0].y = z * (k[0] * s2 * sx2 * cy2 - kap[0] * s2 * sx2 * cy2);
  bV[1][2].y = bV[2][1].y = z * (k[0] * s2 * cx2 * sy2 - kap[0] * s2 * cx2 * sy2);
  if (NU >= 2) {
    // second and third shells (interactions over a full lattice cell)
    Z u2[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) u2[i][j] = cmake(0.0, 0.0);
    u2[0][0].x = k[1] * (2 - 2 * cx * cy) + kap[1] * (4 - 2 * cx * cy) +
                 k[2] * 8 + kap[2] * 16;
    u2[1][1].x = u2[0][0].x;
    u2[2][2].x = k[1] * 2 + kap[1] * (4 - 4 * cx * cy) + k[2] * 8 + kap[2] * 16;
    u2[0][1].x = u2[1][0].x = k[1] * 2 * sx * sy - kap[1] * 2 * sx * sy;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        bU0[i][j] = bU0[i][j] + u2[i][j];
        bU[i][j] = bU[i][j] + u2[i][j];
        bU1[i][j] = bU[i][j];
      }
    bU0[0][0].x -= kap[1] + k[2] * 4 + kap[2] * 8;
    bU0[1][1].x -= kap[1] + k[2] * 4 + kap[2] * 8;
    bU0[2][2].x -= k[1] + k[2] * 4 + kap[2] * 8;
    bU1[0][0].x -= kap[1] + k[2] * 2 / 3 + kap[2] * 10 / 3;
    bU1[1][1].x -= kap[1] + k[2] * 2 / 3 + kap[2] * 10 / 3;
    bU1[2][2].x -= k[1] + k[2] * 8 / 3 + kap[2] * 4 / 3;
    bV[0][0].x += -k[2] * 1 / 3 * cx2 * cy2 * (-10 + 18 * cx + 2 * cy) -
                  kap[2] * 1 / 3 * cx2 * cy2 * (-14 + 6 * cx + 22 * cy);
    bV[1][1].x += -k[2] * 1 / 3 * cx2 * cy2 * (-10 + 2 * cx + 18 * cy) -
                  kap[2] * 1 / 3 * cx2 * cy2 * (-14 + 22 * cx + 6 * cy);
    bV[2][2].x += -k[2] * 2 / 3 * cx2 * cy2 * (-2 + 2 * (cx + cy)) -
                  kap[2] * 2 / 3 * cx2 * cy2 * (-10 + 10 * (cx + cy));
    const double v01 = -k[2] * (-2 - 2 * (cx + cy)) * sx2 * sy2 -
                       kap[2] * (2 + 2 * (cx + cy)) * sx2 * sy2;
    const double v02 =
        z * (-k[2] * 1 / 3 * s2 * cy2 * (-2 - 2 * (3 * cx + cy)) * sx2 -
             kap[2] * 1 / 3 * s2 * cy2 * (2 + 2 * (3 * cx + cy)) * sx2);
    const double v12 =
        z * (-k[2] * 1 / 3 * s2 * cx2 * (-2 - 2 * (cx + 3 * cy)) * sy2 -
             kap[2] * 1 / 3 * s2 * cx2 * (2 + 2 * (cx + 3 * cy)) * sy2);
    bV[0][1].x += v01;
    bV[1][0].x += v01;
    bV[0][2].y += v02;
    bV[2][0].y += v02;
    bV[1][2].y += v12;
    bV[2][1].y += v12;
    bW[0][0].x = -kap[1] - k[2] * 2 / 3 * cx - kap[2] * (4 * cx / 3 + 2 * cy);
    bW[1][1].x = -kap[1] - k[2] * 2 / 3 * cy - kap[2] * (2 * cx + 4 * cy / 3);
    bW[2][2].x = -k[1] - k[2] * 4 / 3 * (cx + cy) - kap[2] * 2 / 3 * (cx + cy);
    bW[0][2].y = bW[2][0].y = z * (k[2] * 2 / 3 * s2 * sx - kap[2] * 2 / 3 * s2 * sx);
    bW[1][2].y = bW[2][1].y = z * (k[2] * 2 / 3 * s2 * sy - kap[2] * 2 / 3 * s2 * sy);
  }
  // Cell assembly (StiffnessKernel::get_dynamical_matrices,
  // surface_stiffness.cpp:127-292).  Blocks between layers a distance r
  // apart: r = 0 the layer's own U (U0, U1 at the surface, else bulk U),
  // r = 1 V, r = 2 W, r >= 3 zero (interactions reach two layers).  Layer
  // shifts (1/2, 1/2) on layer 1 for NU = 2, 4 and layer 3 for NU = 4
  // (fcc100_stiffness.cpp:476-486).
  double xs[NU];
  for (int i = 0; i < NU; i++) xs[i] = 0.0;
  if (NU == 2 || NU == 4) xs[1] = 0.5;
  if (NU == 4) xs[3] = 0.5;
  auto blk = [&](int r, int layer, int a, int b) -> Z {
    if (r == 0) {
      if (layer == 0) return bU0[a][b];
      if (layer == 1 && NU >= 2) return bU1[a][b];
      return bU[a][b];
    }
    if (r == 1) return bV[a][b];
    if (r == 2 && NU >= 2) return bW[a][b];
    return cmake(0.0, 0.0);
  };
  for (int i = 0; i < 3 * NU; i++)
    for (int j = 0; j < 3 * NU; j++) U0[i][j] = U[i][j] = V[i][j] = cmake(0.0, 0.0);
  for (int i = 0; i < NU; i++)
    for (int j = i; j < NU; j++) {
      double sf, cf;
      sincos((xs[j] - xs[i]) * (qx + qy), &sf, &cf);
      const Z fac = cmake(cf, sf);
      const int r = j - i;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          if (r == 0) {
            U0[3 * i + a][3 * j + b] = blk(0, i, a, b);
            U[3 * i + a][3 * j + b] = blk(0, NU, a, b);
          } else {
            const Z m = zmul(fac, blk(r, i, a, b));
            const Z mt = zmul(fac, blk(r, i, b, a));
            U0[3 * i + a][3 * j + b] = m;
            U[3 * i + a][3 * j + b] = m;
            U0[3 * j + a][3 * i + b] = cmake(mt.x, -mt.y);
            U[3 * j + a][3 * i + b] = cmake(mt.x, -mt.y);
          }
        }
    }
  for (int i = 0; i < NU; i++)
    for (int j = 0; j <= i; j++) {
      double sf, cf;
      sincos((xs[j] - xs[i]) * (qx + qy), &sf, &cf);
      const Z fac = cmake(cf, sf);
      const int r = NU + j - i;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
          V[3 * i + a][3 * j + b] = zmul(fac, blk(r, NU, a, b));
    }
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

// ---------------------------------------------------------------------------
// Wave-cooperative small dense algebra for ndof >= 6.  One 64-lane wave
// works on one wavevector: the D x D complex matrices live in LDS and every
// lane takes its share of the elements.  (One thread per wavevector would
// keep 7-9 such matrices per lane -- up to 21 KB of private scratch at
// D = 12 -- which the runtime has to reserve for every wave the device can
// hold.)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int D>
struct WaveMats {
  Z U0[D][D], U[D][D], V[D][D], Vd[D][D], X[D][D], B[D][D], VT[D][D],
      N[D][D], P[D][D];
  Z F[D];
};

// B = X^-1 by Gauss-Jordan on [X | B] with partial pivoting (the first row
// of largest modulus, as zinv); X is destroyed.  The result is uniform
// across the wave.
template <int D>
__device__ bool winv(Z (*X)[D], Z (*B)[D], Z *F, int lane) {
  for (int e = lane; e < D * D; e += 64)
    B[e / D][e % D] = cmake(e / D == e % D ? 1.0 : 0.0, 0.0);
  wave_sync();
  for (int k = 0; k < D; k++) {
    double v = -1.0;
    int r = D;
    if (lane >= k && lane < D) {
      v = zabs2(X[lane][k]);
      r = lane;
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(v, off, 64);
      const int orr = __shfl_xor(r, off, 64);
      if (ov > v || (ov == v && orr < r)) {
        v = ov;
        r = orr;
      }
    }
    if (!(v > 0.0)) return false;
    if (r != k && lane < D) {
      Z t = X[k][lane];
      X[k][lane] = X[r][lane];
      X[r][lane] = t;
      t = B[k][lane];
      B[k][lane] = B[r][lane];
      B[r][lane] = t;
    }
    wave_sync();
    const Z piv = zdiv1(X[k][k]);
    wave_sync();
    if (lane < D) {
      X[k][lane] = zmul(X[k][lane], piv);
      B[k][lane] = zmul(B[k][lane], piv);
      F[lane] = X[lane][k];  // elimination factors (row k is not used)
    }
    wave_sync();
    for (int e = lane; e < D * D; e += 64) {
      const int i = e / D, j = e % D;
      if (i != k) {
        const Z fct = F[i];
        X[i][j] = X[i][j] - zmul(fct, X[k][j]);
        B[i][j] = B[i][j] - zmul(fct, B[k][j]);
      }
    }
    wave_sync();
  }
  return true;
}

// c = a b; c must not alias a or b.
template <int D>
__device__ void wmm(const Z (*a)[D], const Z (*b)[D], Z (*c)[D], int lane) {
  for (int e = lane; e < D * D; e += 64) {
    const int i = e / D, j = e % D;
    Z s = cmake(0.0, 0.0);
    for (int k = 0; k < D; k++) s = s + zmul(a[i][k], b[k][j]);
    c[i][j] = s;
  }
  wave_sync();
}

// transfer_stiffness on the wave's matrices: Phi (into m.N) from m.U0,
// m.U, m.V, same recursion and stopping rule.
template <int D>
__device__ bool wtransfer(int height, WaveMats<D> &m, int lane) {
  if (height == 0) {
    for (int e = lane; e < D * D; e += 64) m.N[e / D][e % D] = m.U0[e / D][e % D];
    wave_sync();
    return true;
  }
  for (int e = lane; e < D * D; e += 64) {
    const int i = e / D, j = e % D;
    m.Vd[i][j] = cmake(-m.V[j][i].x, m.V[j][i].y);  // -conj(V^T)
    m.X[i][j] = m.U[i][j];
  }
  wave_sync();
  if (!winv<D>(m.X, m.B, m.F, lane)) return false;
  wmm<D>(m.B, m.Vd, m.X, lane);
  wmm<D>(m.V, m.X, m.VT, lane);
  const double eps2 = height < 0 ? 1e-16 : 0.0;
  const int maxit = height < 0 ? 100000 : height - 1;
  double dnorm = 1.0;
  int it = 0;
  while (dnorm > eps2 && it < maxit) {
    for (int e = lane; e < D * D; e += 64)
      m.X[e / D][e % D] = m.VT[e / D][e % D] + m.U[e / D][e % D];
    wave_sync();
    if (!winv<D>(m.X, m.B, m.F, lane)) return false;
    wmm<D>(m.B, m.Vd, m.X, lane);
    wmm<D>(m.V, m.X, m.N, lane);
    double sq = 0.0;
    for (int e = lane; e < D * D; e += 64) {
      const int i = e / D, j = e % D;
      sq += zabs2(m.N[i][j] - m.VT[i][j]);
      m.VT[i][j] = m.N[i][j];
    }
    dnorm = wave_sum(sq);
    wave_sync();
    it++;
  }
  if (height < 0 && it >= maxit) return false;
  for (int e = lane; e < D * D; e += 64)
    m.N[e / D][e % D] = m.U0[e / D][e % D] + m.VT[e / D][e % D];
  wave_sync();
  return true;
}

// Packed store of the Hermitian part of a D x D matrix (as store_packed),
// lanes over the elements.
template <int D, typename T>
__device__ void wstore_packed(const Z (*M)[D], double scale, T *__restrict__ out,
                              int nx, int kx, int lane) {
  for (int e = lane; e < D * D; e += 64) {
    const int i = e / D, j = e % D;
    if (i == j) {
      out[(size_t)i * nx + kx] = (T)(M[i][i].x * scale);
    } else if (i < j) {
      const int p = D + 2 * (i * D - i * (i + 1) / 2 + (j - i - 1));
      out[(size_t)p * nx + kx] = (T)(0.5 * (M[i][j].x + M[j][i].x) * scale);
      out[(size_t)(p + 1) * nx + kx] =
          (T)(0.5 * (M[i][j].y - M[j][i].y) * scale);
    }
  }
}

// Reference wavevector convention (gfmd_misc.cpp:42-47).
__device__ __forceinline__ double qcomp(int i, int n) {
  const double tp = 6.283185307179586476925286766559;
  return (i <= n / 2) ? tp * i / n : tp * (i - n) / n;
}

// Per-kernel evaluators of Phi(q) (D x D), used by k_phi_tab.  `periodic`
// marks kernels with Phi(q + 2 pi e) = Phi(q), for which the Nyquist-line
// partner needs no separate evaluation.
struct EvalSC100 {
  static constexpr int D = 3;
  static constexpr bool periodic = true;
  int height;
  __device__ bool operator()(double qx, double qy, Z (&P)[3][3]) const {
    Z U0[3][3], U[3][3], V[3][3];
    sc100_layer(qx, qy, U0, U, V);
    return transfer_stiffness<3>(height, U0, U, V, P);
  }
};

template <int NU>
struct EvalFCC100 {
  static constexpr int D = 3 * NU;
  // the nu = 3 cell carries no layer shifts: its half-angle terms have
  // period 4 pi
  static constexpr bool periodic = NU != 3;
  int height;
  double p[7];
  __device__ bool operator()(double qx, double qy, Z (&P)[D][D]) const {
    Z U0[D][D], U[D][D], V[D][D];
    fcc100_cell<NU>(qx, qy, p, p + 3, p[6], U0, U, V);
    return transfer_stiffness<D>(height, U0, U, V, P);
  }
  // wave version: one lane assembles the cell matrices into LDS
  __device__ void assemble(double qx, double qy, Z (*U0)[D], Z (*U)[D],
                           Z (*V)[D], int lane) const {
    if (lane == 0)
      fcc100_cell<NU>(qx, qy, p, p + 3, p[6],
                      *reinterpret_cast<Z(*)[D][D]>(U0),
                      *reinterpret_cast<Z(*)[D][D]>(U),
                      *reinterpret_cast<Z(*)[D][D]>(V));
    wave_sync();
  }
};

// Isotropic continuum half space: Phi = G^-1 with the closed-form surface
// compliance G(q) (Poisson number nu, shear modulus myu); at q = 0 the
// diagonal compliance (alpha + 1/2, alpha + 1/2, 1) xprd / (gamma pi).
// Reference: IsotropicStiffnessKernel::get_stiffness_matrix
// (src/stiffness_kernels/isotropic_stiffness.cpp:95-167).  Not periodic in
// q (a continuum).
struct EvalIsotropic {
  static constexpr int D = 3;
  static constexpr bool periodic = false;
  double nu, myu, gamma, xprd;
  __device__ bool operator()(double qx, double qy, Z (&P)[3][3]) const {
    const double alpha = 1.0 / (2.0 * (1.0 - nu));
    const double b = sqrt(qx * qx + qy * qy);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) P[i][j] = cmake(0.0, 0.0);
    if (fabs(qx) <= 1e-6 && fabs(qy) <= 1e-6) {
      const double pi = 3.141592653589793238462643383279503;
      P[0][0].x = P[1][1].x = (alpha + 0.5) * xprd / (gamma * pi);
      P[2][2].x = xprd / (gamma * pi);
    } else {
      const double c = (2.0 - 1.0 / alpha) / (2.0 * myu * b * b * b);
      const double s = (1.0 - 1.0 / alpha) / (2.0 * myu * b * b);
      P[0][0].x = 1.0 / (myu * b) - qx * qx * c;
      P[1][1].x = 1.0 / (myu * b) - qy * qy * c;
      P[2][2].x = 1.0 / (2.0 * myu * alpha * b);
      P[0][1].x = P[1][0].x = -qx * qy * c;
      P[0][2].y = -qx * s;
      P[1][2].y = -qy * s;
      P[2][0].y = qx * s;
      P[2][1].y = qy * s;
    }
    return zinv<3>(P);
  }
};

// `ft` kernel: U0, U and V summed from the off-site force-constant blocks
// of every atom pair times exp(i (cx qx + cy qy)) (neighbours in the same
// layer, ab = 0, into U0 and U; those one layer below, ab = -1, into V),
// plus the on-site blocks, then the transfer-matrix stiffness.  Reference:
// FTStiffnessKernel::get_dynamical_matrices (src/stiffness_kernels/
// ft_stiffness.cpp:199-292).  Entries are 12 doubles (cx, cy, ab, 3 x 3
// block) grouped by atom pair (a, b) from off[a NAT + b]; `on` holds the
// on-site blocks of layer 0 then layer 1 (ft_lattice.h).
template <int NAT>
struct EvalFT {
  static constexpr int D = 3 * NAT;
  static constexpr bool periodic = true;  // integer cell offsets
  int height;
  const double *ent;
  const int *off;
  const double *on;
  __device__ bool operator()(double qx, double qy, Z (&P)[D][D]) const {
    Z U0[D][D], U[D][D], V[D][D];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = 0; j < D; j++)
        U0[i][j] = U[i][j] = V[i][j] = cmake(0.0, 0.0);
#pragma unroll
    for (int a = 0; a < NAT; a++)
#pragma unroll
      for (int b = 0; b < NAT; b++) {
        const int e1 = off[a * NAT + b + 1];
        for (int e = off[a * NAT + b]; e < e1; e++) {
          const double *t = ent + 12 * (size_t)e;
          double sn, cs;
          sincos(t[0] * qx + t[1] * qy, &sn, &cs);
          const bool same = t[2] == 0.0;
#pragma unroll
          for (int r = 0; r < 3; r++)
#pragma unroll
            for (int k = 0; k < 3; k++) {
              const double y = t[3 + 3 * r + k];
              const Z v = cmake(cs * y, sn * y);
              if (same) {
                U0[3 * a + r][3 * b + k] = U0[3 * a + r][3 * b + k] + v;
                U[3 * a + r][3 * b + k] = U[3 * a + r][3 * b + k] + v;
              } else {
                V[3 * a + r][3 * b + k] = V[3 * a + r][3 * b + k] + v;
              }
            }
        }
      }
#pragma unroll
    for (int a = 0; a < NAT; a++)
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
          U0[3 * a + r][3 * a + k].x += on[(a * 3 + r) * 3 + k];
          U[3 * a + r][3 * a + k].x += on[((NAT + a) * 3 + r) * 3 + k];
        }
    return transfer_stiffness<D>(height, U0, U, V, P);
  }
  // wave version: every lane sums the entries of its matrix elements
  __device__ void assemble(double qx, double qy, Z (*U0)[D], Z (*U)[D],
                           Z (*V)[D], int lane) const {
    for (int e = lane; e < D * D; e += 64) {
      const int row = e / D, col = e % D;
      const int a = row / 3, r = row % 3, b = col / 3, k = col % 3;
      Z u0 = cmake(0.0, 0.0), v = cmake(0.0, 0.0);
      const int e1 = off[a * NAT + b + 1];
      for (int t = off[a * NAT + b]; t < e1; t++) {
        const double *en = ent + 12 * (size_t)t;
        double sn, cs;
        sincos(en[0] * qx + en[1] * qy, &sn, &cs);
        const double y = en[3 + 3 * r + k];
        if (en[2] == 0.0)
          u0 = u0 + cmake(cs * y, sn * y);
        else
          v = v + cmake(cs * y, sn * y);
      }
      Z uu = u0;
      if (a == b) {
        u0.x += on[(a * 3 + r) * 3 + k];
        uu.x += on[((NAT + a) * 3 + r) * 3 + k];
      }
      U0[row][col] = u0;
      U[row][col] = uu;
      V[row][col] = v;
    }
    wave_sync();
  }
};

// One thread per wavevector of the half spectrum.  On a Nyquist line
// (2 kx = nx or 2 ky = ny) the grid partner q' of q is not -q; the
// reference applies Phi by a complex transform and keeps the real part,
// which is the same as applying (Phi(q) + conj Phi(q'))/2, stored here for
// kernels that are not 2 pi periodic in q.
template <class E, typename T>
__global__ void k_phi_tab(int nx, int ny, int nky, double scale, E ev,
                          T *__restrict__ phi, int *__restrict__ status) {
  constexpr int D = E::D;
  const size_t n = (size_t)nky * nx;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < n;
       L += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(L % nx);
    const int ky = (int)(L / nx);
    Z P[D][D];
    if (!ev(qcomp(kx, nx), qcomp(ky, ny), P)) {
      atomicOr(status, 1);
      continue;
    }
    if (!E::periodic && (2 * kx == nx || 2 * ky == ny)) {
      Z P2[D][D];
      if (!ev(qcomp((nx - kx) % nx, nx), qcomp((ny - ky) % ny, ny), P2)) {
        atomicOr(status, 1);
        continue;
      }
      for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++)
          P[i][j] = cmake(0.5 * (P[i][j].x + P2[i][j].x),
                          0.5 * (P[i][j].y - P2[i][j].y));
    }
    store_packed<D>(P, scale, phi + (size_t)ky * D * D * nx, nx, kx);
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

// Low-pass filter of a height map (FixWallMap::filter_map, src/extras/
// fix_wall_map.cpp:1179-1269): Phi = -A I_3 with A = 0 for the modes with
// sqrt((kx / xprd)^2 + (ky / yprd)^2) > 1 / lambda, kx = min(i, nx - i),
// and A = 1 otherwise; f = -Phi~ u is the filtered map.
template <typename T>
__global__ void k_phi_lowpass(int nx, int ny, int nky, double cutq,
                              double xprd, double yprd, double scale,
                              T *__restrict__ phi) {
  const size_t n = (size_t)nky * nx;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < n;
       L += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(L % nx);
    const int ky = (int)(L / nx);
    const int dx = kx < nx - kx ? kx : nx - kx;
    const int dy = ky < ny - ky ? ky : ny - ky;
    const double dqx = 1.0 / xprd, dqy = 1.0 / yprd;
    const double r = sqrt((dx * dqx) * (dx * dqx) + (dy * dqy) * (dy * dqy));
    const double A = r > cutq ? 0.0 : 1.0;
    T *o = phi + (size_t)ky * 9 * nx + kx;
    for (int e = 0; e < 9; e++)
      o[(size_t)e * nx] = (T)(e < 3 ? -A * scale : 0.0);
  }
}

// full: [nx*ny][d][d] complex128 (reference fill_phi_buffer layout).
// Stores the Hermitian part of Phi_s(q) = (Phi(q) + conj Phi(q'))/2, q' the
// grid partner of q: the operator the reference applies when it keeps the
// real part of its complex inverse transform (equal to Phi(q) for tables
// with Phi(-q) = conj Phi(q)).
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
    const double *mp =
        full + ((size_t)((nx - kx) % nx) * ny + (ny - ky) % ny) * dd * 2;
    // Phi_s entries: re = (Re m + Re mp)/2, im = (Im m - Im mp)/2
    auto sre = [&](int i, int j) {
      return 0.5 * (m[2 * (i * d + j)] + mp[2 * (i * d + j)]);
    };
    auto sim = [&](int i, int j) {
      return 0.5 * (m[2 * (i * d + j) + 1] - mp[2 * (i * d + j) + 1]);
    };
    T *o = phi + (size_t)ky * dd * nx + kx;
    for (int i = 0; i < d; i++) o[(size_t)i * nx] = (T)(sre(i, i) * scale);
    int p = d;
    for (int i = 0; i < d; i++)
      for (int j = i + 1; j < d; j++) {
        const double re = 0.5 * (sre(i, j) + sre(j, i));
        const double im = 0.5 * (sim(i, j) - sim(j, i));
        o[(size_t)p * nx] = (T)(re * scale);
        o[(size_t)(p + 1) * nx] = (T)(im * scale);
        p += 2;
      }
  }
}

// ---------------------------------------------------------------------------
// `nonperiodic` kernel (open boundaries): the real-space Green's function of
// an isotropic half space sampled at the minimum-image offsets of the grid,
// transformed, inverted per q.  Reference: NonperiodicStiffnessKernel::
// create_and_fill_buffer (src/stiffness_kernels/nonperiodic_stiffness.cpp:
// 116-196, 198-270) with LiBerger::sq_uniform_PQ (li_berger.cpp:231-304,
// 453-459) and the plane-geometry predicates (geometry.cpp:42-142).
// ---------------------------------------------------------------------------
namespace lb {

__device__ double angle(double x1, double y1, double x2, double y2) {
  const double l = sqrt(x1 * x1 + y1 * y1) * sqrt(x2 * x2 + y2 * y2);
  return atan2((x2 * y1 - x1 * y2) / l, (x1 * x2 + y1 * y2) / l);
}

__device__ bool cross(double a1x, double a1y, double a2x, double a2y,
                      double b1x, double b1y, double b2x, double b2y) {
  double c1 = (a1x - a2x) * (b1y - a2y) - (a1y - a2y) * (b1x - a2x);
  double c2 = (a1x - a2x) * (b2y - a2y) - (a1y - a2y) * (b2x - a2x);
  if (c1 * c2 > 0) return false;
  c1 = (b1x - b2x) * (a1y - b2y) - (b1y - b2y) * (a1x - b2x);
  c2 = (b1x - b2x) * (a2y - b2y) - (b1y - b2y) * (a2x - b2x);
  return !(c1 * c2 > 0);
}

__device__ bool inside(double bx, double by, double a1x, double a1y,
                       double a2x, double a2y, double a3x, double a3y) {
  const double c1 = (a1x - a2x) * (by - a2y) - (a1y - a2y) * (bx - a2x);
  const double c2 = (a2x - a3x) * (by - a3y) - (a2y - a3y) * (bx - a3x);
  const double c3 = (a3x - a1x) * (by - a1y) - (a3y - a1y) * (bx - a1x);
  return c1 * c2 > 0 && c2 * c3 > 0 && c3 * c1 > 0;
}

__device__ bool overlap(double bx, double by, double cx, double cy,
                        double a1x, double a1y, double a2x, double a2y) {
  if (inside(bx, by, cx, cy, a1x, a1y, a2x, a2y) ||
      inside(cx, cy, bx, by, a1x, a1y, a2x, a2y))
    return true;
  return cross(a1x, a1y, bx, by, a2x, a2y, cx, cy) ||
         cross(a2x, a2y, bx, by, a1x, a1y, cx, cy);
}

__device__ double ltan(double p) { return log(fabs(tan(0.5 * p))); }

// the three displacement components from unit tractions (P, Qx, Qy) on the
// corner triangle (b, a2, a3): out[comp][load], load order Qx, Qy, P
__device__ void corner(double G, double nu, double sgn, double bx, double by,
                       double a2x, double a2y, double a3x, double a3y,
                       double (&o)[3][3]) {
  const double area =
      0.5 * fabs((a2x - bx) * (a3y - by) - (a3x - bx) * (a2y - by));
  if (area <= 0.0) return;
  const double l23 = sqrt((a2x - a3x) * (a2x - a3x) + (a2y - a3y) * (a2y - a3y));
  const double n1 = 2 * area / l23;
  const double d = -angle(1.0, 0.0, a3x - a2x, a3y - a2y);
  const double e1 = angle(a2x - a3x, a2y - a3y, bx - a2x, by - a2y);
  const double e2 = angle(a2x - a3x, a2y - a3y, bx - a3x, by - a3y);
  double sd, cd, s2d, c2d;
  sincos(d, &sd, &cd);
  sincos(2 * d, &s2d, &c2d);
  const double I100 = n1 * (ltan(e1) - ltan(e2));
  const double I201 = -n1 * ((sd * log(fabs(sin(e1))) - e1 * cd) -
                             (sd * log(fabs(sin(e2))) - e2 * cd));
  const double I210 = -n1 * ((cd * log(fabs(sin(e1))) + e1 * sd) -
                             (cd * log(fabs(sin(e2))) + e2 * sd));
  auto f311 = [&](double p) {
    return s2d * ltan(p) + 2 * s2d * cos(p) - 2 * c2d * sin(p);
  };
  auto f302 = [&](double p) { return sd * sd * ltan(p) - cos(2 * d - p); };
  auto f320 = [&](double p) { return cd * cd * ltan(p) + cos(2 * d - p); };
  const double I311 = 0.5 * n1 * (f311(e1) - f311(e2));
  const double I302 = n1 * (f302(e1) - f302(e2));
  const double I320 = n1 * (f320(e1) - f320(e2));
  const double pi = 3.141592653589793238462643383279503;
  const double a = (1 - 2 * nu) / (4 * pi * G), c = (1 - nu) / (2 * pi * G),
               e = nu / (2 * pi * G);
  // load Qx (column 0)
  o[0][0] += sgn * (c * I100 + e * I320);
  o[1][0] += sgn * e * I311;
  o[2][0] += sgn * -a * I210;
  // load Qy (column 1)
  o[0][1] += sgn * e * I311;
  o[1][1] += sgn * (c * I100 + e * I302);
  o[2][1] += sgn * -a * I201;
  // load P (column 2)
  o[0][2] += sgn * a * I210;
  o[1][2] += sgn * a * I201;
  o[2][2] += sgn * c * I100;
}

__device__ void triangle(double G, double nu, const double (&A)[6], double bx,
                         double by, double (&o)[3][3]) {
  for (int k = 0; k < 3; k++) {
    const int i = k, j = (k + 1) % 3, m = (k + 2) % 3;
    const double s = overlap(bx, by, A[2 * m], A[2 * m + 1], A[2 * i],
                             A[2 * i + 1], A[2 * j], A[2 * j + 1])
                         ? 1.0
                         : -1.0;
    corner(G, nu, s, bx, by, A[2 * i], A[2 * i + 1], A[2 * j], A[2 * j + 1], o);
  }
}

}  // namespace lb

// The transform of the Green's function as real matrix products.  Along
// an axis of length n, the h = n/2 + 1 offsets x' = 0 .. n/2 pair each
// point x' with its mirror n - x' (itself at 0 and, for even n, at n/2), and
//   sum_x g(x) e^{+2 pi i k x / n} = sum_x' E(x') cos(2 pi k x' / n)
//                                 + i sum_x' O(x') sin(2 pi k x' / n)
// with E = g(x') + g(-x') and O = g(x') - g(-x') (a self-mirror point: E =
// g(x'), O = 0), for k < h; k > n/2 is k' = n - k with the sine's sign
// flipped.  In two dimensions each component splits into four sectors
// (E or O along x, E or O along y), each transformed by a product with the
// cosine or sine matrix along y and then along x:
//   G~(kx, ky) = Z_EE - sx Z_OO + i (sx Z_OE + Z_EO),  sx = -1 for kx > nx/2.
// Li & Berger's components are even or odd along each axis, so one sector
// carries each; the others hold the rounding asymmetry of g, kept so that
// the table is the reference's full transform (nonperiodic_stiffness.cpp:
// 198-270) to rounding.  Six components: xx, yy, zz, xy, xz, yz; yx = xy,
// zx = -xz, zy = -yz hold exactly in g (corner() above).

// G[6][nx][ny]: displacement from a unit force spread over the cell at the
// origin at the minimum-image offset of each grid point, off-diagonals
// zeroed on the far edge of an even grid (the reference's rule that keeps
// Phi Hermitian)
__global__ void __launch_bounds__(256)
    k_np_green(int nx, int ny, double nu, double mu, double dx, double dy,
               double *__restrict__ G) {
  const size_t n = (size_t)nx * ny;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < n;
       L += (size_t)gridDim.x * blockDim.x) {
    const int ii = (int)(L / ny), jj = (int)(L % ny);
    const int i = ii <= nx / 2 ? ii : ii - nx;
    const int j = jj <= ny / 2 ? jj : jj - ny;
    const double a = 0.5 * dx, b = 0.5 * dy, fac = 1.0 / (dx * dy);
    double o[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const double t1[6] = {-a, -b, a, -b, a, b};
    const double t2[6] = {-a, -b, -a, b, a, b};
    lb::triangle(mu, nu, t1, i * dx, j * dy, o);
    lb::triangle(mu, nu, t2, i * dx, j * dy, o);
    const bool edge = (nx % 2 == 0 && ii == nx / 2) || (ny % 2 == 0 && jj == ny / 2);
    const double off = edge ? 0.0 : fac;
    G[0 * n + L] = fac * o[0][0];
    G[1 * n + L] = fac * o[1][1];
    G[2 * n + L] = fac * o[2][2];
    G[3 * n + L] = off * o[0][1];
    G[4 * n + L] = off * o[0][2];
    G[5 * n + L] = off * o[1][2];
  }
}

// F[6][4][hx][hy]: the sectors EE, EO, OE, OO (x first) of each component
__global__ void k_np_sectors(int nx, int ny, int hx, int hy,
                             const double *__restrict__ G,
                             double *__restrict__ F) {
  const size_t Q = (size_t)hx * hy, n = (size_t)nx * ny;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < 6 * Q;
       L += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(L / Q);
    const size_t r = L - (size_t)c * Q;
    const int x = (int)(r / hy), y = (int)(r % hy);
    // self-mirror offsets: 0 and (even n) n/2
    const bool selfx = x == 0 || (nx % 2 == 0 && 2 * x == nx);
    const bool selfy = y == 0 || (ny % 2 == 0 && 2 * y == ny);
    const int xm = selfx ? x : nx - x, ym = selfy ? y : ny - y;
    const double *g = G + (size_t)c * n;
    const double pp = g[(size_t)x * ny + y];
    // fold along y at x and at its mirror, then along x
    const double ey0 = selfy ? pp : pp + g[(size_t)x * ny + ym];
    const double oy0 = selfy ? 0.0 : pp - g[(size_t)x * ny + ym];
    double ey1 = 0.0, oy1 = 0.0;
    if (!selfx) {
      const double pm = g[(size_t)xm * ny + y];
      ey1 = selfy ? pm : pm + g[(size_t)xm * ny + ym];
      oy1 = selfy ? 0.0 : pm - g[(size_t)xm * ny + ym];
    }
    double *f = F + (size_t)c * 4 * Q + r;
    f[0 * Q] = selfx ? ey0 : ey0 + ey1;  // EE
    f[1 * Q] = selfx ? oy0 : oy0 + oy1;  // EO
    f[2 * Q] = selfx ? 0.0 : ey0 - ey1;  // OE
    f[3 * Q] = selfx ? 0.0 : oy0 - oy1;  // OO
  }
}

// T[k][m] = cos (odd = 0) or sin (odd = 1) of 2 pi (k m mod n) / n for k,
// m < h, from tw[j] = e^{+2 pi i j / n}
__global__ void k_np_trig(int n, int h, int odd, const double2 *__restrict__ tw,
                          double *__restrict__ T) {
  const size_t q = (size_t)h * h;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < q;
       L += (size_t)gridDim.x * blockDim.x) {
    const long long k = (long long)(L / h), m = (long long)(L % h);
    const double2 e = tw[(int)((k * m) % n)];
    T[L] = odd ? e.y : e.x;
  }
}

// C_b = A_b B_b for the 24 sectors (row-major FP64: A M x K, B K x N,
// C M x N).  A 256-thread block computes a 128 x 128 tile of C, 8 x 8
// values per thread (rows ty + 16 i, columns tx + 16 j), stepping K by 16
// with both tiles staged in LDS (A transposed, its rows padded by one
// element).  Every global access is bounds-checked; tiles past the edges
// are zero-filled.
constexpr int NP_SECTORS = 24;
struct NpGemm {
  const double *A[NP_SECTORS];
  const double *B[NP_SECTORS];
  double *C[NP_SECTORS];
};

__global__ void __launch_bounds__(256)
    k_np_dgemm(NpGemm g, int M, int N, int K) {
  constexpr int BM = 128, BN = 128, BK = 16;
  __shared__ double As[BK][BM + 1];
  __shared__ double Bs[BK][BN];
  const int bz = blockIdx.z;
  const double *__restrict__ A = g.A[bz];
  const double *__restrict__ B = g.B[bz];
  double *__restrict__ C = g.C[bz];
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  double acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j < 8; j++) acc[i][j] = 0.0;
  for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const int idx = threadIdx.x + 256 * r;  // < BM * BK = BK * BN = 2048
      const int mm = idx / BK, ka = idx % BK;
      const int gm = m0 + mm, gka = k0 + ka;
      As[ka][mm] = (gm < M && gka < K) ? A[(size_t)gm * K + gka] : 0.0;
      const int kb = idx / BN, nn = idx % BN;
      const int gkb = k0 + kb, gn = n0 + nn;
      Bs[kb][nn] = (gkb < K && gn < N) ? B[(size_t)gkb * N + gn] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk++) {
      double av[8], bv[8];
#pragma unroll
      for (int i = 0; i < 8; i++) av[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 8; j++) bv[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 8; j++) acc[i][j] = fma(av[i], bv[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const int gm = m0 + ty + 16 * i;
    if (gm >= M) continue;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int gn = n0 + tx + 16 * j;
      if (gn < N) C[(size_t)gm * N + gn] = acc[i][j];
    }
  }
}

// Phi = G~^-1 per q of the half spectrum (kx < nx, ky < hy), packed and
// scaled, from the sector transforms Z[6][4][hx][hy]
template <typename T>
__global__ void k_np_invert(int nx, int hx, int hy, double scale,
                            const double *__restrict__ Zs,
                            T *__restrict__ phi, int *__restrict__ status) {
  const size_t n = (size_t)hy * nx;
  const size_t Q = (size_t)hx * hy;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < n;
       L += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(L % nx), ky = (int)(L / nx);
    const bool mir = 2 * kx > nx;
    const size_t off = (size_t)(mir ? nx - kx : kx) * hy + ky;
    const double sx = mir ? -1.0 : 1.0;
    Z g[6];
    for (int c = 0; c < 6; c++) {
      const double *z = Zs + (size_t)c * 4 * Q + off;
      g[c] = cmake(z[0] - sx * z[3 * Q], sx * z[2 * Q] + z[Q]);
    }
    Z P[3][3];
    P[0][0] = g[0];
    P[1][1] = g[1];
    P[2][2] = g[2];
    P[0][1] = g[3];
    P[1][0] = g[3];
    P[0][2] = g[4];
    P[2][0] = cmake(-g[4].x, -g[4].y);
    P[1][2] = g[5];
    P[2][1] = cmake(-g[5].x, -g[5].y);
    if (!zinv<3>(P)) {
      atomicOr(status, 1);
      continue;
    }
    store_packed<3>(P, scale, phi + (size_t)ky * 9 * nx, nx, kx);
  }
}

// Compliance tables for the contact solver from a packed Phi~ table (Phi/N,
// N = nx ny): gz[nky][nx] = -(Phi~^-1)_zz / N^2 for the z component `zc`
// of the surface layer, and (optionally) gfull[nky][d*d][nx] = the packed
// -Phi~^-1 / N^2.  With these tables the force sweep returns G * x with
// G = Phi^-1 the surface compliance.  A singular Phi(q) sets *status.
// With cavg (device, D x D real symmetric) the inverse is of Phi~ + cavg,
// and gz may be NULL: the preconditioner table of prec_gradient.
template <int D, typename T>
__global__ void k_compliance(const T *__restrict__ phi, int nx, int nky,
                             double inv_n2, int zc, T *__restrict__ gz,
                             T *__restrict__ gfull, int *__restrict__ status,
                             const double *__restrict__ cavg) {
  const size_t n = (size_t)nky * nx;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < n;
       L += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(L % nx);
    const int ky = (int)(L / nx);
    const T *ph = phi + (size_t)ky * D * D * nx + kx;
    Z P[D][D];
    for (int i = 0; i < D; i++) P[i][i] = cmake((double)ph[(size_t)i * nx], 0.0);
    int q = D;
    for (int i = 0; i < D; i++)
      for (int j = i + 1; j < D; j++) {
        const double re = (double)ph[(size_t)q * nx];
        const double im = (double)ph[(size_t)(q + 1) * nx];
        P[i][j] = cmake(re, im);
        P[j][i] = cmake(re, -im);
        q += 2;
      }
    if (cavg)
      for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) P[i][j].x += cavg[i * D + j];
    if (!zinv<D>(P)) {
      atomicOr(status, 1);
      continue;
    }
    if (gz) gz[(size_t)ky * nx + kx] = (T)(-P[zc][zc].x * inv_n2);
    if (gfull) {
      for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) P[i][j] = cmake(-P[i][j].x, -P[i][j].y);
      store_packed<D>(P, inv_n2, gfull + (size_t)ky * D * D * nx, nx, kx);
    }
  }
}

// Tabulation for ndof >= 6: one wave (= block) per wavevector, grid-stride.
template <class E, typename T>
__global__ void __launch_bounds__(64)
    k_phi_tab_wave(int nx, int ny, int nky, double scale, E ev,
                   T *__restrict__ phi, int *__restrict__ status) {
  constexpr int D = E::D;
  __shared__ WaveMats<D> m;
  const int lane = threadIdx.x;
  const size_t n = (size_t)nky * nx;
  for (size_t L = blockIdx.x; L < n; L += gridDim.x) {
    const int kx = (int)(L % nx);
    const int ky = (int)(L / nx);
    ev.assemble(qcomp(kx, nx), qcomp(ky, ny), m.U0, m.U, m.V, lane);
    bool ok = wtransfer<D>(ev.height, m, lane);
    if (ok && !E::periodic && (2 * kx == nx || 2 * ky == ny)) {
      // (Phi(q) + conj Phi(q'))/2 on a Nyquist line, as k_phi_tab
      for (int e = lane; e < D * D; e += 64) m.P[e / D][e % D] = m.N[e / D][e % D];
      wave_sync();
      ev.assemble(qcomp((nx - kx) % nx, nx), qcomp((ny - ky) % ny, ny), m.U0,
                  m.U, m.V, lane);
      ok = wtransfer<D>(ev.height, m, lane);
      if (ok) {
        for (int e = lane; e < D * D; e += 64) {
          const int i = e / D, j = e % D;
          m.N[i][j] = cmake(0.5 * (m.P[i][j].x + m.N[i][j].x),
                            0.5 * (m.P[i][j].y - m.N[i][j].y));
        }
        wave_sync();
      }
    }
    if (!ok) {
      if (lane == 0) atomicOr(status, 1);
      wave_sync();
      continue;
    }
    wstore_packed<D>(m.N, scale, phi + (size_t)ky * D * D * nx, nx, kx, lane);
    wave_sync();
  }
}

// Compliance tables for ndof >= 6 (as k_compliance), one wave per q.
template <int D, typename T>
__global__ void __launch_bounds__(64)
    k_compliance_wave(const T *__restrict__ phi, int nx, int nky,
                      double inv_n2, int zc, T *__restrict__ gz,
                      T *__restrict__ gfull, int *__restrict__ status,
                      const double *__restrict__ cavg) {
  __shared__ WaveMats<D> m;
  const int lane = threadIdx.x;
  const size_t n = (size_t)nky * nx;
  for (size_t L = blockIdx.x; L < n; L += gridDim.x) {
    const int kx = (int)(L % nx);
    const int ky = (int)(L / nx);
    const T *ph = phi + (size_t)ky * D * D * nx + kx;
    for (int e = lane; e < D * D; e += 64) {
      const int i = e / D, j = e % D;
      if (i == j) {
        m.X[i][i] = cmake((double)ph[(size_t)i * nx], 0.0);
      } else {
        const int a = i < j ? i : j, b = i < j ? j : i;
        const int p = D + 2 * (a * D - a * (a + 1) / 2 + (b - a - 1));
        const double re = (double)ph[(size_t)p * nx];
        const double im = (double)ph[(size_t)(p + 1) * nx];
        m.X[i][j] = cmake(re, i < j ? im : -im);
      }
      if (cavg) m.X[i][j].x += cavg[e];
    }
    wave_sync();
    if (!winv<D>(m.X, m.B, m.F, lane)) {
      if (lane == 0) atomicOr(status, 1);
      wave_sync();
      continue;
    }
    if (lane == 0 && gz)
      gz[(size_t)ky * nx + kx] = (T)(-m.B[zc][zc].x * inv_n2);
    if (gfull) {
      for (int e = lane; e < D * D; e += 64)
        m.X[e / D][e % D] = cmake(-m.B[e / D][e % D].x, -m.B[e / D][e % D].y);
      wave_sync();
      wstore_packed<D>(m.X, inv_n2, gfull + (size_t)ky * D * D * nx, nx, kx,
                       lane);
    }
    wave_sync();
  }
}

static unsigned nblocks(size_t n);

// One thread per (ky, kx) with kx > nx/2: every packed entry against the
// entry of nx - kx, sign-flipped where exactly one index is an x
// component, relative to the trace at the two wavevectors.
template <typename T>
__global__ void k_phi_mirror_check(const T *__restrict__ phi, int nx, int nky,
                                   int d, double tol, int *__restrict__ bad) {
  const int half = nx / 2 + 1;
  const int w = nx - half;  // kx = half .. nx-1
  const size_t n = (size_t)nky * (w > 0 ? w : 0);
  const int DD = d * d;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < n;
       L += (size_t)gridDim.x * blockDim.x) {
    const int ky = (int)(L / w), kx = half + (int)(L % w), km = nx - kx;
    const T *p = phi + (size_t)ky * DD * nx;
    double tr = 0.0;
    for (int i = 0; i < d; i++)
      tr += fabs((double)p[(size_t)i * nx + kx]) +
            fabs((double)p[(size_t)i * nx + km]);
    const double lim = tol * 0.5 * tr;
    bool ok = true;
    for (int i = 0; i < d; i++)
      ok = ok && fabs((double)p[(size_t)i * nx + kx] -
                      (double)p[(size_t)i * nx + km]) <= lim;
    int e = d;
    for (int i = 0; i < d; i++)
      for (int j = i + 1; j < d; j++, e += 2) {
        const double sg = ((i % 3 == 0) != (j % 3 == 0)) ? -1.0 : 1.0;
        for (int c = 0; c < 2; c++)
          ok = ok && fabs((double)p[(size_t)(e + c) * nx + kx] -
                          sg * (double)p[(size_t)(e + c) * nx + km]) <= lim;
      }
    if (!ok) atomicOr(bad, 1);
  }
}

void launch_phi_mirror_check(const void *phi, int nx, int ny, int d,
                             bool fp64, double tol, int *d_bad,
                             hipStream_t s) {
  const int nky = ny / 2 + 1;
  const size_t n = (size_t)nky * (nx - (nx / 2 + 1));
  if (n == 0) return;
  if (fp64)
    hipLaunchKernelGGL(k_phi_mirror_check<double>, dim3(nblocks(n)),
                       dim3(256), 0, s, (const double *)phi, nx, nky, d, tol,
                       d_bad);
  else
    hipLaunchKernelGGL(k_phi_mirror_check<float>, dim3(nblocks(n)),
                       dim3(256), 0, s, (const float *)phi, nx, nky, d, tol,
                       d_bad);
}

// ndof 3: sets *bad unless Im Phi_xy, Re Phi_xz and Re Phi_yz are exactly
// zero at every stored wavevector (packed entries 4, 5, 7): the structure
// of kernels symmetric under x -> -x and y -> -y separately, whose tables
// the force sweep can read as 6 reals per wavevector.
template <typename T>
__global__ void k_phi_par_check(const T *__restrict__ phi, int nx, int nky,
                                int *__restrict__ bad) {
  const size_t n = (size_t)nky * nx;
  bool ok = true;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < n;
       L += (size_t)gridDim.x * blockDim.x) {
    const int ky = (int)(L / nx), kx = (int)(L % nx);
    const T *p = phi + (size_t)ky * 9 * nx + kx;
    ok = ok && p[(size_t)4 * nx] == (T)0 && p[(size_t)5 * nx] == (T)0 &&
         p[(size_t)7 * nx] == (T)0;
  }
  if (!ok) atomicOr(bad, 1);
}

// The compact table [nky][6][nx]: Phi_xx, Phi_yy, Phi_zz, Re Phi_xy,
// Im Phi_xz, Im Phi_yz (packed entries 0, 1, 2, 3, 6, 8).
template <typename T>
__global__ void k_phi_compact(const T *__restrict__ phi, T *__restrict__ pc,
                              int nx, int nky) {
  const size_t n = (size_t)nky * nx;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < n;
       L += (size_t)gridDim.x * blockDim.x) {
    const int ky = (int)(L / nx), kx = (int)(L % nx);
    const T *p = phi + (size_t)ky * 9 * nx + kx;
    T *q = pc + (size_t)ky * 6 * nx + kx;
    const int src[6] = {0, 1, 2, 3, 6, 8};
#pragma unroll
    for (int e = 0; e < 6; e++) q[(size_t)e * nx] = p[(size_t)src[e] * nx];
  }
}

void launch_phi_par_check(const void *phi, int nx, int ny, bool fp64,
                          int *d_bad, hipStream_t s) {
  const size_t n = (size_t)(ny / 2 + 1) * nx;
  if (fp64)
    hipLaunchKernelGGL(k_phi_par_check<double>, dim3(nblocks(n)), dim3(256),
                       0, s, (const double *)phi, nx, ny / 2 + 1, d_bad);
  else
    hipLaunchKernelGGL(k_phi_par_check<float>, dim3(nblocks(n)), dim3(256), 0,
                       s, (const float *)phi, nx, ny / 2 + 1, d_bad);
}

void launch_phi_compact(const void *phi, void *pc, int nx, int ny, bool fp64,
                        hipStream_t s) {
  const size_t n = (size_t)(ny / 2 + 1) * nx;
  if (fp64)
    hipLaunchKernelGGL(k_phi_compact<double>, dim3(nblocks(n)), dim3(256), 0,
                       s, (const double *)phi, (double *)pc, nx, ny / 2 + 1);
  else
    hipLaunchKernelGGL(k_phi_compact<float>, dim3(nblocks(n)), dim3(256), 0,
                       s, (const float *)phi, (float *)pc, nx, ny / 2 + 1);
}

int launch_compliance(const void *phi, int nx, int ny, int d, bool fp64,
                      int zc, void *gz, void *gfull, int *d_status,
                      hipStream_t s, const double *cavg) {
  const int nky = ny / 2 + 1;
  const size_t n = (size_t)nky * nx;
  const double nn = (double)nx * (double)ny;
  // the preconditioner table (cavg given) is the plain -(Phi~ + cavg)^-1
  const double inv_n2 = cavg ? 1.0 : 1.0 / (nn * nn);
  const unsigned gw = (unsigned)(n < 16384 ? n : 16384);
#define CMPLW(D)                                                              \
  if (d == D) {                                                               \
    if (fp64)                                                                 \
      hipLaunchKernelGGL((k_compliance_wave<D, double>), dim3(gw), dim3(64),  \
                         0, s, (const double *)phi, nx, nky, inv_n2, zc,      \
                         (double *)gz, (double *)gfull, d_status, cavg);      \
    else                                                                      \
      hipLaunchKernelGGL((k_compliance_wave<D, float>), dim3(gw), dim3(64),   \
                         0, s, (const float *)phi, nx, nky, inv_n2, zc,       \
                         (float *)gz, (float *)gfull, d_status, cavg);        \
    return ST_OK;                                                             \
  }
  CMPLW(6) CMPLW(9) CMPLW(12)
#undef CMPLW
#define CMPL(D)                                                               \
  if (d == D) {                                                               \
    if (fp64)                                                                 \
      hipLaunchKernelGGL((k_compliance<D, double>), dim3(nblocks(n)),         \
                         dim3(256), 0, s, (const double *)phi, nx, nky,       \
                         inv_n2, zc, (double *)gz, (double *)gfull, d_status, \
                         cavg);                                               \
    else                                                                      \
      hipLaunchKernelGGL((k_compliance<D, float>), dim3(nblocks(n)),          \
                         dim3(256), 0, s, (const float *)phi, nx, nky,        \
                         inv_n2, zc, (float *)gz, (float *)gfull, d_status,   \
                         cavg);                                               \
    return ST_OK;                                                             \
  }
  CMPL(3)
#undef CMPL
  return ST_UNSUPPORTED;
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
  auto tab = [&](auto ev) {
    using E = decltype(ev);
    if constexpr (E::D >= 6) {
      const unsigned g = (unsigned)(n < 16384 ? n : 16384);
      if (fp64)
        hipLaunchKernelGGL((k_phi_tab_wave<E, double>), dim3(g), dim3(64), 0,
                           s, nx, ny, nky, scale, ev, (double *)phi, d_status);
      else
        hipLaunchKernelGGL((k_phi_tab_wave<E, float>), dim3(g), dim3(64), 0,
                           s, nx, ny, nky, scale, ev, (float *)phi, d_status);
      return ST_OK;
    } else {
      if (fp64)
        hipLaunchKernelGGL((k_phi_tab<E, double>), dim3(nblocks(n)),
                           dim3(256), 0, s, nx, ny, nky, scale, ev,
                           (double *)phi, d_status);
      else
        hipLaunchKernelGGL((k_phi_tab<E, float>), dim3(nblocks(n)),
                           dim3(256), 0, s, nx, ny, nky, scale, ev,
                           (float *)phi, d_status);
      return ST_OK;
    }
  };
  if (k.kind == KIND_SC100) return tab(EvalSC100{k.height});
  if (k.kind == KIND_ISOTROPIC)
    return tab(EvalIsotropic{k.p[0], k.p[1], k.p[2], k.p[3]});
  if (k.kind == KIND_FCC100) {
    auto fcc = [&](auto nu) {
      EvalFCC100<decltype(nu)::value> ev;
      ev.height = k.height;
      for (int i = 0; i < 7; i++) ev.p[i] = k.p[i];
      return tab(ev);
    };
    switch (k.d) {
      case 3: return fcc(std::integral_constant<int, 1>());
      case 6: return fcc(std::integral_constant<int, 2>());
      case 9: return fcc(std::integral_constant<int, 3>());
      case 12: return fcc(std::integral_constant<int, 4>());
    }
    return ST_UNSUPPORTED;
  }
  if (k.kind == KIND_FT) {
    const FTTable &t = *k.ft;
    if (t.nat < 1 || t.nat > 4 || k.d != 3 * t.nat) return ST_UNSUPPORTED;
    double *ent = nullptr, *on = nullptr;
    int *off = nullptr;
    const size_t ne = t.ent.size() ? t.ent.size() : 1;
    int rc = ST_OK;
    if (hipMalloc((void **)&ent, ne * 8) != hipSuccess ||
        hipMalloc((void **)&on, t.onsite.size() * 8) != hipSuccess ||
        hipMalloc((void **)&off, t.off.size() * 4) != hipSuccess)
      rc = ST_HIP;
    if (rc == ST_OK &&
        ((t.ent.size() && hipMemcpyAsync(ent, t.ent.data(), t.ent.size() * 8,
                                         hipMemcpyHostToDevice, s) != hipSuccess) ||
         hipMemcpyAsync(on, t.onsite.data(), t.onsite.size() * 8,
                        hipMemcpyHostToDevice, s) != hipSuccess ||
         hipMemcpyAsync(off, t.off.data(), t.off.size() * 4,
                        hipMemcpyHostToDevice, s) != hipSuccess))
      rc = ST_HIP;
    if (rc == ST_OK) {
      auto run = [&](auto nat) {
        EvalFT<decltype(nat)::value> ev;
        ev.height = k.height;
        ev.ent = ent;
        ev.off = off;
        ev.on = on;
        return tab(ev);
      };
      switch (t.nat) {
        case 1: rc = run(std::integral_constant<int, 1>()); break;
        case 2: rc = run(std::integral_constant<int, 2>()); break;
        case 3: rc = run(std::integral_constant<int, 3>()); break;
        case 4: rc = run(std::integral_constant<int, 4>()); break;
      }
      if (hipStreamSynchronize(s) != hipSuccess) rc = ST_HIP;
    }
    (void)hipFree(ent);
    (void)hipFree(on);
    (void)hipFree(off);
    return rc;
  }
  if (k.kind == KIND_NONPERIODIC) {
    if (nx > 4096 || ny > 4096) return ST_UNSUPPORTED;
    const int hx = nx / 2 + 1, hy = ny / 2 + 1;  // hy = nky
    const size_t N = (size_t)nx * ny, Q = (size_t)hx * hy;
    const size_t SQ = (size_t)NP_SECTORS * Q;
    double *G = nullptr, *F = nullptr, *Y = nullptr, *W = nullptr;
    double2 *tw = nullptr;
    int rc = ST_OK;
    // W: cosine and sine matrices along y (hy x hy), then along x (hx x hx)
    const size_t nw = 2 * ((size_t)hy * hy + (size_t)hx * hx);
    if (hipMalloc((void **)&G, 6 * N * 8) != hipSuccess ||
        hipMalloc((void **)&F, SQ * 8) != hipSuccess ||
        hipMalloc((void **)&Y, SQ * 8) != hipSuccess ||
        hipMalloc((void **)&W, nw * 8) != hipSuccess ||
        hipMalloc((void **)&tw, (size_t)(nx + ny) * 16) != hipSuccess)
      rc = ST_HIP;
    if (rc == ST_OK) {
      // e^{+2 pi i k / n} (the reference transforms with FFTW's sign +1)
      std::vector<double2> h((size_t)nx + ny);
      const long double tp = 6.28318530717958647692528676655900577L;
      for (int i = 0; i < ny; i++)
        h[i] = make_double2((double)cosl(tp * i / ny), (double)sinl(tp * i / ny));
      for (int i = 0; i < nx; i++)
        h[ny + i] =
            make_double2((double)cosl(tp * i / nx), (double)sinl(tp * i / nx));
      if (hipMemcpyAsync(tw, h.data(), h.size() * 16, hipMemcpyHostToDevice,
                         s) != hipSuccess)
        rc = ST_HIP;
      if (rc == ST_OK) {
        const size_t qy = (size_t)hy * hy, qx = (size_t)hx * hx;
        double *Wyc = W, *Wys = W + qy, *Wxc = W + 2 * qy, *Wxs = Wxc + qx;
        hipLaunchKernelGGL(k_np_green, dim3(nblocks(N)), dim3(256), 0, s, nx,
                           ny, k.p[0], k.p[1], k.p[2], k.p[3], G);
        hipLaunchKernelGGL(k_np_sectors, dim3(nblocks(6 * Q)), dim3(256), 0, s,
                           nx, ny, hx, hy, G, F);
        hipLaunchKernelGGL(k_np_trig, dim3(nblocks(qy)), dim3(256), 0, s, ny,
                           hy, 0, tw, Wyc);
        hipLaunchKernelGGL(k_np_trig, dim3(nblocks(qy)), dim3(256), 0, s, ny,
                           hy, 1, tw, Wys);
        hipLaunchKernelGGL(k_np_trig, dim3(nblocks(qx)), dim3(256), 0, s, nx,
                           hx, 0, tw + ny, Wxc);
        hipLaunchKernelGGL(k_np_trig, dim3(nblocks(qx)), dim3(256), 0, s, nx,
                           hx, 1, tw + ny, Wxs);
        // sector q = 2 a + b (a: O along x, b: O along y) of component c
        NpGemm gy, gx;
        for (int c = 0; c < 6; c++)
          for (int q = 0; q < 4; q++) {
            const int bi = 4 * c + q;
            const size_t o = (size_t)bi * Q;
            gy.A[bi] = F + o;  // Y = F Ty: hx x hy times hy x hy
            gy.B[bi] = (q & 1) ? Wys : Wyc;
            gy.C[bi] = Y + o;
            gx.A[bi] = (q & 2) ? Wxs : Wxc;  // Z = Tx Y: hx x hx times hx x hy
            gx.B[bi] = Y + o;
            gx.C[bi] = F + o;  // the sectors are no longer needed
          }
        const dim3 grid((hy + 127) / 128, (hx + 127) / 128, NP_SECTORS);
        hipLaunchKernelGGL(k_np_dgemm, grid, dim3(256), 0, s, gy, hx, hy, hy);
        hipLaunchKernelGGL(k_np_dgemm, grid, dim3(256), 0, s, gx, hx, hy, hx);
        const size_t NH = (size_t)nx * hy;
        if (fp64)
          hipLaunchKernelGGL(k_np_invert<double>, dim3(nblocks(NH)),
                             dim3(256), 0, s, nx, hx, hy, scale, F,
                             (double *)phi, d_status);
        else
          hipLaunchKernelGGL(k_np_invert<float>, dim3(nblocks(NH)), dim3(256),
                             0, s, nx, hx, hy, scale, F, (float *)phi,
                             d_status);
        if (hipGetLastError() != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
          rc = ST_HIP;
      }
    }
    if (G) (void)hipFree(G);
    if (F) (void)hipFree(F);
    if (Y) (void)hipFree(Y);
    if (W) (void)hipFree(W);
    if (tw) (void)hipFree(tw);
    return rc;
  }
  if (k.kind == KIND_LOWPASS) {
    if (k.d != 3) return ST_ARG;
    if (fp64)
      hipLaunchKernelGGL(k_phi_lowpass<double>, dim3(nblocks(n)), dim3(256), 0,
                         s, nx, ny, nky, k.p[0], k.p[1], k.p[2], scale,
                         (double *)phi);
    else
      hipLaunchKernelGGL(k_phi_lowpass<float>, dim3(nblocks(n)), dim3(256), 0,
                         s, nx, ny, nky, k.p[0], k.p[1], k.p[2], scale,
                         (float *)phi);
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
