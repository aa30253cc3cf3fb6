// Generic elastic force evaluation by direct DFTs: any nx, ny (including
// the odd and prime sizes of the reference's tests, e.g. 64x37, 63x63),
// O(N (nx + ny)) work.  Same conventions as the fast FFT path
// (GFMDSolverStatic::post_force, src/solvers/gfmd_solver_static.cpp:145-249):
// half spectrum spec[d][ny/2+1][nx], Hermitian extension in the inverse.
#include "gfmd_internal.h"

namespace gfmd {

// ---------------------------------------------------------------------------
// Generic path: direct DFTs (any nx, ny).
// ---------------------------------------------------------------------------
// spec[c][ky][ix] = sum_iy u[c][ix][iy] w_ny^(ky*iy)
template <typename TI, typename T>
__global__ void k_gen_rows_r2c(const TI *__restrict__ u, C2<T> *__restrict__ spec,
                               const C2<T> *__restrict__ twy, int nx, int ny,
                               int nky, int d, const int *__restrict__ gate) {
  if (gate && *gate == 0) return;
  const size_t n = (size_t)d * nky * nx;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < n;
       L += (size_t)gridDim.x * blockDim.x) {
    const int ix = (int)(L % nx);
    const int ky = (int)((L / nx) % nky);
    const int c = (int)(L / ((size_t)nx * nky));
    const TI *row = u + ((size_t)c * nx + ix) * ny;
    C2<T> acc = cmake((T)0, (T)0);
    int idx = 0;
    for (int iy = 0; iy < ny; iy++) {
      const C2<T> w = twy[idx];
      acc = acc + cscale(w, (T)row[iy]);
      idx += ky;
      if (idx >= ny) idx -= ny;
    }
    spec[L] = acc;
  }
}

// out[c][ky][k] = sum_ix in[c][ky][ix] w_nx^(SIGN k ix)
template <typename T, int SIGN>
__global__ void k_gen_cols(const C2<T> *__restrict__ in, C2<T> *__restrict__ out,
                           const C2<T> *__restrict__ twx, int nx, int nky,
                           int d, const int *__restrict__ gate) {
  if (gate && *gate == 0) return;
  const size_t n = (size_t)d * nky * nx;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < n;
       L += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(L % nx);
    const C2<T> *col = in + (L - k);
    C2<T> acc = cmake((T)0, (T)0);
    int idx = 0;
    for (int ix = 0; ix < nx; ix++) {
      C2<T> w = twx[idx];
      if (SIGN > 0) w = cconj(w);
      acc = acc + cmul(col[ix], w);
      idx += k;
      if (idx >= nx) idx -= nx;
    }
    out[L] = acc;
  }
}

template <typename T>
__global__ void k_gen_apply(C2<T> *__restrict__ q, const T *__restrict__ phi,
                            int nx, int ny, int nky, int d,
                            const double *__restrict__ linf,
                            double *__restrict__ partials, double *__restrict__ u0,
                            const int *__restrict__ gate) {
  __shared__ double red[4];
  if (gate && *gate == 0) return;
  const int dd = d * d;
  const size_t nq = (size_t)nky * nx;
  double e = 0.0;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < nq;
       L += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(L % nx);
    const int ky = (int)(L / nx);
    C2<T> uq[24], F[24];
    for (int c = 0; c < d; c++) uq[c] = q[(size_t)c * nq + L];
    const T *ph = phi + (size_t)ky * dd * nx + kx;
    for (int i = 0; i < d; i++) {
      const T pii = ph[(size_t)i * nx];
      F[i] = cmake(pii * uq[i].x, pii * uq[i].y);
    }
    int p = d;
    for (int i = 0; i < d; i++)
      for (int j = i + 1; j < d; j++) {
        const C2<T> pij = cmake(ph[(size_t)p * nx], ph[(size_t)(p + 1) * nx]);
        p += 2;
        F[i] = F[i] + cmul(pij, uq[j]);
        F[j] = F[j] + cmulc(uq[i], pij);
      }
    T eq = 0;
    for (int i = 0; i < d; i++) eq += F[i].x * uq[i].x + F[i].y * uq[i].y;
    const double w = (ky == 0 || 2 * ky == ny) ? 1.0 : 2.0;
    e += w * (double)eq;
    if (L == 0) {
      for (int c = 0; c < d; c++) u0[c] = (double)uq[c].x;
      if (linf)
        for (int i = 0; i < d / 3; i++) {
          e -= 2.0 * linf[i] * (double)uq[3 * i + 2].x;
        }
    }
    for (int c = 0; c < d; c++) {
      C2<T> o = cmake(-F[c].x, -F[c].y);
      if (L == 0 && linf && c % 3 == 2) o.x += (T)linf[c / 3];
      q[(size_t)c * nq + L] = o;
    }
  }
  const double bs = block_sum(e, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// f[c][ix][iy] = sum over the full ky range with Hermitian extension.
template <typename TO, typename T>
__global__ void k_gen_rows_c2r(const C2<T> *__restrict__ spec, TO *__restrict__ f,
                               const C2<T> *__restrict__ twy, int nx, int ny,
                               int nky, int d, const int *__restrict__ gate) {
  if (gate && *gate == 0) return;
  const size_t n = (size_t)d * nx * ny;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < n;
       L += (size_t)gridDim.x * blockDim.x) {
    const int iy = (int)(L % ny);
    const int ix = (int)((L / ny) % nx);
    const int c = (int)(L / ((size_t)nx * ny));
    const C2<T> *col = spec + (size_t)c * nky * nx + ix;
    T acc = 0;
    int idx = 0;
    for (int ky = 0; ky < nky; ky++) {
      C2<T> a = col[(size_t)ky * nx];
      const bool self = (ky == 0) || (2 * ky == ny);
      if (self) a.y = 0;
      const C2<T> w = cconj(twy[idx]);  // exp(+2 pi i ky iy / ny)
      const T re = a.x * w.x - a.y * w.y;
      acc += self ? re : (T)2 * re;
      idx += iy;
      if (idx >= ny) idx -= ny;
    }
    f[L] = (TO)acc;
  }
}

int force_generic_nparts(int nx, int ny, int d) {
  const size_t nq = (size_t)(ny / 2 + 1) * nx;
  size_t b = (nq + 255) / 256;
  return (int)(b < 1024 ? b : 1024);
}

template <typename TIO, typename T>
static int run_generic(const ForceArgs &a, hipStream_t st) {
  const int nky = a.nky;
  auto nb = [](size_t n) {
    size_t b = (n + 255) / 256;
    return (unsigned)(b < 4096 ? b : 4096);
  };
  const size_t ns = (size_t)a.d * nky * a.nx;
  hipLaunchKernelGGL((k_gen_rows_r2c<TIO, T>), dim3(nb(ns)), dim3(256), 0, st,
                     (const TIO *)a.u, (C2<T> *)a.spec, (const C2<T> *)a.twy,
                     a.nx, a.ny, nky, a.d, a.gate);
  hipLaunchKernelGGL((k_gen_cols<T, -1>), dim3(nb(ns)), dim3(256), 0, st,
                     (const C2<T> *)a.spec, (C2<T> *)a.spec2,
                     (const C2<T> *)a.twx, a.nx, nky, a.d, a.gate);
  const int nparts = force_generic_nparts(a.nx, a.ny, a.d);
  if (nparts > a.nparts) return ST_STATE;
  hipLaunchKernelGGL((k_gen_apply<T>), dim3(nparts), dim3(256), 0, st,
                     (C2<T> *)a.spec2, (const T *)a.phi, a.nx, a.ny, nky, a.d,
                     a.linf, a.partials, a.u0, a.gate);
  launch_energy_final(a.partials, nparts, a.epot, st, a.gate);
  hipLaunchKernelGGL((k_gen_cols<T, +1>), dim3(nb(ns)), dim3(256), 0, st,
                     (const C2<T> *)a.spec2, (C2<T> *)a.spec,
                     (const C2<T> *)a.twx, a.nx, nky, a.d, a.gate);
  const size_t nf = (size_t)a.d * a.nx * a.ny;
  hipLaunchKernelGGL((k_gen_rows_c2r<TIO, T>), dim3(nb(nf)), dim3(256), 0, st,
                     (const C2<T> *)a.spec, (TIO *)a.f, (const C2<T> *)a.twy,
                     a.nx, a.ny, nky, a.d, a.gate);
  return ST_OK;
}

int launch_force_generic(const ForceArgs &a, bool fp64, bool io_double,
                         hipStream_t st) {
  if (a.d > 24 || a.members != 1) return ST_UNSUPPORTED;
  if (fp64) return run_generic<double, double>(a, st);
  if (io_double) return run_generic<double, float>(a, st);
  return run_generic<float, float>(a, st);
}

}  // namespace gfmd
