// Seeded self-affine rough surfaces on the device (the rigid counter-
// surfaces of the contact workloads): white noise -> spectral filter
// sqrt(C(q)) (the force sweep with a diagonal filter "stiffness",
// phi_gen.hip k_phi_filter) -> scaled to a target rms slope.
//
// Random numbers: SplitMix64 (Steele, Lea, Flood, OOPSLA 2014) of
// seed and grid index, two 53-bit uniforms, Box-Muller transform; the
// stream depends only on (seed, index), so surfaces are reproducible and
// independent of launch shape.
#include "gfmd_internal.h"

namespace gfmd {

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ void k_white_noise(double *__restrict__ w, int nx, int ny,
                              unsigned long long seed) {
  const size_t N = (size_t)nx * ny;
  const unsigned long long key = splitmix64(seed);
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
       i += (size_t)gridDim.x * blockDim.x) {
    const unsigned long long r1 = splitmix64(key ^ (2 * i));
    const unsigned long long r2 = splitmix64(key ^ (2 * i + 1));
    const double u1 = ((r1 >> 11) + 1) * (1.0 / 9007199254740993.0);  // (0,1]
    const double u2 = (r2 >> 11) * (1.0 / 9007199254740992.0);        // [0,1)
    w[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
    w[N + i] = 0.0;
    w[2 * N + i] = 0.0;
  }
}

// sum over sites of |grad h|^2 by central differences (periodic)
__global__ void __launch_bounds__(256)
    k_slope_partials(const double *__restrict__ h, int nx, int ny, double a,
                     double *__restrict__ part) {
  __shared__ double red[4];
  const size_t N = (size_t)nx * ny;
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
       i += (size_t)gridDim.x * blockDim.x) {
    const int ix = (int)(i / ny), iy = (int)(i - (size_t)ix * ny);
    const int xp = ix + 1 == nx ? 0 : ix + 1, xm = ix == 0 ? nx - 1 : ix - 1;
    const int yp = iy + 1 == ny ? 0 : iy + 1, ym = iy == 0 ? ny - 1 : iy - 1;
    const double gx = (h[(size_t)xp * ny + iy] - h[(size_t)xm * ny + iy]) / (2 * a);
    const double gy = (h[(size_t)ix * ny + yp] - h[(size_t)ix * ny + ym]) / (2 * a);
    acc += gx * gx + gy * gy;
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_scale(double *__restrict__ h, size_t n,
                        const double *__restrict__ part, int nblk,
                        double target) {
  __shared__ double fac;
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += part[b];
    const double rms = sqrt(s / (double)n);
    fac = rms > 0.0 ? target / rms : 0.0;
  }
  __syncthreads();
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    h[i] *= fac;
}

void launch_white_noise(double *w, int nx, int ny, unsigned long long seed,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_white_noise, dim3(2048), dim3(256), 0, s, w, nx, ny,
                     seed);
}

void launch_slope_partials(const double *h, int nx, int ny, double a,
                           double *part, int nblk, hipStream_t s) {
  hipLaunchKernelGGL(k_slope_partials, dim3(nblk), dim3(256), 0, s, h, nx, ny,
                     a, part);
}

void launch_scale(double *h, size_t n, const double *part, int nblk,
                  double target, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3(1024), dim3(256), 0, s, h, n, part, nblk,
                     target);
}

}  // namespace gfmd
