// Soft wall over a height map (`fix wall/map`) on the device: the map's
// displacement to z0 (FixWallMap::displace_map, src/extras/
// fix_wall_map.cpp:1275-1340) and the wall force on a list of atoms
// (FixWallMap::pre_force, 686-811, with its energy / force / count vector,
// compute_scalar and compute_vector 815-845).  The relaxation and dynamics
// kernels call the same per-atom functions (wall_map.h).
#include "relax.h"
#include "wall_map.h"

namespace gfmd {

// One block: sum, sum of squares, minimum and maximum of the map, then the
// shift za that puts the average (or the minimum / maximum) at z0.
// stats: shifted minimum, shifted maximum, rms height, shifted average, za.
__global__ void __launch_bounds__(1024)
    k_map_stats(const double *__restrict__ v, size_t n, double z0,
                int z0_type, double *__restrict__ stats) {
  __shared__ double red[16][2];
  __shared__ double mm[16][2];
  double s[2] = {0.0, 0.0};  // sum, sum of squares
  double lo = v[0], hi = v[0];
  for (size_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double x = v[i];
    s[0] += x;
    s[1] += x * x;
    lo = fmin(lo, x);
    hi = fmax(hi, x);
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = fmin(lo, __shfl_down(lo, off, 64));
    hi = fmax(hi, __shfl_down(hi, off, 64));
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    mm[wv][0] = lo;
    mm[wv][1] = hi;
  }
  block_sums<2>(s, red);  // its barrier also publishes mm
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int k = 1; k < nw; k++) {
      lo = fmin(lo, mm[k][0]);
      hi = fmax(hi, mm[k][1]);
    }
    const double dn = (double)n;
    const double avg = s[0] / dn;
    const double rms = sqrt(s[1] / dn - avg * avg);
    const double za =
        z0_type == 1 ? z0 - lo : (z0_type == 2 ? z0 - hi : z0 - avg);
    stats[0] = lo + za;
    stats[1] = hi + za;
    stats[2] = rms;
    stats[3] = avg + za;
    stats[4] = za;
  }
}

__global__ void k_map_shift(double *__restrict__ v, size_t n,
                            const double *__restrict__ stats) {
  const double za = stats[4];
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    v[i] += za;
}

void launch_map_displace(double *values, int mx, int my, double z0,
                         int z0_type, double *stats, hipStream_t st) {
  const size_t n = (size_t)mx * my;
  hipLaunchKernelGGL(k_map_stats, dim3(1), dim3(1024), 0, st, values, n, z0,
                     z0_type, stats);
  size_t b = (n + 255) / 256;
  hipLaunchKernelGGL(k_map_shift, dim3((unsigned)(b < 4096 ? b : 4096)),
                     dim3(256), 0, st, values, n, stats);
}

// One block per tile of kTile x kTile spline boxes: bounds of the
// computed spline over the tile.  On a box, f = sum c_ij dx^i dy^j with
// dx, dy in [0, 1), so |f - c_00| <= S = sum_{(i,j) != (0,0)} |c_ij|; the
// margin 1e-12 (|c_00| + S) covers the rounding of the Horner evaluation
// (a few dozen operations of relative error 2^-53 on terms bounded by
// |c_00| + S).
__global__ void __launch_bounds__(kTile * kTile)
    k_map_tile_bounds(WallMapDev w, double *__restrict__ tiles) {
  __shared__ double lo_s[kTile * kTile / 64], hi_s[kTile * kTile / 64];
  const int ntx = (w.mx + kTile - 1) / kTile;
  const int tx = blockIdx.x / w.nty, ty = blockIdx.x - tx * w.nty;
  const int bx = tx * kTile + threadIdx.x / kTile;
  const int by = ty * kTile + threadIdx.x % kTile;
  double lo = INFINITY, hi = -INFINITY;
  if (tx < ntx && bx < w.mx && by < w.my) {
    double cf[16];
    wm_coeffs(w, bx, by, cf);
    double S = 0.0;
    for (int r = 1; r < 16; r++) S += fabs(cf[r]);
    const double m = 1e-12 * (fabs(cf[0]) + S);
    lo = cf[0] - S - m;
    hi = cf[0] + S + m;
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = fmin(lo, __shfl_down(lo, off, 64));
    hi = fmax(hi, __shfl_down(hi, off, 64));
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    lo_s[wv] = lo;
    hi_s[wv] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kTile * kTile / 64; k++) {
      lo = fmin(lo, lo_s[k]);
      hi = fmax(hi, hi_s[k]);
    }
    tiles[2 * (size_t)blockIdx.x] = lo;
    tiles[2 * (size_t)blockIdx.x + 1] = hi;
  }
}

void launch_map_tile_bounds(const WallMapDev &w, double *tiles,
                            hipStream_t st) {
  const int ntx = (w.mx + kTile - 1) / kTile;
  hipLaunchKernelGGL(k_map_tile_bounds, dim3(ntx * w.nty), dim3(kTile * kTile),
                     0, st, w, tiles);
}

enum { W_E = 0, W_FX, W_FY, W_FZ, W_NCON, W_NREP, W_NATT, W_CLOSE, NW };

__global__ void __launch_bounds__(256)
    k_wall_map_force(WallMapDev w, int n, double *__restrict__ x,
                     double *__restrict__ f, double *__restrict__ part) {
  __shared__ double red[4][NW];
  double acc[NW];
  for (int k = 0; k < NW; k++) acc[k] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x) {
    double z = x[3 * i + 2];
    double fx, fy, fz, e, fp;
    if (wm_atom(w, x[3 * i], x[3 * i + 1], z, fx, fy, fz, e, fp,
                acc[W_CLOSE])) {
      x[3 * i + 2] = z;
      f[3 * i] += fx;
      f[3 * i + 1] += fy;
      f[3 * i + 2] += fz;
      acc[W_E] += e;
      acc[W_FX] += fx;
      acc[W_FY] += fy;
      acc[W_FZ] += fz;
      acc[W_NCON] += 1.0;
      if (fp > 0.0)
        acc[W_NREP] += 1.0;
      else
        acc[W_NATT] += 1.0;
    }
  }
  block_sums<NW>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < NW; k++) part[(size_t)k * gridDim.x + blockIdx.x] = acc[k];
}

__global__ void __launch_bounds__(1024)
    k_wall_map_final(const double *__restrict__ part, int nblk,
                     double *__restrict__ out) {
  __shared__ double red[16][NW];
  double tot[NW];
  for (int k = 0; k < NW; k++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x)
      s += part[(size_t)k * nblk + b];
    tot[k] = s;
  }
  block_sums<NW>(tot, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < NW; k++) out[k] = tot[k];
}

void launch_wall_map_force(const WallMapDev &w, int n, double *x, double *f,
                           double *part, int nblk, double *out,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_wall_map_force, dim3(nblk), dim3(256), 0, st, w, n, x,
                     f, part);
  hipLaunchKernelGGL(k_wall_map_final, dim3(1), dim3(1024), 0, st, part, nblk,
                     out);
}

}  // namespace gfmd
