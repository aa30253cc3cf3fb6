// The atom list <-> grid glue of `fix gfmd` on the device, for atoms that
// live in device memory: the layer's centre of mass and its shift in
// lattice units, the gather of displacements (and velocities) onto the
// [ndof][nx][ny] grid, and the scatter of the grid forces back onto the
// atoms with the force sum and the optional removal of the centre-of-mass
// force.
//
// Reference: FixGFMD::pre_force (src/main/fix_gfmd.cpp:659-846: shift from
// the centre of mass, grid-id wrap, minimum-image displacements),
// FixGFMD::post_force / grid_to_list (870-1074: forces onto the atoms,
// force sum, fix_center), get_cm_from_xeq (1895-1926), Group::xcm (LAMMPS:
// mass-weighted centre of unwrapped positions); the reference's own
// device kernels for this step (src/cuda/fix_gfmd_cuda_kernel.cu:89, 187,
// 290).  Everything is one process on one device: no ghost atoms, no
// reverse communication.
#include "fix_glue.h"
#include "relax.h"

namespace gfmd {

// Minimum-image displacement along one axis: the reference's loops
// (fix_gfmd.cpp:766-769), each axis on its own.  A displacement of more
// than 64 periods is first reduced exactly to less than one period
// (fmod), so the loops run at most twice; up to that distance they
// reproduce the reference's result bit for bit.
__device__ __forceinline__ double wrap_half_box(double u, double h,
                                                double box) {
  if (fabs(u) > 64.0 * box) u = fmod(u, box);
  while (u > h) u -= box;
  while (u < -h) u += box;
  return u;
}

// LAMMPS image flags (32-bit imageint): 10 bits per dimension, offset 512.
__device__ __forceinline__ void unpack_image(int img, int &bx, int &by,
                                             int &bz) {
  bx = (img & 1023) - 512;
  by = (img >> 10 & 1023) - 512;
  bz = (img >> 20) - 512;
}

__device__ __forceinline__ bool in_group(const FixAtoms &a, int i) {
  return !a.mask || (a.mask[i] & a.groupbit);
}

// Per-block partial sums of m x (x unwrapped with the image flags, or xeq
// when use_xeq), m, and the group count: part[5][nblk].
__global__ void __launch_bounds__(256)
    k_fix_com(FixAtoms a, FixFrame fr, int use_xeq, double *__restrict__ part) {
  __shared__ double red[4][5];
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
       i += gridDim.x * blockDim.x) {
    if (!in_group(a, i)) continue;
    const double m = a.mass ? a.mass[i] : 1.0;
    double p[3];
    if (use_xeq) {
      for (int c = 0; c < 3; c++) p[c] = a.xeq[3 * (size_t)i + c];
    } else {
      int bx = 0, by = 0, bz = 0;
      if (a.image) unpack_image(a.image[i], bx, by, bz);
      p[0] = a.x[3 * (size_t)i] + bx * fr.box[0];
      p[1] = a.x[3 * (size_t)i + 1] + by * fr.box[1];
      p[2] = a.x[3 * (size_t)i + 2] + bz * fr.box[2];
    }
    for (int c = 0; c < 3; c++) acc[c] += p[c] * m;
    acc[3] += m;
    acc[4] += 1.0;
  }
  block_sums<5>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 5; k++) part[(size_t)k * gridDim.x + blockIdx.x] = acc[k];
}

// One block: the centre of mass; set-up (setup = 1) stores it as the
// equilibrium centre and clears the shift; otherwise the layer's shift in
// lattice units (nearest integer of invcell (xcm - xeqcm)) and its change
// since the last call.
__global__ void __launch_bounds__(1024)
    k_fix_shift(FixFrame fr, const double *__restrict__ part, int nblk,
                int setup, double *__restrict__ st) {
  __shared__ double red[16][5];
  double tot[5];
  for (int k = 0; k < 5; k++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x)
      s += part[(size_t)k * nblk + b];
    tot[k] = s;
  }
  block_sums<5>(tot, red);
  if (threadIdx.x != 0) return;
  const double mt = tot[3];
  double xcm[3];
  for (int c = 0; c < 3; c++) xcm[c] = mt > 0.0 ? tot[c] * (1.0 / mt) : 0.0;
  if (setup) {
    for (int c = 0; c < 3; c++) st[FS_XEQCM + c] = xcm[c];
    st[FS_XSHIFT] = st[FS_YSHIFT] = 0.0;
    st[FS_DXSHIFT] = st[FS_DYSHIFT] = 0.0;
    st[FS_MASSTOTAL] = mt;
    st[FS_NATOMS] = tot[4];
    return;
  }
  double d[3];
  for (int c = 0; c < 3; c++) {
    st[FS_XCM + c] = xcm[c];
    d[c] = xcm[c] - st[FS_XEQCM + c];
  }
  const double sx = rint(fr.invcell[0] * d[0] + fr.invcell[1] * d[1] +
                         fr.invcell[2] * d[2]);
  const double sy = rint(fr.invcell[3] * d[0] + fr.invcell[4] * d[1] +
                         fr.invcell[5] * d[2]);
  st[FS_DXSHIFT] = st[FS_XSHIFT] - sx;
  st[FS_DYSHIFT] = st[FS_YSHIFT] - sy;
  st[FS_XSHIFT] = sx;
  st[FS_YSHIFT] = sy;
}

// Gather: grid id shifted by the layer's shift change (wrapped and stored
// back when it is not zero), u = x - xeq with the minimum image in x and y
// onto the grid, and (optionally) the velocities onto a second grid.
// Partial counts: atoms placed, group atoms whose grid id is outside the
// grid or whose displacement is not within 64 box lengths.
template <typename T>
__global__ void __launch_bounds__(256)
    k_fix_gather(FixAtoms a, FixFrame fr, const double *__restrict__ st,
                 T *__restrict__ u, T *__restrict__ vg,
                 double *__restrict__ part) {
  __shared__ double red[4][2];
  const int dxs = (int)st[FS_DXSHIFT], dys = (int)st[FS_DYSHIFT];
  const size_t N = (size_t)fr.nx * fr.ny;
  const double hx = 0.5 * fr.box[0], hy = 0.5 * fr.box[1];
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
       i += gridDim.x * blockDim.x) {
    if (!in_group(a, i)) continue;
    int *g = a.gid + 3 * (size_t)i;
    int ix = g[0] - dxs, iy = g[1] - dys;
    const int iu = g[2];
    if (dxs != 0 || dys != 0) {
      ix %= fr.nx;
      if (ix < 0) ix += fr.nx;
      iy %= fr.ny;
      if (iy < 0) iy += fr.ny;
      g[0] = ix;
      g[1] = iy;
    }
    if (ix < 0 || ix >= fr.nx || iy < 0 || iy >= fr.ny || iu < 0 ||
        iu >= fr.nl) {
      acc[1] += 1.0;
      continue;
    }
    const double *x = a.x + 3 * (size_t)i, *xe = a.xeq + 3 * (size_t)i;
    double ux = x[0] - xe[0], uy = x[1] - xe[1];
    const double uz = x[2] - xe[2];
    if (!isfinite(ux) || !isfinite(uy)) {
      acc[1] += 1.0;
      continue;
    }
    ux = wrap_half_box(ux, hx, fr.box[0]);
    uy = wrap_half_box(uy, hy, fr.box[1]);
    const size_t site = (size_t)ix * fr.ny + iy;
    u[(size_t)(3 * iu) * N + site] = (T)ux;
    u[(size_t)(3 * iu + 1) * N + site] = (T)uy;
    u[(size_t)(3 * iu + 2) * N + site] = (T)uz;
    if (vg && a.v) {
      const double *v = a.v + 3 * (size_t)i;
      for (int c = 0; c < 3; c++) vg[(size_t)(3 * iu + c) * N + site] = (T)v[c];
    }
    acc[0] += 1.0;
  }
  block_sums<2>(acc, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = acc[0];
    part[(size_t)gridDim.x + blockIdx.x] = acc[1];
  }
}

// Scatter: f_atoms += the grid force at each group atom's grid point, and
// partial sums of those forces (the reference's fsum) and of the atoms
// that got none.
template <typename T>
__global__ void __launch_bounds__(256)
    k_fix_scatter(FixAtoms a, FixFrame fr, const T *__restrict__ fgrid,
                  double *__restrict__ fat, double *__restrict__ part) {
  __shared__ double red[4][4];
  const size_t N = (size_t)fr.nx * fr.ny;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
       i += gridDim.x * blockDim.x) {
    if (!in_group(a, i)) continue;
    const int *g = a.gid + 3 * (size_t)i;
    const int ix = g[0], iy = g[1], iu = g[2];
    if (ix < 0 || ix >= fr.nx || iy < 0 || iy >= fr.ny || iu < 0 ||
        iu >= fr.nl) {
      acc[3] += 1.0;
      continue;
    }
    const size_t site = (size_t)ix * fr.ny + iy;
    for (int c = 0; c < 3; c++) {
      const double fc = (double)fgrid[(size_t)(3 * iu + c) * N + site];
      fat[3 * (size_t)i + c] += fc;
      acc[c] += fc;
    }
  }
  block_sums<4>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; k++) part[(size_t)k * gridDim.x + blockIdx.x] = acc[k];
}

// One block: counters and force sums into the state.
__global__ void __launch_bounds__(1024)
    k_fix_reduce(const double *__restrict__ part, int nblk, int nq, int off,
                 double *__restrict__ st) {
  __shared__ double red[16][4];
  double tot[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < nq; k++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x)
      s += part[(size_t)k * nblk + b];
    tot[k] = s;
  }
  block_sums<4>(tot, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < nq; k++) st[off + k] = tot[k];
}

// fix_center: remove the mean grid force from every group atom.
__global__ void k_fix_center(FixAtoms a, const double *__restrict__ st,
                             int cx, int cy, int cz, double *__restrict__ fat) {
  const double n = st[FS_NATOMS];
  if (!(n > 0.0)) return;  // empty group: nothing to centre
  const double df[3] = {cx ? st[FS_FSUM] / n : 0.0,
                        cy ? st[FS_FSUM + 1] / n : 0.0,
                        cz ? st[FS_FSUM + 2] / n : 0.0};
  const int on[3] = {cx, cy, cz};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
       i += gridDim.x * blockDim.x) {
    if (!in_group(a, i)) continue;
    for (int c = 0; c < 3; c++)
      if (on[c]) fat[3 * (size_t)i + c] -= df[c];
  }
}

static int glue_blocks(int n) {
  const int b = (n + 255) / 256;
  return b < 1 ? 1 : (b > 1024 ? 1024 : b);
}

void launch_fix_com(const FixAtoms &a, const FixFrame &fr, bool setup,
                    double *part, double *st, hipStream_t s) {
  const int nb = glue_blocks(a.n);
  hipLaunchKernelGGL(k_fix_com, dim3(nb), dim3(256), 0, s, a, fr,
                     setup ? 1 : 0, part);
  hipLaunchKernelGGL(k_fix_shift, dim3(1), dim3(1024), 0, s, fr, part, nb,
                     setup ? 1 : 0, st);
}

void launch_fix_gather(const FixAtoms &a, const FixFrame &fr, bool fp64,
                       void *u, void *vg, double *part, double *st,
                       hipStream_t s) {
  const int nb = glue_blocks(a.n);
  if (fp64)
    hipLaunchKernelGGL(k_fix_gather<double>, dim3(nb), dim3(256), 0, s, a, fr,
                       st, (double *)u, (double *)vg, part);
  else
    hipLaunchKernelGGL(k_fix_gather<float>, dim3(nb), dim3(256), 0, s, a, fr,
                       st, (float *)u, (float *)vg, part);
  hipLaunchKernelGGL(k_fix_reduce, dim3(1), dim3(1024), 0, s, part, nb, 2,
                     (int)FS_PLACED, st);
}

void launch_fix_scatter(const FixAtoms &a, const FixFrame &fr, bool fp64,
                        const void *fgrid, double *fat, const int center[3],
                        double *part, double *st, hipStream_t s) {
  const int nb = glue_blocks(a.n);
  if (fp64)
    hipLaunchKernelGGL(k_fix_scatter<double>, dim3(nb), dim3(256), 0, s, a, fr,
                       (const double *)fgrid, fat, part);
  else
    hipLaunchKernelGGL(k_fix_scatter<float>, dim3(nb), dim3(256), 0, s, a, fr,
                       (const float *)fgrid, fat, part);
  hipLaunchKernelGGL(k_fix_reduce, dim3(1), dim3(1024), 0, s, part, nb, 4,
                     (int)FS_FSUM, st);
  if (center[0] || center[1] || center[2])
    hipLaunchKernelGGL(k_fix_center, dim3(nb), dim3(256), 0, s, a, st,
                       center[0], center[1], center[2], fat);
}

}  // namespace gfmd
