// Device-side static relaxation (FIRE + contact); see relax.hip.
#pragma once

#include "gfmd_internal.h"
#include "wall_map.h"

namespace gfmd {

enum {
  CONTACT_NONE = 0,
  CONTACT_LJ_SPHERE = 1,
  CONTACT_HARDWALL = 2,
  CONTACT_WALL_MAP = 3
};

// Scalars of the solver state kept in device memory.
enum {
  S_DT = 0,   // current time step
  S_MIX,      // current mixing factor
  S_CUT,      // steps left before dt may grow again
  S_A,        // mixing: v <- A v + B f
  S_B,
  S_P,        // power v.f of the last step
  S_RES,      // |f_total| (projected), FP64
  S_FEXT,     // |f_ext|
  S_ECON,     // contact energy
  S_ETOT,     // elastic + contact - external work
  S_NCON,     // sites in contact
  S_STEPS,    // FIRE steps taken
  S_INSIDE,   // atoms inside the LJ sphere / too close to the wall map
  S_COUNT
};

// Plain parameters passed by value to the kernels.
struct RelaxDev {
  int nx, ny, d;
  int contact;
  // FIRE (fix_fire.cpp:62-77 names)
  double dt0, dtmax, mix_in, mixdec, incfac, decfac, xlimit, mass;
  int minsteps, adapt_dt;
  // load: normal force per surface site, +z (towards the indenter)
  double pressure;
  // lattice: cell lengths along x and y, position of grid point (0, 0) at
  // rest, periods; atoms 0 .. ncat-1 of a cell feel the sphere, at aoff
  // from the cell's origin
  double ax, ay, x0, y0, z0, lx, ly;
  int ncat;
  double aoff[4][3];
  // LJ sphere: centre, radius, (R + cutoff)^2, 48 e s^12, 24 e s^6,
  // 4 e s^12, 4 e s^6, energy shift at the cutoff
  double sph_c[3], sph_R, sph_rpc2, lj1, lj2, lj3, lj4, lj_off;
  // hard wall: z of the rigid surface per site (T[nx*ny]) + offset
  const void *wall;
  double wall_offset;
  // optional external force field T[d][nx][ny] added to every component
  const void *fext;
  // soft wall over a height map (fix wall/map), on atoms 0 .. ncat-1
  WallMapDev wm;
  // correction problem of a mixed-precision refinement with a nonlinear
  // contact: the FP64 base displacement the correction u is added to, and
  // the contact force at the base, subtracted ([d][nx][ny] each, or null)
  const double *ubase, *fcbase;
  // ... and per site a clearance: while the correction of every atom of the
  // site stays below it (|x|, |y| <= clear_rho and |z| < clear for the wall
  // map, |x| + |y| + |z| < clear for the sphere), each atom is beyond the
  // cutoff both at the base and at base + correction, so its contact force
  // and the base's are exactly zero and neither needs loading (or null)
  const float *clear;
  double clear_rho;
};

// VEC consecutive elements moved as one aligned access (16 bytes for
// 4 x FP32 or 2 x FP64).
template <typename T, int V>
struct alignas(sizeof(T) * V) Vec {
  T e[V];
};
template <typename T, int V>
__device__ __forceinline__ Vec<T, V> vload(const T *p) {
  return *reinterpret_cast<const Vec<T, V> *>(p);
}
template <typename T, int V>
__device__ __forceinline__ void vstore(T *p, const Vec<T, V> &x) {
  *reinterpret_cast<Vec<T, V> *>(p) = x;
}

// Block sums of NQ values at once: one shuffle tree per value, one LDS
// exchange and one barrier for all of them; valid in thread 0.
template <int NQ>
__device__ __forceinline__ void block_sums(double (&acc)[NQ],
                                           double (*red)[NQ]) {
#pragma unroll
  for (int q = 0; q < NQ; q++)
    for (int off = 32; off > 0; off >>= 1)
      acc[q] += __shfl_down(acc[q], off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NQ; q++) red[w][q] = acc[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      double s = 0.0;
      for (int i = 0; i < nw; i++) s += red[i][q];
      acc[q] = s;
    }
  }
}

// Block-wide sums of NQ rows of per-block partials, part[q * nblk + b] for
// q = q0 .. q0 + NQ - 1: every thread issues all of its loads (U per row
// per round, blockDim.x * U blocks per round) before it adds, so a
// reduction of a few thousand partials costs one memory round trip instead
// of one per row and element.  Result valid in thread 0.
template <int NQ, int U>
__device__ __forceinline__ void sum_partials(const double *__restrict__ part,
                                             int nblk, int q0,
                                             double (&tot)[NQ],
                                             double (*red)[NQ]) {
#pragma unroll
  for (int q = 0; q < NQ; q++) tot[q] = 0.0;
  for (int b0 = threadIdx.x; b0 < nblk; b0 += U * blockDim.x) {
    double x[NQ][U];
#pragma unroll
    for (int q = 0; q < NQ; q++)
#pragma unroll
      for (int k = 0; k < U; k++) {
        const int b = b0 + k * blockDim.x;
        x[q][k] = b < nblk ? part[(size_t)(q0 + q) * nblk + b] : 0.0;
      }
#pragma unroll
    for (int q = 0; q < NQ; q++)
#pragma unroll
      for (int k = 0; k < U; k++) tot[q] += x[q][k];
  }
  block_sums<NQ>(tot, red);
}

// Contact force on the surface-layer site (ix, iy) with displacement
// (ux, uy, uz); returns the z-projection flag for hard walls.
__device__ __forceinline__ void contact_force(const RelaxDev &p, int l,
                                              int ix, int iy, double ux,
                                              double uy, double uz,
                                              double &fx, double &fy,
                                              double &fz, double &e,
                                              double &inside) {
  fx = fy = fz = e = 0.0;
  if (p.contact == CONTACT_LJ_SPHERE) {
    // position x = x_eq + u, minimum image in the periodic directions
    double rx = p.x0 + p.ax * ix + p.aoff[l][0] + ux - p.sph_c[0];
    double ry = p.y0 + p.ay * iy + p.aoff[l][1] + uy - p.sph_c[1];
    const double rz = p.z0 + p.aoff[l][2] + uz - p.sph_c[2];
    rx -= p.lx * rint(rx / p.lx);
    ry -= p.ly * rint(ry / p.ly);
    const double rsq = rx * rx + ry * ry + rz * rz;
    if (rsq < p.sph_R * p.sph_R) {
      inside += 1.0;  // fix contact/sphere: atom inside sphere
    } else if (rsq < p.sph_rpc2) {
      const double r = sqrt(rsq);
      const double rinv = 1.0 / (r - p.sph_R);
      const double r2inv = rinv * rinv;
      const double r6inv = r2inv * r2inv * r2inv;
      const double df = r6inv * (p.lj1 * r6inv - p.lj2) * rinv;
      e = r6inv * (p.lj3 * r6inv - p.lj4) - p.lj_off;
      fx = df * rx / r;
      fy = df * ry / r;
      fz = df * rz / r;
    }
  }
}

// Wall-map contact on atom l of site (ix, iy) with displacement
// (ux, uy, uz): FixWallMap::pre_force at the atom's position (the pendist
// rule applied to a copy of z).  Returns whether the atom is within the
// cutoff.
__device__ __forceinline__ bool wall_map_force(const RelaxDev &p, int l,
                                               int ix, int iy, double ux,
                                               double uy, double uz,
                                               double &fx, double &fy,
                                               double &fz, double &e,
                                               double &close) {
#pragma clang fp contract(off)
  const double x = p.x0 + p.ax * ix + p.aoff[l][0] + ux;
  const double y = p.y0 + p.ay * iy + p.aoff[l][1] + uy;
  double z = p.z0 + p.aoff[l][2] + uz;
  double f;
  fx = fy = fz = e = 0.0;
  return wm_atom(p.wm, x, y, z, fx, fy, fz, e, f, close);
}

// The displacement of atom l of site i = (ix, iy), as double: u itself, or
// in a refinement's correction problem the FP64 base plus the correction.
template <typename T>
__device__ __forceinline__ void site_disp(const RelaxDev &p, int l, size_t i,
                                          T ux, T uy, T uz, double &dx,
                                          double &dy, double &dz) {
  dx = (double)ux;
  dy = (double)uy;
  dz = (double)uz;
  if (p.ubase) {
    const size_t N = (size_t)p.nx * p.ny, k = (size_t)(3 * l) * N + i;
    dx = p.ubase[k] + dx;
    dy = p.ubase[k + N] + dy;
    dz = p.ubase[k + 2 * N] + dz;
  }
}

// Whether the correction (ux, uy, uz) of an atom of site i is inside the
// site's clearance (so the atom is out of contact at base and base + u).
template <int CT, typename T>
__device__ __forceinline__ bool site_clear(const RelaxDev &p, size_t i, T ux,
                                           T uy, T uz) {
  if (!p.clear) return false;
  const double c = (double)p.clear[i];
  const double ax = fabs((double)ux), ay = fabs((double)uy),
               az = fabs((double)uz);
  if (CT == CONTACT_LJ_SPHERE) return ax + ay + az < c;
  return ax <= p.clear_rho && ay <= p.clear_rho && az < c;
}

// Contact force (LJ sphere or wall map, CT) on atom l of site i, less the
// base's own contact force in a correction problem.  Returns whether the
// atom is in contact.
template <int CT, typename T>
__device__ __forceinline__ bool site_contact(const RelaxDev &p, int l,
                                             size_t i, int ix, int iy, T ux,
                                             T uy, T uz, double &fx,
                                             double &fy, double &fz,
                                             double &e, double &bad) {
  if (site_clear<CT, T>(p, i, ux, uy, uz)) {
    fx = fy = fz = e = 0.0;
    return false;
  }
  double dx, dy, dz;
  site_disp<T>(p, l, i, ux, uy, uz, dx, dy, dz);
  bool act;
  if (CT == CONTACT_LJ_SPHERE) {
    contact_force(p, l, ix, iy, dx, dy, dz, fx, fy, fz, e, bad);
    act = e != 0.0;
  } else {
    act = wall_map_force(p, l, ix, iy, dx, dy, dz, fx, fy, fz, e, bad);
  }
  if (p.fcbase) {
    const size_t N = (size_t)p.nx * p.ny, k = (size_t)(3 * l) * N + i;
    fx -= p.fcbase[k];
    fy -= p.fcbase[k + N];
    fz -= p.fcbase[k + 2 * N];
  }
  return act;
}

// The pendist rule of the wall map on atom l of site i = (ix, iy): the new
// z displacement (uz itself when the rule does not apply).
template <typename T>
__device__ __forceinline__ T wall_map_reset(const RelaxDev &p, int l,
                                            size_t i, int ix, int iy, T ux,
                                            T uy, T uz) {
#pragma clang fp contract(off)
  if (site_clear<CONTACT_WALL_MAP, T>(p, i, ux, uy, uz)) return uz;
  double dx, dy, dz;
  site_disp<T>(p, l, i, ux, uy, uz, dx, dy, dz);
  const double x = p.x0 + p.ax * ix + p.aoff[l][0] + dx;
  const double y = p.y0 + p.ay * iy + p.aoff[l][1] + dy;
  const double zr = p.z0 + p.aoff[l][2];
  const double z = zr + dz;
  const double zn = wm_reset_z(p.wm, x, y, z);
  if (zn == z) return uz;
  double zb = 0.0;
  if (p.ubase) {
    const size_t N = (size_t)p.nx * p.ny;
    zb = p.ubase[(size_t)(3 * l + 2) * N + i];
  }
  return (T)(zn - zr - zb);
}

struct RelaxBufs {
  void *u, *v, *f, *fel;  // T[d][nx][ny]
  double *state;          // [S_COUNT]
  double *part;           // [16][nblk]
  double *epot;           // elastic energy of the last evaluation
  int nblk;
};

enum {
  STEP_INIT = 0,
  STEP_FORCE = 1,
  STEP_FORCE_KICK = 2,
  STEP_MIX_MOVE = 3,
  STEP_CONTACT_RESET = 4  // wall map: the pendist rule at the start
};

// Damped velocity-Verlet dynamics (dynamics.hip).
struct DynDev {
  double dt, gamma;  // time step, viscous damping rate
  int every;         // energies recorded every `every` states
};
enum {
  DS_FIRST = 0,  // 1 until the first step: v holds v(0), not v(-1/2)
  DS_STEP,       // index n of the next state to record
  DS_KE,         // kinetic energy of the last recorded state
  DS_PE,         // potential energy (elastic + contact - load work)
  DS_NCON,       // sites inside the contact cutoff
  DS_INSIDE,     // atoms inside the LJ sphere / too close to the wall map
  DS_COUNT
};
void launch_dyn_init(double *state, hipStream_t st);
// finish = false: complete v(n) (after the first step), record state n,
// half kick and drift to u(n+1); finish = true: complete v(n) and record.
void launch_dyn_step(const RelaxDev &p, const DynDev &q, const RelaxBufs &b,
                     bool fp64, bool finish, double *hist, hipStream_t st);

int relax_nblocks(int nx, int ny, int ncu);
void launch_convert(const void *src, bool src_fp64, void *dst, bool dst_fp64,
                    size_t n, hipStream_t st);
void launch_relax_step(const RelaxDev &p, const RelaxBufs &b, bool fp64,
                       int which, hipStream_t st);

// Mixed-precision defect correction: from the FP64 state (u64, FP64 elastic
// force fel64, FP64 wall) build the FP32 correction problem: external field
// r = fel64 + pressure (layer 0, z) as T, gap g = wall - u64_z as T, and a
// zero correction.  Then u64 += correction (clamped to the wall).
void launch_refine_setup(const RelaxDev &p64, const double *u64,
                         const double *fel64, const double *wall64,
                         float *r32, float *g32, float *d32, hipStream_t st);
void launch_refine_update(const RelaxDev &p64, double *u64, const float *d32,
                          const double *wall64, hipStream_t st);
// Nonlinear contacts (LJ sphere, wall map): fc = the contact force at the
// FP64 state u64 on the contact atoms' components, and clear = each site's
// clearance for corrections of lateral size up to rho (RelaxDev::clear; a
// negative value where none can be given).  The clearance for rho is the
// largest correction that cannot bring an atom of the site within the
// cutoff: the distance bound at the base less the cutoff, less a margin
// for rounding.
void launch_refine_base(const RelaxDev &p64, const double *u64, double *fc,
                        float *clear, double rho, hipStream_t st);

}  // namespace gfmd
