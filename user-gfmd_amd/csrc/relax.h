// Device-side static relaxation (FIRE + contact); see relax.hip.
#pragma once

#include "gfmd_internal.h"

namespace gfmd {

enum { CONTACT_NONE = 0, CONTACT_LJ_SPHERE = 1, CONTACT_HARDWALL = 2 };

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
  S_INSIDE,   // surface atoms inside the LJ sphere (an error)
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
  // lattice: spacing, position of grid point (0, 0) at rest, periods
  double a, x0, y0, z0, lx, ly;
  // LJ sphere: centre, radius, (R + cutoff)^2, 48 e s^12, 24 e s^6,
  // 4 e s^12, 4 e s^6, energy shift at the cutoff
  double sph_c[3], sph_R, sph_rpc2, lj1, lj2, lj3, lj4, lj_off;
  // hard wall: z of the rigid surface per site (T[nx*ny]) + offset
  const void *wall;
  double wall_offset;
  // optional external force field T[d][nx][ny] added to every component
  const void *fext;
};

struct RelaxBufs {
  void *u, *v, *f, *fel;  // T[d][nx][ny]
  double *state;          // [S_COUNT]
  double *part;           // [16][nblk]
  double *epot;           // elastic energy of the last evaluation
  int nblk;
};

enum { STEP_INIT = 0, STEP_FORCE = 1, STEP_FORCE_KICK = 2, STEP_MIX_MOVE = 3 };

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

}  // namespace gfmd
