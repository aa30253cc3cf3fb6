// Static relaxation on the device: FIRE (reference semantics of
// src/cuda/fix_fire.cpp:117-297, FixFIRE::initial_integrate /
// final_integrate) driving the elastic force evaluation, with the contact
// models of the reference's tests fused into the per-step kernels:
//   - LJ 12-6 wall in the distance to a sphere surface
//     (FixContactSphere::post_force, src/extras/fix_contact_sphere.cpp:
//     207-264), acting on the surface layer;
//   - a rigid hard wall z_wall(x, y) (sphere, flat punch or rough map),
//     enforced by projection (u_z <= z_wall) with the constraint force taken
//     up where the surface touches;
//   - a soft wall over a height map (FixWallMap::pre_force, src/extras/
//     fix_wall_map.cpp:686-811), acting on the contact atoms;
//   - a uniform normal force per surface site (load control through the
//     q = 0 mode).
// All reductions (powers, norms, energies) are FP64 and deterministic
// (fixed-order block partials reduced by one block).
#include "relax.h"

namespace gfmd {

// Per-block partial sums written by k_force_kick.
enum {
  P_VF = 0, P_VV, P_FF, P_RES2, P_ECON, P_UZ, P_NCON, P_FEXT2, P_INSIDE,
  P_WEXT, NPART
};

// After an elastic evaluation fel: total force f = fel + f_ext + f_contact,
// projected at hard-wall contacts, then the FIRE half kick
// (final_integrate) v += dt/2 f/m with the |v| limit, and the partial sums
// of v.f, v.v, f.f, |f|^2 (residual), contact energy, sum u_z, contacts.
// CT: the contact type, a template argument so that the LJ sphere's code
// (and its registers) is only compiled into the variant that uses it.
// VEC consecutive sites per thread, moved with VEC-wide loads and stores
// (16 bytes: 4 FP32 or 2 FP64 sites), so that each wave keeps enough bytes
// in flight; the per-site arithmetic does not depend on VEC.
template <typename T, int CT, int VEC>
__global__ void __launch_bounds__(256)
    k_force_kick(RelaxDev p, const T *__restrict__ fel, const T *__restrict__ u,
                 T *__restrict__ v, T *__restrict__ f,
                 const double *__restrict__ state, double *__restrict__ part,
                 int kick) {
  __shared__ double red[4][NPART];
  const size_t N = (size_t)p.nx * p.ny;
  const size_t NG = N / VEC;
  const int nl = p.d / 3;
  const T *__restrict__ wall = (const T *)p.wall;
  const T *__restrict__ fext = (const T *)p.fext;
  const double dtf = kick ? 0.5 * state[S_DT] / p.mass : 0.0;
  const double vlim2 = p.xlimit > 0 ? (p.xlimit / state[S_DT]) *
                                          (p.xlimit / state[S_DT])
                                    : 0.0;
  double acc[NPART];
  for (int q = 0; q < NPART; q++) acc[q] = 0.0;
  for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < NG;
       g += (size_t)gridDim.x * blockDim.x) {
    const size_t i0 = g * VEC;
    for (int l = 0; l < nl; l++) {
      Vec<T, VEC> fe[3], vv[3], uu[3], xe[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t k = (size_t)(3 * l + c) * N + i0;
        fe[c] = vload<T, VEC>(fel + k);
        vv[c] = vload<T, VEC>(v + k);
        uu[c] = vload<T, VEC>(u + k);
        if (fext) xe[c] = vload<T, VEC>(fext + k);
      }
      Vec<T, VEC> zw;
      if (CT == CONTACT_HARDWALL && l == 0) zw = vload<T, VEC>(wall + i0);
#pragma unroll
      for (int e = 0; e < VEC; e++) {
        const size_t i = i0 + e;
        double F[3], V[3];
        for (int c = 0; c < 3; c++) {
          F[c] = (double)fe[c].e[e];
          V[c] = (double)vv[c].e[e];
          if (fext) {
            const double fx = (double)xe[c].e[e];
            F[c] += fx;
            acc[P_FEXT2] += fx * fx;
            acc[P_WEXT] += fx * (double)uu[c].e[e];
          }
        }
        const T uz = uu[2].e[e];
        if ((CT == CONTACT_LJ_SPHERE || CT == CONTACT_WALL_MAP) &&
            l < p.ncat) {
          const int ix = (int)(i / p.ny), iy = (int)(i - (size_t)ix * p.ny);
          double fx, fy, fz, en;
          if (site_contact<CT, T>(p, l, i, ix, iy, uu[0].e[e], uu[1].e[e], uz,
                                  fx, fy, fz, en, acc[P_INSIDE]))
            acc[P_NCON] += 1.0;
          F[0] += fx;
          F[1] += fy;
          F[2] += fz;
          acc[P_ECON] += en;
        }
        if (l == 0) {
          F[2] += p.pressure;
          acc[P_FEXT2] += p.pressure * p.pressure;
          acc[P_UZ] += (double)uz;
          if (CT == CONTACT_HARDWALL) {
            // at the wall the constraint takes up any push into it
            const T z = (T)((double)zw.e[e] + p.wall_offset);
            if (uz >= z) {
              acc[P_NCON] += 1.0;
              if (F[2] > 0.0) F[2] = 0.0;
              if (V[2] > 0.0) V[2] = 0.0;
            }
          }
        }
        double vsq = 0.0;
        for (int c = 0; c < 3; c++) {
          V[c] += dtf * F[c];
          vsq += V[c] * V[c];
        }
        if (vlim2 > 0.0 && vsq > vlim2) {
          const double sc = sqrt(vlim2 / vsq);
          for (int c = 0; c < 3; c++) V[c] *= sc;
        }
        for (int c = 0; c < 3; c++) {
          fe[c].e[e] = (T)F[c];
          vv[c].e[e] = (T)V[c];
          acc[P_VF] += V[c] * F[c];
          acc[P_VV] += V[c] * V[c];
          acc[P_FF] += F[c] * F[c];
        }
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t k = (size_t)(3 * l + c) * N + i0;
        vstore<T, VEC>(f + k, fe[c]);
        vstore<T, VEC>(v + k, vv[c]);
      }
    }
  }
  acc[P_RES2] = acc[P_FF];
  block_sums<NPART>(acc, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < NPART; q++)
      part[(size_t)q * gridDim.x + blockIdx.x] = acc[q];
}

// One block: reduce the partials, run the FIRE state machine
// (fix_fire.cpp:246-296) and record the step's scalars.  fire = 0 only
// records the scalars (initial force, no final_integrate yet).
__global__ void __launch_bounds__(1024)
    k_fire_update(RelaxDev p, const double *__restrict__ part, int nblk,
                  const double *__restrict__ epot, double *__restrict__ state,
                  int fire) {
  __shared__ double red[16][NPART];
  double tot[NPART];
  sum_partials<NPART, 2>(part, nblk, 0, tot, red);
  if (threadIdx.x != 0) return;
  state[S_RES] = sqrt(tot[P_RES2]);
  state[S_FEXT] = sqrt(tot[P_FEXT2]);
  state[S_ECON] = tot[P_ECON];
  // total potential energy: elastic + contact - work of the external force
  state[S_ETOT] = *epot + tot[P_ECON] - p.pressure * tot[P_UZ] - tot[P_WEXT];
  state[S_NCON] = tot[P_NCON];
  state[S_INSIDE] = tot[P_INSIDE];
  if (!fire) {
    state[S_A] = 1.0;
    state[S_B] = 0.0;
    return;
  }
  const double vf = tot[P_VF], vv = tot[P_VV], ff = tot[P_FF];
  double mix = state[S_MIX];
  const double help = ff > 0.0 ? mix * sqrt(vv / ff) : 0.0;
  state[S_A] = 1.0 - mix;  // v = a v + b f
  state[S_B] = help;
  double dt = state[S_DT];
  double cut = state[S_CUT];
  if (vf < 0.0) {
    state[S_A] = 0.0;
    state[S_B] = 0.0;
    cut = p.minsteps;
    dt = dt * p.decfac;
    mix = p.mix_in;
  } else {
    if (cut < 0) {
      dt = fmin(dt * p.incfac, p.dtmax);
      mix = mix * p.mixdec;
    } else {
      cut = cut - 1;
    }
  }
  if (!p.adapt_dt) dt = p.dt0;  // reference: reset_dt restores update->dt
  state[S_DT] = dt;
  state[S_MIX] = mix;
  state[S_CUT] = cut;
  state[S_P] = vf;
  state[S_STEPS] += 1.0;
}

// FIRE mixing (v = a v + b f, or v = 0 after a power reversal) followed by
// the next step's initial_integrate: v += dt/2 f/m, |v| limit, u += dt v,
// and the hard-wall projection u_z <= z_wall.  WM: the wall map's pendist
// rule on the moved contact atoms (FixWallMap::pre_force resets them before
// the next force evaluation).  VEC consecutive sites per thread.
template <typename T, bool WM, int VEC>
__global__ void __launch_bounds__(256)
    k_mix_move(RelaxDev p, T *__restrict__ u, T *__restrict__ v,
               const T *__restrict__ f, const double *__restrict__ state) {
  const T *__restrict__ wall = (const T *)p.wall;
  const size_t N = (size_t)p.nx * p.ny;
  const size_t NG = N / VEC;
  const int nl = p.d / 3;
  const double A = state[S_A], Bm = state[S_B];
  const double dt = state[S_DT];
  const double dtf = 0.5 * dt / p.mass;
  const double vlim2 = p.xlimit > 0 ? (p.xlimit / dt) * (p.xlimit / dt) : 0.0;
  for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < NG;
       g += (size_t)gridDim.x * blockDim.x) {
    const size_t i0 = g * VEC;
    for (int l = 0; l < nl; l++) {
      Vec<T, VEC> ff[3], vv[3], uu[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t k = (size_t)(3 * l + c) * N + i0;
        ff[c] = vload<T, VEC>(f + k);
        vv[c] = vload<T, VEC>(v + k);
        uu[c] = vload<T, VEC>(u + k);
      }
      Vec<T, VEC> zw;
      if (l == 0 && p.contact == CONTACT_HARDWALL) zw = vload<T, VEC>(wall + i0);
#pragma unroll
      for (int e = 0; e < VEC; e++) {
        double V[3], F[3];
        double vsq = 0.0;
        for (int c = 0; c < 3; c++) {
          F[c] = (double)ff[c].e[e];
          V[c] = A * (double)vv[c].e[e] + Bm * F[c];
          V[c] += dtf * F[c];
          vsq += V[c] * V[c];
        }
        if (vlim2 > 0.0 && vsq > vlim2) {
          const double sc = sqrt(vlim2 / vsq);
          for (int c = 0; c < 3; c++) V[c] *= sc;
        }
        T un[3];
        for (int c = 0; c < 3; c++) un[c] = (T)((double)uu[c].e[e] + dt * V[c]);
        if (l == 0 && p.contact == CONTACT_HARDWALL) {
          const T z = (T)((double)zw.e[e] + p.wall_offset);
          if (un[2] > z) {
            un[2] = z;
            if (V[2] > 0.0) V[2] = 0.0;
          }
        }
        if (WM && l < p.ncat) {
          const size_t i = i0 + e;
          const int ix = (int)(i / p.ny), iy = (int)(i - (size_t)ix * p.ny);
          un[2] = wall_map_reset<T>(p, l, i, ix, iy, un[0], un[1], un[2]);
        }
        for (int c = 0; c < 3; c++) {
          uu[c].e[e] = un[c];
          vv[c].e[e] = (T)V[c];
        }
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t k = (size_t)(3 * l + c) * N + i0;
        vstore<T, VEC>(u + k, uu[c]);
        vstore<T, VEC>(v + k, vv[c]);
      }
    }
  }
}

// The wall map's pendist rule on the contact atoms before the first force
// evaluation (min_setup calls pre_force).
template <typename T>
__global__ void __launch_bounds__(256)
    k_wall_map_reset(RelaxDev p, T *__restrict__ u) {
  const size_t N = (size_t)p.nx * p.ny;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
       i += (size_t)gridDim.x * blockDim.x) {
    const int ix = (int)(i / p.ny), iy = (int)(i - (size_t)ix * p.ny);
    for (int l = 0; l < p.ncat; l++) {
      const size_t kz = (size_t)(3 * l + 2) * N + i;
      u[kz] = wall_map_reset<T>(p, l, i, ix, iy, u[(size_t)(3 * l) * N + i],
                                u[(size_t)(3 * l + 1) * N + i], u[kz]);
    }
  }
}

template <typename T>
__global__ void k_init_state(RelaxDev p, double *state, T *v) {
  const size_t n = (size_t)p.d * p.nx * p.ny;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    v[i] = (T)0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int k = 0; k < S_COUNT; k++) state[k] = 0.0;
    state[S_DT] = p.dt0;
    state[S_MIX] = p.mix_in;
    state[S_CUT] = p.minsteps;
    state[S_RES] = 1e300;
    state[S_A] = 1.0;
  }
}

template <typename A, typename B>
__global__ void k_convert(const A *__restrict__ src, B *__restrict__ dst,
                          size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = (B)src[i];
}

void launch_convert(const void *src, bool src_fp64, void *dst, bool dst_fp64,
                    size_t n, hipStream_t st) {
  size_t b = (n + 255) / 256;
  const unsigned nb = (unsigned)(b < 8192 ? b : 8192);
  if (src_fp64 && dst_fp64)
    hipLaunchKernelGGL((k_convert<double, double>), dim3(nb), dim3(256), 0, st,
                       (const double *)src, (double *)dst, n);
  else if (src_fp64)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(nb), dim3(256), 0, st,
                       (const double *)src, (float *)dst, n);
  else if (dst_fp64)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(nb), dim3(256), 0, st,
                       (const float *)src, (double *)dst, n);
  else
    hipLaunchKernelGGL((k_convert<float, float>), dim3(nb), dim3(256), 0, st,
                       (const float *)src, (float *)dst, n);
}

__global__ void k_refine_setup(RelaxDev p, const double *__restrict__ u64,
                               const double *__restrict__ fel64,
                               const double *__restrict__ wall64,
                               float *__restrict__ r32, float *__restrict__ g32,
                               float *__restrict__ d32) {
  const size_t N = (size_t)p.nx * p.ny;
  const size_t n = (size_t)p.d * N;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    double r = fel64[k];
    if (k >= 2 * N && k < 3 * N) {
      r += p.pressure;
      if (wall64) {
        const size_t i = k - 2 * N;
        g32[i] = (float)(wall64[i] + p.wall_offset - u64[k]);
      }
    }
    r32[k] = (float)r;
    d32[k] = 0.0f;
  }
}

__global__ void k_refine_update(RelaxDev p, double *__restrict__ u64,
                                const float *__restrict__ d32,
                                const double *__restrict__ wall64) {
  const size_t N = (size_t)p.nx * p.ny;
  const size_t n = (size_t)p.d * N;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    double un = u64[k] + (double)d32[k];
    if (wall64 && k >= 2 * N && k < 3 * N) {
      const double zw = wall64[k - 2 * N] + p.wall_offset;
      if (un > zw) un = zw;
    }
    u64[k] = un;
  }
}

// Clearance of a wall-map atom at (x, y, z) for lateral moves up to rho:
// the smallest distance bound over the bound tiles of every spline box the
// moved atom can fall in (one box of slack either side for the rounding of
// x qx), less the cutoff and a rounding margin.  Negative where no bound
// applies (analytic maps, no cutoff, an implausibly wide box range).
static __device__ double wm_clearance(const WallMapDev &w, double x, double y,
                                      double z, double rho) {
#pragma clang fp contract(off)
  if (w.map != WM_TABLE || !w.tiles || !(w.cutoff > 0.0)) return -1.0;
  const double X0 = floor((x - rho) * w.qx) - 1.0;
  const double X1 = floor((x + rho) * w.qx) + 1.0;
  const double Y0 = floor((y - rho) * w.qy) - 1.0;
  const double Y1 = floor((y + rho) * w.qy) + 1.0;
  if (!(fabs(X0) < 1e15 && fabs(X1) < 1e15 && fabs(Y0) < 1e15 &&
        fabs(Y1) < 1e15 && X1 - X0 <= 8.0 && Y1 - Y0 <= 8.0))
    return -1.0;
  double lo = INFINITY, hi = -INFINITY;
  for (long long X = (long long)X0; X <= (long long)X1; X++) {
    long long bx = X % w.mx;
    if (bx < 0) bx += w.mx;
    for (long long Y = (long long)Y0; Y <= (long long)Y1; Y++) {
      long long by = Y % w.my;
      if (by < 0) by += w.my;
      const double *t =
          w.tiles + 2 * ((size_t)(bx / kTile) * w.nty + (size_t)(by / kTile));
      lo = fmin(lo, t[0]);
      hi = fmax(hi, t[1]);
    }
  }
  const double ref = w.norm < 0 ? lo : hi;
  const double b = w.norm < 0 ? lo - z : z - hi;
  return b - w.cutoff - 1e-9 * (1.0 + fabs(z) + fabs(ref));
}

// Contact force at the FP64 base displacement u64 (contact atoms only):
// the term the correction problem of a nonlinear contact subtracts; and
// each site's clearance (the smallest over its contact atoms), rounded
// down to FP32.
template <int CT>
__global__ void k_refine_base(RelaxDev p, const double *__restrict__ u64,
                              double *__restrict__ fc,
                              float *__restrict__ clear, double rho) {
#pragma clang fp contract(off)
  const size_t N = (size_t)p.nx * p.ny;
  double bad = 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N;
       i += (size_t)gridDim.x * blockDim.x) {
    const int ix = (int)(i / p.ny), iy = (int)(i - (size_t)ix * p.ny);
    double cmin = INFINITY;
    for (int l = 0; l < p.ncat; l++) {
      const size_t k = (size_t)(3 * l) * N + i;
      const double ux = u64[k], uy = u64[k + N], uz = u64[k + 2 * N];
      double fx, fy, fz, e;
      site_contact<CT, double>(p, l, i, ix, iy, ux, uy, uz, fx, fy, fz, e,
                               bad);
      fc[k] = fx;
      fc[k + N] = fy;
      fc[k + 2 * N] = fz;
      double c;
      if (CT == CONTACT_LJ_SPHERE) {
        // distance to the nearest image of the centre, less R + cutoff
        double rx = p.x0 + p.ax * ix + p.aoff[l][0] + ux - p.sph_c[0];
        double ry = p.y0 + p.ay * iy + p.aoff[l][1] + uy - p.sph_c[1];
        const double rz = p.z0 + p.aoff[l][2] + uz - p.sph_c[2];
        rx -= p.lx * rint(rx / p.lx);
        ry -= p.ly * rint(ry / p.ly);
        const double r = sqrt(rx * rx + ry * ry + rz * rz);
        const double rc = sqrt(p.sph_rpc2);
        c = r - rc - 1e-9 * (1.0 + r + rc);
      } else {
        c = wm_clearance(p.wm, p.x0 + p.ax * ix + p.aoff[l][0] + ux,
                         p.y0 + p.ay * iy + p.aoff[l][1] + uy,
                         p.z0 + p.aoff[l][2] + uz, rho);
      }
      cmin = c < cmin ? c : cmin;  // a NaN clearance leaves cmin alone ...
      if (c != c) cmin = -1.0;     // ... and then forbids the skip
    }
    float cf = (float)cmin;
    if ((double)cf > cmin) cf = nextafterf(cf, -INFINITY);
    clear[i] = cmin > 0.0 ? cf : -1.0f;
  }
}

void launch_refine_base(const RelaxDev &p64, const double *u64, double *fc,
                        float *clear, double rho, hipStream_t st) {
  RelaxDev p = p64;
  p.ubase = p.fcbase = nullptr;
  p.clear = nullptr;
  if (p.contact == CONTACT_LJ_SPHERE)
    hipLaunchKernelGGL(k_refine_base<CONTACT_LJ_SPHERE>, dim3(4096), dim3(256),
                       0, st, p, u64, fc, clear, rho);
  else if (p.contact == CONTACT_WALL_MAP)
    hipLaunchKernelGGL(k_refine_base<CONTACT_WALL_MAP>, dim3(4096), dim3(256),
                       0, st, p, u64, fc, clear, rho);
}

void launch_refine_setup(const RelaxDev &p64, const double *u64,
                         const double *fel64, const double *wall64,
                         float *r32, float *g32, float *d32, hipStream_t st) {
  hipLaunchKernelGGL(k_refine_setup, dim3(4096), dim3(256), 0, st, p64, u64,
                     fel64, wall64, r32, g32, d32);
}

void launch_refine_update(const RelaxDev &p64, double *u64, const float *d32,
                          const double *wall64, hipStream_t st) {
  hipLaunchKernelGGL(k_refine_update, dim3(4096), dim3(256), 0, st, p64, u64,
                     d32, wall64);
}

// Blocks of the elementwise solver passes (FIRE, dynamics, contact CG):
// up to 8 per CU, so enough waves are in flight to hide the memory latency
// of their 4-byte (FP32) accesses; every pass ends in per-block partial
// sums that one block reduces.
int relax_nblocks(int nx, int ny, int ncu) {
  const size_t N = (size_t)nx * ny;
  size_t b = (N + 255) / 256;
  const size_t cap = (size_t)ncu * 8;
  return (int)(b < cap ? b : cap);
}

template <typename T>
static void launch_steps_t(const RelaxDev &p, const RelaxBufs &b, int which,
                           hipStream_t st) {
  constexpr int VW = 16 / sizeof(T);  // sites per 16-byte access
  const int nb = b.nblk;
  if (which == STEP_INIT) {
    hipLaunchKernelGGL(k_init_state<T>, dim3(nb), dim3(256), 0, st, p,
                       b.state, (T *)b.v);
  } else if (which == STEP_FORCE || which == STEP_FORCE_KICK) {
    const int kick = which == STEP_FORCE_KICK;
    const bool vec = ((size_t)p.nx * p.ny) % VW == 0;
#define FK(CT) (vec ? k_force_kick<T, CT, VW> : k_force_kick<T, CT, 1>)
    auto kern = p.contact == CONTACT_LJ_SPHERE
                    ? FK(CONTACT_LJ_SPHERE)
                    : (p.contact == CONTACT_HARDWALL
                           ? FK(CONTACT_HARDWALL)
                           : (p.contact == CONTACT_WALL_MAP
                                  ? FK(CONTACT_WALL_MAP)
                                  : FK(CONTACT_NONE)));
#undef FK
    hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, st, p, (const T *)b.fel,
                       (const T *)b.u, (T *)b.v, (T *)b.f, b.state, b.part,
                       kick);
    hipLaunchKernelGGL(k_fire_update, dim3(1), dim3(1024), 0, st, p, b.part,
                       nb, b.epot, b.state, kick);
  } else if (which == STEP_MIX_MOVE) {
    const bool vec = ((size_t)p.nx * p.ny) % VW == 0;
    const bool wm = p.contact == CONTACT_WALL_MAP;
    auto kern = vec ? (wm ? k_mix_move<T, true, VW> : k_mix_move<T, false, VW>)
                    : (wm ? k_mix_move<T, true, 1> : k_mix_move<T, false, 1>);
    hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, st, p, (T *)b.u,
                       (T *)b.v, (const T *)b.f, b.state);
  } else if (which == STEP_CONTACT_RESET) {
    if (p.contact == CONTACT_WALL_MAP)
      hipLaunchKernelGGL(k_wall_map_reset<T>, dim3(nb), dim3(256), 0, st, p,
                         (T *)b.u);
  }
}

void launch_relax_step(const RelaxDev &p, const RelaxBufs &b, bool fp64,
                       int which, hipStream_t st) {
  if (fp64) launch_steps_t<double>(p, b, which, st);
  else launch_steps_t<float>(p, b, which, st);
}

}  // namespace gfmd
