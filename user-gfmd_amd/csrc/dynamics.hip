// Damped velocity-Verlet dynamics of the GFMD surface.  The reference runs
// MD with LAMMPS `fix nve` (velocity Verlet) on the `fix gfmd` forces
// (tests/TEST_energy_conservation_fcc100/lammps.in: timestep 0.01, k = 1
// fcc100, total energy within 1e-4); its own dynamic GFMD solver is not
// part of the build (Install.sh:290-296).  Scheme per step, h = dt, with
// optional viscous damping dv/dt = F/m - gamma v:
//   v(n+1/2) = v(n) + h/2 (F(n)/m - gamma v(n))
//   u(n+1)   = u(n) + h v(n+1/2)
//   v(n+1)   = (v(n+1/2) + h/2 F(n+1)/m) / (1 + gamma h/2)
// which is velocity Verlet for gamma = 0.  One fused kernel per step
// completes v(n) from F(n), reduces the kinetic energy and the load terms,
// and makes the next half kick and drift: reads u, v, F(el); writes u, v.
#include "relax.h"

namespace gfmd {

enum { Q_KE = 0, Q_UZ, Q_WEXT, Q_ECON, Q_INSIDE, Q_NCON, NQ_DYN };

// CT: the contact type (the LJ sphere's and the wall map's code and
// registers only in their variants).  The wall map's pendist rule acts on
// the drifted positions (FixWallMap::pre_force precedes the next force).
// VEC consecutive sites per thread with 16-byte accesses.
template <typename T, int CT, int VEC>
__global__ void __launch_bounds__(256)
    k_dyn_step(RelaxDev p, DynDev q, const T *__restrict__ fel,
               T *__restrict__ u, T *__restrict__ v,
               const double *__restrict__ state, double *__restrict__ part,
               int finish) {
  __shared__ double red[4][NQ_DYN];
  const size_t N = (size_t)p.nx * p.ny;
  const size_t NG = N / VEC;
  const int nl = p.d / 3;
  const T *__restrict__ fext = (const T *)p.fext;
  // before the first step v holds v(0); afterwards v(n-1/2), which the
  // force F(n) completes to v(n)
  const bool complete = state[DS_FIRST] == 0.0;
  const double h = q.dt, im = 1.0 / p.mass;
  const double den = 1.0 / (1.0 + 0.5 * q.gamma * h);
  double acc[NQ_DYN];
  for (int k = 0; k < NQ_DYN; k++) acc[k] = 0.0;
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
#pragma unroll
      for (int e = 0; e < VEC; e++) {
        const size_t i = i0 + e;
        double F[3], V[3], U[3];
        for (int c = 0; c < 3; c++) {
          F[c] = (double)fe[c].e[e];
          V[c] = (double)vv[c].e[e];
          U[c] = (double)uu[c].e[e];
          if (fext) {
            const double fx = (double)xe[c].e[e];
            F[c] += fx;
            acc[Q_WEXT] += fx * U[c];
          }
        }
        if (l == 0) {
          F[2] += p.pressure;
          acc[Q_UZ] += U[2];
        }
        if ((CT == CONTACT_LJ_SPHERE || CT == CONTACT_WALL_MAP) &&
            l < p.ncat) {
          const int ix = (int)(i / p.ny), iy = (int)(i - (size_t)ix * p.ny);
          double fx, fy, fz, en;
          if (site_contact<CT, T>(p, l, i, ix, iy, uu[0].e[e], uu[1].e[e],
                                  uu[2].e[e], fx, fy, fz, en, acc[Q_INSIDE]))
            acc[Q_NCON] += 1.0;
          F[0] += fx;
          F[1] += fy;
          F[2] += fz;
          acc[Q_ECON] += en;
        }
        T un[3];
        for (int c = 0; c < 3; c++) {
          if (complete) V[c] = (V[c] + 0.5 * h * F[c] * im) * den;
          acc[Q_KE] += 0.5 * p.mass * V[c] * V[c];
          if (!finish) {
            const double vh = V[c] + 0.5 * h * (F[c] * im - q.gamma * V[c]);
            un[c] = (T)(U[c] + h * vh);
            vv[c].e[e] = (T)vh;
          } else {
            vv[c].e[e] = (T)V[c];
          }
        }
        if (!finish) {
          if (CT == CONTACT_WALL_MAP && l < p.ncat) {
            const int ix = (int)(i / p.ny), iy = (int)(i - (size_t)ix * p.ny);
            un[2] = wall_map_reset<T>(p, l, i, ix, iy, un[0], un[1], un[2]);
          }
          for (int c = 0; c < 3; c++) uu[c].e[e] = un[c];
        }
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const size_t k = (size_t)(3 * l + c) * N + i0;
        vstore<T, VEC>(v + k, vv[c]);
        if (!finish) vstore<T, VEC>(u + k, uu[c]);
      }
    }
  }
  block_sums<NQ_DYN>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < NQ_DYN; k++)
      part[(size_t)k * gridDim.x + blockIdx.x] = acc[k];
}

// One block: energies of state n (elastic energy from the force
// evaluation of u(n), kinetic energy of v(n)) into the history every
// q.every steps, then n <- n + 1.
__global__ void __launch_bounds__(1024)
    k_dyn_record(RelaxDev p, DynDev q, const double *__restrict__ part,
                 int nblk, const double *__restrict__ epot,
                 double *__restrict__ state, double *__restrict__ hist) {
  __shared__ double red[16][NQ_DYN];
  double tot[NQ_DYN];
  sum_partials<NQ_DYN, 2>(part, nblk, 0, tot, red);
  if (threadIdx.x != 0) return;
  const double ke = tot[Q_KE];
  const double pe = *epot + tot[Q_ECON] - p.pressure * tot[Q_UZ] - tot[Q_WEXT];
  const long n = (long)state[DS_STEP];
  if (hist && n % q.every == 0) {
    double *o = hist + 3 * (n / q.every);
    o[0] = ke;
    o[1] = pe;
    o[2] = ke + pe;
  }
  state[DS_KE] = ke;
  state[DS_PE] = pe;
  state[DS_NCON] = tot[Q_NCON];
  state[DS_INSIDE] += tot[Q_INSIDE];
  state[DS_STEP] = (double)(n + 1);
  state[DS_FIRST] = 0.0;
}

__global__ void k_dyn_init(double *state) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    for (int k = 0; k < DS_COUNT; k++) state[k] = 0.0;
    state[DS_FIRST] = 1.0;
  }
}

void launch_dyn_init(double *state, hipStream_t st) {
  hipLaunchKernelGGL(k_dyn_init, dim3(1), dim3(64), 0, st, state);
}

void launch_dyn_step(const RelaxDev &p, const DynDev &q, const RelaxBufs &b,
                     bool fp64, bool finish, double *hist, hipStream_t st) {
  const int nb = b.nblk;
  const int ct = p.contact;
  const size_t N = (size_t)p.nx * p.ny;
#define DK(TT, CT) \
  (N % (16 / sizeof(TT)) == 0 ? k_dyn_step<TT, CT, 16 / sizeof(TT)> \
                              : k_dyn_step<TT, CT, 1>)
  if (fp64) {
    auto kern = ct == CONTACT_LJ_SPHERE
                    ? DK(double, CONTACT_LJ_SPHERE)
                    : (ct == CONTACT_WALL_MAP ? DK(double, CONTACT_WALL_MAP)
                                              : DK(double, CONTACT_NONE));
    hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, st, p, q,
                       (const double *)b.fel, (double *)b.u, (double *)b.v,
                       b.state, b.part, (int)finish);
  } else {
    auto kern = ct == CONTACT_LJ_SPHERE
                    ? DK(float, CONTACT_LJ_SPHERE)
                    : (ct == CONTACT_WALL_MAP ? DK(float, CONTACT_WALL_MAP)
                                              : DK(float, CONTACT_NONE));
    hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, st, p, q,
                       (const float *)b.fel, (float *)b.u, (float *)b.v,
                       b.state, b.part, (int)finish);
  }
#undef DK
  hipLaunchKernelGGL(k_dyn_record, dim3(1), dim3(1024), 0, st, p, q, b.part,
                     nb, b.epot, b.state, hist);
}

}  // namespace gfmd
