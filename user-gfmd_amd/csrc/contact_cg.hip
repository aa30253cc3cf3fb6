// Hard-wall contact by constrained conjugate gradients on the contact
// pressure (Polonsky & Keer, Wear 231 (1999) 206, fixed-approach form).
//
// Frictionless rigid wall above the surface layer: sites satisfy
// u_z <= z_wall, the wall pushes with pressure p >= 0 (in -z), and the
// load pushes every surface site with p0 (in +z).  With the surface
// compliance G = Phi^-1 (finite at q = 0 for a finite-height substrate,
// whose bottom is held), u_z = G_zz * (p0 - p) = c - G_zz * p with the
// uniform c = G_zz(q=0) p0, so the gap is
//     g = (z_wall - c) + G_zz * p  >= 0,   p >= 0,   p g = 0,
// the complementarity problem of  min 1/2 p.G.p + p.h  over p >= 0.  Each
// iteration: u = G*p, g = u + h, conjugate direction t on the contact set
// {p > 0}, r = G*t, step tau = (g.t)/(r.t), projection onto p >= 0, and
// sites with p = 0 and g < 0 (overlap) take pressure -tau g, which also
// restarts the conjugacy.  The convolutions run on the force sweeps with a
// scalar compliance table; the kernels here are the fused elementwise
// passes and their FP64 reductions.
//
// The reference relaxes contacts with FIRE (src/cuda/fix_fire.cpp) and
// CG (src/extras/min_cg_gfmd.cpp, preconditioned by (Phi + C)^-1,
// gfmd_misc.h:38-133); this solver reaches the same unique equilibrium of
// the convex hard-wall problem, checked by the same projected-force
// residual.
#include "contact_cg.h"
#include "relax.h"

namespace gfmd {

enum { Q_GG = 0, Q_NC, Q_GT, Q_RT, Q_DP, Q_PS, Q_OVL, NQ_PK };

// blockIdx.y is the member of an ensemble: its fields start `off` elements
// after the base pointers, its partial sums 8 * gridDim.x after the last
// member's (the reduction kernels use gridDim.x blocks per member)
#define MEMBER_OFFSETS                              \
  const size_t off = (size_t)blockIdx.y * n;        \
  part += (size_t)blockIdx.y * 8 * gridDim.x

// g = u + h; contact-set sums of g^2 and the contact count.
template <typename T>
__global__ void __launch_bounds__(256)
    k_pk_gap(size_t n, const T *__restrict__ u, const T *__restrict__ h,
             const T *__restrict__ p, T *__restrict__ g,
             double *__restrict__ part, const int *__restrict__ gate) {
  __shared__ double red[4][2];
  if (gate && gate[blockIdx.y] == 0) return;
  MEMBER_OFFSETS;
  u += off;
  h += off;
  p += off;
  g += off;
  double acc[2] = {0.0, 0.0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const T gi = u[i] + h[i];
    g[i] = gi;
    if (p[i] > (T)0) {
      acc[0] += (double)gi * (double)gi;
      acc[1] += 1.0;
    }
  }
  block_sums<2>(acc, red);
  if (threadIdx.x == 0) {
    part[(size_t)Q_GG * gridDim.x + blockIdx.x] = acc[0];
    part[(size_t)Q_NC * gridDim.x + blockIdx.x] = acc[1];
  }
}

// t = g + beta t on the contact set, 0 elsewhere.
template <typename T>
__global__ void __launch_bounds__(256)
    k_pk_dir(size_t n, const T *__restrict__ g, const T *__restrict__ p,
             T *__restrict__ t, const double *__restrict__ state,
             const int *__restrict__ gate) {
  if (gate && gate[blockIdx.y] == 0) return;
  const size_t off = blockIdx.y * n;
  g += off;
  p += off;
  t += off;
  state += blockIdx.y * PK_COUNT;
  const T beta = (T)state[PK_BETA];
  const bool empty = state[PK_EMPTY] != 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const bool act = p[i] > (T)0 || (empty && g[i] < (T)0);
    t[i] = act ? g[i] + beta * t[i] : (T)0;
  }
}

// contact-set sums of g.t and r.t (r = G*t)
template <typename T>
__global__ void __launch_bounds__(256)
    k_pk_curv(size_t n, const T *__restrict__ g, const T *__restrict__ t,
              const T *__restrict__ r, const T *__restrict__ p,
              double *__restrict__ part, const int *__restrict__ gate) {
  __shared__ double red[4][2];
  if (gate && gate[blockIdx.y] == 0) return;
  MEMBER_OFFSETS;
  g += off;
  t += off;
  r += off;
  p += off;
  double acc[2] = {0.0, 0.0};
  // t is zero off the direction's set, so the sums over {t != 0} are the
  // sums over the contact set (or the overlap set when that is empty)
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    if (t[i] != (T)0) {
      const double ti = (double)t[i];
      acc[0] += (double)g[i] * ti;
      acc[1] += (double)r[i] * ti;
    }
  }
  block_sums<2>(acc, red);
  if (threadIdx.x == 0) {
    part[(size_t)Q_GT * gridDim.x + blockIdx.x] = acc[0];
    part[(size_t)Q_RT * gridDim.x + blockIdx.x] = acc[1];
  }
}

// p <- max(p - tau t, 0) on the contact set; overlap sites (p = 0, g < 0)
// take -tau g.  Sums of |dp|, p and the overlap count.
template <typename T>
__global__ void __launch_bounds__(256)
    k_pk_update(size_t n, T *__restrict__ p, const T *__restrict__ t,
                const T *__restrict__ g, const double *__restrict__ state,
                double *__restrict__ part, const int *__restrict__ gate) {
  __shared__ double red[4][3];
  if (gate && gate[blockIdx.y] == 0) return;
  MEMBER_OFFSETS;
  p += off;
  t += off;
  g += off;
  state += blockIdx.y * PK_COUNT;
  const T tau = (T)state[PK_TAU];
  double acc[3] = {0.0, 0.0, 0.0};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const T pi = p[i];
    T pn = pi > (T)0 ? pi - tau * t[i] : (T)0;
    if (pn < (T)0) pn = (T)0;
    if (pn == (T)0 && g[i] < (T)0) {
      pn = -tau * g[i];
      acc[2] += 1.0;
    }
    p[i] = pn;
    acc[0] += fabs((double)pn - (double)pi);
    acc[1] += (double)pn;
  }
  block_sums<3>(acc, red);
  if (threadIdx.x == 0) {
    part[(size_t)Q_DP * gridDim.x + blockIdx.x] = acc[0];
    part[(size_t)Q_PS * gridDim.x + blockIdx.x] = acc[1];
    part[(size_t)Q_OVL * gridDim.x + blockIdx.x] = acc[2];
  }
}

// which = 0: after k_pk_gap (beta); 1: after k_pk_curv (tau);
// 2: after k_pk_update (conjugacy flag, convergence measures).
__global__ void __launch_bounds__(256)
    k_pk_state(const double *__restrict__ part, int nblk,
               double *__restrict__ state, int which,
               const int *__restrict__ gate) {
  __shared__ double red2[4][2];
  __shared__ double red3[4][3];
  if (gate && gate[blockIdx.y] == 0) return;
  part += (size_t)blockIdx.y * 8 * nblk;
  state += blockIdx.y * PK_COUNT;
  if (which == 0) {
    double s[2];
    static_assert(Q_NC == Q_GG + 1, "partial rows");
    sum_partials<2, 8>(part, nblk, Q_GG, s, red2);
    const double gg = s[0], nc = s[1];
    if (threadIdx.x == 0) {
      const double gold = state[PK_GOLD];
      state[PK_BETA] = (state[PK_CONJ] != 0.0 && gold > 0.0 && nc > 0.0)
                           ? gg / gold
                           : 0.0;
      state[PK_GG] = gg;
      state[PK_NC] = nc;
      state[PK_EMPTY] = nc == 0.0 ? 1.0 : 0.0;
    }
  } else if (which == 1) {
    double s[2];
    static_assert(Q_RT == Q_GT + 1, "partial rows");
    sum_partials<2, 8>(part, nblk, Q_GT, s, red2);
    if (threadIdx.x == 0) state[PK_TAU] = s[1] > 0.0 ? s[0] / s[1] : 0.0;
  } else {
    double s[3];
    static_assert(Q_PS == Q_DP + 1 && Q_OVL == Q_DP + 2, "partial rows");
    sum_partials<3, 8>(part, nblk, Q_DP, s, red3);
    if (threadIdx.x == 0) {
      state[PK_CONJ] = s[2] == 0.0 ? 1.0 : 0.0;
      state[PK_GOLD] = state[PK_GG];
      state[PK_DP] = s[0];
      state[PK_PSUM] = s[1];
      state[PK_OVL] = s[2];
      state[PK_ITER] += 1.0;
    }
  }
}

template <typename T>
__global__ void k_pk_init(size_t n, T *__restrict__ p, T *__restrict__ t,
                          const T *__restrict__ p_init, double p0,
                          const double *__restrict__ p0m,
                          double *__restrict__ state, int *__restrict__ gate) {
  const int mb = blockIdx.y;
  if (blockIdx.x == 0 && threadIdx.x == 0 && gate) gate[mb] = 1;
  if (p0m) p0 = p0m[mb];
  p += (size_t)mb * n;
  t += (size_t)mb * n;
  if (p_init) p_init += (size_t)mb * n;
  state += mb * PK_COUNT;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    T pi = p_init ? p_init[i] : (T)p0;
    p[i] = pi > (T)0 ? pi : (T)0;
    t[i] = (T)0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int k = 0; k < PK_COUNT; k++) state[k] = 0.0;
}

// h = z_wall + offset - c
template <typename T>
__global__ void k_pk_h(size_t n, const T *__restrict__ wall, double offset,
                       double c, const double *__restrict__ cm,
                       T *__restrict__ h) {
  const int mb = blockIdx.y;
  if (cm) c = cm[mb];
  wall += (size_t)mb * n;
  h += (size_t)mb * n;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    h[i] = (T)((double)wall[i] + offset - c);
}

// Surface traction of the layer-0 z component: p0 - p; zero elsewhere.
template <typename T>
__global__ void k_pk_traction(size_t n, int d, int zc,
                              const T *__restrict__ p, double p0,
                              T *__restrict__ tr) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    for (int c = 0; c < d; c++)
      tr[(size_t)c * n + i] = c == zc ? (T)(p0 - (double)p[i]) : (T)0;
}

// Sites carrying pressure sit on the wall; no site passes it.
template <typename T>
__global__ void k_pk_snap(size_t n, int zc, const T *__restrict__ p,
                          const T *__restrict__ wall, double offset,
                          T *__restrict__ u) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const T zw = (T)((double)wall[i] + offset);
    T &uz = u[(size_t)zc * n + i];
    if (p[i] > (T)0 || uz > zw) uz = zw;
  }
}

static unsigned grid_for(size_t n, int cap) {
  size_t b = (n + 255) / 256;
  return (unsigned)(b < (size_t)cap ? b : (size_t)cap);
}

template <typename T>
static void pk_launch_t(const PKBufs &b, int which, hipStream_t st) {
  const unsigned nb = (unsigned)b.nblk;
  const unsigned M = (unsigned)b.members;
  switch (which) {
    case PK_GAP:
      hipLaunchKernelGGL(k_pk_gap<T>, dim3(nb, M), dim3(256), 0, st, b.n,
                         (const T *)b.u, (const T *)b.h, (const T *)b.p,
                         (T *)b.g, b.part, b.gate);
      hipLaunchKernelGGL(k_pk_state, dim3(1, M), dim3(256), 0, st, b.part,
                         b.nblk, b.state, 0, b.gate);
      hipLaunchKernelGGL(k_pk_dir<T>, dim3(grid_for(b.n, 8192), M), dim3(256),
                         0, st, b.n, (const T *)b.g, (const T *)b.p,
                         (T *)b.t, b.state, b.gate);
      break;
    case PK_CURV_UPDATE:
      hipLaunchKernelGGL(k_pk_curv<T>, dim3(nb, M), dim3(256), 0, st, b.n,
                         (const T *)b.g, (const T *)b.t, (const T *)b.r,
                         (const T *)b.p, b.part, b.gate);
      hipLaunchKernelGGL(k_pk_state, dim3(1, M), dim3(256), 0, st, b.part,
                         b.nblk, b.state, 1, b.gate);
      hipLaunchKernelGGL(k_pk_update<T>, dim3(nb, M), dim3(256), 0, st, b.n,
                         (T *)b.p, (const T *)b.t, (const T *)b.g, b.state,
                         b.part, b.gate);
      hipLaunchKernelGGL(k_pk_state, dim3(1, M), dim3(256), 0, st, b.part,
                         b.nblk, b.state, 2, b.gate);
      break;
  }
}

void launch_pk(const PKBufs &b, bool fp64, int which, hipStream_t st) {
  if (fp64) pk_launch_t<double>(b, which, st);
  else pk_launch_t<float>(b, which, st);
}

void launch_pk_init(const PKBufs &b, bool fp64, const void *p_init, double p0,
                    hipStream_t st, const double *p0m) {
  const dim3 g(grid_for(b.n, 8192), b.members);
  if (fp64)
    hipLaunchKernelGGL(k_pk_init<double>, g, dim3(256), 0, st, b.n,
                       (double *)b.p, (double *)b.t, (const double *)p_init,
                       p0, p0m, b.state, (int *)b.gate);
  else
    hipLaunchKernelGGL(k_pk_init<float>, g, dim3(256), 0, st, b.n,
                       (float *)b.p, (float *)b.t, (const float *)p_init, p0,
                       p0m, b.state, (int *)b.gate);
}

void launch_pk_h(const PKBufs &b, bool fp64, const void *wall, double offset,
                 double c, hipStream_t st, const double *cm) {
  const dim3 g(grid_for(b.n, 8192), b.members);
  if (fp64)
    hipLaunchKernelGGL(k_pk_h<double>, g, dim3(256), 0, st, b.n,
                       (const double *)wall, offset, c, cm, (double *)b.h);
  else
    hipLaunchKernelGGL(k_pk_h<float>, g, dim3(256), 0, st, b.n,
                       (const float *)wall, offset, c, cm, (float *)b.h);
}

void launch_pk_traction(const PKBufs &b, bool fp64, int d, int zc, double p0,
                        void *tr, hipStream_t st) {
  const dim3 g(grid_for(b.n, 8192));
  if (fp64)
    hipLaunchKernelGGL(k_pk_traction<double>, g, dim3(256), 0, st, b.n, d, zc,
                       (const double *)b.p, p0, (double *)tr);
  else
    hipLaunchKernelGGL(k_pk_traction<float>, g, dim3(256), 0, st, b.n, d, zc,
                       (const float *)b.p, p0, (float *)tr);
}

void launch_pk_snap(const PKBufs &b, bool fp64, int zc, const void *wall,
                    double offset, void *u, hipStream_t st) {
  const dim3 g(grid_for(b.n, 8192));
  if (fp64)
    hipLaunchKernelGGL(k_pk_snap<double>, g, dim3(256), 0, st, b.n, zc,
                       (const double *)b.p, (const double *)wall, offset,
                       (double *)u);
  else
    hipLaunchKernelGGL(k_pk_snap<float>, g, dim3(256), 0, st, b.n, zc,
                       (const float *)b.p, (const float *)wall, offset,
                       (float *)u);
}

}  // namespace gfmd
