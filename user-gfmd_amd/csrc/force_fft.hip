// Elastic force evaluation  u(r) -> f(r), E  on the device.
//
// Three sweeps over HBM (reference semantics: GFMDSolverStatic::post_force,
// src/solvers/gfmd_solver_static.cpp:145-249):
//   1. k_rows_r2c : real FFT along y of every row (two real rows per complex
//                   transform), written transposed into the half spectrum
//                   spec[d][nky][nx] through an LDS staging tile.
//   2. k_cols     : per spectral column ky, for all d components at once:
//                   FFT along x -> f(q) = -Phi~(q) u(q) (+ linf at q = 0) with
//                   the FP64 energy partial sum -> inverse FFT along x.
//                   Everything between the load and the store stays in
//                   registers (LDS only for the Stockham exchanges).
//   3. k_rows_c2r : inverse real FFT along y, writes f[d][nx][ny].
// plus a one-block kernel that sums the energy partials in a fixed order.
//
// A generic direct-DFT path (O(N (nx+ny)) work) covers the grids the fast
// path has no instance for (non power-of-two sizes such as 64x37 or 63x63).
#include <cstdio>

#include "gfmd_internal.h"

namespace gfmd {

__device__ __forceinline__ double block_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; i++) s += red[i];
  }
  return s;  // valid in thread 0
}

// ---------------------------------------------------------------------------
// Sweep 1: real FFT along y.  Block = FPB transforms of NT = NY/R threads;
// each transform carries rows (a, b) as z = a + i b; ROUNDS rounds per block
// give TR = 2*FPB*ROUNDS consecutive rows, whose half spectra are staged in
// LDS as [ky][TR] and written to spec[c][ky][ix0 .. ix0+TR).
// ---------------------------------------------------------------------------
template <typename TI, typename T, int NY, int R, int FPB, int ROUNDS>
__global__ void __launch_bounds__(NY / R * FPB)
    k_rows_r2c(const TI *__restrict__ u, C2<T> *__restrict__ spec,
               const C2<T> *__restrict__ twy, int nx) {
  constexpr int NT = NY / R;
  constexpr int TR = 2 * FPB * ROUNDS;
  constexpr int NKY = NY / 2 + 1;
  __shared__ C2<T> xch[FPB * NY];
  __shared__ C2<T> stage[NKY * TR];
  const int c = blockIdx.y;
  const int ix0 = blockIdx.x * TR;
  const int s = threadIdx.x / NT, t = threadIdx.x % NT;
  const TI *uc = u + (size_t)c * nx * NY;
  C2<T> *x = xch + s * NY;

  for (int rnd = 0; rnd < ROUNDS; rnd++) {
    const int rr = 2 * (rnd * FPB + s);
    const TI *ra = uc + (size_t)(ix0 + rr) * NY;
    const TI *rb = ra + NY;
    C2<T> v[1][R];
#pragma unroll
    for (int m = 0; m < R; m++)
      v[0][m] = cmake((T)ra[t + m * NT], (T)rb[t + m * NT]);
    fft_regs<-1, NY, R, 1, 1>(v, x, NY, twy, t);
#pragma unroll
    for (int m = 0; m < R; m++) x[t + m * NT] = v[0][m];
    __syncthreads();
    // Split Z = FFT(a + i b) into the half spectra A, B of the two rows.
#pragma unroll
    for (int m = 0; m <= R / 2; m++) {
      const int k = t + m * NT;
      if (k <= NY / 2) {
        const C2<T> zk = x[k], zn = x[(NY - k) & (NY - 1)];
        const T h = (T)0.5;
        stage[k * TR + rr] = cmake(h * (zk.x + zn.x), h * (zk.y - zn.y));
        stage[k * TR + rr + 1] = cmake(h * (zk.y + zn.y), h * (zn.x - zk.x));
      }
    }
    __syncthreads();
  }
  C2<T> *sc = spec + (size_t)c * NKY * nx + ix0;
  for (int L = threadIdx.x; L < NKY * TR; L += blockDim.x) {
    const int ky = L / TR, r = L - ky * TR;
    sc[(size_t)ky * nx + r] = stage[L];
  }
}

// ---------------------------------------------------------------------------
// Sweep 3: inverse real FFT along y (mirror of sweep 1).
// ---------------------------------------------------------------------------
template <typename TO, typename T, int NY, int R, int FPB, int ROUNDS>
__global__ void __launch_bounds__(NY / R * FPB)
    k_rows_c2r(const C2<T> *__restrict__ spec, TO *__restrict__ f,
               const C2<T> *__restrict__ twy, int nx) {
  constexpr int NT = NY / R;
  constexpr int TR = 2 * FPB * ROUNDS;
  constexpr int NKY = NY / 2 + 1;
  __shared__ C2<T> xch[FPB * NY];
  __shared__ C2<T> stage[NKY * TR];
  const int c = blockIdx.y;
  const int ix0 = blockIdx.x * TR;
  const int s = threadIdx.x / NT, t = threadIdx.x % NT;
  const C2<T> *sc = spec + (size_t)c * NKY * nx + ix0;
  for (int L = threadIdx.x; L < NKY * TR; L += blockDim.x) {
    const int ky = L / TR, r = L - ky * TR;
    stage[L] = sc[(size_t)ky * nx + r];
  }
  __syncthreads();
  TO *fc = f + (size_t)c * nx * NY;
  for (int rnd = 0; rnd < ROUNDS; rnd++) {
    const int rr = 2 * (rnd * FPB + s);
    C2<T> v[1][R];
#pragma unroll
    for (int m = 0; m < R; m++) {
      const int k = t + m * NT;
      C2<T> A, B;
      if (k <= NY / 2) {
        A = stage[k * TR + rr];
        B = stage[k * TR + rr + 1];
        if (k == 0 || k == NY / 2) {
          A.y = 0;
          B.y = 0;
        }
      } else {
        const int kk = NY - k;
        A = cconj(stage[kk * TR + rr]);
        B = cconj(stage[kk * TR + rr + 1]);
      }
      v[0][m] = cmake(A.x - B.y, A.y + B.x);
    }
    fft_regs<+1, NY, R, 1, 1>(v, xch + s * NY, NY, twy, t);
    TO *fa = fc + (size_t)(ix0 + rr) * NY;
    TO *fb = fa + NY;
#pragma unroll
    for (int m = 0; m < R; m++) {
      fa[t + m * NT] = (TO)v[0][m].x;
      fb[t + m * NT] = (TO)v[0][m].y;
    }
  }
}

// ---------------------------------------------------------------------------
// f(q) = -Phi~(q) u(q) for one wavevector, Hermitian-packed Phi:
// P[i] = Phi_ii (i < D), then for each pair i < j in row-major order
// (Re Phi_ij, Im Phi_ij).  Returns u^+ Phi~ u (the energy density).
// ---------------------------------------------------------------------------
template <int D, typename T, typename LOADP>
__device__ __forceinline__ T apply_phi(C2<T> (&u)[D], LOADP loadp) {
  C2<T> F[D];
#pragma unroll
  for (int i = 0; i < D; i++) {
    const T pii = loadp(i);
    F[i] = cmake(pii * u[i].x, pii * u[i].y);
  }
  int p = D;
#pragma unroll
  for (int i = 0; i < D; i++) {
#pragma unroll
    for (int j = i + 1; j < D; j++) {
      const C2<T> pij = cmake(loadp(p), loadp(p + 1));
      p += 2;
      F[i] = F[i] + cmul(pij, u[j]);
      F[j] = F[j] + cmulc(u[i], pij);  // conj(Phi_ij) u_i
    }
  }
  T e = 0;
#pragma unroll
  for (int i = 0; i < D; i++) {
    e += F[i].x * u[i].x + F[i].y * u[i].y;
    u[i] = cmake(-F[i].x, -F[i].y);
  }
  return e;
}

// ---------------------------------------------------------------------------
// Sweep 2: columns.  CPB columns per block, NT = NX/R threads per column.
// ---------------------------------------------------------------------------
template <typename T, int NX, int R, int D, int DG, int CPB>
__global__ void __launch_bounds__(NX / R * CPB)
    k_cols(C2<T> *__restrict__ spec, const T *__restrict__ phi,
           const C2<T> *__restrict__ twx, int nky, int ny,
           const double *__restrict__ linf, double *__restrict__ partials,
           double *__restrict__ u0) {
  constexpr int NT = NX / R;
  constexpr int DD = D * D;
  __shared__ C2<T> xch[CPB * DG * NX];
  __shared__ double red[(NT * CPB + 63) / 64];
  const int s = threadIdx.x / NT, t = threadIdx.x % NT;
  const int col = blockIdx.x * CPB + s;
  const bool valid = col < nky;
  const int colc = valid ? col : nky - 1;

  C2<T> v[D][R];
#pragma unroll
  for (int c = 0; c < D; c++)
#pragma unroll
    for (int m = 0; m < R; m++)
      v[c][m] = spec[((size_t)c * nky + colc) * NX + t + m * NT];

  fft_regs<-1, NX, R, D, DG>(v, xch + s * DG * NX, NX, twx, t);

  const T *ph = phi + (size_t)colc * DD * NX;
  const double wcol = (colc == 0 || 2 * colc == ny) ? 1.0 : 2.0;
  double e = 0.0;
#pragma unroll
  for (int m = 0; m < R; m++) {
    const int kx = t + m * NT;
    C2<T> uq[D];
#pragma unroll
    for (int c = 0; c < D; c++) uq[c] = v[c][m];
    double eg = 0.0;
    const bool gamma = (colc == 0 && kx == 0);
    if (gamma) {
#pragma unroll
      for (int c = 0; c < D; c++) u0[c] = (double)uq[c].x;
      if (linf)
        for (int i = 0; i < D / 3; i++) eg -= 2.0 * linf[i] * (double)uq[3 * i + 2].x;
    }
    const T eq = apply_phi<D>(uq, [&](int k) { return ph[k * NX + kx]; });
    if (gamma && linf)
      for (int i = 0; i < D / 3; i++) uq[3 * i + 2].x += (T)linf[i];
#pragma unroll
    for (int c = 0; c < D; c++) v[c][m] = uq[c];
    e += wcol * (double)eq + eg;
  }

  fft_regs<+1, NX, R, D, DG>(v, xch + s * DG * NX, NX, twx, t);

  if (valid) {
#pragma unroll
    for (int c = 0; c < D; c++)
#pragma unroll
      for (int m = 0; m < R; m++)
        spec[((size_t)c * nky + col) * NX + t + m * NT] = v[c][m];
  } else {
    e = 0.0;
  }
  const double bs = block_sum(e, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// E = 1/2 sum(partials), summed in a fixed order (bitwise reproducible).
__global__ void __launch_bounds__(256)
    k_energy_final(const double *__restrict__ partials, int n,
                   double *__restrict__ epot) {
  __shared__ double red[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += partials[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) *epot = 0.5 * s;
}

// ---------------------------------------------------------------------------
// Fast-path configuration and dispatch.
// ---------------------------------------------------------------------------
template <typename T>
struct RowCfg {
  // rows staged per block: keep the [nky][TR] staging tile near 64 KB.
  static constexpr int tr(int n) {
    return (n <= 1024) ? 8 : (sizeof(T) == 4 ? (n <= 2048 ? 8 : 4)
                                             : (n <= 2048 ? 4 : 2));
  }
};

static constexpr int radix_for(int n) { return n >= 4096 ? 16 : 8; }

template <int N>
struct RowShape {
  static constexpr int R = radix_for(N);
  static constexpr int NT = N / R;
  static constexpr int FPB = NT >= 256 ? 1 : 256 / NT > 32 ? 32 : 256 / NT;
};

template <typename T, int D, int N>
struct ColShape {
  static constexpr int R = (sizeof(T) == 8 && D >= 6) ? 4 : radix_for(N);
  static constexpr int NT = N / R;
  static constexpr int CPB = NT >= 256 ? 1 : 256 / NT;
  static constexpr int per_comp = CPB * N * (int)sizeof(C2<T>);
  static constexpr int DG = (D % 3 == 0 && 3 * per_comp <= 65536) ? 3
                            : (D % 2 == 0 && 2 * per_comp <= 65536) ? 2
                                                                    : 1;
};

template <typename TI, typename T, int NY>
static void launch_r2c(const ForceArgs &a, hipStream_t st) {
  using S = RowShape<NY>;
  constexpr int TRW = RowCfg<T>::tr(NY);
  constexpr int ROUNDS = (TRW / (2 * S::FPB)) > 0 ? TRW / (2 * S::FPB) : 1;
  constexpr int TR = 2 * S::FPB * ROUNDS;
  dim3 grid(a.nx / TR, a.d);
  hipLaunchKernelGGL((k_rows_r2c<TI, T, NY, S::R, S::FPB, ROUNDS>), grid,
                     dim3(S::NT * S::FPB), 0, st, (const TI *)a.u,
                     (C2<T> *)a.spec, (const C2<T> *)a.twy, a.nx);
}

template <typename TO, typename T, int NY>
static void launch_c2r(const ForceArgs &a, hipStream_t st) {
  using S = RowShape<NY>;
  constexpr int TRW = RowCfg<T>::tr(NY);
  constexpr int ROUNDS = (TRW / (2 * S::FPB)) > 0 ? TRW / (2 * S::FPB) : 1;
  constexpr int TR = 2 * S::FPB * ROUNDS;
  dim3 grid(a.nx / TR, a.d);
  hipLaunchKernelGGL((k_rows_c2r<TO, T, NY, S::R, S::FPB, ROUNDS>), grid,
                     dim3(S::NT * S::FPB), 0, st, (const C2<T> *)a.spec,
                     (TO *)a.f, (const C2<T> *)a.twy, a.nx);
}

template <int NY>
static constexpr int row_tr(bool fp64) {
  using S = RowShape<NY>;
  return fp64 ? 2 * S::FPB *
                    ((RowCfg<double>::tr(NY) / (2 * S::FPB)) > 0
                         ? RowCfg<double>::tr(NY) / (2 * S::FPB)
                         : 1)
              : 2 * S::FPB *
                    ((RowCfg<float>::tr(NY) / (2 * S::FPB)) > 0
                         ? RowCfg<float>::tr(NY) / (2 * S::FPB)
                         : 1);
}

template <typename T, int D, int NX>
static int launch_cols(const ForceArgs &a, hipStream_t st) {
  using S = ColShape<T, D, NX>;
  const int nblk = (a.nky + S::CPB - 1) / S::CPB;
  if (nblk > a.nparts) return ST_STATE;
  hipLaunchKernelGGL((k_cols<T, NX, S::R, D, S::DG, S::CPB>), dim3(nblk),
                     dim3(S::NT * S::CPB), 0, st, (C2<T> *)a.spec,
                     (const T *)a.phi, (const C2<T> *)a.twx, a.nky, a.ny,
                     a.linf, a.partials, a.u0);
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, st, a.partials,
                     nblk, a.epot);
  return ST_OK;
}

template <typename T, int NX>
static int cols_nblk(int nky, int d) {
  if (d == 3) return (nky + ColShape<T, 3, NX>::CPB - 1) / ColShape<T, 3, NX>::CPB;
  return (nky + ColShape<T, 6, NX>::CPB - 1) / ColShape<T, 6, NX>::CPB;
}

static bool pow2_in(int n, int lo, int hi) {
  return n >= lo && n <= hi && (n & (n - 1)) == 0;
}

static int row_tr_rt(int ny, bool fp64) {
  switch (ny) {
    case 64: return row_tr<64>(fp64);
    case 128: return row_tr<128>(fp64);
    case 256: return row_tr<256>(fp64);
    case 512: return row_tr<512>(fp64);
    case 1024: return row_tr<1024>(fp64);
    case 2048: return row_tr<2048>(fp64);
    case 4096: return row_tr<4096>(fp64);
  }
  return 0;
}

bool force_fast_supported(int nx, int ny, int d, bool fp64) {
  if (!pow2_in(nx, 64, 4096) || !pow2_in(ny, 64, 4096)) return false;
  if (d != 3 && d != 6) return false;
  const int tr = row_tr_rt(ny, fp64);
  return tr > 0 && nx % tr == 0;
}

int force_fast_nparts(int nx, int ny, int d, bool fp64) {
  const int nky = ny / 2 + 1;
  switch (nx) {
#define NXCASE(N)                                                      \
  case N:                                                              \
    return fp64 ? cols_nblk<double, N>(nky, d) : cols_nblk<float, N>(nky, d);
    NXCASE(64) NXCASE(128) NXCASE(256) NXCASE(512) NXCASE(1024)
    NXCASE(2048) NXCASE(4096)
#undef NXCASE
  }
  return 0;
}

template <typename TI, typename T>
static int dispatch_rows_fwd(const ForceArgs &a, hipStream_t st) {
  switch (a.ny) {
    case 64: launch_r2c<TI, T, 64>(a, st); break;
    case 128: launch_r2c<TI, T, 128>(a, st); break;
    case 256: launch_r2c<TI, T, 256>(a, st); break;
    case 512: launch_r2c<TI, T, 512>(a, st); break;
    case 1024: launch_r2c<TI, T, 1024>(a, st); break;
    case 2048: launch_r2c<TI, T, 2048>(a, st); break;
    case 4096: launch_r2c<TI, T, 4096>(a, st); break;
    default: return ST_UNSUPPORTED;
  }
  return ST_OK;
}

template <typename TO, typename T>
static int dispatch_rows_inv(const ForceArgs &a, hipStream_t st) {
  switch (a.ny) {
    case 64: launch_c2r<TO, T, 64>(a, st); break;
    case 128: launch_c2r<TO, T, 128>(a, st); break;
    case 256: launch_c2r<TO, T, 256>(a, st); break;
    case 512: launch_c2r<TO, T, 512>(a, st); break;
    case 1024: launch_c2r<TO, T, 1024>(a, st); break;
    case 2048: launch_c2r<TO, T, 2048>(a, st); break;
    case 4096: launch_c2r<TO, T, 4096>(a, st); break;
    default: return ST_UNSUPPORTED;
  }
  return ST_OK;
}

template <typename T>
static int dispatch_cols(const ForceArgs &a, hipStream_t st) {
#define COLCASE(N)                                              \
  case N:                                                       \
    return a.d == 3 ? launch_cols<T, 3, N>(a, st)               \
                    : launch_cols<T, 6, N>(a, st);
  switch (a.nx) {
    COLCASE(64) COLCASE(128) COLCASE(256) COLCASE(512) COLCASE(1024)
    COLCASE(2048) COLCASE(4096)
  }
#undef COLCASE
  return ST_UNSUPPORTED;
}

template <typename TIO, typename T>
static int run_fast(const ForceArgs &a, hipStream_t st, hipEvent_t *ev) {
  if (ev) (void)hipEventRecord(ev[0], st);
  int rc = dispatch_rows_fwd<TIO, T>(a, st);
  if (rc) return rc;
  if (ev) (void)hipEventRecord(ev[1], st);
  rc = dispatch_cols<T>(a, st);
  if (rc) return rc;
  if (ev) (void)hipEventRecord(ev[2], st);
  rc = dispatch_rows_inv<TIO, T>(a, st);
  if (ev) (void)hipEventRecord(ev[3], st);
  return rc;
}

int launch_force_fast(const ForceArgs &a, bool fp64, bool io_double,
                      hipStream_t st, hipEvent_t *ev) {
  if (!force_fast_supported(a.nx, a.ny, a.d, fp64)) return ST_UNSUPPORTED;
  if (fp64) return run_fast<double, double>(a, st, ev);
  if (io_double) return run_fast<double, float>(a, st, ev);
  return run_fast<float, float>(a, st, ev);
}

// ---------------------------------------------------------------------------
// Generic path: direct DFTs (any nx, ny).
// ---------------------------------------------------------------------------
// spec[c][ky][ix] = sum_iy u[c][ix][iy] w_ny^(ky*iy)
template <typename TI, typename T>
__global__ void k_gen_rows_r2c(const TI *__restrict__ u, C2<T> *__restrict__ spec,
                               const C2<T> *__restrict__ twy, int nx, int ny,
                               int nky, int d) {
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
                           int d) {
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
                            double *__restrict__ partials, double *__restrict__ u0) {
  __shared__ double red[4];
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
                               int nky, int d) {
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
                     a.nx, a.ny, nky, a.d);
  hipLaunchKernelGGL((k_gen_cols<T, -1>), dim3(nb(ns)), dim3(256), 0, st,
                     (const C2<T> *)a.spec, (C2<T> *)a.spec2,
                     (const C2<T> *)a.twx, a.nx, nky, a.d);
  const int nparts = force_generic_nparts(a.nx, a.ny, a.d);
  if (nparts > a.nparts) return ST_STATE;
  hipLaunchKernelGGL((k_gen_apply<T>), dim3(nparts), dim3(256), 0, st,
                     (C2<T> *)a.spec2, (const T *)a.phi, a.nx, a.ny, nky, a.d,
                     a.linf, a.partials, a.u0);
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, st, a.partials,
                     nparts, a.epot);
  hipLaunchKernelGGL((k_gen_cols<T, +1>), dim3(nb(ns)), dim3(256), 0, st,
                     (const C2<T> *)a.spec2, (C2<T> *)a.spec,
                     (const C2<T> *)a.twx, a.nx, nky, a.d);
  const size_t nf = (size_t)a.d * a.nx * a.ny;
  hipLaunchKernelGGL((k_gen_rows_c2r<TIO, T>), dim3(nb(nf)), dim3(256), 0, st,
                     (const C2<T> *)a.spec, (TIO *)a.f, (const C2<T> *)a.twy,
                     a.nx, a.ny, nky, a.d);
  return ST_OK;
}

int launch_force_generic(const ForceArgs &a, bool fp64, bool io_double,
                         hipStream_t st) {
  if (a.d > 24) return ST_UNSUPPORTED;
  if (fp64) return run_generic<double, double>(a, st);
  if (io_double) return run_generic<double, float>(a, st);
  return run_generic<float, float>(a, st);
}

}  // namespace gfmd
