// Elastic force evaluation  u(r) -> f(r), E  on the device, power-of-two
// grids.  Reference semantics: GFMDSolverStatic::post_force
// (src/solvers/gfmd_solver_static.cpp:145-249): forward transform
// exp(-i q.r), f(q) = -Phi~(q) u(q) with Phi~ = Phi/(nx*ny), energy summed
// over the full grid, Gamma-point linf terms, unnormalised inverse.
//
// Three sweeps over HBM:
//   1. k_rows_r2c : real FFT along y of every row (two real rows per complex
//                   transform), written transposed into the half spectrum
//                   spec[d][ny/2][nx].  Column 0 packs the two real bins of
//                   every row, DC in the real and Nyquist in the imaginary
//                   part, so the spectrum has exactly ny/2 columns.
//   2. k_cols     : one spectral column per block slot, all d components:
//                   FFT along x -> f(q) = -Phi~(q) u(q) (+ linf at q = 0)
//                   with the FP64 energy partial -> inverse FFT along x.
//                   Column 0 is split into its DC and Nyquist columns
//                   (X0 = (W + conj W(-k))/2, Xn = (W - conj W(-k))/2i),
//                   each gets its own Phi, and they are packed again.
//   3. k_rows_c2r : inverse real FFT along y, writes f[d][nx][ny].
// plus a one-block kernel that sums the energy partials in a fixed order.
//
// Row tiles are mapped XCD-aware: consecutive tiles (which share 128-byte
// lines of the transposed spectrum) are given to blocks of the same XCD.
#pragma once

#include <type_traits>

#include "gfmd_internal.h"

namespace gfmd {

// Blocks are dispatched round-robin over the 8 XCDs.  A persistent grid of
// G blocks walks the work in chunks of G units; inside a chunk the blocks of
// one XCD get a contiguous range, so neighbouring row tiles (which share
// 128-byte lines of the transposed spectrum) meet in the same L2.
// A partial last chunk keeps the identity order.
__device__ __forceinline__ int xcd_unit(int it, int G, int ntot) {
  const int base = it - (it % G);
  const int b = it - base;
  if ((G & 7) || base + G > ntot) return it;
  return base + (b & 7) * (G >> 3) + (b >> 3);
}

// ---------------------------------------------------------------------------
// Sweep 1: real FFT along y.  FPB transforms of NT = NY/R threads per block,
// transform s carries rows (2s, 2s+1) of a TR = 2 FPB row tile as z = a + ib.
// The LDS area serves the FFT exchange, then Z, then the staged [ny/2][TR]
// tile.  Persistent: block b handles tiles b, b+G, ...; with PF it
// prefetches the next tile's rows into registers while transforming the
// current one.
// ---------------------------------------------------------------------------
template <typename TI, typename T, int NY, int R, int FPB, bool PF, bool TWC>
__global__ void __launch_bounds__(NY / R * FPB)
    k_rows_r2c(const TI *__restrict__ u, C2<T> *__restrict__ spec,
               const C2<T> *__restrict__ twy, int nx, int ntiles,
               const int *__restrict__ gate, size_t ms_u, size_t ms_s) {
  // blockIdx.y: member of an ensemble (fields ms_* elements apart); a
  // member whose gate is 0 has finished
  if (gate && gate[blockIdx.y] == 0) return;
  u += blockIdx.y * ms_u;
  spec += blockIdx.y * ms_s;
  constexpr int NT = NY / R;
  constexpr int TR = 2 * FPB;
  constexpr int NC = NY / 2;  // spectral columns
  constexpr int NTH = NT * FPB;
  constexpr int AREA = FFTPlan<NY, R>::lds_elems((int)sizeof(C2<T>));
  constexpr int TRP = TR + 1;  // odd row stride of the staged tile
  constexpr int LDSN = FPB * AREA > NC * TRP ? FPB * AREA : NC * TRP;
  __shared__ C2<T> lds[LDSN];
  const int tiles = nx / TR;
  const int G = gridDim.x;
  const int s = threadIdx.x / NT, t = threadIdx.x % NT;
  C2<T> *x = lds + s * AREA;
  // twiddle bases in registers for every tile of the persistent block
  TwCache<NY, R, T> twc;
  if constexpr (TWC) twc.load(twy, t);
  const TwCache<NY, R, T> *wc = TWC ? &twc : nullptr;

  int it = blockIdx.x;
  C2<T> nv[R];
  auto load = [&](int seq) {
    const int lt = xcd_unit(seq, G, ntiles);
    const int c = lt / tiles;
    const int ix0 = (lt - c * tiles) * TR;
    const TI *ra = u + ((size_t)c * nx + ix0 + 2 * s) * NY;
    const TI *rb = ra + NY;
#pragma unroll
    for (int m = 0; m < R; m++)
      nv[m] = cmake((T)ra[t + m * NT], (T)rb[t + m * NT]);
  };
  if (PF && it < ntiles) load(it);
  while (it < ntiles) {
    const int lt = xcd_unit(it, G, ntiles);
    const int c = lt / tiles;
    const int ix0 = (lt - c * tiles) * TR;
    if (!PF) load(it);
    C2<T> v[1][R];
#pragma unroll
    for (int m = 0; m < R; m++) v[0][m] = nv[m];
    it += G;
    if (PF && it < ntiles) load(it);

    fft_regs<-1, NY, R, 1, 1>(v, x, AREA, twy, t, wc);
#pragma unroll
    for (int m = 0; m < R; m++) x[t + m * NT] = v[0][m];
    __syncthreads();
    // Split Z = FFT(a + i b) into the half spectra A, B of the two rows.
    C2<T> A[R / 2 + 1], B[R / 2 + 1];
#pragma unroll
    for (int m = 0; m <= R / 2; m++) {
      const int k = t + m * NT;
      const C2<T> zk = x[k & (NY - 1)], zn = x[(NY - k) & (NY - 1)];
      const T h = (T)0.5;
      A[m] = cmake(h * (zk.x + zn.x), h * (zk.y - zn.y));
      B[m] = cmake(h * (zk.y + zn.y), h * (zn.x - zk.x));
    }
    __syncthreads();
    C2<T> *stage = lds;  // [NC][TRP]
#pragma unroll
    for (int m = 0; m < R / 2; m++) {
      const int k = t + m * NT;
      C2<T> a = A[m], b = B[m];
      if (k == 0) {  // pack the real DC and Nyquist bins (k = NY/2: m = R/2)
        a = cmake(A[0].x, A[R / 2].x);
        b = cmake(B[0].x, B[R / 2].x);
      }
      stage[k * TRP + 2 * s] = a;
      stage[k * TRP + 2 * s + 1] = b;
    }
    __syncthreads();
    C2<T> *sc = spec + (size_t)c * NC * nx + ix0;
#pragma unroll
    for (int j = 0; j < R; j++) {
      const int L = threadIdx.x + j * NTH;
      const int ky = L / TR, r = L - ky * TR;
      sc[(size_t)ky * nx + r] = stage[ky * TRP + r];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Sweep 3: inverse real FFT along y (mirror of sweep 1), persistent; with PF
// the next tile's spectrum is prefetched into registers.
// ---------------------------------------------------------------------------
// Block 0 also sums the column sweep's np energy partials into epot (they
// were written by the previous launch): E = 1/2 sum, in a fixed order.
template <typename TO, typename T, int NY, int R, int FPB, bool PF, bool TWC>
__global__ void __launch_bounds__(NY / R * FPB)
    k_rows_c2r(const C2<T> *__restrict__ spec, TO *__restrict__ f,
               const C2<T> *__restrict__ twy, int nx, int ntiles,
               const int *__restrict__ gate, size_t ms_s, size_t ms_f,
               const double *__restrict__ partials, int np, int ms_p,
               double *__restrict__ epot) {
  if (gate && gate[blockIdx.y] == 0) return;
  spec += blockIdx.y * ms_s;
  f += blockIdx.y * ms_f;
  constexpr int NT = NY / R;
  constexpr int TR = 2 * FPB;
  constexpr int NC = NY / 2;
  constexpr int NTH = NT * FPB;
  constexpr int AREA = FFTPlan<NY, R>::lds_elems((int)sizeof(C2<T>));
  constexpr int TRP = TR + 1;  // odd row stride of the staged tile
  constexpr int LDSN = FPB * AREA > NC * TRP ? FPB * AREA : NC * TRP;
  __shared__ __align__(16) C2<T> lds[LDSN];
  const int tiles = nx / TR;
  const int G = gridDim.x;
  const int s = threadIdx.x / NT, t = threadIdx.x % NT;
  TwCache<NY, R, T> twc;
  if constexpr (TWC) twc.load(twy, t);
  const TwCache<NY, R, T> *wc = TWC ? &twc : nullptr;

  int it = blockIdx.x;
  C2<T> ld[R];
  auto load = [&](int seq) {
    const int lt = xcd_unit(seq, G, ntiles);
    const int c = lt / tiles;
    const int ix0 = (lt - c * tiles) * TR;
    const C2<T> *sc = spec + (size_t)c * NC * nx + ix0;
#pragma unroll
    for (int j = 0; j < R; j++) {
      const int L = threadIdx.x + j * NTH;
      const int ky = L / TR, r = L - ky * TR;
      ld[j] = sc[(size_t)ky * nx + r];
    }
  };
  if (PF && it < ntiles) load(it);
  if (blockIdx.x == 0 && partials) {  // block-uniform
    // the wave sums go to the start of the staging area (some instances
    // use all 160 KB of LDS), which is free until the first tile
    double *red = reinterpret_cast<double *>(lds);
    const double *pp = partials + (size_t)blockIdx.y * ms_p;
    // all four loads of a round are issued before the adds
    double sum = 0.0;
    for (int i0 = threadIdx.x; i0 < np; i0 += 4 * NTH) {
      double x[4];
#pragma unroll
      for (int k = 0; k < 4; k++)
        x[k] = i0 + NTH * k < np ? pp[i0 + NTH * k] : 0.0;
      sum += (x[0] + x[1]) + (x[2] + x[3]);
    }
    sum = block_sum(sum, red);
    if (threadIdx.x == 0) epot[blockIdx.y] = 0.5 * sum;
    __syncthreads();  // red read before the tile is staged over it
  }
  while (it < ntiles) {
    const int lt = xcd_unit(it, G, ntiles);
    const int c = lt / tiles;
    const int ix0 = (lt - c * tiles) * TR;
    if (!PF) load(it);
    // field by field: a whole-struct copy of a 16-byte element into LDS
    // kept ld[] in private scratch memory (one global load in flight at a
    // time, and the scratch lines written back to HBM)
#pragma unroll
    for (int j = 0; j < R; j++) {
      const int L = threadIdx.x + j * NTH;
      const int ky = L / TR, r = L - ky * TR;
      lds[ky * TRP + r] = cmake(ld[j].x, ld[j].y);
    }
    it += G;
    if (PF && it < ntiles) load(it);
    __syncthreads();
    const C2<T> *stage = lds;
    C2<T> v[1][R];
    const C2<T> p0 = stage[2 * s], q0 = stage[2 * s + 1];  // ky = 0 row
#pragma unroll
    for (int m = 0; m < R; m++) {
      const int k = t + m * NT;
      C2<T> A, B;
      if (k == 0) {
        A = cmake(p0.x, (T)0);
        B = cmake(q0.x, (T)0);
      } else if (k == NY / 2) {
        A = cmake(p0.y, (T)0);
        B = cmake(q0.y, (T)0);
      } else if (k < NY / 2) {
        A = stage[k * TRP + 2 * s];
        B = stage[k * TRP + 2 * s + 1];
      } else {
        const int kk = NY - k;
        A = cconj(stage[kk * TRP + 2 * s]);
        B = cconj(stage[kk * TRP + 2 * s + 1]);
      }
      v[0][m] = cmake(A.x - B.y, A.y + B.x);
    }
    __syncthreads();
    fft_regs<+1, NY, R, 1, 1>(v, lds + s * AREA, AREA, twy, t, wc);
    TO *fa = f + ((size_t)c * nx + ix0 + 2 * s) * NY;
    TO *fb = fa + NY;
#pragma unroll
    for (int m = 0; m < R; m++) {
      fa[t + m * NT] = (TO)v[0][m].x;
      fb[t + m * NT] = (TO)v[0][m].y;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// f(q) = -Phi~(q) u(q) for one wavevector, Hermitian-packed Phi:
// P[i] = Phi_ii (i < D), then for each pair i < j in row-major order
// (Re Phi_ij, Im Phi_ij).  Returns u^+ Phi~ u (the energy density).
// Entry e of the wavevector is ph[e * stride + off] with a 32-bit index: ph
// is the same for every thread of the block (a scalar base), off the
// thread's own wavevector offset, so no 64-bit address per entry.
// ---------------------------------------------------------------------------
template <int D, typename T>
__device__ __forceinline__ T apply_phi(C2<T> (&u)[D], const T *__restrict__ ph,
                                       unsigned off, int stride,
                                       bool mir = false) {
  C2<T> F[D];
#pragma unroll
  for (int i = 0; i < D; i++) {
    const T pii = ph[(unsigned)(i * stride) + off];
    F[i] = cmake(pii * u[i].x, pii * u[i].y);
  }
  int p = D;
#pragma unroll
  for (int i = 0; i < D; i++) {
#pragma unroll
    for (int j = i + 1; j < D; j++) {
      C2<T> pij = cmake(ph[(unsigned)(p * stride) + off],
                            ph[(unsigned)((p + 1) * stride) + off]);
      p += 2;
      // x mirror: Phi(-qx) = M Phi(qx) M, M = -1 on every x component
      if ((i % 3 == 0) != (j % 3 == 0) && mir) pij = cmake(-pij.x, -pij.y);
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

// The same for ndof 3 with the compact table (launch_phi_compact): P[0..2]
// the diagonal, P[3] = Phi_xy (real), P[4] = Im Phi_xz, P[5] = Im Phi_yz
// (Phi_xz = i P[4], Phi_yz = i P[5]).
template <typename T>
__device__ __forceinline__ T apply_phi_par(C2<T> (&u)[3],
                                           const T *__restrict__ ph,
                                           unsigned off, int stride,
                                           bool mir = false) {
  const T pxx = ph[off], pyy = ph[(unsigned)stride + off];
  const T pzz = ph[(unsigned)(2 * stride) + off];
  T a = ph[(unsigned)(3 * stride) + off], b = ph[(unsigned)(4 * stride) + off];
  const T c = ph[(unsigned)(5 * stride) + off];
  if (mir) {  // x mirror: Phi_xy and Phi_xz change sign
    a = -a;
    b = -b;
  }
  C2<T> F[3];
  // F_x = Pxx u_x + a u_y + i b u_z
  F[0] = cmake(pxx * u[0].x + a * u[1].x - b * u[2].y,
               pxx * u[0].y + a * u[1].y + b * u[2].x);
  // F_y = a u_x + Pyy u_y + i c u_z
  F[1] = cmake(a * u[0].x + pyy * u[1].x - c * u[2].y,
               a * u[0].y + pyy * u[1].y + c * u[2].x);
  // F_z = -i b u_x - i c u_y + Pzz u_z
  F[2] = cmake(b * u[0].y + c * u[1].y + pzz * u[2].x,
               -b * u[0].x - c * u[1].x + pzz * u[2].y);
  T e = 0;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    e += F[i].x * u[i].x + F[i].y * u[i].y;
    u[i] = cmake(-F[i].x, -F[i].y);
  }
  return e;
}

// The same product for large blocks (ndof 9, 12): one row of the packed
// table at a time, a scheduling barrier between rows so that the loads of
// later rows are not hoisted (the whole 12 x 12 table of a wavevector would
// take 144 registers besides the column's data).
template <int D, typename T>
__device__ __forceinline__ T apply_phi_rows(C2<T> (&u)[D],
                                            const T *__restrict__ ph,
                                            unsigned off, int stride,
                                            bool mir) {
  C2<T> F[D];
#pragma unroll
  for (int i = 0; i < D; i++) {
    const T pii = ph[(unsigned)(i * stride) + off];
    F[i] = cmake(pii * u[i].x, pii * u[i].y);
  }
  int p = D;
#pragma unroll
  for (int i = 0; i < D; i++) {
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = i + 1; j < D; j++) {
      C2<T> pij = cmake(ph[(unsigned)(p * stride) + off],
                            ph[(unsigned)((p + 1) * stride) + off]);
      p += 2;
      if ((i % 3 == 0) != (j % 3 == 0) && mir) pij = cmake(-pij.x, -pij.y);
      F[i] = F[i] + cmul(pij, u[j]);
      F[j] = F[j] + cmulc(u[i], pij);
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

// Phi.u with the full packed table or, for ndof 3 (PAR), the compact one.
template <int D, bool PAR, typename T>
__device__ __forceinline__ T apply_table(C2<T> (&u)[D],
                                         const T *__restrict__ ph,
                                         unsigned off, int stride, bool mir) {
  if constexpr (PAR && D == 3)
    return apply_phi_par<T>(u, ph, off, stride, mir);
  else if constexpr (D >= 9)
    return apply_phi_rows<D>(u, ph, off, stride, mir);
  else
    return apply_phi<D>(u, ph, off, stride, mir);
}

// ---------------------------------------------------------------------------
// Sweep 2: columns.  CPB columns per block, NT = NX/R threads per column,
// one group of CPB columns per block.  phi has ny/2+1 columns (column
// ny/2 = Nyquist).  xbuf: C2<T>[D][NX].  MINW asks the compiler for at
// least MINW waves per SIMD (a register cap).  mirror != 0: the table is
// symmetric under x -> -x (checked when it was written), so wavevectors
// kx > NX/2 read the entry of NX - kx with the x components' signs flipped
// and only half of the table is fetched.
// ---------------------------------------------------------------------------
// Calls f(std::integral_constant<int, m>{}) for m = 0 .. N-1.
template <int N, int M = 0, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (M < N) {
    f(std::integral_constant<int, M>{});
    static_for<N, M + 1>(f);
  }
}

// One column group g: v holds the group's spectrum (thread t of column s:
// elements t + m NT of each component); forward FFT, Phi.u (energy into e),
// inverse FFT, store.  xch: the block's exchange areas.
template <typename T, int NX, int R, int D, int DG, int CPB, bool STRAIGHT,
          bool TWC, bool PAR>
__device__ __forceinline__ void cols_group(
    C2<T> (&v)[D][R], int g, C2<T> *__restrict__ spec,
    const T *__restrict__ phi, const C2<T> *__restrict__ twx, int ncol,
    const double *__restrict__ linf, double *__restrict__ u0,
    C2<T> *__restrict__ xbuf, int mirror, C2<T> *xch,
    const TwCache<NX, R, T> *wc, double &e) {
  constexpr int NT = NX / R;
  // table entries per wavevector: packed Hermitian, or the compact ndof-3
  constexpr int DD = (PAR && D == 3) ? 6 : D * D;
  constexpr int AREA = FFTPlan<NX, R>::lds_elems((int)sizeof(C2<T>));
  const int s = threadIdx.x / NT, t = threadIdx.x % NT;
  const int col = g * CPB + s;
  const bool packed = (col == 0);  // DC + Nyquist column
  C2<T> *specb = spec + (size_t)g * CPB * NX;
  const unsigned lane = (unsigned)(s * NX + t);

  fft_regs<-1, NX, R, D, DG>(v, xch + s * DG * AREA, AREA, twx, t, wc);

  if (g == 0) {  // block-uniform: the group holding the packed column
    if (packed) {
#pragma unroll
      for (int c = 0; c < D; c++)
#pragma unroll
        for (int m = 0; m < R; m++) xbuf[c * NX + t + m * NT] = v[c][m];
    }
    __syncthreads();
  }

  // the block's table base (uniform) and the thread's column offset
  const T *ph = phi + (size_t)g * CPB * DD * NX;
  const unsigned lane_col = (unsigned)(s * DD * NX);
  // Phi.u at the thread's wavevector m, for an ordinary column and for
  // the packed DC/Nyquist column.
  auto common = [&](auto mc) {
    constexpr int m = decltype(mc)::value;
    const int kx = t + m * NT;
    const bool mir = mirror && 2 * kx > NX;
    const int kr = mir ? NX - kx : kx;
    C2<T> uq[D];
#pragma unroll
    for (int c = 0; c < D; c++) uq[c] = v[c][m];
    e += 2.0 * (double)apply_table<D, PAR>(uq, ph, lane_col + kr, NX, mir);
#pragma unroll
    for (int c = 0; c < D; c++) v[c][m] = uq[c];
  };
  auto dcnyq = [&](auto mc) {
    constexpr int m = decltype(mc)::value;
    const int kx = t + m * NT;
    const bool mir = mirror && 2 * kx > NX;
    const int kr = mir ? NX - kx : kx;
    C2<T> uq[D];
#pragma unroll
    for (int c = 0; c < D; c++) uq[c] = v[c][m];
    // W = X0 + i Xn with X0, Xn the x-spectra of the real DC and
    // Nyquist columns; W(-kx) comes from the exchange buffer.
    const int kn = (NX - kx) & (NX - 1);
    C2<T> xn[D];
#pragma unroll
    for (int c = 0; c < D; c++) {
      const C2<T> w = uq[c], wm = xbuf[c * NX + kn];
      const T h = (T)0.5;
      uq[c] = cmake(h * (w.x + wm.x), h * (w.y - wm.y));
      xn[c] = cmake(h * (w.y + wm.y), h * (wm.x - w.x));
    }
    if (kx == 0) {
#pragma unroll
      for (int c = 0; c < D; c++) u0[c] = (double)uq[c].x;
      if (linf)
        for (int i = 0; i < D / 3; i++)
          e -= 2.0 * linf[i] * (double)uq[3 * i + 2].x;
    }
    e += (double)apply_table<D, PAR>(uq, ph, lane_col + kr, NX, mir);
    e += (double)apply_table<D, PAR>(xn, phi + (size_t)ncol * DD * NX, kr,
                                     NX, mir);
    if (kx == 0 && linf)
      for (int i = 0; i < D / 3; i++) uq[3 * i + 2].x += (T)linf[i];
#pragma unroll
    for (int c = 0; c < D; c++)
      v[c][m] = cmake(uq[c].x - xn[c].y, uq[c].y + xn[c].x);
  };
  if constexpr (STRAIGHT) {
    // The packed test sits outside the loop over the wavevectors, so the
    // common path is one straight run of table loads the compiler keeps
    // in flight together (a test per wavevector splits it into R
    // branches, each draining its loads before the next).
    if (!packed)
      static_for<R>(common);
    else
      static_for<R>(dcnyq);
  } else {
    // A block without the packed column (the test is block-uniform, so
    // no per-wavevector branch) takes the table entries of CH wavevectors
    // at a time: one memory round trip per chunk instead of one per
    // wavevector, with no more entries live at once than the register
    // budget allows.  Where even CH = 2 spills (the twiddle state grows
    // with NX, larger blocks hold more per wavevector), CH = 1 and every
    // block takes the per-wavevector path below.
    constexpr int CH = DD <= 6 ? (NX >= 4096 ? 1 : 3)
                               : (DD <= 9 && NX <= 1024 ? 2 : 1);
    if (CH > 1 && g != 0) {
      static_for<(R + CH - 1) / CH>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        __builtin_amdgcn_sched_barrier(0);
        static_for<CH>([&](auto jc) {
          constexpr int m = k * CH + decltype(jc)::value;
          if constexpr (m < R) common(std::integral_constant<int, m>{});
        });
      });
      __builtin_amdgcn_sched_barrier(0);
    } else {
      // one wavevector's table (D(D+1)/2 complex entries) at a time, so
      // its loads do not all stay live at once (larger blocks, or fewer
      // registers per thread)
      static_for<R>([&](auto mc) {
        if (!packed)
          common(mc);
        else
          dcnyq(mc);
      });
    }
  }

  fft_regs<+1, NX, R, D, DG>(v, xch + s * DG * AREA, AREA, twx, t, wc);

#pragma unroll
  for (int c = 0; c < D; c++)
#pragma unroll
    for (int m = 0; m < R; m++)
      (specb + (size_t)c * ncol * NX)[lane + m * NT] = v[c][m];
}

// Loads column group g's spectrum into v (the layout of cols_group).
template <typename T, int NX, int R, int D, int CPB>
__device__ __forceinline__ void cols_load(C2<T> (&v)[D][R], int g,
                                          const C2<T> *__restrict__ spec,
                                          int ncol) {
  constexpr int NT = NX / R;
  const int s = threadIdx.x / NT, t = threadIdx.x % NT;
  // the block's first column (a base that is the same for every thread,
  // per component) and the thread's offset from it: one address register
  // for all of the thread's loads and stores
  const C2<T> *specb = spec + (size_t)g * CPB * NX;
  const unsigned lane = (unsigned)(s * NX + t);
#pragma unroll
  for (int c = 0; c < D; c++)
#pragma unroll
    for (int m = 0; m < R; m++)
      v[c][m] = (specb + (size_t)c * ncol * NX)[lane + m * NT];
}

template <typename T, int NX, int R, int D, int DG, int CPB, bool STRAIGHT,
          bool TWC, bool PAR>
__device__ __forceinline__ void cols_body(
    C2<T> *__restrict__ spec, const T *__restrict__ phi,
    const C2<T> *__restrict__ twx, int ncol, const double *__restrict__ linf,
    double *__restrict__ partials, double *__restrict__ u0,
    C2<T> *__restrict__ xbuf, int mirror) {
  constexpr int NT = NX / R;
  constexpr int AREA = FFTPlan<NX, R>::lds_elems((int)sizeof(C2<T>));
  __shared__ C2<T> xch[CPB * DG * AREA];
  __shared__ double red[(NT * CPB + 63) / 64];
  const int t = threadIdx.x % NT;
  double e = 0.0;
  {
    const int g = blockIdx.x;
    // twiddle bases in registers (TWC): issued with the spectrum loads, one
    // memory round trip for all passes of both transforms
    TwCache<NX, R, T> twc;
    if constexpr (TWC) twc.load(twx, t);
    const TwCache<NX, R, T> *wc = TWC ? &twc : nullptr;
    C2<T> v[D][R];
    cols_load<T, NX, R, D, CPB>(v, g, spec, ncol);
    cols_group<T, NX, R, D, DG, CPB, STRAIGHT, TWC, PAR>(
        v, g, spec, phi, twx, ncol, linf, u0, xbuf, mirror, xch, wc, e);
  }
  const double bs = block_sum(e, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

template <typename T, int NX, int R, int D, int DG, int CPB, int MINW,
          bool STRAIGHT, bool TWC, bool PAR>
__global__ void __launch_bounds__(NX / R * CPB, MINW)
    k_cols(C2<T> *__restrict__ spec, const T *__restrict__ phi,
           const C2<T> *__restrict__ twx, int ncol,
           const double *__restrict__ linf, double *__restrict__ partials,
           double *__restrict__ u0, C2<T> *__restrict__ xbuf,
           const int *__restrict__ gate, size_t ms_s, int ms_p, int mirror) {
  if (gate && gate[blockIdx.y] == 0) return;
  const int mb = blockIdx.y;
  cols_body<T, NX, R, D, DG, CPB, STRAIGHT, TWC, PAR>(
      spec + mb * ms_s, phi, twx, ncol, linf, partials + (size_t)mb * ms_p,
      u0 + mb * D, xbuf + (size_t)mb * D * NX, mirror);
}

// ---------------------------------------------------------------------------
// Sweep 2 split in three for cells of three and four atoms (ndof 9, 12),
// whose fused column kernel cannot hold a column's components, its product
// vectors and the transform's state in registers: the transforms
// along x of every component line on their own (k_cols_fft, forward and
// inverse), and between them the product with the table
// (k_cols_phi: ordinary columns; k_cols_phi_dc: the packed DC/Nyquist
// column, kx paired with nx - kx).  The energy partials of the two product
// kernels follow each other in `partials`.
// ---------------------------------------------------------------------------
// col0 (optional, [d][NX] per member): read column 0 of every component
// from there instead of from spec (the packed column's product output).
template <typename T, int NX, int R, int CPB, int SIGN, bool TWC>
__global__ void __launch_bounds__(NX / R * CPB)
    k_cols_fft(C2<T> *__restrict__ spec, const C2<T> *__restrict__ twx,
               const int *__restrict__ gate, size_t ms_s,
               const C2<T> *__restrict__ col0, int ncol, int d) {
  if (gate && gate[blockIdx.y] == 0) return;
  constexpr int NT = NX / R;
  constexpr int AREA = FFTPlan<NX, R>::lds_elems((int)sizeof(C2<T>));
  __shared__ C2<T> xch[CPB * AREA];
  const int s = threadIdx.x / NT, t = threadIdx.x % NT;
  // line blockIdx.x * CPB + s of spec[d][ncol][NX] (lines are contiguous)
  const int line = blockIdx.x * CPB + s;
  C2<T> *p = spec + blockIdx.y * ms_s + (size_t)line * NX;
  const C2<T> *src = p;
  if (col0 && line % ncol == 0)
    src = col0 + ((size_t)blockIdx.y * d + line / ncol) * NX;
  TwCache<NX, R, T> twc;
  if constexpr (TWC) twc.load(twx, t);
  C2<T> v[1][R];
#pragma unroll
  for (int m = 0; m < R; m++) v[0][m] = src[t + m * NT];
  fft_regs<SIGN, NX, R, 1, 1>(v, xch + s * AREA, AREA, twx, t,
                              TWC ? &twc : nullptr);
#pragma unroll
  for (int m = 0; m < R; m++) p[t + m * NT] = v[0][m];
}

// Phi.u on columns 1 .. ncol-1 (x-spectra in place), one wavevector per
// thread and step; energy partial per block.
template <typename T, int D>
__global__ void __launch_bounds__(256)
    k_cols_phi(C2<T> *__restrict__ spec, const T *__restrict__ phi, int nx,
               int ncol, double *__restrict__ partials,
               const int *__restrict__ gate, size_t ms_s, int ms_p,
               int mirror) {
  __shared__ double red[4];
  if (gate && gate[blockIdx.y] == 0) return;
  spec += blockIdx.y * ms_s;
  constexpr int DD = D * D;
  const size_t npts = (size_t)(ncol - 1) * nx;
  const size_t cs = (size_t)ncol * nx;  // component stride
  double e = 0.0;
  for (size_t L = (size_t)blockIdx.x * 256 + threadIdx.x; L < npts;
       L += (size_t)gridDim.x * 256) {
    const int col = 1 + (int)(L / nx);
    const int kx = (int)(L % nx);
    const bool mir = mirror && 2 * kx > nx;
    const int kr = mir ? nx - kx : kx;
    C2<T> *q = spec + (size_t)col * nx + kx;
    C2<T> u[D];
#pragma unroll
    for (int c = 0; c < D; c++) u[c] = q[c * cs];
    e += 2.0 * (double)apply_table<D, false>(
                   u, phi + (size_t)col * DD * nx, (unsigned)kr, nx, mir);
#pragma unroll
    for (int c = 0; c < D; c++) q[c * cs] = u[c];
  }
  const double bs = block_sum(e, red);
  if (threadIdx.x == 0) partials[(size_t)blockIdx.y * ms_p + blockIdx.x] = bs;
}

// The packed column 0 (W = X0 + i Xn, X0 and Xn the x-spectra of the real
// DC and Nyquist rows), one output component i at one kx per thread (and
// step): X0 = (W(kx) + conj W(-kx))/2 and Xn = (W(kx) - conj W(-kx))/2i
// from all d components, row i of the ky = 0 table applied to X0 and of the
// ky = ny/2 table to Xn (both rows' loads issued together), the Gamma-point
// terms on X0 at kx = 0, and W'_i = X0'_i + i Xn'_i written to out
// ([d][nx] per member; spec's column 0 is still being read by others).
// Packed Hermitian table: diagonal at i, Phi_ij (i < j) at
// d + 2 (i d - i (i + 1)/2 + j - i - 1).
template <typename T, int D>
__global__ void __launch_bounds__(256)
    k_cols_phi_dc(const C2<T> *__restrict__ spec, const T *__restrict__ phi,
                  int nx, int ncol, const double *__restrict__ linf,
                  double *__restrict__ partials, double *__restrict__ u0,
                  C2<T> *__restrict__ out, const int *__restrict__ gate,
                  size_t ms_s, int ms_p, int part0, int mirror) {
  __shared__ double red[4];
  if (gate && gate[blockIdx.y] == 0) return;
  spec += blockIdx.y * ms_s;
  u0 += blockIdx.y * D;
  out += (size_t)blockIdx.y * D * nx;
  constexpr int DD = D * D;
  const size_t cs = (size_t)ncol * nx;
  const T *phin = phi + (size_t)ncol * DD * nx;  // ky = ny/2
  double e = 0.0;
  for (int it = blockIdx.x * 256 + threadIdx.x; it < D * nx;
       it += gridDim.x * 256) {
    const int i = it / nx, kx = it - i * nx;
    const int kn = (nx - kx) & (nx - 1);
    const bool mir = mirror && 2 * kx > nx;
    const unsigned kr = (unsigned)(mir ? nx - kx : kx);
    C2<T> x0[D], xn[D];
    const T h = (T)0.5;
#pragma unroll
    for (int c = 0; c < D; c++) {
      const C2<T> w = spec[c * cs + kx], wm = spec[c * cs + kn];
      x0[c] = cmake(h * (w.x + wm.x), h * (w.y - wm.y));
      xn[c] = cmake(h * (w.y + wm.y), h * (wm.x - w.x));
    }
    // row i of both tables (every array index below is the loop's j)
    C2<T> f0 = cmake((T)0, (T)0), fn = f0, x0i = f0, xni = f0;
#pragma unroll
    for (int j = 0; j < D; j++) {
      if (j == i) {
        const T p0 = phi[(unsigned)(i * nx) + kr];
        const T pn = phin[(unsigned)(i * nx) + kr];
        f0 = f0 + cscale(x0[j], p0);
        fn = fn + cscale(xn[j], pn);
        x0i = x0[j];
        xni = xn[j];
        continue;
      }
      const int a = j < i ? j : i, b = j < i ? i : j;
      const int p = D + 2 * (a * D - a * (a + 1) / 2 + b - a - 1);
      C2<T> q0 = cmake(phi[(unsigned)(p * nx) + kr],
                       phi[(unsigned)((p + 1) * nx) + kr]);
      C2<T> qn = cmake(phin[(unsigned)(p * nx) + kr],
                       phin[(unsigned)((p + 1) * nx) + kr]);
      // x mirror: Phi(-qx) = M Phi(qx) M, M = -1 on every x component
      if ((a % 3 == 0) != (b % 3 == 0) && mir) {
        q0 = cmake(-q0.x, -q0.y);
        qn = cmake(-qn.x, -qn.y);
      }
      if (j < i) {  // Phi_ij = conj(Phi_ji)
        f0 = f0 + cmulc(x0[j], q0);
        fn = fn + cmulc(xn[j], qn);
      } else {
        f0 = f0 + cmul(q0, x0[j]);
        fn = fn + cmul(qn, xn[j]);
      }
    }
    e += (double)(f0.x * x0i.x + f0.y * x0i.y) +
         (double)(fn.x * xni.x + fn.y * xni.y);
    f0 = cmake(-f0.x, -f0.y);
    fn = cmake(-fn.x, -fn.y);
    if (kx == 0) {
      u0[i] = (double)x0i.x;
      if (linf && i % 3 == 2) {
        e -= 2.0 * linf[i / 3] * (double)x0i.x;
        f0.x += (T)linf[i / 3];
      }
    }
    out[(size_t)i * nx + kx] = cmake(f0.x - fn.y, f0.y + fn.x);
  }
  const double bs = block_sum(e, red);
  if (threadIdx.x == 0)
    partials[(size_t)blockIdx.y * ms_p + part0 + blockIdx.x] = bs;
}

// ---------------------------------------------------------------------------
// Shapes, variants and dispatch.
//   row variant 0:  256 (FP32) / 512 (FP64) threads per block, prefetch
//                   (FP32: ~150 VGPRs allow three 4-wave blocks per CU,
//                   where 8-wave blocks of that size fit one)
//   row variant 1:  512 (FP32) / 256 (FP64) threads per block, prefetch
//   row variant 2: 1024 (FP32) / 512 (FP64) threads per block, no prefetch
//                   (fewer registers)
//   col variants: ColShape below (per precision)
// Row sweeps run persistent grids: G = (CUs x b) for the largest b within
// the occupancy that divides the work, so every CU gets the same number of
// tiles.  The column sweep launches one block per column group.
// ---------------------------------------------------------------------------
template <typename T, int N, int V>
struct RowShape {
  static constexpr int R = N >= 4096 ? 16 : 8;
  static constexpr int NT = N / R;
  static constexpr int TPB =
      sizeof(T) == 4 ? (V == 0 ? 256 : 1024 >> (V == 1)) : 512 >> (V == 1);
  static constexpr bool PF = V != 2;
  static constexpr int FPB0 = TPB / NT > 0 ? TPB / NT : 1;
  static constexpr int FPB = FPB0 > 16 ? 16 : FPB0;
  static constexpr int TR = 2 * FPB;
  // twiddle bases in registers, the other twiddles their products
  static constexpr bool TWC = true;
};

template <int N, int R, int CPB, int ES>
constexpr int col_lds_per_comp() {
  return CPB * FFTPlan<N, R>::lds_elems(ES) * ES;
}
template <int D, int PER, int LIMIT>
constexpr int col_dg_max() {
  return (D % 3 == 0 && 3 * PER <= LIMIT)   ? 3
         : (D % 2 == 0 && 2 * PER <= LIMIT) ? 2
                                            : 1;
}

template <typename T, int D, int N, int V>
struct ColShape;

// FP32: V = 1 halves the elements per thread (fewer registers, more
// threads); components per LDS exchange as fit in 64 KB (one per exchange
// for V = 3); register caps: none (V = 0, 1), 3 waves per SIMD (<= 168
// VGPRs, V = 2 and 3).  At 2048 with ndof 3, V = 3 needs 154 VGPRs and
// 18 KB of LDS per block, so a CU holds three blocks instead of the two
// that the 55 KB of the all-components exchange allow.
// ndof 9 and 12 (three- and four-atom cells) run the split column sweep
// (launch_cols_split); their shape here only fixes the number of energy
// partials (ncol / CPB).  Twiddles: each pass's base from the table, its
// powers by at most three products in registers (about 3 ulp; with a
// table load per power instead, each pass waited for its loads after the
// exchange barrier).
template <int D, int N, int V>
struct ColShape<float, D, N, V> {
  static constexpr bool BIG = D >= 9;
  static constexpr int R0 = N >= 4096 ? 16 : 8;
  static constexpr int R =
      BIG ? 4 : ((V == 1 && N / (R0 / 2) <= 1024) ? R0 / 2 : R0);
  static constexpr int NT = N / R;
  static constexpr int CPB = NT >= 256 ? 1 : 256 / NT;
  static constexpr int DG =
      (V == 3 && !BIG)
          ? 1
          : col_dg_max<D, col_lds_per_comp<N, R, CPB, 8>(), 65536>();
  static constexpr int MINW = BIG ? 1 : (V == 2 || V == 3 ? 3 : 1);
  static constexpr bool STRAIGHT = D <= 3;
  static constexpr bool TWC = !BIG;
};

// FP64 (16-byte elements, so a column holds twice the registers):
//   V = 0: 4 elements per thread, all components of a column group per LDS
//          exchange (up to 128 KB of LDS: one block per CU, a third of the
//          barriers and twiddle loads of V = 2)
//   V = 1: 8 elements per thread (three exchanges instead of five at
//          2048), one component per exchange, tables applied one
//          wavevector at a time, <= 256 VGPRs (two 256-thread blocks per CU)
//   V = 2: 4 elements per thread, one component per exchange
//   V = 3: as V = 1 with <= 168 VGPRs (three blocks per CU)
// ndof 9 and 12: the split column sweep, as for FP32 (this shape only
// fixes the partial count).
template <int D, int N, int V>
struct ColShape<double, D, N, V> {
  static constexpr bool BIG = D >= 9;
  static constexpr int R =
      BIG ? (N <= 2048 ? 2 : 4) : ((V == 1 || V == 3) ? 8 : 4);
  static constexpr int NT = N / R;
  static constexpr int CPB = NT >= 256 ? 1 : 256 / NT;
  static constexpr int DG =
      BIG ? col_dg_max<D, col_lds_per_comp<N, R, CPB, 16>(), 65536>()
          : (V == 0 ? col_dg_max<D, col_lds_per_comp<N, R, CPB, 16>(),
                                 131072>()
                    : 1);
  static constexpr int MINW = BIG ? 1 : (V == 1 ? 2 : (V == 3 ? 3 : 1));
  static constexpr bool STRAIGHT = D <= 3 && (V == 0 || V == 2);
  // large blocks keep their registers for the column's data
  static constexpr bool TWC = !BIG;
};

static inline int pick_grid(int units, int ncu, int occ) {
  if (occ < 1) occ = 1;
  for (int b = occ; b >= 1; b--)
    if (units % (ncu * b) == 0) return ncu * b;
  const int cap = ncu * occ;
  if (units <= cap) return units;
  const int rounds = (units + cap - 1) / cap;
  for (int g = (units + rounds - 1) / rounds; g <= cap; g++)
    if (units % g == 0) return g;
  return cap;
}

template <typename K>
static int occupancy(K kernel, int threads) {
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &occ, reinterpret_cast<const void *>(kernel), threads, 0) !=
      hipSuccess)
    occ = 1;
  return occ;
}

template <typename TI, typename T, int NY, int V>
static void launch_r2c(const ForceArgs &a, hipStream_t st) {
  using S = RowShape<T, NY, V>;
  auto kern = k_rows_r2c<TI, T, NY, S::R, S::FPB, S::PF, S::TWC>;
  static const int occ = occupancy(kern, S::NT * S::FPB);
  const int ntiles = (a.nx / S::TR) * a.d;
  hipLaunchKernelGGL(kern, dim3(pick_grid(ntiles, a.ncu, occ), a.members),
                     dim3(S::NT * S::FPB), 0, st, (const TI *)a.u,
                     (C2<T> *)a.spec, (const C2<T> *)a.twy, a.nx, ntiles,
                     a.gate, a.ms_u, a.ms_s);
}

template <typename TO, typename T, int NY, int V>
static void launch_c2r(const ForceArgs &a, hipStream_t st) {
  using S = RowShape<T, NY, V>;
  auto kern = k_rows_c2r<TO, T, NY, S::R, S::FPB, S::PF, S::TWC>;
  static const int occ = occupancy(kern, S::NT * S::FPB);
  const int ntiles = (a.nx / S::TR) * a.d;
  hipLaunchKernelGGL(kern, dim3(pick_grid(ntiles, a.ncu, occ), a.members),
                     dim3(S::NT * S::FPB), 0, st, (const C2<T> *)a.spec,
                     (TO *)a.f, (const C2<T> *)a.twy, a.nx, ntiles,
                     a.gate, a.ms_s, a.ms_f, (const double *)a.partials,
                     a.cols_parts, a.nparts, a.epot);
}

// ndof 9 and 12: transforms and product as separate launches.  The
// product kernels write as many energy partials as the fused kernel would
// (ncol / CPB of the variant's shape), so the c2r sweep sums the same count.
template <typename T, int D, int NX, int V>
static int launch_cols_split(const ForceArgs &a, hipStream_t st) {
  using S = ColShape<T, D, NX, V>;
  constexpr int R = 8;
  constexpr int NT = NX / R;
  constexpr int CPB = NT >= 256 ? 1 : 256 / NT;
  const int ncol = a.ny / 2;
  const int nlines = D * ncol;
  const int parts = ncol / S::CPB;
  // the packed column's kernel: up to one thread per (component, kx), at
  // most half of the partials
  int gdc = (D * NX + 255) / 256;
  if (gdc > parts / 2) gdc = parts / 2;
  const int gmain = parts - gdc;
  if (nlines % CPB || gdc < 1 || parts > a.nparts || ncol % S::CPB)
    return ST_STATE;
  const dim3 gf(nlines / CPB, a.members), bf(NT * CPB);
  // twiddle bases in registers in FP64 only (in FP32 their powers would
  // add rounding)
  constexpr bool TW = sizeof(T) == 8;
  hipLaunchKernelGGL((k_cols_fft<T, NX, R, CPB, -1, TW>), gf, bf, 0, st,
                     (C2<T> *)a.spec, (const C2<T> *)a.twx, a.gate, a.ms_s,
                     (const C2<T> *)nullptr, ncol, D);
  hipLaunchKernelGGL((k_cols_phi<T, D>), dim3(gmain, a.members), dim3(256),
                     0, st, (C2<T> *)a.spec, (const T *)a.phi, a.nx, ncol,
                     a.partials, a.gate, a.ms_s, a.nparts, a.mirror);
  hipLaunchKernelGGL((k_cols_phi_dc<T, D>), dim3(gdc, a.members), dim3(256),
                     0, st, (const C2<T> *)a.spec, (const T *)a.phi, a.nx,
                     ncol, a.linf, a.partials, a.u0, (C2<T> *)a.xbuf, a.gate,
                     a.ms_s, a.nparts, gmain, a.mirror);
  hipLaunchKernelGGL((k_cols_fft<T, NX, R, CPB, +1, TW>), gf, bf, 0, st,
                     (C2<T> *)a.spec, (const C2<T> *)a.twx, a.gate, a.ms_s,
                     (const C2<T> *)a.xbuf, ncol, D);
  if (a.ev_cols_end) (void)hipEventRecord(a.ev_cols_end, st);
  return ST_OK;
}

template <typename T, int D, int NX, int V, bool PAR = false>
static int launch_cols(const ForceArgs &a, hipStream_t st) {
  if constexpr (D >= 9) {
    return launch_cols_split<T, D, NX, V>(a, st);
  } else {
    if constexpr (D == 3 && !PAR) {
      if (a.par) return launch_cols<T, 3, NX, V, true>(a, st);
    }
    using S = ColShape<T, D, NX, V>;
    auto kern = k_cols<T, NX, S::R, D, S::DG, S::CPB, S::MINW, S::STRAIGHT,
                       S::TWC, PAR>;
    const int ncol = a.ny / 2;
    const int G = ncol / S::CPB;
    if (G > a.nparts || ncol % S::CPB) return ST_STATE;
    hipLaunchKernelGGL(kern, dim3(G, a.members), dim3(S::NT * S::CPB), 0, st,
                       (C2<T> *)a.spec, (const T *)a.phi,
                       (const C2<T> *)a.twx, ncol, a.linf, a.partials, a.u0,
                       (C2<T> *)a.xbuf, a.gate, a.ms_s, a.nparts, a.mirror);
    if (a.ev_cols_end) (void)hipEventRecord(a.ev_cols_end, st);
    return ST_OK;
  }
}

static inline bool pow2_in(int n, int lo, int hi) {
  return n >= lo && n <= hi && (n & (n - 1)) == 0;
}

template <typename T, int V>
static int row_tr(int ny) {
  switch (ny) {
    case 64: return RowShape<T, 64, V>::TR;
    case 128: return RowShape<T, 128, V>::TR;
    case 256: return RowShape<T, 256, V>::TR;
    case 512: return RowShape<T, 512, V>::TR;
    case 1024: return RowShape<T, 1024, V>::TR;
    case 2048: return RowShape<T, 2048, V>::TR;
    case 4096: return RowShape<T, 4096, V>::TR;
  }
  return 0;
}

template <typename T, int V>
static int col_cpb(int nx, int d) {
#define CASE(N)                                                         \
  case N:                                                               \
    return d == 3    ? ColShape<T, 3, N, V>::CPB                  \
           : d == 6  ? ColShape<T, 6, N, V>::CPB                  \
           : d == 9  ? ColShape<T, 9, N, V>::CPB                  \
           : d == 12 ? ColShape<T, 12, N, V>::CPB                 \
                     : ColShape<T, 1, N, V>::CPB;
  switch (nx) { CASE(64) CASE(128) CASE(256) CASE(512) CASE(1024) CASE(2048)
                CASE(4096) }
#undef CASE
  return 0;
}

template <typename T>
static int col_cpb_v(int nx, int d, int v) {
  switch (v) {
    case 0: return col_cpb<T, 0>(nx, d);
    case 1: return col_cpb<T, 1>(nx, d);
    case 2: return col_cpb<T, 2>(nx, d);
    case 3: return col_cpb<T, 3>(nx, d);
  }
  return 0;
}

template <typename TI, typename T, int V>
static int dispatch_rows_fwd(const ForceArgs &a, hipStream_t st) {
  switch (a.ny) {
    case 64: launch_r2c<TI, T, 64, V>(a, st); break;
    case 128: launch_r2c<TI, T, 128, V>(a, st); break;
    case 256: launch_r2c<TI, T, 256, V>(a, st); break;
    case 512: launch_r2c<TI, T, 512, V>(a, st); break;
    case 1024: launch_r2c<TI, T, 1024, V>(a, st); break;
    case 2048: launch_r2c<TI, T, 2048, V>(a, st); break;
    case 4096: launch_r2c<TI, T, 4096, V>(a, st); break;
    default: return ST_UNSUPPORTED;
  }
  return ST_OK;
}

template <typename TO, typename T, int V>
static int dispatch_rows_inv(const ForceArgs &a, hipStream_t st) {
  switch (a.ny) {
    case 64: launch_c2r<TO, T, 64, V>(a, st); break;
    case 128: launch_c2r<TO, T, 128, V>(a, st); break;
    case 256: launch_c2r<TO, T, 256, V>(a, st); break;
    case 512: launch_c2r<TO, T, 512, V>(a, st); break;
    case 1024: launch_c2r<TO, T, 1024, V>(a, st); break;
    case 2048: launch_c2r<TO, T, 2048, V>(a, st); break;
    case 4096: launch_c2r<TO, T, 4096, V>(a, st); break;
    default: return ST_UNSUPPORTED;
  }
  return ST_OK;
}

template <typename T, int V>
static int dispatch_cols_v(const ForceArgs &a, hipStream_t st) {
#define COLCASE(N)                                              \
  case N:                                                       \
    return a.d == 3    ? launch_cols<T, 3, N, V>(a, st)         \
           : a.d == 6  ? launch_cols<T, 6, N, V>(a, st)         \
           : a.d == 9  ? launch_cols<T, 9, N, V>(a, st)         \
           : a.d == 12 ? launch_cols<T, 12, N, V>(a, st)        \
                       : launch_cols<T, 1, N, V>(a, st);
  switch (a.nx) {
    COLCASE(64) COLCASE(128) COLCASE(256) COLCASE(512) COLCASE(1024)
    COLCASE(2048) COLCASE(4096)
  }
#undef COLCASE
  return ST_UNSUPPORTED;
}

template <typename T>
static int dispatch_cols(const ForceArgs &a, hipStream_t st) {
  switch (a.col_var) {
    case 0: return dispatch_cols_v<T, 0>(a, st);
    case 1: return dispatch_cols_v<T, 1>(a, st);
    case 2: return dispatch_cols_v<T, 2>(a, st);
    case 3: return dispatch_cols_v<T, 3>(a, st);
  }
  return ST_UNSUPPORTED;
}

// ev (optional, 5 events): before sweep 1, after sweep 1, after the column
// kernel, before sweep 3, after sweep 3 (which also sums the energy).
template <typename TIO, typename T>
int run_fast(const ForceArgs &a0, hipStream_t st, hipEvent_t *ev) {
  ForceArgs a = a0;
  a.ev_cols_end = ev ? ev[2] : nullptr;
  // one energy partial per column block; the c2r sweep sums them
  a.cols_parts = (a.ny / 2) / col_cpb_v<T>(a.nx, a.d, a.col_var);
  if (ev) (void)hipEventRecord(ev[0], st);
  int rc = a.row_var == 1   ? dispatch_rows_fwd<TIO, T, 1>(a, st)
           : a.row_var == 2 ? dispatch_rows_fwd<TIO, T, 2>(a, st)
                            : dispatch_rows_fwd<TIO, T, 0>(a, st);
  if (rc) return rc;
  if (ev) (void)hipEventRecord(ev[1], st);
  rc = dispatch_cols<T>(a, st);
  if (rc) return rc;
  if (ev) (void)hipEventRecord(ev[3], st);
  rc = a.row_var == 1   ? dispatch_rows_inv<TIO, T, 1>(a, st)
       : a.row_var == 2 ? dispatch_rows_inv<TIO, T, 2>(a, st)
                        : dispatch_rows_inv<TIO, T, 0>(a, st);
  if (ev) (void)hipEventRecord(ev[4], st);
  return rc;
}

}  // namespace gfmd
