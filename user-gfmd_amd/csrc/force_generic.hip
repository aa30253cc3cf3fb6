// Generic elastic force evaluation: any nx, ny (including the odd and
// prime sizes of the reference's tests, e.g. 64x37, 63x63), by mixed-radix
// FFTs for sides up to 4096 and by direct DFTs, O(N (nx + ny)) work, beyond.
// Same conventions as the fast FFT path
// (GFMDSolverStatic::post_force, src/solvers/gfmd_solver_static.cpp:145-249):
// half spectrum spec[d][ny/2+1][nx], Hermitian extension in the inverse.
#include "gfmd_internal.h"

namespace gfmd {

// ---------------------------------------------------------------------------
// Generic path: direct DFTs (any nx, ny).
// ---------------------------------------------------------------------------
// spec[c][ky][ix] = sum_iy u[c][ix][iy] w_ny^(ky*iy)
template <typename TI, typename T>
__global__ void k_gen_rows_r2c(const TI *__restrict__ u, C2<T> *__restrict__ spec,
                               const C2<T> *__restrict__ twy, int nx, int ny,
                               int nky, int d, const int *__restrict__ gate) {
  if (gate && *gate == 0) return;
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
                           int d, const int *__restrict__ gate) {
  if (gate && *gate == 0) return;
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

// D > 0: ndof known at compile time (the per-q arrays stay in registers);
// D = 0: any ndof up to 24 (d at run time).
template <typename T, int D>
__global__ __launch_bounds__(256) void k_gen_apply(
    C2<T> *__restrict__ q, const T *__restrict__ phi,
                            int nx, int ny, int nky, int d_run,
                            const double *__restrict__ linf,
                            double *__restrict__ partials, double *__restrict__ u0,
                            const int *__restrict__ gate) {
  __shared__ double red[4];
  if (gate && *gate == 0) return;
  constexpr int DM = D > 0 ? D : 24;
  const int d = D > 0 ? D : d_run;
  const int dd = d * d;
  const size_t nq = (size_t)nky * nx;
  double e = 0.0;
  for (size_t L = blockIdx.x * (size_t)blockDim.x + threadIdx.x; L < nq;
       L += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(L % nx);
    const int ky = (int)(L / nx);
    C2<T> uq[DM], F[DM];
#pragma unroll
    for (int c = 0; c < d; c++) uq[c] = q[(size_t)c * nq + L];
    const T *ph = phi + (size_t)ky * dd * nx + kx;
#pragma unroll
    for (int i = 0; i < d; i++) {
      const T pii = ph[(size_t)i * nx];
      F[i] = cmake(pii * uq[i].x, pii * uq[i].y);
    }
    int p = d;
#pragma unroll
    for (int i = 0; i < d; i++)
#pragma unroll
      for (int j = i + 1; j < d; j++) {
        const C2<T> pij = cmake(ph[(size_t)p * nx], ph[(size_t)(p + 1) * nx]);
        p += 2;
        F[i] = F[i] + cmul(pij, uq[j]);
        F[j] = F[j] + cmulc(uq[i], pij);
      }
    T eq = 0;
#pragma unroll
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
#pragma unroll
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
                               int nky, int d, const int *__restrict__ gate) {
  if (gate && *gate == 0) return;
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

// ---------------------------------------------------------------------------
// Mixed-radix path: any line length up to MR_MAX, O(n sum r_i) per line.
//
// Each line (rows: two real rows packed as one complex line) is held in
// LDS together with the line's n-th roots of unity and transformed by 256
// threads; a row workgroup holds up to four such lines.
// Stockham autosort over the factors of n (radix 4, then 2, 3, 5, 7, ...,
// a remaining prime as one stage of its own): with span s (product of the
// radices already done) and radix R, output i = g s R + a s + k
// (k < s, a < R) is
//   y[i] = sum_t x[g s + k + t n/R] w_n^(t (k + a s) n/(s R)),
// the twiddle and the radix-R DFT in one sum.  Each thread forms its
// outputs in registers, then the line is overwritten in place.
// ---------------------------------------------------------------------------
constexpr int MR_MAX = 4096;
constexpr int MR_THREADS = 256;

struct MrPlan {
  int n, nst;
  int r[24];
};

static MrPlan mr_plan(int n) {
  MrPlan p{};
  p.n = n;
  int m = n;
  while (m % 4 == 0) p.r[p.nst++] = 4, m /= 4;
  while (m % 2 == 0) p.r[p.nst++] = 2, m /= 2;
  for (int f = 3; m > 1 && f * f <= m; f += 2)
    while (m % f == 0) p.r[p.nst++] = f, m /= f;
  if (m > 1) p.r[p.nst++] = m;
  return p;
}

// One Stockham stage (span s, radix R) over L lines of length n in LDS,
// one output per item: y[i] = sum_t x[j + t m] w_n^(t (k + a s) n/(s R)).
// Any R (a line of prime length is one such stage).
template <typename T, int NMAX, int L, int SIGN>
__device__ void mr_stage_sum(C2<T> (*x)[NMAX], const C2<T> *tw, int n, int s,
                             int R) {
  constexpr int PER = NMAX / MR_THREADS;
  constexpr int ITEMS = L * PER;
  const int m = n / R;
  const int sr = s * R;
  const int unit = n / sr;
  C2<T> y[ITEMS];
#pragma unroll
  for (int q = 0; q < ITEMS; q++) {
    const int i = threadIdx.x + (q % PER) * MR_THREADS;
    if (i < n) {
      const C2<T> *xl = x[q / PER];
      const int k = i % s;
      const int a = (i / s) % R;
      const int j = (i / sr) * s + k;
      const int step = (k + a * s) * unit;  // < n
      C2<T> acc = xl[j];
      int e = step;
      for (int t = 1; t < R; t++) {
        C2<T> w = tw[e];
        if (SIGN > 0) w = cconj(w);
        acc = acc + cmul(xl[j + t * m], w);
        e += step;
        if (e >= n) e -= n;
      }
      y[q] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < ITEMS; q++) {
    const int i = threadIdx.x + (q % PER) * MR_THREADS;
    if (i < n) x[q / PER][i] = y[q];
  }
  __syncthreads();
}

// The same stage for a small radix R known at compile time, one butterfly
// per item: the R inputs x[j + t m] read once and twiddled by
// w_n^(k t n/(s R)), an R-point DFT in registers against the R-th roots of
// unity (tw[p n/R]), the R outputs written to j/s s R + k + a s.
template <typename T, int NMAX, int L, int SIGN, int R>
__device__ void mr_stage_bf(C2<T> (*x)[NMAX], const C2<T> *tw, int n, int s) {
  constexpr int BF = (L * NMAX / R + MR_THREADS - 1) / MR_THREADS;
  const int m = n / R;
  const int unit = n / (s * R);
  const int nbf = L * m;
  C2<T> wr[R];
#pragma unroll
  for (int p = 0; p < R; p++) {
    const C2<T> w = tw[p * m];
    wr[p] = SIGN > 0 ? cconj(w) : w;
  }
  C2<T> y[BF][R];
#pragma unroll
  for (int b = 0; b < BF; b++) {
    const int q = threadIdx.x + b * MR_THREADS;
    if (q < nbf) {
      const int line = q / m, j = q - line * m;
      const C2<T> *xl = x[line];
      const int step = (j % s) * unit;  // k t unit < n for t < R
      C2<T> v[R];
      v[0] = xl[j];
#pragma unroll
      for (int t = 1; t < R; t++) {
        C2<T> w = tw[t * step];
        if (SIGN > 0) w = cconj(w);
        v[t] = cmul(xl[j + t * m], w);
      }
#pragma unroll
      for (int a = 0; a < R; a++) {
        C2<T> acc = v[0];
#pragma unroll
        for (int t = 1; t < R; t++) acc = acc + cmul(v[t], wr[(a * t) % R]);
        y[b][a] = acc;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < BF; b++) {
    const int q = threadIdx.x + b * MR_THREADS;
    if (q < nbf) {
      const int line = q / m, j = q - line * m;
      const int k = j % s;
      const int o = (j - k) * R + k;
#pragma unroll
      for (int a = 0; a < R; a++) x[line][o + a * s] = y[b][a];
    }
  }
  __syncthreads();
}

// The transform of L lines of one length (n = p.n <= NMAX), stage by stage
// (one barrier pair per stage for all lines); every thread of the 256 must
// call it.
template <typename T, int NMAX, int L, int SIGN>
__device__ void mr_fft_lines(C2<T> (*x)[NMAX], const C2<T> *tw,
                             const MrPlan &p) {
  const int n = p.n;
  int s = 1;
  for (int st = 0; st < p.nst; st++) {
    const int R = p.r[st];
    switch (R) {
      case 2: mr_stage_bf<T, NMAX, L, SIGN, 2>(x, tw, n, s); break;
      case 3: mr_stage_bf<T, NMAX, L, SIGN, 3>(x, tw, n, s); break;
      case 4: mr_stage_bf<T, NMAX, L, SIGN, 4>(x, tw, n, s); break;
      case 5: mr_stage_bf<T, NMAX, L, SIGN, 5>(x, tw, n, s); break;
      case 7: mr_stage_bf<T, NMAX, L, SIGN, 7>(x, tw, n, s); break;
      default: mr_stage_sum<T, NMAX, L, SIGN>(x, tw, n, s, R);
    }
    s *= R;
  }
}

template <typename T, int NMAX, int SIGN>
__device__ void mr_fft(C2<T> *x, const C2<T> *tw, const MrPlan &p) {
  mr_fft_lines<T, NMAX, 1, SIGN>(reinterpret_cast<C2<T>(*)[NMAX]>(x), tw, p);
}

// P pairs of rows per workgroup (threadIdx.y = pair): rows ra = 2 q and
// rb = ra + 1 of u[d*nx][ny] as one complex line; their half spectra are
// separated by Hermitian symmetry and stored transposed, spec[c][ky][ix],
// the 2P rows of the workgroup side by side (one 16 P- or 32 P-byte run
// per ky).
template <typename TI, typename T, int NMAX, int P>
__global__ __launch_bounds__(MR_THREADS * P) void k_mr_rows_r2c(
    const TI *__restrict__ u, C2<T> *__restrict__ spec,
    const C2<T> *__restrict__ twy, MrPlan p, int nx, int nky, int d,
    const int *__restrict__ gate) {
  __shared__ C2<T> xs[P][NMAX];
  __shared__ C2<T> tw[NMAX];
  if (gate && *gate == 0) return;
  const int ny = p.n;
  const int nrows = d * nx;
  const int r0 = 2 * P * blockIdx.x;
  const int ra = r0 + 2 * threadIdx.y;
  C2<T> *x = xs[threadIdx.y];
  const TI *ua = u + (size_t)ra * ny;
  const T zero = 0;
  for (int i = threadIdx.x; i < ny; i += MR_THREADS) {
    if (threadIdx.y == 0) tw[i] = twy[i];
    const T va = ra < nrows ? (T)ua[i] : zero;
    const T vb = ra + 1 < nrows ? (T)ua[ny + i] : zero;
    x[i] = cmake(va, vb);
  }
  __syncthreads();
  mr_fft<T, NMAX, -1>(x, tw, p);
  // (ky, row) with the row fastest: a wave writes runs of 2P neighbours
  const int t = threadIdx.y * MR_THREADS + threadIdx.x;
  const T h = (T)0.5;
  for (int L = t; L < nky * 2 * P; L += MR_THREADS * P) {
    const int k = L / (2 * P), r = L % (2 * P);
    const int row = r0 + r;
    if (row >= nrows) continue;
    const C2<T> *xq = xs[r >> 1];
    const C2<T> z = xq[k];
    const C2<T> zc = cconj(xq[k ? ny - k : 0]);
    const C2<T> v = (r & 1) ? cmake((z.y - zc.y) * h, (zc.x - z.x) * h)
                            : cmake((z.x + zc.x) * h, (z.y + zc.y) * h);
    spec[((size_t)(row / nx) * nky + k) * nx + row % nx] = v;
  }
}

// One column line spec[line][0..nx) per workgroup, out = DFT_SIGN(in).
template <typename T, int NMAX, int SIGN>
__global__ __launch_bounds__(MR_THREADS) void k_mr_cols(
    const C2<T> *__restrict__ in, C2<T> *__restrict__ out,
    const C2<T> *__restrict__ twx, MrPlan p, const int *__restrict__ gate) {
  __shared__ C2<T> x[NMAX];
  __shared__ C2<T> tw[NMAX];
  if (gate && *gate == 0) return;
  const int nx = p.n;
  const size_t base = (size_t)blockIdx.x * nx;
  for (int i = threadIdx.x; i < nx; i += MR_THREADS) {
    tw[i] = twx[i];
    x[i] = in[base + i];
  }
  __syncthreads();
  mr_fft<T, NMAX, SIGN>(x, tw, p);
  for (int i = threadIdx.x; i < nx; i += MR_THREADS) out[base + i] = x[i];
}

// Inverse of k_mr_rows_r2c: per pair, the two rows' half spectra, extended
// by Hermitian symmetry (the imaginary parts of ky = 0 and ky = ny/2
// dropped), as Za + i Zb; one inverse transform gives row ra in the real
// part and row rb in the imaginary part.  The gather reads the 2P rows of
// the workgroup side by side.
template <typename TO, typename T, int NMAX, int P>
__global__ __launch_bounds__(MR_THREADS * P) void k_mr_rows_c2r(
    const C2<T> *__restrict__ spec, TO *__restrict__ f,
    const C2<T> *__restrict__ twy, MrPlan p, int nx, int nky, int d,
    const int *__restrict__ gate) {
  __shared__ C2<T> xs[P][NMAX];
  __shared__ C2<T> tw[NMAX];
  if (gate && *gate == 0) return;
  const int ny = p.n;
  const int nrows = d * nx;
  const int r0 = 2 * P * blockIdx.x;
  const int t = threadIdx.y * MR_THREADS + threadIdx.x;
  for (int i = t; i < ny; i += MR_THREADS * P) tw[i] = twy[i];
  for (int L = t; L < ny * P; L += MR_THREADS * P) {
    const int k = L / P, q = L % P;
    const int ra = r0 + 2 * q;
    const bool lo = k < nky;
    const int kk = lo ? k : ny - k;
    C2<T> za = cmake((T)0, (T)0), zb = za;
    if (ra < nrows) za = spec[((size_t)(ra / nx) * nky + kk) * nx + ra % nx];
    if (ra + 1 < nrows)
      zb = spec[((size_t)((ra + 1) / nx) * nky + kk) * nx + (ra + 1) % nx];
    if (k == 0 || 2 * k == ny) za.y = zb.y = 0;
    if (!lo) za = cconj(za), zb = cconj(zb);
    xs[q][k] = cmake(za.x - zb.y, za.y + zb.x);
  }
  __syncthreads();
  C2<T> *x = xs[threadIdx.y];
  mr_fft<T, NMAX, +1>(x, tw, p);
  const int ra = r0 + 2 * threadIdx.y;
  TO *fa = f + (size_t)ra * ny;
  for (int i = threadIdx.x; i < ny; i += MR_THREADS) {
    if (ra < nrows) fa[i] = (TO)x[i].x;
    if (ra + 1 < nrows) fa[ny + i] = (TO)x[i].y;
  }
}

// ---------------------------------------------------------------------------
// Fused columns (mixed-radix path): one workgroup per ky holds the D column
// lines spec[c][ky][0..nx) in LDS, transforms them along x, applies Phi(q)
// at every kx (energy, Gamma-point terms), transforms back and stores in
// place: one pass over the spectrum instead of three.  Used when the D lines
// and the twiddles fit 128 KB of LDS.
// ---------------------------------------------------------------------------
template <typename T, int D, int NMAX>
constexpr bool mr_fused_fits() {
  // LDS for the D lines and the twiddles, and at most 24 values per thread
  // in a stage (more spill to private scratch memory)
  return (D + 1) * NMAX * (int)sizeof(C2<T>) <= 128 * 1024 &&
         D * NMAX <= 24 * MR_THREADS;
}

template <typename T, int NMAX, int D>
__global__ __launch_bounds__(MR_THREADS) void k_mr_cols_apply(
    C2<T> *__restrict__ spec, const T *__restrict__ phi,
    const C2<T> *__restrict__ twx, MrPlan p, int ny, int nky,
    const double *__restrict__ linf, double *__restrict__ partials,
    double *__restrict__ u0, const int *__restrict__ gate) {
  __shared__ C2<T> xs[D][NMAX];
  __shared__ C2<T> tw[NMAX];
  __shared__ double red[MR_THREADS / 64];
  if (gate && *gate == 0) return;
  const int nx = p.n;
  const int ky = blockIdx.x;
  constexpr int DD = D * D;
  for (int i = threadIdx.x; i < nx; i += MR_THREADS) {
    tw[i] = twx[i];
#pragma unroll
    for (int c = 0; c < D; c++)
      xs[c][i] = spec[((size_t)c * nky + ky) * nx + i];
  }
  __syncthreads();
  mr_fft_lines<T, NMAX, D, -1>(xs, tw, p);
  // Phi.u at every kx of this ky (the thread's own positions: no barrier
  // needed before the writes)
  const double wgt = (ky == 0 || 2 * ky == ny) ? 1.0 : 2.0;
  double e = 0.0;
  for (int kx = threadIdx.x; kx < nx; kx += MR_THREADS) {
    C2<T> uq[D], F[D];
#pragma unroll
    for (int c = 0; c < D; c++) uq[c] = xs[c][kx];
    const T *ph = phi + (size_t)ky * DD * nx + kx;
#pragma unroll
    for (int i = 0; i < D; i++) {
      const T pii = ph[(size_t)i * nx];
      F[i] = cmake(pii * uq[i].x, pii * uq[i].y);
    }
    int q = D;
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
      for (int j = i + 1; j < D; j++) {
        const C2<T> pij = cmake(ph[(size_t)q * nx], ph[(size_t)(q + 1) * nx]);
        q += 2;
        F[i] = F[i] + cmul(pij, uq[j]);
        F[j] = F[j] + cmulc(uq[i], pij);
      }
    T eq = 0;
#pragma unroll
    for (int i = 0; i < D; i++) eq += F[i].x * uq[i].x + F[i].y * uq[i].y;
    e += wgt * (double)eq;
    const bool g0 = ky == 0 && kx == 0;
    if (g0) {
#pragma unroll
      for (int c = 0; c < D; c++) u0[c] = (double)uq[c].x;
      if (linf)
        for (int i = 0; i < D / 3; i++)
          e -= 2.0 * linf[i] * (double)uq[3 * i + 2].x;
    }
#pragma unroll
    for (int c = 0; c < D; c++) {
      C2<T> o = cmake(-F[c].x, -F[c].y);
      if (g0 && linf && c % 3 == 2) o.x += (T)linf[c / 3];
      xs[c][kx] = o;
    }
  }
  __syncthreads();
  mr_fft_lines<T, NMAX, D, +1>(xs, tw, p);
  for (int i = threadIdx.x; i < nx; i += MR_THREADS)
#pragma unroll
    for (int c = 0; c < D; c++)
      spec[((size_t)c * nky + ky) * nx + i] = xs[c][i];
  const double bs = block_sum(e, red);
  if (threadIdx.x == 0) partials[ky] = bs;
}

bool force_mixed_radix(int nx, int ny) { return nx <= MR_MAX && ny <= MR_MAX; }

// Pairs of rows per workgroup: as many as fit 128 KB of LDS with the
// twiddles, at most 4 (1024 threads).  For 16 KB lines the 128-VGPR budget
// of 1024 threads spills 24-32 bytes per lane, and still runs faster than
// two pairs (FP32 2000^2: 3038 against 2624 evaluations/s).
template <typename T, int NMAX>
constexpr int mr_pairs() {
  constexpr int line = NMAX * (int)sizeof(C2<T>);
  constexpr int fit = (128 * 1024 - line) / line;
  return fit >= 4 ? 4 : fit >= 2 ? 2 : 1;
}

template <typename TIO, typename T, int NMAX>
static void run_mr_rows(const ForceArgs &a, bool forward, hipStream_t st) {
  constexpr int P = mr_pairs<T, NMAX>();
  const MrPlan py = mr_plan(a.ny);
  const unsigned g = (unsigned)((a.d * a.nx + 2 * P - 1) / (2 * P));
  const dim3 blk(MR_THREADS, P);
  if (forward)
    hipLaunchKernelGGL((k_mr_rows_r2c<TIO, T, NMAX, P>), dim3(g), blk, 0, st,
                       (const TIO *)a.u, (C2<T> *)a.spec,
                       (const C2<T> *)a.twy, py, a.nx, a.nky, a.d, a.gate);
  else
    hipLaunchKernelGGL((k_mr_rows_c2r<TIO, T, NMAX, P>), dim3(g), blk, 0, st,
                       (const C2<T> *)a.spec, (TIO *)a.f,
                       (const C2<T> *)a.twy, py, a.nx, a.nky, a.d, a.gate);
}

template <typename TIO, typename T>
static void mr_rows(const ForceArgs &a, bool forward, hipStream_t st) {
  if (a.ny <= 1024) run_mr_rows<TIO, T, 1024>(a, forward, st);
  else if (a.ny <= 2048) run_mr_rows<TIO, T, 2048>(a, forward, st);
  else run_mr_rows<TIO, T, MR_MAX>(a, forward, st);
}

template <typename T, int NMAX, int SIGN>
static void run_mr_cols(const ForceArgs &a, const void *in, void *out,
                        hipStream_t st) {
  hipLaunchKernelGGL((k_mr_cols<T, NMAX, SIGN>), dim3(a.d * a.nky),
                     dim3(MR_THREADS), 0, st, (const C2<T> *)in, (C2<T> *)out,
                     (const C2<T> *)a.twx, mr_plan(a.nx), a.gate);
}

template <typename T, int SIGN>
static void mr_cols(const ForceArgs &a, const void *in, void *out,
                    hipStream_t st) {
  if (a.nx <= 1024) run_mr_cols<T, 1024, SIGN>(a, in, out, st);
  else if (a.nx <= 2048) run_mr_cols<T, 2048, SIGN>(a, in, out, st);
  else run_mr_cols<T, MR_MAX, SIGN>(a, in, out, st);
}

// The fused column kernel for the smallest line size that holds nx, if its
// D lines fit the LDS budget; false: use the separate transform and product
// launches.
template <typename T, int D, int NMAX>
static bool mr_fused_try(const ForceArgs &a, hipStream_t st) {
  if constexpr (mr_fused_fits<T, D, NMAX>()) {
    hipLaunchKernelGGL((k_mr_cols_apply<T, NMAX, D>), dim3(a.nky),
                       dim3(MR_THREADS), 0, st, (C2<T> *)a.spec,
                       (const T *)a.phi, (const C2<T> *)a.twx, mr_plan(a.nx),
                       a.ny, a.nky, a.linf, a.partials, a.u0, a.gate);
    return true;
  }
  return false;
}

// Lines longer than 2048 points keep the separate launches: there one
// 256-thread workgroup per CU holds its lines, too few waves for the
// transforms (FP32, 3000^2: 507 against 585 evaluations/s).
template <typename T, int D>
static bool mr_fused_d(const ForceArgs &a, hipStream_t st) {
  if (a.nx <= 512) return mr_fused_try<T, D, 512>(a, st);
  if (a.nx <= 1024) return mr_fused_try<T, D, 1024>(a, st);
  if (a.nx <= 2048) return mr_fused_try<T, D, 2048>(a, st);
  return false;
}

template <typename T>
static bool mr_fused(const ForceArgs &a, hipStream_t st) {
  switch (a.d) {
    case 1: return mr_fused_d<T, 1>(a, st);
    case 3: return mr_fused_d<T, 3>(a, st);
    case 6: return mr_fused_d<T, 6>(a, st);
    case 9: return mr_fused_d<T, 9>(a, st);
    case 12: return mr_fused_d<T, 12>(a, st);
  }
  return false;
}

int force_generic_nparts(int nx, int ny, int d) {
  const size_t nq = (size_t)(ny / 2 + 1) * nx;
  size_t b = (nq + 255) / 256;
  b = b < 1024 ? b : 1024;
  // the fused column kernel writes one partial per ky
  const size_t nky = (size_t)(ny / 2 + 1);
  return (int)(b > nky ? b : nky);
}

template <typename TIO, typename T>
static int run_generic(const ForceArgs &a, hipStream_t st) {
  const int nky = a.nky;
  auto nb = [](size_t n) {
    size_t b = (n + 255) / 256;
    return (unsigned)(b < 4096 ? b : 4096);
  };
  const size_t ns = (size_t)a.d * nky * a.nx;
  const bool mr = force_mixed_radix(a.nx, a.ny);
  if (!mr) {
    hipLaunchKernelGGL((k_gen_rows_r2c<TIO, T>), dim3(nb(ns)), dim3(256), 0,
                       st, (const TIO *)a.u, (C2<T> *)a.spec,
                       (const C2<T> *)a.twy, a.nx, a.ny, nky, a.d, a.gate);
    hipLaunchKernelGGL((k_gen_cols<T, -1>), dim3(nb(ns)), dim3(256), 0, st,
                       (const C2<T> *)a.spec, (C2<T> *)a.spec2,
                       (const C2<T> *)a.twx, a.nx, nky, a.d, a.gate);
  } else {
    if (a.nky > a.nparts) return ST_STATE;
    mr_rows<TIO, T>(a, true, st);
    if (mr_fused<T>(a, st)) {
      launch_energy_final(a.partials, a.nky, a.epot, st, a.gate);
      mr_rows<TIO, T>(a, false, st);
      return ST_OK;
    }
    mr_cols<T, -1>(a, a.spec, a.spec2, st);
  }
  const int nparts = force_generic_nparts(a.nx, a.ny, a.d);
  if (nparts > a.nparts) return ST_STATE;
  auto apply = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nparts), dim3(256), 0, st, (C2<T> *)a.spec2,
                       (const T *)a.phi, a.nx, a.ny, nky, a.d, a.linf,
                       a.partials, a.u0, a.gate);
  };
  switch (a.d) {
    case 1: apply(k_gen_apply<T, 1>); break;
    case 3: apply(k_gen_apply<T, 3>); break;
    case 6: apply(k_gen_apply<T, 6>); break;
    case 9: apply(k_gen_apply<T, 9>); break;
    case 12: apply(k_gen_apply<T, 12>); break;
    default: apply(k_gen_apply<T, 0>);
  }
  launch_energy_final(a.partials, nparts, a.epot, st, a.gate);
  if (!mr) {
    hipLaunchKernelGGL((k_gen_cols<T, +1>), dim3(nb(ns)), dim3(256), 0, st,
                       (const C2<T> *)a.spec2, (C2<T> *)a.spec,
                       (const C2<T> *)a.twx, a.nx, nky, a.d, a.gate);
    const size_t nf = (size_t)a.d * a.nx * a.ny;
    hipLaunchKernelGGL((k_gen_rows_c2r<TIO, T>), dim3(nb(nf)), dim3(256), 0,
                       st, (const C2<T> *)a.spec, (TIO *)a.f,
                       (const C2<T> *)a.twy, a.nx, a.ny, nky, a.d, a.gate);
  } else {
    mr_cols<T, +1>(a, a.spec2, a.spec, st);
    mr_rows<TIO, T>(a, false, st);
  }
  return ST_OK;
}

int launch_force_generic(const ForceArgs &a, bool fp64, bool io_double,
                         hipStream_t st) {
  if (a.d > 24 || a.members != 1) return ST_UNSUPPORTED;
  if (fp64) return run_generic<double, double>(a, st);
  if (io_double) return run_generic<double, float>(a, st);
  return run_generic<float, float>(a, st);
}

}  // namespace gfmd
