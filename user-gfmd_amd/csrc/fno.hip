// Learned warm start (Fourier neural operator): BF16 MFMA matrix products
// and the fused elementwise passes of its forward and backward.
//
// Every dense product of the network is one batched strided GEMM on the
// matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation): the
// pointwise channel mixing (sites x 32 x 32), the truncated transforms of
// the spectral convolution written as products with DFT matrices (only the
// retained modes are ever formed), and their transposes in the backward.
// The mode-wise complex mixing (544 modes x 32 x 32 per sample) is small
// and runs in FP32.
#include "fno.h"

#include <cmath>
#include <cstdint>

namespace gfmd {
namespace fno {

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ bf16 f2bf(float f) {
  // round to nearest even (finite inputs)
  const uint32_t u = __float_as_uint(f);
  return (bf16)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(bf16 h) {
  return __uint_as_float((uint32_t)h << 16);
}

// ---------------------------------------------------------------------------
// GEMM: 64 x 64 output tile per block of 4 waves (each a 32 x 32 quadrant =
// 2 x 2 MFMA 16x16 tiles), K in steps of 32 through LDS.  The lane map of
// the 16x16x32 MFMA used here: operand A row = lane % 16, operand B column
// = lane % 16, result column = lane % 16 and rows 4 (lane / 16) + j.  The
// k slots of a lane are filled with the same k indices for A and B, so the
// product is independent of how the hardware orders k within a step.
// ---------------------------------------------------------------------------
constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;

__global__ void __launch_bounds__(256)
    k_gemm(int M, int N, int K, const bf16 *__restrict__ A, long long sA,
           long long rsA, long long csA, const bf16 *__restrict__ B,
           long long sB, long long rsB, long long csB, float *__restrict__ Cm,
           long long sC, long long rsC, long long csC, float alpha,
           int accumulate) {
  __shared__ __attribute__((aligned(16))) bf16 As[BM * LDK];
  __shared__ __attribute__((aligned(16))) bf16 Bs[BN * LDK];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  // M tiles on x (up to millions of sites), N tiles on y, batch on z
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const long long bz = blockIdx.z;
  A += bz * sA;
  B += bz * sB;
  Cm += bz * sC;
  // load maps: 8 consecutive elements per thread along the unit-stride
  // dimension when there is one, as one 16-byte load when every row start
  // is 16-byte aligned (base aligned, strides multiples of 8 elements)
  const bool a_kmaj = csA == 1 || rsA != 1;  // k contiguous (or general)
  const bool b_nmaj = csB == 1 || rsB != 1;  // n contiguous (or general)
  auto al8 = [](long long v) { return (v & 7) == 0; };
  const bool a_vec = ((uintptr_t)A & 15) == 0 && al8(sA) &&
                     (a_kmaj ? (csA == 1 && al8(rsA)) : al8(csA));
  const bool b_vec = ((uintptr_t)B & 15) == 0 && al8(sB) &&
                     (b_nmaj ? (csB == 1 && al8(rsB)) : al8(csB));
  f32x4 acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int r = 0; r < 4; r++) acc[i][j][r] = 0.f;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  for (int k0 = 0; k0 < K; k0 += BK) {
    // A tile -> As[m][k]
    if (a_kmaj) {
      const int m = t >> 2, kk = (t & 3) * 8;
      const int gm = m0 + m;
      if (a_vec && gm < M && k0 + kk + 8 <= K) {
        *reinterpret_cast<bf16x8 *>(&As[m * LDK + kk]) =
            *reinterpret_cast<const bf16x8 *>(A + gm * rsA + k0 + kk);
      } else {
        for (int e = 0; e < 8; e++) {
          const int gk = k0 + kk + e;
          As[m * LDK + kk + e] =
              (gm < M && gk < K) ? A[gm * rsA + gk * csA] : (bf16)0;
        }
      }
    } else {
      const int kk = t >> 3, m = (t & 7) * 8;
      const int gk = k0 + kk;
      if (a_vec && gk < K && m0 + m + 8 <= M) {
        const bf16x8 v =
            *reinterpret_cast<const bf16x8 *>(A + (m0 + m) + gk * csA);
        for (int e = 0; e < 8; e++) As[(m + e) * LDK + kk] = (bf16)v[e];
      } else {
        for (int e = 0; e < 8; e++) {
          const int gm = m0 + m + e;
          As[(m + e) * LDK + kk] =
              (gm < M && gk < K) ? A[gm * rsA + gk * csA] : (bf16)0;
        }
      }
    }
    // B tile -> Bs[n][k]
    if (b_nmaj) {
      const int kk = t >> 3, n = (t & 7) * 8;
      const int gk = k0 + kk;
      if (b_vec && gk < K && n0 + n + 8 <= N) {
        const bf16x8 v =
            *reinterpret_cast<const bf16x8 *>(B + gk * rsB + (n0 + n));
        for (int e = 0; e < 8; e++) Bs[(n + e) * LDK + kk] = (bf16)v[e];
      } else {
        for (int e = 0; e < 8; e++) {
          const int gn = n0 + n + e;
          Bs[(n + e) * LDK + kk] =
              (gn < N && gk < K) ? B[gk * rsB + gn * csB] : (bf16)0;
        }
      }
    } else {
      const int n = t >> 2, kk = (t & 3) * 8;
      const int gn = n0 + n;
      if (b_vec && gn < N && k0 + kk + 8 <= K) {
        *reinterpret_cast<bf16x8 *>(&Bs[n * LDK + kk]) =
            *reinterpret_cast<const bf16x8 *>(B + (k0 + kk) + gn * csB);
      } else {
        for (int e = 0; e < 8; e++) {
          const int gk = k0 + kk + e;
          Bs[n * LDK + kk + e] =
              (gn < N && gk < K) ? B[gk * rsB + gn * csB] : (bf16)0;
        }
      }
    }
    __syncthreads();
    const int kq = (lane >> 4) * 8;
    bf16x8 av[2], bv[2];
    for (int i = 0; i < 2; i++)
      av[i] = *reinterpret_cast<const bf16x8 *>(
          &As[(wm + i * 16 + (lane & 15)) * LDK + kq]);
    for (int j = 0; j < 2; j++)
      bv[j] = *reinterpret_cast<const bf16x8 *>(
          &Bs[(wn + j * 16 + (lane & 15)) * LDK + kq]);
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j],
                                                            acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) {
      const int gn = n0 + wn + j * 16 + (lane & 15);
      if (gn >= N) continue;
      for (int r = 0; r < 4; r++) {
        const int gm = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        if (gm >= M) continue;
        float *c = Cm + gm * rsC + gn * csC;
        const float v = alpha * acc[i][j][r];
        *c = accumulate ? *c + v : v;
      }
    }
}

void launch_gemm(const Gemm &g, hipStream_t st) {
  dim3 grid((g.M + BM - 1) / BM, (g.N + BN - 1) / BN, g.batch);
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, st, g.M, g.N, g.K, g.A, g.sA,
                     g.rsA, g.csA, g.B, g.sB, g.rsB, g.csB, g.C, g.sC, g.rsC,
                     g.csC, g.alpha, g.accumulate);
}

// ---------------------------------------------------------------------------
// Elementwise passes
// ---------------------------------------------------------------------------
static unsigned blocks(size_t n, unsigned cap = 8192) {
  const size_t b = (n + 255) / 256;
  return (unsigned)(b < cap ? b : cap);
}
#define GRID_STRIDE(i, n)                                              \
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (n); \
       i += (size_t)gridDim.x * blockDim.x)

__device__ __forceinline__ float gelu(float x) {
  return 0.5f * x * (1.f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_d(float x) {
  return 0.5f * (1.f + erff(x * 0.70710678118654752f)) +
         x * 0.3989422804014327f * expf(-0.5f * x * x);
}

__global__ void k_to_bf16(const float *__restrict__ x, bf16 *__restrict__ y,
                          size_t n) {
  GRID_STRIDE(i, n) y[i] = f2bf(x[i]);
}
void launch_to_bf16(const float *x, bf16 *y, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(n)), dim3(256), 0, st, x, y, n);
}

__global__ void k_lift(const float *__restrict__ a,
                       const float *__restrict__ prm, float *__restrict__ v,
                       bf16 *__restrict__ vb, size_t np) {
  const float *P = prm + OFF_P, *bP = prm + OFF_BP;
  GRID_STRIDE(i, np * C) {
    const size_t p = i / C;
    const int c = (int)(i - p * C);
    float s = bP[c];
    for (int k = 0; k < CIN; k++) s += a[p * CIN + k] * P[k * C + c];
    v[i] = s;
    vb[i] = f2bf(s);
  }
}
void launch_lift(const float *a, const float *prm, float *v, bf16 *vb,
                 size_t np, hipStream_t st) {
  hipLaunchKernelGGL(k_lift, dim3(blocks(np * C)), dim3(256), 0, st, a, prm,
                     v, vb, np);
}

__global__ void k_bias_act(const float *__restrict__ acc,
                           const float *__restrict__ bias,
                           float *__restrict__ z, float *__restrict__ v,
                           bf16 *__restrict__ vb, size_t np, int act) {
  GRID_STRIDE(i, np * C) {
    const float zi = acc[i] + bias[i % C];
    z[i] = zi;
    const float vi = act ? gelu(zi) : zi;
    if (v) v[i] = vi;
    vb[i] = f2bf(vi);
  }
}
void launch_bias_act(const float *acc, const float *bias, float *z, float *v,
                     bf16 *vb, size_t np, bool act, hipStream_t st) {
  hipLaunchKernelGGL(k_bias_act, dim3(blocks(np * C)), dim3(256), 0, st, acc,
                     bias, z, v, vb, np, (int)act);
}

__global__ void k_project(const float *__restrict__ v,
                          const float *__restrict__ prm,
                          float *__restrict__ u, size_t np) {
  const float *Q = prm + OFF_Q, *bQ = prm + OFF_BQ;
  GRID_STRIDE(i, np * COUT) {
    const size_t p = i / COUT;
    const int o = (int)(i - p * COUT);
    float s = bQ[o];
    for (int c = 0; c < C; c++) s += v[p * C + c] * Q[c * COUT + o];
    u[i] = s;
  }
}
void launch_project(const float *v, const float *prm, float *u, size_t np,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_project, dim3(blocks(np * COUT)), dim3(256), 0, st, v,
                     prm, u, np);
}

// block partial sums of NQ values into part[q][block]
template <int NQ>
__device__ __forceinline__ void block_partials(double (&acc)[NQ],
                                               double *__restrict__ part) {
  __shared__ double red[4][NQ];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int q = 0; q < NQ; q++) {
    double s = acc[q];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) red[w][q] = s;
  }
  __syncthreads();
  if (threadIdx.x < NQ) {
    double s = 0.0;
    for (int i = 0; i < 4; i++) s += red[i][threadIdx.x];
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}

// out[q] (+)= sum_b part[q][b] (FP64 sums, FP32 out); one block per q
__global__ void k_finish(const double *__restrict__ part, int nblk,
                         float *__restrict__ out, double *__restrict__ out64,
                         int accumulate, double scale = 1.0) {
  __shared__ double red[4];
  const int q = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x)
    s += part[(size_t)q * nblk + b];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    s = scale * (red[0] + red[1] + red[2] + red[3]);
    if (out) out[q] = accumulate ? out[q] + (float)s : (float)s;
    if (out64) out64[q] = s;
  }
}

constexpr unsigned NRED = 1024;  // blocks of the reducing passes

__global__ void __launch_bounds__(256)
    k_loss(const float *__restrict__ u, const float *__restrict__ t,
           float *__restrict__ du, double *__restrict__ part, size_t n) {
  double acc[1] = {0.0};
  const float sc = 2.f / (float)n;
  GRID_STRIDE(i, n) {
    const float d = u[i] - t[i];
    acc[0] += (double)d * d;
    du[i] = sc * d;
  }
  block_partials<1>(acc, part);
}
void launch_loss(const float *u, const float *t, float *du, double *part,
                 double *loss, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_loss, dim3(NRED), dim3(256), 0, st, u, t, du, part, n);
  // mean over the n values
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, st, part, (int)NRED,
                     (float *)nullptr, loss, 0, 1.0 / (double)n);
}

// dv = du Q^T; partial sums of v^T du (C x COUT) and du (COUT)
__global__ void __launch_bounds__(256)
    k_project_back(const float *__restrict__ v, const float *__restrict__ du,
                   const float *__restrict__ prm, float *__restrict__ dv,
                   double *__restrict__ part, size_t np) {
  const float *Q = prm + OFF_Q;
  double acc[C * COUT + COUT];
  for (int q = 0; q < C * COUT + COUT; q++) acc[q] = 0.0;
  GRID_STRIDE(p, np) {
    float d[COUT];
    for (int o = 0; o < COUT; o++) {
      d[o] = du[p * COUT + o];
      acc[C * COUT + o] += d[o];
    }
    for (int c = 0; c < C; c++) {
      const float vc = v[p * C + c];
      float s = 0.f;
      for (int o = 0; o < COUT; o++) {
        s += d[o] * Q[c * COUT + o];
        acc[c * COUT + o] += (double)vc * d[o];
      }
      dv[p * C + c] = s;
    }
  }
  block_partials<C * COUT + COUT>(acc, part);
}
void launch_project_back(const float *v, const float *du, const float *prm,
                         float *dv, float *grad, double *part, size_t np,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_project_back, dim3(NRED), dim3(256), 0, st, v, du, prm,
                     dv, part, np);
  // gQ (C x COUT) then gbQ (COUT): contiguous in the flat vector
  hipLaunchKernelGGL(k_finish, dim3(C * COUT + COUT), dim3(256), 0, st, part,
                     (int)NRED, grad + OFF_Q, (double *)nullptr, 1, 1.0);
}

// dz = dv gelu'(z) (or dv); bf16 copy; partial sums of dz per channel.
// Threads walk the flat [np][C] arrays (coalesced); the block size and the
// grid-stride step are multiples of C, so thread t only meets channel t % C
// and keeps one accumulator.
static_assert(256 % C == 0, "block size must be a multiple of C");
__global__ void __launch_bounds__(256)
    k_act_back(const float *__restrict__ dv, const float *__restrict__ z,
               float *__restrict__ dz, bf16 *__restrict__ dzb,
               double *__restrict__ part, size_t np, int act) {
  __shared__ double red[256];
  double acc = 0.0;
  GRID_STRIDE(i, np * C) {
    const float g = act ? dv[i] * gelu_d(z[i]) : dv[i];
    dz[i] = g;
    dzb[i] = f2bf(g);
    acc += g;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < C) {
    double s = 0.0;
    for (int k = threadIdx.x; k < 256; k += C) s += red[k];
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}
void launch_act_back(const float *dv, const float *z, float *dz, bf16 *dzb,
                     float *gbias, double *part, size_t np, bool act,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_act_back, dim3(NRED), dim3(256), 0, st, dv, z, dz, dzb,
                     part, np, (int)act);
  hipLaunchKernelGGL(k_finish, dim3(C), dim3(256), 0, st, part, (int)NRED,
                     gbias, (double *)nullptr, 1, 1.0);
}

__global__ void __launch_bounds__(256)
    k_lift_back(const float *__restrict__ a, const float *__restrict__ dv,
                double *__restrict__ part, size_t np) {
  double acc[CIN * C + C];
  for (int q = 0; q < CIN * C + C; q++) acc[q] = 0.0;
  GRID_STRIDE(p, np) {
    float ai[CIN];
    for (int k = 0; k < CIN; k++) ai[k] = a[p * CIN + k];
    for (int c = 0; c < C; c++) {
      const float d = dv[p * C + c];
      for (int k = 0; k < CIN; k++) acc[k * C + c] += (double)ai[k] * d;
      acc[CIN * C + c] += d;
    }
  }
  block_partials<CIN * C + C>(acc, part);
}
void launch_lift_back(const float *a, const float *dv, float *grad,
                      double *part, size_t np, hipStream_t st) {
  hipLaunchKernelGGL(k_lift_back, dim3(NRED), dim3(256), 0, st, a, dv, part,
                     np);
  // gP (CIN x C) then gbP (C): contiguous at the start of the flat vector
  hipLaunchKernelGGL(k_finish, dim3(CIN * C + C), dim3(256), 0, st, part,
                     (int)NRED, grad + OFF_P, (double *)nullptr, 1, 1.0);
}

__global__ void k_sum_rows(const float *__restrict__ part, int rows, int cols,
                           float *__restrict__ out, int accumulate) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  double s = 0.0;
  for (int r = 0; r < rows; r++) s += part[(size_t)r * cols + c];
  out[c] = accumulate ? out[c] + (float)s : (float)s;
}
void launch_sum_rows(const float *part, int rows, int cols, float *out,
                     bool accumulate, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_rows, dim3((cols + 255) / 256), dim3(256), 0, st,
                     part, rows, cols, out, (int)accumulate);
}

// Mode mixing.  Z, O: [b][2 * NKX rows (kx, re/im)][NKY * C]; R: [kx][ky]
// [ci][co][re/im].  One block per (mode, sample), one thread per co.
__global__ void k_mix(const float *__restrict__ Z, const float *__restrict__ R,
                      float *__restrict__ O) {
  const int mode = blockIdx.x, b = blockIdx.y, co = threadIdx.x;
  const int a = mode / NKY, ky = mode - a * NKY;
  const size_t row = (size_t)NKY * C;
  const float *zr = Z + ((size_t)b * 2 * NKX + 2 * a) * row + ky * C;
  const float *zi = zr + row;
  const float *r = R + (size_t)mode * C * C * 2;
  float sr = 0.f, si = 0.f;
  for (int ci = 0; ci < C; ci++) {
    const float xr = zr[ci], xi = zi[ci];
    const float wr = r[(ci * C + co) * 2], wi = r[(ci * C + co) * 2 + 1];
    sr += xr * wr - xi * wi;
    si += xr * wi + xi * wr;
  }
  float *o = O + ((size_t)b * 2 * NKX + 2 * a) * row + ky * C;
  o[co] = sr;
  o[row + co] = si;
}
void launch_mix(const float *Z, const float *R, float *O, int batch,
                hipStream_t st) {
  hipLaunchKernelGGL(k_mix, dim3(NMODE, batch), dim3(C), 0, st, Z, R, O);
}

// dZ = dO R^H (per sample); dR += sum_b conj(Z) dO.  One block per mode,
// C x C threads over (ci, co) for dR, then C threads for dZ.
__global__ void __launch_bounds__(1024)
    k_mix_back(const float *__restrict__ Z, const float *__restrict__ dO,
               const float *__restrict__ R, float *__restrict__ dZ,
               float *__restrict__ dR, int batch) {
  const int mode = blockIdx.x;
  const int a = mode / NKY, ky = mode - a * NKY;
  const size_t row = (size_t)NKY * C;
  const int ci = threadIdx.x / C, co = threadIdx.x % C;
  const float *r = R + (size_t)mode * C * C * 2;
  float gr = 0.f, gi = 0.f;
  for (int b = 0; b < batch; b++) {
    const float *zr = Z + ((size_t)b * 2 * NKX + 2 * a) * row + ky * C;
    const float *gor = dO + ((size_t)b * 2 * NKX + 2 * a) * row + ky * C;
    const float xr = zr[ci], xi = zr[row + ci];
    const float yr = gor[co], yi = gor[row + co];
    // conj(x) y
    gr += xr * yr + xi * yi;
    gi += xr * yi - xi * yr;
  }
  float *g = dR + (size_t)mode * C * C * 2;
  g[(ci * C + co) * 2] += gr;
  g[(ci * C + co) * 2 + 1] += gi;
  if (threadIdx.x < C) {
    const int i = threadIdx.x;
    for (int b = 0; b < batch; b++) {
      const float *gor = dO + ((size_t)b * 2 * NKX + 2 * a) * row + ky * C;
      float sr = 0.f, si = 0.f;
      for (int o = 0; o < C; o++) {
        const float yr = gor[o], yi = gor[row + o];
        const float wr = r[(i * C + o) * 2], wi = r[(i * C + o) * 2 + 1];
        // y conj(w)
        sr += yr * wr + yi * wi;
        si += yi * wr - yr * wi;
      }
      float *dz = dZ + ((size_t)b * 2 * NKX + 2 * a) * row + ky * C;
      dz[i] = sr;
      dz[row + i] = si;
    }
  }
}
void launch_mix_back(const float *Z, const float *dO, const float *R,
                     float *dZ, float *dR, int batch, hipStream_t st) {
  hipLaunchKernelGGL(k_mix_back, dim3(NMODE), dim3(C * C), 0, st, Z, dO, R, dZ,
                     dR, batch);
}

__global__ void k_adam(float *__restrict__ prm, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, size_t n,
                       float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  GRID_STRIDE(i, n) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    prm[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}
void launch_adam(float *prm, const float *grad, float *m, float *v, size_t n,
                 float lr, float b1, float b2, float eps, int step,
                 hipStream_t st) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, st, prm, grad, m,
                     v, n, lr, b1, b2, eps, c1, c2);
}

}  // namespace fno
}  // namespace gfmd
