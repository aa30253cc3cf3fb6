// Learned warm start: Fourier neural operator on the device (fno.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gfmd {
namespace fno {

// Architecture (fixed): C channels, modes kx in [-M, M), ky in [0, M],
// L spectral layers, CIN input and COUT output channels.
constexpr int C = 32, M = 16, L = 4, CIN = 2, COUT = 3;
constexpr int NKX = 2 * M, NKY = M + 1;
constexpr int NMODE = NKX * NKY;           // 544 retained modes
constexpr int ROWS_Y = 2 * NKY;            // re/im rows of the y transform
constexpr size_t R_FLOATS = (size_t)NMODE * C * C * 2;
// flat parameter vector: P, bP, {W_l, b_l, R_l}, Q, bQ
constexpr size_t OFF_P = 0, OFF_BP = CIN * C;
constexpr size_t OFF_L0 = OFF_BP + C;
constexpr size_t LAYER_FLOATS = C * C + C + R_FLOATS;
constexpr size_t OFF_Q = OFF_L0 + L * LAYER_FLOATS;
constexpr size_t OFF_BQ = OFF_Q + C * COUT;
constexpr size_t NPARAM = OFF_BQ + COUT;
inline size_t off_W(int l) { return OFF_L0 + l * LAYER_FLOATS; }
inline size_t off_b(int l) { return off_W(l) + C * C; }
inline size_t off_R(int l) { return off_b(l) + C; }

using bf16 = uint16_t;  // bfloat16 bits

// C[b](MxN) = alpha * A[b](MxK) B[b](KxN) (+ C[b] when accumulate), A, B
// BF16 with element strides (batch, row, column), C FP32.  BF16 MFMA with
// FP32 accumulation.
struct Gemm {
  int M, N, K, batch;
  const bf16 *A;
  long long sA, rsA, csA;
  const bf16 *B;
  long long sB, rsB, csB;
  float *C;
  long long sC, rsC, csC;
  float alpha;
  int accumulate;
};
void launch_gemm(const Gemm &g, hipStream_t st);

// elementwise helpers (n elements)
void launch_to_bf16(const float *x, bf16 *y, size_t n, hipStream_t st);
// lift: v[p][c] = sum_i a[p][i] P[i][c] + bP[c]  (a: CIN per site)
void launch_lift(const float *a, const float *prm, float *v, bf16 *vb,
                 size_t np, hipStream_t st);
// z = acc + bias (per channel); v = gelu(z) (or z when !act); vb = bf16(v)
void launch_bias_act(const float *acc, const float *bias, float *z, float *v,
                     bf16 *vb, size_t np, bool act, hipStream_t st);
// u[p][o] = sum_c v[p][c] Q[c][o] + bQ[o]
void launch_project(const float *v, const float *prm, float *u, size_t np,
                    hipStream_t st);
// loss = mean (u - t)^2 into *loss (FP64, device); du = 2 (u - t) / n
void launch_loss(const float *u, const float *t, float *du, double *part,
                 double *loss, size_t n, hipStream_t st);
// backward of the projection: dv = du Q^T; gQ += v^T du, gbQ += sum du
void launch_project_back(const float *v, const float *du, const float *prm,
                         float *dv, float *grad, double *part, size_t np,
                         hipStream_t st);
// dz = dv * gelu'(z) (or dv); dzb = bf16(dz); gb += sum_p dz
void launch_act_back(const float *dv, const float *z, float *dz, bf16 *dzb,
                     float *gbias, double *part, size_t np, bool act,
                     hipStream_t st);
// backward of the lift: gP += a^T dv, gbP += sum dv
void launch_lift_back(const float *a, const float *dv, float *grad,
                      double *part, size_t np, hipStream_t st);
// column sums of the split-K partial products: out[c] (+)= sum_s part[s][c]
void launch_sum_rows(const float *part, int rows, int cols, float *out,
                     bool accumulate, hipStream_t st);
// spectral mode mixing: O[b][mode] = Z[b][mode] R[mode] (complex, C x C)
// layouts [b][(kx, re/im)][ky * C + c] (x-transform rows)
void launch_mix(const float *Z, const float *R, float *O, int batch,
                hipStream_t st);
// backward: dZ = dO R^H; dR += sum_b conj(Z) dO
void launch_mix_back(const float *Z, const float *dO, const float *R,
                     float *dZ, float *dR, int batch, hipStream_t st);
// Adam on the flat FP32 parameters
void launch_adam(float *prm, const float *grad, float *m, float *v, size_t n,
                 float lr, float b1, float b2, float eps, int step,
                 hipStream_t st);

}  // namespace fno
}  // namespace gfmd
