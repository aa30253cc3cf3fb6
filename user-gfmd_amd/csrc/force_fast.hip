// Fast-path variant table and entry point; the kernels are instantiated
// per compute precision in force_fast_f32.hip and force_fast_f64.hip.
#include "force_fast_impl.h"

namespace gfmd {

// E = 1/2 sum(partials), summed in a fixed order (bitwise reproducible).
__global__ void __launch_bounds__(256)
    k_energy_final(const double *__restrict__ partials, int n,
                   double *__restrict__ epot, const int *__restrict__ gate,
                   int ms_p) {
  __shared__ double red[4];
  // blockIdx.y: member of an ensemble
  if (gate && gate[blockIdx.y] == 0) return;
  partials += (size_t)blockIdx.y * ms_p;
  epot += blockIdx.y;
  // all four loads of a round are issued before the adds (one memory
  // round trip per 1024 partials)
  double s = 0.0;
  for (int i0 = threadIdx.x; i0 < n; i0 += 1024) {
    double x[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
      x[k] = i0 + 256 * k < n ? partials[i0 + 256 * k] : 0.0;
    s += (x[0] + x[1]) + (x[2] + x[3]);
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) *epot = 0.5 * s;
}

void launch_energy_final(const double *partials, int n, double *epot,
                         hipStream_t st, const int *gate, int members,
                         int ms_p) {
  hipLaunchKernelGGL(k_energy_final, dim3(1, members), dim3(256), 0, st,
                     partials, n, epot, gate, ms_p);
}

// exp(-2 pi i k / n) for k < n, formed on the device (sincospi in FP64,
// then rounded to T): the twiddle tables of the allocation-free evaluation,
// which cannot upload host tables.
template <typename T>
__global__ void __launch_bounds__(256) k_twiddles(C2<T> *__restrict__ tw,
                                                  int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    double sn, cs;
    sincospi(-2.0 * (double)k / (double)n, &sn, &cs);
    tw[k] = cmake((T)cs, (T)sn);
  }
}

void launch_twiddles(void *tw, int n, bool fp64, hipStream_t st) {
  const dim3 g((n + 255) / 256);
  if (fp64)
    hipLaunchKernelGGL(k_twiddles<double>, g, dim3(256), 0, st,
                       (C2<double> *)tw, n);
  else
    hipLaunchKernelGGL(k_twiddles<float>, g, dim3(256), 0, st,
                       (C2<float> *)tw, n);
}

int run_fast_f32(const ForceArgs &a, bool io_double, hipStream_t st,
                 hipEvent_t *ev);
int run_fast_f64(const ForceArgs &a, hipStream_t st, hipEvent_t *ev);

bool force_fast_variant_ok(int nx, int ny, int d, bool fp64, int row_var,
                           int col_var) {
  if (!pow2_in(nx, 64, 4096) || !pow2_in(ny, 64, 4096)) return false;
  // d = 1: scalar convolutions (the contact solver's compliance G_zz);
  // 3, 6, 9, 12: one to four atoms per surface cell
  if (d != 1 && d != 3 && d != 6 && d != 9 && d != 12) return false;
  if (row_var < 0 || row_var >= N_ROW_VARIANTS) return false;
  if (col_var < 0 || col_var >= N_COL_VARIANTS) return false;
  int tr;
  if (fp64) tr = row_var == 1 ? row_tr<double, 1>(ny) : row_tr<double, 0>(ny);
  else tr = row_var == 1 ? row_tr<float, 1>(ny) : row_tr<float, 0>(ny);
  const int cpb = fp64 ? col_cpb_v<double>(nx, d, col_var)
                       : col_cpb_v<float>(nx, d, col_var);
  return tr > 0 && nx % tr == 0 && cpb > 0 && (ny / 2) % cpb == 0;
}

bool force_fast_supported(int nx, int ny, int d, bool fp64) {
  return force_fast_variant_ok(nx, ny, d, fp64, 0, 0);
}

int force_fast_nparts(int nx, int ny, int d, bool fp64) {
  int most = 0;
  for (int v = 0; v < N_COL_VARIANTS; v++) {
    const int cpb = fp64 ? col_cpb_v<double>(nx, d, v)
                         : col_cpb_v<float>(nx, d, v);
    if (cpb > 0 && (ny / 2) / cpb > most) most = (ny / 2) / cpb;
  }
  return most;
}

int launch_force_fast(const ForceArgs &a, bool fp64, bool io_double,
                      hipStream_t st, hipEvent_t *ev) {
  if (!force_fast_variant_ok(a.nx, a.ny, a.d, fp64, a.row_var, a.col_var))
    return ST_UNSUPPORTED;
  if (fp64) return run_fast_f64(a, st, ev);
  return run_fast_f32(a, io_double, st, ev);
}

}  // namespace gfmd
