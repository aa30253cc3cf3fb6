// Internal definitions shared by the HIP translation units of libgfmd_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "fft_core.h"

namespace gfmd {

// Status codes (mirrored in include/gfmd_hip.h).
enum {
  ST_OK = 0,
  ST_ARG = 1,
  ST_HIP = 2,
  ST_STATE = 3,
  ST_UNSUPPORTED = 4,
  ST_NONFINITE = 5,
  ST_KERNEL = 6,
  ST_CONTACT = 7,
};

void set_error(const char *fmt, ...);

#define GFMD_HIP_TRY(expr)                                                  \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) {                                                 \
      ::gfmd::set_error("%s failed: %s (%s:%d)", #expr,                     \
                        hipGetErrorString(e_), __FILE__, __LINE__);         \
      return ::gfmd::ST_HIP;                                                \
    }                                                                       \
  } while (0)

// Arguments of one force evaluation, in the layouts described in the
// library header: u, f are [d][nx][ny] (y fastest), the spectrum is the
// half spectrum stored column-major [d][nky][nx] so that transforms along
// x read contiguous memory, and phi is Hermitian-packed [nky][d*d][nx].
struct ForceArgs {
  int nx, ny, nky, d;
  const void *u;       // TI[d][nx][ny]
  void *f;             // TO[d][nx][ny]
  void *spec;          // C2<T>[d][nky][nx]
  void *spec2;         // second spectral buffer (generic path only)
  void *xbuf;          // C2<T>[d][nx] exchange buffer (fast path, column 0)
  const void *phi;     // T[nky][d*d][nx]
  const void *twx;     // C2<T>[nx], exp(-2 pi i n / nx)
  const void *twy;     // C2<T>[ny]
  const double *linf;  // [d/3] (device) or nullptr
  double *partials;    // [nparts] per-block energy partial sums
  int nparts;
  double *u0;          // [d] Re u(q=0) (device)
  double *epot;        // device scalar
  int ncu;             // compute units of the device
  hipEvent_t ev_cols_end = nullptr;  // recorded after the column kernel
  // fast path: the column sweep's partial count (set by the dispatch; the
  // c2r sweep sums them into epot)
  int cols_parts = 0;
  int row_var, col_var;  // kernel variants of the fast path (autotuned)
  // optional device flags [members]: 0 turns every kernel of the
  // evaluation into a no-op for that member (members that have finished)
  const int *gate = nullptr;
  // ensemble of `members` grids (fast path only): member m's fields start
  // m * ms_* elements (spectra, u, f) after the base pointers; partials
  // m * nparts, u0 m * d, xbuf m * d * nx, epot m; phi is shared
  int members = 1;
  size_t ms_u = 0, ms_f = 0, ms_s = 0;
  // phi is symmetric under x -> -x (fast path: half of it is read)
  int mirror = 0;
  // d = 3, fast path: phi is the compact table [nky][6][nx] (Phi_xy real,
  // Phi_xz and Phi_yz imaginary; launch_phi_compact)
  int par = 0;
};

// Number of kernel variants of the fast path per sweep.
constexpr int N_ROW_VARIANTS = 3;
constexpr int N_COL_VARIANTS = 4;
bool force_fast_variant_ok(int nx, int ny, int d, bool fp64, int row_var,
                           int col_var);

// Sum of v over the block, valid in thread 0 (red: one double per wave).
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
  return s;
}

// E = 1/2 sum(partials[0..n)), summed in a fixed order.
void launch_energy_final(const double *partials, int n, double *epot,
                         hipStream_t s, const int *gate = nullptr,
                         int members = 1, int ms_p = 0);

// Dispatch: returns ST_OK or ST_UNSUPPORTED when no fast-path instance
// exists for (nx, ny, d).  io_double selects double u/f buffers when the
// compute type is float.
// ev (optional, 5 events): before sweep 1, after sweep 1, after the column
// kernel, before sweep 3, after sweep 3 (which also sums the energy), on
// stream s.
int launch_force_fast(const ForceArgs &a, bool fp64, bool io_double,
                      hipStream_t s, hipEvent_t *ev = nullptr);
int launch_force_generic(const ForceArgs &a, bool fp64, bool io_double,
                         hipStream_t s);
// Number of energy partials a launch needs (the largest over variants).
int force_fast_nparts(int nx, int ny, int d, bool fp64);
bool force_fast_supported(int nx, int ny, int d, bool fp64);
int force_generic_nparts(int nx, int ny, int d);
// Whether the generic path transforms by the mixed-radix FFT (lines of at
// most 4096 points) rather than by direct DFTs.
bool force_mixed_radix(int nx, int ny);
// Twiddle table exp(-2 pi i k / n), k < n, written on the device.
void launch_twiddles(void *tw, int n, bool fp64, hipStream_t s);

// Phi(q) tabulation on the device (half spectrum, Hermitian packed,
// normalised by 1/(nx*ny) when `normalize`).
struct KernelDesc {
  int kind;         // KIND_* below
  int d;            // degrees of freedom per surface cell
  int height;       // transfer-matrix depth (0: Phi = U0, <0: semi-inf.)
  double p[8];      // kernel parameters
  const struct FTTable *ft = nullptr;  // ft: host tables (ft_lattice.h)
};
enum {
  KIND_SC100 = 1,
  KIND_ISOTROPIC = 2,
  KIND_FILTER = 3,
  KIND_FCC100 = 4,
  KIND_NONPERIODIC = 5,
  KIND_FT = 6,
  KIND_LOWPASS = 7
};
// fcc100: p[0..2] k, p[3..5] kappa (three shells), p[6] surface normal
// isotropic: p[0] Poisson number, p[1] shear modulus, p[2] gamma, p[3] xprd
// nonperiodic: p[0] Poisson number, p[1] shear modulus, p[2] dx, p[3] dy
// lowpass (d = 3): p[0] 1 / lambda, p[1] xprd, p[2] yprd

// Rough-surface generation (roughness.hip): white noise into component 0 of
// a [3][nx][ny] field, and the rms slope (central differences) of a field.
void launch_white_noise(double *w, int nx, int ny, unsigned long long seed,
                        hipStream_t s);
void launch_slope_partials(const double *h, int nx, int ny, double a,
                           double *part, int nblk, hipStream_t s);
void launch_scale(double *h, size_t n, const double *part, int nblk,
                  double target, hipStream_t s);

int launch_phi_generate(const KernelDesc &k, int nx, int ny, int normalize,
                        bool fp64, void *phi, int *d_status, hipStream_t s);
// Sets *d_bad (device int, cleared by the caller) unless the packed table
// satisfies Phi(nx - kx) = M Phi(kx) M, M = -1 on the x component of every
// layer, for every kx > nx/2, to tol times the trace at that wavevector.
void launch_phi_mirror_check(const void *phi, int nx, int ny, int d,
                             bool fp64, double tol, int *d_bad,
                             hipStream_t s);
// ndof 3: *d_bad (device int, cleared by the caller) is set unless
// Im Phi_xy = Re Phi_xz = Re Phi_yz = 0 exactly at every stored q; then
// launch_phi_compact writes the 6-real table the force sweep reads.
void launch_phi_par_check(const void *phi, int nx, int ny, bool fp64,
                          int *d_bad, hipStream_t s);
void launch_phi_compact(const void *phi, void *pc, int nx, int ny, bool fp64,
                        hipStream_t s);
// Compliance tables of the contact solver from the packed Phi~ table:
// gz = -(Phi~^-1)_{zc,zc} / N^2 (one real per q) and, if gfull is not
// NULL, the packed -Phi~^-1 / N^2 (d*d reals per q).
// With cavg (device, d x d real symmetric): gfull = the packed
// -(Phi~ + cavg)^-1 unscaled (the prec_gradient table), gz may be NULL.
int launch_compliance(const void *phi, int nx, int ny, int d, bool fp64,
                      int zc, void *gz, void *gfull, int *d_status,
                      hipStream_t s, const double *cavg = nullptr);
// Pack a host-provided full-grid complex table (already on the device,
// [nx*ny][d][d] complex128) into the half-spectrum layout.
int launch_phi_pack(const double *full, int nx, int ny, int d, double scale,
                    bool fp64, void *phi, hipStream_t s);

}  // namespace gfmd
