// Host-side C++ adapter: see gfmd_solver_hip.h.
#include "gfmd_solver_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace gfmd_host {

GFMDSolverStaticHIP::GFMDSolverStaticHIP(int narg, int *iarg, char **arg)
    : name_("static/hip"), precision_(GFMD_FP64), nx_(0), ny_(0), ndof_(0),
      h_(nullptr) {
  // Option parsing in the style of GFMDSolverStatic's constructor
  // (gfmd_solver_static.cpp:49-64): consume recognised keywords only.
  if (narg > 0 && iarg && arg) {
    while (*iarg < narg) {
      if (!strcmp(arg[*iarg], "fp32")) {
        precision_ = GFMD_FP32;
        (*iarg)++;
      } else if (!strcmp(arg[*iarg], "fp64")) {
        precision_ = GFMD_FP64;
        (*iarg)++;
      } else {
        break;
      }
    }
  }
}

GFMDSolverStaticHIP::~GFMDSolverStaticHIP() {
  if (h_) gfmd_destroy(h_);
}

void GFMDSolverStaticHIP::check(int rc, const char *what) const {
  if (rc != GFMD_OK)
    throw std::runtime_error(std::string("GFMDSolverStaticHIP::") + what +
                             ": " + gfmd_last_error());
}

void GFMDSolverStaticHIP::set_grid_size(int nx, int ny, int ndof) {
  if (h_) {
    gfmd_destroy(h_);
    h_ = nullptr;
  }
  nx_ = nx;
  ny_ = ny;
  ndof_ = ndof;
  u0_.assign(ndof, 0.0);
  check(gfmd_create(nx, ny, ndof, precision_, &h_), "set_grid_size");
}

void GFMDSolverStaticHIP::set_kernel(const char *name, const char *args,
                                     bool normalize) {
  check(gfmd_set_kernel(h_, name, args, normalize ? 1 : 0), "set_kernel");
}

void GFMDSolverStaticHIP::set_kernel_args(const char *name,
                                          const char *args) {
  kname_ = name ? name : "";
  kargs_ = args ? args : "";
}

// Wavevector of grid index i on an n-point axis, as fill_phi_buffer
// computes it (gfmd_misc.cpp:40-47).
static double grid_q(int i, int n) {
  return (i <= n / 2) ? 2.0 * M_PI * i / n : 2.0 * M_PI * (i - n) / n;
}

double GFMDSolverStaticHIP::verify_device_table(StiffnessProvider *kernel,
                                                bool normalize) {
  // Gamma, the zone edges and corners, neighbours of Gamma, and seeded
  // pseudo-random points of the stored half spectrum (0 <= ky <= ny/2)
  std::vector<int> kx = {0, nx_ / 2, 0, nx_ / 2, 1 % nx_, 0, 1 % nx_,
                         nx_ - 1, nx_ - 1, nx_ / 2};
  std::vector<int> ky = {0, 0, ny_ / 2, ny_ / 2, 0, 1 % (ny_ / 2 + 1),
                         1 % (ny_ / 2 + 1), 1 % (ny_ / 2 + 1), ny_ / 2,
                         1 % (ny_ / 2 + 1)};
  unsigned long long st = 0x9e3779b97f4a7c15ULL ^ ((unsigned long long)nx_ << 20) ^ ny_;
  while (kx.size() < 24) {
    st = st * 6364136223846793005ULL + 1442695040888963407ULL;
    kx.push_back((int)((st >> 33) % (unsigned long long)nx_));
    st = st * 6364136223846793005ULL + 1442695040888963407ULL;
    ky.push_back((int)((st >> 33) % (unsigned long long)(ny_ / 2 + 1)));
  }
  const int n = (int)kx.size(), dsq = ndof_ * ndof_;
  std::vector<double> got((size_t)n * dsq * 2);
  check(gfmd_stiffness_at(h_, n, kx.data(), ky.data(), got.data()),
        "set_kernel");
  const double scale = normalize ? 1.0 / ((double)nx_ * ny_) : 1.0;
  std::vector<double> a((size_t)dsq * 2), b((size_t)dsq * 2);
  double err = 0.0, ref = 0.0;
  for (int q = 0; q < n; q++) {
    // stored: (Phi(q) + conj Phi(-q)) / 2, -q taken on the grid
    kernel->get_stiffness_matrix(nx_, ny_, grid_q(kx[q], nx_),
                                 grid_q(ky[q], ny_), a.data());
    const int px = (nx_ - kx[q]) % nx_, py = (ny_ - ky[q]) % ny_;
    kernel->get_stiffness_matrix(nx_, ny_, grid_q(px, nx_), grid_q(py, ny_),
                                 b.data());
    for (int e = 0; e < dsq; e++) {
      const double re = 0.5 * (a[2 * e] + b[2 * e]) * scale;
      const double im = 0.5 * (a[2 * e + 1] - b[2 * e + 1]) * scale;
      const double *g = &got[((size_t)q * dsq + e) * 2];
      err = std::max(err, std::hypot(g[0] - re, g[1] - im));
      ref = std::max(ref, std::hypot(re, im));
    }
  }
  return ref > 0.0 ? err / ref : err;
}

void GFMDSolverStaticHIP::upload_host_table(StiffnessProvider *kernel,
                                            bool normalize) {
  // fill_phi_buffer (gfmd_misc.cpp:32-102) on the host, then one upload.
  const int dsq = ndof_ * ndof_;
  std::vector<double> phi((size_t)nx_ * ny_ * dsq * 2);
  for (int i = 0; i < nx_; i++)
    for (int j = 0; j < ny_; j++)
      kernel->get_stiffness_matrix(nx_, ny_, grid_q(i, nx_), grid_q(j, ny_),
                                   &phi[((size_t)i * ny_ + j) * dsq * 2]);
  // `normalized` tells the library whether the table is final: the raw
  // table is divided by nx*ny there only when normalisation is asked for.
  check(gfmd_set_phi(h_, phi.data(), normalize ? 0 : 1), "set_kernel");
}

void GFMDSolverStaticHIP::set_kernel(StiffnessProvider *kernel,
                                     bool normalize) {
  if (kernel->get_dimension() != ndof_)
    throw std::runtime_error("GFMDSolverStaticHIP::set_kernel: kernel "
                             "dimension does not match the grid");
  source_ = "host";
  note_ = "no kernel arguments recorded";
  if (!kname_.empty()) {
    if (gfmd_set_kernel(h_, kname_.c_str(), kargs_.c_str(),
                        normalize ? 1 : 0) != GFMD_OK) {
      note_ = std::string("device tabulation unavailable: ") +
              gfmd_last_error();
    } else {
      const double tol = precision_ == GFMD_FP32 ? 1e-5 : 1e-9;
      kernel->pre_compute();
      const double dev = verify_device_table(kernel, normalize);
      kernel->post_compute();
      char buf[160];
      snprintf(buf, sizeof(buf), "device table vs kernel at 24 "
               "wavevectors: %.3g relative (tolerance %.0e)", dev, tol);
      note_ = buf;
      if (dev <= tol) source_ = "device";
    }
  }
  if (source_ == "host") {
    kernel->pre_compute();
    upload_host_table(kernel, normalize);
    kernel->post_compute();
  }
  std::vector<double> linf(ndof_ / 3);
  kernel->get_force_at_gamma_point(linf.data());
  check(gfmd_set_linf(h_, linf.data()), "set_kernel");
}

double GFMDSolverStaticHIP::post_force(void *input_buffer,
                                       void *output_buffer, char *) {
  double **u = static_cast<double **>(input_buffer);
  double **f = static_cast<double **>(output_buffer);
  double epot = 0.0;
  check(gfmd_post_force(h_, u[0], f[0], &epot), "post_force");
  check(gfmd_get_u0(h_, u0_.data()), "post_force");
  return epot;
}

void GFMDSolverStaticHIP::prec_gradient(double *cavg, double **g,
                                        double **gP) {
  check(gfmd_prec_gradient(h_, cavg, g[0], gP[0]), "prec_gradient");
}

double GFMDSolverStaticHIP::memory_usage() {
  return h_ ? (double)gfmd_memory_usage(h_) : 0.0;
}

void GFMDSolverStaticHIP::dump_stiffness(std::vector<double> &packed) {
  packed.assign((size_t)(ny_ / 2 + 1) * ndof_ * ndof_ * nx_, 0.0);
  check(gfmd_dump_stiffness(h_, packed.data()), "dump_stiffness");
}

}  // namespace gfmd_host
