// Host-side C++ adapter: see gfmd_solver_hip.h.
#include "gfmd_solver_hip.h"

#include <cmath>
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

void GFMDSolverStaticHIP::set_kernel(StiffnessProvider *kernel,
                                     bool normalize) {
  // fill_phi_buffer (gfmd_misc.cpp:32-102) on the host, then one upload.
  if (kernel->get_dimension() != ndof_)
    throw std::runtime_error("GFMDSolverStaticHIP::set_kernel: kernel "
                             "dimension does not match the grid");
  const int dsq = ndof_ * ndof_;
  std::vector<double> phi((size_t)nx_ * ny_ * dsq * 2);
  for (int i = 0; i < nx_; i++) {
    const double qx = (i <= nx_ / 2) ? 2.0 * M_PI * i / nx_
                                     : 2.0 * M_PI * (i - nx_) / nx_;
    for (int j = 0; j < ny_; j++) {
      const double qy = (j <= ny_ / 2) ? 2.0 * M_PI * j / ny_
                                       : 2.0 * M_PI * (j - ny_) / ny_;
      kernel->get_stiffness_matrix(nx_, ny_, qx, qy,
                                   &phi[((size_t)i * ny_ + j) * dsq * 2]);
    }
  }
  // set_phi expects normalised values when `normalized` is set; pass the raw
  // table and let the library divide by nx*ny when normalisation is asked.
  if (normalize) {
    check(gfmd_set_phi(h_, phi.data(), 0), "set_kernel");
  } else {
    check(gfmd_set_phi(h_, phi.data(), 1), "set_kernel");
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
