// Host-side C++ adapter over the libgfmd_hip C ABI with the reference's
// solver interface (src/main/gfmd_solver.h:34-123,
// src/solvers/gfmd_solver_static.h:32-60): the same method names, argument
// meaning and call order, for a single rank owning the whole grid.
//
// Inside LAMMPS the class derives from LAMMPS_NS::GFMDSolver and reports
// errors through error->one; outside LAMMPS (the tests, other hosts) it
// stands alone and throws std::runtime_error with the library's message.
// Buffers are the fix's double** u_xy / f_xy ([ndof][nxy], one contiguous
// block, fix_gfmd.cpp:613-616).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gfmd_hip.h"

namespace gfmd_host {

// Provider of Phi(q) for kernels the device cannot tabulate itself: the
// reference's StiffnessKernel::get_stiffness_matrix(nx, ny, qx, qy, phi)
// (src/main/surface_stiffness.h:106-117) behind a plain interface.
class StiffnessProvider {
 public:
  virtual ~StiffnessProvider() {}
  virtual int get_dimension() const = 0;
  // phi: dim*dim complex numbers as interleaved (re, im), row-major.
  virtual void get_stiffness_matrix(int nx, int ny, double qx, double qy,
                                    double *phi) = 0;
  virtual void get_force_at_gamma_point(double *linf) const {
    for (int i = 0; i < get_dimension() / 3; i++) linf[i] = 0.0;
  }
  // called around a run of get_stiffness_matrix calls, as fill_phi_buffer
  // does (gfmd_misc.cpp:38, 100)
  virtual void pre_compute() {}
  virtual void post_compute() {}
};

class GFMDSolverStaticHIP {
 public:
  // "fix gfmd ... solver static/hip [fp32]"; iarg advances past the options
  // this solver consumes (the reference's GFMDSolverStatic consumes "dump").
  GFMDSolverStaticHIP(int narg = 0, int *iarg = nullptr,
                      char **arg = nullptr);
  virtual ~GFMDSolverStaticHIP();

  virtual void init() {}
  virtual void set_grid_size(int nx, int ny, int ndof);
  // Kernel given by name and arguments in `fix gfmd` syntax (tabulated on
  // the device when supported, e.g. sc100) ...
  virtual void set_kernel(const char *name, const char *args,
                          bool normalize = true);
  // ... or by the kernel object the fix owns (any StiffnessKernel).  With
  // the kernel's name and argument tokens recorded (set_kernel_args), the
  // table is tabulated on the device and accepted only if it agrees with
  // the kernel's own get_stiffness_matrix at 24 wavevectors (Gamma, zone
  // edges and corners, seeded random points); otherwise, or without
  // recorded tokens, the kernel's table is built on the host
  // (fill_phi_buffer) and uploaded.  The Gamma-point forces always come
  // from the kernel object (gfmd_solver_static.cpp:107).
  virtual void set_kernel(StiffnessProvider *kernel, bool normalize = true);
  // The kernel's name and the argument tokens that FixGFMD's
  // stiffness_kernel_factory consumed (fix_gfmd.cpp:195-197), joined with
  // blanks; continuum kernels (isotropic, nonperiodic) also take
  // "xprd L yprd L" from the fix's domain.  The fix records them around
  // the factory call.
  void set_kernel_args(const char *name, const char *args);
  // Where the last set_kernel(StiffnessProvider*) table came from:
  // "device" (tabulated and verified) or "host" (uploaded), and why.
  const char *kernel_source() const { return source_.c_str(); }
  const char *kernel_note() const { return note_.c_str(); }
  virtual void pre_force(void *, void *) {}
  // u_xy, f_xy: double** [ndof][nxy]; returns the elastic energy.
  virtual double post_force(void *input_buffer, void *output_buffer,
                            char *dump_prefix);

  int get_xlo_loc() const { return 0; }
  int get_xhi_loc() const { return nx_ - 1; }
  int get_ylo_loc() const { return 0; }
  int get_yhi_loc() const { return ny_ - 1; }
  int get_nxy_loc() const { return nx_ * ny_; }
  double *get_u0() { return u0_.data(); }
  const char *get_name() const { return name_.c_str(); }
  // g, gP: double** [ndof][nxy]; cavg: ndof x ndof (GFMDSolver::
  // prec_gradient, gfmd_solver_static.cpp:253-269).
  virtual void prec_gradient(double *cavg, double **g, double **gP);
  virtual double memory_usage();
  virtual void dump_stiffness(std::vector<double> &packed);

  gfmd_solver_t *handle() { return h_; }

 private:
  void check(int rc, const char *what) const;
  // max |stored - expected| / max |expected| over the sampled wavevectors
  double verify_device_table(StiffnessProvider *kernel, bool normalize);
  void upload_host_table(StiffnessProvider *kernel, bool normalize);

  std::string name_;
  int precision_;
  int nx_, ny_, ndof_;
  gfmd_solver_t *h_;
  std::vector<double> u0_;
  std::string kname_, kargs_, source_, note_;
};

}  // namespace gfmd_host
