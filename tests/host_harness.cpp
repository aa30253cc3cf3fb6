// Test harness: drives the host C++ adapter GFMDSolverStaticHIP the way
// LAMMPS' FixGFMD drives a GFMDSolver (fix_gfmd.cpp:404 set_grid_size,
// 527-536 set_kernel, 878-884 post_force on double** [ndof][nxy] buffers
// allocated as one contiguous block with row pointers, LAMMPS_STUB
// memory.h:139-150).
//
//   host_harness nx ny height mode u.bin f.bin
//     mode 0: kernel tabulated on the device ("sc100", "height H")
//     mode 1: kernel provided by the host (StiffnessProvider backed by the
//             CPU oracle's sc100 restatement), uploaded with gfmd_set_phi
//     mode 2: the same provider, with kernel tokens recorded that disagree
//             with it ("sc100", "height 0"): the adapter must detect the
//             mismatch and upload the provider's table
//     mode 3: the same provider with matching tokens ("height H"): the
//             device table is verified and kept
// u.bin: ndof*nx*ny doubles; f.bin receives ndof*nx*ny doubles followed by
// the energy, u0[ndof] and the preconditioned gradient of f
// (prec_gradient with cavg = diag(0.1, 0.2, 0.3) + 0.05 in xy/yx),
// ndof*nx*ny doubles.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../oracle/gfmd_oracle.h"
#include "../user-gfmd_amd/host/gfmd_solver_hip.h"

class OracleSC100 : public gfmd_host::StiffnessProvider {
 public:
  explicit OracleSC100(int height) : height_(height) {}
  int get_dimension() const override { return 3; }
  void get_stiffness_matrix(int, int, double qx, double qy,
                            double *phi) override {
    double p = height_;
    orc_kernel_phi(ORC_KERNEL_SC100, &p, qx, qy,
                   reinterpret_cast<orc_complex *>(phi));
  }

 private:
  int height_;
};

template <typename T>
static T **create2d(int n1, int n2) {
  T *data = (T *)malloc(sizeof(T) * (size_t)n1 * n2);
  T **rows = (T **)malloc(sizeof(T *) * n1);
  for (int i = 0; i < n1; i++) rows[i] = data + (size_t)i * n2;
  return rows;
}

int main(int argc, char **argv) {
  if (argc != 7) {
    fprintf(stderr, "usage: %s nx ny height mode u.bin f.bin\n", argv[0]);
    return 2;
  }
  const int nx = atoi(argv[1]), ny = atoi(argv[2]), height = atoi(argv[3]);
  const int mode = atoi(argv[4]);
  const int ndof = 3, nxy = nx * ny;
  double **u_xy = create2d<double>(ndof, nxy);
  double **f_xy = create2d<double>(ndof, nxy);
  FILE *fu = fopen(argv[5], "rb");
  if (!fu || fread(u_xy[0], sizeof(double), (size_t)ndof * nxy, fu) !=
                 (size_t)ndof * nxy) {
    fprintf(stderr, "cannot read %s\n", argv[5]);
    return 2;
  }
  fclose(fu);
  try {
    char opt0[] = "solver", opt1[] = "static/hip";
    char *arg[] = {opt0, opt1};
    int iarg = 2;  // nothing left for the solver's own options
    gfmd_host::GFMDSolverStaticHIP solver(2, &iarg, arg);
    solver.set_grid_size(nx, ny, ndof);
    if (mode == 0) {
      char args[64];
      snprintf(args, sizeof(args), "height %d", height);
      solver.set_kernel("sc100", args, true);
    } else {
      OracleSC100 k(height);
      if (mode == 2) solver.set_kernel_args("sc100", "height 0");
      if (mode == 3) {
        char args[64];
        snprintf(args, sizeof(args), "height %d", height);
        solver.set_kernel_args("sc100", args);
      }
      solver.set_kernel(&k, true);
      printf("kernel_source=%s (%s)\n", solver.kernel_source(),
             solver.kernel_note());
    }
    solver.pre_force(u_xy, f_xy);
    const double epot = solver.post_force(u_xy, f_xy, nullptr);
    double **g_xy = create2d<double>(ndof, nxy);
    double cavg[9] = {0.1, 0.05, 0.0, 0.05, 0.2, 0.0, 0.0, 0.0, 0.3};
    solver.prec_gradient(cavg, f_xy, g_xy);
    FILE *ff = fopen(argv[6], "wb");
    fwrite(f_xy[0], sizeof(double), (size_t)ndof * nxy, ff);
    fwrite(&epot, sizeof(double), 1, ff);
    fwrite(solver.get_u0(), sizeof(double), ndof, ff);
    fwrite(g_xy[0], sizeof(double), (size_t)ndof * nxy, ff);
    fclose(ff);
    printf("%s nx=%d ny=%d epot=%.17g mem=%.0f\n", solver.get_name(), nx, ny,
           epot, solver.memory_usage());
  } catch (const std::exception &e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
