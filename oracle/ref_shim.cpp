/*
 * Test infrastructure only.  A tiny extern "C" driver around the reference's
 * own LAMMPS-free stiffness-kernel core (StiffnessKernel + fill_phi_buffer),
 * compiled from the reference sources in place by oracle/Makefile.ref into
 * oracle/_ref/libgfmd_ref.so.  It is used by the tests and the fixture
 * generator to pin the oracle's Phi(q) restatement; nothing in the product
 * links it.
 *
 * Construction follows the reference's python wrapper pattern
 * (src/python/py_surface_stiffness.cpp:62-104): a stub Error/Memory/Domain/
 * Force, a whitespace-split argument vector and stiffness_kernel_factory
 * (src/main/surface_stiffness.cpp:440-481).
 */
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "domain.h"
#include "error.h"
#include "force.h"
#include "memory.h"
#include "gfmd_misc.h"
#include "surface_stiffness.h"
#include "li_berger.h"

using namespace LAMMPS_NS;

namespace {
struct RefKernel {
  Error *error;
  Memory *memory;
  Domain *domain;
  Force *force;
  StiffnessKernel *kernel;
  char *argbuf;
};
}  // namespace

extern "C" {

void *ref_kernel_create(const char *name, const char *args, double sx,
                        double sy) {
  RefKernel *k = new RefKernel;
  k->error = new Error();
  k->memory = new Memory(k->error);
  k->domain = new Domain();
  k->domain->set_cell(sx, sy, 1.0);
  k->force = new Force();
  k->argbuf = strdup(args);

  std::vector<char *> argv;
  char *p = k->argbuf;
  while (*p) {
    while (*p && isspace((unsigned char)*p)) *p++ = '\0';
    if (!*p) break;
    argv.push_back(p);
    while (*p && !isspace((unsigned char)*p)) p++;
  }
  int carg = 0;
  char kname[128];
  strncpy(kname, name, sizeof(kname) - 1);
  kname[sizeof(kname) - 1] = '\0';
  k->kernel = stiffness_kernel_factory(kname, (int)argv.size(), &carg,
                                       argv.data(), k->domain, k->force,
                                       k->memory, k->error);
  if (!k->kernel) {
    free(k->argbuf);
    delete k;
    return nullptr;
  }
  return k;
}

int ref_kernel_dim(void *h) {
  return static_cast<RefKernel *>(h)->kernel->get_dimension();
}

int ref_kernel_height(void *h) {
  return static_cast<RefKernel *>(h)->kernel->get_height();
}

int ref_kernel_nu(void *h) {
  return static_cast<RefKernel *>(h)->kernel->get_number_of_atoms();
}

/* Phi(q) at n arbitrary wavevectors; out = n * dim * dim complex128,
   row-major per matrix (MEL(dim, m, i, j) = m[i*dim+j]). */
void ref_kernel_phi(void *h, int n, const double *qx, const double *qy,
                    double *out) {
  StiffnessKernel *k = static_cast<RefKernel *>(h)->kernel;
  int dim = k->get_dimension();
  k->pre_compute();
  for (int i = 0; i < n; i++) {
    k->get_stiffness_matrix(qx[i], qy[i],
                            reinterpret_cast<double_complex *>(out) +
                                (size_t)i * dim * dim);
  }
  k->post_compute();
}

/* The reference's full-grid table (gfmd_misc.cpp:32-102) for a single
   rank owning the whole grid; out = nx*ny*ndof*ndof complex128 with
   idq = ix*ny + iy. */
void ref_fill_phi_buffer(void *h, int ndof, int nx, int ny, int normalize,
                         double *out) {
  RefKernel *r = static_cast<RefKernel *>(h);
  std::vector<double_complex *> rows((size_t)nx * ny);
  double_complex *base = reinterpret_cast<double_complex *>(out);
  for (size_t i = 0; i < rows.size(); i++) rows[i] = base + i * ndof * ndof;
  fill_phi_buffer(ndof, nx, 0, nx - 1, ny, 0, ny - 1, r->kernel, rows.data(),
                  normalize != 0, r->error);
}

void ref_kernel_linf(void *h, double *out) {
  static_cast<RefKernel *>(h)->kernel->get_force_at_gamma_point(out);
}

void ref_kernel_destroy(void *h) {
  RefKernel *k = static_cast<RefKernel *>(h);
  delete k->kernel;
  delete k->force;
  delete k->domain;
  delete k->memory;
  delete k->error;
  free(k->argbuf);
  delete k;
}


// LiBerger::sq_uniform_PQ (li_berger.cpp:453-459): displacement at (Bx, By)
// from uniform tractions (P, Qx, Qy) on the rectangle [-a, a] x [-b, b].
void ref_sq_uniform_pq(double G, double nu, double P, double Qx, double Qy,
                       double a, double b, double Bx, double By, double *u) {
  double ux = 0.0, uy = 0.0, uz = 0.0;
  LiBerger::sq_uniform_PQ(G, nu, P, Qx, Qy, a, b, Bx, By, ux, uy, uz);
  u[0] = ux;
  u[1] = uy;
  u[2] = uz;
}

}  // extern "C"


// ---------------------------------------------------------------------------
// The reference's crystal surface and pair-potential force constants, built
// as FTStiffnessKernel's constructor builds them (ft_stiffness.cpp:47-100):
// crystal_surface_factory on "<surface> <lattice constant> <nu>", then
// FCPairPotential on "pair-potential ..." (which computes the neighbour
// list).  FTStiffnessKernel itself cannot be constructed from these sources
// (force_constants_factory knows no models under -DPYTHON), so the pieces
// are exported one by one: cell, positions, neighbour list with shells,
// cell distances, off-site and on-site force-constant blocks.
// ---------------------------------------------------------------------------
#include "crystal_surface.h"
#include "fc_pair_potential.h"

namespace {
struct RefFT {
  Error *error;
  CrystalSurface *surf;
  ForceConstants *fc;
  char *argbuf;
};
}  // namespace

extern "C" {

void *ref_ft_create(const char *args) {
  RefFT *k = new RefFT;
  k->error = new Error();
  k->argbuf = strdup(args);
  std::vector<char *> argv;
  char *p = k->argbuf;
  while (*p) {
    while (*p && isspace((unsigned char)*p)) *p++ = '\0';
    if (!*p) break;
    argv.push_back(p);
    while (*p && !isspace((unsigned char)*p)) p++;
  }
  int narg = (int)argv.size(), carg = 0;
  k->surf = nullptr;
  k->fc = nullptr;
  if (narg < 4) return nullptr;
  carg = 1;
  k->surf = crystal_surface_factory(argv[0], narg, &carg, argv.data(),
                                    k->error);
  if (!k->surf || carg >= narg || strcmp(argv[carg], "pair-potential"))
    return nullptr;
  carg++;
  k->fc = new FCPairPotential(narg, &carg, argv.data(), k->surf, nullptr,
                              k->error);
  k->fc->pre_compute(k->surf, k->error);
  return k;
}

int ref_ft_nat(void *h) {
  return static_cast<RefFT *>(h)->surf->get_number_of_atoms();
}

int ref_ft_nneigh(void *h) {
  return static_cast<RefFT *>(h)->surf->get_number_of_neighbors();
}

void ref_ft_cell(void *h, double *cell) {
  const double *c = static_cast<RefFT *>(h)->surf->get_cell();
  for (int i = 0; i < 9; i++) cell[i] = c[i];
}

void ref_ft_positions(void *h, double *pos) {
  CrystalSurface *s = static_cast<RefFT *>(h)->surf;
  for (int i = 0; i < 3 * s->get_number_of_atoms(); i++)
    pos[i] = s->get_positions()[i];
}

/* Per neighbour: ints (indexi, indexj, x, y, ab, shell), doubles (r[3],
   rnorm, cell distance[3]), off-site blocks for ab = 0 and 1 (9 each). */
void ref_ft_neighbors(void *h, int *ints, double *dbl, double *D0,
                      double *D1) {
  RefFT *k = static_cast<RefFT *>(h);
  int nn = k->surf->get_number_of_neighbors();
  const CrystalSurface::lattice_neighbor_t *n = k->surf->get_neighbors();
  for (int m = 0; m < nn; m++, n++) {
    int *I = ints + 6 * m;
    I[0] = n->indexi;
    I[1] = n->indexj;
    I[2] = n->x;
    I[3] = n->y;
    I[4] = n->ab;
    I[5] = n->neighbor_shell;
    double *d = dbl + 7 * m;
    for (int c = 0; c < 3; c++) d[c] = n->r[c];
    d[3] = n->rnorm;
    k->surf->cell_distance(n->indexi, *n, d + 4);
    k->fc->offsite(k->surf, n, 0, D0 + 9 * m, k->error);
    k->fc->offsite(k->surf, n, 1, D1 + 9 * m, k->error);
  }
}

/* On-site blocks (force_constants.cpp:79-97) for layer ab, nat x 9. */
void ref_ft_onsite(void *h, int ab, double *D) {
  RefFT *k = static_cast<RefFT *>(h);
  for (int i = 0; i < k->surf->get_number_of_atoms(); i++)
    k->fc->onsite(k->surf, i, ab, D + 9 * i, k->error);
}

void ref_ft_destroy(void *h) {
  RefFT *k = static_cast<RefFT *>(h);
  delete k->fc;
  delete k->surf;
  delete k->error;
  free(k->argbuf);
  delete k;
}

}  // extern "C"


// ---------------------------------------------------------------------------
// The reference's Table2D (the periodic bicubic spline behind fix wall/map's
// map files), constructed as FixWallMap does (interp = true; lowmem as
// given) and evaluated at n points with its eight-output eval:
// out[6 i ..] = f, df/dx, df/dy, d2f/dx2, d2f/dy2, d2f/dxdy.
// ---------------------------------------------------------------------------
#include "table2d.h"

extern "C" {

void ref_table2d_eval(int nx, int ny, const double *values, int interp,
                      int lowmem, int n, const double *x, const double *y,
                      double *out) {
  Error err;
  Memory mem(&err);
  double **v = nullptr;
  mem.create(v, nx, ny, "ref_table2d_eval:v");
  for (int i = 0; i < nx * ny; i++) v[0][i] = values[i];
  Table2D t(nx, ny, v, interp != 0, lowmem != 0, &err, &mem);
  for (int i = 0; i < n; i++) {
    double *o = out + 6 * i;
    t.eval(x[i], y[i], o[0], o[1], o[2], o[3], o[4], o[5]);
  }
  mem.destroy(v);
}

}  // extern "C"
