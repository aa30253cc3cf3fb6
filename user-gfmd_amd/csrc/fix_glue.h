// Device glue between a `fix gfmd` atom list and the solver grid
// (fix_glue.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace gfmd {

// The atoms (device arrays; the C ABI's gfmd_atoms_t).
struct FixAtoms {
  int n;
  const int *mask;  // [n] or null: every atom is in the group
  int groupbit;
  int *gid;         // [n][3]: grid ix, iy, atom index in the cell
  const double *x, *xeq, *v;  // [n][3] (v may be null)
  const int *image;           // [n] LAMMPS image flags or null
  const double *mass;         // [n] or null: unit masses
};

// The grid and the cell (by value into the kernels).
struct FixFrame {
  int nx, ny, nl;      // grid and atoms per cell (ndof / 3)
  double box[3];       // periodic box lengths
  double invcell[6];   // rows 0 and 1 of the inverse surface cell
};

// Device state of the glue (doubles; integers stored exactly).
enum {
  FS_XEQCM = 0,       // [3] centre of mass at equilibrium
  FS_XCM = 3,         // [3] centre of mass of the last pre_force
  FS_XSHIFT = 6,      // layer shift in lattice units, x and y
  FS_YSHIFT = 7,
  FS_DXSHIFT = 8,     // change of the shift at the last pre_force
  FS_DYSHIFT = 9,
  FS_PLACED = 10,     // atoms gathered onto the grid
  FS_UNPLACED = 11,   // group atoms that could not be (bad id / image)
  FS_FSUM = 12,       // [3] sum of the grid forces given to the atoms
  FS_UNFORCED = 15,   // group atoms that received no force
  FS_MASSTOTAL = 16,
  FS_NATOMS = 17,     // group atoms
  FS_COUNT = 18
};

// part: device scratch of at least 5 * 1024 doubles.
void launch_fix_com(const FixAtoms &a, const FixFrame &fr, bool setup,
                    double *part, double *st, hipStream_t s);
void launch_fix_gather(const FixAtoms &a, const FixFrame &fr, bool fp64,
                       void *u, void *vg, double *part, double *st,
                       hipStream_t s);
void launch_fix_scatter(const FixAtoms &a, const FixFrame &fr, bool fp64,
                        const void *fgrid, double *fat, const int center[3],
                        double *part, double *st, hipStream_t s);

}  // namespace gfmd
