// Host tables of the `ft` stiffness kernel: crystal surface cell, neighbour
// list with shells, force constants, on-site blocks and Gamma-point forces,
// packed for the device evaluator of Phi(q).
//
// Reference: FTStiffnessKernel (src/stiffness_kernels/ft_stiffness.cpp:
// 47-292), CrystalSurface (src/main/crystal_surface.cpp:166-553), the
// sc100 / fcc100 / fcc111 primitive cells (src/surfaces/*_surface.cpp),
// ForceConstants::onsite (src/main/force_constants.cpp:79-97),
// FCPairPotential (src/force_constants/fc_pair_potential.cpp:32-181) and
// FCLJCut (src/force_constants/fc_lj_cut.cpp:87-170).
#pragma once

#include <string>
#include <vector>

namespace gfmd {

struct FTTable {
  int nat = 0;     // atoms per surface cell (ndof = 3 nat)
  int height = 0;  // transfer-matrix depth ("height", default 0)
  // Off-site entries grouped by atom pair: those of (i, j) are
  // off[i*nat+j] .. off[i*nat+j+1]-1, in neighbour-list order.  Each entry
  // holds 12 doubles: cell offset cx, cy (as computed, near-integer), layer
  // ab (0 or -1), the 3 x 3 force-constant block row-major.  Neighbours
  // more than one layer away and all-zero blocks are left out (the
  // reference skips the former and adds nothing for the latter).
  std::vector<int> off;
  std::vector<double> ent;
  std::vector<double> onsite;  // [2][nat][9]: layer 0 (U0), layer 1 (U)
  std::vector<double> linf;    // [nat] Gamma-point force along z
  double cell[9] = {};         // row-major, lattice vectors as columns
  std::vector<double> pos;     // [nat][3] atom positions
};

// Builds the tables from the `ft` arguments, e.g.
//   fcc111 1.0 2 pair-potential 1 1.0 height 128
//   sc100 1.3 1 pair-potential 2x 1.0 0.0 1.0 0.0 height 64
//   fcc100 1.0 1 lj/cut 1.0 1.0 1.5 height 32
// Returns "" on success, otherwise the error message.
std::string ft_build(const std::vector<std::string> &tok, FTTable &t);

// In-place inverse of the row-major n x n matrix a by Gauss-Jordan
// elimination with full pivoting (the operation order of the reference's
// GaussJordan, src/mathutils/linearalgebra.cpp:275-350); false if singular.
bool gauss_jordan_inverse(int n, double *a);

}  // namespace gfmd
