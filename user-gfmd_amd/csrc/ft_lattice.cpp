// Host side of the `ft` stiffness kernel: surface cell geometry, neighbour
// list, force constants and the packed tables of ft_lattice.h.
//
// The supercell wraps scaled positions into [0, 1) with floor(); an atom
// whose exact scaled coordinate is an integer lands on either side of the
// cell depending on rounding, which moves it by a lattice vector and
// changes the phase convention of Phi(q).  The cell inverse (Gauss-Jordan
// with full pivoting, src/mathutils/linearalgebra.cpp:275-350) and the
// matrix-vector products (linearalgebra.h:197-207, mat.h:237-246) keep the
// reference's operation order, and contraction into fused multiply-adds is
// off in this file, so the positions come out bit for bit as the reference
// computes them.
#pragma clang fp contract(off)

#include "ft_lattice.h"

#include <algorithm>
#include <array>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace gfmd {

// Gauss-Jordan elimination with full pivoting: the pivot is the largest
// |a| among unused rows and columns (the last one on ties), it is swapped
// onto the diagonal, its row scaled by 1/pivot and its column eliminated
// from every other row; the column interchanges are undone in reverse.
bool gauss_jordan_inverse(int n, double *a) {
  std::vector<int> used(n, 0), rows(n), cols(n);
  for (int step = 0; step < n; step++) {
    double big = 0.0;
    int pr = 0, pc = 0;
    for (int r = 0; r < n; r++) {
      if (used[r] == 1) continue;
      for (int c = 0; c < n; c++) {
        if (used[c] == 0) {
          if (std::fabs(a[n * r + c]) >= big) {
            big = std::fabs(a[n * r + c]);
            pr = r;
            pc = c;
          }
        } else if (used[c] > 1) {
          return false;
        }
      }
    }
    used[pc]++;
    if (pr != pc)
      for (int c = 0; c < n; c++) std::swap(a[n * pr + c], a[n * pc + c]);
    rows[step] = pr;
    cols[step] = pc;
    if (a[n * pc + pc] == 0.0) return false;
    const double inv = 1.0 / a[n * pc + pc];
    a[n * pc + pc] = 1.0;
    for (int c = 0; c < n; c++) a[n * pc + c] *= inv;
    for (int r = 0; r < n; r++) {
      if (r == pc) continue;
      const double f = a[n * r + pc];
      a[n * r + pc] = 0.0;
      for (int c = 0; c < n; c++) a[n * r + c] -= a[n * pc + c] * f;
    }
  }
  for (int step = n - 1; step >= 0; step--) {
    const int r = rows[step], c = cols[step];
    if (r != c)
      for (int k = 0; k < n; k++) std::swap(a[n * k + r], a[n * k + c]);
  }
  return true;
}

namespace {

using Vec3 = std::array<double, 3>;

struct Mat3 {
  double m[9];  // row-major; lattice vectors are the columns
};

Vec3 matvec(const Mat3 &a, const Vec3 &v) {
  Vec3 out;
  for (int i = 0; i < 3; i++) {
    double s = 0.0;
    for (int j = 0; j < 3; j++) s += a.m[3 * i + j] * v[j];
    out[i] = s;
  }
  return out;
}

bool invert3(const Mat3 &in, Mat3 &out) {
  std::memcpy(out.m, in.m, sizeof out.m);
  return gauss_jordan_inverse(3, out.m);
}

struct Neighbor {
  int i, j, x, y, ab, shell;
  Vec3 r;  // r_i - r_j - cell.(x, y, ab)
  double rn;
};

struct Surface {
  Mat3 cell, inv;
  std::vector<Vec3> pos;
  std::vector<Neighbor> nb;

  bool set_cell(const Mat3 &c) {
    if (c.m[1] != 0.0 || c.m[3] != 0.0 || c.m[6] != 0.0 || c.m[7] != 0.0)
      return false;
    cell = c;
    return invert3(cell, inv);
  }
};

// Primitive cells in nearest-neighbour distances (lattice vectors as
// columns) and scaled atom positions: sc100 (sc100_surface.cpp:86-110),
// fcc100 (fcc100_surface.cpp:86-110), fcc111 (fcc111_surface.cpp:67-94).
bool primitive(const std::string &name, Mat3 &cell, std::vector<Vec3> &sp) {
  if (name == "sc100") {
    cell = {{1, 0, 0, 0, 1, 0, 0, 0, 1}};
    sp = {{0.0, 0.0, 0.0}};
  } else if (name == "fcc100") {
    const double h = 1.0 / std::sqrt(2.0);
    cell = {{1, 0, 0.5, 0, 1, 0.5, 0, 0, h}};
    sp = {{0.0, 0.0, 0.0}};
  } else if (name == "fcc111") {
    const double s3 = std::sqrt(3.0);
    cell = {{1, 0, 0.5, 0, s3, 1.0 / (2.0 * s3), 0, 0, std::sqrt(2.0 / 3.0)}};
    sp = {{0.0, 0.0, 0.0}, {0.5, 0.5, 0.0}};
  } else {
    return false;
  }
  return true;
}

// `nu` primitive cells stacked along z (CrystalSurface::supercell(1, 1, nu)
// followed by reverse_indices, so atom 0 is in the top layer), then every
// length scaled by the lattice constant (set_cell_and_scale_atoms).
std::string make_surface(const std::string &name, double lattice, int nu,
                         Surface &s) {
  Mat3 cell;
  std::vector<Vec3> sp;
  if (!primitive(name, cell, sp))
    return "Could not find crystal surface '" + name + "'.";
  if (!s.set_cell(cell)) return "singular surface cell";
  for (const Vec3 &v : sp) s.pos.push_back(matvec(cell, v));
  if (nu > 1) {
    const Mat3 old = s.cell;
    Mat3 nc = old;
    for (int r = 0; r < 3; r++) nc.m[3 * r + 2] *= nu;
    // smallest shift of the stacking vector in x and y
    nc.m[2] -= std::floor(nc.m[2] / nc.m[0]) * nc.m[0];
    nc.m[5] -= std::floor(nc.m[5] / nc.m[4]) * nc.m[4];
    if (!s.set_cell(nc)) return "singular surface cell";
    std::vector<Vec3> out;
    for (int z = 0; z < nu; z++) {
      const Vec3 shift = matvec(old, Vec3{0.0, 0.0, (double)z});
      for (const Vec3 &p : s.pos) {
        Vec3 q;
        for (int k = 0; k < 3; k++) q[k] = p[k] + shift[k];
        Vec3 f = matvec(s.inv, q);
        for (int k = 0; k < 3; k++) f[k] -= std::floor(f[k]);
        out.push_back(matvec(s.cell, f));
      }
    }
    s.pos.assign(out.rbegin(), out.rend());
  }
  for (Vec3 &p : s.pos) p = matvec(s.inv, p);
  Mat3 scaled;
  for (int k = 0; k < 9; k++) scaled.m[k] = lattice * s.cell.m[k];
  if (!s.set_cell(scaled)) return "singular surface cell";
  for (Vec3 &p : s.pos) p = matvec(s.cell, p);
  return "";
}

// All pairs closer than rcut (CrystalSurface::compute_neighbor_list,
// crystal_surface.cpp:166-221): atoms i, j and cell offsets within
// floor(rcut / |a_k|^2) + 1 cells along each lattice vector (the squared
// length, as there), then the neighbour shells: the distinct distances
// (0.01 tolerance) in ascending order (identify_neighbor_shells).
void neighbor_list(Surface &s, double rcut) {
  const double *c = s.cell.m;
  int rng[3];
  for (int k = 0; k < 3; k++) {
    const double l = c[k] * c[k] + c[3 + k] * c[3 + k] + c[6 + k] * c[6 + k];
    rng[k] = (int)std::lrint(std::floor(rcut / l)) + 1;
  }
  const int nat = (int)s.pos.size();
  s.nb.clear();
  for (int i = 0; i < nat; i++)
    for (int j = 0; j < nat; j++)
      for (int x = -rng[0]; x <= rng[0]; x++)
        for (int y = -rng[1]; y <= rng[1]; y++)
          for (int z = -rng[2]; z <= rng[2]; z++) {
            if (i == j && x == 0 && y == 0 && z == 0) continue;
            const Vec3 t = matvec(s.cell, Vec3{(double)x, (double)y, (double)z});
            Neighbor n;
            for (int k = 0; k < 3; k++)
              n.r[k] = s.pos[i][k] - s.pos[j][k] - t[k];
            double acc = 0.0;
            for (int k = 0; k < 3; k++) acc += n.r[k] * n.r[k];
            n.rn = std::sqrt(acc);
            if (n.rn < rcut) {
              n.i = i;
              n.j = j;
              n.x = x;
              n.y = y;
              n.ab = z;
              s.nb.push_back(n);
            }
          }
  std::vector<double> seen;
  for (const Neighbor &n : s.nb) {
    bool found = false;
    for (double v : seen)
      if (std::fabs(n.rn - v) < 0.01) {
        found = true;
        break;
      }
    if (!found) seen.push_back(n.rn);
  }
  std::sort(seen.begin(), seen.end());
  for (Neighbor &n : s.nb)
    for (size_t k = 0; k < seen.size(); k++)
      if (std::fabs(n.rn - seen[k]) < 0.01) {
        n.shell = (int)k;
        break;
      }
}

// D = -k e e^T - kappa (1 - e e^T) for the unit vector e along r.
void pair_block(const Vec3 &r, double k, double kappa, double *D) {
  double acc = 0.0;
  for (int a = 0; a < 3; a++) acc += r[a] * r[a];
  const double nrm = std::sqrt(acc);
  const double e[3] = {r[0] / nrm, r[1] / nrm, r[2] / nrm};
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      const double ee = e[a] * e[b];
      D[3 * a + b] = -k * ee - kappa * ((a == b ? 1.0 : 0.0) - ee);
    }
}

// strtol / strtod with the reference's checks (read_integer, read_double,
// linearalgebra.cpp:617-677): a number must start the token; trailing
// characters are ignored ("1.0" reads as the integer 1).
bool to_int(const std::string &s, int &v) {
  char *end;
  errno = 0;
  const long x = std::strtol(s.c_str(), &end, 10);
  if (end == s.c_str() || errno) return false;
  v = (int)x;
  return true;
}
bool to_double(const std::string &s, double &v) {
  char *end;
  v = std::strtod(s.c_str(), &end);
  return end != s.c_str();
}

}  // namespace

std::string ft_build(const std::vector<std::string> &tok, FTTable &t) {
  const int narg = (int)tok.size();
  int p = 0;
  if (p >= narg) return "Missing crystal surface.";
  const std::string sname = tok[p++];
  double lattice;
  int nu;
  if (p >= narg || !to_double(tok[p], lattice))
    return "ft: lattice constant expected";
  if (lattice < 0.0 || lattice > 100.0)
    return "ft: lattice constant must lie in [0, 100]";
  p++;
  if (p >= narg || !to_int(tok[p], nu))
    return "ft: number of lattice cells expected";
  if (nu < 1 || nu > 100) return "ft: number of lattice cells must lie in [1, 100]";
  p++;
  Surface s;
  std::string err = make_surface(sname, lattice, nu, s);
  if (!err.empty()) return err;
  const int nat = (int)s.pos.size();

  if (p >= narg)
    return "Missing interaction model (force constants definition).";
  const std::string model = tok[p++];
  double k[3] = {0, 0, 0}, kappa[3] = {0, 0, 0};
  double eps = 0, sigma = 0, rc = 0;
  bool lj = false;
  if (model == "pair-potential") {
    // FCPairPotential (fc_pair_potential.cpp:32-140): "<n>[x]" then n
    // spring constants (each followed by kappa with the x suffix), then
    // the keyword "linf" whose values it stores but never applies.
    if (p >= narg)
      return "Missing argument: Number number of interacting neighbors.";
    std::string nk_s = tok[p++];
    const bool x = !nk_s.empty() && nk_s.back() == 'x';
    if (x) nk_s.pop_back();
    int nk;
    if (!to_int(nk_s, nk)) return "Error maximum neighbor shell to integer.";
    if (nk < 0 || nk > 3) return "ft: pair-potential takes at most 3 shells";
    for (int i = 0; i < nk; i++) {
      if (p >= narg || !to_double(tok[p], k[i]))
        return "Error converting spring constant to number.";
      p++;
      if (x) {
        if (p >= narg || !to_double(tok[p], kappa[i]))
          return "Error converting spring constant to number.";
        p++;
      }
    }
    while (p < narg - 1 && tok[p] == "linf") {
      int nl;
      if (!to_int(tok[p + 1], nl) || nl < 0)
        return "Error maximum number of layers to integer.";
      double sum = 0.0;
      for (int i = 0; i < nl; i++) {
        double v;
        if (p + 2 + i >= narg || !to_double(tok[p + 2 + i], v))
          return "ft: missing linf values";
        sum += v;
      }
      if (std::fabs(sum) > 1e-12)
        return "Linear force contributions must sum to zero.";
      p += 2 + nl;
    }
    const double *c = s.cell.m;
    neighbor_list(s, 2.01 * std::max(c[0], c[4]));
  } else if (model == "lj/cut") {
    // FCLJCut (fc_lj_cut.cpp:29-170) takes epsilon, sigma and the cutoff
    // of the LAMMPS pair style lj/cut/gf; here they are arguments.
    if (p + 2 >= narg || !to_double(tok[p], eps) ||
        !to_double(tok[p + 1], sigma) || !to_double(tok[p + 2], rc))
      return "ft: lj/cut needs epsilon sigma cutoff";
    if (!(sigma > 0.0) || !(rc > 0.0))
      return "ft: lj/cut needs sigma > 0 and cutoff > 0";
    p += 3;
    lj = true;
    neighbor_list(s, rc);
  } else {
    return "Could not find interaction model '" + model + "'.";
  }
  int height = 0;
  while (p < narg - 1 && tok[p] == "height") {
    if (!to_int(tok[p + 1], height))
      return "Error converting height argument to number.";
    p += 2;
  }
  if (p < narg) return "ft: unexpected argument '" + tok[p] + "'";

  // 12-6 Lennard-Jones V'(r) and V''(r) inside the cutoff, zero outside
  auto d1 = [&](double r) {
    if (r >= rc) return 0.0;
    const double s6 = std::pow(sigma / r, 6);
    return 24.0 * eps * s6 * (1.0 - 2.0 * s6) / r;
  };
  auto d2 = [&](double r) {
    if (r >= rc) return 0.0;
    const double s6 = std::pow(sigma / r, 6);
    return 24.0 * eps * s6 * (26.0 * s6 - 7.0) / (r * r);
  };
  const size_t nn = s.nb.size();
  std::vector<double> D(9 * nn);
  for (size_t m = 0; m < nn; m++) {
    const Neighbor &n = s.nb[m];
    double kk, kap;
    if (lj) {
      kk = d2(n.rn);
      kap = d1(n.rn) / n.rn;
    } else {
      kk = n.shell < 3 ? k[n.shell] : 0.0;
      kap = n.shell < 3 ? kappa[n.shell] : 0.0;
    }
    pair_block(n.r, kk, kap, &D[9 * m]);
  }

  t = FTTable();
  t.nat = nat;
  t.height = height;
  std::memcpy(t.cell, s.cell.m, sizeof t.cell);
  for (const Vec3 &v : s.pos) t.pos.insert(t.pos.end(), v.begin(), v.end());
  // on-site blocks: minus the sum over the atom's neighbours at or below
  // layer ab (ForceConstants::onsite); the force-constant blocks of both
  // models do not depend on the layer argument
  t.onsite.assign(2 * 9 * nat, 0.0);
  for (int ab = 0; ab < 2; ab++)
    for (size_t m = 0; m < nn; m++)
      if (s.nb[m].ab <= ab)
        for (int e = 0; e < 9; e++)
          t.onsite[(size_t)(ab * nat + s.nb[m].i) * 9 + e] -= D[9 * m + e];
  // Gamma-point force: lj/cut -sum V'(r) e_z over neighbours at or below
  // the atom's layer (FCLJCut::linear); zero for pair-potential
  t.linf.assign(nat, 0.0);
  if (lj)
    for (const Neighbor &n : s.nb)
      if (n.ab <= 0) t.linf[n.i] -= d1(n.rn) * (n.r[2] / n.rn);
  // off-site entries with the cell distance of each neighbour
  // (CrystalSurface::cell_distance, crystal_surface.cpp:410-427)
  t.off.assign(nat * nat + 1, 0);
  std::vector<std::vector<double>> blk(nat * nat);
  for (size_t m = 0; m < nn; m++) {
    const Neighbor &n = s.nb[m];
    Vec3 r;
    for (int k2 = 0; k2 < 3; k2++)
      r[k2] = s.pos[n.i][k2] - s.pos[n.j][k2] - n.r[k2];
    const Vec3 cx = matvec(s.inv, r);
    const int ab = (int)std::nearbyint(cx[2]);
    if (std::fabs(ab - cx[2]) > 1e-6) return "ft: neighbour off the lattice";
    if (ab < -1 || ab > 0) continue;  // V (ab = -1), U0 and U (ab = 0)
    bool zero = true;
    for (int e = 0; e < 9; e++) zero = zero && D[9 * m + e] == 0.0;
    if (zero) continue;
    std::vector<double> &b = blk[n.i * nat + n.j];
    b.push_back(cx[0]);
    b.push_back(cx[1]);
    b.push_back((double)ab);
    for (int e = 0; e < 9; e++) b.push_back(D[9 * m + e]);
  }
  for (int q = 0; q < nat * nat; q++) {
    t.off[q + 1] = t.off[q] + (int)(blk[q].size() / 12);
    t.ent.insert(t.ent.end(), blk[q].begin(), blk[q].end());
  }
  return "";
}

}  // namespace gfmd
