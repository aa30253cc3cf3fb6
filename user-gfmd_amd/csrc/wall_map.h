// Soft wall over a height map (`fix wall/map`), evaluated on the device:
// the map z_map(x, y) with its slopes (flat plane, leading Fourier terms of
// a (100) or (111) surface, a periodic bicubic spline of a height table,
// or the raw table), and the wall potential V(dz) of the normal distance
// dz = normal (z - z_map(x, y)).
//
// Reference: FixWallMap (src/extras/fix_wall_map.cpp: potentials 363-686,
// pre_force 686-811, maps 829-1012 and 1363-1398) and Table2D
// (src/mathutils/table2d.h:94-138, table2d.cpp:203-262).
//
// The functions below keep the reference's operation order and are
// compiled without floating-point contraction, so that the spline and the
// power-law potentials agree bit for bit with a plain IEEE evaluation;
// the (100)/(111) maps and the exponential potentials call the device's
// cos/sin/exp, which agree with the host's to an ulp or so.
#pragma once

#include <hip/hip_runtime.h>

namespace gfmd {

enum { WM_FLAT = 0, WM_100 = 1, WM_111 = 2, WM_TABLE = 3, WM_RAW = 4 };
enum {
  WP_LJ93 = 0,
  WP_LJ93_SMOOTH = 1,
  WP_LJ126 = 2,
  WP_EXP = 3,
  WP_EXPF = 4,
  WP_DUGDALE = 5
};

struct WallMapDev {
  int map;       // WM_*
  int pot;       // WP_*
  int norm;      // surface normal, +1 or -1
  int mx, my;    // table size
  // flat / 100 / 111: z0, amplitude h0, wave numbers qx, qy, shift x0, y0;
  // table / raw: qx = mx / xprd, qy = my / yprd (table points per length)
  double z0, h0, qx, qy, x0, y0;
  const double *values;  // [mx][my] heights, shifted to z0
  // table: per tile of 16 x 16 spline boxes, a lower and an upper bound of
  // the spline's computed values ([ceil(mx/16)][ceil(my/16)][2]); nty =
  // ceil(my / 16).  Atoms whose distance bound already puts them beyond
  // the cutoff skip the spline (their force is zero either way).
  const double *tiles;
  int nty;
  // potential: cutoff (< 0: everywhere, 0: nowhere), 0.01 cutoff, inner
  // cutoff (lj93/smooth), pendist, epsilon, sigma, energy offset and
  // coefficients
  double cutoff, mcutoff, inner, pendist, eps, sigma, offset;
  double c[6], mc[3], s[5];
};

// Inverse of Table2D's 16 x 16 spline system (table2d.cpp:97-124), columns
// 0-11 (columns 12-15 multiply the cross derivatives, which are zero), as
// the reference's Gauss-Jordan elimination with full pivoting leaves it
// (entries within an ulp or two of the exact integers).  The host repeats
// the elimination at set-up and checks these values bit for bit.
constexpr double kSplineInv[16][12] = {
    {0x1.0000000000000p+0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0x1.0000000000000p+0, 0.0, 0.0, 0.0},
    {-0x1.7ffffffffffffp+1, 0.0, 0.0, 0x1.7ffffffffffffp+1, 0.0, 0.0, 0.0, 0.0, -0x1.0000000000000p+1, 0.0, 0.0, -0x1.ffffffffffffep-1},
    {0x1.ffffffffffffep+0, 0.0, 0.0, -0x1.ffffffffffffep+0, 0.0, 0.0, 0.0, 0.0, 0x1.0000000000000p+0, 0.0, 0.0, 0x1.ffffffffffffep-1},
    {0.0, 0.0, 0.0, 0.0, 0x1.0000000000000p+0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.0, 0.0, 0.0, 0.0, -0x1.7ffffffffffffp+1, 0.0, 0.0, 0x1.7ffffffffffffp+1, 0.0, 0.0, 0.0, 0.0},
    {0.0, 0.0, 0.0, 0.0, 0x1.ffffffffffffep+0, 0.0, 0.0, -0x1.ffffffffffffep+0, 0.0, 0.0, 0.0, 0.0},
    {-0x1.7ffffffffffffp+1, 0x1.7ffffffffffffp+1, 0.0, 0.0, -0x1.0000000000000p+1, -0x1.ffffffffffffep-1, -0x1.0000000000000p-53, 0.0, 0.0, 0.0, -0x1.0000000000000p-53, 0x1.0000000000000p-53},
    {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, -0x1.7ffffffffffffp+1, 0x1.7ffffffffffffp+1, 0.0, 0.0},
    {0x1.2000000000002p+3, -0x1.2000000000002p+3, 0x1.2000000000002p+3, -0x1.2000000000002p+3, 0x1.7ffffffffffffp+2, 0x1.8000000000003p+1, -0x1.8000000000003p+1, -0x1.7ffffffffffffp+2, 0x1.7ffffffffffffp+2, -0x1.7ffffffffffffp+2, -0x1.8000000000003p+1, 0x1.8000000000003p+1},
    {-0x1.8000000000003p+2, 0x1.8000000000003p+2, -0x1.8000000000003p+2, 0x1.8000000000003p+2, -0x1.fffffffffffffp+1, -0x1.0000000000003p+1, 0x1.0000000000003p+1, 0x1.fffffffffffffp+1, -0x1.8000000000000p+1, 0x1.8000000000000p+1, 0x1.8000000000003p+1, -0x1.8000000000003p+1},
    {0x1.fffffffffffffp+0, -0x1.fffffffffffffp+0, 0x1.0000000000000p-52, -0x1.0000000000000p-52, 0x1.0000000000000p+0, 0x1.ffffffffffffep-1, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
    {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0x1.ffffffffffffep+0, -0x1.ffffffffffffep+0, 0.0, 0.0},
    {-0x1.8000000000003p+2, 0x1.8000000000003p+2, -0x1.8000000000003p+2, 0x1.8000000000003p+2, -0x1.8000000000000p+1, -0x1.8000000000003p+1, 0x1.8000000000003p+1, 0x1.8000000000000p+1, -0x1.fffffffffffffp+1, 0x1.fffffffffffffp+1, 0x1.0000000000003p+1, -0x1.0000000000003p+1},
    {0x1.0000000000003p+2, -0x1.0000000000003p+2, 0x1.0000000000003p+2, -0x1.0000000000003p+2, 0x1.0000000000000p+1, 0x1.0000000000002p+1, -0x1.0000000000002p+1, -0x1.0000000000000p+1, 0x1.0000000000000p+1, -0x1.0000000000000p+1, -0x1.0000000000002p+1, 0x1.0000000000002p+1}};

constexpr int kTile = 16;  // spline boxes per bound tile along x and y

// Box (bx, by) and normalised coordinates (dx, dy) of (x, y) on the table
// (eval_table2d: (x qx, y qy), floor, periodic wrap).
__device__ __forceinline__ void wm_box(const WallMapDev &w, double x,
                                       double y, int &bx, int &by,
                                       double &dx, double &dy) {
#pragma clang fp contract(off)
  const double X = x * w.qx, Y = y * w.qy;
  const double fx0 = floor(X), fy0 = floor(Y);
  dx = X - fx0;
  dy = Y - fy0;
  bx = (int)fx0 % w.mx;
  by = (int)fy0 % w.my;
  if (bx < 0) bx += w.mx;
  if (by < 0) by += w.my;
}

// Spline coefficients c[4 i + j] of box (bx, by): corner values (corners
// (0,0), (1,0), (1,1), (0,1)) and central-difference slopes as the
// right-hand side, A^-1 applied row by row, each sum from zero in column
// order (terms with a zero matrix entry add nothing and are left out).
__device__ __forceinline__ void wm_coeffs(const WallMapDev &w, int bx, int by,
                                          double (&cf)[16]) {
#pragma clang fp contract(off)
  const int bx1 = bx + 1 == w.mx ? 0 : bx + 1;
  const int by1 = by + 1 == w.my ? 0 : by + 1;
  const int xs[4] = {bx, bx1, bx1, bx};
  const int ys[4] = {by, by, by1, by1};
  double B[12];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int i = xs[r], j = ys[r];
    const int ip = i + 1 == w.mx ? 0 : i + 1, im = i == 0 ? w.mx - 1 : i - 1;
    const int jp = j + 1 == w.my ? 0 : j + 1, jm = j == 0 ? w.my - 1 : j - 1;
    const double *v = w.values;
    B[r] = v[(size_t)i * w.my + j];
    B[r + 4] = (v[(size_t)ip * w.my + j] - v[(size_t)im * w.my + j]) / 2;
    B[r + 8] = (v[(size_t)i * w.my + jp] - v[(size_t)i * w.my + jm]) / 2;
  }
#pragma unroll
  for (int r = 0; r < 16; r++) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < 12; k++)
      if (kSplineInv[r][k] != 0.0) t = t + kSplineInv[r][k] * B[k];
    cf[r] = t;
  }
}

// Value and slopes (per unit of the table grid) of the spline at (dx, dy)
// of its box: the Horner scheme of Table2D::eval.
__device__ __forceinline__ void wm_spline(const double (&cf)[16], double dx,
                                          double dy, double &f, double &fx,
                                          double &fy) {
#pragma clang fp contract(off)
  f = 0.0;
  fx = 0.0;
  fy = 0.0;
#pragma unroll
  for (int i = 3; i >= 0; i--) {
    double sf = 0.0, sfy = 0.0;
#pragma unroll
    for (int j = 3; j >= 0; j--) {
      const double cij = cf[4 * i + j];
      sf = sf * dy + cij;
      if (j > 0) sfy = sfy * dy + j * cij;
    }
    f = f * dx + sf;
    if (i > 0) fx = fx * dx + i * sf;
    fy = fy * dx + sfy;
  }
}

// z_map, dz_map/dx, dz_map/dy at (x, y).
__device__ __forceinline__ void wm_height(const WallMapDev &w, double x,
                                          double y, double &z, double &zx,
                                          double &zy) {
#pragma clang fp contract(off)
  if (w.map == WM_FLAT) {
    z = w.z0;
    zx = zy = 0.0;
  } else if (w.map == WM_100) {
    const double dx = x - w.x0, dy = y - w.y0;
    const double cx = cos(w.qx * dx), cy = cos(w.qy * dy);
    const double sx = sin(w.qx * dx), sy = sin(w.qy * dy);
    z = w.z0 + w.h0 * (cx + cy);
    zx = -w.h0 * w.qx * sx;
    zy = -w.h0 * w.qy * sy;
  } else if (w.map == WM_111) {
    const double dx = x - w.x0, dy = y - w.y0;
    const double a1 = w.qy * dy;
    const double a2 = w.qx * dx + 0.5 * w.qy * dy;
    const double a3 = w.qx * dx - 0.5 * w.qy * dy;
    const double c1 = cos(a1), c2 = cos(a2), c3 = cos(a3);
    const double s1 = sin(a1), s2 = sin(a2), s3 = sin(a3);
    z = w.z0 + w.h0 * (c1 + c2 + c3);
    zx = -w.h0 * w.qx * (s2 + s3);
    zy = -w.h0 * w.qy * (s1 + 0.5 * s2 - 0.5 * s3);
  } else if (w.map == WM_RAW) {
    // eval_raw_data: truncation toward zero, then periodic wrap
    int ix = (int)(x * w.qx) % w.mx, iy = (int)(y * w.qy) % w.my;
    if (ix < 0) ix += w.mx;
    if (iy < 0) iy += w.my;
    z = w.values[(size_t)ix * w.my + iy];
    zx = zy = 0.0;
  } else {
    int bx, by;
    double dx, dy, cf[16];
    wm_box(w, x, y, bx, by, dx, dy);
    wm_coeffs(w, bx, by, cf);
    wm_spline(cf, dx, dy, z, zx, zy);
    zx = zx * w.qx;
    zy = zy * w.qy;
  }
}

// A lower bound of the computed distance dz of an atom at (x, y, z) from a
// table map: the tile's spline bound for the atom's box (monotone rounding
// keeps fl(lo - z) <= fl(z_map - z) for lo <= z_map).  -inf where no bound
// applies.
__device__ __forceinline__ double wm_dz_bound(const WallMapDev &w, double x,
                                              double y, double z) {
#pragma clang fp contract(off)
  if (w.map != WM_TABLE || !w.tiles) return -INFINITY;
  int bx, by;
  double dx, dy;
  wm_box(w, x, y, bx, by, dx, dy);
  const double *t =
      w.tiles + 2 * ((size_t)(bx / kTile) * w.nty + (size_t)(by / kTile));
  return w.norm < 0 ? t[0] - z : z - t[1];
}

// V(dz) and f = -dV/ddz; close += 1 where lj126 is asked for below
// 0.01 cutoff (the reference stops with "Atom too close").
__device__ __forceinline__ void wm_potential(const WallMapDev &w, double dz,
                                             double &e, double &f,
                                             double &close) {
#pragma clang fp contract(off)
  if (w.pot == WP_LJ93 || w.pot == WP_LJ93_SMOOTH) {
    if (dz < w.mcutoff) {
      const double ddz = dz - w.mcutoff;
      e = w.mc[0] + w.mc[1] * ddz + 0.5 * w.mc[2] * ddz * ddz;
      f = -w.mc[1] - w.mc[2] * ddz;
    } else if (w.pot == WP_LJ93 || dz < w.inner) {
      const double rinv = 1.0 / dz;
      const double r2inv = rinv * rinv;
      const double r4inv = r2inv * r2inv;
      const double r10inv = r4inv * r4inv * r2inv;
      e = w.c[2] * r4inv * r4inv * rinv - w.c[3] * r2inv * rinv - w.offset;
      f = w.c[0] * r10inv - w.c[1] * r4inv;
    } else {
      const double r = dz - w.inner;
      const double r2 = r * r;
      e = w.s[0] - w.s[1] * r - w.s[2] * r2 / 2.0 - w.s[3] * r2 * r / 3.0 -
          w.s[4] * r2 * r2 / 4.0;
      f = w.s[1] + w.s[2] * r + w.s[3] * r2 + w.s[4] * r2 * r;
    }
  } else if (w.pot == WP_LJ126) {
    if (dz < w.mcutoff) close += 1.0;
    const double rinv = 1.0 / dz;
    const double r2inv = rinv * rinv;
    const double r6inv = r2inv * r2inv * r2inv;
    e = r6inv * (w.c[2] * r6inv - w.c[3]) - w.offset;
    f = r6inv * (w.c[0] * r6inv - w.c[1]) * rinv;
  } else if (w.pot == WP_EXP) {
    e = w.eps * exp(-dz * w.c[0]);
    f = e * w.c[0];
    e -= w.offset;
  } else if (w.pot == WP_EXPF) {
    const double t = w.eps * exp(w.c[0] * (w.cutoff - dz));
    f = t - w.eps;
    e = w.sigma * t + w.eps * dz - w.offset;
  } else {  // dugdale
    if (dz > 0.0) {
      e = w.c[0] * (w.cutoff - dz);
      f = w.c[0];
    } else {
      e = 1e12;
      f = 0.0;
    }
  }
}

// One atom of FixWallMap::pre_force at (x, y, z): returns false outside
// the cutoff; otherwise the force (fx, fy, fz), the energy e and the
// potential's own force f (its sign classifies the atom as repulsive or
// attractive).  z is replaced by the pendist position where the atom is
// closer than pendist.
__device__ __forceinline__ bool wm_atom(const WallMapDev &w, double x,
                                        double y, double &z, double &fx,
                                        double &fy, double &fz, double &e,
                                        double &f, double &close) {
#pragma clang fp contract(off)
  if (w.cutoff > 0.0 && wm_dz_bound(w, x, y, z) >= w.cutoff) return false;
  double zm, zx, zy;
  wm_height(w, x, y, zm, zx, zy);
  double dz = w.norm * (z - zm);
  if (!(w.cutoff < 0.0 || (w.cutoff > 0.0 && dz < w.cutoff))) return false;
  if (dz < w.pendist) {
    z = zm + w.norm * w.pendist;
    dz = w.pendist;
  }
  wm_potential(w, dz, e, f, close);
  fz = f * w.norm;
  fx = -fz * zx;
  fy = -fz * zy;
  return true;
}

// Where the pendist rule moves an atom at (x, y, z): the new z (z itself
// when it does not apply).
__device__ __forceinline__ double wm_reset_z(const WallMapDev &w, double x,
                                             double y, double z) {
#pragma clang fp contract(off)
  const double lb = wm_dz_bound(w, x, y, z);
  if (lb >= w.pendist || (w.cutoff > 0.0 && lb >= w.cutoff)) return z;
  double zm, zx, zy;
  wm_height(w, x, y, zm, zx, zy);
  const double dz = w.norm * (z - zm);
  if (!(w.cutoff < 0.0 || (w.cutoff > 0.0 && dz < w.cutoff))) return z;
  return dz < w.pendist ? zm + w.norm * w.pendist : z;
}

}  // namespace gfmd

// ---------------------------------------------------------------------------
// Host set-up (wall_map.cpp) and launchers (wall_map.hip).
// ---------------------------------------------------------------------------
#include <string>
#include <vector>

namespace gfmd {

struct WallMapSpec {
  WallMapDev dev{};            // device pointers left unset
  std::vector<double> values;  // table heights as given or read (mx * my)
  double ainv[16 * 12];        // spline inverse, columns 0-11
  double lambda = -1.0;        // > 0: low-pass filter before displacement
  int z0_type = 0;             // 0: average at z0, 1: bottom, 2: top
};

// Parses the arguments of `fix wall/map` after the style name, e.g.
//   flat 2.0 -1 lj93 1.0 1.0 2.5
//   100 1.0 0.2 3.0 -1 exp 1.0 0.5 4.0
//   rough.map top 3.0 -1 lj93/smooth 1.0 1.0 2.0 3.0 lambda 8 pendist -0.5
// A map name other than flat / 100 / 111 is a file in the reference's
// format unless `map` (mx * my heights, x slow) is given, which then
// replaces the file.  xprd, yprd: the periodic box lengths.  Returns "" or
// the error message.
std::string wall_map_parse(const char *args, const double *map, int mx,
                           int my, double xprd, double yprd,
                           WallMapSpec &out);

// values[mx*my] (device) shifted so that the average (z0_type 0), the
// minimum (1) or the maximum (2) sits at z0; stats (device [5]) receive
// the shifted minimum and maximum, the rms height, the shifted average and
// the shift.
void launch_map_displace(double *values, int mx, int my, double z0,
                         int z0_type, double *stats, hipStream_t st);

// Tile bounds of the spline (w.values, w.tiles set; table maps).
void launch_map_tile_bounds(const WallMapDev &w, double *tiles,
                            hipStream_t st);

// FixWallMap::pre_force on n atoms: x [n][3] (pendist resets written
// back), f [n][3] accumulated; out (device [8]): energy, force sums x, y,
// z, atoms within the cutoff, repulsive, attractive, lj126 atoms too
// close.  part: device scratch of 8 * nblk doubles.
void launch_wall_map_force(const WallMapDev &w, int n, double *x, double *f,
                           double *part, int nblk, double *out,
                           hipStream_t st);

}  // namespace gfmd

// The object behind gfmd_wall_map_t.
struct gfmd_wall_map {
  gfmd::WallMapDev dev{};  // with device pointers into buf
  double *buf = nullptr;   // device: stats [8], tile bounds, values
  double info[6] = {};     // min, max, rms height, average, mx, my
  double xprd = 0.0, yprd = 0.0;
  hipStream_t stream = nullptr;
  double *work = nullptr;  // gfmd_wall_map_force scratch
  size_t work_cap = 0;     // atoms the scratch holds
};
