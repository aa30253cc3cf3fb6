// Host set-up of the soft wall over a height map (`fix wall/map`): the
// arguments, the potential's coefficients, the map tables and the inverse
// of the bicubic spline system.
//
// Reference: FixWallMap constructor and keywords (src/extras/
// fix_wall_map.cpp:60-300), init_* of the potentials (363-650), the maps
// (829-1012), read_map (1079-1145); Table2D's system (src/mathutils/
// table2d.cpp:97-124).
#pragma clang fp contract(off)

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>

#include "ft_lattice.h"
#include "wall_map.h"

namespace gfmd {
namespace {

const char kSyntax[] =
    "fix wall/map: Illegal fix command, syntax is 'fix ID group-ID wall/map "
    "plane|<mapfile> [bot|top] <z0> <surface normal> "
    "<lj93|lj93/smooth|lj126|exp|expf|dugdale> [<epsilon> <sigma> <cutoff>] "
    "[lambda <short-wavelength cutoff>] [pendist <distance from surface after "
    "penetration>] [vel <x-velocity> <y-velocity>]'";

std::string fmt(const char *f, const std::string &a) {
  char buf[512];
  std::snprintf(buf, sizeof buf, f, a.c_str());
  return buf;
}

// strtod with the reference's check (nothing converted is an error)
bool to_double(const std::string &s, double &v) {
  char *end = nullptr;
  v = std::strtod(s.c_str(), &end);
  return end != s.c_str();
}

// the fix's round(): (int)(x < 0 ? x - 0.5 : x + 0.5)
int round_half(double x) { return (int)(x < 0 ? (x - 0.5) : (x + 0.5)); }

// read_map: the column count of the first line (a column starts where an
// alphanumeric character, '-' or 'e' follows a blank) sets a square map,
// read row by row with fscanf.
std::string read_map(const std::string &fn, std::vector<double> &v, int &n) {
  FILE *f = std::fopen(fn.c_str(), "r");
  if (!f) return fmt("Could not open contact map file '%s'.", fn);
  n = 0;
  int c = std::fgetc(f);
  int last = ' ';
  while (c != EOF && c != '\n') {
    if ((std::isalnum(c) || c == '-' || c == 'e') && last == ' ') n++;
    last = c;
    c = std::fgetc(f);
  }
  if (c == EOF) {
    std::fclose(f);
    return "fix contact/map: Could not read map file.";
  }
  if (n <= 0) {
    std::fclose(f);
    return "fix contact/map: Something wrong: nx_ <= 0";
  }
  std::rewind(f);
  v.resize((size_t)n * n);
  for (size_t i = 0; i < v.size(); i++) {
    if (std::fscanf(f, "%lf", &v[i]) != 1) {
      std::fclose(f);
      return "fix contact/map: Error read map file.";
    }
  }
  std::fclose(f);
  return "";
}

// The 16 x 16 bicubic system of Table2D (rows: values, d/dx, d/dy,
// d2/dxdy at the corners (0,0), (1,0), (1,1), (0,1); column 4 p1 + p2:
// the monomial x^p1 y^p2), inverted by Gauss-Jordan.
bool spline_inverse(double *ainv12) {
  const int ix1[4] = {0, 1, 1, 0}, ix2[4] = {0, 0, 1, 1};
  double A[16 * 16];
  for (int r = 0; r < 4; r++) {
    const int n1 = ix1[r], n2 = ix2[r];
    for (int p1 = 0; p1 <= 3; p1++)
      for (int p2 = 0; p2 <= 3; p2++) {
        const int p1m = p1 > 0 ? p1 - 1 : 0, p2m = p2 > 0 ? p2 - 1 : 0;
        const int c = 4 * p1 + p2;
        A[r * 16 + c] = 1.0 * (std::pow(n1, p1) * std::pow(n2, p2));
        A[(r + 4) * 16 + c] = 1.0 * (p1 * std::pow(n1, p1m) * std::pow(n2, p2));
        A[(r + 8) * 16 + c] = 1.0 * (std::pow(n1, p1) * p2 * std::pow(n2, p2m));
        A[(r + 12) * 16 + c] =
            1.0 * (p1 * std::pow(n1, p1m) * p2 * std::pow(n2, p2m));
      }
  }
  if (!gauss_jordan_inverse(16, A)) return false;
  for (int r = 0; r < 16; r++)
    for (int k = 0; k < 12; k++) ainv12[r * 12 + k] = A[r * 16 + k];
  return true;
}

struct Parser {
  std::vector<std::string> a;
  size_t i = 0;
  WallMapDev &w;
  explicit Parser(WallMapDev &w_) : w(w_) {}

  std::string eps_sigma() {
    if (i >= a.size()) return kSyntax;
    if (!to_double(a[i], w.eps))
      return "fix contact/map: Could not convert first potential parameter "
             "to floating point value.";
    i++;
    if (i >= a.size()) return kSyntax;
    if (!to_double(a[i], w.sigma))
      return "fix contact/map: Could not convert second potential parameter "
             "to floating point value.";
    i++;
    return "";
  }
  std::string cutoff() {
    if (i >= a.size()) return kSyntax;
    if (!to_double(a[i], w.cutoff))
      return "fix contact/map: Could not convert potential cutoff parameter "
             "to floating point value.";
    i++;
    w.mcutoff = 0.01 * w.cutoff;
    return "";
  }
  std::string lj93() {
    std::string e;
    if (!(e = eps_sigma()).empty() || !(e = cutoff()).empty()) return e;
    const double ep = w.eps, s = w.sigma;
    w.c[0] = 6.0 / 5.0 * ep * std::pow(s, 9.0);
    w.c[1] = 3.0 * ep * std::pow(s, 3.0);
    w.c[2] = 2.0 / 15.0 * ep * std::pow(s, 9.0);
    w.c[3] = ep * std::pow(s, 3.0);
    w.c[4] = 12.0 * ep * std::pow(s, 9.0);
    w.c[5] = 12.0 * ep * std::pow(s, 3.0);
    if (w.cutoff > 0.0) {
      const double rinv = 1.0 / w.cutoff;
      const double r2inv = rinv * rinv;
      const double r4inv = r2inv * r2inv;
      w.offset = w.c[2] * r4inv * r4inv * rinv - w.c[3] * r2inv * rinv;
    }
    const double m = w.mcutoff;
    w.mc[0] = w.c[2] / std::pow(m, 9) - w.c[3] / std::pow(m, 3) - w.offset;
    w.mc[1] = -w.c[0] / std::pow(m, 10) + w.c[1] / std::pow(m, 4);
    w.mc[2] = w.c[4] / std::pow(m, 11) - w.c[5] / std::pow(m, 5);
    return "";
  }
  std::string potential(const std::string &name) {
    std::string e;
    if (name == "lj93") {
      w.pot = WP_LJ93;
      return lj93();
    }
    if (name == "lj126") {
      w.pot = WP_LJ126;
      if (!(e = eps_sigma()).empty() || !(e = cutoff()).empty()) return e;
      const double ep = w.eps, s = w.sigma;
      w.c[0] = 48.0 * ep * std::pow(s, 12.0);
      w.c[1] = 24.0 * ep * std::pow(s, 6.0);
      w.c[2] = 4.0 * ep * std::pow(s, 12.0);
      w.c[3] = 4.0 * ep * std::pow(s, 6.0);
      if (w.cutoff > 0.0) {
        const double r2inv = 1.0 / (w.cutoff * w.cutoff);
        const double r6inv = r2inv * r2inv * r2inv;
        w.offset = r6inv * (w.c[2] * r6inv - w.c[3]);
      }
      return "";
    }
    if (name == "lj93/smooth") {
      w.pot = WP_LJ93_SMOOTH;
      if (!(e = lj93()).empty()) return e;
      w.inner = w.cutoff;
      if (i >= a.size()) return kSyntax;
      if (!to_double(a[i], w.cutoff))
        return "fix contact/map: Could not convert second potential cutoff "
               "parameter to floating point value.";
      i++;
      const double rinv = 1.0 / w.inner;
      const double r2inv = rinv * rinv;
      const double r4inv = r2inv * r2inv;
      const double r10inv = r4inv * r4inv * r2inv;
      // force and its first derivative continuous at the inner cutoff
      const double s1 = w.c[0] * r10inv - w.c[1] * r4inv;
      const double s2 =
          -10 * w.c[0] * r10inv * rinv + 4 * w.c[1] * r4inv * rinv;
      const double dr = w.cutoff - w.inner;
      if (dr <= 0.0)
        return "fix contact/map: lj93/smooth outer cutoff must be bigger "
               "than inner cutoff.";
      const double s3 = -(3 * s1 + 2 * s2 * dr) / (dr * dr);
      const double s4 = (2 * s1 + s2 * dr) / (dr * dr * dr);
      const double s0 = s1 * dr + s2 * dr * dr / 2.0 +
                        s3 * dr * dr * dr / 3.0 + s4 * dr * dr * dr * dr / 4.0;
      w.s[0] = s0;
      w.s[1] = s1;
      w.s[2] = s2;
      w.s[3] = s3;
      w.s[4] = s4;
      w.offset -= s0;  // energy continuous
      return "";
    }
    if (name == "exp") {
      w.pot = WP_EXP;
      if (!(e = eps_sigma()).empty() || !(e = cutoff()).empty()) return e;
      w.c[0] = 1.0 / w.sigma;
      if (w.cutoff > 0.0) w.offset = w.eps * std::exp(-w.cutoff * w.c[0]);
      return "";
    }
    if (name == "expf") {
      w.pot = WP_EXPF;
      if (!(e = eps_sigma()).empty() || !(e = cutoff()).empty()) return e;
      w.c[0] = 1.0 / w.sigma;
      if (w.cutoff > 0.0) w.offset = w.eps * (w.sigma + w.cutoff);
      return "";
    }
    if (name == "dugdale") {
      w.pot = WP_DUGDALE;
      if (!(e = eps_sigma()).empty()) return e;
      w.cutoff = w.sigma;
      w.c[0] = w.eps / w.cutoff;
      return "";
    }
    return fmt("fix contact/map: Unknown potential type '%s'.", name);
  }
};

}  // namespace

std::string wall_map_parse(const char *args, const double *map, int mx,
                           int my, double xprd, double yprd,
                           WallMapSpec &out) {
  out = WallMapSpec{};
  WallMapDev &w = out.dev;
  w.cutoff = -1.0;
  w.pendist = -1.0;
  w.norm = 1;
  Parser P(w);
  {
    std::istringstream is(args ? args : "");
    std::string t;
    while (is >> t) P.a.push_back(t);
  }
  auto &a = P.a;
  size_t &i = P.i;
  if (a.size() < 4) return kSyntax;
  if (!(xprd > 0) || !(yprd > 0)) return "fix wall/map: box lengths must be positive";
  const std::string name = a[i++];
  bool have_table = false;
  if (name == "flat") {
    w.map = WM_FLAT;
  } else if (name == "100" || name == "111") {
    double a0, h;
    if (i + 2 > a.size()) return kSyntax;
    if (!to_double(a[i], a0))
      return fmt("fix contact/map: Could not convert lattice constant (= '%s') "
                 "parameter for (111) surface to floating point value.", a[i]);
    i++;
    if (!to_double(a[i], h))
      return fmt("fix contact/map: Could not convert amplitude (= '%s') "
                 "parameter for (111) surface to floating point value.", a[i]);
    i++;
    if (name == "100") {
      w.map = WM_100;
      w.h0 = h / 2;
      w.qx = 2 * M_PI * round_half(xprd / a0) / xprd;
      // the y wave number divides by xprd, as the reference does
      w.qy = 2 * M_PI * round_half(yprd / a0) / xprd;
    } else {
      w.map = WM_111;
      w.h0 = h / 3;
      const double sqrt3 = std::sqrt(3.0);
      w.qx = 2 * M_PI * round_half(xprd / (a0 * sqrt3)) / xprd;
      w.qy = 2 * M_PI * round_half(yprd / a0) / yprd;
    }
  } else {
    w.map = WM_TABLE;
    have_table = true;
    if (map) {
      if (mx < 2 || my < 2) return "fix wall/map: map must be at least 2 x 2";
      out.values.assign(map, map + (size_t)mx * my);
    } else {
      int n = 0;
      std::string e = read_map(name, out.values, n);
      if (!e.empty()) return e;
      mx = my = n;
      if (n < 2) return "fix wall/map: map must be at least 2 x 2";
    }
    w.mx = mx;
    w.my = my;
  }
  out.z0_type = 0;
  if (i < a.size() && a[i] == "bot") {
    out.z0_type = 1;
    i++;
  } else if (i < a.size() && a[i] == "top") {
    out.z0_type = 2;
    i++;
  }
  if (i >= a.size()) return kSyntax;
  if (!to_double(a[i], w.z0))
    return fmt("fix contact/map: Could not convert z0 (= '%s') parameter to "
               "floating point value. Or did you mean to specify 'bot' or "
               "'top'?", a[i]);
  i++;
  if (i >= a.size()) return kSyntax;
  if (a[i] == "1")
    w.norm = 1;
  else if (a[i] == "-1")
    w.norm = -1;
  else
    return fmt("fix contact/map: Surface normal must be 1 or -1, but is "
               "'%s'.", a[i]);
  i++;
  if (i >= a.size()) return kSyntax;
  const std::string pot = a[i++];
  std::string e = P.potential(pot);
  if (!e.empty()) return e;
  bool raw = false;
  while (i < a.size()) {
    const std::string k = a[i];
    if (k == "lambda" || k == "pendist") {
      i++;
      if (i >= a.size())
        return fmt("fix contact/map: Expected value after '%s' flag.", k);
      double v;
      if (!to_double(a[i], v)) {
        char buf[512];
        std::snprintf(buf, sizeof buf,
                      "fix contact/map: Could not convert '%s' (= '%s') "
                      "parameter to floating point value.",
                      k.c_str(), a[i].c_str());
        return buf;
      }
      if (k == "lambda")
        out.lambda = v;
      else
        w.pendist = v;
    } else if (k == "vel") {
      return "fix wall/map: 'vel' (a map moving with time) is not supported "
             "by the static and dynamic solvers";
    } else if (k == "lowmem" || k == "dump") {
      // spline coefficients are formed on the fly; no map dump
    } else if (k == "disable_interpolation") {
      raw = true;
    } else {
      return fmt("fix contact/map: Unknown keyword '%s'.", k);
    }
    i++;
  }
  if (out.lambda > 0.0 && !have_table)
    return "Map filtering can only be used with map data file.";
  if (have_table) {
    w.qx = w.mx / xprd;
    w.qy = w.my / yprd;
    if (raw) w.map = WM_RAW;
    if (!spline_inverse(out.ainv)) return "fix wall/map: singular spline system";
    // the device evaluates the spline with the compiled copy of this inverse
    // (equal values; the sign of a zero entry does not matter, zero terms
    // are left out of the sums)
    for (int r = 0; r < 16; r++)
      for (int k = 0; k < 12; k++)
        if (!(out.ainv[r * 12 + k] == kSplineInv[r][k]))
          return "fix wall/map: spline inverse differs from the compiled "
                 "table";
  }
  return "";
}

}  // namespace gfmd
