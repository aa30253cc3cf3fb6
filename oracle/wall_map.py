"""Oracle restatement of `fix wall/map` (test infrastructure).

Used only by the tests; the product evaluates the wall on the device
(user-gfmd_amd/csrc/wall_map.h) from tables set up in
user-gfmd_amd/csrc/wall_map.cpp.

`fix ID group wall/map <map> [bot|top] <z0> <normal> <potential> <params>
[lambda L] [pendist d]`: a soft wall z_map(x, y) acting on every atom of
the group through a potential V(dz) of the normal distance
dz = normal (z - z_map(x, y)); the force is f_z = -normal V'(dz) and the
lateral force f_x = -f_z dz_map/dx, f_y = -f_z dz_map/dy.

Reference:
  FixWallMap constructor and keywords (src/extras/fix_wall_map.cpp:60-300)
  potentials lj93, lj126, lj93/smooth, exp, expf, dugdale (363-686)
  pre_force: the force loop and the pendist reset (686-811)
  maps flat, 100, 111 (829-1012), read_map (1079-1145), filter_map
    (1179-1269), displace_map (1275-1340), eval_table2d / eval_raw_data
    (1363-1398)
  Table2D: the periodic bicubic spline (src/mathutils/table2d.h:57-138,
    table2d.cpp:31-264), its 16 x 16 system inverted by Gauss-Jordan with
    full pivoting (linearalgebra.cpp:275-350) and the coefficients formed
    by mat_mul_vec (linearalgebra.h:197-207)

The map and spline arithmetic keeps the reference's operation order (numpy
evaluates elementwise in IEEE double without contraction), so that the
device, compiled without contraction, can be compared bit for bit.
"""
import math
import re

import numpy as np

from ft import gauss_jordan

SYNTAX = ("fix wall/map: Illegal fix command, syntax is 'fix ID group-ID "
          "wall/map plane|<mapfile> [bot|top] <z0> <surface normal> "
          "<lj93|lj93/smooth|lj126|exp|expf|dugdale> "
          "[<epsilon> <sigma> <cutoff>] [lambda <short-wavelength cutoff>] "
          "[pendist <distance from surface after penetration>] "
          "[vel <x-velocity> <y-velocity>]'")

_FLOAT = re.compile(r"\s*[+-]?((\d+\.?\d*|\.\d+)([eE][+-]?\d+)?|inf(inity)?"
                    r"|nan)", re.IGNORECASE)


class WallMapError(ValueError):
    pass


def _strtod(tok):
    """strtod: the longest leading float; None when nothing converts."""
    m = _FLOAT.match(tok)
    return float(m.group(0)) if m else None


def _round(x):
    """The fix's round(): (int)(x < 0 ? x - 0.5 : x + 0.5)."""
    return int(x - 0.5) if x < 0 else int(x + 0.5)


# ---------------------------------------------------------------------------
# Table2D
# ---------------------------------------------------------------------------
_IX1 = (0, 1, 1, 0)
_IX2 = (0, 0, 1, 1)


def _ipow(n, p):
    return float(n) ** p  # pow(0, 0) = 1, as pow() of the C library


def spline_matrix():
    """The 16 x 16 bicubic system (table2d.cpp:97-116): rows 0-3 values,
    4-7 d/dx, 8-11 d/dy, 12-15 d2/dxdy at the corners (0,0), (1,0), (1,1),
    (0,1); column 4 p1 + p2 is the monomial x^p1 y^p2."""
    A = [[0.0] * 16 for _ in range(16)]
    for r in range(4):
        n1, n2 = _IX1[r], _IX2[r]
        for p1 in range(4):
            for p2 in range(4):
                p1m, p2m = max(p1 - 1, 0), max(p2 - 1, 0)
                c = 4 * p1 + p2
                A[r][c] = 1.0 * (_ipow(n1, p1) * _ipow(n2, p2))
                A[r + 4][c] = 1.0 * (p1 * _ipow(n1, p1m) * _ipow(n2, p2))
                A[r + 8][c] = 1.0 * (_ipow(n1, p1) * p2 * _ipow(n2, p2m))
                A[r + 12][c] = 1.0 * (p1 * _ipow(n1, p1m) * p2 *
                                      _ipow(n2, p2m))
    return A


def spline_inverse():
    """A^-1 by the reference's Gauss-Jordan (16 x 16, row-major)."""
    flat = [x for row in spline_matrix() for x in row]
    inv = gauss_jordan(flat, 16)
    return np.array(inv).reshape(16, 16)


_AINV = None


def ainv():
    global _AINV
    if _AINV is None:
        _AINV = spline_inverse()
    return _AINV


class Table2D:
    """Periodic bicubic interpolation of values[nx][ny] on the unit grid:
    corner values, central-difference first derivatives (interp) or zero,
    zero cross derivatives; coefficients A^-1 B per box."""

    def __init__(self, values, interp=True):
        self.v = np.ascontiguousarray(values, dtype=np.float64)
        self.nx, self.ny = self.v.shape
        self.interp = interp

    def coefficients(self, bx, by):
        """coeff[..., i, j] of boxes (bx, by) (integer arrays, wrapped):
        compute_spline_coefficients (table2d.h:94-138)."""
        v, nx, ny = self.v, self.nx, self.ny
        bx = np.asarray(bx)
        by = np.asarray(by)
        B = [None] * 16
        zero = np.zeros(bx.shape)
        for r in range(4):
            x1 = (_IX1[r] + bx) % nx
            x2 = (_IX2[r] + by) % ny
            B[r] = v[x1, x2]
            if self.interp:
                B[r + 4] = (v[(x1 + 1) % nx, x2] - v[(x1 - 1) % nx, x2]) / 2
                B[r + 8] = (v[x1, (x2 + 1) % ny] - v[x1, (x2 - 1) % ny]) / 2
            else:
                B[r + 4] = zero
                B[r + 8] = zero
            B[r + 12] = zero
        A = ainv()
        c = np.empty(bx.shape + (4, 4))
        for row in range(16):
            t = np.zeros(bx.shape)
            for k in range(16):
                t = t + A[row, k] * B[k]
            c[..., row // 4, row % 4] = t
        return c

    def eval(self, x, y):
        """f, df/dx, df/dy at (x, y) in grid units (table2d.cpp:203-262,
        first derivatives of the eight-output form)."""
        x, y = np.broadcast_arrays(np.asarray(x, dtype=np.float64),
                                   np.asarray(y, dtype=np.float64))
        b1 = np.floor(x)
        b2 = np.floor(y)
        dx = x - b1
        dy = y - b2
        c = self.coefficients(b1.astype(np.int64) % self.nx,
                              b2.astype(np.int64) % self.ny)
        f = np.zeros(x.shape)
        fx = np.zeros(x.shape)
        fy = np.zeros(x.shape)
        for i in range(3, -1, -1):
            sf = np.zeros(x.shape)
            sfy = np.zeros(x.shape)
            for j in range(3, -1, -1):
                cij = c[..., i, j]
                sf = sf * dy + cij
                if j > 0:
                    sfy = sfy * dy + j * cij
            f = f * dx + sf
            if i > 0:
                fx = fx * dx + i * sf
            fy = fy * dx + sfy
        return f, fx, fy


# ---------------------------------------------------------------------------
# The fix
# ---------------------------------------------------------------------------
def read_map_text(text):
    """read_map (fix_wall_map.cpp:1079-1145): the column count of the first
    line (a column starts where an alphanumeric, '-' or 'e' follows a
    blank) sets a square nx x nx map, read row by row."""
    nl = text.find("\n")
    if nl < 0:
        raise WallMapError("fix contact/map: Could not read map file.")
    n, last = 0, " "
    for ch in text[:nl]:
        if (ch.isalnum() or ch in "-e") and last == " ":
            n += 1
        last = ch
    if n <= 0:
        raise WallMapError("fix contact/map: Something wrong: nx_ <= 0")
    vals = []
    for tok in text.split():
        v = _strtod(tok)
        if v is None:
            break
        vals.append(v)
        if len(vals) == n * n:
            break
    if len(vals) < n * n:
        raise WallMapError("fix contact/map: Error read map file.")
    return np.array(vals).reshape(n, n)


def filter_map(h, lam, xprd, yprd):
    """filter_map (fix_wall_map.cpp:1179-1269): zero every Fourier mode with
    sqrt((kx/xprd)^2 + (ky/yprd)^2) > 1/lambda, kx = min(i, nx - i)."""
    nx, ny = h.shape
    kx = np.minimum(np.arange(nx), nx - np.arange(nx)) * (1.0 / xprd)
    ky = np.minimum(np.arange(ny), ny - np.arange(ny)) * (1.0 / yprd)
    r = np.sqrt((kx * kx)[:, None] + (ky * ky)[None, :])
    H = np.fft.fft2(h)
    H[r > 1.0 / lam] = 0.0
    return np.fft.ifft2(H).real


def displace_map(h, z0, z0_type):
    """displace_map (fix_wall_map.cpp:1275-1340): shift so that the average
    (or the minimum for 'bot', the maximum for 'top') sits at z0; returns
    the shifted map, the rms height and the new average."""
    n = h.size
    za = float(np.sum(h))
    h2 = float(np.sum(h * h))
    rms = math.sqrt(h2 / n - (za / n) * (za / n))
    if z0_type == "bot":
        s = z0 - float(h.min())
    elif z0_type == "top":
        s = z0 - float(h.max())
    else:
        s = z0 - za / n
    return h + s, rms, za / n + s


class WallMap:
    """The fix with its arguments after `wall/map`; xprd, yprd: the periodic
    box lengths.  map_values replaces reading a map file (the first token
    is then a label)."""

    def __init__(self, args, xprd, yprd, map_values=None, map_text=None):
        a = args.split()
        if len(a) < 4:
            raise WallMapError(SYNTAX)
        self.xprd, self.yprd = xprd, yprd
        self.cutoff = -1.0
        self.offset = 0.0
        self.lam = -1.0
        self.pendist = -1.0
        self.x0 = self.y0 = 0.0
        self.values = None
        self.raw = False
        i = 0
        name = a[i]
        i += 1
        if name == "flat":
            self.kind = "flat"
        elif name in ("100", "111"):
            self.kind = name
            if i + 2 > len(a):
                raise WallMapError(SYNTAX)
            a0 = _strtod(a[i])
            if a0 is None:
                raise WallMapError("fix contact/map: Could not convert "
                                   "lattice constant (= '%s')" % a[i])
            h0 = _strtod(a[i + 1])
            if h0 is None:
                raise WallMapError("fix contact/map: Could not convert "
                                   "amplitude (= '%s')" % a[i + 1])
            i += 2
            if name == "100":
                self.h0 = h0 / 2
                self.qx = 2 * math.pi * _round(xprd / a0) / xprd
                # the y wave number is scaled by xprd, as the reference does
                self.qy = 2 * math.pi * _round(yprd / a0) / xprd
            else:
                self.h0 = h0 / 3
                s3 = math.sqrt(3.0)
                self.qx = 2 * math.pi * _round(xprd / (a0 * s3)) / xprd
                self.qy = 2 * math.pi * _round(yprd / a0) / yprd
        else:
            self.kind = "table"
            if map_values is None:
                if map_text is None:
                    raise WallMapError("Could not open contact map file "
                                       "'%s'." % name)
                map_values = read_map_text(map_text)
            self.values = np.array(map_values, dtype=np.float64)
        self.z0_type = "average"
        if i < len(a) and a[i] in ("bot", "top"):
            self.z0_type = a[i]
            i += 1
        if i >= len(a):
            raise WallMapError(SYNTAX)
        z0 = _strtod(a[i])
        if z0 is None:
            raise WallMapError(
                "fix contact/map: Could not convert z0 (= '%s') parameter to "
                "floating point value. Or did you mean to specify 'bot' or "
                "'top'?" % a[i])
        self.z0 = z0
        i += 1
        if i >= len(a):
            raise WallMapError(SYNTAX)
        if a[i] == "1":
            self.norm = 1
        elif a[i] == "-1":
            self.norm = -1
        else:
            raise WallMapError("fix contact/map: Surface normal must be 1 or "
                               "-1, but is '%s'." % a[i])
        i += 1
        if i >= len(a):
            raise WallMapError(SYNTAX)
        pot = a[i]
        i += 1
        i = self._init_potential(pot, a, i)
        disable = False
        while i < len(a):
            k = a[i]
            if k in ("lambda", "pendist"):
                i += 1
                if i >= len(a):
                    raise WallMapError("fix contact/map: Expected value after "
                                       "'%s' flag." % k)
                v = _strtod(a[i])
                if v is None:
                    raise WallMapError(
                        "fix contact/map: Could not convert '%s' (= '%s') "
                        "parameter to floating point value." % (k, a[i]))
                if k == "lambda":
                    self.lam = v
                else:
                    self.pendist = v
            elif k == "lowmem":
                pass
            elif k == "disable_interpolation":
                disable = True
            else:
                raise WallMapError("fix contact/map: Unknown keyword '%s'."
                                   % k)
            i += 1
        if self.lam > 0.0:
            if self.values is None:
                raise WallMapError("Map filtering can only be used with map "
                                   "data file.")
            self.values = filter_map(self.values, self.lam, xprd, yprd)
        if self.values is not None:
            self.values, self.rms, self.avg = displace_map(
                self.values, self.z0, self.z0_type)
            self.mqx = self.values.shape[0] / xprd
            self.mqy = self.values.shape[1] / yprd
            self.raw = disable
            self.table = Table2D(self.values, True)

    # -- potentials ---------------------------------------------------------
    def _eps_sigma(self, a, i):
        if i + 2 > len(a):
            raise WallMapError(SYNTAX)
        e, s = _strtod(a[i]), _strtod(a[i + 1])
        if e is None:
            raise WallMapError("fix contact/map: Could not convert first "
                               "potential parameter to floating point value.")
        if s is None:
            raise WallMapError("fix contact/map: Could not convert second "
                               "potential parameter to floating point value.")
        self.eps, self.sigma = e, s
        return i + 2

    def _cut(self, a, i):
        if i >= len(a):
            raise WallMapError(SYNTAX)
        c = _strtod(a[i])
        if c is None:
            raise WallMapError("fix contact/map: Could not convert potential "
                               "cutoff parameter to floating point value.")
        self.cutoff = c
        self.mcutoff = 0.01 * c
        return i + 1

    def _lj93(self, a, i):
        i = self._eps_sigma(a, i)
        i = self._cut(a, i)
        e, s = self.eps, self.sigma
        self.c1 = 6.0 / 5.0 * e * math.pow(s, 9.0)
        self.c2 = 3.0 * e * math.pow(s, 3.0)
        self.c3 = 2.0 / 15.0 * e * math.pow(s, 9.0)
        self.c4 = e * math.pow(s, 3.0)
        self.c5 = 12.0 * e * math.pow(s, 9.0)
        self.c6 = 12.0 * e * math.pow(s, 3.0)
        if self.cutoff > 0.0:
            rinv = 1.0 / self.cutoff
            r2 = rinv * rinv
            r4 = r2 * r2
            self.offset = self.c3 * r4 * r4 * rinv - self.c4 * r2 * rinv
        m = self.mcutoff
        self.mc0 = self.c3 / math.pow(m, 9) - self.c4 / math.pow(m, 3) - \
            self.offset
        self.mc1 = -self.c1 / math.pow(m, 10) + self.c2 / math.pow(m, 4)
        self.mc2 = self.c5 / math.pow(m, 11) - self.c6 / math.pow(m, 5)
        return i

    def _init_potential(self, pot, a, i):
        self.pot = pot
        if pot == "lj93":
            return self._lj93(a, i)
        if pot == "lj126":
            i = self._eps_sigma(a, i)
            i = self._cut(a, i)
            e, s = self.eps, self.sigma
            self.c1 = 48.0 * e * math.pow(s, 12.0)
            self.c2 = 24.0 * e * math.pow(s, 6.0)
            self.c3 = 4.0 * e * math.pow(s, 12.0)
            self.c4 = 4.0 * e * math.pow(s, 6.0)
            if self.cutoff > 0.0:
                r2 = 1.0 / (self.cutoff * self.cutoff)
                r6 = r2 * r2 * r2
                self.offset = r6 * (self.c3 * r6 - self.c4)
            return i
        if pot == "lj93/smooth":
            i = self._lj93(a, i)
            self.inner = self.cutoff
            if i >= len(a):
                raise WallMapError(SYNTAX)
            c = _strtod(a[i])
            if c is None:
                raise WallMapError("fix contact/map: Could not convert second "
                                   "potential cutoff parameter to floating "
                                   "point value.")
            self.cutoff = c
            i += 1
            rinv = 1.0 / self.inner
            r2 = rinv * rinv
            r4 = r2 * r2
            r10 = r4 * r4 * r2
            s1 = self.c1 * r10 - self.c2 * r4
            s2 = -10 * self.c1 * r10 * rinv + 4 * self.c2 * r4 * rinv
            dr = self.cutoff - self.inner
            if dr <= 0.0:
                raise WallMapError("fix contact/map: lj93/smooth outer cutoff "
                                   "must be bigger than inner cutoff.")
            s3 = -(3 * s1 + 2 * s2 * dr) / (dr * dr)
            s4 = (2 * s1 + s2 * dr) / (dr * dr * dr)
            s0 = s1 * dr + s2 * dr * dr / 2.0 + s3 * dr * dr * dr / 3.0 + \
                s4 * dr * dr * dr * dr / 4.0
            self.s = (s0, s1, s2, s3, s4)
            self.offset -= s0
            return i
        if pot == "exp":
            i = self._eps_sigma(a, i)
            i = self._cut(a, i)
            self.c1 = 1.0 / self.sigma
            if self.cutoff > 0.0:
                self.offset = self.eps * math.exp(-self.cutoff * self.c1)
            return i
        if pot == "expf":
            i = self._eps_sigma(a, i)
            i = self._cut(a, i)
            self.c1 = 1.0 / self.sigma
            if self.cutoff > 0.0:
                self.offset = self.eps * (self.sigma + self.cutoff)
            return i
        if pot == "dugdale":
            i = self._eps_sigma(a, i)
            self.cutoff = self.sigma
            self.c1 = self.eps / self.cutoff
            return i
        raise WallMapError("fix contact/map: Unknown potential type '%s'."
                           % pot)

    def potential(self, dz):
        """e, f = -dV/ddz, d2e at the distances dz (array); plus a mask of
        lj126 distances below 0.01 cutoff (the reference stops there)."""
        dz = np.asarray(dz, dtype=np.float64)
        close = np.zeros(dz.shape, bool)
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            if self.pot in ("lj93", "lj93/smooth"):
                ddz = dz - self.mcutoff
                e_q = self.mc0 + self.mc1 * ddz + 0.5 * self.mc2 * ddz * ddz
                f_q = -self.mc1 - self.mc2 * ddz
                rinv = 1.0 / dz
                r2 = rinv * rinv
                r4 = r2 * r2
                r10 = r4 * r4 * r2
                e_l = self.c3 * r4 * r4 * rinv - self.c4 * r2 * rinv - \
                    self.offset
                f_l = self.c1 * r10 - self.c2 * r4
                d_l = self.c5 * r10 * rinv - self.c6 * r4 * rinv
                q = dz < self.mcutoff
                e = np.where(q, e_q, e_l)
                f = np.where(q, f_q, f_l)
                d2 = np.where(q, self.mc2, d_l)
                if self.pot == "lj93/smooth":
                    s0, s1, s2, s3, s4 = self.s
                    r = dz - self.inner
                    rr = r * r
                    e_s = s0 - s1 * r - s2 * rr / 2.0 - s3 * rr * r / 3.0 - \
                        s4 * rr * rr / 4.0
                    f_s = s1 + s2 * r + s3 * rr + s4 * rr * r
                    d_s = -s2 - 2.0 * s3 * r - 3.0 * s4 * rr
                    o = ~q & (dz >= self.inner)
                    e = np.where(o, e_s, e)
                    f = np.where(o, f_s, f)
                    d2 = np.where(o, d_s, d2)
                return e, f, d2, close
            if self.pot == "lj126":
                close = dz < self.mcutoff
                rinv = 1.0 / dz
                r2 = rinv * rinv
                r6 = r2 * r2 * r2
                e = r6 * (self.c3 * r6 - self.c4) - self.offset
                f = r6 * (self.c1 * r6 - self.c2) * rinv
                return e, f, np.zeros(dz.shape), close
            if self.pot == "exp":
                e = self.eps * np.exp(-dz * self.c1)
                f = e * self.c1
                d2 = f * self.c1
                return e - self.offset, f, d2, close
            if self.pot == "expf":
                t = self.eps * np.exp(self.c1 * (self.cutoff - dz))
                return (self.sigma * t + self.eps * dz - self.offset,
                        t - self.eps, self.c1 * t, close)
            # dugdale
            pos = dz > 0.0
            e = np.where(pos, self.c1 * (self.cutoff - dz), 1e12)
            f = np.where(pos, self.c1, 0.0)
            return e, f, np.zeros(dz.shape), close

    # -- maps ---------------------------------------------------------------
    def height(self, x, y):
        """z_map, dz/dx, dz/dy at (x, y)."""
        x, y = np.broadcast_arrays(np.asarray(x, dtype=np.float64),
                                   np.asarray(y, dtype=np.float64))
        if self.kind == "flat":
            z = np.full(x.shape, self.z0)
            return z, np.zeros(x.shape), np.zeros(x.shape)
        if self.kind == "100":
            dx, dy = x - self.x0, y - self.y0
            cx, cy = np.cos(self.qx * dx), np.cos(self.qy * dy)
            sx, sy = np.sin(self.qx * dx), np.sin(self.qy * dy)
            return (self.z0 + self.h0 * (cx + cy), -self.h0 * self.qx * sx,
                    -self.h0 * self.qy * sy)
        if self.kind == "111":
            dx, dy = x - self.x0, y - self.y0
            a2 = self.qx * dx + 0.5 * self.qy * dy
            a3 = self.qx * dx - 0.5 * self.qy * dy
            c1, c2, c3 = np.cos(self.qy * dy), np.cos(a2), np.cos(a3)
            s1, s2, s3 = np.sin(self.qy * dy), np.sin(a2), np.sin(a3)
            return (self.z0 + self.h0 * (c1 + c2 + c3),
                    -self.h0 * self.qx * (s2 + s3),
                    -self.h0 * self.qy * (s1 + 0.5 * s2 - 0.5 * s3))
        if self.raw:
            nx, ny = self.values.shape
            ix = np.trunc(x * self.mqx).astype(np.int64) % nx
            iy = np.trunc(y * self.mqy).astype(np.int64) % ny
            return self.values[ix, iy], np.zeros(x.shape), np.zeros(x.shape)
        f, fx, fy = self.table.eval(x * self.mqx, y * self.mqy)
        return f, fx * self.mqx, fy * self.mqy

    # -- pre_force ----------------------------------------------------------
    def pre_force(self, pos):
        """Forces on atoms at pos (n, 3); pos[:, 2] is reset in place where
        the pendist rule applies.  Returns f (n, 3), the fix's vector
        (energy, fx, fy, fz sums, ncon, nrep, natt) and the number of lj126
        atoms that the reference rejects as too close."""
        x, y = pos[:, 0], pos[:, 1]
        z, zx, zy = self.height(x, y)
        dz = self.norm * (pos[:, 2] - z)
        act = np.ones(dz.shape, bool) if self.cutoff < 0.0 else \
            (dz < self.cutoff if self.cutoff > 0.0 else np.zeros(dz.shape,
                                                                 bool))
        reset = act & (dz < self.pendist)
        pos[:, 2] = np.where(reset, z + self.norm * self.pendist, pos[:, 2])
        dz = np.where(reset, self.pendist, dz)
        e, f, _, close = self.potential(np.where(act, dz, 1.0))
        e = np.where(act, e, 0.0)
        f = np.where(act, f, 0.0)
        fz = f * self.norm
        out = np.zeros(pos.shape)
        out[:, 0] = np.where(act, -fz * zx, 0.0)
        out[:, 1] = np.where(act, -fz * zy, 0.0)
        out[:, 2] = fz
        vec = np.array([np.sum(e), np.sum(out[:, 0]), np.sum(out[:, 1]),
                        np.sum(out[:, 2]), np.sum(act),
                        np.sum(act & (f > 0.0)), np.sum(act & ~(f > 0.0))],
                       dtype=np.float64)
        return out, vec, int(np.sum(act & close))


def surface_positions(u, lattice, cell_xy=None, offsets=None):
    """Atom positions of the contact atoms: (x0 + cx ix + off_x + u_x, ...)
    for u (d, nx, ny); returns (natoms, nx, ny, 3)."""
    a, x0, y0, z0 = lattice
    cx, cy = cell_xy if cell_xy is not None else (a, a)
    offsets = offsets if offsets is not None else [(0.0, 0.0, 0.0)]
    _, nx, ny = u.shape
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    out = np.empty((len(offsets), nx, ny, 3))
    for l, (ox, oy, oz) in enumerate(offsets):
        out[l, ..., 0] = x0 + cx * X + ox + u[3 * l]
        out[l, ..., 1] = y0 + cy * Y + oy + u[3 * l + 1]
        out[l, ..., 2] = z0 + oz + u[3 * l + 2]
    return out


def fire_relax_wall_map(phi, u, lattice, wm, fire, steps, pressure=0.0,
                        cell_xy=None, offsets=None):
    """Reference-order FIRE (setup: pre_force then forces; per step:
    initial_integrate, pre_force (pendist reset, wall force), elastic force,
    final_integrate) with the wall acting on the listed atoms of each
    cell and a uniform normal force on layer 0; returns u, v, the fix's
    vector of the last evaluation."""
    from oracle import static_force
    u = u.copy()
    v = np.zeros_like(u)
    offsets = offsets if offsets is not None else [(0.0, 0.0, 0.0)]
    a, x0, y0, z0 = lattice
    nl = len(offsets)

    def force(u):
        pos = surface_positions(u, lattice, cell_xy, offsets)
        flat = pos.reshape(-1, 3).copy()
        fw, vec, close = wm.pre_force(flat)
        if close:
            raise RuntimeError("fix contact/map: Atom too close")
        moved = (flat[:, 2] != pos.reshape(-1, 3)[:, 2]).reshape(pos.shape[:3])
        znew = flat[:, 2].reshape(pos.shape[:3])
        for l in range(nl):
            u[3 * l + 2] = np.where(moved[l], znew[l] - (z0 + offsets[l][2]),
                                    u[3 * l + 2])
        f, _, _ = static_force(phi, u)
        fw = fw.reshape(nl, u.shape[1], u.shape[2], 3)
        for l in range(nl):
            for c in range(3):
                f[3 * l + c] += fw[l, ..., c]
        f[2] += pressure
        return f, vec

    f, vec = force(u)
    for _ in range(steps):
        fire.initial(u, v, f)
        f, vec = force(u)
        fire.final(v, f)
    return u, v, vec
