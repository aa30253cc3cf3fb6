"""Oracle restatement of the `ft` stiffness kernel (test infrastructure).

Used only by the tests and the fixture generator; the product builds its
own tables in user-gfmd_amd/csrc/ft_lattice.cpp.

`ft <surface> <lattice constant> <nu> <force constants> ... [height H]`:
a crystal surface unit cell (sc100, fcc100, fcc111; nu primitive cells
stacked along z), a neighbour list with neighbour shells, per-pair force
constants D = -k e e^T - kappa (1 - e e^T), on-site blocks from the sum
rule over neighbours at or below a layer, and U0 / U / V assembled with the
phases exp(i (cx qx + cy qy)) of the neighbour's cell offset; Phi(q) then
follows from the transfer-matrix stiffness of the oracle.

Reference:
  FTStiffnessKernel (src/stiffness_kernels/ft_stiffness.cpp:47-116 args,
    128-191 tables, 199-292 U0/U/V)
  CrystalSurface (src/main/crystal_surface.cpp:166-246 neighbour list and
    shells, 368-427 transform/scale/cell distance, 430-553 supercell,
    290-364 reverse_indices)
  SC100Surface / FCC100Surface / FCC111Surface primitive cells
    (src/surfaces/sc100_surface.cpp:86-110, fcc100_surface.cpp:86-110,
    fcc111_surface.cpp:67-94)
  ForceConstants::onsite (src/main/force_constants.cpp:79-97)
  FCPairPotential (src/force_constants/fc_pair_potential.cpp:32-181)
  FCLJCut (src/force_constants/fc_lj_cut.cpp:87-170)

Floating-point note: the supercell wraps scaled positions into [0, 1) with
floor(); an atom whose exact scaled coordinate is an integer lands on
either side depending on rounding, which moves it by a lattice vector and
changes the phase convention of Phi.  The cell inverse (Gauss-Jordan with
full pivoting, linearalgebra.cpp:275-350) and the matrix-vector products
therefore keep the reference's operation order here.
"""
import math

import numpy as np


def gauss_jordan(m, n=3):
    """Inverse of an n x n matrix (row-major list of n*n) by Gauss-Jordan
    elimination with full pivoting: pivot = the largest |a| over the rows
    and columns not yet used (the last one on ties), row swap onto the
    diagonal, scaling by 1/pivot, elimination of the pivot column from every
    other row, and the column swaps undone in reverse order at the end."""
    a = list(m)
    used = [0] * n
    rows, cols = [0] * n, [0] * n
    for step in range(n):
        big = 0.0
        pr = pc = 0
        for r in range(n):
            if used[r] == 1:
                continue
            for c in range(n):
                if used[c] == 0 and abs(a[r * n + c]) >= big:
                    big = abs(a[r * n + c])
                    pr, pc = r, c
                elif used[c] > 1:
                    raise ValueError("singular cell")
        used[pc] += 1
        if pr != pc:
            for c in range(n):
                a[pr * n + c], a[pc * n + c] = a[pc * n + c], a[pr * n + c]
        rows[step], cols[step] = pr, pc
        if a[pc * n + pc] == 0.0:
            raise ValueError("singular cell")
        inv = 1.0 / a[pc * n + pc]
        a[pc * n + pc] = 1.0
        for c in range(n):
            a[pc * n + c] *= inv
        for r in range(n):
            if r != pc:
                f = a[r * n + pc]
                a[r * n + pc] = 0.0
                for c in range(n):
                    a[r * n + c] -= a[pc * n + c] * f
    for step in range(n - 1, -1, -1):
        r, c = rows[step], cols[step]
        if r != c:
            for k in range(n):
                a[k * n + r], a[k * n + c] = a[k * n + c], a[k * n + r]
    return a


def matvec(m, v):
    """Row-major 3 x 3 times 3-vector, each row summed left to right from
    zero (mat_mul_vec, linearalgebra.h:197-207)."""
    out = []
    for i in range(3):
        s = 0.0
        for j in range(3):
            s += m[3 * i + j] * v[j]
        out.append(s)
    return out


def _primitive(name):
    """Cell (row-major, lattice vectors as columns, in nearest-neighbour
    distances) and scaled atom positions of one primitive surface cell."""
    if name == "sc100":
        return [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0], [(0.0, 0.0, 0.0)]
    if name == "fcc100":
        h = 1.0 / math.sqrt(2.0)
        return [1.0, 0.0, 0.5, 0.0, 1.0, 0.5, 0.0, 0.0, h], [(0.0, 0.0, 0.0)]
    if name == "fcc111":
        s3 = math.sqrt(3.0)
        return ([1.0, 0.0, 0.5, 0.0, s3, 1.0 / (2.0 * s3), 0.0, 0.0,
                 math.sqrt(2.0 / 3.0)], [(0.0, 0.0, 0.0), (0.5, 0.5, 0.0)])
    raise ValueError("unknown crystal surface %r" % name)


class Surface:
    """Surface unit cell: `nu` primitive cells stacked along z (atom 0 on
    top after the index reversal), scaled to the lattice constant."""

    def __init__(self, name, lattice, nu):
        if not 0 < lattice <= 100:
            raise ValueError("lattice constant out of range")
        if not 1 <= nu <= 100:
            raise ValueError("number of lattice cells out of range")
        self.name = name
        cell, scaled = _primitive(name)
        self._set_cell(cell)
        self.pos = [matvec(cell, list(s)) for s in scaled]
        if nu > 1:
            self._supercell_z(nu)
            self.pos.reverse()
        self._scale([lattice * c for c in self.cell])
        self.nat = len(self.pos)
        self.neigh = []

    def _set_cell(self, cell):
        if cell[1] != 0.0 or cell[3] != 0.0 or cell[6] != 0.0 or cell[7] != 0.0:
            raise ValueError("cell must be aligned along x and y")
        self.cell = list(cell)
        self.inv = gauss_jordan(self.cell)

    def _supercell_z(self, nz):
        old = list(self.cell)
        new = list(old)
        for r in range(3):
            new[3 * r + 2] *= nz
        # smallest shift of the stacking vector in x and y
        new[2] -= math.floor(new[2] / new[0]) * new[0]
        new[5] -= math.floor(new[5] / new[4]) * new[4]
        self._set_cell(new)
        out = []
        for z in range(nz):
            shift = matvec(old, [0.0, 0.0, float(z)])
            for p in self.pos:
                q = [p[k] + shift[k] for k in range(3)]
                s = matvec(self.inv, q)
                s = [x - math.floor(x) for x in s]
                out.append(matvec(self.cell, s))
        self.pos = out

    def _scale(self, cell):
        self.pos = [matvec(self.inv, p) for p in self.pos]
        self._set_cell(cell)
        self.pos = [matvec(self.cell, p) for p in self.pos]

    # -- neighbours --------------------------------------------------------
    def neighbor_list(self, rcut):
        """All (i, j, cell offset) with |r_i - r_j - cell.d| < rcut, in the
        order atom i, atom j, x, y, z; search range floor(rcut / |a_k|^2)
        + 1 cells per direction (the squared length, as the reference
        computes it); shells = distinct distances (0.01 tolerance) sorted."""
        c = self.cell
        lens = [c[k] ** 2 + c[3 + k] ** 2 + c[6 + k] ** 2 for k in range(3)]
        rng = [int(math.floor(rcut / l)) + 1 for l in lens]
        out = []
        for i in range(self.nat):
            for j in range(self.nat):
                for x in range(-rng[0], rng[0] + 1):
                    for y in range(-rng[1], rng[1] + 1):
                        for z in range(-rng[2], rng[2] + 1):
                            if i == j and x == 0 and y == 0 and z == 0:
                                continue
                            t = matvec(c, [float(x), float(y), float(z)])
                            r = [self.pos[i][k] - self.pos[j][k] - t[k]
                                 for k in range(3)]
                            rn = math.sqrt(r[0] * r[0] + r[1] * r[1] +
                                           r[2] * r[2])
                            if rn < rcut:
                                out.append({"i": i, "j": j, "r": r, "rn": rn,
                                            "x": x, "y": y, "ab": z})
        seen = []
        for n in out:
            if not any(abs(n["rn"] - s) < 0.01 for s in seen):
                seen.append(n["rn"])
        seen.sort()
        for n in out:
            n["shell"] = next(k for k, s in enumerate(seen)
                              if abs(n["rn"] - s) < 0.01)
        self.neigh = out
        return out

    def cell_distance(self, n):
        """Cell offset of neighbour n in lattice coordinates (near-integer
        doubles): inverse cell times r_i - r_j - r."""
        i, j = n["i"], n["j"]
        r = [self.pos[i][k] - self.pos[j][k] - n["r"][k] for k in range(3)]
        return matvec(self.inv, r)


def _pair_block(r, rn, k, kappa):
    e = np.array(r) / rn
    ee = np.outer(e, e)
    return -k * ee - kappa * (np.eye(3) - ee)


class PairPotential:
    """`pair-potential <n>[x] k1 [kappa1] ... [linf nu l1 ...]`: springs per
    neighbour shell (up to three).  The parsed linf values are not used by
    the kernel: FCPairPotential::linear takes two arguments and does not
    override the three-argument virtual, so the Gamma-point force stays 0
    (fc_pair_potential.cpp:183-192 vs force_constants.cpp:103-107)."""

    def __init__(self, surface, nshell, k, kappa=None):
        self.k = list(k) + [0.0] * (3 - len(k))
        kap = list(kappa or []) + [0.0] * (3 - len(kappa or []))
        self.kappa = kap
        c = surface.cell
        surface.neighbor_list(2.01 * max(c[0], c[4]))

    def offsite(self, n):
        s = n["shell"]
        k, kap = (self.k[s], self.kappa[s]) if s < 3 else (0.0, 0.0)
        return _pair_block(n["r"], n["rn"], k, kap)

    def linear(self, surface):
        return np.zeros(surface.nat)


class LJCut:
    """`lj/cut` with 12-6 Lennard-Jones epsilon, sigma, cutoff (the
    coefficients the reference reads from the LAMMPS pair style):
    k = V''(r), kappa = V'(r) / r inside the cutoff, neighbour list out to
    the cutoff, and a Gamma-point force -sum V'(r) e_z over neighbours at
    or below the atom's layer."""

    def __init__(self, surface, eps, sigma, rc):
        self.eps, self.sigma, self.rc = eps, sigma, rc
        surface.neighbor_list(rc)

    def d1(self, r):
        if r >= self.rc:
            return 0.0
        s6 = (self.sigma / r) ** 6
        return 24.0 * self.eps * s6 * (1.0 - 2.0 * s6) / r

    def d2(self, r):
        if r >= self.rc:
            return 0.0
        s6 = (self.sigma / r) ** 6
        return 24.0 * self.eps * s6 * (26.0 * s6 - 7.0) / (r * r)

    def offsite(self, n):
        rn = n["rn"]
        return _pair_block(n["r"], rn, self.d2(rn), self.d1(rn) / rn)

    def linear(self, surface):
        out = np.zeros(surface.nat)
        for n in surface.neigh:
            if n["ab"] <= 0:
                out[n["i"]] += self.d1(n["rn"]) * n["r"][2] / n["rn"]
        return -out


def _strtol(tok):
    """Leading base-10 integer of tok, as strtol reads it ("1.0" -> 1)."""
    m = 0
    while m < len(tok) and (tok[m].isdigit() or (m == 0 and tok[m] in "+-")):
        m += 1
    if m == 0 or tok[:m] in "+-":
        raise ValueError("integer expected, got %r" % tok)
    return int(tok[:m])


def parse(args):
    """(surface, force constants, height) from the `ft` argument string."""
    a = args.split()
    if not a:
        raise ValueError("missing crystal surface")
    name = a[0]
    lattice = float(a[1])
    nu = _strtol(a[2])
    surf = Surface(name, lattice, nu)
    p = 3
    model = a[p]
    p += 1
    if model == "pair-potential":
        tok = a[p]
        x = tok.endswith("x")
        nk = _strtol(tok[:-1] if x else tok)
        p += 1
        k, kap = [], []
        for _ in range(nk):
            k.append(float(a[p]))
            p += 1
            if x:
                kap.append(float(a[p]))
                p += 1
        while p < len(a) - 1 and a[p] == "linf":
            nl = _strtol(a[p + 1])
            p += 2 + nl
        fc = PairPotential(surf, nk, k, kap)
    elif model == "lj/cut":
        fc = LJCut(surf, float(a[p]), float(a[p + 1]), float(a[p + 2]))
        p += 3
    else:
        raise ValueError("unknown interaction model %r" % model)
    height = 0
    while p < len(a) - 1 and a[p] == "height":
        height = _strtol(a[p + 1])
        p += 2
    return surf, fc, height


class FTKernel:
    """Phi(q) of `ft <args>` (3 nat x 3 nat), Gamma-point forces linf."""

    def __init__(self, args):
        self.surf, self.fc, self.height = parse(args)
        s = self.surf
        self.dim = 3 * s.nat
        self.pairs = []
        for n in s.neigh:
            cx = s.cell_distance(n)
            ab = int(round(cx[2]))
            assert abs(ab - cx[2]) < 1e-6
            self.pairs.append((n["i"], n["j"], cx[0], cx[1], ab,
                               self.fc.offsite(n)))
        # on-site blocks: minus the sum over neighbours at or below layer ab
        self.on = np.zeros((2, s.nat, 3, 3))
        for n in s.neigh:
            D = self.fc.offsite(n)
            for ab in (0, 1):
                if n["ab"] <= ab:
                    self.on[ab, n["i"]] -= D
        self.linf = self.fc.linear(s)

    def matrices(self, qx, qy):
        d = self.dim
        U0 = np.zeros((d, d), np.complex128)
        U = np.zeros((d, d), np.complex128)
        V = np.zeros((d, d), np.complex128)
        for i, j, cx, cy, ab, D in self.pairs:
            ph = np.exp(1j * (cx * qx + cy * qy))
            bi, bj = slice(3 * i, 3 * i + 3), slice(3 * j, 3 * j + 3)
            if ab == 0:
                U0[bi, bj] += ph * D
                U[bi, bj] += ph * D
            elif ab == -1:
                V[bi, bj] += ph * D
        for i in range(self.surf.nat):
            b = slice(3 * i, 3 * i + 3)
            U0[b, b] += self.on[0, i]
            U[b, b] += self.on[1, i]
        return U0, U, V

    def phi(self, qx, qy):
        import oracle as O
        U0, U, V = self.matrices(qx, qy)
        return O.transfer_stiffness(self.height, U0, U, V)

    def fill_phi(self, nx, ny, normalize=True):
        """Full-grid table (nx*ny, d, d), index ix*ny + iy, reference
        wavevector convention, divided by nx*ny when normalize.  U0, U, V
        are summed for all wavevectors at once (all-zero blocks add
        nothing and are skipped); the transfer matrix runs per q."""
        import oracle as O
        tp = 2.0 * math.pi
        qx1 = np.array([tp * i / nx if i <= nx // 2 else tp * (i - nx) / nx
                        for i in range(nx)])
        qy1 = np.array([tp * i / ny if i <= ny // 2 else tp * (i - ny) / ny
                        for i in range(ny)])
        qx = np.repeat(qx1, ny)
        qy = np.tile(qy1, nx)
        d = self.dim
        M = np.zeros((3, nx * ny, d, d), np.complex128)  # U0, U, V
        for i, j, cx, cy, ab, D in self.pairs:
            if ab not in (0, -1) or not D.any():
                continue
            ph = np.exp(1j * (cx * qx + cy * qy))[:, None, None] * D
            bi, bj = slice(3 * i, 3 * i + 3), slice(3 * j, 3 * j + 3)
            if ab == 0:
                M[0][:, bi, bj] += ph
                M[1][:, bi, bj] += ph
            else:
                M[2][:, bi, bj] += ph
        for i in range(self.surf.nat):
            b = slice(3 * i, 3 * i + 3)
            M[0][:, b, b] += self.on[0, i]
            M[1][:, b, b] += self.on[1, i]
        out = np.zeros((nx * ny, d, d), np.complex128)
        for n in range(nx * ny):
            out[n] = O.transfer_stiffness(self.height, M[0][n], M[1][n],
                                          M[2][n])
        if normalize:
            out /= nx * ny
        return out
