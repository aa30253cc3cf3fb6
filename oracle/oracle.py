"""ctypes front end of the CPU FP64 oracle (oracle/libgfmd_oracle.so) and of
the reference's own stiffness core built into oracle/_ref/libgfmd_ref.so.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, always as the reference result; the product package
never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgfmd_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libgfmd_ref.so")

KERNEL_SC100 = 1
KERNEL_FCC100 = 2
KERNEL_ISOTROPIC = 3

_lib = None
_ref = None

_dp = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile the oracle (and, when the reference tree is present, the
    reference stiffness core)."""
    subprocess.check_call(["make", "-s", "-C", HERE])
    repo = os.path.dirname(HERE)
    if os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-f", "oracle/Makefile.ref"],
                              cwd=repo)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_fft_c2c.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
        L.orc_fft2d.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_int]
        L.orc_invert.argtypes = [ctypes.c_int, ctypes.c_void_p]
        L.orc_sc100_layer.argtypes = [ctypes.c_double, ctypes.c_double,
                                      ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]
        L.orc_transfer_stiffness.argtypes = [ctypes.c_int, ctypes.c_int] + \
            [ctypes.c_void_p] * 4
        L.orc_kernel_phi.argtypes = [ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_double, ctypes.c_double,
                                     ctypes.c_void_p]
        L.orc_fill_phi.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_void_p]
        L.orc_static_force.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p]
        L.orc_static_force.restype = ctypes.c_double
        _lib = L
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def fft_c2c(x, sign):
    x = _c(x, np.complex128).copy()
    lib().orc_fft_c2c(x.size, x.ctypes.data, sign)
    return x


def fft2d(x, sign):
    x = _c(x, np.complex128).copy()
    nx, ny = x.shape
    lib().orc_fft2d(nx, ny, x.ctypes.data, sign)
    return x


def invert(m):
    m = _c(m, np.complex128).copy()
    if lib().orc_invert(m.shape[0], m.ctypes.data):
        raise np.linalg.LinAlgError("singular")
    return m


def sc100_layer(qx, qy):
    U0 = np.zeros((3, 3), np.complex128)
    U = np.zeros((3, 3), np.complex128)
    V = np.zeros((3, 3), np.complex128)
    lib().orc_sc100_layer(qx, qy, U0.ctypes.data, U.ctypes.data, V.ctypes.data)
    return U0, U, V


def transfer_stiffness(height, U0, U, V):
    d = U0.shape[0]
    phi = np.zeros((d, d), np.complex128)
    U0, U, V = (_c(a, np.complex128) for a in (U0, U, V))
    if lib().orc_transfer_stiffness(d, height, U0.ctypes.data, U.ctypes.data,
                                    V.ctypes.data, phi.ctypes.data):
        raise RuntimeError("transfer matrix failed")
    return phi


def kernel_dim(kernel, params):
    return 3 * int(params[1]) if kernel == KERNEL_FCC100 else 3


def isotropic_params(nu, myu, gamma, xprd):
    """IsotropicStiffnessKernel arguments plus the domain length xprd used
    by its q = 0 entry."""
    return [nu, myu, gamma, xprd]


def fcc100_params(height, nu, k, kappa=(0.0, 0.0, 0.0), normal=1):
    k = list(k) + [0.0] * (3 - len(k))
    kappa = list(kappa) + [0.0] * (3 - len(kappa))
    return [height, nu] + k + kappa + [normal]


def kernel_phi(kernel, params, qx, qy):
    p = _c(params, np.float64)
    d = kernel_dim(kernel, params)
    phi = np.zeros((d, d), np.complex128)
    if lib().orc_kernel_phi(kernel, p.ctypes.data, qx, qy, phi.ctypes.data):
        raise RuntimeError("kernel failed")
    return phi


def fill_phi(kernel, params, d, nx, ny, normalize=True):
    p = _c(params, np.float64)
    phi = np.zeros((nx * ny, d, d), np.complex128)
    if lib().orc_fill_phi(kernel, p.ctypes.data, d, nx, ny, int(normalize),
                          phi.ctypes.data):
        raise RuntimeError("fill_phi failed (non-Hermitian or singular)")
    return phi


def static_force(phi, u, linf=None):
    """u: (d, nx, ny) float64; phi: (nx*ny, d, d) normalised full grid.
    Returns (f, epot, u0)."""
    u = _c(u, np.float64)
    d, nx, ny = u.shape
    phi = _c(phi, np.complex128)
    assert phi.shape == (nx * ny, d, d)
    f = np.zeros_like(u)
    u0 = np.zeros(d)
    lp = None
    if linf is not None:
        linf = _c(linf, np.float64)
        lp = linf.ctypes.data
    e = lib().orc_static_force(nx, ny, d, phi.ctypes.data, lp, u.ctypes.data,
                               f.ctypes.data, u0.ctypes.data)
    return f, e, u0


# ---------------------------------------------------------------------------
# fix gfmd's atom list <-> grid glue, one process: FixGFMD::pre_force
# (src/main/fix_gfmd.cpp:659-846), post_force and grid_to_list (870-1074),
# get_cm_from_xeq (1895-1926) and Group::xcm (LAMMPS: mass-weighted centre
# of the positions unwrapped with the image flags, 10 bits per dimension).
# ---------------------------------------------------------------------------
def _unwrap(x, image, box):
    if image is None:
        return x.copy()
    img = np.asarray(image, dtype=np.int64)
    b = np.stack([(img & 1023) - 512, (img >> 10 & 1023) - 512,
                  (img >> 20) - 512], axis=1)
    return x + b * np.asarray(box)[None, :]


class FixGlue:
    """State of the glue: setup(xeq, mass, mask) then pre_force /
    post_force per step; invcell is the 3 x 3 inverse surface cell."""

    def __init__(self, nx, ny, ndof, box, invcell):
        self.nx, self.ny, self.nl = nx, ny, ndof // 3
        self.box = np.asarray(box, dtype=np.float64)
        self.invcell = np.asarray(invcell, dtype=np.float64).reshape(3, 3)
        self.xshift = self.yshift = 0

    def setup(self, xeq, mass=None, mask=None):
        g = np.ones(len(xeq), bool) if mask is None else mask
        m = np.ones(len(xeq)) if mass is None else np.asarray(mass)
        self.masstotal = float(np.sum(m[g]))
        self.natoms = int(np.sum(g))
        self.xeqcm = (xeq[g] * m[g, None]).sum(axis=0) / self.masstotal
        self.xshift = self.yshift = 0

    def pre_force(self, gid, x, xeq, v=None, image=None, mass=None,
                  mask=None):
        """Returns u [ndof][nx][ny] (zeros where no atom), the velocity grid
        (or None), the number of atoms placed; gid is updated in place."""
        n = len(x)
        g = np.ones(n, bool) if mask is None else mask
        m = np.ones(n) if mass is None else np.asarray(mass)
        xu = _unwrap(x, image, self.box)
        xcm = (xu[g] * m[g, None]).sum(axis=0) / np.sum(m[g])
        d = xcm - self.xeqcm
        cx = int(np.rint(self.invcell[0] @ d))
        cy = int(np.rint(self.invcell[1] @ d))
        dxs, dys = self.xshift - cx, self.yshift - cy
        N = self.nx * self.ny
        u = np.zeros((3 * self.nl, self.nx, self.ny))
        vg = np.zeros_like(u) if v is not None else None
        placed = 0
        hx, hy = self.box[0] / 2, self.box[1] / 2
        for i in np.nonzero(g)[0]:
            ix, iy, iu = gid[i][0] - dxs, gid[i][1] - dys, gid[i][2]
            if dxs != 0 or dys != 0:
                ix %= self.nx
                iy %= self.ny
                gid[i][0], gid[i][1] = ix, iy
            if not (0 <= ix < self.nx and 0 <= iy < self.ny and
                    0 <= iu < self.nl):
                continue
            ux = x[i][0] - xeq[i][0]
            uy = x[i][1] - xeq[i][1]
            uz = x[i][2] - xeq[i][2]
            while ux > hx:
                ux -= self.box[0]
            while ux < -hx:
                ux += self.box[0]
            while uy > hy:
                uy -= self.box[1]
            while uy < -hy:
                uy += self.box[1]
            u[3 * iu, ix, iy] = ux
            u[3 * iu + 1, ix, iy] = uy
            u[3 * iu + 2, ix, iy] = uz
            if vg is not None:
                vg[3 * iu:3 * iu + 3, ix, iy] = v[i]
            placed += 1
        self.xshift, self.yshift = cx, cy
        self.xcm = xcm
        return u, vg, placed

    def post_force(self, gid, fgrid, f, fix_center=(0, 0, 0), mask=None):
        """f (n, 3) += the grid force at each atom's grid point; fix_center
        removes fsum / natoms per flagged direction.  Returns fsum."""
        g = np.ones(len(f), bool) if mask is None else mask
        fsum = np.zeros(3)
        for i in np.nonzero(g)[0]:
            ix, iy, iu = gid[i]
            if not (0 <= ix < self.nx and 0 <= iy < self.ny and
                    0 <= iu < self.nl):
                continue
            fa = fgrid[3 * iu:3 * iu + 3, ix, iy]
            f[i] += fa
            fsum += fa
        for c in range(3):
            if fix_center[c]:
                f[g, c] -= fsum[c] / self.natoms
        return fsum


# ---------------------------------------------------------------------------
# Preconditioned gradient: GFMDSolverStatic::prec_gradient
# (src/solvers/gfmd_solver_static.cpp:253-269) with
# precondition_gradient<DEF_G> (src/main/gfmd_misc.h:38-133) on the full
# complex grid: forward transform of each component (fft_forward,
# gfmd_solver_fft.cpp:37-89), (Phi(q) + cavg)^-1 times g(q) per wavevector,
# real part of the unnormalised inverse transform (fft_reverse, 91-135).
# ---------------------------------------------------------------------------
def prec_gradient(phi, g, cavg=None, full=False):
    """phi: (nx*ny, d, d) table as stored (index kx*ny + ky); g (d, nx, ny).
    The reference's general branch (d > 3) writes back only components
    0-2 of each product (gfmd_misc.h:118-121), leaving the others
    untransformed; full=True preconditions every component."""
    d, nx, ny = g.shape
    G = np.fft.fft2(g, axes=(1, 2)).reshape(d, -1).T  # (nxy, d)
    M = phi.copy()
    if cavg is not None:
        M = M + np.asarray(cavg, dtype=np.float64)[None, :, :]
    F = np.linalg.solve(M, G[:, :, None])[:, :, 0]
    if d > 3 and not full:
        F[:, 3:] = G[:, 3:]
    out = np.fft.ifft2(F.T.reshape(d, nx, ny), axes=(1, 2))
    return out.real * (nx * ny)


# ---------------------------------------------------------------------------
# FIRE state machine (fix_fire.cpp:59-306), vectorised restatement in numpy
# mirroring orc_fire_* of the C oracle.
# ---------------------------------------------------------------------------
class Fire:
    def __init__(self, dt, adapt_dt=False, minsteps=10, incfac=1.2,
                 decfac=0.5, mix_in=0.1, mixdec=0.99, dtmax=0.001,
                 xlimit=0.0):
        self.dt = self.dt0 = dt
        self.adapt_dt = adapt_dt
        self.minsteps, self.incfac, self.decfac = minsteps, incfac, decfac
        self.mix_in, self.mixdec, self.dtmax = mix_in, mixdec, dtmax
        self.xlimit = xlimit
        self.cut = minsteps
        self.mix = mix_in

    def _kick(self, v, f, mass):
        v += 0.5 * self.dt / mass * f
        if self.xlimit > 0:
            d = v.shape[0]
            vv = v.reshape(d // 3, 3, -1)
            vsq = (vv ** 2).sum(axis=1, keepdims=True)
            lim = (self.xlimit / self.dt) ** 2
            scale = np.where(vsq > lim, np.sqrt(lim / np.maximum(vsq, 1e-300)),
                             1.0)
            vv *= scale

    def initial(self, x, v, f, mass=1.0):
        self._kick(v, f, mass)
        x += self.dt * v

    def final(self, v, f, mass=1.0):
        self._kick(v, f, mass)
        vf = float(np.sum(v * f))
        vv = float(np.sum(v * v))
        ff = float(np.sum(f * f))
        help_ = self.mix * np.sqrt(vv / ff) if ff > 0 else 0.0
        v *= (1 - self.mix)
        v += help_ * f
        if vf < 0:
            v[...] = 0
            self.cut = self.minsteps
            self.dt *= self.decfac
            self.mix = self.mix_in
        else:
            if self.cut < 0:
                self.dt = min(self.dt * self.incfac, self.dtmax)
                self.mix *= self.mixdec
            else:
                self.cut -= 1
        if not self.adapt_dt:
            self.dt = self.dt0
        return vf


# ---------------------------------------------------------------------------
# LJ 12-6 wall in the distance to a sphere surface: restatement of
# FixContactSphere::post_force (src/extras/fix_contact_sphere.cpp:207-264;
# coefficients from the constructor, 124-140).  Acts on the surface layer.
# ---------------------------------------------------------------------------
def contact_sphere(u, lattice, sphere):
    """u: (3, nx, ny) surface-layer displacements; lattice (a, x0, y0, z0);
    sphere (cx, cy, cz, R, eps, sigma, cutoff).  Returns (f (3,nx,ny),
    energy, ncontact); raises if an atom is inside the sphere."""
    a, x0, y0, z0 = lattice
    cx, cy, cz, R, eps, sig, cut = sphere
    _, nx, ny = u.shape
    c1, c2 = 48 * eps * sig ** 12, 24 * eps * sig ** 6
    c3, c4 = 4 * eps * sig ** 12, 4 * eps * sig ** 6
    r6c = (1.0 / cut ** 2) ** 3
    offset = r6c * (c3 * r6c - c4)
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    rx = x0 + a * X + u[0] - cx
    ry = y0 + a * Y + u[1] - cy
    rz = z0 + u[2] - cz
    lx, ly = a * nx, a * ny
    rx = rx - lx * np.rint(rx / lx)
    ry = ry - ly * np.rint(ry / ly)
    rsq = rx * rx + ry * ry + rz * rz
    if np.any(rsq < R * R):
        raise RuntimeError("fix contact/sphere: atom inside sphere")
    m = rsq < (R + cut) ** 2
    f = np.zeros_like(u)
    r = np.sqrt(rsq[m])
    rinv = 1.0 / (r - R)
    r6 = rinv ** 6
    df = r6 * (c1 * r6 - c2) * rinv
    e = np.sum(r6 * (c3 * r6 - c4) - offset)
    f[0][m] = df * rx[m] / r
    f[1][m] = df * ry[m] / r
    f[2][m] = df * rz[m] / r
    return f, e, int(m.sum())


def fire_relax_sphere(phi, u, lattice, sphere, fire, steps, pressure=0.0):
    """Reference-order FIRE (setup force, then initial_integrate / force /
    final_integrate per step) with the elastic force of static_force plus
    the LJ sphere and a uniform normal force on the surface layer; returns
    u, v after `steps` steps."""
    u = u.copy()
    v = np.zeros_like(u)

    def force(u):
        f, e, _ = static_force(phi, u)
        fc, ec, _ = contact_sphere(u[:3], lattice, sphere)
        f[:3] += fc
        f[2] += pressure
        return f

    f = force(u)
    for _ in range(steps):
        fire.initial(u, v, f)
        f = force(u)
        fire.final(v, f)
    return u, v


def fire_relax_hardwall(phi, u, wall, pressure, fire, max_steps, tol=0.0,
                        check_every=1):
    """FIRE in the reference's order (fix_fire.cpp:117-297: setup force,
    then initial_integrate / force / final_integrate per step) with the
    elastic force of static_force, a uniform normal force `pressure` per
    surface site and a rigid hard wall u_z <= wall (surface layer):
    after each move u_z is clamped to the wall and an upward v_z there is
    zeroed; in the force, a site at the wall keeps no push into it (f_z > 0
    and v_z > 0 are zeroed before the kick).  Stops after max_steps, or at
    the first multiple of check_every where |f| / |f_ext| <= tol.
    Returns u, v, steps, relative residual."""
    u = u.copy()
    v = np.zeros_like(u)
    fext = abs(pressure) * np.sqrt(u[2].size)

    def force(u, v):
        f, _, _ = static_force(phi, u)
        f[2] += pressure
        touch = u[2] >= wall
        f[2][touch & (f[2] > 0)] = 0.0
        v[2][touch & (v[2] > 0)] = 0.0
        return f

    f = force(u, v)
    res = np.sqrt(np.sum(f * f)) / fext
    n = 0
    while n < max_steps:
        fire.initial(u, v, f)
        over = u[2] > wall
        u[2][over] = wall[over]
        v[2][over & (v[2] > 0)] = 0.0
        f = force(u, v)
        fire.final(v, f)
        n += 1
        if tol > 0 and n % check_every == 0:
            res = np.sqrt(np.sum(f * f)) / fext
            if res <= tol:
                break
    res = np.sqrt(np.sum(f * f)) / fext
    return u, v, n, res


# ---------------------------------------------------------------------------
# Reference stiffness core (compiled from the reference sources in place).
# ---------------------------------------------------------------------------
def ref_available():
    return os.path.exists(REF_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        L = ctypes.CDLL(REF_PATH)
        L.ref_kernel_create.restype = ctypes.c_void_p
        L.ref_kernel_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p,
                                        ctypes.c_double, ctypes.c_double]
        for fn in ("ref_kernel_dim", "ref_kernel_height", "ref_kernel_nu",
                   "ref_kernel_destroy"):
            getattr(L, fn).argtypes = [ctypes.c_void_p]
        L.ref_kernel_phi.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p]
        L.ref_fill_phi_buffer.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_void_p]
        L.ref_kernel_linf.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _ref = L
    return _ref


class RefKernel:
    """A reference StiffnessKernel built by stiffness_kernel_factory."""

    def __init__(self, name, args, sx=1.0, sy=1.0):
        L = ref_lib()
        self.h = L.ref_kernel_create(name.encode(), args.encode(), sx, sy)
        if not self.h:
            raise RuntimeError("unknown reference kernel %r" % name)
        self.dim = L.ref_kernel_dim(self.h)
        self.height = L.ref_kernel_height(self.h)
        self.nu = L.ref_kernel_nu(self.h)

    def phi(self, qx, qy):
        qx = _c(np.atleast_1d(qx), np.float64)
        qy = _c(np.atleast_1d(qy), np.float64)
        out = np.zeros((qx.size, self.dim, self.dim), np.complex128)
        ref_lib().ref_kernel_phi(self.h, qx.size, qx.ctypes.data,
                                 qy.ctypes.data, out.ctypes.data)
        return out

    def fill_phi(self, nx, ny, normalize=True):
        out = np.zeros((nx * ny, self.dim, self.dim), np.complex128)
        ref_lib().ref_fill_phi_buffer(self.h, self.dim, nx, ny, int(normalize),
                                      out.ctypes.data)
        return out

    def linf(self):
        out = np.zeros(self.nu)
        ref_lib().ref_kernel_linf(self.h, out.ctypes.data)
        return out

    def __del__(self):
        try:
            ref_lib().ref_kernel_destroy(self.h)
        except Exception:
            pass


class RefFT:
    """The reference's crystal surface and pair-potential force constants
    for `ft <args>` (args without "height"), exported piece by piece by
    oracle/ref_shim.cpp: cell, positions, neighbour list (indexi, indexj, x,
    y, ab, shell; r, rnorm, cell distance), off-site blocks for layers 0
    and 1, on-site blocks."""

    def __init__(self, args):
        L = ref_lib()
        L.ref_ft_create.restype = ctypes.c_void_p
        L.ref_ft_create.argtypes = [ctypes.c_char_p]
        for fn in ("ref_ft_nat", "ref_ft_nneigh", "ref_ft_destroy"):
            getattr(L, fn).argtypes = [ctypes.c_void_p]
        L.ref_ft_cell.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.ref_ft_positions.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.ref_ft_neighbors.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 4
        L.ref_ft_onsite.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                    ctypes.c_void_p]
        h = L.ref_ft_create(args.encode())
        if not h:
            raise RuntimeError("reference ft surface failed for %r" % args)
        nat, nn = L.ref_ft_nat(h), L.ref_ft_nneigh(h)
        self.nat = nat
        self.cell = np.zeros(9)
        L.ref_ft_cell(h, self.cell.ctypes.data)
        self.positions = np.zeros(3 * nat)
        L.ref_ft_positions(h, self.positions.ctypes.data)
        self.ints = np.zeros((nn, 6), np.int32)
        self.dbl = np.zeros((nn, 7))
        self.D0 = np.zeros((nn, 9))
        self.D1 = np.zeros((nn, 9))
        L.ref_ft_neighbors(h, self.ints.ctypes.data, self.dbl.ctypes.data,
                           self.D0.ctypes.data, self.D1.ctypes.data)
        self.on0 = np.zeros((nat, 9))
        self.on1 = np.zeros((nat, 9))
        L.ref_ft_onsite(h, 0, self.on0.ctypes.data)
        L.ref_ft_onsite(h, 1, self.on1.ctypes.data)
        L.ref_ft_destroy(h)


def dynamics(phi, u, v, nsteps, dt, mass=1.0, damping=0.0, pressure=0.0,
             linf=None):
    """Damped velocity Verlet on the oracle's static force (FP64):
      v(n+1/2) = v(n) + dt/2 (F(n)/m - damping v(n)); u(n+1) = u(n) + dt v(n+1/2)
      v(n+1) = (v(n+1/2) + dt/2 F(n+1)/m) / (1 + damping dt/2)
    F = elastic force + pressure on the z component of layer 0.  With
    damping = 0 this is LAMMPS fix nve (velocity Verlet), which the
    reference's MD tests run on the fix gfmd forces
    (tests/TEST_energy_conservation_fcc100/lammps.in).  Returns u, v and
    the (nsteps + 1, 3) kinetic/potential/total energies."""
    u = np.array(u, dtype=np.float64, copy=True)
    v = np.array(v, dtype=np.float64, copy=True)

    def force(x):
        f, e, _ = static_force(phi, x, linf)
        f[2] += pressure
        return f, e - pressure * x[2].sum()

    f, pe = force(u)
    hist = np.zeros((nsteps + 1, 3))
    ke = 0.5 * mass * np.sum(v * v)
    hist[0] = ke, pe, ke + pe
    for n in range(nsteps):
        vh = v + 0.5 * dt * (f / mass - damping * v)
        u = u + dt * vh
        f, pe = force(u)
        v = (vh + 0.5 * dt * f / mass) / (1.0 + 0.5 * damping * dt)
        ke = 0.5 * mass * np.sum(v * v)
        hist[n + 1] = ke, pe, ke + pe
    return u, v, hist


def hardwall_contact(phi, wall, pressure, zc=2):
    """Hard-wall contact restated as the complementarity problem of the
    contact pressure p on the surface layer, solved exactly (dense,
    non-negative least squares through a Cholesky factor): small grids only.

    phi: full-grid table (nx*ny, d, d) normalised by nx*ny (fill_phi);
    wall: (nx, ny) heights of the rigid surface (u_z <= wall); pressure: load
    per site in +z on component zc.  Equilibrium: u = G t with G = Phi^-1 and
    the traction t_zc = pressure - p; gap g = wall - u_z >= 0, p >= 0,
    p g = 0.  Returns (u (d, nx, ny), p (nx, ny))."""
    from scipy.optimize import nnls
    nx, ny = wall.shape
    n = nx * ny
    d = phi.shape[1]
    P = phi.reshape(nx, ny, d, d) * n          # physical Phi(q)
    G = np.linalg.inv(P)                        # compliance per q
    gzz = np.real(np.fft.ifft2(G[:, :, zc, zc]))  # real-space kernel
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ix, iy = ix.ravel(), iy.ravel()
    Gm = gzz[(ix[:, None] - ix[None, :]) % nx, (iy[:, None] - iy[None, :]) % ny]
    Gm = 0.5 * (Gm + Gm.T)
    c = np.real(G[0, 0, zc, zc]) * pressure     # uniform approach of the load
    h = wall.ravel() - c
    L = np.linalg.cholesky(Gm)
    p, _ = nnls(L.T, -np.linalg.solve(L, h), maxiter=50 * n)
    t = np.zeros((d, nx, ny))
    t[zc] = pressure - p.reshape(nx, ny)
    tq = np.fft.fft2(t, axes=(1, 2))
    uq = np.einsum("xyij,jxy->ixy", G, tq)
    u = np.real(np.fft.ifft2(uq, axes=(1, 2)))
    return u, p.reshape(nx, ny)


def sq_uniform_pq(G, nu, P, Qx, Qy, a, b, bx, by):
    """Li & Berger surface displacement at (bx, by) from uniform tractions
    on the rectangle [-a, a] x [-b, b] (the C restatement)."""
    u = np.zeros(3)
    lib().orc_sq_uniform_pq.argtypes = [ctypes.c_double] * 9 + [ctypes.c_void_p]
    lib().orc_sq_uniform_pq(G, nu, P, Qx, Qy, a, b, bx, by, u.ctypes.data)
    return u


def nonperiodic_green(nx, ny, nu, mu, xprd=None, yprd=None):
    """Real-space Green's function G[3][3][nx][ny] of the `nonperiodic`
    kernel: displacement at the minimum-image offset of each grid point from
    a unit force spread uniformly over one grid cell (Li & Berger), with the
    off-diagonal entries zeroed on the far edge of an even grid.  Reference:
    NonperiodicStiffnessKernel::create_and_fill_buffer
    (src/stiffness_kernels/nonperiodic_stiffness.cpp:116-196)."""
    xprd = float(nx if xprd is None else xprd)
    yprd = float(ny if yprd is None else yprd)
    dx, dy = xprd / nx, yprd / ny
    fac = 1.0 / (dx * dy)
    G = np.zeros((3, 3, nx, ny))
    for ii in range(nx):
        i = ii if ii <= nx // 2 else ii - nx
        for jj in range(ny):
            j = jj if jj <= ny // 2 else jj - ny
            # columns: response to Qx, Qy, P (load order of the reference)
            G[:, 0, ii, jj] = sq_uniform_pq(mu, nu, 0, fac, 0, dx / 2, dy / 2,
                                            i * dx, j * dy)
            G[:, 1, ii, jj] = sq_uniform_pq(mu, nu, 0, 0, fac, dx / 2, dy / 2,
                                            i * dx, j * dy)
            G[:, 2, ii, jj] = sq_uniform_pq(mu, nu, fac, 0, 0, dx / 2, dy / 2,
                                            i * dx, j * dy)
            if (nx % 2 == 0 and ii == nx // 2) or (ny % 2 == 0 and jj == ny // 2):
                for a in range(3):
                    for b in range(3):
                        if a != b:
                            G[a, b, ii, jj] = 0.0
    return G


def nonperiodic_phi(nx, ny, nu, mu, normalize=True, xprd=None, yprd=None):
    """Full-grid table (nx*ny, 3, 3) of the `nonperiodic` kernel: the
    Green's function transformed with exponent +1 (unnormalised, as FFTW's
    backward transform) and inverted per wavevector (nonperiodic_
    stiffness.cpp:198-270), divided by nx*ny when normalising (fill_phi_
    buffer)."""
    G = nonperiodic_green(nx, ny, nu, mu, xprd, yprd)
    Gq = np.fft.ifft2(G, axes=(2, 3)) * (nx * ny)
    Gq = np.moveaxis(Gq, (2, 3), (0, 1)).reshape(nx * ny, 3, 3)
    phi = np.linalg.inv(Gq)
    if normalize:
        phi = phi / (nx * ny)
    return phi


class RefTable2D:
    """The reference's Table2D (periodic bicubic spline of fix wall/map)
    on values[nx][ny]: eval(x, y) -> (n, 6) f, df/dx, df/dy, d2f/dx2,
    d2f/dy2, d2f/dxdy."""

    def __init__(self, values, interp=True, lowmem=False):
        self.v = _c(values, np.float64)
        self.interp, self.lowmem = interp, lowmem
        L = ref_lib()
        L.ref_table2d_eval.argtypes = [ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p]

    def eval(self, x, y):
        x = _c(np.atleast_1d(x), np.float64)
        y = _c(np.atleast_1d(y), np.float64)
        out = np.zeros((x.size, 6))
        nx, ny = self.v.shape
        ref_lib().ref_table2d_eval(nx, ny, self.v.ctypes.data,
                                   int(self.interp), int(self.lowmem), x.size,
                                   x.ctypes.data, y.ctypes.data,
                                   out.ctypes.data)
        return out
