"""Python binding of libgfmd_hip.so (the C ABI in include/gfmd_hip.h).

Mirrors the reference's solver interface (src/main/gfmd_solver.h:34-123):
``Solver(nx, ny, ndof)`` plays GFMDSolver::set_grid_size, ``set_kernel``
plays GFMDSolverStatic::set_kernel, ``post_force`` plays
GFMDSolverStatic::post_force and ``get_u0`` GFMDSolver::get_u0.

There is no CPU fallback: if the HIP library is missing the import of the
library fails with an error.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG_ROOT, "libgfmd_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_ROOT), "include", "gfmd_hip.h")

OK = 0
ERR_CONTACT = 7
ERR_ARG = 1
ERR_HIP = 2
ERR_STATE = 3
ERR_UNSUPPORTED = 4
ERR_NONFINITE = 5
ERR_KERNEL = 6

FP64 = 0
FP32 = 1

_lib = None


CONTACT_NONE = 0
CONTACT_LJ_SPHERE = 1
CONTACT_HARDWALL = 2
CONTACT_WALL_MAP = 3


class FireParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("dt", "dtmax", "incfac", "decfac", "mix_in", "mixdec",
                 "xlimit", "mass")] + [("minsteps", ctypes.c_int),
                                       ("adapt_dt", ctypes.c_int)]


class Contact(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("pressure", ctypes.c_double),
                ("lattice", ctypes.c_double * 4),
                ("sphere", ctypes.c_double * 7),
                ("wall", ctypes.POINTER(ctypes.c_double)),
                ("wall_offset", ctypes.c_double),
                ("cell_xy", ctypes.c_double * 2),
                ("contact_atoms", ctypes.c_int),
                ("atom_offset", (ctypes.c_double * 3) * 4),
                ("wall_map", ctypes.c_void_p)]


class Atoms(ctypes.Structure):
    """gfmd_atoms_t: device pointers (ints) of a fix gfmd atom list."""
    _fields_ = [("n", ctypes.c_int), ("mask", ctypes.c_void_p),
                ("groupbit", ctypes.c_int), ("gid", ctypes.c_void_p),
                ("x", ctypes.c_void_p), ("xeq", ctypes.c_void_p),
                ("v", ctypes.c_void_p), ("image", ctypes.c_void_p),
                ("mass", ctypes.c_void_p)]


class FixState(ctypes.Structure):
    _fields_ = [("xeqcm", ctypes.c_double * 3), ("xcm", ctypes.c_double * 3),
                ("xshift", ctypes.c_int), ("yshift", ctypes.c_int),
                ("placed", ctypes.c_int), ("unplaced", ctypes.c_int),
                ("fsum", ctypes.c_double * 3), ("unforced", ctypes.c_int),
                ("masstotal", ctypes.c_double), ("natoms", ctypes.c_int)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        for k in ("xeqcm", "xcm", "fsum"):
            d[k] = list(d[k])
        return d


class RelaxResult(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int), ("converged", ctypes.c_int),
                ("residual", ctypes.c_double), ("fext_norm", ctypes.c_double),
                ("energy", ctypes.c_double),
                ("contact_energy", ctypes.c_double),
                ("ncontact", ctypes.c_double), ("seconds", ctypes.c_double),
                ("device_seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class DynParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("mass", ctypes.c_double),
                ("damping", ctypes.c_double), ("record_every", ctypes.c_int)]


class DynResult(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int), ("ekin", ctypes.c_double),
                ("epot", ctypes.c_double), ("ncontact", ctypes.c_int),
                ("seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def fire_params(**kw):
    """FIRE parameters: the reference defaults, overridden by kw."""
    p = FireParams()
    load().gfmd_fire_defaults(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def contact(type=CONTACT_NONE, pressure=0.0, lattice=(1.0, 0.0, 0.0, 0.0),
            sphere=None, wall=None, wall_offset=0.0, cell_xy=None,
            atom_offsets=None, wall_map=None):
    """Contact description (gfmd_contact_t).  For cells with several atoms
    in contact (LJ sphere, wall map), cell_xy = (cell length x, cell length
    y) and atom_offsets = [(dx, dy, dz), ...] of atoms 0, 1, ... from the
    site's origin.  wall_map: a WallMap (type CONTACT_WALL_MAP)."""
    c = Contact()
    c._wall_map = wall_map
    if wall_map is not None:
        c.wall_map = wall_map.handle
    if cell_xy is not None:
        c.cell_xy[0], c.cell_xy[1] = cell_xy
    if atom_offsets is not None:
        c.contact_atoms = len(atom_offsets)
        for l, off in enumerate(atom_offsets):
            for k in range(3):
                c.atom_offset[l][k] = off[k]
    c.type = type
    c.pressure = pressure
    for i, x in enumerate(lattice):
        c.lattice[i] = x
    if sphere is not None:
        for i, x in enumerate(sphere):
            c.sphere[i] = x
    c._wall = None
    if wall is not None:
        c._wall = np.ascontiguousarray(wall, dtype=np.float64)
        c.wall = c._wall.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    c.wall_offset = wall_offset
    return c


class GFMDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gfmd status %d: %s" % (code, msg))
        self.code = code


def load():
    """Load libgfmd_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    # A process must hold one HIP runtime.  PyTorch ships its own
    # libamdhip64 (same SONAME, loaded by path); importing it first makes
    # this library bind to that copy instead of pulling in a second runtime
    # that would hide the GPU from torch.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError("libgfmd_hip.so not built at %s (run "
                          "__graft_entry__.build() or make -C user-gfmd_amd)"
                          % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.gfmd_last_error.restype = ctypes.c_char_p
    L.gfmd_version.restype = ctypes.c_char_p
    L.gfmd_create.argtypes = [ip, ip, ip, ip, ctypes.POINTER(vp)]
    L.gfmd_destroy.argtypes = [vp]
    L.gfmd_set_kernel.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, ip]
    L.gfmd_set_phi.argtypes = [vp, vp, ip]
    L.gfmd_set_linf.argtypes = [vp, vp]
    L.gfmd_phi_mirror.argtypes = [vp, ctypes.POINTER(ip)]
    L.gfmd_phi_entries.argtypes = [vp, ctypes.POINTER(ip)]
    L.gfmd_ft_cell.argtypes = [ctypes.c_char_p, ctypes.POINTER(ip), vp, vp,
                               vp, ip]
    L.gfmd_post_force.argtypes = [vp, vp, vp, ctypes.POINTER(dp)]
    L.gfmd_post_force_device.argtypes = [vp, vp, vp, vp]
    L.gfmd_get_u0.argtypes = [vp, vp]
    L.gfmd_stream.argtypes = [vp]
    L.gfmd_stream.restype = vp
    L.gfmd_synchronize.argtypes = [vp]
    L.gfmd_get_info.argtypes = [vp] + [ctypes.POINTER(ip)] * 5
    L.gfmd_memory_usage.argtypes = [vp]
    L.gfmd_memory_usage.restype = ctypes.c_size_t
    L.gfmd_dump_stiffness.argtypes = [vp, vp]
    L.gfmd_stiffness_at.argtypes = [vp, ip, vp, vp, vp]
    L.gfmd_post_force_noalloc.argtypes = [ip, ip, ip, ip, vp, ip, vp, vp, vp,
                                          vp, vp, vp]
    L.gfmd_time_force_device.argtypes = [vp, vp, vp, vp, ip,
                                         ctypes.POINTER(dp)]
    L.gfmd_time_force_graph.argtypes = [vp, vp, vp, vp, ip, ip,
                                        ctypes.POINTER(dp),
                                        ctypes.POINTER(dp)]
    L.gfmd_autotune.argtypes = [vp, ip, ctypes.POINTER(ip),
                                ctypes.POINTER(ip), ctypes.POINTER(dp)]
    L.gfmd_set_variant.argtypes = [vp, ip, ip]
    L.gfmd_fire_defaults.argtypes = [vp]
    L.gfmd_fire_defaults.restype = None
    L.gfmd_relax.argtypes = [vp, vp, vp, vp, ip, dp, ip, vp]
    L.gfmd_prec_gradient.argtypes = [vp, vp, vp, vp]
    L.gfmd_fix_setup.argtypes = [vp, vp, vp, vp]
    L.gfmd_fix_pre_force.argtypes = [vp, vp, vp, vp]
    L.gfmd_fix_post_force.argtypes = [vp, vp, vp, vp, vp]
    L.gfmd_fix_state.argtypes = [vp, vp]
    L.gfmd_prec_gradient_device.argtypes = [vp, vp, vp, vp]
    L.gfmd_wall_map_create.argtypes = [ctypes.c_char_p, vp, ip, ip, dp, dp,
                                       ctypes.POINTER(vp)]
    L.gfmd_wall_map_destroy.argtypes = [vp]
    L.gfmd_wall_map_info.argtypes = [vp, vp]
    L.gfmd_wall_map_force.argtypes = [vp, ip, vp, vp, vp]
    L.gfmd_dynamics.argtypes = [vp, vp, vp, vp, vp, ip, vp, vp]
    L.gfmd_contact_solve.argtypes = [vp, vp, vp, vp, vp, ip, dp, vp, vp]
    L.gfmd_contact_solve_batch.argtypes = [ip, vp, vp, vp, vp, ip, dp, vp]
    L.gfmd_fno_param_count.restype = ctypes.c_size_t
    L.gfmd_fno_create.argtypes = [ip, ip, ip, ctypes.POINTER(vp)]
    L.gfmd_fno_destroy.argtypes = [vp]
    L.gfmd_fno_set_params.argtypes = [vp, vp]
    L.gfmd_fno_get_params.argtypes = [vp, vp]
    L.gfmd_fno_forward.argtypes = [vp, ip, vp, vp]
    L.gfmd_fno_forward_device.argtypes = [vp, ip, vp, vp]
    L.gfmd_fno_gradients.argtypes = [vp, ip, vp, vp, vp, vp]
    L.gfmd_fno_train_step.argtypes = [vp, ip, vp, vp, ctypes.c_float, vp]
    L.gfmd_fno_synchronize.argtypes = [vp]
    L.gfmd_fno_gemm_test.argtypes = [ip, ip, ip, ip, vp, vp, vp, vp, vp, vp,
                                     ip]
    L.gfmd_ensemble_create.argtypes = [vp, vp, ip, ctypes.POINTER(vp)]
    L.gfmd_ensemble_destroy.argtypes = [vp]
    L.gfmd_ensemble_contact_solve.argtypes = [ip, vp, vp, vp, ip, dp, vp]
    L.gfmd_rough_surface.argtypes = [ip, ip, dp, dp, dp, dp, dp,
                                     ctypes.c_ulonglong, vp]
    L.gfmd_relax_refine.argtypes = [vp, vp, vp, vp, vp, ip, dp, dp, ip, ip,
                                    vp, ctypes.POINTER(ip)]
    _lib = L
    return L


def post_force_noalloc(nx, ny, ndof, precision, phi_ptr, phi_mirror,
                       linf_ptr, u_ptr, f_ptr, epot_ptr, u0_ptr=None,
                       stream=None):
    """gfmd_post_force_noalloc: one evaluation on caller-owned device
    buffers (table, u, f, energy), no handle and no device allocation."""
    check(load().gfmd_post_force_noalloc(
        int(nx), int(ny), int(ndof), int(precision), ctypes.c_void_p(phi_ptr),
        int(bool(phi_mirror)), ctypes.c_void_p(linf_ptr),
        ctypes.c_void_p(u_ptr), ctypes.c_void_p(f_ptr),
        ctypes.c_void_p(epot_ptr), ctypes.c_void_p(u0_ptr),
        ctypes.c_void_p(stream)))


def check(rc):
    if rc != OK:
        raise GFMDError(rc, load().gfmd_last_error().decode())


def header_functions():
    """Names of the functions declared in include/gfmd_hip.h."""
    import re
    src = open(HEADER_PATH).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gfmd_[a-z0-9_]+)\s*\(", src)))


def ft_cell(args):
    """Surface cell of `ft <args>` (host only): {'nat', 'cell' (3, 3) with
    the lattice vectors as columns, 'positions' (nat, 3), 'linf' (nat)}."""
    import numpy as np
    L = load()
    nat = ctypes.c_int()
    check(L.gfmd_ft_cell(args.encode(), ctypes.byref(nat), None, None, None,
                         0))
    cell = np.zeros(9)
    pos = np.zeros(3 * nat.value)
    linf = np.zeros(nat.value)
    check(L.gfmd_ft_cell(args.encode(), ctypes.byref(nat), cell.ctypes.data,
                         pos.ctypes.data, linf.ctypes.data, nat.value))
    return {"nat": nat.value, "cell": cell.reshape(3, 3),
            "positions": pos.reshape(-1, 3), "linf": linf}


class WallMap:
    """`fix wall/map` set up on the device (gfmd_wall_map_create): args are
    the fix's arguments after `wall/map`; heights (nx, ny) replace the map
    file named in args; xprd, yprd: the periodic box."""

    def __init__(self, args, xprd, yprd, heights=None):
        L = load()
        self._h = ctypes.c_void_p()
        mp, mx, my = None, 0, 0
        if heights is not None:
            self._heights = np.ascontiguousarray(heights, dtype=np.float64)
            mx, my = self._heights.shape
            mp = self._heights.ctypes.data
        check(L.gfmd_wall_map_create(args.encode(), mp, mx, my, float(xprd),
                                     float(yprd), ctypes.byref(self._h)))
        self.xprd, self.yprd = xprd, yprd

    @property
    def handle(self):
        return self._h.value

    def info(self):
        """{'min', 'max', 'rms', 'average', 'nx', 'ny'} of the shifted map."""
        out = np.zeros(6)
        check(load().gfmd_wall_map_info(self._h, out.ctypes.data))
        return dict(zip(("min", "max", "rms", "average", "nx", "ny"), out))

    def force(self, x, f=None):
        """pre_force on atoms x (n, 3): returns (x after the pendist rule,
        f + wall force, fix vector (energy, fx, fy, fz, ncon, nrep, natt))."""
        x = np.array(x, dtype=np.float64, order="C", copy=True)
        n = x.shape[0]
        f = np.zeros((n, 3)) if f is None else \
            np.array(f, dtype=np.float64, order="C", copy=True)
        fmap = np.zeros(7)
        check(load().gfmd_wall_map_force(self._h, n, x.ctypes.data,
                                         f.ctypes.data, fmap.ctypes.data))
        return x, f, fmap

    def close(self):
        if self._h:
            load().gfmd_wall_map_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rough_surface(nx, ny, hurst=0.8, rms_slope=0.1, lambda_short=4.0,
                  lambda_long=0.0, a=1.0, seed=1):
    """Seeded self-affine rough surface (nx, ny) generated on the device."""
    out = np.zeros((nx, ny))
    check(load().gfmd_rough_surface(nx, ny, hurst, rms_slope, lambda_short,
                                    lambda_long, a, seed, out.ctypes.data))
    return out


def contact_solve_batch(solvers64, contacts, solvers32=None, max_iter=20000,
                        tol=1e-8):
    """Ensemble of hard-wall contacts in one graph per iteration.  Returns a
    list of (u, result dict)."""
    k = len(solvers64)
    assert len(contacts) == k and (solvers32 is None or len(solvers32) == k)
    h64 = (ctypes.c_void_p * k)(*[s.handle for s in solvers64])
    h32 = None
    if solvers32 is not None:
        h32 = (ctypes.c_void_p * k)(*[s.handle for s in solvers32])
    cs = (Contact * k)(*contacts)
    us = [np.zeros((s.ndof, s.nx, s.ny)) for s in solvers64]
    up = (ctypes.c_void_p * k)(*[x.ctypes.data for x in us])
    res = (RelaxResult * k)()
    check(load().gfmd_contact_solve_batch(k, h64, h32, cs, up, int(max_iter),
                                          float(tol), res))
    return [(us[i], res[i].as_dict()) for i in range(k)]


class Ensemble:
    """`members` surfaces of the grid of s64 sharing its stiffness (s32: an
    FP32 handle of the same grid for the FP32 stage, or None)."""

    def __init__(self, s64, members, s32=None):
        self.s64, self.s32, self.members = s64, s32, int(members)
        h = ctypes.c_void_p()
        check(load().gfmd_ensemble_create(
            s64.handle, s32.handle if s32 is not None else None,
            self.members, ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            load().gfmd_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ensemble_contact_solve(ensembles, contacts, max_iter=20000, tol=1e-8):
    """All members of all ensembles in one graph per iteration.  contacts:
    flat list over the ensembles' members.  Returns a list of (u, result)."""
    k = len(ensembles)
    total = sum(e.members for e in ensembles)
    assert len(contacts) == total
    eh = (ctypes.c_void_p * k)(*[e.handle for e in ensembles])
    cs = (Contact * total)(*contacts)
    us = []
    for e in ensembles:
        us += [np.zeros((e.s64.ndof, e.s64.nx, e.s64.ny))
               for _ in range(e.members)]
    up = (ctypes.c_void_p * total)(*[x.ctypes.data for x in us])
    res = (RelaxResult * total)()
    check(load().gfmd_ensemble_contact_solve(k, eh, cs, up, int(max_iter),
                                             float(tol), res))
    return [(us[i], res[i].as_dict()) for i in range(total)]


class FNO:
    """The learned warm start's network on the device (fixed architecture:
    32 channels, 4 spectral layers, modes kx in [-16, 16), ky in [0, 16]).
    Arrays are channels-last float32: inputs (B, nx, ny, 2), outputs
    (B, nx, ny, 3)."""

    def __init__(self, nx, ny, max_batch):
        self.nx, self.ny, self.max_batch = int(nx), int(ny), int(max_batch)
        h = ctypes.c_void_p()
        check(load().gfmd_fno_create(self.nx, self.ny, self.max_batch,
                                     ctypes.byref(h)))
        self._h = h

    @staticmethod
    def param_count():
        return int(load().gfmd_fno_param_count())

    def set_params(self, vec):
        vec = np.ascontiguousarray(vec, dtype=np.float32)
        assert vec.size == self.param_count()
        check(load().gfmd_fno_set_params(self._h, vec.ctypes.data))

    def get_params(self):
        out = np.zeros(self.param_count(), dtype=np.float32)
        check(load().gfmd_fno_get_params(self._h, out.ctypes.data))
        return out

    def _in(self, a, ch):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.shape[1:] == (self.nx, self.ny, ch) and a.shape[0] <= \
            self.max_batch
        return a

    def forward(self, a):
        a = self._in(a, 2)
        out = np.zeros(a.shape[:3] + (3,), dtype=np.float32)
        check(load().gfmd_fno_forward(self._h, a.shape[0], a.ctypes.data,
                                      out.ctypes.data))
        return out

    def gradients(self, a, t):
        a, t = self._in(a, 2), self._in(t, 3)
        loss = ctypes.c_double()
        g = np.zeros(self.param_count(), dtype=np.float32)
        check(load().gfmd_fno_gradients(self._h, a.shape[0], a.ctypes.data,
                                        t.ctypes.data, ctypes.byref(loss),
                                        g.ctypes.data))
        return loss.value, g

    def train_step(self, a, t, lr):
        a, t = self._in(a, 2), self._in(t, 3)
        loss = ctypes.c_double()
        check(load().gfmd_fno_train_step(self._h, a.shape[0], a.ctypes.data,
                                         t.ctypes.data, float(lr),
                                         ctypes.byref(loss)))
        return loss.value

    def train_step_device(self, a_ptr, t_ptr, batch, lr):
        """Device pointers; asynchronous (no loss returned)."""
        check(load().gfmd_fno_train_step(self._h, int(batch),
                                         ctypes.c_void_p(a_ptr),
                                         ctypes.c_void_p(t_ptr), float(lr),
                                         None))

    def forward_device(self, a_ptr, out_ptr, batch):
        check(load().gfmd_fno_forward_device(self._h, int(batch),
                                             ctypes.c_void_p(a_ptr),
                                             ctypes.c_void_p(out_ptr)))

    def synchronize(self):
        check(load().gfmd_fno_synchronize(self._h))

    def close(self):
        if self._h:
            load().gfmd_fno_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """One GFMD surface on the GPU (an nx x ny grid, ndof dofs per point)."""

    def __init__(self, nx, ny, ndof=3, precision=FP64):
        L = load()
        h = ctypes.c_void_p()
        check(L.gfmd_create(nx, ny, ndof, precision, ctypes.byref(h)))
        self._h = h
        self.nx, self.ny, self.ndof, self.precision = nx, ny, ndof, precision

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h is not None and self._h.value:
            load().gfmd_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        v = [ctypes.c_int() for _ in range(5)]
        check(load().gfmd_get_info(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(["nx", "ny", "ndof", "precision", "fast_path"],
                        [x.value for x in v]))

    def memory_usage(self):
        return load().gfmd_memory_usage(self._h)

    def set_kernel(self, name, args="", normalize=True):
        check(load().gfmd_set_kernel(self._h, name.encode(), args.encode(),
                                     int(normalize)))

    def set_phi(self, phi, normalized=True):
        """phi: (nx*ny, ndof, ndof) complex128 full grid (fill_phi_buffer
        layout)."""
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        assert phi.shape == (self.nx * self.ny, self.ndof, self.ndof)
        check(load().gfmd_set_phi(self._h, phi.ctypes.data, int(normalized)))

    def phi_mirror(self):
        """True when the stored table is symmetric under x -> -x (the column
        sweep then reads half of it)."""
        m = ctypes.c_int()
        check(load().gfmd_phi_mirror(self._h, ctypes.byref(m)))
        return bool(m.value)

    def phi_entries(self):
        """Reals per wavevector the force sweep reads from the table (6 for
        the compact ndof-3 form, else ndof^2)."""
        m = ctypes.c_int()
        check(load().gfmd_phi_entries(self._h, ctypes.byref(m)))
        return m.value

    def dump_stiffness(self):
        """Device Phi table as (ny//2+1, ndof*ndof, nx) float64 (packed)."""
        out = np.zeros((self.ny // 2 + 1, self.ndof * self.ndof, self.nx))
        check(load().gfmd_dump_stiffness(self._h, out.ctypes.data))
        return out

    def stiffness_at(self, kx, ky):
        """Stored Phi at wavevectors (kx[i], ky[i]) (0 <= ky <= ny/2) as
        (count, ndof, ndof) complex128, scaled as stored."""
        kx = np.ascontiguousarray(kx, dtype=np.int32)
        ky = np.ascontiguousarray(ky, dtype=np.int32)
        d = self.ndof
        out = np.zeros((len(kx), d, d), np.complex128)
        check(load().gfmd_stiffness_at(self._h, len(kx), kx.ctypes.data,
                                       ky.ctypes.data, out.ctypes.data))
        return out

    def time_force_device(self, u_ptr, f_ptr, epot_ptr, iters):
        """Mean per-sweep durations [ms]: rows r2c, columns, rows c2r,
        whole evaluation (HIP events on the handle's stream)."""
        ms = (ctypes.c_double * 4)()
        check(load().gfmd_time_force_device(
            self._h, ctypes.c_void_p(u_ptr), ctypes.c_void_p(f_ptr),
            ctypes.c_void_p(epot_ptr), int(iters), ms))
        return list(ms)

    def time_force_graph(self, u_ptr, f_ptr, epot_ptr, iters, per_graph=1):
        """(ms per evaluation eager, ms per evaluation as graph replays)."""
        me, mg = ctypes.c_double(), ctypes.c_double()
        check(load().gfmd_time_force_graph(
            self._h, u_ptr, f_ptr, epot_ptr, int(iters), int(per_graph),
            ctypes.byref(me), ctypes.byref(mg)))
        return me.value, mg.value

    def autotune(self, iters=20):
        """Pick the fastest fast-path variants; returns (row, col, ms)."""
        r, c, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(load().gfmd_autotune(self._h, int(iters), ctypes.byref(r),
                                   ctypes.byref(c), ctypes.byref(ms)))
        return r.value, c.value, ms.value

    def set_variant(self, row_var, col_var):
        check(load().gfmd_set_variant(self._h, int(row_var), int(col_var)))

    def prec_gradient(self, g, cavg=None):
        """gP = F^-1[(Phi~(q) + cavg)^-1 F[g]] for g (ndof, nx, ny)."""
        g = np.ascontiguousarray(g, dtype=np.float64)
        assert g.shape == (self.ndof, self.nx, self.ny)
        gp = np.empty_like(g)
        cp = None
        if cavg is not None:
            self._cavg = np.ascontiguousarray(cavg, dtype=np.float64)
            assert self._cavg.shape == (self.ndof, self.ndof)
            cp = self._cavg.ctypes.data
        check(load().gfmd_prec_gradient(self._h, cp, g.ctypes.data,
                                        gp.ctypes.data))
        return gp

    def relax(self, u, contact=None, fire=None, max_steps=100000, tol=1e-10,
              check_every=50):
        """FIRE relaxation of u (ndof, nx, ny) float64 (returned updated
        copy) -> (u, result dict)."""
        u = np.array(u, dtype=np.float64, order="C", copy=True)
        assert u.shape == (self.ndof, self.nx, self.ny)
        fire = fire if fire is not None else fire_params()
        res = RelaxResult()
        cptr = ctypes.byref(contact) if contact is not None else None
        check(load().gfmd_relax(self._h, cptr, ctypes.byref(fire),
                                u.ctypes.data, int(max_steps), float(tol),
                                int(check_every), ctypes.byref(res)))
        return u, res.as_dict()

    def relax_refine(self, s32, u, contact=None, fire=None, max_outer=20,
                     tol=1e-10, inner_tol=1e-4, inner_max_steps=100000,
                     check_every=50):
        """Mixed precision: this (FP64) handle runs the outer corrections,
        s32 (an FP32 Solver of the same grid) the inner FIRE solves.
        Returns (u, result dict with 'outer')."""
        u = np.array(u, dtype=np.float64, order="C", copy=True)
        fire = fire if fire is not None else fire_params()
        res = RelaxResult()
        outer = ctypes.c_int()
        cptr = ctypes.byref(contact) if contact is not None else None
        check(load().gfmd_relax_refine(
            self._h, s32.handle, cptr, ctypes.byref(fire), u.ctypes.data,
            int(max_outer), float(tol), float(inner_tol),
            int(inner_max_steps), int(check_every), ctypes.byref(res),
            ctypes.byref(outer)))
        d = res.as_dict()
        d["outer"] = outer.value
        return u, d

    def dynamics(self, u, v, nsteps, dt, mass=1.0, damping=0.0,
                 record_every=1, contact=None):
        """Damped velocity Verlet from (u, v), each (ndof, nx, ny) float64.
        Returns (u, v, history (nrec, 3): kinetic, potential, total energy
        every record_every states, result dict)."""
        u = np.array(u, dtype=np.float64, order="C", copy=True)
        v = np.array(v, dtype=np.float64, order="C", copy=True)
        assert u.shape == v.shape == (self.ndof, self.nx, self.ny)
        prm = DynParams(float(dt), float(mass), float(damping),
                        int(record_every))
        hist = np.zeros((int(nsteps) // int(record_every) + 1, 3))
        res = DynResult()
        cptr = ctypes.byref(contact) if contact is not None else None
        check(load().gfmd_dynamics(self._h, cptr, ctypes.byref(prm),
                                   u.ctypes.data, v.ctypes.data, int(nsteps),
                                   hist.ctypes.data, ctypes.byref(res)))
        return u, v, hist, res.as_dict()

    def contact_solve(self, contact, s32=None, p_init=None, max_iter=20000,
                      tol=1e-8):
        """Hard-wall contact by constrained CG (this handle FP64, s32 an
        optional FP32 handle of the same grid).  Returns (u, result dict
        with 'iters32', 'iters64')."""
        u = np.zeros((self.ndof, self.nx, self.ny))
        res = RelaxResult()
        its = (ctypes.c_int * 2)()
        pin = None
        if p_init is not None:
            p_init = np.ascontiguousarray(p_init, dtype=np.float64)
            assert p_init.shape == (self.nx, self.ny)
            pin = p_init.ctypes.data
        check(load().gfmd_contact_solve(
            self._h, s32.handle if s32 is not None else None,
            ctypes.byref(contact), u.ctypes.data, pin, int(max_iter),
            float(tol), ctypes.byref(res), its))
        d = res.as_dict()
        d["iters32"], d["iters64"] = its[0], its[1]
        return u, d

    def set_linf(self, linf):
        linf = np.ascontiguousarray(linf, dtype=np.float64)
        assert linf.shape == (self.ndof // 3,)
        check(load().gfmd_set_linf(self._h, linf.ctypes.data))

    def post_force(self, u):
        """u: (ndof, nx, ny) float64 host array -> (f, epot)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.shape == (self.ndof, self.nx, self.ny)
        f = np.empty_like(u)
        e = ctypes.c_double()
        check(load().gfmd_post_force(self._h, u.ctypes.data, f.ctypes.data,
                                     ctypes.byref(e)))
        return f, e.value

    def post_force_device(self, u_ptr, f_ptr, epot_ptr):
        """Device pointers (ints); asynchronous on the handle's stream."""
        check(load().gfmd_post_force_device(self._h, ctypes.c_void_p(u_ptr),
                                            ctypes.c_void_p(f_ptr),
                                            ctypes.c_void_p(epot_ptr)))

    def prec_gradient_device(self, g_ptr, gp_ptr, cavg=None):
        """Device pointers (ints) of the handle's I/O type; asynchronous
        once the table for this cavg exists."""
        cp = None
        if cavg is not None:
            self._cavg = np.ascontiguousarray(cavg, dtype=np.float64)
            cp = self._cavg.ctypes.data
        check(load().gfmd_prec_gradient_device(
            self._h, cp, ctypes.c_void_p(g_ptr), ctypes.c_void_p(gp_ptr)))

    def fix_setup(self, atoms, box, invcell):
        """atoms: Atoms (device pointers); box (3,); invcell (3, 3)."""
        b = np.ascontiguousarray(box, dtype=np.float64)
        c = np.ascontiguousarray(invcell, dtype=np.float64).reshape(9)
        check(load().gfmd_fix_setup(self._h, ctypes.byref(atoms),
                                    b.ctypes.data, c.ctypes.data))

    def fix_pre_force(self, atoms, u_ptr, vgrid_ptr=None):
        check(load().gfmd_fix_pre_force(self._h, ctypes.byref(atoms),
                                        ctypes.c_void_p(u_ptr),
                                        ctypes.c_void_p(vgrid_ptr)))

    def fix_post_force(self, atoms, fgrid_ptr, fatoms_ptr,
                       fix_center=(0, 0, 0)):
        fc = (ctypes.c_int * 3)(*[int(v) for v in fix_center])
        check(load().gfmd_fix_post_force(self._h, ctypes.byref(atoms),
                                         ctypes.c_void_p(fgrid_ptr),
                                         ctypes.c_void_p(fatoms_ptr), fc))

    def fix_state(self):
        st = FixState()
        check(load().gfmd_fix_state(self._h, ctypes.byref(st)))
        return st.as_dict()

    def get_u0(self):
        u0 = np.zeros(self.ndof)
        check(load().gfmd_get_u0(self._h, u0.ctypes.data))
        return u0

    def stream(self):
        return load().gfmd_stream(self._h)

    def synchronize(self):
        check(load().gfmd_synchronize(self._h))
