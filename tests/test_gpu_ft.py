"""`ft` stiffness kernel on the device: host tables (surface cell,
neighbour list, force constants), Phi(q) by the device evaluator and the
transfer matrix, forces through the fast (ndof 3, 6) and generic (ndof 9,
12) paths, against the CPU FP64 oracle (oracle/ft.py, pinned to the
reference's crystal surface and force constants and to its sc100 / fcc100
kernels).  Cases include the kernels of the reference's Hertz tests on
fcc111 (TEST_Hertz_fcc111_64x37_2: `ft fcc111 1 1.0 pair-potential 1 1.0
height 128`; TEST_Hertz_fcc111_two_layers_64x37: `ft fcc111 1.0 2 ...`) on
their 64 x 37 grid.

Tolerances: Phi table and FP64 forces 1e-11 relative (max-norm), energy
1e-11; FP32 forces 2e-5 relative to the FP64 path, energy 1e-5.
"""
import numpy as np
import pytest

import ft
import gfmd_hip
import oracle as O
from test_gpu_fcc100 import packed

pytestmark = pytest.mark.gpu

LJ_SIGMA = 2.0 ** (-1.0 / 6.0)  # nearest-neighbour distance 1 at the minimum

CASES = [
    ("sc100 1.0 1 pair-potential 2 1.0 1.0 height 16", 16, 12),
    ("sc100 1.3 2 pair-potential 2 1.0 1.0 height 4", 12, 10),
    ("fcc100 1.0 2 pair-potential 2x 1.0 0.1 0.5 0.05 height 6", 16, 16),
    ("fcc100 1.0 3 pair-potential 2 1.0 0.5 height 2", 8, 6),
    ("fcc111 1 1.0 pair-potential 1 1.0 height 128", 16, 9),
    ("fcc111 1.0 2 pair-potential 1 1.0 height 128", 16, 9),
    ("fcc111 1.2 2 pair-potential 3x 1.0 0.1 0.5 0.2 0.3 0.05 height 0", 8, 8),
    ("fcc100 1.0 1 lj/cut 1.0 %r 1.5 height 8" % LJ_SIGMA, 12, 12),
    ("fcc111 1.0 1 lj/cut 1.0 %r 1.3 height 3" % LJ_SIGMA, 10, 10),
]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("args,nx,ny", CASES)
def test_table_matches_oracle(args, nx, ny):
    k = ft.FTKernel(args)
    s = gfmd_hip.Solver(nx, ny, k.dim, gfmd_hip.FP64)
    s.set_kernel("ft", args)
    ref = packed(k.fill_phi(nx, ny), nx, ny, k.dim)
    assert rel(s.dump_stiffness(), ref) < 1e-11


@pytest.mark.parametrize("args,nx,ny", [
    ("fcc111 1 1.0 pair-potential 1 1.0 height 128", 64, 37),
    ("fcc111 1.0 2 pair-potential 1 1.0 height 128", 64, 37),
    ("fcc111 1 1.0 pair-potential 1 1.0 height 64", 64, 64),
    ("fcc100 1.0 3 pair-potential 2 1.0 0.5 height 5", 32, 24),
    ("fcc100 1.0 1 lj/cut 1.0 %r 1.5 height 8" % LJ_SIGMA, 32, 32)])
def test_forces_match_oracle(args, nx, ny):
    k = ft.FTKernel(args)
    phi = k.fill_phi(nx, ny)
    s = gfmd_hip.Solver(nx, ny, k.dim, gfmd_hip.FP64)
    s.set_kernel("ft", args)
    u = 0.01 * np.random.default_rng(3).standard_normal((k.dim, nx, ny))
    linf = k.linf if np.any(k.linf) else None
    f_ref, e_ref, _ = O.static_force(phi, u, linf)
    f, e = s.post_force(u)
    assert rel(f, f_ref) < 1e-11
    assert abs(e - e_ref) < 1e-11 * abs(e_ref)


def test_fp32_forces_within_bound():
    n, args = 256, "fcc111 1 1.0 pair-potential 1 1.0 height 256"
    u = 0.02 * np.random.default_rng(4).standard_normal((6, n, n))
    out = {}
    for prec in (gfmd_hip.FP64, gfmd_hip.FP32):
        s = gfmd_hip.Solver(n, n, 6, prec)
        s.set_kernel("ft", args)
        out[prec] = s.post_force(u)
    f64, e64 = out[gfmd_hip.FP64]
    f32, e32 = out[gfmd_hip.FP32]
    assert rel(f32, f64) < 2e-5
    assert abs(e32 - e64) < 1e-5 * abs(e64)


def test_ndof_mismatch_and_unsupported_cells():
    s = gfmd_hip.Solver(8, 8, 3, gfmd_hip.FP64)
    with pytest.raises(gfmd_hip.GFMDError, match="degrees of freedom"):
        s.set_kernel("ft", "fcc111 1 1.0 pair-potential 1 1.0")
    s = gfmd_hip.Solver(8, 8, 24, gfmd_hip.FP64)
    with pytest.raises(gfmd_hip.GFMDError, match="up to 4"):
        s.set_kernel("ft", "fcc111 1.0 4 pair-potential 1 1.0")


def _hertz_fcc111(f_top, pos, nx, ny, cell_y, R=100.0, E=1.54):
    """eval.py of TEST_Hertz_fcc111_*: pressure f / A0 (A0 = sqrt(3)/2, two
    atoms per cell) of the top-layer atoms against the Hertz profile of
    their total load, sum of squared deviations.  Distances from the
    sphere's axis use each atom's own position (minimum image)."""
    lx, ly = nx * 1.0, ny * cell_y
    r, f = [], []
    for a, fz in enumerate(f_top):
        x = np.arange(nx)[:, None] + pos[a][0] + 0 * np.arange(ny)[None, :]
        y = cell_y * np.arange(ny)[None, :] + pos[a][1] + 0 * x
        x = x - lx * np.rint(x / lx)
        y = y - ly * np.rint(y / ly)
        r.append(np.sqrt(x * x + y * y).ravel())
        f.append(fz.ravel())
    r, f = np.concatenate(r), np.concatenate(f)
    N = f.sum()
    a = R * (3.0 / 4 * (N / (E * R ** 2))) ** (1.0 / 3)
    p0 = 3 * N / (2 * np.pi * a * a)
    A0 = np.sqrt(3.0) / 2
    pa = np.where(r < a, p0 * np.sqrt(np.maximum(1 - (r / a) ** 2, 0)), 0.0)
    return np.sum((f / A0 - pa) ** 2), N, a


@pytest.mark.parametrize("args,zc,bound", [
    # TEST_Hertz_fcc111_64x37_2: sphere (0, 0, 99.5), criterion 0.002
    ("fcc111 1 1.0 pair-potential 1 1.0 height 128", 99.5, 0.002),
    # TEST_Hertz_fcc111_two_layers_64x37: sphere (0, 0, 100.3), 0.001
    ("fcc111 1.0 2 pair-potential 1 1.0 height 128", 100.3, 0.001)])
def test_hertz_fcc111_reference_cases(args, zc, bound):
    """The reference's fcc111 Hertz tests: 64 x 37 cells, the LJ sphere of
    fix contact/sphere (R 100, eps 1.38888888888889, sigma
    0.890898718140339, cutoff 1) on every atom at its own position, all
    atoms displaced by -2 along z, minimised to |f| <= 1e-6 (the reference
    minimises with ftol 1e-6); z forces of the two top-layer atoms (degrees
    of freedom 2 and 5) against Hertz with E* = 1.54."""
    nx, ny = 64, 37
    cell = gfmd_hip.ft_cell(args)
    nat, pos = cell["nat"], cell["positions"]
    cx, cy = cell["cell"][0, 0], cell["cell"][1, 1]
    s = gfmd_hip.Solver(nx, ny, 3 * nat, gfmd_hip.FP64)
    s.set_kernel("ft", args)
    u0 = np.zeros((3 * nat, nx, ny))
    u0[2::3] = -2.0
    c = gfmd_hip.contact(gfmd_hip.CONTACT_LJ_SPHERE, lattice=(1.0, 0, 0, 0),
                         sphere=(0.0, 0.0, zc, 100.0, 1.38888888888889,
                                 0.890898718140339, 1.0),
                         cell_xy=(cx, cy), atom_offsets=[tuple(p) for p in pos])
    fire = gfmd_hip.fire_params(dt=0.05, dtmax=0.1, adapt_dt=1)
    u, res = s.relax(u0, c, fire, max_steps=400000, tol=1e-6, check_every=100)
    assert res["converged"], res
    f, _ = s.post_force(u)
    r, N, a = _hertz_fcc111([f[2], f[5]], pos, nx, ny, cy)
    print("%s: Hertz residual %.3g (bound %g), load %.3f, radius %.2f, "
          "%d FIRE steps" % (args, r, bound, N, a, res["steps"]), flush=True)
    assert N > 1.0 and a > 3.0
    assert r < bound


@pytest.mark.parametrize("args,lattice,E,A0,xlimit", [
    # TEST_Hertz_sc100_128x128_a0_1.3 (metal units, lattice sc 1.3, atoms
    # moved by 0.5): E* = (8/3) / 1.3, pressure f / 1.3^2
    ("sc100 1.3 1 pair-potential 2 1.0 1.0 height 128",
     (1.3, 0.5, 0.5, 0.5), 8.0 / 3 / 1.3, 1.69, 0.0),
    # TEST_Hertz_fcc100_128x128: E* = 1.39
    ("fcc100 1 1.0 pair-potential 1 1.0 height 128",
     (1.0, 0.5, 0.5, 0.5), 1.39, 1.0, 0.0),
    # TEST_Hertz_fcc100_two_layers_128x128: top atom (0.5, 0.5, 0.707) of
    # the cell at (ix, iy, 0.5), E* = 1.39.  The top layer starts 0.29
    # below the sphere's surface (z_eq 1.207 - 2 vs -0.5), where the LJ
    # force is ~1e8; the reference's CG line search absorbs that, FIRE
    # needs its displacement limit (LAMMPS' dmax) for the first steps.
    ("fcc100 1.0 2 pair-potential 1 1.0 height 64",
     (1.0, 0.0, 0.0, 0.5), 1.39, 1.0, 0.05)])
def test_hertz_square_cells_reference_cases(args, lattice, E, A0, xlimit):
    """The reference's Hertz tests on sc100 (lattice constant 1.3) and
    fcc100 (one and two layers): 128 x 128 cells, the LJ sphere (0, 0,
    99.5), R 100, on every atom, all atoms displaced by -2, minimised to
    |f| <= 1e-6; eval.py: r from (i + 1/2) a0, residual sum (f / A0 -
    p_Hertz)^2 < 1e-2 on the top atom's z force."""
    n = 128
    cell = gfmd_hip.ft_cell(args)
    nat, pos = cell["nat"], cell["positions"]
    s = gfmd_hip.Solver(n, n, 3 * nat, gfmd_hip.FP64)
    s.set_kernel("ft", args)
    u0 = np.zeros((3 * nat, n, n))
    u0[2::3] = -2.0
    c = gfmd_hip.contact(gfmd_hip.CONTACT_LJ_SPHERE, lattice=lattice,
                         sphere=(0.0, 0.0, 99.5, 100.0, 1.38888888888889,
                                 0.890898718140339, 1.0),
                         cell_xy=(cell["cell"][0, 0], cell["cell"][1, 1]),
                         atom_offsets=[tuple(p) for p in pos])
    fire = gfmd_hip.fire_params(dt=0.05, dtmax=0.1, adapt_dt=1,
                                xlimit=xlimit)
    u, res = s.relax(u0, c, fire, max_steps=400000, tol=1e-6, check_every=100)
    assert res["converged"], res
    f, _ = s.post_force(u)
    a0 = lattice[0]
    x = np.arange(n) + 0.5
    x = np.where(x > n / 2, x - n, x) * a0
    r = np.sqrt(x[:, None] ** 2 + x[None, :] ** 2)
    N = f[2].sum()
    a = 100.0 * (3.0 / 4 * (N / (E * 100.0 ** 2))) ** (1.0 / 3)
    p0 = 3 * N / (2 * np.pi * a * a)
    pa = np.where(r < a, p0 * np.sqrt(np.maximum(1 - (r / a) ** 2, 0)), 0.0)
    resid = np.sum((f[2] / A0 - pa) ** 2)
    print("%s: Hertz residual %.3g, load %.3f, radius %.2f, %d FIRE steps"
          % (args, resid, N, a, res["steps"]), flush=True)
    assert N > 1.0 and a > 3.0
    assert resid < 1e-2


@pytest.mark.parametrize("args", [
    "fcc100 1.0 2 pair-potential 1 1.0 height 8",
    "fcc111 1.0 2 pair-potential 1 1.0 height 16"])
def test_contact_solve_multi_atom_cells_match_oracle(args):
    """Hard-wall contact (constrained CG, compliance tables of ndof 6 and
    12 from the wave-cooperative kernels) on atom 0 of each cell against
    the oracle's exact complementarity solution, 16 x 16 cells, rigid
    paraboloid R = 10 at load 0.05 per cell."""
    n = 16
    k = ft.FTKernel(args)
    s = gfmd_hip.Solver(n, n, k.dim, gfmd_hip.FP64)
    s.set_kernel("ft", args)
    X, Y = np.meshgrid(np.arange(n) - n / 2, np.arange(n) - n / 2,
                       indexing="ij")
    wall = (X ** 2 + Y ** 2) / 20.0
    c = gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL, pressure=0.05, wall=wall)
    u, res = s.contact_solve(c, tol=1e-11)
    assert res["converged"], res
    u_ref, p_ref = O.hardwall_contact(k.fill_phi(n, n), wall, 0.05)
    assert (p_ref > 0).sum() > 3
    assert rel(u, u_ref) < 1e-8


@pytest.mark.parametrize("args,nx,ny", [
    ("fcc111 1.0 2 pair-potential 1 1.0 height 128", 64, 64),
    ("fcc111 1.0 2 pair-potential 1 1.0 height 32", 128, 64),
    ("fcc100 1.0 3 pair-potential 2 1.0 0.5 height 5", 64, 128),
    ("fcc111 1.0 2 pair-potential 1 1.0 height 1", 4096, 64)])
def test_large_cells_fast_path_every_variant(args, nx, ny):
    """Three- and four-atom cells (ndof 9, 12) on power-of-two grids of 64
    to 4096 points per side run the FFT sweeps: every variant against the oracle in FP64 (1e-11, as the
    other ft kernels), and the FP32 path within 5e-6."""
    k = ft.FTKernel(args)
    phi = k.fill_phi(nx, ny)
    u = 0.01 * np.random.default_rng(5).standard_normal((k.dim, nx, ny))
    f_ref, e_ref, _ = O.static_force(phi, u)
    s = gfmd_hip.Solver(nx, ny, k.dim, gfmd_hip.FP64)
    s.set_kernel("ft", args)
    assert s.info()["fast_path"] == 1
    for rv in range(3):
        for cv in range(4):
            s.set_variant(rv, cv)
            f, e = s.post_force(u)
            assert rel(f, f_ref) < 1e-11, (rv, cv)
            assert abs(e - e_ref) < 1e-11 * abs(e_ref), (rv, cv)
    s32 = gfmd_hip.Solver(nx, ny, k.dim, gfmd_hip.FP32)
    s32.set_kernel("ft", args)
    assert s32.info()["fast_path"] == 1
    f, e = s32.post_force(u)
    print("ndof %d %dx%d FP32: force %.2e energy %.2e"
          % (k.dim, nx, ny, rel(f, f_ref), abs(e - e_ref) / abs(e_ref)))
    assert rel(f, f_ref) < 5e-6
    assert abs(e - e_ref) < 5e-6 * abs(e_ref)
