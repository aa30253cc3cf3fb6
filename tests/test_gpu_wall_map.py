"""The soft wall over a height map (`fix wall/map`) on the device against
the oracle (oracle/wall_map.py): the per-atom wall force (pre_force) for
every map kind and potential, the map set-up (file, filter, shift), FIRE
trajectories with the wall fused into the solver step, a flat-wall
equilibrium solved in closed form, and energy conservation in dynamics."""
import os
import tempfile

import numpy as np
import pytest

import ft
import gfmd_hip
import oracle as O
import wall_map as W

pytestmark = pytest.mark.gpu

POTS = ["lj93 0.8 0.9 2.5", "lj93/smooth 0.8 0.9 1.6 2.6",
        "lj126 0.5 0.8 2.4 pendist 0.3", "exp 0.7 0.4 2.5",
        "expf 0.2 0.4 2.5", "dugdale 0.6 1.5"]
# pendist: the wall places atoms that come closer than 0.3 at 0.3


def _atoms(rng, n, xprd, yprd, zlo, zhi):
    return np.column_stack([rng.uniform(-0.5 * xprd, 1.5 * xprd, n),
                            rng.uniform(-0.5 * yprd, 1.5 * yprd, n),
                            rng.uniform(zlo, zhi, n)])


def _compare(wd, wo, x, exact):
    xd, fd, vd = wd.force(x)
    xo = x.copy()
    fo, vo, close = wo.pre_force(xo)
    assert close == 0
    if exact:
        assert np.array_equal(xd, xo)
        assert np.array_equal(fd, fo)
    else:
        assert np.abs(xd - xo).max() <= 1e-13 * max(1.0, np.abs(xo).max())
        assert np.abs(fd - fo).max() <= 1e-13 * np.abs(fo).max()
    assert np.array_equal(vd[4:], vo[4:])  # atoms in range, rep., att.
    assert np.abs(vd[:4] - vo[:4]).max() <= 1e-12 * max(1.0, np.abs(vo).max())
    return vd


@pytest.mark.parametrize("pot", POTS)
def test_table_map_force_matches_oracle(pot):
    """Bicubic map: positions and forces bit for bit for the power-law
    potentials (the device keeps the reference's operation order without
    contraction), to 1e-13 for those that call exp."""
    rng = np.random.default_rng(11)
    xprd, yprd = 12.0, 10.0
    h = rng.normal(scale=0.3, size=(24, 20))
    args = "rough.map 2.0 -1 " + pot
    if "pendist" not in pot and not pot.startswith("dugdale"):
        args += " pendist 0.1"
    wd = gfmd_hip.WallMap(args, xprd, yprd, heights=h)
    wo = W.WallMap(args, xprd, yprd, map_values=h)
    x = _atoms(rng, 3000, xprd, yprd, -0.6, 2.4)
    exact = pot.split()[0] in ("lj93", "lj93/smooth", "lj126", "dugdale")
    v = _compare(wd, wo, x, exact)
    assert v[4] > 500 and v[5] > 0
    info = wd.info()
    assert abs(info["average"] - 2.0) < 1e-13
    assert abs(info["rms"] - wo.rms) < 1e-13 * wo.rms
    assert (info["nx"], info["ny"]) == (24, 20)


@pytest.mark.parametrize("kind", ["flat", "100 1.0 0.4", "111 1.0 0.3",
                                  "raw"])
def test_analytic_and_raw_maps_match_oracle(kind):
    rng = np.random.default_rng(12)
    xprd, yprd = 12.0, 12.0
    h = None
    if kind == "raw":
        h = rng.normal(scale=0.3, size=(12, 12))
        args = "rough.map 2.0 -1 lj93 0.8 0.9 2.5 disable_interpolation"
    else:
        args = kind + " 2.0 -1 lj93 0.8 0.9 2.5"
    wd = gfmd_hip.WallMap(args, xprd, yprd, heights=h)
    wo = W.WallMap(args, xprd, yprd, map_values=h)
    x = _atoms(rng, 2000, xprd, yprd, 0.2, 2.2)
    _compare(wd, wo, x, kind in ("flat", "raw"))


def test_filtered_map_from_file():
    """Map read from a file in the reference's format, low-pass filtered
    (lambda) and placed with its top at z0: the same as the oracle's
    filter_map + displace_map; the file and the in-memory table agree."""
    rng = np.random.default_rng(13)
    n, L = 32, 16.0
    h = rng.normal(scale=0.2, size=(n, n))
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "rough.map")
        with open(fn, "w") as f:
            for row in h:
                f.write(" ".join("%.17g" % v for v in row) + "\n")
        args = fn + " top 1.5 -1 lj93/smooth 0.8 0.9 1.6 2.6 lambda 2.0"
        wf = gfmd_hip.WallMap(args, L, L)
        wa = gfmd_hip.WallMap(args, L, L, heights=h)
        wo = W.WallMap(args, L, L, map_values=h)
    inf, ina = wf.info(), wa.info()
    assert inf == ina
    assert abs(inf["max"] - 1.5) < 1e-13
    assert abs(inf["min"] - wo.values.min()) < 1e-12
    assert abs(inf["average"] - wo.avg) < 1e-12
    assert abs(inf["rms"] - wo.rms) < 1e-12 * wo.rms
    # the filter keeps a fraction of the modes
    assert wo.rms < 0.5 * h.std()
    x = _atoms(rng, 2000, L, L, 0.0, 1.4)
    xf, ff, vf = wf.force(x)
    xa, fa, va = wa.force(x)
    assert np.array_equal(ff, fa) and np.array_equal(vf, va)
    xo = x.copy()
    fo, vo, _ = wo.pre_force(xo)
    # filtered heights differ at FFT rounding
    assert np.abs(ff - fo).max() < 1e-9 * np.abs(fo).max()


def test_lj126_too_close_is_an_error():
    w = gfmd_hip.WallMap("flat 1.0 -1 lj126 0.5 0.8 2.4", 8.0, 8.0)
    x = np.array([[1.0, 1.0, 0.99], [2.0, 2.0, 0.5]])  # dz 0.01 < 0.024
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        w.force(x)
    assert ei.value.code == 7 and "Atom too close" in str(ei.value)


def _rough(n, seed=21, scale=0.25):
    rng = np.random.default_rng(seed)
    h = rng.normal(scale=scale, size=(n, n))
    return W.filter_map(h, 4.0, float(n), float(n))


def test_fire_with_wall_map_matches_oracle():
    """60 FIRE steps under a rough lj93/smooth wall (bicubic map) and a load,
    FP64: the device's fused steps follow the oracle's reference-order FIRE
    to rounding."""
    nx = ny = 32
    h = _rough(nx)
    args = "rough.map 1.2 -1 lj93/smooth 0.05 0.8 1.4 2.2 pendist 0.2"
    wd = gfmd_hip.WallMap(args, nx, ny, heights=h)
    wo = W.WallMap(args, nx, ny, map_values=h)
    lattice = (1.0, 0.5, 0.5, 0.0)
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 16")
    fire = gfmd_hip.fire_params(dt=0.1, dtmax=0.3, adapt_dt=1)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_WALL_MAP, pressure=0.002,
                         lattice=lattice, wall_map=wd)
    u0 = np.zeros((3, nx, ny))
    u, res = s.relax(u0, c, fire, max_steps=60, tol=0.0, check_every=7)
    assert res["steps"] == 60
    phi = O.fill_phi(O.KERNEL_SC100, [16], 3, nx, ny, normalize=True)
    F = O.Fire(dt=0.1, adapt_dt=True, dtmax=0.3)
    uo, _, vec = W.fire_relax_wall_map(phi, u0, lattice, wo, F, 60,
                                       pressure=0.002)
    assert np.abs(u - uo).max() < 1e-9 * np.abs(uo).max()
    assert res["ncontact"] == vec[4]
    assert abs(res["contact_energy"] - vec[0]) < 1e-8 * abs(vec[0])


def test_fire_two_atom_cells_match_oracle():
    """fcc(111) surface cells (ft kernel, two atoms in the top layer), both
    atoms under the wall at their own positions."""
    n = 16
    args = "fcc111 1.0 1 pair-potential 1 1.0 height 4"
    k = ft.FTKernel(args)
    cell = gfmd_hip.ft_cell(args)
    cx, cy = cell["cell"][0, 0], cell["cell"][1, 1]
    offs = [tuple(p) for p in cell["positions"]]
    L = (n * cx, n * cy)
    h = _rough(24, seed=22)
    wargs = "rough.map 1.0 -1 exp 0.02 0.3 2.0"
    wd = gfmd_hip.WallMap(wargs, L[0], L[1], heights=h)
    wo = W.WallMap(wargs, L[0], L[1], map_values=h)
    lattice = (1.0, 0.0, 0.0, 0.0)
    s = gfmd_hip.Solver(n, n, k.dim, gfmd_hip.FP64)
    s.set_kernel("ft", args)
    fire = gfmd_hip.fire_params(dt=0.05, dtmax=0.2, adapt_dt=1)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_WALL_MAP, pressure=0.01,
                         lattice=lattice, wall_map=wd, cell_xy=(cx, cy),
                         atom_offsets=offs)
    u0 = np.zeros((k.dim, n, n))
    u, res = s.relax(u0, c, fire, max_steps=40, tol=0.0, check_every=40)
    F = O.Fire(dt=0.05, adapt_dt=True, dtmax=0.2)
    uo, _, vec = W.fire_relax_wall_map(k.fill_phi(n, n), u0, lattice, wo, F,
                                       40, pressure=0.01, cell_xy=(cx, cy),
                                       offsets=offs)
    assert np.abs(u - uo).max() < 1e-9 * np.abs(uo).max()
    assert res["ncontact"] == vec[4] == 2 * n * n


def test_flat_wall_equilibrium_closed_form():
    """Flat lj93 wall above a loaded surface: every site sits where the
    load, the wall force and the uniform-mode elastic force balance,
    p = f_lj93(z0 - u) + K u with K = Phi_zz(q = 0); the scalar root is
    found by bisection on the host."""
    nx = ny = 64
    z0, p = 1.0, 0.01
    args = "flat %r -1 lj93 1.0 1.0 2.5" % z0
    wd = gfmd_hip.WallMap(args, nx, ny)
    wo = W.WallMap(args, nx, ny)
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 32")
    one = np.zeros((3, nx, ny))
    one[2] = 1.0
    K = -s.post_force(one)[0][2].mean()
    assert K > 0

    def g(u):
        return p - wo.potential(np.array([z0 - u]))[1][0] - K * u
    lo, hi = -1.0, z0 - 0.6  # g(lo) > 0 > g(hi)
    assert g(lo) > 0 > g(hi)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if g(mid) > 0 else (lo, mid)
    fire = gfmd_hip.fire_params(dt=0.05, dtmax=0.2, adapt_dt=1)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_WALL_MAP, pressure=p, wall_map=wd)
    u, res = s.relax(np.zeros((3, nx, ny)), c, fire, max_steps=100000,
                     tol=1e-12, check_every=50)
    assert res["converged"]
    assert np.abs(u[2] - lo).max() < 1e-9
    assert np.abs(u[:2]).max() < 1e-12


def test_rough_wall_fp32_matches_fp64():
    nx = ny = 128
    h = _rough(nx, seed=23, scale=0.4)
    # the map's highest point 0.84 above the surface; xlimit (fix fire's
    # largest step) keeps the first kicks from launching atoms
    args = "rough.map 1.5 -1 lj93/smooth 0.05 0.8 1.4 2.2"
    wd = gfmd_hip.WallMap(args, nx, ny, heights=h)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_WALL_MAP, pressure=0.005,
                         wall_map=wd)
    fire = gfmd_hip.fire_params(dt=0.1, dtmax=0.3, adapt_dt=1, xlimit=0.05)
    out = {}
    for prec, tol in ((gfmd_hip.FP64, 1e-9), (gfmd_hip.FP32, 1e-4)):
        s = gfmd_hip.Solver(nx, ny, 3, prec)
        s.set_kernel("sc100", "height 64")
        u, res = s.relax(np.zeros((3, nx, ny)), c, fire, max_steps=200000,
                         tol=tol, check_every=50)
        assert res["converged"], res
        out[prec] = (u, res)
    u64, r64 = out[gfmd_hip.FP64]
    u32, r32 = out[gfmd_hip.FP32]
    assert np.abs(u32 - u64).max() < 1e-3 * np.abs(u64).max()
    assert r64["ncontact"] > 0.2 * nx * ny
    # force balance at the FP64 solution: wall + load + elastic = 0 per site
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 64")
    f_el = s.post_force(u64)[0]
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    pos = np.column_stack([X.ravel() + u64[0].ravel(),
                           Y.ravel() + u64[1].ravel(), u64[2].ravel()])
    _, fw, _ = wd.force(pos)
    tot = f_el.reshape(3, -1).T + fw
    tot[:, 2] += 0.005
    assert np.sqrt((tot ** 2).sum()) <= 2e-9 * 0.005 * nx


def test_refine_wall_map_reaches_fp64_solution():
    """FP32 inner FIRE + FP64 corrections under the rough wall reach the
    FP64 residual 1e-9 (FP32 FIRE alone stalls near 1e-4) and the FP64
    FIRE state."""
    nx = ny = 128
    h = _rough(nx, seed=23, scale=0.4)
    wd = gfmd_hip.WallMap("rough.map 1.5 -1 lj93/smooth 0.05 0.8 1.4 2.2",
                          nx, ny, heights=h)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_WALL_MAP, pressure=0.005,
                         wall_map=wd)
    fire = gfmd_hip.fire_params(dt=0.1, dtmax=0.3, adapt_dt=1, xlimit=0.05)
    s64 = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s64.set_kernel("sc100", "height 64")
    s32 = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP32)
    s32.set_kernel("sc100", "height 64")
    u64, r64 = s64.relax(np.zeros((3, nx, ny)), c, fire, max_steps=200000,
                         tol=1e-9, check_every=50)
    um, rm = s64.relax_refine(s32, np.zeros((3, nx, ny)), c, fire,
                              max_outer=40, tol=1e-9, inner_tol=1e-3)
    print(r64, rm)
    assert rm["converged"] and rm["outer"] >= 1
    assert rm["residual"] / rm["fext_norm"] <= 1e-9
    assert np.abs(um - u64).max() < 1e-6 * np.abs(u64).max()


def test_dynamics_with_wall_map_conserves_energy():
    """NVE under an exp wall over a bicubic map: the fused wall force is the
    gradient of the energy the kernels report."""
    nx = ny = 64
    h = _rough(nx, seed=24, scale=0.2)
    wd = gfmd_hip.WallMap("rough.map 0.8 -1 exp 0.02 0.25 2.0", nx, ny,
                          heights=h)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_WALL_MAP, wall_map=wd)
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 16")
    u0 = np.zeros((3, nx, ny))
    v0 = np.zeros((3, nx, ny))
    v0[2] = 0.05 * np.exp(-((np.arange(nx)[:, None] - 32) ** 2 +
                            (np.arange(ny)[None, :] - 32) ** 2) / 50.0)
    u, v, hist, res = s.dynamics(u0, v0, 3000, 0.02, record_every=10,
                                 contact=c)
    e = hist[:, 2]
    assert np.abs(e - e.mean()).max() < 1e-4 * hist[:, 0].max()
    assert res["ncontact"] > 0


def test_wall_map_errors():
    s = gfmd_hip.Solver(16, 16, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 4")
    w = gfmd_hip.WallMap("flat 1.0 -1 lj93 1 1 2", 20.0, 16.0)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_WALL_MAP, wall_map=w)
    with pytest.raises(gfmd_hip.GFMDError, match="wall map spans"):
        s.relax(np.zeros((3, 16, 16)), c, max_steps=1)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_WALL_MAP)
    with pytest.raises(gfmd_hip.GFMDError, match="needs a gfmd_wall_map_t"):
        s.relax(np.zeros((3, 16, 16)), c, max_steps=1)


def test_refine_wall_map_sparse_contact():
    """Refinement under a soft wall that touches few atoms: most atoms stay
    beyond the cutoff, so the correction passes take the clearance path
    (no loads of the FP64 base for them); the result still meets the FP64
    residual and matches the FP64 FIRE state."""
    n = 256
    h = gfmd_hip.rough_surface(n, n, hurst=0.8, rms_slope=0.1,
                               lambda_short=8.0, seed=3)
    wd = gfmd_hip.WallMap("rough.map bot 1.0 -1 lj93/smooth 0.01 0.8 1.4 2.2",
                          float(n), float(n), heights=h)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_WALL_MAP, pressure=1e-3,
                         wall_map=wd)
    fire = gfmd_hip.fire_params(dt=0.1, dtmax=0.3, adapt_dt=1, xlimit=0.05)
    s64 = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    s64.set_kernel("sc100", "height %d" % n)
    s32 = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP32)
    s32.set_kernel("sc100", "height %d" % n)
    z = np.zeros((3, n, n))
    u64, r64 = s64.relax(z, c, fire, max_steps=100000, tol=1e-9,
                         check_every=50)
    um, rm = s64.relax_refine(s32, z, c, fire, max_outer=40, tol=1e-9,
                              inner_tol=1e-2, check_every=50)
    print(r64, rm)
    assert r64["converged"] and rm["converged"] and rm["outer"] >= 2
    assert 0 < rm["ncontact"] < 0.25 * n * n
    assert rm["residual"] / rm["fext_norm"] <= 1e-9
    assert np.abs(um - u64).max() < 1e-6 * np.abs(u64).max()
