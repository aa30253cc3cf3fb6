"""FIRE relaxation on the device vs the oracle, and the reference's Hertz
contact check (tests/TEST_Hertz_sc100_128x128: lammps.in + eval.py)
reproduced with the device solver."""
import numpy as np
import pytest

import gfmd_hip
import oracle as O

pytestmark = pytest.mark.gpu


def hertz_residual(fz, R=100.0, E=8.0 / 3):
    """TEST_Hertz_sc100_128x128/eval.py:33-85 on a z-force map."""
    nx, ny = fz.shape
    x = np.arange(nx) + 0.5
    x = np.where(x > nx / 2, x - nx, x)
    y = np.arange(ny) + 0.5
    y = np.where(y > ny / 2, y - ny, y)
    r = np.sqrt((x ** 2).reshape(-1, 1) + (y ** 2).reshape(1, -1))
    N = np.sum(fz)
    a = R * (3.0 / 4 * (N / (E * R ** 2))) ** (1.0 / 3)
    p0 = 3 * N / (2 * np.pi * a * a)
    fa = np.where(r < a, p0 * np.sqrt(np.maximum(1 - (r / a) ** 2, 0)), 0.0)
    return np.sum((fz - fa) ** 2), N, a


# Reference Hertz geometry: sites at (ix+0.5, iy+0.5), x_eq at z = 0.5,
# atoms displaced by -2 before minimisation, sphere (0, 0, 99.5), R = 100,
# LJ eps 1.38888888888889, sigma 0.890898718140339, cutoff 1.0.
LATTICE = (1.0, 0.5, 0.5, 0.5)
SPHERE = (0.0, 0.0, 99.5, 100.0, 1.38888888888889, 0.890898718140339, 1.0)


@pytest.mark.parametrize("adapt", [0, 1])
def test_fire_steps_match_oracle(adapt):
    """Same number of FIRE steps on the device and in the oracle (LJ sphere,
    load, FP64): trajectories agree to rounding."""
    nx, ny, height = 32, 32, 16
    sphere = (0.0, 0.0, 40.5, 40.0, 1.0, 0.8, 1.0)
    u0 = np.zeros((3, nx, ny))
    u0[2] = -1.2
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height %d" % height)
    fire = gfmd_hip.fire_params(dt=0.05, dtmax=0.2, adapt_dt=adapt)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_LJ_SPHERE, pressure=0.002,
                         lattice=LATTICE, sphere=sphere)
    steps = 60
    u, res = s.relax(u0, c, fire, max_steps=steps, tol=0.0, check_every=7)
    assert res["steps"] == steps
    phi = O.fill_phi(O.KERNEL_SC100, [height], 3, nx, ny, normalize=True)
    F = O.Fire(dt=0.05, adapt_dt=bool(adapt), dtmax=0.2)
    uo, _ = O.fire_relax_sphere(phi, u0, LATTICE, sphere, F, steps,
                                pressure=0.002)
    assert np.abs(u - uo).max() < 1e-9 * np.abs(uo).max()


@pytest.mark.parametrize("prec", [gfmd_hip.FP64, gfmd_hip.FP32])
def test_fire_steps_match_oracle_odd_grid(prec):
    """A grid whose site count is not a multiple of the passes' vector
    width (31 x 17): the one-site-per-thread passes, against the oracle."""
    nx, ny, height = 31, 17, 8
    sphere = (15.5, 8.5, 8.5, 8.0, 1.0, 0.8, 1.0)
    u0 = np.zeros((3, nx, ny))
    u0[2] = -0.9
    s = gfmd_hip.Solver(nx, ny, 3, prec)
    s.set_kernel("sc100", "height %d" % height)
    fire = gfmd_hip.fire_params(dt=0.02, dtmax=0.05, adapt_dt=1)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_LJ_SPHERE, pressure=0.002,
                         lattice=LATTICE, sphere=sphere)
    u, res = s.relax(u0, c, fire, max_steps=60, tol=0.0, check_every=60)
    assert res["ncontact"] > 10
    phi = O.fill_phi(O.KERNEL_SC100, [height], 3, nx, ny, normalize=True)
    F = O.Fire(dt=0.02, adapt_dt=True, dtmax=0.05)
    uo, _ = O.fire_relax_sphere(phi, u0, LATTICE, sphere, F, 60,
                                pressure=0.002)
    tol = 1e-9 if prec == gfmd_hip.FP64 else 1e-4
    assert np.abs(u - uo).max() < tol * np.abs(uo).max()


@pytest.mark.parametrize("prec,tol", [(gfmd_hip.FP64, 1e-8),
                                      (gfmd_hip.FP32, 5e-4)])
def test_hertz_sc100_128_lj_sphere(prec, tol):
    """The reference's Hertz test: minimise, then compare the GF z-force map
    with the Hertz pressure profile; criterion sum (f - f_Hertz)^2 < 1e-2."""
    nx = ny = 128
    s = gfmd_hip.Solver(nx, ny, 3, prec)
    s.set_kernel("sc100", "height 128")
    u0 = np.zeros((3, nx, ny))
    u0[2] = -2.0
    # LJ wall stiffness ~270 near contact: stable dt < 2/sqrt(278) ~ 0.12
    fire = gfmd_hip.fire_params(dt=0.05, dtmax=0.1, adapt_dt=1)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_LJ_SPHERE, lattice=LATTICE,
                         sphere=SPHERE)
    u, res = s.relax(u0, c, fire, max_steps=200000, tol=tol, check_every=100)
    print(prec, res)
    assert res["converged"]
    f64 = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    f64.set_kernel("sc100", "height 128")
    f, e = f64.post_force(u)
    r, N, a = hertz_residual(f[2])
    print("Hertz residual %.3e, load %.4f, contact radius %.2f" % (r, N, a))
    assert N > 1.0 and a > 3.0
    assert r < 1e-2


def test_hardwall_sphere_load_control_kkt():
    """Rigid sphere (hard wall), load per site 1e-2 on 128x128 (the CPU
    config's problem at a small size): converged state satisfies the
    contact conditions and the total reaction equals the load."""
    nx = ny = 128
    R, p0 = 100.0, 1e-2
    X, Y = np.meshgrid(np.arange(nx) - nx / 2 + 0.5, np.arange(ny) - ny / 2 + 0.5,
                       indexing="ij")
    wall = (X ** 2 + Y ** 2) / (2 * R)  # indenter surface above z = 0
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 128")
    fire = gfmd_hip.fire_params(dt=0.1, dtmax=0.5, adapt_dt=1)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL, pressure=p0, wall=wall)
    u, res = s.relax(np.zeros((3, nx, ny)), c, fire, max_steps=200000,
                     tol=1e-10, check_every=100)
    print(res)
    assert res["converged"] and res["residual"] / res["fext_norm"] <= 1e-10
    f, _ = s.post_force(u)
    gap = wall - u[2]
    assert gap.min() > -1e-12
    # contact pressure the wall exerts (downwards): p = f_el,z + p0
    pz = f[2] + p0
    touch = gap < 1e-9
    assert touch.sum() > 10
    assert np.all(pz[touch] > -1e-8)
    assert np.abs(pz[~touch]).max() < 1e-8
    assert np.abs(f[0]).max() < 1e-8 and np.abs(f[1]).max() < 1e-8
    assert res["ncontact"] == touch.sum()
    # the wall carries what the slab's base does not: sum p = N p0 + sum f_z
    assert abs(pz[touch].sum() - (p0 * nx * ny + f[2].sum())) < 1e-7


@pytest.mark.parametrize("adapt", [0, 1])
def test_fire_hardwall_steps_match_oracle(adapt):
    """FIRE against a rigid hard-wall sphere under load on the device and
    in the oracle's restatement (projection after each move, no push into
    the wall in the force): 80 steps agree to rounding (FP64)."""
    nx, ny, height = 32, 32, 16
    wall = _sphere_wall(nx, ny, 20.0)
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height %d" % height)
    fire = gfmd_hip.fire_params(dt=0.1, dtmax=0.5, adapt_dt=adapt)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL, pressure=1e-2, wall=wall)
    u0 = np.zeros((3, nx, ny))
    u, res = s.relax(u0, c, fire, max_steps=80, tol=0.0, check_every=80)
    assert res["steps"] == 80 and res["ncontact"] >= 4
    phi = O.fill_phi(O.KERNEL_SC100, [height], 3, nx, ny, normalize=True)
    F = O.Fire(dt=0.1, adapt_dt=bool(adapt), dtmax=0.5)
    uo, _, n, _ = O.fire_relax_hardwall(phi, u0, wall, 1e-2, F, 80)
    assert n == 80
    assert np.abs(u - uo).max() < 1e-9 * np.abs(uo).max()


def test_hardwall_sphere_cpu_config_matches_oracle():
    """The CPU configuration (sc100 height 128 on 128^2, rigid sphere R =
    100, load 1e-2 per site, FP64 FIRE to relative residual 1e-10) on the
    device and in the oracle's restatement: both converge, to the same
    state."""
    nx = ny = 128
    wall = _sphere_wall(nx, ny, 100.0)
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 128")
    fire = gfmd_hip.fire_params(dt=0.1, dtmax=0.5, adapt_dt=1)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL, pressure=1e-2, wall=wall)
    u0 = np.zeros((3, nx, ny))
    u, res = s.relax(u0, c, fire, max_steps=200000, tol=1e-10,
                     check_every=100)
    assert res["converged"]
    phi = O.fill_phi(O.KERNEL_SC100, [128], 3, nx, ny, normalize=True)
    F = O.Fire(dt=0.1, adapt_dt=True, dtmax=0.5)
    uo, _, n, r = O.fire_relax_hardwall(phi, u0, wall, 1e-2, F, 200000,
                                        tol=1e-10, check_every=100)
    print("C1: device %d steps, oracle %d steps (residual %.2e)"
          % (res["steps"], n, r))
    assert r <= 1e-10
    assert np.abs(u - uo).max() < 1e-8 * np.abs(uo).max()


def _sphere_wall(nx, ny, R):
    X, Y = np.meshgrid(np.arange(nx) - nx / 2 + 0.5,
                       np.arange(ny) - ny / 2 + 0.5, indexing="ij")
    return (X ** 2 + Y ** 2) / (2 * R)


def test_refine_reaches_fp64_solution():
    """FP32 inner FIRE + FP64 outer corrections reach the FP64 residual
    1e-10 and the FP64-only solution."""
    nx = ny = 128
    wall = _sphere_wall(nx, ny, 100.0)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL, pressure=1e-2, wall=wall)
    fire = gfmd_hip.fire_params(dt=0.1, dtmax=0.5, adapt_dt=1)
    s64 = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s64.set_kernel("sc100", "height 128")
    s32 = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP32)
    s32.set_kernel("sc100", "height 128")
    u64, r64 = s64.relax(np.zeros((3, nx, ny)), c, fire, max_steps=200000,
                         tol=1e-10, check_every=50)
    um, rm = s64.relax_refine(s32, np.zeros((3, nx, ny)), c, fire,
                              max_outer=30, tol=1e-10, inner_tol=1e-4)
    print(r64, rm)
    assert rm["converged"] and rm["residual"] / rm["fext_norm"] <= 1e-10
    assert np.abs(um - u64).max() < 1e-7 * np.abs(u64).max()


def test_refine_lj_sphere_hertz():
    """A nonlinear contact in the mixed-precision refinement: the
    reference's Hertz case (LJ sphere, 128^2) to the FP64 residual 1e-8 with
    FP32 inner FIRE; the FP64 FIRE state and the Hertz criterion."""
    nx = ny = 128
    u0 = np.zeros((3, nx, ny))
    u0[2] = -2.0
    fire = gfmd_hip.fire_params(dt=0.05, dtmax=0.1, adapt_dt=1)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_LJ_SPHERE, lattice=LATTICE,
                         sphere=SPHERE)
    s64 = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s64.set_kernel("sc100", "height 128")
    s32 = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP32)
    s32.set_kernel("sc100", "height 128")
    u64, r64 = s64.relax(u0, c, fire, max_steps=200000, tol=1e-8,
                         check_every=100)
    um, rm = s64.relax_refine(s32, u0, c, fire, max_outer=40, tol=1e-8,
                              inner_tol=1e-3)
    print(r64, rm)
    assert r64["converged"] and rm["converged"] and rm["outer"] >= 1
    assert rm["residual"] <= 1e-8
    assert np.abs(um - u64).max() < 1e-6 * np.abs(u64).max()
    f, _ = s64.post_force(um)
    assert hertz_residual(f[2])[0] < 1e-2


def test_lj_sphere_inside_is_an_error():
    """An atom inside the sphere is the reference's error (fix_contact_sphere
    .cpp:231-233), reported as a status code."""
    nx = ny = 64
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 8")
    c = gfmd_hip.contact(gfmd_hip.CONTACT_LJ_SPHERE, lattice=LATTICE,
                         sphere=SPHERE)
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        s.relax(np.zeros((3, nx, ny)), c, max_steps=10, tol=0.0)
    assert ei.value.code == gfmd_hip.ERR_CONTACT
