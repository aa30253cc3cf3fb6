"""Damped velocity-Verlet dynamics on the device (one captured hipGraph per
step) against the CPU FP64 oracle integrator, and the reference's NVE
energy-conservation check (tests/TEST_energy_conservation_fcc100: total
energy within 1e-4 of its mean over 100000 steps).

Tolerances: FP64 trajectories (u, v) and energies within 1e-10 relative of
the oracle after 60 steps; FP32 NVE drift bound 1e-4 as the reference's.
"""
import numpy as np
import pytest

import gfmd_hip
import oracle as O

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(5)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("nx,ny,damping", [(16, 12, 0.0), (16, 12, 0.3),
                                           (64, 64, 0.0), (128, 64, 0.7)])
def test_dynamics_fp64_matches_oracle(nx, ny, damping):
    print("dynamics", nx, ny, damping, flush=True)
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 8")
    phi = O.fill_phi(O.KERNEL_SC100, [8], 3, nx, ny, normalize=True)
    u0 = 0.02 * RNG.standard_normal((3, nx, ny))
    v0 = 0.01 * RNG.standard_normal((3, nx, ny))
    n, dt, m, p = 60, 0.05, 1.2, 1e-3
    c = gfmd_hip.contact(pressure=p)
    u, v, h, res = s.dynamics(u0, v0, n, dt, mass=m, damping=damping,
                              record_every=1, contact=c)
    ur, vr, hr = O.dynamics(phi, u0, v0, n, dt, mass=m, damping=damping,
                            pressure=p)
    assert rel(u, ur) < 1e-10
    assert rel(v, vr) < 1e-10
    assert np.abs(h - hr).max() < 1e-10 * np.abs(hr).max()
    assert res["steps"] == n
    assert abs(res["ekin"] - hr[-1, 0]) < 1e-10 * abs(hr[-1, 0])


def test_dynamics_record_every_and_zero_steps():
    nx = ny = 16
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 4")
    phi = O.fill_phi(O.KERNEL_SC100, [4], 3, nx, ny, normalize=True)
    u0 = 0.02 * RNG.standard_normal((3, nx, ny))
    v0 = np.zeros_like(u0)
    _, _, h, _ = s.dynamics(u0, v0, 25, 0.05, record_every=5)
    _, _, hr = O.dynamics(phi, u0, v0, 25, 0.05)
    assert h.shape == (6, 3)
    assert np.abs(h - hr[::5]).max() < 1e-10 * np.abs(hr).max()
    u, v, h0, res = s.dynamics(u0, v0, 0, 0.05)
    assert np.array_equal(u, u0) and np.array_equal(v, v0)
    f, e = s.post_force(u0)
    assert abs(h0[0, 1] - e) < 1e-12 * abs(e) and h0[0, 0] == 0.0


@pytest.mark.parametrize("prec", [gfmd_hip.FP64, gfmd_hip.FP32])
def test_nve_energy_conservation_reference_case(prec):
    """fcc100 1 1.0 on 10 x 10 sites, displace_atoms random 0.1, timestep
    0.01, 100000 steps: |E - <E>| / <E> <= 1e-4."""
    nx = ny = 10
    s = gfmd_hip.Solver(nx, ny, 3, prec)
    s.set_kernel("fcc100", "1 1.0")
    rng = np.random.default_rng(12472634)
    u0 = rng.uniform(-0.1, 0.1, (3, nx, ny))
    u, v, h, res = s.dynamics(u0, np.zeros_like(u0), 100000, 0.01,
                              record_every=100)
    e = h[:, 2]
    drift = np.abs(e - e.mean()).max() / e.mean()
    print("NVE drift", prec, drift, "seconds", res["seconds"], flush=True)
    assert drift < 1e-4
    assert h[:, 0].max() > 0.1 * e.mean()


def test_nve_fp32_2048_drift_and_precision():
    """sc100 height 2048 at 2048^2, timestep 0.01 (the reference's): FP32
    energies within 1e-5 of the FP64 run, and both within the 1e-4 drift
    bound (at dt = 0.05 the Verlet fluctuation (w dt)^2 / 8 of the top
    modes alone is ~1e-3, so the timestep matters, not the precision)."""
    nx = ny = 2048
    u0 = 0.01 * RNG.standard_normal((3, nx, ny))
    out = {}
    for prec in (gfmd_hip.FP64, gfmd_hip.FP32):
        s = gfmd_hip.Solver(nx, ny, 3, prec)
        s.set_kernel("sc100", "height 2048")
        u, v, h, res = s.dynamics(u0, np.zeros_like(u0), 300, 0.01,
                                  record_every=10)
        e = h[:, 2]
        drift = np.abs(e - e.mean()).max() / e.mean()
        print("2048^2 NVE prec", prec, "drift", drift, flush=True)
        assert drift < 1e-4
        out[prec] = (u, h)
    (u64, h64), (u32, h32) = out[gfmd_hip.FP64], out[gfmd_hip.FP32]
    err_e = np.abs(h32 - h64).max() / np.abs(h64).max()
    err_u = rel(u32, u64)
    print("2048^2 NVE FP32 vs FP64: energy", err_e, "u", err_u, flush=True)
    assert err_e < 1e-5
    assert err_u < 1e-4


def test_dynamics_lj_sphere_conserves_energy():
    """Smooth LJ indenter (fix contact/sphere form) in NVE."""
    nx = ny = 64
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 16")
    u0 = np.zeros((3, nx, ny))
    v0 = np.zeros((3, nx, ny))
    v0[2] = 0.05 * np.exp(-((np.arange(nx)[:, None] - 32) ** 2 +
                            (np.arange(ny)[None, :] - 32) ** 2) / 50.0)
    # sphere of radius 20 centred above (32, 32) with its lowest point
    # 0.6 above the surface; LJ eps = 1e-3, sigma = 0.3 in the gap,
    # cutoff 1.0: the central sites start inside the attractive range
    c = gfmd_hip.contact(type=gfmd_hip.CONTACT_LJ_SPHERE,
                         sphere=(32.0, 32.0, 20.6, 20.0, 1e-3, 0.3, 1.0))
    u, v, h, res = s.dynamics(u0, v0, 4000, 0.02, record_every=10, contact=c)
    e = h[:, 2]
    assert np.abs(e - e.mean()).max() < 1e-4 * np.abs(e).max() + 1e-12
    assert res["ncontact"] > 0


def test_dynamics_errors():
    s = gfmd_hip.Solver(16, 16, 3, gfmd_hip.FP64)
    u = np.zeros((3, 16, 16))
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        s.dynamics(u, u, 10, 0.01)
    assert ei.value.code == gfmd_hip.ERR_STATE
    s.set_kernel("sc100", "")
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        s.dynamics(u, u, 10, 0.0)
    assert ei.value.code == gfmd_hip.ERR_ARG
    wall = np.zeros((16, 16))
    c = gfmd_hip.contact(type=gfmd_hip.CONTACT_HARDWALL, wall=wall)
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        s.dynamics(u, u, 10, 0.01, contact=c)
    assert ei.value.code == gfmd_hip.ERR_UNSUPPORTED
