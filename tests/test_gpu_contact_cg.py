"""Hard-wall contact by constrained CG on the device (gfmd_contact_solve)
against the exact complementarity solution of the oracle (small grids),
against FIRE (the reference's relaxation, same acceptance test), and the
FP32 -> FP64 hand-off.

Tolerances: solved to relative residual 1e-10 (FP64 projected-force
residual, the FIRE definition); displacements within 1e-7 of the oracle's
(relative to max |u|), within 1e-6 of FIRE's.
"""
import numpy as np
import pytest

import gfmd_hip
import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def rough_wall(n, m, seed, amp=1.0):
    h = gfmd_hip.rough_surface(n, m, hurst=0.8, rms_slope=0.1,
                               lambda_short=4.0, seed=seed) * amp
    return h.max() - h


@pytest.mark.parametrize("kernel,args,prm,d,nx,ny,p0", [
    ("sc100", "height 8", [8], 3, 16, 16, 0.01),
    ("sc100", "height 5", [5], 3, 24, 20, 0.03),
    ("fcc100", "2 1.0 0.3 0.1 height 6",
     O.fcc100_params(6, 2, [1.0, 0.3, 0.1]), 6, 16, 12, 0.02),
])
def test_contact_solve_matches_exact_oracle(kernel, args, prm, d, nx, ny, p0):
    print("contact oracle", kernel, args, nx, ny, flush=True)
    kid = O.KERNEL_SC100 if kernel == "sc100" else O.KERNEL_FCC100
    phi = O.fill_phi(kid, prm, d, nx, ny, normalize=True)
    rng = np.random.default_rng(nx * ny)
    wall = 0.2 * rng.random((nx, ny))
    u_ref, p_ref = O.hardwall_contact(phi, wall, p0)
    assert 0 < (p_ref > 0).sum() < nx * ny
    s = gfmd_hip.Solver(nx, ny, d, gfmd_hip.FP64)
    s.set_kernel(kernel, args)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL, pressure=p0, wall=wall)
    u, r = s.contact_solve(c, tol=1e-10)
    assert r["converged"] and r["residual"] / r["fext_norm"] <= 1e-10
    assert rel(u, u_ref) < 1e-7
    assert r["ncontact"] == (p_ref > 0).sum()


def test_contact_solve_matches_fire_fast_path():
    n, p0 = 128, 5e-3
    wall = rough_wall(n, n, 1015)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL, pressure=p0, wall=wall)
    s = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height %d" % n)
    assert s.info()["fast_path"] == 1
    u, r = s.contact_solve(c, tol=1e-10)
    fire = gfmd_hip.fire_params(dt=0.1, dtmax=0.4, adapt_dt=1)
    uf, rf = s.relax(np.zeros((3, n, n)), c, fire, max_steps=200000,
                     tol=1e-10, check_every=20)
    print("128^2: CG iterations", r["steps"], "FIRE steps", rf["steps"],
          flush=True)
    assert r["converged"] and rf["converged"]
    assert rel(u, uf) < 1e-6
    assert r["ncontact"] == rf["ncontact"]


def test_contact_solve_mixed_precision_and_warm_start():
    n, p0 = 512, 5e-3
    wall = rough_wall(n, n, 77)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL, pressure=p0, wall=wall)
    s64 = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    s64.set_kernel("sc100", "height %d" % n)
    s32 = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP32)
    s32.set_kernel("sc100", "height %d" % n)
    u64, r64 = s64.contact_solve(c, tol=1e-8)
    um, rm = s64.contact_solve(c, s32=s32, tol=1e-8)
    print("512^2 FP64", r64["steps"], r64["seconds"], "mixed",
          rm["iters32"], rm["iters64"], rm["seconds"], flush=True)
    assert r64["converged"] and rm["converged"]
    assert rm["iters32"] > 0
    assert rel(um, u64) < 1e-6
    # warm start from the converged pressures: a handful of iterations
    f, _ = s64.post_force(u64)
    p = np.maximum(f[2] + p0, 0.0) * (u64[2] >= wall - 1e-12)
    uw, rw = s64.contact_solve(c, p_init=p, tol=1e-8)
    assert rw["converged"] and rw["steps"] <= 20
    assert rel(uw, u64) < 1e-6
    # with an FP32 handle: a start below the FP32 floor skips the FP32
    # stage (one FP64 check at most); one that meets tol is returned as is
    ux, rx = s64.contact_solve(c, s32=s32, p_init=p, tol=1e-8)
    assert rx["converged"] and rx["iters32"] == 0 and rx["steps"] <= 5
    assert rel(ux, u64) < 1e-6
    uy, ry = s64.contact_solve(c, s32=s32, p_init=p, tol=1e-3)
    assert ry["converged"] and ry["steps"] == 0
    assert ry["residual"] / ry["fext_norm"] <= 1e-3
    # a poor start still runs the FP32 stage
    uz, rz = s64.contact_solve(c, s32=s32, p_init=np.full((n, n), p0),
                               tol=1e-8)
    assert rz["converged"] and rz["iters32"] > 0
    assert rel(uz, u64) < 1e-6


def test_contact_solve_errors():
    s = gfmd_hip.Solver(16, 16, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 4")
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        s.contact_solve(gfmd_hip.contact(gfmd_hip.CONTACT_NONE, pressure=1e-3))
    assert ei.value.code == gfmd_hip.ERR_ARG
    s32 = gfmd_hip.Solver(16, 8, 3, gfmd_hip.FP32)
    s32.set_kernel("sc100", "height 4")
    c = gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL, pressure=1e-3,
                         wall=np.zeros((16, 16)))
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        s.contact_solve(c, s32=s32)
    assert ei.value.code == gfmd_hip.ERR_ARG


@pytest.mark.parametrize("mixed", [False, True])
def test_contact_batch_matches_single_solves(mixed):
    """Uneven ensemble (fast and generic grids, ndof 3 and 6) in one graph
    per iteration: each member equals its own single solve."""
    specs = [(64, 64, "sc100", "height 64", 3, 1001),
             (128, 128, "sc100", "height 128", 3, 1002),
             (96, 80, "sc100", "height 32", 3, 1003),
             (256, 256, "sc100", "height 256", 3, 1004),
             (64, 64, "fcc100", "2 1.0 0.3 0.1 height 16", 6, 1005),
             (128, 128, "sc100", "height 128", 3, 1006)]
    s64, s32, cs = [], [], []
    for nx, ny, k, a, d, seed in specs:
        h = gfmd_hip.Solver(nx, ny, d, gfmd_hip.FP64)
        h.set_kernel(k, a)
        s64.append(h)
        g = gfmd_hip.Solver(nx, ny, d, gfmd_hip.FP32)
        g.set_kernel(k, a)
        s32.append(g)
        cs.append(gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL, pressure=5e-3,
                                   wall=rough_wall(nx, ny, seed)))
    out = gfmd_hip.contact_solve_batch(s64, cs, s32 if mixed else None,
                                       tol=1e-10)
    for i, (u, r) in enumerate(out):
        assert r["converged"], (i, r)
        us, rs = s64[i].contact_solve(cs[i], tol=1e-10)
        assert rel(u, us) < 1e-7, i
        assert r["ncontact"] == rs["ncontact"]


@pytest.mark.parametrize("mixed", [False, True])
def test_ensemble_matches_single_solves(mixed):
    """Ensembles (members of one grid sharing the stiffness, one launch per
    sweep for all members) equal single solves member by member: the same
    kernels and reductions per member, so to round-off identity."""
    groups = [(64, 3), (128, 2)]
    ens, cs, protos = [], [], []
    for n, members in groups:
        a = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
        a.set_kernel("sc100", "height %d" % n)
        b = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP32)
        b.set_kernel("sc100", "height %d" % n)
        ens.append(gfmd_hip.Ensemble(a, members, b if mixed else None))
        protos.append((a, b))
        for m in range(members):
            cs.append(gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL,
                                       pressure=(3 + m) * 1e-3,
                                       wall=rough_wall(n, n, 2000 + 7 * m + n)))
    out = gfmd_hip.ensemble_contact_solve(ens, cs, tol=1e-10)
    i = 0
    for (a, b), (n, members) in zip(protos, groups):
        for m in range(members):
            u, r = out[i]
            assert r["converged"], (i, r)
            us, rs = a.contact_solve(cs[i], s32=b if mixed else None,
                                     tol=1e-10)
            assert rel(u, us) < 1e-12, i
            assert r["steps"] == rs["steps"]
            i += 1
    for e in ens:
        e.close()


def test_contact_solve_from_zero_pressure():
    """A start without any pressure (empty contact set) must progress:
    the first steps use the overlap set, then the usual iteration."""
    n, p0 = 128, 5e-3
    wall = rough_wall(n, n, 31)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL, pressure=p0, wall=wall)
    s = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height %d" % n)
    u, r = s.contact_solve(c, tol=1e-10)
    uz, rz = s.contact_solve(c, p_init=np.zeros((n, n)), tol=1e-10,
                             max_iter=2000)
    assert rz["converged"], rz
    assert rel(uz, u) < 1e-7
