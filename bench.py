"""Benchmark of the elastic-force hot path on one MI355X.

Workload (the metric's headline configuration): sc100 stiffness kernel,
height 2048, on a 2048 x 2048 periodic grid, ndof = 3; one step = one
elastic force + energy evaluation u(r) -> f(r), E with u, f resident in
HBM.  `value` is evaluations per second of the FP64 path (the reference's
precision: FP64 FFT, Phi.u and energy); the FP32 path (FP32 FFT and Phi.u,
FP64 energy) is reported beside it under "fp32".

Also reported: the roofline of each path's dominant kernel (algorithmic
bytes per launch over its mean duration, HIP events around each launch on
the solver's stream), the error of both paths against the CPU oracle on the
same table and field, the CPU baseline (the FP64 oracle, a C restatement of
the reference's GFMDSolverStatic::post_force, on one host core and on all
host cores, bounded samples) and the CPU configuration's time to solution
(128^2 FIRE under a rigid sphere to 1e-10 on one core), and the solver legs
(time to 1e-8 residual contact, two-layer fcc(100), two-layer fcc(111),
grids that are not powers of two, flat punch, ensemble, learned warm
start).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                    [--no-solve] [--only LEG]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "user-gfmd_amd"))

METRIC = ("elastic force evals/s at 2048²; time to 1e-8 residual contact "
          "vs plain FP64 GPU")
PEAK_HBM_GBS = 8000.0


def committed_traffic(key, size):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, calibrated against a copy) and recorded in profiles/."""
    path = os.path.join(ROOT, "profiles", "force_eval_sc100_%d.md" % size)
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "TRAFFIC" and parts[1] == key:
            return float(parts[2])
    return None


def alg_bytes(nx, ny, d, w, wphi, phi_compression=1, entries=None):
    """Algorithmic HBM bytes per sweep of one evaluation (three-sweep
    row-column model: rows r2c, columns FFT + Phi.u + inverse, rows c2r);
    the stiffness table counts `entries` reals per wavevector (d^2 packed,
    6 in the compact ndof-3 form) and 1/phi_compression of that (2 when its
    x-mirror symmetry lets the column sweep read half of it)."""
    n = nx * ny
    nh = nx * (ny // 2 + 1)
    entries = entries or d * d
    return {
        "rows_r2c": d * (n * w + 2 * nh * w),
        "cols": d * 4 * nh * w + nh * entries * wphi // phi_compression,
        "rows_c2r": d * (2 * nh * w + n * w),
    }


def gpu_bench(nx, ny, height, precision, steps, warmup, u_host, graph=False):
    """`steps` timed evaluations (after `warmup`) of one precision on the
    device field u_host (numpy FP64, converted to the handle's type), u and
    f resident.  Returns wall time per evaluation, per-sweep event means,
    the result f (host), the energy and, for FP64, the stiffness table."""
    import numpy as np
    import torch
    import gfmd_hip

    dt = torch.float32 if precision == gfmd_hip.FP32 else torch.float64
    s = gfmd_hip.Solver(nx, ny, 3, precision)
    t0 = time.time()
    s.set_kernel("sc100", "height %d" % height)
    t_phi = time.time() - t0
    variant = s.autotune(20)
    u = torch.from_numpy(u_host).to(dt).cuda()
    f = torch.empty_like(u)
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    for _ in range(warmup):
        s.post_force_device(u.data_ptr(), f.data_ptr(), e.data_ptr())
    s.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.post_force_device(u.data_ptr(), f.data_ptr(), e.data_ptr())
    s.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = s.time_force_device(u.data_ptr(), f.data_ptr(), e.data_ptr(), steps)
    # the same evaluations launched eagerly and as replays of a captured
    # graph of 10 (HIP events; the replay's output equals the eager one)
    gms = (s.time_force_graph(u.data_ptr(), f.data_ptr(), e.data_ptr(),
                              steps - steps % 10, 10) if graph else None)
    energy = e.item()
    assert np.isfinite(energy)
    comp = 2 if s.phi_mirror() else 1
    entries = s.phi_entries()
    table = s.dump_stiffness() if precision == gfmd_hip.FP64 else None
    f_host = f.double().cpu().numpy()
    s.close()
    w = 4 if precision == gfmd_hip.FP32 else 8
    return {"wall_s": wall, "ms_per_eval": 1e3 * wall / steps, "sweep_ms": ms,
            "bytes": alg_bytes(nx, ny, 3, w, w, comp, entries),
            "phi_setup_s": t_phi, "phi_compression": comp,
            "phi_entries": entries, "f": f_host, "table": table,
            "energy": energy, "variant": variant[:2], "graph_ms": gms}


def roofline(r, nx, tag):
    """Roofline object of the dominant sweep of one precision: algorithmic
    bytes per launch over the launch's mean duration (HIP events
    bracketing each launch on the solver's stream), against 8 TB/s."""
    names = ["rows_r2c", "cols", "rows_c2r"]
    sw = r["sweep_ms"]
    dom = max(range(3), key=lambda i: sw[i])
    b = r["bytes"][names[dom]]
    achieved = b / (sw[dom] * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
            "traffic": committed_traffic(names[dom] + "_" + tag, nx),
            "kernel": names[dom], "kernel_ms": sw[dom], "alg_bytes": b,
            "duration_source": "HIP events around each launch, mean of the "
                               "timed evaluations"}


def time_to_solution(n, tol=1e-8, pressure=5e-3, seed=1015):
    """Contact of the sc100 (height n) surface with a rigid self-affine rough
    surface (H 0.8, rms slope 0.1, short-wavelength cutoff 4a, seeded) under
    a normal load of `pressure` per site, from u = 0 to relative residual
    `tol` (FP64 projected-force residual for every leg).  Legs: plain FP64
    FIRE (the reference's relaxation), FP32 FIRE with FP64 refinement, FP64
    constrained CG on the contact pressure, and CG with an FP32 stage handed
    to FP64.  Times: 'device_s' with inputs resident (the solve), 'call_s'
    including the host transfers of the wall map and the result."""
    import numpy as np
    import gfmd_hip
    h = gfmd_hip.rough_surface(n, n, hurst=0.8, rms_slope=0.1,
                               lambda_short=4.0, seed=seed)
    wall = h.max() - h  # rigid surface above; its lowest point touches z = 0
    c = gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL, pressure=pressure,
                         wall=wall)
    fire = gfmd_hip.fire_params(dt=0.1, dtmax=0.4, adapt_dt=1)
    s64 = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    s64.set_kernel("sc100", "height %d" % n)
    s64.autotune(10)
    s32 = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP32)
    s32.set_kernel("sc100", "height %d" % n)
    s32.autotune(10)
    u0 = np.zeros((3, n, n))

    def leg(r, u, extra=None):
        d = {"device_s": r["device_seconds"], "call_s": r["seconds"],
             "iterations": r["steps"],
             "residual": r["residual"] / r["fext_norm"],
             "converged": bool(r["converged"]),
             "max_rel_diff_u_vs_fire64": float(np.abs(u - u64).max() /
                                               np.abs(u64).max())}
        d.update(extra or {})
        return d

    u64, r64 = s64.relax(u0, c, fire, max_steps=500000, tol=tol,
                         check_every=20)
    um, rm = s64.relax_refine(s32, u0, c, fire, max_outer=60, tol=tol,
                              inner_tol=1e-2, check_every=20)
    s64.contact_solve(c, tol=tol)                 # first call: tables, graph
    s64.contact_solve(c, s32=s32, tol=tol)
    uc, rc = s64.contact_solve(c, tol=tol)
    ux, rx = s64.contact_solve(c, s32=s32, tol=tol)
    out = {"workload": "sc100 height %d, %dx%d, rigid rough surface "
                       "(H 0.8, rms slope 0.1, cutoff 4a, seed %d), hard "
                       "wall, load %g per site, from u = 0 to relative "
                       "residual %g" % (n, n, n, seed, pressure, tol),
           "fire_fp64": leg(r64, u64),
           "fire_fp32_refine": leg(rm, um, {"outer": rm["outer"]}),
           "cg_fp64": leg(rc, uc),
           "cg_fp32_fp64": leg(rx, ux, {"iters_fp32": rx["iters32"],
                                        "iters_fp64": rx["iters64"]}),
           "contact_fraction": rx["ncontact"] / (n * n)}
    out["speedup_vs_fire_fp64"] = {
        k: r64["device_seconds"] / out[k]["device_s"]
        for k in ("fire_fp32_refine", "cg_fp64", "cg_fp32_fp64")}
    return out


def fcc100_two_layer(n=1024, steps=200, warmup=20, sigma=0.02, seed=2):
    """Two-layer fcc(100) surface, nearest-neighbour springs of the 12-6 LJ
    potential (eps = sigma = 1, r_nn = 2^(1/6): k = V''(r_nn) = 57.146,
    kappa = 0), `ft fcc100 1.0 2 pair-potential 1 57.146 height n`, ndof 6
    (a 6 x 6 Hermitian Phi per wavevector), n x n periodic grid,
    displacements Gaussian (sigma 0.02, seeded).  Evaluations per second in
    FP32 and FP64 (u, f resident), and the FP32 result against the FP64
    one: max and rms relative force error, relative energy error."""
    import numpy as np
    import torch
    import gfmd_hip
    args = "fcc100 1.0 2 pair-potential 1 57.146 height %d" % n
    u = sigma * np.random.default_rng(seed).standard_normal((6, n, n))
    out = {"workload": "ft %s, %dx%d periodic, ndof 6, u ~ N(0, %g^2) "
                       "seed %d" % (args, n, n, sigma, seed)}
    res = {}
    for prec, name, dt in ((gfmd_hip.FP32, "fp32", torch.float32),
                           (gfmd_hip.FP64, "fp64", torch.float64)):
        s = gfmd_hip.Solver(n, n, 6, prec)
        s.set_kernel("ft", args)
        s.autotune(10)
        du = torch.from_numpy(u).to(dt).cuda()
        df = torch.empty_like(du)
        de = torch.zeros(1, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        for _ in range(warmup):
            s.post_force_device(du.data_ptr(), df.data_ptr(), de.data_ptr())
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.post_force_device(du.data_ptr(), df.data_ptr(), de.data_ptr())
        s.synchronize()
        wall = time.perf_counter() - t0
        res[name] = (df.double().cpu().numpy(), de.item())
        out[name] = {"evals_per_s": steps / wall,
                     "ms_per_eval": 1e3 * wall / steps,
                     "phi_x_mirror": s.phi_mirror()}
        s.close()
    f64, e64 = res["fp64"]
    f32, e32 = res["fp32"]
    d = f32 - f64
    out["fp32_vs_fp64"] = {
        "max_rel_force": float(np.abs(d).max() / np.abs(f64).max()),
        "rms_rel_force": float(np.sqrt(np.mean(d * d) / np.mean(f64 * f64))),
        "rel_energy": abs(e32 - e64) / abs(e64)}
    return out


def fcc111_two_layer(n=1024, steps=100, warmup=10, sigma=0.02, seed=4):
    """Two-layer fcc(111) surface (four atoms per cell, ndof 12, a 12 x 12
    Hermitian Phi per wavevector), nearest-neighbour springs,
    `ft fcc111 1.0 2 pair-potential 1 1.0 height n`, n x n periodic grid,
    displacements Gaussian (sigma 0.02, seeded).  Evaluations per second of
    the FFT sweeps in FP32 and FP64 (u, f resident), the column kernel's
    roofline (algorithmic bytes over its HIP-event duration), and the FP32
    result against the FP64 one."""
    import numpy as np
    import torch
    import gfmd_hip
    args = "fcc111 1.0 2 pair-potential 1 1.0 height %d" % n
    d = 12
    u = sigma * np.random.default_rng(seed).standard_normal((d, n, n))
    out = {"workload": "ft %s, %dx%d periodic, ndof %d, u ~ N(0, %g^2) "
                       "seed %d" % (args, n, n, d, sigma, seed)}
    res = {}
    for prec, name, dt, w in ((gfmd_hip.FP32, "fp32", torch.float32, 4),
                              (gfmd_hip.FP64, "fp64", torch.float64, 8)):
        s = gfmd_hip.Solver(n, n, d, prec)
        s.set_kernel("ft", args)
        fast = s.info()["fast_path"]
        s.autotune(5)
        du = torch.from_numpy(u).to(dt).cuda()
        df = torch.empty_like(du)
        de = torch.zeros(1, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        for _ in range(warmup):
            s.post_force_device(du.data_ptr(), df.data_ptr(), de.data_ptr())
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            s.post_force_device(du.data_ptr(), df.data_ptr(), de.data_ptr())
        s.synchronize()
        wall = time.perf_counter() - t0
        ms = s.time_force_device(du.data_ptr(), df.data_ptr(), de.data_ptr(),
                                 steps)
        comp = 2 if s.phi_mirror() else 1
        b = alg_bytes(n, n, d, w, w, comp, s.phi_entries())
        res[name] = (df.double().cpu().numpy(), de.item())
        out[name] = {"fast_path": fast == 1,
                     "evals_per_s": steps / wall,
                     "ms_per_eval": 1e3 * wall / steps,
                     "sweep_ms": dict(zip(["rows_r2c", "cols", "rows_c2r",
                                           "eval"], ms)),
                     "phi_x_mirror": bool(comp == 2),
                     "roofline_cols": {
                         "alg_bytes": b["cols"],
                         "achieved_GBs": b["cols"] / (ms[1] * 1e-3) / 1e9,
                         "frac": b["cols"] / (ms[1] * 1e-3) / 1e9 /
                         PEAK_HBM_GBS},
                     "eval_alg_GBs": sum(b.values()) / (ms[3] * 1e-3) / 1e9}
        s.close()
    f64, e64 = res["fp64"]
    f32, e32 = res["fp32"]
    dd = f32 - f64
    out["fp32_vs_fp64"] = {
        "max_rel_force": float(np.abs(dd).max() / np.abs(f64).max()),
        "rel_energy": abs(e32 - e64) / abs(e64)}
    return out


def mixed_radix(grids=((63, 63), (64, 37), (1000, 1000), (2000, 2000),
                        (3000, 3000)), steps=50, warmup=5, seed=6):
    """Grids that are not powers of two (the mixed-radix FFT path: one
    workgroup per line in LDS, Stockham over the line's factors): sc100,
    height = nx, ndof 3, u ~ N(0, 0.01^2); evaluations per second with u, f
    resident, FP32 and FP64, and the algorithmic rate of the three-sweep
    model (what the power-of-two path moves; this path makes two more passes
    over the spectrum)."""
    import numpy as np
    import torch
    import gfmd_hip
    out = {}
    for nx, ny in grids:
        u = 0.01 * np.random.default_rng(seed).standard_normal((3, nx, ny))
        row = {}
        for prec, name, dt, w in ((gfmd_hip.FP32, "fp32", torch.float32, 4),
                                  (gfmd_hip.FP64, "fp64", torch.float64, 8)):
            s = gfmd_hip.Solver(nx, ny, 3, prec)
            s.set_kernel("sc100", "height %d" % nx)
            du = torch.from_numpy(u).to(dt).cuda()
            df = torch.empty_like(du)
            de = torch.zeros(1, dtype=torch.float64, device="cuda")
            for _ in range(warmup):
                s.post_force_device(du.data_ptr(), df.data_ptr(),
                                    de.data_ptr())
            s.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                s.post_force_device(du.data_ptr(), df.data_ptr(),
                                    de.data_ptr())
            s.synchronize()
            wall = (time.perf_counter() - t0) / steps
            b = sum(alg_bytes(nx, ny, 3, w, w, 1, 9).values())
            row[name] = {"path": s.info()["fast_path"],
                         "evals_per_s": 1.0 / wall, "ms_per_eval": 1e3 * wall,
                         "eval_alg_GBs": b / wall / 1e9}
            s.close()
        out["%dx%d" % (nx, ny)] = row
        print("mixed radix %dx%d: FP32 %.0f evals/s, FP64 %.0f evals/s"
              % (nx, ny, row["fp32"]["evals_per_s"],
                 row["fp64"]["evals_per_s"]), file=sys.stderr, flush=True)
    return out


def flat_punch(n=4096, patch=2048, radius=512.0, pressure=1e-3, tol=1e-8,
               nu=0.3, mu=1.0, fire_steps=200000):
    """Open-boundary surface (`nonperiodic nu mu`, isotropic continuum with
    Li & Berger's Green's function) of patch x patch sites in an n x n grid
    whose padding sites are free, pressed by a rigid flat cylindrical punch
    of `radius` at total load pressure * patch^2.  The punch is a hard wall
    at -delta over r <= radius (free elsewhere): the solution is linear in
    delta, so the solve runs at Sneddon's delta for that load and the result
    is scaled to the load exactly.  Relative residual = |f_total| /
    |wall reaction| (the absolute tolerance uses the lower bound
    load / sqrt(punch sites) of |wall reaction|; the value reported uses
    the solution's own).  Legs as for the rough contact: FP64 FIRE, FP32
    FIRE with FP64 refinement, FP64 constrained CG, CG with an FP32 stage;
    every displacement is compared with the FP64 CG solution."""
    import time
    import numpy as np
    import gfmd_hip
    args = "%g %g" % (nu, mu)
    e_star = 2 * mu * (1 + nu) / (1 - nu * nu)
    load = pressure * patch * patch
    delta0 = load / (2 * radius * e_star)
    i = np.arange(n) - n // 2
    r = np.sqrt(i[:, None] ** 2 + i[None, :] ** 2)
    punch = r <= radius
    wall = np.where(punch, -delta0, 1e30)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL, pressure=0.0, wall=wall)
    s64 = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    t0 = time.perf_counter()                      # the table build alone
    s64.set_kernel("nonperiodic", args)
    t_phi = time.perf_counter() - t0
    s64.autotune(5)
    s32 = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP32)
    s32.set_kernel("nonperiodic", args)
    s32.autotune(5)
    tol_abs = tol * load / np.sqrt(punch.sum())

    s64.contact_solve(c, tol=tol_abs)             # first call: tables, graph
    s64.contact_solve(c, s32=s32, tol=tol_abs)
    uc, rc = s64.contact_solve(c, tol=tol_abs)
    ux, rx = s64.contact_solve(c, s32=s32, tol=tol_abs)
    f, _ = s64.post_force(uc)
    p = np.where(punch, f[2], 0.0)
    pnorm = float(np.sqrt(np.sum(p * p)))
    f0 = float(p.sum())
    scale = load / f0
    # Sneddon's pressure under the punch, p(r) = F / (2 pi a sqrt(a^2 - r^2))
    inner = r < 0.9 * radius
    ps = load / (2 * np.pi * radius * np.sqrt(radius ** 2 - r[inner] ** 2))
    prof = float(np.sqrt(np.mean((p[inner] * scale - ps) ** 2)) /
                 np.sqrt(np.mean(ps ** 2)))

    u0 = np.zeros((3, n, n))
    u0[2] = np.minimum(0.0, wall)
    fire = gfmd_hip.fire_params(dt=0.1, dtmax=0.4, adapt_dt=1)
    uf, rf = s64.relax(u0, c, fire, max_steps=fire_steps, tol=tol_abs,
                       check_every=20)
    um, rm = s64.relax_refine(s32, u0, c, fire, max_outer=60, tol=tol_abs,
                              inner_tol=1e-2, check_every=20)

    def leg(res, u, extra=None):
        d = {"device_s": res["device_seconds"], "call_s": res["seconds"],
             "iterations": res["steps"],
             "residual": res["residual"] / pnorm,
             "converged": bool(res["converged"]),
             "max_rel_diff_u_vs_cg64": float(np.abs(u - uc).max() /
                                             np.abs(uc).max())}
        d.update(extra or {})
        return d

    out = {"workload": "nonperiodic %s, %dx%d surface patch in a %dx%d grid "
                       "(padding free), rigid flat punch radius %g at load "
                       "%g per patch site (%g total), from u = 0 to relative "
                       "residual %g" % (args, patch, patch, n, n, radius,
                                        pressure, load, tol),
           "phi_setup_s": t_phi,
           "fire_fp64": leg(rf, uf),
           "fire_fp32_refine": leg(rm, um, {"outer": rm["outer"]}),
           "cg_fp64": leg(rc, uc),
           "cg_fp32_fp64": leg(rx, ux, {"iters_fp32": rx["iters32"],
                                        "iters_fp64": rx["iters64"]}),
           "punch_sites": int(punch.sum()),
           "delta": delta0 * scale,
           "delta_over_sneddon": scale,
           "pressure_profile_rel_rms_vs_sneddon_r_lt_0.9a": prof}
    out["speedup_vs_fire_fp64"] = {
        k: rf["device_seconds"] / out[k]["device_s"]
        for k in ("fire_fp32_refine", "cg_fp64", "cg_fp32_fp64")}
    return out


def wall_map(n=2048, pressure=1e-3, gap=1.0, tol=1e-8, tol32=1e-3,
             fire_steps=100000):
    """Soft rough contact (`fix wall/map`): a seeded self-affine map (the
    rough-surface generator, H = 0.8, rms slope 0.1, shortest wavelength 8)
    as a bicubic wall above an sc100 surface of n x n sites (height n), its
    lowest point `gap` above the surface (placement `bot`, normal -1),
    lj93/smooth (epsilon 0.01, sigma 0.8, cutoffs 1.4 / 2.2), load
    `pressure` per site.  FIRE (xlimit 0.05) in FP64 to relative residual
    tol and in FP32 to tol32; the wall is fused into the per-step kernels,
    so a step costs one force evaluation plus the two solver passes."""
    import numpy as np
    import gfmd_hip
    h = gfmd_hip.rough_surface(n, n, hurst=0.8, rms_slope=0.1,
                               lambda_short=8.0, seed=3)
    args = "rough.map bot %g -1 lj93/smooth 0.01 0.8 1.4 2.2" % gap
    w = gfmd_hip.WallMap(args, float(n), float(n), heights=h)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_WALL_MAP, pressure=pressure,
                         wall_map=w)
    fire = gfmd_hip.fire_params(dt=0.1, dtmax=0.3, adapt_dt=1, xlimit=0.05)
    out = {"workload": "sc100 height %d, %dx%d periodic, wall/map %s over a "
                       "self-affine map (H 0.8, rms slope 0.1, lambda_s 8), "
                       "load %g per site, FIRE from u = 0" %
                       (n, n, n, args, pressure),
           "map": w.info()}
    u64 = None
    for name, prec, t in (("fp64", gfmd_hip.FP64, tol),
                          ("fp32", gfmd_hip.FP32, tol32)):
        s = gfmd_hip.Solver(n, n, 3, prec)
        s.set_kernel("sc100", "height %d" % n)
        s.autotune(5)
        s.relax(np.zeros((3, n, n)), c, fire, max_steps=20, tol=0.0,
                check_every=20)  # first call: buffers, launch set-up
        print("wall_map %s: FIRE to %g" % (name, t), file=sys.stderr,
              flush=True)
        u, r = s.relax(np.zeros((3, n, n)), c, fire, max_steps=fire_steps,
                       tol=t, check_every=100)
        d = {"tol": t, "converged": bool(r["converged"]),
             "steps": r["steps"], "device_s": r["device_seconds"],
             "ms_per_step": 1e3 * r["device_seconds"] / max(r["steps"], 1),
             "residual": r["residual"] / r["fext_norm"],
             "ncontact": r["ncontact"]}
        if u64 is None:
            u64 = u
        else:
            d["max_rel_diff_u_vs_fp64"] = float(np.abs(u - u64).max() /
                                                np.abs(u64).max())
        out[name] = d
        print("wall_map %s: %s" % (name, d), file=sys.stderr, flush=True)
        if prec == gfmd_hip.FP64:
            s64 = s
    # FP32 inner FIRE + FP64 corrections to the FP64 tolerance
    s32 = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP32)
    s32.set_kernel("sc100", "height %d" % n)
    s32.autotune(5)
    print("wall_map fp32_refine: to %g" % tol, file=sys.stderr, flush=True)
    u, r = s64.relax_refine(s32, np.zeros((3, n, n)), c, fire, max_outer=60,
                            tol=tol, inner_tol=1e-2, check_every=50)
    d = {"tol": tol, "converged": bool(r["converged"]), "outer": r["outer"],
         "inner_steps": r["steps"], "device_s": r["device_seconds"],
         "residual": r["residual"] / r["fext_norm"],
         "max_rel_diff_u_vs_fp64": float(np.abs(u - u64).max() /
                                         np.abs(u64).max())}
    d["speedup_vs_fp64"] = out["fp64"]["device_s"] / d["device_s"]
    out["fp32_refine"] = d
    print("wall_map fp32_refine: %s" % d, file=sys.stderr, flush=True)
    return out


def ensemble(tol=1e-8, pressure=5e-3, per_size=16,
             sizes=(256, 512, 1024, 2048)):
    """Batched ensemble: per_size members of each grid size (sc100, height
    = grid size), each against its own rigid self-affine rough surface (H
    0.8, rms slope 0.1, cutoff 4a, seed 1000 + member index) at load
    `pressure` per site, FP32 stage + FP64 finish.  One ensemble per size
    (members share the stiffness, each sweep is one launch over all
    members), all four in one captured graph per iteration, sweep variants
    autotuned per shape.  Compared with the same members solved one after
    another on the prototype handles."""
    import numpy as np
    import gfmd_hip
    ens, cs, protos = [], [], []
    i = 0
    for n in sizes:
        a = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
        a.set_kernel("sc100", "height %d" % n)
        a.autotune(5)
        b = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP32)
        b.set_kernel("sc100", "height %d" % n)
        b.autotune(5)
        protos.append((a, b))
        ens.append(gfmd_hip.Ensemble(a, per_size, b))
        for _ in range(per_size):
            h = gfmd_hip.rough_surface(n, n, hurst=0.8, rms_slope=0.1,
                                       lambda_short=4.0, seed=1000 + i)
            cs.append(gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL,
                                       pressure=pressure, wall=h.max() - h))
            i += 1
    gfmd_hip.ensemble_contact_solve(ens, cs, tol=tol)   # first: graphs
    out = gfmd_hip.ensemble_contact_solve(ens, cs, tol=tol)
    batch_s = out[0][1]["device_seconds"]
    seq_s, diff, k = 0.0, 0.0, 0
    seq_size = {}
    for j, n in enumerate(sizes):
        a, b = protos[j]
        a.contact_solve(cs[k], s32=b, tol=tol)            # per-handle graph
        seq_size[n] = 0.0
        for _ in range(per_size):
            u, r = a.contact_solve(cs[k], s32=b, tol=tol)
            seq_s += r["device_seconds"]
            seq_size[n] += r["device_seconds"]
            diff = max(diff, float(np.abs(u - out[k][0]).max() /
                                   np.abs(u).max()))
            k += 1
    res = [r["residual"] / r["fext_norm"] for _, r in out]
    for e in ens:
        e.close()
    return {"workload": "%d sc100 surfaces (%s x %d each), rigid rough "
                        "surfaces (H 0.8, rms slope 0.1, cutoff 4a, seeds "
                        "1000..%d), hard wall, load %g per site, to relative "
                        "residual %g, FP32 stage + FP64 finish"
                        % (len(cs), "/".join("%d^2" % n for n in sizes),
                           per_size, 999 + len(cs), pressure, tol),
            "members": len(cs),
            "batch_device_s": batch_s,
            "sequential_device_s": seq_s,
            # one after another, per grid size: the largest members' own
            # time is a floor for any batching of this mix
            "sequential_device_s_per_size": {"%d^2" % n: t
                                             for n, t in seq_size.items()},
            "batch_speedup": seq_s / batch_s,
            "all_converged": all(r["converged"] for _, r in out),
            "max_residual": max(res),
            "max_iterations": max(r["steps"] for _, r in out),
            "max_rel_diff_u_batch_vs_single": diff}


def grid_sizes(sizes=(128, 512, 1024, 4096), steps=200, warmup=20,
               cpu_sizes=(128, 1024), cpu=True):
    """The headline evaluation (sc100, height = grid size, ndof 3) at other
    grid sizes: FP64 and FP32 evaluations per second on the device (u, f
    resident, `steps` after `warmup`, autotuned; and per evaluation by HIP
    events, launched eagerly and as replays of a captured graph of 10
    evaluations), and the FP64 oracle on
    one host core at cpu_sizes (median of 20 evaluations after 3 warm-ups,
    10 at 1024^2 and above)."""
    import numpy as np
    import gfmd_hip
    out = {}
    for n in sizes:
        u = 0.01 * np.random.default_rng(n).standard_normal((3, n, n))
        row = {}
        for name, prec in (("fp64", gfmd_hip.FP64), ("fp32", gfmd_hip.FP32)):
            r = gpu_bench(n, n, n, prec, steps, warmup, u, graph=True)
            row[name] = {"evals_per_s": 1e3 / r["ms_per_eval"],
                         "ms_per_eval": r["ms_per_eval"],
                         "events_ms_per_eval_eager": r["graph_ms"][0],
                         "events_ms_per_eval_graph": r["graph_ms"][1],
                         "alg_GBs": sum(r["bytes"].values()) /
                         (r["sweep_ms"][3] * 1e-3) / 1e9}
        out["%d^2" % n] = row
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        c = {}
        for n in cpu_sizes:
            phi = O.fill_phi(O.KERNEL_SC100, [0], 3, n, n, normalize=True)
            u = 0.01 * np.random.default_rng(1).standard_normal((3, n, n))
            for _ in range(3):
                O.static_force(phi, u)
            ts = []
            for _ in range(20 if n < 1024 else 10):
                t0 = time.perf_counter()
                O.static_force(phi, u)
                ts.append(time.perf_counter() - t0)
            c["%d^2" % n] = {"evals_per_s": 1.0 / float(np.median(ts)),
                             "evaluations": len(ts)}
        out["cpu"] = {"kind": "port", "cores": 1, "cpu_model": cpu_model()[0],
                      "sample": "the oracle's FP64 evaluation (own FFT), "
                                "median per evaluation", "sizes": c}
    return out


def aux_paths(n=2048, cpu=True):
    """The paths around the force evaluation, each timed at n^2 (sc100
    height n, ndof 3) with its algorithmic bytes:
    - damped velocity Verlet (`gfmd_dynamics`, one captured graph per step:
      the three force sweeps, then `k_dyn_step` reading f, v, u and writing
      v, u): device time per step from the difference of a 10- and a
      110-step call (their copies and graph capture cancel);
    - the gradient preconditioner (`gfmd_prec_gradient_device`: the force
      sweeps with the inverse table, full packed 9 reals per q, half read):
      100 calls on device buffers between two synchronisations;
    - fix gfmd's atom <-> grid glue (`gfmd_fix_pre_force`: centre-of-mass
      pass and gather of u and v; `gfmd_fix_post_force`: scatter-add of the
      grid force), one atom per site, FP64, 100 calls each; atoms in grid
      order and in a random order.
    CPU beside them (one core, the oracle): dynamics and the preconditioner
    at 512^2, per site; the glue has no CPU baseline (the oracle's is a
    per-atom Python loop)."""
    import time
    import numpy as np
    import torch
    import gfmd_hip
    out = {"grid": [n, n]}
    N = n * n
    rng = np.random.default_rng(11)
    u0 = 0.01 * rng.standard_normal((3, n, n))
    v0 = 0.01 * rng.standard_normal((3, n, n))
    # dynamics, FP64 and FP32
    dyn = {}
    for name, prec, es in (("fp64", gfmd_hip.FP64, 8), ("fp32", gfmd_hip.FP32, 4)):
        s = gfmd_hip.Solver(n, n, 3, prec)
        s.set_kernel("sc100", "height %d" % n)
        s.autotune(5)
        s.dynamics(u0, v0, 10, 0.05, record_every=10)     # warm-up
        t = {}
        for k in (10, 110):
            _, _, _, r = s.dynamics(u0, v0, k, 0.05, record_every=k)
            t[k] = r["seconds"]
        ms = (t[110] - t[10]) / 100 * 1e3
        # the three sweeps move the field or the half spectrum six times,
        # 3 es bytes per site each (u in; spectrum out, in, out, in; f out):
        # 18 es N, plus half of the compact table (6 reals per q): 1.5 es N;
        # k_dyn_step reads f, v, u and writes v, u: 15 es N
        eval_b = 19.5 * es * N
        step_b = 15 * es * N
        dyn[name] = {"ms_per_step": ms, "steps_per_s": 1e3 / ms,
                     "alg_bytes_per_step": eval_b + step_b,
                     "alg_GBs": (eval_b + step_b) / (ms * 1e-3) / 1e9}
        s.close()
    out["dynamics"] = dyn
    # preconditioner, FP64
    s = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height %d" % n)
    s.autotune(5)
    g = torch.tensor(u0, device="cuda")
    gp = torch.empty_like(g)
    s.prec_gradient_device(g.data_ptr(), gp.data_ptr())
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(100):
        s.prec_gradient_device(g.data_ptr(), gp.data_ptr())
    s.synchronize()
    ms = (time.perf_counter() - t0) / 100 * 1e3
    # r2c (u read, spectrum written), columns (spectrum read and written,
    # half of the 9-real table), c2r (spectrum read, gP written): 48 bytes
    # per site each, plus the table
    tab = 9 * 8 * n * (n // 2 + 1) / 2
    b = 3 * 48 * N + tab
    out["prec_gradient"] = {"fp64": {"ms_per_call": ms, "alg_bytes": b,
                                     "alg_GBs": b / (ms * 1e-3) / 1e9}}
    s.close()
    # atom <-> grid glue, FP64, one atom per site
    glue = {}
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    xeq0 = np.column_stack([X.ravel() * 1.0, Y.ravel() * 1.0, np.zeros(N)])
    gid0 = np.column_stack([X.ravel(), Y.ravel(), np.zeros(N)]).astype(np.int32)
    for order in ("grid", "random"):
        perm = np.arange(N) if order == "grid" else rng.permutation(N)
        xeq, gid = xeq0[perm], gid0[perm]
        x = xeq + 0.05 * rng.standard_normal((N, 3))
        img = np.full(N, 512 | (512 << 10) | (512 << 20), np.int32)
        dv = {"gid": torch.tensor(gid, device="cuda"),
              "x": torch.tensor(x, device="cuda"),
              "xeq": torch.tensor(xeq, device="cuda"),
              "v": torch.tensor(rng.standard_normal((N, 3)), device="cuda"),
              "img": torch.tensor(img, device="cuda"),
              "mass": torch.ones(N, dtype=torch.float64, device="cuda")}
        A = gfmd_hip.Atoms(N, None, 0, dv["gid"].data_ptr(), dv["x"].data_ptr(),
                           dv["xeq"].data_ptr(), dv["v"].data_ptr(),
                           dv["img"].data_ptr(), dv["mass"].data_ptr())
        s = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
        s.fix_setup(A, (n * 1.0, n * 1.0, 100.0), np.eye(3))
        ug = torch.zeros((3, n, n), dtype=torch.float64, device="cuda")
        vg = torch.zeros_like(ug)
        fa = torch.zeros((N, 3), dtype=torch.float64, device="cuda")
        res = {}
        for name, fn in (("pre_force", lambda: s.fix_pre_force(
                              A, ug.data_ptr(), vg.data_ptr())),
                         ("post_force", lambda: s.fix_post_force(
                              A, ug.data_ptr(), fa.data_ptr()))):
            fn()
            s.synchronize()
            t0 = time.perf_counter()
            for _ in range(100):
                fn()
            s.synchronize()
            res[name] = (time.perf_counter() - t0) / 100 * 1e3
        # bytes per atom: centre of mass (mass, x, image) + gather (gid, x,
        # xeq, v read; u, v grid written); scatter (gid, f grid, f atoms
        # read; f atoms written)
        bpre, bpost = (8 + 24 + 4 + 12 + 24 + 24 + 24 + 48) * N, (12 + 24 + 48) * N
        glue[order] = {"pre_force_ms": res["pre_force"],
                       "post_force_ms": res["post_force"],
                       "pre_force_GBs": bpre / (res["pre_force"] * 1e-3) / 1e9,
                       "post_force_GBs": bpost / (res["post_force"] * 1e-3) / 1e9,
                       "alg_bytes": {"pre_force": bpre, "post_force": bpost}}
        s.close()
        del dv
    glue["atoms"] = N
    out["fix_glue"] = glue
    if cpu:
        # the CPU baseline: the oracle restatement, timed beside the device
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        m = 512
        phi = O.fill_phi(O.KERNEL_SC100, [m], 3, m, m, normalize=True)
        uc = 0.01 * rng.standard_normal((3, m, m))
        vc = np.zeros_like(uc)
        O.dynamics(phi, uc, vc, 1, 0.05)
        t0 = time.perf_counter()
        O.dynamics(phi, uc, vc, 10, 0.05)
        td = (time.perf_counter() - t0) / 10
        t0 = time.perf_counter()
        for _ in range(5):
            O.prec_gradient(phi, uc)
        tp = (time.perf_counter() - t0) / 5
        out["cpu"] = {"kind": "port", "cores": 1, "cpu_model": cpu_model()[0],
                      "sample": "10 damped-Verlet steps and 5 preconditioner "
                                "calls at %d^2 (oracle: C force evaluation "
                                "with its own FFT; numpy FFT and solve for "
                                "the preconditioner)" % m,
                      "dynamics_ms_per_step": td * 1e3,
                      "prec_gradient_ms_per_call": tp * 1e3,
                      "dynamics_ns_per_site_step": td / (m * m) * 1e9,
                      "prec_gradient_ns_per_site": tp / (m * m) * 1e9}
    return out


def warm_start_transfer(ws, n, held, tol):
    """The network trained on ws's grid applied to n x n problems of the
    same statistics (held-out seeds >= 2 10^6): its 16 retained modes are
    then 8 times longer waves at 4096^2 than at 512^2, and its output is
    rescaled by the ratio of rms heights (at a fixed rms slope the heights,
    and so the displacements, grow with the box).  Cold and warm solves to
    `tol`, as in WarmStart.evaluate."""
    import numpy as np
    from gfmd_hip import warmstart as W
    big = W.WarmStart(n, max_batch=1)
    big.net.set_params(ws.net.get_params())
    rms = lambda m, seeds: float(np.mean(
        [np.sqrt(np.mean(W.problem(m, sd)[0] ** 2)) for sd in seeds]))
    ratio = rms(n, range(2000000, 2000004)) / rms(ws.n, range(1, 9))
    big.scale = (ws.scale * ratio).astype(np.float32)
    rows = big.evaluate(list(range(2000000, 2000000 + held)), tol=tol)
    mean = lambda k: float(np.mean([r[k] for r in rows]))
    out = {"grid": [n, n], "held_out": held, "output_rescale": ratio,
           "iters_cold": mean("iters_cold"), "iters_warm": mean("iters_warm"),
           "device_s_cold": mean("device_s_cold"),
           "device_s_warm": mean("device_s_warm"),
           "net_and_start_s": mean("net_and_start_s"),
           "prediction_residual": mean("prediction_residual"),
           "prediction_rel_error": mean("prediction_rel_error"),
           "all_converged": all(r["converged"] for r in rows)}
    print("warm start at %d^2: cold %.1f iterations, warm %.1f, prediction "
          "error %.3f" % (n, out["iters_cold"], out["iters_warm"],
                          out["prediction_rel_error"]),
          file=sys.stderr, flush=True)
    for h in (big.s64, big.s32, big.net):
        h.close()
    return out


def warm_start(n=512, ntrain=4096, epochs=6, held=8, tol=1e-8, big_n=4096,
               big_held=4):
    """Learned warm start: ntrain contact problems (seeds 1..ntrain, H in
    [0.6, 0.9], load log-uniform in [1e-3, 1e-2]) solved by the ensemble
    solver to 1e-6 as training data, the FNO trained on the GPU (BF16 MFMA
    kernels, Adam), then `held` held-out problems (seeds >= 10^6) solved to
    `tol` from the network's prediction and from the uniform start; the same
    network on `big_held` problems of big_n x big_n (warm_start_transfer)."""
    import time
    import numpy as np
    from gfmd_hip import warmstart as W
    t0 = time.perf_counter()
    ws = W.WarmStart(n, max_batch=16)
    a, u, gen_dev = ws.generate(list(range(1, ntrain + 1)), tol=1e-6)
    t_gen = time.perf_counter() - t0
    print("warm start: %d training problems solved" % ntrain,
          file=sys.stderr, flush=True)
    ws.set_scale(u)
    t0 = time.perf_counter()
    losses = ws.train(a, u, epochs, lr=1e-3,
                      log=lambda m: print("warm start: " + m,
                                          file=sys.stderr, flush=True))
    t_train = time.perf_counter() - t0
    del a, u
    rows = ws.evaluate(list(range(1000000, 1000000 + held)), tol=tol)
    mean = lambda k: float(np.mean([r[k] for r in rows]))
    big = warm_start_transfer(ws, big_n, big_held, tol) if big_n else None
    return {"workload": "sc100 height %d, %dx%d, rough surfaces (H 0.6-0.9, "
                        "rms slope 0.1, cutoff 4a), load 1e-3..1e-2; network "
                        "trained on %d GPU solutions, %d epochs; %d held-out "
                        "problems to relative residual %g"
                        % (n, n, n, ntrain, epochs, held, tol),
            "transfer": big,
            "data_solve_device_s": gen_dev, "data_wall_s": t_gen,
            "train_wall_s": t_train, "train_steps": epochs * (ntrain // 16),
            "loss_first_last_epoch": [losses[0], losses[-1]],
            "iters_cold": mean("iters_cold"), "iters_warm": mean("iters_warm"),
            "iters_exact_start": mean("iters_exact_start"),
            "device_s_cold": mean("device_s_cold"),
            "device_s_warm": mean("device_s_warm"),
            "net_and_start_s": mean("net_and_start_s"),
            "prediction_residual": mean("prediction_residual"),
            "prediction_rel_error": mean("prediction_rel_error"),
            "all_converged": all(r["converged"] for r in rows)}


def copy_bandwidth(nbytes=1 << 30, reps=20):
    """Measured device copy bandwidth (read + write bytes / time) of a plain
    torch copy of `nbytes`; also the known-size copy kernel the HBM-counter
    runs are calibrated against."""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        b.copy_(a)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / reps
    return 2 * nbytes / (ms * 1e-3) / 1e9


def cpu_baseline(nx, ny, budget_s=20.0):
    """FP64 oracle on one host core: per-component full complex 2D FFTs,
    full-grid Phi table, mat_mul_vec loop (the reference's CPU algorithm).
    Phi values come from sc100 height 0 (the per-evaluation cost does not
    depend on them); the table build is not timed."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    phi = O.fill_phi(O.KERNEL_SC100, [0], 3, nx, ny, normalize=True)
    u = 0.01 * np.random.default_rng(1).standard_normal((3, nx, ny))
    O.static_force(phi, u)  # warm-up (plans, page faults)
    n = 0
    t0 = time.perf_counter()
    while True:
        O.static_force(phi, u)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 20:
            break
    name, avail = cpu_model()
    return {"value": n / el, "unit": "evals/s", "cores": 1, "kind": "port",
            "sample": "%d FP64 evaluations at %dx%d, ndof 3, one core, "
                      "oracle/gfmd_oracle.c (own mixed-radix Stockham FFT; "
                      "no FFTW or MKL on this image), %.1f s"
                      % (n, nx, ny, el),
            "cpu_model": name, "host_cores_available": avail}


def cpu_model():
    """Host CPU model and the cores this process may use."""
    name = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                name = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return name, avail


def full_table(packed, nx, ny, d):
    """Full-grid (nx*ny, d, d) complex table from the device's packed half
    spectrum (ky <= ny/2; the rest by Phi(-q) = conj Phi(q))."""
    import numpy as np
    nky = ny // 2 + 1
    H = np.zeros((nx, nky, d, d), np.complex128)
    P = packed.transpose(2, 0, 1)
    for i in range(d):
        H[:, :, i, i] = P[:, :, i]
    p = d
    for i in range(d):
        for j in range(i + 1, d):
            H[:, :, i, j] = P[:, :, p] + 1j * P[:, :, p + 1]
            H[:, :, j, i] = np.conj(H[:, :, i, j])
            p += 2
    full = np.zeros((nx, ny, d, d), np.complex128)
    full[:, :nky] = H
    full[:, nky:] = np.conj(H[(-np.arange(nx)) % nx][:, ny - np.arange(nky, ny)])
    return full.reshape(nx * ny, d, d)


def oracle_error(r64, r32, u_host):
    """The CPU oracle (a C restatement of GFMDSolverStatic::post_force) on
    the device's FP64 table and the same u: max relative force error and
    relative energy error of the device's FP64 and FP32 results."""
    import ctypes
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    d, nx, ny = u_host.shape
    phi = full_table(r64["table"], nx, ny, d)
    L = O.lib()
    L.orc_static_force_mt.restype = ctypes.c_double
    L.orc_static_force_mt.argtypes = [ctypes.c_int] * 3 + \
        [ctypes.c_void_p] * 3 + [ctypes.c_int]
    nt = max(1, min(16, os.cpu_count() or 1))
    u = np.ascontiguousarray(u_host)
    f = np.zeros_like(u)
    e = L.orc_static_force_mt(nx, ny, d, phi.ctypes.data, u.ctypes.data,
                              f.ctypes.data, nt)
    out = {"workload": "oracle evaluation (orc_static_force_mt, %d threads) "
                       "on the device's FP64 table, same u" % nt}
    for name, r in (("fp64", r64), ("fp32", r32)):
        out[name] = {"max_rel_force": float(np.abs(r["f"] - f).max() /
                                            np.abs(f).max()),
                     "rel_energy": abs(r["energy"] - e) / abs(e)}
    return out


def cpu_c1_time_to_solution(n=128, tol=1e-10, pressure=1e-2, R=100.0):
    """The CPU configuration: sc100 height n on an n x n grid under a rigid
    hard-wall sphere (radius R) at load `pressure` per site, FP64 FIRE from
    u = 0 to relative residual tol, timed on one host core with the
    oracle's restatement (oracle.fire_relax_hardwall: the reference's FIRE
    order with the C oracle's force evaluation), beside the device's FP64
    FIRE and constrained CG on the same problem."""
    import numpy as np
    import gfmd_hip
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    X, Y = np.meshgrid(np.arange(n) - n / 2 + 0.5, np.arange(n) - n / 2 + 0.5,
                       indexing="ij")
    wall = (X ** 2 + Y ** 2) / (2 * R)
    u0 = np.zeros((3, n, n))
    phi = O.fill_phi(O.KERNEL_SC100, [n], 3, n, n, normalize=True)
    F = O.Fire(dt=0.1, adapt_dt=True, dtmax=0.5)
    t0 = time.perf_counter()
    uo, _, steps, res = O.fire_relax_hardwall(phi, u0, wall, pressure, F,
                                              200000, tol=tol,
                                              check_every=100)
    cpu_s = time.perf_counter() - t0
    s = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height %d" % n)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL, pressure=pressure,
                         wall=wall)
    fire = gfmd_hip.fire_params(dt=0.1, dtmax=0.5, adapt_dt=1)
    s.relax(u0, c, fire, max_steps=20, tol=0.0)  # first call: buffers
    ug, rg = s.relax(u0, c, fire, max_steps=200000, tol=tol, check_every=100)
    s.contact_solve(c, tol=tol)                  # first call: tables, graph
    uc, rc = s.contact_solve(c, tol=tol)
    name, avail = cpu_model()
    return {"workload": "sc100 height %d, %dx%d, rigid sphere R %g (hard "
                        "wall), load %g per site, FP64 FIRE (dt 0.1, dtmax "
                        "0.5, adaptive) from u = 0 to relative residual %g"
                        % (n, n, n, R, pressure, tol),
            "cpu_s": cpu_s, "cpu_steps": steps, "cpu_residual": res,
            "cpu_cores": 1, "cpu_model": name,
            "cpu_kind": "port (oracle/oracle.py fire_relax_hardwall over "
                        "oracle/gfmd_oracle.c, own FFT)",
            "gpu_fp64_s": rg["device_seconds"], "gpu_fp64_steps": rg["steps"],
            "gpu_fp64_residual": rg["residual"] / rg["fext_norm"],
            "gpu_cg_fp64_s": rc["device_seconds"],
            "gpu_cg_fp64_iterations": rc["steps"],
            "max_rel_diff_u_gpu_vs_cpu": float(np.abs(ug - uo).max() /
                                               np.abs(uo).max()),
            "max_rel_diff_u_cg_vs_cpu": float(np.abs(uc - uo).max() /
                                              np.abs(uo).max())}


def cpu_threads_baseline(nx, ny, budget_s=10.0):
    """The same FP64 oracle evaluation with OpenMP threads over rows,
    columns and wavevectors (orc_static_force_mt), on the host cores the
    box gives this process (OMP_NUM_THREADS, at most 16), at nx x ny and,
    for the size dependence, at 128^2 and 1024^2 (1 thread and all)."""
    import ctypes
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    L = O.lib()
    L.orc_static_force_mt.restype = ctypes.c_double
    L.orc_static_force_mt.argtypes = [ctypes.c_int] * 3 + \
        [ctypes.c_void_p] * 3 + [ctypes.c_int]
    nt = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nt = max(1, min(nt, 16, os.cpu_count() or 1))

    def rate(n, threads, budget, cap):
        phi = O.fill_phi(O.KERNEL_SC100, [0], 3, n, n, normalize=True)
        u = 0.01 * np.random.default_rng(1).standard_normal((3, n, n))
        f = np.zeros_like(u)
        args = (n, n, 3, phi.ctypes.data, u.ctypes.data, f.ctypes.data,
                threads)
        L.orc_static_force_mt(*args)  # warm-up
        k, t0 = 0, time.perf_counter()
        while True:
            L.orc_static_force_mt(*args)
            k += 1
            el = time.perf_counter() - t0
            if el > budget or k >= cap:
                return k / el, k, el

    v, k, el = rate(nx, nt, budget_s, 20)
    out = {"value": v, "unit": "evals/s", "cores": nt, "kind": "port",
           "sample": "%d FP64 evaluations at %dx%d, ndof 3, %d OpenMP threads, "
                     "oracle/gfmd_oracle.c orc_static_force_mt, %.1f s"
                     % (k, nx, ny, nt, el)}
    sizes = {}
    for n in (128, 1024):
        r1 = rate(n, 1, 2.0, 20)[0]
        rn = rate(n, nt, 2.0, 20)[0]
        sizes["%dx%d" % (n, n)] = {"evals_per_s_1_core": r1,
                                   "evals_per_s_%d_cores" % nt: rn}
    out["sizes"] = sizes
    return out


def cpu_library_fft_baseline(nx, ny, budget_s=10.0):
    """The same evaluation with a library FFT on the host: scipy.fft's
    pocketfft (no FFTW or MKL on this image) r2c/c2r transforms over the
    host cores the box gives this process (OMP_NUM_THREADS, at most 16),
    the Hermitian 3 x 3 product on the half spectrum in numpy (one
    thread), FP64.  The table's values do not change the cost (a random
    Hermitian table)."""
    import numpy as np
    import scipy.fft as sf
    nt = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nt = max(1, min(nt, 16, os.cpu_count() or 1))
    rng = np.random.default_rng(1)
    nky = ny // 2 + 1
    u = 0.01 * rng.standard_normal((3, nx, ny))
    dia = 1.0 + rng.random((3, nx, nky))
    off = 0.1 * (rng.standard_normal((3, nx, nky)) +
                 1j * rng.standard_normal((3, nx, nky)))

    def evaluate():
        U = sf.rfft2(u, axes=(1, 2), workers=nt)
        F = dia * U
        for k, (i, j) in enumerate(((0, 1), (0, 2), (1, 2))):
            F[i] += off[k] * U[j]
            F[j] += np.conj(off[k]) * U[i]
        w = np.full(nky, 2.0)
        w[0] = 1.0
        if ny % 2 == 0:
            w[-1] = 1.0
        e = 0.5 * float(np.sum(w * np.real(np.conj(U) * F).sum(axis=(0, 1))))
        return sf.irfft2(-F, s=(nx, ny), axes=(1, 2), workers=nt), e

    evaluate()
    k, t0 = 0, time.perf_counter()
    while True:
        evaluate()
        k += 1
        el = time.perf_counter() - t0
        if el > budget_s or k >= 20:
            break
    return {"value": k / el, "unit": "evals/s", "cores": nt, "kind": "port",
            "sample": "%d FP64 evaluations at %dx%d, ndof 3, scipy.fft "
                      "(pocketfft) r2c/c2r with %d workers, numpy Phi.u, "
                      "%.1f s" % (k, nx, ny, nt, el)}


def variants(n=2048, iters=50, height=None):
    """Autotuner landscape: per-sweep mean times (HIP events on the
    solver's stream) of every compiled fast-path variant (row variant x
    column variant) for sc100 at n x n, ndof 3, in FP32 and FP64, u and f
    resident; the variant the autotuner picks is the fastest whole
    evaluation."""
    import torch
    import gfmd_hip
    height = height or n
    out = {"workload": "sc100 height %d, %dx%d periodic, ndof 3, %d "
                       "evaluations per variant" % (height, n, n, iters)}
    for prec, name, dt in ((gfmd_hip.FP32, "fp32", torch.float32),
                           (gfmd_hip.FP64, "fp64", torch.float64)):
        s = gfmd_hip.Solver(n, n, 3, prec)
        s.set_kernel("sc100", "height %d" % height)
        g = torch.Generator(device="cuda").manual_seed(1)
        u = 0.01 * torch.randn((3, n, n), dtype=dt, device="cuda",
                               generator=g)
        f = torch.empty_like(u)
        e = torch.zeros(1, dtype=torch.float64, device="cuda")
        rows = {}
        for rv in range(3):
            for cv in range(4):
                try:
                    s.set_variant(rv, cv)
                except gfmd_hip.GFMDError:
                    continue
                ms = s.time_force_device(u.data_ptr(), f.data_ptr(),
                                         e.data_ptr(), iters)
                rows["%d,%d" % (rv, cv)] = dict(
                    zip(["rows_r2c", "cols", "rows_c2r", "eval"], ms))
                print("variants %s (%d,%d): %s" % (name, rv, cv, ms),
                      file=sys.stderr, flush=True)
        out[name] = rows
        s.close()
    return out


LEGS = {
    "variants": lambda a: variants(a.size),
    "fcc100_two_layer": lambda a: fcc100_two_layer(),
    "fcc111_two_layer": lambda a: fcc111_two_layer(),
    "mixed_radix": lambda a: mixed_radix(),
    "time_to_solution": lambda a: time_to_solution(a.size),
    "flat_punch": lambda a: flat_punch(),
    "wall_map": lambda a: wall_map(),
    "ensemble": lambda a: ensemble(),
    "warm_start": lambda a: warm_start(),
    "aux_paths": lambda a: aux_paths(cpu=not a.no_cpu),
    "grid_sizes": lambda a: grid_sizes(cpu=not a.no_cpu),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--no-solve", action="store_true",
                    help="skip the time-to-residual contact leg")
    ap.add_argument("--only", choices=sorted(LEGS),
                    help="run one solver leg alone and print its JSON "
                         "(for profiling; not the benchmark line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only (the force evaluation of "
              "one surface on one MI355X); --gpus %d is not supported"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.only:
        print(json.dumps(LEGS[args.only](args)), flush=True)
        return
    import gfmd_hip

    import numpy as np
    nx = ny = args.size
    height = nx
    u_host = 0.01 * np.random.default_rng(1).standard_normal((3, nx, ny))
    r64 = gpu_bench(nx, ny, height, gfmd_hip.FP64, args.steps, args.warmup,
                    u_host)
    print("fp64: %.4f ms/eval, sweeps %s" % (r64["ms_per_eval"],
                                             r64["sweep_ms"]),
          file=sys.stderr, flush=True)
    r32 = gpu_bench(nx, ny, height, gfmd_hip.FP32, args.steps, args.warmup,
                    u_host)
    print("fp32: %.4f ms/eval, sweeps %s" % (r32["ms_per_eval"],
                                             r32["sweep_ms"]),
          file=sys.stderr, flush=True)

    names = ["rows_r2c", "cols", "rows_c2r"]

    def eval_gbs(r):
        return sum(r["bytes"].values()) / (r["sweep_ms"][3] * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": 1e3 / r64["ms_per_eval"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_eval"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "sc100 height %d, %dx%d periodic grid, ndof 3, "
                               "one FP64 elastic force+energy evaluation per "
                               "step (u, f resident in HBM; u ~ N(0, 0.01^2) "
                               "seed 1)" % (height, nx, ny),
                   "grid": [nx, ny], "ndof": 3, "kernel": "sc100",
                   "height": height},
        "roofline": roofline(r64, nx, "fp64"),
        "eval_alg_GBs": eval_gbs(r64),
        "eval_alg_frac": eval_gbs(r64) / PEAK_HBM_GBS,
        "sweep_ms": dict(zip(names + ["eval"], r64["sweep_ms"])),
        "fp32": {"value": 1e3 / r32["ms_per_eval"], "unit": "evals/s",
                 "ms_per_step": r32["ms_per_eval"],
                 "note": "FP32 FFT and Phi.u, FP64 energy; secondary: the "
                         "headline is the FP64 path",
                 "sweep_ms": dict(zip(names + ["eval"], r32["sweep_ms"])),
                 "roofline": roofline(r32, nx, "fp32"),
                 "eval_alg_GBs": eval_gbs(r32),
                 "eval_alg_frac": eval_gbs(r32) / PEAK_HBM_GBS,
                 "fp32_vs_fp64_device": {
                     "max_rel_force": float(
                         np.abs(r32["f"] - r64["f"]).max() /
                         np.abs(r64["f"]).max()),
                     "rel_energy": abs(r32["energy"] - r64["energy"]) /
                     abs(r64["energy"])}},
        "phi_setup_s": {"fp32": r32["phi_setup_s"], "fp64": r64["phi_setup_s"]},
        "phi_compression": r64["phi_compression"],
        "phi_entries_per_q": r64["phi_entries"],
        "autotuned_variant": {"fp32": r32["variant"], "fp64": r64["variant"]},
    }
    out["copy_GBs"] = copy_bandwidth()
    if not args.no_solve:
        tts = time_to_solution(nx)
        print("time to 1e-8 (device): FIRE FP64 %.3f s, CG FP64 %.4f s, "
              "CG FP32->FP64 %.4f s" % (tts["fire_fp64"]["device_s"],
                                        tts["cg_fp64"]["device_s"],
                                        tts["cg_fp32_fp64"]["device_s"]),
              file=sys.stderr)
        out["time_to_1e-8"] = tts
        gs = grid_sizes(cpu=not args.no_cpu)
        print("other sizes, FP64 evals/s: %s" % ", ".join(
            "%s %.0f" % (k, v["fp64"]["evals_per_s"]) for k, v in gs.items()
            if k != "cpu"), file=sys.stderr, flush=True)
        out["grid_sizes"] = gs
        c2 = fcc100_two_layer()
        print("fcc100 two layers 1024^2: FP32 %.0f evals/s, FP64 %.0f evals/s, "
              "FP32 max rel force error %.2g" % (c2["fp32"]["evals_per_s"],
                                                 c2["fp64"]["evals_per_s"],
                                                 c2["fp32_vs_fp64"]["max_rel_force"]),
              file=sys.stderr, flush=True)
        out["fcc100_two_layer"] = c2
        c12 = fcc111_two_layer()
        print("fcc111 two layers 1024^2 (ndof 12): FP32 %.0f evals/s, FP64 "
              "%.0f evals/s" % (c12["fp32"]["evals_per_s"],
                                c12["fp64"]["evals_per_s"]),
              file=sys.stderr, flush=True)
        out["fcc111_two_layer"] = c12
        out["mixed_radix"] = mixed_radix()
        fp = flat_punch()
        print("flat punch 4096^2: FIRE FP64 %.3f s, CG FP64 %.4f s, "
              "CG FP32->FP64 %.4f s" % (fp["fire_fp64"]["device_s"],
                                        fp["cg_fp64"]["device_s"],
                                        fp["cg_fp32_fp64"]["device_s"]),
              file=sys.stderr, flush=True)
        out["flat_punch"] = fp
        wm = wall_map()
        out["wall_map"] = wm
        ens = ensemble()
        print("ensemble of %d: batch %.3f s, one after another %.3f s"
              % (ens["members"], ens["batch_device_s"],
                 ens["sequential_device_s"]), file=sys.stderr)
        out["ensemble"] = ens
        wsr = warm_start()
        print("warm start %dx%d: iterations cold %.1f, network start %.1f, "
              "exact start %.1f" % (512, 512, wsr["iters_cold"],
                                    wsr["iters_warm"],
                                    wsr["iters_exact_start"]),
              file=sys.stderr)
        out["warm_start"] = wsr
        ax = aux_paths(nx, cpu=not args.no_cpu)
        print("dynamics %d^2: FP64 %.3f ms/step, FP32 %.3f ms/step; "
              "preconditioner FP64 %.3f ms; glue pre/post %.3f / %.3f ms "
              "(grid order)" % (nx, ax["dynamics"]["fp64"]["ms_per_step"],
                                ax["dynamics"]["fp32"]["ms_per_step"],
                                ax["prec_gradient"]["fp64"]["ms_per_call"],
                                ax["fix_glue"]["grid"]["pre_force_ms"],
                                ax["fix_glue"]["grid"]["post_force_ms"]),
              file=sys.stderr, flush=True)
        out["aux_paths"] = ax
    if not args.no_cpu:
        cb = cpu_baseline(nx, ny)
        out["cpu_baseline"] = cb
        out["cpu_baseline_threads"] = cpu_threads_baseline(nx, ny)
        out["cpu_baseline_library_fft"] = cpu_library_fft_baseline(nx, ny)
        print("cpu library FFT: %.3f evals/s on %d threads"
              % (out["cpu_baseline_library_fft"]["value"],
                 out["cpu_baseline_library_fft"]["cores"]), file=sys.stderr,
              flush=True)
        print("cpu: %.3f evals/s on 1 core, %.3f on %d threads"
              % (cb["value"], out["cpu_baseline_threads"]["value"],
                 out["cpu_baseline_threads"]["cores"]), file=sys.stderr,
              flush=True)
        # the oracle's evaluation on the device's FP64 table and the same u:
        # error of both device paths at the headline size
        err = oracle_error(r64, r32, u_host)
        out["vs_oracle"] = err
        out["fp32"]["vs_oracle"] = err["fp32"]
        print("vs oracle at %d^2: FP64 force %.2e energy %.2e; FP32 force "
              "%.2e energy %.2e" % (nx, err["fp64"]["max_rel_force"],
                                   err["fp64"]["rel_energy"],
                                   err["fp32"]["max_rel_force"],
                                   err["fp32"]["rel_energy"]),
              file=sys.stderr, flush=True)
        c1 = cpu_c1_time_to_solution()
        out["cpu_baseline_c1"] = c1
        print("C1 128^2 FIRE to 1e-10: CPU %.2f s (%d steps), GPU FP64 "
              "%.4f s" % (c1["cpu_s"], c1["cpu_steps"], c1["gpu_fp64_s"]),
              file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
