"""Learned warm start for hard-wall contact of rough surfaces.

A Fourier neural operator (gfmd_hip.FNO, BF16 MFMA kernels) is trained on
contact solutions that the GPU solver itself produces, then predicts the
relaxed displacement of a new surface; its prediction gives the starting
pressures of the constrained-CG contact solve, which refines to the same
residual as a cold start and judges the prediction (its relative force
residual).  Problems: sc100 (height = grid size) against a rigid
self-affine rough surface (rms slope 0.1, short-wavelength cutoff 4a,
Hurst exponent and load drawn per seed: H in [0.6, 0.9], load per site
log-uniform in [1e-3, 1e-2]).
"""
import time

import numpy as np

from . import (CONTACT_HARDWALL, FNO, FP32, FP64, Ensemble, Solver, contact,
               ensemble_contact_solve, rough_surface)

C, M, L, CIN, COUT = 32, 16, 4, 2, 3


def draw(seed):
    """Hurst exponent and load of problem `seed`."""
    rng = np.random.default_rng(seed)
    return float(rng.uniform(0.6, 0.9)), float(10 ** rng.uniform(-3.0, -2.0))


def problem(n, seed):
    """(heights h, wall map, load p0) of problem `seed` on an n x n grid."""
    hurst, p0 = draw(seed)
    h = rough_surface(n, n, hurst=hurst, rms_slope=0.1, lambda_short=4.0,
                      seed=seed)
    return h, h.max() - h, p0


def features(h, p0):
    """Network inputs (n, n, 2): height below the highest summit in units
    of the rms height, and log10 of the load (shifted)."""
    rms = float(np.sqrt(np.mean(h * h))) or 1.0
    a = np.empty(h.shape + (2,), dtype=np.float32)
    a[..., 0] = (h - h.max()) / rms
    a[..., 1] = np.log10(p0) + 2.5
    return a


def init_params(seed=0):
    """Initial flat parameters (layout of gfmd_fno_set_params)."""
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal(CIN * C) / np.sqrt(CIN), np.zeros(C)]
    nmode = 2 * M * (M + 1)
    for _ in range(L):
        parts += [rng.standard_normal(C * C) / np.sqrt(C), np.zeros(C),
                  rng.random(nmode * C * C * 2) / (C * C)]
    parts += [rng.standard_normal(C * COUT) / np.sqrt(C), np.zeros(COUT)]
    vec = np.concatenate(parts).astype(np.float32)
    assert vec.size == FNO.param_count()
    return vec


class WarmStart:
    """Solver handles, the network and the output scaling for n x n grids."""

    def __init__(self, n, max_batch=16, seed=0):
        self.n = n
        self.s64 = Solver(n, n, 3, FP64)
        self.s64.set_kernel("sc100", "height %d" % n)
        self.s64.autotune(5)
        self.s32 = Solver(n, n, 3, FP32)
        self.s32.set_kernel("sc100", "height %d" % n)
        self.s32.autotune(5)
        self.net = FNO(n, n, max_batch)
        self.net.set_params(init_params(seed))
        self.scale = np.ones(COUT, dtype=np.float32)
        self.max_batch = max_batch

    # -- data ---------------------------------------------------------------
    def generate(self, seeds, tol=1e-6, group=16):
        """Solve the problems `seeds` with ensembles of `group` members;
        returns (inputs (k, n, n, 2), displacements (k, n, n, 3)) float32
        and the solve time (device, all groups)."""
        n = self.n
        a = np.empty((len(seeds), n, n, CIN), dtype=np.float32)
        u = np.empty((len(seeds), n, n, COUT), dtype=np.float32)
        dev = 0.0
        for g0 in range(0, len(seeds), group):
            ss = seeds[g0:g0 + group]
            ens = Ensemble(self.s64, len(ss), self.s32)
            cs = []
            for i, s in enumerate(ss):
                h, wall, p0 = problem(n, s)
                a[g0 + i] = features(h, p0)
                cs.append(contact(CONTACT_HARDWALL, pressure=p0, wall=wall))
            out = ensemble_contact_solve([ens], cs, tol=tol)
            dev += out[0][1]["device_seconds"]
            for i, (ui, _) in enumerate(out):
                u[g0 + i] = np.moveaxis(ui, 0, -1)
            ens.close()
        return a, u, dev

    def set_scale(self, u):
        self.scale = np.sqrt(np.mean(u.astype(np.float64) ** 2,
                                     axis=(0, 1, 2))).astype(np.float32)
        self.scale[self.scale == 0] = 1.0

    # -- training -----------------------------------------------------------
    def train(self, a, u, epochs, batch=None, lr=1e-3, seed=0, log=None):
        """Adam on (a, u / scale), shuffled per epoch; the data stay on the
        device.  Returns the last loss of every epoch."""
        import torch
        batch = batch or self.max_batch
        rng = np.random.default_rng(seed)
        da = torch.from_numpy(a).cuda()
        dt = torch.from_numpy(u / self.scale).cuda()
        k = a.shape[0] // batch * batch
        losses = []
        for ep in range(epochs):
            perm = torch.from_numpy(rng.permutation(a.shape[0])[:k]).cuda()
            ea = da.index_select(0, perm).contiguous()
            et = dt.index_select(0, perm).contiguous()
            torch.cuda.synchronize()
            sa, st = ea[0].numel() * 4, et[0].numel() * 4
            for i in range(0, k - batch, batch):
                self.net.train_step_device(ea.data_ptr() + i * sa,
                                           et.data_ptr() + i * st, batch, lr)
            i = k - batch
            loss = self.net.train_step(ea[i:i + batch].cpu().numpy(),
                                       et[i:i + batch].cpu().numpy(), lr)
            losses.append(loss)
            if log:
                log("epoch %d loss %.4g" % (ep, loss))
            del ea, et
        return losses

    # -- use ----------------------------------------------------------------
    def predict(self, h, p0):
        """Displacement (3, n, n) float64 predicted for one problem."""
        out = self.net.forward(features(h, p0)[None])[0] * self.scale
        return np.moveaxis(out.astype(np.float64), -1, 0)

    def start_pressure(self, u, wall, p0):
        """Starting pressures from a displacement: the wall pressure that
        balances elastic force and load, p = max(0, f_el,z + p0), on the
        sites the displacement puts at (or past) the wall."""
        f, _ = self.s64.post_force(u)
        p = np.maximum(f[2] + p0, 0.0)
        return np.where(u[2] >= wall - 1e-3 * np.std(u[2]), p, 0.0)

    def solve(self, wall, p0, p_init=None, tol=1e-8):
        c = contact(CONTACT_HARDWALL, pressure=p0, wall=wall)
        return self.s64.contact_solve(c, s32=self.s32, p_init=p_init, tol=tol)

    def evaluate(self, seeds, tol=1e-8):
        """Cold vs warm start on held-out problems.  Per problem: iterations
        and device time of both solves, the network's inference + starting
        pressure time, and the prediction's error judged by GFMD: relative
        force residual (FP64, projected at the wall) and the relative
        distance to the converged displacement."""
        from . import fire_params
        rows = []
        fire = fire_params()
        for s in seeds:
            h, wall, p0 = problem(self.n, s)
            c = contact(CONTACT_HARDWALL, pressure=p0, wall=wall)
            self.solve(wall, p0, tol=tol)  # warm-up of this problem's tables
            u_cold, rc = self.solve(wall, p0, tol=tol)
            t0 = time.perf_counter()
            up = self.predict(h, p0)
            pin = self.start_pressure(up, wall, p0)
            t_net = time.perf_counter() - t0
            u_warm, rw = self.solve(wall, p0, p_init=pin, tol=tol)
            # the bound a perfect network would reach: start from the
            # converged pressures
            f, _ = self.s64.post_force(u_cold)
            pex = np.where(u_cold[2] >= wall - 1e-12,
                           np.maximum(f[2] + p0, 0.0), 0.0)
            _, rx = self.solve(wall, p0, p_init=pex, tol=tol)
            upc = up.copy()
            upc[2] = np.minimum(upc[2], wall)
            _, r0 = self.s64.relax(upc, c, fire, max_steps=0, tol=tol)
            rows.append({
                "seed": s, "load": p0,
                "iters_cold": rc["steps"], "iters_warm": rw["steps"],
                "iters_exact_start": rx["steps"],
                "device_s_cold": rc["device_seconds"],
                "device_s_warm": rw["device_seconds"],
                "net_and_start_s": t_net,
                "converged": bool(rc["converged"] and rw["converged"]),
                "prediction_residual": r0["residual"] / r0["fext_norm"],
                "prediction_rel_error": float(np.linalg.norm(up - u_cold) /
                                              np.linalg.norm(u_cold)),
                "warm_vs_cold_rel_diff": float(np.abs(u_warm - u_cold).max() /
                                               np.abs(u_cold).max())})
        return rows
