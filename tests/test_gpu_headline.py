"""Parity at the headline sizes, through the C ABI.

* Random-field FP64 forces and energy at 1024^2 and 2048^2 (sc100, height =
  grid size) against the CPU oracle, for every compiled fast-path variant,
  within 1e-12 relative.  The oracle's full-grid table cannot be tabulated
  on the CPU in seconds at these sizes (a transfer-matrix recursion of
  2048 steps per wavevector), so the device table is pinned first: 64
  sampled wavevectors (Gamma, zone edges and corners, seeded random points)
  against oracle.kernel_phi within 1e-12; the oracle's FFT + Phi.u
  evaluation then runs on that table.
* FP32 forces and energy at 2048^2 against the same oracle evaluation:
  forces within 3e-6 relative (max-norm), energy within 1e-6.
* The two-layer fcc(100) configuration (nearest-neighbour LJ springs,
  k = 57.146, 1024^2, ndof 6: a 6 x 6 Hermitian Phi per wavevector) the
  same way: the device table at 64 wavevectors against oracle.kernel_phi,
  then FP64 forces and energy within 1e-12 and FP32 within 2e-6 / 1e-6 of
  the oracle's evaluation on that table.
* The two-layer fcc(111) `ft` cell (four atoms, ndof 12, 1024^2) the same
  way: its table at 64 wavevectors against the oracle's ft restatement
  within 1e-12, FP64 forces and energy within 1e-11 and FP32 within 5e-6 /
  1e-6 of the oracle's evaluation.
* The open-boundary flat punch at full size (2048^2 patch in a 4096^2 grid,
  radius 512, load 1e-3 per patch site, FP32 stage handed to FP64 to
  relative residual 1e-8): indentation within 0.1 % of Sneddon's, pressure
  profile within 0.5 % rms of Sneddon's for r < 0.9 a, displacement within
  1e-8 of the FP64 solve.
* The 64-member ensemble of rough contacts (16 each of 256^2, 512^2,
  1024^2, 2048^2): every member converges to 1e-8, one member per size
  equals its single solve to 1e-12.
"""
import numpy as np
import pytest

import gfmd_hip
import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def sample_wavevectors(nx, ny, count, seed):
    kx = [0, nx // 2, 0, nx // 2, 1, nx - 1]
    ky = [0, 0, ny // 2, ny // 2, 1, ny // 2]
    rng = np.random.default_rng(seed)
    kx += list(rng.integers(0, nx, count - len(kx)))
    ky += list(rng.integers(0, ny // 2 + 1, count - len(ky)))
    return np.array(kx), np.array(ky)


def grid_q(i, n):
    return 2 * np.pi * (i if i <= n // 2 else i - n) / n


def full_table(packed, nx, ny, d):
    """Full-grid (nx*ny, d, d) complex table from the device's packed half
    spectrum (ky <= ny/2; the rest by Phi(-q) = conj Phi(q))."""
    nky = ny // 2 + 1
    H = np.zeros((nx, nky, d, d), np.complex128)
    P = packed.transpose(2, 0, 1)  # (nx, nky, d*d)
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
    kxm = (-np.arange(nx)) % nx
    kym = ny - np.arange(nky, ny)
    full[:, nky:] = np.conj(H[kxm][:, kym])
    return full.reshape(nx * ny, d, d)


@pytest.mark.parametrize("n", [1024, 2048])
def test_random_field_fp64_every_variant_matches_oracle(n):
    s = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height %d" % n)
    assert s.info()["fast_path"] == 1
    # the device table at 64 wavevectors against the oracle's kernel
    kx, ky = sample_wavevectors(n, n, 64, seed=n)
    got = s.stiffness_at(kx, ky)
    ref = np.array([O.kernel_phi(O.KERNEL_SC100, [n], grid_q(a, n),
                                 grid_q(b, n)) for a, b in zip(kx, ky)])
    ref /= n * n
    assert rel(got, ref) < 1e-12
    # the oracle's evaluation on that table
    phi = full_table(s.dump_stiffness(), n, n, 3)
    u = 0.01 * np.random.default_rng(n + 1).standard_normal((3, n, n))
    linf = np.array([0.003])
    s.set_linf(linf)
    f_ref, e_ref, u0_ref = O.static_force(phi, u, linf)
    del phi
    seen = 0
    for rv in range(3):
        for cv in range(4):
            s.set_variant(rv, cv)
            f, e = s.post_force(u)
            print("fp64 %d^2 variant (%d,%d): force %.2e energy %.2e"
                  % (n, rv, cv, rel(f, f_ref), abs(e - e_ref) / abs(e_ref)),
                  flush=True)
            assert rel(f, f_ref) < 1e-12, (rv, cv)
            assert abs(e - e_ref) < 1e-12 * abs(e_ref), (rv, cv)
            seen += 1
    assert seen == 12
    assert np.allclose(s.get_u0(), u0_ref, rtol=1e-12, atol=1e-14)


def test_fcc100_two_layer_1024_against_oracle():
    n, args = 1024, "2 57.146 0 0 height 64"
    prm = O.fcc100_params(64, 2, [57.146, 0, 0])
    s64 = gfmd_hip.Solver(n, n, 6, gfmd_hip.FP64)
    s64.set_kernel("fcc100", args)
    assert s64.info()["fast_path"] == 1
    kx, ky = sample_wavevectors(n, n, 64, seed=7)
    got = s64.stiffness_at(kx, ky)
    ref = np.array([O.kernel_phi(O.KERNEL_FCC100, prm, grid_q(a, n),
                                 grid_q(b, n)) for a, b in zip(kx, ky)])
    ref /= n * n
    assert rel(got, ref) < 1e-12
    phi = full_table(s64.dump_stiffness(), n, n, 6)
    u = 0.02 * np.random.default_rng(2).standard_normal((6, n, n))
    f_ref, e_ref, _ = O.static_force(phi, u)
    del phi
    f64, e64 = s64.post_force(u)
    s32 = gfmd_hip.Solver(n, n, 6, gfmd_hip.FP32)
    s32.set_kernel("fcc100", args)
    f32, e32 = s32.post_force(u)
    print("fcc100 two layers 1024^2 vs oracle: FP64 force %.2e energy %.2e, "
          "FP32 force %.2e energy %.2e"
          % (rel(f64, f_ref), abs(e64 - e_ref) / abs(e_ref), rel(f32, f_ref),
             abs(e32 - e_ref) / abs(e_ref)), flush=True)
    assert rel(f64, f_ref) < 1e-12
    assert abs(e64 - e_ref) < 1e-12 * abs(e_ref)
    assert rel(f32, f_ref) < 2e-6
    assert abs(e32 - e_ref) < 1e-6 * abs(e_ref)


def test_fcc111_two_layer_1024_against_oracle():
    import ft
    n, args = 1024, "fcc111 1.0 2 pair-potential 1 1.0 height 1024"
    k = ft.FTKernel(args)
    assert k.dim == 12
    s64 = gfmd_hip.Solver(n, n, 12, gfmd_hip.FP64)
    s64.set_kernel("ft", args)
    assert s64.info()["fast_path"] == 1
    kx, ky = sample_wavevectors(n, n, 64, seed=9)
    got = s64.stiffness_at(kx, ky)
    ref = np.array([k.phi(grid_q(a, n), grid_q(b, n))
                    for a, b in zip(kx, ky)]) / (n * n)
    assert rel(got, ref) < 1e-12
    phi = full_table(s64.dump_stiffness(), n, n, 12)
    u = 0.02 * np.random.default_rng(4).standard_normal((12, n, n))
    f_ref, e_ref, _ = O.static_force(phi, u)
    del phi
    f64, e64 = s64.post_force(u)
    s64.close()
    s32 = gfmd_hip.Solver(n, n, 12, gfmd_hip.FP32)
    s32.set_kernel("ft", args)
    f32, e32 = s32.post_force(u)
    print("fcc111 two layers 1024^2 vs oracle: FP64 force %.2e energy %.2e, "
          "FP32 force %.2e energy %.2e"
          % (rel(f64, f_ref), abs(e64 - e_ref) / abs(e_ref), rel(f32, f_ref),
             abs(e32 - e_ref) / abs(e_ref)), flush=True)
    assert rel(f64, f_ref) < 1e-11
    assert abs(e64 - e_ref) < 1e-11 * abs(e_ref)
    assert rel(f32, f_ref) < 5e-6
    assert abs(e32 - e_ref) < 1e-6 * abs(e_ref)


def test_random_field_fp32_2048_bound_against_oracle():
    n = 2048
    s64 = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    s64.set_kernel("sc100", "height %d" % n)
    phi = full_table(s64.dump_stiffness(), n, n, 3)
    s64.close()
    u = 0.01 * np.random.default_rng(11).standard_normal((3, n, n))
    f_ref, e_ref, _ = O.static_force(phi, u)
    del phi
    s = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP32)
    s.set_kernel("sc100", "height %d" % n)
    worst = 0.0
    for rv in range(3):
        for cv in range(4):
            s.set_variant(rv, cv)
            f, e = s.post_force(u)
            ef, ee = rel(f, f_ref), abs(e - e_ref) / abs(e_ref)
            print("fp32 2048^2 variant (%d,%d): force %.2e energy %.2e"
                  % (rv, cv, ef, ee), flush=True)
            worst = max(worst, ef)
            assert ef < 3e-6, (rv, cv)
            assert ee < 1e-6, (rv, cv)
    print("fp32 2048^2 worst force error %.2e" % worst)


def test_flat_punch_full_size_mixed_precision():
    n, patch, a = 4096, 2048, 512.0
    load = 1e-3 * patch * patch
    nu, mu = 0.3, 1.0
    e_star = 2 * mu * (1 + nu) / (1 - nu * nu)
    delta0 = load / (2 * a * e_star)
    i = np.arange(n) - n // 2
    r = np.sqrt(i[:, None] ** 2 + i[None, :] ** 2)
    punch = r <= a
    wall = np.where(punch, -delta0, 1e30)
    c = gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL, pressure=0.0, wall=wall)
    s64 = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    s64.set_kernel("nonperiodic", "%g %g" % (nu, mu))
    s32 = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP32)
    s32.set_kernel("nonperiodic", "%g %g" % (nu, mu))
    tol = 1e-8 * load / np.sqrt(punch.sum())
    u64, r64 = s64.contact_solve(c, tol=tol)
    um, rm = s64.contact_solve(c, s32=s32, tol=tol)
    print("punch 4096^2: FP64 %d iterations; FP32 -> FP64 %d + %d"
          % (r64["steps"], rm["iters32"], rm["iters64"]), flush=True)
    assert r64["converged"] and rm["converged"]
    assert rm["iters32"] > 0
    assert np.abs(um - u64).max() < 1e-8 * np.abs(u64).max()
    f, _ = s64.post_force(u64)
    p = np.where(punch, f[2], 0.0)
    pnorm = np.sqrt(np.sum(p * p))
    assert r64["residual"] <= 1e-8 * pnorm * 1.0001
    scale = load / p.sum()
    # Sneddon: delta = F / (2 a E*); the solve ran at delta0 for F = load
    assert abs(scale - 1) < 1e-3, scale
    inner = r < 0.9 * a
    ps = load / (2 * np.pi * a * np.sqrt(a * a - r[inner] ** 2))
    prof = np.sqrt(np.mean((p[inner] * scale - ps) ** 2) / np.mean(ps ** 2))
    print("punch 4096^2: delta/Sneddon %.5f, profile rms %.4f" % (scale, prof))
    assert prof < 5e-3, prof
    # the padding and the free surface carry no force (to the residual)
    assert np.abs(f[2][r > a + 1]).max() < 1e-8 * pnorm


def test_ensemble_full_mix_64_members():
    sizes, per, pressure, tol = (256, 512, 1024, 2048), 16, 5e-3, 1e-8
    ens, cs, protos = [], [], []
    k = 0
    for n in sizes:
        a = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
        a.set_kernel("sc100", "height %d" % n)
        b = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP32)
        b.set_kernel("sc100", "height %d" % n)
        protos.append((a, b))
        ens.append(gfmd_hip.Ensemble(a, per, b))
        for _ in range(per):
            h = gfmd_hip.rough_surface(n, n, hurst=0.8, rms_slope=0.1,
                                       lambda_short=4.0, seed=1000 + k)
            cs.append(gfmd_hip.contact(gfmd_hip.CONTACT_HARDWALL,
                                       pressure=pressure, wall=h.max() - h))
            k += 1
    out = gfmd_hip.ensemble_contact_solve(ens, cs, tol=tol)
    assert len(out) == 64
    for u, r in out:
        assert r["converged"], r
        assert r["residual"] <= tol * r["fext_norm"] * 1.0001
    for j, n in enumerate(sizes):
        a, b = protos[j]
        m = j * per + 5
        u, r = a.contact_solve(cs[m], s32=b, tol=tol)
        assert r["converged"]
        d = np.abs(u - out[m][0]).max() / np.abs(u).max()
        print("ensemble member %d (%d^2): batch vs single %.2e, %d "
              "iterations" % (m, n, d, r["steps"]), flush=True)
        assert d <= 1e-12
    for e in ens:
        e.close()
