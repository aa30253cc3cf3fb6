"""fcc100 stiffness kernel on the device: tabulation from the reference's
argument syntax, forces through both the fast (ndof 3, 6) and the generic
(ndof 9, 12) paths, against the CPU FP64 oracle (itself pinned to the
reference's FCC100StiffnessKernel by golden vectors).

Tolerances:
  Phi table, FP64 forces : 1e-11 relative (max-norm), energy 1e-11.
  FP32 forces (two-layer nearest-neighbour kernel, k = 57.146, 1024^2):
                           2e-5 relative to the FP64 path, energy 1e-5.
"""
import numpy as np
import pytest

import gfmd_hip
import oracle as O

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(11)

# (argument string, oracle parameters without height, height)
CASES = [
    ("1 1.0 height 16", O.fcc100_params(16, 1, [1.0])),
    ("2 57.146 0 0 height 32", O.fcc100_params(32, 2, [57.146, 0, 0])),
    ("x 2 1.0 0.1 0.3 0.05 0.2 0.02 height 8",
     O.fcc100_params(8, 2, [1.0, 0.3, 0.2], [0.1, 0.05, 0.02])),
    ("2 1.0 0.5 0.25 height 12 normal -1",
     O.fcc100_params(12, 2, [1.0, 0.5, 0.25], normal=-1)),
    ("x 1 1.0 0.2 height 4", O.fcc100_params(4, 1, [1.0], [0.2])),
    ("2 1.0 0.5 0.25", O.fcc100_params(0, 2, [1.0, 0.5, 0.25])),
    ("x 3 1.0 0.1 0.4 0.05 0.2 0.03 height 6",
     O.fcc100_params(6, 3, [1.0, 0.4, 0.2], [0.1, 0.05, 0.03])),
    ("4 1.0 0.3 0.1 height 5 normal -1",
     O.fcc100_params(5, 4, [1.0, 0.3, 0.1], normal=-1)),
]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def packed(full, nx, ny, d):
    """Device layout of Phi_s(q) = (Phi(q) + conj Phi(q'))/2 (q' the grid
    partner of q), the operator a real-part inverse transform applies."""
    nky = ny // 2 + 1
    out = np.zeros((nky, d * d, nx))
    F = full.reshape(nx, ny, d, d)
    P = F[(-np.arange(nx)) % nx][:, (-np.arange(ny)) % ny]
    F = 0.5 * (F + P.conj())
    for ky in range(nky):
        M = F[:, ky]
        for i in range(d):
            out[ky, i] = M[:, i, i].real
        p = d
        for i in range(d):
            for j in range(i + 1, d):
                h = 0.5 * (M[:, i, j] + M[:, j, i].conj())
                out[ky, p], out[ky, p + 1] = h.real, h.imag
                p += 2
    return out


def test_fcc100_three_layers_not_2pi_periodic():
    """The reference's nu = 3 cell carries no layer shifts
    (fcc100_stiffness.cpp:476-486), so Phi has period 4 pi in q and the
    Nyquist lines hold Phi(q') != conj Phi(q).  Forces must still equal the
    reference's complex-transform result (the oracle's static_force)."""
    prm = CASES[6][1]
    a = O.kernel_phi(O.KERNEL_FCC100, prm, np.pi, 0.7)
    b = O.kernel_phi(O.KERNEL_FCC100, prm, np.pi, -0.7)
    assert np.abs(a - b.conj()).max() > 1e-3 * np.abs(a).max()
    nx, ny = 8, 10
    s = gfmd_hip.Solver(nx, ny, 9, gfmd_hip.FP64)
    s.set_kernel("fcc100", CASES[6][0])
    phi = O.fill_phi(O.KERNEL_FCC100, prm, 9, nx, ny, normalize=True)
    u = 0.01 * RNG.standard_normal((9, nx, ny))
    f_ref, e_ref, _ = O.static_force(phi, u)
    f, e = s.post_force(u)
    assert rel(f, f_ref) < 1e-11
    assert abs(e - e_ref) < 1e-11 * abs(e_ref)
    # the same table through the host route
    t = gfmd_hip.Solver(nx, ny, 9, gfmd_hip.FP64)
    t.set_phi(phi, normalized=True)
    f2, e2 = t.post_force(u)
    assert rel(f2, f_ref) < 1e-11
    assert abs(e2 - e_ref) < 1e-11 * abs(e_ref)


@pytest.mark.parametrize("args,prm", CASES)
def test_fcc100_table_matches_oracle(args, prm):
    nx, ny = 32, 24
    d = O.kernel_dim(O.KERNEL_FCC100, prm)
    print("fcc100", args, "d", d, flush=True)
    s = gfmd_hip.Solver(nx, ny, d, gfmd_hip.FP64)
    s.set_kernel("fcc100", args)
    ref = O.fill_phi(O.KERNEL_FCC100, prm, d, nx, ny, normalize=True)
    assert rel(s.dump_stiffness(), packed(ref, nx, ny, d)) < 1e-11


@pytest.mark.parametrize("args,prm,nx,ny", [
    (CASES[1][0], CASES[1][1], 128, 64),      # fast path, ndof 6
    (CASES[2][0], CASES[2][1], 64, 128),      # fast path, ndof 6, kappa
    (CASES[0][0], CASES[0][1], 64, 64),       # fast path, ndof 3
    (CASES[6][0], CASES[6][1], 24, 20),       # generic path, ndof 9
    (CASES[6][0], CASES[6][1], 15, 13),       # generic path, odd sizes
    (CASES[7][0], CASES[7][1], 16, 18),       # generic path, ndof 12
])
def test_fcc100_forces_fp64_match_oracle(args, prm, nx, ny):
    d = O.kernel_dim(O.KERNEL_FCC100, prm)
    print("fcc100 forces", args, nx, ny, flush=True)
    s = gfmd_hip.Solver(nx, ny, d, gfmd_hip.FP64)
    s.set_kernel("fcc100", args)
    phi = O.fill_phi(O.KERNEL_FCC100, prm, d, nx, ny, normalize=True)
    u = 0.01 * RNG.standard_normal((d, nx, ny))
    f_ref, e_ref, u0_ref = O.static_force(phi, u)
    f, e = s.post_force(u)
    assert rel(f, f_ref) < 1e-11
    assert abs(e - e_ref) < 1e-11 * abs(e_ref)
    assert np.allclose(s.get_u0(), u0_ref, rtol=1e-11, atol=1e-14)


def test_fcc100_linf_keyword():
    """'linf' values become the Gamma-point forces of each layer."""
    nx, ny = 64, 64
    prm = O.fcc100_params(10, 2, [1.0, 0.2, 0.1])
    s = gfmd_hip.Solver(nx, ny, 6, gfmd_hip.FP64)
    s.set_kernel("fcc100", "2 1.0 0.2 0.1 height 10 linf 0.25 -0.25")
    phi = O.fill_phi(O.KERNEL_FCC100, prm, 6, nx, ny, normalize=True)
    u = 0.01 * RNG.standard_normal((6, nx, ny))
    f_ref, e_ref, _ = O.static_force(phi, u, np.array([0.25, -0.25]))
    f, e = s.post_force(u)
    assert rel(f, f_ref) < 1e-11
    assert abs(e - e_ref) < 1e-11 * abs(e_ref)


def test_fcc100_two_layer_1024_fp32_vs_fp64():
    """Two-layer nearest-neighbour fcc100 (k = V''(2^(1/6)) = 57.146 of a
    unit LJ pair) at 1024^2: FP32 path within the stated bound of FP64, and
    the FP64 path against the plane-wave closed form through the oracle's
    Phi(q)."""
    nx = ny = 1024
    args = "2 57.146 0 0 height 64"
    prm = O.fcc100_params(64, 2, [57.146, 0, 0])
    s64 = gfmd_hip.Solver(nx, ny, 6, gfmd_hip.FP64)
    s64.set_kernel("fcc100", args)
    s32 = gfmd_hip.Solver(nx, ny, 6, gfmd_hip.FP32)
    s32.set_kernel("fcc100", args)
    assert s64.info()["fast_path"] == 1 and s32.info()["fast_path"] == 1
    u = 0.01 * RNG.standard_normal((6, nx, ny))
    f64, e64 = s64.post_force(u)
    f32, e32 = s32.post_force(u)
    err = rel(f32, f64)
    print("fcc100 1024^2 fp32 vs fp64: force", err, "energy",
          abs(e32 - e64) / e64, flush=True)
    assert err < 2e-5
    assert abs(e32 - e64) < 1e-5 * abs(e64)
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    i0, j0 = 5, 301
    qx, qy = 2 * np.pi * i0 / nx, 2 * np.pi * j0 / ny
    A = 0.01 * (RNG.standard_normal(6) + 1j * RNG.standard_normal(6))
    ph = np.exp(1j * (qx * X + qy * Y))
    w = np.real(A[:, None, None] * ph[None])
    f, e = s64.post_force(w)
    P = O.kernel_phi(O.KERNEL_FCC100, prm, qx, qy)
    fref = -np.real((P @ A)[:, None, None] * ph[None])
    assert rel(f, fref) < 1e-11
    assert abs(e - nx * ny / 4 * np.real(A.conj() @ P @ A)) < 1e-11 * abs(e)


@pytest.mark.parametrize("args", [
    "",                       # number of layers missing
    "two 1.0",                # layers not a number
    "0 1.0",                  # fewer than one layer
    "2 1.0 0.5",              # third spring constant missing
    "x 1 1.0",                # kappa missing
    "2 1.0 0 0 height h",     # height not a number
    "2 1.0 0 0 normal 2",     # normal not +-1
    "2 1.0 0 0 linf 0.1 0.2",  # linf does not sum to zero
    "2 1.0 0 0 bogus 1",      # unknown keyword left over
])
def test_fcc100_argument_errors(args):
    s = gfmd_hip.Solver(16, 16, 6, gfmd_hip.FP64)
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        s.set_kernel("fcc100", args)
    assert ei.value.code == gfmd_hip.ERR_ARG, str(ei.value)


def test_fcc100_layer_count_must_match_ndof():
    s = gfmd_hip.Solver(16, 16, 3, gfmd_hip.FP64)
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        s.set_kernel("fcc100", "2 1.0 0 0")
    assert ei.value.code == gfmd_hip.ERR_ARG
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        gfmd_hip.Solver(16, 16, 15, gfmd_hip.FP64).set_kernel(
            "fcc100", "5 1.0 0 0")
    assert ei.value.code == gfmd_hip.ERR_UNSUPPORTED
