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
