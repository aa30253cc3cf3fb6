"""The mixed-radix FFT path (grids that are not powers of two in [64, 4096],
lines of at most 4096 points) against the CPU oracle, through the C ABI.

The reference's FFT3d takes any grid (src/solvers/gfmd_solver_fft.cpp:72-80)
and its tests run 64x37 and 63x63 grids.  Cases: every radix class of the
line plans (4, 2, 3, 5, 7, a lone prime: 37, 2039), both LDS sizes (lines
up to 1024 points and up to 4096), odd row counts (the last row of the
two-row packing alone), a side above 4096 (direct DFTs), and a square
grid of a size the benchmark times (1000^2).

Tolerances: FP64 forces and energy within 1e-12 relative (max-norm), FP32
forces within 2e-6 and energy within 1e-6 relative; a line of prime length
p is one radix-p stage whose outputs are sums of p terms in sequence (an
FP32 transform of 2039 points alone errs by ~1.2e-6 of its largest
output), so there, and on the direct-DFT path, the FP32 force bound is
1e-5.
"""
import numpy as np
import pytest

import gfmd_hip
import oracle as O

pytestmark = pytest.mark.gpu

CASES = [(63, 63, 3), (64, 37, 3), (96, 100, 3), (210, 11, 3), (37, 64, 12),
         (1000, 48, 3), (3000, 6, 3), (3, 2039, 3), (49, 125, 6),
         (5000, 4, 3), (1000, 1000, 3)]


def largest_prime(n):
    p, f = 1, 2
    while n > 1:
        while n % f == 0:
            p, n = f, n // f
        f += 1
    return p


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def table(nx, ny, d):
    """sc100 tables of different heights on the diagonal blocks, the z
    components of block 0 and block b > 0 coupled by a real, even term
    (Hermitian per q, real in real space)."""
    heights = (8, 3, 5, 2)
    phi = np.zeros((nx * ny, d, d), np.complex128)
    for b in range(d // 3):
        phi[:, 3 * b:3 * b + 3, 3 * b:3 * b + 3] = O.fill_phi(
            O.KERNEL_SC100, [heights[b]], 3, nx, ny, normalize=True)
    for b in range(1, d // 3):
        c = 0.1 * phi[:, 2, 2].real / b
        phi[:, 2, 3 * b + 2] = c
        phi[:, 3 * b + 2, 2] = c
    return phi


@pytest.mark.parametrize("nx,ny,d", CASES)
def test_mixed_radix_matches_oracle(nx, ny, d):
    phi = table(nx, ny, d)
    u = 0.01 * np.random.default_rng(nx * 7919 + ny).standard_normal(
        (d, nx, ny))
    f_ref, e_ref, u0_ref = O.static_force(phi, u)
    path = 2 if max(nx, ny) <= 4096 else 0
    long_sums = path == 0 or max(largest_prime(nx), largest_prime(ny)) > 64
    f32 = 1e-5 if long_sums else 2e-6
    for prec, tol_f, tol_e in ((gfmd_hip.FP64, 1e-12, 1e-12),
                               (gfmd_hip.FP32, f32, 1e-6)):
        s = gfmd_hip.Solver(nx, ny, d, prec)
        assert s.info()["fast_path"] == path
        s.set_phi(phi, normalized=True)
        f, e = s.post_force(u)
        ef, ee = rel(f, f_ref), abs(e - e_ref) / abs(e_ref)
        print("%dx%d ndof %d %s: force %.2e energy %.2e"
              % (nx, ny, d, "fp64" if prec == gfmd_hip.FP64 else "fp32",
                 ef, ee), flush=True)
        assert ef < tol_f
        assert ee < tol_e
        if prec == gfmd_hip.FP64:
            assert np.allclose(s.get_u0(), u0_ref, rtol=1e-12, atol=1e-14)
        s.close()
