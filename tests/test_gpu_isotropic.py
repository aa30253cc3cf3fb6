"""Isotropic continuum kernel on the device against the CPU FP64 oracle
(itself pinned to the reference's IsotropicStiffnessKernel by golden
vectors and the Boussinesq/Cerruti closed forms).

Tolerances: Phi table and FP64 forces 1e-12 relative (max-norm), energy
1e-12; FP32 forces 2e-6, energy 1e-6.
"""
import numpy as np
import pytest

import gfmd_hip
import oracle as O

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(23)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def packed(full, nx, ny, d):
    """Device layout of (Phi(q) + conj Phi(q'))/2, q' the grid partner."""
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


@pytest.mark.parametrize("n,args,L", [(32, "0.3 1.0 1.0", 32.0),
                                      (31, "0.5 2.0 0.7", 31.0),
                                      (16, "0.25 0.8 3.0 xprd 40 yprd 40",
                                       40.0)])
def test_isotropic_table_matches_oracle(n, args, L):
    nu, myu, gamma = (float(x) for x in args.split()[:3])
    prm = O.isotropic_params(nu, myu, gamma, L)
    s = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    s.set_kernel("isotropic", args)
    ref = O.fill_phi(O.KERNEL_ISOTROPIC, prm, 3, n, n, normalize=True)
    assert rel(s.dump_stiffness(), packed(ref, n, n, 3)) < 1e-12


@pytest.mark.parametrize("n,prec", [(64, gfmd_hip.FP64), (256, gfmd_hip.FP64),
                                    (33, gfmd_hip.FP64), (128, gfmd_hip.FP32)])
def test_isotropic_forces_match_oracle(n, prec):
    prm = O.isotropic_params(0.3, 1.0, 1.0, float(n))
    s = gfmd_hip.Solver(n, n, 3, prec)
    s.set_kernel("isotropic", "0.3 1.0 1.0")
    phi = O.fill_phi(O.KERNEL_ISOTROPIC, prm, 3, n, n, normalize=True)
    u = 0.01 * RNG.standard_normal((3, n, n))
    f_ref, e_ref, _ = O.static_force(phi, u)
    f, e = s.post_force(u)
    tf, te = (1e-12, 1e-12) if prec == gfmd_hip.FP64 else (2e-6, 1e-6)
    assert rel(f, f_ref) < tf
    assert abs(e - e_ref) < te * abs(e_ref)


def test_isotropic_argument_errors():
    s = gfmd_hip.Solver(32, 16, 3, gfmd_hip.FP64)
    for args, code in [("0.3 1.0", gfmd_hip.ERR_ARG),
                       ("0.3 x 1.0", gfmd_hip.ERR_ARG),
                       ("0.3 1.0 1.0", gfmd_hip.ERR_ARG),  # 32 x 16 domain
                       ("0.3 1.0 1.0 xprd 0", gfmd_hip.ERR_ARG),
                       ("0.3 1.0 1.0 height 2", gfmd_hip.ERR_ARG)]:
        with pytest.raises(gfmd_hip.GFMDError) as ei:
            s.set_kernel("isotropic", args)
        assert ei.value.code == code, args
    s.set_kernel("isotropic", "0.3 1.0 1.0 xprd 32 yprd 32")
