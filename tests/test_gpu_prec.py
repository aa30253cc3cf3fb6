"""Preconditioned gradient (GFMDSolverStatic::prec_gradient) on the device
against the oracle, the force round trip, the table cache and errors."""
import numpy as np
import pytest

import gfmd_hip
import oracle as O

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(31)
C3 = np.array([[0.3, 0.05, 0.0], [0.05, 0.2, 0.01], [0.0, 0.01, 0.4]])


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("nx,ny", [(32, 24), (64, 64)])
def test_sc100_matches_oracle(nx, ny):
    """FP64 to 1e-12, FP32 to 1e-5; cavg coupling x and y (the table then
    loses the x -> -x symmetry the force path can exploit)."""
    g = RNG.normal(size=(3, nx, ny))
    phi = O.fill_phi(O.KERNEL_SC100, [16], 3, nx, ny, normalize=True)
    ref = O.prec_gradient(phi, g, C3)
    for prec, tol in ((gfmd_hip.FP64, 1e-12), (gfmd_hip.FP32, 1e-5)):
        s = gfmd_hip.Solver(nx, ny, 3, prec)
        s.set_kernel("sc100", "height 16")
        assert rel(s.prec_gradient(g, C3), ref) < tol
        # cavg = 0 and a diagonal cavg (x -> -x symmetric table)
        assert rel(s.prec_gradient(g), O.prec_gradient(phi, g)) < tol
        D = np.diag([0.1, 0.2, 0.3])
        assert rel(s.prec_gradient(g, D), O.prec_gradient(phi, g, D)) < tol


def test_two_layer_cells():
    """ndof 6 (fcc100, two layers): every component preconditioned;
    components 0-2 are the reference's (its general branch leaves the
    others untransformed)."""
    nx, ny = 32, 32
    p = O.fcc100_params(8, 2, [1.0, 0.3, 0.2], [0.1, 0.05, 0.02])
    phi = O.fill_phi(O.KERNEL_FCC100, p, 6, nx, ny, normalize=True)
    s = gfmd_hip.Solver(nx, ny, 6, gfmd_hip.FP64)
    s.set_kernel("fcc100", "x 2 1.0 0.1 0.3 0.05 0.2 0.02 height 8")
    g = RNG.normal(size=(6, nx, ny))
    C = np.diag(np.linspace(0.1, 0.6, 6))
    C[2, 5] = C[5, 2] = 0.02
    got = s.prec_gradient(g, C)
    assert rel(got, O.prec_gradient(phi, g, C, full=True)) < 1e-12
    ref = O.prec_gradient(phi, g, C)
    assert rel(got[:3], ref[:3]) < 1e-12


def test_round_trip_through_the_force():
    """prec_gradient(f(u)) = -(nx ny)^2 u (cavg = 0): the fast path at
    512^2, FP64, with the x -> -x symmetric table."""
    n = 512
    s = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 128")
    assert s.phi_mirror()
    u = RNG.normal(size=(3, n, n))
    f, _ = s.post_force(u)
    gp = s.prec_gradient(f)
    assert rel(gp, -float(n * n) ** 2 * u) < 1e-10


def test_table_cache_and_device_api():
    import torch
    nx = ny = 64
    g = RNG.normal(size=(3, nx, ny))
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 16")
    a = s.prec_gradient(g, C3)
    b = s.prec_gradient(g, 2 * C3)
    assert rel(s.prec_gradient(g, C3), a) == 0.0
    assert rel(b, a) > 1e-3
    s.set_kernel("sc100", "height 4")  # a new Phi rebuilds the table
    phi = O.fill_phi(O.KERNEL_SC100, [4], 3, nx, ny, normalize=True)
    assert rel(s.prec_gradient(g, C3), O.prec_gradient(phi, g, C3)) < 1e-12
    for prec, dt, tol in ((gfmd_hip.FP64, torch.float64, 1e-13),
                          (gfmd_hip.FP32, torch.float32, 1e-6)):
        s = gfmd_hip.Solver(nx, ny, 3, prec)
        s.set_kernel("sc100", "height 16")
        host = s.prec_gradient(g, C3)
        gd = torch.tensor(g, dtype=dt, device="cuda")
        pd = torch.empty_like(gd)
        torch.cuda.synchronize()
        s.prec_gradient_device(gd.data_ptr(), pd.data_ptr(), C3)
        s.synchronize()
        assert rel(pd.cpu().double().numpy(), host) < tol


def test_errors():
    s = gfmd_hip.Solver(16, 16, 3, gfmd_hip.FP64)
    g = np.zeros((3, 16, 16))
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        s.prec_gradient(g)
    assert ei.value.code == gfmd_hip.ERR_STATE
    s.set_kernel("sc100", "height 4")
    C = C3.copy()
    C[0, 1] = 0.2
    with pytest.raises(gfmd_hip.GFMDError, match="symmetric") as ei:
        s.prec_gradient(g, C)
    assert ei.value.code == gfmd_hip.ERR_ARG
    s.set_phi(np.zeros((16 * 16, 3, 3), np.complex128))
    with pytest.raises(gfmd_hip.GFMDError, match="singular") as ei:
        s.prec_gradient(g)
    assert ei.value.code == gfmd_hip.ERR_KERNEL
