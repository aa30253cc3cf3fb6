"""The allocation-free force evaluation (gfmd_post_force_noalloc) against
the handle's gfmd_post_force_device on the same table and field, through
the C ABI: two calls in a row, the second on a smaller grid with another
ndof, in FP64 (1e-12 relative) and FP32 (1e-6); sizes beyond the static
scratch and non-power-of-two grids are rejected."""
import numpy as np
import pytest

import gfmd_hip

pytestmark = pytest.mark.gpu


def _case(torch, n_x, n_y, ndof, prec, kernel, args, linf, seed):
    dt = torch.float64 if prec == gfmd_hip.FP64 else torch.float32
    s = gfmd_hip.Solver(n_x, n_y, ndof, prec)
    s.set_kernel(kernel, args)
    lin = None
    if linf is not None:
        s.set_linf(np.array(linf))
        lin = torch.tensor(linf, dtype=torch.float64, device="cuda")
    phi = torch.tensor(s.dump_stiffness(), dtype=dt, device="cuda")
    u = torch.tensor(0.01 * np.random.default_rng(seed).standard_normal(
        (ndof, n_x, n_y)), dtype=dt, device="cuda")
    f_ref = torch.empty_like(u)
    e_ref = torch.zeros(1, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    s.post_force_device(u.data_ptr(), f_ref.data_ptr(), e_ref.data_ptr())
    s.synchronize()
    u0_ref = s.get_u0()
    return s, phi, lin, u, f_ref, e_ref, u0_ref


@pytest.mark.parametrize("prec", [gfmd_hip.FP64, gfmd_hip.FP32])
def test_noalloc_matches_handle_twice(prec):
    import torch
    tol = 1e-12 if prec == gfmd_hip.FP64 else 1e-6
    cases = [(256, 512, 3, "sc100", "height 64", None, 1),
             (128, 64, 6, "fcc100", "2 1.0 0 0 height 8", [0.01, -0.01], 2)]
    for nx, ny, ndof, kern, args, linf, seed in cases:
        s, phi, lin, u, f_ref, e_ref, u0_ref = _case(
            torch, nx, ny, ndof, prec, kern, args, linf, seed)
        f = torch.full_like(u, float("nan"))
        e = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
        u0 = torch.full((ndof,), float("nan"), dtype=torch.float64,
                        device="cuda")
        gfmd_hip.post_force_noalloc(
            nx, ny, ndof, prec, phi.data_ptr(), s.phi_mirror(),
            lin.data_ptr() if lin is not None else None, u.data_ptr(),
            f.data_ptr(), e.data_ptr(), u0.data_ptr())
        torch.cuda.synchronize()
        fr = f_ref.double().cpu().numpy()
        d = np.abs(f.double().cpu().numpy() - fr).max() / np.abs(fr).max()
        de = abs(e.item() - e_ref.item()) / abs(e_ref.item())
        print("noalloc %dx%d ndof %d prec %d: force %.2e energy %.2e"
              % (nx, ny, ndof, prec, d, de), flush=True)
        assert d < tol and de < tol
        assert np.allclose(u0.cpu().numpy(), u0_ref, rtol=tol, atol=1e-14)
        s.close()


def test_noalloc_rejects_what_it_cannot_hold():
    import torch
    t = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = t.data_ptr()
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        gfmd_hip.post_force_noalloc(4096, 4096, 3, gfmd_hip.FP64, p, 0, None,
                                    p, p + 8, p + 16)
    assert ei.value.code == gfmd_hip.ERR_ARG
    assert "scratch" in str(ei.value)
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        gfmd_hip.post_force_noalloc(100, 64, 3, gfmd_hip.FP64, p, 0, None,
                                    p, p + 8, p + 16)
    assert ei.value.code == gfmd_hip.ERR_UNSUPPORTED
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        gfmd_hip.post_force_noalloc(64, 64, 3, gfmd_hip.FP64, p, 0, None,
                                    p, p, p + 16)
    assert ei.value.code == gfmd_hip.ERR_ARG
