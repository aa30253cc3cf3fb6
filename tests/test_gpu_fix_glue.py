"""fix gfmd's atom list <-> grid glue on the device (gfmd_fix_*) against the
oracle (oracle.FixGlue): gather bit for bit (grid ids, displacements,
velocities), the layer shift, the full step atoms -> grid -> elastic force
-> atoms, fix_center, and the error counts."""
import numpy as np
import pytest

import gfmd_hip
import oracle as O

pytestmark = pytest.mark.gpu


def _atoms_case(nx, ny, nl, seed, a=1.0):
    """A shuffled list of nl atoms per cell on an nx x ny grid, displaced
    at random (some across the periodic boundary), with image flags."""
    rng = np.random.default_rng(seed)
    X, Y, U = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nl),
                          indexing="ij")
    n = nx * ny * nl
    xeq = np.column_stack([a * X.ravel() + 0.3 * U.ravel(), a * Y.ravel(),
                           -0.5 * U.ravel()])
    gid = np.column_stack([X.ravel(), Y.ravel(), U.ravel()]).astype(np.int32)
    perm = rng.permutation(n)
    xeq, gid = xeq[perm], gid[perm]
    x = xeq + rng.normal(scale=0.05, size=(n, 3))
    # a tenth of the atoms wrapped back through the x or y boundary
    img = np.full(n, 512 | (512 << 10) | (512 << 20), np.int32)
    k = rng.choice(n, n // 10, replace=False)
    half = k[: len(k) // 2]
    x[half, 0] -= nx * a
    img[half] = 513 | (512 << 10) | (512 << 20)
    rest = k[len(k) // 2:]
    x[rest, 1] += ny * a
    img[rest] = 512 | (511 << 10) | (512 << 20)
    v = rng.normal(size=(n, 3))
    mass = rng.uniform(0.5, 2.0, n)
    return x, xeq, gid, v, img, mass


def _dev(arr, torch, dtype=None):
    t = torch.tensor(np.ascontiguousarray(arr), device="cuda")
    return t if dtype is None else t.to(dtype)


def _atoms(t, n, mask=None, groupbit=0):
    return gfmd_hip.Atoms(n, mask.data_ptr() if mask is not None else None,
                          groupbit, t["gid"].data_ptr(), t["x"].data_ptr(),
                          t["xeq"].data_ptr(), t["v"].data_ptr(),
                          t["img"].data_ptr(), t["mass"].data_ptr())


@pytest.mark.parametrize("nl,prec", [(1, gfmd_hip.FP64), (2, gfmd_hip.FP64),
                                     (1, gfmd_hip.FP32)])
def test_gather_and_scatter_match_oracle(nl, prec):
    import torch
    nx, ny = 32, 24
    box = (nx * 1.0, ny * 1.0, 50.0)
    x, xeq, gid, v, img, mass = _atoms_case(nx, ny, nl, 3)
    n = len(x)
    mask = np.full(n, 1, np.int32)
    out = np.random.default_rng(4).choice(n, 7, replace=False)
    mask[out] = 2                      # not in the group (groupbit 1)
    dt = torch.float64 if prec == gfmd_hip.FP64 else torch.float32
    t = {"gid": _dev(gid, torch), "x": _dev(x, torch), "xeq": _dev(xeq, torch),
         "v": _dev(v, torch), "img": _dev(img, torch),
         "mass": _dev(mass, torch)}
    mk = _dev(mask, torch)
    A = _atoms(t, n, mk, 1)
    s = gfmd_hip.Solver(nx, ny, 3 * nl, prec)
    s.fix_setup(A, box, np.eye(3))
    ug = torch.zeros((3 * nl, nx, ny), dtype=dt, device="cuda")
    vg = torch.zeros_like(ug)
    s.fix_pre_force(A, ug.data_ptr(), vg.data_ptr())
    st = s.fix_state()
    g = O.FixGlue(nx, ny, 3 * nl, box, np.eye(3))
    sel = mask == 1
    g.setup(xeq, mass=mass, mask=sel)
    go = gid.copy()
    uo, vo, placed = g.pre_force(go, x, xeq, v=v, image=img, mass=mass,
                                 mask=sel)
    assert st["placed"] == placed == n - 7 and st["unplaced"] == 0
    assert st["natoms"] == n - 7
    assert np.allclose(st["xeqcm"], g.xeqcm, rtol=1e-14, atol=1e-14)
    assert np.allclose(st["xcm"], g.xcm, rtol=1e-14, atol=1e-14)
    # excluded atoms leave their grid points zero in both
    ud = ug.cpu().double().numpy()
    vd = vg.cpu().double().numpy()
    if prec == gfmd_hip.FP64:
        assert np.array_equal(ud, uo) and np.array_equal(vd, vo)
    else:
        assert np.array_equal(ud, uo.astype(np.float32).astype(np.float64))
    assert np.array_equal(t["gid"].cpu().numpy(), go)
    # the step: elastic force of the gathered grid, scattered to the atoms
    s.set_kernel("sc100" if nl == 1 else "ft",
                 "height 8" if nl == 1 else
                 "fcc111 1.0 1 pair-potential 1 1.0 height 4")
    fg = torch.empty_like(ug)
    e = torch.zeros(1, dtype=torch.float64, device="cuda")
    s.post_force_device(ug.data_ptr(), fg.data_ptr(), e.data_ptr())
    fa = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    s.fix_post_force(A, fg.data_ptr(), fa.data_ptr(), (1, 1, 0))
    st = s.fix_state()
    fgn = fg.cpu().double().numpy()
    fo = np.zeros((n, 3))
    fsum = g.post_force(go, fgn, fo, fix_center=(1, 1, 0), mask=sel)
    assert st["unforced"] == 0
    assert np.allclose(st["fsum"], fsum, rtol=1e-12, atol=1e-12)
    assert np.abs(fa.cpu().numpy() - fo).max() < 1e-12 * max(1.0, np.abs(fo).max())
    assert np.all(fa.cpu().numpy()[out] == 0.0)


def test_layer_slide_shifts_grid_ids():
    import torch
    nx, ny = 16, 16
    x, xeq, gid, v, img, mass = _atoms_case(nx, ny, 1, 5)
    n = len(x)
    t = {"gid": _dev(gid, torch), "x": _dev(xeq, torch),
         "xeq": _dev(xeq, torch), "v": _dev(v, torch),
         "img": _dev(np.full(n, 512 | (512 << 10) | (512 << 20), np.int32),
                     torch),
         "mass": _dev(mass, torch)}
    A = _atoms(t, n)
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    box = (16.0, 16.0, 40.0)
    s.fix_setup(A, box, np.eye(3))
    ug = torch.zeros((3, nx, ny), dtype=torch.float64, device="cuda")
    s.fix_pre_force(A, ug.data_ptr())
    assert s.fix_state()["xshift"] == 0
    t["x"] += torch.tensor([0.0, -1.0, 0.0], dtype=torch.float64,
                           device="cuda")           # one spacing along -y
    s.fix_pre_force(A, ug.data_ptr())
    st = s.fix_state()
    assert (st["xshift"], st["yshift"]) == (0, -1)
    assert np.array_equal(t["gid"].cpu().numpy()[:, 1], (gid[:, 1] - 1) % ny)
    assert np.allclose(ug[1].cpu().numpy(), -1.0)


def test_errors_and_counts():
    import torch
    nx = ny = 8
    x, xeq, gid, v, img, mass = _atoms_case(nx, ny, 1, 6)
    gid[0] = (nx, 0, 0)        # outside the grid
    gid[1] = (0, 0, 1)         # no second atom in a one-atom cell
    t = {"gid": _dev(gid, torch), "x": _dev(x, torch), "xeq": _dev(xeq, torch),
         "v": _dev(v, torch), "img": _dev(img, torch),
         "mass": _dev(mass, torch)}
    A = _atoms(t, len(x))
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    ug = torch.zeros((3, nx, ny), dtype=torch.float64, device="cuda")
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        s.fix_pre_force(A, ug.data_ptr())
    assert ei.value.code == gfmd_hip.ERR_STATE
    s.fix_setup(A, (8.0, 8.0, 20.0), np.eye(3))
    s.fix_pre_force(A, ug.data_ptr())
    st = s.fix_state()
    assert st["unplaced"] == 2 and st["placed"] == len(x) - 2
    fa = torch.zeros((len(x), 3), dtype=torch.float64, device="cuda")
    s.set_kernel("sc100", "height 4")
    s.fix_post_force(A, ug.data_ptr(), fa.data_ptr())
    assert s.fix_state()["unforced"] == 2


def test_wrap_counts_each_axis_separately():
    """Atoms many periods away: 62 boxes in x and 3 in y for one atom (65
    wraps in all), 100 in x for another.  Every atom is placed; up to 64
    periods per axis the displacement equals the reference's wrap loops
    bit for bit, beyond that within rounding of the box."""
    import torch
    nx, ny = 16, 12
    box = (nx * 1.0, ny * 1.0, 50.0)
    x, xeq, gid, v, img, mass = _atoms_case(nx, ny, 1, 9)
    n = len(x)
    img = np.full(n, 512 | (512 << 10) | (512 << 20), np.int32)
    x[3, 0] -= 62 * nx
    x[3, 1] += 3 * ny
    img[3] = (512 + 62) | ((512 - 3) << 10) | (512 << 20)
    x[8, 0] += 100 * nx
    img[8] = (512 - 100) | (512 << 10) | (512 << 20)
    t = {"gid": _dev(gid, torch), "x": _dev(x, torch), "xeq": _dev(xeq, torch),
         "v": _dev(v, torch), "img": _dev(img, torch),
         "mass": _dev(mass, torch)}
    A = _atoms(t, n)
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.fix_setup(A, box, np.eye(3))
    ug = torch.zeros((3, nx, ny), dtype=torch.float64, device="cuda")
    s.fix_pre_force(A, ug.data_ptr())
    st = s.fix_state()
    g = O.FixGlue(nx, ny, 3, box, np.eye(3))
    g.setup(xeq, mass=mass)
    go = gid.copy()
    uo, _, placed = g.pre_force(go, x, xeq, image=img, mass=mass)
    assert st["placed"] == placed == n and st["unplaced"] == 0
    ud = ug.cpu().numpy()
    ix8, iy8 = go[8][0], go[8][1]
    far = np.zeros_like(ud, bool)
    far[:, ix8, iy8] = True
    assert np.array_equal(ud[~far], uo[~far])
    assert np.abs(ud[far] - uo[far]).max() < 1e-12 * nx
