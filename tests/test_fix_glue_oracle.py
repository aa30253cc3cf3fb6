"""The oracle's fix gfmd glue (oracle.FixGlue), on the CPU: hand-worked
displacements with the periodic wrap, the layer shift after a slide by one
lattice spacing, image-flag unwrapping in the centre of mass, fix_center."""
import numpy as np

import oracle as O


def _lattice(nx, ny, a=1.0, z=0.0):
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    xeq = np.column_stack([a * X.ravel(), a * Y.ravel(),
                           np.full(nx * ny, z)])
    gid = np.column_stack([X.ravel(), Y.ravel(), np.zeros(nx * ny, int)])
    return xeq, gid


def test_hand_displacements_and_wrap():
    glue = O.FixGlue(2, 2, 3, (2.0, 2.0, 10.0), np.eye(3))
    xeq, gid = _lattice(2, 2)
    glue.setup(xeq)
    x = xeq.copy()
    x[0] += (0.1, -0.2, 0.3)
    x[3] += (-1.95, 0.05, 0.0)       # crossed the x boundary: u_x = 0.05
    u, _, placed = glue.pre_force(gid, x, xeq)
    assert placed == 4
    assert np.allclose(u[:, 0, 0], (0.1, -0.2, 0.3))
    assert np.allclose(u[:, 1, 1], (0.05, 0.05, 0.0))
    assert (glue.xshift, glue.yshift) == (0, 0)


def test_layer_slide_moves_the_grid_ids():
    nx, ny = 4, 3
    glue = O.FixGlue(nx, ny, 3, (4.0, 3.0, 10.0), np.eye(3))
    xeq, gid = _lattice(nx, ny)
    glue.setup(xeq)
    g0 = gid.copy()
    x = xeq + (1.0, 0.0, 0.0)        # the whole layer one spacing along x
    u, _, placed = glue.pre_force(gid, x, xeq)
    assert (glue.xshift, glue.yshift) == (1, 0) and placed == nx * ny
    assert np.array_equal(gid[:, 0], (g0[:, 0] + 1) % nx)
    assert np.allclose(u[0], 1.0) and np.allclose(u[1:], 0.0)
    # a second call without further motion leaves the ids alone
    g1 = gid.copy()
    glue.pre_force(gid, x, xeq)
    assert np.array_equal(gid, g1)


def test_centre_of_mass_unwraps_images():
    glue = O.FixGlue(2, 1, 3, (2.0, 1.0, 5.0), np.eye(3))
    xeq = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    gid = np.array([[0, 0, 0], [1, 0, 0]])
    glue.setup(xeq, mass=np.array([1.0, 3.0]))
    assert np.allclose(glue.xeqcm, (0.75, 0.0, 0.0))
    x = np.array([[1.9, 0.0, 0.0], [1.0, 0.0, 0.0]])
    img = np.array([(511 << 0) | (512 << 10) | (512 << 20),
                    512 | (512 << 10) | (512 << 20)])  # atom 0 one box left
    glue.pre_force(gid, x, xeq, image=img, mass=np.array([1.0, 3.0]))
    assert np.allclose(glue.xcm, (((1.9 - 2.0) * 1 + 3.0) / 4.0, 0.0, 0.0))


def test_fix_center_removes_the_mean_force():
    glue = O.FixGlue(3, 2, 3, (3.0, 2.0, 9.0), np.eye(3))
    xeq, gid = _lattice(3, 2)
    glue.setup(xeq)
    fg = np.random.default_rng(1).normal(size=(3, 3, 2))
    f = np.zeros((6, 3))
    fsum = glue.post_force(gid, fg, f, fix_center=(1, 0, 1))
    assert np.allclose(fsum, fg.reshape(3, -1).sum(axis=1))
    assert abs(f[:, 0].sum()) < 1e-13 and abs(f[:, 2].sum()) < 1e-13
    assert np.isclose(f[:, 1].sum(), fsum[1])
