"""The `fix wall/map` oracle (oracle/wall_map.py) and the library's argument
handling, on the CPU.

What pins the oracle:
  - the reference's own Table2D (golden vectors from the reference sources,
    tests/golden/table2d_reference.npz): value and slopes bit for bit;
  - the spline system's Gauss-Jordan inverse against the exact integer
    inverse of the bicubic system, and A A^-1 = I;
  - the spline's defining properties: node values, central-difference node
    slopes, C1 continuity across box edges, constants reproduced;
  - every potential: f = -dV/ddz by finite differences, continuity of the
    quadratic extension below 0.01 cutoff and of the smooth tail of
    lj93/smooth (value, force and force slope), zero energy at the cutoff
    (and zero force for lj93/smooth), against hand-derived values of the
    12-6 and 9-3 forms;
  - the (100) and (111) maps: slopes by finite differences and the lattice
    periodicity;
  - the wall force: -grad of the total wall energy (finite differences)
    over a bicubic map, which pins the sign and chain-rule scaling of the
    lateral forces;
  - the low-pass filter against a brute-force DFT, and the map shift.
"""
import os

import numpy as np
import pytest

import wall_map as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(os.path.join(ROOT, "tests", "golden",
                                     "table2d_reference.npz"),
                        allow_pickle=False))


def test_table2d_matches_reference(gold):
    for c in range(int(gold["count"])):
        t = W.Table2D(gold["t%d_values" % c])
        f, fx, fy = t.eval(gold["t%d_x" % c], gold["t%d_y" % c])
        ref = gold["t%d_ref" % c]
        assert np.array_equal(f, ref[:, 0])
        assert np.array_equal(fx, ref[:, 1])
        assert np.array_equal(fy, ref[:, 2])


def test_spline_inverse():
    A = np.array(W.spline_matrix())
    Ai = W.spline_inverse()
    exact = np.linalg.inv(A)
    assert np.abs(exact - np.round(exact)).max() < 1e-12
    assert np.abs(Ai - np.round(exact)).max() < 1e-13
    assert np.abs(A @ Ai - np.eye(16)).max() < 1e-13


def test_spline_properties():
    rng = np.random.default_rng(3)
    v = rng.normal(size=(9, 7))
    t = W.Table2D(v)
    X, Y = np.meshgrid(np.arange(9.0), np.arange(7.0), indexing="ij")
    f, fx, fy = t.eval(X, Y)
    assert np.abs(f - v).max() < 1e-14
    cx = (np.roll(v, -1, 0) - np.roll(v, 1, 0)) / 2
    cy = (np.roll(v, -1, 1) - np.roll(v, 1, 1)) / 2
    assert np.abs(fx - cx).max() < 1e-13
    assert np.abs(fy - cy).max() < 1e-13
    # C1 across box edges: both sides of x = 3 and y = 4
    e = 1e-9
    y = rng.uniform(0, 7, 50)
    a = t.eval(3 - e, y)
    b = t.eval(3 + e, y)
    for k in range(3):
        assert np.abs(a[k] - b[k]).max() < 1e-7
    x = rng.uniform(0, 9, 50)
    a = t.eval(x, 4 - e)
    b = t.eval(x, 4 + e)
    for k in range(3):
        assert np.abs(a[k] - b[k]).max() < 1e-7
    # periodic, and constants reproduced with zero slope
    p = t.eval(x + 9, y - 14)
    q = t.eval(x, y)
    assert np.abs(p[0] - q[0]).max() < 1e-13
    c = W.Table2D(np.full((5, 6), 2.5)).eval(x, y)
    assert np.abs(c[0] - 2.5).max() < 1e-14
    assert np.abs(c[1]).max() < 1e-14 and np.abs(c[2]).max() < 1e-14


POTS = ["lj93 1.3 0.9 2.5", "lj93/smooth 1.3 0.9 1.8 3.0", "lj126 0.8 1.1 3.0",
        "exp 0.7 0.4 3.0", "expf 0.7 0.4 3.0", "dugdale 0.6 0.5"]


@pytest.mark.parametrize("pot", POTS)
def test_potential_force_is_minus_slope(pot):
    w = W.WallMap("flat 0.0 1 " + pot, 10.0, 10.0)
    lo = 0.6 if pot.startswith("lj126") else 0.05
    dz = np.linspace(lo, 2.9, 57)
    if pot.startswith("dugdale"):
        dz = np.linspace(0.01, 0.49, 31)
    e, f, d2, _ = w.potential(dz)
    h = 1e-6
    ep = w.potential(dz + h)[0]
    em = w.potential(dz - h)[0]
    assert np.allclose(f, -(ep - em) / (2 * h), rtol=1e-6,
                       atol=1e-7 * np.abs(f).max())
    if not pot.startswith(("lj126", "dugdale")):
        fp = w.potential(dz + h)[1]
        fm = w.potential(dz - h)[1]
        assert np.allclose(d2, -(fp - fm) / (2 * h), rtol=1e-5,
                           atol=1e-6 * np.abs(d2).max())


def test_potential_values_and_joins():
    # lj93: V = eps (2/15 (s/z)^9 - (s/z)^3) - V(cutoff)
    w = W.WallMap("flat 0.0 1 lj93 1.3 0.9 2.5", 10.0, 10.0)
    V = lambda z: 1.3 * (2.0 / 15 * (0.9 / z) ** 9 - (0.9 / z) ** 3)
    z = np.array([0.5, 0.8, 1.7])
    assert np.allclose(w.potential(z)[0], V(z) - V(2.5), rtol=1e-13)
    assert abs(w.potential(np.array([2.5]))[0][0]) < 1e-15
    # quadratic extension joins at 0.01 cutoff: value, force, slope
    m = 0.025
    a = w.potential(np.array([m - 1e-12, m + 1e-12]))
    for k in range(3):
        assert abs(a[k][0] - a[k][1]) < 1e-6 * abs(a[k][1])
    # lj126: V = 4 eps ((s/z)^12 - (s/z)^6) - V(cutoff)
    w = W.WallMap("flat 0.0 1 lj126 0.8 1.1 3.0", 10.0, 10.0)
    V = lambda z: 4 * 0.8 * ((1.1 / z) ** 12 - (1.1 / z) ** 6)
    assert np.allclose(w.potential(z + 0.5)[0], V(z + 0.5) - V(3.0),
                       rtol=1e-13)
    assert w.potential(np.array([0.02]))[3][0]  # too close
    # lj93/smooth: value, force and slope continuous at the inner cutoff,
    # energy and force zero at the outer one
    w = W.WallMap("flat 0.0 1 lj93/smooth 1.3 0.9 1.8 3.0", 10.0, 10.0)
    a = w.potential(np.array([1.8 - 1e-10, 1.8 + 1e-10]))
    for k in range(3):
        assert abs(a[k][0] - a[k][1]) < 1e-8 * max(1.0, abs(a[k][1]))
    e, f, d2, _ = w.potential(np.array([3.0]))
    assert abs(e[0]) < 1e-14 and abs(f[0]) < 1e-14 and abs(d2[0]) < 1e-13
    # exp / expf zero energy at the cutoff; dugdale constant force
    for pot in ("exp 0.7 0.4 3.0", "expf 0.7 0.4 3.0"):
        w = W.WallMap("flat 0.0 1 " + pot, 10.0, 10.0)
        assert abs(w.potential(np.array([3.0]))[0][0]) < 1e-15
    w = W.WallMap("flat 0.0 1 expf 0.7 0.4 3.0", 10.0, 10.0)
    assert abs(w.potential(np.array([3.0]))[1][0]) < 1e-15
    w = W.WallMap("flat 0.0 1 dugdale 0.6 0.5", 10.0, 10.0)
    e, f, _, _ = w.potential(np.array([0.1, 0.4, -0.1]))
    assert np.allclose(f, [1.2, 1.2, 0.0]) and e[2] == 1e12
    assert np.allclose(e[:2], 1.2 * (0.5 - np.array([0.1, 0.4])))


@pytest.mark.parametrize("kind", ["100", "111"])
def test_lattice_maps(kind):
    w = W.WallMap("%s 1.0 0.3 0.5 1 lj93 1 1 2" % kind, 12.0, 12.0)
    rng = np.random.default_rng(4)
    x, y = rng.uniform(0, 12, 40), rng.uniform(0, 12, 40)
    z, zx, zy = w.height(x, y)
    h = 1e-6
    assert np.allclose(zx, (w.height(x + h, y)[0] - w.height(x - h, y)[0]) /
                       (2 * h), atol=1e-8)
    assert np.allclose(zy, (w.height(x, y + h)[0] - w.height(x, y - h)[0]) /
                       (2 * h), atol=1e-8)
    # lattice periodicity, the periods rounded to fit the box: (100) a0
    # along x and y; (111) the triangular lattice (2 pi/qx, 0),
    # (pi/qx, 2 pi/qy) with 2 pi/qx = 12/7 (near sqrt(3) a0)
    px, py = 2 * np.pi / w.qx, 2 * np.pi / w.qy
    assert np.isclose(py, 1.0)
    if kind == "100":
        assert np.isclose(px, 1.0)
        assert np.allclose(w.height(x, y + py)[0], z, atol=1e-12)
    else:
        assert np.isclose(px, 12.0 / 7)
        assert np.allclose(w.height(x + px / 2, y + py)[0], z, atol=1e-12)
    assert np.allclose(w.height(x + px, y)[0], z, atol=1e-12)
    assert abs(np.mean(w.height(*np.meshgrid(np.linspace(0, 12, 240,
                                                         endpoint=False),
                                             np.linspace(0, 12, 240,
                                                         endpoint=False)))[0])
               - 0.5) < 1e-12


def test_wall_force_is_minus_energy_gradient():
    rng = np.random.default_rng(5)
    hmap = rng.normal(scale=0.3, size=(16, 12))
    w = W.WallMap("table 2.0 -1 lj93/smooth 1.0 1.0 1.5 2.5", 8.0, 6.0,
                  map_values=hmap)
    pos = np.column_stack([rng.uniform(0, 8, 60), rng.uniform(0, 6, 60),
                           rng.uniform(0.0, 1.2, 60)])
    f, vec, _ = w.pre_force(pos.copy())
    h = 1e-6
    for k in range(3):
        dp = pos.copy()
        dm = pos.copy()
        dp[:, k] += h
        dm[:, k] -= h
        ep = np.array([w.pre_force(dp[i:i + 1].copy())[1][0]
                       for i in range(60)])
        em = np.array([w.pre_force(dm[i:i + 1].copy())[1][0]
                       for i in range(60)])
        assert np.allclose(f[:, k], -(ep - em) / (2 * h), rtol=1e-5,
                           atol=1e-6 * np.abs(f).max())
    assert vec[4] == vec[5] + vec[6]  # within the cutoff: repulsive or not
    assert np.isclose(vec[3], f[:, 2].sum())


def test_pendist_reset():
    w = W.WallMap("flat 1.0 -1 lj93 1.0 1.0 2.5 pendist 0.2", 10.0, 10.0)
    pos = np.array([[1.0, 1.0, 0.9], [2.0, 2.0, 0.5], [3.0, 3.0, 1.3]])
    f, vec, _ = w.pre_force(pos)
    # dz = -(z - 1): 0.1 (< 0.2: reset to z = 0.8), 0.5, -0.3 (reset)
    assert np.allclose(pos[:, 2], [0.8, 0.5, 0.8])
    assert f[0, 2] == f[2, 2]


def test_filter_against_brute_force_dft():
    rng = np.random.default_rng(6)
    nx, ny, xprd, yprd, lam = 6, 5, 3.0, 2.5, 0.9
    h = rng.normal(size=(nx, ny))
    got = W.filter_map(h, lam, xprd, yprd)
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    out = np.zeros((nx, ny))
    for kx in range(nx):
        for ky in range(ny):
            dx, dy = min(kx, nx - kx), min(ky, ny - ky)
            if np.hypot(dx / xprd, dy / yprd) > 1 / lam:
                continue
            ph = np.exp(-2j * np.pi * (kx * X / nx + ky * Y / ny))
            c = np.sum(h * ph)
            out += (c * np.conj(ph)).real / (nx * ny)
    assert np.abs(got - out).max() < 1e-13
    assert np.abs(W.filter_map(got, lam, xprd, yprd) - got).max() < 1e-13
    assert np.abs(W.filter_map(h, 1e-3, xprd, yprd) - h).max() < 1e-13


def test_displacement():
    rng = np.random.default_rng(7)
    h = rng.normal(size=(8, 8))
    for t, ref in (("average", 0.0), ("bot", None), ("top", None)):
        s, rms, avg = W.displace_map(h, 1.5, t)
        if t == "average":
            assert abs(s.mean() - 1.5) < 1e-14
        elif t == "bot":
            assert abs(s.min() - 1.5) < 1e-14
        else:
            assert abs(s.max() - 1.5) < 1e-14
        assert abs(rms - h.std()) < 1e-13
        assert abs(avg - s.mean()) < 1e-13


def test_map_text_reader():
    txt = "1 -2.5 3e-1\n4 5 6\n7 8 9\n"
    v = W.read_map_text(txt)
    assert v.shape == (3, 3) and v[0, 1] == -2.5 and v[0, 2] == 0.3
    with pytest.raises(W.WallMapError, match="Error read map"):
        W.read_map_text("1 2 3\n4 5\n")


@pytest.mark.parametrize("args,msg", [
    ("flat 1.0 -1", "Illegal fix command"),
    ("flat abc -1 lj93 1 1 2", "Could not convert z0"),
    ("flat 1.0 2 lj93 1 1 2", "Surface normal must be 1 or -1"),
    ("flat 1.0 1 lj99 1 1 2", "Unknown potential type 'lj99'"),
    ("flat 1.0 1 lj93 1 1 2 foo", "Unknown keyword 'foo'"),
    ("flat 1.0 1 lj93 1 1 2 lambda 3", "Map filtering can only be used"),
    ("flat 1.0 1 lj93/smooth 1 1 2 1.5", "outer cutoff must be bigger"),
    ("flat 1.0 1 lj93 x 1 2", "Could not convert first potential"),
    ("flat 1.0 1 lj93 1 1 2 pendist", "Expected value after 'pendist'"),
    ("/nonexistent/map.txt 1.0 1 lj93 1 1 2",
     "Could not open contact map file")])
def test_library_rejects_bad_arguments(args, msg):
    """The library parses on the host before touching the device, with the
    reference's messages (as does the oracle)."""
    import gfmd_hip
    with pytest.raises(gfmd_hip.GFMDError, match=msg):
        gfmd_hip.WallMap(args, 10.0, 10.0)
    if "nonexistent" not in args:
        with pytest.raises(W.WallMapError, match=msg):
            W.WallMap(args, 10.0, 10.0)


def test_library_accepts_table_maps():
    """A valid table map passes the host set-up (argument parsing and the
    check of the compiled spline inverse against the host's Gauss-Jordan);
    without a device the call then stops at the first HIP call (status 2),
    never with an argument error (status 1)."""
    import gfmd_hip
    try:
        gfmd_hip.WallMap("rough.map 1.0 -1 lj93 1 1 2 lambda 3", 8.0, 8.0,
                         heights=np.zeros((8, 8)))
    except gfmd_hip.GFMDError as e:
        assert e.code == 2, str(e)
