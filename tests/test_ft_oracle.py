"""The `ft` stiffness kernel's oracle (oracle/ft.py) and the library's host
tables (gfmd_ft_cell), on the CPU.

What pins the oracle:
  - the reference's own crystal surface and pair-potential force constants
    (golden fixtures from the reference sources, tests/golden/ft_reference.
    npz): cell, positions and neighbour list bit for bit, force-constant
    and on-site blocks to 1e-15;
  - the reference's sc100 and fcc100 kernels, which `ft` with the same
    springs reproduces (its Hertz tests use `ft sc100` / `ft fcc100`):
    Phi(q) to 1e-12 relative, one- and two-atom cells, with kappa;
  - a layer identity no single-cell check can fake: a two-cell surface of
    height H, with its lower cell eliminated (Schur complement), is the
    one-cell surface of height 2H + 1 (eigenvalues, 1e-12), for sc100 and
    the two-atom fcc111 cell;
  - lj/cut (its LAMMPS pair style cannot be built here): V' and V'' by
    finite differences of the 12-6 potential, and the Gamma-point force
    -sum_{ab <= 0} V' e_z equal to +sum_{ab > 0} V' e_z (the full
    neighbour shell of a bulk atom carries no net force).
"""
import os

import numpy as np
import pytest

import ft

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(os.path.join(ROOT, "tests", "golden",
                                     "ft_reference.npz"), allow_pickle=False))


def test_geometry_matches_reference(gold):
    for c, args in enumerate(gold["geom_args"]):
        k = ft.FTKernel(str(args))
        s = k.surf
        assert np.array_equal(np.array(s.cell), gold["g%d_cell" % c]), args
        assert np.array_equal(np.array(s.pos).ravel(),
                              gold["g%d_positions" % c]), args
        ints = np.array([[n["i"], n["j"], n["x"], n["y"], n["ab"],
                          n["shell"]] for n in s.neigh])
        assert np.array_equal(ints, gold["g%d_ints" % c]), args
        D = np.array([k.fc.offsite(n).ravel() for n in s.neigh])
        nz = gold["g%d_nonzero" % c]
        assert np.array_equal(np.nonzero(np.abs(D).sum(axis=1))[0], nz)
        assert np.abs(D[nz] - gold["g%d_D" % c]).max() < 1e-15
        assert np.abs(k.on[0].reshape(-1, 9) - gold["g%d_on0" % c]).max() < 1e-14
        assert np.abs(k.on[1].reshape(-1, 9) - gold["g%d_on1" % c]).max() < 1e-14


def test_ft_reproduces_reference_kernels(gold):
    Q = gold["q"]
    for c, args in enumerate(gold["phi_args"]):
        k = ft.FTKernel(str(args))
        ref = gold["p%d_phi" % c]
        got = np.array([k.phi(qx, qy) for qx, qy in Q])
        assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max(), args


def _schur(P, keep):
    drop = [i for i in range(P.shape[0]) if i not in keep]
    A = P[np.ix_(keep, keep)]
    B = P[np.ix_(keep, drop)]
    C = P[np.ix_(drop, keep)]
    D = P[np.ix_(drop, drop)]
    return A - B @ np.linalg.solve(D, C)


@pytest.mark.parametrize("surf,fc,nat", [
    ("sc100", "pair-potential 2 1.0 1.0", 1),
    ("fcc100", "pair-potential 1 1.0", 1),
    ("fcc111", "pair-potential 1 1.0", 2),
    ("fcc111", "pair-potential 2x 1.0 0.1 0.3 0.05", 2)])
def test_two_cells_reduce_to_one_cell(surf, fc, nat):
    rng = np.random.default_rng(5)
    H = 3
    k2 = ft.FTKernel("%s 1.0 2 %s height %d" % (surf, fc, H))
    k1 = ft.FTKernel("%s 1.0 1 %s height %d" % (surf, fc, 2 * H + 1))
    keep = list(range(3 * nat))
    for qx, qy in rng.uniform(-np.pi, np.pi, size=(6, 2)):
        a = np.linalg.eigvalsh(_schur(k2.phi(qx, qy), keep))
        b = np.linalg.eigvalsh(k1.phi(qx, qy))
        assert np.abs(a - b).max() < 1e-12 * np.abs(b).max()


def test_lj_derivatives_and_gamma_force():
    eps, sigma, rc = 1.3, 0.9, 2.1
    k = ft.FTKernel("fcc111 1.0 2 lj/cut %r %r %r height 2" % (eps, sigma, rc))
    fc = k.fc
    V = lambda r: 4 * eps * ((sigma / r) ** 12 - (sigma / r) ** 6)
    for r in (0.95, 1.0, 1.2, 1.7):
        h = 1e-5
        assert abs(fc.d1(r) - (V(r + h) - V(r - h)) / (2 * h)) < 1e-7
        assert abs(fc.d2(r) - (fc.d1(r + h) - fc.d1(r - h)) / (2 * h)) < 1e-5
    assert fc.d1(rc) == 0.0 and fc.d2(rc + 0.1) == 0.0
    above = np.zeros(k.surf.nat)
    for n in k.surf.neigh:
        if n["ab"] > 0:
            above[n["i"]] += fc.d1(n["rn"]) * n["r"][2] / n["rn"]
    assert np.abs(k.linf - above).max() < 1e-12
    assert np.abs(k.linf).max() > 1e-3
    # the pair potential's linf keyword is parsed but never applied
    kp = ft.FTKernel("fcc100 1.0 2 pair-potential 1 1.0 linf 2 0.1 -0.1 "
                     "height 4")
    assert np.all(kp.linf == 0.0)


def test_phi_hermitian_and_periodic():
    k = ft.FTKernel("fcc111 1.0 2 pair-potential 2 1.0 0.3 height 5")
    for qx, qy in [(0.3, -1.1), (np.pi, 0.2), (2.0, 2.5)]:
        P = k.phi(qx, qy)
        assert np.abs(P - P.conj().T).max() < 1e-12 * np.abs(P).max()
        assert np.abs(k.phi(-qx, -qy) - P.conj()).max() < 1e-12 * np.abs(P).max()
        assert np.abs(k.phi(qx + 2 * np.pi, qy) - P).max() < 1e-12 * np.abs(P).max()


def test_library_host_geometry_matches_reference(gold):
    """gfmd_ft_cell runs on the host (no device): the library's own surface
    cell is the reference's, bit for bit."""
    import gfmd_hip
    for c, args in enumerate(gold["geom_args"]):
        got = gfmd_hip.ft_cell(str(args))
        assert np.array_equal(got["cell"].ravel(), gold["g%d_cell" % c])
        assert np.array_equal(got["positions"].ravel(),
                              gold["g%d_positions" % c])
        assert np.all(got["linf"] == 0.0)
    lj = "fcc100 1.0 2 lj/cut 1.0 0.8908987181403393 1.5 height 3"
    assert np.abs(gfmd_hip.ft_cell(lj)["linf"] - ft.FTKernel(lj).linf).max() < 1e-14


@pytest.mark.parametrize("args,msg", [
    ("fcc110 1.0 1 pair-potential 1 1.0", "Could not find crystal surface"),
    ("fcc111 1.0 1", "Missing interaction model"),
    ("fcc111 1.0 1 eam 1", "Could not find interaction model"),
    ("sc100 1.0 1 pair-potential 2 1.0", "spring constant"),
    ("sc100 1.0 0 pair-potential 1 1.0", "number of lattice cells"),
    ("sc100 1.0 1 pair-potential 1 1.0 linf 2 0.1 0.2 height 3",
     "sum to zero"),
    ("sc100 1.0 1 lj/cut 1.0 1.0", "epsilon sigma cutoff")])
def test_library_rejects_bad_arguments(args, msg):
    import gfmd_hip
    with pytest.raises(gfmd_hip.GFMDError, match=msg):
        gfmd_hip.ft_cell(args)
