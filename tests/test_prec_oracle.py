"""The oracle's preconditioned gradient (oracle.prec_gradient, restating
GFMDSolverStatic::prec_gradient), on the CPU.

What pins it:
  - the round trip through the force: prec_gradient(f(u)) = -(nx ny)^2 u
    with cavg = 0, since f = -F^-1[Phi~ F[u]] and both transforms are
    unnormalised (a wrong sign, transpose or normalisation fails it);
  - a uniform gradient only touches q = 0: gP = nx ny (Phi~(0) + cavg)^-1 g;
  - the operator is self-adjoint (<a, P b> = <P a, b>), and the reference's
    ndof > 3 branch leaves components 3.. untransformed (nx ny g).
"""
import numpy as np
import pytest

import oracle as O


def _case(kind):
    if kind == "sc100":
        return O.fill_phi(O.KERNEL_SC100, [8], 3, 12, 10, normalize=True), 3
    p = O.fcc100_params(8, 2, [1.0, 0.3, 0.2], [0.1, 0.05, 0.02])
    return O.fill_phi(O.KERNEL_FCC100, p, 6, 12, 10, normalize=True), 6


@pytest.mark.parametrize("kind", ["sc100", "fcc100"])
def test_round_trip_through_the_force(kind):
    phi, d = _case(kind)
    rng = np.random.default_rng(1)
    u = rng.normal(size=(d, 12, 10))
    f, _, _ = O.static_force(phi, u)
    N = 12 * 10
    gp = O.prec_gradient(phi, f, full=True)
    assert np.abs(gp + N * N * u).max() < 1e-9 * N * N * np.abs(u).max()
    ref = O.prec_gradient(phi, f)  # the reference's branch
    assert np.abs(ref[:3] - gp[:3]).max() < 1e-9 * N * N
    if d > 3:
        assert np.abs(ref[3:] - N * f[3:]).max() < 1e-9 * N * np.abs(f).max()


def test_uniform_gradient_and_self_adjoint():
    phi, d = _case("sc100")
    N = 120
    C = np.array([[0.3, 0.05, 0.0], [0.05, 0.2, 0.01], [0.0, 0.01, 0.4]])
    g0 = np.array([0.7, -0.2, 1.1])
    g = np.broadcast_to(g0[:, None, None], (3, 12, 10)).copy()
    gp = O.prec_gradient(phi, g, C)
    want = N * np.linalg.solve(phi[0].real + C, g0)
    assert np.abs(gp - want[:, None, None]).max() < 1e-12 * np.abs(want).max()
    rng = np.random.default_rng(2)
    a = rng.normal(size=(3, 12, 10))
    b = rng.normal(size=(3, 12, 10))
    l = np.sum(a * O.prec_gradient(phi, b, C))
    r = np.sum(O.prec_gradient(phi, a, C) * b)
    assert abs(l - r) < 1e-11 * abs(l)
