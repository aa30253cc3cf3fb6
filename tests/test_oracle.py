"""Pins the CPU FP64 oracle (oracle/gfmd_oracle.c) before the HIP path is
compared with it: golden vectors from the reference's own
stiffness core, closed forms, brute force and invariants."""
import numpy as np
import pytest

import oracle as O

RNG = np.random.default_rng(1234)


# ---------------------------------------------------------------------------
# FFT
# ---------------------------------------------------------------------------
def brute_dft(x, sign):
    n = x.size
    k = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return W @ x


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 12, 16, 37, 63, 64,
                               100, 128, 210])
def test_fft_c2c_brute_force(n):
    x = RNG.standard_normal(n) + 1j * RNG.standard_normal(n)
    for sign in (-1, +1):
        got = O.fft_c2c(x, sign)
        ref = brute_dft(x, sign)
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n", [2048, 4096, 3 * 5 * 7 * 8])
def test_fft_c2c_numpy(n):
    x = RNG.standard_normal(n) + 1j * RNG.standard_normal(n)
    assert np.allclose(O.fft_c2c(x, -1), np.fft.fft(x), rtol=0, atol=1e-11)
    assert np.allclose(O.fft_c2c(x, +1), n * np.fft.ifft(x), rtol=0, atol=1e-11)


@pytest.mark.parametrize("shape", [(8, 6), (5, 7), (64, 37), (1, 4), (16, 1)])
def test_fft2d_numpy(shape):
    x = RNG.standard_normal(shape) + 1j * RNG.standard_normal(shape)
    assert np.allclose(O.fft2d(x, -1), np.fft.fft2(x), rtol=0, atol=1e-11)
    assert np.allclose(O.fft2d(x, +1), x.size * np.fft.ifft2(x), rtol=0,
                       atol=1e-11)


def test_invert():
    for d in (1, 3, 6, 12):
        A = RNG.standard_normal((d, d)) + 1j * RNG.standard_normal((d, d))
        assert np.allclose(O.invert(A) @ A, np.eye(d), atol=1e-12)
    with pytest.raises(np.linalg.LinAlgError):
        O.invert(np.zeros((3, 3)))


# ---------------------------------------------------------------------------
# Stiffness Phi(q)
# ---------------------------------------------------------------------------
def test_phi_golden_reference(golden):
    """Oracle Phi(q) == the reference core's Phi(q) (sc100 at several
    heights, including the semi-infinite iteration)."""
    for i, name in enumerate(golden["case_names"]):
        height = int(golden["case%d_height" % i])
        qx, qy = golden["case%d_qx" % i], golden["case%d_qy" % i]
        ref = golden["case%d_phi" % i]
        for k in range(qx.size):
            got = O.kernel_phi(O.KERNEL_SC100, [height], qx[k], qy[k])
            scale = np.abs(ref[k]).max()
            assert np.abs(got - ref[k]).max() <= 1e-12 * scale, (name, k)


def test_fill_phi_golden_reference(golden):
    nx, ny = int(golden["fill_nx"]), int(golden["fill_ny"])
    got = O.fill_phi(O.KERNEL_SC100, [int(golden["fill_height"])], 3, nx, ny)
    ref = golden["fill_phi"]
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


def test_phi_height0_is_U0():
    for qx, qy in [(0.3, -1.2), (np.pi, 0.1), (0.0, 0.0)]:
        U0, U, V = O.sc100_layer(qx, qy)
        assert np.array_equal(O.kernel_phi(O.KERNEL_SC100, [0], qx, qy), U0)


def test_phi_height1_closed_form():
    """height 1: Phi = U0 - V U^-1 V^+ (one transfer-matrix step)."""
    for qx, qy in [(0.3, -1.2), (2.0, 0.7), (0.0, 0.0)]:
        U0, U, V = O.sc100_layer(qx, qy)
        ref = U0 - V @ np.linalg.inv(U) @ V.conj().T
        got = O.kernel_phi(O.KERNEL_SC100, [1], qx, qy)
        assert np.allclose(got, ref, rtol=0, atol=1e-13)


def test_phi_fixed_point_semi_infinite():
    """height < 0 solves X = V (U + X)^-1 (-V^+) to the reference's
    tolerance; check the fixed-point residual directly."""
    qx, qy = 0.4, 1.1
    U0, U, V = O.sc100_layer(qx, qy)
    phi = O.kernel_phi(O.KERNEL_SC100, [-1], qx, qy)
    X = phi - U0
    res = X + V @ np.linalg.inv(U + X) @ V.conj().T
    assert np.abs(res).max() < 1e-7


def test_sc100_layer_matrices():
    """U, U0, V of the sc100 lattice (nn + 2nn central springs, k = 1) from
    a direct lattice sum over the 18 neighbours."""
    qx, qy = 0.37, -1.3
    U0, U, V = O.sc100_layer(qx, qy)
    nbrs = [np.array(v, float) for v in
            [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
             (1, 1, 0), (1, -1, 0), (-1, 1, 0), (-1, -1, 0), (1, 0, 1),
             (-1, 0, 1), (0, 1, 1), (0, -1, 1), (1, 0, -1), (-1, 0, -1),
             (0, 1, -1), (0, -1, -1)]]

    def block(dz, exclude_up=False):
        M = np.zeros((3, 3), complex)
        for r in nbrs:
            if exclude_up and r[2] > 0:
                continue
            e = r / np.linalg.norm(r)
            D = -np.outer(e, e)
            if r[2] == dz:
                M += D * np.exp(1j * (qx * r[0] + qy * r[1]))
        return M

    def onsite(exclude_up=False):
        S = np.zeros((3, 3))
        for r in nbrs:
            if exclude_up and r[2] > 0:
                continue
            e = r / np.linalg.norm(r)
            S += np.outer(e, e)
        return S

    Ubulk = onsite() + block(0)
    Usurf = onsite(exclude_up=True) + block(0)
    # the V block couples a layer to the one below (dz = -1): V = sum over
    # bonds to the lower layer with phase exp(i q.r)
    Vlow = block(-1)
    assert np.allclose(U, Ubulk, atol=1e-14)
    assert np.allclose(U0, Usurf, atol=1e-14)
    assert np.allclose(V, Vlow, atol=1e-14) or np.allclose(V, Vlow.conj().T,
                                                           atol=1e-14)


def test_phi_hermitian_and_inversion_symmetry():
    for qx, qy in [(0.3, 1.2), (-2.0, 0.4), (np.pi, 1.0)]:
        P = O.kernel_phi(O.KERNEL_SC100, [64], qx, qy)
        Pm = O.kernel_phi(O.KERNEL_SC100, [64], -qx, -qy)
        assert np.allclose(P, P.conj().T, atol=1e-13)
        assert np.allclose(Pm, P.conj(), atol=1e-13)
        assert np.all(np.linalg.eigvalsh(0.5 * (P + P.conj().T)) > 0)


def test_sc100_long_wavelength_modulus():
    """Semi-infinite sc100: the normal stiffness with free tangential
    displacements, 1/(Phi^-1)_zz, tends to E* |q| / 2 with E* = 8/3 (the
    contact modulus the reference's Hertz tests use,
    TEST_Hertz_sc100_128x128/eval.py:33-39), in every direction."""
    for q, ang in [(4e-3, 0.0), (4e-3, np.pi / 4), (4e-3, 1.0)]:
        P = O.kernel_phi(O.KERNEL_SC100, [-1], q * np.cos(ang),
                         q * np.sin(ang))
        kzz = 1.0 / np.linalg.inv(P)[2, 2].real
        assert abs(kzz / q - 0.5 * 8.0 / 3.0) < 2e-3


# ---------------------------------------------------------------------------
# Static force evaluation
# ---------------------------------------------------------------------------
def _setup(nx, ny, height=8):
    phi = O.fill_phi(O.KERNEL_SC100, [height], 3, nx, ny, normalize=True)
    return phi


@pytest.mark.parametrize("shape", [(8, 8), (6, 5), (7, 4)])
def test_force_plane_wave(shape):
    """u = Re(A e^{i q0.r}) => f = -Re(Phi(q0) A e^{i q0.r}),
    E = N/4 Re(A^+ Phi(q0) A) (q0 not self-conjugate)."""
    nx, ny = shape
    phi = _setup(nx, ny)
    i0, j0 = 1, 2 % ny
    qx, qy = 2 * np.pi * i0 / nx, 2 * np.pi * j0 / ny
    A = np.array([0.3 + 0.1j, -0.2 + 0.05j, 0.4 - 0.3j])
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ph = np.exp(1j * (qx * X + qy * Y))
    u = np.real(A[:, None, None] * ph[None])
    f, e, _ = O.static_force(phi, u)
    P = O.kernel_phi(O.KERNEL_SC100, [8], qx if i0 <= nx // 2 else qx - 2 * np.pi,
                     qy)
    fref = -np.real((P @ A)[:, None, None] * ph[None])
    assert np.allclose(f, fref, atol=1e-13)
    assert abs(e - nx * ny / 4 * np.real(A.conj() @ P @ A)) < 1e-12


def test_force_energy_identity_and_linearity():
    nx, ny = 8, 6
    phi = _setup(nx, ny)
    u = RNG.standard_normal((3, nx, ny))
    v = RNG.standard_normal((3, nx, ny))
    fu, eu, _ = O.static_force(phi, u)
    fv, ev, _ = O.static_force(phi, v)
    assert abs(eu - (-0.5 * np.sum(u * fu))) < 1e-12 * abs(eu)
    fw, _, _ = O.static_force(phi, 2.0 * u - 3.0 * v)
    assert np.allclose(fw, 2.0 * fu - 3.0 * fv, atol=1e-12)
    assert eu > 0


def test_force_is_minus_energy_gradient():
    nx, ny = 6, 5
    phi = _setup(nx, ny)
    linf = np.array([0.37])
    u = 0.1 * RNG.standard_normal((3, nx, ny))
    f, e, _ = O.static_force(phi, u, linf)
    h = 1e-6
    for idx in [(0, 1, 2), (1, 4, 0), (2, 3, 3), (2, 0, 0)]:
        up, um = u.copy(), u.copy()
        up[idx] += h
        um[idx] -= h
        ep = O.static_force(phi, up, linf)[1]
        em = O.static_force(phi, um, linf)[1]
        assert abs(-(ep - em) / (2 * h) - f[idx]) < 1e-7


def test_force_translation_sum_rule():
    """sum_r f = -Phi(0) sum_r u (q = 0 mode)."""
    nx, ny = 6, 4
    phi = _setup(nx, ny)
    u = RNG.standard_normal((3, nx, ny))
    f, _, u0 = O.static_force(phi, u)
    P0 = O.kernel_phi(O.KERNEL_SC100, [8], 0.0, 0.0)
    assert np.allclose(f.sum(axis=(1, 2)), -np.real(P0 @ u.sum(axis=(1, 2))),
                       atol=1e-12)
    assert np.allclose(u0, u.sum(axis=(1, 2)), atol=1e-12)


def test_force_gamma_linf():
    """Gamma-point linear force: E = E0 - linf sum u_z, f_z = f0_z + linf
    (gfmd_solver_static.cpp:193-231)."""
    nx, ny = 6, 4
    phi = _setup(nx, ny)
    u = RNG.standard_normal((3, nx, ny))
    f0, e0, _ = O.static_force(phi, u)
    f1, e1, _ = O.static_force(phi, u, np.array([0.25]))
    assert abs(e1 - (e0 - 0.25 * u[2].sum())) < 1e-12
    assert np.allclose(f1[:2], f0[:2], atol=1e-13)
    assert np.allclose(f1[2], f0[2] + 0.25, atol=1e-13)


def test_force_real_space_brute_force():
    """Tiny grid: assemble the real-space dynamical matrix column by column
    from unit displacements and check it is symmetric (Hermitian Phi) and
    translation invariant (a circulant)."""
    nx, ny = 4, 3
    phi = _setup(nx, ny)
    n = 3 * nx * ny
    K = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        f, _, _ = O.static_force(phi, e.reshape(3, nx, ny))
        K[:, j] = -f.reshape(-1)
    assert np.allclose(K, K.T, atol=1e-13)
    Kr = K.reshape(3, nx, ny, 3, nx, ny)
    for sx in range(nx):
        for sy in range(ny):
            sh = np.roll(np.roll(Kr, sx, axis=1), sy, axis=2)
            sh = np.roll(np.roll(sh, sx, axis=4), sy, axis=5)
            assert np.allclose(sh, Kr, atol=1e-13)


# ---------------------------------------------------------------------------
# FIRE state machine (src/cuda/fix_fire.cpp:117-297)
# ---------------------------------------------------------------------------
def test_fire_steps_by_hand():
    F = O.Fire(dt=0.1)
    x = np.zeros((3, 1, 2))
    v = np.zeros_like(x)
    f = np.ones_like(x)
    F.initial(x, v, f)
    assert np.allclose(v, 0.05) and np.allclose(x, 0.005)
    P = F.final(v, f)
    # v = 0.1 after the kick; mix: (1-0.1)*0.1 + 0.1*sqrt(vv/ff)*1 = 0.1
    assert P > 0 and np.allclose(v, 0.1)
    assert F.cut == 9 and F.mix == 0.1
    # power negative -> velocities cut, mixing reset
    v[:] = -1.0
    P = F.final(v, f)
    assert P < 0 and np.all(v == 0) and F.cut == 10 and F.mix == 0.1
    # reference quirk: reset_dt restores dt
    assert F.dt == 0.1
    # after minsteps positive steps the mixing decays
    for _ in range(11):
        v[:] = 0.5
        F.final(v, f)
    assert F.cut == -1
    v[:] = 0.5
    F.final(v, f)
    assert abs(F.mix - 0.1 * 0.99) < 1e-15


def test_fire_velocity_limit():
    F = O.Fire(dt=0.1, xlimit=0.01)
    x = np.zeros((3, 1, 1))
    v = np.zeros_like(x)
    f = np.full_like(x, 10.0)
    F.initial(x, v, f)
    assert abs(np.sqrt((v ** 2).sum()) - 0.1) < 1e-14


# ---------------------------------------------------------------------------
# fcc100 (multi-layer) stiffness vs the reference core
# ---------------------------------------------------------------------------
FCC_PARAMS = [
    O.fcc100_params(16, 1, [1.0]),
    O.fcc100_params(32, 2, [57.146, 0, 0]),
    O.fcc100_params(8, 2, [1.0, 0.3, 0.2], [0.1, 0.05, 0.02]),
    O.fcc100_params(12, 2, [1.0, 0.5, 0.25], normal=-1),
    O.fcc100_params(4, 1, [1.0], [0.2]),
    O.fcc100_params(-1, 2, [1.0, 0.0, 0.0]),
    O.fcc100_params(6, 3, [1.0, 0.4, 0.2], [0.1, 0.05, 0.03]),
    O.fcc100_params(5, 4, [1.0, 0.3, 0.1], normal=-1),
]


def test_fcc100_golden_reference(golden):
    names = list(golden["fcc_names"])
    assert len(names) == len(FCC_PARAMS)
    for i, prm in enumerate(FCC_PARAMS):
        qx, qy = golden["fcc%d_qx" % i], golden["fcc%d_qy" % i]
        ref = golden["fcc%d_phi" % i]
        assert ref.shape[1] == O.kernel_dim(O.KERNEL_FCC100, prm)
        for k in range(qx.size):
            got = O.kernel_phi(O.KERNEL_FCC100, prm, qx[k], qy[k])
            scale = np.abs(ref[k]).max()
            assert np.abs(got - ref[k]).max() <= 1e-11 * scale, (names[i], k)


def test_fcc100_properties():
    """Two layers: Hermitian, Phi(-q) = conj Phi(q), positive definite off
    q = 0, and the q = 0 translation mode is soft (finite height)."""
    prm = O.fcc100_params(32, 2, [57.146, 0, 0])
    for qx, qy in [(0.3, 1.2), (-2.0, 0.4), (np.pi, np.pi)]:
        P = O.kernel_phi(O.KERNEL_FCC100, prm, qx, qy)
        Pm = O.kernel_phi(O.KERNEL_FCC100, prm, -qx, -qy)
        assert np.allclose(P, P.conj().T, atol=1e-10 * np.abs(P).max())
        assert np.allclose(Pm, P.conj(), atol=1e-10 * np.abs(P).max())
        assert np.linalg.eigvalsh(0.5 * (P + P.conj().T)).min() > 0


ISO_PARAMS = [O.isotropic_params(0.3, 1.0, 1.0, 32.0),
              O.isotropic_params(0.5, 2.0, 0.7, 64.0),
              O.isotropic_params(0.0, 1.0, 1.0, 16.0),
              O.isotropic_params(0.25, 0.8, 3.0, 128.0)]


def test_isotropic_golden_reference(golden):
    """IsotropicStiffnessKernel (isotropic_stiffness.cpp:95-167) values from
    the reference core, including the q = 0 entry that uses the domain
    length."""
    args = list(golden["iso_args"])
    assert len(args) == len(ISO_PARAMS)
    Q = golden["q"]
    for i, prm in enumerate(ISO_PARAMS):
        assert float(golden["iso%d_L" % i]) == prm[3]
        ref = golden["iso%d_phi" % i]
        for k in range(Q.shape[0]):
            got = O.kernel_phi(O.KERNEL_ISOTROPIC, prm, Q[k, 0], Q[k, 1])
            scale = np.abs(ref[k]).max()
            assert np.abs(got - ref[k]).max() <= 1e-12 * scale, (args[i], k)


def test_isotropic_continuum_properties():
    """Half-space elasticity: Phi is Hermitian positive definite, linear in
    |q| (no length scale), covariant under in-plane rotations, and its
    normal compliance is the Boussinesq value (1 - nu)/(myu |q|)."""
    nu, myu = 0.3, 1.7
    prm = O.isotropic_params(nu, myu, 1.0, 64.0)
    q = np.array([0.37, -0.81])
    P = O.kernel_phi(O.KERNEL_ISOTROPIC, prm, *q)
    assert np.abs(P - P.conj().T).max() < 1e-14 * np.abs(P).max()
    assert np.linalg.eigvalsh(P).min() > 0
    P2 = O.kernel_phi(O.KERNEL_ISOTROPIC, prm, *(2 * q))
    assert np.abs(P2 - 2 * P).max() < 1e-13 * np.abs(P).max()
    th = 0.6
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0],
                  [0, 0, 1]])
    PR = O.kernel_phi(O.KERNEL_ISOTROPIC, prm, *(R[:2, :2] @ q))
    assert np.abs(PR - R @ P @ R.T).max() < 1e-13 * np.abs(P).max()
    G = np.linalg.inv(P)
    b = np.hypot(*q)
    assert abs(G[2, 2] - (1 - nu) / (myu * b)) < 1e-13 / b
    # Cerruti: tangential compliance (1/q - nu qx^2/q^3)/myu
    assert abs(G[0, 0] - (1 / b - nu * q[0] ** 2 / b ** 3) / myu) < 1e-13 / b


def _plane_wave_mode(nx, ny, i0, j0, params):
    """Real field of the lowest eigenvector e of Phi(q0) and its eigenvalue
    (q0 = 2 pi (i0/nx, j0/ny))."""
    qx, qy = 2 * np.pi * i0 / nx, 2 * np.pi * j0 / ny
    lam, vec = np.linalg.eigh(O.kernel_phi(O.KERNEL_SC100, params, qx, qy))
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ph = np.exp(1j * (qx * X + qy * Y))
    return np.real(vec[:, 0][:, None, None] * ph[None]), lam[0]


def test_dynamics_oracle_second_order_damped_oscillator():
    """A plane-wave eigenmode of Phi is a damped harmonic oscillator
    x'' = -(lambda/m) x - gamma x'; the integrator must converge to its
    exact solution with second-order error."""
    nx, ny, m, gamma = 8, 6, 1.3, 0.4
    phi = O.fill_phi(O.KERNEL_SC100, [4], 3, nx, ny, normalize=True)
    mode, lam = _plane_wave_mode(nx, ny, 1, 2, [4])
    w2 = lam / m
    wd = np.sqrt(w2 - gamma ** 2 / 4)
    x0, v0, T = 0.02, -0.01, 3.0

    def exact(t):
        return np.exp(-gamma * t / 2) * (x0 * np.cos(wd * t) +
                                         (v0 + gamma * x0 / 2) / wd *
                                         np.sin(wd * t))

    errs = []
    for dt in (0.05, 0.025, 0.0125):
        n = int(round(T / dt))
        u, v, _ = O.dynamics(phi, x0 * mode, v0 * mode, n, dt, mass=m,
                             damping=gamma)
        amp = np.sum(u * mode) / np.sum(mode * mode)
        assert np.abs(u - amp * mode).max() < 1e-12  # stays in the mode
        errs.append(abs(amp - exact(T)))
    assert errs[-1] < 1e-4 * x0
    for a, b in zip(errs, errs[1:]):
        assert 3.5 < a / b < 4.5


def test_dynamics_oracle_nve_drift_reference_case():
    """tests/TEST_energy_conservation_fcc100: one-layer fcc100 (k = 1) on
    10 x 10 sites, random displacements of 0.1, timestep 0.01: total energy
    within 1e-4 of its mean (here over 2000 steps)."""
    nx = ny = 10
    prm = O.fcc100_params(0, 1, [1.0])
    phi = O.fill_phi(O.KERNEL_FCC100, prm, 3, nx, ny, normalize=True)
    rng = np.random.default_rng(12472634)
    u = rng.uniform(-0.1, 0.1, (3, nx, ny))
    _, _, h = O.dynamics(phi, u, np.zeros_like(u), 2000, 0.01)
    e = h[:, 2]
    assert np.abs(e - e.mean()).max() / e.mean() < 1e-4
    assert h[:, 0].max() > 0.1 * e.mean()  # energy really exchanges


def test_hardwall_contact_oracle_kkt():
    """The exact complementarity solution satisfies what the mathematics
    fixes: p >= 0, gap >= 0, p gap = 0 on every site, force balance of the
    q = 0 mode (sum of wall pressures = load minus the substrate spring
    Phi_zz(0) <u_z>), and zero elastic+load force off the contact."""
    nx, ny, height, p0 = 12, 10, 6, 0.02
    phi = O.fill_phi(O.KERNEL_SC100, [height], 3, nx, ny, normalize=True)
    rng = np.random.default_rng(3)
    wall = 0.3 * rng.random((nx, ny))
    u, p = O.hardwall_contact(phi, wall, p0)
    gap = wall - u[2]
    assert p.min() >= 0 and gap.min() > -1e-10
    assert np.abs(p * gap).max() < 1e-10
    assert 0 < (p > 0).sum() < nx * ny  # partial contact
    f, _, _ = O.static_force(phi, u)
    f[2] += p0
    # elastic + load force equals the wall pressure on the surface layer
    assert np.abs(f[2] - p).max() < 1e-10
    assert np.abs(f[:2]).max() < 1e-10
    phi0 = O.kernel_phi(O.KERNEL_SC100, [height], 0.0, 0.0).real
    assert abs(p.sum() - (p0 * nx * ny - phi0[2, 2] * u[2].sum())) < 1e-9


# --- learned warm start: FNO oracle -------------------------------------
import fno as F  # noqa: E402


def test_fno_spectral_matches_direct_dft():
    """The spectral convolution equals the brute-force definition: a direct
    DFT over the grid, mode-wise mixing on the retained block (and the
    conjugate partners the real field implies), direct inverse."""
    rng = np.random.default_rng(4)
    B, nx, ny, C, M = 2, 6, 8, 3, 2
    v = rng.standard_normal((B, nx, ny, C))
    R = rng.standard_normal((2 * M, M + 1, C, C)) + \
        1j * rng.standard_normal((2 * M, M + 1, C, C))
    got = F.spectral(v, R, M)
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ref = np.zeros_like(v)
    kxs = F.mode_indices(nx, M)
    for b in range(B):
        for a, kx in enumerate(kxs):
            for ky in range(M + 1):
                ph = np.exp(-2j * np.pi * (kx * X / nx + ky * Y / ny))
                Vk = np.einsum("xy,xyc->c", ph, v[b])
                out = Vk @ R[a, ky]
                # a real inverse transform keeps Re for self-conjugate
                # columns and adds the conjugate partner otherwise
                w = 1.0 if ky == 0 or 2 * ky == ny else 2.0
                ref[b] += w * np.real(out[None, None, :] *
                                      np.conj(ph)[:, :, None]) / (nx * ny)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def test_fno_gradients_match_finite_differences():
    rng = np.random.default_rng(9)
    C, M, L = 4, 2, 3
    p = F.init_params(rng, C=C, M=M, L=L)
    for k in p:
        p[k] = p[k] + 0.1 * (rng.standard_normal(p[k].shape) +
                             (1j * rng.standard_normal(p[k].shape)
                              if np.iscomplexobj(p[k]) else 0))
    a = rng.standard_normal((2, 8, 6, 2))
    t = rng.standard_normal((2, 8, 6, 3))
    loss, g = F.loss_and_grads(p, a, t, M=M, L=L)
    eps = 1e-6
    for name in list(p) + ["input"]:
        x = a if name == "input" else p[name]
        flat = x.reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            for part in ((1.0,) if not np.iscomplexobj(x) else (1.0, 1j)):
                old = flat[idx]
                flat[idx] = old + eps * part
                lp = F.loss_and_grads(p, a, t, M=M, L=L)[0]
                flat[idx] = old - eps * part
                lm = F.loss_and_grads(p, a, t, M=M, L=L)[0]
                flat[idx] = old
                fd = (lp - lm) / (2 * eps)
                gv = g[name].reshape(-1)[idx]
                an = gv.real if part == 1.0 else gv.imag
                assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (name, idx,
                                                                 part, fd, an)


def test_fno_flat_layout_round_trip():
    rng = np.random.default_rng(1)
    p = F.init_params(rng, C=8, M=3, L=2)
    q = F.unflatten(F.flatten(p, C=8, M=3, L=2), C=8, M=3, L=2)
    for k in p:
        assert np.array_equal(p[k], q[k])


def test_li_berger_golden_reference(golden):
    """The oracle's Li & Berger restatement against the reference's
    sq_uniform_PQ (li_berger.cpp:453-459) compiled here."""
    for x, ref in zip(golden["lb_in"], golden["lb_out"]):
        got = O.sq_uniform_pq(*x)
        assert np.abs(got - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-12)


def test_li_berger_far_field_is_boussinesq_cerruti():
    """Far from the loaded cell the displacement of a unit force is the
    point-force solution: u_z = (1 - nu) / (2 pi G r) under a normal force,
    u_x = (1 - nu + nu x^2/r^2) / (2 pi G r) under a tangential one."""
    G, nu, r = 1.3, 0.3, 40.0
    uz = O.sq_uniform_pq(G, nu, 1.0, 0, 0, 0.5, 0.5, r, 0.0)[2]
    assert abs(uz - (1 - nu) / (2 * np.pi * G * r)) < 1e-3 * abs(uz)
    ux = O.sq_uniform_pq(G, nu, 0, 1.0, 0, 0.5, 0.5, r, 0.0)[0]
    assert abs(ux - 1.0 / (2 * np.pi * G * r)) < 1e-3 * abs(ux)
    uy = O.sq_uniform_pq(G, nu, 0, 1.0, 0, 0.5, 0.5, 0.0, r)[0]
    assert abs(uy - (1 - nu) / (2 * np.pi * G * r)) < 1e-3 * abs(uy)


def test_nonperiodic_table_symmetries():
    """The edge rule keeps Phi Hermitian; a real G(r) gives Phi(-q) =
    conj Phi(q) exactly."""
    for nx, ny in [(16, 16), (15, 12)]:
        phi = O.nonperiodic_phi(nx, ny, 0.3, 1.0, normalize=False)
        P = phi.reshape(nx, ny, 3, 3)
        herm = np.abs(P - np.conj(np.swapaxes(P, 2, 3))).max()
        assert herm < 1e-10 * np.abs(P).max()
        Pm = P[(-np.arange(nx)) % nx][:, (-np.arange(ny)) % ny]
        assert np.abs(P - np.conj(Pm)).max() < 1e-12 * np.abs(P).max()


def test_nonperiodic_open_boundary_brute_force():
    """A 16 x 16 patch of random point forces t in a 32 x 32 grid (padding
    free): the displacement summed directly over the patch from Li & Berger's
    square-load solution (no images: all patch offsets are below half the
    grid) is what the nonperiodic table turns back into -t.  The table
    inverts u(r) = sum_s G(s - r) t(s), the transpose of the direct
    convolution by reciprocity G_ab(r) = G_ba(-r) (the reference transforms
    G with exponent +1, nonperiodic_stiffness.cpp:132-134, 203-205, while
    forces use exp(-i q r)): normal responses are those of the direct sum,
    the normal-tangential coupling has the opposite sign."""
    n, nu, mu = 32, 0.3, 1.0
    phi = O.nonperiodic_phi(n, n, nu, mu)
    rng = np.random.default_rng(3)
    t = np.zeros((3, n, n))
    t[:, 8:24, 8:24] = rng.standard_normal((3, 16, 16))
    # direct open-boundary sums from the square-load solution
    cache = {}

    def g(dx, dy):  # displacement at offset (dx, dy) from unit loads [b][a]
        if (dx, dy) not in cache:
            cache[(dx, dy)] = np.array([
                O.sq_uniform_pq(mu, nu, 0, 1, 0, 0.5, 0.5, dx, dy),
                O.sq_uniform_pq(mu, nu, 0, 0, 1, 0.5, 0.5, dx, dy),
                O.sq_uniform_pq(mu, nu, 1, 0, 0, 0.5, 0.5, dx, dy)]).T
        return cache[(dx, dy)]

    src = [(i, j) for i in range(8, 24) for j in range(8, 24)]
    u_dir = np.zeros((3, n, n))
    u_mir = np.zeros((3, n, n))
    for i, j in src:
        for sx, sy in src:
            u_dir[:, i, j] += g(i - sx, j - sy) @ t[:, sx, sy]
            u_mir[:, i, j] += g(sx - i, sy - j) @ t[:, sx, sy]
    # the full-grid field the table acts on: minimum-image sum everywhere
    G = O.nonperiodic_green(n, n, nu, mu)
    Gm = G[:, :, (-np.arange(n)) % n][:, :, :, (-np.arange(n)) % n]
    u = np.zeros((3, n, n))
    for sx, sy in src:
        u += np.einsum("abxy,b->axy",
                       np.roll(np.roll(Gm, sx, axis=2), sy, axis=3), t[:, sx, sy])
    p = (slice(None), slice(8, 24), slice(8, 24))
    assert np.abs(u[p] - u_mir[p]).max() < 1e-13 * np.abs(u_mir).max()
    f, _, _ = O.static_force(phi, u)
    assert np.abs(f + t).max() < 1e-12 * np.abs(t).max()
    # normal response of a normal load is the direct sum's; the lateral
    # response of a normal load is reversed
    tz = t.copy()
    tz[:2] = 0.0
    uz_dir = np.zeros((3, n, n))
    uz_mir = np.zeros((3, n, n))
    for i, j in src[::37]:
        for sx, sy in src:
            uz_dir[:, i, j] += g(i - sx, j - sy) @ tz[:, sx, sy]
            uz_mir[:, i, j] += g(sx - i, sy - j) @ tz[:, sx, sy]
    assert np.abs(uz_dir[2] - uz_mir[2]).max() < 1e-13 * np.abs(uz_dir[2]).max()
    assert np.abs(uz_dir[:2] + uz_mir[:2]).max() < 1e-13 * np.abs(uz_dir[:2]).max()


def test_threaded_timing_baseline_is_the_same_evaluation():
    """bench.py's all-cores CPU baseline (orc_static_force_mt) computes the
    oracle's evaluation: forces bit for bit, energy to the reduction order."""
    import ctypes
    L = O.lib()
    L.orc_static_force_mt.restype = ctypes.c_double
    L.orc_static_force_mt.argtypes = [ctypes.c_int] * 3 + \
        [ctypes.c_void_p] * 3 + [ctypes.c_int]
    nx, ny = 48, 30
    phi = O.fill_phi(O.KERNEL_SC100, [6], 3, nx, ny, normalize=True)
    u = np.random.default_rng(9).normal(size=(3, nx, ny))
    f1, e1, _ = O.static_force(phi, u)
    for nt in (1, 4):
        f2 = np.zeros_like(u)
        e2 = L.orc_static_force_mt(nx, ny, 3, phi.ctypes.data, u.ctypes.data,
                                   f2.ctypes.data, nt)
        assert np.array_equal(f1, f2)
        assert abs(e1 - e2) <= 1e-13 * abs(e1)
