"""Parity of the HIP elastic-force path with the CPU FP64 oracle, through
the C ABI (gfmd_hip binding of libgfmd_hip.so).

Tolerances (written here, stated in the design notes):
  FP64 path : forces and energy within 1e-12 relative (max-norm of f
              relative to max|f|; |dE|/|E|).
  FP32 path : forces within 2e-6 relative (max-norm) and energy within
              1e-6 relative at these sizes (FP32 FFT + Phi.u, FP64 energy
              accumulation).
"""
import numpy as np
import pytest

import gfmd_hip
import oracle as O

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(7)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def oracle_phi(nx, ny, height):
    return O.fill_phi(O.KERNEL_SC100, [height], 3, nx, ny, normalize=True)


def phi_packed_from_full(full, nx, ny, d):
    """Device-layout packing of a full-grid table, done in numpy."""
    nky = ny // 2 + 1
    out = np.zeros((nky, d * d, nx))
    F = full.reshape(nx, ny, d, d)
    for ky in range(nky):
        M = F[:, ky]  # (nx, d, d)
        for i in range(d):
            out[ky, i] = M[:, i, i].real
        p = d
        for i in range(d):
            for j in range(i + 1, d):
                h = 0.5 * (M[:, i, j] + M[:, j, i].conj())
                out[ky, p], out[ky, p + 1] = h.real, h.imag
                p += 2
    return out


@pytest.mark.parametrize("nx,ny,height", [(64, 64, 0), (128, 128, 128),
                                          (256, 64, 7), (64, 37, 16),
                                          (63, 63, -1)])
def test_phi_tabulation_matches_oracle(nx, ny, height):
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height %d" % height)
    got = s.dump_stiffness()
    ref = phi_packed_from_full(oracle_phi(nx, ny, height), nx, ny, 3)
    assert rel(got, ref) < 1e-12


@pytest.mark.parametrize("nx,ny", [(64, 64), (128, 128), (256, 128),
                                   (128, 512), (512, 256), (64, 37), (63, 63),
                                   (16, 16)])
def test_force_fp64_matches_oracle(nx, ny):
    height = 32
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height %d" % height)
    phi = oracle_phi(nx, ny, height)
    u = 0.01 * RNG.standard_normal((3, nx, ny))
    f_ref, e_ref, u0_ref = O.static_force(phi, u)
    f, e = s.post_force(u)
    print("fp64", nx, ny, s.info(), rel(f, f_ref), abs(e - e_ref) / e_ref)
    assert rel(f, f_ref) < 1e-12
    assert abs(e - e_ref) < 1e-12 * abs(e_ref)
    assert np.allclose(s.get_u0(), u0_ref, rtol=1e-12, atol=1e-14)


def test_force_linf_gamma_terms():
    nx, ny = 128, 64
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 16")
    linf = np.array([0.125])
    s.set_linf(linf)
    phi = oracle_phi(nx, ny, 16)
    u = 0.05 * RNG.standard_normal((3, nx, ny))
    f_ref, e_ref, _ = O.static_force(phi, u, linf)
    f, e = s.post_force(u)
    assert rel(f, f_ref) < 1e-12
    assert abs(e - e_ref) < 1e-12 * abs(e_ref)


def test_set_phi_host_table_matches_set_kernel():
    nx, ny = 128, 64
    a = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    a.set_kernel("sc100", "height 9")
    b = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    b.set_phi(oracle_phi(nx, ny, 9), normalized=True)
    u = RNG.standard_normal((3, nx, ny))
    fa, ea = a.post_force(u)
    fb, eb = b.post_force(u)
    assert rel(fa, fb) < 1e-13 and abs(ea - eb) < 1e-13 * abs(eb)


def test_set_phi_table_without_pair_symmetry_matches_oracle():
    """A Hermitian table that breaks Phi(-q) = conj Phi(q) at one point is
    accepted (fill_phi_buffer checks only Hermiticity), and the forces are
    the reference's: complex inverse transform, real part."""
    nx, ny = 16, 12
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    phi = oracle_phi(nx, ny, 4)
    phi[5] *= 1.01
    s.set_phi(phi)
    u = RNG.standard_normal((3, nx, ny))
    f_ref, e_ref, _ = O.static_force(phi, u)
    f, e = s.post_force(u)
    assert rel(f, f_ref) < 1e-12
    assert abs(e - e_ref) < 1e-12 * abs(e_ref)


def test_set_phi_rejects_non_hermitian_matrix():
    nx, ny = 16, 16
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    phi = oracle_phi(nx, ny, 4)
    phi[37, 0, 1] += 1e-3 * abs(np.trace(phi[37]))  # Phi_01 != conj Phi_10
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        s.set_phi(phi)
    assert ei.value.code == gfmd_hip.ERR_ARG
    assert "not Hermitian" in str(ei.value)


@pytest.mark.parametrize("nx,ny", [(128, 128), (1024, 256), (64, 37)])
def test_force_fp32_bound(nx, ny):
    height = 64
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP32)
    s.set_kernel("sc100", "height %d" % height)
    phi = oracle_phi(nx, ny, height)
    u = 0.01 * RNG.standard_normal((3, nx, ny))
    f_ref, e_ref, _ = O.static_force(phi, u)
    f, e = s.post_force(u)
    print("fp32", nx, ny, rel(f, f_ref), abs(e - e_ref) / e_ref)
    assert rel(f, f_ref) < 2e-6
    assert abs(e - e_ref) < 1e-6 * abs(e_ref)


def test_force_2048_properties_fp64():
    """Full headline size: size-independent properties (plane wave closed
    form through the device Phi, E = -1/2 u.f, linearity)."""
    nx = ny = 2048
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 2048")
    assert s.info()["fast_path"] == 1
    X, Y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    i0, j0 = 3, 700
    qx, qy = 2 * np.pi * i0 / nx, 2 * np.pi * j0 / ny
    A = np.array([0.02 - 0.01j, 0.015 + 0.0j, -0.03 + 0.02j])
    ph = np.exp(1j * (qx * X + qy * Y))
    u = np.real(A[:, None, None] * ph[None])
    f, e = s.post_force(u)
    P = O.kernel_phi(O.KERNEL_SC100, [2048], qx, qy)
    fref = -np.real((P @ A)[:, None, None] * ph[None])
    assert rel(f, fref) < 1e-11
    assert abs(e - nx * ny / 4 * np.real(A.conj() @ P @ A)) < 1e-11 * abs(e)
    v = 0.01 * RNG.standard_normal((3, nx, ny))
    fv, ev = s.post_force(v)
    assert abs(ev + 0.5 * np.sum(v * fv)) < 1e-11 * ev
    fw, _ = s.post_force(u - 2 * v)
    assert rel(fw, f - 2 * fv) < 1e-11


def test_force_4096_every_row_variant_fp64():
    """4096^2 FP64, every row variant (16 elements per thread at this
    length): a plane wave through the device Phi against the closed form
    (1e-11), and the variants' forces on a random field against each other
    (1e-13)."""
    n = 4096
    s = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height %d" % n)
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    i0, j0 = 1301, 1777
    qx, qy = 2 * np.pi * i0 / n, 2 * np.pi * j0 / n
    A = np.array([0.01 + 0.02j, -0.02 + 0.0j, 0.03 - 0.01j])
    ph = np.exp(1j * (qx * X + qy * Y))
    w = np.real(A[:, None, None] * ph[None])
    del X, Y
    P = O.kernel_phi(O.KERNEL_SC100, [n], qx, qy)
    fw_ref = -np.real((P @ A)[:, None, None] * ph[None])
    del ph
    v = 0.01 * RNG.standard_normal((3, n, n))
    fv0 = None
    for rv in range(3):
        s.set_variant(rv, 1)
        fw, _ = s.post_force(w)
        assert rel(fw, fw_ref) < 1e-11, rv
        fv, ev = s.post_force(v)
        if fv0 is None:
            fv0, ev0 = fv, ev
        else:
            assert rel(fv, fv0) < 1e-13, rv
            assert abs(ev - ev0) < 1e-13 * abs(ev0), rv


def test_device_api_matches_host_api():
    import torch
    nx, ny = 256, 256
    for prec, dt in [(gfmd_hip.FP64, torch.float64), (gfmd_hip.FP32,
                                                     torch.float32)]:
        s = gfmd_hip.Solver(nx, ny, 3, prec)
        s.set_kernel("sc100", "height 64")
        u = 0.01 * RNG.standard_normal((3, nx, ny))
        f_host, e_host = s.post_force(u)
        ud = torch.tensor(u, dtype=dt, device="cuda")
        fd = torch.empty_like(ud)
        ed = torch.zeros(1, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        s.post_force_device(ud.data_ptr(), fd.data_ptr(), ed.data_ptr())
        s.synchronize()
        tol = 1e-13 if prec == gfmd_hip.FP64 else 1e-5
        assert rel(fd.cpu().double().numpy(), f_host) < tol
        assert abs(ed.item() - e_host) < tol * abs(e_host)


def test_error_paths():
    s = gfmd_hip.Solver(64, 64, 3)
    u = np.zeros((3, 64, 64))
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        s.post_force(u)
    assert ei.value.code == gfmd_hip.ERR_STATE
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        s.set_kernel("fcc999", "")
    assert ei.value.code == gfmd_hip.ERR_UNSUPPORTED
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        s.set_kernel("sc100", "height x")
    assert ei.value.code == gfmd_hip.ERR_ARG
    s.set_kernel("sc100", "")
    u[1, 3, 4] = np.nan
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        s.post_force(u)
    assert ei.value.code == gfmd_hip.ERR_NONFINITE
    with pytest.raises(gfmd_hip.GFMDError) as ei:
        gfmd_hip.Solver(64, 64, 6).set_kernel("sc100", "")
    assert ei.value.code == gfmd_hip.ERR_ARG


@pytest.mark.parametrize("prec", [gfmd_hip.FP64, gfmd_hip.FP32])
def test_every_fast_variant_matches_oracle(prec):
    nx, ny, height = 512, 256, 24
    phi = oracle_phi(nx, ny, height)
    u = 0.01 * RNG.standard_normal((3, nx, ny))
    f_ref, e_ref, _ = O.static_force(phi, u, np.array([0.01]))
    tol_f, tol_e = (1e-12, 1e-12) if prec == gfmd_hip.FP64 else (2e-6, 1e-6)
    s = gfmd_hip.Solver(nx, ny, 3, prec)
    s.set_kernel("sc100", "height %d" % height)
    s.set_linf(np.array([0.01]))
    n = 0
    for rv in range(3):
        for cv in range(4):
            s.set_variant(rv, cv)
            f, e = s.post_force(u)
            assert rel(f, f_ref) < tol_f, (rv, cv)
            assert abs(e - e_ref) < tol_e * abs(e_ref), (rv, cv)
            n += 1
    assert n == 12
    with pytest.raises(gfmd_hip.GFMDError):
        s.set_variant(3, 0)
    with pytest.raises(gfmd_hip.GFMDError):
        s.set_variant(0, 4)


def test_autotune_2048_variants_agree():
    nx = ny = 2048
    u = 0.01 * RNG.standard_normal((3, nx, ny))
    s64 = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP64)
    s64.set_kernel("sc100", "height 2048")
    f64, e64 = s64.post_force(u)
    s = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP32)
    s.set_kernel("sc100", "height 2048")
    rv, cv, ms = s.autotune(5)
    assert 0 <= rv < 3 and 0 <= cv < 4 and ms > 0
    for r in range(3):
        for c in range(4):
            s.set_variant(r, c)
            f, e = s.post_force(u)
            assert rel(f, f64) < 5e-6, (r, c)
            assert abs(e - e64) < 2e-6 * abs(e64)
    # a new handle of the same shape starts from the tuned variant
    t = gfmd_hip.Solver(nx, ny, 3, gfmd_hip.FP32)
    t.set_kernel("sc100", "height 2048")
    t.set_variant(rv, cv)


def test_phi_mirror_detection_and_parity():
    """The x-mirror check: sc100 and isotropic tables are symmetric (the
    column sweep reads half the table), the two-layer fcc100 cell with its
    layer shift and a random Hermitian table are not; forces match the
    oracle either way (FP64, 1e-12)."""
    n = 64
    rng = np.random.default_rng(17)
    s = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 16")
    assert s.phi_mirror()
    s32 = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP32)
    s32.set_kernel("sc100", "height 16")
    assert s32.phi_mirror()
    iso = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    iso.set_kernel("isotropic", "0.3 1.0 1.0")
    assert iso.phi_mirror()
    two = gfmd_hip.Solver(n, n, 6, gfmd_hip.FP64)
    two.set_kernel("fcc100", "2 1.0 0 0 height 8")
    assert not two.phi_mirror()
    # a random Hermitian table with Phi(-q) = conj Phi(q) has no mirror
    A = rng.standard_normal((n, n, 3, 3)) + 1j * rng.standard_normal((n, n, 3, 3))
    H = A @ A.conj().transpose(0, 1, 3, 2) + 3 * np.eye(3)
    P = H[(-np.arange(n)) % n][:, (-np.arange(n)) % n]
    H = 0.5 * (H + P.conj())
    r = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    r.set_phi(H.reshape(n * n, 3, 3) / (n * n))
    assert not r.phi_mirror()
    u = 0.01 * rng.standard_normal((3, n, n))
    for h, phi in ((s, O.fill_phi(O.KERNEL_SC100, [16], 3, n, n)),
                   (r, H.reshape(n * n, 3, 3) / (n * n))):
        f_ref, e_ref, _ = O.static_force(phi, u)
        f, e = h.post_force(u)
        assert np.abs(f - f_ref).max() < 1e-12 * np.abs(f_ref).max()
        assert abs(e - e_ref) < 1e-12 * abs(e_ref)


def test_compact_table_detection_and_parity():
    """ndof-3 tables with Phi_xy real and Phi_xz, Phi_yz imaginary at every
    wavevector (sc100, isotropic) are read as 6 reals per q by the force
    sweep; a host table that breaks the structure, the open-boundary table
    and ndof-6 tables keep the packed form.  Forces match the oracle either
    way (FP64, 1e-12)."""
    n = 64
    rng = np.random.default_rng(23)
    s = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    s.set_kernel("sc100", "height 16")
    assert s.phi_entries() == 6
    iso = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP32)
    iso.set_kernel("isotropic", "0.3 1.0 1.0")
    assert iso.phi_entries() == 6
    npd = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    npd.set_kernel("nonperiodic", "0.3 1.0")
    assert npd.phi_entries() == 9
    two = gfmd_hip.Solver(n, n, 6, gfmd_hip.FP64)
    two.set_kernel("fcc100", "2 1.0 0 0 height 8")
    assert two.phi_entries() == 36
    phi = O.fill_phi(O.KERNEL_SC100, [16], 3, n, n)
    h = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    h.set_phi(phi)
    assert h.phi_entries() == 6
    bent = phi.copy()
    bent[:, 0, 2] += 1e-3 * np.abs(bent[:, 0, 2]).max()   # Re Phi_xz != 0
    bent[:, 2, 0] = np.conj(bent[:, 0, 2])
    b = gfmd_hip.Solver(n, n, 3, gfmd_hip.FP64)
    b.set_phi(bent)
    assert b.phi_entries() == 9
    u = 0.01 * rng.standard_normal((3, n, n))
    for hd, table in ((s, phi), (h, phi), (b, bent)):
        f_ref, e_ref, _ = O.static_force(table, u)
        f, e = hd.post_force(u)
        assert rel(f, f_ref) < 1e-12
        assert abs(e - e_ref) < 1e-12 * abs(e_ref)


@pytest.mark.parametrize("prec", [gfmd_hip.FP64, gfmd_hip.FP32])
def test_every_variant_packed_and_six_component_tables(prec):
    """Every compiled row x column variant on the tables outside the
    compact, x-mirrored ndof-3 case: an ndof-3 host table with a constant
    added to Re Phi_xz (Hermitian and real in real space still, but neither
    of the compact form nor x-mirror symmetric: the packed 9-real table read
    in full), and the two-layer fcc(100) table (ndof 6, layer shifts: not
    mirror symmetric), both against the oracle."""
    tol_f, tol_e = (1e-12, 1e-12) if prec == gfmd_hip.FP64 else (2e-6, 1e-6)
    nx, ny = 256, 128
    phi = oracle_phi(nx, ny, 24)
    phi[:, 0, 2] += 1e-3 * np.abs(phi[:, 0, 2]).max()
    phi[:, 2, 0] = np.conj(phi[:, 0, 2])
    prm = O.fcc100_params(8, 2, [1.0, 0.3, 0.2], [0.1, 0.05, 0.02])
    cases = [(3, phi, None), (6, O.fill_phi(O.KERNEL_FCC100, prm, 6, nx, ny,
                                            normalize=True),
                              "x 2 1.0 0.1 0.3 0.05 0.2 0.02 height 8")]
    for d, table, args in cases:
        u = 0.01 * RNG.standard_normal((d, nx, ny))
        f_ref, e_ref, _ = O.static_force(table, u)
        s = gfmd_hip.Solver(nx, ny, d, prec)
        if args is None:
            s.set_phi(table)
        else:
            s.set_kernel("fcc100", args)
        assert s.phi_entries() == d * d and not s.phi_mirror()
        for rv in range(3):
            for cv in range(4):
                s.set_variant(rv, cv)
                f, e = s.post_force(u)
                assert rel(f, f_ref) < tol_f, (d, rv, cv)
                assert abs(e - e_ref) < tol_e * abs(e_ref), (d, rv, cv)
        s.close()


@pytest.mark.parametrize("n,prec", [(256, gfmd_hip.FP64), (128, gfmd_hip.FP32)])
def test_graph_replays_equal_eager_evaluations(n, prec):
    """gfmd_time_force_graph: a captured graph of two evaluations, replayed
    several times, with eager evaluations of another field in between; the
    output after every timing call equals the eager evaluation bit for
    bit."""
    import torch
    dt = torch.float64 if prec == gfmd_hip.FP64 else torch.float32
    s = gfmd_hip.Solver(n, n, 3, prec)
    s.set_kernel("sc100", "height %d" % n)
    u = torch.tensor(0.01 * RNG.standard_normal((3, n, n)), device="cuda").to(dt)
    w = torch.tensor(0.01 * RNG.standard_normal((3, n, n)), device="cuda").to(dt)
    f_ref = torch.empty_like(u)
    e_ref = torch.zeros(1, dtype=torch.float64, device="cuda")
    s.post_force_device(u.data_ptr(), f_ref.data_ptr(), e_ref.data_ptr())
    s.synchronize()
    for rep in range(3):
        f = torch.full_like(u, float("nan"))
        e = torch.zeros(1, dtype=torch.float64, device="cuda")
        me, mg = s.time_force_graph(u.data_ptr(), f.data_ptr(), e.data_ptr(),
                                    6, 2)
        assert me > 0 and mg > 0
        assert torch.equal(f, f_ref), rep
        assert e.item() == e_ref.item(), rep
        # eager work on the same handle between the timing calls
        g = torch.empty_like(w)
        s.post_force_device(w.data_ptr(), g.data_ptr(), e.data_ptr())
        s.synchronize()
    with pytest.raises(gfmd_hip.GFMDError):
        s.time_force_graph(u.data_ptr(), f.data_ptr(), e.data_ptr(), 5, 2)
