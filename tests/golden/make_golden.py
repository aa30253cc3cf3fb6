"""Writes tests/golden/phi_reference.npz: golden stiffness-matrix vectors
computed by the reference's own C++ core (StiffnessKernel,
stiffness_kernel_factory, fill_phi_buffer), compiled from the reference
sources in place by oracle/Makefile.ref into oracle/_ref/libgfmd_ref.so.

Run in the build container (where the reference tree exists):
    make -f oracle/Makefile.ref && python tests/golden/make_golden.py

The fixtures hold inputs (kernel, arguments, wavevectors, grid) and outputs
(complex Phi matrices) only.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

rng = np.random.default_rng(20261202)

# wavevectors: special points (Gamma, zone edges, corners) + random ones
special = np.array([[0.0, 0.0], [np.pi, 0.0], [0.0, np.pi], [np.pi, np.pi],
                    [-np.pi, 0.3], [1e-3, 0.0], [0.0, -1e-3], [0.5, 0.5],
                    [-0.7, 1.9]])
rand = rng.uniform(-np.pi, np.pi, size=(23, 2))
Q = np.vstack([special, rand])

out = {"q": Q}
cases = [("sc100", ""), ("sc100", "height 1"), ("sc100", "height 2"),
         ("sc100", "height 7"), ("sc100", "height 128"),
         ("sc100", "height 2048"), ("sc100", "height -1")]
names = []
for i, (name, args) in enumerate(cases):
    K = oracle.RefKernel(name, args)
    qx, qy = Q[:, 0].copy(), Q[:, 1].copy()
    if args == "height -1":
        # the semi-infinite iteration cannot converge at q = 0
        keep = np.hypot(qx, qy) > 1e-2
        qx, qy = qx[keep], qy[keep]
    out["case%d_qx" % i] = qx
    out["case%d_qy" % i] = qy
    out["case%d_phi" % i] = K.phi(qx, qy)
    out["case%d_height" % i] = np.array(K.height)
    names.append("%s|%s" % (name, args))
out["case_names"] = np.array(names)

# fcc100: one and two layers, nearest-neighbour LJ-derived springs (k =
# V''(2^(1/6)) = 57.146 for eps = sigma = 1), three shells with kappa,
# normal -1, semi-infinite
fcc = [("fcc100", "1 1.0 height 16"),
       ("fcc100", "2 57.146 0 0 height 32"),
       ("fcc100", "x 2 1.0 0.1 0.3 0.05 0.2 0.02 height 8"),
       ("fcc100", "2 1.0 0.5 0.25 height 12 normal -1"),
       ("fcc100", "x 1 1.0 0.2 height 4"),
       ("fcc100", "2 1.0 0.0 0.0 height -1"),
       ("fcc100", "x 3 1.0 0.1 0.4 0.05 0.2 0.03 height 6"),
       ("fcc100", "4 1.0 0.3 0.1 height 5 normal -1")]
fnames = []
for i, (name, args) in enumerate(fcc):
    K = oracle.RefKernel(name, args)
    qx, qy = Q[:, 0].copy(), Q[:, 1].copy()
    if "height -1" in args:
        keep = np.hypot(qx, qy) > 1e-2
        qx, qy = qx[keep], qy[keep]
    out["fcc%d_qx" % i] = qx
    out["fcc%d_qy" % i] = qy
    out["fcc%d_phi" % i] = K.phi(qx, qy)
    out["fcc%d_dim" % i] = np.array(K.dim)
    fnames.append("%s|%s" % (name, args))
out["fcc_names"] = np.array(fnames)

# isotropic continuum: (arguments, square domain length) -- q = 0 uses
# xprd = domain length
iso = [("0.3 1.0 1.0", 32.0), ("0.5 2.0 0.7", 64.0), ("0.0 1.0 1.0", 16.0),
       ("0.25 0.8 3.0", 128.0)]
for i, (args, L) in enumerate(iso):
    K = oracle.RefKernel("isotropic", args, L, L)
    out["iso%d_phi" % i] = K.phi(Q[:, 0].copy(), Q[:, 1].copy())
    out["iso%d_L" % i] = np.array(L)
out["iso_args"] = np.array([a for a, _ in iso])

# Li & Berger uniform-square Green's function (li_berger.cpp:453-459), the
# sample the `nonperiodic` kernel takes at every grid offset: loads
# (P, Qx, Qy), half sides (a, b), moduli (G, nu), field points on the grid
# (including the square's diagonal) and off it
import ctypes  # noqa: E402
ref = oracle.ref_lib()
ref.ref_sq_uniform_pq.argtypes = [ctypes.c_double] * 9 + [ctypes.c_void_p]
lb_in, lb_out = [], []
for t in range(400):
    G = [1.0, 2.5][t % 2]
    nu = [0.3, 0.0, 0.45, 0.5][t % 4]
    P, Qx, Qy = rng.standard_normal(3)
    a, b = [(0.5, 0.5), (0.5, 0.75), (1.0, 0.4)][t % 3]
    if t < 250:
        bx, by = rng.integers(-12, 13, 2) * np.array([2 * a, 2 * b])
    else:
        bx, by = rng.uniform(-6, 6, 2)
    u = np.zeros(3)
    ref.ref_sq_uniform_pq(G, nu, P, Qx, Qy, a, b, bx, by, u.ctypes.data)
    lb_in.append([G, nu, P, Qx, Qy, a, b, bx, by])
    lb_out.append(u)
out["lb_in"] = np.array(lb_in)
out["lb_out"] = np.array(lb_out)

# a full fill_phi_buffer table (normalised) on an odd x even grid
K = oracle.RefKernel("sc100", "height 16")
out["fill_nx"] = np.array(6)
out["fill_ny"] = np.array(5)
out["fill_phi"] = K.fill_phi(6, 5, normalize=True)
out["fill_height"] = np.array(16)

np.savez_compressed(os.path.join(ROOT, "tests", "golden", "phi_reference.npz"),
                    **out)
print("wrote", len(cases), "cases")


# ---------------------------------------------------------------------------
# `ft` kernel (tests/golden/ft_reference.npz).  FTStiffnessKernel cannot be
# instantiated from the reference's -DPYTHON build, so the fixtures hold the
# pieces its crystal surface and pair-potential force constants produce
# (cell, positions, neighbour list, force-constant blocks), plus Phi(q) of
# the reference's sc100 and fcc100 kernels, which `ft` with the same springs
# must reproduce (the reference's own Hertz tests use `ft sc100 ...` and
# `ft fcc100 ...` for these lattices).
# ---------------------------------------------------------------------------
ft_geom = ["sc100 1.0 1 pair-potential 2 1.0 1.0",
           "sc100 1.3 2 pair-potential 2 1.0 1.0",
           "fcc100 1.0 2 pair-potential 2 1.0 -0.1",
           "fcc100 1.0 3 pair-potential 2x 1.0 0.2 0.5 0.1",
           "fcc111 1 1.0 pair-potential 1 1.0",
           "fcc111 1.0 2 pair-potential 1 1.0",
           "fcc111 1.2 2 pair-potential 3x 1.0 0.1 0.5 0.2 0.3 0.05"]
fo = {"geom_args": np.array(ft_geom)}
for c, a in enumerate(ft_geom):
    R = oracle.RefFT(a)
    nz = np.nonzero(np.abs(R.D0).sum(axis=1))[0]
    assert np.array_equal(R.D0, R.D1)
    fo["g%d_cell" % c] = R.cell
    fo["g%d_positions" % c] = R.positions
    fo["g%d_ints" % c] = R.ints.astype(np.int16)
    fo["g%d_nonzero" % c] = nz.astype(np.int32)
    fo["g%d_D" % c] = R.D0[nz]
    fo["g%d_on0" % c] = R.on0
    fo["g%d_on1" % c] = R.on1
ft_phi = [("sc100 1.0 1 pair-potential 2 1.0 1.0 height 16", "sc100", "height 16"),
          ("sc100 1.3 1 pair-potential 2 1.0 1.0", "sc100", ""),
          ("fcc100 1.0 1 pair-potential 1 1.0 height 8", "fcc100", "1 1.0 height 8"),
          ("fcc100 1.0 2 pair-potential 1 1.0 height 8", "fcc100", "2 1.0 0 0 height 8"),
          ("fcc100 1.0 2 pair-potential 2x 1.0 0.1 0.5 0.05 height 6", "fcc100",
           "x 2 1.0 0.1 0.5 0.05 0 0 height 6")]
fo["phi_args"] = np.array([f for f, _, _ in ft_phi])
fo["phi_ref"] = np.array(["%s|%s" % (n, a) for _, n, a in ft_phi])
for c, (_, n, a) in enumerate(ft_phi):
    fo["p%d_phi" % c] = oracle.RefKernel(n, a).phi(Q[:, 0].copy(), Q[:, 1].copy())
fo["q"] = Q
np.savez_compressed(os.path.join(ROOT, "tests", "golden", "ft_reference.npz"), **fo)
print("wrote ft_reference.npz")
