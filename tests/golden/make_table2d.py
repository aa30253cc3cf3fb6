"""Writes tests/golden/table2d_reference.npz: golden vectors of the
reference's Table2D (the periodic bicubic spline behind fix wall/map's map
files, src/mathutils/table2d.cpp), compiled from the reference sources in
place by oracle/Makefile.ref into oracle/_ref/libgfmd_ref.so.

Run in the build container (where the reference tree exists):
    make -f oracle/Makefile.ref && python tests/golden/make_table2d.py

The fixture holds inputs (height tables, evaluation points) and the
reference's outputs (value and first and second derivatives) only.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

rng = np.random.default_rng(20261203)
out = {}
shapes = [(7, 5), (16, 16), (3, 9), (32, 24)]
for c, (nx, ny) in enumerate(shapes):
    v = rng.normal(size=(nx, ny))
    n = 400
    x = rng.uniform(-2 * nx, 3 * nx, n)
    y = rng.uniform(-2 * ny, 3 * ny, n)
    x[:40] = np.round(x[:40])  # grid lines and nodes
    y[20:60] = np.round(y[20:60])
    out["t%d_values" % c] = v
    out["t%d_x" % c] = x
    out["t%d_y" % c] = y
    out["t%d_ref" % c] = oracle.RefTable2D(v).eval(x, y)
    lm = oracle.RefTable2D(v, lowmem=True).eval(x, y)
    assert np.array_equal(lm, out["t%d_ref" % c])
out["count"] = np.array(len(shapes))
np.savez_compressed(os.path.join(ROOT, "tests", "golden",
                                 "table2d_reference.npz"), **out)
print("wrote table2d_reference.npz")
