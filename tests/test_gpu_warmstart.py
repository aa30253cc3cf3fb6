"""Learned warm start end to end on the device, at a small scale: training
data from the ensemble solver, training with the FNO kernels, and warm-
started contact solves that must reach the same solution as cold ones.
The gains (iterations, time) are printed; the assertion on the network is
that training lowers the prediction error it is judged by."""
import numpy as np
import pytest

import gfmd_hip
from gfmd_hip import warmstart as W

pytestmark = pytest.mark.gpu


def test_warm_start_trains_and_keeps_the_solution():
    n = 128
    ws = W.WarmStart(n, max_batch=16)
    a, u, dev = ws.generate(list(range(1, 257)))
    assert np.isfinite(u).all()
    ws.set_scale(u)
    held = list(range(1000000, 1000004))
    before = ws.evaluate(held[:2])
    losses = ws.train(a, u, 8, lr=1e-3)
    after = ws.evaluate(held)
    for r in after:
        print("warm start 128^2:", r, flush=True)
        assert r["converged"]
        assert r["warm_vs_cold_rel_diff"] < 1e-6
    e0 = np.mean([r["prediction_rel_error"] for r in before])
    e1 = np.mean([r["prediction_rel_error"] for r in after])
    print("prediction error untrained %.3g trained %.3g; loss %.3g -> %.3g"
          % (e0, e1, losses[0], losses[-1]), flush=True)
    assert e1 < 0.5 * e0
    assert losses[-1] < losses[0]


def test_warm_start_at_4096_keeps_the_solution():
    """The warm-started path at 4096^2 (network inference on a 4096^2
    surface, starting pressures from its prediction, constrained CG to
    1e-8): the solve converges to the cold solve's solution, and a start
    from the pressures read off the converged displacement (the contact
    sites it puts at the wall) needs a small fraction of the cold
    iterations.  The network here is the untrained one (its weights do not
    change the code path)."""
    n = 4096
    ws = W.WarmStart(n, max_batch=1)
    h, wall, p0 = W.problem(n, 3000000)
    ws.scale = np.full(W.COUT, float(np.sqrt(np.mean(h * h))),
                       dtype=np.float32)
    rows = ws.evaluate([3000000], tol=1e-8)
    r = rows[0]
    print("warm start 4096^2:", r, flush=True)
    assert r["converged"]
    # both solves end at relative residual 1e-8; at 4096^2 the compliance
    # of the longest waves turns that into up to ~1e-6 of the displacement
    assert r["warm_vs_cold_rel_diff"] < 1e-5
    assert r["iters_exact_start"] < r["iters_cold"] / 4
    assert np.isfinite(r["prediction_rel_error"])
    for hd in (ws.s64, ws.s32, ws.net):
        hd.close()
