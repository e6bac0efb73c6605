"""The HIP solver (FP64, through the C ABI) against the reference's own output.

The reference ships three residual histories it printed while solving the
RAE 2822 case on its 3558-cell mesh (unresidual_*.plt, Resd every 10 or 100
steps to 10 decimals), at M = 0.729, alpha = 2.31 deg, CFL = 1.0.  Every
kernel set must reproduce every printed digit over the ranges the oracle
does: no smoothing to step 500, smoothing with local dt to step 160 (after
which the reference's contracted multiply-adds make the chaotic runs
separate) and smoothing with the global time step to step 2000.
"""
import json
import os

import numpy as np
import pytest

from gpueularsolver2d_amd import Solver, read_dat

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rae():
    return read_dat(os.path.join(GOLDEN, "2822-3k.dat.gz"))


@pytest.fixture(scope="module")
def ref():
    with open(os.path.join(GOLDEN, "reference_residuals.json")) as f:
        return {k: np.array(v) for k, v in json.load(f).items()}


def printed_digits_match(hist, table, upto):
    rows = table[table[:, 0] <= upto]
    steps = rows[:, 0].astype(int)
    bad = np.abs(np.round(hist[steps], 10) - rows[:, 1]) > 0.6e-10
    assert not bad.any(), f"first mismatch at step {steps[bad][0]}: {hist[steps[bad][0]]} vs {rows[bad][0, 1]}"
    return len(rows)


CASES = [("unresidual_nosmooth", dict(smoothing=False), 500, 50),
         ("unresidual_smooth", dict(smoothing=True), 160, 16),
         ("unresidual_time2", dict(smoothing=True, global_dt=True), 2000, 20)]


@pytest.mark.parametrize("kernels", ["tiled", "split", "plain"])
@pytest.mark.parametrize("name,opts,steps,points", CASES, ids=[c[0] for c in CASES])
def test_reference_printed_history(rae, ref, kernels, name, opts, steps, points):
    s = Solver(rae, 0.729, 2.31, cfl=1.0, order=2, precision="fp64", kernels=kernels, **opts)
    hist, n, st = s.run(steps)
    assert st == 0 and n == steps
    assert printed_digits_match(hist, ref[name], steps) == points
