"""fvm_run_noalloc (the run entry point that allocates nothing: partial sums
and residual history in the library's per-device scratch array) against
fvm_run, through the C ABI: the same states, histories and step counts bit for
bit, on two calls in a row (the second on a smaller mesh, so the scratch
holds the first call's larger contents), for every kernel set; a run that
continues another; and a request beyond the scratch refused with
FVM_ERR_ARGUMENT before anything runs."""
import numpy as np
import pytest

from gpueularsolver2d_amd import FvmError, Solver, naca0012_mesh

pytestmark = pytest.mark.gpu

FLOW = dict(mach=0.8, alpha_deg=1.25, cfl=1.5)
SCRATCH_BYTES = 4 << 20   # FVM_NOALLOC_SCRATCH_BYTES


@pytest.fixture(scope="module")
def meshes():
    return naca0012_mesh(6000, seed=21), naca0012_mesh(3000, seed=22)


def same(a, b, steps, run_a, run_b):
    ha, na, sa = run_a(steps)
    hb, nb, sb = run_b(steps)
    assert (na, sa) == (nb, sb) == (steps, 0)
    np.testing.assert_array_equal(ha, hb)
    assert np.array_equal(a.state(), b.state())
    assert a.forces() == b.forces()


@pytest.mark.parametrize("precision,kernels", [("fp64", "tiled"), ("fp32", "tiled"),
                                               ("fp32", "split"), ("fp64", "plain")])
def test_noalloc_equals_run_on_two_calls(meshes, precision, kernels):
    big, small = meshes
    for mesh, steps in ((big, 60), (small, 25)):   # the second call is smaller
        print(f"noalloc vs run: {precision} {kernels}, {mesh.ncell} cells, {steps} steps", flush=True)
        a = Solver(mesh, precision=precision, kernels=kernels, **FLOW)
        b = Solver(mesh, precision=precision, kernels=kernels, **FLOW)
        same(a, b, steps, a.run_noalloc, b.run)


def test_noalloc_continues_a_run_and_stops_at_epsl(meshes):
    big, _ = meshes
    a = Solver(big, precision="fp64", **FLOW)
    b = Solver(big, precision="fp64", **FLOW)
    a.run(10)
    b.run(10)
    ha, na, _ = a.run_noalloc(37)
    hb, nb, _ = b.run(37)
    assert na == nb == 37
    np.testing.assert_array_equal(ha[11:], hb[11:])
    assert np.array_equal(a.state(), b.state())
    # the stop rule step < max_steps && Resd > epsl, as fvm_run
    c = Solver(big, precision="fp64", **FLOW)
    d = Solver(big, precision="fp64", **FLOW)
    hc, nc, _ = c.run_noalloc(300, epsl=0.3)
    hd, nd, _ = d.run(300, epsl=0.3)
    assert nc == nd < 300
    np.testing.assert_array_equal(hc, hd)
    assert np.array_equal(c.state(), d.state())


def test_noalloc_rejects_runs_beyond_the_scratch(meshes):
    _, small = meshes
    a = Solver(small, precision="fp32", **FLOW)
    too_many = SCRATCH_BYTES // 8   # the history alone would overflow the scratch
    with pytest.raises(FvmError) as e:
        a.run_noalloc(too_many)
    assert e.value.name == "FVM_ERR_ARGUMENT" and "scratch" in str(e.value)
    # nothing ran: the next run starts from step 0 and equals a fresh one
    b = Solver(small, precision="fp32", **FLOW)
    same(a, b, 12, a.run_noalloc, b.run)
