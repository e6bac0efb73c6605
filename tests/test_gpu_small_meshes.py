"""Edge cases of the tiling and the boundaries on small hand-built meshes.

Annuli between a wall circle and a far-field circle, where every ring of
neighbours soon covers the whole mesh and tiles hold down to one cell, and a
two-triangle square whose edges are all far field.  On the square the
freestream is preserved exactly, so step 1 changes nothing and the residual,
normalised by step 1, is +inf from the first step: the reference's loop goes
on (inf > EPSL) and so must the solver, without reporting a divergence.
"""
import numpy as np
import pytest

from gpueularsolver2d_amd import Batch, Solver
from oracle_ctypes import OracleMesh, OracleSolver
from small_meshes import annulus, square2

pytestmark = pytest.mark.gpu

STEPS = 30
FLOW = dict(mach=0.3, alpha_deg=1.25, cfl=0.8)
MESHES = {"annulus8x1": lambda: annulus(8, 1), "annulus24x3": lambda: annulus(24, 3),
          "annulus48x6": lambda: annulus(48, 6), "square2": square2}
SCHEMES = {"o2-smooth": dict(), "o1": dict(order=1, smoothing=False), "global-dt": dict(global_dt=True)}
SHAPES = [("tiled", 0, 0), ("tiled", 1, 64), ("tiled", 3, 64), ("tiled", 16, 64), ("split", 1, 64),
          ("split", 5, 64), ("plain", 0, 0)]


@pytest.fixture(scope="module")
def oracle_runs():
    out = {}
    for mname, make in MESHES.items():
        m = make()
        om = OracleMesh(m.x, m.y, m.tri)
        for sname, kw in SCHEMES.items():
            o = OracleSolver(om, FLOW["mach"], FLOW["alpha_deg"], cfl=FLOW["cfl"], **kw)
            h, _ = o.run(STEPS)
            out[mname, sname] = (m, o.state(), np.asarray(h[:STEPS + 1]), o.forces())
    return out


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("mname", MESHES)
def test_fp64_bitwise_on_small_meshes(oracle_runs, mname, scheme):
    m, Wo, ho, fo = oracle_runs[mname, scheme]
    for kernels, tile, threads in SHAPES:
        s = Solver(m, precision="fp64", kernels=kernels, tile_cells=tile, threads=threads,
                   **SCHEMES[scheme], **FLOW)
        hist, n, st = s.run(STEPS)
        where = f"{mname} {scheme} {kernels} tile {tile} threads {threads}"
        assert st == 0 and n == STEPS, where
        assert np.array_equal(s.state(), Wo), where
        np.testing.assert_allclose(hist[1:], ho[1:], rtol=1e-12, atol=0, err_msg=where)
        cl, cd = s.forces()
        assert abs(cl - fo[0]) <= 1e-12 * max(1.0, abs(fo[0])), where
        assert abs(cd - fo[1]) <= 1e-12 * max(1.0, abs(fo[1])), where


def test_steady_freestream_is_not_a_divergence(oracle_runs):
    _, Wo, ho, _ = oracle_runs["square2", "o2-smooth"]
    assert np.isinf(ho[1:]).all()
    s = Solver(square2(), precision="fp64", **FLOW)
    hist, n, st = s.run(STEPS)
    assert st == 0 and n == STEPS
    assert np.isinf(hist[1:]).all()
    np.testing.assert_array_equal(s.state(), Wo)  # freestream, exactly
    # FP32 fluxes need not cancel to the last bit, so its residual may be
    # finite; the run must still complete and stay at freestream
    s = Solver(square2(), precision="fp32", **FLOW)
    hist, n, st = s.run(STEPS)
    assert st == 0 and n == STEPS
    np.testing.assert_allclose(s.state(), Wo, rtol=1e-6, atol=1e-6)


def test_fp32_small_meshes_close_to_oracle(oracle_runs):
    for mname in ("annulus8x1", "annulus24x3", "annulus48x6"):
        m, Wo, ho, _ = oracle_runs[mname, "o2-smooth"]
        for tile, threads in ((0, 0), (1, 64), (16, 64)):
            s = Solver(m, precision="fp32", tile_cells=tile, threads=threads, **FLOW)
            hist, n, st = s.run(STEPS)
            assert st == 0 and n == STEPS
            err = np.linalg.norm(s.state()[:, 0] - Wo[:, 0]) / np.linalg.norm(Wo[:, 0])
            assert err < 1e-5, (mname, tile, err)


def test_batch_of_small_and_uneven_meshes(oracle_runs):
    names = ["annulus8x1", "square2", "annulus48x6", "annulus24x3"]
    meshes = [oracle_runs[k, "o2-smooth"][0] for k in names]
    case_mesh = [2, 0, 1, 3, 0, 2]
    b = Batch(meshes, case_mesh, [FLOW] * len(case_mesh), precision="fp64")
    hist, steps, status = b.run(STEPS)
    assert list(steps) == [STEPS] * len(case_mesh) and not status.any()
    for i, mi in enumerate(case_mesh):
        _, Wo, ho, _ = oracle_runs[names[mi], "o2-smooth"]
        assert np.array_equal(b.state(i), Wo), names[mi]
        np.testing.assert_allclose(hist[i, 1:STEPS + 1], ho[1:], rtol=1e-12, atol=0)
