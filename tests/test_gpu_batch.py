"""GPU batch engine and C-ABI behaviour: uneven batches equal single solves,
per-case stopping, resumed runs, state round trip, failure reporting."""
import numpy as np
import pytest

from gpueularsolver2d_amd import Batch, FvmError, Solver, naca0012_mesh
from oracle_ctypes import OracleMesh, OracleSolver

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def meshes():
    return [naca0012_mesh(3000, seed=11), naca0012_mesh(6000, seed=12)]


FLOWS = [dict(mach=0.5, alpha_deg=0.0, cfl=1.5), dict(mach=0.8, alpha_deg=1.25, cfl=1.5),
         dict(mach=0.7, alpha_deg=3.0, cfl=1.2)]


@pytest.mark.parametrize("precision,kernels", [("fp64", "tiled"), ("fp32", "tiled"), ("fp64", "plain"),
                                               ("fp32", "split")])
def test_batch_equals_single_solves(meshes, precision, kernels):
    case_mesh = [0, 1, 1, 0, 1]
    flows = [FLOWS[i % 3] for i in range(len(case_mesh))]
    b = Batch(meshes, case_mesh, flows, precision=precision, kernels=kernels)
    hist, steps, status = b.run(25)
    assert list(steps) == [25] * 5 and not status.any()
    cl, cd = b.forces()
    for i, (m, f) in enumerate(zip(case_mesh, flows)):
        s = Solver(meshes[m], precision=precision, kernels=kernels, **f)
        h, n, st = s.run(25)
        assert np.array_equal(b.state(i), s.state())
        np.testing.assert_array_equal(hist[i, 1:26], h[1:])
        assert (cl[i], cd[i]) == s.forces()


@pytest.mark.parametrize("kernels", ["tiled", "split"])
def test_global_dt_batch_matches_oracle_per_case(meshes, kernels):
    # each case advances with its own minimum over its own tiles; small tiles
    # give every case several of them, and the oracle checks each case alone
    case_mesh = [1, 0, 1]
    flows = [FLOWS[1], FLOWS[2], FLOWS[0]]
    b = Batch(meshes, case_mesh, flows, precision="fp64", kernels=kernels, global_dt=True,
              tile_cells=64, threads=128)
    hist, steps, status = b.run(12)
    assert list(steps) == [12] * 3 and not status.any()
    for i, (m, f) in enumerate(zip(case_mesh, flows)):
        om = OracleMesh(meshes[m].x, meshes[m].y, meshes[m].tri)
        o = OracleSolver(om, f["mach"], f["alpha_deg"], cfl=f["cfl"], global_dt=True)
        ho, _ = o.run(12)
        assert np.array_equal(b.state(i), o.state()), (kernels, i)
        np.testing.assert_allclose(hist[i, 1:13], ho[1:], rtol=1e-12, atol=0)


def test_batch_fp64_matches_oracle(meshes):
    b = Batch(meshes, [1, 0], [FLOWS[1], FLOWS[2]], precision="fp64")
    b.run(15)
    for i, (m, f) in enumerate([(1, FLOWS[1]), (0, FLOWS[2])]):
        om = OracleMesh(meshes[m].x, meshes[m].y, meshes[m].tri)
        o = OracleSolver(om, f["mach"], f["alpha_deg"], cfl=f["cfl"])
        o.run(15)
        assert np.array_equal(b.state(i), o.state())


def test_per_case_stop(meshes):
    # each case stops at its own first step with Resd <= epsl, as a single run would
    b = Batch(meshes, [0, 1, 0], [FLOWS[0], FLOWS[1], FLOWS[2]], precision="fp64",
              steps_per_graph=5)
    hist, steps, status = b.run(400, epsl=0.2)
    for i, (m, f) in enumerate([(0, FLOWS[0]), (1, FLOWS[1]), (0, FLOWS[2])]):
        s = Solver(meshes[m], precision="fp64", **f)
        h, n, st = s.run(400, epsl=0.2)
        assert steps[i] == n
        assert np.array_equal(b.state(i), s.state())
    assert len(set(steps.tolist())) > 1


def test_resume_equals_one_run(meshes):
    # a batch holds 4096 steps of residual history from its creation; the
    # second call needs a longer one (a new buffer under the same captured
    # graphs), the third fits in the one it got
    a = Solver(meshes[0], precision="fp32", **FLOWS[1])
    a.run(30)
    h1a, _, _ = a.run(4150)
    h1, n1, _ = a.run(4200)
    b = Solver(meshes[0], precision="fp32", **FLOWS[1])
    h2, n2, _ = b.run(4200)
    assert n1 == n2 == 4200
    assert np.array_equal(a.state(), b.state())
    np.testing.assert_array_equal(h1a[31:], h2[31:4151])
    np.testing.assert_array_equal(h1[4151:], h2[4151:])


def test_run_lengths_cover_graph_ladder(meshes):
    # a run is whole graphs of steps_per_graph steps, one graph of each even
    # step count below it and at most one direct step; any split of a run
    # into calls gives the same bits as one call
    for precision, kernels in (("fp32", "tiled"), ("fp64", "tiled"), ("fp64", "split"),
                               ("fp64", "plain")):
        a = Solver(meshes[0], precision=precision, kernels=kernels, **FLOWS[2])
        hist = []
        for n in (1, 4, 9, 30, 79, 127, 128):   # increments 1, 3, 5, 21, 49, 48, 1
            h, k, st = a.run(n)
            assert k == n and st == 0
            hist.append(h)
        b = Solver(meshes[0], precision=precision, kernels=kernels, **FLOWS[2])
        hb, _, _ = b.run(128)
        assert np.array_equal(a.state(), b.state()), (precision, kernels)
        prev = 0
        for n, h in zip((1, 4, 9, 30, 79, 127, 128), hist):
            np.testing.assert_array_equal(h[prev + 1:n + 1], hb[prev + 1:n + 1])
            prev = n


def test_state_round_trip_and_restart(meshes):
    s = Solver(meshes[1], precision="fp64", **FLOWS[2])
    s.run(10)
    W = s.state()
    t = Solver(meshes[1], precision="fp64", **FLOWS[2])
    t.set_state(W)
    assert np.array_equal(t.state(), W)
    # the FP32 path stores single precision
    u = Solver(meshes[1], precision="fp32", **FLOWS[2])
    u.set_state(W)
    np.testing.assert_array_equal(u.state(), W.astype(np.float32).astype(np.float64))


def test_nan_reported_not_silent(meshes):
    # a non-physical state (negative density) forced from the host makes the
    # residual non-finite: the run stops and reports FVM_ERR_NAN instead of
    # ending silently as the reference loop does
    s = Solver(meshes[0], precision="fp64", **FLOWS[0])
    W = s.state()
    W[100, 0] = -1.0
    s.set_state(W)
    with pytest.raises(FvmError) as e:
        s.run(50)
    assert e.value.name == "FVM_ERR_NAN"
    h, n, st = s.run(50, check=False)
    assert st == 5 and n < 50


def test_kernel_times_and_config(meshes):
    s = Solver(meshes[1], precision="fp32", tile_cells=64, threads=128, **FLOWS[1])
    s.run(3)
    ms = s.kernel_times(reps=5)
    assert 0 < ms[0] < 10 and ms[1] == 0.0          # one kernel per stage
    ms2 = Solver(meshes[1], precision="fp32", kernels="split", **FLOWS[1]).kernel_times(reps=5)
    assert all(0 < v < 10 for v in ms2)             # flux_tile, update_tile
    cfg = s.config()
    assert cfg["tile_cells"] == 64 and cfg["threads"] == 128 and cfg["tiles"] >= meshes[1].ncell // 64
    W = s.state()
    s.kernel_times(reps=3)  # timing launches leave the state unchanged
    assert np.array_equal(s.state(), W)
    # library-chosen tile size: the tiles fill whole waves of the CUs' slots
    for prec, threads in (("fp32", 512), ("fp64", 512)):
        c = Solver(meshes[1], precision=prec, **FLOWS[1]).config()
        slots = c["blocks_per_cu"] * c["cus"]
        assert c["threads"] == threads and slots > 0
        assert threads // 4 <= c["tile_cells"] <= threads * 5 // 8
        assert c["tiles"] == -(-meshes[1].ncell // c["tile_cells"]) and c["tiles"] <= slots


def test_results_independent_of_graph_length(meshes):
    # The graph length decides which steps close inside a stage launch and
    # which in finalize_kernel; both are the same code built with the
    # precision's options, so histories and states match bit for bit.
    for precision in ("fp32", "fp64"):
        ref = None
        for spg in (2, 6, 20):
            s = Solver(meshes[0], precision=precision, steps_per_graph=spg, **FLOWS[1])
            h, n, st = s.run(50)
            assert n == 50 and st == 0
            if ref is None:
                ref = (h, s.state())
            else:
                np.testing.assert_array_equal(h, ref[0])
                assert np.array_equal(s.state(), ref[1])


def test_oversized_tiles_are_argument_errors(meshes):
    # a tile whose LDS footprint exceeds the CU's 160 KB is rejected at
    # creation with a status, before any launch
    for precision, threads in (("fp32", 0), ("fp64", 0), ("fp32", 64)):
        with pytest.raises(FvmError) as e:
            Solver(meshes[1], precision=precision, tile_cells=6000, threads=threads, **FLOWS[0])
        assert e.value.name == "FVM_ERR_ARGUMENT"
        assert "LDS" in str(e.value) or "too large" in str(e.value)


def test_batch_set_state_restores_a_case(meshes):
    # a case restored into a fresh batch continues exactly as in the batch it
    # came from (the step count is kept: the state alone sets the update)
    flows = [FLOWS[0], FLOWS[1], FLOWS[2]]
    for precision in ("fp64", "fp32"):
        a = Batch(meshes, [0, 1, 0], flows, precision=precision)
        a.run(5)
        W = a.state(2)
        b = Batch(meshes, [0, 1, 0], flows, precision=precision)
        b.set_state(2, W)
        assert np.array_equal(b.state(2), W)
        a.run(15)
        b.run(10)
        assert np.array_equal(a.state(2), b.state(2))
        with pytest.raises(FvmError):
            b.set_state(3, W)


def test_long_graphs_split_runs_into_ladder_chunks(meshes):
    # with 200 steps per graph the even ladder stops at 64 steps (32 graphs),
    # so a run of 199 steps is chunks of 6, 64, 64, 64 and one direct step;
    # the bits equal those of the default 50-step graphs
    a = Solver(meshes[0], precision="fp32", steps_per_graph=200, **FLOWS[2])
    assert a.config()["steps_per_graph"] == 200
    ha, n, st = a.run(199)
    assert n == 199 and st == 0
    hb2, n, st = a.run(401)   # one whole graph and a 2-step chunk
    assert n == 401 and st == 0
    b = Solver(meshes[0], precision="fp32", **FLOWS[2])
    hb, _, _ = b.run(401)
    assert np.array_equal(a.state(), b.state())
    np.testing.assert_array_equal(ha[1:], hb[1:200])
    np.testing.assert_array_equal(hb2[200:], hb[200:])
    with pytest.raises(FvmError) as e:
        Solver(meshes[0], precision="fp32", steps_per_graph=1025, **FLOWS[2])
    assert "steps_per_graph" in str(e.value)
