"""GPU parity on the benchmark's configurations that no other test runs:

* C2, the ~50k-triangle NACA 0012 mesh (seed 2) at M 0.8, alpha 1.25, second
  order with smoothing: the plain FP64 kernels (the reference's structure) and
  the tiled FP64 kernels bit for bit against the oracle, and the FP32 path
  against the oracle over the first steps of the transient;
* C4, the 64-case batch (four meshes of ~3k-90k triangles, 16 cases each):
  the FP64 batch bit for bit against the oracle on one case of each mesh, and
  the FP32 batch against the FP64 batch case by case;
* C5, the 256-case batch of uneven meshes (2k-150k triangles, ~9.5M cells):
  the FP64 batch bit for bit against the oracle on its smallest, median and
  largest case, and the FP32 batch against the FP64 batch case by case.
FP32 tolerances: the error of each conservative relative to the L2 norm of
the whole state, and |dCl|, |dCd|, after 10 steps from the freestream (before
the limiter and smoothing dynamics amplify rounding differences).
"""
import numpy as np
import pytest

from gpueularsolver2d_amd import Batch, Solver, workloads
from oracle_ctypes import LIB_OMP_PATH, OracleMesh, OracleSolver

pytestmark = pytest.mark.gpu


def oracle(mesh, flow, steps, **kw):
    om = OracleMesh(mesh.x, mesh.y, mesh.tri, library=LIB_OMP_PATH)
    s = OracleSolver(om, flow["mach"], flow["alpha_deg"], cfl=flow["cfl"], **kw)
    h, _ = s.run(steps)
    return s, h


def state_err(W, Wref):
    scale = np.linalg.norm(Wref)
    return max(float(np.linalg.norm(W[:, k] - Wref[:, k]) / scale) for k in range(4))


@pytest.fixture(scope="module")
def c2():
    return workloads.config_mesh(workloads.C2)


def test_c2_plain_and_tiled_fp64_bitwise_vs_oracle(c2):
    steps = 12
    flow = workloads.C2["flow"]
    print(f"C2 FP64 vs oracle, {c2.ncell} cells, {steps} steps", flush=True)
    orc, ho = oracle(c2, flow, steps)
    Wo = orc.state()
    for kernels in ("plain", "tiled"):
        s = Solver(c2, precision="fp64", kernels=kernels, **flow)
        h, n, st = s.run(steps)
        assert st == 0 and n == steps
        assert np.array_equal(s.state(), Wo), kernels
        np.testing.assert_allclose(h[1:], ho[1:], rtol=1e-12, atol=0)
        assert s.forces() == pytest.approx(orc.forces(), rel=1e-12, abs=1e-14)


def test_c2_fp32_tracks_oracle(c2):
    flow = workloads.C2["flow"]
    orc, _ = oracle(c2, flow, 10)
    s = Solver(c2, precision="fp32", **flow)
    _, n, st = s.run(10)
    assert st == 0 and n == 10
    err = state_err(s.state(), orc.state())
    (cl, cd), (clo, cdo) = s.forces(), orc.forces()
    print(f"C2 FP32 vs oracle after 10 steps: state {err:.2e}, dCl {cl - clo:.2e}, "
          f"dCd {cd - cdo:.2e}", flush=True)
    # measured on MI355X: 1.2e-6, dCl 7.6e-8, dCd 3.1e-7 (the coarser 50k
    # mesh amplifies rounding more than the 200k one, 1.5e-7 there)
    assert err < 5e-6
    assert abs(cl - clo) < 5e-6 and abs(cd - cdo) < 5e-6


@pytest.fixture(scope="module")
def c4():
    return workloads.batch64()


def test_c4_fp64_batch_bitwise_vs_oracle(c4):
    meshes, case_mesh, flows = c4
    steps = 5
    b = Batch(meshes, case_mesh, flows, precision="fp64")
    print(f"C4 FP64 batch: 64 cases on meshes of {[m.ncell for m in meshes]} cells, "
          f"launch {b.config()}", flush=True)
    _, n, st = b.run(steps)
    assert list(n) == [steps] * 64 and not st.any()
    # one case per mesh, each at a different (Mach, alpha) pair
    for i in (0, 21, 42, 63):
        orc, _ = oracle(meshes[case_mesh[i]], flows[i], steps)
        assert np.array_equal(b.state(i), orc.state()), f"case {i}"


def test_c4_fp32_batch_tracks_fp64_batch(c4):
    meshes, case_mesh, flows = c4
    steps = 10
    out = {}
    for prec in ("fp32", "fp64"):
        b = Batch(meshes, case_mesh, flows, precision=prec)
        _, n, st = b.run(steps)
        assert list(n) == [steps] * 64 and not st.any()
        out[prec] = ([b.state(i) for i in range(64)], b.forces())
        del b
    W32, (cl32, cd32) = out["fp32"]
    W64, (cl64, cd64) = out["fp64"]
    err = np.array([state_err(a, b) for a, b in zip(W32, W64)])
    dcl, dcd = np.abs(cl32 - cl64), np.abs(cd32 - cd64)
    print(f"C4 FP32 vs FP64 after {steps} steps: state max {err.max():.2e} median "
          f"{np.median(err):.2e}; |dCl| max {dcl.max():.2e}; |dCd| max {dcd.max():.2e}", flush=True)
    # the same bars as the 256-case batch below (and the reasons for them)
    assert np.median(err) < 1e-6 and np.percentile(err, 95) < 5e-6
    assert np.median(dcl) < 1e-6 and np.percentile(dcl, 95) < 5e-6
    assert dcl.max() < 5e-4 and dcd.max() < 5e-4


@pytest.fixture(scope="module")
def c5():
    print("generating the 256 meshes of C5", flush=True)
    meshes, flows = workloads.batch256()
    return meshes, flows


def test_c5_fp64_batch_bitwise_vs_oracle(c5):
    meshes, flows = c5
    steps = 5
    b = Batch(meshes, list(range(256)), flows, precision="fp64")
    print(f"C5 FP64 batch: {sum(m.ncell for m in meshes)} cells, launch {b.config()}", flush=True)
    _, n, st = b.run(steps)
    assert list(n) == [steps] * 256 and not st.any()
    order = np.argsort([m.ncell for m in meshes])
    for i in (int(order[0]), int(order[128]), int(order[-1])):
        orc, _ = oracle(meshes[i], flows[i], steps)
        assert np.array_equal(b.state(i), orc.state()), f"case {i} ({meshes[i].ncell} cells)"


def test_c5_fp32_batch_tracks_fp64_batch(c5):
    meshes, flows = c5
    steps = 10
    out = {}
    for prec in ("fp32", "fp64"):
        b = Batch(meshes, list(range(256)), flows, precision=prec)
        _, n, st = b.run(steps)
        assert list(n) == [steps] * 256 and not st.any()
        out[prec] = ([b.state(i) for i in range(256)], b.forces())
        del b
    W32, (cl32, cd32) = out["fp32"]
    W64, (cl64, cd64) = out["fp64"]
    err = np.array([state_err(a, b) for a, b in zip(W32, W64)])
    dcl, dcd = np.abs(cl32 - cl64), np.abs(cd32 - cd64)
    print(f"C5 FP32 vs FP64 after {steps} steps: state max {err.max():.2e} median "
          f"{np.median(err):.2e}; |dCl| max {dcl.max():.2e}; |dCd| max {dcd.max():.2e}", flush=True)
    for i in np.argsort(-dcl)[:6]:
        print(f"  case {i}: {meshes[i].ncell} cells, M {flows[i]['mach']:.4f}, "
              f"alpha {flows[i]['alpha_deg']:.3f}: Cl64 {cl64[i]:.6f} dCl {cl32[i] - cl64[i]:.3e} "
              f"dCd {cd32[i] - cd64[i]:.3e} state {err[i]:.2e}", flush=True)
    # Most cases agree to rounding level.  A few can differ in single cells:
    # the smoothing keeps whichever of the raw residual R and the smoothed R2
    # is smaller in magnitude (gpu_solver.cu:1714-1718), a choice that jumps by
    # |R| + |R2| when they have opposite signs and nearly equal magnitudes,
    # and R, a difference of O(1) fluxes, carries a relative FP32 error far
    # above FP32's epsilon.  Such a near-tie in the energy of one wall cell at
    # step 10 (R = -3.38682e-4, R2 = +3.38696e-4 in the oracle) once moved a
    # case's Cl by 1.7e-4; on the current meshes the largest |dCl| is 9.4e-5
    # (measured on MI355X).
    assert np.median(err) < 1e-6 and np.percentile(err, 95) < 5e-6
    assert np.median(dcl) < 1e-6 and np.percentile(dcl, 95) < 5e-6
    assert dcl.max() < 5e-4 and dcd.max() < 5e-4
