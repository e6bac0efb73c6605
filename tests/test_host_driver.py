"""The reference-compatible host program (inputMesh ... Output over the C
ABI): file handling on the CPU, and on the GPU an end-to-end run on the
reference's own mesh whose residual file reproduces the reference's printed
history digit for digit."""
import gzip
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
DRIVER = os.path.join(ROOT, "gpueularsolver2d_amd", "fvm_solver")


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(DRIVER):
        from gpueularsolver2d_amd import build
        build.build_driver()
    return DRIVER


def write_case(d, maxstep, epsl=-1.0):
    with gzip.open(os.path.join(GOLDEN, "2822-3k.dat.gz"), "rb") as src, \
            open(os.path.join(d, "2822-3k.dat"), "wb") as dst:
        shutil.copyfileobj(src, dst)
    with open(os.path.join(d, "input.dat"), "w") as f:
        f.write(f"0.729 2.31 1.4 101325 288.15 1.0 {epsl} {maxstep}\n")


def test_missing_files_are_errors_not_exit0(driver, tmp_path):
    r = subprocess.run([driver, "--quiet"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 2 and "Cannot open the mesh" in r.stderr
    write_case(str(tmp_path), 10)
    os.remove(tmp_path / "input.dat")
    r = subprocess.run([driver, "--quiet"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 1 and "Cannot open input" in r.stderr
    (tmp_path / "input.dat").write_text("0.5 0 1.4\n")
    r = subprocess.run([driver, "--quiet"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 1


@pytest.fixture(scope="module")
def driver_run(driver, tmp_path_factory):
    """One end-to-end run of the host program: the reference's mesh and case,
    no smoothing, 500 steps, residual every 10 steps."""
    d = tmp_path_factory.mktemp("driver")
    write_case(str(d), 500)
    r = subprocess.run([driver, "--no-smooth", "--resid-every", "10", "--quiet"], cwd=d,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return d


@pytest.mark.gpu
def test_driver_reproduces_reference_residual_file(driver_run):
    rows = [l.split() for l in open(driver_run / "unresidual.plt").read().splitlines()[3:]]
    got = {int(a): float(b) for a, b in rows}
    with open(os.path.join(GOLDEN, "reference_residuals.json")) as f:
        ref = json.load(f)["unresidual_nosmooth"]
    assert len(ref) == 50
    for stepno, val in ref:
        assert abs(got[stepno] - val) < 0.6e-10, (stepno, got[stepno], val)


def reference_postprocessing(mesh, W, gamma=1.4, p_in=101325.0, t_in=288.15, mach=0.729):
    """Node_Computing and Output's wall pressure (gpu_fvm.cu:71-131, 182-185)
    restated in numpy: area-weighted cell-to-node averages of rho, rho u,
    rho v, P, T and A (CalculationPTA, gpu_init.cu:49-61), then u, v and the
    Mach number per node; -Cp = -2 (p - PIN / (ROUIN VIN^2)) at wall nodes."""
    R = 287.06
    x, y, tri = mesh.x, mesh.y, mesh.tri - 1
    a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
    area = 0.5 * ((y[c] - y[a]) * (x[b] - x[a]) - (y[b] - y[a]) * (x[c] - x[a]))
    rho, ru, rv, re = W.T
    u, v = ru / rho, rv / rho
    e = re / rho - 0.5 * (u * u + v * v)
    P = (gamma - 1.0) * rho * e
    A = np.sqrt(gamma * P / rho)
    T = (gamma - 1.0) * e / R
    n = len(x)
    acc = np.zeros((n, 8))
    for q, val in enumerate((rho, ru, rv, re, T, P, A, np.ones_like(rho))):
        for k in range(3):
            np.add.at(acc[:, q], tri[:, k], val * area)
    ra = acc[:, 7]
    ROU = acc[:, 0] / ra
    U = acc[:, 1] / ra / ROU
    V = acc[:, 2] / ra / ROU
    Tn, Pn, An = acc[:, 4] / ra, acc[:, 5] / ra, acc[:, 6] / ra
    Mach = np.sqrt(U * U + V * V) / An
    ain = np.sqrt(gamma * R * t_in)
    vin = mach * ain
    rouin = p_in / (R * t_in)
    cp = -2.0 * (Pn - p_in / (rouin * vin * vin))
    return np.column_stack([x, y, ROU, U, V, Pn, Tn, Mach]), cp


@pytest.mark.gpu
def test_driver_node_averages_and_wall_pressure(driver_run):
    # the same case through the library gives the state the program wrote out
    from gpueularsolver2d_amd import Solver, mesh_topology, read_dat
    mesh = read_dat(os.path.join(GOLDEN, "2822-3k.dat.gz"))
    s = Solver(mesh, 0.729, 2.31, cfl=1.0, smoothing=False, precision="fp64")
    s.run(500)
    nodes, cp = reference_postprocessing(mesh, s.state())
    res = open(driver_run / "unresult.plt").read().splitlines()
    assert res[3].startswith(f"N={mesh.nnode},E={mesh.ncell}")
    got = np.array([[float(v) for v in l.split()] for l in res[6:6 + mesh.nnode]])
    np.testing.assert_allclose(got, nodes, rtol=0, atol=1.5e-10)   # 10 printed decimals
    tri = np.array([[int(v) for v in l.split()] for l in res[6 + mesh.nnode:]])
    assert np.array_equal(tri, mesh.tri)
    # wall pressure: both nodes of every wall edge in wall-id order
    t = mesh_topology(mesh)
    wall = np.flatnonzero(t["elog"] == 2)
    wall = wall[np.argsort(t["wallid"][wall])]
    pts = np.column_stack([t["node1"][wall], t["node2"][wall]]).ravel() - 1
    lines = open(driver_run / "unairfoilpress.plt").read().splitlines()
    assert lines[2].startswith(f"ZONE N={2 * len(wall)} E={len(wall)}")
    got = np.array([[float(v) for v in l.split()] for l in lines[3:3 + 2 * len(wall)]])
    np.testing.assert_allclose(got[:, 0], mesh.x[pts], rtol=0, atol=1.5e-10)
    np.testing.assert_allclose(got[:, 1], cp[pts], rtol=0, atol=1.5e-10)


@pytest.mark.gpu
def test_driver_forces_match_integrated_wall_pressure(driver, tmp_path):
    # The coefficients the program prints (fvm_get_forces: the cells' wall
    # fluxes) against the integral of the wall pressure coefficient it writes
    # in the reference's format (unairfoilpress.plt: node averages of
    # Node_Computing, gpu_fvm.cu:182-185), on a converged subsonic first-order
    # case.  Two discretisations of one surface integral: they agree to a few
    # per cent, while a force of the wrong sign or from the wrong components
    # would be off by its whole size.
    from force_checks import cp_force_coefficients, wall_edges
    from gpueularsolver2d_amd import mesh_topology, naca0012_mesh, write_dat
    mesh = naca0012_mesh(3000, seed=1)
    write_dat(mesh, str(tmp_path / "naca.dat"))
    (tmp_path / "input.dat").write_text("0.5 2.0 1.4 101325 288.15 1.5 1e-7 20000\n")
    r = subprocess.run([driver, "naca.dat", "input.dat", "--order", "1", "--no-smooth", "--quiet"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("steps ")][-1].split()
    cl, cd = float(line[line.index("Cl") + 1]), float(line[line.index("Cd") + 1])
    assert float(line[3]) < 1e-7   # converged
    t = mesh_topology(mesh)
    x, y, tri = np.asarray(mesh.x), np.asarray(mesh.y), np.asarray(mesh.tri) - 1
    cx = np.concatenate([[0.0], x[tri].mean(axis=1)])   # 1-based cell ids
    cy = np.concatenate([[0.0], y[tri].mean(axis=1)])
    a, b, _ = wall_edges(t, x, y, cx, cy)
    lines = open(tmp_path / "unairfoilpress.plt").read().splitlines()
    got = np.array([[float(v) for v in l.split()] for l in lines[3:3 + 2 * len(a)]])
    # the file lists (x, -Cp) for node1 and node2 of each wall edge in wall-id order
    wall = np.flatnonzero(t["elog"] == 2)
    wall = wall[np.argsort(t["wallid"][wall])]
    n1, n2 = t["node1"][wall] - 1, t["node2"][wall] - 1
    np.testing.assert_allclose(got[0::2, 0], x[n1], atol=1.5e-10)
    cp = np.zeros(len(x))
    cp[n1] = -got[0::2, 1]
    cp[n2] = -got[1::2, 1]
    clp, cdp = cp_force_coefficients(x, y, a, b, cp, 2.0)
    print(f"printed Cl {cl:.5f} Cd {cd:.5f}; integrated wall Cp: Cl {clp:.5f} Cd {cdp:.5f}", flush=True)
    assert cl > 0.1
    assert abs(clp - cl) < 0.06 * abs(cl)
    assert abs(cdp - cd) < 0.015


def read_resid(d):
    rows = [l.split() for l in open(os.path.join(d, "unresidual.plt")).read().splitlines()[3:]]
    return {int(a): float(b) for a, b in rows}


@pytest.mark.gpu
def test_driver_global_dt_reproduces_reference_file(driver, tmp_path):
    # the reference's unresidual_time2.plt (global time step, smoothing on),
    # through the host program's --global-dt, to every printed digit
    write_case(str(tmp_path), 2000)
    r = subprocess.run([driver, "--global-dt", "--resid-every", "100", "--quiet"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = read_resid(tmp_path)
    with open(os.path.join(GOLDEN, "reference_residuals.json")) as f:
        ref = json.load(f)["unresidual_time2"]
    assert len(ref) == 20
    for stepno, val in ref:
        assert abs(got[stepno] - val) < 0.6e-10, (stepno, got[stepno], val)


@pytest.mark.gpu
def test_driver_plain_kernels_write_the_same_files(driver, driver_run, tmp_path):
    # the reference's kernel structure (--plain) gives the default kernels'
    # FP64 bits: the same residual, node-average and wall-pressure files
    write_case(str(tmp_path), 500)
    r = subprocess.run([driver, "--no-smooth", "--plain", "--resid-every", "10", "--quiet"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for name in ("unresidual.plt", "unresult.plt", "unairfoilpress.plt"):
        with open(os.path.join(driver_run, name)) as a, open(os.path.join(tmp_path, name)) as b:
            assert a.read() == b.read(), name


@pytest.mark.gpu
def test_driver_fp32_tracks_fp64(driver, driver_run, tmp_path):
    # --fp32: the same run in FP32 bulk; its residual history follows the
    # FP64 one over these 500 transient steps (measured on MI355X: largest
    # relative difference 3.0e-6)
    write_case(str(tmp_path), 500)
    r = subprocess.run([driver, "--no-smooth", "--fp32", "--resid-every", "10", "--quiet"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    a, b = read_resid(driver_run), read_resid(tmp_path)
    assert sorted(a) == sorted(b) and len(a) == 50
    rel = max(abs(b[k] - a[k]) / a[k] for k in a)
    print(f"FP32 vs FP64 residual history, max relative difference {rel:.2e}", flush=True)
    assert rel < 3e-5
