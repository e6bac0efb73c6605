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


@pytest.mark.gpu
def test_driver_reproduces_reference_residual_file(driver, tmp_path):
    write_case(str(tmp_path), 500)
    r = subprocess.run([driver, "--no-smooth", "--resid-every", "10", "--quiet"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [l.split() for l in open(tmp_path / "unresidual.plt").read().splitlines()[3:]]
    got = {int(a): float(b) for a, b in rows}
    with open(os.path.join(GOLDEN, "reference_residuals.json")) as f:
        ref = json.load(f)["unresidual_nosmooth"]
    assert len(ref) == 50
    for stepno, val in ref:
        assert abs(got[stepno] - val) < 0.6e-10, (stepno, got[stepno], val)
    # node averages and wall pressure files are written and finite
    res = open(tmp_path / "unresult.plt").read().splitlines()
    assert res[3].startswith("N=1907,E=3558")
    vals = np.array([[float(v) for v in l.split()] for l in res[6:6 + 1907]])
    assert np.isfinite(vals).all() and (vals[:, 2] > 0).all()
    cp = open(tmp_path / "unairfoilpress.plt").read().splitlines()
    assert len(cp) == 3 + 2 * 128 + 128
