"""Writes the committed fixtures under tests/golden/.

reference_residuals.json
    The residual histories the reference itself printed (its Tecplot files
    unresidual_nosmooth.plt, unresidual_smooth.plt, unresidual_time2.plt,
    3558-cell RAE 2822 mesh 2822-3k.dat), truncated to their first 500 steps.
    Data only: (step, Resd) pairs.
2822-3k.dat.gz, 2822-7k.dat.gz, 2822-5w.dat.gz
    The reference's own mesh files (3558, 13513 and 51800 cells), gzip-compressed.
    The 51800-cell mesh is the one the thesis's per-kernel timings were taken on.
oracle_vectors.npz
    Regression vectors of the FP64 oracle on small cases (residual history,
    final state, Cl/Cd): they pin the oracle against unintended change, not
    against the reference.

Run from the repository root:  python tests/golden/make_fixtures.py
(needs /root/reference for the first two; the last one needs only the oracle).
"""
import gzip
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))


def read_plt(path, max_step):
    out = []
    with open(path) as f:
        for line in f.read().splitlines()[3:]:
            a, b = line.split()
            if int(a) <= max_step:
                out.append([int(a), float(b)])
    return out


def reference_fixtures():
    d = {}
    for name in ("unresidual_nosmooth", "unresidual_smooth", "unresidual_time2"):
        d[name] = read_plt(os.path.join(REF, name + ".plt"), 500 if "time2" not in name else 2000)
    with open(os.path.join(HERE, "reference_residuals.json"), "w") as f:
        json.dump(d, f, indent=0)
    for m in ("2822-3k", "2822-7k", "2822-5w"):
        with open(os.path.join(REF, m + ".dat"), "rb") as src, \
                open(os.path.join(HERE, m + ".dat.gz"), "wb") as raw, \
                gzip.GzipFile(fileobj=raw, mode="wb", compresslevel=9, mtime=0) as dst:
            shutil.copyfileobj(src, dst)


ORACLE_CASES = {
    # name: (mesh, mach, alpha, cfl, order, smoothing, global_dt, steps)
    "rae3k_o2_s1": ("2822-3k", 0.729, 2.31, 1.0, 2, True, False, 30),
    "rae3k_o1_s0": ("2822-3k", 0.5, 0.0, 1.5, 1, False, False, 30),
    "naca3k_o2_s1": ("naca3000s1", 0.8, 1.25, 1.5, 2, True, False, 30),
}


def load_mesh(name):
    from gpueularsolver2d_amd.meshgen import naca0012_mesh, read_dat
    if name.startswith("naca"):
        return naca0012_mesh(3000, seed=1)
    return read_dat(os.path.join(HERE, name + ".dat.gz"))


def oracle_fixtures():
    from oracle_ctypes import OracleMesh, OracleSolver
    arrays = {}
    for key, (mname, M, A, cfl, order, sm, gdt, steps) in ORACLE_CASES.items():
        mesh = load_mesh(mname)
        om = OracleMesh(mesh.x, mesh.y, mesh.tri)
        s = OracleSolver(om, M, A, cfl=cfl, order=order, smoothing=sm, global_dt=gdt)
        hist, _ = s.run(steps)
        arrays[key + "_hist"] = hist
        arrays[key + "_state"] = s.state()
        arrays[key + "_forces"] = np.array(s.forces())
    np.savez_compressed(os.path.join(HERE, "oracle_vectors.npz"), **arrays)


if __name__ == "__main__":
    if os.path.isdir(REF):
        reference_fixtures()
    oracle_fixtures()
    print("fixtures written to", HERE)
