"""The benchmark workloads: synthetic NACA 0012 meshes and flow cases.

One place for the configurations that the benchmark times and the tests
check, so both use the same meshes:

  c1        ~3k triangles (seed 1), M 0.5, alpha 0, first order, no smoothing
  c2        ~50k triangles (seed 2), M 0.8, alpha 1.25, second order + smoothing
  c3        ~200k triangles (seed 3), M 0.8, alpha 1.25, second order + smoothing
  batch64   4 meshes (~3k, 10k, 50k, 90k; seeds 11-14) x 16 (Mach, alpha)
            pairs on M in {0.5, 0.6167, 0.7333, 0.85} x alpha in {0, 1, 2, 3}
  batch256  256 meshes, sizes log-uniform in [2k, 150k] triangles (seed
            1000 + i), M uniform in [0.3, 0.85], alpha uniform in [-2, 4]
            (numpy seed 256)
All use gamma 1.4 and CFL 1.5.
"""
from __future__ import annotations

import multiprocessing as mp
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from .meshgen import naca0012_mesh

C1 = dict(cells=3000, seed=1, flow=dict(mach=0.5, alpha_deg=0.0, cfl=1.5),
          scheme=dict(order=1, smoothing=False))
C2 = dict(cells=50000, seed=2, flow=dict(mach=0.8, alpha_deg=1.25, cfl=1.5), scheme={})
C3 = dict(cells=200000, seed=3, flow=dict(mach=0.8, alpha_deg=1.25, cfl=1.5), scheme={})


def config_mesh(cfg):
    return naca0012_mesh(cfg["cells"], seed=cfg["seed"])


def batch64():
    """(meshes, case_mesh, flows) of the 64-case batch: case i runs on mesh
    i % 4 with the (Mach, alpha) pair (i // 4) % 16."""
    meshes = [naca0012_mesh(n, seed=s) for n, s in ((3000, 11), (10000, 12), (50000, 13),
                                                      (90000, 14))]
    pairs = [dict(mach=float(m), alpha_deg=float(a), cfl=1.5)
             for m in np.linspace(0.5, 0.85, 4) for a in (0.0, 1.0, 2.0, 3.0)]
    case_mesh = [i % 4 for i in range(64)]
    flows = [pairs[(i // 4) % 16] for i in range(64)]
    return meshes, case_mesh, flows


def _mesh_job(args):
    n, seed = args
    return naca0012_mesh(n, seed=seed)


def batch256_params():
    """Sizes, Mach numbers and angles of the 256 cases (numpy seed 256)."""
    rng = np.random.default_rng(256)
    sizes = np.exp(rng.uniform(np.log(2000), np.log(150000), 256)).astype(int)
    mach = rng.uniform(0.3, 0.85, 256)
    alpha = rng.uniform(-2.0, 4.0, 256)
    return sizes, mach, alpha


def _pool_context():
    """Worker processes started fresh (spawn), so that a parent that has
    opened the GPU does not pass its state on; spawn re-imports the parent's
    main module, so a parent run from stdin or a prompt forks instead."""
    main = getattr(sys.modules.get("__main__"), "__file__", None)
    return mp.get_context("spawn" if main and os.path.exists(main) else "fork")


def batch256(workers=16, cases=None):
    """(meshes, flows) of the 256-case batch (every case on its own mesh), or
    of the listed case indices only, generated in a pool of worker
    processes."""
    sizes, mach, alpha = batch256_params()
    idx = list(range(256)) if cases is None else list(cases)
    jobs = [(int(sizes[i]), 1000 + i) for i in idx]
    nw = max(1, min(workers, os.cpu_count() or 1, len(jobs)))
    if nw == 1:
        meshes = [_mesh_job(j) for j in jobs]
    else:
        with ProcessPoolExecutor(max_workers=nw, mp_context=_pool_context()) as ex:
            meshes = list(ex.map(_mesh_job, jobs))
    flows = [dict(mach=float(mach[i]), alpha_deg=float(alpha[i]), cfl=1.5) for i in idx]
    return meshes, flows
