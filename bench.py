#!/usr/bin/env python3
"""Throughput of the explicit FVM Euler solver on one MI355X.

Workload (default): the ~200k-triangle NACA 0012 mesh (seed 3), M = 0.8,
alpha = 1.25 deg, second order with the limiter, residual smoothing, 4-stage
RK, CFL 1.5.  A step is one time step (four RK stages) of the whole mesh; the
metric counts cell-stage updates (4 * cells per step).

Measured in one run:
  value            FP32-bulk tiled path (FP64 geometry and reductions), Mcsu/s,
                   host wall clock over exactly --steps steps, bracketed by
                   device synchronisation
  fp64_tiled / fp64_plain   the same workload in FP64 (tiled kernels; the
                   reference's per-phase kernel structure)
  roofline         dominant kernel (stage_tile, one launch per RK stage):
                   algorithmic bytes per launch / average launch time from HIP
                   events on the solver stream
  accuracy         FP32 vs FP64 (the FP64 path equals the FP64 oracle bit for
                   bit) after --acc-steps steps: |dCl|, |dCd|, density rel. L2
  cpu_baseline     the FP64 CPU oracle on this host, one thread, on a bounded
                   sample of steps of the same mesh
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-stage-updates/s & % HBM peak; batched FP32 vs plain FP64 at Cl err<1e-4"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per cell-stage update of each tiled kernel (four
# neighbour-dependent sweeps, each input read once and each output written
# once, int32 ids, E/C = 1.5): flux_tile = sweeps 1+2 = 45 w_s + 17.5 w_g + 24,
# update_tile = sweeps 3+4 = 25 w_s + 1 w_g + 24 (w_s = state word, w_g =
# stored geometry word).
def flux_bytes(w):
    return 45 * w + 17.5 * w + 24


def update_bytes(w):
    return 25 * w + 1 * w + 24


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:
        pass


def timed_run(solver, warmup, steps):
    solver.run(warmup)
    sync()
    t0 = time.perf_counter()
    hist, n, st = solver.run(warmup + steps)
    sync()
    t1 = time.perf_counter()
    if n != warmup + steps or st != 0:
        raise RuntimeError(f"run stopped at step {n} with status {st}")
    return t1 - t0, hist


def traffic_from_profiles(workload, precision, kernel):
    """HBM bytes per launch measured with rocprofv3 counters (FETCH_SIZE +
    WRITE_SIZE, each scaled by the same counter's reading for a copy kernel
    of known size), recorded in profiles/traffic.json."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get(f"{workload}:{precision}:{kernel}")


def cpu_baseline(mesh, steps):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ctypes import OracleMesh, OracleSolver
    om = OracleMesh(mesh.x, mesh.y, mesh.tri)
    s = OracleSolver(om, 0.8, 1.25, cfl=1.5, order=2, smoothing=True)
    s.time_steps(1)
    sec = s.time_steps(steps)
    return dict(value=4.0 * mesh.ncell * steps / sec / 1e6, unit="Mcsu/s", cores=1, kind="port",
                sample=f"{steps} time steps of the same {mesh.ncell}-cell mesh, FP64 oracle "
                       f"(gcc -O2, one thread), {sec:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cells", type=int, default=200000)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--acc-steps", type=int, default=5000)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-cells", type=int, default=0)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--fp32-only", action="store_true",
                    help="time only the FP32 path (for profiler runs)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="only run one device copy kernel of this many MiB and exit (calibrates "
                         "the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_mb > 0:
        import torch
        x = torch.ones(args.copy_mb << 18, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        torch.mul(x, 1.0, out=y)  # one kernel: reads and writes copy_mb MiB each
        torch.cuda.synchronize()
        print(json.dumps({"copy_mib": args.copy_mb}))
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU only: the solver has no multi-GPU path "
              "(independent cases would be split across GPUs by case; see DESIGN.md)",
              file=sys.stderr)
        return 2

    from gpueularsolver2d_amd import Solver, naca0012_mesh
    mesh = naca0012_mesh(args.cells, seed=args.seed)
    C = mesh.ncell
    flow = dict(mach=0.8, alpha_deg=1.25, cfl=1.5)
    csu = 4.0 * C * args.steps

    s32 = Solver(mesh, precision="fp32", kernels="tiled", tile_cells=args.tile_cells,
                 threads=args.threads, **flow)
    t32, _ = timed_run(s32, args.warmup, args.steps)
    kt = s32.kernel_times(reps=50)
    v32 = csu / t32 / 1e6
    if args.fp32_only:
        print(json.dumps({"metric": METRIC, "value": round(v32, 1), "unit": "Mcsu/s",
                          "steps": args.steps, "kernel_ms": kt}))
        return 0
    s64 = Solver(mesh, precision="fp64", kernels="tiled", **flow)
    t64, _ = timed_run(s64, args.warmup, args.steps)
    kt64 = s64.kernel_times(reps=20)
    plain_steps = max(10, args.steps // 4)
    sp = Solver(mesh, precision="fp64", kernels="plain", **flow)
    tpl, _ = timed_run(sp, 5, plain_steps)

    v64 = csu / t64 / 1e6
    vpl = 4.0 * C * plain_steps / tpl / 1e6

    # accuracy after acc_steps from the same start.  With the limiter the
    # second-order scheme settles into a small limit cycle (the residual
    # stalls near 1e-2..1e-3), so besides the instantaneous difference the
    # coefficients are also averaged over the last `window` steps.
    acc = None
    if args.acc_steps > 0:
        n = max(args.acc_steps, args.warmup + args.steps)
        window, every = 1000, 50
        start = max(n - window, args.warmup + args.steps)
        s32.run(start)
        s64.run(start)
        samples = []
        for k in range(start + every, n + 1, every):
            s32.run(k)
            s64.run(k)
            samples.append(s32.forces() + s64.forces())
        smp = np.array(samples)
        W32, W64 = s32.state(), s64.state()
        cl32, cd32, cl64, cd64 = smp[-1]
        acc = dict(steps=n, cl_fp64=cl64, cd_fp64=cd64, dcl=abs(cl32 - cl64), dcd=abs(cd32 - cd64),
                   rho_rel_l2=float(np.linalg.norm(W32[:, 0] - W64[:, 0]) / np.linalg.norm(W64[:, 0])),
                   window=[int(start), int(n)],
                   dcl_mean=float(abs(smp[:, 0].mean() - smp[:, 2].mean())),
                   dcd_mean=float(abs(smp[:, 1].mean() - smp[:, 3].mean())),
                   cl_fp64_range=float(np.ptp(smp[:, 2])))

    w = 4  # FP32 state and stored geometry
    if kt[1] > 0:   # two kernels per stage: flux_tile dominates
        name, bytes_unit, t_kern = "flux_tile", flux_bytes(w), kt[0]
    else:           # one kernel per stage carries all four sweeps
        name, bytes_unit, t_kern = "stage_tile", flux_bytes(w) + update_bytes(w), kt[0]
    achieved = bytes_unit * C / (t_kern * 1e-3) / 1e9
    roof = dict(bound="hbm", kernel=name, achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
                unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                traffic=traffic_from_profiles(f"naca0012-{C}", "fp32", name),
                kernel_ms=t_kern, bytes_per_csu_model=flux_bytes(w) + update_bytes(w),
                bytes_per_csu_kernel=bytes_unit,
                step_gbs_model=round((flux_bytes(w) + update_bytes(w)) * v32 * 1e6 / 1e9, 1))

    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        cpu = cpu_baseline(mesh, args.cpu_steps)

    out = {
        "metric": METRIC, "value": round(v32, 1), "unit": "Mcsu/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * t32 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"naca0012-{C} (~200k-triangle NACA 0012, seed {args.seed}), "
                               "M 0.8, alpha 1.25, 2nd order + limiter + smoothing, 4-stage RK, CFL 1.5",
                   "cells": C, "precision": "fp32 bulk, fp64 geometry + reductions",
                   "kernels": "tiled", "launch": s32.config()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp64_tiled": {"value": round(v64, 1), "kernel_ms": kt64[0]},
        "fp64_plain": {"value": round(vpl, 1), "steps": plain_steps},
        "fp32_over_fp64_plain": round(v32 / vpl, 2),
        "accuracy": acc,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
