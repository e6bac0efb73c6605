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
                   algorithmic bytes per launch / average launch time, the
                   device time of the timed steps (HIP events on the solver's
                   stream) over the stage launches in them
  config_c1 / c2   the small configurations: 3k cells first order against the
                   oracle (bitwise), 50k cells 5000 steps in FP64 and FP32
  accuracy         FP32 vs FP64 (the FP64 path equals the FP64 oracle bit for
                   bit) on the workload's limit cycle: window means of Cl/Cd,
                   beside FP64 vs FP64 from a 1e-12-perturbed start
  accuracy_converged   FP32 vs FP64 on a converging first-order case
  batch64          64 cases on 4 uneven meshes in one batch (FP32 Mcsu/s); per-case
                   |dCl| against FP64 beside FP64 from a 1e-12-perturbed start
  batch256         256 cases on 256 meshes of 2k-150k cells (~9.5M cells, HBM-resident),
                   the same accuracy figures
  cpu_baseline     the FP64 CPU oracle on this host's cores (OpenMP build), on a
                   bounded sample of steps of the same mesh; one-thread figure beside it
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


def progress(msg):
    """One line per finished part on stderr (the JSON result goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def sync():
    """Solver.run and Batch.run return after the device work has finished
    (the library synchronises its stream before returning), so the timed
    regions need no further synchronisation; kept as the one place to add it."""
    return None


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


def cpu_baseline(mesh, steps, omp_steps):
    """The FP64 oracle (the reference's loops restated in C) on this host's
    cores, on a bounded sample of steps of the same mesh and scheme: the
    all-core build (OpenMP over the per-cell and per-edge loops, the same
    bits) is the reported baseline; the one-thread build is reported beside it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ctypes import LIB_OMP_PATH, LIB_PATH, OracleMesh, OracleSolver, lib
    out = {}
    for path, n in ((LIB_PATH, steps), (LIB_OMP_PATH, omp_steps)):
        om = OracleMesh(mesh.x, mesh.y, mesh.tri, library=path)
        s = OracleSolver(om, 0.8, 1.25, cfl=1.5, order=2, smoothing=True)
        s.time_steps(1)
        sec = s.time_steps(n)
        out[path] = (4.0 * mesh.ncell * n / sec / 1e6, lib(path).orc_threads(), n, sec)
    v1, _, n1, s1 = out[LIB_PATH]
    vN, cores, nN, sN = out[LIB_OMP_PATH]
    return dict(value=vN, unit="Mcsu/s", cores=cores, kind="port",
                sample=f"{nN} time steps of the same {mesh.ncell}-cell mesh, FP64 oracle "
                       f"(gcc -O3 -fopenmp, {cores} threads), {sN:.1f} s",
                single_thread=dict(value=v1, cores=1,
                                   sample=f"{n1} time steps, gcc -O2, one thread, {s1:.1f} s"))


def limit_cycle_accuracy(mesh, flow, t0, t1, every):
    """FP32 against FP64 on the workload itself.  With the limiter the
    second-order transonic solution does not converge: the residual stalls
    near 1e-2 and Cl keeps cycling by a few 1e-4, and two FP64 runs whose
    start states differ by 1e-12 part on that cycle just as FP32 and FP64 do.
    So the coefficients are compared as means over [t0, t1] (sampled every
    `every` steps), beside the same statistics for FP64 against FP64 from a
    perturbed start, which measures the cycle's own sensitivity."""
    from gpueularsolver2d_amd import Solver
    A = Solver(mesh, precision="fp64", **flow)
    B = Solver(mesh, precision="fp64", **flow)
    W = B.state()
    B.set_state(W * (1.0 + 1e-12 * np.random.default_rng(0).standard_normal(W.shape)))
    F = Solver(mesh, precision="fp32", **flow)
    for s in (A, B, F):
        s.run(t0)
    smp, resd = [], []
    for n in range(t0 + every, t1 + 1, every):
        row = []
        for s in (A, B, F):
            h, _, _ = s.run(n)
            row += list(s.forces())
            if s is A:
                resd.append(h[n])
        smp.append(row)
    smp = np.array(smp)  # cl_A cd_A cl_B cd_B cl_F cd_F
    WA, WF = A.state(), F.state()
    stats = lambda i: dict(dcl_mean=float(smp[:, i].mean() - smp[:, 0].mean()),
                           dcd_mean=float(smp[:, i + 1].mean() - smp[:, 1].mean()),
                           dcl_rms=float(np.sqrt(np.mean((smp[:, i] - smp[:, 0]) ** 2))),
                           dcl_max=float(np.abs(smp[:, i] - smp[:, 0]).max()))
    return dict(window=[t0, t1], every=every, samples=len(smp),
                cl_fp64_mean=float(smp[:, 0].mean()), cd_fp64_mean=float(smp[:, 1].mean()),
                cl_fp64_std=float(smp[:, 0].std()),
                resd_fp64_window=[float(min(resd)), float(max(resd))],
                fp32_vs_fp64=dict(stats(4), rho_rel_l2_end=float(
                    np.linalg.norm(WF[:, 0] - WA[:, 0]) / np.linalg.norm(WA[:, 0]))),
                fp64_perturbed_vs_fp64=stats(2))


def converged_accuracy(mesh, max_steps, epsl=1e-7):
    """FP32 against FP64 on a case that converges: first order, M 0.5,
    alpha 1.25, no smoothing, same mesh, run until the FP64 residual drops
    below epsl (both precisions run the same number of steps)."""
    from gpueularsolver2d_amd import Solver
    flow = dict(mach=0.5, alpha_deg=1.25, cfl=1.5, order=1, smoothing=False)
    A = Solver(mesh, precision="fp64", **flow)
    h, n, _ = A.run(max_steps, epsl=epsl)
    F = Solver(mesh, precision="fp32", **flow)
    hF, _, _ = F.run(n)
    (cl, cd), (clF, cdF) = A.forces(), F.forces()
    WA, WF = A.state(), F.state()
    return dict(steps=int(n), resd_fp64=float(h[-1]), resd_fp32=float(hF[-1]), cl_fp64=cl, cd_fp64=cd,
                dcl=float(clF - cl), dcd=float(cdF - cd),
                rho_rel_l2=float(np.linalg.norm(WF[:, 0] - WA[:, 0]) / np.linalg.norm(WA[:, 0])))


def perturbed_fp64_control(meshes, case_mesh, flows, nsteps):
    """The same batch in FP64 from a start perturbed by 1e-12 part, run as
    long.  Its per-case Cl against the unperturbed FP64 batch shows how far
    each case's own trajectory separates over the run (limit cycles and
    transients amplify any difference), the scale for the FP32 differences."""
    from gpueularsolver2d_amd import Batch
    P = Batch(meshes, case_mesh, flows, precision="fp64")
    rng = np.random.default_rng(0)
    for i in range(len(case_mesh)):
        W = P.state(i)
        P.set_state(i, W * (1.0 + 1e-12 * rng.standard_normal(W.shape)))
    _, _, st = P.run(nsteps, check=False)
    cl, cd = P.forces()
    return cl, cd, st


def batch64(steps):
    """64 cases in one batch: four NACA 0012 meshes of uneven size (~3k, 10k,
    50k, 90k triangles, seeds 11-14) x 16 (Mach, alpha) pairs on the grid
    M in {0.5, 0.6167, 0.7333, 0.85} x alpha in {0, 1, 2, 3} deg; second order
    with the limiter and smoothing, CFL 1.5.  FP32 throughput over `steps`
    steps, and the per-case |dCl| against the same batch in FP64 after them
    (instantaneous values; see the limit-cycle comparison for their meaning).
    Cases that diverge (non-finite residual, in FP64 as well) stop at that
    step; only the cell-stage updates done count."""
    from gpueularsolver2d_amd import Batch, naca0012_mesh
    meshes = [naca0012_mesh(n, seed=s) for n, s in ((3000, 11), (10000, 12), (50000, 13), (90000, 14))]
    flows = [dict(mach=float(m), alpha_deg=float(a), cfl=1.5)
             for m in np.linspace(0.5, 0.85, 4) for a in (0.0, 1.0, 2.0, 3.0)]
    case_mesh = [i % 4 for i in range(64)]
    case_flow = [flows[(i // 4) % 16] for i in range(64)]
    cells = sum(meshes[m].ncell for m in case_mesh)
    t0 = time.perf_counter()
    F = Batch(meshes, case_mesh, case_flow, precision="fp32")
    tset = time.perf_counter() - t0
    F.run(20)
    sync()
    t0 = time.perf_counter()
    _, nF, stF = F.run(20 + steps, check=False)
    sync()
    t = time.perf_counter() - t0
    A = Batch(meshes, case_mesh, case_flow, precision="fp64")
    _, nA, stA = A.run(20 + steps, check=False)
    # a case whose residual turns non-finite stops there; the scheme itself
    # (FP64, equal to the oracle) diverges for such cases too
    done = sum((int(nF[i]) - 20) * meshes[m].ncell for i, m in enumerate(case_mesh))
    ok = (stF == 0) & (stA == 0)
    (clF, cdF), (clA, cdA) = F.forces(), A.forces()
    d = np.abs(clF - clA)[ok]
    clP, _, stP = perturbed_fp64_control(meshes, case_mesh, case_flow, 20 + steps)
    dP = np.abs(clP - clA)[ok & (stP == 0)]
    return dict(value=round(4.0 * done / t / 1e6, 1), unit="Mcsu/s", cases=64, cells=int(cells),
                mesh_cells=[m.ncell for m in meshes], steps=steps, launch=F.config(),
                setup_s=round(tset, 2),
                diverged={str(i): dict(mesh_cells=meshes[case_mesh[i]].ncell, **case_flow[i],
                                       fp32_step=int(nF[i]), fp64_step=int(nA[i]))
                          for i in range(64) if not ok[i]},
                dcl_max=float(d.max()), dcl_median=float(np.median(d)),
                dcd_max=float(np.abs(cdF - cdA)[ok].max()),
                fp64_perturbed_dcl_max=float(dP.max()), fp64_perturbed_dcl_median=float(np.median(dP)))


def config_c1(steps=2000):
    """Config C1: the ~3k-triangle NACA 0012 mesh (seed 1), M 0.5, alpha 0,
    first order, no smoothing, CFL 1.5, 2000 steps.  The GPU FP64 run against
    the FP64 oracle on this host's cores (the same arithmetic): whether every
    cell's state is bitwise equal, the largest history difference, Cl, Cd."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ctypes import LIB_OMP_PATH, OracleMesh, OracleSolver
    from gpueularsolver2d_amd import Solver, naca0012_mesh
    m = naca0012_mesh(3000, seed=1)
    kw = dict(mach=0.5, alpha_deg=0.0, cfl=1.5, order=1, smoothing=False)
    s = Solver(m, precision="fp64", **kw)
    h, n, _ = s.run(steps)
    o = OracleSolver(OracleMesh(m.x, m.y, m.tri, library=LIB_OMP_PATH), 0.5, 0.0, cfl=1.5, order=1,
                     smoothing=False)
    ho, _ = o.run(steps)
    cl, cd = s.forces()
    clo, cdo = o.forces()
    return dict(cells=m.ncell, steps=int(n), resd=float(h[-1]), cl=cl, cd=cd,
                state_bitwise_equal_to_oracle=bool(np.array_equal(s.state(), o.state())),
                hist_max_rel_diff=float(np.max(np.abs(h[1:] - ho[1:]) / np.abs(ho[1:]))),
                dcl_vs_oracle=cl - clo, dcd_vs_oracle=cd - cdo)


def config_c2(steps=5000):
    """Config C2: the ~50k-triangle NACA 0012 mesh (seed 2), M 0.8,
    alpha 1.25, second order with the limiter and smoothing, CFL 1.5, 5000
    steps, in FP64 (equal to the oracle) and in FP32: throughput of both and
    the FP32 errors after the 5000 steps (density rel. L2, Cl, Cd).  The flow
    is still in its transient at step 5000 (see the limit-cycle comparison for
    what instantaneous differences mean on such flows)."""
    from gpueularsolver2d_amd import Solver, naca0012_mesh
    m = naca0012_mesh(50000, seed=2)
    flow = dict(mach=0.8, alpha_deg=1.25, cfl=1.5)
    out = dict(cells=m.ncell, steps=steps)
    sol = {}
    for prec in ("fp64", "fp32"):
        s = Solver(m, precision=prec, **flow)
        sync()
        t0 = time.perf_counter()
        h, n, _ = s.run(steps)
        sync()
        t = time.perf_counter() - t0
        out[prec] = dict(value=round(4.0 * m.ncell * steps / t / 1e6, 1), unit="Mcsu/s",
                         resd=float(h[-1]), launch=s.config())
        sol[prec] = s
    (cl, cd), (clF, cdF) = sol["fp64"].forces(), sol["fp32"].forces()
    WA, WF = sol["fp64"].state(), sol["fp32"].state()
    out.update(cl_fp64=cl, cd_fp64=cd, dcl=clF - cl, dcd=cdF - cd,
               rho_rel_l2=float(np.linalg.norm(WF[:, 0] - WA[:, 0]) / np.linalg.norm(WA[:, 0])))
    return out


def _mesh_job(args):
    from gpueularsolver2d_amd import naca0012_mesh
    n, seed = args
    return naca0012_mesh(n, seed=seed)


def batch256_meshes(workers=16):
    """The 256-case workload: meshes, flows (see batch256)."""
    from concurrent.futures import ProcessPoolExecutor
    rng = np.random.default_rng(256)
    sizes = np.exp(rng.uniform(np.log(2000), np.log(150000), 256)).astype(int)
    mach = rng.uniform(0.3, 0.85, 256)
    alpha = rng.uniform(-2.0, 4.0, 256)
    with ProcessPoolExecutor(max_workers=min(workers, os.cpu_count() or 1)) as ex:
        meshes = list(ex.map(_mesh_job, [(int(n), 1000 + i) for i, n in enumerate(sizes)]))
    flows = [dict(mach=float(mach[i]), alpha_deg=float(alpha[i]), cfl=1.5) for i in range(256)]
    return meshes, flows


def batch256(steps, workers=16):
    """256 cases in one batch, every case on its own mesh: sizes log-uniform in
    [2k, 150k] triangles (mesh seed 1000 + i), M uniform in [0.3, 0.85], alpha
    uniform in [-2, 4] deg (drawn with numpy seed 256), second order with the
    limiter and smoothing, CFL 1.5, library-chosen launch shape.  FP32
    throughput over `steps` steps (all state resident in HBM), then the
    worst per-case |dCl| and |dCd| against the same batch in FP64 after the
    same steps.  Meshes are generated in a process pool."""
    from gpueularsolver2d_amd import Batch
    t0 = time.perf_counter()
    meshes, flows = batch256_meshes(workers)
    tgen = time.perf_counter() - t0
    case_mesh = list(range(256))
    cells = sum(m.ncell for m in meshes)
    t0 = time.perf_counter()
    F = Batch(meshes, case_mesh, flows, precision="fp32")
    tset = time.perf_counter() - t0
    F.run(10)
    sync()
    t0 = time.perf_counter()
    _, nF, stF = F.run(10 + steps, check=False)
    sync()
    t = time.perf_counter() - t0
    kt = F.kernel_times(reps=10)
    A = Batch(meshes, case_mesh, flows, precision="fp64")
    _, nA, stA = A.run(10 + steps, check=False)
    done = sum((int(nF[i]) - 10) * meshes[i].ncell for i in range(256))
    ok = (stF == 0) & (stA == 0)
    (clF, cdF), (clA, cdA) = F.forces(), A.forces()
    dcl, dcd = np.abs(clF - clA), np.abs(cdF - cdA)
    worst = int(np.argmax(np.where(ok, dcl, -1.0)))
    clP, _, stP = perturbed_fp64_control(meshes, list(range(256)), flows, 10 + steps)
    dP = np.abs(clP - clA)[ok & (stP == 0)]
    achieved = 402.0 * cells / (kt[0] * 1e-3) / 1e9
    return dict(value=round(4.0 * done / t / 1e6, 1), unit="Mcsu/s", cases=256, cells=int(cells),
                steps=steps, launch=F.config(), stage_tile_ms=kt[0],
                roofline_frac=round(achieved / HBM_PEAK_GBS, 4), achieved_gbs=round(achieved, 1),
                traffic=traffic_from_profiles(f"batch256-{cells}", "fp32", "stage_tile"),
                diverged={str(i): dict(fp32_status=int(stF[i]), fp32_step=int(nF[i]),
                                       fp64_status=int(stA[i]), fp64_step=int(nA[i]))
                          for i in range(256) if not ok[i]},
                dcl_max=float(dcl[ok].max()), dcl_median=float(np.median(dcl[ok])),
                dcd_max=float(dcd[ok].max()),
                fp64_perturbed_dcl_max=float(dP.max()), fp64_perturbed_dcl_median=float(np.median(dP)),
                worst_case=dict(case=worst, mesh_cells=meshes[worst].ncell, **flows[worst],
                                cl_fp64=float(clA[worst])),
                mesh_generation_s=round(tgen, 1), setup_s=round(tset, 2))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cells", type=int, default=200000)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--acc-window", type=lambda v: [int(x) for x in v.split(":")],
                    default=[30000, 70000, 100],
                    help="start:end:every of the averaging window of the limit-cycle comparison "
                         "(end 0 skips it)")
    ap.add_argument("--conv-steps", type=int, default=40000,
                    help="step limit of the converged first-order comparison (0 skips it)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the C1 (3k, first order) and C2 (50k, 5000 steps) lines")
    ap.add_argument("--batch-steps", type=int, default=2000,
                    help="steps of the 64-case batch (0 skips it)")
    ap.add_argument("--batch256-steps", type=int, default=3000,
                    help="steps of the 256-case batch (0 skips it)")
    ap.add_argument("--cpu-steps", type=int, default=100,
                    help="steps of the one-thread CPU sample")
    ap.add_argument("--cpu-omp-steps", type=int, default=200,
                    help="steps of the all-core CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-cells", type=int, default=0)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--fp32-only", action="store_true",
                    help="time only the FP32 path (for profiler runs)")
    ap.add_argument("--fp64-only", action="store_true",
                    help="time only the FP64 one-kernel-per-stage path (for profiler runs)")
    ap.add_argument("--batch256-fp32-only", action="store_true",
                    help="only run the FP32 256-case batch for --batch256-steps steps (for "
                         "profiler runs)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="only run one device copy kernel of this many MiB and exit (calibrates "
                         "the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.batch256_fp32_only:
        # the FP32 256-case batch alone (for profiler runs)
        from gpueularsolver2d_amd import Batch
        meshes, flows = batch256_meshes()
        b = Batch(meshes, list(range(256)), flows, precision="fp32")
        _, n, st = b.run(args.batch256_steps, check=False)
        kt = b.kernel_times(reps=5)
        print(json.dumps({"batch256_steps": args.batch256_steps, "cells": sum(m.ncell for m in meshes),
                          "stage_tile_ms": kt[0], "launch": b.config()}))
        return 0
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
              "(independent cases would be split across GPUs by case)",
              file=sys.stderr)
        return 2

    from gpueularsolver2d_amd import Solver, naca0012_mesh
    mesh = naca0012_mesh(args.cells, seed=args.seed)
    C = mesh.ncell
    flow = dict(mach=0.8, alpha_deg=1.25, cfl=1.5)
    csu = 4.0 * C * args.steps

    if args.fp64_only:
        s64 = Solver(mesh, precision="fp64", kernels="tiled", **flow)
        t64, _ = timed_run(s64, args.warmup, args.steps)
        ms64_region, _ = s64.timing()
        kt64 = s64.kernel_times(reps=20)
        print(json.dumps({"metric": METRIC, "value": round(csu / t64 / 1e6, 1), "unit": "Mcsu/s",
                          "precision": "fp64", "steps": args.steps,
                          "kernel_ms": ms64_region / (4 * args.steps), "kernel_ms_isolated": kt64[0]}))
        return 0
    s32 = Solver(mesh, precision="fp32", kernels="tiled", tile_cells=args.tile_cells,
                 threads=args.threads, **flow)
    t32, _ = timed_run(s32, args.warmup, args.steps)
    ms_region, _ = s32.timing()  # device time of the timed steps, solver stream
    kt = s32.kernel_times(reps=50)
    v32 = csu / t32 / 1e6
    progress(f"fp32 {v32:.1f} Mcsu/s")
    if args.fp32_only:
        print(json.dumps({"metric": METRIC, "value": round(v32, 1), "unit": "Mcsu/s",
                          "steps": args.steps, "kernel_ms": kt}))
        return 0
    s64 = Solver(mesh, precision="fp64", kernels="tiled", **flow)
    t64, _ = timed_run(s64, args.warmup, args.steps)
    ms64_region, _ = s64.timing()
    kt64 = s64.kernel_times(reps=20)
    plain_steps = max(10, args.steps // 4)
    sp = Solver(mesh, precision="fp64", kernels="plain", **flow)
    tpl, _ = timed_run(sp, 5, plain_steps)

    v64 = csu / t64 / 1e6
    vpl = 4.0 * C * plain_steps / tpl / 1e6
    progress(f"fp64 tiled {v64:.1f}, fp64 plain {vpl:.1f} Mcsu/s")

    c1 = config_c1() if not args.no_configs else None
    c2 = config_c2() if not args.no_configs else None
    progress("configs C1, C2 done")
    acc = limit_cycle_accuracy(mesh, flow, *args.acc_window) if args.acc_window[1] > 0 else None
    progress("limit-cycle accuracy done")
    conv = converged_accuracy(mesh, args.conv_steps) if args.conv_steps > 0 else None
    progress("converged-case accuracy done")
    b64 = batch64(args.batch_steps) if args.batch_steps > 0 else None
    progress("64-case batch done")
    b256 = batch256(args.batch256_steps) if args.batch256_steps > 0 else None
    progress("256-case batch done")

    w = 4  # FP32 state and stored geometry
    if kt[1] > 0:   # two kernels per stage: flux_tile dominates; its share of a
        # step cannot be read off the step's events, so its own launches are timed
        name, bytes_unit, t_kern = "flux_tile", flux_bytes(w), kt[0]
    else:           # one kernel per stage carries all four sweeps: the timed
        # steps' device time over their 4 launches per step (the launch gaps and
        # the one close launch per graph included)
        name, bytes_unit = "stage_tile", flux_bytes(w) + update_bytes(w)
        t_kern = ms_region / (4 * args.steps)
    achieved = bytes_unit * C / (t_kern * 1e-3) / 1e9
    roof = dict(bound="hbm", kernel=name, achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
                unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                traffic=traffic_from_profiles(f"naca0012-{C}", "fp32", name),
                kernel_ms=t_kern, kernel_ms_isolated=kt[0],
                bytes_per_csu_model=flux_bytes(w) + update_bytes(w),
                bytes_per_csu_kernel=bytes_unit,
                step_gbs_model=round((flux_bytes(w) + update_bytes(w)) * v32 * 1e6 / 1e9, 1))

    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        cpu = cpu_baseline(mesh, args.cpu_steps, args.cpu_omp_steps)
        progress("cpu baseline done")

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
        "fp64_tiled": {"value": round(v64, 1), "kernel_ms": ms64_region / (4 * args.steps),
                       "kernel_ms_isolated": kt64[0]},
        "fp64_plain": {"value": round(vpl, 1), "steps": plain_steps},
        "fp32_over_fp64_plain": round(v32 / vpl, 2),
        "config_c1": c1,
        "config_c2": c2,
        "accuracy": acc,
        "accuracy_converged": conv,
        "batch64": b64,
        "batch256": b256,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
