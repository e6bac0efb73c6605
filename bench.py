#!/usr/bin/env python3
"""Throughput of the explicit FVM Euler solver on one MI355X.

Workload (the value): the ~200k-triangle NACA 0012 mesh (seed 3), M 0.8,
alpha 1.25 deg, second order with the limiter, residual smoothing, 4-stage
RK, CFL 1.5, FP32 bulk with FP64 geometry and reductions.  A step is one time
step (four RK stages) of the whole mesh; the metric counts cell-stage updates
(4 * cells per step), host wall clock over exactly --steps steps of one
fvm_run call, after --warmup steps; before those, 0.3 s of the same workload
on a second solver instance brings the GPU from its idle clock to its working
one (the host's set-up of the solver leaves it idle).

Measured in the same run:
  roofline         the dominant kernel (stage_tile, one launch per RK stage):
                   its average launch time over the timed steps (HIP events on
                   the solver's stream, launch gaps included) against three byte
                   counts per cell-stage update: the 4-sweep algorithmic model
                   (402 B, frac / frac_model), the counter-measured HBM-side
                   traffic of one launch (frac_measured) and the single-pass
                   compulsory bytes (98 B, frac_compulsory)
  fp64_tiled / fp64_plain   the same workload in FP64 (tiled kernels; the
                   reference's per-phase kernel structure)
  global_dt        the same workload with the global time step, FP32 and FP64
  config_c1        3k cells, first order: the GPU FP64 run against the oracle
  config_c2        50k cells, 5000 steps: FP64, FP32 (autotuned launch shape)
                   and plain FP64 throughput; FP32 accuracy as window means
  accuracy         FP32 vs FP64 on the workload's limit cycle (window means)
  accuracy_converged   FP32 vs FP64 on a converging first-order case
  batch64          64 cases on 4 uneven meshes in one batch (FP32)
  batch256         256 cases on 256 meshes of 2k-150k cells (~9.5M cells,
                   HBM-resident), autotuned launch shape, roofline fractions
  cpu_baseline     the FP64 CPU oracle on this host's cores
Per-case accuracy of the batches and C2: Cl and Cd averaged over a window of
steps after the timed ones, FP32 against FP64, beside the same statistics for
FP64 restarted from a state perturbed by 1e-12 (the flows' own sensitivity).
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-stage-updates/s & % HBM peak; batched FP32 vs plain FP64 at Cl err<1e-4"
HBM_PEAK_GBS = 8000.0
DCL_GATE = 1e-4


# Bytes per cell-stage update (w_s: state word, w_g: stored geometry word).
# Algorithmic model: four neighbour-dependent sweeps per stage, each reading
# its inputs once and writing its outputs once (int32 ids, E/C = 1.5):
# gradient + flux = 45 w_s + 17.5 w_g + 24, smoothing + update = 25 w_s + w_g + 24.
def model_bytes(w):
    return 70 * w + 18.5 * w + 48


# Single pass, every array read once and written once: 18.5 w + 24.
def compulsory_bytes(w):
    return 18.5 * w + 24


def flux_bytes(w):
    return 45 * w + 17.5 * w + 24


def progress(msg):
    """One line per finished part on stderr (the JSON result goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


# Seconds of the same workload, on a second solver instance, run right before
# a timed leg's warm-up steps.  The GPU lowers its clock while the host sets a
# solver up (tens of ms with the device idle), and it takes several ms of
# load to come back: a 20-step run right after set-up paced 43 us per step on
# the 198,578-cell mesh, against 40.3 right after other work and 39.1 in a
# 2000-step run.  The timed region is unchanged: exactly `steps` steps of the
# measured solver, after its `warmup` steps.
DEVICE_WARMUP_S = 0.3


def warm_device(warm, seconds=DEVICE_WARMUP_S):
    """Keeps the device busy for `seconds` with runs of the solver `warm`."""
    if warm is None or seconds <= 0:
        return
    n, t_end = 0, time.perf_counter() + seconds
    while time.perf_counter() < t_end:
        n += 200
        warm.run(n, check=False)


def timed_run(solver, warmup, steps, warm=None):
    """Wall clock of exactly `steps` steps in one run, after `warmup` steps,
    right after DEVICE_WARMUP_S of runs of `warm` (another instance of the
    same workload) when given.  Solver.run returns once the device work is
    done (the library waits for its stream), so the timer brackets the device
    work."""
    warm_device(warm)
    solver.run(warmup)
    t0 = time.perf_counter()
    hist, n, st = solver.run(warmup + steps)
    t1 = time.perf_counter()
    if n != warmup + steps or st != 0:
        raise RuntimeError(f"run stopped at step {n} with status {st}")
    return t1 - t0, hist


def kernel_source_sha():
    """Hash of the sources of the FP32 tiled kernels and their compiler flags
    (include paths left out, so the hash does not depend on where the
    repository lies): the measured traffic in profiles/ applies only to the
    kernel it was measured on."""
    from gpueularsolver2d_amd import build as B
    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "gpueularsolver2d_amd", "csrc")
    for f in ("fvm_tiled.h", "fvm_numerics.h", "fvm_internal.h", "fvm_launch.h",
              "fvm_tiled_inst.inc", "fvm_tiled32.hip"):
        with open(os.path.join(csrc, f), "rb") as fh:
            h.update(fh.read())
    flags = list(B.COMMON) + list(dict(B.SOURCES)["fvm_tiled32.hip"])
    flags = [f for i, f in enumerate(flags) if f != "-I" and (i == 0 or flags[i - 1] != "-I")]
    h.update(" ".join(flags).encode())
    return h.hexdigest()[:16]


def measured_traffic(workload, kernel="stage_tile"):
    """HBM-side bytes per launch measured with rocprofv3 counters (L2 read
    requests by size, 32/64/128 B, plus WRITE_SIZE) on the current kernel
    source, from profiles/traffic.json; None (with the reason) if that file
    holds no entry for this workload or was measured on another source."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None, "no profiles/traffic.json"
    with open(path) as f:
        d = json.load(f)
    e = d.get(f"{workload}:fp32:{kernel}")
    if not isinstance(e, dict):
        return None, f"no entry for {workload}"
    if e.get("kernel_source_sha") != kernel_source_sha():
        return None, "measured on another kernel source"
    return e, None


def roofline(cells, t_kern_ms, workload, w=4):
    """Fractions of the 8 TB/s HBM peak for the stage kernel at t_kern_ms per
    launch over `cells` cells: by the algorithmic model (frac_model, which
    counts intermediate round trips the fused kernel does not make, so it can
    pass 1), by counter-measured traffic (frac_measured) and by the
    compulsory bytes (frac_compulsory)."""
    t = t_kern_ms * 1e-3
    model = model_bytes(w) * cells / t / 1e9
    e, why = measured_traffic(workload)
    out = dict(bound="hbm", kernel="stage_tile", achieved=round(model, 1), peak=HBM_PEAK_GBS,
               unit="GB/s", frac=round(model / HBM_PEAK_GBS, 4),
               traffic=None if e is None else int(e["bytes_per_launch"]),
               kernel_ms=t_kern_ms,
               bytes_per_csu_model=model_bytes(w), frac_model=round(model / HBM_PEAK_GBS, 4),
               bytes_per_csu_compulsory=compulsory_bytes(w),
               frac_compulsory=round(compulsory_bytes(w) * cells / t / 1e9 / HBM_PEAK_GBS, 4),
               frac_model_note="achieved / frac / frac_model use the 4-sweep algorithmic bytes "
                               "(402 B/csu); the fused kernel moves about one read of each array, "
                               "so this is a throughput-equivalent and can exceed 1")
    if e is not None:
        out.update(frac_measured=round(e["bytes_per_launch"] / t / 1e9 / HBM_PEAK_GBS, 4),
                   measured_gbs=round(e["bytes_per_launch"] / t / 1e9, 1),
                   measured_bytes_per_cell=round(e["bytes_per_launch"] / cells, 1),
                   traffic_read=int(e["read_bytes"]), traffic_write=int(e["write_bytes"]))
    else:
        out.update(frac_measured=None, traffic_note=why)
    return out


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


# The reference's own 51,800-cell RAE 2822 mesh (2822-5w.dat), on which the
# thesis timed its final CUDA code: 19.109 ms per step on a Tesla C2050, FP64
# (10.84 Mcsu/s), and 811.75 ms per step for the one-thread CPU code on a
# Phenom 9850.  The flow is the reference's printed case (M 0.729, alpha
# 2.31 deg, CFL 1.0, second order with smoothing).
REF_MESH_5W = os.path.join(ROOT, "tests", "golden", "2822-5w.dat.gz")
REF_FLOW = dict(mach=0.729, alpha_deg=2.31, cfl=1.0)
C2050_MS_PER_STEP = 19.109
PHENOM_MS_PER_STEP = 811.75


def reference_mesh(steps):
    """The three GPU paths on the reference's 51,800-cell mesh, beside the
    thesis's C2050 time for the same mesh (other hardware, FP64)."""
    from gpueularsolver2d_amd import Solver, read_dat
    mesh = read_dat(REF_MESH_5W)
    C = mesh.ncell
    out = dict(mesh="2822-5w.dat (the reference's)", cells=C, steps=steps, flow=REF_FLOW,
               c2050_fp64=dict(ms_per_step=C2050_MS_PER_STEP,
                               mcsu_per_s=round(4.0 * C / C2050_MS_PER_STEP / 1e3, 2),
                               source="thesis Table 8; other hardware"))
    for name, kw in (("fp32", dict(precision="fp32")), ("fp64", dict(precision="fp64")),
                     ("fp64_plain", dict(precision="fp64", kernels="plain"))):
        s = Solver(mesh, **REF_FLOW, **kw)
        t, _ = timed_run(s, 10, steps, warm=Solver(mesh, **REF_FLOW, **kw))
        out[name] = dict(value=round(4.0 * C * steps / t / 1e6, 1), unit="Mcsu/s",
                         ms_per_step=round(1e3 * t / steps, 4),
                         x_c2050=round(C2050_MS_PER_STEP / (1e3 * t / steps), 1))
        del s
    return out


def cpu_reference_mesh(steps):
    """The one-thread oracle on the reference's 51,800-cell mesh, beside the
    thesis's one-thread CPU time for it (part of the CPU baseline)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from gpueularsolver2d_amd import read_dat
    from oracle_ctypes import LIB_PATH, OracleMesh, OracleSolver
    mesh = read_dat(REF_MESH_5W)
    om = OracleMesh(mesh.x, mesh.y, mesh.tri, library=LIB_PATH)
    s = OracleSolver(om, REF_FLOW["mach"], REF_FLOW["alpha_deg"], cfl=REF_FLOW["cfl"])
    s.time_steps(1)
    sec = s.time_steps(steps)
    return dict(cells=mesh.ncell, steps=steps, ms_per_step=round(1e3 * sec / steps, 2),
                value=round(4.0 * mesh.ncell * steps / sec / 1e6, 2), unit="Mcsu/s", cores=1,
                phenom9850_ms_per_step=PHENOM_MS_PER_STEP)


def perturbed(W, rng):
    return W * (1.0 + 1e-12 * rng.standard_normal(W.shape))


def window_compare(runs, start, end, every):
    """Advances every run (name -> object with run(n) and forces()) from step
    `start` to `end`, sampling Cl and Cd every `every` steps; returns per-run
    arrays (samples x cases) and the cases that stopped (diverged) in any run."""
    cl = {k: [] for k in runs}
    cd = {k: [] for k in runs}
    bad = None
    for n in range(start + every, end + 1, every):
        for k, r in runs.items():
            _, steps, st = r.run(n, check=False)
            c, d = r.forces()
            cl[k].append(np.atleast_1d(c))
            cd[k].append(np.atleast_1d(d))
            stopped = np.atleast_1d(st) != 0
            bad = stopped if bad is None else (bad | stopped)
    return ({k: np.array(v) for k, v in cl.items()}, {k: np.array(v) for k, v in cd.items()},
            bad)


def window_stats(cl, cd, ok, ref="fp64", others=("fp32", "fp64_perturbed")):
    """Per-case differences of the window means of Cl and Cd against `ref`,
    over the cases that ran through (ok), and of the last sample."""
    out = {}
    for k in others:
        dm = (cl[k].mean(axis=0) - cl[ref].mean(axis=0))[ok]
        dd = (cd[k].mean(axis=0) - cd[ref].mean(axis=0))[ok]
        inst = np.abs(cl[k][-1] - cl[ref][-1])[ok]
        out[k] = dict(cases=int(ok.sum()), within_gate=int(np.sum(np.abs(dm) < DCL_GATE)),
                      dcl_mean_abs_median=float(np.median(np.abs(dm))),
                      dcl_mean_abs_p95=float(np.percentile(np.abs(dm), 95)),
                      dcl_mean_abs_max=float(np.abs(dm).max()),
                      dcd_mean_abs_max=float(np.abs(dd).max()),
                      dcl_last_abs_median=float(np.median(inst)), dcl_last_abs_max=float(inst.max()))
    out["cl_fp64_std_median"] = float(np.median(cl[ref].std(axis=0)[ok]))
    return out


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
    B.set_state(perturbed(B.state(), np.random.default_rng(0)))
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


def batch_accuracy(meshes, case_mesh, flows, F, start, window, every):
    """The FP32 batch F (at step `start`) against the FP64 batch and the FP64
    batch from a 1e-12-perturbed start, both run to `start` first, then all
    three sampled over [start, start + window]."""
    from gpueularsolver2d_amd import Batch
    A = Batch(meshes, case_mesh, flows, precision="fp64")
    P = Batch(meshes, case_mesh, flows, precision="fp64")
    rng = np.random.default_rng(0)
    for i in range(len(case_mesh)):
        P.set_state(i, perturbed(P.state(i), rng))
    A.run(start, check=False)
    P.run(start, check=False)
    cl, cd, bad = window_compare(dict(fp32=F, fp64=A, fp64_perturbed=P), start, start + window,
                                 every)
    ok = ~bad
    out = dict(window=[start, start + window], every=every, diverged=[int(i) for i in np.flatnonzero(bad)])
    out.update(window_stats(cl, cd, ok))
    return out


def config_c1(steps=2000):
    """Config C1: the ~3k-triangle NACA 0012 mesh (seed 1), M 0.5, alpha 0,
    first order, no smoothing, CFL 1.5, 2000 steps.  The GPU FP64 run against
    the FP64 oracle on this host's cores (the same arithmetic): whether every
    cell's state is bitwise equal, the largest history difference, Cl, Cd."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ctypes import LIB_OMP_PATH, OracleMesh, OracleSolver
    from gpueularsolver2d_amd import Solver, workloads
    c = workloads.C1
    m = workloads.config_mesh(c)
    s = Solver(m, precision="fp64", **c["flow"], **c["scheme"])
    h, n, _ = s.run(steps)
    o = OracleSolver(OracleMesh(m.x, m.y, m.tri, library=LIB_OMP_PATH), c["flow"]["mach"],
                     c["flow"]["alpha_deg"], cfl=c["flow"]["cfl"], **c["scheme"])
    ho, _ = o.run(steps)
    cl, cd = s.forces()
    clo, cdo = o.forces()
    return dict(cells=m.ncell, steps=int(n), resd=float(h[-1]), cl=cl, cd=cd,
                state_bitwise_equal_to_oracle=bool(np.array_equal(s.state(), o.state())),
                hist_max_rel_diff=float(np.max(np.abs(h[1:] - ho[1:]) / np.abs(ho[1:]))),
                dcl_vs_oracle=cl - clo, dcd_vs_oracle=cd - cdo)


def config_c2(steps=5000, window=500, every=5):
    """Config C2: the ~50k-triangle NACA 0012 mesh (seed 2), M 0.8, alpha
    1.25, second order with the limiter and smoothing, CFL 1.5, 5000 steps.
    FP32 runs the launch shape the autotuner picks for this mesh (its table is
    reported); FP64 (tiled, = the oracle bit for bit) and the plain FP64
    kernels run the library's.  Accuracy: Cl and Cd averaged over the `window`
    steps after the timed ones, FP32 against FP64 beside FP64 from a
    perturbed start."""
    from gpueularsolver2d_amd import Solver, autotune, workloads
    c = workloads.C2
    m = workloads.config_mesh(c)
    flow = c["flow"]
    tuned = autotune.tune(m, precision="fp32", flow=flow, steps=40, write_cache=False)
    out = dict(cells=m.ncell, steps=steps,
               autotune=dict(chosen=dict(tile_cells=tuned["tile_cells"], threads=tuned["threads"]),
                             table=tuned["table"]))
    sol = {}
    for name, kw in (("fp64", dict(precision="fp64")),
                     ("fp32", dict(precision="fp32", tile_cells=tuned["tile_cells"],
                                   threads=tuned["threads"])),
                     ("fp32_library_shape", dict(precision="fp32")),
                     ("fp64_plain", dict(precision="fp64", kernels="plain"))):
        s = Solver(m, **kw, **flow)
        n_run = steps if name != "fp64_plain" else steps // 5
        s.run(10)
        t0 = time.perf_counter()
        h, n, _ = s.run(10 + n_run)
        t = time.perf_counter() - t0
        out[name] = dict(value=round(4.0 * m.ncell * n_run / t / 1e6, 1), unit="Mcsu/s",
                         steps=n_run, resd=float(h[-1]), launch=s.config())
        sol[name] = s
    A, F = sol["fp64"], sol["fp32"]
    # accuracy: the FP32 and FP64 runs continue for the window, beside FP64
    # started from a 1e-12-perturbed freestream and run as long
    P = Solver(m, precision="fp64", **flow)
    P.set_state(perturbed(P.state(), np.random.default_rng(0)))
    P.run(10 + steps)
    (cl, cd), (clF, cdF) = A.forces(), F.forces()
    WA, WF = A.state(), F.state()
    out.update(cl_fp64=cl, cd_fp64=cd, dcl=clF - cl, dcd=cdF - cd,
               rho_rel_l2=float(np.linalg.norm(WF[:, 0] - WA[:, 0]) / np.linalg.norm(WA[:, 0])))
    s0 = 10 + steps
    cls, cds, bad = window_compare(dict(fp32=F, fp64=A, fp64_perturbed=P), s0, s0 + window, every)
    w = dict(window=[s0, s0 + window], every=every)
    w.update(window_stats(cls, cds, ~bad))
    out["window_accuracy"] = w
    return out


def batch64(steps, window=500, every=5):
    """64 cases in one batch: four NACA 0012 meshes of uneven size (~3k, 10k,
    50k, 90k triangles, seeds 11-14) x 16 (Mach, alpha) pairs on the grid
    M in {0.5, 0.6167, 0.7333, 0.85} x alpha in {0, 1, 2, 3} deg; second order
    with the limiter and smoothing, CFL 1.5.  FP32 throughput over `steps`
    steps, then the per-case window accuracy.  Cases that diverge (non-finite
    residual, in FP64 as well) stop at that step; only the cell-stage updates
    done count."""
    from gpueularsolver2d_amd import Batch, workloads
    meshes, case_mesh, case_flow = workloads.batch64()
    cells = sum(meshes[m].ncell for m in case_mesh)
    t0 = time.perf_counter()
    F = Batch(meshes, case_mesh, case_flow, precision="fp32")
    tset = time.perf_counter() - t0
    F.run(20)
    t0 = time.perf_counter()
    _, nF, stF = F.run(20 + steps, check=False)
    t = time.perf_counter() - t0
    done = sum((int(nF[i]) - 20) * meshes[m].ncell for i, m in enumerate(case_mesh))
    acc = batch_accuracy(meshes, case_mesh, case_flow, F, 20 + steps, window, every)
    return dict(value=round(4.0 * done / t / 1e6, 1), unit="Mcsu/s", cases=64, cells=int(cells),
                mesh_cells=[m.ncell for m in meshes], steps=steps, launch=F.config(),
                setup_s=round(tset, 2), fp32_stopped=[int(i) for i in np.flatnonzero(stF != 0)],
                window_accuracy=acc)


def batch256(steps, window=500, every=5, workers=16):
    """256 cases in one batch, every case on its own mesh: sizes log-uniform in
    [2k, 150k] triangles (mesh seed 1000 + i), M uniform in [0.3, 0.85], alpha
    uniform in [-2, 4] deg (numpy seed 256), second order with the limiter and
    smoothing, CFL 1.5.  The autotuner times launch shapes on the batch and the
    fastest runs `steps` timed FP32 steps (all state resident in HBM); the
    stage kernel's time gives the roofline fractions; then the per-case
    window accuracy."""
    from gpueularsolver2d_amd import Batch, autotune, workloads
    t0 = time.perf_counter()
    meshes, flows = workloads.batch256(workers)
    tgen = time.perf_counter() - t0
    case_mesh = list(range(256))
    cells = sum(m.ncell for m in meshes)
    t0 = time.perf_counter()
    tuned = autotune.tune_batch(meshes, case_mesh, flows, precision="fp32", steps=10)
    ttune = time.perf_counter() - t0
    F = Batch(meshes, case_mesh, flows, precision="fp32", tile_cells=tuned["tile_cells"],
              threads=tuned["threads"])
    F.run(10)
    t0 = time.perf_counter()
    _, nF, stF = F.run(10 + steps, check=False)
    t = time.perf_counter() - t0
    ms_region, _ = F.timing()
    kt = F.kernel_times(reps=10)
    done = sum((int(nF[i]) - 10) * meshes[i].ncell for i in range(256))
    t_kern = ms_region / (4 * steps)
    roof = roofline(cells, t_kern, f"batch256-{cells}")
    roof["kernel_ms_isolated"] = kt[0]
    acc = batch_accuracy(meshes, case_mesh, flows, F, 10 + steps, window, every)
    return dict(value=round(4.0 * done / t / 1e6, 1), unit="Mcsu/s", cases=256, cells=int(cells),
                steps=steps, launch=F.config(),
                autotune=dict(chosen=dict(tile_cells=tuned["tile_cells"], threads=tuned["threads"]),
                              table=tuned["table"], seconds=round(ttune, 1)),
                roofline=roof, fp32_stopped=[int(i) for i in np.flatnonzero(stF != 0)],
                window_accuracy=acc, mesh_generation_s=round(tgen, 1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--acc-window", type=lambda v: [int(x) for x in v.split(":")],
                    default=[30000, 70000, 20],
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
    ap.add_argument("--window", type=int, default=500,
                    help="steps of the per-case accuracy window of C2 and the batches")
    ap.add_argument("--refmesh-steps", type=int, default=1000,
                    help="steps of the GPU paths on the reference's 51,800-cell mesh (0 skips it)")
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
    ap.add_argument("--mesh", choices=["c3", "2822-5w"], default="c3",
                    help="mesh of the --fp32-only, --fp64-only and --plain-only modes: the "
                         "bench mesh, or the reference's 51,800-cell mesh at its printed case")
    ap.add_argument("--plain-only", action="store_true",
                    help="time only the FP64 plain path, the reference's kernel structure "
                         "(for profiler runs)")
    ap.add_argument("--batch256-fp32-only", action="store_true",
                    help="only run the FP32 256-case batch for --batch256-steps steps at the "
                         "library's launch shape (for profiler runs)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="only run one device copy kernel of this many MiB and exit (checks the "
                         "traffic counters against a known size)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the solver has no multi-GPU path "
              "(independent cases would be split across GPUs by case)", file=sys.stderr)
        return 2
    if args.batch256_fp32_only:
        from gpueularsolver2d_amd import Batch, workloads
        meshes, flows = workloads.batch256()
        b = Batch(meshes, list(range(256)), flows, precision="fp32", tile_cells=args.tile_cells,
                  threads=args.threads)
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

    from gpueularsolver2d_amd import Solver, read_dat, workloads
    cfg = workloads.C3
    mesh = workloads.config_mesh(cfg)
    flow = cfg["flow"]
    if args.mesh == "2822-5w":
        if not (args.plain_only or args.fp32_only or args.fp64_only):
            print("--mesh 2822-5w applies to the --fp32-only, --fp64-only and --plain-only modes",
                  file=sys.stderr)
            return 2
        mesh, flow = read_dat(REF_MESH_5W), REF_FLOW
    C = mesh.ncell
    csu = 4.0 * C * args.steps

    if args.plain_only:
        sp = Solver(mesh, precision="fp64", kernels="plain", **flow)
        tpl, _ = timed_run(sp, args.warmup, args.steps,
                           warm=Solver(mesh, precision="fp64", kernels="plain", **flow))
        print(json.dumps({"metric": METRIC, "value": round(csu / tpl / 1e6, 1), "unit": "Mcsu/s",
                          "precision": "fp64", "kernels": "plain", "steps": args.steps}))
        return 0
    if args.fp64_only:
        s64 = Solver(mesh, precision="fp64", kernels="tiled", **flow)
        t64, _ = timed_run(s64, args.warmup, args.steps,
                           warm=Solver(mesh, precision="fp64", kernels="tiled", **flow))
        ms64_region, _ = s64.timing()
        kt64 = s64.kernel_times(reps=20)
        print(json.dumps({"metric": METRIC, "value": round(csu / t64 / 1e6, 1), "unit": "Mcsu/s",
                          "precision": "fp64", "steps": args.steps,
                          "kernel_ms": ms64_region / (4 * args.steps), "kernel_ms_isolated": kt64[0]}))
        return 0
    s32 = Solver(mesh, precision="fp32", kernels="tiled", tile_cells=args.tile_cells,
                 threads=args.threads, **flow)
    t32, _ = timed_run(s32, args.warmup, args.steps,
                       warm=Solver(mesh, precision="fp32", kernels="tiled", tile_cells=args.tile_cells,
                                   threads=args.threads, **flow))
    ms_region, _ = s32.timing()  # device time of the timed steps, solver stream
    kt = s32.kernel_times(reps=50)
    v32 = csu / t32 / 1e6
    progress(f"fp32 {v32:.1f} Mcsu/s")
    if args.fp32_only:
        print(json.dumps({"metric": METRIC, "value": round(v32, 1), "unit": "Mcsu/s",
                          "steps": args.steps, "kernel_ms": kt}))
        return 0
    s64 = Solver(mesh, precision="fp64", kernels="tiled", **flow)
    t64, _ = timed_run(s64, args.warmup, args.steps,
                       warm=Solver(mesh, precision="fp64", kernels="tiled", **flow))
    ms64_region, _ = s64.timing()
    kt64 = s64.kernel_times(reps=20)
    plain_steps = max(10, args.steps // 4)
    sp = Solver(mesh, precision="fp64", kernels="plain", **flow)
    tpl, _ = timed_run(sp, 5, plain_steps,
                       warm=Solver(mesh, precision="fp64", kernels="plain", **flow))
    v64 = csu / t64 / 1e6
    vpl = 4.0 * C * plain_steps / tpl / 1e6
    progress(f"fp64 tiled {v64:.1f}, fp64 plain {vpl:.1f} Mcsu/s")
    # the global time step (kernel_TimeStep's mode): one kernel per stage, each
    # launch taking the case's minimum of the previous stage's tile minima
    gdt = {}
    gsteps = max(10, args.steps // 4)
    for prec in ("fp32", "fp64"):
        sg = Solver(mesh, precision=prec, global_dt=True, **flow)
        tg, _ = timed_run(sg, 5, gsteps, warm=Solver(mesh, precision=prec, global_dt=True, **flow))
        kt_g = sg.kernel_times(reps=10)
        gdt[prec] = dict(value=round(4.0 * C * gsteps / tg / 1e6, 1), steps=gsteps,
                         stage_tile_ms=kt_g[0])

    # the stage kernel's average launch time over the timed steps: the device
    # time of the steps (HIP events on the solver's stream) over their launches
    t_kern = ms_region / (4 * args.steps)
    roof = roofline(C, t_kern, f"naca0012-{C}")
    roof.update(kernel_ms_isolated=kt[0],
                step_gbs_model=round(model_bytes(4) * v32 * 1e6 / 1e9, 1))

    c1 = config_c1() if not args.no_configs else None
    c2 = config_c2(window=args.window) if not args.no_configs else None
    progress("configs C1, C2 done")
    acc = limit_cycle_accuracy(mesh, flow, *args.acc_window) if args.acc_window[1] > 0 else None
    progress("limit-cycle accuracy done")
    conv = converged_accuracy(mesh, args.conv_steps) if args.conv_steps > 0 else None
    progress("converged-case accuracy done")
    b64 = batch64(args.batch_steps, window=args.window) if args.batch_steps > 0 else None
    progress("64-case batch done")
    b256 = batch256(args.batch256_steps, window=args.window) if args.batch256_steps > 0 else None
    progress("256-case batch done")
    refm = reference_mesh(args.refmesh_steps) if args.refmesh_steps > 0 else None
    progress("reference mesh done")
    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        cpu = cpu_baseline(mesh, args.cpu_steps, args.cpu_omp_steps)
        if args.refmesh_steps > 0:
            cpu["reference_mesh_2822_5w"] = cpu_reference_mesh(20)
        progress("cpu baseline done")

    out = {
        "metric": METRIC, "value": round(v32, 1), "unit": "Mcsu/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * t32 / args.steps,
        "device_warmup_s": DEVICE_WARMUP_S,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"naca0012-{C} (~200k-triangle NACA 0012, seed {cfg['seed']}), "
                               "M 0.8, alpha 1.25, 2nd order + limiter + smoothing, 4-stage RK, CFL 1.5",
                   "cells": C, "precision": "fp32 bulk, fp64 geometry + reductions",
                   "kernels": "tiled", "launch": s32.config()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "reference_mesh_2822_5w": refm,
        "fp64_tiled": {"value": round(v64, 1), "kernel_ms": ms64_region / (4 * args.steps),
                       "kernel_ms_isolated": kt64[0]},
        "fp64_plain": {"value": round(vpl, 1), "steps": plain_steps},
        "fp32_over_fp64_plain": round(v32 / vpl, 2),
        "global_dt": gdt,
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
