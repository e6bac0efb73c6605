/*
 * fvm_abi.h -- C ABI of the MI355X explicit finite-volume Euler solver
 * (libfvm_hip.so).
 *
 * This is the drop-in boundary for the solve path of GPUEularSolver2D.  In
 * the reference, main() (gpu_fvm.cu:27-68) calls inputMesh (gpu_init.cu:5),
 * Initialization (gpu_init.cu:63), CalculationMeshGeo (gpu_init.cu:100),
 * Cuda_Init (gpu_init.cu:249) and ROE_SOLVER (gpu_solver.cu:2171) over global
 * state.  Here the host keeps those entry points (gpueularsolver2d_amd/host/)
 * and ROE_SOLVER becomes: pack the globals into descriptors, fvm_create,
 * fvm_run, copy the residual history and the final state back.
 *
 * Conventions
 *  - Plain C types only; every array is caller-owned and is copied on entry.
 *  - Node ids in fvm_mesh_desc.tri are 1-based and triangles are in file
 *    order, exactly as inputMesh reads them (gpu_init.cu:16-32).  Per-cell
 *    outputs are in that same cell order, 0-based arrays.
 *  - Every call returns an fvm_status; nothing calls exit().  The reference
 *    calls exit(0) on a missing file (gpu_init.cu:10-14), only prints a
 *    warning for a non-positive cell area (gpu_init.cu:112) and stops silently
 *    on a NaN residual (gpu_solver.cu:2200); these become FVM_ERR_MESH,
 *    FVM_ERR_AREA and FVM_ERR_NAN.
 *  - A handle is not thread-safe; use one handle per host thread.  All work
 *    of a handle runs on one HIP stream of its own.
 *  - Results are deterministic: no floating-point atomics anywhere.
 */
#ifndef FVM_ABI_H
#define FVM_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum fvm_status {
    FVM_OK = 0,
    FVM_ERR_ARGUMENT = 1,      /* NULL pointer, bad count or option */
    FVM_ERR_MESH = 2,          /* node id out of range, non-manifold edge, unreadable file */
    FVM_ERR_AREA = 3,          /* a cell with area <= 0 (clockwise or degenerate) */
    FVM_ERR_HIP = 4,           /* HIP runtime error (fvm_last_error has the text) */
    FVM_ERR_NAN = 5,           /* NaN residual: the case diverged */
    FVM_ERR_NOT_CONVERGED = 6, /* reserved: fvm_run reports max_steps exhaustion via steps_done */
    FVM_ERR_NO_DEVICE = 7,     /* no HIP device, or the requested ordinal does not exist */
    FVM_ERR_ALLOC = 8          /* host or device allocation failed */
} fvm_status;

/* Mesh as read by inputMesh (gpu_init.cu:5-33). */
typedef struct fvm_mesh_desc {
    int32_t nnode, ncell;
    const double *x;      /* [nnode] node x */
    const double *y;      /* [nnode] node y */
    const int32_t *tri;   /* [3*ncell] 1-based node ids, counter-clockwise */
} fvm_mesh_desc;

/* The values of input.dat (gpu_init.cu:41) that the solve path uses:
 * MA ALPHA(deg) GAMA PIN TIN CFL.  EPSL and MAXSTEP are fvm_run arguments. */
typedef struct fvm_flow_params {
    double mach, alpha_deg, gamma, p_in, t_in, cfl;
} fvm_flow_params;

typedef enum fvm_precision {
    FVM_FP64 = 0, /* every quantity in FP64 (bitwise equal to the FP64 oracle) */
    FVM_FP32 = 1  /* state, flux, gradient, limiter in FP32; geometry precomputed in FP64;
                     residual and force reductions in FP64 */
} fvm_precision;

typedef enum fvm_kernels {
    FVM_KERNELS_TILED = 0, /* one kernel per RK stage over LDS-staged cell tiles (default), with
                              the local or the global time step */
    FVM_KERNELS_PLAIN = 1, /* the reference's kernel structure: one kernel per phase, edge
                              slots in global memory (FP64 only) */
    FVM_KERNELS_SPLIT = 2  /* two kernels per RK stage (flux, then smoothing + update) over the
                              same tiles */
} fvm_kernels;

/* Scheme options.  The reference hard-codes second order with the limiter,
 * smoothing with eps = 0.8 (gpu_solver.cu:1633) and the local time step;
 * fvm_default_opts sets exactly those. */
typedef struct fvm_scheme_opts {
    int32_t order;          /* 1: first order; 2: Green-Gauss gradient + per-face limiter */
    int32_t smoothing;      /* 0/1: two Jacobi sweeps of implicit residual smoothing */
    double eps_smooth;      /* smoothing coefficient */
    int32_t global_dt;      /* 0: local time step; 1: global minimum time step */
    int32_t precision;      /* fvm_precision */
    int32_t kernels;        /* fvm_kernels */
    int32_t tile_cells;     /* cells per workgroup tile, 0 = chosen by the library */
    int32_t threads;        /* threads per workgroup, 0 = 512 (one kernel per stage) or 256 */
    int32_t steps_per_graph;/* time steps per captured hipGraph, 0 = automatic (50), at most 1024 */
    int32_t device;         /* HIP device ordinal */
} fvm_scheme_opts;

typedef struct fvm_solver fvm_solver;
typedef struct fvm_batch fvm_batch;

/* Fills the reference's defaults: order 2, smoothing on, eps 0.8, local dt,
 * FP64, tiled kernels, autotuned launch shape, device 0. */
void fvm_default_opts(fvm_scheme_opts *opts);

const char *fvm_status_string(fvm_status s);
/* Text of the last error raised on the calling thread (empty if none). */
const char *fvm_last_error(void);

/* Host-only: the reference's edge/cell topology (CalculationMeshGeo,
 * gpu_init.cu:100-228, followed by the interior/far/wall edge reorder of
 * changeEdge, gpu_solver.cu:1832-1906).  counts = {nnode, ncell, nedge,
 * nwall, nfar}.  Array outputs may be NULL (call once with NULLs to size
 * them).  Per cell [3*ncell]: celledge (1-based edge id), neicell (cell id,
 * -1 far field, -2 wall); clog [ncell].  Per edge [nedge]: elog (0 interior,
 * 1 far, 2 wall), left, right (-1 for boundary), farid, wallid, node1, node2.
 * All ids 1-based as in the reference. */
fvm_status fvm_mesh_topology(const fvm_mesh_desc *mesh, int32_t counts[5], int32_t *celledge,
                             int32_t *neicell, int32_t *clog, int32_t *elog, int32_t *left,
                             int32_t *right, int32_t *farid, int32_t *wallid, int32_t *node1,
                             int32_t *node2);

/* One flow case on one mesh.  Replaces ROE_SOLVER's set-up (gpu_solver.cu:
 * 2184-2199: allocation, changeEdge, struct2Array, copyH2D) and Initialization
 * (freestream state, gpu_init.cu:63-93). */
fvm_status fvm_create(const fvm_mesh_desc *mesh, const fvm_flow_params *flow,
                      const fvm_scheme_opts *opts, fvm_solver **out);

/* The iteration loop of ROE_SOLVER (gpu_solver.cu:2200-2434): runs while
 * step < max_steps and Resd[step] > epsl, continuing from the current step.
 * resid_hist (may be NULL) receives Resd[0..steps] and must hold
 * max_steps+1 doubles; resid_hist[0] = 100 as in the reference.
 * steps_done receives the step count reached.  Returns FVM_ERR_NAN if the
 * residual became NaN (the case diverged).  A residual of +inf is not an
 * error: the reference's normalisation by step 1 gives it when step 1 changed
 * nothing, and its loop goes on. */
fvm_status fvm_run(fvm_solver *s, int32_t max_steps, double epsl, double *resid_hist,
                   int32_t *steps_done);

/* fvm_run without allocating: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, results
 * and status codes as fvm_run, bit for bit, with two differences:
 *  - the step closes' per-tile partial sums and the device copy of the
 *    residual history live in a scratch array of FVM_NOALLOC_SCRATCH_BYTES
 *    that the library holds on each device (set up by fvm_create), so a run
 *    may cover at most the steps that fit:
 *      256 + 8 * tiles (rounded up to 256) + 8 * (max_steps + 1) bytes;
 *    a larger request returns FVM_ERR_ARGUMENT and runs nothing;
 *  - the steps are launched directly on the solver's stream, without graphs
 *    (instantiating a graph allocates).
 * Every call on a device shares that device's scratch and keeps the last
 * call's contents there; calls that overlap on one device race, so run one at
 * a time per device.  Nothing in the scratch needs to start at zero. */
#define FVM_NOALLOC_SCRATCH_BYTES (4u << 20)
fvm_status fvm_run_noalloc(fvm_solver *s, int32_t max_steps, double epsl, double *resid_hist,
                           int32_t *steps_done);

/* Final conservative state per cell (reference cell order), FP64. */
fvm_status fvm_get_state(fvm_solver *s, double *rho, double *rhou, double *rhov, double *rhoE);
/* Replaces the state (checkpoint restore); the step count is kept. */
fvm_status fvm_set_state(fvm_solver *s, const double *rho, const double *rhou, const double *rhov,
                         const double *rhoE);

/* Lift and drag coefficients of the current state.  Not computed by the
 * reference: force = sum over wall edges (wall-id order) of the first-order
 * Roe wall-flux momentum times the edge length, minus p_inf * sum(n ds),
 * normalised by 0.5 rho_inf V_inf^2 c = 0.5; FP64 compensated sums. */
fvm_status fvm_get_forces(fvm_solver *s, double *cl, double *cd);

/* Device time of the last fvm_run (HIP events on the solver's stream) and the
 * cell-stage updates it performed (sum over active steps of 4 * ncell). */
fvm_status fvm_get_timing(fvm_solver *s, double *ms, int64_t *cell_stage_updates);

void fvm_destroy(fvm_solver *s);

/* Many flow cases on several meshes of uneven size, advanced together: every
 * kernel launch covers the tiles of all cases, with per-case offset tables,
 * per-case flow constants and per-case convergence.  case_mesh[i] selects the
 * mesh of case i.  Each case stops at its own step exactly as a single
 * fvm_run would. */
fvm_status fvm_batch_create(int32_t nmesh, const fvm_mesh_desc *meshes, int32_t ncase,
                            const int32_t *case_mesh, const fvm_flow_params *flows,
                            const fvm_scheme_opts *opts, fvm_batch **out);
/* resid_hist (may be NULL): ncase rows of max_steps+1; steps_done: [ncase];
 * case_status (may be NULL): [ncase] fvm_status of each case. */
fvm_status fvm_batch_run(fvm_batch *b, int32_t max_steps, double epsl, double *resid_hist,
                         int32_t *steps_done, int32_t *case_status);
fvm_status fvm_batch_get_state(fvm_batch *b, int32_t icase, double *rho, double *rhou,
                               double *rhov, double *rhoE);
/* Replaces case icase's state (checkpoint restore), as fvm_set_state does for
 * a single solver; its step count is kept. */
fvm_status fvm_batch_set_state(fvm_batch *b, int32_t icase, const double *rho, const double *rhou,
                               const double *rhov, const double *rhoE);
fvm_status fvm_batch_get_forces(fvm_batch *b, double *cl, double *cd); /* [ncase] each */
fvm_status fvm_batch_get_timing(fvm_batch *b, double *ms, int64_t *cell_stage_updates);
/* Average device time of one launch of each tiled kernel of an RK stage, from
 * HIP events on the batch's stream around `reps` back-to-back launches.  One
 * kernel per stage: ms[0] = stage_tile (smoothing + update + gradient + flux
 * + gather), ms[1] = 0.  Two kernels per stage: ms[0] = flux_tile (gradient
 * + flux + gather), ms[1] = update_tile (smoothing + RK update).  The
 * launches use the current state and leave it unchanged.  No counterpart in the reference (its per-phase wall-clock timers,
 * gpu_solver.cu:2214-2423, play this role).  Tiled kernels only. */
fvm_status fvm_batch_kernel_times(fvm_batch *b, int32_t reps, double ms[2]);
fvm_status fvm_kernel_times(fvm_solver *s, int32_t reps, double ms[2]);
/* Launch shape in use: {tile_cells, threads, steps_per_graph, total tiles,
 * workgroups of the tile kernel resident per CU, CUs of the device}
 * (the last two are 0 for the plain kernels). */
fvm_status fvm_batch_get_config(fvm_batch *b, int32_t config[6]);
fvm_status fvm_get_config(fvm_solver *s, int32_t config[6]);
void fvm_batch_destroy(fvm_batch *b);

#ifdef __cplusplus
}
#endif
#endif
