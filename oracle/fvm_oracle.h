/*
 * fvm_oracle.h -- FP64 CPU restatement of the GPUEularSolver2D explicit
 * finite-volume Euler path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker that the HIP solver is
 * compared against; it is never linked into, loaded by, or called from the
 * product library.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it.
 *
 * Every function restates one piece of the reference (file:line cited at each
 * definition in fvm_oracle.c).  Arithmetic follows the reference's CUDA kernels
 * operation by operation (same association, reciprocal-multiply where the
 * kernel multiplies by a reciprocal) and is compiled with -ffp-contract=off, so
 * that a correctly rounded FP64 implementation of the same sequence reproduces
 * it bit for bit.
 *
 * Conventions follow the reference: cells, nodes and edges are 1-based; index 0
 * of every array is unused.
 */
#ifndef FVM_ORACLE_H
#define FVM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mesh {
    int nnode, ncell, nedge, nwall, nfar;
    int n_bad_area;           /* cells with area <= 0 (reference only warns) */
    double *x, *y;            /* [nnode+1] */
    int *pt;                  /* [3*(ncell+1)]  Point[i] of cell c at pt[3c+i] */
    double *area, *cx, *cy;   /* [ncell+1] */
    int *celledge, *neicell;  /* [3*(ncell+1)] neighbor[i] of cell c */
    int *clog;                /* [ncell+1] number of interior neighbours */
    int *elog, *left, *right, *farid, *wallid, *node1, *node2; /* [nedge+1] */
    double *vx, *vy, *vn, *midx, *midy;                         /* [nedge+1] */
    int *wall_edge;           /* [nwall+1] edge id (after reorder) of wall id w */
} orc_mesh;

typedef struct orc_flow {
    double mach, alpha_deg, gamma, p_in, t_in, cfl;
} orc_flow;

typedef struct orc_opts {
    int order;        /* 1 = first order, 2 = limited linear reconstruction */
    int smoothing;    /* 0/1: implicit residual smoothing, two Jacobi sweeps */
    double eps;       /* smoothing coefficient (reference: 0.8) */
    int global_dt;    /* 0 = local time step, 1 = global minimum time step */
} orc_opts;

typedef struct orc_solver orc_solver;

/* status codes */
#define ORC_OK 0
#define ORC_ERR_ALLOC (-1)
#define ORC_ERR_INDEX (-2)
#define ORC_ERR_NONMANIFOLD (-4)
#define ORC_ERR_FILE (-5)

/* Mesh: x,y are 0-based arrays of nnode coordinates; tri is 3*ncell 1-based
 * node ids in file order.  literal=1 runs the reference's own O(C*E + C^2)
 * search; literal=0 an equivalent sort-based build (rejects non-manifold
 * edges).  Both then apply the reference's interior/far/wall edge reorder. */
int orc_mesh_build(int nnode, int ncell, const double *x, const double *y,
                   const int *tri, int literal, orc_mesh **out);
int orc_mesh_read_dat(const char *path, orc_mesh **out, int literal);
void orc_mesh_free(orc_mesh *m);
/* Copies the topology into caller arrays (all sized as in orc_mesh, 1-based,
 * any pointer may be NULL).  counts = {nnode, ncell, nedge, nwall, nfar}. */
void orc_mesh_counts(const orc_mesh *m, int counts[5]);
void orc_mesh_export(const orc_mesh *m, int *celledge, int *neicell, int *clog,
                     int *elog, int *left, int *right, int *farid, int *wallid,
                     int *node1, int *node2, double *area, double *cx, double *cy,
                     double *vx, double *vy, double *vn, double *midx, double *midy);

orc_solver *orc_create(const orc_mesh *m, const orc_flow *f, const orc_opts *o);
void orc_destroy(orc_solver *s);
/* Runs until step == maxstep or resd <= epsl, continuing from the current
 * state and step count.  resid[step] gets each step's normalised residual
 * (resid must hold maxstep+1 doubles; resid[0] = 100).  Returns steps run. */
int orc_run(orc_solver *s, int maxstep, double epsl, double *resid);
int orc_step_count(const orc_solver *s);
/* conservative state in reference cell order, 0-based arrays of ncell */
void orc_get_state(const orc_solver *s, double *rho, double *rhou, double *rhov, double *rhoe);
void orc_set_state(orc_solver *s, const double *rho, const double *rhou, const double *rhov,
                   const double *rhoe);
void orc_forces(const orc_solver *s, double *cl, double *cd);
/* freestream quantities after Initialization: {UI, VI, AIN, VIN, ROUIN, PIS, rhoE_inf} */
void orc_freestream(const orc_solver *s, double out[7]);

/* One residual evaluation at the current state (ghosts, gradient, flux, SR,
 * edge-to-cell gather), no smoothing.  Outputs are 0-based per cell:
 * res[4*ncell] (component-major per cell), sr[ncell] (sum of spectral radii),
 * dt[ncell], grad[8*ncell] (d/dx 4 comps then d/dy 4 comps), umax/umin[4*ncell]. */
void orc_eval_residual(orc_solver *s, double *res, double *sr, double *dt, double *grad,
                       double *umax, double *umin);
/* Applies the two smoothing sweeps to res (0-based, 4 per cell) in place. */
void orc_smooth(orc_solver *s, double *res);

/* Point kernels (for known-answer tests). */
void orc_pta(double gamma, double R, const double w[4], double pta[3]);
void orc_roe(const double wl[4], const double wr[4], double nx, double ny, double gamma,
             double flux[4]);
void orc_far_ghost(const orc_solver *s, const double wd[4], double vx, double vy, double out[4]);
void orc_wall_ghost(const double wd[4], double vx, double vy, double vn, double out[4]);
void orc_limiter(const double u[4], const double umax[4], const double umin[4], double delt[4]);
double orc_spectral_radius(const double wl[4], const double wr[4], double vx, double vy,
                           double gamma);

/* Wall-clock of `steps` oracle steps (for the CPU baseline); returns seconds. */
double orc_time_steps(orc_solver *s, int steps);
/* Threads of the per-cell / per-edge loops (1 unless built with -fopenmp). */
int orc_threads(void);

#ifdef __cplusplus
}
#endif
#endif
