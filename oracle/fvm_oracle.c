/*
 * fvm_oracle.c -- FP64 CPU restatement of the GPUEularSolver2D explicit
 * finite-volume Euler path.  TEST INFRASTRUCTURE: the checker for the HIP
 * solver, never part of the product path.
 *
 * Pinning status of each part (how a plausible mistake here would be caught):
 *   - mesh topology (orc_mesh_build): the literal O(C*E) search is pinned by
 *     the reference meshes' edge / wall / far counts and by per-cell closure
 *     (sum of n*ds = 0); the sort-based build is pinned against the literal one.
 *   - point kernels (PTA, Roe, ghosts, limiter, spectral radius): closed forms
 *     and invariants in tests/test_oracle_kats.py.
 *   - gradient, gather, smoothing: constant-field identities, telescoping
 *     conservation and brute-force hand meshes in tests/test_oracle_invariants.py.
 *   - the whole solver: reproduction of the reference's own residual history
 *     files where the flow parameters are recoverable (tests/test_oracle_reference_history.py).
 *   - forces (Cl/Cd): not computed by the reference; pinned only by the
 *     closed-contour identity and symmetry -- parity unpinned against the
 *     reference.
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math (oracle/Makefile).
 */
#include "fvm_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Reference constants: gpu_fvm.cu:8-10, gpu_solver.cu:603 (Harten delta),
 * gpu_solver.cu:831 (dt cap), gpu_solver.cu:34 (block size for the residual
 * partial sums). */
static const double ORC_R = 287.06;
static const double ORC_PI = 3.14159265358979;
static const double ORC_RK[4] = {0.0833, 0.2069, 0.4265, 1.0};
#define ORC_BLOCK 64

typedef struct orc_w {
    double d, du, dv, de, p, t, a;
} orc_w;

struct orc_solver {
    const orc_mesh *m;
    orc_flow flow;
    orc_opts opts;
    double GAMA, R, TIN, AIN, VIN, ROUIN, UI, VI, PIS, CFL, ALPHA, rhoE_inf;
    orc_w *w, *w0, *farw, *wallw;
    double *deltU; /* 8*(C+1): [c*8 + j*4 + k] j=0 d/dx, j=1 d/dy */
    double *umax, *umin; /* 4*(C+1) */
    double *rel, *rer;   /* 4*(E+1) per-edge flux slots (left / right) */
    double *dtl, *dtr;   /* E+1 per-edge spectral radius slots */
    double *res, *r1;    /* 4*(C+1) */
    double *dt, *tsum;   /* C+1 */
    int step;
    double epsm;
};

/* ------------------------------------------------------------------------ */
/* Mesh: gpu_init.cu:95-228 (geometry + topology), gpu_solver.cu:1832-1906   */
/* (edge reorder).                                                          */
/* ------------------------------------------------------------------------ */

/* gpu_init.cu:95-98 */
static double calc_area(double x1, double y1, double x2, double y2, double x3, double y3)
{
    return (0.5 * ((y3 - y1) * (x2 - x1) - (y2 - y1) * (x3 - x1)));
}

static void *xcalloc(size_t n, size_t sz)
{
    return calloc(n ? n : 1, sz);
}

void orc_mesh_free(orc_mesh *m)
{
    if (!m) return;
    free(m->x); free(m->y); free(m->pt); free(m->area); free(m->cx); free(m->cy);
    free(m->celledge); free(m->neicell); free(m->clog);
    free(m->elog); free(m->left); free(m->right); free(m->farid); free(m->wallid);
    free(m->node1); free(m->node2); free(m->vx); free(m->vy); free(m->vn);
    free(m->midx); free(m->midy); free(m->wall_edge);
    free(m);
}

/* Creates edge `ne` from cell icell's local edge (N1, N2), gpu_init.cu:147-158. */
static void make_edge(orc_mesh *m, int ne, int icell, int N1, int N2)
{
    m->left[ne] = icell;
    m->node1[ne] = N1;
    m->node2[ne] = N2;
    m->vx[ne] = m->x[N1] - m->x[N2];
    m->vy[ne] = m->y[N2] - m->y[N1];
    m->midx[ne] = 0.5 * (m->x[N1] + m->x[N2]);
    m->midy[ne] = 0.5 * (m->y[N1] + m->y[N2]);
    m->vn[ne] = sqrt(m->vx[ne] * m->vx[ne] + m->vy[ne] * m->vy[ne]);
}

/* Boundary classification of a new edge with no right cell, gpu_init.cu:187-204. */
static void classify_boundary(orc_mesh *m, int ne)
{
    m->right[ne] = -1;
    if (fabs(m->midx[ne]) < 2.0 && fabs(m->midy[ne]) < 2.0) {
        m->elog[ne] = 2;
        m->nwall++;
        m->wallid[ne] = m->nwall;
    } else {
        m->elog[ne] = 1;
        m->nfar++;
        m->farid[ne] = m->nfar;
    }
}

/* neighbor[] and CLog of one cell, gpu_init.cu:209-222. */
static void cell_neighbors(orc_mesh *m, int icell)
{
    int nn = 3;
    for (int i = 0; i < 3; i++) {
        int e = m->celledge[3 * icell + i];
        int nc = (m->left[e] == icell) ? m->right[e] : m->left[e];
        if (nc == -1) {
            nc *= m->elog[e];
            nn--;
        }
        m->neicell[3 * icell + i] = nc;
    }
    m->clog[icell] = nn;
}

/* Literal restatement of CalculationMeshGeo's search, gpu_init.cu:121-223:
 * linear search of existing edges, then linear search of later cells for the
 * right neighbour.  O(C*E + C^2); for small meshes and for pinning. */
static int topology_literal(orc_mesh *m)
{
    static const int IP[3] = {1, 2, 0};
    int C = m->ncell;
    for (int icell = 1; icell <= C; icell++) {
        for (int i = 0; i < 3; i++) {
            int ie1 = IP[i];
            int ie2 = IP[ie1];
            int N1 = m->pt[3 * icell + ie1];
            int N2 = m->pt[3 * icell + ie2];
            int found = 0, iedge;
            for (iedge = 1; iedge <= m->nedge; iedge++) {
                if ((m->node1[iedge] == N1 && m->node2[iedge] == N2) ||
                    (m->node1[iedge] == N2 && m->node2[iedge] == N1)) {
                    found = 1;
                    break;
                }
            }
            m->celledge[3 * icell + i] = iedge;
            if (!found) {
                m->nedge++;
                make_edge(m, m->nedge, icell, N1, N2);
                int findcell = 0, ci;
                for (ci = icell + 1; ci <= C; ci++) {
                    for (int j = 0; j < 3; j++) {
                        int je1 = IP[j];
                        int je2 = IP[je1];
                        int NN1 = m->pt[3 * ci + je1];
                        int NN2 = m->pt[3 * ci + je2];
                        if ((NN1 == N1 && NN2 == N2) || (NN1 == N2 && NN2 == N1)) {
                            findcell = 1;
                            break;
                        }
                    }
                    if (findcell) break;
                }
                if (findcell) {
                    m->right[iedge] = ci;
                    m->elog[iedge] = 0;
                } else {
                    classify_boundary(m, iedge);
                }
            }
        }
        cell_neighbors(m, icell);
    }
    return ORC_OK;
}

typedef struct {
    uint64_t key;
    int cell, loc;
} orc_half;

static int half_cmp(const void *a, const void *b)
{
    const orc_half *p = (const orc_half *)a, *q = (const orc_half *)b;
    if (p->key != q->key) return p->key < q->key ? -1 : 1;
    if (p->cell != q->cell) return p->cell < q->cell ? -1 : 1;
    return p->loc - q->loc;
}

/* Sort-based build with the same result as topology_literal on manifold
 * meshes: an edge is created by the lowest-numbered cell containing it, in
 * (cell, local edge) scan order; its right cell is the other cell. */
static int topology_sorted(orc_mesh *m)
{
    static const int IP[3] = {1, 2, 0};
    int C = m->ncell;
    size_t nh = (size_t)3 * C;
    orc_half *h = (orc_half *)xcalloc(nh, sizeof(orc_half));
    int *grp = (int *)xcalloc(nh, sizeof(int));      /* (cell,loc) -> group */
    int *gfirst = (int *)xcalloc(nh, sizeof(int));   /* group -> first index in h */
    int *gsize = (int *)xcalloc(nh, sizeof(int));
    int *gedge = (int *)xcalloc(nh, sizeof(int));
    if (!h || !grp || !gfirst || !gsize || !gedge) {
        free(h); free(grp); free(gfirst); free(gsize); free(gedge);
        return ORC_ERR_ALLOC;
    }
    for (int c = 1; c <= C; c++)
        for (int i = 0; i < 3; i++) {
            int N1 = m->pt[3 * c + IP[i]];
            int N2 = m->pt[3 * c + IP[IP[i]]];
            uint64_t lo = (uint64_t)(N1 < N2 ? N1 : N2), hi = (uint64_t)(N1 < N2 ? N2 : N1);
            orc_half *e = &h[(size_t)3 * (c - 1) + i];
            e->key = (lo << 32) | hi;
            e->cell = c;
            e->loc = i;
        }
    qsort(h, nh, sizeof(orc_half), half_cmp);
    int ng = 0;
    int status = ORC_OK;
    for (size_t k = 0; k < nh;) {
        size_t j = k;
        while (j < nh && h[j].key == h[k].key) j++;
        if (j - k > 2 || (j - k == 2 && h[k].cell == h[k + 1].cell)) status = ORC_ERR_NONMANIFOLD;
        gfirst[ng] = (int)k;
        gsize[ng] = (int)(j - k);
        for (size_t t = k; t < j; t++) grp[(size_t)3 * (h[t].cell - 1) + h[t].loc] = ng;
        ng++;
        k = j;
    }
    if (status == ORC_OK) {
        for (int c = 1; c <= C; c++) {
            for (int i = 0; i < 3; i++) {
                int g = grp[(size_t)3 * (c - 1) + i];
                if (gedge[g] == 0) {
                    int N1 = m->pt[3 * c + IP[i]];
                    int N2 = m->pt[3 * c + IP[IP[i]]];
                    int ne = ++m->nedge;
                    gedge[g] = ne;
                    make_edge(m, ne, c, N1, N2);
                    if (gsize[g] == 2) {
                        const orc_half *a = &h[gfirst[g]], *b = &h[gfirst[g] + 1];
                        m->right[ne] = (a->cell == c) ? b->cell : a->cell;
                        m->elog[ne] = 0;
                    } else {
                        classify_boundary(m, ne);
                    }
                }
                m->celledge[3 * c + i] = gedge[g];
            }
            cell_neighbors(m, c);
        }
    }
    free(h); free(grp); free(gfirst); free(gsize); free(gedge);
    return status;
}

/* changeEdge, gpu_solver.cu:1832-1906: stable partition of edges into
 * interior (ELog 0), far field (1), wall (2); celledge is remapped. */
static int change_edge(orc_mesh *m)
{
    int E = m->nedge;
    int *newid = (int *)xcalloc((size_t)E + 1, sizeof(int));
    int *order = (int *)xcalloc((size_t)E + 1, sizeof(int));
    if (!newid || !order) { free(newid); free(order); return ORC_ERR_ALLOC; }
    int j = 1;
    for (int pass = 0; pass < 3; pass++)
        for (int i = 1; i <= E; i++)
            if (m->elog[i] == pass) { newid[i] = j; order[j] = i; j++; }
#define PERMUTE(arr, type)                                                     \
    do {                                                                       \
        type *tmp = (type *)xcalloc((size_t)E + 1, sizeof(type));              \
        for (int k = 1; k <= E; k++) tmp[k] = m->arr[order[k]];                \
        free(m->arr);                                                          \
        m->arr = tmp;                                                          \
    } while (0)
    PERMUTE(elog, int); PERMUTE(left, int); PERMUTE(right, int); PERMUTE(farid, int);
    PERMUTE(wallid, int); PERMUTE(node1, int); PERMUTE(node2, int);
    PERMUTE(vx, double); PERMUTE(vy, double); PERMUTE(vn, double);
    PERMUTE(midx, double); PERMUTE(midy, double);
#undef PERMUTE
    for (int c = 1; c <= m->ncell; c++)
        for (int i = 0; i < 3; i++) m->celledge[3 * c + i] = newid[m->celledge[3 * c + i]];
    m->wall_edge = (int *)xcalloc((size_t)m->nwall + 1, sizeof(int));
    for (int e = 1; e <= E; e++)
        if (m->elog[e] == 2) m->wall_edge[m->wallid[e]] = e;
    free(newid);
    free(order);
    return ORC_OK;
}

int orc_mesh_build(int nnode, int ncell, const double *x, const double *y, const int *tri,
                   int literal, orc_mesh **out)
{
    *out = NULL;
    orc_mesh *m = (orc_mesh *)xcalloc(1, sizeof(orc_mesh));
    if (!m) return ORC_ERR_ALLOC;
    m->nnode = nnode;
    m->ncell = ncell;
    size_t C1 = (size_t)ncell + 1, E1 = (size_t)3 * ncell + 1;
    m->x = (double *)xcalloc((size_t)nnode + 1, sizeof(double));
    m->y = (double *)xcalloc((size_t)nnode + 1, sizeof(double));
    m->pt = (int *)xcalloc(3 * C1, sizeof(int));
    m->area = (double *)xcalloc(C1, sizeof(double));
    m->cx = (double *)xcalloc(C1, sizeof(double));
    m->cy = (double *)xcalloc(C1, sizeof(double));
    m->celledge = (int *)xcalloc(3 * C1, sizeof(int));
    m->neicell = (int *)xcalloc(3 * C1, sizeof(int));
    m->clog = (int *)xcalloc(C1, sizeof(int));
    m->elog = (int *)xcalloc(E1, sizeof(int));
    m->left = (int *)xcalloc(E1, sizeof(int));
    m->right = (int *)xcalloc(E1, sizeof(int));
    m->farid = (int *)xcalloc(E1, sizeof(int));
    m->wallid = (int *)xcalloc(E1, sizeof(int));
    m->node1 = (int *)xcalloc(E1, sizeof(int));
    m->node2 = (int *)xcalloc(E1, sizeof(int));
    m->vx = (double *)xcalloc(E1, sizeof(double));
    m->vy = (double *)xcalloc(E1, sizeof(double));
    m->vn = (double *)xcalloc(E1, sizeof(double));
    m->midx = (double *)xcalloc(E1, sizeof(double));
    m->midy = (double *)xcalloc(E1, sizeof(double));
    if (!m->x || !m->y || !m->pt || !m->area || !m->cx || !m->cy || !m->celledge ||
        !m->neicell || !m->clog || !m->elog || !m->left || !m->right || !m->farid ||
        !m->wallid || !m->node1 || !m->node2 || !m->vx || !m->vy || !m->vn || !m->midx ||
        !m->midy) {
        orc_mesh_free(m);
        return ORC_ERR_ALLOC;
    }
    for (int i = 1; i <= nnode; i++) {
        m->x[i] = x[i - 1];
        m->y[i] = y[i - 1];
    }
    for (int c = 1; c <= ncell; c++)
        for (int i = 0; i < 3; i++) {
            int p = tri[3 * (c - 1) + i];
            if (p < 1 || p > nnode) { orc_mesh_free(m); return ORC_ERR_INDEX; }
            m->pt[3 * c + i] = p;
        }
    /* area and centroid, gpu_init.cu:107-116 */
    for (int c = 1; c <= ncell; c++) {
        int p0 = m->pt[3 * c], p1 = m->pt[3 * c + 1], p2 = m->pt[3 * c + 2];
        m->area[c] = calc_area(m->x[p0], m->y[p0], m->x[p1], m->y[p1], m->x[p2], m->y[p2]);
        if (m->area[c] <= 0.0) m->n_bad_area++;
        m->cx[c] = (m->x[p0] + m->x[p1] + m->x[p2]) / 3.0;
        m->cy[c] = (m->y[p0] + m->y[p1] + m->y[p2]) / 3.0;
    }
    int st = literal ? topology_literal(m) : topology_sorted(m);
    if (st == ORC_OK) st = change_edge(m);
    if (st != ORC_OK) { orc_mesh_free(m); return st; }
    *out = m;
    return ORC_OK;
}

/* Reader for the reference mesh format, gpu_init.cu:16-32: "Nnode Ncell",
 * Nnode x values, Nnode y values, Ncell triangles of 1-based node ids. */
int orc_mesh_read_dat(const char *path, orc_mesh **out, int literal)
{
    *out = NULL;
    FILE *fp = fopen(path, "r");
    if (!fp) return ORC_ERR_FILE;
    int nnode = 0, ncell = 0;
    if (fscanf(fp, "%d %d", &nnode, &ncell) != 2 || nnode < 3 || ncell < 1) {
        fclose(fp);
        return ORC_ERR_FILE;
    }
    double *x = (double *)xcalloc((size_t)nnode, sizeof(double));
    double *y = (double *)xcalloc((size_t)nnode, sizeof(double));
    int *tri = (int *)xcalloc((size_t)3 * ncell, sizeof(int));
    int ok = x && y && tri;
    for (int i = 0; ok && i < nnode; i++) ok = fscanf(fp, "%lf", &x[i]) == 1;
    for (int i = 0; ok && i < nnode; i++) ok = fscanf(fp, "%lf", &y[i]) == 1;
    for (int i = 0; ok && i < 3 * ncell; i++) ok = fscanf(fp, "%d", &tri[i]) == 1;
    fclose(fp);
    int st = ok ? orc_mesh_build(nnode, ncell, x, y, tri, literal, out) : ORC_ERR_FILE;
    free(x); free(y); free(tri);
    return st;
}

void orc_mesh_counts(const orc_mesh *m, int counts[5])
{
    counts[0] = m->nnode; counts[1] = m->ncell; counts[2] = m->nedge;
    counts[3] = m->nwall; counts[4] = m->nfar;
}

#define CP(dst, src, n, type) do { if (dst) memcpy(dst, src, (size_t)(n) * sizeof(type)); } while (0)
void orc_mesh_export(const orc_mesh *m, int *celledge, int *neicell, int *clog, int *elog,
                     int *left, int *right, int *farid, int *wallid, int *node1, int *node2,
                     double *area, double *cx, double *cy, double *vx, double *vy, double *vn,
                     double *midx, double *midy)
{
    size_t C1 = (size_t)m->ncell + 1, E1 = (size_t)m->nedge + 1;
    CP(celledge, m->celledge, 3 * C1, int); CP(neicell, m->neicell, 3 * C1, int);
    CP(clog, m->clog, C1, int);
    CP(elog, m->elog, E1, int); CP(left, m->left, E1, int); CP(right, m->right, E1, int);
    CP(farid, m->farid, E1, int); CP(wallid, m->wallid, E1, int);
    CP(node1, m->node1, E1, int); CP(node2, m->node2, E1, int);
    CP(area, m->area, C1, double); CP(cx, m->cx, C1, double); CP(cy, m->cy, C1, double);
    CP(vx, m->vx, E1, double); CP(vy, m->vy, E1, double); CP(vn, m->vn, E1, double);
    CP(midx, m->midx, E1, double); CP(midy, m->midy, E1, double);
}
#undef CP

/* ------------------------------------------------------------------------ */
/* Point kernels                                                            */
/* ------------------------------------------------------------------------ */

/* kernel_CalculationPTA, gpu_solver.cu:49-60 (= CalculationPTA gpu_init.cu:49-61). */
static void pta(orc_w *W, double GAMA, double R)
{
    double U, V, E;
    U = W->du / W->d;
    V = W->dv / W->d;
    E = W->de / W->d - 0.5 * (U * U + V * V);
    W->p = (GAMA - 1.0) * W->d * E;
    W->a = sqrt(GAMA * W->p / W->d);
    W->t = (GAMA - 1.0) * E / R;
}

void orc_pta(double gamma, double R, const double w[4], double out[3])
{
    orc_w W = {w[0], w[1], w[2], w[3], 0, 0, 0};
    pta(&W, gamma, R);
    out[0] = W.p; out[1] = W.t; out[2] = W.a;
}

/* kernel_Far_Boundary, gpu_solver.cu:62-135: characteristic far-field ghost. */
static void far_boundary(const orc_solver *s, orc_w *w0, const orc_w *DW, double vectorx,
                         double vectory)
{
    double PB, PD, PA, nx, ny, nn, Un, Uncell, AD, ROUD, UB, VB, UD, VD;
    double GAMA = s->GAMA, UI = s->UI, VI = s->VI, R = s->R;
    nn = sqrt(vectorx * vectorx + vectory * vectory);
    nx = vectory / nn;
    ny = vectorx / nn;
    UD = DW->du / DW->d;
    VD = DW->dv / DW->d;
    Un = fabs(UI * nx + VI * ny);
    Uncell = UD * nx + VD * ny;
    if (Un / s->AIN > 1.0) { /* supersonic: compares with dimensional AIN (gpu_solver.cu:79) */
        if (Uncell > 0.0) *w0 = *DW;
        if (Uncell <= 0.0) {
            w0->d = 1.0;
            w0->du = UI;
            w0->dv = VI;
            w0->de = R * s->TIN / (s->VIN * s->VIN) / (GAMA - 1.0) + 0.50;
            pta(w0, GAMA, R);
        }
    } else {
        PA = s->PIS;
        PD = DW->p;
        AD = DW->a;
        ROUD = DW->d;
        if (Uncell > 0.0) {
            PB = PA;
            w0->p = PB;
            w0->d = ROUD + (PB - PD) / AD / AD;
            UB = UD + nx * (PD - PB) / ROUD / AD;
            VB = VD + ny * (PD - PB) / ROUD / AD;
            w0->du = UB * w0->d;
            w0->dv = VB * w0->d;
            w0->de = PB / (GAMA - 1.0) + 0.5 * (UB * UB + VB * VB) * w0->d;
            pta(w0, GAMA, R);
        }
        if (Uncell <= 0.0) {
            PB = 0.5 * (PA + PD - ROUD * AD * (nx * (UI - UD) + ny * (VI - VD)));
            w0->p = PB;
            w0->d = 1.0 + (PB - PA) / AD / AD;
            UB = UI - nx * (PA - PB) / ROUD / AD;
            VB = VI - ny * (PA - PB) / ROUD / AD;
            w0->du = UB * w0->d;
            w0->dv = VB * w0->d;
            w0->de = PB / (GAMA - 1.0) + 0.5 * (UB * UB + VB * VB) * w0->d;
            pta(w0, GAMA, R);
        }
    }
}

/* Wall mirror ghost, kernel_Boundary gpu_solver.cu:188-221. */
static void wall_boundary(orc_w *out, const orc_w *wd_in, double vectorx, double vectory,
                          double vectorn)
{
    orc_w wd = *wd_in;
    double nij0 = vectory / vectorn;
    double nij1 = vectorx / vectorn;
    double UU = wd.du / wd.d;
    double VV = wd.dv / wd.d;
    double UV = UU * nij0 + VV * nij1;
    UU = UU - 2 * UV * nij0;
    VV = VV - 2 * UV * nij1;
    wd.du = wd.d * UU;
    wd.dv = wd.d * VV;
    *out = wd;
}

/* kernel_Limiter, gpu_solver.cu:456-494 (per face, per component). */
static void limiter(const double UI[4], const double umax[4], const double umin[4],
                    double delt2[4])
{
    for (int i = 0; i < 4; i++) {
        double fi = 1.0;
        if (delt2[i] > 0) fi = fmin((umax[i] - UI[i]) / delt2[i], 1.0);
        if (delt2[i] < 0) fi = fmin((umin[i] - UI[i]) / delt2[i], 1.0);
        if (delt2[i] == 0) fi = 1.0;
        delt2[i] *= fi;
    }
}

void orc_limiter(const double u[4], const double umax[4], const double umin[4], double delt[4])
{
    limiter(u, umax, umin, delt);
}

/* kernel_Calc_Function, gpu_solver.cu:552-566 (reciprocal-multiply form). */
static void calc_function(const orc_w *w, double *A, double *U, double *V, double *H,
                          double *P, double GAMA)
{
    double temp = 1.0 / w->d;
    *U = w->du * temp;
    *V = w->dv * temp;
    *P = (GAMA - 1.0) * w->d * (w->de * temp - 0.5 * (*U * *U + *V * *V));
    *A = sqrt(GAMA * *P * temp);
    *H = w->de * temp + *P * temp;
}

/* kernel_ROE_SCHEME, gpu_solver.cu:569-647: Roe flux with Harten fix 0.05. */
static void roe(const orc_w *LW, const orc_w *RW, double nx, double ny, double GAMA,
                double FLUX[4])
{
    double ROUB, UB, VB, HB, QB, VC, AB, DVC, epsilon;
    double A[2], U[2], V[2], H[2], P[2], M[2], FF[3][4], AD[3], WL[4], WR[4];
    double lrou, rrou, sqrtlrou, sqrtrrou;
    calc_function(LW, &A[0], &U[0], &V[0], &H[0], &P[0], GAMA);
    calc_function(RW, &A[1], &U[1], &V[1], &H[1], &P[1], GAMA);
    M[0] = (U[0] * nx + V[0] * ny);
    M[1] = (U[1] * nx + V[1] * ny);
    DVC = M[1] - M[0];
    lrou = LW->d;
    rrou = RW->d;
    sqrtlrou = sqrt(lrou);
    sqrtrrou = sqrt(rrou);
    ROUB = sqrtlrou * sqrtrrou;
    UB = (U[0] * sqrtlrou + U[1] * sqrtrrou) / (sqrtlrou + sqrtrrou);
    VB = (V[0] * sqrtlrou + V[1] * sqrtrrou) / (sqrtlrou + sqrtrrou);
    HB = (H[0] * sqrtlrou + H[1] * sqrtrrou) / (sqrtlrou + sqrtrrou);
    QB = UB * UB + VB * VB;
    AB = sqrt((GAMA - 1.0) * (HB - QB / 2.0));
    VC = UB * nx + VB * ny;
    epsilon = 0.05;
    if (fabs(VC - AB) < epsilon)
        AD[0] = ((VC - AB) * (VC - AB) + epsilon * epsilon) / (2 * epsilon);
    else
        AD[0] = fabs(VC - AB);
    if (fabs(VC) < epsilon)
        AD[1] = (VC * VC + epsilon * epsilon) / (2 * epsilon);
    else
        AD[1] = fabs(VC);
    if (fabs(VC + AB) < epsilon)
        AD[2] = ((VC + AB) * (VC + AB) + epsilon * epsilon) / (2 * epsilon);
    else
        AD[2] = fabs(VC + AB);
    FF[0][0] = AD[0] * ((P[1] - P[0]) - ROUB * AB * DVC) / (2.0 * AB * AB);
    FF[0][1] = FF[0][0] * (UB - AB * nx);
    FF[0][2] = FF[0][0] * (VB - AB * ny);
    FF[0][3] = FF[0][0] * (HB - AB * VC);
    FF[1][0] = AD[1] * ((rrou - lrou) - (P[1] - P[0]) / (AB * AB));
    FF[1][1] = FF[1][0] * UB + AD[1] * ROUB * (U[1] - U[0] - DVC * nx);
    FF[1][2] = FF[1][0] * VB + AD[1] * ROUB * (V[1] - V[0] - DVC * ny);
    FF[1][3] = FF[1][0] * QB * 0.5 + AD[1] * ROUB * (UB * (U[1] - U[0]) + VB * (V[1] - V[0]) - DVC * VC);
    FF[2][0] = AD[2] * ((P[1] - P[0]) + ROUB * AB * DVC) / (2.0 * AB * AB);
    FF[2][1] = FF[2][0] * (UB + AB * nx);
    FF[2][2] = FF[2][0] * (VB + AB * ny);
    FF[2][3] = FF[2][0] * (HB + AB * VC);
    WL[0] = lrou * M[0];
    WL[1] = U[0] * WL[0] + nx * P[0];
    WL[2] = V[0] * WL[0] + ny * P[0];
    WL[3] = WL[0] * H[0];
    WR[0] = rrou * M[1];
    WR[1] = U[1] * WR[0] + nx * P[1];
    WR[2] = V[1] * WR[0] + ny * P[1];
    WR[3] = WR[0] * H[1];
    for (int k = 0; k < 4; k++) FLUX[k] = 0.5 * ((WL[k] + WR[k]) - (FF[0][k] + FF[1][k] + FF[2][k]));
    (void)A;
}

void orc_roe(const double wl[4], const double wr[4], double nx, double ny, double gamma,
             double flux[4])
{
    orc_w L = {wl[0], wl[1], wl[2], wl[3], 0, 0, 0}, R = {wr[0], wr[1], wr[2], wr[3], 0, 0, 0};
    roe(&L, &R, nx, ny, gamma, flux);
}

/* kernel_deltT body, gpu_solver.cu:434-442: spectral radius of one edge from
 * the arithmetic mean of the unreconstructed states. */
static double spectral_radius(const orc_w *L, const orc_w *Rw, double vectorx, double vectory,
                              double GAMA)
{
    double u, v, rou, p, VC, CS, SR, EE;
    rou = (L->d + Rw->d) * 0.5;
    u = (L->du + Rw->du) / rou * 0.5;
    v = (L->dv + Rw->dv) / rou * 0.5;
    EE = (L->de + Rw->de) * 0.5 / rou - (u * u + v * v) * 0.5;
    VC = u * vectory + v * vectorx;
    p = rou * EE * (GAMA - 1.0);
    CS = sqrt((GAMA * p / rou) * (vectorx * vectorx + vectory * vectory));
    SR = fabs(VC) + CS;
    return SR;
}

double orc_spectral_radius(const double wl[4], const double wr[4], double vx, double vy,
                           double gamma)
{
    orc_w L = {wl[0], wl[1], wl[2], wl[3], 0, 0, 0}, R = {wr[0], wr[1], wr[2], wr[3], 0, 0, 0};
    return spectral_radius(&L, &R, vx, vy, gamma);
}

void orc_far_ghost(const orc_solver *s, const double wd[4], double vx, double vy, double out[4])
{
    orc_w D = {wd[0], wd[1], wd[2], wd[3], 0, 0, 0}, G;
    memset(&G, 0, sizeof(G));
    pta(&D, s->GAMA, s->R);
    far_boundary(s, &G, &D, vx, vy);
    out[0] = G.d; out[1] = G.du; out[2] = G.dv; out[3] = G.de;
}

void orc_wall_ghost(const double wd[4], double vx, double vy, double vn, double out[4])
{
    orc_w D = {wd[0], wd[1], wd[2], wd[3], 0, 0, 0}, G;
    wall_boundary(&G, &D, vx, vy, vn);
    out[0] = G.d; out[1] = G.du; out[2] = G.dv; out[3] = G.de;
}

/* ------------------------------------------------------------------------ */
/* Solver                                                                   */
/* ------------------------------------------------------------------------ */

/* Initialization, gpu_init.cu:63-93 (non-dimensional freestream). */
orc_solver *orc_create(const orc_mesh *m, const orc_flow *f, const orc_opts *o)
{
    orc_solver *s = (orc_solver *)xcalloc(1, sizeof(orc_solver));
    if (!s) return NULL;
    s->m = m;
    s->flow = *f;
    s->opts = *o;
    size_t C1 = (size_t)m->ncell + 1, E1 = (size_t)m->nedge + 1;
    s->w = (orc_w *)xcalloc(C1, sizeof(orc_w));
    s->w0 = (orc_w *)xcalloc(C1, sizeof(orc_w));
    s->farw = (orc_w *)xcalloc((size_t)m->nfar + 1, sizeof(orc_w));
    s->wallw = (orc_w *)xcalloc((size_t)m->nwall + 1, sizeof(orc_w));
    s->deltU = (double *)xcalloc(8 * C1, sizeof(double));
    s->umax = (double *)xcalloc(4 * C1, sizeof(double));
    s->umin = (double *)xcalloc(4 * C1, sizeof(double));
    s->rel = (double *)xcalloc(4 * E1, sizeof(double));
    s->rer = (double *)xcalloc(4 * E1, sizeof(double));
    s->dtl = (double *)xcalloc(E1, sizeof(double));
    s->dtr = (double *)xcalloc(E1, sizeof(double));
    s->res = (double *)xcalloc(4 * C1, sizeof(double));
    s->r1 = (double *)xcalloc(4 * C1, sizeof(double));
    s->dt = (double *)xcalloc(C1, sizeof(double));
    s->tsum = (double *)xcalloc(C1, sizeof(double));
    if (!s->w || !s->w0 || !s->farw || !s->wallw || !s->deltU || !s->umax || !s->umin ||
        !s->rel || !s->rer || !s->dtl || !s->dtr || !s->res || !s->r1 || !s->dt || !s->tsum) {
        orc_destroy(s);
        return NULL;
    }
    s->GAMA = f->gamma;
    s->R = ORC_R;
    s->TIN = f->t_in;
    s->CFL = f->cfl;
    s->ALPHA = f->alpha_deg * ORC_PI / 180.0;
    s->AIN = sqrt(s->GAMA * s->R * s->TIN);
    s->VIN = f->mach * s->AIN;
    s->ROUIN = f->p_in / (s->R * s->TIN);
    s->UI = cos(s->ALPHA);
    s->VI = sin(s->ALPHA);
    double rou_U = cos(s->ALPHA), rou_V = sin(s->ALPHA);
    double rou_E = s->R * s->TIN / (s->VIN * s->VIN) / (s->GAMA - 1.0) + 0.50;
    s->rhoE_inf = rou_E;
    for (int c = 1; c <= m->ncell; c++) {
        s->w[c].d = 1.0;
        s->w[c].du = rou_U;
        s->w[c].dv = rou_V;
        s->w[c].de = rou_E;
        pta(&s->w[c], s->GAMA, s->R);
    }
    s->PIS = s->w[1].p;
    s->step = 0;
    s->epsm = 0.0;
    return s;
}

void orc_destroy(orc_solver *s)
{
    if (!s) return;
    free(s->w); free(s->w0); free(s->farw); free(s->wallw); free(s->deltU);
    free(s->umax); free(s->umin); free(s->rel); free(s->rer); free(s->dtl); free(s->dtr);
    free(s->res); free(s->r1); free(s->dt); free(s->tsum);
    free(s);
}

void orc_freestream(const orc_solver *s, double out[7])
{
    out[0] = s->UI; out[1] = s->VI; out[2] = s->AIN; out[3] = s->VIN;
    out[4] = s->ROUIN; out[5] = s->PIS; out[6] = s->rhoE_inf;
}

/* kernel_Boundary, gpu_solver.cu:137-223. */
static void boundary(orc_solver *s)
{
    const orc_mesh *m = s->m;
#pragma omp parallel for schedule(static)
    for (int e = 1; e <= m->nedge; e++) {
        const orc_w *wd = &s->w[m->left[e]];
        if (m->elog[e] == 1) far_boundary(s, &s->farw[m->farid[e]], wd, m->vx[e], m->vy[e]);
        if (m->elog[e] == 2) wall_boundary(&s->wallw[m->wallid[e]], wd, m->vx[e], m->vy[e], m->vn[e]);
    }
}

/* State across local edge i of cell c: neighbour cell, far ghost or wall
 * ghost (gpu_solver.cu:285-310). */
static const orc_w *across(const orc_solver *s, int c, int i)
{
    const orc_mesh *m = s->m;
    int nc = m->neicell[3 * c + i], e = m->celledge[3 * c + i];
    if (nc > 0) return &s->w[nc];
    if (nc == -1) return &s->farw[m->farid[e]];
    return &s->wallw[m->wallid[e]];
}

/* kernel_Gradient, gpu_solver.cu:225-383: Green-Gauss gradient and the
 * neighbourhood min/max. */
static void gradient(orc_solver *s)
{
    const orc_mesh *m = s->m;
#pragma omp parallel for schedule(static)
    for (int c = 1; c <= m->ncell; c++) {
        const orc_w *Wi = &s->w[c];
        double Ui[4] = {Wi->d, Wi->du, Wi->dv, Wi->de};
        double dU[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
        double jmax[4], jmin[4];
        for (int i = 0; i < 3; i++) {
            int N1 = m->pt[3 * c + (i + 1) % 3];
            int N2 = m->pt[3 * c + (i + 2) % 3];
            double vectorx = m->x[N1] - m->x[N2];
            double vectory = m->y[N2] - m->y[N1];
            const orc_w *Wc = across(s, c, i);
            double comp[4] = {Wc->d, Wc->du, Wc->dv, Wc->de};
            for (int j = 0; j < 4; j++) {
                if (i == 0) {
                    jmax[j] = comp[j];
                    jmin[j] = comp[j];
                } else {
                    if (comp[j] < jmin[j]) jmin[j] = comp[j];
                    if (comp[j] > jmax[j]) jmax[j] = comp[j];
                }
            }
            for (int j = 0; j < 4; j++) dU[0][j] += 0.5 * (Ui[j] + comp[j]) * vectory;
            for (int j = 0; j < 4; j++) dU[1][j] += 0.5 * (Ui[j] + comp[j]) * vectorx;
        }
        double area = 1.0 / m->area[c];
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 4; k++) s->deltU[8 * c + 4 * j + k] = dU[j][k] * area;
        for (int j = 0; j < 4; j++) {
            s->umax[4 * c + j] = fmax(Ui[j], jmax[j]);
            s->umin[4 * c + j] = fmin(Ui[j], jmin[j]);
        }
    }
}

/* Limited linear reconstruction of cell c's state at (midx, midy),
 * kernel_LBJinterpolate gpu_solver.cu:496-549 (one side). */
static void reconstruct(const orc_solver *s, int c, double midx, double midy, orc_w *out)
{
    const orc_mesh *m = s->m;
    const orc_w *W = &s->w[c];
    double U[4] = {W->d, W->du, W->dv, W->de};
    double r0 = midx - m->cx[c];
    double r1 = midy - m->cy[c];
    double delt2[4];
    for (int k = 0; k < 4; k++)
        delt2[k] = s->deltU[8 * c + k] * r0 + s->deltU[8 * c + 4 + k] * r1;
    limiter(U, &s->umax[4 * c], &s->umin[4 * c], delt2);
    out->d = U[0] + delt2[0];
    out->du = U[1] + delt2[1];
    out->dv = U[2] + delt2[2];
    out->de = U[3] + delt2[3];
}

/* kernel_Flux2 (gpu_solver.cu:651-764) and kernel_deltT (gpu_solver.cu:385-454). */
static void flux(orc_solver *s)
{
    const orc_mesh *m = s->m;
#pragma omp parallel for schedule(static)
    for (int e = 1; e <= m->nedge; e++) {
        int L = m->left[e], R = m->right[e];
        orc_w LW, RW;
        const orc_w *Rraw;
        if (m->elog[e] == 0) {
            if (s->opts.order >= 2) {
                reconstruct(s, L, m->midx[e], m->midy[e], &LW);
                reconstruct(s, R, m->midx[e], m->midy[e], &RW);
            } else {
                LW = s->w[L];
                RW = s->w[R];
            }
            Rraw = &s->w[R];
        } else {
            LW = s->w[L];
            RW = (m->elog[e] == 1) ? s->farw[m->farid[e]] : s->wallw[m->wallid[e]];
            Rraw = (m->elog[e] == 1) ? &s->farw[m->farid[e]] : &s->wallw[m->wallid[e]];
        }
        double F[4];
        double vectorx = m->vx[e], vectory = m->vy[e], vectorn = m->vn[e];
        roe(&LW, &RW, vectory / vectorn, vectorx / vectorn, s->GAMA, F);
        for (int k = 0; k < 4; k++) F[k] *= vectorn;
        for (int k = 0; k < 4; k++) {
            s->rel[4 * e + k] = F[k];
            s->rer[4 * e + k] = (R != -1) ? F[k] : 0.0;
        }
        double SR = spectral_radius(&s->w[L], Rraw, vectorx, vectory, s->GAMA);
        s->dtl[e] = SR;
        if (R != -1) s->dtr[e] = SR;
    }
}

/* kernel_Flux_e2c, gpu_solver.cu:768-878: ordered gather of the three edge
 * slots of each cell, and the local time step.  The global-time-step variant
 * restates kernel_TimeStep (gpu_solver.cu:880-936) with the host minimum of
 * the commented-out code at gpu_solver.cu:2376-2387. */
static void gather(orc_solver *s)
{
    const orc_mesh *m = s->m;
#pragma omp parallel for schedule(static)
    for (int c = 1; c <= m->ncell; c++) {
        double T = 0;
        double r[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < 3; i++) {
            int e = m->celledge[3 * c + i];
            if (m->left[e] == c) {
                T += s->dtl[e];
                for (int j = 0; j < 4; j++) r[j] += s->rel[4 * e + j];
            }
            if (m->right[e] == c) {
                T += s->dtr[e];
                for (int j = 0; j < 4; j++) r[j] -= s->rer[4 * e + j];
            }
        }
        for (int j = 0; j < 4; j++) s->res[4 * c + j] = r[j];
        s->tsum[c] = T;
        s->dt[c] = fmin(m->area[c] * s->CFL / T, 0.1);
    }
    if (s->opts.global_dt) {
        double dtg = 9999999.0;
        for (int c = 1; c <= m->ncell; c++) dtg = fmin(dtg, m->area[c] * s->CFL / s->tsum[c]);
        for (int c = 1; c <= m->ncell; c++) s->dt[c] = dtg;
    }
}

/* kernel_Smooth1 / kernel_Smooth2, gpu_solver.cu:1633-1720. */
static void smooth_sweeps(orc_solver *s, double *res)
{
    const orc_mesh *m = s->m;
    double eps = s->opts.eps;
#pragma omp parallel for schedule(static)
    for (int c = 1; c <= m->ncell; c++) {
        const int *nb = &m->neicell[3 * c];
        for (int i = 0; i < 4; i++) {
            double data = 0.0;
            if (nb[0] > 0) data += res[4 * nb[0] + i];
            if (nb[1] > 0) data += res[4 * nb[1] + i];
            if (nb[2] > 0) data += res[4 * nb[2] + i];
            s->r1[4 * c + i] = (data * (eps) + res[4 * c + i]) / (1.0 + (eps) * m->clog[c]);
        }
    }
#pragma omp parallel for schedule(static)
    for (int c = 1; c <= m->ncell; c++) {
        const int *nb = &m->neicell[3 * c];
        for (int i = 0; i < 4; i++) {
            double data = 0.0;
            if (nb[0] > 0) data += s->r1[4 * nb[0] + i];
            if (nb[1] > 0) data += s->r1[4 * nb[1] + i];
            if (nb[2] > 0) data += s->r1[4 * nb[2] + i];
            double data2 = (data * (eps) + res[4 * c + i]) / (1.0 + (eps) * m->clog[c]);
            if (fabs(res[4 * c + i]) > fabs(data2)) res[4 * c + i] = data2;
        }
    }
}

/* kernel_RKStep, gpu_solver.cu:938-966. */
static void rk_step(orc_solver *s, double RKc)
{
    const orc_mesh *m = s->m;
#pragma omp parallel for schedule(static)
    for (int c = 1; c <= m->ncell; c++) {
        double RK = RKc;
        RK *= s->dt[c];
        orc_w ww;
        ww.d = s->w0[c].d - RK * s->res[4 * c + 0] / m->area[c];
        ww.du = s->w0[c].du - RK * s->res[4 * c + 1] / m->area[c];
        ww.dv = s->w0[c].dv - RK * s->res[4 * c + 2] / m->area[c];
        ww.de = s->w0[c].de - RK * s->res[4 * c + 3] / m->area[c];
        pta(&ww, s->GAMA, s->R);
        s->w[c] = ww;
    }
}

/* kernel_drou (gpu_solver.cu:968-1020) with 64-wide blocks and its tree
 * reduction, then the host's in-order sum of block partials (:2409-2414). */
static double drou_sum(const orc_solver *s)
{
    int C = s->m->ncell;
    int nblk = C / ORC_BLOCK + 1;
    double total = 0.0;
    for (int b = 0; b < nblk; b++) {
        double r[ORC_BLOCK];
        for (int t = 0; t < ORC_BLOCK; t++) {
            int c = b * ORC_BLOCK + t;
            if (c > 0 && c <= C) {
                double d = s->w[c].d - s->w0[c].d;
                r[t] = d * d;
            } else {
                r[t] = 0.0;
            }
        }
        for (int st = ORC_BLOCK / 2; st >= 1; st /= 2)
            for (int t = 0; t < st; t++) r[t] = r[t] + r[t + st];
        total += r[0];
    }
    return total;
}

/* ROE_SOLVER iteration loop, gpu_solver.cu:2200-2434. */
int orc_run(orc_solver *s, int maxstep, double epsl, double *resid)
{
    const orc_mesh *m = s->m;
    int start = s->step;
    if (resid) resid[0] = 100.0;
    double last = 100.0;
    while (s->step < maxstep && last > epsl) {
        s->step++;
#pragma omp parallel for schedule(static)
        for (int c = 1; c <= m->ncell; c++) s->w0[c] = s->w[c];
        double drou = 0.0;
        for (int IR = 0; IR < 4; IR++) {
            boundary(s);
            if (s->opts.order >= 2) gradient(s);
            flux(s);
            gather(s);
            if (s->opts.smoothing) smooth_sweeps(s, s->res);
            rk_step(s, ORC_RK[IR]);
            if (IR == 3) drou = drou_sum(s);
        }
        if (s->step == 1) s->epsm = drou;
        last = sqrt((drou + 1.0E-32) / s->epsm);
        if (resid) resid[s->step] = last;
    }
    return s->step - start;
}

int orc_step_count(const orc_solver *s)
{
    return s->step;
}

void orc_get_state(const orc_solver *s, double *rho, double *rhou, double *rhov, double *rhoe)
{
    for (int c = 1; c <= s->m->ncell; c++) {
        if (rho) rho[c - 1] = s->w[c].d;
        if (rhou) rhou[c - 1] = s->w[c].du;
        if (rhov) rhov[c - 1] = s->w[c].dv;
        if (rhoe) rhoe[c - 1] = s->w[c].de;
    }
}

void orc_set_state(orc_solver *s, const double *rho, const double *rhou, const double *rhov,
                   const double *rhoe)
{
    for (int c = 1; c <= s->m->ncell; c++) {
        s->w[c].d = rho[c - 1];
        s->w[c].du = rhou[c - 1];
        s->w[c].dv = rhov[c - 1];
        s->w[c].de = rhoe[c - 1];
        pta(&s->w[c], s->GAMA, s->R);
    }
}

void orc_eval_residual(orc_solver *s, double *res, double *sr, double *dt, double *grad,
                       double *umax, double *umin)
{
    const orc_mesh *m = s->m;
    boundary(s);
    if (s->opts.order >= 2) gradient(s);
    flux(s);
    gather(s);
    for (int c = 1; c <= m->ncell; c++) {
        for (int k = 0; k < 4; k++) {
            if (res) res[4 * (c - 1) + k] = s->res[4 * c + k];
            if (umax) umax[4 * (c - 1) + k] = s->umax[4 * c + k];
            if (umin) umin[4 * (c - 1) + k] = s->umin[4 * c + k];
        }
        for (int k = 0; k < 8; k++)
            if (grad) grad[8 * (c - 1) + k] = s->deltU[8 * c + k];
        if (sr) sr[c - 1] = s->tsum[c];
        if (dt) dt[c - 1] = s->dt[c];
    }
}

void orc_smooth(orc_solver *s, double *res)
{
    int C = s->m->ncell;
    double *tmp = (double *)xcalloc(4 * ((size_t)C + 1), sizeof(double));
    if (!tmp) return;
    memcpy(tmp + 4, res, sizeof(double) * 4 * (size_t)C);
    smooth_sweeps(s, tmp);
    memcpy(res, tmp + 4, sizeof(double) * 4 * (size_t)C);
    free(tmp);
}

/* Neumaier-compensated accumulation. */
static void nsum(double *sum, double *comp, double v)
{
    double t = *sum + v;
    if (fabs(*sum) >= fabs(v))
        *comp += (*sum - t) + v;
    else
        *comp += (v - t) + *sum;
    *sum = t;
}

/* Lift and drag coefficients (not computed by the reference).  Force on the
 * body = sum over wall edges, in wall-id order, of the momentum components of
 * the first-order Roe wall flux times the edge length (the wall ghost of
 * kernel_Boundary gpu_solver.cu:188-221, kernel_ROE_SCHEME :569-647), minus
 * p_inf * sum(n ds); normalised by 0.5 rho_inf V_inf^2 c = 0.5. */
void orc_forces(const orc_solver *s, double *cl, double *cd)
{
    const orc_mesh *m = s->m;
    double fx = 0, cfx = 0, fy = 0, cfy = 0, sx = 0, csx = 0, sy = 0, csy = 0;
    for (int w = 1; w <= m->nwall; w++) {
        int e = m->wall_edge[w];
        orc_w G, L = s->w[m->left[e]];
        wall_boundary(&G, &L, m->vx[e], m->vy[e], m->vn[e]);
        double F[4];
        roe(&L, &G, m->vy[e] / m->vn[e], m->vx[e] / m->vn[e], s->GAMA, F);
        nsum(&fx, &cfx, F[1] * m->vn[e]);
        nsum(&fy, &cfy, F[2] * m->vn[e]);
        nsum(&sx, &csx, m->vy[e]);
        nsum(&sy, &csy, m->vx[e]);
    }
    double Fx = (fx + cfx) - s->PIS * (sx + csx);
    double Fy = (fy + cfy) - s->PIS * (sy + csy);
    double q = 0.5;
    if (cl) *cl = (-Fx * s->VI + Fy * s->UI) / q;
    if (cd) *cd = (Fx * s->UI + Fy * s->VI) / q;
}

/* Threads the per-cell and per-edge loops use: 1 in the plain build, the
 * OpenMP team size in the -fopenmp build (liboracle_omp.so, CPU baseline
 * only).  Every parallel loop writes per-cell or per-edge outputs only, and
 * the residual sum stays sequential, so both builds give the same bits. */
int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

double orc_time_steps(orc_solver *s, int steps)
{
    struct timespec t0, t1;
    double *hist = (double *)xcalloc((size_t)s->step + steps + 1, sizeof(double));
    clock_gettime(CLOCK_MONOTONIC, &t0);
    orc_run(s, s->step + steps, -1.0, hist);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(hist);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
