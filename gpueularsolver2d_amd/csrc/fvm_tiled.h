// Tiled kernels of the explicit finite-volume Euler path.  Every launch
// covers every tile of every case of a batch, one workgroup per tile.
//
// One kernel per Runge-Kutta stage (stage_tile, the default): the RK update
// of the previous stage is recomputed for the tile and two rings of
// neighbours (so the updated state never leaves LDS), then the gradient,
// flux and gather of this stage run as in kernel A below.  In order:
//   1. the previous stage's residual of the tile and rings 1..4 into LDS,
//      one load per cell;
//   2. smoothing sweep 1 on the tile and rings 1..3, gathering from LDS
//      (kernel_Smooth1, gpu_solver.cu:1635-1673);
//   3. sweep 2, selection and W = W0 - a dt R / A on the tile and rings 1..2
//      (kernel_Smooth2 :1675-1720, kernel_RKStep :938-966); in the stage that
//      closes a time step the tile's new W0 and its FP64 partial of
//      (rho - rho0)^2 (kernel_drou :968-1020) are written;
//   4.-6. gradient, flux and gather as in kernel A.
// Cells of the rings are updated by several tiles from identical inputs, so
// every tile sees bitwise the same values.
//
// Two kernels per stage (flux_tile + update_tile), used with the global time
// step, which needs a grid-wide minimum between flux and update:
// Kernel A (flux_tile): one workgroup per tile of cells.
//   1. Green-Gauss gradient and neighbourhood bounds of the tile's cells and
//      of its first-ring halo (kernel_Gradient, gpu_solver.cu:225-383, with
//      the ghost states of kernel_Boundary :137-223 computed in place from
//      the boundary cell's own state): FP64 stores them in LDS rows, FP32
//      goes on to each cell's limited face states at its three edges
//      (kernel_Limiter :456-494, kernel_LBJinterpolate :496-549) and stores
//      those.
//   2. Roe flux and spectral radius of every edge touching the tile, into
//      LDS edge slots (FP64) or the two face slots of the edge (FP32)
//      (kernel_Flux2 :651-764, kernel_deltT :385-454).  Edges cut by the tile
//      border are evaluated by both tiles from identical inputs, so both see
//      bitwise the same flux.
//   3. Ordered gather of each cell's three edges and the local time step
//      (kernel_Flux_e2c :768-878); writes R and dt.
// Kernel B (update_tile): one workgroup per tile.
//   1. First smoothing sweep for the tile and its halo into LDS
//      (kernel_Smooth1 :1635-1673).
//   2. Second sweep and selection (kernel_Smooth2 :1675-1720), then the RK
//      update W = W0 - a dt R / A (kernel_RKStep :938-966); in the last stage
//      the tile's FP64 partial sum of (rho - rho0)^2 (kernel_drou :968-1020).
// finalize_kernel: per case, the FP64 residual and the stop test of the
// iteration loop (gpu_solver.cu:2403-2433).
//
// Determinism: every sum has a fixed order; there are no atomics.
#pragma once

#include "fvm_internal.h"
#include "fvm_numerics.h"

namespace fvm {

template <typename R>
struct alignas(16) SVec4 {
  R d, du, dv, de;
};

// Two values read together from LDS: a gradient component (d/dx, d/dy) or
// a bound (max, min) of one cell.
template <typename R>
struct alignas(2 * sizeof(R)) Pair {
  R a, b;
};

// A cell's (1 / area, area), adjacent in its CellGeom record: one load.  FP32
// stores 1 / area only (the second value is 0).
template <typename R>
__device__ __forceinline__ Pair<R> area_pair(const CellGeom<R>* cg, int i) {
  return *reinterpret_cast<const Pair<R>*>(&cg[i].inv_area);
}
template <>
__device__ __forceinline__ Pair<float> area_pair<float>(const CellGeom<float>* cg, int i) {
  return Pair<float>{cg[i].inv_area, 0.0f};
}

// Three signed 16-bit values and a fourth small one in 8 bytes (x: bits
// 0-15 and 16-31, y: bits 0-15, then w in bits 16-31): the tile-local
// neighbour ids of a cell (w = its count of interior neighbours), whose
// ranges the host checks.  Read every launch, so the narrower record
// is HBM bandwidth saved.
__host__ __device__ inline int2 pack_tri16(int a, int b, int c, int w = 0) {
  const unsigned lo = ((unsigned)a & 0xffffu) | (((unsigned)b & 0xffffu) << 16);
  return make_int2((int)lo, (int)(((unsigned)c & 0xffffu) | ((unsigned)w << 16)));
}
__device__ __forceinline__ int4 unpack_tri16(int2 v) {
  return make_int4((v.x << 16) >> 16, v.x >> 16, (v.y << 16) >> 16, v.y >> 16);
}

// LDS index of a neighbour entry for the smoothing sums: its own index, or
// the -0 slot `dummy` for a boundary side (negative codes become large
// unsigned values, so one unsigned minimum selects).
__device__ __forceinline__ int side_slot(int lk, int dummy) {
  return (int)min((unsigned)lk, (unsigned)dummy);
}

// Per tile edge: unit normal (n = (vy, vx)/|v|, the reference's convention),
// length, the tile-local indices of the two cells (right: -1 far field, -2
// wall) and the edge's local index (0..2) in each of them.
template <typename R>
struct EdgeRec;

// FP64: also the reconstruction offsets mid - centroid of both cells, and the
// edge vector itself, because the spectral radius of the reference uses it
// directly (u vy + v vx, vx^2 + vy^2).  l and r carry the local ids in their
// low 16 bits (r signed) and the edge's local index in the left / right cell
// above them.
template <>
struct alignas(16) EdgeRec<double> {
  double nx, ny, vn, vx, vy, rlx, rly, rrx, rry;
  int l, r;
};

// FP32: 12 bytes, the edge vector (its unit normal and length are formed
// from it: n = (vy, vx) / |v|) and, in one word, the two local indices and the
// edge's local index in each of its cells.  No reconstruction offsets: FP32
// forms each cell's face states in the gradient phase from the cell's own
// side vectors (face_states).
template <>
struct alignas(4) EdgeRec<float> {
  float vx, vy;
  int lr;  // bits 0-13: left (unsigned), 14-15: its local edge index; bits 16-29: right
           // (signed: -1 far field, -2 wall), 30-31: its local edge index
};

__host__ __device__ inline EdgeRec<double> make_edge_rec(double nx, double ny, double vn, double vx,
                                                         double vy, double rlx, double rly,
                                                         double rrx, double rry, int l, int r,
                                                         int kl, int kr, double*) {
  return EdgeRec<double>{nx, ny, vn, vx, vy, rlx, rly, rrx, rry,
                         (int)(((unsigned)kl << 16) | ((unsigned)l & 0xffffu)),
                         (int)(((unsigned)kr << 16) | ((unsigned)r & 0xffffu))};
}
__host__ __device__ inline EdgeRec<float> make_edge_rec(double, double, double, double vx, double vy,
                                                        double, double, double, double, int l,
                                                        int r, int kl, int kr, float*) {
  return EdgeRec<float>{(float)vx, (float)vy,
                        (int)(((unsigned)kr << 30) | (((unsigned)r & 0x3fffu) << 16) |
                              ((unsigned)kl << 14) | ((unsigned)l & 0x3fffu))};
}

// Unit normal n = (vy, vx) / |v| and length |v| of a tile edge: stored in
// FP64 (computed in FP64 on the host), formed from the stored edge vector in
// FP32.
template <typename R>
struct EdgeNormal {
  R nx, ny, vn;
};
__device__ __forceinline__ EdgeNormal<double> edge_normal(const EdgeRec<double>& e) {
  return {e.nx, e.ny, e.vn};
}
__device__ __forceinline__ EdgeNormal<float> edge_normal(const EdgeRec<float>& e) {
  const float vn = sqrtf(e.vx * e.vx + e.vy * e.vy);
  const float iv = 1.0f / vn;
  return {e.vy * iv, e.vx * iv, vn};
}

// Spectral radius of a tile edge: FP64 keeps the reference's form on the
// edge vector, FP32 uses the unit normal and length.
__device__ __forceinline__ double edge_sr(const EdgeNormal<double>&, const Vec4<double>& L,
                                          const Vec4<double>& R, double vx, double vy, double g) {
  return spectral_radius(L, R, vx, vy, g);
}
__device__ __forceinline__ float edge_sr(const EdgeNormal<float>& n, const Vec4<float>& L,
                                         const Vec4<float>& R, float, float, float g) {
  return spectral_radius_n(L, R, n.nx, n.ny, n.vn, g);
}

struct EdgeView {
  int l, r, kl, kr;
};
__device__ __forceinline__ EdgeView edge_cells(const EdgeRec<double>& e) {
  return {e.l & 0xffff, (e.r << 16) >> 16, e.l >> 16, (e.r >> 16) & 3};
}
__device__ __forceinline__ EdgeView edge_cells(const EdgeRec<float>& e) {
  return {e.lr & 0x3fff, (e.lr << 2) >> 18, (e.lr >> 14) & 3, (int)((unsigned)e.lr >> 30)};
}
__device__ __forceinline__ double edge_vx(const EdgeRec<double>& e) { return e.vx; }
__device__ __forceinline__ double edge_vy(const EdgeRec<double>& e) { return e.vy; }
__device__ __forceinline__ float edge_vx(const EdgeRec<float>& e) { return e.vx; }
__device__ __forceinline__ float edge_vy(const EdgeRec<float>& e) { return e.vy; }

template <typename R>
struct TiledArgs {
  const TileDesc* tiles;
  int ntiles;
  CaseCtl* ctl;
  const FlowT<R>* flow;       // per case
  const CellGeom<R>* cg;      // per mesh cell (all meshes concatenated)
  const int4* nbr;            // per mesh cell: neighbours (new ids, -1 far, -2 wall), .w = CLog
  const int4* nloc;           // per mesh cell: tile-local index of interior neighbours or -1
  const int* halo;
  const EdgeRec<R>* edges;
  const SVec4<R>* w_in;       // state read by kernel A
  SVec4<R>* res;              // residual written by kernel A / stage_tile
  R* dt;                      // local dt written by kernel A / stage_tile (FP32: dt / A; A*CFL/T in
                              // global mode)
  R* tile_min;                // per tile minimum of A*CFL/T (global dt mode); stage_tile: this
                              // stage's, written
  const R* tile_min_in;       // stage_tile, global dt mode: the previous stage's tile minima
  const R* dt_global;         // per case (global dt mode, two-kernel path)
  const SVec4<R>* w0;         // step-start state (two-kernel path)
  SVec4<R>* w_out;            // kernel B output
  double* partial;            // per tile sum of (rho - rho0)^2
  R rk;                       // RK coefficient of the update applied by this launch
  int last_stage;             // kernel B: 1 in the fourth stage
  int order, smoothing, global_dt;
  int max_local, max_edges;
  // one-kernel-per-stage path
  SVec4<R>* w0a;              // the two step-start state buffers (CaseCtl::w0_cur selects)
  SVec4<R>* w0b;
  const SVec4<R>* r_in;       // residual of the previous stage
  const R* dt_in;             // time step of the previous stage
  const int* rings;           // rings 1..4 of every tile
  const int2* lnb;            // tile-local neighbours of tile + rings 1..3 (Tri16)
  int apply_update;           // 0 only for the first launch of a run
  int close_step;             // 1: the update closes a time step (new W0 + partial)
  int max_r4_local, max_ring_local, max_w_local;
  int parity;                 // step parity within the run: step-start buffer = w0_cur ^ parity
  int close_prev;             // 1: the workgroup of each case's first tile closes the previous step
  int early_exit;             // 1: test a case's active flag before any load (several cases)
  int slot_off;               // stage_tile, FP64: byte offset of the gather slot table in LDS
  const HistDesc* hist;       // residual history
};

__device__ __forceinline__ int tile_of_block(int bid, int ntiles) {
  // Consecutive workgroup ids go to different XCDs (round robin over 8).  Give
  // each XCD a contiguous range of tiles so that neighbouring tiles, which
  // share halo cells, hit the same L2.
  const int x = bid & 7, k = bid >> 3;
  const int q = ntiles >> 3, r = ntiles & 7;
  return x * q + (x < r ? x : r) + k;
}

template <typename R>
__device__ __forceinline__ Vec4<R> ld(const SVec4<R>& s) {
  return Vec4<R>{s.d, s.du, s.dv, s.de};
}
template <typename R>
__device__ __forceinline__ SVec4<R> st(const Vec4<R>& v) {
  return SVec4<R>{v.d, v.du, v.dv, v.de};
}

__device__ __forceinline__ void neumaier(double& s, double& c, double v) {
  const double t = s + v;
  if (fabs(s) >= fabs(v))
    c += (s - t) + v;
  else
    c += (v - t) + s;
  s = t;
}

// Adds the compensated pair (s2, c2) to (s, c): the exact rounding error of
// s + s2 (Knuth's two-sum) joins the compensation.
__device__ __forceinline__ void add_pair(double& s, double& c, double s2, double c2) {
  const double t = s + s2;
  const double bb = t - s;
  const double e = (s - (t - bb)) + (s2 - bb);
  c = (c + c2) + e;
  s = t;
}

// Closes one time step of case c from its tiles' partial sums of
// (rho - rho0)^2 (all threads of one workgroup; blockDim a multiple of 64):
// Resd = sqrt((drou + 1e-32) / drou_of_step_1), the history entry and the
// stop test step < max_steps && Resd > epsl (gpu_solver.cu:2200,
// :2403-2433).  The sum is compensated and its order fixed: thread t adds
// partials t, t + blockDim, ... (Neumaier), a butterfly over each wave
// combines lane pairs lower lane first (both lanes of a pair then hold bitwise
// the same value), then thread 0 adds the waves in order.
__device__ inline void close_case(CaseCtl& c, const double* partial, const HistDesc* hist,
                                  int icase) {
  if (!c.active) return;
  const int n = c.ntile;
  const double* p = partial + c.tile0;
  double s = 0.0, cc = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) neumaier(s, cc, p[k]);
  const int lane = threadIdx.x & 63;
  for (int off = 1; off < 64; off <<= 1) {
    const double s2 = __shfl_xor(s, off, 64), c2 = __shfl_xor(cc, off, 64);
    if (lane & off) {  // this lane is the upper one of the pair
      double ls = s2, lc = c2;
      add_pair(ls, lc, s, cc);
      s = ls;
      cc = lc;
    } else {
      add_pair(s, cc, s2, c2);
    }
  }
  __shared__ double ws[16], wc[16];
  if (lane == 0) {
    ws[threadIdx.x >> 6] = s;
    wc[threadIdx.x >> 6] = cc;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double S = ws[0], C = wc[0];
  for (int k = 1; k < (int)(blockDim.x >> 6); k++) add_pair(S, C, ws[k], wc[k]);
  const double drou = S + C;
  const int step = c.step + 1;
  c.step = step;
  if (step == 1) c.epsm = drou;
  // A step 1 that changed nothing (drou_1 = 0: a steady state such as
  // freestream on a mesh without walls) gives +inf, as IEEE division does.
  // It is written out because the FP32 translation unit, whose kernels close
  // steps too, is built with a reciprocal relaxation that would give NaN.
  const double num = drou + 1.0E-32;
  const double resd = c.epsm != 0.0 ? sqrt(num / c.epsm) : (isnan(num) ? num : HUGE_VAL);
  c.resd = resd;
  const HistDesc h = *hist;
  if (step < h.stride) h.base[(size_t)icase * h.stride + step] = resd;
  // Resd is +inf, not NaN, when step 1 changed nothing (drou = 0, a state
  // that is already steady, e.g. freestream on a mesh without walls); the
  // reference's loop goes on then (inf > EPSL), and so does this one
  if (isnan(resd)) c.status = FVM_ERR_NAN;
  if (!(step < c.max_steps && resd > c.epsl)) c.active = 0;
}

// The step close as a launch of its own, one workgroup per case: used after
// the last step of each launch sequence (the other steps close inside the
// next step's first stage launch) and by the two-kernel and plain paths.  A
// separate launch orders the tiles' partial sums before the close without a
// device-scope fence, which on MI355X would write back a whole L2 per
// workgroup.  It is instantiated in each precision's translation unit (R is
// otherwise unused) so that every close of a run, in a stage kernel or here,
// is the same code built with the same floating-point options.  host_out
// (pinned host memory, may be null) receives each case's control block once
// it is closed, so the host reads a launch sequence's outcome without a copy
// of its own.
constexpr int kFinalizeThreads = 256;
template <typename R>
__global__ void __launch_bounds__(kFinalizeThreads) finalize_kernel(CaseCtl* ctl, const double* partial,
                                                                   const HistDesc* hist,
                                                                   CaseCtl* host_out) {
  close_case(ctl[blockIdx.x], partial, hist, blockIdx.x);
  // close_case's last writes are thread 0's own
  if (host_out && threadIdx.x == 0) {
    host_out[blockIdx.x] = ctl[blockIdx.x];
    __threadfence_system();
  }
}

// ---------------------------------------------------------------------------
// Phases shared by flux_tile and stage_tile
// ---------------------------------------------------------------------------

// FP64: Green-Gauss gradient and neighbourhood bounds of one cell into the
// LDS arrays sG[4][ML] of (d/dx, d/dy) pairs and sB[4][ML] of (max, min)
// pairs, one row per conservative (kernel_Gradient, gpu_solver.cu:225-383).
// The flux reads each pair with one LDS instruction.
// nbk: neighbour codes (>= 0 interior, -1 far field, -2 wall); nbv: the
// states of the interior neighbours (entries of boundary sides are unused).
// Callers load all three sides at once, a boundary side reading the cell
// itself, so the loads overlap instead of waiting on each other.
__device__ __forceinline__ void gradient_cell(int j, const Vec4<double>& Wi, const CellGeom<double>& cg,
                                              const int nbk[3], const Vec4<double> nbv[3],
                                              const FlowT<double>& f, double* sG, double* sB,
                                              int ML) {
  double gx[4] = {0.0, 0.0, 0.0, 0.0}, gy[4] = {0.0, 0.0, 0.0, 0.0};
  double mx[4], mn[4];
  const double ui[4] = {Wi.d, Wi.du, Wi.dv, Wi.de};
  // the sides' states, ghosts on the boundary sides, behind one branch that
  // cells with three interior neighbours (nearly all) skip together
  Vec4<double> cs[3] = {nbv[0], nbv[1], nbv[2]};
  if ((nbk[0] | nbk[1] | nbk[2]) < 0) {
#pragma unroll
    for (int k = 0; k < 3; k++)
      if (nbk[k] < 0) cs[k] = boundary_ghost<double>(-nbk[k], Wi, cg.gx(k), cg.gy(k), f);
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double cc[4] = {cs[k].d, cs[k].du, cs[k].dv, cs[k].de};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (k == 0) {
        mx[q] = cc[q];
        mn[q] = cc[q];
      } else {
        if (cc[q] < mn[q]) mn[q] = cc[q];
        if (cc[q] > mx[q]) mx[q] = cc[q];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) gx[q] += 0.5 * (ui[q] + cc[q]) * cg.gy(k);
#pragma unroll
    for (int q = 0; q < 4; q++) gy[q] += 0.5 * (ui[q] + cc[q]) * cg.gx(k);
  }
  Pair<double>* sG2 = reinterpret_cast<Pair<double>*>(sG);
  Pair<double>* sB2 = reinterpret_cast<Pair<double>*>(sB);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    sG2[q * ML + j] = Pair<double>{gx[q] * cg.inv_area, gy[q] * cg.inv_area};
    sB2[q * ML + j] = Pair<double>{fmax(ui[q], mx[q]), fmin(ui[q], mn[q])};
  }
}

// FP32: the limited reconstruction of one cell at each of its three edges,
// u + limit(grad u . r_k, u, Umax, Umin) per conservative (kernel_Gradient
// :225-383, kernel_Limiter :456-494 and the reconstruction of
// kernel_LBJinterpolate :496-549), into the face slots sFc[k][ML].
// r_k, from the centroid to the midpoint of local edge k (the side opposite
// vertex k, from P[k+1] to P[k+2]), is (P[k+1] + P[k+2]) / 2 - (P0 + P1 + P2) / 3
// = (d[k+2] - d[k+1]) / 6 with the side vectors d[i] = P[i+2] - P[i+1].  The
// cell record holds g[i] = (-d[i].x, d[i].y) * 0.5 / A, so
// r_k = (g[k+1].x - g[k+2].x, g[k+2].y - g[k+1].y) * A / 3, with A from the
// stored 1 / A.  Each face slot belongs to one edge: its flux thread reads the
// state there and then stores its flux in its place (flux_edges).
__device__ __forceinline__ void face_states(int j, const Vec4<float>& Wi, const CellGeom<float>& cg,
                                            const int nbk[3], const Vec4<float> nbv[3],
                                            const FlowT<float>& f, SVec4<float>* sFc, int ML) {
  float gx[4] = {0.0f, 0.0f, 0.0f, 0.0f}, gy[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  float mx[4], mn[4];
  const float ui[4] = {Wi.d, Wi.du, Wi.dv, Wi.de};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const Vec4<float> c =
        nbk[k] >= 0 ? nbv[k] : boundary_ghost<float>(-nbk[k], Wi, cg.gx(k), cg.gy(k), f);
    const float cc[4] = {c.d, c.du, c.dv, c.de};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      mn[q] = k == 0 ? cc[q] : fminf(mn[q], cc[q]);
      mx[q] = k == 0 ? cc[q] : fmaxf(mx[q], cc[q]);
      gx[q] += (ui[q] + cc[q]) * cg.gy(k);
      gy[q] += (ui[q] + cc[q]) * cg.gx(k);
    }
  }
  const float a3 = 1.0f / (3.0f * cg.inv_area);  // A / 3
#pragma unroll
  for (int q = 0; q < 4; q++) {
    gx[q] *= a3;
    gy[q] *= a3;
    mx[q] = fmaxf(ui[q], mx[q]);
    mn[q] = fminf(ui[q], mn[q]);
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
    const float rx = cg.gx(k1) - cg.gx(k2), ry = cg.gy(k2) - cg.gy(k1);
    float v[4];
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = ui[q] + limit(gx[q] * rx + gy[q] * ry, ui[q], mx[q], mn[q]);
    sFc[k * ML + j] = SVec4<float>{v[0], v[1], v[2], v[3]};
  }
}

// LDS arrays of the flux and gather phases (ML: the row stride).
// FP64: G and B the gradient and bound rows [4][ML] of (d/dx, d/dy) and (max,
// min) pairs, F and S the flux and spectral radius of each tile edge, and
// slots the gather table [4][tile cells] of 2 * edge + side.
// FP32: F the face states of each tile and ring-1 cell [3][ML]
// (face_states), which the flux phase replaces with the fluxes signed for
// the face's cell, and S [3][ML] the spectral radius of each face's edge; G,
// B and slots are unused.
template <typename R>
struct FluxLds {
  const SVec4<R>* W;
  R* G;
  R* B;
  SVec4<R>* F;
  R* S;
  unsigned short* slots;
  int ML;
};

// Gradient and bounds (FP64) or face states (FP32) of one tile or ring-1 cell.
template <typename R>
__device__ __forceinline__ void cell_slopes(int j, const Vec4<R>& Wi, const CellGeom<R>& cg,
                                            const int nbk[3], const Vec4<R> nbv[3],
                                            const FlowT<R>& f, const FluxLds<R>& x) {
  if constexpr (sizeof(R) == 4)
    face_states(j, Wi, cg, nbk, nbv, f, x.F, x.ML);
  else
    gradient_cell(j, Wi, cg, nbk, nbv, f, x.G, x.B, x.ML);
}

// Roe flux and spectral radius of every edge of the tile into LDS
// (kernel_Flux2 gpu_solver.cu:651-764 with kernel_LBJinterpolate :496-549,
// kernel_deltT :385-454).  kPre: e0 is this thread's prefetched first record
// (otherwise unused).
// FP64 reconstructs both sides from the gradient and bound rows and stores
// the edge's flux and spectral radius at its index; each edge also enters its
// slot, 2 * edge + side, in the gather table slots[4 * cell + local edge
// index] of each of its tile cells (cells of the tile have local ids below
// d.ncell), so the gather reads no slot table from memory (four entries per
// cell, the fourth unused, so that the gather reads a cell's three with one
// 8-byte access).
// FP32 reads the two face states and stores the flux, signed for each side's
// cell, and the spectral radius in the same two face slots.
template <typename R, bool kPre = true>
__device__ __forceinline__ void flux_edges(const TiledArgs<R>& a, const TileDesc& d,
                                           const FlowT<R>& f, const FluxLds<R>& x,
                                           const EdgeRec<R>& e0) {
  constexpr bool kF32 = sizeof(R) == 4;
  const int s0 = threadIdx.x, ML = x.ML;
  // FP32: the record of this thread's next edge is requested before this
  // edge is computed, so an edge beyond the first blockDim costs no round
  // trip of its own (FP64's 80-byte record would hold 20 more registers)
  constexpr bool kNext = kF32;
  EdgeRec<R> en{};
  if (kNext && kPre && s0 < d.nedge) en = e0;
  for (int s = threadIdx.x; s < d.nedge; s += blockDim.x) {
    EdgeRec<R> e;
    if constexpr (kNext) {
      e = (!kPre && s == s0) ? a.edges[d.edge0 + s] : en;
      if (s + (int)blockDim.x < d.nedge) en = a.edges[d.edge0 + s + blockDim.x];
    } else {
      e = (kPre && s == s0) ? e0 : a.edges[d.edge0 + s];
    }
    const EdgeView ev = edge_cells(e);
    if constexpr (!kF32) {
      if (ev.l < d.ncell) x.slots[4 * ev.l + ev.kl] = (unsigned short)(2 * s);
      if (ev.r >= 0 && ev.r < d.ncell) x.slots[4 * ev.r + ev.kr] = (unsigned short)(2 * s + 1);
    }
    // FP32 face slots of the two sides (the right one only for r >= 0)
    const int fl = ev.kl * ML + ev.l, fr = ev.kr * ML + ev.r;
    const R vx = edge_vx(e), vy = edge_vy(e);
    const EdgeNormal<R> en_ = edge_normal(e);
    const Vec4<R> WL = ld(x.W[ev.l]);
    Vec4<R> L = WL, Rr, Rraw;
    if (ev.r >= 0) {
      Rraw = ld(x.W[ev.r]);
      Rr = Rraw;
      if (a.order >= 2) {
        if constexpr (kF32) {
          L = ld(x.F[fl]);
          Rr = ld(x.F[fr]);
        } else {
          const int l = ev.l, r = ev.r;
          const Pair<R>* sG2 = reinterpret_cast<const Pair<R>*>(x.G);
          const Pair<R>* sB2 = reinterpret_cast<const Pair<R>*>(x.B);
          R dl[4], dr[4];
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const Pair<R> gl = sG2[q * ML + l], gr = sG2[q * ML + r];
            dl[q] = gl.a * e.rlx + gl.b * e.rly;
            dr[q] = gr.a * e.rrx + gr.b * e.rry;
          }
          const R ul[4] = {WL.d, WL.du, WL.dv, WL.de};
          const R ur[4] = {Rraw.d, Rraw.du, Rraw.dv, Rraw.de};
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const Pair<R> bl = sB2[q * ML + l], br = sB2[q * ML + r];
            dl[q] = limit(dl[q], ul[q], bl.a, bl.b);
            dr[q] = limit(dr[q], ur[q], br.a, br.b);
          }
          L = Vec4<R>{ul[0] + dl[0], ul[1] + dl[1], ul[2] + dl[2], ul[3] + dl[3]};
          Rr = Vec4<R>{ur[0] + dr[0], ur[1] + dr[1], ur[2] + dr[2], ur[3] + dr[3]};
        }
      }
    } else if (ev.r == -1) {
      Rraw = far_ghost(WL, vx, vy, f);
      Rr = Rraw;
    } else {
      Rraw = wall_ghost(WL, vx, vy, en_.vn);
      Rr = Rraw;
    }
    Vec4<R> F = roe_flux(L, Rr, en_.nx, en_.ny, f.gamma);
    F.d *= en_.vn;
    F.du *= en_.vn;
    F.dv *= en_.vn;
    F.de *= en_.vn;
    const R sr = edge_sr(en_, WL, Rraw, vx, vy, f.gamma);
    if constexpr (kF32) {
      // both slots are this edge's alone, and their states were read above
      x.F[fl] = st(F);
      x.S[fl] = sr;
      if (ev.r >= 0) {
        x.F[fr] = SVec4<R>{-F.d, -F.du, -F.dv, -F.de};
        x.S[fr] = sr;
      }
    } else {
      x.F[s] = st(F);
      x.S[s] = sr;
    }
  }
}

// Ordered gather of each tile cell's three edges, R and the local time
// step (kernel_Flux_e2c gpu_solver.cu:768-878): r += F on the edge's left
// side, r -= F on its right, over the cell's local edges 0, 1, 2.  Returns
// this thread's minimum of A*CFL/T (global time step mode, which stores no
// per-cell time step).  The local time step is stored as dt in FP64 and as
// dt / A in FP32 (the RK update's only use of dt is dt R / A, so FP32 folds
// the division by the area in here and the update reads no area).  ar0: this
// thread's first cell's (1 / A, A), read at kernel start.  FP64 finds each
// flux through the slots the flux phase entered (2 * edge + side); FP32 reads
// the cell's face slots, which hold the fluxes already signed for the cell.
template <typename R>
__device__ __forceinline__ R gather_cells(const TiledArgs<R>& a, const TileDesc& d, const FlowT<R>& f,
                                          const FluxLds<R>& x, SVec4<R>* Rout, R* dtout,
                                          Pair<R> ar0) {
  const int s0 = threadIdx.x;
  R tmin = R(3.0e38);
  for (int i = threadIdx.x; i < d.ncell; i += blockDim.x) {
    const int g = d.cell0 + i;
    R T = 0;
    R r0 = R(0.0), r1 = R(0.0), r2 = R(0.0), r3 = R(0.0);
    if constexpr (sizeof(R) == 4) {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const SVec4<R> F = x.F[k * x.ML + i];
        T += x.S[k * x.ML + i];
        r0 += F.d; r1 += F.du; r2 += F.dv; r3 += F.de;
      }
    } else {
      const uint2 sp = reinterpret_cast<const uint2*>(x.slots)[i];
      const int slk[3] = {(int)(sp.x & 0xffffu), (int)(sp.x >> 16), (int)(sp.y & 0xffffu)};
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int s = slk[k] >> 1;
        const SVec4<R> F = x.F[s];
        T += x.S[s];
        // sg * F is exact, so r + sg * F rounds as r + F or r - F does
        const R sg = (slk[k] & 1) ? R(-1.0) : R(1.0);
        r0 += sg * F.d; r1 += sg * F.du; r2 += sg * F.dv; r3 += sg * F.de;
      }
    }
    Rout[d.st0 + g] = SVec4<R>{r0, r1, r2, r3};
    const Pair<R> ar = (i == s0) ? ar0 : area_pair(a.cg + d.geo0, g);
    if (a.global_dt) {
      // only the minimum over the case is used (kernel_TimeStep)
      R v;
      if constexpr (sizeof(R) == 4)
        v = f.cfl / (T * ar.a);  // A CFL / T from the stored 1 / A
      else
        v = ar.b * f.cfl / T;
      tmin = fmin(tmin, v);
    } else if constexpr (sizeof(R) == 4) {
      // dt / A = min(A CFL / T, 0.1) / A = min(CFL / T, 0.1 / A)
      dtout[d.st0 + g] = fmin(f.cfl / T, R(0.1) * ar.a);
    } else {
      dtout[d.st0 + g] = fmin(ar.b * f.cfl / T, R(0.1));
    }
  }
  return tmin;
}

// Fixed-order sum of one double per thread over the workgroup (wave
// shuffles, then waves in order); the result is valid in thread 0.
// Its two halves, for a kernel that has a barrier between them anyway: each
// wave's butterfly sum into wsum[wave], then (after the barrier) the waves'
// sums in order, valid in thread 0.
__device__ __forceinline__ void wave_sums(double acc, double* wsum) {
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
}
__device__ __forceinline__ double sum_of_waves(const double* wsum) {
  double s = 0.0;
  for (int k = 0; k < (int)(blockDim.x + 63) / 64; k++) s += wsum[k];
  return s;
}
__device__ __forceinline__ double block_sum(double acc) {
  __shared__ double wsum[16];
  wave_sums(acc, wsum);
  __syncthreads();
  return threadIdx.x == 0 ? sum_of_waves(wsum) : 0.0;
}

// --------------------------------------------------------------------------
// Kernel A: gradient + flux + gather for one tile.
// LDS, FP64: W[ML] | grad[8][ML] | bounds[8][ML] | F[ME] | SR[ME] | gather slots[4][nt] (16-bit)
//      FP32: W[ML] | faces[3][ML] | SR[3][ML]
// --------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(512) flux_tile(TiledArgs<R> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int t = tile_of_block(blockIdx.x, a.ntiles);
  const TileDesc d = a.tiles[t];
  if (!a.ctl[d.icase].active) return;
  const FlowT<R> f = a.flow[d.icase];
  const int nt = d.ncell, nl = d.ncell + d.nhalo;
  const int ML = a.max_local, ME = a.max_edges;
  SVec4<R>* sW = reinterpret_cast<SVec4<R>*>(smem);
  FluxLds<R> x{sW, nullptr, nullptr, nullptr, nullptr, nullptr, ML};
  if constexpr (sizeof(R) == 4) {
    x.F = sW + ML;
    x.S = reinterpret_cast<R*>(x.F + 3 * ML);
  } else {
    x.G = reinterpret_cast<R*>(sW + ML);  // [4][ML] (d/dx, d/dy) pairs
    x.B = x.G + 8 * ML;                   // [4][ML] (max, min) pairs
    x.F = reinterpret_cast<SVec4<R>*>(x.B + 8 * ML);
    x.S = reinterpret_cast<R*>(x.F + ME);
    // [4][ncell] gather slots, 8-byte aligned (read as one 8-byte access per cell)
    x.slots = reinterpret_cast<unsigned short*>(
        smem + ((reinterpret_cast<unsigned char*>(x.S + ME) - smem + 7) & ~7));
  }
  const SVec4<R>* W = a.w_in + d.st0;

  // prefetch this thread's first edge record and its first cell's 1 / A;
  // their latency hides behind the gradient phase
  const int s0 = threadIdx.x;
  EdgeRec<R> e0{};
  if (s0 < d.nedge) e0 = a.edges[d.edge0 + s0];
  Pair<R> ar0{R(0), R(0)};
  if (s0 < nt) ar0 = area_pair(a.cg + d.geo0, d.cell0 + s0);

  for (int j = threadIdx.x; j < nl; j += blockDim.x) {
    const int g = j < nt ? d.cell0 + j : a.halo[d.halo0 + j - nt];
    const Vec4<R> Wi = ld(W[g]);
    sW[j] = st(Wi);
    if (a.order >= 2) {
      const int4 nb = a.nbr[d.geo0 + g];
      const int nbk[3] = {nb.x, nb.y, nb.z};
      Vec4<R> nbv[3];
#pragma unroll
      for (int k = 0; k < 3; k++) nbv[k] = ld(W[nbk[k] >= 0 ? nbk[k] : g]);
      cell_slopes<R>(j, Wi, a.cg[d.geo0 + g], nbk, nbv, f, x);
    }
  }
  __syncthreads();
  flux_edges<R>(a, d, f, x, e0);
  __syncthreads();
  const R tmin = gather_cells<R>(a, d, f, x, a.res, a.dt, ar0);
  if (a.global_dt) {
    const R m = block_min(tmin);
    if (threadIdx.x == 0) a.tile_min[t] = m;
  }
}

// Per-case minimum of the tile minima (global time step mode,
// kernel_TimeStep gpu_solver.cu:880-936 with the host's minimum over its
// blocks): one workgroup per case, each thread the minimum of a strided
// share of the tiles, then block_min.  The starting value 9999999 is the
// reference's.
constexpr int kDtMinThreads = 256;
template <typename R>
__global__ void __launch_bounds__(kDtMinThreads) global_dt_min(const CaseCtl* ctl, const R* tile_min,
                                                               R* dt_global) {
  const CaseCtl& c = ctl[blockIdx.x];
  if (!c.active) return;  // read by every thread; no other launch writes it meanwhile
  R m = R(9999999.0);
  for (int k = threadIdx.x; k < c.ntile; k += blockDim.x) m = fmin(m, tile_min[c.tile0 + k]);
  m = block_min(m);
  if (threadIdx.x == 0) dt_global[blockIdx.x] = m;
}

// --------------------------------------------------------------------------
// Kernel B: smoothing + RK update for one tile.  LDS: R1[nl]
// --------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(512) update_tile(TiledArgs<R> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int t = tile_of_block(blockIdx.x, a.ntiles);
  const TileDesc d = a.tiles[t];
  if (!a.ctl[d.icase].active) return;
  const FlowT<R> f = a.flow[d.icase];
  const int nt = d.ncell, nl = d.ncell + d.nhalo;
  SVec4<R>* sR1 = reinterpret_cast<SVec4<R>*>(smem);
  const SVec4<R>* Rs = a.res + d.st0;
  const R eps = f.eps;
  if (a.smoothing) {
    for (int j = threadIdx.x; j < nl; j += blockDim.x) {
      const int g = j < nt ? d.cell0 + j : a.halo[d.halo0 + j - nt];
      const int4 nb = a.nbr[d.geo0 + g];
      const int nbk[3] = {nb.x, nb.y, nb.z};
      const SVec4<R> ri = Rs[g];
      SVec4<R> v[3];
#pragma unroll
      for (int k = 0; k < 3; k++) v[k] = Rs[nbk[k] >= 0 ? nbk[k] : g];
      R s0 = R(0.0), s1 = R(0.0), s2 = R(0.0), s3 = R(0.0);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        if (nbk[k] >= 0) {
          s0 += v[k].d; s1 += v[k].du; s2 += v[k].dv; s3 += v[k].de;
        }
      }
      const R den = R(1.0) + eps * R(nb.w);
      sR1[j] = SVec4<R>{(s0 * eps + ri.d) / den, (s1 * eps + ri.du) / den,
                        (s2 * eps + ri.dv) / den, (s3 * eps + ri.de) / den};
    }
    __syncthreads();
  }
  double acc = 0.0;
  for (int i = threadIdx.x; i < nt; i += blockDim.x) {
    const int g = d.cell0 + i;
    SVec4<R> ri = Rs[g];
    if (a.smoothing) {
      const int4 nl4 = a.nloc[d.geo0 + g];
      const int lk[3] = {nl4.x, nl4.y, nl4.z};
      SVec4<R> v[3];
#pragma unroll
      for (int k = 0; k < 3; k++) v[k] = sR1[lk[k] >= 0 ? lk[k] : i];
      R s0 = R(0.0), s1 = R(0.0), s2 = R(0.0), s3 = R(0.0);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        if (lk[k] >= 0) {
          s0 += v[k].d; s1 += v[k].du; s2 += v[k].dv; s3 += v[k].de;
        }
      }
      const R den = R(1.0) + eps * R(a.nbr[d.geo0 + g].w);
      const R q0 = (s0 * eps + ri.d) / den, q1 = (s1 * eps + ri.du) / den;
      const R q2 = (s2 * eps + ri.dv) / den, q3 = (s3 * eps + ri.de) / den;
      if (fabs(ri.d) > fabs(q0)) ri.d = q0;
      if (fabs(ri.du) > fabs(q1)) ri.du = q1;
      if (fabs(ri.dv) > fabs(q2)) ri.dv = q2;
      if (fabs(ri.de) > fabs(q3)) ri.de = q3;
    }
    const R dtc = a.global_dt ? a.dt_global[d.icase] : a.dt[d.st0 + g];
    const R rk = a.rk * dtc;
    const SVec4<R> w0 = a.w0[d.st0 + g];
    SVec4<R> wn;
    if constexpr (sizeof(R) == 4) {
      // FP32: the local time step is stored as dt / A (see gather_cells); the
      // global one is divided by the area here, through the stored 1 / A
      const R s = a.global_dt ? rk * a.cg[d.geo0 + g].inv_area : rk;
      wn = SVec4<R>{w0.d - s * ri.d, w0.du - s * ri.du, w0.dv - s * ri.dv, w0.de - s * ri.de};
    } else {
      const R area = a.cg[d.geo0 + g].area;
      wn.d = w0.d - rk * ri.d / area;
      wn.du = w0.du - rk * ri.du / area;
      wn.dv = w0.dv - rk * ri.dv / area;
      wn.de = w0.de - rk * ri.de / area;
    }
    a.w_out[d.st0 + g] = wn;
    if (a.last_stage) {
      const double df = (double)wn.d - (double)w0.d;
      acc += df * df;
    }
  }
  if (a.last_stage) {
    const double s = block_sum(acc);
    if (threadIdx.x == 0) a.partial[t] = s;
  }
}

// --------------------------------------------------------------------------
// One kernel per RK stage (see the header comment).
// LDS: W[MW] | gid[MW4] | union { R[M4 + 1] | R1[MR + 1] ;  flux phase }, the flux phase
//   FP64: grad[8][ML] | bounds[8][ML] | F[ME] | SR[ME], and the gather slots[4][tile cells]
//         (16-bit) at slot_off: ending where gid ends when that leaves W's entries of tile +
//         ring 1 clear, else after SR;
//   FP32: faces[3][ML] | SR[3][ML]
// (ML: the slope row stride, S or max_local; MW4: MW rounded up to 4)
// R holds the previous stage's residual of the tile and rings 1..4, one load
// per cell, so that smoothing sweep 1 gathers from LDS through the 16-bit
// local neighbour table; a cell's CLog (its count of interior neighbours) is
// read off the same entry.  gid holds the cell id of each entry of the tile
// and rings 1..2, read by steps 3-4 instead of the ring table.  The residual
// buffers of steps 1-3 are dead once W is complete, so the gradient and flux
// buffers of steps 4-6 reuse their space (a smaller footprint lets more
// workgroups share a CU).
// --------------------------------------------------------------------------
// S > 0: the slope (FP64) or face (FP32) rows are S apart (S >= max_local), a
// compile-time stride, so a cell's LDS accesses share one address register
// with immediate offsets; S == 0: the stride is max_local.
template <typename R, int S>
__global__ void __launch_bounds__(512) stage_tile(TiledArgs<R> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int t = tile_of_block(blockIdx.x, a.ntiles);
  const TileDesc d = a.tiles[t];
  const CaseCtl& cc = a.ctl[d.icase];
  // The first launch of a run prepares every healthy case; later ones only
  // the cases still iterating.  `active` may change during a launch that
  // closes the previous step (the workgroup of the case's first tile writes
  // it at its end), so thread 0 reads it once and the workgroup acts on that
  // one value, shared through LDS.  With one case the flag is read once the
  // previous stage's residual has been requested (below), so that those loads
  // do not wait for the control block; the host stops launching once the
  // case stops.  With several, a stopped case's tiles leave before any load.
  __shared__ int s_active;
  __shared__ R s_wmin[16];  // global time step: the waves' minima
  if (!a.apply_update) {
    if (cc.status != FVM_OK) return;  // no launch writes it meanwhile
  } else if (a.early_exit) {
    if (threadIdx.x == 0) s_active = cc.active;
    __syncthreads();
    if (!s_active) return;
  }
  const FlowT<R> f = a.flow[d.icase];
  const int nt = d.ncell, NG = nt + d.nhalo, NW = NG + d.n2, NR = NW + d.n3, N4 = NR + d.n4;
  const int M4 = a.max_r4_local, MR = a.max_ring_local, MW = a.max_w_local, ME = a.max_edges;
  const int ML = S > 0 ? S : a.max_local;
  const int MW4 = (MW + 3) & ~3;
  SVec4<R>* sW = reinterpret_cast<SVec4<R>*>(smem);
  int* sGid = reinterpret_cast<int*>(sW + MW);
  SVec4<R>* sR = reinterpret_cast<SVec4<R>*>(sGid + MW4);
  SVec4<R>* sR1 = sR + M4 + 1;
  FluxLds<R> x{sW, nullptr, nullptr, nullptr, nullptr, nullptr, ML};
  if constexpr (sizeof(R) == 4) {
    x.F = sR;
    x.S = reinterpret_cast<R*>(x.F + 3 * ML);
  } else {
    x.G = reinterpret_cast<R*>(sR);
    x.B = x.G + 8 * ML;
    x.F = reinterpret_cast<SVec4<R>*>(x.B + 8 * ML);
    x.S = reinterpret_cast<R*>(x.F + ME);
    // [4][ncell] gather slots, 8-byte aligned (one 8-byte read per cell):
    // where the host found room, over the tail of W (rings 2, which only the
    // gradient reads) and the cell ids (read by sweep 2 and the gradient),
    // both dead in the flux and gather phases that use the table; else after SR
    x.slots = reinterpret_cast<unsigned short*>(smem + a.slot_off);
  }
  const int cur = cc.w0_cur ^ a.parity;
  const SVec4<R>* W0 = (cur ? a.w0b : a.w0a) + d.st0;
  SVec4<R>* W0n = (cur ? a.w0a : a.w0b) + d.st0;
  const int2* lnb = a.lnb + d.lnb0;
  const CellGeom<R>* cg = a.cg + d.geo0;
  auto gid = [&](int j) { return j < nt ? d.cell0 + j : a.rings[d.ring0 + j - nt]; };

  // this thread's first edge record is loaded here, its latency hidden by
  // the sweeps, in FP32; the 80-byte FP64 record would hold 20 registers
  // through them and cost a wave per SIMD, so FP64 loads it in the flux loop
  constexpr bool kPreEdge = sizeof(R) == 4;
  const int s0 = threadIdx.x;
  EdgeRec<R> e0{};
  if (kPreEdge && s0 < d.nedge) e0 = a.edges[d.edge0 + s0];
  Pair<R> ar0{R(0), R(0)};
  if (s0 < nt) ar0 = area_pair(cg, d.cell0 + s0);

  // this thread's first local-neighbour entry, read here with the other
  // prefetches and used by sweeps 1 and 2 and the gradient (a thread takes
  // the same entry in each)
  int2 lp0 = make_int2(0, 0);
  const bool have_lp0 = a.apply_update && a.smoothing && s0 < NR;
  if (have_lp0) lp0 = lnb[s0];
  double acc = 0.0;
  if (a.apply_update) {
    const SVec4<R>* Rin = a.r_in + d.st0;
    const R eps = f.eps;
    // 1. the previous stage's residual of the tile and rings 1..4 (1..2
    //    without smoothing)
    const int n1 = a.smoothing ? N4 : NW;
    // Global time step: a thread's first two of the case's tile minima of the
    // previous stage are requested here, ahead of the residual staging, and
    // combined after it (further ones, for cases of more than 2 blockDim
    // tiles, after the staging barrier).  FP32 also requests the 1 / A of its
    // two staged entries with their residuals, for the update.
    R gm0 = R(9999999.0), gm1 = R(9999999.0);  // the reference's initial minimum
    R ia0 = R(1), ia1 = R(1);
    if (a.global_dt) {
      const int n = cc.ntile, t0 = cc.tile0, k2 = s0 + (int)blockDim.x;
      if (s0 < n) gm0 = a.tile_min_in[t0 + s0];
      if (k2 < n) gm1 = a.tile_min_in[t0 + k2];
    }
    // The smoothing sums `if (neighbour is a cell) q += R[neighbour]` read a
    // boundary side from a slot past the end of R (and of R1) that holds -0:
    // adding -0 leaves every q unchanged (+0 + -0 = +0, -0 + -0 = -0, and q
    // is never denormal under FP32 flush-to-zero), so the sums keep the
    // reference's bits and order without a select per component.
    if (threadIdx.x == 0) {
      sR[M4] = SVec4<R>{R(-0.0), R(-0.0), R(-0.0), R(-0.0)};
      sR1[MR] = SVec4<R>{R(-0.0), R(-0.0), R(-0.0), R(-0.0)};
    }
    // A thread's entries s0 and s0 + blockDim are staged together: both ring
    // ids are requested, then both residuals, so that the entries beyond the
    // first blockDim (rings 3-4 of larger tiles) cost no round trips of their
    // own.  Entries beyond 2 blockDim (very large tiles) follow one by one.
    {
      const int j = s0, j2 = s0 + (int)blockDim.x;
      // an entry past the end reads the tile's first cell (in bounds, unused),
      // so that both residual loads are unconditional and in flight together
      int rid = d.cell0, rid2 = d.cell0;
      if (j < n1 && j >= nt) rid = a.rings[d.ring0 + j - nt];
      if (j2 < n1 && j2 >= nt) rid2 = a.rings[d.ring0 + j2 - nt];
      const int g = j < nt ? d.cell0 + j : rid;
      const int g2 = j2 < nt ? d.cell0 + j2 : rid2;
      const SVec4<R> r = Rin[g];
      const SVec4<R> r2 = Rin[g2];
      if constexpr (sizeof(R) == 4) {
        if (a.global_dt) {
          ia0 = area_pair(cg, g).a;
          ia1 = area_pair(cg, g2).a;
        }
      }
      // both loads are issued here, before the stores that use them (left to
      // itself the compiler moves each load into its store's branch, where
      // the second waits for the first)
      __asm__ volatile("" : : "v"(r.d), "v"(r.du), "v"(r.dv), "v"(r.de), "v"(r2.d), "v"(r2.du),
                       "v"(r2.dv), "v"(r2.de));
      if (j < n1) {
        sR[j] = r;
        if (j < NW) sGid[j] = g;
      }
      if (j2 < n1) {
        sR[j2] = r2;
        if (j2 < NW) sGid[j2] = g2;
      }
    }
    for (int j = s0 + 2 * (int)blockDim.x; j < n1; j += blockDim.x) {
      const int g = gid(j);
      sR[j] = Rin[g];
      if (j < NW) sGid[j] = g;
    }
    if (!a.early_exit && threadIdx.x == 0) s_active = cc.active;
    __syncthreads();
    if (!s_active) return;
    if (a.global_dt) {
      const int n = cc.ntile, t0 = cc.tile0;
      for (int k = s0 + 2 * (int)blockDim.x; k < n; k += blockDim.x)
        gm1 = fmin(gm1, a.tile_min_in[t0 + k]);
    }
    // 2. smoothing sweep 1 on the tile and rings 1..3
    if (a.smoothing) {
      auto sweep1 = [&](int j, int2 lp) {
        const int4 ln = unpack_tri16(lp);
        const int lk[3] = {ln.x, ln.y, ln.z};
        const SVec4<R> ri = sR[j];
        SVec4<R> v[3];
#pragma unroll
        for (int k = 0; k < 3; k++) v[k] = sR[side_slot(lk[k], M4)];
        R q0 = R(0.0), q1 = R(0.0), q2 = R(0.0), q3 = R(0.0);
#pragma unroll
        for (int k = 0; k < 3; k++) {
          q0 += v[k].d; q1 += v[k].du; q2 += v[k].dv; q3 += v[k].de;
        }
        const R den = R(1.0) + eps * R(ln.w);  // ln.w: CLog
        sR1[j] = SVec4<R>{(q0 * eps + ri.d) / den, (q1 * eps + ri.du) / den,
                          (q2 * eps + ri.dv) / den, (q3 * eps + ri.de) / den};
      };
      // the neighbour entry of s0 + blockDim is requested before entry s0
      // (whose entry was read at kernel start) is computed
      {
        const int j2 = s0 + (int)blockDim.x;
        int2 lp2 = make_int2(0, 0);
        if (j2 < NR) lp2 = lnb[j2];
        if (s0 < NR) sweep1(s0, lp0);
        if (j2 < NR) sweep1(j2, lp2);
      }
      for (int j = s0 + 2 * (int)blockDim.x; j < NR; j += blockDim.x) sweep1(j, lnb[j]);
    }
    // Global time step (kernel_TimeStep, gpu_solver.cu:880-936): the
    // previous stage's minimum of A*CFL/T over the case, from its tiles'
    // minima (requested above, combined here: each wave's minimum, then the
    // waves' minima through LDS after the barrier).  A minimum does not depend on
    // the order it is taken in, so every tile of the case gets the same value.
    if (a.global_dt) {
      R m = fmin(gm0, gm1);
      m = wave_min(m);
      if ((threadIdx.x & 63) == 0) s_wmin[threadIdx.x >> 6] = m;
    }
    if (a.smoothing || a.global_dt) __syncthreads();
    R dtg = R(0);
    if (a.global_dt) {
      dtg = s_wmin[0];
      for (int k = 1; k < (int)(blockDim.x >> 6); k++) dtg = fmin(dtg, s_wmin[k]);
    }
    // 3. sweep 2, selection and the RK update on the tile and rings 1..2
    for (int j = threadIdx.x; j < NW; j += blockDim.x) {
      const int g = sGid[j];
      // FP32 local time step: dt / A; global: the case's dt, and FP32 then
      // multiplies by 1 / A
      R dtv;
      if (a.global_dt) {
        dtv = dtg;
        if constexpr (sizeof(R) == 4)
          dtv *= j == s0 ? ia0 : (j == s0 + (int)blockDim.x ? ia1 : area_pair(cg, g).a);
      } else {
        dtv = a.dt_in[d.st0 + g];
      }
      R area = R(1);
      if constexpr (sizeof(R) == 8) {
        if (j == s0 && j < nt)
          area = ar0.b;  // this thread's tile cell, read at kernel start
        else
          area = cg[g].area;
      }
      const SVec4<R> w0 = W0[g];
      SVec4<R> ri = sR[j];
      if (a.smoothing) {
        int2 lp;
        if (j == s0)
          lp = lp0;
        else
          lp = lnb[j];
        const int4 ln = unpack_tri16(lp);
        const int lk[3] = {ln.x, ln.y, ln.z};
        SVec4<R> v[3];
#pragma unroll
        for (int k = 0; k < 3; k++) v[k] = sR1[side_slot(lk[k], MR)];
        R q0 = R(0.0), q1 = R(0.0), q2 = R(0.0), q3 = R(0.0);
#pragma unroll
        for (int k = 0; k < 3; k++) {
          q0 += v[k].d; q1 += v[k].du; q2 += v[k].dv; q3 += v[k].de;
        }
        const R den = R(1.0) + eps * R(ln.w);  // ln.w: CLog
        const R p0 = (q0 * eps + ri.d) / den, p1 = (q1 * eps + ri.du) / den;
        const R p2 = (q2 * eps + ri.dv) / den, p3 = (q3 * eps + ri.de) / den;
        if (fabs(ri.d) > fabs(p0)) ri.d = p0;
        if (fabs(ri.du) > fabs(p1)) ri.du = p1;
        if (fabs(ri.dv) > fabs(p2)) ri.dv = p2;
        if (fabs(ri.de) > fabs(p3)) ri.de = p3;
      }
      const R rk = a.rk * dtv;
      SVec4<R> wn;
      if constexpr (sizeof(R) == 4) {
        wn = SVec4<R>{w0.d - rk * ri.d, w0.du - rk * ri.du, w0.dv - rk * ri.dv, w0.de - rk * ri.de};
      } else {
        wn.d = w0.d - rk * ri.d / area;
        wn.du = w0.du - rk * ri.du / area;
        wn.dv = w0.dv - rk * ri.dv / area;
        wn.de = w0.de - rk * ri.de / area;
      }
      sW[j] = wn;
      if (a.close_step && j < nt) {
        W0n[g] = wn;
        const double df = (double)wn.d - (double)w0.d;
        acc += df * df;
      }
    }
  } else {
    for (int j = threadIdx.x; j < NW; j += blockDim.x) {
      const int g = gid(j);
      sW[j] = W0[g];
      sGid[j] = g;
    }
  }
  // the tile's FP64 partial of (rho - rho0)^2 is complete once the update
  // is: its wave sums go before this barrier, the tile's sum after it
  __shared__ double wpart[16];
  const bool partial_here = a.close_step && a.apply_update;
  if (partial_here) wave_sums(acc, wpart);
  __syncthreads();
  if (partial_here && threadIdx.x == 0) a.partial[t] = sum_of_waves(wpart);
  // 4. gradient and bounds (FP64) or face states (FP32) of the tile and ring 1
  //    from the updated state
  if (a.order >= 2) {
    for (int j = threadIdx.x; j < NG; j += blockDim.x) {
      int2 lp;
      if (j == s0 && have_lp0)
        lp = lp0;
      else
        lp = lnb[j];
      const int4 ln = unpack_tri16(lp);
      const CellGeom<R> cgj = cg[sGid[j]];
      const int nbk[3] = {ln.x, ln.y, ln.z};
      Vec4<R> nbv[3];
#pragma unroll
      for (int k = 0; k < 3; k++) nbv[k] = ld(sW[nbk[k] >= 0 ? nbk[k] : j]);
      cell_slopes<R>(j, ld(sW[j]), cgj, nbk, nbv, f, x);
    }
  }
  __syncthreads();
  // 5. flux and spectral radius of the tile's edges
  flux_edges<R, kPreEdge>(a, d, f, x, e0);
  __syncthreads();
  // 6. ordered gather: this stage's residual and time step (global time
  //    step: the tile's minimum of A*CFL/T)
  const R tmin = gather_cells<R>(a, d, f, x, a.res, a.dt, ar0);
  if (a.global_dt) {
    const R m = block_min(tmin);
    if (threadIdx.x == 0) a.tile_min[t] = m;
  }
  // the previous step's close, by the workgroup of the case's first tile once
  // its own tile is done.  The other tiles of this launch read `active` only
  // (a tile that sees the old value computes a residual that the next launch,
  // seeing the case stopped, ignores) and take the step-start buffer from the
  // launch's parity, not from the control block.
  if (a.close_prev && t == cc.tile0) {
    __syncthreads();
    close_case(a.ctl[d.icase], a.partial, a.hist, d.icase);
  }
}

}  // namespace fvm
