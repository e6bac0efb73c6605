// Internal data structures of libfvm_hip.so: host topology, tiled device
// layout, and the kernel argument blocks shared by the launch code and the
// kernels.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fvm_abi.h"

namespace fvm {

// Error carrier used inside the library; the C-ABI turns it into a status.
struct Error {
  fvm_status status;
  std::string msg;
};
void set_last_error(const std::string& msg);

// ---------------------------------------------------------------------------
// Topology and FP64 geometry in the reference's conventions (0-based storage;
// edge order after the interior / far / wall partition).
// ---------------------------------------------------------------------------
struct Topology {
  int nnode = 0, ncell = 0, nedge = 0, nwall = 0, nfar = 0;
  std::vector<double> x, y;          // [nnode]
  std::vector<int> pt;               // [3*ncell] 0-based node ids
  std::vector<double> area, cx, cy;  // [ncell]
  std::vector<int> celledge;         // [3*ncell] 0-based edge id
  std::vector<int> neicell;          // [3*ncell] 0-based cell, or -1 far / -2 wall
  std::vector<int> clog;             // [ncell]
  std::vector<int> elog, left, right, farid, wallid, node1, node2;  // [nedge]; right -1 on boundary,
                                                                     // farid/wallid 1-based, nodes 0-based
  std::vector<double> vx, vy, vn, midx, midy;                        // [nedge]
  std::vector<int> wall_edge;        // [nwall] edge id in wall-id order
};

// Throws Error.  Rejects out-of-range node ids, non-manifold edges and
// non-positive cell areas.
Topology build_topology(const fvm_mesh_desc& mesh);

// ---------------------------------------------------------------------------
// Tiled layout used by the tiled kernels.  Cells are renumbered by recursive
// coordinate bisection into tiles of at most tile_cells cells; tile t owns
// the contiguous new ids [tile_cell0[t], tile_cell0[t+1]).  Renumbering is a
// storage permutation only: every edge keeps the reference's left/right
// orientation and every cell keeps its local edge order, so the arithmetic is
// unchanged.
// ---------------------------------------------------------------------------
struct TiledMesh {
  int ncell = 0, ntile = 0, tile_cells = 0;
  std::vector<int> new_to_old, old_to_new;  // [ncell]
  // per cell, new numbering
  std::vector<double> gvx, gvy;   // [3*ncell] local edge vectors (reference gradient form)
  std::vector<double> inv_area;   // [ncell] 1.0 / area
  std::vector<double> area;       // [ncell]
  std::vector<int> nbr;           // [3*ncell] new id, or -1 far / -2 wall
  std::vector<int> clog;          // [ncell]
  std::vector<int> slot;          // [3*ncell] 2*(tile edge index) + side (0 left, 1 right)
  std::vector<int> nloc;          // [3*ncell] tile-local index of interior neighbour, -1 otherwise
  // per tile
  std::vector<int> tile_cell0;    // [ntile+1]
  std::vector<int> tile_halo0;    // [ntile+1] into halo
  std::vector<int> halo;          // new ids of first-ring cells outside their tile
  std::vector<int> tile_edge0;    // [ntile+1] into the tile edge records
  // tile edge records
  std::vector<int> e_l, e_r;      // tile-local left / right (right -1 far, -2 wall)
  std::vector<int> e_kl, e_kr;    // the edge's local index (0..2) in its left / right cell
  std::vector<double> e_nx, e_ny, e_vn, e_vx, e_vy, e_rlx, e_rly, e_rrx, e_rry;
  int max_local = 0, max_edges = 0;  // largest tile-local cell and edge counts
  int max_tile = 0;                  // largest tile's own cell count
  // Rings for the one-kernel-per-stage path: ring 1 is the halo above, ring 2
  // the interior neighbours of ring 1 outside the tile and ring 1, rings 3 and
  // 4 likewise.  The tile-local index space is
  // [tile | ring 1 | ring 2 | ring 3 | ring 4].
  std::vector<int> tile_ring0;    // [ntile+1] into rings (rings 1..4 of each tile)
  std::vector<int> ring_n1, ring_n2, ring_n3, ring_n4;  // [ntile]
  std::vector<int> rings;         // new ids of ring 1..4 cells, per tile
  std::vector<int> tile_lnb0;     // [ntile+1] into lnb (one entry per tile/ring-1..3 cell)
  std::vector<int> lnb;           // 3 per entry: tile-local index of each interior neighbour,
                                  // -1 far field, -2 wall
  int max_r4_local = 0;           // largest tile + rings 1..4 count
  int max_ring_local = 0;         // largest tile + rings 1..3 count
  int max_w_local = 0;            // largest tile + rings 1..2 count
  // wall edges in wall-id order (forces)
  std::vector<int> w_cell;        // new id of the wall edge's cell
  std::vector<double> w_vx, w_vy, w_vn;
};

TiledMesh build_tiled(const Topology& topo, int tile_cells);

// ---------------------------------------------------------------------------
// Device records and kernel arguments.
// ---------------------------------------------------------------------------
// Per cell: the gradient's three local edge vectors (gx, gy of side k) and
// the area.  FP64 stores all three; FP32 stores two, scaled by 0.5 / area:
// the three vectors of a triangle's sides sum to zero (x1 - x2 + x2 - x0 +
// x0 - x1), so the third is minus the sum of the first two.  20 instead of
// 32 bytes per cell, read for the tile and its first ring every launch.
template <typename R>
struct alignas(16) CellGeom {
  R vx[3], vy[3], inv_area, area;
  __host__ __device__ R gx(int k) const { return vx[k]; }
  __host__ __device__ R gy(int k) const { return vy[k]; }
  __host__ __device__ void set(int k, double x, double y) {
    vx[k] = (R)x;
    vy[k] = (R)y;
  }
  __host__ __device__ void set_area(double a) { area = (R)a; }
};
// FP32 keeps 1 / A only (20 bytes): its time step is formed as dt / A =
// min(CFL / T, 0.1 / A), and A CFL / T, the global time step's operand, as
// CFL / (T / A).
template <>
struct alignas(4) CellGeom<float> {
  float vx[2], vy[2], inv_area;
  __host__ __device__ float gx(int k) const { return k < 2 ? vx[k] : -(vx[0] + vx[1]); }
  __host__ __device__ float gy(int k) const { return k < 2 ? vy[k] : -(vy[0] + vy[1]); }
  __host__ __device__ void set(int k, double x, double y) {
    if (k < 2) {
      vx[k] = (float)x;
      vy[k] = (float)y;
    }
  }
  __host__ __device__ void set_area(double) {}
};

template <typename R>
struct EdgeGeom;                // per tile edge (see fvm_tiled.h)

struct TileDesc {               // one (case, mesh tile) pair of a batch
  int icase;
  int geo0;                     // offset of the case's mesh in the per-cell geometry arrays
  int st0;                      // offset of the case in the per-cell state arrays
  int cell0, ncell;             // mesh-local new ids owned by the tile
  int halo0, nhalo;             // range in the halo array (global)
  int edge0, nedge;             // range in the edge record array (global)
  int ring0, n2, n3, n4;        // rings 1..4 list offset (global); ring sizes 2, 3 and 4
  int lnb0;                     // local-neighbour table offset (global)
};

struct CaseCtl {                // per-case run control, device resident
  int active;                   // 1 while step < max_steps and resd > epsl
  int step;                     // steps done
  int status;                   // fvm_status of the case
  int tile0, ntile;             // the case's tiles in the batch tile list
  int max_steps;
  double epsl;
  double epsm;                  // Resd normaliser (drou of step 1)
  double resd;                  // last Resd
  int w0_cur;                   // which of the two step-start state buffers is current
  int pad_;
};

// Where each step's Resd goes: row icase of a buffer of `stride` doubles per
// case, device resident.  A step close reads it when it writes, so a longer
// history (a new buffer) changes this record and not the captured graphs.
struct HistDesc {
  double* base;
  int stride;
  int pad_;
};

}  // namespace fvm
