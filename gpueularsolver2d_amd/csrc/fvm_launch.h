// Launch wrappers of the tiled kernels, instantiated once per working
// precision (fvm_tiled64.hip, fvm_tiled32.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "fvm_tiled.h"

namespace fvm {

template <typename R>
hipError_t launch_flux_tile(const TiledArgs<R>& a, int threads, size_t lds, hipStream_t s);
template <typename R>
hipError_t launch_update_tile(const TiledArgs<R>& a, int threads, size_t lds, hipStream_t s);
template <typename R>
hipError_t launch_stage_tile(const TiledArgs<R>& a, int threads, size_t lds, hipStream_t s);
// finalize_kernel over ncase control blocks, built with R's options.
template <typename R>
hipError_t launch_finalize(CaseCtl* ctl, const double* partial, const HistDesc* hist,
                           CaseCtl* host_out, int ncase, hipStream_t s);
template <typename R>
hipError_t launch_global_dt_min(const CaseCtl* ctl, const R* tile_min, R* dt_global, int ncase,
                                hipStream_t s);
// Largest dynamic LDS the two tile kernels may use (bytes).
template <typename R>
hipError_t set_tile_lds_limit(size_t bytes);
// Workgroups of one tile kernel that a CU holds at once with this workgroup
// size and dynamic LDS (registers and LDS both counted).
enum TileKernel { kFluxTile = 0, kUpdateTile = 1, kStageTile = 2 };
// max_local selects the stage kernel's instantiation (see stage_slope_stride).
template <typename R>
hipError_t tile_blocks_per_cu(TileKernel k, int threads, size_t lds, int max_local, int* blocks);

// FP64: max_tile is the largest tile's cell count (the gather slot table,
// four 16-bit entries per tile cell, follows the edge arrays, 8-byte
// aligned).  FP32: W, then three face states and three spectral radii per
// cell of the tile and ring 1.
template <typename R>
inline size_t flux_tile_lds(int max_local, int max_edges, int max_tile) {
  if constexpr (sizeof(R) == 4)
    return (size_t)max_local * (4 * sizeof(SVec4<R>) + 3 * sizeof(R));
  const size_t sr_end = (size_t)max_local * (sizeof(SVec4<R>) + 16 * sizeof(R)) +
                        (size_t)max_edges * (sizeof(SVec4<R>) + sizeof(R));
  return ((sr_end + 7) & ~size_t(7)) + (size_t)max_tile * 4 * 2;
}
template <typename R>
inline size_t update_tile_lds(int max_local) {
  return (size_t)max_local * sizeof(SVec4<R>);
}
// Slope row stride of the stage kernel instantiation used for tiles whose
// largest local count (tile + ring 1) is max_local: the smallest compiled
// stride that holds it, or 0 (the runtime stride max_local).
inline int stage_slope_stride(int max_local) {
  for (int s : {192, 320, 448})
    if (max_local <= s) return s;
  return 0;
}

// max_local: the slope (FP64) or face (FP32) row stride (stage_slope_stride's
// value, or max_local when that is 0); max_tile: the largest tile's cell count.
//
// FP64: the gather slot table (8 bytes per tile cell) goes where it costs no
// LDS if it can: ending at the end of the cell-id array, over W's ring-2
// entries and the cell ids, which the flux and gather phases do not read; it
// must then start at or after W's tile + ring 1 entries (max_local of them,
// the true count here).  Otherwise it follows SR in the union.  FP32 has no
// slot table: its flux phase holds three face slots and three spectral radii
// per row entry.
struct StageLds {
  size_t bytes;     // dynamic LDS per workgroup
  size_t slot_off;  // byte offset of the gather slot table (FP64)
};
template <typename R>
inline StageLds stage_tile_lds(int max_r4_local, int max_ring_local, int max_w_local, int max_local,
                               int true_max_local, int max_edges, int max_tile) {
  // W[MW] | gid[MW rounded up to 4] | union { R[M4 + 1] | R1[MR + 1] ; grad + bounds + F + SR }
  const size_t head = (size_t)max_w_local * sizeof(SVec4<R>) + (size_t)((max_w_local + 3) & ~3) * sizeof(int);
  const size_t res = (size_t)(max_r4_local + max_ring_local + 2) * sizeof(SVec4<R>);
  if constexpr (sizeof(R) == 4) {
    const size_t faces = (size_t)max_local * 3 * (sizeof(SVec4<R>) + sizeof(R));
    return StageLds{head + (res > faces ? res : faces), 0};
  }
  const size_t table = (size_t)max_tile * 4 * 2;
  const size_t flux_end = (size_t)max_local * 16 * sizeof(R) + (size_t)max_edges * (sizeof(SVec4<R>) + sizeof(R));
  if (head >= table && head - table >= (size_t)true_max_local * sizeof(SVec4<R>)) {
    const size_t off = (head - table) & ~size_t(7);
    if (off >= (size_t)true_max_local * sizeof(SVec4<R>))
      return StageLds{head + (res > flux_end ? res : flux_end), off};
  }
  const size_t off = head + ((flux_end + 7) & ~size_t(7));
  const size_t union_bytes = ((flux_end + 7) & ~size_t(7)) + table;
  return StageLds{head + (res > union_bytes ? res : union_bytes), off};
}

}  // namespace fvm
