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

template <typename R>
inline size_t flux_tile_lds(int max_local, int max_edges) {
  return (size_t)max_local * (sizeof(SVec4<R>) + 16 * sizeof(R)) +
         (size_t)max_edges * (sizeof(SVec4<R>) + sizeof(R));
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

// max_local: the slope row stride (stage_slope_stride's value, or max_local
// when that is 0).
template <typename R>
inline size_t stage_tile_lds(int max_r4_local, int max_ring_local, int max_w_local, int max_local,
                             int max_edges) {
  // W[MW] | gid[MW rounded up to 4] | union { R[M4 + 1] | R1[MR + 1] ; grad + bounds + F + SR }
  const size_t res = (size_t)(max_r4_local + max_ring_local + 2) * sizeof(SVec4<R>);
  const size_t flux = flux_tile_lds<R>(max_local, max_edges) - (size_t)max_local * sizeof(SVec4<R>);
  return (size_t)max_w_local * sizeof(SVec4<R>) + (size_t)((max_w_local + 3) & ~3) * sizeof(int) +
         (res > flux ? res : flux);
}

}  // namespace fvm
