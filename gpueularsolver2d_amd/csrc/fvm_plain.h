// Plain FP64 path: the reference's kernel structure (one kernel per phase,
// per-edge flux slots in global memory, reference cell and edge numbering),
// written for gfx950.  It is the FP64 baseline the tiled kernels are measured
// against, and a second, independent GPU implementation of the same
// arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include "fvm_internal.h"
#include "fvm_numerics.h"

namespace fvm {

struct PlainDev {
  int C = 0, E = 0, nfar = 0, nwall = 0, nblk64 = 0;
  // state: the conservatives of the reference's W_S (P, T and A, which it also
  // stores, are recomputed where they are needed)
  double *w[4] = {};
  double *w0[4] = {};
  double *farw[4] = {}, *wallw[4] = {};
  // cells
  double *gvx[3] = {}, *gvy[3] = {};   // per-cell edge vectors from node coordinates
  int *celledge[3] = {}, *neicell[3] = {}, *clog = nullptr;
  double *area = nullptr, *cx = nullptr, *cy = nullptr;
  double *grad[8] = {}, *umax[4] = {}, *umin[4] = {};
  // edges (interior | far | wall order)
  int *elog = nullptr, *left = nullptr, *right = nullptr, *farid = nullptr, *wallid = nullptr;
  double *vx = nullptr, *vy = nullptr, *vn = nullptr, *midx = nullptr, *midy = nullptr;
  double *rel[4] = {}, *rer[4] = {}, *dtl = nullptr, *dtr = nullptr;
  double *res[4] = {}, *r1[4] = {}, *dt = nullptr, *tsum = nullptr;
  double *blockmin = nullptr, *dtg = nullptr;
  double *partial = nullptr;            // per 64-cell block
};

// Launches one time step (W0 copy, four stages) of case `icase` on `stream`.
void plain_launch_step(const PlainDev& p, CaseCtl* ctl, int icase, const FlowT<double>* flow,
                       int order, int smoothing, int global_dt, hipStream_t stream);

}  // namespace fvm
