// Plain FP64 kernels: the reference's phase structure, one thread per cell or
// per edge (gpu_solver.cu:36-1020, 1633-1728; driver :2200-2433).
#include "fvm_plain.h"

namespace fvm {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ Vec4<double> load_w(const PlainDev& p, int c) {
  return Vec4<double>{p.w[0][c], p.w[1][c], p.w[2][c], p.w[3][c]};
}

// kernel_Boundary, gpu_solver.cu:137-223
__global__ void k_boundary(PlainDev p, const CaseCtl* ctl, const FlowT<double>* flow) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (!ctl->active || e >= p.E) return;
  const int lg = p.elog[e];
  if (lg == 0) return;
  const Vec4<double> wd = load_w(p, p.left[e]);
  if (lg == 1) {
    const Vec4<double> g = far_ghost(wd, p.vx[e], p.vy[e], *flow);
    const int id = p.farid[e] - 1;
    p.farw[0][id] = g.d; p.farw[1][id] = g.du; p.farw[2][id] = g.dv; p.farw[3][id] = g.de;
  } else {
    const Vec4<double> g = wall_ghost(wd, p.vx[e], p.vy[e], p.vn[e]);
    const int id = p.wallid[e] - 1;
    p.wallw[0][id] = g.d; p.wallw[1][id] = g.du; p.wallw[2][id] = g.dv; p.wallw[3][id] = g.de;
  }
}

__device__ __forceinline__ Vec4<double> across(const PlainDev& p, int c, int k) {
  const int nc = p.neicell[k][c];
  if (nc >= 0) return load_w(p, nc);
  const int e = p.celledge[k][c];
  if (nc == -1) {
    const int id = p.farid[e] - 1;
    return Vec4<double>{p.farw[0][id], p.farw[1][id], p.farw[2][id], p.farw[3][id]};
  }
  const int id = p.wallid[e] - 1;
  return Vec4<double>{p.wallw[0][id], p.wallw[1][id], p.wallw[2][id], p.wallw[3][id]};
}

// kernel_Gradient, gpu_solver.cu:225-383
__global__ void k_gradient(PlainDev p, const CaseCtl* ctl) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (!ctl->active || c >= p.C) return;
  const Vec4<double> wi = load_w(p, c);
  const double ui[4] = {wi.d, wi.du, wi.dv, wi.de};
  double gx[4] = {0.0, 0.0, 0.0, 0.0}, gy[4] = {0.0, 0.0, 0.0, 0.0}, mx[4], mn[4];
  for (int k = 0; k < 3; k++) {
    const Vec4<double> cv = across(p, c, k);
    const double cc[4] = {cv.d, cv.du, cv.dv, cv.de};
    const double vxk = p.gvx[k][c], vyk = p.gvy[k][c];
    for (int q = 0; q < 4; q++) {
      if (k == 0) {
        mx[q] = cc[q];
        mn[q] = cc[q];
      } else {
        if (cc[q] < mn[q]) mn[q] = cc[q];
        if (cc[q] > mx[q]) mx[q] = cc[q];
      }
    }
    for (int q = 0; q < 4; q++) gx[q] += 0.5 * (ui[q] + cc[q]) * vyk;
    for (int q = 0; q < 4; q++) gy[q] += 0.5 * (ui[q] + cc[q]) * vxk;
  }
  const double ia = 1.0 / p.area[c];
  for (int q = 0; q < 4; q++) {
    p.grad[q][c] = gx[q] * ia;
    p.grad[4 + q][c] = gy[q] * ia;
    p.umax[q][c] = fmax(ui[q], mx[q]);
    p.umin[q][c] = fmin(ui[q], mn[q]);
  }
}

__device__ __forceinline__ Vec4<double> reconstruct(const PlainDev& p, int c, double midx,
                                                    double midy) {
  const Vec4<double> w = load_w(p, c);
  const double u[4] = {w.d, w.du, w.dv, w.de};
  const double r0 = midx - p.cx[c], r1 = midy - p.cy[c];
  double out[4];
  for (int q = 0; q < 4; q++) {
    const double dl = p.grad[q][c] * r0 + p.grad[4 + q][c] * r1;
    out[q] = u[q] + limit(dl, u[q], p.umax[q][c], p.umin[q][c]);
  }
  return Vec4<double>{out[0], out[1], out[2], out[3]};
}

// kernel_Flux2, gpu_solver.cu:651-764
__global__ void k_flux(PlainDev p, const CaseCtl* ctl, const FlowT<double>* flow, int order) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (!ctl->active || e >= p.E) return;
  const int lg = p.elog[e], L = p.left[e], Rc = p.right[e];
  Vec4<double> LW, RW;
  if (lg == 0) {
    if (order >= 2) {
      LW = reconstruct(p, L, p.midx[e], p.midy[e]);
      RW = reconstruct(p, Rc, p.midx[e], p.midy[e]);
    } else {
      LW = load_w(p, L);
      RW = load_w(p, Rc);
    }
  } else {
    LW = load_w(p, L);
    if (lg == 1) {
      const int id = p.farid[e] - 1;
      RW = Vec4<double>{p.farw[0][id], p.farw[1][id], p.farw[2][id], p.farw[3][id]};
    } else {
      const int id = p.wallid[e] - 1;
      RW = Vec4<double>{p.wallw[0][id], p.wallw[1][id], p.wallw[2][id], p.wallw[3][id]};
    }
  }
  const double vectorn = p.vn[e];
  Vec4<double> F = roe_flux(LW, RW, p.vy[e] / vectorn, p.vx[e] / vectorn, flow->gamma);
  F.d *= vectorn;
  F.du *= vectorn;
  F.dv *= vectorn;
  F.de *= vectorn;
  p.rel[0][e] = F.d; p.rel[1][e] = F.du; p.rel[2][e] = F.dv; p.rel[3][e] = F.de;
  if (Rc != -1) {
    p.rer[0][e] = F.d; p.rer[1][e] = F.du; p.rer[2][e] = F.dv; p.rer[3][e] = F.de;
  } else {
    p.rer[0][e] = 0; p.rer[1][e] = 0; p.rer[2][e] = 0; p.rer[3][e] = 0;
  }
}

// kernel_deltT, gpu_solver.cu:385-454
__global__ void k_deltT(PlainDev p, const CaseCtl* ctl, const FlowT<double>* flow) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (!ctl->active || e >= p.E) return;
  const int lg = p.elog[e], Rc = p.right[e];
  Vec4<double> RW;
  if (lg == 0) {
    RW = load_w(p, Rc);
  } else if (lg == 1) {
    const int id = p.farid[e] - 1;
    RW = Vec4<double>{p.farw[0][id], p.farw[1][id], p.farw[2][id], p.farw[3][id]};
  } else {
    const int id = p.wallid[e] - 1;
    RW = Vec4<double>{p.wallw[0][id], p.wallw[1][id], p.wallw[2][id], p.wallw[3][id]};
  }
  const double SR = spectral_radius(load_w(p, p.left[e]), RW, p.vx[e], p.vy[e], flow->gamma);
  p.dtl[e] = SR;
  if (Rc != -1) p.dtr[e] = SR;
}

// kernel_Flux_e2c, gpu_solver.cu:768-878
__global__ void k_e2c(PlainDev p, const CaseCtl* ctl, const FlowT<double>* flow) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (!ctl->active || c >= p.C) return;
  double T = 0;
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < 3; k++) {
    const int e = p.celledge[k][c];
    if (p.left[e] == c) {
      T += p.dtl[e];
      for (int q = 0; q < 4; q++) r[q] += p.rel[q][e];
    }
    if (p.right[e] == c) {
      T += p.dtr[e];
      for (int q = 0; q < 4; q++) r[q] -= p.rer[q][e];
    }
  }
  for (int q = 0; q < 4; q++) p.res[q][c] = r[q];
  p.tsum[c] = T;
  p.dt[c] = fmin(p.area[c] * flow->cfl / T, 0.1);
}

// Global time step: per-block minimum of A*CFL/T (kernel_TimeStep,
// gpu_solver.cu:880-936), then the minimum over blocks.
__global__ void k_dtmin_blocks(PlainDev p, const CaseCtl* ctl, const FlowT<double>* flow) {
  __shared__ double s[kThreads];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (!ctl->active) return;
  s[threadIdx.x] = c < p.C ? p.area[c] * flow->cfl / p.tsum[c] : 9999999.0;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) s[threadIdx.x] = fmin(s[threadIdx.x], s[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) p.blockmin[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kThreads) k_dtmin_final(PlainDev p, const CaseCtl* ctl, int nblocks) {
  if (!ctl->active) return;
  double m = 9999999.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) m = fmin(m, p.blockmin[b]);
  m = block_min(m);
  if (threadIdx.x == 0) *p.dtg = m;
}

// kernel_Smooth1, gpu_solver.cu:1635-1673
__global__ void k_smooth1(PlainDev p, const CaseCtl* ctl, const FlowT<double>* flow) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (!ctl->active || c >= p.C) return;
  const int n0 = p.neicell[0][c], n1 = p.neicell[1][c], n2 = p.neicell[2][c];
  const double eps = flow->eps;
  for (int q = 0; q < 4; q++) {
    double data = 0.0;
    if (n0 >= 0) data += p.res[q][n0];
    if (n1 >= 0) data += p.res[q][n1];
    if (n2 >= 0) data += p.res[q][n2];
    p.r1[q][c] = (data * eps + p.res[q][c]) / (1.0 + eps * p.clog[c]);
  }
}

// kernel_Smooth2, gpu_solver.cu:1675-1720
__global__ void k_smooth2(PlainDev p, const CaseCtl* ctl, const FlowT<double>* flow) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (!ctl->active || c >= p.C) return;
  const int n0 = p.neicell[0][c], n1 = p.neicell[1][c], n2 = p.neicell[2][c];
  const double eps = flow->eps;
  for (int q = 0; q < 4; q++) {
    double data = 0.0;
    if (n0 >= 0) data += p.r1[q][n0];
    if (n1 >= 0) data += p.r1[q][n1];
    if (n2 >= 0) data += p.r1[q][n2];
    const double d2 = (data * eps + p.res[q][c]) / (1.0 + eps * p.clog[c]);
    if (fabs(p.res[q][c]) > fabs(d2)) p.res[q][c] = d2;
  }
}

// kernel_RKStep, gpu_solver.cu:938-966.  The reference also stores P, T and A
// of the new state here, and its far-field ghost reads P and A back
// (gpu_solver.cu:96-98); far_ghost recomputes them from the cell's
// conservatives with the same formula (kernel_CalculationPTA), which gives
// the same bits, so this path stores only the conservatives.
__global__ void k_rkstep(PlainDev p, const CaseCtl* ctl, double RKc, int global_dt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (!ctl->active || c >= p.C) return;
  double RK = RKc;
  RK *= global_dt ? *p.dtg : p.dt[c];
  p.w[0][c] = p.w0[0][c] - RK * p.res[0][c] / p.area[c];
  p.w[1][c] = p.w0[1][c] - RK * p.res[1][c] / p.area[c];
  p.w[2][c] = p.w0[2][c] - RK * p.res[2][c] / p.area[c];
  p.w[3][c] = p.w0[3][c] - RK * p.res[3][c] / p.area[c];
}

// kernel_drou, gpu_solver.cu:968-1020: 64-cell blocks over the reference's
// 1-based ids (block b holds ids 64b .. 64b+63, id 0 unused), tree in LDS.
__global__ void __launch_bounds__(64) k_drou(PlainDev p, const CaseCtl* ctl) {
  __shared__ double r[64];
  const int c = blockIdx.x * 64 + threadIdx.x - 1;
  if (!ctl->active) return;
  if (c >= 0 && c < p.C) {
    const double d = p.w[0][c] - p.w0[0][c];
    r[threadIdx.x] = d * d;
  } else {
    r[threadIdx.x] = 0.0;
  }
  __syncthreads();
  for (int s = 32; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) r[threadIdx.x] = r[threadIdx.x] + r[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) p.partial[blockIdx.x] = r[0];
}

__global__ void k_copy_w0(PlainDev p, const CaseCtl* ctl) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (!ctl->active || c >= p.C) return;
  for (int q = 0; q < 4; q++) p.w0[q][c] = p.w[q][c];
}

}  // namespace

void plain_launch_step(const PlainDev& p, CaseCtl* ctl_all, int icase, const FlowT<double>* flow_all,
                       int order, int smoothing, int global_dt, hipStream_t s) {
  static const double RK[4] = {0.0833, 0.2069, 0.4265, 1.0};  // gpu_fvm.cu:10
  const CaseCtl* ctl = ctl_all + icase;
  const FlowT<double>* flow = flow_all + icase;
  const int gc = (p.C + kThreads - 1) / kThreads, ge = (p.E + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(k_copy_w0, dim3(gc), dim3(kThreads), 0, s, p, ctl);
  for (int ir = 0; ir < 4; ir++) {
    hipLaunchKernelGGL(k_boundary, dim3(ge), dim3(kThreads), 0, s, p, ctl, flow);
    if (order >= 2) hipLaunchKernelGGL(k_gradient, dim3(gc), dim3(kThreads), 0, s, p, ctl);
    hipLaunchKernelGGL(k_flux, dim3(ge), dim3(kThreads), 0, s, p, ctl, flow, order);
    hipLaunchKernelGGL(k_deltT, dim3(ge), dim3(kThreads), 0, s, p, ctl, flow);
    hipLaunchKernelGGL(k_e2c, dim3(gc), dim3(kThreads), 0, s, p, ctl, flow);
    if (global_dt) {
      hipLaunchKernelGGL(k_dtmin_blocks, dim3(gc), dim3(kThreads), 0, s, p, ctl, flow);
      hipLaunchKernelGGL(k_dtmin_final, dim3(1), dim3(kThreads), 0, s, p, ctl, gc);
    }
    if (smoothing) {
      hipLaunchKernelGGL(k_smooth1, dim3(gc), dim3(kThreads), 0, s, p, ctl, flow);
      hipLaunchKernelGGL(k_smooth2, dim3(gc), dim3(kThreads), 0, s, p, ctl, flow);
    }
    hipLaunchKernelGGL(k_rkstep, dim3(gc), dim3(kThreads), 0, s, p, ctl, RK[ir], global_dt);
  }
  hipLaunchKernelGGL(k_drou, dim3(p.nblk64), dim3(64), 0, s, p, ctl);
}

}  // namespace fvm
