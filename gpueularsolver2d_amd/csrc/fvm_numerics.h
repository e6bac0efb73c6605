// Point numerics of the explicit finite-volume Euler path, shared by every
// kernel.  Templated on the working type T (double or float).
//
// For T = double every expression keeps the association of the reference's
// CUDA kernels (gpu_solver.cu, lines cited per function), and the library is
// compiled with -ffp-contract=off, so each result is the correctly rounded
// value of the same operation sequence and matches the FP64 CPU restatement
// bit for bit.  For T = float the same sequence runs in single precision.
#pragma once

#include <hip/hip_runtime.h>

namespace fvm {

template <typename T>
struct Vec4 {
  T d, du, dv, de;
};

// Per-case flow constants (reference globals gpu_fvm.cu:5-9, set by
// Initialization gpu_init.cu:63-93), converted to the working type.
template <typename T>
struct FlowT {
  T gamma, R, ui, vi, ain, tin, vin, pis, cfl, eps;
};

// The value of v in another lane of the wave, by a DPP lane pattern (kCtrl)
// applied to the rows in kRowMask; lanes left without a source keep their
// own value.  DPP moves register values between lanes in the VALU, where a
// shuffle goes through the LDS crossbar (ds_bpermute).
template <int kCtrl, int kRowMask>
__device__ __forceinline__ float dpp_from(float v) {
  const int b = __float_as_int(v);
  return __int_as_float(__builtin_amdgcn_update_dpp(b, b, kCtrl, kRowMask, 0xf, false));
}
template <int kCtrl, int kRowMask>
__device__ __forceinline__ double dpp_from(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = (int)b, hi = (int)(b >> 32);
  const int l2 = __builtin_amdgcn_update_dpp(lo, lo, kCtrl, kRowMask, 0xf, false);
  const int h2 = __builtin_amdgcn_update_dpp(hi, hi, kCtrl, kRowMask, 0xf, false);
  return __longlong_as_double(((long long)h2 << 32) | (unsigned)l2);
}
__device__ __forceinline__ float lane63(float v) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ double lane63(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Minimum over the 64 lanes of a wave (all active), returned in every lane:
// neighbours, then pairs of a quad (quad_perm [1,0,3,2], [2,3,0,1]), the
// mirrored lane of each half row and of each row of 16 (after which every
// lane holds its row's minimum), then row 0's into row 1 and row 2's into
// row 3 (row_bcast:15), and lane 31's into rows 2 and 3 (row_bcast:31), which
// leaves the wave's minimum in lane 63.
template <typename R>
__device__ __forceinline__ R wave_min(R v) {
  v = fmin(v, dpp_from<0xb1, 0xf>(v));
  v = fmin(v, dpp_from<0x4e, 0xf>(v));
  v = fmin(v, dpp_from<0x141, 0xf>(v));
  v = fmin(v, dpp_from<0x140, 0xf>(v));
  v = fmin(v, dpp_from<0x142, 0xa>(v));
  v = fmin(v, dpp_from<0x143, 0xc>(v));
  return lane63(v);
}

// Minimum over the workgroup (blockDim a multiple of 64, at most 1024): each
// wave's minimum, then thread 0 takes the waves' minima.  The minimum of a
// set does not depend on the order it is taken in (the values are positive
// time steps; fmin skips a NaN whatever its position), so the result is the
// same bits as a sequential scan.  Valid in thread 0.
template <typename R>
__device__ __forceinline__ R block_min(R v) {
  __shared__ R wmin[16];
  v = wave_min(v);
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) v = fmin(v, wmin[k]);
  return v;
}

// kernel_CalculationPTA, gpu_solver.cu:49-60: pressure and sound speed.
template <typename T>
__device__ __forceinline__ void pressure_sound(const Vec4<T>& w, T gamma, T& P, T& A) {
  T U = w.du / w.d;
  T V = w.dv / w.d;
  T E = w.de / w.d - T(0.5) * (U * U + V * V);
  P = (gamma - T(1.0)) * w.d * E;
  A = sqrt(gamma * P / w.d);
}

// Wall mirror ghost, kernel_Boundary gpu_solver.cu:188-221.
template <typename T>
__device__ __forceinline__ Vec4<T> wall_ghost(const Vec4<T>& wd, T vectorx, T vectory, T vectorn) {
  T n0 = vectory / vectorn;
  T n1 = vectorx / vectorn;
  T UU = wd.du / wd.d;
  T VV = wd.dv / wd.d;
  T UV = UU * n0 + VV * n1;
  UU = UU - T(2) * UV * n0;
  VV = VV - T(2) * UV * n1;
  Vec4<T> g;
  g.d = wd.d;
  g.du = wd.d * UU;
  g.dv = wd.d * VV;
  g.de = wd.de;
  return g;
}

// Characteristic far-field ghost, kernel_Far_Boundary gpu_solver.cu:62-135.
// PD and AD are the domain cell's pressure and sound speed (kernel_RKStep
// stores them with the same formula as pressure_sound).  The supersonic test
// compares a non-dimensional normal speed with the dimensional AIN, as the
// reference does (gpu_solver.cu:79).
template <typename T>
__device__ __forceinline__ Vec4<T> far_ghost(const Vec4<T>& DW, T vectorx, T vectory,
                                             const FlowT<T>& f) {
  T nn = sqrt(vectorx * vectorx + vectory * vectory);
  T nx = vectory / nn;
  T ny = vectorx / nn;
  T UD = DW.du / DW.d;
  T VD = DW.dv / DW.d;
  T Un = fabs(f.ui * nx + f.vi * ny);
  T Uncell = UD * nx + VD * ny;
  const T gm1 = f.gamma - T(1.0);
  Vec4<T> g;
  if (Un / f.ain > T(1.0)) {
    if (Uncell > T(0.0)) {
      g = DW;
    } else {
      g.d = T(1.0);
      g.du = f.ui;
      g.dv = f.vi;
      g.de = f.R * f.tin / (f.vin * f.vin) / gm1 + T(0.50);
    }
    return g;
  }
  T PD, AD;
  pressure_sound(DW, f.gamma, PD, AD);
  T PA = f.pis;
  T ROUD = DW.d;
  T PB, UB, VB;
  if (Uncell > T(0.0)) {
    PB = PA;
    g.d = ROUD + (PB - PD) / AD / AD;
    UB = UD + nx * (PD - PB) / ROUD / AD;
    VB = VD + ny * (PD - PB) / ROUD / AD;
  } else {
    PB = T(0.5) * (PA + PD - ROUD * AD * (nx * (f.ui - UD) + ny * (f.vi - VD)));
    g.d = T(1.0) + (PB - PA) / AD / AD;
    UB = f.ui - nx * (PA - PB) / ROUD / AD;
    VB = f.vi - ny * (PA - PB) / ROUD / AD;
  }
  g.du = UB * g.d;
  g.dv = VB * g.d;
  g.de = PB / gm1 + T(0.5) * (UB * UB + VB * VB) * g.d;
  return g;
}

// Ghost across a boundary edge of type 1 (far field) or 2 (wall).
template <typename T>
__device__ __forceinline__ Vec4<T> boundary_ghost(int type, const Vec4<T>& wd, T vx, T vy,
                                                  const FlowT<T>& f) {
  if (type == 1) return far_ghost(wd, vx, vy, f);
  return wall_ghost(wd, vx, vy, T(sqrt(vx * vx + vy * vy)));
}

// kernel_Limiter, gpu_solver.cu:456-494: per face and component,
// delta *= min((Umax-U)/delta, 1) for delta > 0, min((Umin-U)/delta, 1) for
// delta < 0, unchanged for delta == 0.
// The two cases share one division: its numerator is picked by the sign of
// delta, and delta = 0 (for which the division's value is unused) keeps
// fi = 1.  Each case performs the same operations as the reference's
// branches, so the result is the same bit for bit.
template <typename T>
__device__ __forceinline__ T limit(T delt, T u, T umax, T umin) {
  const T num = delt > T(0) ? umax - u : umin - u;
  const T q = fmin(num / delt, T(1.0));
  return delt * (delt != T(0) ? q : T(1.0));
}

// FP32: the same limited slope without a division.  For delta > 0,
// delta * min((Umax-U)/delta, 1) = min(Umax-U, delta); for delta < 0 the
// product with a negative delta turns the min into max(Umin-U, delta).  The
// bounds hold U (Umin <= U <= Umax), so both cases, and delta = 0, are the
// clamp of delta to [Umin-U, Umax-U]: one median of three.  A zero slope
// may come out with the other sign, which changes U + slope only for U = -0.
template <>
__device__ __forceinline__ float limit<float>(float delt, float u, float umax, float umin) {
  return __builtin_amdgcn_fmed3f(delt, umin - u, umax - u);
}

// kernel_Calc_Function, gpu_solver.cu:552-566 (reciprocal-multiply form).
template <typename T>
__device__ __forceinline__ void calc_function(const Vec4<T>& w, T gamma, T& U, T& V, T& H, T& P) {
  T temp = T(1.0) / w.d;
  U = w.du * temp;
  V = w.dv * temp;
  P = (gamma - T(1.0)) * w.d * (w.de * temp - T(0.5) * (U * U + V * V));
  H = w.de * temp + P * temp;
}

// kernel_ROE_SCHEME, gpu_solver.cu:569-647: Roe flux difference splitting with
// the Harten entropy fix (delta = 0.05).  Returns the flux per unit length.
template <typename T>
__device__ __forceinline__ Vec4<T> roe_flux(const Vec4<T>& LW, const Vec4<T>& RW, T nx, T ny,
                                            T gamma) {
  T U0, V0, H0, P0, U1, V1, H1, P1;
  calc_function(LW, gamma, U0, V0, H0, P0);
  calc_function(RW, gamma, U1, V1, H1, P1);
  T M0 = (U0 * nx + V0 * ny);
  T M1 = (U1 * nx + V1 * ny);
  T DVC = M1 - M0;
  T lrou = LW.d, rrou = RW.d;
  T sl = sqrt(lrou), sr = sqrt(rrou);
  T ROUB = sl * sr;
  T UB = (U0 * sl + U1 * sr) / (sl + sr);
  T VB = (V0 * sl + V1 * sr) / (sl + sr);
  T HB = (H0 * sl + H1 * sr) / (sl + sr);
  T QB = UB * UB + VB * VB;
  T AB = sqrt((gamma - T(1.0)) * (HB - QB / T(2.0)));
  T VC = UB * nx + VB * ny;
  const T eps = T(0.05);
  T AD0, AD1, AD2;
  if (fabs(VC - AB) < eps)
    AD0 = ((VC - AB) * (VC - AB) + eps * eps) / (T(2) * eps);
  else
    AD0 = fabs(VC - AB);
  if (fabs(VC) < eps)
    AD1 = (VC * VC + eps * eps) / (T(2) * eps);
  else
    AD1 = fabs(VC);
  if (fabs(VC + AB) < eps)
    AD2 = ((VC + AB) * (VC + AB) + eps * eps) / (T(2) * eps);
  else
    AD2 = fabs(VC + AB);
  T dP = P1 - P0;
  T F00 = AD0 * (dP - ROUB * AB * DVC) / (T(2.0) * AB * AB);
  T F10 = AD1 * ((rrou - lrou) - dP / (AB * AB));
  T F20 = AD2 * (dP + ROUB * AB * DVC) / (T(2.0) * AB * AB);
  T F01 = F00 * (UB - AB * nx);
  T F02 = F00 * (VB - AB * ny);
  T F03 = F00 * (HB - AB * VC);
  T F11 = F10 * UB + AD1 * ROUB * (U1 - U0 - DVC * nx);
  T F12 = F10 * VB + AD1 * ROUB * (V1 - V0 - DVC * ny);
  T F13 = F10 * QB * T(0.5) + AD1 * ROUB * (UB * (U1 - U0) + VB * (V1 - V0) - DVC * VC);
  T F21 = F20 * (UB + AB * nx);
  T F22 = F20 * (VB + AB * ny);
  T F23 = F20 * (HB + AB * VC);
  T WL0 = lrou * M0;
  T WL1 = U0 * WL0 + nx * P0;
  T WL2 = V0 * WL0 + ny * P0;
  T WL3 = WL0 * H0;
  T WR0 = rrou * M1;
  T WR1 = U1 * WR0 + nx * P1;
  T WR2 = V1 * WR0 + ny * P1;
  T WR3 = WR0 * H1;
  Vec4<T> F;
  F.d = T(0.5) * ((WL0 + WR0) - (F00 + F10 + F20));
  F.du = T(0.5) * ((WL1 + WR1) - (F01 + F11 + F21));
  F.dv = T(0.5) * ((WL2 + WR2) - (F02 + F12 + F22));
  F.de = T(0.5) * ((WL3 + WR3) - (F03 + F13 + F23));
  return F;
}

// FP32 form of the same Roe flux: identical formulas, with shared
// reciprocals (one of sqrt(rho_L)+sqrt(rho_R), one of the squared Roe sound
// speed) and the entropy-fix division by 2*delta folded into a constant.
template <>
__device__ __forceinline__ Vec4<float> roe_flux<float>(const Vec4<float>& LW, const Vec4<float>& RW,
                                                       float nx, float ny, float gamma) {
  const float gm1 = gamma - 1.0f;
  const float rl = 1.0f / LW.d, rr = 1.0f / RW.d;
  const float U0 = LW.du * rl, V0 = LW.dv * rl, U1 = RW.du * rr, V1 = RW.dv * rr;
  const float P0 = gm1 * (LW.de - 0.5f * LW.d * (U0 * U0 + V0 * V0));
  const float P1 = gm1 * (RW.de - 0.5f * RW.d * (U1 * U1 + V1 * V1));
  const float H0 = (LW.de + P0) * rl, H1 = (RW.de + P1) * rr;
  const float M0 = U0 * nx + V0 * ny, M1 = U1 * nx + V1 * ny, DVC = M1 - M0;
  const float sl = sqrtf(LW.d), sr = sqrtf(RW.d);
  const float ROUB = sl * sr;
  const float is = 1.0f / (sl + sr);
  const float wl = sl * is, wr = sr * is;
  const float UB = U0 * wl + U1 * wr, VB = V0 * wl + V1 * wr, HB = H0 * wl + H1 * wr;
  const float QB = UB * UB + VB * VB;
  const float AB2 = gm1 * (HB - 0.5f * QB);
  const float AB = sqrtf(AB2);
  const float iAB2 = 1.0f / AB2;
  const float VC = UB * nx + VB * ny;
  const float e = 0.05f, e2 = 0.05f * 0.05f, i2e = 1.0f / (2.0f * 0.05f);
  const float l0 = VC - AB, l2 = VC + AB;
  const float AD0 = fabsf(l0) < e ? (l0 * l0 + e2) * i2e : fabsf(l0);
  const float AD1 = fabsf(VC) < e ? (VC * VC + e2) * i2e : fabsf(VC);
  const float AD2 = fabsf(l2) < e ? (l2 * l2 + e2) * i2e : fabsf(l2);
  const float dP = P1 - P0, rad = ROUB * AB * DVC, h = 0.5f * iAB2;
  const float F00 = AD0 * (dP - rad) * h;
  const float F20 = AD2 * (dP + rad) * h;
  const float F10 = AD1 * ((RW.d - LW.d) - dP * iAB2);
  const float ar = AD1 * ROUB;
  const float dU = U1 - U0, dV = V1 - V0;
  const float WL0 = LW.d * M0, WR0 = RW.d * M1;
  Vec4<float> F;
  F.d = 0.5f * ((WL0 + WR0) - (F00 + F10 + F20));
  F.du = 0.5f * ((U0 * WL0 + nx * P0 + U1 * WR0 + nx * P1) -
                 (F00 * (UB - AB * nx) + F10 * UB + ar * (dU - DVC * nx) + F20 * (UB + AB * nx)));
  F.dv = 0.5f * ((V0 * WL0 + ny * P0 + V1 * WR0 + ny * P1) -
                 (F00 * (VB - AB * ny) + F10 * VB + ar * (dV - DVC * ny) + F20 * (VB + AB * ny)));
  F.de = 0.5f * ((WL0 * H0 + WR0 * H1) -
                 (F00 * (HB - AB * VC) + F10 * QB * 0.5f + ar * (UB * dU + VB * dV - DVC * VC) +
                  F20 * (HB + AB * VC)));
  return F;
}

// kernel_deltT body, gpu_solver.cu:434-442: spectral radius of an edge from the
// arithmetic mean of the unreconstructed left and right states.
template <typename T>
__device__ __forceinline__ T spectral_radius(const Vec4<T>& L, const Vec4<T>& R, T vx, T vy,
                                             T gamma) {
  T rou = (L.d + R.d) * T(0.5);
  T u = (L.du + R.du) / rou * T(0.5);
  T v = (L.dv + R.dv) / rou * T(0.5);
  T EE = (L.de + R.de) * T(0.5) / rou - (u * u + v * v) * T(0.5);
  T VC = u * vy + v * vx;
  T p = rou * EE * (gamma - T(1.0));
  T CS = sqrt((gamma * p / rou) * (vx * vx + vy * vy));
  return fabs(VC) + CS;
}

// FP32: the edge vector enters as (vx, vy) = (ny, nx) |v| (see EdgeRec<float>),
// so |u vy + v vx| + sqrt(gamma p / rho) |v| is formed from the unit normal.
__device__ __forceinline__ float spectral_radius_n(const Vec4<float>& L, const Vec4<float>& R,
                                                   float nx, float ny, float vn, float gamma) {
  const float rou = (L.d + R.d) * 0.5f;
  const float ir = 1.0f / rou;
  const float u = (L.du + R.du) * 0.5f * ir;
  const float v = (L.dv + R.dv) * 0.5f * ir;
  const float EE = (L.de + R.de) * 0.5f * ir - (u * u + v * v) * 0.5f;
  const float a2 = gamma * (gamma - 1.0f) * EE;  // gamma p / rho with p = rho EE (gamma - 1)
  return (fabsf(u * nx + v * ny) + sqrtf(a2)) * vn;
}

}  // namespace fvm
