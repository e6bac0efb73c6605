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
template <typename T>
__device__ __forceinline__ T limit(T delt, T u, T umax, T umin) {
  T fi = T(1.0);
  if (delt > T(0)) fi = fmin((umax - u) / delt, T(1.0));
  if (delt < T(0)) fi = fmin((umin - u) / delt, T(1.0));
  return delt * fi;
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

}  // namespace fvm
