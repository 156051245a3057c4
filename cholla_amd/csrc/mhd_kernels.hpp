// mhd_kernels.hpp -- one VL step of ideal MHD (PLMC + HLLD + constrained transport) for gfx950.
//
// The reference step (src/integrators/VL_3D_cuda.cu:38-360 with MHD) is PCM, HLLD x3,
// CT electric fields, half update, half face-field update, PLMC x3, HLLD x3, CT,
// full update, full face-field update: 14 launches that stream the 7-field interface
// arrays Q_L/Q_R (6 of them) and the flux arrays through HBM.  Here:
//
//   MhdFlux1AllKernel  PCM states + HLLD on the three faces i+1/2 of every cell, one
//                      launch (no interface arrays: the PCM states are the cells
//                      themselves); in the device-timed step it also reduces the step's
//                      inverse dt (mhdInverseCrossingTime) from the state it reads
//   MhdFlux1MarchKernel the same by x-y tiles marching in z, the plane staged in LDS
//   MhdCtUpdateKernel  CT edge fields (S&G 2009 eq. 22-24) and the dt/2 update of the
//                      hydro and face fields in one launch: x-y tiles marching in z, the
//                      flux values, cell terms and edge fields in LDS (never in HBM)
//   MhdFlux2TileKernel x faces: PLMC or PPMC of each cell once + HLLD, 64 x 4 face tiles
//                      with the stencil's primitives staged in LDS
//   MhdFlux2MarchKernel y and z faces: the same per face, each lane marching along the
//                      direction with the stencil's primitives in a register window
//   MhdCtUpdateKernel  CT on the half-step fluxes and state, fused with the dt update of
//                      the real cells' hydro fields (density and temperature floors fused)
//                      and of their face fields
//
// Only the cells and faces the real cells depend on are computed, every read in
// bounds; the reference's edge cells that read past the grid do not feed real cells.
//
// Layout: U[f*n + i + j*nx + k*nx*ny], f = density, momentum x/y/z, Energy, Bx, By, Bz
// (face-centred: B?[i] on the face i+1/2).  Flux arrays F_d (d = 0,1,2): 7 fields,
// density, momentum x/y/z (global order), Energy, then the fluxes of the two
// transverse fields in cyclic order (grid_enum fluxX_magnetic_z, fluxX_magnetic_y,
// src/grid/grid_enum.h:80-100).  Electric fields E: 3 fields on cell edges.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "hydro_physics.hpp"
#include "vl3d_kernels.hpp"

namespace chl
{
namespace mhd
{

constexpr double kHlldSmall = 1.0e-8;  // mhd::internal::_hlldSmallNumber, hlld_cuda.h:56
constexpr int kBx = 5;                 // grid_enum::magnetic_x

template <typename T>
__device__ __forceinline__ T Dot(T a1, T a2, T a3, T b1, T b2, T b3)
{
  return a1 * b1 + ((a2 * b2) + (a3 * b3));  // math_utils::dotProduct
}

template <typename T>
__device__ __forceinline__ T MagEnergy(T bx, T by, T bz)
{
  return T(0.5) * Dot(bx, by, bz, bx, by, bz);
}

// mhd::utils, src/utils/mhd_utilities.h:50-151
template <class P>
__device__ __forceinline__ typename P::T Magnetosonic(typename P::T d, typename P::T p, typename P::T bx,
                                                      typename P::T by, typename P::T bz, typename P::T g,
                                                      typename P::T choice)
{
  using T = typename P::T;
  T bx2 = bx * bx;
  T b2  = bx2 + ((by * by) + (bz * bz));
  T t1  = g * p + b2;
  T t2  = (t1 * t1) - T(4.) * g * p * bx2;
  if constexpr (!P::kExact) {
    // (gp + b^2)^2 - 4 gp bx^2 >= (gp - bx^2)^2 >= 0 and t1 >= t2 in exact arithmetic; the
    // rounding of the reordered forms (FP32 above all) must not turn either into a NaN
    t2 = Fmax(t2, T(0));
    t2 = Sqrt<P>(t2);
    return Sqrt<P>(Fmax(Div1<P>(t1 + choice * t2, T(2.0) * Fmax(d, T(kTinyNumber))), T(0)));
  }
  t2 = Sqrt<P>(t2);
  return Sqrt<P>(Div1<P>(t1 + choice * t2, T(2.0) * Fmax(d, T(kTinyNumber))));
}

template <class P>
__device__ __forceinline__ typename P::T Alfven(typename P::T bx, typename P::T d)
{
  using T = typename P::T;
  return Div1<P>(Fabs(bx), Sqrt<P>(Fmax(d, T(kTinyNumber))));
}

// hydro_utilities::Calc_Pressure_Primitive / Calc_Energy_Primitive with MHD (hydro_utilities.h:34-45, 106-118)
template <typename T>
__device__ __forceinline__ T PressurePrimM(T E, T d, T vx, T vy, T vz, T g, T bx, T by, T bz)
{
  T p = E - T(0.5) * d * Dot(vx, vy, vz, vx, vy, vz);
  p -= MagEnergy(bx, by, bz);
  return Fmax((g - T(1.)) * p, T(kTinyNumber));
}
template <typename T>
__device__ __forceinline__ T EnergyPrimM(T P, T d, T vx, T vy, T vz, T g, T bx, T by, T bz)
{
  T e = (Fmax(P, T(kTinyNumber)) / (g - T(1.))) + T(0.5) * d * Dot(vx, vy, vz, vx, vy, vz);
  e += MagEnergy(bx, by, bz);
  return e;
}

// mhd::utils::cellCenteredMagneticFields, mhd_utilities.h:177-207
__device__ __forceinline__ void CellB(const double *U, size_t id, int x, int y, int z, size_t n, int nx, int ny,
                                      double &bx, double &by, double &bz)
{
  const double *B = U + kBx * n;
  bx = (x > 0) ? 0.5 * (B[id] + B[id - 1]) : B[id];
  by = (y > 0) ? 0.5 * (B[n + id] + B[n + id - size_t(nx)]) : B[n + id];
  bz = (z > 0) ? 0.5 * (B[2 * n + id] + B[2 * n + id - size_t(nx) * ny]) : B[2 * n + id];
}

template <typename T>
struct PrimM {
  T d, vx, vy, vz, p, bx, by, bz;
};
template <typename T>
struct ConsM {  // rotated interface state: d, m (normal, t1, t2), E, B transverse
  T d, mx, my, mz, E, by, bz;
};

// Rotate a global (x, y, z) triple into the frame of dir: (normal, o2, o3)
template <typename T>
__device__ __forceinline__ void RotM(int dir, T x, T y, T z, T &a, T &b, T &c)
{
  if (dir == 0) {
    a = x, b = y, c = z;
  } else if (dir == 1) {
    a = y, b = z, c = x;
  } else {
    a = z, b = x, c = y;
  }
}

// reconstruction::Load_Data with MHD (reconstruction.h:139-212); pressure recovered in FP64
template <class P>
__device__ __forceinline__ PrimM<typename P::T> LoadDataM(const double *U, int x, int y, int z, int nx, int ny,
                                                          size_t n, int dir, double g)
{
  using T         = typename P::T;
  const size_t id = size_t(x) + size_t(y) * nx + size_t(z) * size_t(nx) * ny;
  double mx, my, mz, cx, cy, cz;
  RotM(dir, U[n + id], U[2 * n + id], U[3 * n + id], mx, my, mz);
  CellB(U, id, x, y, z, n, nx, ny, cx, cy, cz);
  double bx, by, bz;
  RotM(dir, cx, cy, cz, bx, by, bz);
  const double d = U[id];
  double vx, vy, vz;
  if constexpr (P::kExact) {
    vx = mx / d, vy = my / d, vz = mz / d;
  } else {
    const double inv = 1.0 / d;
    vx = mx * inv, vy = my * inv, vz = mz * inv;
  }
  const double p = PressurePrimM(U[4 * n + id], d, vx, vy, vz, g, bx, by, bz);
  return PrimM<T>{T(d), T(vx), T(vy), T(vz), T(p), T(bx), T(by), T(bz)};
}

// reconstruction::Compute_Eigenvectors with MHD (reconstruction.h:310-370)
template <typename T>
struct EigenM {
  T cf, cs, af, as, by, bz, nfs, sign, qf, qs, Af, As, qpf, qps, Apf, Aps, sqrt_d;
};

template <class P>
__device__ __forceinline__ EigenM<typename P::T> ComputeEigenM(const PrimM<typename P::T> &w, typename P::T cs,
                                                               typename P::T cs2, typename P::T g)
{
  using T = typename P::T;
  EigenM<T> e;
  e.cf         = Magnetosonic<P>(w.d, w.p, w.bx, w.by, w.bz, g, T(1.0));
  e.cs         = Magnetosonic<P>(w.d, w.p, w.bx, w.by, w.bz, g, T(-1.0));
  const T cf2 = e.cf * e.cf, css = e.cs * e.cs;
  const T denom = (cf2 - css), num2 = (cf2 - cs2);
  if (denom <= T(0.0) or num2 <= T(0.0)) {
    e.af = T(1.0), e.as = T(0.0);
  } else {
    const T num1 = (cs2 - css);
    if (num1 <= T(0.0)) {
      e.af = T(0.0), e.as = T(1.0);
    } else {
      e.af = Sqrt<P>(Div1<P>(num1, denom));
      e.as = Sqrt<P>(Div1<P>(num2, denom));
    }
  }
  const T bden = Div1<P>(T(1.0), Sqrt<P>(w.by * w.by + w.bz * w.bz));  // rhypot
  const bool fin = __builtin_isfinite(bden);
  e.by           = fin ? w.by * bden : T(1.0);
  e.bz           = fin ? w.bz * bden : T(0.0);
  e.sign         = __builtin_copysign(T(1.0), w.bx);
  e.nfs          = Div1<P>(T(0.5), cs2);
  e.qpf          = e.sign * e.nfs * e.af * e.cf;
  e.qps          = e.sign * e.nfs * e.as * e.cs;
  e.qf           = e.sign * e.af * e.cf;
  e.qs           = e.sign * e.as * e.cs;
  e.sqrt_d       = Sqrt<P>(w.d);
  e.Af           = e.af * cs * e.sqrt_d;
  e.As           = e.as * cs * e.sqrt_d;
  e.Apf          = Div1<P>(T(0.5) * e.af, cs * e.sqrt_d);
  e.Aps          = Div1<P>(T(0.5) * e.as, cs * e.sqrt_d);
  return e;
}

template <typename T>
struct Char7 {
  T a0, a1, a2, a3, a4, a5, a6;
};

// reconstruction::Primitive_To_Characteristic with MHD (reconstruction.h:378-425)
template <class P>
__device__ __forceinline__ Char7<typename P::T> PrimToCharM(const PrimM<typename P::T> &w,
                                                            const PrimM<typename P::T> &s,
                                                            const EigenM<typename P::T> &e, typename P::T cs2,
                                                            typename P::T inv_d, typename P::T inv_cs2)
{
  using T     = typename P::T;
  const T isd = Div1<P>(T(1.0), e.sqrt_d);  // rsqrt
  const T pd  = Div<P>(s.p, w.d, inv_d);
  Char7<T> c;
  c.a0 = e.nfs * e.af * (pd - e.cf * s.vx) + e.qps * (e.by * s.vy + e.bz * s.vz) + e.Aps * (e.by * s.by + e.bz * s.bz);
  c.a1 = T(0.5) * (e.by * (s.bz * e.sign * isd + s.vz) - e.bz * (s.by * e.sign * isd + s.vy));
  c.a2 = e.nfs * e.as * (pd - e.cs * s.vx) - e.qpf * (e.by * s.vy + e.bz * s.vz) - e.Apf * (e.by * s.by + e.bz * s.bz);
  c.a3 = s.d - Div<P>(s.p, cs2, inv_cs2);
  c.a4 = e.nfs * e.as * (pd + e.cs * s.vx) + e.qpf * (e.by * s.vy + e.bz * s.vz) - e.Apf * (e.by * s.by + e.bz * s.bz);
  c.a5 = T(0.5) * (e.by * (s.bz * e.sign * isd - s.vz) - e.bz * (s.by * e.sign * isd - s.vy));
  c.a6 = e.nfs * e.af * (pd + e.cf * s.vx) - e.qps * (e.by * s.vy + e.bz * s.vz) + e.Aps * (e.by * s.by + e.bz * s.bz);
  return c;
}

// reconstruction::Characteristic_To_Primitive with MHD (reconstruction.h:450-490)
template <class P>
__device__ __forceinline__ PrimM<typename P::T> CharToPrimM(const PrimM<typename P::T> &w,
                                                            const Char7<typename P::T> &a,
                                                            const EigenM<typename P::T> &e, typename P::T cs2)
{
  using T = typename P::T;
  PrimM<T> o;
  o.d  = w.d * (e.af * (a.a0 + a.a6) + e.as * (a.a2 + a.a4)) + a.a3;
  o.vx = e.cf * e.af * (a.a6 - a.a0) + e.cs * e.as * (a.a4 - a.a2);
  o.vy = e.by * (e.qs * (a.a0 - a.a6) + e.qf * (a.a4 - a.a2)) + e.bz * (a.a5 - a.a1);
  o.vz = e.bz * (e.qs * (a.a0 - a.a6) + e.qf * (a.a4 - a.a2)) + e.by * (a.a1 - a.a5);
  o.p  = w.d * cs2 * (e.af * (a.a0 + a.a6) + e.as * (a.a2 + a.a4));
  o.bx = T(0.0);
  o.by = e.by * (e.As * (a.a0 + a.a6) - e.Af * (a.a2 + a.a4)) - e.bz * e.sign * e.sqrt_d * (a.a5 + a.a1);
  o.bz = e.bz * (e.As * (a.a0 + a.a6) - e.Af * (a.a2 + a.a4)) + e.by * e.sign * e.sqrt_d * (a.a5 + a.a1);
  return o;
}

template <typename T>
__device__ __forceinline__ T MonoM(T l, T r, T c, T v)  // Monotonize_Characteristic_Return_Primitive
{
  if (l * r > T(0.0)) {
    T la = T(2.0) * Fmin(Fabs(l), Fabs(r));
    T lb = Fmin(Fabs(c), Fabs(v));
    return __builtin_copysign(Fmin(la, lb), c);
  }
  return T(0.0);
}
template <typename T>
__device__ __forceinline__ T VlM(T l, T r)  // Van_Leer_Slope
{
  return (l * r > T(0.0)) ? T(2.0) * l * r / (l + r) : T(0.0);
}
template <typename T>
__device__ __forceinline__ void LimitM(T &l, T &r, T vm, T v0, T vp)  // Plm_Limit_Interfaces
{
  r = Fmax(Fmin(v0, vm), r);
  r = Fmin(Fmax(v0, vm), r);
  l = Fmax(Fmin(v0, vp), l);
  l = Fmin(Fmax(v0, vp), l);
}

// PLMC_cuda with MHD for one cell (plmc_cuda.cu:24-300, VL build) from the primitives of the
// cell (c0) and its neighbours along dir (cm, cp): the interface states at +1/2 (Lp) and
// -1/2 (Rm).
template <class P>
__device__ __forceinline__ void PlmcPrimsM(const PrimM<typename P::T> &cm, const PrimM<typename P::T> &c0,
                                           const PrimM<typename P::T> &cp, double g, PrimM<typename P::T> &Lp,
                                           PrimM<typename P::T> &Rm)
{
  using T           = typename P::T;
  const T gg        = T(g);
  const T inv_d     = Div1<P>(T(1.0), c0.d);
  const T cs        = Sqrt<P>(Div<P>(gg * c0.p, c0.d, inv_d));  // Calc_Sound_Speed
  const T cs2       = cs * cs;
  const T inv_cs2   = Div1<P>(T(1.0), cs2);
  const EigenM<T> e = ComputeEigenM<P>(c0, cs, cs2, gg);
  auto slope = [](const PrimM<T> &l, const PrimM<T> &r, T c) {
    return PrimM<T>{c * (r.d - l.d), c * (r.vx - l.vx), c * (r.vy - l.vy), c * (r.vz - l.vz),
                    c * (r.p - l.p), T(0),             c * (r.by - l.by), c * (r.bz - l.bz)};
  };
  const PrimM<T> dL = slope(cm, c0, T(1.0)), dR = slope(c0, cp, T(1.0)), dC = slope(cm, cp, T(0.5));
  const PrimM<T> dG{VlM(dL.d, dR.d), VlM(dL.vx, dR.vx), VlM(dL.vy, dR.vy), VlM(dL.vz, dR.vz),
                    VlM(dL.p, dR.p), T(0),              VlM(dL.by, dR.by), VlM(dL.bz, dR.bz)};
  const Char7<T> aL = PrimToCharM<P>(c0, dL, e, cs2, inv_d, inv_cs2);
  const Char7<T> aR = PrimToCharM<P>(c0, dR, e, cs2, inv_d, inv_cs2);
  Char7<T> aC;
  if constexpr (P::kExact) {  // the reference projects the centred slope itself
    aC = PrimToCharM<P>(c0, dC, e, cs2, inv_d, inv_cs2);
  } else {  // dC = (dL + dR) / 2 and the projection is linear: one projection fewer
    aC = Char7<T>{T(0.5) * (aL.a0 + aR.a0), T(0.5) * (aL.a1 + aR.a1), T(0.5) * (aL.a2 + aR.a2),
                  T(0.5) * (aL.a3 + aR.a3), T(0.5) * (aL.a4 + aR.a4), T(0.5) * (aL.a5 + aR.a5),
                  T(0.5) * (aL.a6 + aR.a6)};
  }
  const Char7<T> aG = PrimToCharM<P>(c0, dG, e, cs2, inv_d, inv_cs2);
  Char7<T> am;
  am.a0 = MonoM(aL.a0, aR.a0, aC.a0, aG.a0);
  am.a1 = MonoM(aL.a1, aR.a1, aC.a1, aG.a1);
  am.a2 = MonoM(aL.a2, aR.a2, aC.a2, aG.a2);
  am.a3 = MonoM(aL.a3, aR.a3, aC.a3, aG.a3);
  am.a4 = MonoM(aL.a4, aR.a4, aC.a4, aG.a4);
  am.a5 = MonoM(aL.a5, aR.a5, aC.a5, aG.a5);
  am.a6 = MonoM(aL.a6, aR.a6, aC.a6, aG.a6);
  const PrimM<T> dm = CharToPrimM<P>(c0, am, e, cs2);
  auto iface = [](T q, T s, T sign) { return q + sign * T(0.5) * s; };  // Calc_Interface_Linear
  Lp = PrimM<T>{iface(c0.d, dm.d, T(1)),   iface(c0.vx, dm.vx, T(1)), iface(c0.vy, dm.vy, T(1)),
                iface(c0.vz, dm.vz, T(1)), iface(c0.p, dm.p, T(1)),   T(0),
                iface(c0.by, dm.by, T(1)), iface(c0.bz, dm.bz, T(1))};
  Rm = PrimM<T>{iface(c0.d, dm.d, T(-1)),   iface(c0.vx, dm.vx, T(-1)), iface(c0.vy, dm.vy, T(-1)),
                iface(c0.vz, dm.vz, T(-1)), iface(c0.p, dm.p, T(-1)),   T(0),
                iface(c0.by, dm.by, T(-1)), iface(c0.bz, dm.bz, T(-1))};
  LimitM(Lp.d, Rm.d, cm.d, c0.d, cp.d);
  LimitM(Lp.vx, Rm.vx, cm.vx, c0.vx, cp.vx);
  LimitM(Lp.vy, Rm.vy, cm.vy, c0.vy, cp.vy);
  LimitM(Lp.vz, Rm.vz, cm.vz, c0.vz, cp.vz);
  LimitM(Lp.p, Rm.p, cm.p, c0.p, cp.p);
  LimitM(Lp.by, Rm.by, cm.by, c0.by, cp.by);
  LimitM(Lp.bz, Rm.bz, cm.bz, c0.bz, cp.bz);
  Rm.d = Fmax(Rm.d, T(kTinyNumber));
  Lp.d = Fmax(Lp.d, T(kTinyNumber));
  Rm.p = Fmax(Rm.p, T(kTinyNumber));
  Lp.p = Fmax(Lp.p, T(kTinyNumber));
}

// PLMC with MHD for cell (x,y,z) along dir, its stencil loaded (Load_Data) here
template <class P>
__device__ __forceinline__ void PlmcCellM(const double *U, int x, int y, int z, int nx, int ny, size_t n, int dir,
                                          double g, PrimM<typename P::T> &Lp, PrimM<typename P::T> &Rm)
{
  const int di = dir == 0, dj = dir == 1, dk = dir == 2;
  PlmcPrimsM<P>(LoadDataM<P>(U, x - di, y - dj, z - dk, nx, ny, n, dir, g), LoadDataM<P>(U, x, y, z, nx, ny, n, dir, g),
                LoadDataM<P>(U, x + di, y + dj, z + dk, nx, ny, n, dir, g), g, Lp, Rm);
}

// PPMC_VL with MHD for one cell (ppmc_cuda.cu:540-695) from the primitives of the cell (c0)
// and its neighbours along dir (cm2, cm1, cp1, cp2): the 5-cell stencil projected on the
// cell's seven MHD characteristics, PPM_Single_Variable on each (the reduced FP32 form under
// the mixed policy), projected back, density and pressure floored: the interface states at
// +1/2 (Lp) and -1/2 (Rm).
template <class P>
__device__ __forceinline__ void PpmcPrimsM(const PrimM<typename P::T> &cm2, const PrimM<typename P::T> &cm1,
                                           const PrimM<typename P::T> &c0, const PrimM<typename P::T> &cp1,
                                           const PrimM<typename P::T> &cp2, double g, PrimM<typename P::T> &Lp,
                                           PrimM<typename P::T> &Rm)
{
  using T           = typename P::T;
  const T gg        = T(g);
  const T inv_d     = Div1<P>(T(1.0), c0.d);
  const T cs        = Sqrt<P>(Div<P>(gg * c0.p, c0.d, inv_d));  // Calc_Sound_Speed
  const T cs2       = cs * cs;
  const T inv_cs2   = Div1<P>(T(1.0), cs2);
  const EigenM<T> e = ComputeEigenM<P>(c0, cs, cs2, gg);
  const Char7<T> a0  = PrimToCharM<P>(c0, c0, e, cs2, inv_d, inv_cs2);
  const Char7<T> am1 = PrimToCharM<P>(c0, cm1, e, cs2, inv_d, inv_cs2);
  const Char7<T> am2 = PrimToCharM<P>(c0, cm2, e, cs2, inv_d, inv_cs2);
  const Char7<T> ap1 = PrimToCharM<P>(c0, cp1, e, cs2, inv_d, inv_cs2);
  const Char7<T> ap2 = PrimToCharM<P>(c0, cp2, e, cs2, inv_d, inv_cs2);
  Char7<T> L, R;
  auto ppm = [](T im2, T im1, T i0, T ip1, T ip2, T &l, T &r) {
    if constexpr (P::kId == PolicyMixed::kId) {
      PpmSingleReduced(im2, im1, i0, ip1, ip2, l, r);
    } else {
      PpmSingle<P>(im2, im1, i0, ip1, ip2, l, r);
    }
  };
  ppm(am2.a0, am1.a0, a0.a0, ap1.a0, ap2.a0, L.a0, R.a0);
  ppm(am2.a1, am1.a1, a0.a1, ap1.a1, ap2.a1, L.a1, R.a1);
  ppm(am2.a2, am1.a2, a0.a2, ap1.a2, ap2.a2, L.a2, R.a2);
  ppm(am2.a3, am1.a3, a0.a3, ap1.a3, ap2.a3, L.a3, R.a3);
  ppm(am2.a4, am1.a4, a0.a4, ap1.a4, ap2.a4, L.a4, R.a4);
  ppm(am2.a5, am1.a5, a0.a5, ap1.a5, ap2.a5, L.a5, R.a5);
  ppm(am2.a6, am1.a6, a0.a6, ap1.a6, ap2.a6, L.a6, R.a6);
  Lp   = CharToPrimM<P>(c0, L, e, cs2);
  Rm   = CharToPrimM<P>(c0, R, e, cs2);
  Rm.d = Fmax(Rm.d, T(kTinyNumber));
  Lp.d = Fmax(Lp.d, T(kTinyNumber));
  Rm.p = Fmax(Rm.p, T(kTinyNumber));
  Lp.p = Fmax(Lp.p, T(kTinyNumber));
}

// PPMC with MHD for cell (x,y,z) along dir, its stencil loaded (Load_Data) here
template <class P>
__device__ __forceinline__ void PpmcCellM(const double *U, int x, int y, int z, int nx, int ny, size_t n, int dir,
                                          double g, PrimM<typename P::T> &Lp, PrimM<typename P::T> &Rm)
{
  const int di = dir == 0, dj = dir == 1, dk = dir == 2;
  PpmcPrimsM<P>(LoadDataM<P>(U, x - 2 * di, y - 2 * dj, z - 2 * dk, nx, ny, n, dir, g),
                LoadDataM<P>(U, x - di, y - dj, z - dk, nx, ny, n, dir, g), LoadDataM<P>(U, x, y, z, nx, ny, n, dir, g),
                LoadDataM<P>(U, x + di, y + dj, z + dk, nx, ny, n, dir, g),
                LoadDataM<P>(U, x + 2 * di, y + 2 * dj, z + 2 * dk, nx, ny, n, dir, g), g, Lp, Rm);
}

// Stage-2 reconstruction of the MHD build (CHL_RECON_*): ghost cells it needs, and the
// stencil half-width of the cell reconstruction (PCM forms its states from the cells).
__host__ __device__ constexpr int MhdReconGhosts(int recon) { return recon == 0 ? 3 : (recon == 2 ? 5 : 4); }

// reconstruction::Write_Data with MHD (reconstruction.h:918-970), kept in registers:
// conserved interface state with the face field bx for the energy
template <class P>
__device__ __forceinline__ ConsM<typename P::T> WriteDataM(const PrimM<typename P::T> &s, typename P::T bx,
                                                           typename P::T g)
{
  using T = typename P::T;
  return ConsM<T>{s.d, s.d * s.vx, s.d * s.vy, s.d * s.vz, EnergyPrimM(s.p, s.d, s.vx, s.vy, s.vz, g, bx, s.by, s.bz),
                  s.by, s.bz};
}

// ---------------------------------------------------------------------------- HLLD (hlld_cuda.cu)
template <typename T>
struct HStateM {
  T d, vx, vy, vz, E, by, bz, pg, pt;
};
template <typename T>
struct HStarM {
  T d, vy, vz, E, by, bz;
};

// mhd::internal::loadState (hlld_cuda.cu:170-214)
template <class P>
__device__ __forceinline__ HStateM<typename P::T> LoadStateM(const ConsM<typename P::T> &q, typename P::T bx,
                                                             typename P::T g)
{
  using T = typename P::T;
  HStateM<T> s;
  s.d       = Fmax(q.d, T(kTinyNumber));
  const T inv = Div1<P>(T(1.0), s.d);
  s.vx      = Div<P>(q.mx, s.d, inv);
  s.vy      = Div<P>(q.my, s.d, inv);
  s.vz      = Div<P>(q.mz, s.d, inv);
  s.E       = Fmax(q.E, T(kTinyNumber));
  s.by = q.by, s.bz = q.bz;
  s.pg = PressurePrimM(s.E, s.d, s.vx, s.vy, s.vz, g, bx, s.by, s.bz);
  s.pt = Fmax(s.pg + MagEnergy(bx, s.by, s.bz), T(kTinyNumber));  // computeTotalPressure
  return s;
}

template <typename T>
__device__ __forceinline__ ConsM<T> NonStarFluxM(const HStateM<T> &s, T bx)  // hlld_cuda.cu:262-283
{
  ConsM<T> f;
  f.d  = s.d * s.vx;
  f.mx = f.d * s.vx + s.pt - bx * bx;
  f.my = f.d * s.vy - bx * s.by;
  f.mz = f.d * s.vz - bx * s.bz;
  f.by = s.by * s.vx - bx * s.vy;
  f.bz = s.bz * s.vx - bx * s.vz;
  f.E  = s.vx * (s.E + s.pt) - bx * (s.vx * bx + ((s.vy * s.by) + (s.vz * s.bz)));
  return f;
}

template <class P>
__device__ __forceinline__ HStarM<typename P::T> StarStateM(const HStateM<typename P::T> &s, typename P::T M,
                                                            typename P::T side, typename P::T bx,
                                                            typename P::T ptStar)  // hlld_cuda.cu:327-372
{
  using T = typename P::T;
  HStarM<T> o;
  o.d = Div1<P>(s.d * (side - s.vx), (side - M));
  if (Fabs(s.d * (side - s.vx) * (side - M) - (bx * bx)) < ptStar * T(kHlldSmall)) {
    o.vy = s.vy, o.vz = s.vz, o.by = s.by, o.bz = s.bz;
  } else {
    const T denom = s.d * (side - s.vx) * (side - M) - (bx * bx);
    const T inv   = Div1<P>(T(1.0), denom);
    T coef        = Div<P>(bx * (M - s.vx), denom, inv);
    o.vy          = s.vy - s.by * coef;
    o.vz          = s.vz - s.bz * coef;
    T tp          = (side - s.vx);
    tp            = tp * tp;
    coef          = Div<P>((s.d * tp - (bx * bx)), denom, inv);
    o.by          = s.by * coef;
    o.bz          = s.bz * coef;
  }
  o.E = Div1<P>((s.E * (side - s.vx) - s.pt * s.vx + ptStar * M +
                 bx * (Dot(s.vx, s.vy, s.vz, bx, s.by, s.bz) - Dot(M, o.vy, o.vz, bx, o.by, o.bz))),
                (side - M));
  return o;
}

template <typename T>
__device__ __forceinline__ ConsM<T> StarFluxM(const HStarM<T> &st, const HStateM<T> &s, const ConsM<T> &f, T M,
                                              T side)  // hlld_cuda.cu:376-397
{
  ConsM<T> o;
  o.d  = f.d + side * (st.d - s.d);
  o.mx = f.mx + side * (st.d * M - s.d * s.vx);
  o.my = f.my + side * (st.d * st.vy - s.d * s.vy);
  o.mz = f.mz + side * (st.d * st.vz - s.d * s.vz);
  o.E  = f.E + side * (st.E - s.E);
  o.by = f.by + side * (st.by - s.by);
  o.bz = f.bz + side * (st.bz - s.bz);
  return o;
}

// Calculate_HLLD_Fluxes_CUDA for one interface (hlld_cuda.cu:31-164), rotated frame.
// Returns false when no branch of the reference assigns a flux (non-finite speeds).
template <class P>
__device__ __forceinline__ bool HlldM(const ConsM<typename P::T> &qL, const ConsM<typename P::T> &qR, typename P::T bx,
                                      typename P::T g, ConsM<typename P::T> &out)
{
  using T              = typename P::T;
  const HStateM<T> sL = LoadStateM<P>(qL, bx, g), sR = LoadStateM<P>(qR, bx, g);
  T SL, SR;
  {
    const T cL = Magnetosonic<P>(sL.d, sL.pg, bx, sL.by, sL.bz, g, T(1.0));
    const T cR = Magnetosonic<P>(sR.d, sR.pg, bx, sR.by, sR.bz, g, T(1.0));
    const T cM = Fmax(cL, cR);
    SL         = Fmin(sL.vx, sR.vx) - cM;
    SR         = Fmax(sL.vx, sR.vx) + cM;
  }
  const ConsM<T> fL = NonStarFluxM(sL, bx);
  if (SL > T(0.0)) {
    out = fL;
    return true;
  }
  const ConsM<T> fR = NonStarFluxM(sR, bx);
  if (SR < T(0.0)) {
    out = fR;
    return true;
  }
  const T rd = SR - sR.vx, ld = SL - sL.vx;
  const T M  = Div1<P>((rd * sR.d * sR.vx - ld * sL.d * sL.vx - sR.pt + sL.pt), (rd * sR.d - ld * sL.d));
  const T ptStar = Div1<P>((sR.d * sL.pt * (SR - sR.vx) - sL.d * sR.pt * (SL - sL.vx) +
                            sL.d * sR.d * (SR - sR.vx) * (SL - sL.vx) * (sR.vx - sL.vx)),
                           (sR.d * (SR - sR.vx) - sL.d * (SL - sL.vx)));
  const HStarM<T> stL = StarStateM<P>(sL, M, SL, bx, ptStar);
  const T SLs         = M + T(-1.0) * Alfven<P>(bx, stL.d);
  if (SLs > T(0.0) and SL <= T(0.0)) {
    out = StarFluxM(stL, sL, fL, M, SL);
    return true;
  }
  const HStarM<T> stR = StarStateM<P>(sR, M, SR, bx, ptStar);
  const T SRs         = M + T(1.0) * Alfven<P>(bx, stR.d);
  if (SRs <= T(0.0) and SR >= T(0.0)) {
    out = StarFluxM(stR, sR, fR, M, SR);
    return true;
  }
  // computeDoubleStarState (hlld_cuda.cu:401-465)
  T vy, vz, by, bz, EL, ER;
  if (T(0.5) * (bx * bx) < T(kHlldSmall) * ptStar) {
    if (M >= T(0.0)) {
      vy = stL.vy, vz = stL.vz, by = stL.by, bz = stL.bz, EL = stL.E, ER = stR.E;
    } else {
      vy = stR.vy, vz = stR.vz, by = stR.by, bz = stR.bz, ER = stR.E, EL = stL.E;
    }
  } else {
    const T sqL = Sqrt<P>(stL.d), sqR = Sqrt<P>(stR.d);
    const T inv = Div1<P>(T(1.0), (sqL + sqR));
    const T sgn = __builtin_copysign(T(1.0), bx);
    vy          = inv * (sqL * stL.vy + sqR * stR.vy + sgn * (stR.by - stL.by));
    vz          = inv * (sqL * stL.vz + sqR * stR.vz + sgn * (stR.bz - stL.bz));
    by          = inv * (sqL * stR.by + sqR * stL.by + sgn * (sqL * sqR) * (stR.vy - stL.vy));
    bz          = inv * (sqL * stR.bz + sqR * stL.bz + sgn * (sqL * sqR) * (stR.vz - stL.vz));
    const T vb  = Dot(M, vy, vz, bx, by, bz);
    EL          = stL.E - sqL * sgn * (Dot(M, stL.vy, stL.vz, bx, stL.by, stL.bz) - vb);
    ER          = stR.E + sqR * sgn * (Dot(M, stR.vy, stR.vz, bx, stR.by, stR.bz) - vb);
  }
  // computeDoubleStarFluxes (hlld_cuda.cu:469-505)
  auto dstar = [&](const HStarM<T> &st, const HStateM<T> &s, const ConsM<T> &f, T side, T sideStar, T dE) {
    const T diff = sideStar - side;
    ConsM<T> o;
    o.d  = f.d - side * s.d - diff * st.d + sideStar * st.d;
    o.mx = f.mx - side * (s.d * s.vx) - diff * (st.d * M) + sideStar * (st.d * M);
    o.my = f.my - side * (s.d * s.vy) - diff * (st.d * st.vy) + sideStar * (st.d * vy);
    o.mz = f.mz - side * (s.d * s.vz) - diff * (st.d * st.vz) + sideStar * (st.d * vz);
    o.E  = f.E - side * s.E - diff * st.E + sideStar * dE;
    o.by = f.by - side * s.by - diff * st.by + sideStar * by;
    o.bz = f.bz - side * s.bz - diff * st.bz + sideStar * bz;
    return o;
  };
  if (M > T(0.0) and SLs <= T(0.0)) {
    out = dstar(stL, sL, fL, SL, SLs, EL);
    return true;
  }
  if (SRs > T(0.0) and M <= T(0.0)) {
    out = dstar(stR, sR, fR, SR, SRs, ER);
    return true;
  }
  return false;
}

// mhd::internal::returnFluxes (hlld_cuda.cu:287-304): rotated flux to the global slots
template <typename TF, typename T>
__device__ __forceinline__ void StoreFluxM(TF *F, size_t id, size_t n, int dir, const ConsM<T> &f)
{
  T gx, gy, gz;
  if (dir == 0) {
    gx = f.mx, gy = f.my, gz = f.mz;
  } else if (dir == 1) {
    gy = f.mx, gz = f.my, gx = f.mz;
  } else {
    gz = f.mx, gx = f.my, gy = f.mz;
  }
  F[id]         = TF(f.d);
  F[n + id]     = TF(gx);
  F[2 * n + id] = TF(gy);
  F[3 * n + id] = TF(gz);
  F[4 * n + id] = TF(f.E);
  F[5 * n + id] = TF(f.by);
  F[6 * n + id] = TF(f.bz);
}

// ---------------------------------------------------------------------------- CT (ct_electric_fields.cu/.h)
// Reads outside [0, total) return 0; within the regions the step kernels compute none occurs.
template <typename TF>
__device__ __forceinline__ double AtF(const TF *a, long i, long total)
{
  return (i >= 0 && i < total) ? double(a[i]) : 0.0;
}

// mhd::internal::_ctSlope (ct_electric_fields.h:66-119)
template <typename TF>
__device__ __forceinline__ double CtSlopeM(const TF *flux, const double *U, double fluxSign, int ctDir, int cq1,
                                           int cq2, int fq1, int fq2, int x, int y, int z, int nx, int ny, long n,
                                           long nF, long nU)
{
  const int m1 = (ctDir + 1) % 3, m2 = (ctDir + 2) % 3;
  const int xc = x - int(cq1 == 0) - int(cq2 == 0);
  const int yc = y - int(cq1 == 1) - int(cq2 == 1);
  const int zc = z - int(cq1 == 2) - int(cq2 == 2);
  auto lin     = [&](long i, long j, long k) { return i + j * long(nx) + k * long(nx) * long(ny); };
  const long ic    = lin(xc, yc, zc);
  const long iflux = lin(x - int(fq1 == 0) - int(fq2 == 0), y - int(fq1 == 1) - int(fq2 == 1),
                         z - int(fq1 == 2) - int(fq2 == 2));
  const long ib2 = lin(xc - int(m1 == 0), yc - int(m1 == 1), zc - int(m1 == 2));
  const long ib3 = lin(xc - int(m2 == 0), yc - int(m2 == 1), zc - int(m2 == 2));
  const double d   = AtF(U, ic, nU);
  const double mm2 = AtF(U, ic + (m1 + 1) * n, nU);
  const double mm3 = AtF(U, ic + (m2 + 1) * n, nU);
  const double b2  = 0.5 * (AtF(U, ic + (m1 + kBx) * n, nU) + AtF(U, ib2 + (m1 + kBx) * n, nU));
  const double b3  = 0.5 * (AtF(U, ic + (m2 + kBx) * n, nU) + AtF(U, ib3 + (m2 + kBx) * n, nU));
  const double ec  = (mm3 * b2 - mm2 * b3) / d;
  const double ef  = fluxSign * AtF(flux, iflux + (int(fluxSign == 1) + kBx) * n, nF);
  return ef - ec;
}

// The cell-centred electric field term of _ctSlope (ct_electric_fields.h:102-114) of cell
// ic for the edge direction ctDir: (m3 * B2 - m2 * B3) / d with the face fields averaged to
// the cell centre; the same expression, operands and bounds guards as CtSlopeM.
__device__ __forceinline__ double CellEmf(const double *U, int ctDir, long ic, int nx, int ny, long n, long nU)
{
  const int m1 = (ctDir + 1) % 3, m2 = (ctDir + 2) % 3;
  const long ib2   = ic - (m1 == 0 ? 1 : (m1 == 1 ? long(nx) : long(nx) * long(ny)));
  const long ib3   = ic - (m2 == 0 ? 1 : (m2 == 1 ? long(nx) : long(nx) * long(ny)));
  const double d   = AtF(U, ic, nU);
  const double mm2 = AtF(U, ic + (m1 + 1) * n, nU);
  const double mm3 = AtF(U, ic + (m2 + 1) * n, nU);
  const double b2  = 0.5 * (AtF(U, ic + (m1 + kBx) * n, nU) + AtF(U, ib2 + (m1 + kBx) * n, nU));
  const double b3  = 0.5 * (AtF(U, ic + (m2 + kBx) * n, nU) + AtF(U, ib3 + (m2 + kBx) * n, nU));
  return (mm3 * b2 - mm2 * b3) / d;
}

// mhd::Calculate_CT_Electric_Fields for edges with every index in [lo, hi] (ct_electric_fields.cu:20-279),
// one lane per edge with bounds-guarded loads, in the shape of the reference's kernel: the
// reference-vector entry point (chl_ct_electric_fields) runs it over whole grids.  The step
// computes its edges in MhdCtUpdateKernel (CtEdge), with the same arithmetic per edge.
template <typename TF>
__global__ void __launch_bounds__(256) MhdEmfKernel(const TF *Fx, const TF *Fy, const TF *Fz, const double *U,
                                                    double *E, int nx, int ny, int nz, int lo, int hix, int hiy,
                                                    int hiz)
{
  const size_t n  = size_t(nx) * ny * nz;
  const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const int x = int(id % nx), y = int((id / nx) % ny), z = int(id / (size_t(nx) * ny));
  if (x < lo || y < lo || z < lo || x > hix || y > hiy || z > hiz) return;
  const long nF = 7 * long(n), nU = 8 * long(n);
  auto F = [&](const TF *a, int i, int j, int k, int f) {
    return AtF(a, long(i) + long(j) * nx + long(k) * long(nx) * ny + long(f) * long(n), nF);
  };
  auto S = [&](const TF *fl, double sgn, int ctd, int a, int b, int c, int d, int i, int j, int k) {
    return CtSlopeM(fl, U, sgn, ctd, a, b, c, d, i, j, k, nx, ny, long(n), nF, nU);
  };
  auto pick = [](double s, double pos, double neg) {
    return s > 0.0 ? pos : (s < 0.0 ? neg : 0.5 * (pos + neg));
  };
  const int i = x, j = y, k = z;
  {  // x electric field
    const double syp = pick(F(Fz, i, j, k - 1, 0), S(Fy, -1, 0, 2, -1, 1, 2, i, j, k), S(Fy, -1, 0, -1, -1, 1, -1, i, j, k));
    const double syn = pick(F(Fz, i, j - 1, k - 1, 0), S(Fy, -1, 0, 1, 2, 1, 2, i, j, k), S(Fy, -1, 0, 1, -1, 1, -1, i, j, k));
    const double szp = pick(F(Fy, i, j - 1, k, 0), S(Fz, 1, 0, 1, -1, 1, 2, i, j, k), S(Fz, 1, 0, -1, -1, 2, -1, i, j, k));
    const double szn = pick(F(Fy, i, j - 1, k - 1, 0), S(Fz, 1, 0, 1, 2, 1, 2, i, j, k), S(Fz, 1, 0, 2, -1, -1, 2, i, j, k));
    const double fyp = +F(Fz, i, j, k - 1, 6), fyn = +F(Fz, i, j - 1, k - 1, 6);
    const double fzp = -F(Fy, i, j - 1, k, 5), fzn = -F(Fy, i, j - 1, k - 1, 5);
    E[id] = 0.25 * (+fyp + fyn + fzp + fzn + syp + syn + szp + szn);
  }
  {  // y electric field
    const double sxp = pick(F(Fz, i, j, k - 1, 0), S(Fx, 1, 1, 2, -1, 0, 2, i, j, k), S(Fx, 1, 1, -1, -1, 0, -1, i, j, k));
    const double sxn = pick(F(Fz, i - 1, j, k - 1, 0), S(Fx, 1, 1, 0, 2, 0, 2, i, j, k), S(Fx, 1, 1, 0, -1, 0, -1, i, j, k));
    const double szp = pick(F(Fx, i - 1, j, k, 0), S(Fz, -1, 1, 0, -1, 0, 2, i, j, k), S(Fz, -1, 1, -1, -1, 2, -1, i, j, k));
    const double szn = pick(F(Fx, i - 1, j, k - 1, 0), S(Fz, -1, 1, 0, 2, 0, 2, i, j, k), S(Fz, -1, 1, 2, -1, 2, -1, i, j, k));
    const double fxp = -F(Fz, i, j, k - 1, 5), fxn = -F(Fz, i - 1, j, k - 1, 5);
    const double fzp = +F(Fx, i - 1, j, k, 6), fzn = +F(Fx, i - 1, j, k - 1, 6);
    E[n + id] = 0.25 * (+fxp + fxn + fzp + fzn + sxp + sxn + szp + szn);
  }
  {  // z electric field
    const double syp = pick(F(Fx, i - 1, j, k, 0), S(Fy, 1, 2, 0, -1, 0, 1, i, j, k), S(Fy, 1, 2, -1, -1, 1, -1, i, j, k));
    const double syn = pick(F(Fx, i - 1, j - 1, k, 0), S(Fy, 1, 2, 0, 1, 0, 1, i, j, k), S(Fy, 1, 2, 1, -1, 1, -1, i, j, k));
    const double sxp = pick(F(Fy, i, j - 1, k, 0), S(Fx, -1, 2, 1, -1, 0, 1, i, j, k), S(Fx, -1, 2, -1, -1, 0, -1, i, j, k));
    const double sxn = pick(F(Fy, i - 1, j - 1, k, 0), S(Fx, -1, 2, 0, 1, 0, 1, i, j, k), S(Fx, -1, 2, 0, -1, 0, -1, i, j, k));
    const double fxp = +F(Fy, i, j - 1, k, 6), fxn = +F(Fy, i - 1, j - 1, k, 6);
    const double fyp = -F(Fx, i - 1, j, k, 5), fyn = -F(Fx, i - 1, j - 1, k, 5);
    E[2 * n + id] = 0.25 * (+fxp + fxn + fyp + fyn + sxp + sxn + syp + syn);
  }
}

// One z plane of cells staged for the CT edges: for SX x SY cells of the plane, fields 0, 5
// and 6 of each direction's fluxes (the upwind mass flux and the two transverse field fluxes)
// and the three cell-centred terms of the slopes (CellEmf).
template <typename TF, int SX, int SY>
struct CtPlane {
  TF f[3][3][SY][SX];
  double ec[3][SY][SX];
};

// Stage the cell cid at [cy][cx] of P: its flux fields 0, 5 and 6 and its cell terms, from loads of the cell's density,
// momenta and face fields and of the face fields below it, which are CellEmf's operands in
// its order (CellEmf guards reads outside the array; the cells staged here have every one of
// them inside).  own (or null): the cell's 8 fields, for an update of that same state.
template <typename TF, int SX, int SY>
__device__ __forceinline__ void CtStageCell(CtPlane<TF, SX, SY> &P, int cx, int cy, const TF *const Fl[3],
                                            const double *U, size_t cid, size_t n, size_t sy, size_t sz, double *own)
{
#pragma unroll
  for (int d = 0; d < 3; d++) {
    P.f[d][0][cy][cx] = Fl[d][cid];
    P.f[d][1][cy][cx] = Fl[d][size_t(kBx) * n + cid];
    P.f[d][2][cy][cx] = Fl[d][size_t(kBx + 1) * n + cid];
  }
  const double d     = U[cid];
  const double m[3]  = {U[n + cid], U[2 * n + cid], U[3 * n + cid]};
  const double b[3]  = {U[5 * n + cid], U[6 * n + cid], U[7 * n + cid]};
  const double bm[3] = {U[5 * n + cid - 1], U[6 * n + cid - sy], U[7 * n + cid - sz]};
#pragma unroll
  for (int q = 0; q < 3; q++) {  // CellEmf(U, q, cid): (m3 b2 - m2 b3) / d
    const int m1 = (q + 1) % 3, m2 = (q + 2) % 3;
    const double b2 = 0.5 * (b[m1] + bm[m1]);
    const double b3 = 0.5 * (b[m2] + bm[m2]);
    P.ec[q][cy][cx] = (m[m2] * b2 - m[m1] * b3) / d;
  }
  if (own) {
    own[0] = d, own[1] = m[0], own[2] = m[1], own[3] = m[2], own[4] = U[4 * n + cid];
    own[5] = b[0], own[6] = b[1], own[7] = b[2];
  }
}

// The three electric fields of the edge whose own cell is staged at [ey + 1][ex + 1] of
// plane k (pl[k & 1]; plane k - 1 in pl[(k - 1) & 1]).  The arithmetic of MhdEmfKernel
// (ct_electric_fields.cu:20-279): slopes, upwind selection and sums in the same order.
template <typename TF, int SX, int SY>
__device__ __forceinline__ void CtEdge(const CtPlane<TF, SX, SY> *pl, int k, int ex, int ey, double &Ex, double &Ey,
                                       double &Ez)
{
  // (ox, oy, oz): the cell (i - ox, j - oy, k - oz) of the edge's 2 x 2 x 2 block
  auto flx = [&](int dir, int ox, int oy, int oz, int fi) -> double {
    return double(pl[(k - oz) & 1].f[dir][fi][ey + 1 - oy][ex + 1 - ox]);
  };
  // _ctSlope with the cell-centred term read back; a..d as in MhdEmfKernel's S(): the axes
  // (or -1) whose one-cell shifts locate the cell term (a, b) and the flux (c, d)
  auto slope = [&](int dir, double sgn, int ctd, int a, int b, int c, int d) -> double {
    const double ef = sgn * flx(dir, (c == 0) + (d == 0), (c == 1) + (d == 1), (c == 2) + (d == 2), 1 + int(sgn == 1));
    const int ox = (a == 0) + (b == 0), oy = (a == 1) + (b == 1), oz = (a == 2) + (b == 2);
    return ef - pl[(k - oz) & 1].ec[ctd][ey + 1 - oy][ex + 1 - ox];
  };
  auto up = [](double sgn_flux, double pos, double neg) -> double {
    return sgn_flux > 0.0 ? pos : (sgn_flux < 0.0 ? neg : 0.5 * (pos + neg));
  };
  {  // x electric field
    const double syp = up(flx(2, 0, 0, 1, 0), slope(1, -1, 0, 2, -1, 1, 2), slope(1, -1, 0, -1, -1, 1, -1));
    const double syn = up(flx(2, 0, 1, 1, 0), slope(1, -1, 0, 1, 2, 1, 2), slope(1, -1, 0, 1, -1, 1, -1));
    const double szp = up(flx(1, 0, 1, 0, 0), slope(2, 1, 0, 1, -1, 1, 2), slope(2, 1, 0, -1, -1, 2, -1));
    const double szn = up(flx(1, 0, 1, 1, 0), slope(2, 1, 0, 1, 2, 1, 2), slope(2, 1, 0, 2, -1, -1, 2));
    const double fyp = +flx(2, 0, 0, 1, 2), fyn = +flx(2, 0, 1, 1, 2);
    const double fzp = -flx(1, 0, 1, 0, 1), fzn = -flx(1, 0, 1, 1, 1);
    Ex = 0.25 * (+fyp + fyn + fzp + fzn + syp + syn + szp + szn);
  }
  {  // y electric field
    const double sxp = up(flx(2, 0, 0, 1, 0), slope(0, 1, 1, 2, -1, 0, 2), slope(0, 1, 1, -1, -1, 0, -1));
    const double sxn = up(flx(2, 1, 0, 1, 0), slope(0, 1, 1, 0, 2, 0, 2), slope(0, 1, 1, 0, -1, 0, -1));
    const double szp = up(flx(0, 1, 0, 0, 0), slope(2, -1, 1, 0, -1, 0, 2), slope(2, -1, 1, -1, -1, 2, -1));
    const double szn = up(flx(0, 1, 0, 1, 0), slope(2, -1, 1, 0, 2, 0, 2), slope(2, -1, 1, 2, -1, 2, -1));
    const double fxp = -flx(2, 0, 0, 1, 1), fxn = -flx(2, 1, 0, 1, 1);
    const double fzp = +flx(0, 1, 0, 0, 2), fzn = +flx(0, 1, 0, 1, 2);
    Ey = 0.25 * (+fxp + fxn + fzp + fzn + sxp + sxn + szp + szn);
  }
  {  // z electric field
    const double syp = up(flx(0, 1, 0, 0, 0), slope(1, 1, 2, 0, -1, 0, 1), slope(1, 1, 2, -1, -1, 1, -1));
    const double syn = up(flx(0, 1, 1, 0, 0), slope(1, 1, 2, 0, 1, 0, 1), slope(1, 1, 2, 1, -1, 1, -1));
    const double sxp = up(flx(1, 0, 1, 0, 0), slope(0, -1, 2, 1, -1, 0, 1), slope(0, -1, 2, -1, -1, 0, -1));
    const double sxn = up(flx(1, 1, 1, 0, 0), slope(0, -1, 2, 0, 1, 0, 1), slope(0, -1, 2, 0, -1, 0, -1));
    const double fxp = +flx(1, 0, 1, 0, 2), fxn = +flx(1, 1, 1, 0, 2);
    const double fyp = -flx(0, 1, 0, 0, 1), fyn = -flx(0, 1, 1, 0, 1);
    Ez = 0.25 * (+fxp + fxn + fyp + fyn + sxp + sxn + syp + syn);
  }
}

// ---------------------------------------------------------------------------- step arguments
struct MhdArgs {
  double *U;          // state (8 fields, in/out)
  double *Uh;         // half-step state (8 fields)
  void *F[3];         // flux arrays (7 fields each, P::T)
  int nx, ny, nz, ng;
  double dx, dy, dz, gamma;
  const double *dt_dev;  // device dt (device-timed step), or null
  double dt;
  unsigned long long *dti;
  double density_floor, U_floor;
  int flags;  // kFlag* bits of vl3d_kernels.hpp
  int recon;  // stage-2 reconstruction: CHL_RECON_PCM, _PLMC or _PPMC (0, 1, 2)
  // kFlagFoldDt: the device clock DtAdvanceKernel advances between the stage-1 fluxes (which
  // reduce dti) and the rest of the step; dt_dev points at clk_dt
  double *clk_dt, *clk_t;
  int *clk_nstep;
  double clk_tout, clk_cfl;
  RepairState rep;  // cells the full update leaves to RepairKernel (vl3d_kernels.hpp)
};

__device__ __forceinline__ double MhdDt(const MhdArgs &a) { return a.dt_dev ? *a.dt_dev : a.dt; }

// Region launches: a 1D grid of 256-lane workgroups over the box [lo, lo + W) x [lo, lo + H)
// x [lo, lo + D).  A row block is 256 consecutive cells of the region's x-y plane (row-major,
// rows of W cells: no lane idles at the row ends of a region whose width is not a multiple of
// 64).  MakeMhdSlab gives each lane one cell and orders the workgroups z-fastest within slabs of
// kSlab planes, inside each XCD's contiguous run of items (XcdItem): a cell's z neighbours are
// loaded by workgroups running at about the same time on the same L2, and its y neighbours
// by the next row block's.  MakeMhdMarch gives each lane a column of kchunk planes to visit in
// turn (the reductions, whose workgroups each fold into one atomic).
struct MhdRegion {
  int lo, W, H, D;
  int kz;     // planes per slab (MakeMhdSlab) or per column (MakeMhdMarch)
  int nxy;    // row blocks per plane
  int items;  // workgroups
};
constexpr int kSlab = 8;

inline MhdRegion MakeMhdRegion(int lo, int hix, int hiy, int hiz, int kz)
{
  MhdRegion m;
  m.lo = lo, m.W = hix - lo + 1, m.H = hiy - lo + 1, m.D = hiz - lo + 1, m.kz = kz;
  m.nxy   = (m.W * m.H + 255) / 256;
  m.items = m.nxy * ((m.D + kz - 1) / kz) * kz;
  return m;
}
inline MhdRegion MakeMhdSlab(int lo, int hix, int hiy, int hiz) { return MakeMhdRegion(lo, hix, hiy, hiz, kSlab); }
inline MhdRegion MakeMhdMarch(int lo, int hix, int hiy, int hiz, int kchunk)
{
  MhdRegion m = MakeMhdRegion(lo, hix, hiy, hiz, kchunk);
  m.items     = m.nxy * ((m.D + kchunk - 1) / kchunk);
  return m;
}
inline unsigned MhdRegionBlocks(const MhdRegion &m) { return unsigned((m.items + 7) / 8 * 8); }

// (x, y) of this lane in row block bxy; false past the plane's last cell.
__device__ __forceinline__ bool RowCell(const MhdRegion &m, int bxy, int &x, int &y)
{
  const int idx = bxy * 256 + int(threadIdx.x);
  if (idx >= m.W * m.H) return false;
  const int r = idx / m.W;
  y = m.lo + r;
  x = m.lo + (idx - r * m.W);
  return true;
}

// The cell of this lane in a slab launch; false if it has none.
__device__ __forceinline__ bool SlabCell(const MhdRegion &m, int &x, int &y, int &z)
{
  const int q = XcdItem(int(blockIdx.x), m.items);
  if (q >= m.items) return false;
  const int zc = q % m.kz, r = q / m.kz;
  const int bxy = r % m.nxy, slab = r / m.nxy;
  z = m.lo + slab * m.kz + zc;
  if (z >= m.lo + m.D) return false;
  return RowCell(m, bxy, x, y);
}

// This lane's column (x, y) and z range [z0, z1) in a march launch; false if it has none.
__device__ __forceinline__ bool MarchCell(const MhdRegion &m, int &x, int &y, int &z0, int &z1)
{
  const int q = XcdItem(int(blockIdx.x), m.items);
  if (q >= m.items) return false;
  const int bxy = q % m.nxy, bz = q / m.nxy;
  z0 = m.lo + bz * m.kz;
  z1 = min(z0 + m.kz, m.lo + m.D);
  return RowCell(m, bxy, x, y);
}

// ---------------------------------------------------------------------------- inverse dt
// mhdInverseCrossingTime (hydro_cuda.cu:420-435) with the cell-centred field
// of a cell's density, momentum, energy and cell-centred field
__device__ __forceinline__ double MhdInverseCrossingTimeOf(double d, double mx, double my, double mz, double E,
                                                           double bx, double by, double bz, double dx, double dy,
                                                           double dz, double g)
{
  const double dinv = 1.0 / d;
  const double vx = mx * dinv, vy = my * dinv, vz = mz * dinv;
  const double gasP = PressurePrimM(E, d, vx, vy, vz, g, bx, by, bz);
  const double cf   = Magnetosonic<PolicyExact>(d, gasP, bx, by, bz, g, 1.0);
  double m          = Fmax((Fabs(vx) + cf) / dx, (Fabs(vy) + cf) / dy);
  m                 = Fmax(m, (Fabs(vz) + cf) / dz);
  return Fmax(m, 0.0);
}

__device__ __forceinline__ double MhdInverseCrossingTime(const double *U, size_t id, int x, int y, int z, size_t n,
                                                         int nx, int ny, double dx, double dy, double dz, double g)
{
  double bx, by, bz;
  CellB(U, id, x, y, z, n, nx, ny, bx, by, bz);
  return MhdInverseCrossingTimeOf(U[id], U[n + id], U[2 * n + id], U[3 * n + id], U[4 * n + id], bx, by, bz, dx, dy,
                                  dz, g);
}

// Block maximum of non-negative doubles folded into *out with one atomic per block; every
// lane of the block calls it.
__device__ __forceinline__ void BlockMaxAtomic(double m, unsigned long long *out)
{
  __shared__ double red[1024 / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = Fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = red[0];
    for (int w = 1; w < int(blockDim.x) / 64; w++) b = Fmax(b, red[w]);
    atomicMax(out, (unsigned long long)__double_as_longlong(b));
  }
}

// The three stage-1 faces of cell (x, y, z) = id.
template <class P>
__device__ __forceinline__ void Flux1Cell(const MhdArgs &a, int x, int y, int z, size_t id, size_t n, size_t sy,
                                          size_t sz)
{
  using T      = typename P::T;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const double *U = a.U;
  const T g       = T(a.gamma);
  struct Cell {
    double d, m[3], E, b[3];
  };
  auto load = [&](int xx, int yy, int zz, size_t cid) {
    Cell c;
    c.d = U[cid], c.m[0] = U[n + cid], c.m[1] = U[2 * n + cid], c.m[2] = U[3 * n + cid], c.E = U[4 * n + cid];
    CellB(U, cid, xx, yy, zz, n, nx, ny, c.b[0], c.b[1], c.b[2]);
    return c;
  };
  auto pcm = [&](int dir, const Cell &c) {
    double mx, my, mz, bx, by, bz;
    RotM(dir, c.m[0], c.m[1], c.m[2], mx, my, mz);
    RotM(dir, c.b[0], c.b[1], c.b[2], bx, by, bz);
    return ConsM<T>{T(c.d), T(mx), T(my), T(mz), T(c.E), T(by), T(bz)};
  };
  const Cell own = load(x, y, z, id);
  auto face = [&](int dir, int xn, int yn, int zn, size_t nid) {
    const ConsM<T> qL = pcm(dir, own);
    const ConsM<T> qR = pcm(dir, load(xn, yn, zn, nid));
    ConsM<T> f;
    if (HlldM<P>(qL, qR, T(U[(kBx + dir) * n + id]), g, f)) StoreFluxM(reinterpret_cast<T *>(a.F[dir]), id, n, dir, f);
  };
  if (x <= nx - 2) face(0, x + 1, y, z, id + 1);
  if (y <= ny - 2) face(1, x, y + 1, z, id + sy);
  if (z <= nz - 2) face(2, x, y, z + 1, id + sz);
}

// Stage-1 fluxes of all three directions in one launch over the whole grid (MakeMhdSlab): each lane loads its cell once, forms the PCM state of every face it is the
// left cell of (i+1/2, j+1/2, k+1/2) by rotation, and solves the three faces with its upper
// neighbours: the PCM states (the cells with their cell-centred fields) and HLLD on the face
// i+1/2 of every cell whose face lies inside the grid.
//
// With kFlagFoldDt the launch also reduces the inverse crossing time of the real cells of the
// state it reads into *a.dti (which holds 0 or that same maximum), for the step's own dt: the
// device-timed step then needs no separate pass over the state.  The dt is not known yet, so
// the fluxes are computed even in a step that the output time will make empty.
template <class P>
__global__ void __launch_bounds__(256) MhdFlux1AllKernel(MhdArgs a, MhdRegion m)
{
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t n  = size_t(nx) * ny * nz;
  const size_t sy = size_t(nx), sz = size_t(nx) * ny;
  int x, y, z;
  if (a.flags & kFlagFoldDt) {
    double dti = 0.0;
    if (SlabCell(m, x, y, z)) {  // the slabs cover the whole grid
      const size_t id = size_t(x) + size_t(y) * sy + size_t(z) * sz;
      Flux1Cell<P>(a, x, y, z, id, n, sy, sz);
      const int ng = a.ng;
      if (x >= ng && y >= ng && z >= ng && x < nx - ng && y < ny - ng && z < nz - ng)
        dti = MhdInverseCrossingTime(a.U, id, x, y, z, n, nx, ny, a.dx, a.dy, a.dz, a.gamma);
    }
    BlockMaxAtomic(dti, a.dti);  // every lane gets here
    return;
  }
  if ((a.flags & kFlagReduceDt) && blockIdx.x == 0 && threadIdx.x == 0) *a.dti = 0ull;  // MhdCalcDtKernel fills it
  if (!SlabCell(m, x, y, z)) return;
  if (MhdDt(a) == 0.0) return;
  Flux1Cell<P>(a, x, y, z, size_t(x) + size_t(y) * sy + size_t(z) * sz, n, sy, sz);
}

// The same stage-1 fluxes (and, with kFlagFoldDt, the same inverse-dt reduction) by TX x TY
// workgroups that march a chunk of kz planes in z.  Each lane loads its cell of the next
// plane once: it is the z neighbour of the current plane's cell and then the next plane's own
// cell.  The x and y neighbours come from the current plane staged in LDS (the tile plus one
// column and one row of halo cells), as the PCM states in the arithmetic type.  A cell's
// values, rotations and HLLD are those of Flux1Cell.  Compiled for 4 waves per SIMD (the
// mixed form fits 128 VGPRs without spills; the FP64 forms keep MhdFlux1AllKernel).
template <class P, int TX, int TY>
__global__ void __launch_bounds__(TX *TY, 4) MhdFlux1MarchKernel(MhdArgs a, int kz)
{
  using T      = typename P::T;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t n  = size_t(nx) * ny * nz;
  const size_t sy = size_t(nx), sz = size_t(nx) * ny;
  const bool fold = (a.flags & kFlagFoldDt) != 0;
  if (!fold) {
    if ((a.flags & kFlagReduceDt) && blockIdx.x == 0 && threadIdx.x == 0) *a.dti = 0ull;  // MhdCalcDtKernel fills it
    if (MhdDt(a) == 0.0) return;
  }
  const int gx = (nx + TX - 1) / TX, gy = (ny + TY - 1) / TY, gz = (nz + kz - 1) / kz;
  const int q = XcdItem(int(blockIdx.x), gx * gy * gz);
  if (q >= gx * gy * gz) return;  // uniform per workgroup
  const int tid = int(threadIdx.x), tx = tid % TX, ty = tid / TX;
  const int x0 = (q % gx) * TX, y0 = ((q / gx) % gy) * TY;
  const int x = x0 + tx, y = y0 + ty;
  const int z0 = (q / (gx * gy)) * kz, z1 = min(z0 + kz, nz);
  const bool in = x < nx && y < ny;
  const double *U = a.U;
  const T g       = T(a.gamma);
  const int ng = a.ng;
  const bool real_xy = x >= ng && y >= ng && x < nx - ng && y < ny - ng;
  double dti = 0.0;
  // a cell in the arithmetic type: d, m x/y/z, E, B x/y/z (global orientation); its inverse
  // crossing time, from the FP64 values, is folded into dti as it is loaded
  struct Cell {
    T v[8];
  };
  auto load = [&](int xx, int yy, int zz, bool with_dt) {
    const size_t cid = size_t(xx) + size_t(yy) * sy + size_t(zz) * sz;
    const double d = U[cid], mx = U[n + cid], my = U[2 * n + cid], mz = U[3 * n + cid], E = U[4 * n + cid];
    double bx, by, bz;
    CellB(U, cid, xx, yy, zz, n, nx, ny, bx, by, bz);
    if (with_dt) dti = Fmax(dti, MhdInverseCrossingTimeOf(d, mx, my, mz, E, bx, by, bz, a.dx, a.dy, a.dz, a.gamma));
    return Cell{{T(d), T(mx), T(my), T(mz), T(E), T(bx), T(by), T(bz)}};
  };
  auto real = [&](int zz) { return fold && real_xy && zz >= ng && zz < nz - ng; };
  __shared__ T pl[8][TY + 1][TX + 1];  // the current plane and its halo
  auto stage = [&](const Cell &c, int r, int col) {
#pragma unroll
    for (int f = 0; f < 8; f++) pl[f][r][col] = c.v[f];
  };
  auto pcm = [&](int dir, const Cell &c) {
    T mx, my, mz, bx, by, bz;
    RotM(dir, c.v[1], c.v[2], c.v[3], mx, my, mz);
    RotM(dir, c.v[5], c.v[6], c.v[7], bx, by, bz);
    return ConsM<T>{c.v[0], mx, my, mz, c.v[4], by, bz};
  };
  auto staged = [&](int r, int col) {
    Cell c;
#pragma unroll
    for (int f = 0; f < 8; f++) c.v[f] = pl[f][r][col];
    return c;
  };
  auto solve = [&](int dir, size_t id, const ConsM<T> &qL, const ConsM<T> &qR) {
    ConsM<T> f;
    if (HlldM<P>(qL, qR, T(U[(kBx + dir) * n + id]), g, f)) StoreFluxM(reinterpret_cast<T *>(a.F[dir]), id, n, dir, f);
  };
  // halo lanes: the column x0 + TX (TY cells), then the row y0 + TY (TX cells)
  static_assert(TX + TY <= TX * TY, "one halo cell per lane");
  const bool is_h = tid < TX + TY;
  const int hx = tid < TY ? x0 + TX : x0 + (tid - TY), hy = tid < TY ? y0 + tid : y0 + TY;
  const int hr = tid < TY ? tid : TY, hc = tid < TY ? TX : tid - TY;
  const bool has_h = is_h && hx < nx && hy < ny;
  Cell own;
  if (in) own = load(x, y, z0, real(z0));
  for (int z = z0; z < z1; z++) {
    if (in) stage(own, ty, tx);
    if (has_h) stage(load(hx, hy, z, false), hr, hc);
    Cell up;
    if (in && z + 1 < nz) up = load(x, y, z + 1, z + 1 < z1 && real(z + 1));
    __syncthreads();
    if (in) {
      const size_t id = size_t(x) + size_t(y) * sy + size_t(z) * sz;
      if (x <= nx - 2) solve(0, id, pcm(0, own), pcm(0, staged(ty, tx + 1)));
      if (y <= ny - 2) solve(1, id, pcm(1, own), pcm(1, staged(ty + 1, tx)));
      if (z <= nz - 2) solve(2, id, pcm(2, own), pcm(2, up));
    }
    __syncthreads();
    own = up;
  }
  if (fold) BlockMaxAtomic(dti, a.dti);  // every lane of the workgroup gets here
}

// Stage-2 fluxes of the PCM build (pcm_cuda.cu:256-420 on the half-step state, then HLLD): the
// faces i+1/2, j+1/2, k+1/2 of every cell with each index in [lo, hi] (MakeMhdSlab), the
// interface states being the cells themselves as in stage 1.
template <class P>
__global__ void __launch_bounds__(256) MhdFlux2PcmKernel(MhdArgs a, MhdRegion m)
{
  int x, y, z;
  if (!SlabCell(m, x, y, z)) return;
  if (MhdDt(a) == 0.0) return;
  const size_t n  = size_t(a.nx) * a.ny * a.nz;
  const size_t sy = size_t(a.nx), sz = size_t(a.nx) * a.ny;
  MhdArgs h = a;
  h.U       = a.Uh;  // the cells of the half-step state
  Flux1Cell<P>(h, x, y, z, size_t(x) + size_t(y) * sy + size_t(z) * sz, n, sy, sz);
}

// The edge fields the face-field update of one cell reads (Update_Magnetic_Field_3D,
// magnetic_update.cu:21-81), named by field and by the offset of the edge's index from the
// cell's: ey_x = Ey[i+1, j, k], ex_yz = Ex[i, j+1, k+1], ...  Edges of plane k: ey_x, ez_x,
// ez_xy, ez_y, ex_y; of plane k+1: ey_xz, ex_yz, ex_z, ey_z.
struct EdgeSet {
  double ey_x, ez_x, ez_xy, ez_y, ex_y;
  double ey_xz, ex_yz, ex_z, ey_z;
};

// The state and flux values one cell's update reads: its 8 fields and the 5 hydro fluxes of
// its lower (fm) and upper (fp) faces per direction.
template <typename TF>
struct CellIn {
  double u[8];
  TF fm[3][5], fp[3][5];
};

// Half update: hydro fields of cells [1, n-2] (Update_Conserved_Variables_3D_half,
// VL_3D_cuda.cu:379-495, with the density floor) and their face fields
// (Update_Magnetic_Field_3D, magnetic_update.cu:21-81, from U into Uh).
template <class P>
__device__ __forceinline__ void HalfCell(const MhdArgs &a, double dt, int x, int y, int z,
                                         const CellIn<typename P::T> &c, const EdgeSet &e)
{
  const size_t n  = size_t(a.nx) * a.ny * a.nz;
  const size_t id = size_t(x) + size_t(y) * a.nx + size_t(z) * size_t(a.nx) * a.ny;
  const double dtodx = dt / a.dx, dtody = dt / a.dy, dtodz = dt / a.dz;
  double *Uh = a.Uh;
  double uh[5];
#pragma unroll
  for (int f = 0; f < 5; f++) {
    uh[f] = c.u[f] + dtodx * (double(c.fm[0][f]) - double(c.fp[0][f])) +
            dtody * (double(c.fm[1][f]) - double(c.fp[1][f])) + dtodz * (double(c.fm[2][f]) - double(c.fp[2][f]));
  }
  if ((a.flags & kFlagDensityFloor) && uh[0] < a.density_floor) {
    const double d0 = uh[0];
    uh[0]           = a.density_floor;
#pragma unroll
    for (int f = 1; f < 5; f++) uh[f] *= (a.density_floor / d0);
  }
#pragma unroll
  for (int f = 0; f < 5; f++) Uh[size_t(f) * n + id] = uh[f];
  // face fields whose edge fields were computed (edges [2, n-1]): own index >= 1, others >= 2
  if (y >= 2 && z >= 2) Uh[5 * n + id] = c.u[5] + dtodz * (e.ey_xz - e.ey_x) + dtody * (e.ez_x - e.ez_xy);
  if (x >= 2 && z >= 2) Uh[6 * n + id] = c.u[6] + dtodx * (e.ez_xy - e.ez_y) + dtodz * (e.ex_y - e.ex_yz);
  if (x >= 2 && y >= 2) Uh[7 * n + id] = c.u[7] + dtody * (e.ex_yz - e.ex_z) + dtodx * (e.ey_z - e.ey_xz);
}

// Stage-2 fluxes along x with each cell's primitives (Load_Data) recovered once and each
// cell's reconstruction (RK: 1 PLMC, 2 PPMC) computed once: a workgroup takes 64 (x) by 4
// (y) faces of one z plane, stages the primitives of the 64 + 1 + 2 R cells of each row in
// LDS (one array per field), reconstructs the cells of its faces plus the next cell along x
// from that stencil, keeps each cell's face-(i+1/2) state in registers and passes its
// face-(i-1/2) state to the neighbour through LDS.  The arithmetic of each reconstruction
// and flux is the per-cell one.  The fast FP64 form is held to 3 waves per SIMD (168 VGPRs,
// no spills; the exact form would spill).
template <class P, int RK = 1>
__global__ void __launch_bounds__(256, (P::kId == PolicyFast::kId ? 3 : 1)) MhdFlux2TileKernel(MhdArgs a, int lo, int hix, int hiy, int hiz)
{
  using T          = typename P::T;
  constexpr int TW = 64, TH = 4;           // faces: x by y
  constexpr int R  = RK == 2 ? 2 : 1;      // stencil half-width
  constexpr int SW = TW + 1 + 2 * R;       // staged cells per row: x0 - R .. x0 + TW + R
  __shared__ T pr[8][TH][SW];              // the staged primitives, field-major
  __shared__ PrimM<T> rm[TH][TW + 1];      // each cell's face-(i-1/2) state
  if (MhdDt(a) == 0.0) return;
  const int nx = a.nx, ny = a.ny;
  const size_t n = size_t(nx) * ny * a.nz;
  const int tx = threadIdx.x % TW, ty = threadIdx.x / TW;
  // tiles in XCD-contiguous order (XcdItem): x fastest, then y, then z
  const int gx = (hix - lo + TW) / TW, gs = (hiy - lo + TH) / TH, gt = hiz - lo + 1;
  const int q = XcdItem(int(blockIdx.x), gx * gs * gt);
  if (q >= gx * gs * gt) return;  // uniform per workgroup
  const int x0 = lo + (q % gx) * TW, s0 = lo + ((q / gx) % gs) * TH, z = lo + q / (gx * gs);
  const double *Uh = a.Uh;
  // stage the cells some reconstructed cell reads: x <= hix + 1 + R, y <= hiy
  for (int i = int(threadIdx.x); i < TH * SW; i += 256) {
    const int r = i / SW, c = i % SW, x = x0 - R + c, y = s0 + r;
    if (x <= hix + 1 + R && y <= hiy) {
      const PrimM<T> w = LoadDataM<P>(Uh, x, y, z, nx, ny, n, 0, a.gamma);
      pr[0][r][c] = w.d, pr[1][r][c] = w.vx, pr[2][r][c] = w.vy, pr[3][r][c] = w.vz;
      pr[4][r][c] = w.p, pr[5][r][c] = w.bx, pr[6][r][c] = w.by, pr[7][r][c] = w.bz;
    }
  }
  __syncthreads();
  auto cell = [&](int r, int c) {
    return PrimM<T>{pr[0][r][c], pr[1][r][c], pr[2][r][c], pr[3][r][c],
                    pr[4][r][c], pr[5][r][c], pr[6][r][c], pr[7][r][c]};
  };
  auto reconstruct = [&](int r, int c, PrimM<T> &Lp, PrimM<T> &Rm) {  // the cell at column c + R
    if constexpr (RK == 2) {
      PpmcPrimsM<P>(cell(r, c), cell(r, c + 1), cell(r, c + 2), cell(r, c + 3), cell(r, c + 4), a.gamma, Lp, Rm);
    } else {
      PlmcPrimsM<P>(cell(r, c), cell(r, c + 1), cell(r, c + 2), a.gamma, Lp, Rm);
    }
  };
  PrimM<T> Lp, Rm;
  const int fx = x0 + tx, fs = s0 + ty;
  const bool mine = fx <= hix + 1 && fs <= hiy;  // a cell some face of the region reads
  if (mine) {
    reconstruct(ty, tx, Lp, Rm);
    rm[ty][tx] = Rm;
  }
  // the next cell along x
  if (threadIdx.x < TH && x0 + TW <= hix + 1 && s0 + int(threadIdx.x) <= hiy) {
    PrimM<T> l2, r2;
    reconstruct(int(threadIdx.x), TW, l2, r2);
    rm[threadIdx.x][TW] = r2;
  }
  __syncthreads();
  if (!mine || fx > hix) return;
  const size_t id   = size_t(fx) + size_t(fs) * nx + size_t(z) * size_t(nx) * ny;
  const PrimM<T> Rn = rm[ty][tx + 1];
  const T g         = T(a.gamma);
  const T bx        = T(Uh[kBx * n + id]);
  const ConsM<T> qL = WriteDataM<P>(Lp, bx, g), qR = WriteDataM<P>(Rn, bx, g);
  ConsM<T> f;
  if (HlldM<P>(qL, qR, bx, g, f)) StoreFluxM(reinterpret_cast<T *>(a.F[0]), id, n, 0, f);
}

// Stage-2 fluxes along DIR = y or z by a march along DIR: each lane owns a column (x and the
// third axis) of a 64 x 4 block and walks a chunk of kMarch faces.  Each cell's primitives
// (Load_Data) are recovered once and slide through the reconstruction stencil's window, each
// cell is reconstructed once, and its face-(i+1/2) state carries to the next face: a chunk of
// K faces loads K + 2 R cells and reconstructs K + 1 (a 64 x 4 face tile across DIR would
// reconstruct 5 rows for 4, and each reconstruction recovers its 2 R + 1 cells' primitives).
// Each face's flux is the tile kernel's: HLLD of the two cells' interface states, stored at
// the lower cell.
constexpr int kMarch = 16;
template <class P, int DIR, int RK = 1>
__global__ void __launch_bounds__(256) MhdFlux2MarchKernel(MhdArgs a, int lo, int hix, int hiy, int hiz)
{
  static_assert(DIR == 1 || DIR == 2, "the x faces use MhdFlux2TileKernel");
  using T = typename P::T;
  if (MhdDt(a) == 0.0) return;
  const int nx = a.nx, ny = a.ny;
  const size_t n = size_t(nx) * ny * a.nz;
  const int his = DIR == 2 ? hiz : hiy, hit = DIR == 2 ? hiy : hiz;  // march axis and third axis bounds
  const int tx = threadIdx.x % 64, tt = threadIdx.x / 64;
  // blocks in XCD-contiguous order: x fastest, then the third axis, then the chunk along DIR
  const int gx = (hix - lo + 64) / 64, gt = (hit - lo + 4) / 4, gs = (his - lo + kMarch) / kMarch;
  const int q = XcdItem(int(blockIdx.x), gx * gt * gs);
  if (q >= gx * gt * gs) return;
  const int x = lo + (q % gx) * 64 + tx, t = lo + ((q / gx) % gt) * 4 + tt;
  const int s0 = lo + (q / (gx * gt)) * kMarch, s1 = min(s0 + kMarch - 1, his);
  if (x > hix || t > hit) return;  // no barrier below
  const double *Uh = a.Uh;
  const T g        = T(a.gamma);
  auto load = [&](int s) {
    return DIR == 2 ? LoadDataM<P>(Uh, x, t, s, nx, ny, n, DIR, a.gamma)
                    : LoadDataM<P>(Uh, x, s, t, nx, ny, n, DIR, a.gamma);
  };
  // the window holds cells s - R .. s + R of the cell s being reconstructed
  constexpr int R = RK == 2 ? 2 : 1;
  PrimM<T> w[2 * R + 1];
  for (int i = 0; i < 2 * R + 1; i++) w[i] = load(s0 - R + i);
  auto reconstruct = [&](PrimM<T> &Lp, PrimM<T> &Rm) {
    if constexpr (RK == 2) {
      PpmcPrimsM<P>(w[0], w[1], w[2], w[3], w[4], a.gamma, Lp, Rm);
    } else {
      PlmcPrimsM<P>(w[0], w[1], w[2], a.gamma, Lp, Rm);
    }
  };
  PrimM<T> Lprev, Rskip;
  reconstruct(Lprev, Rskip);
  for (int s = s0; s <= s1; s++) {
    for (int i = 0; i < 2 * R; i++) w[i] = w[i + 1];
    w[2 * R] = load(s + 1 + R);
    PrimM<T> Lnext, Rnext;
    reconstruct(Lnext, Rnext);
    const int y = DIR == 2 ? t : s, z = DIR == 2 ? s : t;
    const size_t id   = size_t(x) + size_t(y) * nx + size_t(z) * size_t(nx) * ny;
    const T bx        = T(Uh[(kBx + DIR) * n + id]);
    const ConsM<T> qL = WriteDataM<P>(Lprev, bx, g), qR = WriteDataM<P>(Rnext, bx, g);
    ConsM<T> f;
    if (HlldM<P>(qL, qR, bx, g, f)) StoreFluxM(reinterpret_cast<T *>(a.F[DIR]), id, n, DIR, f);
    Lprev = Lnext;
  }
}

// Full update (Update_Conserved_Variables_3D, hydro_cuda.cu:172-402, and
// Update_Magnetic_Field_3D in place): hydro fields of the real cells with the density
// floor and the temperature floor (Temperature_Floor_Kernel, hydro_cuda.cu:1117-1153);
// face fields whose own index is in [ng-1, n-ng-1] and other indices in [ng, n-ng-1].
template <class P>
__device__ __forceinline__ void FullCell(const MhdArgs &a, double dt, int x, int y, int z,
                                         const CellIn<typename P::T> &c, const EdgeSet &e)
{
  const int nx   = a.nx, ny = a.ny, nz = a.nz, ng = a.ng;
  const size_t n = size_t(nx) * ny * nz;
  const size_t id = size_t(x) + size_t(y) * nx + size_t(z) * size_t(nx) * ny;
  const bool rx = x >= ng, ry = y >= ng, rz = z >= ng;
  const double dtodx = dt / a.dx, dtody = dt / a.dy, dtodz = dt / a.dz;
  double *U       = a.U;
  if (rx && ry && rz) {
    double un[5];
#pragma unroll
    for (int f = 0; f < 5; f++) {
      un[f] = c.u[f] + (dtodx * (double(c.fm[0][f]) - double(c.fp[0][f])) +
                        dtody * (double(c.fm[1][f]) - double(c.fp[1][f])) +
                        dtodz * (double(c.fm[2][f]) - double(c.fp[2][f])));
    }
    int mark = 0;
    if ((a.flags & kFlagDensityFloor) && un[0] < a.density_floor) {
      if (un[0] > 0) {
        const double d0 = un[0];
        un[0]           = a.density_floor;
#pragma unroll
        for (int f = 1; f < 5; f++) un[f] *= (a.density_floor / d0);
      } else {
        mark = kRepairDensity;  // Average_Cell_Single_Field, done by RepairKernel in cell order
      }
    }
    if (mark == 0 && (a.flags & kFlagCrashRepair) && CrashedCell(un[0], un[4])) mark = kRepairAll;
    if (mark) {
      MarkRepair(a.rep, id, mark);
    } else if (a.flags & kFlagTemperatureFloor) {
      const double dinv = 1.0 / un[0];
      const double vx = un[1] * dinv, vy = un[2] * dinv, vz = un[3] * dinv;
      const double Ekin = 0.5 * un[0] * (vx * vx + vy * vy + vz * vz);
      if ((un[4] - Ekin) / un[0] < a.U_floor) un[4] = Ekin + un[0] * a.U_floor;
    }
#pragma unroll
    for (int f = 0; f < 5; f++) U[size_t(f) * n + id] = un[f];
  }
  if (ry && rz) U[5 * n + id] = c.u[5] + dtodz * (e.ey_xz - e.ey_x) + dtody * (e.ez_x - e.ez_xy);
  if (rx && rz) U[6 * n + id] = c.u[6] + dtodx * (e.ez_xy - e.ez_y) + dtodz * (e.ex_y - e.ex_yz);
  if (rx && ry) U[7 * n + id] = c.u[7] + dtody * (e.ex_yz - e.ex_z) + dtodx * (e.ey_z - e.ey_xz);
}

// CT edge fields and the update of one stage in one launch.  FULL false: the half update
// (HalfCell, U -> Uh over the cells [clo, ch*]) from the edges of the stage-1 fluxes and U;
// FULL true: the full update (FullCell, U in place) from the edges of the stage-2 fluxes and
// the half-step state.  A CTX x CTY tile of cells marches in z over a chunk of kchunk planes;
// at plane k it stages the cells of plane k + 1 (CtStageCell), computes the (CTX+1) x (CTY+1)
// edges of plane k + 1 into LDS (CtEdge, the arithmetic of MhdEmfKernel), and updates its
// cells of plane k from the edges of plane k (kept in registers since the previous plane) and
// of plane k + 1.  The edge fields never reach HBM and each cell's state and fluxes are read
// once per stage.  Edges outside [elo, eh*] are not computed and read as 0: no update that
// runs uses them.  FULL: the tile's lowest column, row and plane also update the face fields
// FullCell updates on the cells just below [clo, ch*] (Bx of (clo-1, j, k), By of (i, clo-1,
// k), Bz of (i, j, clo-1)), whose edges are the tile's lowest ones; then the cells [clo, ch*]
// are the real cells, and the tiles divide them evenly.
template <class P, int CTX, int CTY, bool FULL>
__global__ void __launch_bounds__(CTX *CTY) MhdCtUpdateKernel(MhdArgs a, int clo, int chx, int chy, int chz,
                                                              int elo, int ehx, int ehy, int ehz, int kchunk)
{
  using TF         = typename P::T;
  constexpr int NT = CTX * CTY, EX = CTX + 1, EY = CTY + 1;
  __shared__ CtPlane<TF, CTX + 2, CTY + 2> pl[2];
  __shared__ double eds[3][EY][EX];
  const double dt = FULL ? MhdDt(a) : 0.5 * MhdDt(a);
  if (dt == 0.0) return;
  // tiles in XCD-contiguous order (XcdItem): x fastest, then y, then the z chunk
  const int gx = (chx - clo + CTX) / CTX, gy = (chy - clo + CTY) / CTY, gz = (chz - clo + kchunk) / kchunk;
  const int q = XcdItem(int(blockIdx.x), gx * gy * gz);
  if (q >= gx * gy * gz) return;
  const int bx = q % gx, by = (q / gx) % gy, bz = q / (gx * gy);
  const int tx = int(threadIdx.x) % CTX, ty = int(threadIdx.x) / CTX;
  const int i0 = clo + bx * CTX, j0 = clo + by * CTY;
  const int k0 = clo + bz * kchunk, k1 = min(k0 + kchunk - 1, chz);
  const int nx = a.nx, ny = a.ny;
  const size_t n        = size_t(nx) * ny * a.nz;
  const TF *const Fl[3] = {reinterpret_cast<const TF *>(a.F[0]), reinterpret_cast<const TF *>(a.F[1]),
                           reinterpret_cast<const TF *>(a.F[2])};
  const double *Us = FULL ? a.Uh : a.U;  // the state of the edges' cell terms
  const size_t sy = size_t(nx), sz = size_t(nx) * ny;
  // staged cell (cx, cy) = (i0 - 1 + cx, j0 - 1 + cy); edge (ex, ey) = (i0 + ex, j0 + ey),
  // whose own cell is staged at [ey + 1][ex + 1].  Staged: the cells the computed edges read
  // and the lower x / y neighbours of the updated cells (whose mass fluxes the update reads
  // back), every index in [slo, eh*].  Each thread stages its own cell (tx + 1, ty + 1), and
  // the first NH threads one cell of the tile's rim each.
  constexpr int SX = CTX + 2, SY = CTY + 2, NH = 2 * SX + 2 * (SY - 2);
  static_assert(NH <= NT, "one rim cell per thread");
  const int slo = min(elo, clo) - 1;
  auto in_box = [&](int gi, int gj, int kk) {
    return gi >= slo && gj >= slo && kk >= slo && gi <= ehx && gj <= ehy && kk <= ehz;
  };
  auto stage = [&](int kk, double *own) {
    CtPlane<TF, SX, SY> &Pl = pl[kk & 1];
    const size_t base       = size_t(kk) * sz;
    {
      const int gi = i0 + tx, gj = j0 + ty;
      if (in_box(gi, gj, kk)) CtStageCell(Pl, tx + 1, ty + 1, Fl, Us, size_t(gi) + size_t(gj) * sy + base, n, sy, sz, own);
    }
    const int t = int(threadIdx.x);
    if (t < NH) {
      int cx, cy;
      if (t < SX) {
        cx = t, cy = 0;
      } else if (t < 2 * SX) {
        cx = t - SX, cy = SY - 1;
      } else {
        const int r = t - 2 * SX;
        cx = (r & 1) ? SX - 1 : 0, cy = 1 + (r >> 1);
      }
      const int gi = i0 - 1 + cx, gj = j0 - 1 + cy;
      if (in_box(gi, gj, kk)) CtStageCell(Pl, cx, cy, Fl, Us, size_t(gi) + size_t(gj) * sy + base, n, sy, sz, nullptr);
    }
  };
  auto edges = [&](int ek) {
    for (int c = int(threadIdx.x); c < EX * EY; c += NT) {
      const int ey = c / EX, ex = c - ey * EX;
      const int gi = i0 + ex, gj = j0 + ey;
      double vx = 0.0, vy = 0.0, vz = 0.0;
      if (gi >= elo && gj >= elo && ek >= elo && gi <= ehx && gj <= ehy && ek <= ehz) CtEdge(pl, ek, ex, ey, vx, vy, vz);
      eds[0][ey][ex] = vx, eds[1][ey][ex] = vy, eds[2][ey][ex] = vz;
    }
  };
  const int i = i0 + tx, j = j0 + ty;
  const bool mine = i <= chx && j <= chy;
  const bool left_x = FULL && tx == 0 && i0 == clo && mine, left_y = FULL && ty == 0 && j0 == clo && mine;
  EdgeSet e{};
  double lx_ey = 0.0, lx_ez = 0.0, lx_ez1 = 0.0;  // Ey(i0, j, k), Ez(i0, j, k), Ez(i0, j+1, k)
  double ly_ez = 0.0, ly_ez1 = 0.0, ly_ex = 0.0;  // Ez(i, j0, k), Ez(i+1, j0, k), Ex(i, j0, k)
  auto read_k = [&]() {  // this thread's edges of the plane in LDS, as plane k
    e.ey_x = eds[1][ty][tx + 1], e.ez_x = eds[2][ty][tx + 1], e.ez_xy = eds[2][ty + 1][tx + 1];
    e.ez_y = eds[2][ty + 1][tx], e.ex_y = eds[0][ty + 1][tx];
    if constexpr (FULL) {
      lx_ey = eds[1][ty][0], lx_ez = eds[2][ty][0], lx_ez1 = eds[2][ty + 1][0];
      ly_ez = eds[2][0][tx], ly_ez1 = eds[2][0][tx + 1], ly_ex = eds[0][0][tx];
    }
  };
  auto read_k1 = [&](EdgeSet &s) {  // ... as plane k + 1
    s.ey_xz = eds[1][ty][tx + 1], s.ex_yz = eds[0][ty + 1][tx], s.ex_z = eds[0][ty][tx], s.ey_z = eds[1][ty][tx];
  };
  CellIn<TF> lc{};  // the face field below the tile (FULL: Bz of (i, j, clo - 1), then Bx / By)
  const bool left_z = FULL && k0 == clo && mine;
  if (left_z) lc.u[7] = a.U[7 * n + size_t(i) + size_t(j) * sy + size_t(clo - 1) * sz];
  // the half update's cell state comes from its own staging (own_k: plane k, own_n: k + 1);
  // fz_lo: the hydro fluxes of the face below the cell, carried from the plane below
  double own_k[8] = {}, own_n[8] = {};
  TF fz_lo[5] = {};
  if (mine) {
    const size_t id0 = size_t(i) + size_t(j) * sy + size_t(k0) * sz;
#pragma unroll
    for (int f = 0; f < 5; f++) fz_lo[f] = Fl[2][size_t(f) * n + id0 - sz];
  }
  stage(k0 - 1, nullptr);
  stage(k0, FULL ? nullptr : own_k);
  __syncthreads();
  edges(k0);
  __syncthreads();
  read_k();
  if (left_z) {  // Bz of (i, j, clo - 1): the edges of plane clo
    EdgeSet zl{};
    read_k1(zl);
    FullCell<P>(a, dt, i, j, clo - 1, lc, zl);
  }
  for (int k = k0; k <= k1; k++) {
    stage(k + 1, FULL ? nullptr : own_n);
    __syncthreads();
    edges(k + 1);
    __syncthreads();
    read_k1(e);
    if (mine) {
      // the cell's update inputs: the state (staged, or loaded for the full update, whose
      // edges use the half-step state), the mass fluxes from the staged plane k, the other
      // fluxes loaded, and the z-face below from the plane below
      const size_t id          = size_t(i) + size_t(j) * sy + size_t(k) * sz;
      const CtPlane<TF, SX, SY> &Pk = pl[k & 1];
      CellIn<TF> c;
#pragma unroll
      for (int f = 0; f < 8; f++) c.u[f] = FULL ? a.U[size_t(f) * n + id] : own_k[f];
      c.fp[0][0] = Pk.f[0][0][ty + 1][tx + 1], c.fm[0][0] = Pk.f[0][0][ty + 1][tx];
      c.fp[1][0] = Pk.f[1][0][ty + 1][tx + 1], c.fm[1][0] = Pk.f[1][0][ty][tx + 1];
      c.fp[2][0] = Pk.f[2][0][ty + 1][tx + 1], c.fm[2][0] = fz_lo[0];
#pragma unroll
      for (int f = 1; f < 5; f++) {
        const size_t o = size_t(f) * n + id;
        c.fp[0][f] = Fl[0][o], c.fm[0][f] = Fl[0][o - 1];
        c.fp[1][f] = Fl[1][o], c.fm[1][f] = Fl[1][o - sy];
        c.fp[2][f] = Fl[2][o], c.fm[2][f] = fz_lo[f];
      }
#pragma unroll
      for (int f = 0; f < 5; f++) fz_lo[f] = c.fp[2][f];
      if constexpr (FULL) {
        const double bl_x = left_x ? a.U[5 * n + id - 1] : 0.0;
        const double bl_y = left_y ? a.U[6 * n + id - sy] : 0.0;
        FullCell<P>(a, dt, i, j, k, c, e);
        if (left_x) {  // Bx of (clo - 1, j, k)
          EdgeSet l{};
          l.ey_x = lx_ey, l.ez_x = lx_ez, l.ez_xy = lx_ez1, l.ey_xz = eds[1][ty][0];
          lc.u[5] = bl_x;
          FullCell<P>(a, dt, clo - 1, j, k, lc, l);
        }
        if (left_y) {  // By of (i, clo - 1, k)
          EdgeSet l{};
          l.ez_y = ly_ez, l.ez_xy = ly_ez1, l.ex_y = ly_ex, l.ex_yz = eds[0][0][tx];
          lc.u[6] = bl_y;
          FullCell<P>(a, dt, i, clo - 1, k, lc, l);
        }
      } else {
        HalfCell<P>(a, dt, i, j, k, c, e);
      }
    }
    if constexpr (!FULL) {
#pragma unroll
      for (int f = 0; f < 8; f++) own_k[f] = own_n[f];
    }
    read_k();  // plane k + 1 is the next plane k; LDS is rewritten only after the next barrier
  }
}

}  // namespace mhd

// Host launchers (mhd_exact.hip / mhd_fast.hip)
hipError_t LaunchMhdStep(int policy, const mhd::MhdArgs &a, hipStream_t s);
hipError_t LaunchMhdCalcDt(const double *U, int nx, int ny, int nz, int ng, double dx, double dy, double dz,
                           double gamma, unsigned long long *dti, int num_cu, hipStream_t s);
hipError_t LaunchMhdDivergence(const double *U, int nx, int ny, int nz, double dx, double dy, double dz,
                               unsigned long long *out, int num_cu, hipStream_t s);
// reference-shaped stage kernels (FP64, reference operation order) for the unit vectors
hipError_t LaunchHlldGrid(const double *QL, const double *QR, const double *magX, double *F, long n_cells,
                          double gamma, int dir, hipStream_t s);
hipError_t LaunchMhdPlmcGrid(const double *U, double *QL, double *QR, int nx, int ny, int nz, double gamma, int dir,
                             hipStream_t s);
hipError_t LaunchMhdPpmcGrid(const double *U, double *QL, double *QR, int nx, int ny, int nz, double gamma, int dir,
                             hipStream_t s);
hipError_t LaunchCtGrid(const double *Fx, const double *Fy, const double *Fz, const double *U, double *E, int nx,
                        int ny, int nz, hipStream_t s);
hipError_t LaunchMagneticUpdateGrid(const double *src, double *dst, const double *E, int nx, int ny, int nz,
                                    double dt, double dx, double dy, double dz, hipStream_t s);

}  // namespace chl
