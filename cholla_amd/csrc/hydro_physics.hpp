// hydro_physics.hpp -- device physics of the VL hydro step for gfx950.
//
// One source, three arithmetic policies:
//   PolicyExact  FP64, every expression in the reference's operation order and
//                association; the translation unit that instantiates it is
//                compiled with -ffp-contract=off so results are bit-identical to
//                a strict IEEE restatement of the reference kernels.
//   PolicyFast   FP64, divisions by a common denominator replaced by one
//                reciprocal and products; a product and a sum in one expression
//                fuse into an FMA (-ffp-contract=on, so every kernel that inlines
//                a tile body rounds it the same way).
//   PolicyMixed  FP32 reconstruction / characteristic projection / Riemann
//                solve with hardware rcp and sqrt; the conserved state, its
//                flux-difference accumulation, pressure recovery and dt stay FP64
//                (handled by the callers, which convert at the boundary).
//                Reconstruction runs on deviations from the cell-centred state and
//                the pressure is carried as an FP32 pair (high part + low part), so
//                the FP32 rounding in a face value scales with the local variation,
//                not with the size of the state (see SplitPressure below).
//
// States are held "rotated": vx / mx is the component normal to the sweep
// direction, (vy, vz) the two transverse ones in the reference's cyclic order
// (o1, o2, o3) = (x,y,z), (y,z,x), (z,x,y)  (src/riemann_solvers/hllc_cuda.cu:54-69,
// src/reconstruction/ppmc_cuda.cu:557-574).
#pragma once

#include <hip/hip_runtime.h>

namespace chl
{

constexpr double kTinyNumber = 1.0e-20;  // TINY_NUMBER, src/global/global.h:25
// Dual energy (the DE build option): GasEnergy is field 5 (no scalars, no MHD) and the
// internal energy switch thresholds are DE_ETA_1 / DE_ETA_2 (src/global/global.h:62-73).
constexpr int kGasEnergy = 5;
constexpr double kDeEta1 = 0.001;
constexpr double kDeEta2 = 0.035;

// Riemann solvers (the reference's -DHLLC / -DHLL build options).
constexpr int kRiemannHllc = 0;
constexpr int kRiemannHll  = 1;
constexpr int kRiemannRoe  = 2;
constexpr int kRiemannExact = 3;

struct PolicyExact {
  using T                      = double;
  static constexpr bool kExact = true;
  static constexpr int kId     = 0;
  static constexpr int kSolver = kRiemannHllc;
  static constexpr bool kGrav  = false;
};
struct PolicyFast {
  using T                      = double;
  static constexpr bool kExact = false;
  static constexpr int kId     = 1;
  static constexpr int kSolver = kRiemannHllc;
  static constexpr bool kGrav  = false;
};
struct PolicyMixed {
  using T                      = float;
  static constexpr bool kExact = false;
  static constexpr int kId     = 2;
  static constexpr int kSolver = kRiemannHllc;
  static constexpr bool kGrav  = false;
};
// An arithmetic form with another Riemann solver in place of HLLC; everything else as Base.
template <class Base, int S>
struct SolverPolicy : Base {
  static constexpr int kSolver = S;
};
// An arithmetic form whose full update adds the static-gravity source (STATIC_GRAV).
template <class Base>
struct GravPolicy : Base {
  static constexpr bool kGrav = true;
};
template <class Base>
using HllPolicy = SolverPolicy<Base, kRiemannHll>;
template <class Base>
using RoePolicy = SolverPolicy<Base, kRiemannRoe>;
template <class Base>
using ExactPolicy = SolverPolicy<Base, kRiemannExact>;

// The mixed form's pressure treatment ("split pressure").  A uniform pressure p0 adds
// exactly p0 to the normal-momentum flux of every face solver here and cancels in a
// cell's flux difference, so the FP32 solve carries only the pressure relative to a
// reference: each face takes the high FP32 part of its left cell's FP64 pressure as p0,
// the solver's momentum flux is returned less p0, and the update adds the FP64
// difference of the two faces' references back.  The state's FP32 rounding then enters
// the small momenta through local pressure differences only: an O(p) rounding in the
// normal-momentum flux of a nearly static direction was the error that dominated.
template <class P>
constexpr bool SplitPressure = P::kId == PolicyMixed::kId;

template <typename T>
struct Prim {
  T d, vx, vy, vz, p;
};
// A reconstructed interface value as the stage-2 tiles exchange it.  Under the split
// pressure p holds the face pressure less the cell's reference ph (the high FP32 part
// of the cell's pressure) and ph that reference; otherwise it is a plain Prim.
template <typename T, bool Split>
struct IPrim : Prim<T> {
};
template <typename T>
struct IPrim<T, true> : Prim<T> {
  T ph;
};
template <class P>
using IPrimOf = IPrim<typename P::T, SplitPressure<P>>;
template <typename T>
struct Cons {
  T d, mx, my, mz, E;
};
template <typename T>
struct Char5 {
  T a0, a1, a2, a3, a4;
};

// ---------------------------------------------------------------------------- scalar helpers
template <class P>
__device__ __forceinline__ typename P::T Sqrt(typename P::T x)
{
  if constexpr (P::kId == PolicyMixed::kId) {
    return __builtin_amdgcn_sqrtf(x);
  } else {
    return __builtin_sqrt(x);
  }
}

// Reciprocal used only by the non-exact policies (dead code under PolicyExact).
template <class P>
__device__ __forceinline__ typename P::T Rcp(typename P::T x)
{
  if constexpr (P::kId == PolicyMixed::kId) {
    return __builtin_amdgcn_rcpf(x);
  } else {
    return 1.0 / x;
  }
}

// a / b, or a * inv_b where the caller precomputed inv_b = Rcp(b)
template <class P>
__device__ __forceinline__ typename P::T Div(typename P::T a, typename P::T b, typename P::T inv_b)
{
  if constexpr (P::kExact) {
    return a / b;
  } else {
    return a * inv_b;
  }
}

// a / b where no shared reciprocal exists
template <class P>
__device__ __forceinline__ typename P::T Div1(typename P::T a, typename P::T b)
{
  if constexpr (P::kExact) {
    return a / b;
  } else {
    return a * Rcp<P>(b);
  }
}

// fmax / fmin / fabs in the operands' own precision: the unsuffixed builtins are the
// double ones, which would widen FP32 operands (a conversion per operand and an FP64
// min/max).
template <typename T>
__device__ __forceinline__ T Fmax(T a, T b)
{
  if constexpr (sizeof(T) == 4) {
    return __builtin_fmaxf(a, b);
  } else {
    return __builtin_fmax(a, b);
  }
}
template <typename T>
__device__ __forceinline__ T Fmin(T a, T b)
{
  if constexpr (sizeof(T) == 4) {
    return __builtin_fminf(a, b);
  } else {
    return __builtin_fmin(a, b);
  }
}
template <typename T>
__device__ __forceinline__ T Fabs(T a)
{
  if constexpr (sizeof(T) == 4) {
    return __builtin_fabsf(a);
  } else {
    return __builtin_fabs(a);
  }
}
template <typename T>
__device__ __forceinline__ T SignOf(T a)  // SIGN macro, src/global/global.h:144
{
  return (a < T(0)) ? T(-1) : T(1);
}

// math_utils::SquareMagnitude, src/utils/math_utilities.h:78-97
template <typename T>
__device__ __forceinline__ T SquareMag(T a, T b, T c)
{
  return a * a + ((b * b) + (c * c));
}

// ---------------------------------------------------------------------------- EOS
// hydro_utilities::Calc_Pressure_Primitive, src/utils/hydro_utilities.h:34-45
template <typename T>
__device__ __forceinline__ T PressureFromCons(T E, T d, T vx, T vy, T vz, T gamma)
{
  T pressure = E - T(0.5) * d * SquareMag(vx, vy, vz);
  return Fmax((gamma - T(1.)) * pressure, T(kTinyNumber));
}

// hydro_utilities::Get_Pressure_From_DE, src/utils/hydro_utilities.h:136-151
template <typename T>
__device__ __forceinline__ T PressureFromDE(T E, T U_total, T U_advected, T gamma)
{
  T U = (U_total / E > T(kDeEta1)) ? U_total : U_advected;
  return Fmax(U * (gamma - T(1.0)), T(kTinyNumber));
}

// reconstruction::Load_Data, src/reconstruction/reconstruction.h:139-212: velocities
// by division, pressure by Calc_Pressure_Primitive.  Always evaluated in double:
// the mixed form recovers pressure in FP64 and narrows afterwards.
__device__ __forceinline__ Prim<double> LoadPrim(double d, double mx, double my, double mz, double E, double gamma)
{
  Prim<double> q;
  q.d  = d;
  q.vx = mx / d;
  q.vy = my / d;
  q.vz = mz / d;
  q.p  = PressureFromCons(E, d, q.vx, q.vy, q.vz, gamma);
  return q;
}

// hydro_utilities::Calc_Energy_Primitive, src/utils/hydro_utilities.h:106-118
template <class P>
__device__ __forceinline__ typename P::T EnergyFromPrim(const Prim<typename P::T> &s, typename P::T gamma,
                                                        typename P::T inv_gm1)
{
  using T = typename P::T;
  return Div<P>(Fmax(s.p, T(kTinyNumber)), gamma - T(1.), inv_gm1) + T(0.5) * s.d * SquareMag(s.vx, s.vy, s.vz);
}

// ---------------------------------------------------------------------------- characteristic maps (hydro)
template <class P>
struct SoundSpeed {
  using T = typename P::T;
  T c, c2;
  T inv_2c, inv_2c2, inv_c2, inv_d;
  __device__ __forceinline__ SoundSpeed(const Prim<T> &q, T gamma)
  {
    inv_d   = Rcp<P>(q.d);
    c       = Sqrt<P>(Div<P>(gamma * q.p, q.d, inv_d));  // Calc_Sound_Speed, hydro_utilities.h:209-212
    c2      = c * c;
    inv_2c  = Rcp<P>(T(2.0) * c);
    inv_2c2 = Rcp<P>(T(2.0) * c2);
    inv_c2  = Rcp<P>(c2);
  }
};

// Primitive_To_Characteristic (hydro branch), src/reconstruction/reconstruction.h:423-431
template <class P>
__device__ __forceinline__ Char5<typename P::T> PrimToChar(const Prim<typename P::T> &q, const Prim<typename P::T> &s,
                                                           const SoundSpeed<P> &cs)
{
  using T = typename P::T;
  Char5<T> a;
  a.a0 = Div<P>(-q.d * s.vx, T(2.0) * cs.c, cs.inv_2c) + Div<P>(s.p, T(2.0) * cs.c2, cs.inv_2c2);
  a.a1 = s.d - Div<P>(s.p, cs.c2, cs.inv_c2);
  a.a2 = s.vy;
  a.a3 = s.vz;
  a.a4 = Div<P>(q.d * s.vx, T(2.0) * cs.c, cs.inv_2c) + Div<P>(s.p, T(2.0) * cs.c2, cs.inv_2c2);
  return a;
}

// Characteristic_To_Primitive (hydro branch), reconstruction.h:482-488
template <class P>
__device__ __forceinline__ Prim<typename P::T> CharToPrim(const Prim<typename P::T> &q, const Char5<typename P::T> &a,
                                                          const SoundSpeed<P> &cs)
{
  using T = typename P::T;
  Prim<T> o;
  o.d  = a.a0 + a.a1 + a.a4;
  o.vx = Div<P>(cs.c, q.d, cs.inv_d) * (a.a4 - a.a0);
  o.vy = a.a2;
  o.vz = a.a3;
  o.p  = cs.c2 * (a.a0 + a.a4);
  return o;
}

// ---------------------------------------------------------------------------- PPM limiter
// reconstruction::PPM_Single_Variable, src/reconstruction/reconstruction.h:794-901.
// Every branch of the reference is evaluated and the result selected, so a wave
// never splits on the limiter's data-dependent tests; each candidate is computed
// with the reference's own expression, so the selected value is the reference's.
// SIGN(a) == SIGN(b) is (a < 0) == (b < 0), and SIGN(a) * m is (a < 0 ? -m : m).
template <class P>
__device__ __forceinline__ void PpmSingle(typename P::T im2, typename P::T im1, typename P::T i0, typename P::T ip1,
                                          typename P::T ip2, typename P::T &L_iph, typename P::T &R_imh)
{
  using T      = typename P::T;
  const T C2   = T(1.25);
  const T inv6 = T(1.0 / 6.0);
  auto neg     = [](T x) { return x < T(0); };
  auto limit_iface = [&](T c_i, T c_im1, T iface, T s2_i, T s2_im1) -> T {
    T s2c     = T(3.0) * (c_im1 + c_i - T(2.0) * iface);
    bool same = (neg(s2c) == neg(s2_im1)) && (neg(s2c) == neg(s2_i));
    T mag     = Fmin(C2 * Fabs(s2_im1), Fmin(C2 * Fabs(s2_i), Fabs(s2c)));
    T limited = same ? (neg(s2c) ? -mag : mag) : T(0.0);
    T dl      = iface - c_im1;
    T dr      = c_i - iface;
    T extr    = T(0.5) * (c_im1 + c_i) - Div<P>(limited, T(6.0), inv6);
    return (dl * dr < T(0.0)) ? extr : iface;
  };

  T slope_left    = (i0 - im1);
  T slope_right   = (ip1 - i0);
  T slope_avg_im1 = T(0.5) * slope_left + T(0.5) * (im1 - im2);
  T slope_avg_i   = T(0.5) * slope_right + T(0.5) * slope_left;
  T slope_avg_ip1 = T(0.5) * (ip2 - ip1) + T(0.5) * slope_right;

  R_imh = T(0.5) * (im1 + i0) + Div<P>(slope_avg_im1 - slope_avg_i, T(6.0), inv6);
  L_iph = T(0.5) * (i0 + ip1) + Div<P>(slope_avg_i - slope_avg_ip1, T(6.0), inv6);

  T s2_im1 = im2 + i0 - T(2.0) * im1;
  T s2_i   = im1 + ip1 - T(2.0) * i0;
  T s2_ip1 = i0 + ip2 - T(2.0) * ip1;

  R_imh = limit_iface(i0, im1, R_imh, s2_i, s2_im1);
  L_iph = limit_iface(ip1, i0, L_iph, s2_ip1, s2_i);

  T s2_face   = T(6.0) * (R_imh + L_iph - T(2.0) * i0);
  bool smooth = neg(s2_im1) == neg(s2_i) && neg(s2_im1) == neg(s2_ip1) && neg(s2_im1) == neg(s2_face);
  T mag2      = Fmin(Fmin(C2 * Fabs(s2_im1), C2 * Fabs(s2_i)), Fmin(C2 * Fabs(s2_ip1), Fabs(s2_face)));
  T s2_limited = smooth ? (neg(s2_face) ? -mag2 : mag2) : T(0.0);

  T cell_max = Fmax(Fabs(im2), Fabs(im1));
  cell_max   = Fmax(cell_max, Fabs(i0));
  cell_max   = Fmax(cell_max, Fabs(ip1));
  cell_max   = Fmax(cell_max, Fabs(ip2));

  // The 1e-12 round-off guards are the reference's; in FP32 they are below the
  // format's resolution and act as "exactly zero" tests, which keeps the same
  // branch structure.
  T ratio = Div1<P>(s2_limited, s2_face);
  T rho   = (Fabs(s2_face) > T(1.0e-12) * cell_max) ? ratio : T(0.0);

  T sfl         = i0 - R_imh;
  T sfr         = L_iph - i0;
  bool extremum = (sfl * sfr) <= T(0.0) || ((ip1 - i0) * (i0 - im1)) <= T(0.0);
  bool smooth_x = rho <= (T(1.0) - T(1.0e-12));
  T R_ext       = smooth_x ? i0 - rho * sfl : R_imh;
  T L_ext       = smooth_x ? i0 + rho * sfr : L_iph;
  T R_mono      = (Fabs(sfl) >= T(2.0) * Fabs(sfr)) ? i0 - T(2.0) * sfr : R_imh;
  T L_mono      = (Fabs(sfr) >= T(2.0) * Fabs(sfl)) ? i0 + T(2.0) * sfl : L_iph;
  R_imh         = extremum ? R_ext : R_mono;
  L_iph         = extremum ? L_ext : L_mono;
}

// The same limiter for the mixed form, algebraically reduced (valid in exact
// arithmetic, rounded differently): the averaged-slope interface values of
// reconstruction.h:830-839 are the fourth-order interpolants
//   R_{i-1/2} = 7/12 (a_{i-1} + a_i) - 1/12 (a_{i-2} + a_{i+1}),
//   L_{i+1/2} = 7/12 (a_i + a_{i+1}) - 1/12 (a_{i-1} + a_{i+2}),
// and "SIGN agrees for all of x, y, z" is min3 >= 0 or max3 < 0.  Data-dependent
// choices stay ternaries so a wave that agrees runs one side only.
template <typename T>
__device__ __forceinline__ void PpmSingleReduced(T im2, T im1, T i0, T ip1, T ip2, T &L_iph, T &R_imh)
{
  const T C2   = T(1.25);
  const T sixth = T(1.0 / 6.0);
  T sL = i0 - im1, sR = ip1 - i0;
  T a_m = im1 + i0, a_p = i0 + ip1;
  T Rv = T(7.0 / 12.0) * a_m - T(1.0 / 12.0) * (im2 + ip1);
  T Lv = T(7.0 / 12.0) * a_p - T(1.0 / 12.0) * (im1 + ip2);
  T s2m = im2 + i0 - T(2.0) * im1;
  T s2c = im1 + ip1 - T(2.0) * i0;
  T s2p = i0 + ip2 - T(2.0) * ip1;
  auto lim = [&](T asum, T c_left, T c_right, T iface, T s2a, T s2b) -> T {
    T s2f   = T(3.0) * asum - T(6.0) * iface;
    T mn    = Fmin(Fmin(s2f, s2a), s2b);
    T mx    = Fmax(Fmax(s2f, s2a), s2b);
    T mag   = Fmin(C2 * Fmin(Fabs(s2a), Fabs(s2b)), Fabs(s2f));
    T limit = mn >= T(0) ? mag : (mx < T(0) ? -mag : T(0));
    bool ext = (iface - c_left) * (c_right - iface) < T(0);
    return ext ? T(0.5) * asum - limit * sixth : iface;
  };
  T R = lim(a_m, im1, i0, Rv, s2c, s2m);
  T L = lim(a_p, i0, ip1, Lv, s2p, s2c);
  T s2face = T(6.0) * (R + L) - T(12.0) * i0;
  T mn4    = Fmin(Fmin(Fmin(s2m, s2c), s2p), s2face);
  T mx4    = Fmax(Fmax(Fmax(s2m, s2c), s2p), s2face);
  T mag2   = Fmin(C2 * Fmin(Fmin(Fabs(s2m), Fabs(s2c)), Fabs(s2p)), Fabs(s2face));
  T s2lim  = mn4 >= T(0) ? mag2 : (mx4 < T(0) ? -mag2 : T(0));
  T cmax   = Fmax(Fmax(Fmax(Fabs(im2), Fabs(im1)), Fmax(Fabs(i0), Fabs(ip1))), Fabs(ip2));
  T rho    = Fabs(s2face) > T(1.0e-12) * cmax ? s2lim * __builtin_amdgcn_rcpf(s2face) : T(0);
  T sfl = i0 - R, sfr = L - i0;
  if ((sfl * sfr) <= T(0) || (sR * sL) <= T(0)) {
    if (rho <= T(1.0) - T(1.0e-12)) {
      R = i0 - rho * sfl;
      L = i0 + rho * sfr;
    }
  } else {
    if (Fabs(sfl) >= T(2.0) * Fabs(sfr)) R = i0 - T(2.0) * sfr;
    if (Fabs(sfr) >= T(2.0) * Fabs(sfl)) L = i0 + T(2.0) * sfl;
  }
  R_imh = R;
  L_iph = L;
}

// PpmSingleReduced on deviations from the centre value (i0 = 0), as the mixed form's
// reconstruction hands it the stencil; the same expressions with i0 substituted.
template <typename T>
__device__ __forceinline__ void PpmSingleCentred(T im2, T im1, T ip1, T ip2, T &L_iph, T &R_imh)
{
  const T C2    = T(1.25);
  const T sixth = T(1.0 / 6.0);
  T Rv  = T(7.0 / 12.0) * im1 - T(1.0 / 12.0) * (im2 + ip1);
  T Lv  = T(7.0 / 12.0) * ip1 - T(1.0 / 12.0) * (im1 + ip2);
  T s2m = im2 - T(2.0) * im1;
  T s2c = im1 + ip1;
  T s2p = ip2 - T(2.0) * ip1;
  auto lim = [&](T asum, T iface, T s2a, T s2b, bool ext) -> T {
    T s2f   = T(3.0) * asum - T(6.0) * iface;
    T mn    = Fmin(Fmin(s2f, s2a), s2b);
    T mx    = Fmax(Fmax(s2f, s2a), s2b);
    T mag   = Fmin(C2 * Fmin(Fabs(s2a), Fabs(s2b)), Fabs(s2f));
    T limit = mn >= T(0) ? mag : (mx < T(0) ? -mag : T(0));
    return ext ? T(0.5) * asum - limit * sixth : iface;
  };
  // (iface - c_left) * (c_right - iface) < 0 with the centre at 0
  T R = lim(im1, Rv, s2c, s2m, (Rv - im1) * Rv > T(0));
  T L = lim(ip1, Lv, s2p, s2c, Lv * (ip1 - Lv) < T(0));
  T s2face = T(6.0) * (R + L);
  T mn4    = Fmin(Fmin(Fmin(s2m, s2c), s2p), s2face);
  T mx4    = Fmax(Fmax(Fmax(s2m, s2c), s2p), s2face);
  T mag2   = Fmin(C2 * Fmin(Fmin(Fabs(s2m), Fabs(s2c)), Fabs(s2p)), Fabs(s2face));
  T s2lim  = mn4 >= T(0) ? mag2 : (mx4 < T(0) ? -mag2 : T(0));
  T cmax   = Fmax(Fmax(Fabs(im2), Fabs(im1)), Fmax(Fabs(ip1), Fabs(ip2)));
  T rho    = Fabs(s2face) > T(1.0e-12) * cmax ? s2lim * __builtin_amdgcn_rcpf(s2face) : T(0);
  // sfl = -R, sfr = L
  if ((R * L) >= T(0) || (ip1 * im1) >= T(0)) {
    if (rho <= T(1.0) - T(1.0e-12)) {
      R = rho * R;
      L = rho * L;
    }
  } else {
    const T R0 = R, L0 = L, aR = Fabs(R0), aL = Fabs(L0);
    if (aR >= T(2.0) * aL) R = -T(2.0) * L0;
    if (aL >= T(2.0) * aR) L = -T(2.0) * R0;
  }
  R_imh = R;
  L_iph = L;
}

// ---------------------------------------------------------------------------- PPMC (VL)
// PPMC_VL body, src/reconstruction/ppmc_cuda.cu:603-684: project the 5-point stencil
// on cell i's characteristics, limit each characteristic field with PPM, project
// back and floor density/pressure.  s[0..4] = cells i-2..i+2 (rotated primitives).
//
// fallback: the positivity fallback, an option with no counterpart in the reference.
// When an interface density or pressure of the cell comes out non-positive, both of
// the cell's interfaces take its cell-centred values before the floors; the
// reference floors each quantity at TINY_NUMBER on its own, which can pair a 1e-20
// density with a finite pressure (an unbounded sound speed, and an HLLC flux that
// injects energy on the scale of p/1e-20).  Returns whether the fallback fired.
//
// ReconPpmcCore is the projection, limiting and back-projection alone, with the
// eigenvectors taken from q: the map is linear in the stencil values (and PPM shifts
// with them), so the mixed form hands it deviations from the centre cell with the
// centre's absolute state as q, and adds the centre back afterwards.
template <class P>
__device__ __forceinline__ void ReconPpmcCore(const Prim<typename P::T> s[5], const Prim<typename P::T> &q,
                                              typename P::T gamma, Prim<typename P::T> &L_iph,
                                              Prim<typename P::T> &R_imh)
{
  using T = typename P::T;
  SoundSpeed<P> cs(q, gamma);
  Char5<T> a[5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    if (SplitPressure<P> && m == 2) continue;  // deviations: the centre projects to zero
    a[m] = PrimToChar<P>(q, s[m], cs);
  }
  Char5<T> L, R;
  if constexpr (SplitPressure<P>) {
    PpmSingleCentred(a[0].a0, a[1].a0, a[3].a0, a[4].a0, L.a0, R.a0);
    PpmSingleCentred(a[0].a1, a[1].a1, a[3].a1, a[4].a1, L.a1, R.a1);
    PpmSingleCentred(a[0].a2, a[1].a2, a[3].a2, a[4].a2, L.a2, R.a2);
    PpmSingleCentred(a[0].a3, a[1].a3, a[3].a3, a[4].a3, L.a3, R.a3);
    PpmSingleCentred(a[0].a4, a[1].a4, a[3].a4, a[4].a4, L.a4, R.a4);
  } else if constexpr (P::kId == PolicyMixed::kId) {
    PpmSingleReduced(a[0].a0, a[1].a0, a[2].a0, a[3].a0, a[4].a0, L.a0, R.a0);
    PpmSingleReduced(a[0].a1, a[1].a1, a[2].a1, a[3].a1, a[4].a1, L.a1, R.a1);
    PpmSingleReduced(a[0].a2, a[1].a2, a[2].a2, a[3].a2, a[4].a2, L.a2, R.a2);
    PpmSingleReduced(a[0].a3, a[1].a3, a[2].a3, a[3].a3, a[4].a3, L.a3, R.a3);
    PpmSingleReduced(a[0].a4, a[1].a4, a[2].a4, a[3].a4, a[4].a4, L.a4, R.a4);
  } else {
    PpmSingle<P>(a[0].a0, a[1].a0, a[2].a0, a[3].a0, a[4].a0, L.a0, R.a0);
    PpmSingle<P>(a[0].a1, a[1].a1, a[2].a1, a[3].a1, a[4].a1, L.a1, R.a1);
    PpmSingle<P>(a[0].a2, a[1].a2, a[2].a2, a[3].a2, a[4].a2, L.a2, R.a2);
    PpmSingle<P>(a[0].a3, a[1].a3, a[2].a3, a[3].a3, a[4].a3, L.a3, R.a3);
    PpmSingle<P>(a[0].a4, a[1].a4, a[2].a4, a[3].a4, a[4].a4, L.a4, R.a4);
  }
  L_iph = CharToPrim<P>(q, L, cs);
  R_imh = CharToPrim<P>(q, R, cs);
}

template <class P>
__device__ __forceinline__ bool ReconPpmc(const Prim<typename P::T> s[5], typename P::T gamma,
                                          Prim<typename P::T> &L_iph, Prim<typename P::T> &R_imh,
                                          bool fallback = false)
{
  using T = typename P::T;
  const Prim<T> &q = s[2];
  ReconPpmcCore<P>(s, q, gamma, L_iph, R_imh);
  bool fb = false;
  if (fallback) {
    fb = !(L_iph.d > T(0) && R_imh.d > T(0) && L_iph.p > T(0) && R_imh.p > T(0));
    if (fb) L_iph = R_imh = q;
  }
  R_imh.d = Fmax(R_imh.d, T(kTinyNumber));
  L_iph.d = Fmax(L_iph.d, T(kTinyNumber));
  R_imh.p = Fmax(R_imh.p, T(kTinyNumber));
  L_iph.p = Fmax(L_iph.p, T(kTinyNumber));
  return fb;
}

// ---------------------------------------------------------------------------- PLMC (VL)
template <typename T>
__device__ __forceinline__ T VlSlope(T l, T r, T inv_sum)  // Van_Leer_Slope, reconstruction.h:266-272
{
  return (l * r > T(0.0)) ? T(2.0) * l * r / (l + r) : T(0.0);
}

// Monotonize lambda, reconstruction.h:519-527
template <typename T>
__device__ __forceinline__ T MonotonizeOne(T left, T right, T centered, T van_leer)
{
  if (left * right > T(0.0)) {
    T lim_a = T(2.0) * Fmin(Fabs(left), Fabs(right));
    T lim_b = Fmin(Fabs(centered), Fabs(van_leer));
    return __builtin_copysign(Fmin(lim_a, lim_b), centered);
  }
  return T(0.0);
}

// Interface_Values_PLM, src/reconstruction/plmp_cuda.cu:325-365: the Van Leer limited slope
// of one primitive variable; lo / hi are the values at the cell's left / right face.
template <class P>
__device__ __forceinline__ void PlmpInterface(typename P::T qm, typename P::T q0, typename P::T qp, typename P::T &lo,
                                              typename P::T &hi)
{
  using T = typename P::T;
  const T dL = q0 - qm, dR = qp - q0, dC = T(0.5) * (qp - qm);
  const T dG = dL * dR > T(0.0) ? Div1<P>(T(2.0) * dL * dR, dL + dR) : T(0.0);
  const T a = Fmin(Fabs(dL), Fabs(dR));
  const T b = Fmin(Fabs(dC), Fabs(dG));
  const T dm = SignOf(dC) * Fmin(T(2.0) * a, b);  // sgn_CUDA, global_cuda.h:41-48
  lo = q0 - T(0.5) * dm;
  hi = q0 + T(0.5) * dm;
}

// PLMP_cuda body in the VL build (plmp_cuda.cu:199-224; the half-step evolution at
// :231-284 is compiled out): each primitive limited on its own, then the density and
// pressure floors.  s[0..2] = cells i-1, i, i+1.
template <class P>
__device__ __forceinline__ void ReconPlmpCore(const Prim<typename P::T> s[3], Prim<typename P::T> &L_iph,
                                              Prim<typename P::T> &R_imh)
{
  PlmpInterface<P>(s[0].d, s[1].d, s[2].d, R_imh.d, L_iph.d);
  PlmpInterface<P>(s[0].vx, s[1].vx, s[2].vx, R_imh.vx, L_iph.vx);
  PlmpInterface<P>(s[0].vy, s[1].vy, s[2].vy, R_imh.vy, L_iph.vy);
  PlmpInterface<P>(s[0].vz, s[1].vz, s[2].vz, R_imh.vz, L_iph.vz);
  PlmpInterface<P>(s[0].p, s[1].p, s[2].p, R_imh.p, L_iph.p);
}

template <class P>
__device__ __forceinline__ void ReconPlmp(const Prim<typename P::T> s[3], Prim<typename P::T> &L_iph,
                                          Prim<typename P::T> &R_imh)
{
  using T = typename P::T;
  ReconPlmpCore<P>(s, L_iph, R_imh);
  R_imh.d = Fmax(R_imh.d, T(kTinyNumber));
  L_iph.d = Fmax(L_iph.d, T(kTinyNumber));
  R_imh.p = Fmax(R_imh.p, T(kTinyNumber));
  L_iph.p = Fmax(L_iph.p, T(kTinyNumber));
}

// PPMP_cuda body in the VL build (ppmp_cuda.cu:300-352, 640-650; STEEPENING and FLATTENING
// undefined): Calculate_Slope (the PLMP slope, :687-720) on cells i-1, i, i+1 of each
// primitive, Interface_Values_PPM (:722-755), then the density and pressure floors.
// s[0..4] = cells i-2 .. i+2.
template <class P>
__device__ __forceinline__ typename P::T LimitedSlope(typename P::T qm, typename P::T q0, typename P::T qp)
{
  using T    = typename P::T;
  const T dL = q0 - qm, dR = qp - q0, dC = T(0.5) * (qp - qm);
  const T dG = dL * dR > T(0.0) ? Div1<P>(T(2.0) * dL * dR, dL + dR) : T(0.0);
  return SignOf(dC) * Fmin(T(2.0) * Fmin(Fabs(dL), Fabs(dR)), Fmin(Fabs(dC), Fabs(dG)));
}

template <class P>
__device__ __forceinline__ void PpmpInterfaces(const typename P::T q[5], typename P::T &lo, typename P::T &hi)
{
  using T    = typename P::T;
  const T sm = LimitedSlope<P>(q[0], q[1], q[2]), s0 = LimitedSlope<P>(q[1], q[2], q[3]);
  const T sp = LimitedSlope<P>(q[2], q[3], q[4]);
  const T qm = q[1], qi = q[2], qp = q[3];
  T L = T(0.5) * (qi + qm) - (T(1.0) / T(6.0)) * (s0 - sm);
  T R = T(0.5) * (qp + qi) - (T(1.0) / T(6.0)) * (sp - s0);
  if ((R - qi) * (qi - L) <= T(0)) L = R = qi;
  if (T(6.0) * (R - L) * (qi - T(0.5) * (L + R)) > (R - L) * (R - L)) L = T(3.0) * qi - T(2.0) * R;
  if (T(6.0) * (R - L) * (qi - T(0.5) * (L + R)) < -(R - L) * (R - L)) R = T(3.0) * qi - T(2.0) * L;
  L  = Fmin(Fmax(qi, qm), Fmax(Fmin(qi, qm), L));
  R  = Fmin(Fmax(qi, qp), Fmax(Fmin(qi, qp), R));
  lo = L, hi = R;
}

template <class P>
__device__ __forceinline__ void ReconPpmpCore(const Prim<typename P::T> s[5], Prim<typename P::T> &L_iph,
                                              Prim<typename P::T> &R_imh)
{
  using T = typename P::T;
  T q[5];
#define CHL_PPMP_VAR(field)                                   \
  for (int m = 0; m < 5; m++) q[m] = s[m].field;              \
  PpmpInterfaces<P>(q, R_imh.field, L_iph.field);
  CHL_PPMP_VAR(d)
  CHL_PPMP_VAR(vx)
  CHL_PPMP_VAR(vy)
  CHL_PPMP_VAR(vz)
  CHL_PPMP_VAR(p)
#undef CHL_PPMP_VAR
}

template <class P>
__device__ __forceinline__ void ReconPpmp(const Prim<typename P::T> s[5], Prim<typename P::T> &L_iph,
                                          Prim<typename P::T> &R_imh)
{
  using T = typename P::T;
  ReconPpmpCore<P>(s, L_iph, R_imh);
  R_imh.d = Fmax(R_imh.d, T(kTinyNumber));
  L_iph.d = Fmax(L_iph.d, T(kTinyNumber));
  R_imh.p = Fmax(R_imh.p, T(kTinyNumber));
  L_iph.p = Fmax(L_iph.p, T(kTinyNumber));
}

// PLMC_cuda body (VL build: no characteristic tracing), src/reconstruction/plmc_cuda.cu:73-291.
// s[0..2] = cells i-1, i, i+1.  ReconPlmcCore leaves out the final floors and takes the
// eigenvectors from qe (the centre cell's absolute state; s may hold deviations from it,
// as for ReconPpmcCore: slopes, limiters and interface limits all shift with the values).
template <class P>
__device__ __forceinline__ void ReconPlmcCore(const Prim<typename P::T> s[3], const Prim<typename P::T> &qe,
                                              typename P::T gamma, Prim<typename P::T> &L_iph,
                                              Prim<typename P::T> &R_imh)
{
  using T = typename P::T;
  const Prim<T> &im = s[0], &q = s[1], &ip = s[2];
  SoundSpeed<P> cs(qe, gamma);
  // Compute_Slope (coef 1 and 0.5), reconstruction.h:224-251
  Prim<T> dL{T(1.0) * (q.d - im.d), T(1.0) * (q.vx - im.vx), T(1.0) * (q.vy - im.vy), T(1.0) * (q.vz - im.vz),
             T(1.0) * (q.p - im.p)};
  Prim<T> dR{T(1.0) * (ip.d - q.d), T(1.0) * (ip.vx - q.vx), T(1.0) * (ip.vy - q.vy), T(1.0) * (ip.vz - q.vz),
             T(1.0) * (ip.p - q.p)};
  Prim<T> dC{T(0.5) * (ip.d - im.d), T(0.5) * (ip.vx - im.vx), T(0.5) * (ip.vy - im.vy), T(0.5) * (ip.vz - im.vz),
             T(0.5) * (ip.p - im.p)};
  Prim<T> dG{VlSlope(dL.d, dR.d, T(0)), VlSlope(dL.vx, dR.vx, T(0)), VlSlope(dL.vy, dR.vy, T(0)),
             VlSlope(dL.vz, dR.vz, T(0)), VlSlope(dL.p, dR.p, T(0))};
  Char5<T> aL = PrimToChar<P>(qe, dL, cs), aR = PrimToChar<P>(qe, dR, cs), aG = PrimToChar<P>(qe, dG, cs);
  Char5<T> aC;
  if constexpr (P::kExact) {  // the reference projects the centred slope itself
    aC = PrimToChar<P>(qe, dC, cs);
  } else {  // dC = (dL + dR) / 2 and the projection is linear: one projection fewer
    aC = Char5<T>{T(0.5) * (aL.a0 + aR.a0), T(0.5) * (aL.a1 + aR.a1), T(0.5) * (aL.a2 + aR.a2),
                  T(0.5) * (aL.a3 + aR.a3), T(0.5) * (aL.a4 + aR.a4)};
  }
  Char5<T> m;
  m.a0       = MonotonizeOne(aL.a0, aR.a0, aC.a0, aG.a0);
  m.a1       = MonotonizeOne(aL.a1, aR.a1, aC.a1, aG.a1);
  m.a2       = MonotonizeOne(aL.a2, aR.a2, aC.a2, aG.a2);
  m.a3       = MonotonizeOne(aL.a3, aR.a3, aC.a3, aG.a3);
  m.a4       = MonotonizeOne(aL.a4, aR.a4, aC.a4, aG.a4);
  Prim<T> dm = CharToPrim<P>(qe, m, cs);
  // Calc_Interface_Linear, reconstruction.h:645-673
  const T hp = T(1.0) * T(0.5), hm = T(-1.0) * T(0.5);
  L_iph = {q.d + hp * dm.d, q.vx + hp * dm.vx, q.vy + hp * dm.vy, q.vz + hp * dm.vz, q.p + hp * dm.p};
  R_imh = {q.d + hm * dm.d, q.vx + hm * dm.vx, q.vy + hm * dm.vy, q.vz + hm * dm.vz, q.p + hm * dm.p};
  // Plm_Limit_Interfaces, reconstruction.h:686-723
  auto lim = [](T &l_iph, T &r_imh, T v_imo, T v_i, T v_ipo) {
    r_imh = Fmax(Fmin(v_i, v_imo), r_imh);
    r_imh = Fmin(Fmax(v_i, v_imo), r_imh);
    l_iph = Fmax(Fmin(v_i, v_ipo), l_iph);
    l_iph = Fmin(Fmax(v_i, v_ipo), l_iph);
  };
  lim(L_iph.d, R_imh.d, im.d, q.d, ip.d);
  lim(L_iph.vx, R_imh.vx, im.vx, q.vx, ip.vx);
  lim(L_iph.vy, R_imh.vy, im.vy, q.vy, ip.vy);
  lim(L_iph.vz, R_imh.vz, im.vz, q.vz, ip.vz);
  lim(L_iph.p, R_imh.p, im.p, q.p, ip.p);
}

template <class P>
__device__ __forceinline__ void ReconPlmc(const Prim<typename P::T> s[3], typename P::T gamma,
                                          Prim<typename P::T> &L_iph, Prim<typename P::T> &R_imh)
{
  using T = typename P::T;
  ReconPlmcCore<P>(s, s[1], gamma, L_iph, R_imh);
  R_imh.d = Fmax(R_imh.d, T(kTinyNumber));
  L_iph.d = Fmax(L_iph.d, T(kTinyNumber));
  R_imh.p = Fmax(R_imh.p, T(kTinyNumber));
  L_iph.p = Fmax(L_iph.p, T(kTinyNumber));
}

// A passive variable (the specific GasEnergy under dual energy) through the same
// reconstructions: PLMC limits its primitive slopes (Monotonize_Characteristic_Return_Primitive,
// reconstruction.h:549-551), interpolates (Calc_Interface_Linear :664) and limits the
// interfaces (Plm_Limit_Interfaces :714-715); PPMC applies PPM_Single_Variable to it
// (ppmc_cuda.cu:668-672).
template <class P>
__device__ __forceinline__ void ReconPlmcScalar(const typename P::T g[3], typename P::T &L_iph, typename P::T &R_imh)
{
  using T  = typename P::T;
  T dL     = T(1.0) * (g[1] - g[0]);
  T dR     = T(1.0) * (g[2] - g[1]);
  T dC     = T(0.5) * (g[2] - g[0]);
  T dG     = VlSlope(dL, dR, T(0));
  T dm     = MonotonizeOne(dL, dR, dC, dG);
  const T hp = T(1.0) * T(0.5), hm = T(-1.0) * T(0.5);
  L_iph    = g[1] + hp * dm;
  R_imh    = g[1] + hm * dm;
  R_imh    = Fmax(Fmin(g[1], g[0]), R_imh);
  R_imh    = Fmin(Fmax(g[1], g[0]), R_imh);
  L_iph    = Fmax(Fmin(g[1], g[2]), L_iph);
  L_iph    = Fmin(Fmax(g[1], g[2]), L_iph);
}

// PLMP and PPMP reconstruct the specific GasEnergy as they do each primitive
// (plmp_cuda.cu:204-206 with Interface_Values_PLM; ppmp_cuda.cu:353-361 with Calculate_Slope
// and Interface_Values_PPM).
template <class P>
__device__ __forceinline__ void ReconPlmpScalar(const typename P::T g[3], typename P::T &L_iph, typename P::T &R_imh)
{
  PlmpInterface<P>(g[0], g[1], g[2], R_imh, L_iph);
}

template <class P>
__device__ __forceinline__ void ReconPpmpScalar(const typename P::T g[5], typename P::T &L_iph, typename P::T &R_imh)
{
  PpmpInterfaces<P>(g, R_imh, L_iph);
}

template <class P>
__device__ __forceinline__ void ReconPpmcScalar(const typename P::T g[5], typename P::T &L_iph, typename P::T &R_imh)
{
  if constexpr (P::kId == PolicyMixed::kId) {
    PpmSingleReduced(g[0], g[1], g[2], g[3], g[4], L_iph, R_imh);
  } else {
    PpmSingle<P>(g[0], g[1], g[2], g[3], g[4], L_iph, R_imh);
  }
}

// ---------------------------------------------------------------------------- HLLC
// Riemann input for one side: conserved and primitive forms of the same state.
template <typename T>
struct RState {
  T d, mx, my, mz, E;  // conserved (rotated)
  T vx, vy, vz, p;     // primitive
  T inv_d;             // 1/d (non-exact policies only)
  T dge;               // conserved GasEnergy (dual energy only)
  T pd, p0;            // split pressure only: p less the face reference p0, and p0
};

// The pressure a solver puts in the normal-momentum flux and in pressure differences:
// the pressure itself, or under the split pressure its part above the face reference
// (the solvers then return that momentum flux less p0).
template <class P>
__device__ __forceinline__ typename P::T MomP(const RState<typename P::T> &s)
{
  if constexpr (SplitPressure<P>) {
    return s.pd;
  } else {
    return s.p;
  }
}

// Side state from a conserved interface value, as Calculate_HLLC_Fluxes_CUDA
// derives it (hllc_cuda.cu:104-131): velocities by division, pressure from the
// energy with the sum ((vx^2 + vy^2) + vz^2), floored at TINY_NUMBER.
// With dual energy the pressure is Get_Pressure_From_DE of the energy, the
// internal energy left after the kinetic part and the interface GasEnergy
// (hllc_cuda.cu:107-112, 125-130).
template <class P, bool DE = false>
__device__ __forceinline__ RState<typename P::T> RStateFromCons(typename P::T d, typename P::T mx, typename P::T my,
                                                                typename P::T mz, typename P::T E, typename P::T gamma,
                                                                typename P::T dge = typename P::T(0))
{
  using T = typename P::T;
  RState<T> s;
  s.d = d, s.mx = mx, s.my = my, s.mz = mz, s.E = E;
  s.inv_d = Rcp<P>(d);
  s.vx    = Div<P>(mx, d, s.inv_d);
  s.vy    = Div<P>(my, d, s.inv_d);
  s.vz    = Div<P>(mz, d, s.inv_d);
  if constexpr (DE) {
    T E_kin = T(0.5) * d * (s.vx * s.vx + s.vy * s.vy + s.vz * s.vz);
    s.p     = PressureFromDE(E, E - E_kin, dge, gamma);
    s.dge   = dge;
  } else {
    s.p = (E - T(0.5) * d * (s.vx * s.vx + s.vy * s.vy + s.vz * s.vz)) * (gamma - T(1.0));
  }
  s.p = Fmax(s.p, T(kTinyNumber));
  return s;
}

// Side state from a reconstructed primitive interface value.  PolicyExact goes
// through the reference's conserved round trip (Write_Data then the solver's own
// primitive recovery, reconstruction.h:918-958 + hllc_cuda.cu:104-131); the other
// policies use the primitive values directly.
// With dual energy, ge is the reconstructed specific GasEnergy of the interface
// (Write_Data stores density * gas_energy, reconstruction.h:960-962).  The non-exact
// policies apply the DE_ETA_1 switch to the primitive state: the internal energy
// p/(gamma-1) against the total energy, else the pressure of the GasEnergy.
template <class P, bool DE = false>
__device__ __forceinline__ RState<typename P::T> RStateFromPrim(const Prim<typename P::T> &s, typename P::T gamma,
                                                                typename P::T inv_gm1,
                                                                typename P::T ge = typename P::T(0))
{
  using T = typename P::T;
  T d = s.d, mx = s.d * s.vx, my = s.d * s.vy, mz = s.d * s.vz;
  T E = EnergyFromPrim<P>(s, gamma, inv_gm1);
  if constexpr (P::kExact) {
    return RStateFromCons<P, DE>(d, mx, my, mz, E, gamma, DE ? s.d * ge : T(0));
  } else {
    RState<T> r;
    r.d = d, r.mx = mx, r.my = my, r.mz = mz, r.E = E;
    r.vx = s.vx, r.vy = s.vy, r.vz = s.vz, r.p = s.p;
    r.inv_d = Rcp<P>(d);
    if constexpr (DE) {
      r.dge  = s.d * ge;
      T th   = Div<P>(Fmax(s.p, T(kTinyNumber)), gamma - T(1.), inv_gm1);
      bool ok = th > T(kDeEta1) * E;  // (E - E_kin) / E > eta_1 with E > 0
      r.p    = ok ? s.p : Fmax(r.dge * (gamma - T(1.0)), T(kTinyNumber));
    }
    return r;
  }
}

// Calculate_HLLC_Fluxes_CUDA body, src/riemann_solvers/hllc_cuda.cu:141-305 (etah = 0).
// Returns the flux in rotated order; with dual energy the GasEnergy flux goes to
// *fge (hllc_cuda.cu:183-198, 218-235, 251-303).
template <class P, bool DE = false>
__device__ __forceinline__ Cons<typename P::T> Hllc(const RState<typename P::T> &l, const RState<typename P::T> &r,
                                                    typename P::T gamma, typename P::T *fge = nullptr)
{
  using T    = typename P::T;
  const T g1 = gamma - T(1.0);
  T Hl       = Div<P>(l.E + l.p, l.d, l.inv_d);
  T Hr       = Div<P>(r.E + r.p, r.d, r.inv_d);

  T sqrtdl = Sqrt<P>(l.d), sqrtdr = Sqrt<P>(r.d);
  T den     = sqrtdl + sqrtdr;
  T inv_den = Rcp<P>(den);
  T vx      = Div<P>(sqrtdl * l.vx + sqrtdr * r.vx, den, inv_den);
  T vy      = Div<P>(sqrtdl * l.vy + sqrtdr * r.vy, den, inv_den);
  T vz      = Div<P>(sqrtdl * l.vz + sqrtdr * r.vz, den, inv_den);
  T H       = Div<P>(sqrtdl * Hl + sqrtdr * Hr, den, inv_den);
  T vsq     = (vx * vx + vy * vy + vz * vz);
  T asq     = g1 * (H - T(0.5) * vsq);
  // H - vsq/2 >= 0 holds for Roe averages in exact arithmetic; where both sides carry a
  // (floored) near-zero pressure it is a difference of two nearly equal kinetic terms,
  // and its rounding can go negative.  The non-exact forms take the root of max(asq, 0).
  if constexpr (!P::kExact) asq = Fmax(asq, T(0));
  T a       = Sqrt<P>(asq);
  T lambda_m = vx - a, lambda_p = vx + a;
  T cfl = Sqrt<P>(Div<P>(gamma * l.p, l.d, l.inv_d));
  T cfr = Sqrt<P>(Div<P>(gamma * r.p, r.d, r.inv_d));
  if constexpr (P::kId == PolicyMixed::kId) {
    // FP32: a sound speed below the resolution of the velocity (a floored interface
    // pressure) would put Sl on vxl or Sr on vxr, leave the star-state denominators at
    // zero and turn the flux into 0/0; keep the signal speeds at 2^-20 |v| at least.
    cfl = Fmax(cfl, T(0x1p-20) * Fabs(l.vx));
    cfr = Fmax(cfr, T(0x1p-20) * Fabs(r.vx));
  }
  T Sl  = Fmin(lambda_m, l.vx - cfl);
  T Sr  = Fmax(lambda_p, r.vx + cfr);
  // H-correction hook with etah = 0 (hllc_cuda.cu:52,174-175)
  Sl = ((Sl < T(0)) ? T(-1) : T(1)) * Fmax(Fabs(Sl), T(0));
  Sr = ((Sr < T(0)) ? T(-1) : T(1)) * Fmax(Fabs(Sr), T(0));

  // momentum-flux pressures (the pressures themselves outside the split-pressure form)
  const T pml = MomP<P>(l), pmr = MomP<P>(r);
  Cons<T> fl{l.mx, l.mx * l.vx + pml, l.my * l.vx, l.mz * l.vx, (l.E + l.p) * l.vx};
  Cons<T> fr{r.mx, r.mx * r.vx + pmr, r.my * r.vx, r.mz * r.vx, (r.E + r.p) * r.vx};
  T f_ge_l = T(0), f_ge_r = T(0);
  if constexpr (DE) f_ge_l = l.dge * l.vx, f_ge_r = r.dge * r.vx;
  if (Sl > T(0.0)) {
    if constexpr (DE) *fge = f_ge_l;
    return fl;
  }
  if (Sr < T(0.0)) {
    if constexpr (DE) *fge = f_ge_r;
    return fr;
  }

  T Sm = Div1<P>(r.d * r.vx * (Sr - r.vx) - l.d * l.vx * (Sl - l.vx) + pml - pmr,
                 r.d * (Sr - r.vx) - l.d * (Sl - l.vx));
  T psm     = l.d * (l.vx - Sl) * (l.vx - Sm) + pml;  // star pressure (less p0 when split)
  T ps      = SplitPressure<P> ? psm + l.p0 : psm;
  T inv_lsm = Rcp<P>(Sl - Sm), inv_rsm = Rcp<P>(Sr - Sm);
  T dls     = Div<P>(l.d * (Sl - l.vx), Sl - Sm, inv_lsm);
  T mxls    = Div<P>(l.mx * (Sl - l.vx) + psm - pml, Sl - Sm, inv_lsm);
  T myls = dls * l.vy, mzls = dls * l.vz;
  T Els  = Div<P>(l.E * (Sl - l.vx) - l.p * l.vx + ps * Sm, Sl - Sm, inv_lsm);
  T drs  = Div<P>(r.d * (Sr - r.vx), Sr - Sm, inv_rsm);
  T mxrs = Div<P>(r.mx * (Sr - r.vx) + psm - pmr, Sr - Sm, inv_rsm);
  T myrs = drs * r.vy, mzrs = drs * r.vz;
  T Ers  = Div<P>(r.E * (Sr - r.vx) - r.p * r.vx + ps * Sm, Sr - Sm, inv_rsm);
  T aSm  = Fabs(Sm);
  Cons<T> f;
  f.d  = T(0.5) * (fl.d + fr.d + (Sr - aSm) * drs + (Sl + aSm) * dls - Sl * l.d - Sr * r.d);
  f.mx = T(0.5) * (fl.mx + fr.mx + (Sr - aSm) * mxrs + (Sl + aSm) * mxls - Sl * l.mx - Sr * r.mx);
  f.my = T(0.5) * (fl.my + fr.my + (Sr - aSm) * myrs + (Sl + aSm) * myls - Sl * l.my - Sr * r.my);
  f.mz = T(0.5) * (fl.mz + fr.mz + (Sr - aSm) * mzrs + (Sl + aSm) * mzls - Sl * l.mz - Sr * r.mz);
  f.E  = T(0.5) * (fl.E + fr.E + (Sr - aSm) * Ers + (Sl + aSm) * Els - Sl * l.E - Sr * r.E);
  if constexpr (DE) {
    T gel = Div<P>(l.dge, l.d, l.inv_d), ger = Div<P>(r.dge, r.d, r.inv_d);
    T gels = dls * gel, gers = drs * ger;
    *fge   = T(0.5) * (f_ge_l + f_ge_r + (Sr - aSm) * gers + (Sl + aSm) * gels - Sl * l.dge - Sr * r.dge);
  }
  return f;
}

// Calculate_HLL_Fluxes_CUDA body, src/riemann_solvers/hll_cuda.cu:141-266: Davis wave
// speeds from the two sides' sound speeds (no Roe average), the upwind physical flux for a
// supersonic face, else the HLL average.  With dual energy the GasEnergy flux goes to *fge.
template <class P, bool DE = false>
__device__ __forceinline__ Cons<typename P::T> Hll(const RState<typename P::T> &l, const RState<typename P::T> &r,
                                                   typename P::T gamma, typename P::T *fge = nullptr)
{
  using T = typename P::T;
  T cfl   = Sqrt<P>(Div<P>(gamma * l.p, l.d, l.inv_d));
  T cfr   = Sqrt<P>(Div<P>(gamma * r.p, r.d, r.inv_d));
  if constexpr (P::kId == PolicyMixed::kId) {
    // as in Hllc: keep the speeds apart where FP32 cannot resolve a floored sound speed
    cfl = Fmax(cfl, T(0x1p-20) * Fabs(l.vx));
    cfr = Fmax(cfr, T(0x1p-20) * Fabs(r.vx));
  }
  const T Sl = Fmin(r.vx - cfr, l.vx - cfl);
  const T Sr = Fmax(l.vx + cfl, r.vx + cfr);
  Cons<T> fl{l.mx, l.mx * l.vx + MomP<P>(l), l.my * l.vx, l.mz * l.vx, (l.E + l.p) * l.vx};
  Cons<T> fr{r.mx, r.mx * r.vx + MomP<P>(r), r.my * r.vx, r.mz * r.vx, (r.E + r.p) * r.vx};
  T f_ge_l = T(0), f_ge_r = T(0);
  if constexpr (DE) f_ge_l = l.dge * l.vx, f_ge_r = r.dge * r.vx;
  if (Sl > T(0.0)) {
    if constexpr (DE) *fge = f_ge_l;
    return fl;
  }
  if (Sr < T(0.0)) {
    if constexpr (DE) *fge = f_ge_r;
    return fr;
  }
  const T den = Sr - Sl, inv = Rcp<P>(den);
  auto avg    = [&](T a_l, T a_r, T u_l, T u_r) { return Div<P>((Sr * a_l) - (Sl * a_r) + Sl * Sr * (u_r - u_l), den, inv); };
  Cons<T> f;
  f.d  = avg(fl.d, fr.d, l.d, r.d);
  f.mx = avg(fl.mx, fr.mx, l.mx, r.mx);
  f.my = avg(fl.my, fr.my, l.my, r.my);
  f.mz = avg(fl.mz, fr.mz, l.mz, r.mz);
  f.E  = avg(fl.E, fr.E, l.E, r.E);
  if constexpr (DE) *fge = avg(f_ge_l, f_ge_r, l.dge, r.dge);
  return f;
}

// Calculate_Roe_Fluxes_CUDA body, src/riemann_solvers/roe_cuda.cu:160-379 (hydro; its etah
// is 0, so each eigenvalue enters as |lambda|): upwind where the Roe-averaged eigenvalues are
// one-signed, else the Roe flux, or HLLE with Einfeldt speeds where an intermediate state of
// the linearised solution has non-positive density or negative pressure.  The second of
// those tests adds the fourth wave's coefficient to the density only (roe_cuda.cu:330-334);
// it is kept so, as the switch decides which flux comes out.  Expressions keep the
// reference's association so PolicyExact reproduces its bits.
//
// With dual energy the GasEnergy flux goes to *fge (roe_cuda.cu:168-225, 344-397): the
// upwind side's mass flux times its specific GasEnergy, the HLLE average of
// dge (v - b), or the Roe mass flux times the upwind specific GasEnergy.
template <class P, bool DE = false>
__device__ __forceinline__ Cons<typename P::T> Roe(const RState<typename P::T> &l, const RState<typename P::T> &r,
                                                   typename P::T gamma, typename P::T *fge = nullptr)
{
  using T    = typename P::T;
  const T g1 = gamma - T(1.0);
  const T Hl = Div1<P>(l.E + l.p, l.d), Hr = Div1<P>(r.E + r.p, r.d);
  const T sl = Sqrt<P>(l.d), sr = Sqrt<P>(r.d);
  const T inv_s = Rcp<P>(sl + sr);
  const T vx  = Div<P>(sl * l.vx + sr * r.vx, sl + sr, inv_s);
  const T vy  = Div<P>(sl * l.vy + sr * r.vy, sl + sr, inv_s);
  const T vz  = Div<P>(sl * l.vz + sr * r.vz, sl + sr, inv_s);
  const T H   = Div<P>(sl * Hl + sr * Hr, sl + sr, inv_s);
  const T vsq = (vx * vx + vy * vy + vz * vz);
  const T asq = g1 * Fmax((H - T(0.5) * vsq), T(kTinyNumber));
  const T a   = Sqrt<P>(asq);
  const T lm = vx - a, l0 = vx, lp = vx + a;
  const T pml = MomP<P>(l), pmr = MomP<P>(r);
  Cons<T> fl{l.mx, l.mx * l.vx + pml, l.mx * l.vy, l.mx * l.vz, (l.E + l.p) * l.vx};
  Cons<T> fr{r.mx, r.mx * r.vx + pmr, r.mx * r.vy, r.mx * r.vz, (r.E + r.p) * r.vx};
  T gel = T(0), ger = T(0);
  if constexpr (DE) gel = Div1<P>(l.dge, l.d), ger = Div1<P>(r.dge, r.d);
  if (lm >= T(0.0)) {
    if constexpr (DE) *fge = l.mx * gel;
    return fl;
  }
  if (lp <= T(0.0)) {
    if constexpr (DE) *fge = r.mx * ger;
    return fr;
  }
  const T dd = r.d - l.d, dmx = r.mx - l.mx, dmy = r.my - l.my, dmz = r.mz - l.mz;
  T dE = r.E - l.E;
  if constexpr (SplitPressure<P>) {
    // the energy jump from the pressure jump relative to the face reference, so the
    // O(p) rounding of the two energies does not enter the wave strengths
    dE = Div1<P>(pmr - pml, g1) + (T(0.5) * r.d * SquareMag(r.vx, r.vy, r.vz) -
                                   T(0.5) * l.d * SquareMag(l.vx, l.vy, l.vz));
  }
  const T Na = Div1<P>(T(0.5), asq);
  const T inv_asq = Rcp<P>(asq);
  T s0 = T(0.0), s1 = T(0.0), s2 = T(0.0), s3 = T(0.0), s4 = T(0.0);
  const T a0 = dd * Na * (T(0.5) * g1 * vsq + vx * a) - dmx * Na * (g1 * vx + a) - dmy * Na * g1 * vy -
               dmz * Na * g1 * vz + dE * Na * g1;
  T c = a0 * Fabs(lm);
  s0 += c, s1 += c * (vx - a), s2 += c * vy, s3 += c * vz, s4 += c * (H - vx * a);
  const T a1 = -dd * vy + dmy;
  c          = a1 * Fabs(l0);
  s2 += c, s4 += c * vy;
  const T a2 = -dd * vz + dmz;
  c          = a2 * Fabs(l0);
  s3 += c, s4 += c * vz;
  const T a3 = dd * (T(1.0) - Na * g1 * vsq) + Div<P>(dmx * g1 * vx, asq, inv_asq) + Div<P>(dmy * g1 * vy, asq, inv_asq) +
               Div<P>(dmz * g1 * vz, asq, inv_asq) - Div<P>(dE * g1, asq, inv_asq);
  c = a3 * Fabs(l0);
  s0 += c, s1 += c * vx, s2 += c * vy, s3 += c * vz, s4 += c * T(0.5) * vsq;
  const T a4 = dd * Na * (T(0.5) * g1 * vsq - vx * a) - dmx * Na * (g1 * vx - a) - dmy * Na * g1 * vy -
               dmz * Na * g1 * vz + dE * Na * g1;
  c = a4 * Fabs(lp);
  s0 += c, s1 += c * (vx + a), s2 += c * vy, s3 += c * vz, s4 += c * (H + vx * a);
  bool hlle = false;
  T t0 = l.d + a0, t1 = l.mx + a0 * (vx - a), t2 = l.my + a0 * vy, t3 = l.mz + a0 * vz, t4 = l.E + a0 * (H - vx * a);
  if (l0 > lm) hlle = hlle || t0 <= T(0.0) || t4 - Div1<P>(T(0.5) * (t1 * t1 + t2 * t2 + t3 * t3), t0) < T(0.0);
  t0 += a3 + a4, t1 += a3 * vx, t2 += a1 + a3 * vy, t3 += a2 + a3 * vz, t4 += a1 * vy + a2 * vz + a3 * T(0.5) * vsq;
  if (lp > l0) hlle = hlle || t0 <= T(0.0) || t4 - Div1<P>(T(0.5) * (t1 * t1 + t2 * t2 + t3 * t3), t0) < T(0.0);
  Cons<T> f;
  if (hlle) {
    const T cfl = Sqrt<P>(Div1<P>(gamma * l.p, l.d)), cfr = Sqrt<P>(Div1<P>(gamma * r.p, r.d));
    const T bm = Fmin(Fmin(lm, l.vx - cfl), T(0.0)), bp = Fmax(Fmax(lp, r.vx + cfr), T(0.0));
    const T tmp = Div1<P>(T(0.5) * (bp + bm), bp - bm);
    auto hlle_avg = [&](T hl, T hr) { return T(0.5) * (hl + hr) + (hl - hr) * tmp; };
    f.d  = hlle_avg(l.mx - bm * l.d, r.mx - bp * r.d);
    f.mx = hlle_avg(l.mx * (l.vx - bm) + pml, r.mx * (r.vx - bp) + pmr);
    f.my = hlle_avg(l.my * (l.vx - bm), r.my * (r.vx - bp));
    f.mz = hlle_avg(l.mz * (l.vx - bm), r.mz * (r.vx - bp));
    f.E  = hlle_avg(l.E * (l.vx - bm) + l.p * l.vx, r.E * (r.vx - bp) + r.p * r.vx);
    if constexpr (DE) *fge = hlle_avg(l.dge * (l.vx - bm), r.dge * (r.vx - bp));
  } else {
    f.d  = T(0.5) * (fl.d + fr.d - s0);
    f.mx = T(0.5) * (fl.mx + fr.mx - s1);
    f.my = T(0.5) * (fl.my + fr.my - s2);
    f.mz = T(0.5) * (fl.mz + fr.mz - s3);
    f.E  = T(0.5) * (fl.E + fr.E - s4);
    if constexpr (DE) *fge = f.d >= T(0.0) ? f.d * gel : f.d * ger;
  }
  return f;
}

// ---------------------------------------------------------------------------- exact solver
// powf as exact_cuda.cu calls it on doubles: both arguments narrowed to float and an FP32
// result, evaluated as the correctly rounded float of the double pow of the narrowed
// arguments (any conforming powf differs from it by float rounding only).
__device__ __forceinline__ double PowF(double x, double y)
{
  return double(float(::pow(double(float(x)), double(float(y)))));
}

// guessp_CUDA, prefun_CUDA, starpv_CUDA and sample_CUDA, exact_cuda.cu:137-334: the star
// state by Newton iteration from the PVRS / two-shock guess (at most 21 updates, relative
// change <= 1e-6), then the solution sampled on the interface.  Always FP64.
struct ExactSideState {
  double d, v, p, c;
};

__device__ __forceinline__ void ExactPressureFn(double &f, double &fd, double p, const ExactSideState &k, double gamma)
{
  if (p <= k.p) {
    f  = (2.0 / (gamma - 1.0)) * k.c * (PowF(p / k.p, (gamma - 1.0) / (2.0 * gamma)) - 1.0);
    fd = (1.0 / (k.d * k.c)) * PowF((p / k.p), -((gamma + 1.0) / (2.0 * gamma)));
  } else {
    const double qrt = sqrt(((2.0 / (gamma + 1.0)) / k.d) / ((((gamma - 1.0) / (gamma + 1.0)) * k.p) + p));
    f                = (p - k.p) * qrt;
    fd               = (1.0 - 0.5 * (p - k.p) / ((((gamma - 1.0) / (gamma + 1.0)) * k.p) + p)) * qrt;
  }
}

__device__ __forceinline__ void ExactStar(double &pm, double &vm, const ExactSideState &L, const ExactSideState &R,
                                          double gamma)
{
  const double TOL = 1.0e-6;
  double ppv       = 0.5 * (L.p + R.p) + 0.125 * (L.v - R.v) * (L.d + R.d) * (L.c + R.c);
  if (ppv < 0.0) ppv = 0.0;
  const double gl = sqrt((2.0 / ((gamma + 1.0) * L.d)) / (((gamma - 1.0) / (gamma + 1.0)) * L.p + ppv));
  const double gr = sqrt((2.0 / ((gamma + 1.0) * R.d)) / (((gamma - 1.0) / (gamma + 1.0)) * R.p + ppv));
  double pold     = (gl * L.p + gr * R.p - (R.v - L.v)) / (gl + gr);
  if (pold < 0.0) pold = TOL;
  double fl = 0.0, fld = 0.0, fr = 0.0, frd = 0.0;
  for (int i = 0; i <= 20; i++) {
    ExactPressureFn(fl, fld, pold, L, gamma);
    ExactPressureFn(fr, frd, pold, R, gamma);
    pm                  = pold - (fl + fr + R.v - L.v) / (fld + frd);
    const double change = 2.0 * Fabs((pm - pold) / (pm + pold));
    if (change <= TOL) break;
    if (pm < 0.0) pm = TOL;
    pold = pm;
  }
  vm = 0.5 * (L.v + R.v + fr - fl);
}

__device__ __forceinline__ void ExactSample(double pm, double vm, const ExactSideState &L, const ExactSideState &R,
                                            double gamma, double &d, double &v, double &p)
{
  const double gm1 = gamma - 1.0, gp1 = gamma + 1.0;
  if (vm >= 0) {
    if (pm <= L.p) {  // left rarefaction
      if (L.v - L.c >= 0) {
        d = L.d, v = L.v, p = L.p;
      } else if (vm - L.c * PowF(pm / L.p, gm1 / (2.0 * gamma)) < 0) {
        d = L.d * PowF(pm / L.p, 1.0 / gamma), v = vm, p = pm;
      } else {
        const double c = (2.0 / gp1) * (L.c + (gm1 / 2.0) * L.v);
        v = c, d = L.d * PowF(c / L.c, 2.0 / gm1), p = L.p * PowF(c / L.c, 2.0 * gamma / gm1);
      }
    } else {  // left shock
      const double sl = L.v - L.c * sqrt((gp1 / (2.0 * gamma)) * (pm / L.p) + (gm1 / (2.0 * gamma)));
      if (sl >= 0) {
        d = L.d, v = L.v, p = L.p;
      } else {
        d = L.d * (pm / L.p + (gm1 / gp1)) / ((pm / L.p) * (gm1 / gp1) + 1.0), v = vm, p = pm;
      }
    }
  } else {
    if (pm > R.p) {  // right shock
      const double sr = R.v + R.c * sqrt((gp1 / (2.0 * gamma)) * (pm / R.p) + (gm1 / (2.0 * gamma)));
      if (sr <= 0) {
        d = R.d, v = R.v, p = R.p;
      } else {
        d = R.d * (pm / R.p + (gm1 / gp1)) / ((pm / R.p) * (gm1 / gp1) + 1.0), v = vm, p = pm;
      }
    } else {  // right rarefaction
      if (R.v + R.c <= 0) {
        d = R.d, v = R.v, p = R.p;
      } else if (vm + R.c * PowF(pm / R.p, gm1 / (2.0 * gamma)) >= 0) {
        d = R.d * PowF(pm / R.p, (1.0 / gamma)), v = vm, p = pm;
      } else {
        const double c = (2.0 / gp1) * (R.c - (gm1 / 2.0) * R.v);
        v = -c, d = R.d * PowF(c / R.c, 2.0 / gm1), p = R.p * PowF(c / R.c, 2.0 * gamma / gm1);
      }
    }
  }
}

// Calculate_Exact_Fluxes_CUDA body, exact_cuda.cu:82-135: the Godunov flux of the sampled
// state, transverse velocities upwinded by the sampled velocity.  The reference's vacuum
// warning (a printf, :95-98) is left out.  With dual energy the GasEnergy flux, the sampled
// mass flux times the upwind side's specific GasEnergy, goes to *fge (exact_cuda.cu:122-133).
template <class P, bool DE = false>
__device__ __forceinline__ Cons<typename P::T> ExactFlux(const RState<typename P::T> &l,
                                                         const RState<typename P::T> &r, typename P::T gamma_t,
                                                         typename P::T *fge = nullptr)
{
  using T            = typename P::T;
  const double gamma = double(gamma_t);
  ExactSideState L{double(l.d), double(l.vx), double(l.p), 0.0}, R{double(r.d), double(r.vx), double(r.p), 0.0};
  double p0 = 0.0;  // the momentum flux is returned less p0 (split pressure only)
  if constexpr (SplitPressure<P>) {
    p0  = double(l.p0);
    L.p = Fmax(p0 + double(l.pd), kTinyNumber);
    R.p = Fmax(p0 + double(r.pd), kTinyNumber);
  }
  L.c = sqrt(gamma * L.p / L.d);
  R.c = sqrt(gamma * R.p / R.d);
  double pm = 0.0, vm = 0.0, ds, vs, ps;
  ExactStar(pm, vm, L, R, gamma);
  ExactSample(pm, vm, L, R, gamma, ds, vs, ps);
  const bool left = vs >= 0;
  const double vy = left ? double(l.vy) : double(r.vy), vz = left ? double(l.vz) : double(r.vz);
  const double Es = (ps / (gamma - 1.0)) + 0.5 * ds * (vs * vs + vy * vy + vz * vz);
  const double fm = SplitPressure<P> ? ds * vs * vs + (ps - p0) : ds * vs * vs + ps;
  if constexpr (DE) {
    const double ge = left ? double(l.dge) / double(l.d) : double(r.dge) / double(r.d);
    *fge            = T(ds * vs * ge);
  }
  return Cons<T>{T(ds * vs), T(fm), T(ds * vs * vy), T(ds * vs * vz), T((Es + ps) * vs)};
}

// The solver of the form P (HLLC unless P is a SolverPolicy).
template <class P, bool DE = false>
__device__ __forceinline__ Cons<typename P::T> RiemannFlux(const RState<typename P::T> &l,
                                                           const RState<typename P::T> &r, typename P::T gamma,
                                                           typename P::T *fge = nullptr)
{
  if constexpr (P::kSolver == kRiemannHll) {
    return Hll<P, DE>(l, r, gamma, fge);
  } else if constexpr (P::kSolver == kRiemannRoe) {
    return Roe<P, DE>(l, r, gamma, fge);
  } else if constexpr (P::kSolver == kRiemannExact) {
    return ExactFlux<P, DE>(l, r, gamma, fge);
  } else {
    return Hllc<P, DE>(l, r, gamma, fge);
  }
}

// ---------------------------------------------------------------------------- dt
// hydroInverseCrossingTime, src/hydro/hydro_cuda.cu:404-418, FP64 in every form; the
// non-exact policies multiply by precomputed 1/dx, 1/dy, 1/dz instead of dividing.
template <class P>
__device__ __forceinline__ double InverseCrossingTimeP(double E, double d, double d_inv, double vx, double vy,
                                                       double vz, double dx, double dy, double dz, double idx,
                                                       double idy, double idz, double gamma)
{
  double P_ = (E - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0);
  double cs = __builtin_sqrt(d_inv * gamma * P_);
  double ax, ay, az;
  if constexpr (P::kExact) {
    ax = (Fabs(vx) + cs) / dx, ay = (Fabs(vy) + cs) / dy, az = (Fabs(vz) + cs) / dz;
  } else {
    ax = (Fabs(vx) + cs) * idx, ay = (Fabs(vy) + cs) * idy, az = (Fabs(vz) + cs) * idz;
  }
  double m = Fmax(ax, ay);
  m        = Fmax(m, az);
  m        = Fmax(m, 0.0);
  return m;
}

__device__ __forceinline__ double InverseCrossingTime(double E, double d, double d_inv, double vx, double vy, double vz,
                                                      double dx, double dy, double dz, double gamma)
{
  double P  = (E - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0);
  double cs = __builtin_sqrt(d_inv * gamma * P);
  double m  = Fmax((Fabs(vx) + cs) / dx, (Fabs(vy) + cs) / dy);
  m         = Fmax(m, (Fabs(vz) + cs) / dz);
  m         = Fmax(m, 0.0);
  return m;
}

}  // namespace chl
