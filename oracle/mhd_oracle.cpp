// mhd_oracle.cpp -- CPU restatement of Cholla's MHD hot path (FP64): PCM and
// PLMC reconstruction with the MHD characteristic projection, the HLLD Riemann
// solver, constrained-transport edge electric fields, the face-centred magnetic
// field update, the MHD CFL reduction, the magnetic divergence, and the VL step
// and driver loop that connect them.
//
// TEST INFRASTRUCTURE ONLY, under the same rules as cholla_oracle.cpp: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
// and only as the checker; the product path never links, loads or calls it.
//
// Every function restates the reference function it names (file:line under
// src/ of cholla-hydro/cholla) in the same operation order and association,
// compiled with -ffp-contract=off.  Two device intrinsics have no host
// counterpart and are restated from their definitions: rhypot(a, b) as
// 1 / sqrt(a*a + b*b) and rsqrt(a) as 1 / sqrt(a) (both correctly rounded
// steps; CUDA documents its own as within 2 ulp of the exact value).
//
// Layout (MHD build, no scalars, no dual energy; src/grid/grid_enum.h:20-111):
//   U: 8 fields  density, momentum_x, momentum_y, momentum_z, Energy,
//                magnetic_x, magnetic_y, magnetic_z, field-major;
//      magnetic_?[i] is the face-centred field on the face i+1/2 of cell i.
//   interface (Q) and flux (F) arrays: 7 fields, the 5 hydro fields then two
//      magnetic slots (Q_x: By, Bz; Q_y: Bz, Bx; Q_z: Bx, By -- the transverse
//      fields in cyclic order; F: the fluxes of those two fields, slot 5 holding
//      the flux of the first, slot 6 the flux of the second).
//   electric fields: 3 fields, E_x, E_y, E_z on cell edges.
// Scratch arrays persist between calls (the reference allocates them once and
// never clears them again), so cells the kernels never write keep their values.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

#ifdef _OPENMP
  #include <omp.h>
#endif

extern "C" {
void orc_update_half_3d(const double *U, double *Uh, const double *Fx, const double *Fy, const double *Fz, int nx,
                        int ny, int nz, double dx, double dy, double dz, double dt, double gamma, int n_fields,
                        int density_floor_on, double density_floor);
int orc_get_dual_energy();
void orc_update_3d(double *U, const double *Fx, const double *Fy, const double *Fz, int nx, int ny, int nz,
                   int n_ghost, double dx, double dy, double dz, double dt, int n_fields, int density_floor_on,
                   double density_floor, double gamma);
void orc_set_crash_repair(int on);
int orc_get_crash_repair();
void orc_temperature_floor(double *U, int nx, int ny, int nz, int n_ghost, double U_floor);
void orc_set_boundaries_mhd(double *c, int nx, int ny, int nz, int n_fields, int n_ghost, const int *flags_in);
void orc_ppm_single_variable(double im2, double im1, double i0, double ip1, double ip2, double *L_iph, double *R_imh);
}

namespace
{

constexpr double kTiny      = 1.0e-20;  // TINY_NUMBER, src/global/global.h:25
constexpr double kHlldSmall = 1.0e-8;   // mhd::internal::_hlldSmallNumber, src/riemann_solvers/hlld_cuda.h:56
constexpr double kCfl       = 0.3;      // C_cfl, src/grid/grid3D.cpp:157
constexpr int kBx = 5, kBy = 6, kBz = 7;  // grid_enum::magnetic_x..z

inline long Lin(long i, long j, long k, long nx, long ny) { return i + j * nx + k * nx * ny; }
inline size_t Idx(int i, int j, int k, int nx, int ny)
{
  return size_t(i) + size_t(j) * size_t(nx) + size_t(k) * size_t(nx) * size_t(ny);
}

// math_utils::dotProduct / SquareMagnitude, src/utils/math_utilities.h:78-97
inline double Dot(double a1, double a2, double a3, double b1, double b2, double b3)
{
  return a1 * b1 + ((a2 * b2) + (a3 * b3));
}
inline double SquareMagnitude(double a, double b, double c) { return Dot(a, b, c, a, b, c); }

// mhd::utils, src/utils/mhd_utilities.h:50-151
inline double MagneticEnergy(double bx, double by, double bz) { return 0.5 * SquareMagnitude(bx, by, bz); }
inline double TotalPressure(double p, double bx, double by, double bz)
{
  double pt = p + MagneticEnergy(bx, by, bz);
  return std::fmax(pt, kTiny);
}
inline double MagnetosonicSpeed(double d, double p, double bx, double by, double bz, double gamma, double choice)
{
  double bx2 = bx * bx;
  double b2  = bx2 + ((by * by) + (bz * bz));
  double t1  = gamma * p + b2;
  double t2  = (t1 * t1) - 4. * gamma * p * bx2;
  t2         = std::sqrt(t2);
  return std::sqrt((t1 + choice * t2) / (2.0 * std::fmax(d, kTiny)));
}
inline double FastSpeed(double d, double p, double bx, double by, double bz, double g)
{
  return MagnetosonicSpeed(d, p, bx, by, bz, g, 1.0);
}
inline double SlowSpeed(double d, double p, double bx, double by, double bz, double g)
{
  return MagnetosonicSpeed(d, p, bx, by, bz, g, -1.0);
}
inline double AlfvenSpeed(double bx, double d) { return std::fabs(bx) / std::sqrt(std::fmax(d, kTiny)); }

// hydro_utilities::Calc_Pressure_Primitive / Calc_Energy_Primitive with MHD,
// src/utils/hydro_utilities.h:34-45, 106-118
inline double PressurePrim(double E, double d, double vx, double vy, double vz, double g, double bx, double by,
                           double bz)
{
  double p = E - 0.5 * d * SquareMagnitude(vx, vy, vz);
  p -= MagneticEnergy(bx, by, bz);
  return std::fmax((g - 1.) * p, kTiny);
}
inline double EnergyPrim(double P, double d, double vx, double vy, double vz, double g, double bx, double by, double bz)
{
  double e = (std::fmax(P, kTiny) / (g - 1.)) + 0.5 * d * SquareMagnitude(vx, vy, vz);
  e += MagneticEnergy(bx, by, bz);
  return e;
}

// mhd::utils::cellCenteredMagneticFields, src/utils/mhd_utilities.h:177-207
inline void CellCenteredB(const double *U, size_t id, int xid, int yid, int zid, size_t n_cells, int nx, int ny,
                          double &bx, double &by, double &bz)
{
  bx = (xid > 0) ? 0.5 * (U[kBx * n_cells + id] + U[kBx * n_cells + Idx(xid - 1, yid, zid, nx, ny)])
                 : U[kBx * n_cells + id];
  by = (yid > 0) ? 0.5 * (U[kBy * n_cells + id] + U[kBy * n_cells + Idx(xid, yid - 1, zid, nx, ny)])
                 : U[kBy * n_cells + id];
  bz = (zid > 0) ? 0.5 * (U[kBz * n_cells + id] + U[kBz * n_cells + Idx(xid, yid, zid - 1, nx, ny)])
                 : U[kBz * n_cells + id];
}

struct PrimM {
  double d, vx, vy, vz, p, bx, by, bz;
};
struct CharM {
  double a0, a1, a2, a3, a4, a5, a6;
};
struct Eigen {
  double cf, cs, cf2, cs2, af, as, by, bz, nfs, sign, qf, qs, Af, As, qpf, qps, Apf, Aps;
};

// reconstruction::Load_Data with MHD, src/reconstruction/reconstruction.h:139-212
PrimM LoadData(const double *U, int xid, int yid, int zid, int nx, int ny, size_t n_cells, int dir, double g)
{
  const int o1 = 1 + dir, o2 = 1 + (dir + 1) % 3, o3 = 1 + (dir + 2) % 3;
  size_t id = Idx(xid, yid, zid, nx, ny);
  PrimM w;
  w.d  = U[id];
  w.vx = U[o1 * n_cells + id] / w.d;
  w.vy = U[o2 * n_cells + id] / w.d;
  w.vz = U[o3 * n_cells + id] / w.d;
  double cx, cy, cz;
  CellCenteredB(U, id, xid, yid, zid, n_cells, nx, ny, cx, cy, cz);
  if (dir == 0) w.bx = cx, w.by = cy, w.bz = cz;
  if (dir == 1) w.bx = cy, w.by = cz, w.bz = cx;
  if (dir == 2) w.bx = cz, w.by = cx, w.bz = cy;
  w.p = PressurePrim(U[4 * n_cells + id], w.d, w.vx, w.vy, w.vz, g, w.bx, w.by, w.bz);
  return w;
}

// reconstruction::Compute_Slope / Van_Leer_Slope, reconstruction.h:224-300
PrimM Slope(const PrimM &l, const PrimM &r, double c = 1.0)
{
  return PrimM{c * (r.d - l.d),   c * (r.vx - l.vx), c * (r.vy - l.vy), c * (r.vz - l.vz),
               c * (r.p - l.p),   0.0,               c * (r.by - l.by), c * (r.bz - l.bz)};
}
inline double VlSlope(double l, double r) { return (l * r > 0.0) ? 2.0 * l * r / (l + r) : 0.0; }
PrimM VanLeer(const PrimM &l, const PrimM &r)
{
  return PrimM{VlSlope(l.d, r.d), VlSlope(l.vx, r.vx), VlSlope(l.vy, r.vy), VlSlope(l.vz, r.vz),
               VlSlope(l.p, r.p), 0.0,                 VlSlope(l.by, r.by), VlSlope(l.bz, r.bz)};
}

// reconstruction::Compute_Eigenvectors, reconstruction.h:310-370 (Stone et al. 2008, app. A)
Eigen ComputeEigen(const PrimM &w, double cs, double cs2, double g)
{
  Eigen e;
  e.cf  = FastSpeed(w.d, w.p, w.bx, w.by, w.bz, g);
  e.cs  = SlowSpeed(w.d, w.p, w.bx, w.by, w.bz, g);
  e.cf2 = e.cf * e.cf;
  e.cs2 = e.cs * e.cs;
  const double denom = (e.cf2 - e.cs2), num2 = (e.cf2 - cs2);
  if (denom <= 0.0 or num2 <= 0.0) {
    e.af = 1.0, e.as = 0.0;
  } else {
    const double num1 = (cs2 - e.cs2);
    if (num1 <= 0.0) {
      e.af = 0.0, e.as = 1.0;
    } else {
      e.af = std::sqrt(num1 / denom);
      e.as = std::sqrt(num2 / denom);
    }
  }
  const double bden = 1.0 / std::sqrt(w.by * w.by + w.bz * w.bz);  // rhypot
  e.by              = std::isfinite(bden) ? w.by * bden : 1.0;
  e.bz              = std::isfinite(bden) ? w.bz * bden : 0.0;
  e.sign            = std::copysign(1.0, w.bx);
  e.nfs             = 0.5 / cs2;
  e.qpf             = e.sign * e.nfs * e.af * e.cf;
  e.qps             = e.sign * e.nfs * e.as * e.cs;
  e.qf              = e.sign * e.af * e.cf;
  e.qs              = e.sign * e.as * e.cs;
  e.Af              = e.af * cs * std::sqrt(w.d);
  e.As              = e.as * cs * std::sqrt(w.d);
  e.Apf             = 0.5 * e.af / (cs * std::sqrt(w.d));
  e.Aps             = 0.5 * e.as / (cs * std::sqrt(w.d));
  return e;
}

// reconstruction::Primitive_To_Characteristic with MHD, reconstruction.h:378-425
CharM PrimToChar(const PrimM &w, const PrimM &s, const Eigen &e, double cs2)
{
  CharM c;
  const double isd = 1.0 / std::sqrt(w.d);  // rsqrt
  c.a0 = e.nfs * e.af * (s.p / w.d - e.cf * s.vx) + e.qps * (e.by * s.vy + e.bz * s.vz) +
         e.Aps * (e.by * s.by + e.bz * s.bz);
  c.a1 = 0.5 * (e.by * (s.bz * e.sign * isd + s.vz) - e.bz * (s.by * e.sign * isd + s.vy));
  c.a2 = e.nfs * e.as * (s.p / w.d - e.cs * s.vx) - e.qpf * (e.by * s.vy + e.bz * s.vz) -
         e.Apf * (e.by * s.by + e.bz * s.bz);
  c.a3 = s.d - s.p / cs2;
  c.a4 = e.nfs * e.as * (s.p / w.d + e.cs * s.vx) + e.qpf * (e.by * s.vy + e.bz * s.vz) -
         e.Apf * (e.by * s.by + e.bz * s.bz);
  c.a5 = 0.5 * (e.by * (s.bz * e.sign * isd - s.vz) - e.bz * (s.by * e.sign * isd - s.vy));
  c.a6 = e.nfs * e.af * (s.p / w.d + e.cf * s.vx) - e.qps * (e.by * s.vy + e.bz * s.vz) +
         e.Aps * (e.by * s.by + e.bz * s.bz);
  return c;
}

// reconstruction::Characteristic_To_Primitive with MHD, reconstruction.h:450-490
PrimM CharToPrim(const PrimM &w, const CharM &a, const Eigen &e, double cs2)
{
  PrimM o;
  o.d  = w.d * (e.af * (a.a0 + a.a6) + e.as * (a.a2 + a.a4)) + a.a3;
  o.vx = e.cf * e.af * (a.a6 - a.a0) + e.cs * e.as * (a.a4 - a.a2);
  o.vy = e.by * (e.qs * (a.a0 - a.a6) + e.qf * (a.a4 - a.a2)) + e.bz * (a.a5 - a.a1);
  o.vz = e.bz * (e.qs * (a.a0 - a.a6) + e.qf * (a.a4 - a.a2)) + e.by * (a.a1 - a.a5);
  o.p  = w.d * cs2 * (e.af * (a.a0 + a.a6) + e.as * (a.a2 + a.a4));
  o.bx = 0.0;
  o.by = e.by * (e.As * (a.a0 + a.a6) - e.Af * (a.a2 + a.a4)) - e.bz * e.sign * std::sqrt(w.d) * (a.a5 + a.a1);
  o.bz = e.bz * (e.As * (a.a0 + a.a6) - e.Af * (a.a2 + a.a4)) + e.by * e.sign * std::sqrt(w.d) * (a.a5 + a.a1);
  return o;
}

// reconstruction::Monotonize_Characteristic_Return_Primitive, reconstruction.h:512-560
inline double Mono(double l, double r, double c, double v)
{
  if (l * r > 0.0) {
    double la = 2.0 * std::fmin(std::fabs(l), std::fabs(r));
    double lb = std::fmin(std::fabs(c), std::fabs(v));
    return std::copysign(std::fmin(la, lb), c);
  }
  return 0.0;
}

// reconstruction::Calc_Interface_Linear, reconstruction.h:645-680
PrimM InterfaceLinear(const PrimM &w, const PrimM &s, double sign)
{
  auto f = [&](double q, double d) { return q + sign * 0.5 * d; };
  return PrimM{f(w.d, s.d), f(w.vx, s.vx), f(w.vy, s.vy), f(w.vz, s.vz), f(w.p, s.p), 0.0, f(w.by, s.by),
               f(w.bz, s.bz)};
}

// reconstruction::Plm_Limit_Interfaces, reconstruction.h:686-725
inline void Limit(double &l, double &r, double vm, double v0, double vp)
{
  r = std::fmax(std::fmin(v0, vm), r);
  r = std::fmin(std::fmax(v0, vm), r);
  l = std::fmax(std::fmin(v0, vp), l);
  l = std::fmin(std::fmax(v0, vp), l);
}

// reconstruction::Write_Data with MHD, reconstruction.h:918-970
void WriteData(const PrimM &s, double *Q, const double *U, size_t id, size_t n_cells, int dir, double g)
{
  const int o1 = 1 + dir, o2 = 1 + (dir + 1) % 3, o3 = 1 + (dir + 2) % 3;
  Q[id]               = s.d;
  Q[o1 * n_cells + id] = s.d * s.vx;
  Q[o2 * n_cells + id] = s.d * s.vy;
  Q[o3 * n_cells + id] = s.d * s.vz;
  Q[5 * n_cells + id]  = s.by;
  Q[6 * n_cells + id]  = s.bz;
  const double bx      = U[(kBx + dir) * n_cells + id];
  Q[4 * n_cells + id]  = EnergyPrim(s.p, s.d, s.vx, s.vy, s.vz, g, bx, s.by, s.bz);
}

// mhd::internal HLLD state/flux types, src/riemann_solvers/hlld_cuda.h
struct HState {
  double d, vx, vy, vz, E, by, bz, pg, pt;
};
struct HFlux {
  double d, mx, my, mz, E, by, bz;
};
struct HStar {
  double d, vy, vz, E, by, bz;
};
struct HDStar {
  double vy, vz, by, bz, EL, ER;
};
struct HSpeeds {
  double L, R, M, LStar, RStar;
};

// mhd::internal::loadState, hlld_cuda.cu:170-214 (no scalars, no DE)
HState LoadState(const double *Q, double bx, double g, size_t id, size_t n, int dir)
{
  const int o1 = 1 + dir, o2 = 1 + (dir + 1) % 3, o3 = 1 + (dir + 2) % 3;
  HState s;
  s.d  = std::fmax(Q[id], kTiny);
  s.vx = Q[o1 * n + id] / s.d;
  s.vy = Q[o2 * n + id] / s.d;
  s.vz = Q[o3 * n + id] / s.d;
  s.E  = std::fmax(Q[4 * n + id], kTiny);
  s.by = Q[5 * n + id];
  s.bz = Q[6 * n + id];
  s.pg = PressurePrim(s.E, s.d, s.vx, s.vy, s.vz, g, bx, s.by, s.bz);
  s.pt = TotalPressure(s.pg, bx, s.by, s.bz);
  return s;
}

// mhd::internal::nonStarFluxes, hlld_cuda.cu:262-283
HFlux NonStarFlux(const HState &s, double bx)
{
  HFlux f;
  f.d  = s.d * s.vx;
  f.mx = f.d * s.vx + s.pt - bx * bx;
  f.my = f.d * s.vy - bx * s.by;
  f.mz = f.d * s.vz - bx * s.bz;
  f.by = s.by * s.vx - bx * s.vy;
  f.bz = s.bz * s.vx - bx * s.vz;
  f.E  = s.vx * (s.E + s.pt) - bx * (s.vx * bx + ((s.vy * s.by) + (s.vz * s.bz)));
  return f;
}

// mhd::internal::computeStarState, hlld_cuda.cu:327-372
HStar StarState(const HState &s, const HSpeeds &sp, double side, double bx, double ptStar)
{
  HStar o;
  o.d = s.d * (side - s.vx) / (side - sp.M);
  if (std::fabs(s.d * (side - s.vx) * (side - sp.M) - (bx * bx)) < ptStar * kHlldSmall) {
    o.vy = s.vy, o.vz = s.vz, o.by = s.by, o.bz = s.bz;
  } else {
    const double denom = s.d * (side - s.vx) * (side - sp.M) - (bx * bx);
    double coef        = bx * (sp.M - s.vx) / denom;
    o.vy               = s.vy - s.by * coef;
    o.vz               = s.vz - s.bz * coef;
    double tp          = (side - s.vx);
    tp                 = tp * tp;
    coef               = (s.d * tp - (bx * bx)) / denom;
    o.by               = s.by * coef;
    o.bz               = s.bz * coef;
  }
  o.E = (s.E * (side - s.vx) - s.pt * s.vx + ptStar * sp.M +
         bx * (Dot(s.vx, s.vy, s.vz, bx, s.by, s.bz) - Dot(sp.M, o.vy, o.vz, bx, o.by, o.bz))) /
        (side - sp.M);
  return o;
}

// mhd::internal::starFluxes, hlld_cuda.cu:376-397
HFlux StarFlux(const HStar &st, const HState &s, const HFlux &f, const HSpeeds &sp, double side)
{
  HFlux o;
  o.d  = f.d + side * (st.d - s.d);
  o.mx = f.mx + side * (st.d * sp.M - s.d * s.vx);
  o.my = f.my + side * (st.d * st.vy - s.d * s.vy);
  o.mz = f.mz + side * (st.d * st.vz - s.d * s.vz);
  o.E  = f.E + side * (st.E - s.E);
  o.by = f.by + side * (st.by - s.by);
  o.bz = f.bz + side * (st.bz - s.bz);
  return o;
}

// mhd::internal::computeDoubleStarState, hlld_cuda.cu:401-465.  With Bx ~ 0 the
// reference sets only the energy of the side it selects; the other side's energy
// is left undefined there and is set to that side's star energy here.
HDStar DoubleStar(const HStar &L, const HStar &R, double bx, double ptStar, const HSpeeds &sp)
{
  HDStar o;
  if (0.5 * (bx * bx) < kHlldSmall * ptStar) {
    if (sp.M >= 0.0) {
      o.vy = L.vy, o.vz = L.vz, o.by = L.by, o.bz = L.bz, o.EL = L.E, o.ER = R.E;
    } else {
      o.vy = R.vy, o.vz = R.vz, o.by = R.by, o.bz = R.bz, o.ER = R.E, o.EL = L.E;
    }
  } else {
    const double sL = std::sqrt(L.d), sR = std::sqrt(R.d);
    const double inv = 1.0 / (sL + sR);
    const double sgn = std::copysign(1.0, bx);
    o.vy = inv * (sL * L.vy + sR * R.vy + sgn * (R.by - L.by));
    o.vz = inv * (sL * L.vz + sR * R.vz + sgn * (R.bz - L.bz));
    o.by = inv * (sL * R.by + sR * L.by + sgn * (sL * sR) * (R.vy - L.vy));
    o.bz = inv * (sL * R.bz + sR * L.bz + sgn * (sL * sR) * (R.vz - L.vz));
    const double vb = Dot(sp.M, o.vy, o.vz, bx, o.by, o.bz);
    o.EL = L.E - sL * sgn * (Dot(sp.M, L.vy, L.vz, bx, L.by, L.bz) - vb);
    o.ER = R.E + sR * sgn * (Dot(sp.M, R.vy, R.vz, bx, R.by, R.bz) - vb);
  }
  return o;
}

// mhd::internal::computeDoubleStarFluxes, hlld_cuda.cu:469-505
HFlux DoubleStarFlux(const HDStar &dd, double dE, const HStar &st, const HState &s, const HFlux &f, const HSpeeds &sp,
                     double side, double sideStar)
{
  const double diff = sideStar - side;
  HFlux o;
  o.d  = f.d - side * s.d - diff * st.d + sideStar * st.d;
  o.mx = f.mx - side * (s.d * s.vx) - diff * (st.d * sp.M) + sideStar * (st.d * sp.M);
  o.my = f.my - side * (s.d * s.vy) - diff * (st.d * st.vy) + sideStar * (st.d * dd.vy);
  o.mz = f.mz - side * (s.d * s.vz) - diff * (st.d * st.vz) + sideStar * (st.d * dd.vz);
  o.E  = f.E - side * s.E - diff * st.E + sideStar * dE;
  o.by = f.by - side * s.by - diff * st.by + sideStar * dd.by;
  o.bz = f.bz - side * s.bz - diff * st.bz + sideStar * dd.bz;
  return o;
}

// mhd::internal::returnFluxes, hlld_cuda.cu:287-304
void ReturnFlux(double *F, size_t id, size_t n, int dir, const HFlux &f)
{
  const int o1 = 1 + dir, o2 = 1 + (dir + 1) % 3, o3 = 1 + (dir + 2) % 3;
  F[id]          = f.d;
  F[o1 * n + id] = f.mx;
  F[o2 * n + id] = f.my;
  F[o3 * n + id] = f.mz;
  F[4 * n + id]  = f.E;
  F[5 * n + id]  = f.by;  // fluxX_magnetic_z slot
  F[6 * n + id]  = f.bz;  // fluxX_magnetic_y slot
}

// One interface of mhd::Calculate_HLLD_Fluxes_CUDA, hlld_cuda.cu:31-164
void HlldCell(const double *QL, const double *QR, double bx, double *F, size_t id, size_t n, double g, int dir)
{
  const HState sL = LoadState(QL, bx, g, id, n, dir), sR = LoadState(QR, bx, g, id, n, dir);
  HSpeeds sp{};
  // approximateLRWaveSpeeds, hlld_cuda.cu:218-236
  {
    double cL = FastSpeed(sL.d, sL.pg, bx, sL.by, sL.bz, g);
    double cR = FastSpeed(sR.d, sR.pg, bx, sR.by, sR.bz, g);
    double cM = std::fmax(cL, cR);
    sp.L      = std::fmin(sL.vx, sR.vx) - cM;
    sp.R      = std::fmax(sL.vx, sR.vx) + cM;
  }
  HFlux fL = NonStarFlux(sL, bx);
  if (sp.L > 0.0) return ReturnFlux(F, id, n, dir, fL);
  HFlux fR = NonStarFlux(sR, bx);
  if (sp.R < 0.0) return ReturnFlux(F, id, n, dir, fR);
  // approximateMiddleWaveSpeed, hlld_cuda.cu:240-256
  {
    const double rd = sp.R - sR.vx, ld = sp.L - sL.vx;
    sp.M = (rd * sR.d * sR.vx - ld * sL.d * sL.vx - sR.pt + sL.pt) / (rd * sR.d - ld * sL.d);
  }
  // starTotalPressure, hlld_cuda.cu:308-323
  const double ptStar = (sR.d * sL.pt * (sp.R - sR.vx) - sL.d * sR.pt * (sp.L - sL.vx) +
                         sL.d * sR.d * (sp.R - sR.vx) * (sp.L - sL.vx) * (sR.vx - sL.vx)) /
                        (sR.d * (sp.R - sR.vx) - sL.d * (sp.L - sL.vx));
  const HStar stL = StarState(sL, sp, sp.L, bx, ptStar);
  sp.LStar        = sp.M + -1.0 * AlfvenSpeed(bx, stL.d);  // approximateStarWaveSpeed, hlld_cuda.cu:260-266
  if (sp.LStar > 0.0 and sp.L <= 0.0) return ReturnFlux(F, id, n, dir, StarFlux(stL, sL, fL, sp, sp.L));
  const HStar stR = StarState(sR, sp, sp.R, bx, ptStar);
  sp.RStar        = sp.M + 1.0 * AlfvenSpeed(bx, stR.d);
  if (sp.RStar <= 0.0 and sp.R >= 0.0) return ReturnFlux(F, id, n, dir, StarFlux(stR, sR, fR, sp, sp.R));
  const HDStar dd = DoubleStar(stL, stR, bx, ptStar, sp);
  if (sp.M > 0.0 and sp.LStar <= 0.0)
    return ReturnFlux(F, id, n, dir, DoubleStarFlux(dd, dd.EL, stL, sL, fL, sp, sp.L, sp.LStar));
  if (sp.RStar > 0.0 and sp.M <= 0.0)
    return ReturnFlux(F, id, n, dir, DoubleStarFlux(dd, dd.ER, stR, sR, fR, sp, sp.R, sp.RStar));
}

// mhd::internal::_ctSlope, src/mhd/ct_electric_fields.h:66-119.  Indices are formed
// as the reference forms them; a read outside the array returns 0 (such reads only
// occur at the outer edge of the grid, whose edge fields no real cell uses).
double CtSlope(const double *flux, const double *U, double fluxSign, int ctDir, int cq1, int cq2, int fq1, int fq2,
               int xid, int yid, int zid, int nx, int ny, long n_cells, long n_flux, long n_u)
{
  auto at = [](const double *a, long i, long n) { return (i >= 0 && i < n) ? a[i] : 0.0; };
  const int m1 = (ctDir + 1) % 3, m2 = (ctDir + 2) % 3;
  const int xc = xid - int(cq1 == 0) - int(cq2 == 0);
  const int yc = yid - int(cq1 == 1) - int(cq2 == 1);
  const int zc = zid - int(cq1 == 2) - int(cq2 == 2);
  const long ic = Lin(xc, yc, zc, nx, ny);
  const long iflux = Lin(xid - int(fq1 == 0) - int(fq2 == 0), yid - int(fq1 == 1) - int(fq2 == 1),
                         zid - int(fq1 == 2) - int(fq2 == 2), nx, ny);
  const long ib2 = Lin(xc - int(m1 == 0), yc - int(m1 == 1), zc - int(m1 == 2), nx, ny);
  const long ib3 = Lin(xc - int(m2 == 0), yc - int(m2 == 1), zc - int(m2 == 2), nx, ny);
  const double d  = at(U, ic, n_u);
  const double m2v = at(U, ic + (m1 + 1) * n_cells, n_u);
  const double m3v = at(U, ic + (m2 + 1) * n_cells, n_u);
  const double b2 = 0.5 * (at(U, ic + (m1 + kBx) * n_cells, n_u) + at(U, ib2 + (m1 + kBx) * n_cells, n_u));
  const double b3 = 0.5 * (at(U, ic + (m2 + kBx) * n_cells, n_u) + at(U, ib3 + (m2 + kBx) * n_cells, n_u));
  const double ec = (m3v * b2 - m2v * b3) / d;
  const double ef = fluxSign * at(flux, iflux + (int(fluxSign == 1) + kBx) * n_cells, n_flux);
  return ef - ec;
}

struct MhdScratch {
  std::vector<double> Uh, QLx, QRx, QLy, QRy, QLz, QRz, Fx, Fy, Fz, E;
  size_t n = 0;
  void Resize(size_t n_cells)
  {
    if (n == n_cells) return;
    n = n_cells;
    for (auto *v : {&QLx, &QRx, &QLy, &QRy, &QLz, &QRz, &Fx, &Fy, &Fz}) v->assign(7 * n_cells, 0.0);
    Uh.assign(8 * n_cells, 0.0);
    E.assign(3 * n_cells, 0.0);
  }
};
MhdScratch g_mhd;

}  // namespace

extern "C" {

// ---------------------------------------------------------------- utilities (for the unit vectors)
double orc_mhd_fast_speed(double d, double p, double bx, double by, double bz, double g)
{
  return FastSpeed(d, p, bx, by, bz, g);
}
double orc_mhd_slow_speed(double d, double p, double bx, double by, double bz, double g)
{
  return SlowSpeed(d, p, bx, by, bz, g);
}
double orc_mhd_alfven_speed(double bx, double d) { return AlfvenSpeed(bx, d); }
double orc_mhd_total_pressure(double p, double bx, double by, double bz) { return TotalPressure(p, bx, by, bz); }
double orc_mhd_pressure_prim(double E, double d, double vx, double vy, double vz, double g, double bx, double by,
                             double bz)
{
  return PressurePrim(E, d, vx, vy, vz, g, bx, by, bz);
}
double orc_mhd_energy_prim(double P, double d, double vx, double vy, double vz, double g, double bx, double by,
                           double bz)
{
  return EnergyPrim(P, d, vx, vy, vz, g, bx, by, bz);
}
void orc_mhd_cell_centered_b(const double *U, int xid, int yid, int zid, int nx, int ny, int nz, double *out)
{
  CellCenteredB(U, Idx(xid, yid, zid, nx, ny), xid, yid, zid, size_t(nx) * ny * nz, nx, ny, out[0], out[1], out[2]);
}
// Load_Data: out = d, vx, vy, vz, p, bx, by, bz (rotated to dir)
void orc_mhd_load_data(const double *U, int xid, int yid, int zid, int nx, int ny, int nz, int dir, double g,
                       double *out)
{
  PrimM w = LoadData(U, xid, yid, zid, nx, ny, size_t(nx) * ny * nz, dir, g);
  double v[8] = {w.d, w.vx, w.vy, w.vz, w.p, w.bx, w.by, w.bz};
  std::memcpy(out, v, sizeof(v));
}
// Compute_Eigenvectors: out = 18 values in the EigenVecs member order of reconstruction.h:51-63
void orc_mhd_eigenvectors(const double *prim, double cs, double cs2, double g, double *out)
{
  PrimM w{prim[0], prim[1], prim[2], prim[3], prim[4], prim[5], prim[6], prim[7]};
  Eigen e = ComputeEigen(w, cs, cs2, g);
  double v[18] = {e.cf, e.cs, e.cf2, e.cs2, e.af, e.as, e.by, e.bz, e.nfs, e.sign,
                  e.qf, e.qs, e.Af,  e.As,  e.qpf, e.qps, e.Apf, e.Aps};
  std::memcpy(out, v, sizeof(v));
}
// Primitive_To_Characteristic / Characteristic_To_Primitive for one cell (prim: 8 values,
// slope: 8 values, char: 7 values)
void orc_mhd_prim_to_char(const double *prim, const double *slope, double cs, double cs2, double g, double *out)
{
  PrimM w{prim[0], prim[1], prim[2], prim[3], prim[4], prim[5], prim[6], prim[7]};
  PrimM s{slope[0], slope[1], slope[2], slope[3], slope[4], slope[5], slope[6], slope[7]};
  Eigen e = ComputeEigen(w, cs, cs2, g);
  CharM c = PrimToChar(w, s, e, cs2);
  double v[7] = {c.a0, c.a1, c.a2, c.a3, c.a4, c.a5, c.a6};
  std::memcpy(out, v, sizeof(v));
}
void orc_mhd_char_to_prim(const double *prim, const double *ch, double cs, double cs2, double g, double *out)
{
  PrimM w{prim[0], prim[1], prim[2], prim[3], prim[4], prim[5], prim[6], prim[7]};
  CharM c{ch[0], ch[1], ch[2], ch[3], ch[4], ch[5], ch[6]};
  Eigen e = ComputeEigen(w, cs, cs2, g);
  PrimM o = CharToPrim(w, c, e, cs2);
  double v[8] = {o.d, o.vx, o.vy, o.vz, o.p, o.bx, o.by, o.bz};
  std::memcpy(out, v, sizeof(v));
}

// mhdInverseCrossingTime, src/hydro/hydro_cuda.cu:420-435
double orc_mhd_inverse_crossing_time(double E, double d, double d_inv, double vx, double vy, double vz, double bx,
                                     double by, double bz, double dx, double dy, double dz, double g)
{
  double gasP = PressurePrim(E, d, vx, vy, vz, g, bx, by, bz);
  double cf   = FastSpeed(d, gasP, bx, by, bz, g);
  double m    = std::fmax((std::fabs(vx) + cf) / dx, (std::fabs(vy) + cf) / dy);
  m           = std::fmax(m, (std::fabs(vz) + cf) / dz);
  return std::fmax(m, 0.0);
}

// ---------------------------------------------------------------- stage kernels over whole grids
// PCM_Reconstruction_3D with MHD, src/reconstruction/pcm_cuda.cu:256-420
void orc_mhd_pcm_3d(const double *U, double *QLx, double *QRx, double *QLy, double *QRy, double *QLz, double *QRz,
                    int nx, int ny, int nz)
{
  const size_t n = size_t(nx) * ny * nz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const size_t id = Idx(i, j, k, nx, ny);
        double c[5];
        for (int f = 0; f < 5; f++) c[f] = U[f * n + id];
        double bx, by, bz;
        CellCenteredB(U, id, i, j, k, n, nx, ny, bx, by, bz);
        const double tx[2] = {by, bz}, ty[2] = {bz, bx}, tz[2] = {bx, by};
        auto put = [&](double *Q, size_t at, const double *t) {
          for (int f = 0; f < 5; f++) Q[f * n + at] = c[f];
          Q[5 * n + at] = t[0];
          Q[6 * n + at] = t[1];
        };
        put(QLx, id, tx);
        put(QLy, id, ty);
        put(QLz, id, tz);
        if (i > 0) put(QRx, Idx(i - 1, j, k, nx, ny), tx);
        if (j > 0) put(QRy, Idx(i, j - 1, k, nx, ny), ty);
        if (k > 0) put(QRz, Idx(i, j, k - 1, nx, ny), tz);
      }
}

// PLMC_cuda with MHD (VL), src/reconstruction/plmc_cuda.cu:24-300
void orc_mhd_plmc(const double *U, double *QL, double *QR, int nx, int ny, int nz, double g, int dir)
{
  const size_t n = size_t(nx) * ny * nz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        // reconstruction::Thread_Guard<2>, reconstruction.h:95-119
        bool guard = i < 1 || i >= nx - 2;
        if (ny > 1) guard = guard || j < 1 || j >= ny - 2;
        if (nz > 1) guard = guard || k < 1 || k >= nz - 2;
        if (guard) continue;
        const int di = dir == 0, dj = dir == 1, dk = dir == 2;
        const PrimM c0 = LoadData(U, i, j, k, nx, ny, n, dir, g);
        const PrimM cm = LoadData(U, i - di, j - dj, k - dk, nx, ny, n, dir, g);
        const PrimM cp = LoadData(U, i + di, j + dj, k + dk, nx, ny, n, dir, g);
        const double cs  = std::sqrt(g * c0.p / c0.d);  // Calc_Sound_Speed, hydro_utilities.h:209-212
        const double cs2 = cs * cs;
        const Eigen e    = ComputeEigen(c0, cs, cs2, g);
        const PrimM dL = Slope(cm, c0), dR = Slope(c0, cp), dC = Slope(cm, cp, 0.5), dG = VanLeer(dL, dR);
        const CharM aL = PrimToChar(c0, dL, e, cs2), aR = PrimToChar(c0, dR, e, cs2);
        const CharM aC = PrimToChar(c0, dC, e, cs2), aG = PrimToChar(c0, dG, e, cs2);
        CharM am;
        am.a0 = Mono(aL.a0, aR.a0, aC.a0, aG.a0);
        am.a1 = Mono(aL.a1, aR.a1, aC.a1, aG.a1);
        am.a2 = Mono(aL.a2, aR.a2, aC.a2, aG.a2);
        am.a3 = Mono(aL.a3, aR.a3, aC.a3, aG.a3);
        am.a4 = Mono(aL.a4, aR.a4, aC.a4, aG.a4);
        am.a5 = Mono(aL.a5, aR.a5, aC.a5, aG.a5);
        am.a6 = Mono(aL.a6, aR.a6, aC.a6, aG.a6);
        const PrimM dm = CharToPrim(c0, am, e, cs2);
        PrimM Lp = InterfaceLinear(c0, dm, 1.0), Rm = InterfaceLinear(c0, dm, -1.0);
        Limit(Lp.d, Rm.d, cm.d, c0.d, cp.d);
        Limit(Lp.vx, Rm.vx, cm.vx, c0.vx, cp.vx);
        Limit(Lp.vy, Rm.vy, cm.vy, c0.vy, cp.vy);
        Limit(Lp.vz, Rm.vz, cm.vz, c0.vz, cp.vz);
        Limit(Lp.p, Rm.p, cm.p, c0.p, cp.p);
        Limit(Lp.by, Rm.by, cm.by, c0.by, cp.by);
        Limit(Lp.bz, Rm.bz, cm.bz, c0.bz, cp.bz);
        Rm.d = std::fmax(Rm.d, kTiny);
        Lp.d = std::fmax(Lp.d, kTiny);
        Rm.p = std::fmax(Rm.p, kTiny);
        Lp.p = std::fmax(Lp.p, kTiny);
        WriteData(Lp, QL, U, Idx(i, j, k, nx, ny), n, dir, g);
        WriteData(Rm, QR, U, Idx(i - di, j - dj, k - dk, nx, ny), n, dir, g);
      }
}

// PPMC_VL with MHD, src/reconstruction/ppmc_cuda.cu:540-695: the 5-cell stencil projected on
// cell i's seven MHD characteristics, PPM_Single_Variable on each, projected back, density and
// pressure floored, written as conserved states (Q_L at i, Q_R at i-1).
void orc_mhd_ppmc(const double *U, double *QL, double *QR, int nx, int ny, int nz, double g, int dir)
{
  const size_t n = size_t(nx) * ny * nz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        // reconstruction::Thread_Guard<3>, reconstruction.h:95-119
        bool guard = i < 2 || i >= nx - 3;
        if (ny > 1) guard = guard || j < 2 || j >= ny - 3;
        if (nz > 1) guard = guard || k < 2 || k >= nz - 3;
        if (guard) continue;
        const int di = dir == 0, dj = dir == 1, dk = dir == 2;
        const PrimM c0  = LoadData(U, i, j, k, nx, ny, n, dir, g);
        const PrimM cm1 = LoadData(U, i - di, j - dj, k - dk, nx, ny, n, dir, g);
        const PrimM cp1 = LoadData(U, i + di, j + dj, k + dk, nx, ny, n, dir, g);
        const PrimM cm2 = LoadData(U, i - 2 * di, j - 2 * dj, k - 2 * dk, nx, ny, n, dir, g);
        const PrimM cp2 = LoadData(U, i + 2 * di, j + 2 * dj, k + 2 * dk, nx, ny, n, dir, g);
        const double cs  = std::sqrt(g * c0.p / c0.d);  // Calc_Sound_Speed, hydro_utilities.h:209-212
        const double cs2 = cs * cs;
        const Eigen e    = ComputeEigen(c0, cs, cs2, g);
        const CharM a0 = PrimToChar(c0, c0, e, cs2), am1 = PrimToChar(c0, cm1, e, cs2);
        const CharM am2 = PrimToChar(c0, cm2, e, cs2), ap1 = PrimToChar(c0, cp1, e, cs2);
        const CharM ap2 = PrimToChar(c0, cp2, e, cs2);
        CharM L, R;
        orc_ppm_single_variable(am2.a0, am1.a0, a0.a0, ap1.a0, ap2.a0, &L.a0, &R.a0);
        orc_ppm_single_variable(am2.a1, am1.a1, a0.a1, ap1.a1, ap2.a1, &L.a1, &R.a1);
        orc_ppm_single_variable(am2.a2, am1.a2, a0.a2, ap1.a2, ap2.a2, &L.a2, &R.a2);
        orc_ppm_single_variable(am2.a3, am1.a3, a0.a3, ap1.a3, ap2.a3, &L.a3, &R.a3);
        orc_ppm_single_variable(am2.a4, am1.a4, a0.a4, ap1.a4, ap2.a4, &L.a4, &R.a4);
        orc_ppm_single_variable(am2.a5, am1.a5, a0.a5, ap1.a5, ap2.a5, &L.a5, &R.a5);
        orc_ppm_single_variable(am2.a6, am1.a6, a0.a6, ap1.a6, ap2.a6, &L.a6, &R.a6);
        PrimM Lp = CharToPrim(c0, L, e, cs2), Rm = CharToPrim(c0, R, e, cs2);
        Rm.d = std::fmax(Rm.d, kTiny);
        Lp.d = std::fmax(Lp.d, kTiny);
        Rm.p = std::fmax(Rm.p, kTiny);
        Lp.p = std::fmax(Lp.p, kTiny);
        WriteData(Lp, QL, U, Idx(i, j, k, nx, ny), n, dir, g);
        WriteData(Rm, QR, U, Idx(i - di, j - dj, k - dk, nx, ny), n, dir, g);
      }
}

// mhd::Calculate_HLLD_Fluxes_CUDA over a grid, hlld_cuda.cu:31-164.  magX: the face
// field normal to the interfaces (the state's magnetic_x/y/z array for dir 0/1/2).
void orc_hlld(const double *QL, const double *QR, const double *magX, double *F, long n_cells, double g, int dir)
{
#pragma omp parallel for schedule(static)
  for (long id = 0; id < n_cells; id++) HlldCell(QL, QR, magX[id], F, size_t(id), size_t(n_cells), g, dir);
}

double orc_ct_slope(const double *flux, const double *U, double fluxSign, int ctDir, int cq1, int cq2, int fq1,
                    int fq2, int xid, int yid, int zid, int nx, int ny, long n_cells, long n_flux_total,
                    long n_u_total)
{
  return CtSlope(flux, U, fluxSign, ctDir, cq1, cq2, fq1, fq2, xid, yid, zid, nx, ny, n_cells, n_flux_total,
                 n_u_total);
}

// mhd::Calculate_CT_Electric_Fields, src/mhd/ct_electric_fields.cu:20-279.
// n_u_fields: fields in U (8 for the MHD grid).
void orc_ct_electric_fields(const double *Fx, const double *Fy, const double *Fz, const double *U, double *E, int nx,
                            int ny, int nz, int n_u_fields)
{
  const long n = long(nx) * ny * nz, nF = 7 * n, nU = long(n_u_fields) * n;
  auto F = [&](const double *a, int i, int j, int k, int f) {
    long q = Lin(i, j, k, nx, ny) + long(f) * n;
    return (q >= 0 && q < nF) ? a[q] : 0.0;
  };
  auto S = [&](const double *flux, double sgn, int ctd, int a, int b, int c, int d, int i, int j, int k) {
    return CtSlope(flux, U, sgn, ctd, a, b, c, d, i, j, k, nx, ny, n, nF, nU);
  };
  // S&G 2009 eq. 23: upwinded slope, or the mean of the two when the mass flux is 0
  auto up = [](double sgnv, double pos, double neg) { return sgnv > 0.0 ? pos : (sgnv < 0.0 ? neg : 0.5 * (pos + neg)); };
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k < nz; k++)
    for (int j = 1; j < ny; j++)
      for (int i = 1; i < nx; i++) {
        const size_t id = Idx(i, j, k, nx, ny);
        double syp, syn, szp, szn, sxp, sxn;
        // ---- x electric field
        {
          double s = F(Fz, i, j, k - 1, 0);
          syp = s > 0 ? S(Fy, -1, 0, 2, -1, 1, 2, i, j, k) : (s < 0 ? S(Fy, -1, 0, -1, -1, 1, -1, i, j, k)
                                                                 : up(0, S(Fy, -1, 0, 2, -1, 1, 2, i, j, k),
                                                                      S(Fy, -1, 0, -1, -1, 1, -1, i, j, k)));
          s   = F(Fz, i, j - 1, k - 1, 0);
          syn = s > 0 ? S(Fy, -1, 0, 1, 2, 1, 2, i, j, k) : (s < 0 ? S(Fy, -1, 0, 1, -1, 1, -1, i, j, k)
                                                                : up(0, S(Fy, -1, 0, 1, 2, 1, 2, i, j, k),
                                                                     S(Fy, -1, 0, 1, -1, 1, -1, i, j, k)));
          s   = F(Fy, i, j - 1, k, 0);
          szp = s > 0 ? S(Fz, 1, 0, 1, -1, 1, 2, i, j, k) : (s < 0 ? S(Fz, 1, 0, -1, -1, 2, -1, i, j, k)
                                                                : up(0, S(Fz, 1, 0, 1, -1, 1, 2, i, j, k),
                                                                     S(Fz, 1, 0, -1, -1, 2, -1, i, j, k)));
          s   = F(Fy, i, j - 1, k - 1, 0);
          szn = s > 0 ? S(Fz, 1, 0, 1, 2, 1, 2, i, j, k) : (s < 0 ? S(Fz, 1, 0, 2, -1, -1, 2, i, j, k)
                                                               : up(0, S(Fz, 1, 0, 1, 2, 1, 2, i, j, k),
                                                                    S(Fz, 1, 0, 2, -1, -1, 2, i, j, k)));
          const double fyp = +F(Fz, i, j, k - 1, 6), fyn = +F(Fz, i, j - 1, k - 1, 6);  // fluxZ_magnetic_x
          const double fzp = -F(Fy, i, j - 1, k, 5), fzn = -F(Fy, i, j - 1, k - 1, 5);  // fluxY_magnetic_x
          E[0 * n + id] = 0.25 * (+fyp + fyn + fzp + fzn + syp + syn + szp + szn);
        }
        // ---- y electric field
        {
          double s = F(Fz, i, j, k - 1, 0);
          sxp = s > 0 ? S(Fx, 1, 1, 2, -1, 0, 2, i, j, k) : (s < 0 ? S(Fx, 1, 1, -1, -1, 0, -1, i, j, k)
                                                                : up(0, S(Fx, 1, 1, 2, -1, 0, 2, i, j, k),
                                                                     S(Fx, 1, 1, -1, -1, 0, -1, i, j, k)));
          s   = F(Fz, i - 1, j, k - 1, 0);
          sxn = s > 0 ? S(Fx, 1, 1, 0, 2, 0, 2, i, j, k) : (s < 0 ? S(Fx, 1, 1, 0, -1, 0, -1, i, j, k)
                                                               : up(0, S(Fx, 1, 1, 0, 2, 0, 2, i, j, k),
                                                                    S(Fx, 1, 1, 0, -1, 0, -1, i, j, k)));
          s   = F(Fx, i - 1, j, k, 0);
          szp = s > 0 ? S(Fz, -1, 1, 0, -1, 0, 2, i, j, k) : (s < 0 ? S(Fz, -1, 1, -1, -1, 2, -1, i, j, k)
                                                                 : up(0, S(Fz, -1, 1, 0, -1, 0, 2, i, j, k),
                                                                      S(Fz, -1, 1, -1, -1, 2, -1, i, j, k)));
          s   = F(Fx, i - 1, j, k - 1, 0);
          szn = s > 0 ? S(Fz, -1, 1, 0, 2, 0, 2, i, j, k) : (s < 0 ? S(Fz, -1, 1, 2, -1, 2, -1, i, j, k)
                                                                : up(0, S(Fz, -1, 1, 0, 2, 0, 2, i, j, k),
                                                                     S(Fz, -1, 1, 2, -1, 2, -1, i, j, k)));
          const double fxp = -F(Fz, i, j, k - 1, 5), fxn = -F(Fz, i - 1, j, k - 1, 5);  // fluxZ_magnetic_y
          const double fzp = +F(Fx, i - 1, j, k, 6), fzn = +F(Fx, i - 1, j, k - 1, 6);  // fluxX_magnetic_y
          E[1 * n + id] = 0.25 * (+fxp + fxn + fzp + fzn + sxp + sxn + szp + szn);
        }
        // ---- z electric field
        {
          double s = F(Fx, i - 1, j, k, 0);
          syp = s > 0 ? S(Fy, 1, 2, 0, -1, 0, 1, i, j, k) : (s < 0 ? S(Fy, 1, 2, -1, -1, 1, -1, i, j, k)
                                                                : up(0, S(Fy, 1, 2, 0, -1, 0, 1, i, j, k),
                                                                     S(Fy, 1, 2, -1, -1, 1, -1, i, j, k)));
          s   = F(Fx, i - 1, j - 1, k, 0);
          syn = s > 0 ? S(Fy, 1, 2, 0, 1, 0, 1, i, j, k) : (s < 0 ? S(Fy, 1, 2, 1, -1, 1, -1, i, j, k)
                                                               : up(0, S(Fy, 1, 2, 0, 1, 0, 1, i, j, k),
                                                                    S(Fy, 1, 2, 1, -1, 1, -1, i, j, k)));
          s   = F(Fy, i, j - 1, k, 0);
          sxp = s > 0 ? S(Fx, -1, 2, 1, -1, 0, 1, i, j, k) : (s < 0 ? S(Fx, -1, 2, -1, -1, 0, -1, i, j, k)
                                                                 : up(0, S(Fx, -1, 2, 1, -1, 0, 1, i, j, k),
                                                                      S(Fx, -1, 2, -1, -1, 0, -1, i, j, k)));
          s   = F(Fy, i - 1, j - 1, k, 0);
          sxn = s > 0 ? S(Fx, -1, 2, 0, 1, 0, 1, i, j, k) : (s < 0 ? S(Fx, -1, 2, 0, -1, 0, -1, i, j, k)
                                                                : up(0, S(Fx, -1, 2, 0, 1, 0, 1, i, j, k),
                                                                     S(Fx, -1, 2, 0, -1, 0, -1, i, j, k)));
          const double fxp = +F(Fy, i, j - 1, k, 6), fxn = +F(Fy, i - 1, j - 1, k, 6);  // fluxY_magnetic_z
          const double fyp = -F(Fx, i - 1, j, k, 5), fyn = -F(Fx, i - 1, j - 1, k, 5);  // fluxX_magnetic_z
          E[2 * n + id] = 0.25 * (+fxp + fxn + fyp + fyn + sxp + sxn + syp + syn);
        }
      }
}

// mhd::Update_Magnetic_Field_3D, src/mhd/magnetic_update.cu:21-81 (src may equal dst)
void orc_update_magnetic(const double *src, double *dst, const double *E, int nx, int ny, int nz, double dt, double dx,
                         double dy, double dz)
{
  const size_t n = size_t(nx) * ny * nz;
  const double dtodx = dt / dx, dtody = dt / dy, dtodz = dt / dz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k < nz - 1; k++)
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        const size_t id = Idx(i, j, k, nx, ny);
        const double ex1 = E[0 * n + Idx(i, j + 1, k, nx, ny)], ex2 = E[0 * n + Idx(i, j, k + 1, nx, ny)],
                     ex3 = E[0 * n + Idx(i, j + 1, k + 1, nx, ny)];
        const double ey1 = E[1 * n + Idx(i + 1, j, k, nx, ny)], ey2 = E[1 * n + Idx(i, j, k + 1, nx, ny)],
                     ey3 = E[1 * n + Idx(i + 1, j, k + 1, nx, ny)];
        const double ez1 = E[2 * n + Idx(i + 1, j, k, nx, ny)], ez2 = E[2 * n + Idx(i, j + 1, k, nx, ny)],
                     ez3 = E[2 * n + Idx(i + 1, j + 1, k, nx, ny)];
        dst[kBx * n + id] = src[kBx * n + id] + dtodz * (ey3 - ey1) + dtody * (ez1 - ez3);
        dst[kBy * n + id] = src[kBy * n + id] + dtodx * (ez3 - ez2) + dtodz * (ex1 - ex3);
        dst[kBz * n + id] = src[kBz * n + id] + dtody * (ex3 - ex2) + dtodx * (ey2 - ey3);
      }
}

// Calc_dt_3D with MHD, src/hydro/hydro_cuda.cu:514-560: max inverse crossing time over real cells
double orc_mhd_calc_dti(const double *U, int nx, int ny, int nz, int n_ghost, double dx, double dy, double dz,
                        double g)
{
  const size_t n = size_t(nx) * ny * nz;
  double m = -std::numeric_limits<double>::max();
#pragma omp parallel for collapse(2) reduction(max : m) schedule(static)
  for (int k = n_ghost; k < nz - n_ghost; k++)
    for (int j = n_ghost; j < ny - n_ghost; j++)
      for (int i = n_ghost; i < nx - n_ghost; i++) {
        const size_t id = Idx(i, j, k, nx, ny);
        const double d = U[id], dinv = 1.0 / d;
        const double vx = U[n + id] * dinv, vy = U[2 * n + id] * dinv, vz = U[3 * n + id] * dinv;
        double bx, by, bz;
        CellCenteredB(U, id, i, j, k, n, nx, ny, bx, by, bz);
        m = std::fmax(m, orc_mhd_inverse_crossing_time(U[4 * n + id], d, dinv, vx, vy, vz, bx, by, bz, dx, dy, dz, g));
      }
  return m;
}

// mhd::calculateMagneticDivergence, src/mhd/magnetic_divergence.cu:33-75: max |div B|
double orc_mhd_divergence(const double *U, int nx, int ny, int nz, double dx, double dy, double dz)
{
  const size_t n = size_t(nx) * ny * nz;
  double m = 0.0;
#pragma omp parallel for collapse(2) reduction(max : m) schedule(static)
  for (int k = 2; k < nz; k++)
    for (int j = 2; j < ny; j++)
      for (int i = 2; i < nx; i++) {
        const size_t id = Idx(i, j, k, nx, ny);
        double div = ((U[kBx * n + id] - U[kBx * n + Idx(i - 1, j, k, nx, ny)]) / dx) +
                     ((U[kBy * n + id] - U[kBy * n + Idx(i, j - 1, k, nx, ny)]) / dy) +
                     ((U[kBz * n + id] - U[kBz * n + Idx(i, j, k - 1, nx, ny)]) / dz);
        m = std::max(m, std::fabs(div));
      }
  return m;
}

// Reconstruction of the MHD step's second stage (the build's PCM / PLMC / PPMC option):
// 0 PCM, 1 PLMC (the default), 2 PPMC.
static int g_mhd_recon = 1;
void orc_mhd_set_reconstruction(int recon) { g_mhd_recon = recon; }
int orc_mhd_get_reconstruction() { return g_mhd_recon; }

// VL_Algorithm_3D_CUDA with MHD (PCM, PLMC or PPMC + HLLD + CT), src/integrators/VL_3D_cuda.cu:38-360
void orc_mhd_vl3d(double *U, int nx, int ny, int nz, int n_ghost, double dx, double dy, double dz, double dt, double g,
                  int density_floor_on, double density_floor)
{
  const size_t n = size_t(nx) * ny * nz;
  MhdScratch &s  = g_mhd;
  s.Resize(n);
  double *Uh = s.Uh.data();
  orc_mhd_pcm_3d(U, s.QLx.data(), s.QRx.data(), s.QLy.data(), s.QRy.data(), s.QLz.data(), s.QRz.data(), nx, ny, nz);
  orc_hlld(s.QLx.data(), s.QRx.data(), U + kBx * n, s.Fx.data(), long(n), g, 0);
  orc_hlld(s.QLy.data(), s.QRy.data(), U + kBy * n, s.Fy.data(), long(n), g, 1);
  orc_hlld(s.QLz.data(), s.QRz.data(), U + kBz * n, s.Fz.data(), long(n), g, 2);
  orc_ct_electric_fields(s.Fx.data(), s.Fy.data(), s.Fz.data(), U, s.E.data(), nx, ny, nz, 8);
  orc_update_half_3d(U, Uh, s.Fx.data(), s.Fy.data(), s.Fz.data(), nx, ny, nz, dx, dy, dz, 0.5 * dt, g, 5,
                     density_floor_on, density_floor);
  orc_update_magnetic(U, Uh, s.E.data(), nx, ny, nz, 0.5 * dt, dx, dy, dz);
  if (g_mhd_recon == 0) {  // VL_3D_cuda.cu:226-229 (PCM), 239-247 (PLMC), 257-265 (PPMC)
    orc_mhd_pcm_3d(Uh, s.QLx.data(), s.QRx.data(), s.QLy.data(), s.QRy.data(), s.QLz.data(), s.QRz.data(), nx, ny,
                   nz);
  } else {
    auto recon = g_mhd_recon == 2 ? orc_mhd_ppmc : orc_mhd_plmc;
    recon(Uh, s.QLx.data(), s.QRx.data(), nx, ny, nz, g, 0);
    recon(Uh, s.QLy.data(), s.QRy.data(), nx, ny, nz, g, 1);
    recon(Uh, s.QLz.data(), s.QRz.data(), nx, ny, nz, g, 2);
  }
  orc_hlld(s.QLx.data(), s.QRx.data(), Uh + kBx * n, s.Fx.data(), long(n), g, 0);
  orc_hlld(s.QLy.data(), s.QRy.data(), Uh + kBy * n, s.Fy.data(), long(n), g, 1);
  orc_hlld(s.QLz.data(), s.QRz.data(), Uh + kBz * n, s.Fz.data(), long(n), g, 2);
  orc_ct_electric_fields(s.Fx.data(), s.Fy.data(), s.Fz.data(), Uh, s.E.data(), nx, ny, nz, 8);
  orc_update_3d(U, s.Fx.data(), s.Fy.data(), s.Fz.data(), nx, ny, nz, n_ghost, dx, dy, dz, dt, 5, density_floor_on,
                density_floor, g);
  orc_update_magnetic(U, U, s.E.data(), nx, ny, nz, dt, dx, dy, dz);
}

// The driver loop of src/main.cpp:245-379 for the MHD build (as orc_run for hydro):
// dt from the MHD inverse crossing time, VL step, temperature floor, boundaries.
int orc_mhd_run(double *U, int nx, int ny, int nz, int n_ghost, double dx, double dy, double dz, double g,
                const int *flags, double *t_io, double tout, int max_steps, int density_floor_on, double density_floor,
                int temperature_floor_on, double U_floor, double *dt_log)
{
  if (orc_get_dual_energy()) return -1;  // the MHD restatement has no dual-energy branches
  double t   = *t_io;
  double dti = orc_mhd_calc_dti(U, nx, ny, nz, n_ghost, dx, dy, dz, g);
  int steps  = 0;
  while (t < tout && steps < max_steps) {
    double dt = kCfl / dti;
    if (t + dt > tout) dt = tout - t;
    const int crash0 = orc_get_crash_repair();
    orc_set_crash_repair(!(density_floor_on && temperature_floor_on));
    orc_mhd_vl3d(U, nx, ny, nz, n_ghost, dx, dy, dz, dt, g, density_floor_on, density_floor);
    orc_set_crash_repair(crash0);
    if (temperature_floor_on) orc_temperature_floor(U, nx, ny, nz, n_ghost, U_floor);
    dti = orc_mhd_calc_dti(U, nx, ny, nz, n_ghost, dx, dy, dz, g);
    if (dt_log) dt_log[steps] = dt;
    t += dt;
    steps++;
    orc_set_boundaries_mhd(U, nx, ny, nz, 8, n_ghost, flags);
  }
  *t_io = t;
  return steps;
}

// Fills an array with std::uniform_real_distribution<double>(a, b) draws of a
// 32-bit std::mt19937(seed) engine (the generator of the reference's _ctSlope test,
// src/mhd/ct_electric_fields_tests.cu:183-194).
void orc_fill_uniform32(uint32_t seed, double a, double b, long n, double *out)
{
  std::mt19937 prng(seed);
  std::uniform_real_distribution<double> dist(a, b);
  for (long i = 0; i < n; i++) out[i] = dist(prng);
}

}  // extern "C"
