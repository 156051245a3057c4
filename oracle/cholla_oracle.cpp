// cholla_oracle.cpp -- CPU restatement of Cholla's hydro hot path (FP64).
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product
// path. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it, and only to check or to time a CPU baseline; the product path
// (cholla_amd/) never links, loads or calls it.
//
// Every stage restates the reference kernel it names (file:line under
// src/ of cholla-hydro/cholla) in the same operation order and association,
// unfused, over full field-major arrays, exactly as the reference's stage
// kernels see them.  It is compiled with -ffp-contract=off so every product
// and sum rounds as written (IEEE double, no fused multiply-add).
//
// Layout: U[f*n_cells + i + j*nx + k*nx*ny], field order density, momentum_x,
// momentum_y, momentum_z, Energy (src/grid/grid_enum.h:20-30), and GasEnergy as
// field 5 when the dual-energy option (the reference's -DDE build) is switched on
// with orc_set_dual_energy (grid_enum.h:60-72 with no scalars and no MHD).
//
// Parity pins (which golden vector or invariant checks which function) are
// listed in tests/test_oracle.py; every function here is checked by at least
// one of them.

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

namespace
{

constexpr double kTiny = 1.0e-20;  // TINY_NUMBER, src/global/global.h:25
constexpr double kCfl  = 0.3;      // C_cfl, src/grid/grid3D.cpp:157
constexpr double kEta1 = 0.001;    // DE_ETA_1 (non-cosmology), src/global/global.h:62-69
constexpr double kEta2 = 0.035;    // DE_ETA_2, src/global/global.h:71-73
constexpr int kGE      = 5;        // grid_enum::GasEnergy with DE, no scalars, no MHD

// The DE build option as a run-time switch (every stage below follows its #ifdef DE
// branches when set).
int g_de = 0;

// PPMC positivity fallback, an option of this implementation with no counterpart in
// the reference: where PPMC_VL's interface density or pressure of a cell comes out
// non-positive (the reference floors each at TINY_NUMBER independently, ppmc_cuda.cu:681-684,
// which can pair a 1e-20 density with a finite pressure and so an unbounded sound
// speed), both interfaces of that cell in that sweep take the cell-centred values.
int g_pos_fallback = 0;

// The Riemann solver build option (-DHLLC or -DHLL) as a run-time switch: 0 HLLC, 1 HLL.
int g_riemann = 0;
// gamma of the 1D step in progress (the 1D half update's DE term needs it)
static double g_gamma_1d = 5.0 / 3.0;
static double gamma_1d() { return g_gamma_1d; }

inline size_t Idx(int i, int j, int k, int nx, int ny)
{
  return size_t(i) + size_t(j) * size_t(nx) + size_t(k) * size_t(nx) * size_t(ny);
}

// math_utils::dotProduct / SquareMagnitude, src/utils/math_utilities.h:78-97
inline double SquareMagnitude(double a, double b, double c) { return a * a + ((b * b) + (c * c)); }

// SIGN macro, src/global/global.h:144
inline double SignOf(double a) { return (a < 0.) ? -1. : 1.; }

struct Prim {
  double d, vx, vy, vz, p;
  double ge = 0.0;  // specific gas energy GasEnergy/d (DE only)
};
struct Char {
  double a0, a1, a2, a3, a4;
};

// reconstruction::Thread_Guard, src/reconstruction/reconstruction.h:95-119
bool ThreadGuard(int order, int nx, int ny, int nz, int xid, int yid, int zid)
{
  bool out = xid < order - 1 || xid >= nx - order;
  if (ny > 1) out = yid < order - 1 || yid >= ny - order || out;
  if (nz > 1) {
    out = zid < order - 1 || zid >= nz - order || out;
  } else {
    out = zid >= nz || out;
  }
  return out;
}

// hydro_utilities::Calc_Pressure_Primitive, src/utils/hydro_utilities.h:34-45
inline double PressurePrimitive(double E, double d, double vx, double vy, double vz, double gamma)
{
  double pressure = E - 0.5 * d * SquareMagnitude(vx, vy, vz);
  return std::fmax((gamma - 1.) * pressure, kTiny);
}

// hydro_utilities::Calc_Energy_Primitive, src/utils/hydro_utilities.h:106-118
inline double EnergyPrimitive(double P, double d, double vx, double vy, double vz, double gamma)
{
  return (std::fmax(P, kTiny) / (gamma - 1.)) + 0.5 * d * SquareMagnitude(vx, vy, vz);
}

// hydro_utilities::Get_Pressure_From_DE, src/utils/hydro_utilities.h:136-151
inline double PressureFromDE(double E, double U_total, double U_advected, double gamma)
{
  double U = (U_total / E > kEta1) ? U_total : U_advected;
  double P = U * (gamma - 1.0);
  return std::fmax(P, kTiny);
}

// hydro_utilities::Calc_Kinetic_Energy_From_Velocity, src/utils/hydro_utilities.h:162-166
inline double KineticEnergy(double d, double vx, double vy, double vz) { return 0.5 * d * SquareMagnitude(vx, vy, vz); }

// reconstruction::Load_Data, src/reconstruction/reconstruction.h:139-212 (hydro; DE branch :179-191)
Prim LoadData(const double *U, int xid, int yid, int zid, int nx, int ny, size_t n_cells, int o1, int o2, int o3,
              double gamma)
{
  size_t id = Idx(xid, yid, zid, nx, ny);
  Prim q;
  q.d  = U[id];
  q.vx = U[o1 * n_cells + id] / q.d;
  q.vy = U[o2 * n_cells + id] / q.d;
  q.vz = U[o3 * n_cells + id] / q.d;
  if (g_de) {
    double E  = U[4 * n_cells + id];
    double ge = U[kGE * n_cells + id];
    double Ek = KineticEnergy(q.d, q.vx, q.vy, q.vz);
    q.p       = PressureFromDE(E, E - Ek, ge, gamma);
    q.ge      = ge / q.d;
  } else {
    q.p = PressurePrimitive(U[4 * n_cells + id], q.d, q.vx, q.vy, q.vz, gamma);
  }
  return q;
}

// reconstruction::Compute_Slope, reconstruction.h:224-251
Prim Slope(const Prim &l, const Prim &r, double coef = 1.0)
{
  return {coef * (r.d - l.d),   coef * (r.vx - l.vx), coef * (r.vy - l.vy),
          coef * (r.vz - l.vz), coef * (r.p - l.p),   coef * (r.ge - l.ge)};
}

// reconstruction::Van_Leer_Slope, reconstruction.h:262-296
inline double VlSlope(double l, double r) { return (l * r > 0.0) ? 2.0 * l * r / (l + r) : 0.0; }
Prim VanLeer(const Prim &l, const Prim &r)
{
  return {VlSlope(l.d, r.d),   VlSlope(l.vx, r.vx), VlSlope(l.vy, r.vy),
          VlSlope(l.vz, r.vz), VlSlope(l.p, r.p),   VlSlope(l.ge, r.ge)};
}

// reconstruction::Primitive_To_Characteristic (hydro branch), reconstruction.h:423-431
Char PrimToChar(const Prim &q, const Prim &s, double c, double c2)
{
  Char a;
  a.a0 = -q.d * s.vx / (2.0 * c) + s.p / (2.0 * c2);
  a.a1 = s.d - s.p / (c2);
  a.a2 = s.vy;
  a.a3 = s.vz;
  a.a4 = q.d * s.vx / (2.0 * c) + s.p / (2.0 * c2);
  return a;
}

// reconstruction::Characteristic_To_Primitive (hydro branch), reconstruction.h:482-488
Prim CharToPrim(const Prim &q, const Char &a, double c, double c2)
{
  Prim o;
  o.d  = a.a0 + a.a1 + a.a4;
  o.vx = c / q.d * (a.a4 - a.a0);
  o.vy = a.a2;
  o.vz = a.a3;
  o.p  = c2 * (a.a0 + a.a4);
  return o;
}

// Monotonize lambda of Monotonize_Characteristic_Return_Primitive, reconstruction.h:519-527
inline double MonotonizeOne(double left, double right, double centered, double van_leer)
{
  if (left * right > 0.0) {
    double lim_a = 2.0 * std::fmin(std::fabs(left), std::fabs(right));
    double lim_b = std::fmin(std::fabs(centered), std::fabs(van_leer));
    return std::copysign(std::fmin(lim_a, lim_b), centered);
  }
  return 0.0;
}

// reconstruction::Monotonize_Characteristic_Return_Primitive, reconstruction.h:512-559
Prim MonotonizeReturnPrim(const Prim &q, const Char &L, const Char &R, const Char &C, const Char &G, double c,
                          double c2)
{
  Char m;
  m.a0 = MonotonizeOne(L.a0, R.a0, C.a0, G.a0);
  m.a1 = MonotonizeOne(L.a1, R.a1, C.a1, G.a1);
  m.a2 = MonotonizeOne(L.a2, R.a2, C.a2, G.a2);
  m.a3 = MonotonizeOne(L.a3, R.a3, C.a3, G.a3);
  m.a4 = MonotonizeOne(L.a4, R.a4, C.a4, G.a4);
  return CharToPrim(q, m, c, c2);
}

// reconstruction::Calc_Interface_Linear, reconstruction.h:645-673
Prim InterfaceLinear(const Prim &q, const Prim &s, double sign)
{
  auto f = [&](double state, double slope) { return state + sign * 0.5 * slope; };
  return {f(q.d, s.d), f(q.vx, s.vx), f(q.vy, s.vy), f(q.vz, s.vz), f(q.p, s.p), f(q.ge, s.ge)};
}

// reconstruction::Plm_Limit_Interfaces, reconstruction.h:686-723
void PlmLimit(Prim &L, Prim &R, const Prim &im, const Prim &i, const Prim &ip)
{
  auto lim = [](double &l_iph, double &r_imh, double v_imo, double v_i, double v_ipo) {
    r_imh = std::fmax(std::fmin(v_i, v_imo), r_imh);
    r_imh = std::fmin(std::fmax(v_i, v_imo), r_imh);
    l_iph = std::fmax(std::fmin(v_i, v_ipo), l_iph);
    l_iph = std::fmin(std::fmax(v_i, v_ipo), l_iph);
  };
  lim(L.d, R.d, im.d, i.d, ip.d);
  lim(L.vx, R.vx, im.vx, i.vx, ip.vx);
  lim(L.vy, R.vy, im.vy, i.vy, ip.vy);
  lim(L.vz, R.vz, im.vz, i.vz, ip.vz);
  lim(L.p, R.p, im.p, i.p, ip.p);
  if (g_de) lim(L.ge, R.ge, im.ge, i.ge, ip.ge);
}

// reconstruction::PPM_Single_Variable, reconstruction.h:794-901
void PpmSingle(double im2, double im1, double i0, double ip1, double ip2, double &L_iph, double &R_imh)
{
  const double C2   = 1.25;
  auto limit_iface = [&](double c_i, double c_im1, double iface, double s2_i, double s2_im1) -> double {
    double s2c     = 3.0 * (c_im1 + c_i - 2.0 * iface);
    double limited = 0.0;
    if (SignOf(s2c) == SignOf(s2_im1) && SignOf(s2c) == SignOf(s2_i)) {
      limited = SignOf(s2c) * std::fmin(C2 * std::fabs(s2_im1), std::fmin(C2 * std::fabs(s2_i), std::fabs(s2c)));
    }
    double dl = iface - c_im1;
    double dr = c_i - iface;
    if (dl * dr < 0.0) return 0.5 * (c_im1 + c_i) - limited / 6.0;
    return iface;
  };

  double slope_left    = (i0 - im1);
  double slope_right   = (ip1 - i0);
  double slope_avg_im1 = 0.5 * slope_left + 0.5 * (im1 - im2);
  double slope_avg_i   = 0.5 * slope_right + 0.5 * slope_left;
  double slope_avg_ip1 = 0.5 * (ip2 - ip1) + 0.5 * slope_right;

  R_imh = 0.5 * (im1 + i0) + (slope_avg_im1 - slope_avg_i) / 6.0;
  L_iph = 0.5 * (i0 + ip1) + (slope_avg_i - slope_avg_ip1) / 6.0;

  double s2_im1 = im2 + i0 - 2.0 * im1;
  double s2_i   = im1 + ip1 - 2.0 * i0;
  double s2_ip1 = i0 + ip2 - 2.0 * ip1;

  R_imh = limit_iface(i0, im1, R_imh, s2_i, s2_im1);
  L_iph = limit_iface(ip1, i0, L_iph, s2_ip1, s2_i);

  double s2_face    = 6.0 * (R_imh + L_iph - 2.0 * i0);
  double s2_limited = 0.0;
  if (SignOf(s2_im1) == SignOf(s2_i) && SignOf(s2_im1) == SignOf(s2_ip1) && SignOf(s2_im1) == SignOf(s2_face)) {
    s2_limited = SignOf(s2_face) * std::fmin(std::fmin(C2 * std::fabs(s2_im1), C2 * std::fabs(s2_i)),
                                             std::fmin(C2 * std::fabs(s2_ip1), std::fabs(s2_face)));
  }

  double cell_max = std::fmax(std::fabs(im2), std::fabs(im1));
  cell_max        = std::fmax(cell_max, std::fabs(i0));
  cell_max        = std::fmax(cell_max, std::fabs(ip1));
  cell_max        = std::fmax(cell_max, std::fabs(ip2));

  double rho = (std::fabs(s2_face) > (1.0e-12) * cell_max) ? s2_limited / s2_face : 0.0;

  double sfl = i0 - R_imh;
  double sfr = L_iph - i0;
  if ((sfl * sfr) <= 0.0 || ((ip1 - i0) * (i0 - im1)) <= 0.0) {
    if (rho <= (1.0 - (1.0e-12))) {
      R_imh = i0 - rho * sfl;
      L_iph = i0 + rho * sfr;
    }
  } else {
    if (std::fabs(sfl) >= 2.0 * std::fabs(sfr)) R_imh = i0 - 2.0 * sfr;
    if (std::fabs(sfr) >= 2.0 * std::fabs(sfl)) L_iph = i0 + 2.0 * sfl;
  }
}

// reconstruction::Write_Data (hydro), reconstruction.h:918-968
void WriteData(const Prim &s, double *Q, size_t id, size_t n_cells, int o1, int o2, int o3, double gamma)
{
  Q[id]               = s.d;
  Q[o1 * n_cells + id] = s.d * s.vx;
  Q[o2 * n_cells + id] = s.d * s.vy;
  Q[o3 * n_cells + id] = s.d * s.vz;
  Q[4 * n_cells + id]  = EnergyPrimitive(s.p, s.d, s.vx, s.vy, s.vz, gamma);
  if (g_de) Q[kGE * n_cells + id] = s.d * s.ge;  // reconstruction.h:960-962
}

void DirOrder(int dir, int &o1, int &o2, int &o3)
{
  if (dir == 0) {
    o1 = 1, o2 = 2, o3 = 3;
  } else if (dir == 1) {
    o1 = 2, o2 = 3, o3 = 1;
  } else {
    o1 = 3, o2 = 1, o3 = 2;
  }
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Scalar helpers exposed for the known-answer tests
// ---------------------------------------------------------------------------
void orc_ppm_single_variable(double im2, double im1, double i0, double ip1, double ip2, double *L_iph, double *R_imh)
{
  PpmSingle(im2, im1, i0, ip1, ip2, *L_iph, *R_imh);
}

int orc_thread_guard(int order, int nx, int ny, int nz, int xid, int yid, int zid)
{
  return ThreadGuard(order, nx, ny, nz, xid, yid, zid) ? 1 : 0;
}

// Load_Data on a field-major grid; out = {d, vx, vy, vz, p}
void orc_load_data(const double *U, int xid, int yid, int zid, int nx, int ny, int nz, int dir, double gamma,
                   double *out)
{
  int o1, o2, o3;
  DirOrder(dir, o1, o2, o3);
  Prim q = LoadData(U, xid, yid, zid, nx, ny, size_t(nx) * ny * nz, o1, o2, o3, gamma);
  out[0] = q.d, out[1] = q.vx, out[2] = q.vy, out[3] = q.vz, out[4] = q.p;
}

// Compute_Slope / Van_Leer_Slope / Calc_Interface_Linear on 5-vectors {d, vx, vy, vz, p}
void orc_compute_slope(const double *l, const double *r, double coef, double *out)
{
  Prim s = Slope({l[0], l[1], l[2], l[3], l[4]}, {r[0], r[1], r[2], r[3], r[4]}, coef);
  out[0] = s.d, out[1] = s.vx, out[2] = s.vy, out[3] = s.vz, out[4] = s.p;
}
void orc_van_leer_slope(const double *l, const double *r, double *out)
{
  Prim s = VanLeer({l[0], l[1], l[2], l[3], l[4]}, {r[0], r[1], r[2], r[3], r[4]});
  out[0] = s.d, out[1] = s.vx, out[2] = s.vy, out[3] = s.vz, out[4] = s.p;
}
void orc_interface_linear(const double *q, const double *s, double sign, double *out)
{
  Prim o = InterfaceLinear({q[0], q[1], q[2], q[3], q[4]}, {s[0], s[1], s[2], s[3], s[4]}, sign);
  out[0] = o.d, out[1] = o.vx, out[2] = o.vy, out[3] = o.vz, out[4] = o.p;
}
// Monotonize_Characteristic_Return_Primitive (hydro); chars are 5-vectors a0..a4
void orc_monotonize_char_return_prim(const double *q, const double *aL, const double *aR, const double *aC,
                                     const double *aG, double c, double c2, double *out)
{
  auto ch = [](const double *a) { return Char{a[0], a[1], a[2], a[3], a[4]}; };
  Prim o  = MonotonizeReturnPrim({q[0], q[1], q[2], q[3], q[4]}, ch(aL), ch(aR), ch(aC), ch(aG), c, c2);
  out[0] = o.d, out[1] = o.vx, out[2] = o.vy, out[3] = o.vz, out[4] = o.p;
}
// Plm_Limit_Interfaces; L, R updated in place
void orc_plm_limit_interfaces(double *L, double *R, const double *im, const double *i0, const double *ip)
{
  Prim l{L[0], L[1], L[2], L[3], L[4]}, r{R[0], R[1], R[2], R[3], R[4]};
  PlmLimit(l, r, {im[0], im[1], im[2], im[3], im[4]}, {i0[0], i0[1], i0[2], i0[3], i0[4]},
           {ip[0], ip[1], ip[2], ip[3], ip[4]});
  L[0] = l.d, L[1] = l.vx, L[2] = l.vy, L[3] = l.vz, L[4] = l.p;
  R[0] = r.d, R[1] = r.vx, R[2] = r.vy, R[3] = r.vz, R[4] = r.p;
}
// Write_Data into a field-major interface array
void orc_write_data(const double *s, double *Q, long id, long n_cells, int dir, double gamma)
{
  int o1, o2, o3;
  DirOrder(dir, o1, o2, o3);
  WriteData({s[0], s[1], s[2], s[3], s[4]}, Q, size_t(id), size_t(n_cells), o1, o2, o3, gamma);
}

// hydroInverseCrossingTime, src/hydro/hydro_cuda.cu:404-418
double orc_hydro_inverse_crossing_time(double E, double d, double d_inv, double vx, double vy, double vz, double dx,
                                       double dy, double dz, double gamma)
{
  double P  = (E - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0);
  double cs = std::sqrt(d_inv * gamma * P);
  double m  = std::fmax((std::fabs(vx) + cs) / dx, (std::fabs(vy) + cs) / dy);
  m         = std::fmax(m, (std::fabs(vz) + cs) / dz);
  m         = std::fmax(m, 0.0);
  return m;
}

// std::mt19937_64 + std::uniform_real_distribution<double>, the generator the
// reference's unit tests draw their inputs from (e.g. plmc_cuda_tests.cu:34-48)
void orc_fill_uniform(uint64_t seed, double a, double b, long n, double *out)
{
  std::mt19937_64 prng(seed);
  std::uniform_real_distribution<double> dist(a, b);
  for (long i = 0; i < n; i++) out[i] = dist(prng);
}

// ---------------------------------------------------------------------------
// Stage kernels (one call = one reference kernel launch over the whole grid)
// ---------------------------------------------------------------------------

// PCM_Reconstruction_3D, src/reconstruction/pcm_cuda.cu:256-... (hydro)
void orc_pcm_3d(const double *U, double *QLx, double *QRx, double *QLy, double *QRy, double *QLz, double *QRz, int nx,
                int ny, int nz, int n_fields)
{
  size_t n_cells = size_t(nx) * ny * nz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        size_t id = Idx(i, j, k, nx, ny);
        for (int f = 0; f < n_fields; f++) {
          double v               = U[f * n_cells + id];
          QLx[f * n_cells + id] = v;
          QLy[f * n_cells + id] = v;
          QLz[f * n_cells + id] = v;
          if (i > 0) QRx[f * n_cells + Idx(i - 1, j, k, nx, ny)] = v;
          if (j > 0) QRy[f * n_cells + Idx(i, j - 1, k, nx, ny)] = v;
          if (k > 0) QRz[f * n_cells + Idx(i, j, k - 1, nx, ny)] = v;
        }
      }
}

// PCM_Reconstruction_1D, pcm_cuda.cu:13-97
void orc_pcm_1d(const double *U, double *QL, double *QR, int n_cells, int n_fields)
{
  for (int x = 0; x < n_cells - 1; x++)
    for (int f = 0; f < n_fields; f++) {
      QL[size_t(f) * n_cells + x] = U[size_t(f) * n_cells + x];
      QR[size_t(f) * n_cells + x] = U[size_t(f) * n_cells + x + 1];
    }
}

// PLMC_cuda (VL build: no characteristic tracing), src/reconstruction/plmc_cuda.cu:24-300
void orc_plmc(const double *U, double *QL, double *QR, int nx, int ny, int nz, double gamma, int dir)
{
  size_t n_cells = size_t(nx) * ny * nz;
  int o1, o2, o3;
  DirOrder(dir, o1, o2, o3);
  int ix = dir == 0, iy = dir == 1, iz = dir == 2;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        if (ThreadGuard(2, nx, ny, nz, i, j, k)) continue;
        Prim ci  = LoadData(U, i, j, k, nx, ny, n_cells, o1, o2, o3, gamma);
        Prim cim = LoadData(U, i - ix, j - iy, k - iz, nx, ny, n_cells, o1, o2, o3, gamma);
        Prim cip = LoadData(U, i + ix, j + iy, k + iz, nx, ny, n_cells, o1, o2, o3, gamma);
        double c  = std::sqrt(gamma * ci.p / ci.d);  // Calc_Sound_Speed, hydro_utilities.h:209-212
        double c2 = c * c;
        Prim dL   = Slope(cim, ci);
        Prim dR   = Slope(ci, cip);
        Prim dC   = Slope(cim, cip, 0.5);
        Prim dG   = VanLeer(dL, dR);
        Char aL = PrimToChar(ci, dL, c, c2), aR = PrimToChar(ci, dR, c, c2), aC = PrimToChar(ci, dC, c, c2),
             aG = PrimToChar(ci, dG, c, c2);
        Prim dm = MonotonizeReturnPrim(ci, aL, aR, aC, aG, c, c2);
        if (g_de) dm.ge = MonotonizeOne(dL.ge, dR.ge, dC.ge, dG.ge);  // reconstruction.h:549-551
        Prim L  = InterfaceLinear(ci, dm, 1.0);
        Prim R  = InterfaceLinear(ci, dm, -1.0);
        PlmLimit(L, R, cim, ci, cip);
        R.d = std::fmax(R.d, kTiny);
        L.d = std::fmax(L.d, kTiny);
        R.p = std::fmax(R.p, kTiny);
        L.p = std::fmax(L.p, kTiny);
        WriteData(L, QL, Idx(i, j, k, nx, ny), n_cells, o1, o2, o3, gamma);
        WriteData(R, QR, Idx(i - ix, j - iy, k - iz, nx, ny), n_cells, o1, o2, o3, gamma);
      }
}

// Interface_Values_PLM, src/reconstruction/plmp_cuda.cu:325-365: Van Leer limited slope of
// one primitive variable; *left and *right are the values at the cell's left and right faces.
// (the slope itself is Calculate_Slope of ppmp_cuda.cu:687-720, the same steps)
static double LimitedSlope(double qm, double q0, double qp)
{
  double dL = q0 - qm, dR = qp - q0, dC = 0.5 * (qp - qm);
  double dG = dL * dR > 0.0 ? 2.0 * dL * dR / (dL + dR) : 0.0;
  double a = std::fmin(std::fabs(dL), std::fabs(dR));
  double b = std::fmin(std::fabs(dC), std::fabs(dG));
  return (dC < 0 ? -1.0 : 1.0) * std::fmin(2.0 * a, b);  // sgn_CUDA, global_cuda.h:41-48
}
static void PlmpInterface(double qm, double q0, double qp, double *left, double *right)
{
  const double dm = LimitedSlope(qm, q0, qp);
  *left           = q0 - 0.5 * dm;
  *right          = q0 + 0.5 * dm;
}

// The primitive stencil value PLMP_cuda and PPMP_cuda load (plmp_cuda.cu:108-132,
// ppmp_cuda.cu:155-174): velocities by division, the pressure from the left-to-right sum of
// squared velocities, or under DE Get_Pressure_From_DE with that kinetic energy
// (plmp_cuda.cu:116-121), floored; q[5] the specific GasEnergy dge / d under DE.
static void LoadPlmpPrim(const double *U, size_t id, size_t n_cells, int o1, int o2, int o3, double gamma, double q[6])
{
  double d = U[id], vx = U[o1 * n_cells + id] / d, vy = U[o2 * n_cells + id] / d, vz = U[o3 * n_cells + id] / d;
  double p;
  if (g_de) {
    double E = U[4 * n_cells + id], E_kin = 0.5 * d * (vx * vx + vy * vy + vz * vz);
    double dge = U[kGE * n_cells + id];
    p          = PressureFromDE(E, E - E_kin, dge, gamma);
    q[5]       = dge / d;
  } else {
    p = (U[4 * n_cells + id] - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0);
  }
  q[0] = d, q[1] = vx, q[2] = vy, q[3] = vz, q[4] = std::fmax(p, kTiny);
}

// The interface value PLMP_cuda and PPMP_cuda write (plmp_cuda.cu:217-230, ppmp_cuda.cu:640-683):
// conserved variables from the floored primitives, the GasEnergy as d * ge under DE.
static void WritePlmpIface(const double q[6], double *Q, size_t id, size_t n_cells, int o1, int o2, int o3,
                           double gamma)
{
  Q[id]                = q[0];
  Q[o1 * n_cells + id] = q[0] * q[1];
  Q[o2 * n_cells + id] = q[0] * q[2];
  Q[o3 * n_cells + id] = q[0] * q[3];
  Q[4 * n_cells + id]  = q[4] / (gamma - 1.0) + 0.5 * q[0] * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (g_de) Q[kGE * n_cells + id] = q[0] * q[5];
}

// PLMP_cuda in the VL build (no half-step evolution, plmp_cuda.cu:231-284 compiled out),
// src/reconstruction/plmp_cuda.cu:20-320, hydro without scalars: primitives as LoadPlmpPrim
// forms them, each limited on its own (the specific GasEnergy too under DE,
// plmp_cuda.cu:204-206), the density and pressure floors, then conserved variables.
void orc_plmp(const double *U, double *QL, double *QR, int nx, int ny, int nz, double gamma, int dir)
{
  size_t n_cells = size_t(nx) * ny * nz;
  int o1, o2, o3;
  DirOrder(dir, o1, o2, o3);
  int ix = dir == 0, iy = dir == 1, iz = dir == 2;
  const int nv = g_de ? 6 : 5;
  // plmp_cuda.cu:80-103: one cell off each end along dir, the full range across it
  int xs = ix, xe = ix ? nx - 2 : nx, ys = iy, ye = iy ? ny - 2 : ny, zs = iz, ze = iz ? nz - 2 : nz;
  auto load = [&](int i, int j, int k, double q[6]) {
    LoadPlmpPrim(U, Idx(i, j, k, nx, ny), n_cells, o1, o2, o3, gamma, q);
  };
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = zs; k < ze; k++)
    for (int j = ys; j < ye; j++)
      for (int i = xs; i < xe; i++) {
        double qm[6], q0[6], qp[6], lf[6], rf[6];
        load(i - ix, j - iy, k - iz, qm);
        load(i, j, k, q0);
        load(i + ix, j + iy, k + iz, qp);
        for (int v = 0; v < nv; v++) PlmpInterface(qm[v], q0[v], qp[v], &lf[v], &rf[v]);
        lf[0] = std::fmax(lf[0], kTiny), rf[0] = std::fmax(rf[0], kTiny);
        lf[4] = std::fmax(lf[4], kTiny), rf[4] = std::fmax(rf[4], kTiny);
        WritePlmpIface(lf, QR, Idx(i - ix, j - iy, k - iz, nx, ny), n_cells, o1, o2, o3, gamma);  // face i-1/2
        WritePlmpIface(rf, QL, Idx(i, j, k, nx, ny), n_cells, o1, o2, o3, gamma);                 // face i+1/2
      }
}

// Interface_Values_PPM, src/reconstruction/ppmp_cuda.cu:722-755: parabolic interfaces from
// the limited slopes of cells i-1, i, i+1, then the monotonicity constraints in order
static void PpmpInterfaces(double qm, double q0, double qp, double sm, double s0, double sp, double *left,
                           double *right)
{
  double L = 0.5 * (q0 + qm) - (1.0 / 6.0) * (s0 - sm);
  double R = 0.5 * (qp + q0) - (1.0 / 6.0) * (sp - s0);
  if ((R - q0) * (q0 - L) <= 0) L = R = q0;
  if (6.0 * (R - L) * (q0 - 0.5 * (L + R)) > (R - L) * (R - L)) L = 3.0 * q0 - 2.0 * R;
  if (6.0 * (R - L) * (q0 - 0.5 * (L + R)) < -(R - L) * (R - L)) R = 3.0 * q0 - 2.0 * L;
  L = std::fmax(std::fmin(q0, qm), L);
  L = std::fmin(std::fmax(q0, qm), L);
  R = std::fmax(std::fmin(q0, qp), R);
  R = std::fmin(std::fmax(q0, qp), R);
  *left = L, *right = R;
}

// PPMP_cuda in the VL build (ppmp_cuda.cu:28-683; STEEPENING and FLATTENING are not defined
// there and the CTU tracing at :474-640 is compiled out), hydro without scalars: primitives
// as LoadPlmpPrim forms them, Calculate_Slope (:687-720, the PLMP slope) on cells i-1, i,
// i+1, parabolic interfaces (the specific GasEnergy too under DE, :353-361), floors,
// conserved variables.
void orc_ppmp(const double *U, double *QL, double *QR, int nx, int ny, int nz, double gamma, int dir)
{
  size_t n_cells = size_t(nx) * ny * nz;
  int o1, o2, o3;
  DirOrder(dir, o1, o2, o3);
  int ix = dir == 0, iy = dir == 1, iz = dir == 2;
  const int nv = g_de ? 6 : 5;
  // ppmp_cuda.cu:125-149: two cells off the low end and three off the high end along dir
  int xs = 2 * ix, xe = ix ? nx - 3 : nx, ys = 2 * iy, ye = iy ? ny - 3 : ny, zs = 2 * iz, ze = iz ? nz - 3 : nz;
  auto slope = LimitedSlope;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = zs; k < ze; k++)
    for (int j = ys; j < ye; j++)
      for (int i = xs; i < xe; i++) {
        double c[5][6], lf[6], rf[6];
        for (int m = 0; m < 5; m++)
          LoadPlmpPrim(U, Idx(i + (m - 2) * ix, j + (m - 2) * iy, k + (m - 2) * iz, nx, ny), n_cells, o1, o2, o3,
                       gamma, c[m]);
        for (int v = 0; v < nv; v++) {
          double sm = slope(c[0][v], c[1][v], c[2][v]), s0 = slope(c[1][v], c[2][v], c[3][v]);
          double sp = slope(c[2][v], c[3][v], c[4][v]);
          PpmpInterfaces(c[1][v], c[2][v], c[3][v], sm, s0, sp, &lf[v], &rf[v]);
        }
        lf[0] = std::fmax(lf[0], kTiny), rf[0] = std::fmax(rf[0], kTiny);
        lf[4] = std::fmax(lf[4], kTiny), rf[4] = std::fmax(rf[4], kTiny);
        WritePlmpIface(lf, QR, Idx(i - ix, j - iy, k - iz, nx, ny), n_cells, o1, o2, o3, gamma);
        WritePlmpIface(rf, QL, Idx(i, j, k, nx, ny), n_cells, o1, o2, o3, gamma);
      }
}

// PPMC_VL, src/reconstruction/ppmc_cuda.cu:540-695 (hydro, no scalars; DE branch :668-672)
void orc_ppmc_vl(const double *U, double *QL, double *QR, int nx, int ny, int nz, double gamma, int dir)
{
  size_t n_cells = size_t(nx) * ny * nz;
  int o1, o2, o3;
  DirOrder(dir, o1, o2, o3);
  int ix = dir == 0, iy = dir == 1, iz = dir == 2;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        if (ThreadGuard(3, nx, ny, nz, i, j, k)) continue;
        Prim c0  = LoadData(U, i, j, k, nx, ny, n_cells, o1, o2, o3, gamma);
        Prim cm1 = LoadData(U, i - ix, j - iy, k - iz, nx, ny, n_cells, o1, o2, o3, gamma);
        Prim cp1 = LoadData(U, i + ix, j + iy, k + iz, nx, ny, n_cells, o1, o2, o3, gamma);
        Prim cm2 = LoadData(U, i - 2 * ix, j - 2 * iy, k - 2 * iz, nx, ny, n_cells, o1, o2, o3, gamma);
        Prim cp2 = LoadData(U, i + 2 * ix, j + 2 * iy, k + 2 * iz, nx, ny, n_cells, o1, o2, o3, gamma);
        double c  = std::sqrt(gamma * c0.p / c0.d);
        double c2 = c * c;
        Char a0 = PrimToChar(c0, c0, c, c2), am1 = PrimToChar(c0, cm1, c, c2), am2 = PrimToChar(c0, cm2, c, c2),
             ap1 = PrimToChar(c0, cp1, c, c2), ap2 = PrimToChar(c0, cp2, c, c2);
        Char L, R;
        PpmSingle(am2.a0, am1.a0, a0.a0, ap1.a0, ap2.a0, L.a0, R.a0);
        PpmSingle(am2.a1, am1.a1, a0.a1, ap1.a1, ap2.a1, L.a1, R.a1);
        PpmSingle(am2.a2, am1.a2, a0.a2, ap1.a2, ap2.a2, L.a2, R.a2);
        PpmSingle(am2.a3, am1.a3, a0.a3, ap1.a3, ap2.a3, L.a3, R.a3);
        PpmSingle(am2.a4, am1.a4, a0.a4, ap1.a4, ap2.a4, L.a4, R.a4);
        Prim pL = CharToPrim(c0, L, c, c2);
        Prim pR = CharToPrim(c0, R, c, c2);
        if (g_de) PpmSingle(cm2.ge, cm1.ge, c0.ge, cp1.ge, cp2.ge, pL.ge, pR.ge);
        if (g_pos_fallback && !(pL.d > 0. && pR.d > 0. && pL.p > 0. && pR.p > 0.)) pL = pR = c0;
        pR.d    = std::fmax(pR.d, kTiny);
        pL.d    = std::fmax(pL.d, kTiny);
        pR.p    = std::fmax(pR.p, kTiny);
        pL.p    = std::fmax(pL.p, kTiny);
        WriteData(pL, QL, Idx(i, j, k, nx, ny), n_cells, o1, o2, o3, gamma);
        WriteData(pR, QR, Idx(i - ix, j - iy, k - iz, nx, ny), n_cells, o1, o2, o3, gamma);
      }
}

// Calculate_HLLC_Fluxes_CUDA, src/riemann_solvers/hllc_cuda.cu:19-307 (no scalars; DE
// branches :85-138, :183-198, :218-235, :251-303)
void orc_hllc(const double *QL, const double *QR, double *F, int nx, int ny, int nz, double gamma, int dir)
{
  size_t n_cells = size_t(nx) * ny * nz;
  int o1, o2, o3;
  DirOrder(dir, o1, o2, o3);
  const double g1 = gamma - 1.0;
#pragma omp parallel for schedule(static)
  for (long t = 0; t < long(n_cells); t++) {
    double dl = QL[t], mxl = QL[o1 * n_cells + t], myl = QL[o2 * n_cells + t], mzl = QL[o3 * n_cells + t],
           El = QL[4 * n_cells + t];
    double dr = QR[t], mxr = QR[o1 * n_cells + t], myr = QR[o2 * n_cells + t], mzr = QR[o3 * n_cells + t],
           Er = QR[4 * n_cells + t];

    double dgel = g_de ? QL[kGE * n_cells + t] : 0.0, dger = g_de ? QR[kGE * n_cells + t] : 0.0;

    double vxl = mxl / dl, vyl = myl / dl, vzl = mzl / dl;
    double pl;
    if (g_de) {
      double E_kin = 0.5 * dl * (vxl * vxl + vyl * vyl + vzl * vzl);
      pl           = PressureFromDE(El, El - E_kin, dgel, gamma);
    } else {
      pl = (El - 0.5 * dl * (vxl * vxl + vyl * vyl + vzl * vzl)) * (gamma - 1.0);
    }
    pl         = std::fmax(pl, kTiny);
    double gel = g_de ? dgel / dl : 0.0;
    double vxr = mxr / dr, vyr = myr / dr, vzr = mzr / dr;
    double pr;
    if (g_de) {
      double E_kin = 0.5 * dr * (vxr * vxr + vyr * vyr + vzr * vzr);
      pr           = PressureFromDE(Er, Er - E_kin, dger, gamma);
    } else {
      pr = (Er - 0.5 * dr * (vxr * vxr + vyr * vyr + vzr * vzr)) * (gamma - 1.0);
    }
    pr         = std::fmax(pr, kTiny);
    double ger = g_de ? dger / dr : 0.0;

    double Hl = (El + pl) / dl;
    double Hr = (Er + pr) / dr;

    double sqrtdl = std::sqrt(dl), sqrtdr = std::sqrt(dr);
    double vx  = (sqrtdl * vxl + sqrtdr * vxr) / (sqrtdl + sqrtdr);
    double vy  = (sqrtdl * vyl + sqrtdr * vyr) / (sqrtdl + sqrtdr);
    double vz  = (sqrtdl * vzl + sqrtdr * vzr) / (sqrtdl + sqrtdr);
    double H   = (sqrtdl * Hl + sqrtdr * Hr) / (sqrtdl + sqrtdr);
    double vsq = (vx * vx + vy * vy + vz * vz);
    double asq = g1 * (H - 0.5 * vsq);
    double a   = std::sqrt(asq);
    double lambda_m = vx - a, lambda_p = vx + a;
    double cfl = std::sqrt(gamma * pl / dl), cfr = std::sqrt(gamma * pr / dr);
    double Sl = std::fmin(lambda_m, vxl - cfl);
    double Sr = std::fmax(lambda_p, vxr + cfr);
    // H-correction hook with etah = 0 (hllc_cuda.cu:52,174-175): sgn(S)*fmax(|S|, 0)
    Sl = (Sl < 0 ? -1.0 : 1.0) * std::fmax(std::fabs(Sl), 0.0);
    Sr = (Sr < 0 ? -1.0 : 1.0) * std::fmax(std::fabs(Sr), 0.0);

    double f_d_l = mxl, f_mx_l = mxl * vxl + pl, f_my_l = myl * vxl, f_mz_l = mzl * vxl, f_E_l = (El + pl) * vxl;
    double f_d_r = mxr, f_mx_r = mxr * vxr + pr, f_my_r = myr * vxr, f_mz_r = mzr * vxr, f_E_r = (Er + pr) * vxr;

    double f_ge_l = dgel * vxl, f_ge_r = dger * vxr;

    double fd, fmx, fmy, fmz, fE, fge = 0.0;
    if (Sl > 0.0) {
      fd = f_d_l, fmx = f_mx_l, fmy = f_my_l, fmz = f_mz_l, fE = f_E_l, fge = f_ge_l;
    } else if (Sr < 0.0) {
      fd = f_d_r, fmx = f_mx_r, fmy = f_my_r, fmz = f_mz_r, fE = f_E_r, fge = f_ge_r;
    } else {
      double Sm = (dr * vxr * (Sr - vxr) - dl * vxl * (Sl - vxl) + pl - pr) / (dr * (Sr - vxr) - dl * (Sl - vxl));
      double ps = dl * (vxl - Sl) * (vxl - Sm) + pl;
      double dls  = dl * (Sl - vxl) / (Sl - Sm);
      double mxls = (mxl * (Sl - vxl) + ps - pl) / (Sl - Sm);
      double myls = dls * vyl, mzls = dls * vzl;
      double Els  = (El * (Sl - vxl) - pl * vxl + ps * Sm) / (Sl - Sm);
      double drs  = dr * (Sr - vxr) / (Sr - Sm);
      double mxrs = (mxr * (Sr - vxr) + ps - pr) / (Sr - Sm);
      double myrs = drs * vyr, mzrs = drs * vzr;
      double Ers  = (Er * (Sr - vxr) - pr * vxr + ps * Sm) / (Sr - Sm);
      double aSm  = std::fabs(Sm);
      fd  = 0.5 * (f_d_l + f_d_r + (Sr - aSm) * drs + (Sl + aSm) * dls - Sl * dl - Sr * dr);
      fmx = 0.5 * (f_mx_l + f_mx_r + (Sr - aSm) * mxrs + (Sl + aSm) * mxls - Sl * mxl - Sr * mxr);
      fmy = 0.5 * (f_my_l + f_my_r + (Sr - aSm) * myrs + (Sl + aSm) * myls - Sl * myl - Sr * myr);
      fmz = 0.5 * (f_mz_l + f_mz_r + (Sr - aSm) * mzrs + (Sl + aSm) * mzls - Sl * mzl - Sr * mzr);
      fE  = 0.5 * (f_E_l + f_E_r + (Sr - aSm) * Ers + (Sl + aSm) * Els - Sl * El - Sr * Er);
      if (g_de) {
        double gels = dls * gel, gers = drs * ger;
        fge         = 0.5 * (f_ge_l + f_ge_r + (Sr - aSm) * gers + (Sl + aSm) * gels - Sl * dgel - Sr * dger);
      }
    }
    F[t]               = fd;
    F[o1 * n_cells + t] = fmx;
    F[o2 * n_cells + t] = fmy;
    F[o3 * n_cells + t] = fmz;
    F[4 * n_cells + t]  = fE;
    if (g_de) F[kGE * n_cells + t] = fge;
  }
}

// Calculate_HLL_Fluxes_CUDA, src/riemann_solvers/hll_cuda.cu:19-268 (no scalars; DE
// branches :87-89, :100-112, :118-130, :176-178, :191-193, :206-208, :221-223, :235-237,
// :253-255): Davis wave-speed estimates from the two sides' sound speeds, no Roe average.
void orc_hll(const double *QL, const double *QR, double *F, int nx, int ny, int nz, double gamma, int dir)
{
  size_t n_cells = size_t(nx) * ny * nz;
  int o1, o2, o3;
  DirOrder(dir, o1, o2, o3);
#pragma omp parallel for schedule(static)
  for (long t = 0; t < long(n_cells); t++) {
    double dl = QL[t], mxl = QL[o1 * n_cells + t], myl = QL[o2 * n_cells + t], mzl = QL[o3 * n_cells + t],
           El = QL[4 * n_cells + t];
    double dr = QR[t], mxr = QR[o1 * n_cells + t], myr = QR[o2 * n_cells + t], mzr = QR[o3 * n_cells + t],
           Er = QR[4 * n_cells + t];
    double dgel = g_de ? QL[kGE * n_cells + t] : 0.0, dger = g_de ? QR[kGE * n_cells + t] : 0.0;

    double vxl = mxl / dl, vyl = myl / dl, vzl = mzl / dl;
    double pl;
    if (g_de) {
      double E_kin = 0.5 * dl * (vxl * vxl + vyl * vyl + vzl * vzl);
      pl           = PressureFromDE(El, El - E_kin, dgel, gamma);
    } else {
      pl = (El - 0.5 * dl * (vxl * vxl + vyl * vyl + vzl * vzl)) * (gamma - 1.0);
    }
    pl         = std::fmax(pl, kTiny);
    double vxr = mxr / dr, vyr = myr / dr, vzr = mzr / dr;
    double pr;
    if (g_de) {
      double E_kin = 0.5 * dr * (vxr * vxr + vyr * vyr + vzr * vzr);
      pr           = PressureFromDE(Er, Er - E_kin, dger, gamma);
    } else {
      pr = (Er - 0.5 * dr * (vxr * vxr + vyr * vyr + vzr * vzr)) * (gamma - 1.0);
    }
    pr = std::fmax(pr, kTiny);

    double cfl = std::sqrt(gamma * pl / dl), cfr = std::sqrt(gamma * pr / dr);
    double Sl = std::fmin(vxr - cfr, vxl - cfl);
    double Sr = std::fmax(vxl + cfl, vxr + cfr);

    double f_d_l = mxl, f_mx_l = mxl * vxl + pl, f_my_l = myl * vxl, f_mz_l = mzl * vxl, f_E_l = (El + pl) * vxl;
    double f_d_r = mxr, f_mx_r = mxr * vxr + pr, f_my_r = myr * vxr, f_mz_r = mzr * vxr, f_E_r = (Er + pr) * vxr;
    double f_ge_l = dgel * vxl, f_ge_r = dger * vxr;

    double fd, fmx, fmy, fmz, fE, fge;
    if (Sl > 0.0) {
      fd = f_d_l, fmx = f_mx_l, fmy = f_my_l, fmz = f_mz_l, fE = f_E_l, fge = f_ge_l;
    } else if (Sr < 0.0) {
      fd = f_d_r, fmx = f_mx_r, fmy = f_my_r, fmz = f_mz_r, fE = f_E_r, fge = f_ge_r;
    } else {
      fd  = ((Sr * f_d_l) - (Sl * f_d_r) + Sl * Sr * (dr - dl)) / (Sr - Sl);
      fmx = ((Sr * f_mx_l) - (Sl * f_mx_r) + Sl * Sr * (mxr - mxl)) / (Sr - Sl);
      fmy = ((Sr * f_my_l) - (Sl * f_my_r) + Sl * Sr * (myr - myl)) / (Sr - Sl);
      fmz = ((Sr * f_mz_l) - (Sl * f_mz_r) + Sl * Sr * (mzr - mzl)) / (Sr - Sl);
      fE  = ((Sr * f_E_l) - (Sl * f_E_r) + Sl * Sr * (Er - El)) / (Sr - Sl);
      fge = ((Sr * f_ge_l) - (Sl * f_ge_r) + Sl * Sr * (dger - dgel)) / (Sr - Sl);
    }
    F[t]               = fd;
    F[o1 * n_cells + t] = fmx;
    F[o2 * n_cells + t] = fmy;
    F[o3 * n_cells + t] = fmz;
    F[4 * n_cells + t]  = fE;
    if (g_de) F[kGE * n_cells + t] = fge;
  }
}

// Calculate_Roe_Fluxes_CUDA, src/riemann_solvers/roe_cuda.cu:19-400 (no scalars; etah is 0
// there, so the eigenvalue floor is |lambda|): the upwind flux where the Roe-averaged
// eigenvalues are one-signed, else the Roe flux, or HLLE where an intermediate state of the
// linearised solution has non-positive density or negative pressure.
void orc_roe(const double *QL, const double *QR, double *F, int nx, int ny, int nz, double gamma, int dir)
{
  size_t n_cells = size_t(nx) * ny * nz;
  int o1, o2, o3;
  DirOrder(dir, o1, o2, o3);
  const double g1 = gamma - 1.0;
#pragma omp parallel for schedule(static)
  for (long t = 0; t < long(n_cells); t++) {
    const double dl = QL[t], mxl = QL[o1 * n_cells + t], myl = QL[o2 * n_cells + t], mzl = QL[o3 * n_cells + t],
                 El = QL[4 * n_cells + t];
    const double dr = QR[t], mxr = QR[o1 * n_cells + t], myr = QR[o2 * n_cells + t], mzr = QR[o3 * n_cells + t],
                 Er = QR[4 * n_cells + t];
    const double dgel = g_de ? QL[kGE * n_cells + t] : 0.0, dger = g_de ? QR[kGE * n_cells + t] : 0.0;
    // primitives (roe_cuda.cu:118-160)
    const double vxl = mxl / dl, vyl = myl / dl, vzl = mzl / dl;
    const double vxr = mxr / dr, vyr = myr / dr, vzr = mzr / dr;
    double pl, pr;
    if (g_de) {
      pl = PressureFromDE(El, El - 0.5 * dl * (vxl * vxl + vyl * vyl + vzl * vzl), dgel, gamma);
      pr = PressureFromDE(Er, Er - 0.5 * dr * (vxr * vxr + vyr * vyr + vzr * vzr), dger, gamma);
    } else {
      pl = (El - 0.5 * dl * (vxl * vxl + vyl * vyl + vzl * vzl)) * (gamma - 1.0);
      pr = (Er - 0.5 * dr * (vxr * vxr + vyr * vyr + vzr * vzr)) * (gamma - 1.0);
    }
    pl = std::fmax(pl, kTiny), pr = std::fmax(pr, kTiny);
    const double gel = dgel / dl, ger = dger / dr;
    // Roe averages (roe_cuda.cu:163-184; Toro 11.118)
    const double Hl = (El + pl) / dl, Hr = (Er + pr) / dr;
    const double sl = std::sqrt(dl), sr = std::sqrt(dr);
    const double vx = (sl * vxl + sr * vxr) / (sl + sr), vy = (sl * vyl + sr * vyr) / (sl + sr);
    const double vz = (sl * vzl + sr * vzr) / (sl + sr), H = (sl * Hl + sr * Hr) / (sl + sr);
    const double vsq = (vx * vx + vy * vy + vz * vz);
    const double asq = g1 * std::fmax((H - 0.5 * vsq), kTiny);
    const double a   = std::sqrt(asq);
    const double lm = vx - a, l0 = vx, lp = vx + a;
    double fl[6] = {mxl, mxl * vxl + pl, mxl * vyl, mxl * vzl, (El + pl) * vxl, mxl * gel};
    double fr[6] = {mxr, mxr * vxr + pr, mxr * vyr, mxr * vzr, (Er + pr) * vxr, mxr * ger};
    double out[6];
    if (lm >= 0.0) {  // roe_cuda.cu:216-247
      for (int q = 0; q < 6; q++) out[q] = fl[q];
    } else if (lp <= 0.0) {
      for (int q = 0; q < 6; q++) out[q] = fr[q];
    } else {
      const double dd = dr - dl, dmx = mxr - mxl, dmy = myr - myl, dmz = mzr - mzl, dE = Er - El;
      const double Na = 0.5 / asq;
      double sum[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      // projections on the left eigenvectors, each weighted by |lambda| (roe_cuda.cu:261-309)
      const double a0 = dd * Na * (0.5 * g1 * vsq + vx * a) - dmx * Na * (g1 * vx + a) - dmy * Na * g1 * vy -
                        dmz * Na * g1 * vz + dE * Na * g1;
      double c = a0 * std::fabs(lm);
      sum[0] += c, sum[1] += c * (vx - a), sum[2] += c * vy, sum[3] += c * vz, sum[4] += c * (H - vx * a);
      const double a1 = -dd * vy + dmy;
      c               = a1 * std::fabs(l0);
      sum[2] += c, sum[4] += c * vy;
      const double a2 = -dd * vz + dmz;
      c               = a2 * std::fabs(l0);
      sum[3] += c, sum[4] += c * vz;
      const double a3 = dd * (1.0 - Na * g1 * vsq) + dmx * g1 * vx / asq + dmy * g1 * vy / asq + dmz * g1 * vz / asq -
                        dE * g1 / asq;
      c = a3 * std::fabs(l0);
      sum[0] += c, sum[1] += c * vx, sum[2] += c * vy, sum[3] += c * vz, sum[4] += c * 0.5 * vsq;
      const double a4 = dd * Na * (0.5 * g1 * vsq - vx * a) - dmx * Na * (g1 * vx - a) - dmy * Na * g1 * vy -
                        dmz * Na * g1 * vz + dE * Na * g1;
      c = a4 * std::fabs(lp);
      sum[0] += c, sum[1] += c * (vx + a), sum[2] += c * vy, sum[3] += c * vz, sum[4] += c * (H + vx * a);
      // positivity of the two intermediate states (roe_cuda.cu:311-343).  The second test
      // adds the fourth wave's coefficient to the density only (roe_cuda.cu:330-334); kept
      // as the reference has it, so the switch is not mirror-symmetric.
      int hlle = 0;
      double t0 = dl + a0, t1 = mxl + a0 * (vx - a), t2 = myl + a0 * vy, t3 = mzl + a0 * vz,
             t4 = El + a0 * (H - vx * a);
      if (l0 > lm) {
        if (t0 <= 0.0) hlle = 1;
        if (t4 - 0.5 * (t1 * t1 + t2 * t2 + t3 * t3) / t0 < 0.0) hlle = 2;
      }
      t0 += a3 + a4, t1 += a3 * vx, t2 += a1 + a3 * vy, t3 += a2 + a3 * vz, t4 += a1 * vy + a2 * vz + a3 * 0.5 * vsq;
      if (lp > l0) {
        if (t0 <= 0.0) hlle = 1;
        if (t4 - 0.5 * (t1 * t1 + t2 * t2 + t3 * t3) / t0 < 0.0) hlle = 2;
      }
      if (hlle != 0) {  // HLLE (roe_cuda.cu:346-379)
        const double cfl = std::sqrt(gamma * pl / dl), cfr = std::sqrt(gamma * pr / dr);
        const double bm = std::fmin(std::fmin(lm, vxl - cfl), 0.0), bp = std::fmax(std::fmax(lp, vxr + cfr), 0.0);
        const double hl[6] = {mxl - bm * dl,           mxl * (vxl - bm) + pl, myl * (vxl - bm),
                              mzl * (vxl - bm),        El * (vxl - bm) + pl * vxl, dgel * (vxl - bm)};
        const double hr[6] = {mxr - bp * dr,           mxr * (vxr - bp) + pr, myr * (vxr - bp),
                              mzr * (vxr - bp),        Er * (vxr - bp) + pr * vxr, dger * (vxr - bp)};
        const double tmp = 0.5 * (bp + bm) / (bp - bm);
        for (int q = 0; q < 6; q++) out[q] = 0.5 * (hl[q] + hr[q]) + (hl[q] - hr[q]) * tmp;
      } else {  // roe_cuda.cu:381-397
        for (int q = 0; q < 5; q++) out[q] = 0.5 * (fl[q] + fr[q] - sum[q]);
        out[5] = out[0] >= 0.0 ? out[0] * gel : out[0] * ger;
      }
    }
    F[t]                = out[0];
    F[o1 * n_cells + t] = out[1];
    F[o2 * n_cells + t] = out[2];
    F[o3 * n_cells + t] = out[3];
    F[4 * n_cells + t]  = out[4];
    if (g_de) F[kGE * n_cells + t] = out[5];
  }
}

// powf as exact_cuda.cu calls it on doubles: both arguments narrowed to float and an FP32
// result.  Evaluated as the correctly rounded float of the double pow of the narrowed
// arguments, so any conforming powf differs from it by float rounding only.
static double PowF(double x, double y) { return double(float(std::pow(double(float(x)), double(float(y))))); }

// guessp_CUDA, exact_cuda.cu:137-158: the PVRS guess, then the two-shock guess
static double ExactGuess(double dl, double vxl, double pl, double cl, double dr, double vxr, double pr, double cr,
                         double gamma)
{
  const double TOL = 1.0e-6;
  double ppv       = 0.5 * (pl + pr) + 0.125 * (vxl - vxr) * (dl + dr) * (cl + cr);
  if (ppv < 0.0) ppv = 0.0;
  double gl = std::sqrt((2.0 / ((gamma + 1.0) * dl)) / (((gamma - 1.0) / (gamma + 1.0)) * pl + ppv));
  double gr = std::sqrt((2.0 / ((gamma + 1.0) * dr)) / (((gamma - 1.0) / (gamma + 1.0)) * pr + ppv));
  double pm = (gl * pl + gr * pr - (vxr - vxl)) / (gl + gr);
  return pm < 0.0 ? TOL : pm;
}

// prefun_CUDA, exact_cuda.cu:160-178: the pressure function of one side and its derivative
static void ExactSide(double *f, double *fd, double p, double dk, double pk, double ck, double gamma)
{
  if (p <= pk) {
    *f  = (2.0 / (gamma - 1.0)) * ck * (PowF(p / pk, (gamma - 1.0) / (2.0 * gamma)) - 1.0);
    *fd = (1.0 / (dk * ck)) * PowF((p / pk), -((gamma + 1.0) / (2.0 * gamma)));
  } else {
    double qrt = std::sqrt(((2.0 / (gamma + 1.0)) / dk) / ((((gamma - 1.0) / (gamma + 1.0)) * pk) + p));
    *f         = (p - pk) * qrt;
    *fd        = (1.0 - 0.5 * (p - pk) / ((((gamma - 1.0) / (gamma + 1.0)) * pk) + p)) * qrt;
  }
}

// starpv_CUDA, exact_cuda.cu:180-215: Newton iteration for the star pressure (at most 21
// updates, relative change <= 1e-6), then the star velocity
static void ExactStar(double *p, double *v, double dl, double vxl, double pl, double cl, double dr, double vxr,
                      double pr, double cr, double gamma)
{
  const int nriter = 20;
  const double TOL = 1.0e-6;
  double fl = 0, fld = 0, fr = 0, frd = 0;
  double pold = ExactGuess(dl, vxl, pl, cl, dr, vxr, pr, cr, gamma);
  for (int i = 0; i <= nriter; i++) {
    ExactSide(&fl, &fld, pold, dl, pl, cl, gamma);
    ExactSide(&fr, &frd, pold, dr, pr, cr, gamma);
    *p            = pold - (fl + fr + vxr - vxl) / (fld + frd);
    double change = 2.0 * std::fabs((*p - pold) / (*p + pold));
    if (change <= TOL) break;
    if (*p < 0.0) *p = TOL;
    pold = *p;
  }
  *v = 0.5 * (vxl + vxr + fr - fl);
}

// sample_CUDA, exact_cuda.cu:217-334: the solution on the interface (x/t = 0)
static void ExactSample(double pm, double vm, double *d, double *v, double *p, double dl, double vxl, double pl,
                        double cl, double dr, double vxr, double pr, double cr, double gamma)
{
  const double gm1 = gamma - 1.0, gp1 = gamma + 1.0;
  if (vm >= 0) {
    if (pm <= pl) {  // left rarefaction
      if (vxl - cl >= 0) {
        *d = dl, *v = vxl, *p = pl;
      } else if (vm - cl * PowF(pm / pl, gm1 / (2.0 * gamma)) < 0) {
        *d = dl * PowF(pm / pl, 1.0 / gamma), *v = vm, *p = pm;
      } else {
        double c = (2.0 / gp1) * (cl + (gm1 / 2.0) * vxl);
        *v       = c;
        *d       = dl * PowF(c / cl, 2.0 / gm1);
        *p       = pl * PowF(c / cl, 2.0 * gamma / gm1);
      }
    } else {  // left shock
      double sl = vxl - cl * std::sqrt((gp1 / (2.0 * gamma)) * (pm / pl) + (gm1 / (2.0 * gamma)));
      if (sl >= 0) {
        *d = dl, *v = vxl, *p = pl;
      } else {
        *d = dl * (pm / pl + (gm1 / gp1)) / ((pm / pl) * (gm1 / gp1) + 1.0), *v = vm, *p = pm;
      }
    }
  } else {
    if (pm > pr) {  // right shock
      double sr = vxr + cr * std::sqrt((gp1 / (2.0 * gamma)) * (pm / pr) + (gm1 / (2.0 * gamma)));
      if (sr <= 0) {
        *d = dr, *v = vxr, *p = pr;
      } else {
        *d = dr * (pm / pr + (gm1 / gp1)) / ((pm / pr) * (gm1 / gp1) + 1.0), *v = vm, *p = pm;
      }
    } else {  // right rarefaction
      if (vxr + cr <= 0) {
        *d = dr, *v = vxr, *p = pr;
      } else if (vm + cr * PowF(pm / pr, gm1 / (2.0 * gamma)) >= 0) {
        *d = dr * PowF(pm / pr, (1.0 / gamma)), *v = vm, *p = pm;
      } else {
        double c = (2.0 / gp1) * (cr - (gm1 / 2.0) * vxr);
        *v       = -c;
        *d       = dr * PowF(c / cr, 2.0 / gm1);
        *p       = pr * PowF(c / cr, 2.0 * gamma / gm1);
      }
    }
  }
}

// Calculate_Exact_Fluxes_CUDA, src/riemann_solvers/exact_cuda.cu:20-135 (no scalars): the
// Godunov flux of the exact solution sampled on the interface; transverse velocities (and
// the specific GasEnergy) upwinded with the sampled normal velocity.
void orc_exact(const double *QL, const double *QR, double *F, int nx, int ny, int nz, double gamma, int dir)
{
  size_t n_cells = size_t(nx) * ny * nz;
  int o1, o2, o3;
  DirOrder(dir, o1, o2, o3);
#pragma omp parallel for schedule(static)
  for (long t = 0; t < long(n_cells); t++) {
    const double dl = QL[t], vxl = QL[o1 * n_cells + t] / dl, vyl = QL[o2 * n_cells + t] / dl,
                 vzl = QL[o3 * n_cells + t] / dl;
    const double dr = QR[t], vxr = QR[o1 * n_cells + t] / dr, vyr = QR[o2 * n_cells + t] / dr,
                 vzr = QR[o3 * n_cells + t] / dr;
    const double El = QL[4 * n_cells + t], Er = QR[4 * n_cells + t];
    const double dgel = g_de ? QL[kGE * n_cells + t] : 0.0, dger = g_de ? QR[kGE * n_cells + t] : 0.0;
    double pl, pr;
    if (g_de) {
      pl = PressureFromDE(El, El - 0.5 * dl * (vxl * vxl + vyl * vyl + vzl * vzl), dgel, gamma);
      pr = PressureFromDE(Er, Er - 0.5 * dr * (vxr * vxr + vyr * vyr + vzr * vzr), dger, gamma);
    } else {
      pl = (El - 0.5 * dl * (vxl * vxl + vyl * vyl + vzl * vzl)) * (gamma - 1.0);
      pr = (Er - 0.5 * dr * (vxr * vxr + vyr * vyr + vzr * vzr)) * (gamma - 1.0);
    }
    pl = std::fmax(pl, kTiny), pr = std::fmax(pr, kTiny);
    const double cl = std::sqrt(gamma * pl / dl), cr = std::sqrt(gamma * pr / dr);
    double pm, vm, ds, vs, ps;
    ExactStar(&pm, &vm, dl, vxl, pl, cl, dr, vxr, pr, cr, gamma);
    ExactSample(pm, vm, &ds, &vs, &ps, dl, vxl, pl, cl, dr, vxr, pr, cr, gamma);
    const bool left = vs >= 0;
    const double vy = left ? vyl : vyr, vz = left ? vzl : vzr, ge = left ? dgel / dl : dger / dr;
    const double Es = (ps / (gamma - 1.0)) + 0.5 * ds * (vs * vs + vy * vy + vz * vz);
    F[t]                = ds * vs;
    F[o1 * n_cells + t] = ds * vs * vs + ps;
    F[o2 * n_cells + t] = ds * vs * vy;
    F[o3 * n_cells + t] = ds * vs * vz;
    F[4 * n_cells + t]  = (Es + ps) * vs;
    if (g_de) F[kGE * n_cells + t] = ds * vs * ge;
  }
}

// The solver the build selects (g_riemann).
static void RiemannFluxes(const double *QL, const double *QR, double *F, int nx, int ny, int nz, double gamma,
                          int dir)
{
  if (g_riemann == 1) {
    orc_hll(QL, QR, F, nx, ny, nz, gamma, dir);
  } else if (g_riemann == 2) {
    orc_roe(QL, QR, F, nx, ny, nz, gamma, dir);
  } else if (g_riemann == 3) {
    orc_exact(QL, QR, F, nx, ny, nz, gamma, dir);
  } else {
    orc_hllc(QL, QR, F, nx, ny, nz, gamma, dir);
  }
}

// The DE pressure of a cell and the centred velocity difference terms of the pdV
// source, as Update_Conserved_Variables_3D_half (VL_3D_cuda.cu:409-441) and
// Partial_Update_Advected_Internal_Energy_3D (hydro_cuda.cu:779-810) compute them:
// returns 0.5 * P * (dtodx (vx_imo - vx_ipo) + dtody (vy_jmo - vy_jpo) + dtodz (vz_kmo - vz_kpo)).
static double PdvTerm(const double *U, int i, int j, int k, int nx, int ny, size_t n_cells, double dtodx,
                      double dtody, double dtodz, double gamma)
{
  size_t id    = Idx(i, j, k, nx, ny);
  double d     = U[id];
  double d_inv = 1.0 / d;
  double vx = U[n_cells + id] * d_inv, vy = U[2 * n_cells + id] * d_inv, vz = U[3 * n_cells + id] * d_inv;
  double E     = U[4 * n_cells + id];
  double GE    = U[kGE * n_cells + id];
  double E_kin = KineticEnergy(d, vx, vy, vz);
  double P     = PressureFromDE(E, E - E_kin, GE, gamma);
  P            = std::fmax(P, kTiny);
  size_t imo = Idx(i - 1, j, k, nx, ny), ipo = Idx(i + 1, j, k, nx, ny);
  size_t jmo = Idx(i, j - 1, k, nx, ny), jpo = Idx(i, j + 1, k, nx, ny);
  size_t kmo = Idx(i, j, k - 1, nx, ny), kpo = Idx(i, j, k + 1, nx, ny);
  double vx_imo = U[n_cells + imo] / U[imo], vx_ipo = U[n_cells + ipo] / U[ipo];
  double vy_jmo = U[2 * n_cells + jmo] / U[jmo], vy_jpo = U[2 * n_cells + jpo] / U[jpo];
  double vz_kmo = U[3 * n_cells + kmo] / U[kmo], vz_kpo = U[3 * n_cells + kpo] / U[kpo];
  return 0.5 * P * (dtodx * (vx_imo - vx_ipo) + dtody * (vy_jmo - vy_jpo) + dtodz * (vz_kmo - vz_kpo));
}

// Update_Conserved_Variables_3D_half, src/integrators/VL_3D_cuda.cu:379-495
// (density floor branch of :479-493 when density_floor_on; DE branches :400-441, :471-478)
void orc_update_half_3d(const double *U, double *Uh, const double *Fx, const double *Fy, const double *Fz, int nx,
                        int ny, int nz, double dx, double dy, double dz, double dt, double gamma, int n_fields,
                        int density_floor_on, double density_floor)
{
  size_t n_cells = size_t(nx) * ny * nz;
  double dtodx = dt / dx, dtody = dt / dy, dtodz = dt / dz;
  const int n_cons = g_de ? n_fields - 1 : n_fields;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k < nz - 1; k++)
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        size_t id = Idx(i, j, k, nx, ny), imo = Idx(i - 1, j, k, nx, ny), jmo = Idx(i, j - 1, k, nx, ny),
               kmo = Idx(i, j, k - 1, nx, ny);
        for (int f = 0; f < n_cons; f++) {
          size_t o = size_t(f) * n_cells;
          Uh[o + id] = U[o + id] + dtodx * (Fx[o + imo] - Fx[o + id]) + dtody * (Fy[o + jmo] - Fy[o + id]) +
                       dtodz * (Fz[o + kmo] - Fz[o + id]);
        }
        if (g_de) {
          size_t o   = size_t(kGE) * n_cells;
          double pdv = PdvTerm(U, i, j, k, nx, ny, n_cells, dtodx, dtody, dtodz, gamma);
          Uh[o + id] = U[o + id] + dtodx * (Fx[o + imo] - Fx[o + id]) + dtody * (Fy[o + jmo] - Fy[o + id]) +
                       dtodz * (Fz[o + kmo] - Fz[o + id]) + pdv;
        }
        if (density_floor_on && Uh[id] < density_floor) {
          double d0 = Uh[id];
          Uh[id]    = density_floor;
          for (int f = 1; f < 5; f++) Uh[f * n_cells + id] *= (density_floor / d0);
          if (g_de) Uh[kGE * n_cells + id] *= (density_floor / d0);
        }
      }
}

// Partial_Update_Advected_Internal_Energy_3D, src/hydro/hydro_cuda.cu:754-833: the
// full-step pdV source added to GasEnergy of the real cells before the update
void orc_partial_update_ge_3d(double *U, int nx, int ny, int nz, int n_ghost, double dx, double dy, double dz,
                              double dt, double gamma)
{
  size_t n_cells = size_t(nx) * ny * nz;
  double dtodx = dt / dx, dtody = dt / dy, dtodz = dt / dz;
  // the term reads only density, momenta, Energy and the cell's own GasEnergy, so
  // computing every term first and adding afterwards is the kernel's result
  std::vector<double> add(n_cells, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = n_ghost; k < nz - n_ghost; k++)
    for (int j = n_ghost; j < ny - n_ghost; j++)
      for (int i = n_ghost; i < nx - n_ghost; i++)
        add[Idx(i, j, k, nx, ny)] = PdvTerm(U, i, j, k, nx, ny, n_cells, dtodx, dtody, dtodz, gamma);
  for (int k = n_ghost; k < nz - n_ghost; k++)
    for (int j = n_ghost; j < ny - n_ghost; j++)
      for (int i = n_ghost; i < nx - n_ghost; i++) {
        size_t id = Idx(i, j, k, nx, ny);
        U[kGE * n_cells + id] += add[id];
      }
}

// Average_Cell_Single_Field, src/hydro/hydro_cuda.cu:1155-1177 (density only)
static void AverageCellSingleField(int f, int i, int j, int k, int nx, int ny, size_t n_cells, double *U)
{
  double vl = U[f * n_cells + Idx(i - 1, j, k, nx, ny)], vr = U[f * n_cells + Idx(i + 1, j, k, nx, ny)];
  double vd = U[f * n_cells + Idx(i, j - 1, k, nx, ny)], vu = U[f * n_cells + Idx(i, j + 1, k, nx, ny)];
  double vb = U[f * n_cells + Idx(i, j, k - 1, nx, ny)], vt = U[f * n_cells + Idx(i, j, k + 1, nx, ny)];
  U[f * n_cells + Idx(i, j, k, nx, ny)] = (vl + vr + vd + vu + vb + vt) / 6;
}

// Average_Cell_All_Fields, src/hydro/hydro_cuda.cu:1179-1268 (no scalars, printf left out):
// the mean density, momentum-weighted velocity and pressure of the cells of the 3x3x3 block
// with positive density and pressure, summed k, j, i with i fastest; with dual energy
// GasEnergy takes the mean internal energy.
static void AverageCellAllFields(int i, int j, int k, int nx, int ny, size_t n_cells, double gamma, double *U)
{
  double d_av = 0.0, vx_av = 0.0, vy_av = 0.0, vz_av = 0.0, P_av = 0.0;
  int N = 0;
  for (int kk = k - 1; kk <= k + 1; kk++)
    for (int jj = j - 1; jj <= j + 1; jj++)
      for (int ii = i - 1; ii <= i + 1; ii++) {
        const size_t n = Idx(ii, jj, kk, nx, ny);
        const double d = U[n], mx = U[n_cells + n], my = U[2 * n_cells + n], mz = U[3 * n_cells + n];
        const double P = (U[4 * n_cells + n] - (0.5 / d) * (mx * mx + my * my + mz * mz)) * (gamma - 1.0);
        if (d > 0.0 && P > 0.0) {
          d_av += d;
          vx_av += mx;
          vy_av += my;
          vz_av += mz;
          P_av += P / (gamma - 1.0);
          N++;
        }
      }
  P_av  = P_av / N;
  vx_av = vx_av / d_av;
  vy_av = vy_av / d_av;
  vz_av = vz_av / d_av;
  d_av  = d_av / N;
  const size_t id   = Idx(i, j, k, nx, ny);
  U[id]             = d_av;
  U[n_cells + id]   = d_av * vx_av;
  U[2 * n_cells + id] = d_av * vy_av;
  U[3 * n_cells + id] = d_av * vz_av;
  U[4 * n_cells + id] = P_av / (gamma - 1.0) + 0.5 * d_av * (vx_av * vx_av + vy_av * vy_av + vz_av * vz_av);
  if (g_de) U[kGE * n_cells + id] = P_av / (gamma - 1.0);
}

// A build without both DENSITY_FLOOR and TEMPERATURE_FLOOR checks each updated cell for a
// negative or NaN density or energy and repairs it (hydro_cuda.cu:372-382).
static int g_crash_repair = 0;

// Update_Conserved_Variables_3D, src/hydro/hydro_cuda.cu:172-402 (no gravity, no scalars)
// STATIC_GRAV build: the field id (Parameters::custom_grav, 0 = none) and the grid's place in
// the domain (x_off.., xbound..), set by orc_set_static_gravity
static int g_grav           = 0;
static int g_grav_off[3]    = {0, 0, 0};
static double g_grav_bnd[3] = {0.0, 0.0, 0.0};
constexpr double kGN        = 4.49451e-18;  // GN, kpc^3 / M_sun / kyr^2 (src/global/global.h:29)
constexpr double kMsunCgs   = 1.98847e33;   // MSUN_CGS (src/global/global.h:35)

static double GravPos(int axis, int idx, int n_ghost, double h)
{
  return (g_grav_off[axis] + idx - n_ghost + 0.5) * h + g_grav_bnd[axis];
}

// calc_g_3D, src/gravity/static_grav.h:126-210: case 1, NFW halo + Miyamoto-Nagai disk
static void StaticG3D(double x, double y, double z, double *gx, double *gy, double *gz)
{
  *gx = *gy = *gz = 0.0;
  if (g_grav != 1) return;
  const double r_disk = std::sqrt(x * x + y * y), r_halo = std::sqrt(x * x + y * y + z * z);
  const double M_vir = 1.0e12, M_d = 6.5e10, R_d = 3.5, z_d = 3.5 / 5.0, R_vir = 261., c_vir = 20.0;
  const double M_h = M_vir - M_d, R_h = R_vir / c_vir;
  const double phi_0_h = kGN * M_h / (std::log(1.0 + c_vir) - c_vir / (1.0 + c_vir));
  const double xx      = r_halo / R_h;
  const double a_halo  = -phi_0_h * (std::log(1 + xx) - xx / (1 + xx)) / (r_halo * r_halo);
  const double a_halo_r = a_halo * (r_disk / r_halo), a_halo_z = a_halo * (z / r_halo);
  const double zs       = std::sqrt(z * z + z_d * z_d), b = R_d + zs;
  const double a_disk_r = -kGN * M_d * r_disk * std::pow(r_disk * r_disk + b * b, -1.5);
  const double a_disk_z = -kGN * M_d * z * b / (std::pow(r_disk * r_disk + b * b, 1.5) * zs);
  *gx = (x / r_disk) * (a_disk_r + a_halo_r);
  *gy = (y / r_disk) * (a_disk_r + a_halo_r);
  *gz = a_disk_z + a_halo_z;
}

// calc_g_2D, static_grav.h:60-124: 1 Gresho vortex, 2 Rayleigh-Taylor, 3 Keplerian disk,
// 4 Kuzmin disk + NFW halo
static void StaticG2D(double x, double y, double *gx, double *gy)
{
  *gx = *gy = 0.0;
  const double r = std::sqrt(x * x + y * y), phi = std::atan2(y, x);
  if (g_grav == 1) {
    if (r < 0.2) {
      *gx = -std::cos(phi) * 25.0 * r, *gy = -std::sin(phi) * 25.0 * r;
    } else if (r >= 0.2 && r < 0.4) {
      *gx = -std::cos(phi) * (4.0 - 20.0 * r + 25.0 * r * r) / r;
      *gy = -std::sin(phi) * (4.0 - 20.0 * r + 25.0 * r * r) / r;
    }
  } else if (g_grav == 2) {
    *gx = 0, *gy = -1;
  } else if (g_grav == 3) {
    const double M = 1 * kMsunCgs;
    *gx = -std::cos(phi) * kGN * M / (r * r), *gy = -std::sin(phi) * kGN * M / (r * r);
  } else if (g_grav == 4) {
    const double M_vir = 1.0e12, M_d = 6.5e10, M_h = M_vir - M_d, R_vir = 261, c_vir = 20, R_s = R_vir / c_vir;
    const double R_d = 3.5, xx = r / R_s;
    const double a_d = kGN * M_d * r * std::pow(r * r + R_d * R_d, -1.5);
    const double a_h = kGN * M_h * (std::log(1 + xx) - xx / (1 + xx)) /
                       ((std::log(1 + c_vir) - c_vir / (1 + c_vir)) * r * r);
    const double a = a_d + a_h;
    *gx = -std::cos(phi) * a, *gy = -std::sin(phi) * a;
  }
}

// calc_g_1D, static_grav.h:16-58: case 1, the vertical field of the same halo + disk at
// cylindrical radius 13.92 kpc
static double StaticG1D(double x)
{
  if (g_grav != 1) return 0.0;
  const double r_disk = 13.9211647546, r_halo = std::sqrt(x * x + r_disk * r_disk);
  const double M_vir = 1.0e12, M_d = 6.5e10, M_h = M_vir - M_d, R_vir = 261, c_vir = 20.0, R_h = R_vir / c_vir;
  const double R_d = 3.5, z_d = 3.5 / 5.0;
  const double phi_0_h = kGN * M_h / (std::log(1.0 + c_vir) - c_vir / (1.0 + c_vir)), xx = r_halo / R_h;
  const double a_halo  = -phi_0_h * (std::log(1 + xx) - xx / (1 + xx)) / (r_halo * r_halo);
  const double zs = std::sqrt(x * x + z_d * z_d), b = R_d + zs;
  const double a_disk_z = -kGN * M_d * x * b / (std::pow(r_disk * r_disk + b * b, 1.5) * zs);
  return (x / r_halo) * a_halo + a_disk_z;
}

// The time-centred source of the full update (hydro_cuda.cu:301-315 and the 1D / 2D forms at
// :59-67, :142-153): momentum += dt g (d + d_n) / 2, energy += dt g . (d + d_n)(v + v_n) / 4, with
// d, v at the start of the step and d_n, v_n after the flux update and floor.
static void GravSource(double *U, size_t n_cells, size_t id, double dt, const double s[4], const double g[3], int dims)
{
  const double d = s[0], d_inv = 1.0 / d, d_n = U[id], d_inv_n = 1.0 / d_n;
  double v[3], vn[3], e = 0.0;
  for (int a = 0; a < dims; a++) {
    v[a]  = s[1 + a] * d_inv;
    vn[a] = U[(1 + a) * n_cells + id] * d_inv_n;
  }
  for (int a = 0; a < dims; a++) U[(1 + a) * n_cells + id] += 0.5 * dt * g[a] * (d + d_n);
  for (int a = 0; a < dims; a++) e = a == 0 ? 0.25 * dt * g[0] * (d + d_n) * (v[0] + vn[0])
                                            : e + 0.25 * dt * g[a] * (d + d_n) * (v[a] + vn[a]);
  U[4 * n_cells + id] += e;
}

void orc_set_static_gravity(int custom_grav, int x_off, int y_off, int z_off, double xbound, double ybound,
                            double zbound)
{
  g_grav     = custom_grav;
  g_grav_off[0] = x_off, g_grav_off[1] = y_off, g_grav_off[2] = z_off;
  g_grav_bnd[0] = xbound, g_grav_bnd[1] = ybound, g_grav_bnd[2] = zbound;
}

// The fields at one position (for the tests): dims 1, 2 or 3
void orc_static_gravity_field(int dims, double x, double y, double z, double *g)
{
  g[0] = g[1] = g[2] = 0.0;
  if (dims == 1) g[0] = StaticG1D(x);
  if (dims == 2) StaticG2D(x, y, &g[0], &g[1]);
  if (dims == 3) StaticG3D(x, y, z, &g[0], &g[1], &g[2]);
}

void orc_update_3d(double *U, const double *Fx, const double *Fy, const double *Fz, int nx, int ny, int nz,
                   int n_ghost, double dx, double dy, double dz, double dt, int n_fields, int density_floor_on,
                   double density_floor, double gamma)
{
  size_t n_cells = size_t(nx) * ny * nz;
  double dtodx = dt / dx, dtody = dt / dy, dtodz = dt / dz;
  std::vector<uint8_t> repair(n_cells, 0);
  std::vector<double> start;
  if (g_grav) start.assign(U, U + 4 * n_cells);  // d and momenta at the start of the step
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = n_ghost; k < nz - n_ghost; k++)
    for (int j = n_ghost; j < ny - n_ghost; j++)
      for (int i = n_ghost; i < nx - n_ghost; i++) {
        size_t id = Idx(i, j, k, nx, ny), imo = Idx(i - 1, j, k, nx, ny), jmo = Idx(i, j - 1, k, nx, ny),
               kmo = Idx(i, j, k - 1, nx, ny);
        for (int f = 0; f < n_fields; f++) {
          size_t o = size_t(f) * n_cells;
          U[o + id] += dtodx * (Fx[o + imo] - Fx[o + id]) + dtody * (Fy[o + jmo] - Fy[o + id]) +
                       dtodz * (Fz[o + kmo] - Fz[o + id]);
        }
        if (density_floor_on && U[id] < density_floor) {
          if (U[id] > 0) {
            double d0 = U[id];
            U[id]     = density_floor;
            for (int f = 1; f < 5; f++) U[f * n_cells + id] *= (density_floor / d0);
            if (g_de) U[kGE * n_cells + id] *= (density_floor / d0);  // hydro_cuda.cu:290-292
          } else {
            repair[id] = 1;
          }
        }
      }
  // Negative-density repair (hydro_cuda.cu:293-297). The reference averages
  // the six neighbours while they are being updated concurrently (a race); here
  // it runs after the sweep on the updated neighbours, serially in cell order.
  for (int k = n_ghost; k < nz - n_ghost; k++)
    for (int j = n_ghost; j < ny - n_ghost; j++)
      for (int i = n_ghost; i < nx - n_ghost; i++)
        if (repair[Idx(i, j, k, nx, ny)]) AverageCellSingleField(0, i, j, k, nx, ny, n_cells, U);
  if (g_grav) {  // STATIC_GRAV, hydro_cuda.cu:301-315
    for (int k = n_ghost; k < nz - n_ghost; k++)
      for (int j = n_ghost; j < ny - n_ghost; j++)
        for (int i = n_ghost; i < nx - n_ghost; i++) {
          const size_t id = Idx(i, j, k, nx, ny);
          const double s4[4] = {start[id], start[n_cells + id], start[2 * n_cells + id], start[3 * n_cells + id]};
          double g[3];
          StaticG3D(GravPos(0, i, n_ghost, dx), GravPos(1, j, n_ghost, dy), GravPos(2, k, n_ghost, dz), &g[0], &g[1],
                    &g[2]);
          GravSource(U, n_cells, id, dt, s4, g, 3);
        }
  }
  // The crash repair (hydro_cuda.cu:372-382), after the floor and the gravity source; like
  // the density average it runs after the sweep, serially in cell order.
  if (g_crash_repair) {
    for (int k = n_ghost; k < nz - n_ghost; k++)
      for (int j = n_ghost; j < ny - n_ghost; j++)
        for (int i = n_ghost; i < nx - n_ghost; i++) {
          const size_t id = Idx(i, j, k, nx, ny);
          const double d = U[id], E = U[4 * n_cells + id];
          if (d < 0.0 || d != d || E < 0.0 || E != E) AverageCellAllFields(i, j, k, nx, ny, n_cells, gamma, U);
        }
  }
}

// Update_Conserved_Variables_1D_half, src/integrators/VL_1D_cuda.cu:158-...
void orc_update_half_1d(const double *U, double *Uh, const double *F, int n_cells, double dx, double dt, int n_fields)
{
  double dtodx = dt / dx;
  for (int id = 1; id < n_cells - 1; id++)
    for (int f = 0; f < n_fields; f++) {
      size_t o   = size_t(f) * n_cells;
      Uh[o + id] = U[o + id] + dtodx * (F[o + id - 1] - F[o + id]);
      if (g_de && f == kGE) {
        // the pdV source at the half step (hydro_cuda.cu:222-235, 259-263): pressure from
        // the total energy, velocities by the inverse density, neighbours by division
        const double d = U[id], d_inv = 1.0 / d;
        const double vx = U[n_cells + id] * d_inv, vy = U[2 * n_cells + id] * d_inv, vz = U[3 * n_cells + id] * d_inv;
        const double P      = (U[4 * n_cells + id] - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma_1d() - 1.0);
        const double vx_imo = U[n_cells + id - 1] / U[id - 1], vx_ipo = U[n_cells + id + 1] / U[id + 1];
        Uh[o + id]          = U[o + id] + dtodx * (F[o + id - 1] - F[o + id]) + 0.5 * P * (dtodx * (vx_imo - vx_ipo));
      }
    }
}

// Partial_Update_Advected_Internal_Energy_1D, src/hydro/hydro_cuda.cu:663-700: the pdV source
// of the full step from the state at its start (DE pressure, neighbours' centred velocities)
void orc_partial_update_ge_1d(double *U, int nx, int n_ghost, double dx, double dt, double gamma)
{
  const size_t n_cells = size_t(nx);
  const double dtodx   = dt / dx;
  for (int id = n_ghost; id < nx - n_ghost; id++) {
    const double d = U[id], d_inv = 1.0 / d;
    const double vx = U[n_cells + id] * d_inv, vy = U[2 * n_cells + id] * d_inv, vz = U[3 * n_cells + id] * d_inv;
    const double E = U[4 * n_cells + id], GE = U[kGE * n_cells + id];
    const double E_kin = 0.5 * d * (vx * vx + vy * vy + vz * vz);
    const double P     = std::fmax(PressureFromDE(E, E - E_kin, GE, gamma), kTiny);
    const double vx_imo = U[n_cells + id - 1] / U[id - 1], vx_ipo = U[n_cells + id + 1] / U[id + 1];
    U[kGE * n_cells + id] += 0.5 * P * (dtodx * (vx_imo - vx_ipo));
  }
}

// Select_Internal_Energy_1D, src/hydro/hydro_cuda.cu:835-885 (neighbours clamped to real cells)
void orc_select_internal_energy_1d(double *U, int nx, int n_ghost)
{
  const size_t n_cells = size_t(nx);
  std::vector<double> sel(n_cells, 0.0);
  for (int id = n_ghost; id < nx - n_ghost; id++) {
    const int imo = std::max(id - 1, n_ghost), ipo = std::min(id + 1, nx - n_ghost - 1);
    const double d = U[id], d_inv = 1.0 / d;
    const double vx = U[n_cells + id] * d_inv, vy = U[2 * n_cells + id] * d_inv, vz = U[3 * n_cells + id] * d_inv;
    const double E = U[4 * n_cells + id], U_advected = U[kGE * n_cells + id];
    const double U_total = E - 0.5 * d * (vx * vx + vy * vy + vz * vz);
    double Emax          = std::fmax(U[4 * n_cells + imo], E);
    Emax                 = std::fmax(Emax, U[4 * n_cells + ipo]);
    const double u       = ((U_total / E > kEta1) || (U_total / Emax > kEta2)) ? U_total : U_advected;
    sel[id]              = std::fmax(u, kTiny);
  }
  for (int id = n_ghost; id < nx - n_ghost; id++) U[kGE * n_cells + id] = sel[id];
}

// Sync_Energies_1D, src/hydro/hydro_cuda.cu:1018-1040
void orc_sync_energies_1d(double *U, int nx, int n_ghost)
{
  const size_t n_cells = size_t(nx);
  for (int id = n_ghost; id < nx - n_ghost; id++) {
    const double d = U[id], d_inv = 1.0 / d;
    const double vx = U[n_cells + id] * d_inv, vy = U[2 * n_cells + id] * d_inv, vz = U[3 * n_cells + id] * d_inv;
    U[4 * n_cells + id] = 0.5 * d * (vx * vx + vy * vy + vz * vz) + U[kGE * n_cells + id];
  }
}

// Update_Conserved_Variables_1D, src/hydro/hydro_cuda.cu:20-82
void orc_update_1d(double *U, const double *F, int n_cells, int n_ghost, double dx, double dt, int n_fields)
{
  double dtodx = dt / dx;
  for (int id = n_ghost; id < n_cells - n_ghost; id++) {
    const double s4[4] = {U[id], U[n_cells + id], U[2 * n_cells + id], U[3 * n_cells + id]};
    for (int f = 0; f < n_fields; f++) {
      size_t o = size_t(f) * n_cells;
      U[o + id] += dtodx * (F[o + id - 1] - F[o + id]);
    }
    if (g_grav) {  // STATIC_GRAV, hydro_cuda.cu:59-67
      const double g[3] = {StaticG1D(GravPos(0, id, n_ghost, dx)), 0.0, 0.0};
      GravSource(U, size_t(n_cells), size_t(id), dt, s4, g, 1);
    }
  }
}

// Temperature_Floor_Kernel, src/hydro/hydro_cuda.cu:1117-1153 (DE branch :1146-1151)
void orc_temperature_floor(double *U, int nx, int ny, int nz, int n_ghost, double U_floor)
{
  size_t n_cells = size_t(nx) * ny * nz;
  // The reference kernel's real-cell guard uses n_ghost in every direction
  // (hydro_cuda.cu:1130-1131), so on a 1D or 2D grid it touches no cell.
  if (ny == 1 || nz == 1) return;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = n_ghost; k < nz - n_ghost; k++)
    for (int j = n_ghost; j < ny - n_ghost; j++)
      for (int i = n_ghost; i < nx - n_ghost; i++) {
        size_t id   = Idx(i, j, k, nx, ny);
        double d    = U[id];
        double dinv = 1.0 / d;
        double vx = U[n_cells + id] * dinv, vy = U[2 * n_cells + id] * dinv, vz = U[3 * n_cells + id] * dinv;
        double E    = U[4 * n_cells + id];
        double Ekin = 0.5 * d * (vx * vx + vy * vy + vz * vz);
        double u    = (E - Ekin) / d;
        if (u < U_floor) U[4 * n_cells + id] = Ekin + d * U_floor;
        if (g_de) {
          u = U[kGE * n_cells + id] / d;
          if (u < U_floor) U[kGE * n_cells + id] = d * U_floor;
        }
      }
}

// Select_Internal_Energy_3D, src/hydro/hydro_cuda.cu:949-1016: GasEnergy of every
// real cell becomes the internal energy the dual-energy condition selects
void orc_select_internal_energy_3d(double *U, int nx, int ny, int nz, int n_ghost)
{
  size_t n_cells = size_t(nx) * ny * nz;
  std::vector<double> sel(n_cells, 0.0);
  // every cell reads Energy only, which this kernel does not write
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = n_ghost; k < nz - n_ghost; k++)
    for (int j = n_ghost; j < ny - n_ghost; j++)
      for (int i = n_ghost; i < nx - n_ghost; i++) {
        size_t id    = Idx(i, j, k, nx, ny);
        size_t imo   = Idx(std::max(i - 1, n_ghost), j, k, nx, ny);
        size_t ipo   = Idx(std::min(i + 1, nx - n_ghost - 1), j, k, nx, ny);
        size_t jmo   = Idx(i, std::max(j - 1, n_ghost), k, nx, ny);
        size_t jpo   = Idx(i, std::min(j + 1, ny - n_ghost - 1), k, nx, ny);
        size_t kmo   = Idx(i, j, std::max(k - 1, n_ghost), nx, ny);
        size_t kpo   = Idx(i, j, std::min(k + 1, nz - n_ghost - 1), nx, ny);
        double d     = U[id];
        double d_inv = 1.0 / d;
        double vx = U[n_cells + id] * d_inv, vy = U[2 * n_cells + id] * d_inv, vz = U[3 * n_cells + id] * d_inv;
        double E          = U[4 * n_cells + id];
        double U_advected = U[kGE * n_cells + id];
        double U_total    = E - 0.5 * d * (vx * vx + vy * vy + vz * vz);
        const double *En  = U + 4 * n_cells;
        double Emax       = std::fmax(En[imo], E);
        Emax              = std::fmax(Emax, En[ipo]);
        Emax              = std::fmax(Emax, En[jmo]);
        Emax              = std::fmax(Emax, En[jpo]);
        Emax              = std::fmax(Emax, En[kmo]);
        Emax              = std::fmax(Emax, En[kpo]);
        double u          = ((U_total / E > kEta1) || (U_total / Emax > kEta2)) ? U_total : U_advected;
        sel[id]           = std::fmax(u, kTiny);
      }
  for (int k = n_ghost; k < nz - n_ghost; k++)
    for (int j = n_ghost; j < ny - n_ghost; j++)
      for (int i = n_ghost; i < nx - n_ghost; i++) {
        size_t id             = Idx(i, j, k, nx, ny);
        U[kGE * n_cells + id] = sel[id];
      }
}

// Sync_Energies_3D, src/hydro/hydro_cuda.cu:1070-1099: Energy = kinetic + selected internal
void orc_sync_energies_3d(double *U, int nx, int ny, int nz, int n_ghost)
{
  size_t n_cells = size_t(nx) * ny * nz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = n_ghost; k < nz - n_ghost; k++)
    for (int j = n_ghost; j < ny - n_ghost; j++)
      for (int i = n_ghost; i < nx - n_ghost; i++) {
        size_t id    = Idx(i, j, k, nx, ny);
        double d     = U[id];
        double d_inv = 1.0 / d;
        double vx = U[n_cells + id] * d_inv, vy = U[2 * n_cells + id] * d_inv, vz = U[3 * n_cells + id] * d_inv;
        double u = U[kGE * n_cells + id];
        U[4 * n_cells + id] = 0.5 * d * (vx * vx + vy * vy + vz * vz) + u;
      }
}

// Calc_dt_3D + Calc_dt_1D, src/hydro/hydro_cuda.cu:437-560, returns max inverse dt
double orc_calc_dti(const double *U, int nx, int ny, int nz, int n_ghost, double dx, double dy, double dz,
                    double gamma)
{
  size_t n_cells = size_t(nx) * ny * nz;
  double max_dti = -std::numeric_limits<double>::max();
  if (nx > 1 && ny == 1 && nz == 1) {
    for (int id = n_ghost; id < nx - n_ghost; id++) {
      double d = U[id], dinv = 1.0 / d;
      double vx = U[n_cells + id] * dinv, vy = U[2 * n_cells + id] * dinv, vz = U[3 * n_cells + id] * dinv;
      double P = (U[4 * n_cells + id] - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0);
      P        = std::fmax(P, kTiny);
      double cs = std::sqrt(dinv * gamma * P);
      max_dti   = std::fmax(max_dti, (std::fabs(vx) + cs) / dx);
    }
    return max_dti;
  }
  if (ny > 1 && nz == 1) {  // Calc_dt_2D, hydro_cuda.cu:474-510
    for (int j = n_ghost; j < ny - n_ghost; j++)
      for (int i = n_ghost; i < nx - n_ghost; i++) {
        size_t id = Idx(i, j, 0, nx, ny);
        double d = U[id], dinv = 1.0 / d;
        double vx = U[n_cells + id] * dinv, vy = U[2 * n_cells + id] * dinv, vz = U[3 * n_cells + id] * dinv;
        double P  = (U[4 * n_cells + id] - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0);
        P         = std::fmax(P, 1.0e-20);
        double cs = std::sqrt(dinv * gamma * P);
        max_dti   = std::fmax(max_dti, std::fmax((std::fabs(vx) + cs) / dx, (std::fabs(vy) + cs) / dy));
      }
    return max_dti;
  }
#pragma omp parallel for collapse(2) reduction(max : max_dti) schedule(static)
  for (int k = n_ghost; k < nz - n_ghost; k++)
    for (int j = n_ghost; j < ny - n_ghost; j++)
      for (int i = n_ghost; i < nx - n_ghost; i++) {
        size_t id = Idx(i, j, k, nx, ny);
        double d = U[id], dinv = 1.0 / d;
        double vx = U[n_cells + id] * dinv, vy = U[2 * n_cells + id] * dinv, vz = U[3 * n_cells + id] * dinv;
        double E = U[4 * n_cells + id];
        double v = orc_hydro_inverse_crossing_time(E, d, dinv, vx, vy, vz, dx, dy, dz, gamma);
        // fmax semantics: a NaN crossing time never wins (hydro_cuda.cu:551)
        max_dti = std::fmax(max_dti, v);
      }
  return max_dti;
}

// FindIndex, src/grid/cuda_boundaries.cu:240-302
static int FindIndex(int ig, int n, int flag, int face, int n_ghost, double *a)
{
  if (face == 0) {
    switch (flag) {
      case 2:
        *a = -1.0;
        return 2 * n_ghost - ig - 1;
      case 3:
        return n_ghost;
      case 4:
        return -1;
      case 5:
        return ig;
      default:
        return ig + n - 2 * n_ghost;
    }
  }
  switch (flag) {
    case 2:
      *a = -1.0;
      return 2 * (n - n_ghost) - ig - 1;
    case 3:
      return n - n_ghost - 1;
    case 4:
      return -1;
    case 5:
      return ig;
    default:
      return ig - n + 2 * n_ghost;
  }
}

// SetBoundaryMapping, src/grid/cuda_boundaries.cu:161-238
static long BoundaryMapping(int ig, int jg, int kg, double *a, const int *flags, int nx, int ny, int nz, int n_ghost)
{
  long idx = 0;
  int ir, jr, kr;
  if (nx > 1) {
    if (ig < n_ghost)
      ir = FindIndex(ig, nx, flags[0], 0, n_ghost, &a[0]);
    else if (ig >= nx - n_ghost)
      ir = FindIndex(ig, nx, flags[1], 1, n_ghost, &a[0]);
    else
      ir = ig;
    if (ir < 0) return -1;
    idx += ir;
  }
  if (ny > 1) {
    if (jg < n_ghost)
      jr = FindIndex(jg, ny, flags[2], 0, n_ghost, &a[1]);
    else if (jg >= ny - n_ghost)
      jr = FindIndex(jg, ny, flags[3], 1, n_ghost, &a[1]);
    else
      jr = jg;
    if (jr < 0) return -1;
    idx += long(nx) * jr;
  }
  if (nz > 1) {
    if (kg < n_ghost)
      kr = FindIndex(kg, nz, flags[4], 0, n_ghost, &a[2]);
    else if (kg >= nz - n_ghost)
      kr = FindIndex(kg, nz, flags[5], 1, n_ghost, &a[2]);
    else
      kr = kg;
    if (kr < 0) return -1;
    idx += long(nx) * ny * kr;
  }
  return idx;
}

// SetGhostCellsKernel over one face slab, src/grid/cuda_boundaries.cu:71-159.  diode: the
// transmissive outflow diode of :121-147, which the MHD build compiles out (#ifndef MHD).
static void SetGhostSlab(double *c, int nx, int ny, int nz, int n_fields, int n_ghost, const int *flags, int isize,
                         int jsize, int ksize, int imin, int jmin, int kmin, int dir, bool diode)
{
  size_t n_cells = size_t(nx) * ny * nz;
  for (int kk = 0; kk < ksize; kk++)
    for (int jj = 0; jj < jsize; jj++)
      for (int ii = 0; ii < isize; ii++) {
        int i = ii + imin, j = jj + jmin, k = kk + kmin;
        size_t g = Idx(i, j, k, nx, ny);
        double a[3] = {1., 1., 1.};
        long idx     = BoundaryMapping(i, j, k, a, flags, nx, ny, nz, n_ghost);
        if (idx < 0) continue;
        for (int f = 0; f < n_fields; f++) c[g + f * n_cells] = c[idx + f * n_cells];
        if (flags[0] == 2 || flags[1] == 2) c[g + n_cells] *= a[0];
        if (flags[2] == 2 || flags[3] == 2) c[g + 2 * n_cells] *= a[1];
        if (flags[4] == 2 || flags[5] == 2) c[g + 3 * n_cells] *= a[2];
        if (diode && flags[dir] == 3) {  // hydro outflow diode, cuda_boundaries.cu:121-147
          size_t mom = g + size_t(dir / 2 + 1) * n_cells;
          if (dir % 2 == 0) {
            if (c[mom] > 0.0) {
              c[g + 4 * n_cells] -= 0.5 * (c[mom] * c[mom]) / c[g];
              c[mom] = 0.0;
            }
          } else {
            if (c[mom] < 0.0) {
              c[g + 4 * n_cells] -= 0.5 * (c[mom] * c[mom]) / c[g];
              c[mom] = 0.0;
            }
          }
        }
      }
}

void orc_set_ghost_cells(double *c, int nx, int ny, int nz, int n_fields, int n_ghost, const int *flags, int isize,
                         int jsize, int ksize, int imin, int jmin, int kmin, int dir)
{
  SetGhostSlab(c, nx, ny, nz, n_fields, n_ghost, flags, isize, jsize, ksize, imin, jmin, kmin, dir, true);
}

// the same slab as the MHD build fills it: no outflow diode
void orc_set_ghost_cells_mhd(double *c, int nx, int ny, int nz, int n_fields, int n_ghost, const int *flags,
                             int isize, int jsize, int ksize, int imin, int jmin, int kmin, int dir)
{
  SetGhostSlab(c, nx, ny, nz, n_fields, n_ghost, flags, isize, jsize, ksize, imin, jmin, kmin, dir, false);
}

// Set_Boundary_Conditions (single process) -> Set_Boundaries(dir) for dir 0..5 with
// the slab extents of Set_Boundary_Extents, src/grid/boundary_conditions.cpp:96-111, 413-505
static void SetBoundaries(double *c, int nx, int ny, int nz, int n_fields, int n_ghost, const int *flags_in,
                          bool diode)
{
  int flags[6] = {0, 0, 0, 0, 0, 0};
  if (nx > 1) flags[0] = flags_in[0], flags[1] = flags_in[1];
  if (ny > 1) flags[2] = flags_in[2], flags[3] = flags_in[3];
  if (nz > 1) flags[4] = flags_in[4], flags[5] = flags_in[5];
  int jl = 0, ju = ny, kl = 0, ku = nz;
  if (ny > 1) jl = n_ghost, ju = ny - n_ghost;
  if (nz > 1) kl = n_ghost, ku = nz - n_ghost;
  int ndirs = nz > 1 ? 6 : (ny > 1 ? 4 : 2);
  for (int dir = 0; dir < ndirs; dir++) {
    if (flags[dir] == 4 || flags[dir] == 5) continue;  // custom / MPI faces are not set here
    int imin[3], imax[3];
    if (dir < 2) {
      imin[0] = dir == 0 ? 0 : nx - n_ghost, imax[0] = dir == 0 ? n_ghost : nx;
      imin[1] = jl, imax[1] = ju, imin[2] = kl, imax[2] = ku;
    } else if (dir < 4) {
      imin[0] = 0, imax[0] = nx;
      imin[1] = dir == 2 ? 0 : ny - n_ghost, imax[1] = dir == 2 ? n_ghost : ny;
      imin[2] = kl, imax[2] = ku;
    } else {
      imin[0] = 0, imax[0] = nx, imin[1] = 0, imax[1] = ny;
      imin[2] = dir == 4 ? 0 : nz - n_ghost, imax[2] = dir == 4 ? n_ghost : nz;
    }
    SetGhostSlab(c, nx, ny, nz, n_fields, n_ghost, flags, imax[0] - imin[0], imax[1] - imin[1], imax[2] - imin[2],
                 imin[0], imin[1], imin[2], dir, diode);
  }
}

void orc_set_boundaries(double *c, int nx, int ny, int nz, int n_fields, int n_ghost, const int *flags_in)
{
  SetBoundaries(c, nx, ny, nz, n_fields, n_ghost, flags_in, true);
}

// the MHD build's boundary fill (8 fields, no outflow diode)
void orc_set_boundaries_mhd(double *c, int nx, int ny, int nz, int n_fields, int n_ghost, const int *flags_in)
{
  SetBoundaries(c, nx, ny, nz, n_fields, n_ghost, flags_in, false);
}

// ---------------------------------------------------------------------------
// Integrators
// ---------------------------------------------------------------------------

enum { ORC_PCM = 0, ORC_PLMC = 1, ORC_PPMC = 2, ORC_PLMP = 3, ORC_PPMP = 4 };

struct Scratch {
  size_t n = 0;
  std::vector<double> Uh, QLx, QRx, QLy, QRy, QLz, QRz, Fx, Fy, Fz;
  void Resize(size_t n_total)
  {
    if (n == n_total) return;
    n = n_total;
    for (auto *v : {&Uh, &QLx, &QRx, &QLy, &QRy, &QLz, &QRz, &Fx, &Fy, &Fz}) v->assign(n_total, 0.0);
  }
};
static Scratch g_scratch;

// ---------------------------------------------------------------------------- error-budget model
// A model of the mixed form's rounding, for the accuracy analysis only: after each Riemann
// solve of orc_vl3d every face flux F gets a deterministic pseudo-random relative error
// u * scale * 2^-23 * |R|, u uniform in [-1/2, 1/2], where R is F itself (mode 1: an FP32
// evaluation of the flux, rounding on the flux's own magnitude) or F - F_phys(left cell)
// (mode 2: a flux evaluated in FP32 as a deviation from an FP64 reference flux of the cell
// left of the face).  Off (scale 0) unless orc_set_flux_rounding_model turns it on.
static double g_model_scale = 0.0;
static int g_model_mode     = 0;
static uint64_t g_model_call = 1;

static inline double ModelU(uint64_t key)  // splitmix64 -> [-1/2, 1/2)
{
  uint64_t z = key + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return double(z >> 11) * 0x1p-53 - 0.5;
}

// The physical flux of cell id's state along dir, in global field order.
static void PhysicalFlux(const double *U, size_t id, size_t n, int dir, double gamma, double f[5])
{
  const int o1 = 1 + dir, o2 = 1 + (dir + 1) % 3, o3 = 1 + (dir + 2) % 3;
  const double d = U[id], mn = U[o1 * n + id], m2 = U[o2 * n + id], m3 = U[o3 * n + id], E = U[4 * n + id];
  const double u = mn / d, v = m2 / d, w = m3 / d;
  const double p = (gamma - 1.0) * (E - 0.5 * d * (u * u + v * v + w * w));
  f[0] = mn, f[o1] = mn * u + p, f[o2] = mn * v, f[o3] = mn * w, f[4] = (E + p) * u;
}

static void ModelFluxRounding(double *F, const double *U, int nx, int ny, int nz, int dir, double gamma)
{
  if (g_model_scale == 0.0) return;
  const size_t n      = size_t(nx) * ny * nz;
  const uint64_t base = (g_model_call++) << 36;
  for (size_t id = 0; id < n; id++) {
    double f0[5] = {0, 0, 0, 0, 0};
    if (g_model_mode == 2) PhysicalFlux(U, id, n, dir, gamma, f0);
    for (int q = 0; q < 5; q++) {
      const double r = std::fabs(F[q * n + id] - f0[q]);
      F[q * n + id] += g_model_scale * 0x1p-23 * r * ModelU(base + id * 8 + q);
    }
  }
}

// VL_Algorithm_3D_CUDA, src/integrators/VL_3D_cuda.cu:38-360 (hydro; HLLC)
void orc_vl3d(double *U, int nx, int ny, int nz, int n_ghost, double dx, double dy, double dz, double dt,
              double gamma, int n_fields, int recon, int density_floor_on, double density_floor)
{
  size_t n_cells = size_t(nx) * ny * nz;
  Scratch &s     = g_scratch;
  s.Resize(n_cells * n_fields);
  orc_pcm_3d(U, s.QLx.data(), s.QRx.data(), s.QLy.data(), s.QRy.data(), s.QLz.data(), s.QRz.data(), nx, ny, nz,
             n_fields);
  RiemannFluxes(s.QLx.data(), s.QRx.data(), s.Fx.data(), nx, ny, nz, gamma, 0);
  RiemannFluxes(s.QLy.data(), s.QRy.data(), s.Fy.data(), nx, ny, nz, gamma, 1);
  RiemannFluxes(s.QLz.data(), s.QRz.data(), s.Fz.data(), nx, ny, nz, gamma, 2);
  ModelFluxRounding(s.Fx.data(), U, nx, ny, nz, 0, gamma);
  ModelFluxRounding(s.Fy.data(), U, nx, ny, nz, 1, gamma);
  ModelFluxRounding(s.Fz.data(), U, nx, ny, nz, 2, gamma);
  orc_update_half_3d(U, s.Uh.data(), s.Fx.data(), s.Fy.data(), s.Fz.data(), nx, ny, nz, dx, dy, dz, 0.5 * dt,
                     gamma, n_fields, density_floor_on, density_floor);
  if (recon == ORC_PLMC) {
    orc_plmc(s.Uh.data(), s.QLx.data(), s.QRx.data(), nx, ny, nz, gamma, 0);
    orc_plmc(s.Uh.data(), s.QLy.data(), s.QRy.data(), nx, ny, nz, gamma, 1);
    orc_plmc(s.Uh.data(), s.QLz.data(), s.QRz.data(), nx, ny, nz, gamma, 2);
  } else if (recon == ORC_PPMC) {
    orc_ppmc_vl(s.Uh.data(), s.QLx.data(), s.QRx.data(), nx, ny, nz, gamma, 0);
    orc_ppmc_vl(s.Uh.data(), s.QLy.data(), s.QRy.data(), nx, ny, nz, gamma, 1);
    orc_ppmc_vl(s.Uh.data(), s.QLz.data(), s.QRz.data(), nx, ny, nz, gamma, 2);
  } else if (recon == ORC_PLMP) {  // VL_3D_cuda.cu:230-236
    orc_plmp(s.Uh.data(), s.QLx.data(), s.QRx.data(), nx, ny, nz, gamma, 0);
    orc_plmp(s.Uh.data(), s.QLy.data(), s.QRy.data(), nx, ny, nz, gamma, 1);
    orc_plmp(s.Uh.data(), s.QLz.data(), s.QRz.data(), nx, ny, nz, gamma, 2);
  } else if (recon == ORC_PPMP) {  // VL_3D_cuda.cu:244-252
    orc_ppmp(s.Uh.data(), s.QLx.data(), s.QRx.data(), nx, ny, nz, gamma, 0);
    orc_ppmp(s.Uh.data(), s.QLy.data(), s.QRy.data(), nx, ny, nz, gamma, 1);
    orc_ppmp(s.Uh.data(), s.QLz.data(), s.QRz.data(), nx, ny, nz, gamma, 2);
  } else {
    orc_pcm_3d(s.Uh.data(), s.QLx.data(), s.QRx.data(), s.QLy.data(), s.QRy.data(), s.QLz.data(), s.QRz.data(), nx,
               ny, nz, n_fields);
  }
  RiemannFluxes(s.QLx.data(), s.QRx.data(), s.Fx.data(), nx, ny, nz, gamma, 0);
  RiemannFluxes(s.QLy.data(), s.QRy.data(), s.Fy.data(), nx, ny, nz, gamma, 1);
  RiemannFluxes(s.QLz.data(), s.QRz.data(), s.Fz.data(), nx, ny, nz, gamma, 2);
  ModelFluxRounding(s.Fx.data(), s.Uh.data(), nx, ny, nz, 0, gamma);
  ModelFluxRounding(s.Fy.data(), s.Uh.data(), nx, ny, nz, 1, gamma);
  ModelFluxRounding(s.Fz.data(), s.Uh.data(), nx, ny, nz, 2, gamma);
  if (g_de) orc_partial_update_ge_3d(U, nx, ny, nz, n_ghost, dx, dy, dz, dt, gamma);  // VL_3D_cuda.cu:314-324
  orc_update_3d(U, s.Fx.data(), s.Fy.data(), s.Fz.data(), nx, ny, nz, n_ghost, dx, dy, dz, dt, n_fields,
                density_floor_on, density_floor, gamma);
  if (g_de) {  // VL_3D_cuda.cu:349-357
    orc_select_internal_energy_3d(U, nx, ny, nz, n_ghost);
    orc_sync_energies_3d(U, nx, ny, nz, n_ghost);
  }
}

// Update_Conserved_Variables_2D_half, VL_2D_cuda.cu:183-256 (no DE): cells 1..n-2 in x and y,
// the y term added after the x term
void orc_update_half_2d(const double *U, double *Uh, const double *Fx, const double *Fy, int nx, int ny, double dx,
                        double dy, double dt, int n_fields)
{
  const size_t n_cells = size_t(nx) * ny;
  const double dtodx = dt / dx, dtody = dt / dy;
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      const size_t id = Idx(i, j, 0, nx, ny), imo = id - 1, jmo = id - nx;
      for (int f = 0; f < n_fields; f++) {
        const size_t o = size_t(f) * n_cells;
        Uh[o + id]     = U[o + id] + dtodx * (Fx[o + imo] - Fx[o + id]) + dtody * (Fy[o + jmo] - Fy[o + id]);
        if (g_de && f == kGE) {  // the half-step pdV source (VL_2D_cuda.cu:203-218, 248-252)
          const double d = U[id], d_inv = 1.0 / d;
          const double vx = U[n_cells + id] * d_inv, vy = U[2 * n_cells + id] * d_inv, vz = U[3 * n_cells + id] * d_inv;
          const double P  = (U[4 * n_cells + id] - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma_1d() - 1.0);
          const size_t ipo = id + 1, jpo = id + nx;
          const double vx_imo = U[n_cells + imo] / U[imo], vx_ipo = U[n_cells + ipo] / U[ipo];
          const double vy_jmo = U[2 * n_cells + jmo] / U[jmo], vy_jpo = U[2 * n_cells + jpo] / U[jpo];
          Uh[o + id] = U[o + id] + dtodx * (Fx[o + imo] - Fx[o + id]) + dtody * (Fy[o + jmo] - Fy[o + id]) +
                       0.5 * P * (dtodx * (vx_imo - vx_ipo) + dtody * (vy_jmo - vy_jpo));
        }
      }
    }
}

// Partial_Update_Advected_Internal_Energy_2D, hydro_cuda.cu:702-750
void orc_partial_update_ge_2d(double *U, int nx, int ny, int n_ghost, double dx, double dy, double dt, double gamma)
{
  const size_t n_cells = size_t(nx) * ny;
  const double dtodx = dt / dx, dtody = dt / dy;
  for (int j = n_ghost; j < ny - n_ghost; j++)
    for (int i = n_ghost; i < nx - n_ghost; i++) {
      const size_t id = Idx(i, j, 0, nx, ny), imo = id - 1, ipo = id + 1, jmo = id - nx, jpo = id + nx;
      const double d = U[id], d_inv = 1.0 / d;
      const double vx = U[n_cells + id] * d_inv, vy = U[2 * n_cells + id] * d_inv, vz = U[3 * n_cells + id] * d_inv;
      const double E = U[4 * n_cells + id], GE = U[kGE * n_cells + id];
      const double E_kin = 0.5 * d * (vx * vx + vy * vy + vz * vz);
      const double P     = std::fmax(PressureFromDE(E, E - E_kin, GE, gamma), kTiny);
      const double vx_imo = U[n_cells + imo] / U[imo], vx_ipo = U[n_cells + ipo] / U[ipo];
      const double vy_jmo = U[2 * n_cells + jmo] / U[jmo], vy_jpo = U[2 * n_cells + jpo] / U[jpo];
      U[kGE * n_cells + id] += 0.5 * P * (dtodx * (vx_imo - vx_ipo) + dtody * (vy_jmo - vy_jpo));
    }
}

// Select_Internal_Energy_2D and Sync_Energies_2D, hydro_cuda.cu:887-935, 1042-1068
void orc_select_sync_2d(double *U, int nx, int ny, int n_ghost)
{
  const size_t n_cells = size_t(nx) * ny;
  std::vector<double> sel(n_cells, 0.0);
  for (int j = n_ghost; j < ny - n_ghost; j++)
    for (int i = n_ghost; i < nx - n_ghost; i++) {
      const size_t id  = Idx(i, j, 0, nx, ny);
      const size_t imo = Idx(std::max(i - 1, n_ghost), j, 0, nx, ny), ipo = Idx(std::min(i + 1, nx - n_ghost - 1), j, 0, nx, ny);
      const size_t jmo = Idx(i, std::max(j - 1, n_ghost), 0, nx, ny), jpo = Idx(i, std::min(j + 1, ny - n_ghost - 1), 0, nx, ny);
      const double d = U[id], d_inv = 1.0 / d;
      const double vx = U[n_cells + id] * d_inv, vy = U[2 * n_cells + id] * d_inv, vz = U[3 * n_cells + id] * d_inv;
      const double E = U[4 * n_cells + id], U_adv = U[kGE * n_cells + id];
      const double U_tot = E - 0.5 * d * (vx * vx + vy * vy + vz * vz);
      const double *En   = U + 4 * n_cells;
      double Emax        = std::fmax(En[imo], E);
      Emax               = std::fmax(Emax, En[ipo]);
      Emax               = std::fmax(Emax, En[jmo]);
      Emax               = std::fmax(Emax, En[jpo]);
      const double u     = ((U_tot / E > kEta1) || (U_tot / Emax > kEta2)) ? U_tot : U_adv;
      sel[id]            = std::fmax(u, kTiny);
    }
  for (int j = n_ghost; j < ny - n_ghost; j++)
    for (int i = n_ghost; i < nx - n_ghost; i++) {
      const size_t id = Idx(i, j, 0, nx, ny);
      U[kGE * n_cells + id] = sel[id];
      const double d = U[id], d_inv = 1.0 / d;
      const double vx = U[n_cells + id] * d_inv, vy = U[2 * n_cells + id] * d_inv, vz = U[3 * n_cells + id] * d_inv;
      U[4 * n_cells + id] = 0.5 * d * (vx * vx + vy * vy + vz * vz) + sel[id];
    }
}

// Update_Conserved_Variables_2D, hydro_cuda.cu:84-160 (no gravity, no DE): real cells, the
// x and y terms summed before they are added
void orc_update_2d(double *U, const double *Fx, const double *Fy, int nx, int ny, int n_ghost, double dx, double dy,
                   double dt, int n_fields)
{
  const size_t n_cells = size_t(nx) * ny;
  const double dtodx = dt / dx, dtody = dt / dy;
  for (int j = n_ghost; j < ny - n_ghost; j++)
    for (int i = n_ghost; i < nx - n_ghost; i++) {
      const size_t id = Idx(i, j, 0, nx, ny), imo = id - 1, jmo = id - nx;
      const double s4[4] = {U[id], U[n_cells + id], U[2 * n_cells + id], U[3 * n_cells + id]};
      for (int f = 0; f < n_fields; f++) {
        const size_t o = size_t(f) * n_cells;
        U[o + id] += dtodx * (Fx[o + imo] - Fx[o + id]) + dtody * (Fy[o + jmo] - Fy[o + id]);
      }
      if (g_grav) {  // STATIC_GRAV, hydro_cuda.cu:142-153
        double g[3] = {0.0, 0.0, 0.0};
        StaticG2D(GravPos(0, i, n_ghost, dx), GravPos(1, j, n_ghost, dy), &g[0], &g[1]);
        GravSource(U, n_cells, id, dt, s4, g, 2);
      }
    }
}

// VL_Algorithm_2D_CUDA, src/integrators/VL_2D_cuda.cu:24-168 (no DE, no scalars)
void orc_vl2d(double *U, int nx, int ny, int n_ghost, double dx, double dy, double dt, double gamma, int n_fields,
              int recon)
{
  const size_t n_cells = size_t(nx) * ny;
  Scratch &s           = g_scratch;
  s.Resize(n_cells * n_fields);
  orc_pcm_3d(U, s.QLx.data(), s.QRx.data(), s.QLy.data(), s.QRy.data(), s.QLz.data(), s.QRz.data(), nx, ny, 1,
             n_fields);
  RiemannFluxes(s.QLx.data(), s.QRx.data(), s.Fx.data(), nx, ny, 1, gamma, 0);
  RiemannFluxes(s.QLy.data(), s.QRy.data(), s.Fy.data(), nx, ny, 1, gamma, 1);
  g_gamma_1d = gamma;  // the half update's DE term needs it
  orc_update_half_2d(U, s.Uh.data(), s.Fx.data(), s.Fy.data(), nx, ny, dx, dy, 0.5 * dt, n_fields);
  if (recon == ORC_PLMC) {
    orc_plmc(s.Uh.data(), s.QLx.data(), s.QRx.data(), nx, ny, 1, gamma, 0);
    orc_plmc(s.Uh.data(), s.QLy.data(), s.QRy.data(), nx, ny, 1, gamma, 1);
  } else if (recon == ORC_PPMC) {
    orc_ppmc_vl(s.Uh.data(), s.QLx.data(), s.QRx.data(), nx, ny, 1, gamma, 0);
    orc_ppmc_vl(s.Uh.data(), s.QLy.data(), s.QRy.data(), nx, ny, 1, gamma, 1);
  } else if (recon == ORC_PLMP) {
    orc_plmp(s.Uh.data(), s.QLx.data(), s.QRx.data(), nx, ny, 1, gamma, 0);
    orc_plmp(s.Uh.data(), s.QLy.data(), s.QRy.data(), nx, ny, 1, gamma, 1);
  } else if (recon == ORC_PPMP) {
    orc_ppmp(s.Uh.data(), s.QLx.data(), s.QRx.data(), nx, ny, 1, gamma, 0);
    orc_ppmp(s.Uh.data(), s.QLy.data(), s.QRy.data(), nx, ny, 1, gamma, 1);
  } else {
    orc_pcm_3d(s.Uh.data(), s.QLx.data(), s.QRx.data(), s.QLy.data(), s.QRy.data(), s.QLz.data(), s.QRz.data(), nx,
               ny, 1, n_fields);
  }
  RiemannFluxes(s.QLx.data(), s.QRx.data(), s.Fx.data(), nx, ny, 1, gamma, 0);
  RiemannFluxes(s.QLy.data(), s.QRy.data(), s.Fy.data(), nx, ny, 1, gamma, 1);
  if (g_de) orc_partial_update_ge_2d(U, nx, ny, n_ghost, dx, dy, dt, gamma);  // VL_2D_cuda.cu:140-144
  orc_update_2d(U, s.Fx.data(), s.Fy.data(), nx, ny, n_ghost, dx, dy, dt, n_fields);
  if (g_de) orc_select_sync_2d(U, nx, ny, n_ghost);  // VL_2D_cuda.cu:148-152
}

// VL_Algorithm_1D_CUDA, src/integrators/VL_1D_cuda.cu:30-146 (HLLC)
void orc_vl1d(double *U, int nx, int n_ghost, double dx, double dt, double gamma, int n_fields, int recon)
{
  size_t n_cells = size_t(nx);
  Scratch &s     = g_scratch;
  s.Resize(n_cells * n_fields);
  g_gamma_1d = gamma;
  orc_pcm_1d(U, s.QLx.data(), s.QRx.data(), nx, n_fields);
  RiemannFluxes(s.QLx.data(), s.QRx.data(), s.Fx.data(), nx, 1, 1, gamma, 0);
  orc_update_half_1d(U, s.Uh.data(), s.Fx.data(), nx, dx, 0.5 * dt, n_fields);
  if (recon == ORC_PLMC) {
    orc_plmc(s.Uh.data(), s.QLx.data(), s.QRx.data(), nx, 1, 1, gamma, 0);
  } else if (recon == ORC_PPMC) {
    orc_ppmc_vl(s.Uh.data(), s.QLx.data(), s.QRx.data(), nx, 1, 1, gamma, 0);
  } else if (recon == ORC_PLMP) {  // VL_1D_cuda.cu:98-101
    orc_plmp(s.Uh.data(), s.QLx.data(), s.QRx.data(), nx, 1, 1, gamma, 0);
  } else if (recon == ORC_PPMP) {  // VL_1D_cuda.cu:102-105
    orc_ppmp(s.Uh.data(), s.QLx.data(), s.QRx.data(), nx, 1, 1, gamma, 0);
  } else {
    orc_pcm_1d(s.Uh.data(), s.QLx.data(), s.QRx.data(), nx, n_fields);
  }
  RiemannFluxes(s.QLx.data(), s.QRx.data(), s.Fx.data(), nx, 1, 1, gamma, 0);
  if (g_de) orc_partial_update_ge_1d(U, nx, n_ghost, dx, dt, gamma);  // VL_1D_cuda.cu:126-132
  orc_update_1d(U, s.Fx.data(), nx, n_ghost, dx, dt, n_fields);
  if (g_de) {  // VL_1D_cuda.cu:137-142
    orc_select_internal_energy_1d(U, nx, n_ghost);
    orc_sync_energies_1d(U, nx, n_ghost);
  }
}

// The driver loop of src/main.cpp:245-379 restricted to the hydro path:
//   dti = Calc_Inverse_Timestep()
//   while t < tout and step < max_steps:
//     dt = C_cfl / dti; clamp to tout (main.cpp:258-261)
//     Update_Hydro_Grid: VL step, temperature floor, dti of the new state (grid3D.cpp:474-585)
//     t += dt; n_step++; Set_Boundary_Conditions_Grid
// The state must already hold its ghost cells (main.cpp:190 sets them before the loop).
// Returns the number of steps taken; *t_io is advanced in place.
int orc_run(double *U, int nx, int ny, int nz, int n_ghost, int n_fields, double dx, double dy, double dz,
            double gamma, int recon, const int *flags, double *t_io, double tout, int max_steps, int density_floor_on,
            double density_floor, int temperature_floor_on, double U_floor, double *dt_log)
{
  double t   = *t_io;
  double dti = orc_calc_dti(U, nx, ny, nz, n_ghost, dx, dy, dz, gamma);
  int steps  = 0;
  bool one_d = ny == 1 && nz == 1;
  if (g_de && n_fields != 6) return -1;  // DE: GasEnergy as field 5
  while (t < tout && steps < max_steps) {
    double dt = kCfl / dti;
    if (t + dt > tout) dt = tout - t;
    if (one_d) {
      orc_vl1d(U, nx, n_ghost, dx, dt, gamma, n_fields, recon);
    } else if (nz == 1) {  // 2D: no temperature floor (its guard, hydro_cuda.cu:1130-1131, admits no cell)
      orc_vl2d(U, nx, ny, n_ghost, dx, dy, dt, gamma, n_fields, recon);
    } else {
      const int crash0 = g_crash_repair;
      g_crash_repair   = !(density_floor_on && temperature_floor_on);
      orc_vl3d(U, nx, ny, nz, n_ghost, dx, dy, dz, dt, gamma, n_fields, recon, density_floor_on, density_floor);
      g_crash_repair = crash0;
      if (temperature_floor_on) orc_temperature_floor(U, nx, ny, nz, n_ghost, U_floor);
    }
    dti = orc_calc_dti(U, nx, ny, nz, n_ghost, dx, dy, dz, gamma);
    if (dt_log) dt_log[steps] = dt;
    t += dt;
    steps++;
    orc_set_boundaries(U, nx, ny, nz, n_fields, n_ghost, flags);
  }
  *t_io = t;
  return steps;
}

// The DE build option (dual energy) as a run-time switch for every stage above.
void orc_set_dual_energy(int on) { g_de = on ? 1 : 0; }
void orc_set_positivity_fallback(int on) { g_pos_fallback = on ? 1 : 0; }
// The error-budget model of the mixed form's rounding (see ModelFluxRounding); scale 0 = off.
void orc_set_flux_rounding_model(double scale, int mode)
{
  g_model_scale = scale;
  g_model_mode  = mode;
  g_model_call  = 1;
}
// orc_vl3d called on its own: whether the build lacks one of the floors (crash repair)
void orc_set_crash_repair(int on) { g_crash_repair = on ? 1 : 0; }
int orc_get_crash_repair() { return g_crash_repair; }
void orc_set_riemann(int solver) { g_riemann = (solver >= 1 && solver <= 3) ? solver : 0; }
int orc_get_dual_energy() { return g_de; }

// hydro_utilities::Get_Pressure_From_DE, for the known-answer tests
double orc_pressure_from_de(double E, double U_total, double U_advected, double gamma)
{
  return PressureFromDE(E, U_total, U_advected, gamma);
}

int orc_num_threads()
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
