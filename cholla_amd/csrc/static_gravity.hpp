// static_gravity.hpp -- the STATIC_GRAV build option: the analytic accelerations of
// src/gravity/static_grav.h (calc_g_1D/2D/3D) and the time-centred source the full update
// adds (hydro_cuda.cu:59-67, 142-153, 301-315).  Always FP64.
#pragma once
#include <hip/hip_runtime.h>

namespace chl
{

constexpr double kGravG    = 4.49451e-18;  // GN, kpc^3 / M_sun / kyr^2 (src/global/global.h:29)
constexpr double kMsunCgs  = 1.98847e33;   // MSUN_CGS (src/global/global.h:35)

// Where the field sits: Parameters::custom_grav (0 = none) and the grid's offsets and lower
// bounds; a cell centre is (off + i - n_ghost + 0.5) * dx + bound (static_grav.h:133-135).
struct GravSpec {
  int id;
  int off[3];
  double bound[3];
  __device__ __forceinline__ double Pos(int axis, int idx, int ng, double h) const
  {
    return (off[axis] + idx - ng + 0.5) * h + bound[axis];
  }
};

// calc_g_3D case 1 (static_grav.h:126-210): NFW halo + Miyamoto-Nagai disk
__device__ inline void GravField3D(int id, double x, double y, double z, double &gx, double &gy, double &gz)
{
  gx = gy = gz = 0.0;
  if (id != 1) return;
  const double r_disk = sqrt(x * x + y * y), r_halo = sqrt(x * x + y * y + z * z);
  const double M_vir = 1.0e12, M_d = 6.5e10, R_d = 3.5, z_d = 3.5 / 5.0, R_vir = 261., c_vir = 20.0;
  const double M_h = M_vir - M_d, R_h = R_vir / c_vir;
  const double phi_0_h  = kGravG * M_h / (log(1.0 + c_vir) - c_vir / (1.0 + c_vir));
  const double xx       = r_halo / R_h;
  const double a_halo   = -phi_0_h * (log(1 + xx) - xx / (1 + xx)) / (r_halo * r_halo);
  const double a_halo_r = a_halo * (r_disk / r_halo), a_halo_z = a_halo * (z / r_halo);
  const double zs = sqrt(z * z + z_d * z_d), b = R_d + zs;
  const double a_disk_r = -kGravG * M_d * r_disk * pow(r_disk * r_disk + b * b, -1.5);
  const double a_disk_z = -kGravG * M_d * z * b / (pow(r_disk * r_disk + b * b, 1.5) * zs);
  gx = (x / r_disk) * (a_disk_r + a_halo_r);
  gy = (y / r_disk) * (a_disk_r + a_halo_r);
  gz = a_disk_z + a_halo_z;
}

// calc_g_2D (static_grav.h:60-124): 1 Gresho vortex, 2 Rayleigh-Taylor, 3 Keplerian disk,
// 4 Kuzmin disk + NFW halo
__device__ inline void GravField2D(int id, double x, double y, double &gx, double &gy)
{
  gx = gy = 0.0;
  const double r = sqrt(x * x + y * y), phi = atan2(y, x);
  if (id == 1) {
    if (r < 0.2) {
      gx = -cos(phi) * 25.0 * r, gy = -sin(phi) * 25.0 * r;
    } else if (r >= 0.2 && r < 0.4) {
      gx = -cos(phi) * (4.0 - 20.0 * r + 25.0 * r * r) / r;
      gy = -sin(phi) * (4.0 - 20.0 * r + 25.0 * r * r) / r;
    }
  } else if (id == 2) {
    gx = 0, gy = -1;
  } else if (id == 3) {
    const double M = 1 * kMsunCgs;
    gx = -cos(phi) * kGravG * M / (r * r), gy = -sin(phi) * kGravG * M / (r * r);
  } else if (id == 4) {
    const double M_vir = 1.0e12, M_d = 6.5e10, M_h = M_vir - M_d, R_vir = 261, c_vir = 20, R_s = R_vir / c_vir;
    const double R_d = 3.5, xx = r / R_s;
    const double a_d = kGravG * M_d * r * pow(r * r + R_d * R_d, -1.5);
    const double a_h =
        kGravG * M_h * (log(1 + xx) - xx / (1 + xx)) / ((log(1 + c_vir) - c_vir / (1 + c_vir)) * r * r);
    const double a = a_d + a_h;
    gx = -cos(phi) * a, gy = -sin(phi) * a;
  }
}

// calc_g_1D case 1 (static_grav.h:16-58): the vertical field of the same halo + disk at
// cylindrical radius 13.92 kpc
__device__ inline double GravField1D(int id, double x)
{
  if (id != 1) return 0.0;
  const double r_disk = 13.9211647546, r_halo = sqrt(x * x + r_disk * r_disk);
  const double M_vir = 1.0e12, M_d = 6.5e10, M_h = M_vir - M_d, R_vir = 261, c_vir = 20.0, R_h = R_vir / c_vir;
  const double R_d = 3.5, z_d = 3.5 / 5.0;
  const double phi_0_h = kGravG * M_h / (log(1.0 + c_vir) - c_vir / (1.0 + c_vir)), xx = r_halo / R_h;
  const double a_halo  = -phi_0_h * (log(1 + xx) - xx / (1 + xx)) / (r_halo * r_halo);
  const double zs = sqrt(x * x + z_d * z_d), b = R_d + zs;
  const double a_disk_z = -kGravG * M_d * x * b / (pow(r_disk * r_disk + b * b, 1.5) * zs);
  return (x / r_halo) * a_halo + a_disk_z;
}

// The source on one cell: s = density and momenta at the start of the step, un = the state
// after the flux update and floor (updated in place); DIMS momentum components.
template <int DIMS>
__device__ __forceinline__ void GravSource(double *un, const double *s, const double *g, double dt)
{
  const double d = s[0], d_inv = 1.0 / d, d_n = un[0], d_inv_n = 1.0 / d_n;
  double v[3], vn[3];
#pragma unroll
  for (int a = 0; a < DIMS; a++) v[a] = s[1 + a] * d_inv, vn[a] = un[1 + a] * d_inv_n;
#pragma unroll
  for (int a = 0; a < DIMS; a++) un[1 + a] += 0.5 * dt * g[a] * (d + d_n);
  double e = 0.25 * dt * g[0] * (d + d_n) * (v[0] + vn[0]);
#pragma unroll
  for (int a = 1; a < DIMS; a++) e = e + 0.25 * dt * g[a] * (d + d_n) * (v[a] + vn[a]);
  un[4] += e;
}

}  // namespace chl
