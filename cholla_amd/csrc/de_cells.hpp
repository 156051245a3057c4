// de_cells.hpp -- per-cell bodies of the dual-energy selection and energy sync, shared by
// the single-box kernels (support_kernels.hip) and the ensemble kernels.
//
// They run in FP64 with the reference's operation order in every arithmetic form and are
// compiled without FMA contraction wherever they are inlined (the pragma below), so the
// single-box, split, persistent and per-box-loop launches produce the same bits.
#pragma once

#include <hip/hip_runtime.h>

#include "hydro_physics.hpp"
#include "vl3d_kernels.hpp"

namespace chl
{

// Select_Internal_Energy_3D body, src/hydro/hydro_cuda.cu:949-1016, for real cell (i, j, k):
// GasEnergy becomes U_total = E - kinetic where the thermal fraction exceeds DE_ETA_1 (or
// DE_ETA_2 of the largest energy among the cell and its six neighbours, clamped to the
// real cells), else it keeps the advected value; floored at TINY_NUMBER.  Reads Energy only.
__device__ __forceinline__ void DeSelectCell(double *U, size_t nc, int nx, int ny, int nz, int ng, int i, int j, int k)
{
#pragma clang fp contract(off)
  const size_t id  = Lin(i, j, k, nx, ny);
  const double *En = U + 4 * nc;
  const size_t imo = Lin(max(i - 1, ng), j, k, nx, ny), ipo = Lin(min(i + 1, nx - ng - 1), j, k, nx, ny);
  const size_t jmo = Lin(i, max(j - 1, ng), k, nx, ny), jpo = Lin(i, min(j + 1, ny - ng - 1), k, nx, ny);
  const size_t kmo = Lin(i, j, max(k - 1, ng), nx, ny), kpo = Lin(i, j, min(k + 1, nz - ng - 1), nx, ny);
  double d = U[id], d_inv = 1.0 / d;
  double vx = U[nc + id] * d_inv, vy = U[2 * nc + id] * d_inv, vz = U[3 * nc + id] * d_inv;
  double E          = En[id];
  double U_advected = U[kGasEnergy * nc + id];
  double U_total    = E - 0.5 * d * (vx * vx + vy * vy + vz * vz);
  double Emax       = Fmax(En[imo], E);
  Emax              = Fmax(Emax, En[ipo]);
  Emax              = Fmax(Emax, En[jmo]);
  Emax              = Fmax(Emax, En[jpo]);
  Emax              = Fmax(Emax, En[kmo]);
  Emax              = Fmax(Emax, En[kpo]);
  double u          = ((U_total / E > kDeEta1) || (U_total / Emax > kDeEta2)) ? U_total : U_advected;
  U[kGasEnergy * nc + id] = Fmax(u, kTinyNumber);
}

// Sync_Energies_3D body (hydro_cuda.cu:1070-1099) for real cell (i, j, k): Energy =
// kinetic + selected internal energy; then, as Update_Hydro_Grid continues
// (grid3D.cpp:497-577), the temperature floor of E and GasEnergy
// (hydro_cuda.cu:1132-1151) and the periodic ghost images of the cell.  Returns the
// cell's inverse crossing time (hydroInverseCrossingTime, hydro_cuda.cu:404-418) when
// kFlagReduceDt is set, else 0.
__device__ __forceinline__ double DeSyncCell(double *U, size_t nc, int nx, int ny, int nz, int ng, int i, int j, int k,
                                             int flags, double U_floor, double dx, double dy, double dz, double gamma)
{
#pragma clang fp contract(off)
  const size_t id = Lin(i, j, k, nx, ny);
  double un[6];
#pragma unroll
  for (int q = 0; q < 6; q++) un[q] = U[q * nc + id];
  double d_inv = 1.0 / un[0];
  double vx = un[1] * d_inv, vy = un[2] * d_inv, vz = un[3] * d_inv;
  un[4] = 0.5 * un[0] * (vx * vx + vy * vy + vz * vz) + un[kGasEnergy];
  if (flags & kFlagTemperatureFloor) {
    double Ekin = 0.5 * un[0] * (vx * vx + vy * vy + vz * vz);
    double u    = (un[4] - Ekin) / un[0];
    if (u < U_floor) un[4] = Ekin + un[0] * U_floor;
    u = un[kGasEnergy] / un[0];
    if (u < U_floor) un[kGasEnergy] = un[0] * U_floor;
  }
  U[4 * nc + id]          = un[4];
  U[kGasEnergy * nc + id] = un[kGasEnergy];
  if (flags & kFlagPeriodicGhosts) WritePeriodicImages<6>(U, nc, nx, ny, nz, ng, i, j, k, un);
  double m = 0.0;
  if (flags & kFlagReduceDt) {
    double P  = (un[4] - 0.5 * un[0] * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0);
    double cs = __builtin_sqrt(d_inv * gamma * P);
    m         = Fmax((Fabs(vx) + cs) / dx, (Fabs(vy) + cs) / dy);
    m         = Fmax(m, (Fabs(vz) + cs) / dz);
    m         = Fmax(m, 0.0);
  }
  return m;
}

// Real cell of a 1D index over the real cells of an nx x ny x nz box.
__device__ __forceinline__ bool RealCell(long t, int nx, int ny, int nz, int ng, int &i, int &j, int &k)
{
  const long rx = nx - 2 * ng, ry = ny - 2 * ng, rz = nz - 2 * ng;
  if (t >= rx * ry * rz) return false;
  i = int(t % rx) + ng;
  long q = t / rx;
  j = int(q % ry) + ng;
  k = int(q / ry) + ng;
  return true;
}

}  // namespace chl
