// cholla_dropin.hpp -- the reference's own C++ entry points of the hydro hot path,
// implemented on top of the C ABI (cholla_amd.h) by cholla_amd/csrc/cholla_dropin.cpp.
//
// A Cholla build that compiles cholla_dropin.cpp instead of src/integrators/VL_3D_cuda.cu,
// VL_1D_cuda.cu and the corresponding parts of src/hydro/hydro_cuda.cu and
// src/grid/cuda_boundaries.cu, and links libcholla_amd.so, keeps every caller unchanged:
// Grid3D::Execute_Hydro_Integrator (src/grid/grid3D.cpp:432-461), Grid3D::Update_Hydro_Grid
// (:502, :577), Grid3D::FreeMemory (:633-643) and Grid3D::Set_Boundaries
// (src/grid/boundary_conditions.cpp:413).
//
// Build options read at compile time, as the reference's make.type files set them:
//   PLMC or PPMC, DENSITY_FLOOR, TEMPERATURE_FLOOR, MHD (with PLMC: the HLLD + CT integrator on
//   8-field states, builds/make.type.mhd), and CHL_PRECISION (0 = FP64 in the reference's
//   operation order [default], 1 = FP64 fast, 2 = mixed FP32/FP64), STATIC_GRAV.
#pragma once

#ifndef TYPEDEF_DEFINED_REAL
typedef double Real;  // PRECISION == 2 (src/global/global.h:17-22)
#endif

// The adiabatic index the integrators read (src/global/global.h:147); Cholla defines it.
extern Real gama;

// src/integrators/VL_3D_cuda.h:9-13
void VL_Algorithm_3D_CUDA(Real *d_conserved, Real *d_grav_potential, int nx, int ny, int nz, int x_off, int y_off,
                          int z_off, int n_ghost, Real dx, Real dy, Real dz, Real xbound, Real ybound, Real zbound,
                          Real dt, int n_fields, int custom_grav, Real density_floor, Real *host_grav_potential);
void Free_Memory_VL_3D();

// src/integrators/VL_1D_cuda.h (VL_1D_cuda.cu:30, :147)
void VL_Algorithm_1D_CUDA(Real *d_conserved, int nx, int x_off, int n_ghost, Real dx, Real xbound, Real dt,
                          int n_fields, int custom_grav);
void Free_Memory_VL_1D();

// src/hydro/hydro_cuda.h:69-70, :87
Real Calc_dt_GPU(Real *dev_conserved, int nx, int ny, int nz, int n_ghost, int n_fields, Real dx, Real dy, Real dz,
                 Real gamma);
void Apply_Temperature_Floor(Real *dev_conserved, int nx, int ny, int nz, int n_ghost, int n_fields, Real U_floor);

// src/grid/cuda_boundaries.h:15-16 (under MHD without the outflow diode, cuda_boundaries.cu:121-147)
void SetGhostCells(Real *c_head, int nx, int ny, int nz, int n_fields, int n_cells, int n_ghost, int flags[],
                   int isize, int jsize, int ksize, int imin, int jmin, int kmin, int dir);

namespace mhd
{
// src/mhd/magnetic_divergence.h: Real checkMagneticDivergence(Grid3D const &G), called every
// step by src/main.cpp:375-378.  Written as a template over the grid type so this header needs
// no Cholla header; Cholla's Grid3D (G.C.device, G.H.nx .. G.H.dz) instantiates it at the call
// site unchanged.  Prints the maximum |div B| and exits above 1e-9 (magnetic_divergence.cu:79-123).
Real CheckMagneticDivergenceOfState(const Real *dev_conserved, int nx, int ny, int nz, Real dx, Real dy, Real dz);
template <class Grid>
Real checkMagneticDivergence(Grid const &G)
{
  return CheckMagneticDivergenceOfState(G.C.device, G.H.nx, G.H.ny, G.H.nz, G.H.dx, G.H.dy, G.H.dz);
}
}  // namespace mhd
