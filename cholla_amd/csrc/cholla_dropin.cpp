// cholla_dropin.cpp -- the reference's C++ entry points of the hydro hot path, on top of
// the C ABI.  Compiled into the host program (Cholla's binary, or the test driver), not
// into libcholla_amd.so.  Error behaviour follows the reference: a failing call prints the
// message and exits the process, as GPU_Error_Check -> chexit does (src/utils/gpu.hpp:139-155,
// src/utils/error_handling.cpp:11-29).
#include <cstdio>
#include <cstdlib>

#include "../../include/cholla_amd.h"
#include "../../include/cholla_dropin.hpp"

#ifndef CHL_PRECISION
  #define CHL_PRECISION 0
#endif

namespace
{

#if defined(MHD) && (defined(PLMP) || defined(PPMP))
  // src/utils/error_handling.cpp:98-101: MHD admits PCM, PLMC and PPMC only
  #error "cholla_dropin: MHD supports PCM, PLMC and PPMC reconstruction"
#endif
#if defined(PPMC)
constexpr int kRecon = CHL_RECON_PPMC;
#elif defined(PLMP)
constexpr int kRecon = CHL_RECON_PLMP;
#elif defined(PPMP)
constexpr int kRecon = CHL_RECON_PPMP;
#elif defined(PCM)
constexpr int kRecon = CHL_RECON_PCM;
#else
constexpr int kRecon = CHL_RECON_PLMC;  // builds/make.type.hydro:5 default
#endif
#if defined(HLL)
constexpr int kRiemann = CHL_RIEMANN_HLL;  // -DHLL in place of -DHLLC
#elif defined(ROE)
constexpr int kRiemann = CHL_RIEMANN_ROE;
#elif defined(EXACT)
constexpr int kRiemann = CHL_RIEMANN_EXACT;
#else
constexpr int kRiemann = CHL_RIEMANN_HLLC;
#endif
#ifdef DENSITY_FLOOR
constexpr int kDensityFloor = 1;
#else
constexpr int kDensityFloor = 0;
#endif

void Check(int rc, const char *where)
{
  if (rc != CHL_OK) {
    std::fprintf(stderr, "cholla_amd: %s failed (%d): %s\n", where, rc, chl_last_error());
    std::exit(-1);
  }
}

// The reference keeps its VL scratch in file-scope globals allocated on the first call
// (src/global/global_cuda.h:18-32, VL_3D_cuda.cu:58-133).  Here one handle per integrator,
// re-created when the box shape changes.
struct Slot {
  chl_hydro *h = nullptr;
  int nx = 0, ny = 0, nz = 0, ng = 0, nf = 0;
};
Slot g_vl3d, g_vl1d;

#ifdef MHD
// The MHD build's VL integrator (-DMHD -DHLLD -DVL with -DPLMC as in builds/make.type.mhd,
// or -DPCM / -DPPMC).
struct MhdSlot {
  chl_mhd *h = nullptr;
  int nx = 0, ny = 0, nz = 0, ng = 0;
};
MhdSlot g_mhd;

chl_mhd *AcquireMhd(int nx, int ny, int nz, int ng)
{
  MhdSlot &s = g_mhd;
  if (s.h && s.nx == nx && s.ny == ny && s.nz == nz && s.ng == ng) return s.h;
  if (s.h) Check(chl_mhd_destroy(s.h), "chl_mhd_destroy");
  chl_mhd_config c{};
  c.nx = nx, c.ny = ny, c.nz = nz, c.n_ghost = ng, c.precision = CHL_PRECISION;
  c.density_floor = kDensityFloor, c.temperature_floor = 0;
  for (int q = 0; q < 6; q++) c.bc_flags[q] = CHL_BC_PERIODIC;
  c.reconstruction = kRecon;
  Check(chl_mhd_create(&c, &s.h), "chl_mhd_create");
  s.nx = nx, s.ny = ny, s.nz = nz, s.ng = ng;
  return s.h;
}
#endif

chl_hydro *Acquire(Slot &s, int nx, int ny, int nz, int ng, int nf)
{
  if (s.h && s.nx == nx && s.ny == ny && s.nz == nz && s.ng == ng && s.nf == nf) return s.h;
  if (s.h) Check(chl_hydro_destroy(s.h), "chl_hydro_destroy");
  chl_hydro_config c{};
  c.nx = nx, c.ny = ny, c.nz = nz, c.n_ghost = ng, c.n_fields = nf;
  c.reconstruction = kRecon, c.riemann = kRiemann, c.precision = CHL_PRECISION;
  c.density_floor = kDensityFloor, c.temperature_floor = 0;  // the floor is its own call here
  for (int q = 0; q < 6; q++) c.bc_flags[q] = CHL_BC_PERIODIC;  // unused: VL never fills ghosts
  c.tile1 = -1, c.tile2 = -1;
#ifdef DE
  c.dual_energy = 1;  // the DE build: GasEnergy is the last of nf == 6 fields
#endif
  Check(chl_hydro_create(&c, &s.h), "chl_hydro_create");
  s.nx = nx, s.ny = ny, s.nz = nz, s.ng = ng, s.nf = nf;
  return s.h;
}

}  // namespace

void VL_Algorithm_3D_CUDA(Real *d_conserved, Real *d_grav_potential, int nx, int ny, int nz, int x_off, int y_off,
                          int z_off, int n_ghost, Real dx, Real dy, Real dz, Real xbound, Real ybound, Real zbound,
                          Real dt, int n_fields, int custom_grav, Real density_floor, Real *host_grav_potential)
{
  (void)d_grav_potential, (void)x_off, (void)y_off, (void)z_off, (void)xbound, (void)ybound, (void)zbound;
  (void)custom_grav, (void)host_grav_potential;  // self-gravity (GRAVITY) is outside this path
#ifdef MHD
  (void)n_fields;  // 8: the hydro fields and the three face-centred magnetic fields
  chl_mhd *m = AcquireMhd(nx, ny, nz, n_ghost);
  Check(chl_mhd_vl_step(m, d_conserved, dx, dy, dz, dt, gama, density_floor), "VL_Algorithm_3D_CUDA");
#else
  chl_hydro *h = Acquire(g_vl3d, nx, ny, nz, n_ghost, n_fields);
  #ifdef STATIC_GRAV
  Check(chl_hydro_set_static_gravity(h, custom_grav, x_off, y_off, z_off, xbound, ybound, zbound),
        "VL_Algorithm_3D_CUDA");
  #endif
  Check(chl_hydro_vl_step(h, d_conserved, dx, dy, dz, dt, gama, density_floor), "VL_Algorithm_3D_CUDA");
#endif
}

// The reference also frees the caller's state here through an alias
// (VL_3D_cuda.cu:365); the state belongs to Grid3D, so only the scratch is released.
void Free_Memory_VL_3D()
{
  if (g_vl3d.h) Check(chl_hydro_destroy(g_vl3d.h), "Free_Memory_VL_3D");
  g_vl3d = Slot{};
#ifdef MHD
  if (g_mhd.h) Check(chl_mhd_destroy(g_mhd.h), "Free_Memory_VL_3D");
  g_mhd = MhdSlot{};
#endif
}

void VL_Algorithm_1D_CUDA(Real *d_conserved, int nx, int x_off, int n_ghost, Real dx, Real xbound, Real dt,
                          int n_fields, int custom_grav)
{
  (void)x_off, (void)xbound, (void)custom_grav;
  chl_hydro *h = Acquire(g_vl1d, nx, 1, 1, n_ghost, n_fields);
#ifdef STATIC_GRAV
  Check(chl_hydro_set_static_gravity(h, custom_grav, x_off, 0, 0, xbound, 0.0, 0.0), "VL_Algorithm_1D_CUDA");
#endif
  Check(chl_hydro_vl_step(h, d_conserved, dx, dx, dx, dt, gama, 0.0), "VL_Algorithm_1D_CUDA");
}

void Free_Memory_VL_1D()
{
  if (g_vl1d.h) Check(chl_hydro_destroy(g_vl1d.h), "Free_Memory_VL_1D");
  g_vl1d = Slot{};
}

Real Calc_dt_GPU(Real *dev_conserved, int nx, int ny, int nz, int n_ghost, int n_fields, Real dx, Real dy, Real dz,
                 Real gamma)
{
  Real dti = 0;
#ifdef MHD
  (void)n_fields;  // mhdInverseCrossingTime with the cell-centred field (hydro_cuda.cu:420-435, 546-551)
  Check(chl_mhd_calc_dt(dev_conserved, nx, ny, nz, n_ghost, dx, dy, dz, gamma, &dti), "Calc_dt_GPU");
#else
  Check(chl_calc_dt(dev_conserved, nx, ny, nz, n_ghost, n_fields, dx, dy, dz, gamma, &dti), "Calc_dt_GPU");
#endif
  return dti;
}

void Apply_Temperature_Floor(Real *dev_conserved, int nx, int ny, int nz, int n_ghost, int n_fields, Real U_floor)
{
  Check(chl_apply_temperature_floor(dev_conserved, nx, ny, nz, n_ghost, n_fields, U_floor),
        "Apply_Temperature_Floor");
}

void SetGhostCells(Real *c_head, int nx, int ny, int nz, int n_fields, int n_cells, int n_ghost, int flags[],
                   int isize, int jsize, int ksize, int imin, int jmin, int kmin, int dir)
{
#ifdef MHD
  // the MHD build's fill has no outflow diode (src/grid/cuda_boundaries.cu:121-147 is #ifndef MHD)
  Check(chl_mhd_set_ghost_cells(c_head, nx, ny, nz, n_fields, n_cells, n_ghost, flags, isize, jsize, ksize, imin, jmin,
                                kmin, dir),
        "SetGhostCells");
#else
  Check(chl_set_ghost_cells(c_head, nx, ny, nz, n_fields, n_cells, n_ghost, flags, isize, jsize, ksize, imin, jmin,
                            kmin, dir),
        "SetGhostCells");
#endif
}

namespace mhd
{
// The body of mhd::checkMagneticDivergence (src/mhd/magnetic_divergence.cu:79-123) on the
// grid's state: the maximum |div B| over the grid, printed; above 1e-9, or negative, the
// process exits as chexit(-1) does.
Real CheckMagneticDivergenceOfState(const Real *dev_conserved, int nx, int ny, int nz, Real dx, Real dy, Real dz)
{
  Real div = 0;
  Check(chl_mhd_divergence(dev_conserved, nx, ny, nz, dx, dy, dz, &div), "checkMagneticDivergence");
  const Real limit = 1.0E-9;
  if (div > limit) {
    std::printf(
        "The magnetic divergence has exceeded the maximum allowed value. "
        "Divergence = %7.4e, the maximum allowed divergence = %7.4e\n",
        div, limit);
    std::exit(-1);
  } else if (div < 0.0) {
    std::printf("The magnetic divergence is negative. Divergence = %7.4e\n", div);
    std::exit(-1);
  }
  std::printf("Global maximum magnetic divergence = %7.4e\n", div);
  return div;
}
}  // namespace mhd
