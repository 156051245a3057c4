// support_kernels.hpp -- launchers for the non-fused parts of the hot path.
#pragma once

#include <hip/hip_runtime.h>
#include "static_gravity.hpp"

namespace chl
{

struct GhostFlags {
  int f[6];
  int diode;  // 1: transmissive faces apply the hydro outflow diode; 0 for MHD states
};

// mhd: the state is the MHD build's (8 fields), whose fill has no outflow diode
// (src/grid/cuda_boundaries.cu:121-147 is compiled under #ifndef MHD).
hipError_t LaunchGhostFill(double *U, int nx, int ny, int nz, int n_fields, int ng, const int flags[6], hipStream_t s,
                           bool mhd = false);
hipError_t LaunchGhostSlab(double *U, int nx, int ny, int nz, int n_fields, int ng, const int flags[6], int isize,
                           int jsize, int ksize, int imin, int jmin, int kmin, int dir, hipStream_t s,
                           bool mhd = false);
hipError_t LaunchCalcDt(const double *U, int nx, int ny, int nz, int ng, double dx, double dy, double dz,
                        double gamma, unsigned long long *dti, int num_cu, hipStream_t s);
hipError_t LaunchTemperatureFloor(double *U, int nx, int ny, int nz, int ng, double U_floor, int n_fields,
                                  hipStream_t s);
// Dual energy: Select_Internal_Energy_3D, then Sync_Energies_3D fused with the
// temperature floor, the inverse-dt reduction and periodic ghost images (flags: the
// kFlag* bits of vl3d_kernels.hpp).
hipError_t LaunchDeSelect(double *U, int nx, int ny, int nz, int ng, hipStream_t s);
hipError_t LaunchGasEnergyInit(double *U, long n_cells, hipStream_t s);
hipError_t LaunchDeSync(double *U, int nx, int ny, int nz, int ng, int flags, double U_floor, double dx, double dy,
                        double dz, double gamma, unsigned long long *dti, hipStream_t s);
// take: read the accumulator and zero it in one atomic exchange (for a reduction that follows
// in the same stream)
hipError_t LaunchDtAdvance(unsigned long long *dti, double *dt_dev, double *t_dev, double tout, double cfl,
                           int *n_step, hipStream_t s, bool take = false);
hipError_t LaunchTurbulenceIc(double *U, int n, int ng, int n_modes, const double *modes, double vscale, double rho,
                              double p, double gamma, hipStream_t s);
// dti: reduce the next inverse dt into it (stage 2 with one block, or the DE energy sync);
// dti_reset: zero it first, from stage 1 (both may be the same pointer).
// form: the arithmetic form's id (PolicyExact::kId, PolicyFast::kId, PolicyMixed::kId)
hipError_t LaunchVl1d(double *U, double *W, int n, int ng, int recon, int riemann, bool de, int form, double dx,
                      const double *dt_dev, double dt, double gamma, unsigned long long *dti,
                      unsigned long long *dti_reset, const GravSpec &grav, hipStream_t s);
hipError_t LaunchVl2d(double *U, double *scratch, int nx, int ny, int ng, int recon, int riemann, bool de, int form,
                      double dx, double dy, const double *dt_dev, double dt, double gamma, unsigned long long *dti,
                      const GravSpec &grav, hipStream_t s);
hipError_t LaunchReconGrid(int recon, const double *U, double *QL, double *QR, int nx, int ny, int nz, double gamma,
                           int dir, hipStream_t s);
hipError_t LaunchRiemannGrid(int riemann, const double *QL, const double *QR, double *F, int nx, int ny, int nz,
                             double gamma, int dir, hipStream_t s);

}  // namespace chl
