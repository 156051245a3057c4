// capi_internal.hpp -- helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/cholla_amd.h"
#include "vl3d_kernels.hpp"

namespace chl_capi
{
// Records msg as the text of chl_last_error() (per thread) and returns code.
int Fail(int code, const std::string &msg);

constexpr double kCfl = 0.3;  // C_cfl, src/grid/grid3D.cpp:157

inline int ReconRadius(int recon)
{
  if (recon == CHL_RECON_PPMC || recon == CHL_RECON_PPMP) return 2;
  return (recon == CHL_RECON_PLMC || recon == CHL_RECON_PLMP) ? 1 : 0;
}
}  // namespace chl_capi

// The integrator handle of one box (opaque to C callers).
struct chl_hydro {
  chl_hydro_config cfg{};
  size_t n_cells = 0;
  bool one_d     = false;
  bool two_d     = false;  // nz == 1 < ny: the unfused 2D step, W holds its scratch
  chl::GravSpec grav{};    // static gravity (chl_hydro_set_static_gravity); id 0 = none
  bool tile2_auto = false; // the caller left tile2 to the library (-1)
  int num_cu     = 256;
  void *W        = nullptr;  // half-step primitives (+ the pdV sources with dual energy)
  size_t W_bytes = 0;
  double *pdv    = nullptr;  // inside W: dual energy's full-step pdV source per cell
  // 3D boxes: the repair marks, counts, side buffer and gravity records (chl::RepairState)
  void *rep_buf = nullptr;
  chl::RepairState rep{};
  // device scalars: [0] dti bits, [1] dt, [2] t, [3] n_step (int in a double slot)
  unsigned char *scal = nullptr;
  hipStream_t stream  = nullptr;
  hipGraphExec_t graph = nullptr;
  // key of the captured graph
  double *g_U = nullptr;
  double g_key[7]{};

  unsigned long long *dti() { return reinterpret_cast<unsigned long long *>(scal); }
  double *dt() { return reinterpret_cast<double *>(scal + 8); }
  double *t() { return reinterpret_cast<double *>(scal + 16); }
  int *nstep() { return reinterpret_cast<int *>(scal + 24); }

  bool PeriodicFusable() const
  {
    for (int q = 0; q < 6; q++)
      if (cfg.bc_flags[q] != CHL_BC_PERIODIC) return false;
    int ng = cfg.n_ghost;
    return cfg.nx - 2 * ng >= ng && cfg.ny - 2 * ng >= ng && cfg.nz - 2 * ng >= ng;
  }
};

namespace chl_capi
{
// Fill in the library's defaults (tiles, z-chunks) of a hydro config and check it.
int PrepareConfig(chl_hydro_config &c);
// Bytes of the half-step scratch W (and, with dual energy, the pdV sources after w_main).
size_t HalfStepBytes(const chl_hydro_config &c, size_t *w_main);
// Bytes of a 3D box's repair scratch, and h->rep pointed into one of that size at base.
size_t RepairBytes(const chl_hydro_config &c);
void PlaceRepair(chl_hydro *h, void *base);
// The same for a box of nx x ny x nz cells (ghosts included) with room for grav_cap static
// gravity records (0: none); RepairAt lays the state out at base.
size_t RepairBytesFor(int nx, int ny, int nz, int ng, int grav_cap);
chl::RepairState RepairAt(void *base, int nx, int ny, int nz, int ng, int grav_cap);
// Enqueue one VL step of a 3D box on h->stream (chl_hydro_vl_step's work, no synchronize).
int EnqueueVlStep(chl_hydro *h, double *U, double dx, double dy, double dz, double dt, double gamma,
                  double density_floor);
}  // namespace chl_capi

#define CHL_HIP(expr)                                                                                          \
  do {                                                                                                         \
    hipError_t e_ = (expr);                                                                                    \
    if (e_ != hipSuccess) return chl_capi::Fail(CHL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CHL_TRY(expr)              \
  do {                             \
    int rc_ = (expr);              \
    if (rc_ != CHL_OK) return rc_; \
  } while (0)
