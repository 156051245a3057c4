// dropin_driver.cpp -- a host program that drives the hydro step exactly the way Cholla's
// main loop and Grid3D do, through the reference-signature entry points of
// cholla_dropin.hpp only.  It is the drop-in check: if Cholla's callers work unchanged, so
// does this program, which restates them:
//   main loop            src/main.cpp:245-290
//   Grid3D::set_dt       src/grid/grid3D.cpp:373-411
//   Update_Hydro_Grid    src/grid/grid3D.cpp:474-585 (integrator, temperature floor, dti)
//   Execute_Hydro_Integrator  src/grid/grid3D.cpp:413-471
//   Set_Boundary_Conditions / Set_Boundaries / Set_Boundary_Extents
//                        src/grid/boundary_conditions.cpp:96-111, 165-413, 419-505
//   MHD: mhd::checkMagneticDivergence after every step (src/main.cpp:375-378)
//   STATIC_GRAV: Parameters::custom_grav from the optional third argument (default 0)
//
// I/O (little-endian, native):  in-file = header + state, out-file = header + state + run info
//   usage:  dropin_driver in.bin out.bin [custom_grav]
//   header: int32 nx, ny, nz, n_ghost, n_fields, max_steps, flags[6];
//           float64 dx, dy, dz, gamma, t, tout, U_floor, density_floor
//   state:  float64 x n_fields*nx*ny*nz (field-major, ghosts included)
//   run info (out only): int32 n_steps; float64 t
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/cholla_amd.h"
#include "../../include/cholla_dropin.hpp"

Real gama;  // src/global/global.cpp

namespace
{
constexpr Real kCfl = 0.3;  // C_cfl, src/grid/grid3D.cpp:157

struct Header {
  int32_t nx, ny, nz, n_ghost, n_fields, max_steps, flags[6];
  double dx, dy, dz, gamma, t, tout, U_floor, density_floor;
};

struct Conserved {
  Real *device = nullptr;  // Grid3D::Conserved::device
};

struct Grid {
  Header H;
  Conserved C;
  Real dt         = 0;
  int custom_grav = 0;

  void set_dt(Real dti) { dt = kCfl / dti; }

  void Execute_Hydro_Integrator()
  {
    if (H.nx > 1 && H.ny == 1 && H.nz == 1) {
      VL_Algorithm_1D_CUDA(C.device, H.nx, 0, H.n_ghost, H.dx, 0.0, dt, H.n_fields, custom_grav);
    } else {
      VL_Algorithm_3D_CUDA(C.device, nullptr, H.nx, H.ny, H.nz, 0, 0, 0, H.n_ghost, H.dx, H.dy, H.dz, 0.0, 0.0, 0.0,
                           dt, H.n_fields, custom_grav, H.density_floor, nullptr);
    }
  }

  Real Update_Hydro_Grid()
  {
    Execute_Hydro_Integrator();
#ifdef TEMPERATURE_FLOOR
    Apply_Temperature_Floor(C.device, H.nx, H.ny, H.nz, H.n_ghost, H.n_fields, H.U_floor);
#endif
    // grid3D.cpp:119 passes n_cells where n_fields is declared
    return Calc_dt_GPU(C.device, H.nx, H.ny, H.nz, H.n_ghost, H.nx * H.ny * H.nz, H.dx, H.dy, H.dz, gama);
  }

  void Set_Boundary_Extents(int dir, int *imin, int *imax)
  {
    int jl = 0, ju = H.ny, kl = 0, ku = H.nz;
    if (H.ny > 1) jl = H.n_ghost, ju = H.ny - H.n_ghost;
    if (H.nz > 1) kl = H.n_ghost, ku = H.nz - H.n_ghost;
    if (dir < 2) {
      imin[0] = dir == 0 ? 0 : H.nx - H.n_ghost, imax[0] = dir == 0 ? H.n_ghost : H.nx;
      imin[1] = jl, imax[1] = ju, imin[2] = kl, imax[2] = ku;
    } else if (dir < 4) {
      imin[0] = 0, imax[0] = H.nx;
      imin[1] = dir == 2 ? 0 : H.ny - H.n_ghost, imax[1] = dir == 2 ? H.n_ghost : H.ny;
      imin[2] = kl, imax[2] = ku;
    } else {
      imin[0] = 0, imax[0] = H.nx, imin[1] = 0, imax[1] = H.ny;
      imin[2] = dir == 4 ? 0 : H.nz - H.n_ghost, imax[2] = dir == 4 ? H.n_ghost : H.nz;
    }
  }

  void Set_Boundaries(int dir, int flags[])
  {
    if (flags[dir] == 4) return;
    int imin[3], imax[3];
    Set_Boundary_Extents(dir, imin, imax);
    SetGhostCells(C.device, H.nx, H.ny, H.nz, H.n_fields, H.nx * H.ny * H.nz, H.n_ghost, flags, imax[0] - imin[0],
                  imax[1] - imin[1], imax[2] - imin[2], imin[0], imin[1], imin[2], dir);
  }

  void Set_Boundary_Conditions_Grid()
  {
    int flags[6] = {0, 0, 0, 0, 0, 0};
    if (H.nx > 1) flags[0] = H.flags[0], flags[1] = H.flags[1];
    if (H.ny > 1) flags[2] = H.flags[2], flags[3] = H.flags[3];
    if (H.nz > 1) flags[4] = H.flags[4], flags[5] = H.flags[5];
    if (H.nx > 1) Set_Boundaries(0, flags), Set_Boundaries(1, flags);
    if (H.ny > 1) Set_Boundaries(2, flags), Set_Boundaries(3, flags);
    if (H.nz > 1) Set_Boundaries(4, flags), Set_Boundaries(5, flags);
  }
};

void Die(const char *msg)
{
  std::fprintf(stderr, "dropin_driver: %s (%s)\n", msg, chl_last_error());
  std::exit(1);
}
}  // namespace

int main(int argc, char **argv)
{
  if (argc != 3 && argc != 4) {
    std::fprintf(stderr, "usage: %s in.bin out.bin [custom_grav]\n", argv[0]);
    return 2;
  }
  Grid G;
  if (argc == 4) G.custom_grav = std::atoi(argv[3]);
  FILE *f = std::fopen(argv[1], "rb");
  if (!f || std::fread(&G.H, sizeof(Header), 1, f) != 1) Die("cannot read header");
  size_t n = size_t(G.H.n_fields) * G.H.nx * G.H.ny * G.H.nz;
  std::vector<Real> host(n);
  if (std::fread(host.data(), sizeof(Real), n, f) != n) Die("cannot read state");
  std::fclose(f);
  gama = G.H.gamma;
  if (chl_malloc(reinterpret_cast<void **>(&G.C.device), n * sizeof(Real)) != CHL_OK) Die("chl_malloc");
  if (chl_memcpy_htod(G.C.device, host.data(), n * sizeof(Real)) != CHL_OK) Die("upload");

  // main.cpp:190 sets the ghost cells before the loop, :245 computes the first dti
  G.Set_Boundary_Conditions_Grid();
  Real dti   = Calc_dt_GPU(G.C.device, G.H.nx, G.H.ny, G.H.nz, G.H.n_ghost, G.H.nx * G.H.ny * G.H.nz, G.H.dx, G.H.dy,
                           G.H.dz, gama);
  Real t     = G.H.t;
  int32_t ns = 0;
  while (t < G.H.tout && ns < G.H.max_steps) {
    G.set_dt(dti);
    if (t + G.dt > G.H.tout) G.dt = G.H.tout - t;  // main.cpp:258-261
    dti = G.Update_Hydro_Grid();
    t += G.dt;  // Update_Time
    ns++;
    G.Set_Boundary_Conditions_Grid();
#ifdef MHD
    mhd::checkMagneticDivergence(G);
#endif
  }
  if (chl_memcpy_dtoh(host.data(), G.C.device, n * sizeof(Real)) != CHL_OK) Die("download");
  Free_Memory_VL_3D();
  Free_Memory_VL_1D();
  chl_free(G.C.device);
  FILE *o = std::fopen(argv[2], "wb");
  if (!o) Die("cannot write output");
  std::fwrite(&G.H, sizeof(Header), 1, o);
  std::fwrite(host.data(), sizeof(Real), n, o);
  std::fwrite(&ns, sizeof(ns), 1, o);
  std::fwrite(&t, sizeof(t), 1, o);
  std::fclose(o);
  std::printf("dropin_driver: %d steps, t = %.17g\n", ns, t);
  return 0;
}
