// capi.cpp -- the extern "C" boundary of libcholla_amd.so (declared in include/cholla_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <vector>

#include "../../include/cholla_amd.h"
#include "capi_internal.hpp"
#include "snapshot_kernels.hpp"
#include "support_kernels.hpp"
#include "vl3d_launch.hpp"

namespace chl_capi
{
namespace
{
thread_local std::string g_err;
}
int Fail(int code, const std::string &msg)
{
  g_err = msg;
  return code;
}
}  // namespace chl_capi

namespace
{
using chl_capi::Fail;
using chl_capi::kCfl;
using chl_capi::ReconRadius;
}  // namespace


namespace
{

int CheckConfig(const chl_hydro_config &c)
{
  if (c.dual_energy != 0 && c.dual_energy != 1) return Fail(CHL_ERR_INVALID, "dual_energy must be 0 or 1");
  if (c.n_fields != 5 + c.dual_energy)
    return Fail(CHL_ERR_UNSUPPORTED, "n_fields must be 5, or 6 with dual_energy (no passive scalars)");
  if (c.riemann < CHL_RIEMANN_HLLC || c.riemann > CHL_RIEMANN_EXACT)
    return Fail(CHL_ERR_UNSUPPORTED, "riemann must be HLLC, HLL, ROE or EXACT");
  if (c.reconstruction < CHL_RECON_PCM || c.reconstruction > CHL_RECON_PPMP)
    return Fail(CHL_ERR_UNSUPPORTED, "reconstruction must be PCM, PLMC, PPMC, PLMP or PPMP");
  const bool simple_recon = c.reconstruction == CHL_RECON_PCM || c.reconstruction == CHL_RECON_PLMP ||
                            c.reconstruction == CHL_RECON_PPMP;
  if (c.precision < CHL_PREC_FP64_EXACT || c.precision > CHL_PREC_MIXED) return Fail(CHL_ERR_INVALID, "bad precision");
  if (c.positivity_fallback != 0 && c.positivity_fallback != 1)
    return Fail(CHL_ERR_INVALID, "positivity_fallback must be 0 or 1");
  if (c.positivity_fallback && (c.reconstruction != CHL_RECON_PPMC || c.nz == 1))
    return Fail(CHL_ERR_UNSUPPORTED, "the positivity fallback is built for PPMC on 3D boxes only");
  const int ng = c.n_ghost, r = ReconRadius(c.reconstruction);
  if (ng < r + 2)
    return Fail(CHL_ERR_INVALID, "n_ghost too small for the reconstruction (PCM needs 2, PLMC and PLMP 3, PPMC 4)");
  if (c.nx < 2 * ng + 1) return Fail(CHL_ERR_INVALID, "nx must exceed 2*n_ghost");
  bool one_d = c.ny == 1 && c.nz == 1;
  bool two_d = c.nz == 1 && c.ny > 1;
  if (two_d) {
    if (c.ny < 2 * ng + 1) return Fail(CHL_ERR_INVALID, "2D boxes need ny > 2*n_ghost");
  } else if (!one_d) {
    if (c.ny < 2 * ng + 1 || c.nz < 2 * ng + 1) return Fail(CHL_ERR_INVALID, "3D boxes need ny, nz > 2*n_ghost");
    if (c.tile1 >= chl::kNumTiles || c.tile2 >= chl::kNumTiles) return Fail(CHL_ERR_INVALID, "tile index out of range");
    if (c.dual_energy && (c.tile1 >= chl::kNumDeTiles || c.tile2 >= chl::kNumDeTiles))
      return Fail(CHL_ERR_UNSUPPORTED, "dual energy is built for tile shapes 0 and 1 only");
    if (c.riemann != CHL_RIEMANN_HLLC && (c.tile1 >= chl::kNumDeTiles || c.tile2 >= chl::kNumDeTiles))
      return Fail(CHL_ERR_UNSUPPORTED, "the HLL, Roe and exact solvers are built for tile shapes 0 and 1 only");
    if (simple_recon && (c.tile1 >= chl::kNumDeTiles || c.tile2 >= chl::kNumDeTiles))
      return Fail(CHL_ERR_UNSUPPORTED, "PCM, PLMP and PPMP are built for tile shapes 0 and 1 only");
    if (c.tile_order != 0 && c.tile_order != 1) return Fail(CHL_ERR_INVALID, "tile_order must be 0 or 1");
  }
  // 2^31 guard on per-axis products used in 32-bit grid arithmetic of launch shapes
  if (double(c.nx) * c.ny * c.nz * 5 > 9.0e15) return Fail(CHL_ERR_INVALID, "box too large");
  for (int q = 0; q < 6; q++)
    if (c.bc_flags[q] < 1 || c.bc_flags[q] > 3)
      return Fail(CHL_ERR_UNSUPPORTED, "boundary flags must be 1 (periodic), 2 (reflective) or 3 (transmissive)");
  return CHL_OK;
}

int SolverKind(int riemann)
{
  if (riemann == CHL_RIEMANN_HLL) return chl::kRiemannHll;
  if (riemann == CHL_RIEMANN_ROE) return chl::kRiemannRoe;
  if (riemann == CHL_RIEMANN_EXACT) return chl::kRiemannExact;
  return chl::kRiemannHllc;
}

int DefaultChunk(int n_real) { return std::max(8, std::min(n_real, 32)); }

// Gravity records a step keeps for repaired cells (static gravity only).
constexpr int kRepairGravCap = 4096;

// The repair scratch of a 3D box: counts, gravity records, the side buffer of the six
// boundary layers, one mark byte per 4096 cells and one per cell.  RepairBytes sizes it,
// PlaceRepair points h->rep into it.
size_t RepairLayout(int nx, int ny, int nz, int ng_, int grav_cap, size_t off[4])
{
  const size_t ng = size_t(ng_), Lx = nx - 2 * ng, Ly = ny - 2 * ng, Lz = nz - 2 * ng;
  const size_t n_cells = size_t(nx) * ny * nz;
  const size_t chunk = size_t(1) << chl::kRepairChunkShift;
  const size_t n_chunks = (n_cells + chunk - 1) / chunk;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  off[0] = 256;                                                            // gravity records
  off[1] = off[0] + up(size_t(grav_cap) * 5 * sizeof(double));            // side buffer
  off[2] = off[1] + up(2 * (Ly * Lz + Lx * Lz + Lx * Ly) * sizeof(double));  // chunk marks
  off[3] = off[2] + up(n_chunks);                                          // cell marks
  return off[3] + up(n_chunks * chunk);
}

hipError_t AllocRepair(chl_hydro *h)
{
  const size_t bytes = chl_capi::RepairBytes(h->cfg);
  hipError_t e       = hipMalloc(&h->rep_buf, bytes);
  if (e != hipSuccess) return e;
  chl_capi::PlaceRepair(h, h->rep_buf);
  return hipMemsetAsync(h->rep_buf, 0, bytes, h->stream);
}

// Arguments of the fused stage kernels for this box (kchunk is set per stage).
chl::StepArgs MakeStepArgs(chl_hydro *h, double *U, double dx, double dy, double dz, double gamma, double dt_host,
                           const double *dt_dev, double density_floor, double U_floor, int flags)
{
  const chl_hydro_config &c = h->cfg;
  chl::StepArgs a{};
  a.U = U, a.W = h->W, a.pdv = h->pdv, a.nx = c.nx, a.ny = c.ny, a.nz = c.nz, a.ng = c.n_ghost;
  a.dx = dx, a.dy = dy, a.dz = dz, a.gamma = gamma, a.dt = dt_host, a.dt_dev = dt_dev, a.dti = h->dti();
  a.density_floor = density_floor, a.U_floor = U_floor;
  a.flags     = flags;
  a.xcd_order = c.tile_order == 0;
  a.de        = c.dual_energy;
  a.riemann   = SolverKind(c.riemann);
  a.grav      = h->grav;
  a.rep       = h->rep;
  return a;
}

int LaunchStage1(chl_hydro *h, chl::StepArgs &a, int tile = -1, int kchunk = -1)
{
  const chl_hydro_config &c = h->cfg;
  a.kchunk     = kchunk > 0 ? kchunk : c.kchunk1;
  const int t  = tile >= 0 ? tile : c.tile1;
  const int R1 = ReconRadius(c.reconstruction) + 1;
  CHL_HIP(c.precision == CHL_PREC_FP64_EXACT ? chl::LaunchStage1Exact(a, R1, t, h->stream)
                                             : chl::LaunchStage1Fast(c.precision, a, R1, t, h->stream));
  return CHL_OK;
}

int LaunchStage2(chl_hydro *h, chl::StepArgs &a, int tile = -1, int kchunk = -1)
{
  const chl_hydro_config &c = h->cfg;
  a.kchunk    = kchunk > 0 ? kchunk : c.kchunk2;
  const int t = tile >= 0 ? tile : c.tile2;
  CHL_HIP(c.precision == CHL_PREC_FP64_EXACT ? chl::LaunchStage2Exact(a, c.reconstruction, t, h->stream)
                                             : chl::LaunchStage2Fast(c.precision, a, c.reconstruction, t, h->stream));
  return CHL_OK;
}

// Dual energy: the internal-energy selection and energy sync after stage 2
// (VL_3D_cuda.cu:349-357), carrying the floor / dt / ghost work stage 2 leaves out.
int LaunchDualEnergy(chl_hydro *h, const chl::StepArgs &a)
{
  if (!h->cfg.dual_energy) return CHL_OK;
  const chl_hydro_config &c = h->cfg;
  CHL_HIP(chl::LaunchDeSelect(a.U, c.nx, c.ny, c.nz, c.n_ghost, h->stream));
  CHL_HIP(chl::LaunchDeSync(a.U, c.nx, c.ny, c.nz, c.n_ghost, a.flags, a.U_floor, a.dx, a.dy, a.dz, a.gamma, a.dti,
                            h->stream));
  return CHL_OK;
}

// Enqueue one VL step.  mode: 0 = VL only, 1 = VL + floors + dti, (fill) ghost refresh.
int EnqueueStep(chl_hydro *h, double *U, double dx, double dy, double dz, double gamma, double dt_host,
                const double *dt_dev, double density_floor, double U_floor, bool update_grid, bool fill)
{
  const chl_hydro_config &c = h->cfg;
  const int ng              = c.n_ghost;
  if (h->one_d) {
    const int n = c.nx;
    // the dt reduction rides on the last kernel of the step: the one-block stage 2, or with
    // dual energy the energy sync (any number of blocks)
    // (stage 1 zeroes the accumulator in-kernel, so a captured step holds no memset node)
    bool fuse_dt = update_grid && (n <= 1024 || c.dual_energy);
    CHL_HIP(chl::LaunchVl1d(U, reinterpret_cast<double *>(h->W), n, ng, c.reconstruction, SolverKind(c.riemann),
                            c.dual_energy != 0, c.precision, dx, dt_dev, dt_host, gamma, fuse_dt ? h->dti() : nullptr,
                            update_grid ? h->dti() : nullptr, h->grav, h->stream));
    if (update_grid && !fuse_dt)
      CHL_HIP(chl::LaunchCalcDt(U, n, 1, 1, ng, dx, dy, dz, gamma, h->dti(), h->num_cu, h->stream));
    if (fill) CHL_HIP(chl::LaunchGhostFill(U, c.nx, c.ny, c.nz, c.n_fields, ng, c.bc_flags, h->stream));
    return CHL_OK;
  }
  if (h->two_d) {  // no temperature floor: the reference's guard admits no 2D cell (hydro_cuda.cu:1130-1131)
    CHL_HIP(chl::LaunchVl2d(U, reinterpret_cast<double *>(h->W), c.nx, c.ny, ng, c.reconstruction,
                            SolverKind(c.riemann), c.dual_energy != 0, c.precision, dx, dy, dt_dev, dt_host, gamma,
                            update_grid ? h->dti() : nullptr, h->grav, h->stream));
    if (fill) CHL_HIP(chl::LaunchGhostFill(U, c.nx, c.ny, c.nz, c.n_fields, ng, c.bc_flags, h->stream));
    return CHL_OK;
  }
  // A build without both floors repairs crashed cells from the 3x3x3 block of the update
  // alone (Average_Cell_All_Fields, hydro_cuda.cu:375-381): there the temperature floor, the
  // dt reduction and the ghost fill run as their own kernels after the repairs, not fused
  // into stage 2 (the dual-energy kernels, which carry them otherwise, already follow).
  const bool crash   = !(c.density_floor && c.temperature_floor);
  const bool unfused = crash && !c.dual_energy;
  int flags          = 0;
  if (c.density_floor) flags |= chl::kFlagDensityFloor;
  if (crash) flags |= chl::kFlagCrashRepair;
  if (update_grid && c.temperature_floor && !unfused) flags |= chl::kFlagTemperatureFloor;
  if (update_grid && !unfused) flags |= chl::kFlagReduceDt;
  if (update_grid) flags |= chl::kFlagResetDti;
  if (c.positivity_fallback) flags |= chl::kFlagPositivityFallback;
  if (c.reconstruction == CHL_RECON_PCM) flags |= chl::kFlagConservedW;
  const bool fused_ghosts = fill && h->PeriodicFusable() && !unfused;
  if (fused_ghosts) flags |= chl::kFlagPeriodicGhosts;
  chl::StepArgs a = MakeStepArgs(h, U, dx, dy, dz, gamma, dt_host, dt_dev, density_floor, U_floor, flags);
  // stage 2 overwrites ghost cells with new periodic images: a density repair next to one
  // reads the pre-step value from the side buffer instead
  if (!(fused_ghosts && c.density_floor && !c.dual_energy)) a.rep.side = nullptr;
  if (a.rep.side) CHL_HIP(chl::LaunchRepairSide(a, h->stream));
  CHL_TRY(LaunchStage1(h, a));
  CHL_TRY(LaunchStage2(h, a));
  CHL_HIP(c.precision == CHL_PREC_FP64_EXACT ? chl::LaunchRepairExact(a, h->stream)
                                             : chl::LaunchRepairFast(c.precision, a, h->stream));
  CHL_TRY(LaunchDualEnergy(h, a));
  if (unfused && update_grid) {
    if (c.temperature_floor)
      CHL_HIP(chl::LaunchTemperatureFloor(U, c.nx, c.ny, c.nz, ng, U_floor, c.n_fields, h->stream));
    CHL_HIP(chl::LaunchCalcDt(U, c.nx, c.ny, c.nz, ng, dx, dy, dz, gamma, h->dti(), h->num_cu, h->stream));
  }
  if (fill && !fused_ghosts) CHL_HIP(chl::LaunchGhostFill(U, c.nx, c.ny, c.nz, c.n_fields, ng, c.bc_flags, h->stream));
  return CHL_OK;
}

int ReadDti(chl_hydro *h, double *dti_out)
{
  unsigned long long bits = 0;
  CHL_HIP(hipMemcpyAsync(&bits, h->dti(), 8, hipMemcpyDeviceToHost, h->stream));
  CHL_HIP(hipStreamSynchronize(h->stream));
  double v;
  std::memcpy(&v, &bits, 8);
  *dti_out = v;
  return CHL_OK;
}

int InitialDti(chl_hydro *h, const double *U, double dx, double dy, double dz, double gamma)
{
  const chl_hydro_config &c = h->cfg;
  CHL_HIP(hipMemsetAsync(h->dti(), 0, 8, h->stream));
  CHL_HIP(chl::LaunchCalcDt(U, c.nx, c.ny, c.nz, c.n_ghost, dx, dy, dz, gamma, h->dti(), h->num_cu, h->stream));
  return CHL_OK;
}

int SetTimeState(chl_hydro *h, double t0)
{
  unsigned char buf[16];
  double z = 0.0;
  int n0   = 0;
  std::memcpy(buf, &t0, 8);
  std::memcpy(buf + 8, &n0, 4);
  std::memset(buf + 12, 0, 4);
  (void)z;
  CHL_HIP(hipMemcpyAsync(h->t(), buf, 16, hipMemcpyHostToDevice, h->stream));
  return CHL_OK;
}

// One device-timed step: dt from dti on the device, then the fused step with ghost refresh.
int EnqueueDeviceStep(chl_hydro *h, double *U, double dx, double dy, double dz, double gamma, double density_floor,
                      double U_floor, double tout)
{
  CHL_HIP(chl::LaunchDtAdvance(h->dti(), h->dt(), h->t(), tout, kCfl, h->nstep(), h->stream));
  return EnqueueStep(h, U, dx, dy, dz, gamma, 0.0, h->dt(), density_floor, U_floor, true, true);
}

int EnsureGraph(chl_hydro *h, double *U, double dx, double dy, double dz, double gamma, double density_floor,
                double U_floor, double tout)
{
  const double key[7] = {dx, dy, dz, gamma, density_floor, U_floor, tout};
  if (h->graph && h->g_U == U && std::memcmp(key, h->g_key, sizeof(key)) == 0) return CHL_OK;
  if (h->graph) {
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  hipGraph_t g;
  CHL_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc        = EnqueueDeviceStep(h, U, dx, dy, dz, gamma, density_floor, U_floor, tout);
  hipError_t e2 = hipStreamEndCapture(h->stream, &g);
  if (rc != CHL_OK) return rc;
  CHL_HIP(e2);
  hipError_t e3 = hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  CHL_HIP(e3);
  h->g_U = U;
  std::memcpy(h->g_key, key, sizeof(key));
  return CHL_OK;
}

}  // namespace

size_t chl_capi::RepairBytes(const chl_hydro_config &c)
{
  return RepairBytesFor(c.nx, c.ny, c.nz, c.n_ghost, kRepairGravCap);
}

size_t chl_capi::RepairBytesFor(int nx, int ny, int nz, int ng, int grav_cap)
{
  size_t off[4];
  return RepairLayout(nx, ny, nz, ng, grav_cap, off);
}

chl::RepairState chl_capi::RepairAt(void *base, int nx, int ny, int nz, int ng, int grav_cap)
{
  size_t off[4];
  (void)RepairLayout(nx, ny, nz, ng, grav_cap, off);
  unsigned char *b = static_cast<unsigned char *>(base);
  chl::RepairState r{};
  r.count    = reinterpret_cast<unsigned *>(b);
  r.grav     = grav_cap > 0 ? reinterpret_cast<double *>(b + off[0]) : nullptr;
  r.grav_cap = grav_cap;
  r.side     = reinterpret_cast<double *>(b + off[1]);
  r.chunk    = b + off[2];
  r.mark     = b + off[3];
  return r;
}

void chl_capi::PlaceRepair(chl_hydro *h, void *base)
{
  const chl_hydro_config &c = h->cfg;
  h->rep                    = RepairAt(base, c.nx, c.ny, c.nz, c.n_ghost, kRepairGravCap);
}

size_t chl_capi::HalfStepBytes(const chl_hydro_config &c, size_t *w_main)
{
  const size_t n    = size_t(c.nx) * c.ny * c.nz;
  const size_t elem = c.precision == CHL_PREC_MIXED ? sizeof(float) : sizeof(double);
  // W: 5 half-step fields, the pressure's low part in the mixed form, GasEnergy with DE
  const size_t w_fields = 5 + size_t(c.precision == CHL_PREC_MIXED) + size_t(c.dual_energy);
  const size_t main_    = (w_fields * n * elem + 255) / 256 * 256;  // pdV sources start aligned
  if (w_main) *w_main = main_;
  return main_ + (c.dual_energy ? n * sizeof(double) : 0);
}

int chl_capi::PrepareConfig(chl_hydro_config &c)
{
  if (c.tile1 < 0) c.tile1 = 0;
  const bool narrow = c.dual_energy || c.riemann != CHL_RIEMANN_HLLC || c.reconstruction == CHL_RECON_PCM ||
                      c.reconstruction == CHL_RECON_PLMP || c.reconstruction == CHL_RECON_PPMP;  // kernels built for tile shapes 0 and 1 only
  if (c.tile2 < 0) c.tile2 = narrow ? 0 : 3;
  CHL_TRY(CheckConfig(c));
  if (c.kchunk1 <= 0) c.kchunk1 = DefaultChunk(c.nz - 2 * c.n_ghost + 2 * (ReconRadius(c.reconstruction) + 1));
  if (c.kchunk2 <= 0) c.kchunk2 = DefaultChunk(c.nz - 2 * c.n_ghost);
  return CHL_OK;
}

int chl_capi::EnqueueVlStep(chl_hydro *h, double *U, double dx, double dy, double dz, double dt, double gamma,
                            double density_floor)
{
  return EnqueueStep(h, U, dx, dy, dz, gamma, dt, nullptr, density_floor, 0.0, false, false);
}

extern "C" {

const char *chl_version(void) { return "cholla_amd 0.1 (gfx950)"; }
const char *chl_last_error(void) { return chl_capi::g_err.c_str(); }

int chl_device_count(int *n)
{
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *n = c;
  return CHL_OK;
}

int chl_device_name(char *buf, int len)
{
  int dev;
  CHL_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  CHL_HIP(hipGetDeviceProperties(&p, dev));
  std::snprintf(buf, size_t(len), "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
  return CHL_OK;
}

int chl_malloc(void **ptr, size_t bytes)
{
  CHL_HIP(hipMalloc(ptr, bytes));
  return CHL_OK;
}
int chl_free(void *ptr)
{
  CHL_HIP(hipFree(ptr));
  return CHL_OK;
}
int chl_memcpy_htod(void *dst, const void *src, size_t bytes)
{
  CHL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return CHL_OK;
}
int chl_memcpy_dtoh(void *dst, const void *src, size_t bytes)
{
  CHL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return CHL_OK;
}
int chl_memcpy_dtod(void *dst, const void *src, size_t bytes)
{
  CHL_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return CHL_OK;
}
int chl_synchronize(void)
{
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int chl_hydro_create(const chl_hydro_config *cfg, chl_hydro **out)
{
  if (!cfg || !out) return Fail(CHL_ERR_INVALID, "null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return Fail(CHL_ERR_NO_DEVICE, "no HIP device");
  chl_hydro_config c = *cfg;
  CHL_TRY(chl_capi::PrepareConfig(c));
  auto *h    = new chl_hydro();
  h->tile2_auto = cfg->tile2 < 0;
  h->one_d   = c.ny == 1 && c.nz == 1;
  h->two_d   = c.nz == 1 && c.ny > 1;
  h->n_cells = size_t(c.nx) * c.ny * c.nz;
  h->cfg = c;
  int dev;
  hipDeviceProp_t p;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) {
    delete h;
    return Fail(CHL_ERR_HIP, "hipGetDeviceProperties failed");
  }
  h->num_cu        = p.multiProcessorCount;
  size_t w_main    = 0;
  h->W_bytes       = chl_capi::HalfStepBytes(c, &w_main);
  if (h->two_d) h->W_bytes = 42 * h->n_cells * sizeof(double);  // half state, 4 interface, 2 flux arrays of 6 fields
  if (h->one_d)  // the 1D integrator's half state is FP64 in every form
    h->W_bytes = std::max(h->W_bytes, (5 + size_t(c.dual_energy)) * h->n_cells * sizeof(double));
  hipError_t e     = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&h->W, h->W_bytes);
  if (e == hipSuccess && c.dual_energy)
    h->pdv = reinterpret_cast<double *>(static_cast<unsigned char *>(h->W) + w_main);
  if (e == hipSuccess) e = hipMemsetAsync(h->W, 0, h->W_bytes, h->stream);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&h->scal), 64);
  if (e == hipSuccess) e = hipMemsetAsync(h->scal, 0, 64, h->stream);
  if (e == hipSuccess && !h->one_d && !h->two_d) e = AllocRepair(h);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    chl_hydro_destroy(h);
    return Fail(CHL_ERR_HIP, std::string("allocating integrator scratch: ") + hipGetErrorString(e));
  }
  *out = h;
  return CHL_OK;
}

int chl_hydro_destroy(chl_hydro *h)
{
  if (!h) return CHL_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->graph) (void)hipGraphExecDestroy(h->graph);
  if (h->W) (void)hipFree(h->W);
  if (h->rep_buf) (void)hipFree(h->rep_buf);
  if (h->scal) (void)hipFree(h->scal);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return CHL_OK;
}

int chl_hydro_get_config(const chl_hydro *h, chl_hydro_config *out)
{
  if (!h || !out) return Fail(CHL_ERR_INVALID, "null argument");
  *out = h->cfg;
  return CHL_OK;
}

int chl_hydro_vl_step(chl_hydro *h, double *U, double dx, double dy, double dz, double dt, double gamma,
                      double density_floor)
{
  if (!h || !U) return Fail(CHL_ERR_INVALID, "null argument");
  CHL_TRY(EnqueueStep(h, U, dx, dy, dz, gamma, dt, nullptr, density_floor, 0.0, false, false));
  CHL_HIP(hipStreamSynchronize(h->stream));
  return CHL_OK;
}

int chl_hydro_update_grid(chl_hydro *h, double *U, double dx, double dy, double dz, double dt, double gamma,
                          double density_floor, double U_floor, int fill_ghosts, double *dti_out)
{
  if (!h || !U) return Fail(CHL_ERR_INVALID, "null argument");
  CHL_TRY(EnqueueStep(h, U, dx, dy, dz, gamma, dt, nullptr, density_floor, U_floor, true, fill_ghosts != 0));
  if (dti_out) return ReadDti(h, dti_out);
  CHL_HIP(hipStreamSynchronize(h->stream));
  return CHL_OK;
}

int chl_hydro_run(chl_hydro *h, double *U, double dx, double dy, double dz, double gamma, double density_floor,
                  double U_floor, double t0, double tout, int max_steps, int use_graph, double *t_out, int *steps_out)
{
  if (!h || !U) return Fail(CHL_ERR_INVALID, "null argument");
  CHL_TRY(InitialDti(h, U, dx, dy, dz, gamma));
  CHL_TRY(SetTimeState(h, t0));
  if (use_graph) CHL_TRY(EnsureGraph(h, U, dx, dy, dz, gamma, density_floor, U_floor, tout));
  const int check_every = 32;
  double t              = t0;
  int done              = 0;
  while (done < max_steps && t < tout) {
    int batch = std::min(check_every, max_steps - done);
    for (int s = 0; s < batch; s++) {
      if (use_graph) {
        CHL_HIP(hipGraphLaunch(h->graph, h->stream));
      } else {
        CHL_TRY(EnqueueDeviceStep(h, U, dx, dy, dz, gamma, density_floor, U_floor, tout));
      }
    }
    done += batch;
    CHL_HIP(hipMemcpyAsync(&t, h->t(), 8, hipMemcpyDeviceToHost, h->stream));
    CHL_HIP(hipStreamSynchronize(h->stream));
  }
  int ns = 0;
  CHL_HIP(hipMemcpy(&ns, h->nstep(), 4, hipMemcpyDeviceToHost));
  if (t_out) *t_out = t;
  if (steps_out) *steps_out = ns;
  return CHL_OK;
}

int chl_hydro_time_steps(chl_hydro *h, double *U, double dx, double dy, double dz, double gamma, int n_steps,
                         int use_graph, float *ms_out, double *dt_last_out)
{
  if (!h || !U || n_steps < 0) return Fail(CHL_ERR_INVALID, "bad argument");
  const double tout = std::numeric_limits<double>::infinity();
  CHL_TRY(InitialDti(h, U, dx, dy, dz, gamma));
  CHL_TRY(SetTimeState(h, 0.0));
  if (use_graph) CHL_TRY(EnsureGraph(h, U, dx, dy, dz, gamma, 0.0, 0.0, tout));
  hipEvent_t e0, e1;
  CHL_HIP(hipEventCreate(&e0));
  CHL_HIP(hipEventCreate(&e1));
  CHL_HIP(hipStreamSynchronize(h->stream));
  CHL_HIP(hipEventRecord(e0, h->stream));
  for (int s = 0; s < n_steps; s++) {
    if (use_graph) {
      CHL_HIP(hipGraphLaunch(h->graph, h->stream));
    } else {
      CHL_TRY(EnqueueDeviceStep(h, U, dx, dy, dz, gamma, 0.0, 0.0, tout));
    }
  }
  CHL_HIP(hipEventRecord(e1, h->stream));
  CHL_HIP(hipEventSynchronize(e1));
  float ms = 0;
  CHL_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms_out) *ms_out = ms;
  if (dt_last_out) CHL_HIP(hipMemcpy(dt_last_out, h->dt(), 8, hipMemcpyDeviceToHost));
  return CHL_OK;
}

int chl_hydro_kernel_times(chl_hydro *h, double *U, double dx, double dy, double dz, double gamma, int n_steps,
                           float *ms1, float *ms2, float *ms3)
{
  if (!h || !U || h->one_d || h->two_d) return Fail(CHL_ERR_INVALID, "bad argument (3D boxes only)");
  const chl_hydro_config &c = h->cfg;
  const double tout         = std::numeric_limits<double>::infinity();
  CHL_TRY(InitialDti(h, U, dx, dy, dz, gamma));
  CHL_TRY(SetTimeState(h, 0.0));
  hipEvent_t ev[5];
  for (auto &e : ev) CHL_HIP(hipEventCreate(&e));
  double acc[3] = {0, 0, 0};
  int flags = chl::kFlagReduceDt;
  if (c.density_floor) flags |= chl::kFlagDensityFloor;
  if (c.temperature_floor) flags |= chl::kFlagTemperatureFloor;
  if (c.positivity_fallback) flags |= chl::kFlagPositivityFallback;
  if (c.reconstruction == CHL_RECON_PCM) flags |= chl::kFlagConservedW;
  const bool fused_ghosts = h->PeriodicFusable();
  if (fused_ghosts) flags |= chl::kFlagPeriodicGhosts;
  chl::StepArgs a = MakeStepArgs(h, U, dx, dy, dz, gamma, 0.0, h->dt(), 0.0, 0.0, flags);
  if (!(fused_ghosts && c.density_floor && !c.dual_energy)) a.rep.side = nullptr;
  for (int s = 0; s < n_steps; s++) {
    CHL_HIP(hipEventRecord(ev[0], h->stream));
    CHL_HIP(chl::LaunchDtAdvance(h->dti(), h->dt(), h->t(), tout, kCfl, h->nstep(), h->stream));
    if (a.rep.side) CHL_HIP(chl::LaunchRepairSide(a, h->stream));
    CHL_HIP(hipEventRecord(ev[1], h->stream));
    CHL_TRY(LaunchStage1(h, a));
    CHL_HIP(hipEventRecord(ev[2], h->stream));
    CHL_TRY(LaunchStage2(h, a));
    CHL_HIP(hipEventRecord(ev[3], h->stream));
    CHL_HIP(c.precision == CHL_PREC_FP64_EXACT ? chl::LaunchRepairExact(a, h->stream)
                                               : chl::LaunchRepairFast(c.precision, a, h->stream));
    CHL_TRY(LaunchDualEnergy(h, a));
    CHL_HIP(hipEventRecord(ev[4], h->stream));
    if (!fused_ghosts)
      CHL_HIP(chl::LaunchGhostFill(U, c.nx, c.ny, c.nz, c.n_fields, c.n_ghost, c.bc_flags, h->stream));
    CHL_HIP(hipEventSynchronize(ev[4]));
    float t01, t12, t23, t34;
    CHL_HIP(hipEventElapsedTime(&t01, ev[0], ev[1]));
    CHL_HIP(hipEventElapsedTime(&t12, ev[1], ev[2]));
    CHL_HIP(hipEventElapsedTime(&t23, ev[2], ev[3]));
    CHL_HIP(hipEventElapsedTime(&t34, ev[3], ev[4]));
    acc[0] += t12, acc[1] += t23, acc[2] += t01 + t34;
  }
  CHL_HIP(hipStreamSynchronize(h->stream));
  for (auto &e : ev) (void)hipEventDestroy(e);
  if (ms1) *ms1 = float(acc[0] / std::max(1, n_steps));
  if (ms2) *ms2 = float(acc[1] / std::max(1, n_steps));
  if (ms3) *ms3 = float(acc[2] / std::max(1, n_steps));
  return CHL_OK;
}

int chl_hydro_set_static_gravity(chl_hydro *h, int custom_grav, int x_off, int y_off, int z_off, double xbound,
                                 double ybound, double zbound)
{
  if (!h) return Fail(CHL_ERR_INVALID, "null handle");
  const chl_hydro_config &c = h->cfg;
  const int max_id          = h->two_d ? 4 : 1;  // calc_g_2D has four fields, calc_g_1D / calc_g_3D one
  if (custom_grav < 0 || custom_grav > max_id) return Fail(CHL_ERR_INVALID, "custom_grav out of range for this box");
  const bool grav3d = custom_grav != 0 && !h->one_d && !h->two_d;
  if (grav3d && c.tile2 >= chl::kNumDeTiles) {
    if (!h->tile2_auto) return Fail(CHL_ERR_UNSUPPORTED, "3D static gravity is built for stage-2 tile shapes 0 and 1");
    // the library's default stage-2 tile has no gravity kernel: take the 16x16 one
    h->cfg.tile2 = 1;
    if (h->graph) {
      (void)hipGraphExecDestroy(h->graph);
      h->graph = nullptr;
    }
  }
  chl::GravSpec g{custom_grav, {x_off, y_off, z_off}, {xbound, ybound, zbound}};
  const chl::GravSpec &o = h->grav;
  const bool same = o.id == g.id && o.off[0] == g.off[0] && o.off[1] == g.off[1] && o.off[2] == g.off[2] &&
                    o.bound[0] == g.bound[0] && o.bound[1] == g.bound[1] && o.bound[2] == g.bound[2];
  if (same) return CHL_OK;
  h->grav = g;
  if (h->graph) {  // the captured step holds the old arguments
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  return CHL_OK;
}

int chl_hydro_autotune(chl_hydro *h, const double *U, double dx, double dy, double dz, double gamma, int steps,
                       int *tile1, int *tile2, int *kchunk1, int *kchunk2)
{
  if (!h || !U || h->one_d || h->two_d) return Fail(CHL_ERR_INVALID, "bad argument (3D boxes only)");
  chl_hydro_config &c = h->cfg;
  const size_t bytes  = size_t(c.n_fields) * h->n_cells * sizeof(double);
  double *S           = nullptr;
  CHL_HIP(hipMalloc(&S, bytes));
  hipEvent_t e0, e1;
  CHL_HIP(hipEventCreate(&e0));
  CHL_HIP(hipEventCreate(&e1));
  // a small fixed dt keeps every candidate on the same, physical state
  double dt_fixed = 0;
  {
    double dti = 0;
    CHL_TRY(chl_calc_dt(U, c.nx, c.ny, c.nz, c.n_ghost, c.n_fields, dx, dy, dz, gamma, &dti));
    dt_fixed = 0.1 * kCfl / dti;
  }
  chl::StepArgs a = MakeStepArgs(h, S, dx, dy, dz, gamma, dt_fixed, nullptr, 0.0, 0.0,
                                 chl::kFlagReduceDt | (c.density_floor ? chl::kFlagDensityFloor : 0) |
                                     (c.temperature_floor ? chl::kFlagTemperatureFloor : 0) |
                                     (c.positivity_fallback ? chl::kFlagPositivityFallback : 0) |
                                     (c.reconstruction == CHL_RECON_PCM ? chl::kFlagConservedW : 0));
  a.rep = chl::RepairState{};  // timing runs on a scratch copy: nothing is marked or repaired
  const int R1     = ReconRadius(c.reconstruction) + 1;
  const int nz1    = c.nz - 2 * c.n_ghost + 2 * R1, nz2 = c.nz - 2 * c.n_ghost;
  const bool narrow = c.dual_energy || c.riemann != CHL_RIEMANN_HLLC || c.reconstruction == CHL_RECON_PCM ||
                     c.reconstruction == CHL_RECON_PLMP || c.reconstruction == CHL_RECON_PPMP || h->grav.id != 0;
  const int ntiles  = narrow ? chl::kNumDeTiles : chl::kNumTiles;
  std::vector<int> chunks = {8, 16, 32, 64, 128};
  float best1 = 1e30f, best2 = 1e30f;
  int b_t1 = c.tile1, b_t2 = c.tile2, b_k1 = c.kchunk1, b_k2 = c.kchunk2;
  int rc = CHL_OK;
  for (int t = 0; t < ntiles && rc == CHL_OK; t++) {
    for (int kc : chunks) {
      if (kc > 2 * nz1 && kc != chunks.front()) continue;
      const int k1 = std::min(kc, nz1);
      float ms     = 0;
      for (int rep = 0; rep < 2 && rc == CHL_OK; rep++) {  // rep 0 warms up
        if (hipMemcpyAsync(S, U, bytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess) rc = CHL_ERR_HIP;
        (void)hipEventRecord(e0, h->stream);
        for (int s = 0; s < steps && rc == CHL_OK; s++) rc = LaunchStage1(h, a, t, k1);
        (void)hipEventRecord(e1, h->stream);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
      }
      if (ms < best1) best1 = ms, b_t1 = t, b_k1 = k1;
    }
  }
  for (int t = 0; t < ntiles && rc == CHL_OK; t++) {
    for (int kc : chunks) {
      if (kc > 2 * nz2 && kc != chunks.front()) continue;
      const int k2 = std::min(kc, nz2);
      float ms     = 0;
      for (int rep = 0; rep < 2 && rc == CHL_OK; rep++) {
        if (hipMemcpyAsync(S, U, bytes, hipMemcpyDeviceToDevice, h->stream) != hipSuccess) rc = CHL_ERR_HIP;
        if (rc == CHL_OK) rc = LaunchStage1(h, a, b_t1, b_k1);
        (void)hipEventRecord(e0, h->stream);
        for (int s = 0; s < steps && rc == CHL_OK; s++) rc = LaunchStage2(h, a, t, k2);
        (void)hipEventRecord(e1, h->stream);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
      }
      if (ms < best2) best2 = ms, b_t2 = t, b_k2 = k2;
    }
  }
  (void)hipStreamSynchronize(h->stream);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(S);
  if (rc != CHL_OK) return rc;
  c.tile1 = b_t1, c.tile2 = b_t2, c.kchunk1 = b_k1, c.kchunk2 = b_k2;
  if (h->graph) {
    (void)hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  if (tile1) *tile1 = b_t1;
  if (tile2) *tile2 = b_t2;
  if (kchunk1) *kchunk1 = b_k1;
  if (kchunk2) *kchunk2 = b_k2;
  return CHL_OK;
}

int chl_calc_dt(const double *U, int nx, int ny, int nz, int n_ghost, int n_fields, double dx, double dy, double dz,
                double gamma, double *dti_out)
{
  (void)n_fields;  // unused in 3D, as in the reference (grid3D.cpp:119 passes n_cells here)
  if (!U || !dti_out) return Fail(CHL_ERR_INVALID, "null argument");
  unsigned long long *d = nullptr;
  CHL_HIP(hipMalloc(&d, 8));
  int num_cu = 256, dev;
  hipDeviceProp_t p;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) num_cu = p.multiProcessorCount;
  hipError_t e = hipMemset(d, 0, 8);
  if (e == hipSuccess) e = chl::LaunchCalcDt(U, nx, ny, nz, n_ghost, dx, dy, dz, gamma, d, num_cu, nullptr);
  unsigned long long bits = 0;
  if (e == hipSuccess) e = hipMemcpy(&bits, d, 8, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  CHL_HIP(e);
  std::memcpy(dti_out, &bits, 8);
  return CHL_OK;
}

int chl_apply_temperature_floor(double *U, int nx, int ny, int nz, int n_ghost, int n_fields, double U_floor)
{
  if (!U) return Fail(CHL_ERR_INVALID, "null argument");
  CHL_HIP(chl::LaunchTemperatureFloor(U, nx, ny, nz, n_ghost, U_floor, n_fields, nullptr));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

}  // extern "C"

namespace
{
int SetGhostCellsImpl(double *c_head, int nx, int ny, int nz, int n_fields, int n_cells, int n_ghost,
                      const int flags[6], int isize, int jsize, int ksize, int imin, int jmin, int kmin, int dir,
                      bool mhd)
{
  if (!c_head || !flags) return Fail(CHL_ERR_INVALID, "null argument");
  if (long(nx) * ny * nz != long(n_cells)) return Fail(CHL_ERR_INVALID, "n_cells != nx*ny*nz");
  if (dir < 0 || dir > 5) return Fail(CHL_ERR_INVALID, "dir must be 0..5");
  if (isize < 0 || jsize < 0 || ksize < 0 || imin < 0 || jmin < 0 || kmin < 0 || imin + isize > nx ||
      jmin + jsize > ny || kmin + ksize > nz)
    return Fail(CHL_ERR_INVALID, "slab extents outside the grid");
  if (flags[dir] == 4 || flags[dir] == 5) return CHL_OK;
  CHL_HIP(chl::LaunchGhostSlab(c_head, nx, ny, nz, n_fields, n_ghost, flags, isize, jsize, ksize, imin, jmin, kmin,
                               dir, nullptr, mhd));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int SetBoundariesImpl(double *U, int nx, int ny, int nz, int n_fields, int n_ghost, const int flags_in[6], bool mhd)
{
  if (!U || !flags_in) return Fail(CHL_ERR_INVALID, "null argument");
  int flags[6] = {0, 0, 0, 0, 0, 0};
  if (nx > 1) flags[0] = flags_in[0], flags[1] = flags_in[1];
  if (ny > 1) flags[2] = flags_in[2], flags[3] = flags_in[3];
  if (nz > 1) flags[4] = flags_in[4], flags[5] = flags_in[5];
  CHL_HIP(chl::LaunchGhostFill(U, nx, ny, nz, n_fields, n_ghost, flags, nullptr, mhd));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}
}  // namespace

extern "C" {

int chl_set_ghost_cells(double *c_head, int nx, int ny, int nz, int n_fields, int n_cells, int n_ghost,
                        const int flags[6], int isize, int jsize, int ksize, int imin, int jmin, int kmin, int dir)
{
  return SetGhostCellsImpl(c_head, nx, ny, nz, n_fields, n_cells, n_ghost, flags, isize, jsize, ksize, imin, jmin,
                           kmin, dir, false);
}

int chl_mhd_set_ghost_cells(double *c_head, int nx, int ny, int nz, int n_fields, int n_cells, int n_ghost,
                            const int flags[6], int isize, int jsize, int ksize, int imin, int jmin, int kmin, int dir)
{
  return SetGhostCellsImpl(c_head, nx, ny, nz, n_fields, n_cells, n_ghost, flags, isize, jsize, ksize, imin, jmin,
                           kmin, dir, true);
}

int chl_set_boundaries(double *U, int nx, int ny, int nz, int n_fields, int n_ghost, const int flags_in[6])
{
  return SetBoundariesImpl(U, nx, ny, nz, n_fields, n_ghost, flags_in, false);
}

int chl_mhd_set_boundaries(double *U, int nx, int ny, int nz, int n_ghost, const int flags_in[6])
{
  return SetBoundariesImpl(U, nx, ny, nz, 8, n_ghost, flags_in, true);
}

}  // extern "C"

namespace
{
// The dataset region of one field (see the snapshot entry points in cholla_amd.h).
int SnapshotRegion(int nx, int ny, int nz, int ng, int nxr, int nyr, int nzr, int mhd_direction, chl::SnapshotShape &s)
{
  const bool one_d = ny == 1 && nz == 1;
  if (!one_d && (ny == 1 || nz == 1)) return Fail(CHL_ERR_UNSUPPORTED, "2D grids are not built");
  if (mhd_direction < -1 || mhd_direction > 2 || (one_d && mhd_direction >= 0))
    return Fail(CHL_ERR_INVALID, "bad mhd_direction");
  if (one_d) nyr = nzr = 1;
  s = {nx,
       ny,
       nz,
       nxr,
       nyr,
       nzr,
       ng - int(mhd_direction == 0),
       one_d ? 0 : ng - int(mhd_direction == 1),
       one_d ? 0 : ng - int(mhd_direction == 2)};
  if (!s.Valid()) return Fail(CHL_ERR_INVALID, "dataset region outside the grid");
  return CHL_OK;
}
}  // namespace

extern "C" {

int chl_pack_real_cells(const double *d_field, double *d_buf, int nx, int ny, int nz, int n_ghost, int nx_real,
                        int ny_real, int nz_real, int mhd_direction)
{
  if (!d_field || !d_buf) return Fail(CHL_ERR_INVALID, "null argument");
  chl::SnapshotShape s;
  CHL_TRY(SnapshotRegion(nx, ny, nz, n_ghost, nx_real, ny_real, nz_real, mhd_direction, s));
  CHL_HIP(chl::LaunchSnapshotPack(d_field, d_buf, s, nullptr));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int chl_pack_real_cells_f32(const double *d_field, float *d_buf, int nx, int ny, int nz, int n_ghost, int nx_real,
                            int ny_real, int nz_real, int mhd_direction)
{
  if (!d_field || !d_buf) return Fail(CHL_ERR_INVALID, "null argument");
  chl::SnapshotShape s;
  CHL_TRY(SnapshotRegion(nx, ny, nz, n_ghost, nx_real, ny_real, nz_real, mhd_direction, s));
  CHL_HIP(chl::LaunchSnapshotPack32(d_field, d_buf, s, nullptr));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int chl_unpack_real_cells(const double *d_buf, double *d_field, int nx, int ny, int nz, int n_ghost, int nx_real,
                          int ny_real, int nz_real, int mhd_direction)
{
  if (!d_field || !d_buf) return Fail(CHL_ERR_INVALID, "null argument");
  chl::SnapshotShape s;
  CHL_TRY(SnapshotRegion(nx, ny, nz, n_ghost, nx_real, ny_real, nz_real, mhd_direction, s));
  CHL_HIP(chl::LaunchSnapshotUnpack(d_buf, d_field, s, nullptr));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int chl_reconstruct(int recon, const double *U, double *QL, double *QR, int nx, int ny, int nz, double gamma, int dir)
{
  if (!U || !QL || !QR || dir < 0 || dir > 2) return Fail(CHL_ERR_INVALID, "bad argument");
  if (recon != CHL_RECON_PLMC && recon != CHL_RECON_PPMC) return Fail(CHL_ERR_UNSUPPORTED, "PLMC or PPMC only");
  CHL_HIP(chl::LaunchReconGrid(recon, U, QL, QR, nx, ny, nz, gamma, dir, nullptr));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int chl_hllc_fluxes(const double *QL, const double *QR, double *F, int nx, int ny, int nz, double gamma, int dir)
{
  if (!QL || !QR || !F || dir < 0 || dir > 2) return Fail(CHL_ERR_INVALID, "bad argument");
  CHL_HIP(chl::LaunchRiemannGrid(chl::kRiemannHllc, QL, QR, F, nx, ny, nz, gamma, dir, nullptr));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int chl_hll_fluxes(const double *QL, const double *QR, double *F, int nx, int ny, int nz, double gamma, int dir)
{
  if (!QL || !QR || !F || dir < 0 || dir > 2) return Fail(CHL_ERR_INVALID, "bad argument");
  CHL_HIP(chl::LaunchRiemannGrid(chl::kRiemannHll, QL, QR, F, nx, ny, nz, gamma, dir, nullptr));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int chl_roe_fluxes(const double *QL, const double *QR, double *F, int nx, int ny, int nz, double gamma, int dir)
{
  if (!QL || !QR || !F || dir < 0 || dir > 2) return Fail(CHL_ERR_INVALID, "bad argument");
  CHL_HIP(chl::LaunchRiemannGrid(chl::kRiemannRoe, QL, QR, F, nx, ny, nz, gamma, dir, nullptr));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int chl_exact_fluxes(const double *QL, const double *QR, double *F, int nx, int ny, int nz, double gamma, int dir)
{
  if (!QL || !QR || !F || dir < 0 || dir > 2) return Fail(CHL_ERR_INVALID, "bad argument");
  CHL_HIP(chl::LaunchRiemannGrid(chl::kRiemannExact, QL, QR, F, nx, ny, nz, gamma, dir, nullptr));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int chl_ic_kelvin_helmholtz(double *U, int n, int ng, double gamma, uint64_t seed)
{
  if (!U || n < ng || ng < 1) return Fail(CHL_ERR_INVALID, "bad argument");
  const int N     = n + 2 * ng;
  const size_t nc = size_t(N) * N * N;
  const double dx = 1.0 / n;
  std::mt19937_64 prng(seed);
  std::uniform_real_distribution<double> noise(-1.0, 1.0);
  for (int k = ng; k < N - ng; k++)
    for (int j = ng; j < N - ng; j++)
      for (int i = ng; i < N - ng; i++) {
        // cell centres as Grid3D::Get_Position (src/grid/grid3D.cpp:84-86)
        double x = 0.0 + dx * (i - ng) + 0.5 * dx;
        double y = 0.0 + dx * (j - ng) + 0.5 * dx;
        double u = noise(prng);
        bool in  = std::fabs(y - 0.5) < 0.25;
        double d = in ? 2.0 : 1.0, vx = in ? 0.5 : -0.5;
        double vy = 0.01 * std::sin(4.0 * M_PI * x) + 1.0e-4 * u, vz = 0.0, p = 2.5;
        size_t id     = size_t(i) + size_t(j) * N + size_t(k) * N * N;
        U[id]         = d;
        U[nc + id]    = d * vx;
        U[2 * nc + id] = d * vy;
        U[3 * nc + id] = d * vz;
        // Calc_Energy_Primitive, src/utils/hydro_utilities.h:106-118
        U[4 * nc + id] = (std::fmax(p, 1.0e-20) / (gamma - 1.)) + 0.5 * d * (vx * vx + ((vy * vy) + (vz * vz)));
      }
  // periodic ghost cells (flag 1 on every face)
  auto wrap = [&](int a) { return a < ng ? a + n : (a >= N - ng ? a - n : a); };
  for (int k = 0; k < N; k++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        int ir = wrap(i), jr = wrap(j), kr = wrap(k);
        if (ir == i && jr == j && kr == k) continue;
        size_t g = size_t(i) + size_t(j) * N + size_t(k) * N * N;
        size_t s = size_t(ir) + size_t(jr) * N + size_t(kr) * N * N;
        for (int f = 0; f < 5; f++) U[f * nc + g] = U[f * nc + s];
      }
  return CHL_OK;
}

}  // extern "C"
