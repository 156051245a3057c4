// mhd.cpp -- C ABI of the MHD integrator (chl_mhd_*), the MHD stage entry points and the
// MHD problem setup (declared in include/cholla_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <vector>

#include "../../include/cholla_amd.h"
#include "capi_internal.hpp"
#include "mhd_kernels.hpp"
#include "support_kernels.hpp"
#include "vl3d_launch.hpp"

namespace
{
using chl_capi::Fail;
using chl_capi::kCfl;
}  // namespace

struct chl_mhd {
  chl_mhd_config cfg{};
  size_t n_cells     = 0;
  int num_cu         = 256;
  double *Uh         = nullptr;
  void *F            = nullptr;  // 3 x 7 fields of the flux element type
  size_t F_elem      = 8;
  unsigned char *scal = nullptr;  // [0] dti bits, [1] dt, [2] t, [3] n_step; [4] divergence bits
  void *rep_buf      = nullptr;  // the repair marks and counts (chl::RepairState)
  chl::RepairState rep{};
  hipStream_t stream = nullptr;
  hipGraphExec_t graph = nullptr;
  double *g_U        = nullptr;
  double g_key[7]{};

  unsigned long long *dti() { return reinterpret_cast<unsigned long long *>(scal); }
  double *dt() { return reinterpret_cast<double *>(scal + 8); }
  double *t() { return reinterpret_cast<double *>(scal + 16); }
  int *nstep() { return reinterpret_cast<int *>(scal + 24); }
  unsigned long long *divb() { return reinterpret_cast<unsigned long long *>(scal + 32); }
  void DropGraph()
  {
    if (graph) (void)hipGraphExecDestroy(graph);
    graph = nullptr;
  }
};

namespace
{

int CheckMhdConfig(const chl_mhd_config &c)
{
  if (c.reconstruction != CHL_RECON_PCM && c.reconstruction != CHL_RECON_PLMC && c.reconstruction != CHL_RECON_PPMC)
    return Fail(CHL_ERR_UNSUPPORTED, "MHD reconstruction must be PCM, PLMC or PPMC (src/utils/error_handling.cpp:98-101)");
  if (c.n_ghost < chl::mhd::MhdReconGhosts(c.reconstruction))
    return Fail(CHL_ERR_INVALID, "MHD needs n_ghost >= 3 (PCM), 4 (PLMC) or 5 (PPMC): the reconstruction's ghosts + 1 "
                                 "for the face fields (src/grid/grid3D.cpp:53-76)");
  if (c.precision < CHL_PREC_FP64_EXACT || c.precision > CHL_PREC_MIXED) return Fail(CHL_ERR_INVALID, "bad precision");
  const int ng = c.n_ghost;
  if (c.nx < 3 * ng || c.ny < 3 * ng || c.nz < 3 * ng)
    return Fail(CHL_ERR_INVALID, "MHD boxes need at least n_ghost real cells per axis (3D only)");
  if (double(c.nx) * c.ny * c.nz * 8 > 9.0e15) return Fail(CHL_ERR_INVALID, "box too large");
  for (int q = 0; q < 6; q++)
    if (c.bc_flags[q] < 1 || c.bc_flags[q] > 3)
      return Fail(CHL_ERR_UNSUPPORTED, "boundary flags must be 1 (periodic), 2 (reflective) or 3 (transmissive)");
  return CHL_OK;
}

chl::mhd::MhdArgs MakeArgs(chl_mhd *h, double *U, double dx, double dy, double dz, double gamma, double dt_host,
                           const double *dt_dev, double density_floor, double U_floor, bool update_grid)
{
  const chl_mhd_config &c = h->cfg;
  chl::mhd::MhdArgs a{};
  a.U = U, a.Uh = h->Uh;
  for (int d = 0; d < 3; d++) a.F[d] = static_cast<unsigned char *>(h->F) + size_t(d) * 7 * h->n_cells * h->F_elem;
  a.nx = c.nx, a.ny = c.ny, a.nz = c.nz, a.ng = c.n_ghost;
  a.dx = dx, a.dy = dy, a.dz = dz, a.gamma = gamma, a.dt = dt_host, a.dt_dev = dt_dev, a.dti = h->dti();
  a.density_floor = density_floor, a.U_floor = U_floor;
  // without both floors the crash repair reads the 3x3x3 block as the update left it, so the
  // temperature floor runs after the repairs (RepairStep) instead of in the full update
  const bool crash = !(c.density_floor && c.temperature_floor);
  int flags        = 0;
  if (c.density_floor) flags |= chl::kFlagDensityFloor;
  if (crash) flags |= chl::kFlagCrashRepair;
  if (update_grid && c.temperature_floor && !crash) flags |= chl::kFlagTemperatureFloor;
  if (update_grid) flags |= chl::kFlagReduceDt;
  a.flags = flags;
  a.recon = c.reconstruction;
  a.rep   = h->rep;
  return a;
}

// The repairs the full update marked (RepairKernel over the five hydro fields, as
// Average_Cell_Single_Field / Average_Cell_All_Fields act on them), then, in a build without
// both floors, the temperature floor.
int RepairStep(chl_mhd *h, const chl::mhd::MhdArgs &m, bool update_grid)
{
  const chl_mhd_config &c = h->cfg;
  chl::StepArgs a{};
  a.U = m.U, a.nx = m.nx, a.ny = m.ny, a.nz = m.nz, a.ng = m.ng;
  a.dx = m.dx, a.dy = m.dy, a.dz = m.dz, a.gamma = m.gamma, a.dt = m.dt, a.dt_dev = m.dt_dev, a.dti = m.dti;
  a.density_floor = m.density_floor, a.U_floor = m.U_floor;
  a.flags = m.flags & (chl::kFlagDensityFloor | chl::kFlagCrashRepair | chl::kFlagTemperatureFloor);
  a.rep   = m.rep;
  a.rep.side = nullptr;  // the MHD ghost fill follows the step: ghosts hold pre-step values
  CHL_HIP(c.precision == CHL_PREC_FP64_EXACT ? chl::LaunchRepairExact(a, h->stream)
                                             : chl::LaunchRepairFast(c.precision, a, h->stream));
  if (update_grid && c.temperature_floor && (m.flags & chl::kFlagCrashRepair))
    CHL_HIP(chl::LaunchTemperatureFloor(m.U, c.nx, c.ny, c.nz, c.n_ghost, m.U_floor, 8, h->stream));
  return CHL_OK;
}

// The step with its own dt: the stage-1 flux launch reduces dti from the state it reads
// (the state the reference's Calc_dt reads at the end of the previous step: the repairs and
// floors have run, and the ghost refresh changes no real cell), the device clock turns it into
// dt and clears it, then the rest of the step, the repairs and the ghost refresh.
int EnqueueDeviceStep(chl_mhd *h, double *U, double dx, double dy, double dz, double gamma, double density_floor,
                      double U_floor, double tout)
{
  const chl_mhd_config &c = h->cfg;
  chl::mhd::MhdArgs a = MakeArgs(h, U, dx, dy, dz, gamma, 0.0, h->dt(), density_floor, U_floor, true);
  a.flags |= chl::kFlagFoldDt;
  a.clk_dt = h->dt(), a.clk_t = h->t(), a.clk_nstep = h->nstep(), a.clk_tout = tout, a.clk_cfl = kCfl;
  CHL_HIP(chl::LaunchMhdStep(c.precision, a, h->stream));
  CHL_TRY(RepairStep(h, a, true));
  CHL_HIP(chl::LaunchGhostFill(U, c.nx, c.ny, c.nz, 8, c.n_ghost, c.bc_flags, h->stream, true));
  return CHL_OK;
}

// Device clock at t0 with dti = 0, which the first step's flux launch fills.
int InitialDti(chl_mhd *h, double t0)
{
  unsigned char buf[32] = {};
  std::memcpy(buf + 16, &t0, 8);
  CHL_HIP(hipMemcpyAsync(h->scal, buf, 32, hipMemcpyHostToDevice, h->stream));
  return CHL_OK;
}

int EnsureGraph(chl_mhd *h, double *U, double dx, double dy, double dz, double gamma, double density_floor,
                double U_floor, double tout)
{
  const double key[7] = {dx, dy, dz, gamma, density_floor, U_floor, tout};
  if (h->graph && h->g_U == U && std::memcmp(key, h->g_key, sizeof(key)) == 0) return CHL_OK;
  h->DropGraph();
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

int NumCu()
{
  int dev, cu = 256;
  hipDeviceProp_t p;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) cu = p.multiProcessorCount;
  return cu;
}

}  // namespace

extern "C" {

int chl_mhd_create(const chl_mhd_config *cfg, chl_mhd **out)
{
  if (!cfg || !out) return Fail(CHL_ERR_INVALID, "null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return Fail(CHL_ERR_NO_DEVICE, "no HIP device");
  CHL_TRY(CheckMhdConfig(*cfg));
  auto *h    = new chl_mhd();
  h->cfg     = *cfg;
  h->n_cells = size_t(cfg->nx) * cfg->ny * cfg->nz;
  h->num_cu  = NumCu();
  h->F_elem  = cfg->precision == CHL_PREC_MIXED ? sizeof(float) : sizeof(double);
  const size_t n = h->n_cells;
  hipError_t e   = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&h->Uh, 8 * n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&h->F, 21 * n * h->F_elem);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&h->scal), 64);
  const size_t rep_bytes = chl_capi::RepairBytesFor(cfg->nx, cfg->ny, cfg->nz, cfg->n_ghost, 0);
  if (e == hipSuccess) e = hipMalloc(&h->rep_buf, rep_bytes);
  if (e == hipSuccess) e = hipMemsetAsync(h->rep_buf, 0, rep_bytes, h->stream);
  if (e == hipSuccess) h->rep = chl_capi::RepairAt(h->rep_buf, cfg->nx, cfg->ny, cfg->nz, cfg->n_ghost, 0);
  // cleared once: cells and faces outside the computed regions keep these values
  if (e == hipSuccess) e = hipMemsetAsync(h->Uh, 0, 8 * n * sizeof(double), h->stream);
  if (e == hipSuccess) e = hipMemsetAsync(h->F, 0, 21 * n * h->F_elem, h->stream);
  if (e == hipSuccess) e = hipMemsetAsync(h->scal, 0, 64, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) {
    chl_mhd_destroy(h);
    return Fail(CHL_ERR_HIP, std::string("allocating the MHD integrator scratch: ") + hipGetErrorString(e));
  }
  *out = h;
  return CHL_OK;
}

int chl_mhd_destroy(chl_mhd *h)
{
  if (!h) return CHL_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  h->DropGraph();
  if (h->Uh) (void)hipFree(h->Uh);
  if (h->F) (void)hipFree(h->F);
  if (h->scal) (void)hipFree(h->scal);
  if (h->rep_buf) (void)hipFree(h->rep_buf);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return CHL_OK;
}

int chl_mhd_vl_step(chl_mhd *h, double *U, double dx, double dy, double dz, double dt, double gamma,
                    double density_floor)
{
  if (!h || !U) return Fail(CHL_ERR_INVALID, "null argument");
  chl::mhd::MhdArgs a = MakeArgs(h, U, dx, dy, dz, gamma, dt, nullptr, density_floor, 0.0, false);
  CHL_HIP(chl::LaunchMhdStep(h->cfg.precision, a, h->stream));
  CHL_TRY(RepairStep(h, a, false));
  CHL_HIP(hipStreamSynchronize(h->stream));
  return CHL_OK;
}

int chl_mhd_run(chl_mhd *h, double *U, double dx, double dy, double dz, double gamma, double density_floor,
                double U_floor, double t0, double tout, int max_steps, int use_graph, double *t_out, int *steps_out)
{
  if (!h || !U) return Fail(CHL_ERR_INVALID, "null argument");
  CHL_TRY(InitialDti(h, t0));
  if (use_graph) CHL_TRY(EnsureGraph(h, U, dx, dy, dz, gamma, density_floor, U_floor, tout));
  double t = t0;
  int done = 0;
  while (done < max_steps && t < tout) {
    const int batch = std::min(32, max_steps - done);
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

int chl_mhd_time_steps(chl_mhd *h, double *U, double dx, double dy, double dz, double gamma, int n_steps,
                       int use_graph, float *ms_out, double *dt_last_out)
{
  if (!h || !U || n_steps < 0) return Fail(CHL_ERR_INVALID, "bad argument");
  const double tout = std::numeric_limits<double>::infinity();
  CHL_TRY(InitialDti(h, 0.0));
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

int chl_mhd_calc_dt(const double *U, int nx, int ny, int nz, int n_ghost, double dx, double dy, double dz,
                    double gamma, double *dti_out)
{
  if (!U || !dti_out) return Fail(CHL_ERR_INVALID, "null argument");
  unsigned long long *d = nullptr;
  CHL_HIP(hipMalloc(&d, 8));
  hipError_t e = hipMemset(d, 0, 8);
  if (e == hipSuccess) e = chl::LaunchMhdCalcDt(U, nx, ny, nz, n_ghost, dx, dy, dz, gamma, d, NumCu(), nullptr);
  unsigned long long bits = 0;
  if (e == hipSuccess) e = hipMemcpy(&bits, d, 8, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  CHL_HIP(e);
  std::memcpy(dti_out, &bits, 8);
  return CHL_OK;
}

int chl_mhd_divergence(const double *U, int nx, int ny, int nz, double dx, double dy, double dz, double *max_div)
{
  if (!U || !max_div) return Fail(CHL_ERR_INVALID, "null argument");
  unsigned long long *d = nullptr;
  CHL_HIP(hipMalloc(&d, 8));
  hipError_t e = hipMemset(d, 0, 8);
  if (e == hipSuccess) e = chl::LaunchMhdDivergence(U, nx, ny, nz, dx, dy, dz, d, NumCu(), nullptr);
  unsigned long long bits = 0;
  if (e == hipSuccess) e = hipMemcpy(&bits, d, 8, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  CHL_HIP(e);
  std::memcpy(max_div, &bits, 8);
  return CHL_OK;
}

int chl_hlld_fluxes(const double *QL, const double *QR, const double *magX, double *F, long n_cells, double gamma,
                    int dir)
{
  if (!QL || !QR || !magX || !F || n_cells < 0 || dir < 0 || dir > 2) return Fail(CHL_ERR_INVALID, "bad argument");
  CHL_HIP(chl::LaunchHlldGrid(QL, QR, magX, F, n_cells, gamma, dir, nullptr));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int chl_mhd_reconstruct_plmc(const double *U, double *QL, double *QR, int nx, int ny, int nz, double gamma, int dir)
{
  if (!U || !QL || !QR || dir < 0 || dir > 2) return Fail(CHL_ERR_INVALID, "bad argument");
  CHL_HIP(chl::LaunchMhdPlmcGrid(U, QL, QR, nx, ny, nz, gamma, dir, nullptr));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int chl_mhd_reconstruct_ppmc(const double *U, double *QL, double *QR, int nx, int ny, int nz, double gamma, int dir)
{
  if (!U || !QL || !QR || dir < 0 || dir > 2) return Fail(CHL_ERR_INVALID, "bad argument");
  CHL_HIP(chl::LaunchMhdPpmcGrid(U, QL, QR, nx, ny, nz, gamma, dir, nullptr));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int chl_ct_electric_fields(const double *Fx, const double *Fy, const double *Fz, const double *U, double *E, int nx,
                           int ny, int nz)
{
  if (!Fx || !Fy || !Fz || !U || !E) return Fail(CHL_ERR_INVALID, "null argument");
  CHL_HIP(chl::LaunchCtGrid(Fx, Fy, Fz, U, E, nx, ny, nz, nullptr));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int chl_update_magnetic_field(const double *src, double *dst, const double *E, int nx, int ny, int nz, double dt,
                              double dx, double dy, double dz)
{
  if (!src || !dst || !E) return Fail(CHL_ERR_INVALID, "null argument");
  CHL_HIP(chl::LaunchMagneticUpdateGrid(src, dst, E, nx, ny, nz, dt, dx, dy, dz, nullptr));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int chl_magnetic_field_from_vector_potential(double *U, int nx, int ny, int nz, double dx, double dy, double dz,
                                             const double *A)
{
  if (!U || !A || nx < 1 || ny < 1 || nz < 1) return Fail(CHL_ERR_INVALID, "bad argument");
  const size_t n = size_t(nx) * ny * nz;
  auto at        = [&](int i, int j, int k) { return size_t(i) + size_t(j) * nx + size_t(k) * size_t(nx) * ny; };
  for (int k = 1; k < nz; k++)
    for (int j = 1; j < ny; j++)
      for (int i = 1; i < nx; i++) {
        const size_t id = at(i, j, k), xm = at(i - 1, j, k), ym = at(i, j - 1, k), zm = at(i, j, k - 1);
        U[5 * n + id] = (A[id + 2 * n] - A[ym + 2 * n]) / dy - (A[id + 1 * n] - A[zm + 1 * n]) / dz;
        U[6 * n + id] = (A[id + 0 * n] - A[zm + 0 * n]) / dz - (A[id + 2 * n] - A[xm + 2 * n]) / dx;
        U[7 * n + id] = (A[id + 1 * n] - A[xm + 1 * n]) / dx - (A[id + 0 * n] - A[ym + 0 * n]) / dy;
      }
  return CHL_OK;
}

int chl_ic_orszag_tang(double *U, int n, int ng, double gamma, uint64_t seed, double amplitude)
{
  if (!U || n < ng || ng < 1) return Fail(CHL_ERR_INVALID, "bad argument");
  const int N      = n + 2 * ng;
  const size_t nc  = size_t(N) * N * N;
  const double dx  = 1.0 / n;
  const double B0  = 1.0 / std::sqrt(4.0 * M_PI);
  const double rho = 25.0 / (36.0 * M_PI), v0 = 1.0, p0 = 5.0 / (12.0 * M_PI);
  std::fill(U, U + 8 * nc, 0.0);
  // z-dependent perturbation modes of the potential (A_x, A_y)(z): amplitude/(2 pi m)
  // a_m cos(2 pi m z + phi_m), m = 1..4, drawn a_x, phi_x, a_y, phi_y per m
  std::mt19937_64 prng(seed);
  std::uniform_real_distribution<double> amp(-1.0, 1.0), phase(0.0, 2.0 * M_PI);
  double ax[4], px[4], ay[4], py[4];
  for (int m = 0; m < 4; m++) ax[m] = amp(prng), px[m] = phase(prng), ay[m] = amp(prng), py[m] = phase(prng);
  auto pos = [&](int i) { return 0.0 + dx * (i - ng) + 0.5 * dx; };  // Grid3D::Get_Position, grid3D.cpp:84-86
  std::vector<double> A(3 * nc, 0.0);
  for (int k = 0; k < N; k++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        const size_t id = size_t(i) + size_t(j) * N + size_t(k) * N * N;
        const double x = pos(i), y = pos(j), z = pos(k);
        // Orszag_Tang_Vortex (src/grid/initial_conditions.cpp:1926-1952)
        A[id + 2 * nc] = B0 / (4.0 * M_PI) * (std::cos(4.0 * M_PI * x) + 2.0 * std::cos(2.0 * M_PI * y));
        double sx = 0, sy = 0;
        for (int m = 0; m < 4; m++) {
          sx += ax[m] / (m + 1) * std::cos(2.0 * M_PI * (m + 1) * z + px[m]);
          sy += ay[m] / (m + 1) * std::cos(2.0 * M_PI * (m + 1) * z + py[m]);
        }
        A[id + 0 * nc] = amplitude / (2.0 * M_PI) * sx;
        A[id + 1 * nc] = amplitude / (2.0 * M_PI) * sy;
      }
  CHL_TRY(chl_magnetic_field_from_vector_potential(U, N, N, N, dx, dx, dx, A.data()));
  for (int k = ng - 1; k < N - ng; k++)
    for (int j = ng - 1; j < N - ng; j++)
      for (int i = ng - 1; i < N - ng; i++) {
        const size_t id = size_t(i) + size_t(j) * N + size_t(k) * N * N;
        const double x = pos(i), y = pos(j);
        // cellCenteredMagneticFields (src/utils/mhd_utilities.h:177-207)
        const double bx = 0.5 * (U[5 * nc + id] + U[5 * nc + id - 1]);
        const double by = 0.5 * (U[6 * nc + id] + U[6 * nc + id - N]);
        const double bz = 0.5 * (U[7 * nc + id] + U[7 * nc + id - size_t(N) * N]);
        const double mx = rho * v0 * std::sin(2.0 * M_PI * y), my = -rho * v0 * std::sin(2.0 * M_PI * x), mz = 0.0;
        U[id]          = rho;
        U[nc + id]     = mx;
        U[2 * nc + id] = my;
        U[3 * nc + id] = mz;
        // Calc_Energy_Conserved with MHD (src/utils/hydro_utilities.h:120-133)
        U[4 * nc + id] = (std::fmax(p0, 1.0e-20) / (gamma - 1.)) + (0.5 / rho) * (mx * mx + ((my * my) + (mz * mz))) +
                         0.5 * (bx * bx + ((by * by) + (bz * bz)));
      }
  // periodic ghost cells, all eight fields (SetGhostCells with flag 1 on every face)
  auto wrap = [&](int a) { return a < ng ? a + n : (a >= N - ng ? a - n : a); };
  for (int k = 0; k < N; k++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        const int ir = wrap(i), jr = wrap(j), kr = wrap(k);
        if (ir == i && jr == j && kr == k) continue;
        const size_t g = size_t(i) + size_t(j) * N + size_t(k) * N * N;
        const size_t s = size_t(ir) + size_t(jr) * N + size_t(kr) * N * N;
        for (int f = 0; f < 8; f++) U[f * nc + g] = U[f * nc + s];
      }
  return CHL_OK;
}

}  // extern "C"
