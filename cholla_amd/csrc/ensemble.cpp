// ensemble.cpp -- C ABI of the ensemble integrator (chl_ensemble_*) and of the
// turbulence initial conditions of the ensemble benchmark.
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
#include "ensemble_kernels.hpp"
#include "support_kernels.hpp"
#include "vl3d_launch.hpp"

namespace
{
using chl_capi::Fail;
using chl_capi::kCfl;
using chl_capi::ReconRadius;

constexpr int kDefaultTile1   = 2;  // 64x4
constexpr int kDefaultTile2   = 1;  // 16x16
constexpr int kDefaultDeTile  = 1;  // 16x16, both stages, with dual energy
constexpr int kDefaultKchunk1 = 16;
constexpr int kDefaultKchunk2 = 32;
// Sleeps of a waiting stage-2 tile before it gives up (s_sleep 2 = 128 clocks each:
// about half a second, three orders of magnitude above a stage-1 tile's run time).
constexpr long long kSpinLimit = 1ll << 23;

// Persistent pairs compiled with dual energy: both tiles among the dual-energy shapes.
bool DePairIdx(int pair) { return chl::kPairs[pair][0] < chl::kNumDeTiles && chl::kPairs[pair][1] < chl::kNumDeTiles; }

int PairIndex(int tile1, int tile2)
{
  for (int p = 0; p < chl::kNumPairs; p++)
    if (chl::kPairs[p][0] == tile1 && chl::kPairs[p][1] == tile2) return p;
  return -1;
}

}  // namespace

struct chl_ensemble {
  chl_ensemble_config cfg{};
  int n = 0;
  std::vector<int> dims;         // 3 per box, totals incl. ghosts
  std::vector<double *> U;       // bound states
  std::vector<double> dxyz;      // 3 per box
  std::vector<double> tout;      // per box
  std::vector<int> order;        // boxes by decreasing cell count
  std::vector<chl::BoxDesc> hbox;
  chl::BoxDesc *box   = nullptr;
  chl::BoxClock *clk  = nullptr;
  chl::WorkItem *items = nullptr;  // [stage-1 list | stage-2 list | persistent queue]
  size_t items_cap    = 0;
  int n1 = 0, n2 = 0, nq = 0;
  unsigned char *W    = nullptr;
  size_t W_bytes      = 0;
  unsigned int *head  = nullptr;  // persistent queue head; status word right after it
  hipStream_t stream  = nullptr;
  hipGraphExec_t graph = nullptr;
  int num_cu          = 256;
  int per_cu          = 1;
  int qoff[chl::kQueueXcds + 1] = {};  // persistent queue of XCD x: [qoff[x], qoff[x + 1])
  bool bound = false, started = false;
  double gamma = 5.0 / 3.0, density_floor = 0.0, U_floor = 0.0;
  long max_face = 0;  // the largest side buffer of a box (cells), for RepairSide launches

  // head buffer: the persistent queues' heads (XCD x at x * kHeadStride), then the status word
  static constexpr size_t kHeadBytes = 4096;
  int *status() { return reinterpret_cast<int *>(head + chl::kQueueXcds * chl::kHeadStride); }
  size_t Cells(int b) const { return size_t(dims[3 * b]) * dims[3 * b + 1] * dims[3 * b + 2]; }
  void DropGraph()
  {
    if (graph) (void)hipGraphExecDestroy(graph);
    graph = nullptr;
  }
};

namespace
{

int CheckEnsembleConfig(const chl_ensemble_config &c)
{
  if (c.dual_energy != 0 && c.dual_energy != 1) return Fail(CHL_ERR_INVALID, "dual_energy must be 0 or 1");
  if (c.n_fields != 5 + c.dual_energy)
    return Fail(CHL_ERR_UNSUPPORTED, "n_fields must be 5, or 6 with dual_energy (no passive scalars)");
  if (c.riemann < CHL_RIEMANN_HLLC || c.riemann > CHL_RIEMANN_EXACT)
    return Fail(CHL_ERR_UNSUPPORTED, "riemann must be HLLC, HLL, ROE or EXACT");
  if (c.reconstruction < CHL_RECON_PCM || c.reconstruction > CHL_RECON_PPMP)
    return Fail(CHL_ERR_UNSUPPORTED, "reconstruction must be PCM, PLMC, PPMC, PLMP or PPMP");
  // the HLLC + PLMC/PPMC kernels are compiled for every tile shape and both launches; the
  // other solvers and reconstructions for the split launch on tile shapes 0 and 1, dual energy
  // with the characteristic reconstructions only (as for single boxes)
  const bool simple_recon = c.reconstruction == CHL_RECON_PCM || c.reconstruction == CHL_RECON_PLMP ||
                            c.reconstruction == CHL_RECON_PPMP;
  if (c.riemann != CHL_RIEMANN_HLLC || simple_recon) {
    if (c.launch != CHL_LAUNCH_SPLIT)
      return Fail(CHL_ERR_UNSUPPORTED, "the HLL, Roe and exact solvers and PCM, PLMP and PPMP run the split launch");
    if (c.tile1 >= chl::kNumDeTiles || c.tile2 >= chl::kNumDeTiles)
      return Fail(CHL_ERR_UNSUPPORTED,
                  "the HLL, Roe and exact solvers and PCM, PLMP and PPMP are built for tile shapes 0 and 1 only");
  }
  if (c.precision < CHL_PREC_FP64_EXACT || c.precision > CHL_PREC_MIXED) return Fail(CHL_ERR_INVALID, "bad precision");
  if (c.positivity_fallback != 0 && c.positivity_fallback != 1)
    return Fail(CHL_ERR_INVALID, "positivity_fallback must be 0 or 1");
  if (c.positivity_fallback && c.reconstruction != CHL_RECON_PPMC)
    return Fail(CHL_ERR_UNSUPPORTED, "the positivity fallback is built for PPMC only");
  if (c.n_ghost < ReconRadius(c.reconstruction) + 2)
    return Fail(CHL_ERR_INVALID, "n_ghost too small for the reconstruction (PCM needs 2, PLMC and PLMP 3, PPMC and PPMP 4)");
  if (c.launch != CHL_LAUNCH_SPLIT && c.launch != CHL_LAUNCH_PERSISTENT) return Fail(CHL_ERR_INVALID, "bad launch");
  if (c.tile1 < 0 || c.tile1 >= chl::kNumTiles || c.tile2 < 0 || c.tile2 >= chl::kNumTiles)
    return Fail(CHL_ERR_INVALID, "tile index out of range");
  if (c.launch == CHL_LAUNCH_PERSISTENT && PairIndex(c.tile1, c.tile2) < 0)
    return Fail(CHL_ERR_UNSUPPORTED, "this (tile1, tile2) pair is not compiled for the persistent launch");
  if (c.dual_energy && (c.tile1 >= chl::kNumDeTiles || c.tile2 >= chl::kNumDeTiles))
    return Fail(CHL_ERR_UNSUPPORTED, "dual energy is built for tile shapes 0 and 1 only");
  if (c.kchunk1 < 1 || c.kchunk2 < 1) return Fail(CHL_ERR_INVALID, "kchunk must be positive");
  if (!c.density_floor || !c.temperature_floor)
    return Fail(CHL_ERR_UNSUPPORTED, "ensembles run the default hydro build (density and temperature floors): a build "
                                     "without both needs the crash repair's unfused floor, dt and fill");
  return CHL_OK;
}

int TileCount(int extent, int tile) { return (extent + tile - 1) / tile; }

// Tile counts, work lists, scratch layout and box table for the current config.
// U_override (may be null) replaces the bound states in the uploaded box table.
int BuildPlan(chl_ensemble *e, const std::vector<double *> *U_override = nullptr)
{
  const chl_ensemble_config &c = e->cfg;
  const int ng = c.n_ghost, R1 = ReconRadius(c.reconstruction) + 1;
  const chl::TileShape s1 = chl::kTiles[c.tile1], s2 = chl::kTiles[c.tile2];
  const size_t elem       = c.precision == CHL_PREC_MIXED ? sizeof(float) : sizeof(double);
  const size_t w_fields   = 5 + size_t(c.precision == CHL_PREC_MIXED) + size_t(c.dual_energy);
  e->hbox.assign(e->n, chl::BoxDesc{});
  std::vector<size_t> w_off(e->n);
  size_t w_total = 0;
  for (int b = 0; b < e->n; b++) {
    chl::BoxDesc &d = e->hbox[b];
    d.nx = e->dims[3 * b], d.ny = e->dims[3 * b + 1], d.nz = e->dims[3 * b + 2];
    const int o1z = d.nz - 2 * ng + 2 * R1, o2z = d.nz - 2 * ng;
    d.kchunk1 = std::min(c.kchunk1, o1z);
    d.kchunk2 = std::min(c.kchunk2, o2z);
    d.t1[0]   = TileCount(d.nx - 2 * ng + 2 * R1, s1.tx);
    d.t1[1]   = TileCount(d.ny - 2 * ng + 2 * R1, s1.ty);
    d.t1[2]   = TileCount(o1z, d.kchunk1);
    d.t2[0]   = TileCount(d.nx - 2 * ng, s2.tx);
    d.t2[1]   = TileCount(d.ny - 2 * ng, s2.ty);
    d.t2[2]   = TileCount(o2z, d.kchunk2);
    d.items1  = d.t1[0] * d.t1[1] * d.t1[2];
    d.items2  = d.t2[0] * d.t2[1] * d.t2[2];
    d.U       = U_override ? (*U_override)[b] : (e->U.empty() ? nullptr : e->U[b]);
    d.dx = e->dxyz[3 * b], d.dy = e->dxyz[3 * b + 1], d.dz = e->dxyz[3 * b + 2];
    d.tout   = e->tout[b];
    w_off[b] = w_total;
    // W: 5 half-step fields (one more, the pressure's low part, in the mixed form; one
    // more, GasEnergy, with dual energy), then dual energy's pdV sources in FP64
    w_total += (w_fields * e->Cells(b) * elem + 255) / 256 * 256;
    if (c.dual_energy) w_total += (e->Cells(b) * sizeof(double) + 255) / 256 * 256;
  }
  // then each box's repair state (marks, counts, side buffer; no gravity records)
  std::vector<size_t> r_off(e->n);
  e->max_face = 0;
  for (int b = 0; b < e->n; b++) {
    const chl::BoxDesc &d = e->hbox[b];
    r_off[b]              = w_total;
    w_total += chl_capi::RepairBytesFor(d.nx, d.ny, d.nz, ng, 0);
    const long Lx = d.nx - 2 * ng, Ly = d.ny - 2 * ng, Lz = d.nz - 2 * ng;
    e->max_face   = std::max(e->max_face, 2 * (Ly * Lz + Lx * Lz + Lx * Ly));
  }
  if (w_total > e->W_bytes) {
    if (e->W) (void)hipFree(e->W);
    e->W       = nullptr;
    e->W_bytes = 0;
    CHL_HIP(hipMalloc(reinterpret_cast<void **>(&e->W), w_total));
    e->W_bytes = w_total;
  }
  // the repair state starts at zero (the region may have held another layout's data)
  CHL_HIP(hipMemsetAsync(e->W + r_off[0], 0, w_total - r_off[0], e->stream));
  for (int b = 0; b < e->n; b++) {
    const chl::BoxDesc &d = e->hbox[b];
    e->hbox[b].rep = chl_capi::RepairAt(e->W + r_off[b], d.nx, d.ny, d.nz, ng, 0);
    if (!c.density_floor) e->hbox[b].rep.side = nullptr;
    e->hbox[b].W   = e->W + w_off[b];
    e->hbox[b].pdv = c.dual_energy ? reinterpret_cast<double *>(
                                         e->W + w_off[b] + ((w_fields * e->Cells(b) * elem + 255) / 256 * 256))
                                   : nullptr;
  }

  std::vector<chl::WorkItem> l1, l2, q;
  auto push = [](std::vector<chl::WorkItem> &v, int b, int stage, int count) {
    for (int l = 0; l < count; l++) v.push_back(chl::WorkItem{b, stage, l, 0});
  };
  for (int b : e->order) {
    const chl::BoxDesc &d = e->hbox[b];
    push(l1, b, 0, d.items1);
    push(l2, b, 1, d.t2[0] * d.t2[1] * d.t2[2]);
  }
  // Persistent queues, one per XCD.  The boxes are dealt to the XCDs largest first, each
  // to the XCD with the least work so far (cells), so every phase of a box runs on one XCD.
  // Within a queue: the stage-1 items of its boxes in order; each later phase of a box
  // (stage 2, and with dual energy the selection and the sync) goes in once at least
  // `gap` further items have been queued after the box's previous phase, gap being the
  // XCD's resident workgroups.  A phase item is then taken only after a whole generation
  // of workgroups has started since the last item it depends on, so it rarely waits, also
  // for small boxes whose phases hold a few items each.  Every item of a box's phase sits
  // after all items of its previous phase in the same queue, and stage-1 items never
  // wait: each queue drains without deadlock.
  if (c.launch == CHL_LAUNCH_PERSISTENT) {
    int pc = 0;
    CHL_HIP(chl::EnsemblePersistentSlots(c.precision, c.reconstruction, PairIndex(c.tile1, c.tile2), c.dual_energy,
                                         &pc));
    e->per_cu = std::max(pc, 1);
  }
  {
    std::vector<std::vector<int>> owned(chl::kQueueXcds);
    std::vector<double> load(chl::kQueueXcds, 0.0);
    for (int b : e->order) {  // order: largest box first
      const int x = int(std::min_element(load.begin(), load.end()) - load.begin());
      owned[x].push_back(b);
      load[x] += double(e->Cells(b));
    }
    const int last_phase = c.dual_energy ? 3 : 1;
    const size_t gap     = size_t(e->per_cu) * size_t(std::max(e->num_cu / chl::kQueueXcds, 1));
    struct Pending {
      size_t ready_at;
      int b, phase;
    };
    for (int x = 0; x < chl::kQueueXcds; x++) {
      e->qoff[x] = int(q.size());
      const size_t base = q.size();
      std::vector<Pending> pend;  // FIFO; ready_at grows along it
      size_t head = 0;
      auto queue_phase = [&](Pending p) {
        push(q, p.b, p.phase, e->hbox[p.b].items2);
        if (p.phase < last_phase) pend.push_back({q.size() - base + gap, p.b, p.phase + 1});
      };
      for (int b : owned[x]) {
        push(q, b, 0, e->hbox[b].items1);
        pend.push_back({q.size() - base + gap, b, 1});
        while (head < pend.size() && pend[head].ready_at <= q.size() - base) queue_phase(pend[head++]);
      }
      while (head < pend.size()) queue_phase(pend[head++]);
    }
    e->qoff[chl::kQueueXcds] = int(q.size());
  }
  e->n1 = int(l1.size()), e->n2 = int(l2.size()), e->nq = int(q.size());
  const size_t need = l1.size() + l2.size() + q.size();
  if (need > e->items_cap) {
    if (e->items) (void)hipFree(e->items);
    e->items     = nullptr;
    e->items_cap = 0;
    CHL_HIP(hipMalloc(reinterpret_cast<void **>(&e->items), need * sizeof(chl::WorkItem)));
    e->items_cap = need;
  }
  std::vector<chl::WorkItem> all;
  all.reserve(need);
  all.insert(all.end(), l1.begin(), l1.end());
  all.insert(all.end(), l2.begin(), l2.end());
  all.insert(all.end(), q.begin(), q.end());
  CHL_HIP(hipMemcpyAsync(e->items, all.data(), need * sizeof(chl::WorkItem), hipMemcpyHostToDevice, e->stream));
  CHL_HIP(hipMemcpyAsync(e->box, e->hbox.data(), e->n * sizeof(chl::BoxDesc), hipMemcpyHostToDevice, e->stream));
  CHL_HIP(hipStreamSynchronize(e->stream));
  e->DropGraph();
  return CHL_OK;
}

chl::EnsembleArgs Args(chl_ensemble *e, int list)
{
  const chl_ensemble_config &c = e->cfg;
  chl::EnsembleArgs a{};
  a.box = e->box, a.clk = e->clk, a.n_boxes = e->n, a.ng = c.n_ghost, a.R1 = ReconRadius(c.reconstruction) + 1;
  a.gamma = e->gamma, a.density_floor = e->density_floor, a.U_floor = e->U_floor;
  int flags = chl::kFlagReduceDt | chl::kFlagPeriodicGhosts;
  if (c.density_floor) flags |= chl::kFlagDensityFloor;
  if (c.temperature_floor) flags |= chl::kFlagTemperatureFloor;
  if (c.positivity_fallback) flags |= chl::kFlagPositivityFallback;
  if (c.reconstruction == CHL_RECON_PCM) flags |= chl::kFlagConservedW;  // stage 1 stores momenta for PCM
  a.flags = flags, a.head = e->head, a.status = e->status(), a.spin_limit = kSpinLimit;
  for (int x = 0; x <= chl::kQueueXcds; x++) a.qoff[x] = e->qoff[x];
  a.de      = c.dual_energy;
  a.riemann = c.riemann;  // CHL_RIEMANN_* and kRiemann* share their values
  if (list == 0) a.items = e->items, a.n_items = e->n1;
  if (list == 1) a.items = e->items + e->n1, a.n_items = e->n2;
  if (list == 2) a.items = e->items + e->n1 + e->n2, a.n_items = e->nq;
  return a;
}

int EnqueueStage(chl_ensemble *e, int stage)
{
  const chl_ensemble_config &c = e->cfg;
  if (stage == 1) CHL_HIP(chl::LaunchEnsembleStage1(c.precision, Args(e, 0), c.tile1, e->stream));
  if (stage == 2) CHL_HIP(chl::LaunchEnsembleStage2(c.precision, Args(e, 1), c.reconstruction, c.tile2, e->stream));
  return CHL_OK;
}

int EnqueueStep(chl_ensemble *e)
{
  const chl_ensemble_config &c = e->cfg;
  CHL_HIP(chl::LaunchEnsembleClock(e->clk, e->box, e->n, kCfl, e->head, e->stream));
  if (c.density_floor) CHL_HIP(chl::LaunchEnsembleRepairSide(Args(e, 1), e->max_face, e->stream));
  if (c.launch == CHL_LAUNCH_SPLIT) {
    CHL_TRY(EnqueueStage(e, 1));
    CHL_TRY(EnqueueStage(e, 2));
    CHL_HIP(chl::LaunchEnsembleRepair(c.precision, Args(e, 1), e->stream));
    if (c.dual_energy) {
      CHL_HIP(chl::LaunchEnsembleDe(0, Args(e, 1), c.tile2, e->stream));
      CHL_HIP(chl::LaunchEnsembleDe(1, Args(e, 1), c.tile2, e->stream));
    }
  } else {
    // every XCD gets its resident workgroups (dispatch deals workgroups round-robin), so
    // every queue is served
    const int grid = e->per_cu * e->num_cu;
    CHL_HIP(chl::LaunchEnsemblePersistent(c.precision, Args(e, 2), c.reconstruction, PairIndex(c.tile1, c.tile2),
                                          grid, e->stream));
    // the repairs follow the whole launch (with dual energy, after the selection and sync
    // the launch carries: a repaired cell's selection saw its unrepaired density)
    CHL_HIP(chl::LaunchEnsembleRepair(c.precision, Args(e, 1), e->stream));
  }
  return CHL_OK;
}

int EnsureGraph(chl_ensemble *e)
{
  if (e->graph) return CHL_OK;
  hipGraph_t g;
  CHL_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  int rc        = EnqueueStep(e);
  hipError_t e2 = hipStreamEndCapture(e->stream, &g);
  if (rc != CHL_OK) return rc;
  CHL_HIP(e2);
  hipError_t e3 = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  CHL_HIP(e3);
  return CHL_OK;
}

// Reports (and clears) a dependency wait of the persistent launch that gave up, and a
// persistent queue that the last step left undrained.
int CheckStatus(chl_ensemble *e, bool stepped)
{
  int st = 0;
  CHL_HIP(hipMemcpyAsync(&st, e->status(), sizeof(int), hipMemcpyDeviceToHost, e->stream));
  CHL_HIP(hipStreamSynchronize(e->stream));
  if (e->cfg.launch == CHL_LAUNCH_PERSISTENT && stepped) {
    std::vector<unsigned int> heads(chl::kQueueXcds * chl::kHeadStride);
    CHL_HIP(hipMemcpyAsync(heads.data(), e->head, heads.size() * sizeof(unsigned int), hipMemcpyDeviceToHost,
                           e->stream));
    CHL_HIP(hipStreamSynchronize(e->stream));
    for (int x = 0; x < chl::kQueueXcds; x++)
      if (heads[x * chl::kHeadStride] < unsigned(e->qoff[x + 1] - e->qoff[x]))
        return Fail(CHL_ERR_HIP, "persistent ensemble launch: the queue of XCD " + std::to_string(x) +
                                     " was not drained; the step's result is not valid");
  }
  if (st != 0) {
    CHL_HIP(hipMemsetAsync(e->status(), 0, sizeof(int), e->stream));
    CHL_HIP(hipStreamSynchronize(e->stream));
    return Fail(CHL_ERR_HIP, "persistent ensemble launch: a stage-2 tile stopped waiting for its box's stage 1; "
                             "the step's result is not valid");
  }
  return CHL_OK;
}

// Steps [enqueue n steps] timed with events on the library stream.
template <class F>
int TimeOnStream(chl_ensemble *e, F &&enqueue, float *ms_out)
{
  hipEvent_t e0, e1;
  CHL_HIP(hipEventCreate(&e0));
  CHL_HIP(hipEventCreate(&e1));
  int rc = CHL_OK;
  if (hipEventRecord(e0, e->stream) != hipSuccess) rc = Fail(CHL_ERR_HIP, "hipEventRecord failed");
  if (rc == CHL_OK) rc = enqueue();
  float ms = 0.f;
  if (rc == CHL_OK && (hipEventRecord(e1, e->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                       hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
    rc = Fail(CHL_ERR_HIP, "timing events failed");
  (void)hipStreamSynchronize(e->stream);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms_out) *ms_out = ms;
  return rc;
}

double Uniform01(std::mt19937_64 &g) { return double(g() >> 11) * 0x1.0p-53; }

}  // namespace

extern "C" {

int chl_ensemble_persistent_pairs(int *pairs, int max_pairs, int *n_pairs)
{
  if (n_pairs) *n_pairs = chl::kNumPairs;
  for (int p = 0; p < std::min(max_pairs, chl::kNumPairs) && pairs; p++) {
    pairs[2 * p]     = chl::kPairs[p][0];
    pairs[2 * p + 1] = chl::kPairs[p][1];
  }
  return CHL_OK;
}

int chl_ensemble_create(const chl_ensemble_config *cfg, int n_boxes, const int *dims, chl_ensemble **out)
{
  if (!cfg || !dims || !out || n_boxes < 1) return Fail(CHL_ERR_INVALID, "bad argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return Fail(CHL_ERR_NO_DEVICE, "no HIP device");
  chl_ensemble_config c = *cfg;
  // the other solvers and reconstructions have kernels for tile shapes 0 and 1 only
  const bool narrow = c.riemann != CHL_RIEMANN_HLLC || c.reconstruction == CHL_RECON_PCM ||
                      c.reconstruction == CHL_RECON_PLMP || c.reconstruction == CHL_RECON_PPMP;
  if (c.tile1 < 0) c.tile1 = c.dual_energy || narrow ? kDefaultDeTile : kDefaultTile1;
  if (c.tile2 < 0) c.tile2 = c.dual_energy || narrow ? kDefaultDeTile : kDefaultTile2;
  if (c.kchunk1 <= 0) c.kchunk1 = kDefaultKchunk1;
  if (c.kchunk2 <= 0) c.kchunk2 = kDefaultKchunk2;
  CHL_TRY(CheckEnsembleConfig(c));
  const int ng = c.n_ghost;
  for (int b = 0; b < n_boxes; b++)
    for (int a = 0; a < 3; a++) {
      // every periodic image of a real cell must land in a ghost layer (fused ghost refresh)
      if (dims[3 * b + a] < 3 * ng)
        return Fail(CHL_ERR_INVALID, "box " + std::to_string(b) + ": each axis needs at least n_ghost real cells");
      if (dims[3 * b + a] > 65536) return Fail(CHL_ERR_INVALID, "box " + std::to_string(b) + " too large");
    }
  auto *e = new chl_ensemble();
  e->cfg  = c;
  e->n    = n_boxes;
  e->dims.assign(dims, dims + 3 * n_boxes);
  e->dxyz.assign(3 * n_boxes, 1.0);
  e->tout.assign(n_boxes, std::numeric_limits<double>::infinity());
  e->order.resize(n_boxes);
  for (int b = 0; b < n_boxes; b++) e->order[b] = b;
  std::stable_sort(e->order.begin(), e->order.end(), [&](int a, int b) { return e->Cells(a) > e->Cells(b); });
  int dev;
  hipDeviceProp_t p;
  hipError_t er = hipGetDevice(&dev);
  if (er == hipSuccess) er = hipGetDeviceProperties(&p, dev);
  if (er == hipSuccess) e->num_cu = p.multiProcessorCount;
  if (er == hipSuccess) er = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (er == hipSuccess) er = hipMalloc(reinterpret_cast<void **>(&e->box), n_boxes * sizeof(chl::BoxDesc));
  if (er == hipSuccess) er = hipMalloc(reinterpret_cast<void **>(&e->clk), n_boxes * sizeof(chl::BoxClock));
  if (er == hipSuccess) er = hipMalloc(reinterpret_cast<void **>(&e->head), chl_ensemble::kHeadBytes);
  if (er == hipSuccess) er = hipMemsetAsync(e->clk, 0, n_boxes * sizeof(chl::BoxClock), e->stream);
  if (er == hipSuccess) er = hipMemsetAsync(e->head, 0, chl_ensemble::kHeadBytes, e->stream);
  if (er == hipSuccess) er = hipStreamSynchronize(e->stream);
  if (er != hipSuccess) {
    chl_ensemble_destroy(e);
    return Fail(CHL_ERR_HIP, std::string("creating the ensemble: ") + hipGetErrorString(er));
  }
  int rc = BuildPlan(e);
  if (rc != CHL_OK) {
    std::string msg = chl_last_error();
    chl_ensemble_destroy(e);
    return Fail(rc, msg);
  }
  *out = e;
  return CHL_OK;
}

int chl_ensemble_destroy(chl_ensemble *e)
{
  if (!e) return CHL_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  e->DropGraph();
  if (e->W) (void)hipFree(e->W);
  if (e->items) (void)hipFree(e->items);
  if (e->box) (void)hipFree(e->box);
  if (e->clk) (void)hipFree(e->clk);
  if (e->head) (void)hipFree(e->head);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return CHL_OK;
}

int chl_ensemble_get_config(const chl_ensemble *e, chl_ensemble_config *out)
{
  if (!e || !out) return Fail(CHL_ERR_INVALID, "null argument");
  *out = e->cfg;
  return CHL_OK;
}

int chl_ensemble_bind(chl_ensemble *e, double *const *d_states, const double *dxyz)
{
  if (!e || !d_states || !dxyz) return Fail(CHL_ERR_INVALID, "null argument");
  for (int b = 0; b < e->n; b++) {
    if (!d_states[b]) return Fail(CHL_ERR_INVALID, "null state of box " + std::to_string(b));
    if (!(dxyz[3 * b] > 0 && dxyz[3 * b + 1] > 0 && dxyz[3 * b + 2] > 0))
      return Fail(CHL_ERR_INVALID, "cell widths must be positive");
  }
  e->U.assign(d_states, d_states + e->n);
  e->dxyz.assign(dxyz, dxyz + 3 * e->n);
  e->bound   = true;
  e->started = false;
  return BuildPlan(e);
}

int chl_ensemble_start(chl_ensemble *e, double gamma, double density_floor, double U_floor, const double *t0,
                       const double *tout)
{
  if (!e) return Fail(CHL_ERR_INVALID, "null argument");
  if (!e->bound) return Fail(CHL_ERR_INVALID, "bind the box states first");
  e->gamma = gamma, e->density_floor = density_floor, e->U_floor = U_floor;
  for (int b = 0; b < e->n; b++) e->tout[b] = tout ? tout[b] : std::numeric_limits<double>::infinity();
  CHL_TRY(BuildPlan(e));
  std::vector<chl::BoxClock> hc(e->n);
  for (int b = 0; b < e->n; b++) {
    hc[b] = chl::BoxClock{};
    hc[b].t = t0 ? t0[b] : 0.0;
  }
  CHL_HIP(hipMemcpyAsync(e->clk, hc.data(), e->n * sizeof(chl::BoxClock), hipMemcpyHostToDevice, e->stream));
  for (int b = 0; b < e->n; b++) {
    const chl::BoxDesc &d = e->hbox[b];
    CHL_HIP(chl::LaunchCalcDt(e->U[b], d.nx, d.ny, d.nz, e->cfg.n_ghost, d.dx, d.dy, d.dz, gamma, &e->clk[b].dti,
                              e->num_cu, e->stream));
  }
  CHL_HIP(hipStreamSynchronize(e->stream));
  e->started = true;
  return CHL_OK;
}

int chl_ensemble_advance(chl_ensemble *e, int n_steps, int use_graph, float *ms_out)
{
  if (!e || n_steps < 0) return Fail(CHL_ERR_INVALID, "bad argument");
  if (!e->started) return Fail(CHL_ERR_INVALID, "start the ensemble first");
  if (use_graph) CHL_TRY(EnsureGraph(e));
  auto enqueue = [&]() -> int {
    for (int s = 0; s < n_steps; s++) {
      if (use_graph) {
        CHL_HIP(hipGraphLaunch(e->graph, e->stream));
      } else {
        CHL_TRY(EnqueueStep(e));
      }
    }
    return CHL_OK;
  };
  CHL_TRY(TimeOnStream(e, enqueue, ms_out));
  return CheckStatus(e, n_steps > 0);
}

int chl_ensemble_advance_looped(chl_ensemble *e, int n_steps, float *ms_out)
{
  if (!e || n_steps < 0) return Fail(CHL_ERR_INVALID, "bad argument");
  if (!e->started) return Fail(CHL_ERR_INVALID, "start the ensemble first");
  const chl_ensemble_config &c = e->cfg;
  const int R1 = ReconRadius(c.reconstruction) + 1, pol = c.precision;
  auto enqueue = [&]() -> int {
    for (int s = 0; s < n_steps; s++) {
      for (int b : e->order) {
        const chl::BoxDesc &d = e->hbox[b];
        chl::BoxClock *k      = e->clk + b;
        CHL_HIP(chl::LaunchDtAdvance(&k->dti, &k->dt, &k->t, d.tout, kCfl, &k->nstep, e->stream));
        chl::StepArgs a{};
        a.U = d.U, a.W = d.W, a.pdv = d.pdv, a.nx = d.nx, a.ny = d.ny, a.nz = d.nz, a.ng = c.n_ghost;
        a.dx = d.dx, a.dy = d.dy, a.dz = d.dz, a.gamma = e->gamma, a.dt = 0.0, a.dt_dev = &k->dt, a.dti = &k->dti;
        a.density_floor = e->density_floor, a.U_floor = e->U_floor;
        a.flags = Args(e, 0).flags;
        a.de     = c.dual_energy;
        a.riemann = c.riemann;
        a.rep    = d.rep;
        a.kchunk = d.kchunk1;
        if (a.rep.side) CHL_HIP(chl::LaunchRepairSide(a, e->stream));
        CHL_HIP(pol == CHL_PREC_FP64_EXACT ? chl::LaunchStage1Exact(a, R1, c.tile1, e->stream)
                                           : chl::LaunchStage1Fast(pol, a, R1, c.tile1, e->stream));
        a.kchunk = d.kchunk2;
        CHL_HIP(pol == CHL_PREC_FP64_EXACT ? chl::LaunchStage2Exact(a, c.reconstruction, c.tile2, e->stream)
                                           : chl::LaunchStage2Fast(pol, a, c.reconstruction, c.tile2, e->stream));
        CHL_HIP(pol == CHL_PREC_FP64_EXACT ? chl::LaunchRepairExact(a, e->stream)
                                           : chl::LaunchRepairFast(pol, a, e->stream));
        if (c.dual_energy) {
          CHL_HIP(chl::LaunchDeSelect(d.U, d.nx, d.ny, d.nz, c.n_ghost, e->stream));
          CHL_HIP(chl::LaunchDeSync(d.U, d.nx, d.ny, d.nz, c.n_ghost, a.flags, a.U_floor, a.dx, a.dy, a.dz, a.gamma,
                                    a.dti, e->stream));
        }
      }
    }
    return CHL_OK;
  };
  return TimeOnStream(e, enqueue, ms_out);
}

int chl_ensemble_clocks(chl_ensemble *e, double *t, double *dt, int *n_steps)
{
  if (!e) return Fail(CHL_ERR_INVALID, "null argument");
  std::vector<chl::BoxClock> hc(e->n);
  CHL_HIP(hipMemcpyAsync(hc.data(), e->clk, e->n * sizeof(chl::BoxClock), hipMemcpyDeviceToHost, e->stream));
  CHL_HIP(hipStreamSynchronize(e->stream));
  for (int b = 0; b < e->n; b++) {
    if (t) t[b] = hc[b].t;
    if (dt) dt[b] = hc[b].dt;
    if (n_steps) n_steps[b] = hc[b].nstep;
  }
  return CHL_OK;
}

int chl_ensemble_autotune(chl_ensemble *e, int steps, chl_ensemble_config *chosen)
{
  if (!e || steps < 1) return Fail(CHL_ERR_INVALID, "bad argument");
  if (!e->started) return Fail(CHL_ERR_INVALID, "start the ensemble first");
  const chl_ensemble_config orig = e->cfg;
  // scratch copy of every state, and the clocks to put back
  std::vector<chl::BoxClock> saved(e->n);
  CHL_HIP(hipMemcpyAsync(saved.data(), e->clk, e->n * sizeof(chl::BoxClock), hipMemcpyDeviceToHost, e->stream));
  size_t total = 0;
  std::vector<size_t> off(e->n);
  const size_t nf = size_t(orig.n_fields);
  for (int b = 0; b < e->n; b++) off[b] = total, total += nf * e->Cells(b);
  double *S = nullptr;
  CHL_HIP(hipMalloc(&S, total * sizeof(double)));
  std::vector<double *> scratch(e->n);
  int rc = CHL_OK;
  for (int b = 0; b < e->n && rc == CHL_OK; b++) {
    scratch[b] = S + off[b];
    if (hipMemcpyAsync(scratch[b], e->U[b], nf * e->Cells(b) * sizeof(double), hipMemcpyDeviceToDevice, e->stream) !=
        hipSuccess)
      rc = Fail(CHL_ERR_HIP, "scratch copy failed");
  }
  auto plan = [&](const chl_ensemble_config &c) {
    e->cfg = c;
    return BuildPlan(e, &scratch);
  };
  // times `steps` repetitions of f after one untimed one
  auto timed = [&](auto &&f, float *ms) -> int {
    CHL_TRY(f());
    CHL_HIP(hipStreamSynchronize(e->stream));
    return TimeOnStream(
        e,
        [&]() -> int {
          for (int s = 0; s < steps; s++) CHL_TRY(f());
          return CHL_OK;
        },
        ms);
  };
  chl_ensemble_config best = orig;
  float best_ms            = std::numeric_limits<float>::infinity();
  if (rc == CHL_OK) rc = plan(orig);
  // one clock step so the boxes carry a non-zero dt
  if (rc == CHL_OK && chl::LaunchEnsembleClock(e->clk, e->box, e->n, kCfl, e->head, e->stream) != hipSuccess)
    rc = Fail(CHL_ERR_HIP, "clock launch failed");
  const int chunks[] = {8, 16, 32, 64};
  // split launch: stage 1 then stage 2, each on its own
  chl_ensemble_config sc = orig;
  sc.launch              = CHL_LAUNCH_SPLIT;
  float b1 = std::numeric_limits<float>::infinity(), b2 = b1;
  // the other solvers and reconstructions: tile shapes 0 and 1, split launch only
  const bool narrow = orig.riemann != CHL_RIEMANN_HLLC || orig.reconstruction == CHL_RECON_PCM ||
                      orig.reconstruction == CHL_RECON_PLMP || orig.reconstruction == CHL_RECON_PPMP;
  const int ntiles = orig.dual_energy || narrow ? chl::kNumDeTiles : chl::kNumTiles;
  if (narrow) sc.tile1 = std::min(sc.tile1, chl::kNumDeTiles - 1), sc.tile2 = std::min(sc.tile2, chl::kNumDeTiles - 1);
  for (int t = 0; t < ntiles && rc == CHL_OK; t++)
    for (int kc : chunks) {
      chl_ensemble_config c = sc;
      c.tile1 = t, c.kchunk1 = kc;
      float ms = 0;
      if ((rc = plan(c)) != CHL_OK) break;
      if ((rc = timed([&] { return EnqueueStage(e, 1); }, &ms)) != CHL_OK) break;
      if (ms < b1) b1 = ms, sc.tile1 = t, sc.kchunk1 = kc;
    }
  for (int t = 0; t < ntiles && rc == CHL_OK; t++)
    for (int kc : chunks) {
      chl_ensemble_config c = sc;
      c.tile2 = t, c.kchunk2 = kc;
      float ms = 0;
      if ((rc = plan(c)) != CHL_OK) break;
      if ((rc = timed([&] { return EnqueueStage(e, 2); }, &ms)) != CHL_OK) break;
      if (ms < b2) b2 = ms, sc.tile2 = t, sc.kchunk2 = kc;
    }
  if (rc == CHL_OK && (rc = plan(sc)) == CHL_OK) {
    float ms = 0;
    rc = timed([&] { return EnqueueStep(e); }, &ms);
    if (rc == CHL_OK && ms < best_ms) best_ms = ms, best = sc;
  }
  // persistent launch: whole steps
  const int kc_pairs[5][2] = {{8, 16}, {8, 32}, {16, 16}, {16, 32}, {32, 32}};
  for (int p = 0; p < chl::kNumPairs && rc == CHL_OK && !narrow; p++)
    for (const auto &kc : kc_pairs) {
      if (orig.dual_energy && !DePairIdx(p)) continue;
      chl_ensemble_config c = orig;
      c.launch = CHL_LAUNCH_PERSISTENT, c.tile1 = chl::kPairs[p][0], c.tile2 = chl::kPairs[p][1];
      c.kchunk1 = kc[0], c.kchunk2 = kc[1];
      float ms = 0;
      if ((rc = plan(c)) != CHL_OK) break;
      if ((rc = timed([&] { return EnqueueStep(e); }, &ms)) != CHL_OK) break;
      if ((rc = CheckStatus(e, true)) != CHL_OK) break;
      if (ms < best_ms) best_ms = ms, best = c;
    }
  const std::string err = rc == CHL_OK ? std::string() : std::string(chl_last_error());
  // put the clocks and states back
  (void)hipStreamSynchronize(e->stream);
  (void)hipMemcpy(e->clk, saved.data(), e->n * sizeof(chl::BoxClock), hipMemcpyHostToDevice);
  (void)hipFree(S);
  e->cfg     = rc == CHL_OK ? best : orig;
  int rc2    = BuildPlan(e);
  if (rc != CHL_OK) return Fail(rc, err);
  if (rc2 != CHL_OK) return rc2;
  if (chosen) *chosen = e->cfg;
  return CHL_OK;
}

int chl_init_gas_energy(double *d_U, int nx, int ny, int nz)
{
  if (!d_U || nx < 1 || ny < 1 || nz < 1) return Fail(CHL_ERR_INVALID, "bad argument");
  CHL_HIP(chl::LaunchGasEnergyInit(d_U, long(nx) * ny * nz, nullptr));
  CHL_HIP(hipDeviceSynchronize());
  return CHL_OK;
}

int chl_ensemble_layout(uint64_t seed, int n_boxes, int *edges, double *mach)
{
  if (n_boxes < 0 || (n_boxes > 0 && (!edges || !mach))) return Fail(CHL_ERR_INVALID, "bad argument");
  static const int kEdges[5] = {32, 48, 64, 96, 128};
  std::mt19937_64 g(seed);
  for (int b = 0; b < n_boxes; b++) {
    edges[b] = kEdges[g() % 5];
    mach[b]  = 0.5 + 2.5 * Uniform01(g);
  }
  return CHL_OK;
}

int chl_turbulence_modes(uint64_t seed, double *modes, int max_modes, int *n_modes)
{
  if (!n_modes) return Fail(CHL_ERR_INVALID, "null argument");
  std::mt19937_64 g(seed);
  auto normal_pair = [&](double &z0, double &z1) {
    const double u1 = 1.0 - Uniform01(g), u2 = Uniform01(g);  // u1 in (0, 1]
    const double r  = std::sqrt(-2.0 * std::log(u1));
    z0 = r * std::cos(2.0 * M_PI * u2);
    z1 = r * std::sin(2.0 * M_PI * u2);
  };
  int m = 0;
  for (int kx = 0; kx <= 4; kx++)
    for (int ky = -4; ky <= 4; ky++)
      for (int kz = -4; kz <= 4; kz++) {
        const int k2 = kx * kx + ky * ky + kz * kz;
        if (k2 < 1 || k2 > 16) continue;
        if (!(kx > 0 || (kx == 0 && (ky > 0 || (ky == 0 && kz > 0))))) continue;
        double v[6];
        normal_pair(v[0], v[1]);
        normal_pair(v[2], v[3]);
        normal_pair(v[4], v[5]);
        if (modes && m < max_modes) {
          const double sigma = 1.0 / double(k2);  // |k|^-2: shell spectrum E(k) ~ k^2 |k|^-4 = k^-2
          double *md         = modes + 9 * m;
          md[0] = kx, md[1] = ky, md[2] = kz;
          for (int c = 0; c < 2; c++) {
            double a[3] = {sigma * v[3 * c], sigma * v[3 * c + 1], sigma * v[3 * c + 2]};
            const double kd = (kx * a[0] + ky * a[1] + kz * a[2]) / k2;
            md[3 + 3 * c] = a[0] - kd * kx;
            md[4 + 3 * c] = a[1] - kd * ky;
            md[5 + 3 * c] = a[2] - kd * kz;
          }
        }
        m++;
      }
  *n_modes = m;
  if (modes && m > max_modes) return Fail(CHL_ERR_INVALID, "max_modes too small");
  return CHL_OK;
}

int chl_ic_turbulence(double *d_U, int n, int ng, double gamma, double rho, double p, double mach, uint64_t seed)
{
  if (!d_U || n < 1 || ng < 0 || ng > n || !(rho > 0) || !(p > 0) || !(gamma > 1) || !(mach >= 0))
    return Fail(CHL_ERR_INVALID, "bad argument");
  int m = 0;
  CHL_TRY(chl_turbulence_modes(seed, nullptr, 0, &m));
  std::vector<double> modes(9 * size_t(m));
  CHL_TRY(chl_turbulence_modes(seed, modes.data(), m, &m));
  // rms |v| of the unscaled field over a periodic grid of more than 8 cells per axis:
  // the modes are orthogonal there, each contributing (|a|^2 + |b|^2) / 2
  double ms = 0;
  for (int q = 0; q < m; q++)
    for (int c = 3; c < 9; c++) ms += 0.5 * modes[9 * q + c] * modes[9 * q + c];
  const double vscale = mach * std::sqrt(gamma * p / rho) / std::sqrt(ms);
  double *d_modes     = nullptr;
  CHL_HIP(hipMalloc(&d_modes, modes.size() * sizeof(double)));
  hipError_t er = hipMemcpy(d_modes, modes.data(), modes.size() * sizeof(double), hipMemcpyHostToDevice);
  if (er == hipSuccess) er = chl::LaunchTurbulenceIc(d_U, n, ng, m, d_modes, vscale, rho, p, gamma, nullptr);
  if (er == hipSuccess) er = hipDeviceSynchronize();
  (void)hipFree(d_modes);
  CHL_HIP(er);
  return CHL_OK;
}

}  // extern "C"
