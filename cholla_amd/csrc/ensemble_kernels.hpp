// ensemble_kernels.hpp -- one VL step over an ensemble of independent periodic boxes.
//
// The reference advances one box per process and launches ~20 kernels per step
// for it (src/integrators/VL_3D_cuda.cu:38-360).  An ensemble of small boxes
// launched that way leaves most of the GPU idle: a 40^3 box is a few workgroups.
// Here every box's tiles go into one work list and a step is:
//
//   EnsembleClockKernel      per-box dt = C_cfl / dti (clamped to the box's output
//                            time), t += dt, dti reset (the reference's set_dt /
//                            Update_Time, src/grid/grid3D.cpp:373-411, 587-600)
//   then either
//   EnsembleStage1Kernel +   one launch per stage over the tiles of all boxes
//   EnsembleStage2Kernel     ("split" launch)
//   or
//   EnsemblePersistentKernel one launch per step: resident workgroups pull items
//                            from their XCD's queue, which holds the stage-1
//                            tiles of that XCD's boxes and, a generation of
//                            workgroups later, each box's stage-2 tiles; a
//                            stage-2 tile waits until its box's stage-1 tiles
//                            are all done ("persistent" launch).
//
// Each box keeps its own dt: the tile bodies are the single-box ones
// (Stage1Tile / Stage2Tile) fed with that box's StepArgs.
//
// With dual energy two more phases follow stage 2 per box, over the stage-2 tiles: the
// internal-energy selection, then the energy sync with the floor, the inverse-dt
// reduction and the ghost images (DeSelectCell / DeSyncCell).  The split launch gives
// each its own kernel; the persistent queue carries them as stages 2 and 3, each item
// waiting for its box's previous phase.
#pragma once

#include <hip/hip_runtime.h>

#include "de_cells.hpp"
#include "vl3d_kernels.hpp"

namespace chl
{

// One box of the ensemble, resident on the device.
struct BoxDesc {
  double *U;           // conserved state, 5 (6 with dual energy) x cells (in/out, owned by the caller)
  void *W;             // half-step primitives of this box
  double *pdv;         // dual energy: full-step pdV source per cell, else null
  int nx, ny, nz;      // totals, ghosts included
  int t1[3], t2[3];    // tile counts (x, y, z-chunks) of stage 1 and stage 2
  int kchunk1, kchunk2;
  int items1;          // t1[0] * t1[1] * t1[2]
  int items2;          // t2[0] * t2[1] * t2[2]
  double dx, dy, dz;
  double tout;         // output time of the box (dt is clamped to reach it)
  RepairState rep;     // the box's repair marks (see RepairKernel)
};

// Time state of one box.
struct BoxClock {
  double dt, t;
  unsigned long long dti;  // bits of the non-negative inverse dt (atomicMax target)
  int nstep;
  int done1;  // stage-1 tiles finished in the current step (persistent launch)
  int done2;  // dual energy: stage-2 tiles finished
  int done3;  // dual energy: selection tiles finished
  int pad_;
};

// A work item: box and the tile's linear index l = bx + tx * (by + ty * bz).
struct WorkItem {
  int box;
  int stage;  // 0, 1; 2 and 3 with dual energy (persistent queue only)
  int local;
  int pad_;
};

// The persistent launch keeps one queue per XCD (see EnsemblePersistentKernel).
constexpr int kQueueXcds  = 8;
constexpr int kHeadStride = 32;  // queue heads 128 bytes apart

struct EnsembleArgs {
  const BoxDesc *box;
  BoxClock *clk;
  const WorkItem *items;  // stage-1 list, stage-2 list or the persistent queue
  int n_items;
  int n_boxes;
  int ng;
  int R1;  // stage-1 region growth: reconstruction radius + 1
  double gamma, density_floor, U_floor;
  int flags;
  unsigned int *head;  // persistent queue heads, XCD x at head[x * kHeadStride]
  int *status;         // set non-zero if a dependency wait gave up
  int qoff[kQueueXcds + 1];  // persistent queue of XCD x: items[qoff[x] .. qoff[x + 1])
  long long spin_limit;
  int de;              // dual energy (host-side dispatch only; the kernels are templated on it)
  int riemann;         // kRiemann* (host-side dispatch only: the solver is part of the policy)
};

__device__ __forceinline__ StepArgs BoxStepArgs(const EnsembleArgs &e, int b, int kchunk)
{
  const BoxDesc &d = e.box[b];
  StepArgs a{};
  a.U = d.U, a.W = d.W, a.pdv = d.pdv;
  a.nx = d.nx, a.ny = d.ny, a.nz = d.nz, a.ng = e.ng;
  a.dx = d.dx, a.dy = d.dy, a.dz = d.dz;
  a.gamma         = e.gamma;
  a.dt            = 0.0;
  a.dt_dev        = &e.clk[b].dt;
  a.dti           = &e.clk[b].dti;
  a.density_floor = e.density_floor;
  a.U_floor       = e.U_floor;
  a.flags         = e.flags;
  a.kchunk        = kchunk;
  a.rep           = d.rep;
  return a;
}

__device__ __forceinline__ void TileCoords(int l, const int t[3], int &bx, int &by, int &bz)
{
  bx = l % t[0];
  int r = l / t[0];
  by = r % t[1];
  bz = r / t[1];
}

// Per-box dt from the previous step's reduction, as DtAdvanceKernel does for one box.
__global__ void EnsembleClockKernel(BoxClock *clk, const BoxDesc *box, int n_boxes, double cfl, unsigned int *head);

template <class P, int TX, int TY, bool DE = false>
__global__ void __launch_bounds__(TX *TY, (kStage1MinWaves<P, DE>)) EnsembleStage1Kernel(EnsembleArgs e)
{
  __shared__ Stage1Smem<P, TX, TY, DE> sm;
  const int q = XcdItem(blockIdx.x, e.n_items);
  if (q >= e.n_items) return;
  const WorkItem it = e.items[q];
  const BoxDesc &d  = e.box[it.box];
  int bx, by, bz;
  TileCoords(it.local, d.t1, bx, by, bz);
  StepArgs a = BoxStepArgs(e, it.box, d.kchunk1);
  Stage1Tile<P, TX, TY, DE>(a, e.R1, bx, by, bz, sm);
}

template <class P, int RK, int TX, int TY, bool DE = false>
__global__ void __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(Stage2Occupancy<P>::kMinWaves)))
EnsembleStage2Kernel(EnsembleArgs e)
{
  __shared__ Stage2Smem<P, RK, TX, TY, DE> sm;
  const int q = XcdItem(blockIdx.x, e.n_items);
  if (q >= e.n_items) return;
  const WorkItem it = e.items[q];
  const BoxDesc &d  = e.box[it.box];
  int bx, by, bz;
  TileCoords(it.local, d.t2, bx, by, bz);
  StepArgs a = BoxStepArgs(e, it.box, d.kchunk2);
  Stage2Tile<P, RK, TX, TY, DE>(a, bx, by, bz, sm);
}

// The repairs stage 2 marked (RepairBox), one workgroup per box, after the step's stage 2;
// a box without marks returns after reading its count.
template <class P, bool DE>
__global__ void __launch_bounds__(kRepairThreads) EnsembleRepairKernel(EnsembleArgs e)
{
  StepArgs a = BoxStepArgs(e, blockIdx.x, 1);
  RepairBox<P, DE>(a);
}

// The side buffers of the boxes (RepairSideCell), before stage 2 rewrites their ghosts:
// blockIdx.y is the box.
__global__ void EnsembleRepairSideKernel(EnsembleArgs e);

// Dual-energy phase PH (0: selection, 1: sync) over the real cells of stage-2 tile
// (bx, by, bz) of box d; returns this lane's inverse-dt maximum (sync only).
template <int TX, int TY, int PH>
__device__ __forceinline__ double DeTile(const EnsembleArgs &e, const BoxDesc &d, int bx, int by, int bz)
{
  const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX, ng = e.ng;
  const int i = ng + bx * TX + tx, j = ng + by * TY + ty;
  const int k0 = ng + bz * d.kchunk2, k1 = min(k0 + d.kchunk2, d.nz - ng);
  double m = 0.0;
  if (i < d.nx - ng && j < d.ny - ng) {
    const size_t nc = size_t(d.nx) * d.ny * d.nz;
    for (int k = k0; k < k1; k++) {
      if constexpr (PH == 0) {
        DeSelectCell(d.U, nc, d.nx, d.ny, d.nz, ng, i, j, k);
      } else {
        m = Fmax(m, DeSyncCell(d.U, nc, d.nx, d.ny, d.nz, ng, i, j, k, e.flags, e.U_floor, d.dx, d.dy, d.dz,
                               e.gamma));
      }
    }
  }
  return m;
}

// Workgroup maximum of v (wave64 shuffles, then LDS) into *out with one atomic.
template <int NT>
__device__ __forceinline__ void BlockMaxAtomic(double v, unsigned long long *out, double *red)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = Fmax(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < NT / 64; w++) m = Fmax(m, red[w]);
    atomicMax(out, (unsigned long long)__double_as_longlong(m));
  }
}

// Split launch of a dual-energy phase over the stage-2 tiles of every box.
template <int TX, int TY, int PH>
__global__ void __launch_bounds__(TX *TY) EnsembleDeKernel(EnsembleArgs e)
{
  __shared__ double red[TX * TY / 64];
  const int q = XcdItem(blockIdx.x, e.n_items);
  if (q >= e.n_items) return;
  const WorkItem it = e.items[q];
  const BoxDesc &d  = e.box[it.box];
  int bx, by, bz;
  TileCoords(it.local, d.t2, bx, by, bz);
  const double m = DeTile<TX, TY, PH>(e, d, bx, by, bz);
  if (PH == 1 && (e.flags & kFlagReduceDt)) BlockMaxAtomic<TX * TY>(m, &e.clk[it.box].dti, red);
}

template <class A, class B>
union SmemPair {
  A s1;
  B s2;
};

// Persistent-launch dependencies.  Every phase of a box runs on one XCD (its queue's),
// so what one tile of a box writes and another reads only has to cross that XCD's L2,
// which all its CUs share.  The finishing side: every lane waits until its stores have
// been acknowledged by the L2 (the vector L1 is write-through), the workgroup meets, one
// lane counts the tile; a device-scope release would also write the XCD's whole L2 back to
// HBM for every tile.  The waiting side: one lane spins on the count (bounded: after
// spin_limit sleeps it sets *status and gives up, as does every later wait once one has),
// the workgroup meets and acquires at device scope, which drops the CU's L1 lines: the CU
// may hold lines of the box read earlier in the same launch (a stage-1 tile reads the state
// that the box's stage-2 tiles then overwrite).  A workgroup-scope L1 invalidate is not
// enough: outside thread-group-split mode the hardware treats it as a no-op.
__device__ __forceinline__ void SignalTileDone(int *count)
{
  // The wait intrinsic is declared as touching no memory, so nothing stops the compiler
  // from moving a load or store of the tile past it; a signal fence (a compiler barrier,
  // no instruction) keeps every memory access of the tile ahead of the wait.
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(count, 1);
}
__device__ __forceinline__ void WaitTiles(const EnsembleArgs &e, int *count, int need)
{
  if (threadIdx.x == 0) {
    long long n = 0;
    while (__hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
      __builtin_amdgcn_s_sleep(2);
      if (++n > e.spin_limit) {
        atomicOr(e.status, 1);
        break;
      }
      // once one wait has given up, the others stop waiting too
      if ((n & 1023) == 0 && __hip_atomic_load(e.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// XCD of the calling workgroup (the XCC_ID hardware register).
__device__ __forceinline__ int XcdId() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) % kQueueXcds; }

// One whole step in one launch.  Each workgroup serves the queue of the XCD it runs
// on; the host splits the boxes over the XCD queues by cell count.  Deadlock freedom:
// items are taken from a queue in order, and every stage-2 item of box b sits after all
// stage-1 items of b in the same queue, so when a workgroup waits on box b, each stage-1
// item of b is already held by a running workgroup of that XCD, and stage-1 items never
// wait.  The wait is still bounded (spin_limit sleeps); if it gives up, *status is set
// and the host reports it, as it does a queue that no workgroup drained.
// The persistent kernel holds both stage bodies; left to itself the register
// allocator sizes it for two waves per SIMD, one fewer than the mixed form's
// stage kernels run at.
template <class P>
struct PersistentOccupancy {
  static constexpr int kMinWaves = P::kId == PolicyMixed::kId ? 3 : 2;
};

template <class P, int RK, int TX1, int TY1, int TX2, int TY2, bool DE = false>
__global__ void __launch_bounds__(TX2 *TY2) __attribute__((amdgpu_waves_per_eu(PersistentOccupancy<P>::kMinWaves)))
EnsemblePersistentKernel(EnsembleArgs e)
{
  static_assert(TX1 * TY1 == TX2 * TY2, "both stages run on the same workgroup");
  __shared__ SmemPair<Stage1Smem<P, TX1, TY1, DE>, Stage2Smem<P, RK, TX2, TY2, DE>> sm;
  __shared__ double s_red[TX2 * TY2 / 64];
  __shared__ int s_item;
  const int tid = threadIdx.x;
  const int xcd = XcdId();
  unsigned int *qhead = e.head + xcd * kHeadStride;
  const int q_lo = e.qoff[xcd], q_hi = e.qoff[xcd + 1];
  for (;;) {
    __syncthreads();  // the previous item is finished with LDS and s_item
    if (tid == 0) s_item = q_lo + int(atomicAdd(qhead, 1u));
    __syncthreads();
    const int q = __builtin_amdgcn_readfirstlane(s_item);
    if (q >= q_hi) break;
    const WorkItem it = e.items[q];
    const int b       = __builtin_amdgcn_readfirstlane(it.box);
    const BoxDesc &d  = e.box[b];
    int bx, by, bz;
    const int stage = __builtin_amdgcn_readfirstlane(it.stage);
    if (stage == 0) {
      TileCoords(it.local, d.t1, bx, by, bz);
      StepArgs a = BoxStepArgs(e, b, d.kchunk1);
      Stage1Tile<P, TX1, TY1, DE>(a, e.R1, bx, by, bz, sm.s1);
      SignalTileDone(&e.clk[b].done1);  // the box's W (and pdV sources), device-wide
    } else if (stage == 1) {
      WaitTiles(e, &e.clk[b].done1, d.items1);
      TileCoords(it.local, d.t2, bx, by, bz);
      StepArgs a = BoxStepArgs(e, b, d.kchunk2);
      Stage2Tile<P, RK, TX2, TY2, DE>(a, bx, by, bz, sm.s2);
      if constexpr (DE) SignalTileDone(&e.clk[b].done2);
    } else if constexpr (DE) {
      TileCoords(it.local, d.t2, bx, by, bz);
      if (stage == 2) {
        WaitTiles(e, &e.clk[b].done2, d.items2);  // the box's new state
        DeTile<TX2, TY2, 0>(e, d, bx, by, bz);
        SignalTileDone(&e.clk[b].done3);
      } else {
        WaitTiles(e, &e.clk[b].done3, d.items2);  // every selected GasEnergy of the box
        const double m = DeTile<TX2, TY2, 1>(e, d, bx, by, bz);
        if (e.flags & kFlagReduceDt) BlockMaxAtomic<TX2 * TY2>(m, &e.clk[b].dti, s_red);
      }
    }
  }
}

// Tile pairs (stage-1 tile, stage-2 tile; indices into kTiles) compiled for the
// persistent launch.  Both stages of a pair use 256 lanes.
constexpr int kNumPairs           = 4;
constexpr int kPairs[kNumPairs][2] = {{2, 1}, {2, 0}, {0, 0}, {1, 1}};

// Host launchers, instantiated per policy in ensemble_exact.hip / ensemble_fast.hip.
hipError_t LaunchEnsembleStage1(int policy, const EnsembleArgs &e, int tile, hipStream_t s);
hipError_t LaunchEnsembleStage2(int policy, const EnsembleArgs &e, int recon, int tile, hipStream_t s);
hipError_t LaunchEnsemblePersistent(int policy, const EnsembleArgs &e, int recon, int pair, int grid, hipStream_t s);
// Resident workgroups per CU of the persistent kernel (the grid is that times the CU count).
hipError_t EnsemblePersistentSlots(int policy, int recon, int pair, int de, int *per_cu);
// Dual-energy phase (0: selection, 1: sync) of the split launch over e's item list (the
// stage-2 tiles); tile: the stage-2 tile shape (one of the dual-energy tiles).
hipError_t LaunchEnsembleDe(int phase, const EnsembleArgs &e, int tile, hipStream_t s);
hipError_t LaunchEnsembleClock(BoxClock *clk, const BoxDesc *box, int n_boxes, double cfl, unsigned int *head,
                               hipStream_t s);
// The repairs of every box after stage 2 (EnsembleRepairKernel), and the boxes' side buffers
// before it (max_face_cells: the largest side buffer of a box, in cells).
hipError_t LaunchEnsembleRepair(int policy, const EnsembleArgs &e, hipStream_t s);
hipError_t LaunchEnsembleRepairSide(const EnsembleArgs &e, long max_face_cells, hipStream_t s);

}  // namespace chl
