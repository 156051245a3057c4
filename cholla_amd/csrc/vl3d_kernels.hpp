// vl3d_kernels.hpp -- fused kernels of one van Leer (VL) step on a 3D box, for gfx950.
//
// The reference runs one VL step as ~20 unfused launches that stream ~225
// field-arrays per cell through HBM (src/integrators/VL_3D_cuda.cu:38-360).  Here
// a step is two kernels, both "2.5D marches": a workgroup owns a TX x TY column
// of cells and walks up the z axis; the z stencil lives in each lane's registers,
// the x/y stencils of the current z-plane are staged in LDS, and interface states
// and face fluxes are exchanged through LDS so each face is solved exactly once.
//
//   Stage1Kernel  PCM states + HLLC on every face of the half-step region, the
//                 dt/2 conserved update (Update_Conserved_Variables_3D_half,
//                 VL_3D_cuda.cu:379-495), then the primitive variables of the
//                 half state (Load_Data, reconstruction.h:139-212) written once
//                 per cell as W (FP64, or FP32 in the mixed form).
//   Stage2Kernel  PLMC/PPMC on W + HLLC per face + the full dt update of the real
//                 cells in place (hydro_cuda.cu:172-402), with the density floor,
//                 temperature floor (hydro_cuda.cu:1117-1153), the next-step
//                 inverse-dt reduction (Calc_dt_3D, hydro_cuda.cu:514-560) and,
//                 for all-periodic boxes, the ghost-cell images
//                 (cuda_boundaries.cu:71-159 with flag 1) fused into the epilogue.
//
// Index convention: U[f*n_cells + i + j*nx + k*nx*ny], 64-bit offsets.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "hydro_physics.hpp"
#include "static_gravity.hpp"

namespace chl
{

enum ReconKind : int { kPCM = 0, kPLMC = 1, kPPMC = 2, kPLMP = 3, kPPMP = 4 };

template <int R>
struct ReconTraits;
template <>
struct ReconTraits<kPLMC> {
  static constexpr int kRadius = 1;
};
template <>
struct ReconTraits<kPPMC> {
  static constexpr int kRadius = 2;
};
template <>
struct ReconTraits<kPLMP> {
  static constexpr int kRadius = 1;
};
template <>
struct ReconTraits<kPPMP> {
  static constexpr int kRadius = 2;
};
template <>
struct ReconTraits<kPCM> {
  static constexpr int kRadius = 0;
};

// Cells whose update ends in one of the reference's repairs: the average of the six
// neighbours' densities for a non-positive density under the density floor
// (Average_Cell_Single_Field, hydro_cuda.cu:293-297, 1155-1177), and the average of the
// healthy cells of the 3x3x3 block when the density or energy comes out negative or NaN in
// a build without both floors (Average_Cell_All_Fields, hydro_cuda.cu:375-381, 1179-1268).
// The reference runs them inside the update kernel on neighbours other threads are
// updating at the same time; here stage 2 only marks such a cell and RepairKernel, the
// next launch, repairs the marks in cell order from the finished update.
enum : unsigned char { kRepairDensity = 1, kRepairAll = 2 };
constexpr int kRepairChunkShift = 12;  // marks are found through one byte per 4096 cells

struct RepairState {
  unsigned char *mark;   // per cell: 0, kRepairDensity or kRepairAll
  unsigned char *chunk;  // per 4096 cells: 1 if any cell in it is marked
  unsigned *count;       // [0] marked cells this step, [1] gravity records, [2] records that did not fit
  double *side;          // fused periodic ghosts: pre-step density of the six boundary layers
  double *grav;          // static gravity: (cell id, d, mx, my, mz) of marked cells at the step's start
  int grav_cap;          // records grav holds
};

// Parameters shared by both stages.
struct StepArgs {
  double *U;         // conserved state, n_fields x n_cells (in/out)
  void *W;           // half-step primitives, 5 (6 with dual energy) x n_cells of P::T
  double *pdv;       // dual energy: full-step pdV source of each real cell (n_cells), else null
  int nx, ny, nz;    // total cells per axis, ghosts included
  int ng;            // ghost depth
  double dx, dy, dz;
  double gamma;
  double dt;               // full step (used when dt_dev is null)
  const double *dt_dev;    // device-resident dt (device time-stepping), or null
  unsigned long long *dti; // inverse-dt max accumulator (bits of a non-negative double)
  double density_floor;
  double U_floor;
  int flags;               // kFlag* bits
  int kchunk;              // z cells per workgroup
  int gx, gy, gz;          // tile grid of the launch (set by the launcher)
  int xcd_order;           // 1: give each XCD a contiguous run of tiles (see TileOf)
  int de;                  // dual energy (host-side dispatch only; the kernels are templated on it)
  int riemann;             // kRiemannHllc / kRiemannHll (host-side dispatch only, as de)
  GravSpec grav;           // static gravity (kernels under a GravPolicy only); grav.id 0 = none
  RepairState rep;         // marked cells (kFlagDensityFloor / kFlagCrashRepair), see RepairKernel
};

enum : int {
  kFlagDensityFloor       = 1,
  kFlagTemperatureFloor   = 2,
  kFlagPeriodicGhosts     = 4,  // write periodic ghost images from the stage-2 epilogue
  kFlagReduceDt           = 8,
  kFlagPositivityFallback = 16,  // PPMC positivity fallback (see ReconPpmc)
  kFlagConservedW         = 32,  // PCM: stage 1 stores momenta (exact form) or the total-energy pressure (DE)
  kFlagCrashRepair        = 64,  // a build without both floors: mark negative / NaN cells (kRepairAll)
  kFlagResetDti           = 128, // stage 1 zeroes the inverse-dt accumulator (a later kernel fills it)
  kFlagFoldDt             = 256, // MHD: the stage-1 flux launch reduces the step's own inverse dt
};

__device__ __forceinline__ size_t Lin(int i, int j, int k, int nx, int ny)
{
  return size_t(i) + size_t(j) * size_t(nx) + size_t(k) * size_t(nx) * size_t(ny);
}

__device__ __forceinline__ double StepDt(const StepArgs &a) { return a.dt_dev ? *a.dt_dev : a.dt; }

// Workgroups are handed to the 8 XCDs round-robin by id; this gives each XCD a
// contiguous run of work items, so neighbouring tiles (which share halo cells)
// are resident in the same L2.  The grid is padded to a multiple of 8.
__device__ __forceinline__ int XcdItem(int g, int n_items)
{
  const int per = (n_items + 7) / 8;
  return (g & 7) * per + (g >> 3);
}

// Tile (bx, by, bz) of this workgroup in a 1D launch over the gx x gy x gz tile grid.
__device__ __forceinline__ bool TileOf(const StepArgs &a, int &bx, int &by, int &bz)
{
  const int T = a.gx * a.gy * a.gz;
  int q       = int(blockIdx.x);
  if (a.xcd_order) q = XcdItem(q, T);
  if (q >= T) return false;
  bx    = q % a.gx;
  int r = q / a.gx;
  by    = r % a.gy;
  bz    = r / a.gy;
  return true;
}

// Rotate a global-order (x,y,z) triple into the sweep frame of dir.
template <int DIR, typename T>
__device__ __forceinline__ void Rot(T gx, T gy, T gz, T &n, T &t1, T &t2)
{
  if constexpr (DIR == 0) {
    n = gx, t1 = gy, t2 = gz;
  } else if constexpr (DIR == 1) {
    n = gy, t1 = gz, t2 = gx;
  } else {
    n = gz, t1 = gx, t2 = gy;
  }
}
// Back to global order: rotated (n, t1, t2) -> (x, y, z)
template <int DIR, typename T>
__device__ __forceinline__ void Unrot(T n, T t1, T t2, T &gx, T &gy, T &gz)
{
  if constexpr (DIR == 0) {
    gx = n, gy = t1, gz = t2;
  } else if constexpr (DIR == 1) {
    gy = n, gz = t1, gx = t2;
  } else {
    gz = n, gx = t1, gy = t2;
  }
}

// ============================================================================ stage 1
// Per-cell values staged for the PCM Riemann problems: the conserved state
// plus the velocities (divisions done once per cell) and the direction-free
// pressure used by the non-exact policies; with dual energy also the conserved
// GasEnergy.
// The low FP32 part of the FP64 pressure, carried by the mixed form (the only form
// with T = float) for its split-pressure faces: p + plo is the FP64 pressure to
// about 2^-48 relative.
template <typename T>
struct PressureLo {
};
template <>
struct PressureLo<float> {
  float plo;
};

template <typename T, bool DE = false>
struct PcmCell : PressureLo<T> {
  T d, mx, my, mz, E, vx, vy, vz, p;
};
template <typename T>
struct PcmCell<T, true> : PressureLo<T> {
  T d, mx, my, mz, E, vx, vy, vz, p, ge;
};

// Number of conserved fields a stage carries: 5, or 6 with GasEnergy.
template <bool DE>
constexpr int kNumCons = DE ? 6 : 5;

// W (the half-step primitives) of a form: density, three velocities, pressure (energy
// under PolicyExact), the pressure's low part under the split pressure, then the
// specific GasEnergy with dual energy.
template <class P>
constexpr int kWGasEnergy = SplitPressure<P> ? 6 : 5;
template <class P, bool DE>
constexpr int kWFields = kWGasEnergy<P> + (DE ? 1 : 0);

template <class P, int DIR, bool DE = false>
__device__ __forceinline__ RState<typename P::T> PcmSide(const PcmCell<typename P::T, DE> &c, typename P::T gamma)
{
  using T = typename P::T;
  RState<T> s;
  s.d = c.d, s.E = c.E;
  Rot<DIR>(c.mx, c.my, c.mz, s.mx, s.my, s.mz);
  Rot<DIR>(c.vx, c.vy, c.vz, s.vx, s.vy, s.vz);
  if constexpr (P::kExact) {
    // hllc_cuda.cu:107-113: the kinetic sum follows the rotated component order
    if constexpr (DE) {
      T E_kin = T(0.5) * s.d * (s.vx * s.vx + s.vy * s.vy + s.vz * s.vz);
      s.p     = PressureFromDE(s.E, s.E - E_kin, c.ge, gamma);
    } else {
      s.p = (s.E - T(0.5) * s.d * (s.vx * s.vx + s.vy * s.vy + s.vz * s.vz)) * (gamma - T(1.0));
    }
    s.p = Fmax(s.p, T(kTinyNumber));
  } else {
    s.p = c.p;
  }
  if constexpr (DE) s.dge = c.ge;
  s.inv_d = Rcp<P>(s.d);
  return s;
}

template <class P, bool DE = false>
__device__ __forceinline__ PcmCell<typename P::T, DE> MakePcmCell(double d, double mx, double my, double mz, double E,
                                                                  double ge, double gamma)
{
  using T = typename P::T;
  PcmCell<T, DE> c;
  if constexpr (P::kExact) {
    c.d = d, c.mx = mx, c.my = my, c.mz = mz, c.E = E;
    c.vx = mx / d, c.vy = my / d, c.vz = mz / d;
    c.p = 0;
  } else {
    // FP64 primitive recovery, then narrowing (a no-op for PolicyFast)
    double inv = 1.0 / d;
    double vx = mx * inv, vy = my * inv, vz = mz * inv;
    double p;
    if constexpr (DE) {
      double E_kin = 0.5 * d * SquareMag(vx, vy, vz);
      p            = PressureFromDE(E, E - E_kin, ge, gamma);
    } else {
      p = PressureFromCons(E, d, vx, vy, vz, gamma);
    }
    c.d = T(d), c.mx = T(mx), c.my = T(my), c.mz = T(mz), c.E = T(E);
    c.vx = T(vx), c.vy = T(vy), c.vz = T(vz), c.p = T(p);
    if constexpr (SplitPressure<P>) c.plo = float(p - double(c.p));
  }
  if constexpr (DE) c.ge = T(ge);
  return c;
}

template <class P, int DIR>
__device__ __forceinline__ void FluxGlobal(const Cons<typename P::T> &f, typename P::T out[5])
{
  out[0] = f.d;
  Unrot<DIR>(f.mx, f.my, f.mz, out[1], out[2], out[3]);
  out[4] = f.E;
}

// The Riemann solve (HLLC, or HLL under an HllPolicy) on one face; the flux in global
// component order into out[0..4] and, with dual energy, the GasEnergy flux into out[5].
template <class P, int DIR, bool DE>
__device__ __forceinline__ void FaceFlux(const RState<typename P::T> &l, const RState<typename P::T> &r,
                                         typename P::T gamma, typename P::T *out)
{
  using T = typename P::T;
  T fge   = T(0);
  Cons<T> f = RiemannFlux<P, DE>(l, r, gamma, &fge);
  FluxGlobal<P, DIR>(f, out);
  if constexpr (DE) out[kGasEnergy] = fge;
}

// The flux through the face between stage-1 cells a (left) and b (right).  Under the
// split pressure the face reference is a.p (the high part of a's pressure): the left
// state's momentum pressure is a.plo and the right state's (b.p - a.p) + b.plo, and
// the normal-momentum flux comes back less a.p (the update adds the references back).
template <class P, int DIR, bool DE>
__device__ __forceinline__ void PcmFaceFlux(const PcmCell<typename P::T, DE> &a, const PcmCell<typename P::T, DE> &b,
                                            typename P::T gamma, typename P::T *out)
{
  RState<typename P::T> l = PcmSide<P, DIR, DE>(a, gamma), r = PcmSide<P, DIR, DE>(b, gamma);
  if constexpr (SplitPressure<P>) {
    l.p0 = r.p0 = a.p;
    l.pd        = a.plo;
    r.pd        = (b.p - a.p) + b.plo;
  }
  FaceFlux<P, DIR, DE>(l, r, gamma, out);
}

// Dual energy: the DE pressure of a cell and its centred velocity-difference source
// 0.5 P (dtodx (vx_imo - vx_ipo) + dtody (vy_jmo - vy_jpo) + dtodz (vz_kmo - vz_kpo)),
// as Update_Conserved_Variables_3D_half (VL_3D_cuda.cu:409-441, 471-478) and
// Partial_Update_Advected_Internal_Energy_3D (hydro_cuda.cu:779-810) form it, in FP64.
// u: the cell's conserved values; d*: the neighbours' velocity differences
// (vx_imo - vx_ipo, vy_jmo - vy_jpo, vz_kmo - vz_kpo).
__device__ __forceinline__ double DePressure(const double u[6], double gamma)
{
  double d_inv = 1.0 / u[0];
  double vx = u[1] * d_inv, vy = u[2] * d_inv, vz = u[3] * d_inv;
  double E_kin = 0.5 * u[0] * SquareMag(vx, vy, vz);
  return Fmax(PressureFromDE(u[4], u[4] - E_kin, u[5], gamma), kTinyNumber);
}
__device__ __forceinline__ double PdvTerm(double P, double dtodx, double dtody, double dtodz, double dvx, double dvy,
                                          double dvz)
{
  return 0.5 * P * (dtodx * dvx + dtody * dvy + dtodz * dvz);
}

// LDS of one stage-1 tile.
template <class P, int TX, int TY, bool DE = false>
struct Stage1Smem {
  using T = typename P::T;
  PcmCell<T, DE> plane[TY + 2][TX + 2];  // staged plane with a 1-cell halo
  T fx[kNumCons<DE>][TY][TX + 1];
  T fy[kNumCons<DE>][TY + 1][TX];
};

// The stage-1 halo prefetch (Stage1Tile) and the waves per SIMD the stage-1 kernels are
// compiled for.  With HLLC and no DE: the mixed form fits 4 waves (127 VGPRs) with the
// prefetch; the exact form 3 (36 bytes of spills cost less than the third wave gives back).
// The fast FP64 form, DE and the other solvers keep the compiler's allocation and no
// prefetch: its registers push them into spills or a wave lower.
template <class P, bool DE>
constexpr bool kStage1Prefetch = P::kId != PolicyFast::kId && !DE && P::kSolver == kRiemannHllc;
template <class P, bool DE>
constexpr int kStage1MinWaves = !kStage1Prefetch<P, DE> ? 1 : (P::kId == PolicyMixed::kId ? 4 : 3);

// Stage 1 of one tile (bx, by) x z-chunk bz of the half-step region [lo, hi) per
// axis, lo = ng - R1, hi = n - ng + R1.  Called by every lane of a TX*TY workgroup.
// With dual energy (DE) it also advances GasEnergy with its pdV source, stores the
// half-step specific GasEnergy in W and the full-step pdV source of every real cell
// in a.pdv (Partial_Update_Advected_Internal_Energy_3D reads the state at the start
// of the step, which only this stage still sees).
template <class P, int TX, int TY, bool DE = false>
__device__ __forceinline__ void Stage1Tile(const StepArgs &a, int R1, int bx, int by, int bz,
                                           Stage1Smem<P, TX, TY, DE> &sm)
{
  using T            = typename P::T;
  constexpr int NT   = TX * TY;
  constexpr int NFX  = (TX + 1) * TY;
  constexpr int NFY  = TX * (TY + 1);
  constexpr int NC   = kNumCons<DE>;
  auto &plane        = sm.plane;
  auto &fx           = sm.fx;
  auto &fy           = sm.fy;

  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t nc = size_t(nx) * ny * nz;
  const int lo_x = a.ng - R1, hi_x = nx - a.ng + R1;
  const int lo_y = a.ng - R1, hi_y = ny - a.ng + R1;
  const int lo_z = a.ng - R1, hi_z = nz - a.ng + R1;
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int i0 = lo_x + bx * TX, j0 = lo_y + by * TY;
  const int i = i0 + tx, j = j0 + ty;
  const int k0 = lo_z + bz * a.kchunk;
  const int k1 = min(k0 + a.kchunk, hi_z);
  const bool own = (i < hi_x) && (j < hi_y);
  const T gamma  = T(a.gamma);
  const double dt_full = StepDt(a);
  const double dt    = 0.5 * dt_full;
  const double dtodx = dt / a.dx, dtody = dt / a.dy, dtodz = dt / a.dz;
  const double *U    = a.U;
  T *W               = reinterpret_cast<T *>(a.W);

  if (dt == 0.0) return;  // a device-timed step past the output time (uniform per box)

  // Lanes past the region's end still stage their own cell: it is the right
  // neighbour of the last owned cell.  Only the index is clamped, to stay in memory.
  const int ic = min(i, nx - 1), jc = min(j, ny - 1);
  auto load_cell = [&](int ii, int jj, int kk) {
    size_t id = Lin(ii, jj, kk, nx, ny);
    return MakePcmCell<P, DE>(U[id], U[nc + id], U[2 * nc + id], U[3 * nc + id], U[4 * nc + id],
                              DE ? U[kGasEnergy * nc + id] : 0.0, a.gamma);
  };

  PcmCell<T, DE> c_lo  = load_cell(ic, jc, k0 - 1);  // cell k-1
  PcmCell<T, DE> c_mid = load_cell(ic, jc, k0);      // cell k
  double u_mid[NC];
  T fz_lo[NC];
  PcmFaceFlux<P, 2, DE>(c_lo, c_mid, gamma, fz_lo);

  // The plane's halo, x sides (TY rows x 2) and y sides (TX cols x 2), no corners: one cell
  // per lane of the first 2 (TX + TY).  Under kStage1Prefetch its conserved values for the
  // next plane are loaded while the current plane's faces are solved, and converted once the
  // next plane starts (KH 512^3: mixed 3.94 -> 3.56 ms, exact 10.71 -> 10.28 ms); otherwise
  // they are loaded after the first barrier.
  static_assert(2 * TY + 2 * TX <= NT, "one halo cell per lane");
  constexpr bool kPrefetch = kStage1Prefetch<P, DE>;
  const bool has_h = tid < 2 * TY + 2 * TX;
  int hpi = 0, hpj = 0;
  if (tid < 2 * TY) {
    hpj = 1 + (tid >> 1);
    hpi = (tid & 1) ? TX + 1 : 0;
  } else {
    const int g = tid - 2 * TY;
    hpi         = 1 + (g >> 1);
    hpj         = (g & 1) ? TY + 1 : 0;
  }
  const int hgi = min(i0 - 1 + hpi, nx - 1), hgj = min(j0 - 1 + hpj, ny - 1);
  double h_raw[NC];
  auto load_halo = [&](int kk) {
    const size_t id = Lin(hgi, hgj, kk, nx, ny);
#pragma unroll
    for (int f = 0; f < NC; f++) h_raw[f] = U[size_t(f) * nc + id];  // GasEnergy (DE) is field 5
  };
  if (kPrefetch && has_h) load_halo(k0);

  for (int k = k0; k < k1; k++) {
    {
      size_t id = Lin(ic, jc, k, nx, ny);
#pragma unroll
      for (int f = 0; f < NC; f++) u_mid[f] = U[f * nc + id];
    }
    PcmCell<T, DE> c_hi = load_cell(ic, jc, k + 1);
    __syncthreads();  // previous plane fully consumed
    plane[ty + 1][tx + 1] = c_mid;
    if (has_h) {
      if constexpr (!kPrefetch) load_halo(k);
      plane[hpj][hpi] = MakePcmCell<P, DE>(h_raw[0], h_raw[1], h_raw[2], h_raw[3], h_raw[4],
                                           DE ? h_raw[NC - 1] : 0.0, a.gamma);
      if (kPrefetch && k + 1 < k1) load_halo(k + 1);
    }
    __syncthreads();
    // x faces: face fi sits between plane cells (fi, fj+1) and (fi+1, fj+1)
    for (int it = tid; it < NFX; it += NT) {
      int fj = it / (TX + 1), fi = it - fj * (TX + 1);
      T g[NC];
      PcmFaceFlux<P, 0, DE>(plane[fj + 1][fi], plane[fj + 1][fi + 1], gamma, g);
#pragma unroll
      for (int q = 0; q < NC; q++) fx[q][fj][fi] = g[q];
    }
    for (int it = tid; it < NFY; it += NT) {
      int fj = it / TX, fi = it - fj * TX;
      T g[NC];
      PcmFaceFlux<P, 1, DE>(plane[fj][fi + 1], plane[fj + 1][fi + 1], gamma, g);
#pragma unroll
      for (int q = 0; q < NC; q++) fy[q][fj][fi] = g[q];
    }
    T fz_hi[NC];
    PcmFaceFlux<P, 2, DE>(c_mid, c_hi, gamma, fz_hi);
    __syncthreads();
    if (own) {
      double uh[NC];
#pragma unroll
      for (int q = 0; q < NC; q++) {
        // (((U + dtodx dFx) + dtody dFy) + dtodz dFz), VL_3D_cuda.cu:444-461
        uh[q] = u_mid[q] + dtodx * (double(fx[q][ty][tx]) - double(fx[q][ty][tx + 1])) +
                dtody * (double(fy[q][ty][tx]) - double(fy[q][ty + 1][tx])) +
                dtodz * (double(fz_lo[q]) - double(fz_hi[q]));
      }
      if constexpr (SplitPressure<P>) {
        // the face references (left cells' pressures) the normal-momentum fluxes left out
        uh[1] += dtodx * (double(plane[ty + 1][tx].p) - double(plane[ty + 1][tx + 1].p));
        uh[2] += dtody * (double(plane[ty][tx + 1].p) - double(plane[ty + 1][tx + 1].p));
        uh[3] += dtodz * (double(c_lo.p) - double(c_mid.p));
      }
      if constexpr (DE) {
        // velocity differences of the neighbours at the start of the step, each a
        // division as in VL_3D_cuda.cu:434-439 (FP64; the mixed form's staged
        // velocities are FP32 and are not used for this source)
        double dvx, dvy, dvz;
        if constexpr (P::kExact) {
          dvx = plane[ty + 1][tx].vx - plane[ty + 1][tx + 2].vx;
          dvy = plane[ty][tx + 1].vy - plane[ty + 2][tx + 1].vy;
          dvz = c_lo.vz - c_hi.vz;
        } else {
          const size_t id = Lin(i, j, k, nx, ny), sy = size_t(nx), sz = size_t(nx) * ny;
          dvx = U[nc + id - 1] / U[id - 1] - U[nc + id + 1] / U[id + 1];
          dvy = U[2 * nc + id - sy] / U[id - sy] - U[2 * nc + id + sy] / U[id + sy];
          dvz = U[3 * nc + id - sz] / U[id - sz] - U[3 * nc + id + sz] / U[id + sz];
        }
        const double Pde = DePressure(u_mid, a.gamma);
        uh[kGasEnergy] += PdvTerm(Pde, dtodx, dtody, dtodz, dvx, dvy, dvz);
        if (i >= a.ng && i < nx - a.ng && j >= a.ng && j < ny - a.ng && k >= a.ng && k < nz - a.ng) {
          const double fx_ = dt_full / a.dx, fy_ = dt_full / a.dy, fz_ = dt_full / a.dz;
          a.pdv[Lin(i, j, k, nx, ny)] = PdvTerm(Pde, fx_, fy_, fz_, dvx, dvy, dvz);
        }
      }
      if ((a.flags & kFlagDensityFloor) && uh[0] < a.density_floor) {  // VL_3D_cuda.cu:479-493
        double d0 = uh[0];
        uh[0]     = a.density_floor;
#pragma unroll
        for (int q = 1; q < NC; q++) uh[q] *= (a.density_floor / d0);
      }
      size_t id = Lin(i, j, k, nx, ny);
      // the pressure's low part (split pressure): p - float(p), in W's field 5
      auto store_p = [&](double p) {
        W[4 * nc + id] = T(p);
        if constexpr (SplitPressure<P>) W[5 * nc + id] = T(p - double(T(p)));
      };
      if constexpr (DE) {
        // Load_Data under DE (reconstruction.h:179-191): the exact form keeps E and
        // GasEnergy and recovers p and ge per sweep; the others store p and ge
        if constexpr (P::kExact) {
          // PCM hands the conserved half state itself to the Riemann solver (pcm_cuda.cu)
          const bool cons_w = (a.flags & kFlagConservedW) != 0;
          W[id]          = T(uh[0]);
          W[nc + id]     = cons_w ? T(uh[1]) : T(uh[1] / uh[0]);
          W[2 * nc + id] = cons_w ? T(uh[2]) : T(uh[2] / uh[0]);
          W[3 * nc + id] = cons_w ? T(uh[3]) : T(uh[3] / uh[0]);
          W[4 * nc + id] = T(uh[4]);
          W[5 * nc + id] = T(uh[5]);
        } else {
          double vx = uh[1] / uh[0], vy = uh[2] / uh[0], vz = uh[3] / uh[0];
          double E_kin   = 0.5 * uh[0] * SquareMag(vx, vy, vz);
          W[id]          = T(uh[0]);
          W[nc + id]     = T(vx);
          W[2 * nc + id] = T(vy);
          W[3 * nc + id] = T(vz);
          if (a.flags & kFlagConservedW) {
            // PCM hands the conserved state to the solver (pcm_cuda.cu): the total-energy
            // pressure re-forms the cell's own E, and the face state's DE switch then picks the
            // pressure Get_Pressure_From_DE would
            store_p((uh[4] - E_kin) * (a.gamma - 1.0));
          } else {
            store_p(PressureFromDE(uh[4], uh[4] - E_kin, uh[5], a.gamma));
          }
          W[kWGasEnergy<P> * nc + id] = T(uh[5] / uh[0]);
        }
      } else {
        Prim<double> w = LoadPrim(uh[0], uh[1], uh[2], uh[3], uh[4], a.gamma);
        // PCM hands the conserved half state itself to the Riemann solver (pcm_cuda.cu)
        const bool cons_w = P::kExact && (a.flags & kFlagConservedW);
        W[id]          = T(w.d);
        W[nc + id]     = cons_w ? T(uh[1]) : T(w.vx);
        W[2 * nc + id] = cons_w ? T(uh[2]) : T(w.vy);
        W[3 * nc + id] = cons_w ? T(uh[3]) : T(w.vz);
        if constexpr (P::kExact) {
          W[4 * nc + id] = T(uh[4]);  // see WCell
        } else {
          store_p(w.p);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < NC; q++) fz_lo[q] = fz_hi[q];
    c_lo  = c_mid;
    c_mid = c_hi;
  }
}


template <class P, int TX, int TY, bool DE = false>
__global__ void __launch_bounds__(TX *TY, (kStage1MinWaves<P, DE>)) Stage1Kernel(StepArgs a, int R1)
{
  __shared__ Stage1Smem<P, TX, TY, DE> sm;
  // the inverse-dt accumulator is reset here; stage 2 (next launch) fills it
  if ((a.flags & (kFlagReduceDt | kFlagResetDti)) && blockIdx.x == 0 && threadIdx.x == 0) *a.dti = 0ull;
  int bx, by, bz;
  if (!TileOf(a, bx, by, bz)) return;
  Stage1Tile<P, TX, TY, DE>(a, R1, bx, by, bz, sm);
}

// ============================================================================ stage 2
// Half-step cell as stored in W.  The fifth slot holds the pressure, except under
// PolicyExact where it holds the energy: Load_Data recovers pressure from the
// velocities in the sweep's rotated order (reconstruction.h:194-201 with
// SquareMagnitude, math_utilities.h:78-97), so its last bit depends on the sweep
// direction and is recomputed per sweep.  With dual energy the sixth slot holds the
// specific GasEnergy, or under PolicyExact the conserved GasEnergy (Load_Data
// divides it by the density, reconstruction.h:191).
// The mixed form also stores the pressure's low part (PressureLo) as field 5, which
// moves the GasEnergy slot to field 6.
template <typename T, bool DE = false>
struct WCell : PressureLo<T> {
  T d, vx, vy, vz, p_or_E;
};
template <typename T>
struct WCell<T, true> : PressureLo<T> {
  T d, vx, vy, vz, p_or_E, g;
};

// The split-pressure form of ReconDir: the stencil as deviations from its centre cell c
// (the pressure's from the FP32 pairs, so they resolve differences far below the
// pressure's own FP32 rounding), the reconstruction's linear core on those with c's
// eigenvectors, then c added back.  The interface pressures stay relative to c's high
// part, which each interface carries as ph; the pressure floor is applied where the
// face states are formed (FaceStates).
template <class P, int RK, int DIR, bool DE = false>
__device__ __forceinline__ void ReconDirSplit(const WCell<float, DE> *s, float gamma, bool pos_fallback,
                                              IPrim<float, true> &L, IPrim<float, true> &R, float *geL, float *geR)
{
  using T          = float;
  constexpr int r  = ReconTraits<RK>::kRadius;
  constexpr int NS = 2 * r + 1;
  const WCell<T, DE> &c = s[r];
  Prim<T> qc;  // the centre's absolute state (rotated)
  qc.d = c.d;
  Rot<DIR>(c.vx, c.vy, c.vz, qc.vx, qc.vy, qc.vz);
  qc.p = c.p_or_E;
  Prim<T> st[NS];
  T ge[NS];
#pragma unroll
  for (int m = 0; m < NS; m++) {
    st[m].d = s[m].d - c.d;
    Rot<DIR>(s[m].vx - c.vx, s[m].vy - c.vy, s[m].vz - c.vz, st[m].vx, st[m].vy, st[m].vz);
    st[m].p = (s[m].p_or_E - c.p_or_E) + (s[m].plo - c.plo);
    if constexpr (DE) ge[m] = s[m].g;
  }
  Prim<T> dl{}, dr{};
  if constexpr (RK == kPLMP) {
    ReconPlmpCore<P>(st, dl, dr);
  } else if constexpr (RK == kPPMP) {
    ReconPpmpCore<P>(st, dl, dr);
  } else if constexpr (RK == kPPMC) {
    ReconPpmcCore<P>(st, qc, gamma, dl, dr);
  } else if constexpr (RK == kPLMC) {
    ReconPlmcCore<P>(st, qc, gamma, dl, dr);
  }
  if constexpr (DE) {
    if constexpr (RK == kPPMC) {
      ReconPpmcScalar<P>(ge, *geL, *geR);
    } else if constexpr (RK == kPLMC) {
      ReconPlmcScalar<P>(ge, *geL, *geR);
    } else if constexpr (RK == kPLMP) {
      ReconPlmpScalar<P>(ge, *geL, *geR);
    } else if constexpr (RK == kPPMP) {
      ReconPpmpScalar<P>(ge, *geL, *geR);
    } else if constexpr (RK == kPCM) {
      *geL = *geR = ge[0];
    }
  }
  // PPMC's positivity fallback (see ReconPpmc): the cell-centred values on both sides
  if (RK == kPPMC && pos_fallback &&
      !(qc.d + dl.d > T(0) && qc.d + dr.d > T(0) && qc.p + dl.p > T(0) && qc.p + dr.p > T(0))) {
    dl = dr = Prim<T>{};
    if constexpr (DE) *geL = *geR = ge[r];
  }
  L.d = Fmax(qc.d + dl.d, T(kTinyNumber)), R.d = Fmax(qc.d + dr.d, T(kTinyNumber));
  L.vx = qc.vx + dl.vx, L.vy = qc.vy + dl.vy, L.vz = qc.vz + dl.vz;
  R.vx = qc.vx + dr.vx, R.vy = qc.vy + dr.vy, R.vz = qc.vz + dr.vz;
  L.p = dl.p + c.plo, R.p = dr.p + c.plo;
  L.ph = R.ph = c.p_or_E;
}

template <class P, int RK, int DIR, bool DE = false>
__device__ __forceinline__ void ReconDirPlain(const WCell<typename P::T, DE> *s, typename P::T gamma,
                                              bool pos_fallback, Prim<typename P::T> &L, Prim<typename P::T> &R,
                                              typename P::T *geL, typename P::T *geR)
{
  using T          = typename P::T;
  constexpr int NS = 2 * ReconTraits<RK>::kRadius + 1;
  Prim<T> st[NS];
  T ge[NS];
#pragma unroll
  for (int m = 0; m < NS; m++) {
    st[m].d = s[m].d;
    Rot<DIR>(s[m].vx, s[m].vy, s[m].vz, st[m].vx, st[m].vy, st[m].vz);
    if constexpr (RK == kPCM) {
      st[m].p = s[m].p_or_E;  // exact form: (d, momenta, E) carried through as they are
      if constexpr (DE) ge[m] = s[m].g;  // exact form: the conserved GasEnergy, else the specific one
    } else if constexpr (P::kExact && (RK == kPLMP || RK == kPPMP)) {
      // plmp_cuda.cu:113-120 (ppmp_cuda.cu:160-171): the squared velocities summed left to right;
      // under DE the pressure from Get_Pressure_From_DE with that kinetic energy
      const T v2 = (st[m].vx * st[m].vx + st[m].vy * st[m].vy) + st[m].vz * st[m].vz;
      if constexpr (DE) {
        const T E_kin = T(0.5) * st[m].d * v2;
        st[m].p       = Fmax(PressureFromDE(s[m].p_or_E, s[m].p_or_E - E_kin, s[m].g, gamma), T(kTinyNumber));
        ge[m]         = s[m].g / st[m].d;
      } else {
        st[m].p = Fmax((s[m].p_or_E - T(0.5) * st[m].d * v2) * (gamma - T(1.0)), T(kTinyNumber));
      }
    } else if constexpr (P::kExact) {
      if constexpr (DE) {
        T E_kin  = T(0.5) * st[m].d * SquareMag(st[m].vx, st[m].vy, st[m].vz);
        st[m].p  = PressureFromDE(s[m].p_or_E, s[m].p_or_E - E_kin, s[m].g, gamma);
        ge[m]    = s[m].g / st[m].d;
      } else {
        st[m].p = PressureFromCons(s[m].p_or_E, st[m].d, st[m].vx, st[m].vy, st[m].vz, gamma);
      }
    } else {
      st[m].p = s[m].p_or_E;
      if constexpr (DE) ge[m] = s[m].g;
    }
  }
  if constexpr (RK == kPCM) {
    L = R = st[0];
    if constexpr (DE) *geL = *geR = ge[0];
  } else if constexpr (RK == kPLMP) {
    ReconPlmp<P>(st, L, R);
    if constexpr (DE) ReconPlmpScalar<P>(ge, *geL, *geR);
  } else if constexpr (RK == kPPMP) {
    ReconPpmp<P>(st, L, R);
    if constexpr (DE) ReconPpmpScalar<P>(ge, *geL, *geR);
  } else if constexpr (RK == kPPMC) {
    bool fb = ReconPpmc<P>(st, gamma, L, R, pos_fallback);
    if constexpr (DE) {
      ReconPpmcScalar<P>(ge, *geL, *geR);
      if (fb) *geL = *geR = ge[2];
    }
  } else {
    ReconPlmc<P>(st, gamma, L, R);
    if constexpr (DE) ReconPlmcScalar<P>(ge, *geL, *geR);
  }
}

// The interface values of the cell at the centre of the stencil s[0..2r] (its left
// value R at i-1/2 and right value L at i+1/2), rotated into the sweep frame of DIR.
template <class P, int RK, int DIR, bool DE = false>
__device__ __forceinline__ void ReconDir(const WCell<typename P::T, DE> *s, typename P::T gamma, bool pos_fallback,
                                         IPrimOf<P> &L, IPrimOf<P> &R, typename P::T *geL = nullptr,
                                         typename P::T *geR = nullptr)
{
  if constexpr (SplitPressure<P>) {
    ReconDirSplit<P, RK, DIR, DE>(s, gamma, pos_fallback, L, R, geL, geR);
  } else {
    ReconDirPlain<P, RK, DIR, DE>(s, gamma, pos_fallback, L, R, geL, geR);
  }
}

// The Riemann input of one interface value: the conserved variables the reference's
// reconstruction writes to dev_bounds, as the Riemann solver reads them back.  PCM and
// PLMP form them in their own order under the exact form (pcm_cuda.cu copies the state;
// plmp_cuda.cu:217-225 sums the squared velocities left to right).
template <class P, int RK, bool DE = false>
__device__ __forceinline__ RState<typename P::T> IfaceState(const Prim<typename P::T> &q, typename P::T gamma,
                                                            typename P::T inv_gm1,
                                                            typename P::T ge = typename P::T(0))
{
  using T = typename P::T;
  if constexpr (P::kExact && RK == kPCM) {
    // q holds (d, momenta, E) and ge the conserved GasEnergy
    return RStateFromCons<P, DE>(q.d, q.vx, q.vy, q.vz, q.p, gamma, DE ? ge : T(0));
  } else if constexpr (P::kExact && (RK == kPLMP || RK == kPPMP)) {  // ppmp_cuda.cu:645-650 likewise
    const T E = q.p / (gamma - T(1.0)) + T(0.5) * q.d * ((q.vx * q.vx + q.vy * q.vy) + q.vz * q.vz);
    return RStateFromCons<P, DE>(q.d, q.d * q.vx, q.d * q.vy, q.d * q.vz, E, gamma, DE ? q.d * ge : T(0));
  } else {
    return RStateFromPrim<P, DE>(q, gamma, inv_gm1, ge);
  }
}

// The two Riemann inputs of one face from its left and right interface values.  Under
// the split pressure the face reference p0 is the left cell's high pressure part; both
// sides' momentum pressures are expressed relative to it (held at >= -p0, so p0 + pd is
// never negative) and the absolute pressure, floored at TINY_NUMBER, feeds the rest.
template <class P, int RK, bool DE = false>
__device__ __forceinline__ void FaceStates(const IPrimOf<P> &L, const IPrimOf<P> &R, typename P::T gamma,
                                           typename P::T inv_gm1, typename P::T geL, typename P::T geR,
                                           RState<typename P::T> &l, RState<typename P::T> &r)
{
  using T = typename P::T;
  if constexpr (SplitPressure<P>) {
    const T p0 = L.ph;
    Prim<T> a  = L, b = R;
    a.p        = Fmax(L.ph + L.p, T(kTinyNumber));
    b.p        = Fmax(R.ph + R.p, T(kTinyNumber));
    l          = RStateFromPrim<P, DE>(a, gamma, inv_gm1, geL);
    r          = RStateFromPrim<P, DE>(b, gamma, inv_gm1, geR);
    l.p0 = r.p0 = p0;
    l.pd        = Fmax(L.p, -p0);
    r.pd        = Fmax(R.p + (R.ph - p0), -p0);
    if constexpr (DE) {  // the GasEnergy pressure replaced the interface pressure
      if (l.p != a.p) l.pd = l.p - p0;
      if (r.p != b.p) r.pd = r.p - p0;
    }
  } else {
    l = IfaceState<P, RK, DE>(L, gamma, inv_gm1, geL);
    r = IfaceState<P, RK, DE>(R, gamma, inv_gm1, geR);
  }
}

// FaceFlux of the face between two reconstructed interface values.
template <class P, int RK, int DIR, bool DE>
__device__ __forceinline__ void IfaceFlux(const IPrimOf<P> &L, const IPrimOf<P> &R, typename P::T gamma,
                                          typename P::T inv_gm1, typename P::T geL, typename P::T geR,
                                          typename P::T *out)
{
  RState<typename P::T> l, r;
  FaceStates<P, RK, DE>(L, R, gamma, inv_gm1, geL, geR, l, r);
  FaceFlux<P, DIR, DE>(l, r, gamma, out);
}

// LDS footprint of stage 2 (bytes) for a given tile; used by the host to pick tiles.
template <class P, int RK, int TX, int TY>
struct Stage2Shape {
  using T                 = typename P::T;
  static constexpr int H  = ReconTraits<RK>::kRadius + 1;  // plane halo
  static constexpr int PX = TX + 2 * H, PY = TY + 2 * H;
  static constexpr int NQX = (TX + 2) * TY, NQY = TX * (TY + 2);
  static constexpr int NQ  = NQX > NQY ? NQX : NQY;
  static constexpr int NFX = (TX + 1) * TY, NFY = TX * (TY + 1);
  static constexpr int NF  = NFX > NFY ? NFX : NFY;
};

// Minimum waves per SIMD requested from the register allocator.  Left to itself it
// gives the FP64 forms a 256-VGPR budget (one wave per SIMD), which leaves the long
// FP64 divide/sqrt chains and LDS round trips without latency cover.
template <class P>
struct Stage2Occupancy {
  static constexpr int kMinWaves = P::kId == PolicyMixed::kId ? 3 : 2;
};

// Interface values of the passive GasEnergy (dual energy only).
template <typename T, int N, bool DE>
struct GeIfaces {
  __device__ T *L() { return nullptr; }
  __device__ T *R() { return nullptr; }
};
template <typename T, int N>
struct GeIfaces<T, N, true> {
  T gL[N], gR[N];
  __device__ T *L() { return gL; }
  __device__ T *R() { return gR; }
};

// Whether a stage-2 tile runs its x and y work as one list (both directions'
// reconstructions in one pass, both directions' fluxes in the next): the lists then
// fill whole waves (a 16x16 tile has 288 + 288 reconstructions, 9 full wave64 passes
// instead of 2 x 5 with a half-empty wave each) and the plane takes three barriers
// instead of five.  It holds both directions' interface values and fluxes in LDS at
// once, so it is taken only where the tiles that keep the kernel at its minimum
// occupancy still fit in a CU's 160 KiB.
template <class P, int RK, int TX, int TY, bool DE = false>
struct Stage2Merge {
  using T                    = typename P::T;
  using S                    = Stage2Shape<P, RK, TX, TY>;
  static constexpr size_t kBytes = sizeof(WCell<T, DE>) * S::PY * S::PX +
                                   size_t(S::NQX + S::NQY) * (2 * sizeof(IPrimOf<P>) + (DE ? 2 * sizeof(T) : 0)) +
                                   size_t(kNumCons<DE>) * (S::NFX + S::NFY) * sizeof(T) + 8 * (TX * TY / 64);
  static constexpr int kWaves  = TX * TY / 64;
  static constexpr int kGroups = (4 * Stage2Occupancy<P>::kMinWaves + kWaves - 1) / kWaves;
  static constexpr bool value  = kBytes * kGroups <= 160 * 1024;
};

// LDS of one stage-2 tile.  Merged lists keep the x items first, then the y items.
template <class P, int RK, int TX, int TY, bool DE = false>
struct Stage2Smem {
  using T = typename P::T;
  using S = Stage2Shape<P, RK, TX, TY>;
  static constexpr bool kMerge = Stage2Merge<P, RK, TX, TY, DE>::value;
  static constexpr int NQS     = kMerge ? S::NQX + S::NQY : S::NQ;
  static constexpr int NFS     = kMerge ? S::NFX + S::NFY : S::NF;
  WCell<T, DE> plane[S::PY][S::PX];
  IPrimOf<P> qL[NQS];
  IPrimOf<P> qR[NQS];
  GeIfaces<T, NQS, DE> qg;
  T fl[kNumCons<DE>][NFS];
  double red[TX * TY / 64];
};

// The y sweep's view of a cell: the velocity triple rotated into (n, t1, t2) = (vy, vz,
// vx) where the x sweep's is (vx, vy, vz), so one reconstruction and one Riemann code
// path serve both directions of a merged list (runtime selects, no divergence).
template <typename T, bool DE>
__device__ __forceinline__ void RotateForY(WCell<T, DE> &c, bool y)
{
  const T vx = c.vx, vy = c.vy, vz = c.vz;
  c.vx = y ? vy : vx;
  c.vy = y ? vz : vy;
  c.vz = y ? vx : vz;
}

// ---------------------------------------------------------------------------- repair marks
// The crash test of hydro_cuda.cu:375-376: negative or NaN density or energy.
__device__ __forceinline__ bool CrashedCell(double d, double E) { return d < 0.0 || d != d || E < 0.0 || E != E; }

__device__ __forceinline__ void MarkRepair(const RepairState &r, size_t id, int kind)
{
  if (!r.mark) return;  // a launch without repair scratch leaves the cell as the update left it
  r.mark[id]                          = (unsigned char)kind;
  r.chunk[id >> kRepairChunkShift]    = 1;
  atomicAdd(r.count, 1u);
}

// Static gravity adds its source after the repair (hydro_cuda.cu:301-315) from the cell's
// state at the start of the step, which the marked cell leaves here.
__device__ __forceinline__ void RecordRepairGrav(const RepairState &r, size_t id, const double s[4])
{
  if (!r.mark) return;
  unsigned slot = atomicAdd(r.count + 1, 1u);
  if (int(slot) >= r.grav_cap) {
    atomicAdd(r.count + 2, 1u);
    return;
  }
  double *g = r.grav + 5 * size_t(slot);
  g[0] = double(id), g[1] = s[0], g[2] = s[1], g[3] = s[2], g[4] = s[3];
}

// Side buffer of the fused periodic fill: face f (x-lo, x-hi, y-lo, y-hi, z-lo, z-hi) holds
// the densities of the ghost layer next to the real cells on that face as they were at the
// start of the step, before stage 2 overwrote the ghosts with the new periodic images
// (RepairSideKernel copies them).  Lx, Ly, Lz: real cells per axis.
__device__ __forceinline__ size_t RepairSideIndex(int face, int a, int b, int Lx, int Ly, int Lz)
{
  const size_t fx = size_t(Ly) * Lz, fy = size_t(Lx) * Lz, fz = size_t(Lx) * Ly;
  const size_t base[6] = {0, fx, 2 * fx, 2 * fx + fy, 2 * fx + 2 * fy, 2 * fx + 2 * fy + fz};
  const size_t row     = face < 2 ? size_t(Ly) : size_t(Lx);
  return base[face] + size_t(b) * row + a;  // (a, b) = (j, k) on x faces, (i, k) on y, (i, j) on z
}

// Every periodic image of real cell (i, j, k) inside the ghost layers gets the
// cell's values (the host enables this only when every axis holds >= ng real cells,
// so each image lands in a ghost layer and maps back to this cell).  Per axis: the
// cell itself, then up to two images (+L, -L); selects instead of small arrays keep
// the coordinates in registers.
template <int NF>
__device__ __forceinline__ void WritePeriodicImages(double *U, size_t nc, int nx, int ny, int nz, int ng, int i, int j,
                                                    int k, const double *un)
{
  const int Lx = nx - 2 * ng, Ly = ny - 2 * ng, Lz = nz - 2 * ng;
  const bool xp = i < 2 * ng, xm = i >= nx - 2 * ng;
  const bool yp = j < 2 * ng, ym = j >= ny - 2 * ng;
  const bool zp = k < 2 * ng, zm = k >= nz - 2 * ng;
  const int x1 = xp ? i + Lx : i - Lx, y1 = yp ? j + Ly : j - Ly, z1 = zp ? k + Lz : k - Lz;
  const int x2 = i - Lx, y2 = j - Ly, z2 = k - Lz;  // second image only when both apply
  const int nxi = 1 + xp + xm, nyi = 1 + yp + ym, nzi = 1 + zp + zm;
  for (int c = 1; c < nxi * nyi * nzi; c++) {
    int ax = c % nxi, ay = (c / nxi) % nyi, az = c / (nxi * nyi);
    const int gx = ax == 0 ? i : (ax == 1 ? x1 : x2);
    const int gy = ay == 0 ? j : (ay == 1 ? y1 : y2);
    const int gz = az == 0 ? k : (az == 1 ? z1 : z2);
    size_t g = Lin(gx, gy, gz, nx, ny);
#pragma unroll
    for (int q = 0; q < NF; q++) U[q * nc + g] = un[q];
  }
}

// Stage 2 of one tile (bx, by) x z-chunk bz of the real cells.  Called by every
// lane of a TX*TY workgroup; the block's inverse-dt maximum goes to *a.dti.
// With dual energy (DE) it also updates GasEnergy (its flux differences on top of
// the pdV source stage 1 stored) and leaves the temperature floor, the inverse-dt
// reduction and the ghost images to the internal-energy selection that follows
// (DeSelectKernel / DeSyncKernel).
template <class P, int RK, int TX, int TY, bool DE = false>
__device__ __forceinline__ void Stage2Tile(const StepArgs &a, int bx, int by, int bz,
                                           Stage2Smem<P, RK, TX, TY, DE> &sm)
{
  using T          = typename P::T;
  using S          = Stage2Shape<P, RK, TX, TY>;
  using SM         = Stage2Smem<P, RK, TX, TY, DE>;
  constexpr int NT = TX * TY;
  constexpr int r  = ReconTraits<RK>::kRadius;
  constexpr int NS = 2 * r + 1;
  constexpr int H  = S::H;
  constexpr int NC = kNumCons<DE>;
  auto &plane      = sm.plane;
  auto &qL         = sm.qL;
  auto &qR         = sm.qR;
  T *qgL           = sm.qg.L();
  T *qgR           = sm.qg.R();
  auto &fl         = sm.fl;
  auto &red        = sm.red;

  const int nx = a.nx, ny = a.ny, nz = a.nz, ng = a.ng;
  const size_t nc = size_t(nx) * ny * nz;
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int i0 = ng + bx * TX, j0 = ng + by * TY;
  const int i = i0 + tx, j = j0 + ty;
  const int k0 = ng + bz * a.kchunk;
  const int k1 = min(k0 + a.kchunk, nz - ng);
  const bool own = (i < nx - ng) && (j < ny - ng);
  // Lanes past the last real cell still stage their own column of W: those
  // cells are the stencil of the last owned cells.  Only the index is clamped.
  const int ic = min(i, nx - 1), jc = min(j, ny - 1);
  const T gamma      = T(a.gamma);
  const T inv_gm1    = Rcp<P>(gamma - T(1.));
  const double dt    = StepDt(a);
  const double dtodx = dt / a.dx, dtody = dt / a.dy, dtodz = dt / a.dz;
  const T *W         = reinterpret_cast<const T *>(a.W);
  double *U          = a.U;
  const double inv_dx = 1.0 / a.dx, inv_dy = 1.0 / a.dy, inv_dz = 1.0 / a.dz;
  if (dt == 0.0) return;  // a device-timed step past the output time (uniform per box)
  const bool pf = (a.flags & kFlagPositivityFallback) != 0;

  auto loadW = [&](int ii, int jj, int kk) {
    size_t id = Lin(ii, jj, kk, nx, ny);
    WCell<T, DE> c;
    c.d = W[id], c.vx = W[nc + id], c.vy = W[2 * nc + id], c.vz = W[3 * nc + id], c.p_or_E = W[4 * nc + id];
    if constexpr (SplitPressure<P>) c.plo = W[5 * nc + id];
    if constexpr (DE) c.g = W[kWGasEnergy<P> * nc + id];
    return c;
  };

  // z window of this lane's column: win[m] = W(c - r + m) for the cell c whose z
  // interfaces are reconstructed next.  Prologue: cells k0-1 and k0 give the face k0-1/2.
  WCell<T, DE> win[NS];
#pragma unroll
  for (int m = 0; m < NS; m++) win[m] = loadW(ic, jc, k0 - 1 - r + m);
  IPrimOf<P> qLz, qRz;
  T gLz = T(0), gRz = T(0);
  ReconDir<P, RK, 2, DE>(win, gamma, pf, qLz, qRz, &gLz, &gRz);  // cell k0-1
  IPrimOf<P> qL_lo = qLz;                                    // left state of face k0-1/2
  T gL_lo          = gLz;
#pragma unroll
  for (int m = 0; m < NS - 1; m++) win[m] = win[m + 1];
  win[NS - 1] = loadW(ic, jc, k0 + r);
  ReconDir<P, RK, 2, DE>(win, gamma, pf, qLz, qRz, &gLz, &gRz);  // cell k0
  T fz_lo[NC];                                               // flux through face k-1/2 of the current cell
  IfaceFlux<P, RK, 2, DE>(qL_lo, qRz, gamma, inv_gm1, gL_lo, gRz, fz_lo);
  // split pressure: the reference pressures of faces k-1/2 and k+1/2 (their left cells')
  T zref_lo = T(0), zref_hi = T(0);
  if constexpr (SplitPressure<P>) zref_lo = qL_lo.ph;
  qL_lo = qLz;  // left state of face k0+1/2
  gL_lo = gLz;
  double dti_max = 0.0;

  for (int k = k0; k < k1; k++) {
    // Stage plane k (the window centre) into LDS.  Every read of the previous plane
    // and of the interface/flux buffers finished before the last barrier of the
    // previous iteration.
    plane[ty + H][tx + H] = win[r];
    for (int h = tid; h < 2 * H * TY + 2 * H * TX; h += NT) {
      int pi, pj;
      if (h < 2 * H * TY) {
        int row = h / (2 * H), c = h - row * 2 * H;
        pj = H + row;
        pi = c < H ? c : TX + c;
      } else {
        int g = h - 2 * H * TY;
        int col = g / (2 * H), c = g - col * 2 * H;
        pi = H + col;
        pj = c < H ? c : TY + c;
      }
      int gi = min(i0 - H + pi, nx - 1), gj = min(j0 - H + pj, ny - 1);
      plane[pj][pi] = loadW(gi, gj, k);
    }
    // z: reconstruct cell k+1 and solve face k+1/2 while the halo loads land
#pragma unroll
    for (int m = 0; m < NS - 1; m++) win[m] = win[m + 1];
    win[NS - 1] = loadW(ic, jc, k + 1 + r);
    ReconDir<P, RK, 2, DE>(win, gamma, pf, qLz, qRz, &gLz, &gRz);
    T fz_hi[NC];
    IfaceFlux<P, RK, 2, DE>(qL_lo, qRz, gamma, inv_gm1, gL_lo, gRz, fz_hi);
    if constexpr (SplitPressure<P>) zref_hi = qL_lo.ph;
    qL_lo = qLz;
    gL_lo = gLz;
    __syncthreads();
    double acc[NC];  // dtodx dFx (then + dtody dFy at the update) of the lane's cell
    if constexpr (SM::kMerge) {
      // split pressure: the x and y faces' references (cells i-1, i and j-1, j), read
      // before the plane is rewritten at the top of the next iteration
      double ref_x = 0.0, ref_y = 0.0;
      if constexpr (SplitPressure<P>) {
        const double pc = double(plane[ty + H][tx + H].p_or_E);
        ref_x           = dtodx * (double(plane[ty + H][tx + H - 1].p_or_E) - pc);
        ref_y           = dtody * (double(plane[ty + H - 1][tx + H].p_or_E) - pc);
      }
      // x reconstruction of cells -1..TX of every row, then y reconstruction of cells
      // -1..TY of every column, as one list; y items in the y sweep's frame
      for (int it = tid; it < S::NQX + S::NQY; it += NT) {
        const bool y = it >= S::NQX;
        int row, col, stride;
        if (!y) {
          int rj = it / (TX + 2), ri = it - rj * (TX + 2);
          row = rj + H, col = ri - 1 + H - r, stride = 1;
        } else {
          int iy = it - S::NQX, rj = iy / TX, ri = iy - rj * TX;
          row = rj - 1 + H - r, col = ri + H, stride = S::PX;
        }
        const WCell<T, DE> *base = &plane[row][col];
        WCell<T, DE> st[NS];
#pragma unroll
        for (int m = 0; m < NS; m++) {
          st[m] = base[m * stride];
          RotateForY(st[m], y);
        }
        if constexpr (DE) {
          ReconDir<P, RK, 0, DE>(st, gamma, pf, qL[it], qR[it], &qgL[it], &qgR[it]);
        } else {
          ReconDir<P, RK, 0>(st, gamma, pf, qL[it], qR[it]);
        }
      }
      __syncthreads();
      // x faces, then y faces; the fluxes stay in each sweep's frame
      for (int it = tid; it < S::NFX + S::NFY; it += NT) {
        int lq, rq;
        if (it < S::NFX) {
          int fj = it / (TX + 1), fi = it - fj * (TX + 1);
          lq = fj * (TX + 2) + fi, rq = lq + 1;
        } else {
          int iy = it - S::NFX, fj = iy / TX, fi = iy - fj * TX;
          lq = S::NQX + fj * TX + fi, rq = lq + TX;
        }
        T g[NC];
        if constexpr (DE) {
          IfaceFlux<P, RK, 0, DE>(qL[lq], qR[rq], gamma, inv_gm1, qgL[lq], qgR[rq], g);
        } else {
          IfaceFlux<P, RK, 0, DE>(qL[lq], qR[rq], gamma, inv_gm1, T(0), T(0), g);
        }
#pragma unroll
        for (int q = 0; q < NC; q++) fl[q][it] = g[q];
      }
      __syncthreads();
      {
        int fa = ty * (TX + 1) + tx;
#pragma unroll
        for (int q = 0; q < NC; q++) acc[q] = dtodx * (double(fl[q][fa]) - double(fl[q][fa + 1]));
        if constexpr (SplitPressure<P>) {
          acc[1] += ref_x;
          acc[2] += ref_y;
        }
      }
    } else {
      // x reconstruction of cells -1..TX of every row
      for (int it = tid; it < S::NQX; it += NT) {
        int rj = it / (TX + 2), ri = it - rj * (TX + 2);  // plane column of cell = ri - 1 + H
        WCell<T, DE> st[NS];
#pragma unroll
        for (int m = 0; m < NS; m++) st[m] = plane[rj + H][ri - 1 + H - r + m];
        if constexpr (DE) {
          ReconDir<P, RK, 0, DE>(st, gamma, pf, qL[it], qR[it], &qgL[it], &qgR[it]);
        } else {
          ReconDir<P, RK, 0>(st, gamma, pf, qL[it], qR[it]);
        }
      }
      __syncthreads();
      for (int it = tid; it < S::NFX; it += NT) {
        int fj = it / (TX + 1), fi = it - fj * (TX + 1);  // between recon cells fi and fi+1 of row fj
        int ql = fj * (TX + 2) + fi;
        T g[NC];
        if constexpr (DE) {
          IfaceFlux<P, RK, 0, DE>(qL[ql], qR[ql + 1], gamma, inv_gm1, qgL[ql], qgR[ql + 1], g);
        } else {
          IfaceFlux<P, RK, 0, DE>(qL[ql], qR[ql + 1], gamma, inv_gm1, T(0), T(0), g);
        }
#pragma unroll
        for (int q = 0; q < NC; q++) fl[q][it] = g[q];
      }
      __syncthreads();
      {
        int fa = ty * (TX + 1) + tx;
#pragma unroll
        for (int q = 0; q < NC; q++) acc[q] = dtodx * (double(fl[q][fa]) - double(fl[q][fa + 1]));
        if constexpr (SplitPressure<P>) {
          // the x and y faces' references (cells i-1, i and j-1, j), read here: the plane
          // is rewritten after the last barrier of this iteration
          const double pc = double(plane[ty + H][tx + H].p_or_E);
          acc[1] += dtodx * (double(plane[ty + H][tx + H - 1].p_or_E) - pc);
          acc[2] += dtody * (double(plane[ty + H - 1][tx + H].p_or_E) - pc);
        }
      }
      // y reconstruction of cells -1..TY of every column
      for (int it = tid; it < S::NQY; it += NT) {
        int rj = it / TX, ri = it - rj * TX;
        WCell<T, DE> st[NS];
#pragma unroll
        for (int m = 0; m < NS; m++) st[m] = plane[rj - 1 + H - r + m][ri + H];
        if constexpr (DE) {
          ReconDir<P, RK, 1, DE>(st, gamma, pf, qL[it], qR[it], &qgL[it], &qgR[it]);
        } else {
          ReconDir<P, RK, 1>(st, gamma, pf, qL[it], qR[it]);
        }
      }
      __syncthreads();
      for (int it = tid; it < S::NFY; it += NT) {
        int fj = it / TX, fi = it - fj * TX;  // between recon rows fj and fj+1 of column fi
        const int lq = fj * TX + fi, rq = (fj + 1) * TX + fi;
        T g[NC];
        if constexpr (DE) {
          IfaceFlux<P, RK, 1, DE>(qL[lq], qR[rq], gamma, inv_gm1, qgL[lq], qgR[rq], g);
        } else {
          IfaceFlux<P, RK, 1, DE>(qL[lq], qR[rq], gamma, inv_gm1, T(0), T(0), g);
        }
#pragma unroll
        for (int q = 0; q < NC; q++) fl[q][it] = g[q];
      }
    }
    __syncthreads();
    if (own) {
      const size_t id = Lin(i, j, k, nx, ny);
      double un[NC];
      double ustart[4];  // density and momenta at the start of the step (static gravity)
      {
        // y fluxes: global order in the separate phases; merged, after the x faces and in
        // the y sweep's frame (d, my, mz, mx, E[, GasEnergy])
        const int fa = (SM::kMerge ? S::NFX : 0) + ty * TX + tx;
        // U += ((dtodx dFx + dtody dFy) + dtodz dFz), hydro_cuda.cu:241-257; GasEnergy
        // first takes the pdV source (the separate Partial_Update kernel, VL_3D_cuda.cu:314-324)
        if constexpr (SplitPressure<P>) acc[3] += dtodz * (double(zref_lo) - double(zref_hi));  // z faces' references
#pragma unroll
        for (int q = 0; q < NC; q++) {
          const int qy = SM::kMerge ? (q == 1 ? 3 : (q == 2 ? 1 : (q == 3 ? 2 : q))) : q;
          double axy = acc[q] + dtody * (double(fl[qy][fa]) - double(fl[qy][fa + TX]));
          double u0  = U[q * nc + id];
          if constexpr (P::kGrav) {
            if (q < 4) ustart[q] = u0;
          }
          if (DE && q == kGasEnergy) u0 = u0 + a.pdv[id];
          un[q] = u0 + (axy + dtodz * (double(fz_lo[q]) - double(fz_hi[q])));
        }
      }
      int mark = 0;
      if ((a.flags & kFlagDensityFloor) && un[0] < a.density_floor) {  // hydro_cuda.cu:279-299
        if (un[0] > 0) {
          double d0 = un[0];
          un[0]     = a.density_floor;
#pragma unroll
          for (int q = 1; q < NC; q++) un[q] *= (a.density_floor / d0);
        } else {
          mark = kRepairDensity;  // Average_Cell_Single_Field, done by RepairKernel
        }
      }
      if constexpr (P::kGrav) {  // STATIC_GRAV, hydro_cuda.cu:301-315 (after the repair)
        if (mark == 0) {
          double g[3];
          GravField3D(a.grav.id, a.grav.Pos(0, i, ng, a.dx), a.grav.Pos(1, j, ng, a.dy), a.grav.Pos(2, k, ng, a.dz),
                      g[0], g[1], g[2]);
          GravSource<3>(un, ustart, g, dt);
        } else {
          RecordRepairGrav(a.rep, id, ustart);
        }
      }
      if (mark == 0 && (a.flags & kFlagCrashRepair) && CrashedCell(un[0], un[4])) mark = kRepairAll;
      if (mark) {
        // the update as it stands; the repair, floors, dt and images follow in RepairKernel
#pragma unroll
        for (int q = 0; q < NC; q++) U[q * nc + id] = un[q];
        MarkRepair(a.rep, id, mark);
      } else if constexpr (DE) {
#pragma unroll
        for (int q = 0; q < NC; q++) U[q * nc + id] = un[q];
      } else {
        double dinv = 1.0 / un[0];
        double vx = un[1] * dinv, vy = un[2] * dinv, vz = un[3] * dinv;
        if (a.flags & kFlagTemperatureFloor) {  // Temperature_Floor_Kernel, hydro_cuda.cu:1132-1146
          double Ekin = 0.5 * un[0] * (vx * vx + vy * vy + vz * vz);
          double uint = P::kExact ? (un[4] - Ekin) / un[0] : (un[4] - Ekin) * dinv;
          if (uint < a.U_floor) un[4] = Ekin + un[0] * a.U_floor;
        }
#pragma unroll
        for (int q = 0; q < 5; q++) U[q * nc + id] = un[q];
        if (a.flags & kFlagReduceDt) {
          dti_max = Fmax(dti_max, InverseCrossingTimeP<P>(un[4], un[0], dinv, vx, vy, vz, a.dx, a.dy, a.dz, inv_dx,
                                                          inv_dy, inv_dz, a.gamma));
        }
        if (a.flags & kFlagPeriodicGhosts) WritePeriodicImages<5>(U, nc, nx, ny, nz, ng, i, j, k, un);
      }
    }
#pragma unroll
    for (int q = 0; q < NC; q++) fz_lo[q] = fz_hi[q];
    zref_lo = zref_hi;
  }

  if (!DE && (a.flags & kFlagReduceDt)) {
    // wave max (wave64 cross-lane), then one LDS pass and one atomic per workgroup
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) dti_max = Fmax(dti_max, __shfl_xor(dti_max, off, 64));
    if ((tid & 63) == 0) red[tid >> 6] = dti_max;
    __syncthreads();
    if (tid == 0) {
      double m = red[0];
      for (int w = 1; w < NT / 64; w++) m = Fmax(m, red[w]);
      atomicMax(a.dti, (unsigned long long)__double_as_longlong(m));
    }
  }
}

template <class P, int RK, int TX, int TY, bool DE = false>
__global__ void __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(Stage2Occupancy<P>::kMinWaves)))
Stage2Kernel(StepArgs a)
{
  __shared__ Stage2Smem<P, RK, TX, TY, DE> sm;
  int bx, by, bz;
  if (!TileOf(a, bx, by, bz)) return;
  Stage2Tile<P, RK, TX, TY, DE>(a, bx, by, bz, sm);
}

// ============================================================================ repairs
constexpr int kRepairThreads = 1024;  // one workgroup; 4 cells per lane cover a mark chunk

// Exclusive prefix sum of v over the workgroup (wave64 scan, then the wave totals).
__device__ __forceinline__ int RepairBlockScan(int v, int *s_wave, int &total)
{
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) s_wave[w] = x;
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int q = 0; q < kRepairThreads / 64; q++) {
      int t    = s_wave[q];
      s_wave[q] = acc;
      acc += t;
    }
    s_wave[kRepairThreads / 64] = acc;
  }
  __syncthreads();
  total = s_wave[kRepairThreads / 64];
  return s_wave[w] + x - v;
}

// Density of the neighbour (i+di, j+dj, k+dk) of real cell (i, j, k) for the density
// average: the state in memory, except that a ghost neighbour of a box whose periodic images
// stage 2 wrote takes the pre-step density its image held (the side buffer).
__device__ __forceinline__ double RepairNeighbourDensity(const StepArgs &a, int i, int j, int k, int di, int dj, int dk)
{
  const int ng = a.ng, Lx = a.nx - 2 * ng, Ly = a.ny - 2 * ng, Lz = a.nz - 2 * ng;
  const int ii = i + di, jj = j + dj, kk = k + dk;
  if (a.rep.side) {
    const int ri = i - ng, rj = j - ng, rk = k - ng;
    if (ii < ng) return a.rep.side[RepairSideIndex(0, rj, rk, Lx, Ly, Lz)];
    if (ii >= a.nx - ng) return a.rep.side[RepairSideIndex(1, rj, rk, Lx, Ly, Lz)];
    if (jj < ng) return a.rep.side[RepairSideIndex(2, ri, rk, Lx, Ly, Lz)];
    if (jj >= a.ny - ng) return a.rep.side[RepairSideIndex(3, ri, rk, Lx, Ly, Lz)];
    if (kk < ng) return a.rep.side[RepairSideIndex(4, ri, rj, Lx, Ly, Lz)];
    if (kk >= a.nz - ng) return a.rep.side[RepairSideIndex(5, ri, rj, Lx, Ly, Lz)];
  }
  return a.U[Lin(ii, jj, kk, a.nx, a.ny)];
}

// Average_Cell_All_Fields, hydro_cuda.cu:1179-1268: the mean density, momentum-weighted
// velocity and pressure of the cells of the 3x3x3 block with positive density and
// pressure (the cell itself included), in the reference's summation order; with dual
// energy GasEnergy takes the mean internal energy.  (Its printf diagnostics are left out.)
template <bool DE>
__device__ void RepairAllFields(const StepArgs &a, int i, int j, int k)
{
  const size_t nc = size_t(a.nx) * a.ny * a.nz, id = Lin(i, j, k, a.nx, a.ny);
  const double gamma = a.gamma;
  double d_av = 0.0, vx_av = 0.0, vy_av = 0.0, vz_av = 0.0, P_av = 0.0;
  int N = 0;
  for (int kk = k - 1; kk <= k + 1; kk++)
    for (int jj = j - 1; jj <= j + 1; jj++)
      for (int ii = i - 1; ii <= i + 1; ii++) {
        const size_t n = Lin(ii, jj, kk, a.nx, a.ny);
        const double d = a.U[n], mx = a.U[nc + n], my = a.U[2 * nc + n], mz = a.U[3 * nc + n];
        const double P = (a.U[4 * nc + n] - (0.5 / d) * (mx * mx + my * my + mz * mz)) * (gamma - 1.0);
        if (d > 0.0 && P > 0.0) {
          d_av += d;
          vx_av += mx;
          vy_av += my;
          vz_av += mz;
          P_av += P / (gamma - 1.0);
          N++;
        }
      }
  P_av  = P_av / N;
  vx_av = vx_av / d_av;
  vy_av = vy_av / d_av;
  vz_av = vz_av / d_av;
  d_av  = d_av / N;
  a.U[id]          = d_av;
  a.U[nc + id]     = d_av * vx_av;
  a.U[2 * nc + id] = d_av * vy_av;
  a.U[3 * nc + id] = d_av * vz_av;
  a.U[4 * nc + id] = P_av / (gamma - 1.0) + 0.5 * d_av * (vx_av * vx_av + vy_av * vy_av + vz_av * vz_av);
  if constexpr (DE) a.U[kGasEnergy * nc + id] = P_av / (gamma - 1.0);
}

// The repairs stage 2 marked, from the finished update, in cell order (the order the
// oracle's serial loops use), by one workgroup:
//   pass 0  density marks: the six-neighbour density average (Average_Cell_Single_Field,
//           hydro_cuda.cu:1155-1177), static gravity's source, then the crash test, which
//           turns the mark into an all-field mark when it fails;
//   pass 1  all-field marks: Average_Cell_All_Fields;
//   pass 2  what stage 2 fuses into its epilogue for unmarked cells (temperature floor,
//           inverse-dt reduction, periodic images; dual energy leaves them to its own
//           kernels), and the marks are cleared.
// A step without marks returns after reading the count.
template <class P, bool DE>
__device__ void RepairBox(const StepArgs &a)
{
  constexpr int NT = kRepairThreads, NC = kNumCons<DE>, CH = 1 << kRepairChunkShift;
  static_assert(CH == 4 * NT, "one mark chunk is four cells per lane");
  __shared__ unsigned s_count;
  __shared__ int s_wave[NT / 64 + 1];
  __shared__ unsigned s_chunks[NT];
  __shared__ unsigned s_cells[CH];
  const int tid = threadIdx.x;
  if (!a.rep.count) return;  // no repair scratch
  if (tid == 0) s_count = a.rep.count[0];
  __syncthreads();
  if (s_count == 0) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz, ng = a.ng;
  const size_t nc = size_t(nx) * ny * nz, n_chunks = (nc + CH - 1) / CH;
  const double dt = StepDt(a);
  for (int pass = 0; pass < 3; pass++) {
    for (size_t c0 = 0; c0 < n_chunks; c0 += NT) {
      const size_t c = c0 + tid;
      const int has  = (c < n_chunks && a.rep.chunk[c]) ? 1 : 0;
      int n_has;
      const int at = RepairBlockScan(has, s_wave, n_has);
      if (has) s_chunks[at] = unsigned(c - c0);
      __syncthreads();
      for (int h = 0; h < n_has; h++) {
        const size_t base = (c0 + s_chunks[h]) * CH;
        int m = 0;
        unsigned char mk[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const size_t id = base + 4 * size_t(tid) + q;
          mk[q]           = id < nc ? a.rep.mark[id] : 0;
          const bool take = pass == 0 ? mk[q] == kRepairDensity : (pass == 1 ? mk[q] == kRepairAll : mk[q] != 0);
          m += take ? 1 : 0;
        }
        int n_cells;
        int o = RepairBlockScan(m, s_wave, n_cells);
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const bool take = pass == 0 ? mk[q] == kRepairDensity : (pass == 1 ? mk[q] == kRepairAll : mk[q] != 0);
          if (take) s_cells[o++] = unsigned(4 * tid + q);
        }
        __syncthreads();
        if (pass < 2) {
          if (tid == 0) {
            for (int e = 0; e < n_cells; e++) {
              const size_t id = base + s_cells[e];
              const int i = int(id % nx), j = int((id / nx) % ny), k = int(id / (size_t(nx) * ny));
              if (pass == 1) {
                RepairAllFields<DE>(a, i, j, k);
                continue;
              }
              // hydro_cuda.cu:1165-1176, the same summation order
              const double v_l = RepairNeighbourDensity(a, i, j, k, -1, 0, 0);
              const double v_r = RepairNeighbourDensity(a, i, j, k, 1, 0, 0);
              const double v_d = RepairNeighbourDensity(a, i, j, k, 0, -1, 0);
              const double v_u = RepairNeighbourDensity(a, i, j, k, 0, 1, 0);
              const double v_b = RepairNeighbourDensity(a, i, j, k, 0, 0, -1);
              const double v_t = RepairNeighbourDensity(a, i, j, k, 0, 0, 1);
              a.U[id]          = (v_l + v_r + v_d + v_u + v_b + v_t) / 6;
              if (a.grav.id != 0) {  // hydro_cuda.cu:301-315, from the recorded start of the step
                const unsigned nrec = min(a.rep.count[1], unsigned(a.rep.grav_cap));
                for (unsigned r = 0; r < nrec; r++) {
                  const double *g = a.rep.grav + 5 * size_t(r);
                  if (size_t(g[0]) != id) continue;
                  double un[NC], gg[3];
#pragma unroll
                  for (int q = 0; q < NC; q++) un[q] = a.U[q * nc + id];
                  GravField3D(a.grav.id, a.grav.Pos(0, i, ng, a.dx), a.grav.Pos(1, j, ng, a.dy),
                              a.grav.Pos(2, k, ng, a.dz), gg[0], gg[1], gg[2]);
                  GravSource<3>(un, g + 1, gg, dt);
#pragma unroll
                  for (int q = 0; q < NC; q++) a.U[q * nc + id] = un[q];
                  break;
                }
              }
              if ((a.flags & kFlagCrashRepair) && CrashedCell(a.U[id], a.U[4 * nc + id])) a.rep.mark[id] = kRepairAll;
            }
          }
        } else {
          for (int e = tid; e < n_cells; e += NT) {
            const size_t id = base + s_cells[e];
            a.rep.mark[id]  = 0;
            if constexpr (!DE) {
              const int i = int(id % nx), j = int((id / nx) % ny), k = int(id / (size_t(nx) * ny));
              double un[5];
#pragma unroll
              for (int q = 0; q < 5; q++) un[q] = a.U[q * nc + id];
              double dinv = 1.0 / un[0];
              double vx = un[1] * dinv, vy = un[2] * dinv, vz = un[3] * dinv;
              if (a.flags & kFlagTemperatureFloor) {  // as the stage-2 epilogue
                double Ekin = 0.5 * un[0] * (vx * vx + vy * vy + vz * vz);
                double uint = P::kExact ? (un[4] - Ekin) / un[0] : (un[4] - Ekin) * dinv;
                if (uint < a.U_floor) un[4] = Ekin + un[0] * a.U_floor;
                a.U[4 * nc + id] = un[4];
              }
              if (a.flags & kFlagReduceDt) {
                double m_ = InverseCrossingTimeP<P>(un[4], un[0], dinv, vx, vy, vz, a.dx, a.dy, a.dz, 1.0 / a.dx,
                                                    1.0 / a.dy, 1.0 / a.dz, a.gamma);
                atomicMax(a.dti, (unsigned long long)__double_as_longlong(m_));
              }
              if (a.flags & kFlagPeriodicGhosts) WritePeriodicImages<5>(a.U, nc, nx, ny, nz, ng, i, j, k, un);
            }
          }
        }
        __syncthreads();
      }
      __syncthreads();
    }
  }
  for (size_t c = tid; c < n_chunks; c += NT) a.rep.chunk[c] = 0;
  if (tid == 0) a.rep.count[0] = 0, a.rep.count[1] = 0;
}

template <class P, bool DE>
__global__ void __launch_bounds__(kRepairThreads) RepairKernel(StepArgs a)
{
  RepairBox<P, DE>(a);
}

// One cell t of the side buffer (RepairSideIndex): the depth-1 ghost layer's density on the
// six faces, copied before stage 2 rewrites the ghosts.
__device__ __forceinline__ void RepairSideCell(const double *U, double *side, int nx, int ny, int nz, int ng, long t)
{
  const int Lx = nx - 2 * ng, Ly = ny - 2 * ng, Lz = nz - 2 * ng;
  const long fx = long(Ly) * Lz, fy = long(Lx) * Lz, fz = long(Lx) * Ly;
  int face, a, b;
  if (t < 2 * fx) {
    face = int(t / fx), t -= face * fx, a = int(t % Ly), b = int(t / Ly);
  } else if ((t -= 2 * fx) < 2 * fy) {
    face = 2 + int(t / fy), t -= (face - 2) * fy, a = int(t % Lx), b = int(t / Lx);
  } else if ((t -= 2 * fy) < 2 * fz) {
    face = 4 + int(t / fz), t -= (face - 4) * fz, a = int(t % Lx), b = int(t / Lx);
  } else {
    return;
  }
  int i, j, k;
  if (face < 2) {
    i = face == 0 ? ng - 1 : nx - ng, j = ng + a, k = ng + b;
  } else if (face < 4) {
    i = ng + a, j = face == 2 ? ng - 1 : ny - ng, k = ng + b;
  } else {
    i = ng + a, j = ng + b, k = face == 4 ? ng - 1 : nz - ng;
  }
  side[RepairSideIndex(face, a, b, Lx, Ly, Lz)] = U[Lin(i, j, k, nx, ny)];
}

}  // namespace chl
