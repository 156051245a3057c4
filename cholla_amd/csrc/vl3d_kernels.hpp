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
};

enum : int {
  kFlagDensityFloor       = 1,
  kFlagTemperatureFloor   = 2,
  kFlagPeriodicGhosts     = 4,  // write periodic ghost images from the stage-2 epilogue
  kFlagReduceDt           = 8,
  kFlagPositivityFallback = 16,  // PPMC positivity fallback (see ReconPpmc)
  kFlagConservedW         = 32,  // exact form, PCM: stage 1 stores momenta, not velocities
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
template <typename T, bool DE = false>
struct PcmCell {
  T d, mx, my, mz, E, vx, vy, vz, p;
};
template <typename T>
struct PcmCell<T, true> {
  T d, mx, my, mz, E, vx, vy, vz, p, ge;
};

// Number of conserved fields a stage carries: 5, or 6 with GasEnergy.
template <bool DE>
constexpr int kNumCons = DE ? 6 : 5;

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
  FaceFlux<P, 2, DE>(PcmSide<P, 2, DE>(c_lo, gamma), PcmSide<P, 2, DE>(c_mid, gamma), gamma, fz_lo);

  for (int k = k0; k < k1; k++) {
    {
      size_t id = Lin(ic, jc, k, nx, ny);
#pragma unroll
      for (int f = 0; f < NC; f++) u_mid[f] = U[f * nc + id];
    }
    PcmCell<T, DE> c_hi = load_cell(ic, jc, k + 1);
    __syncthreads();  // previous plane fully consumed
    plane[ty + 1][tx + 1] = c_mid;
    // halo: x sides (PY-2 rows x 2) and y sides (PX-2 cols x 2), no corners
    for (int h = tid; h < 2 * TY + 2 * TX; h += NT) {
      int pi, pj;
      if (h < 2 * TY) {
        pj = 1 + (h >> 1);
        pi = (h & 1) ? TX + 1 : 0;
      } else {
        int g = h - 2 * TY;
        pi    = 1 + (g >> 1);
        pj    = (g & 1) ? TY + 1 : 0;
      }
      int gi = min(i0 - 1 + pi, nx - 1), gj = min(j0 - 1 + pj, ny - 1);
      plane[pj][pi] = load_cell(gi, gj, k);
    }
    __syncthreads();
    // x faces: face fi sits between plane cells (fi, fj+1) and (fi+1, fj+1)
    for (int it = tid; it < NFX; it += NT) {
      int fj = it / (TX + 1), fi = it - fj * (TX + 1);
      T g[NC];
      FaceFlux<P, 0, DE>(PcmSide<P, 0, DE>(plane[fj + 1][fi], gamma), PcmSide<P, 0, DE>(plane[fj + 1][fi + 1], gamma),
                         gamma, g);
#pragma unroll
      for (int q = 0; q < NC; q++) fx[q][fj][fi] = g[q];
    }
    for (int it = tid; it < NFY; it += NT) {
      int fj = it / TX, fi = it - fj * TX;
      T g[NC];
      FaceFlux<P, 1, DE>(PcmSide<P, 1, DE>(plane[fj][fi + 1], gamma), PcmSide<P, 1, DE>(plane[fj + 1][fi + 1], gamma),
                         gamma, g);
#pragma unroll
      for (int q = 0; q < NC; q++) fy[q][fj][fi] = g[q];
    }
    T fz_hi[NC];
    FaceFlux<P, 2, DE>(PcmSide<P, 2, DE>(c_mid, gamma), PcmSide<P, 2, DE>(c_hi, gamma), gamma, fz_hi);
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
      if constexpr (DE) {
        // Load_Data under DE (reconstruction.h:179-191): the exact form keeps E and
        // GasEnergy and recovers p and ge per sweep; the others store p and ge
        if constexpr (P::kExact) {
          W[id]          = T(uh[0]);
          W[nc + id]     = T(uh[1] / uh[0]);
          W[2 * nc + id] = T(uh[2] / uh[0]);
          W[3 * nc + id] = T(uh[3] / uh[0]);
          W[4 * nc + id] = T(uh[4]);
          W[5 * nc + id] = T(uh[5]);
        } else {
          double vx = uh[1] / uh[0], vy = uh[2] / uh[0], vz = uh[3] / uh[0];
          double E_kin   = 0.5 * uh[0] * SquareMag(vx, vy, vz);
          W[id]          = T(uh[0]);
          W[nc + id]     = T(vx);
          W[2 * nc + id] = T(vy);
          W[3 * nc + id] = T(vz);
          W[4 * nc + id] = T(PressureFromDE(uh[4], uh[4] - E_kin, uh[5], a.gamma));
          W[5 * nc + id] = T(uh[5] / uh[0]);
        }
      } else {
        Prim<double> w = LoadPrim(uh[0], uh[1], uh[2], uh[3], uh[4], a.gamma);
        // PCM hands the conserved half state itself to the Riemann solver (pcm_cuda.cu)
        const bool cons_w = P::kExact && (a.flags & kFlagConservedW);
        W[id]          = T(w.d);
        W[nc + id]     = cons_w ? T(uh[1]) : T(w.vx);
        W[2 * nc + id] = cons_w ? T(uh[2]) : T(w.vy);
        W[3 * nc + id] = cons_w ? T(uh[3]) : T(w.vz);
        W[4 * nc + id] = P::kExact ? T(uh[4]) : T(w.p);  // see WCell
      }
    }
#pragma unroll
    for (int q = 0; q < NC; q++) fz_lo[q] = fz_hi[q];
    c_lo  = c_mid;
    c_mid = c_hi;
  }
}

template <class P, int TX, int TY, bool DE = false>
__global__ void __launch_bounds__(TX *TY) Stage1Kernel(StepArgs a, int R1)
{
  __shared__ Stage1Smem<P, TX, TY, DE> sm;
  // the inverse-dt accumulator is reset here; stage 2 (next launch) fills it
  if ((a.flags & kFlagReduceDt) && blockIdx.x == 0 && threadIdx.x == 0) *a.dti = 0ull;
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
template <typename T, bool DE = false>
struct WCell {
  T d, vx, vy, vz, p_or_E;
};
template <typename T>
struct WCell<T, true> {
  T d, vx, vy, vz, p_or_E, g;
};

template <class P, int RK, int DIR, bool DE = false>
__device__ __forceinline__ void ReconDir(const WCell<typename P::T, DE> *s, typename P::T gamma, bool pos_fallback,
                                         Prim<typename P::T> &L, Prim<typename P::T> &R, typename P::T *geL = nullptr,
                                         typename P::T *geR = nullptr)
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
    } else if constexpr (P::kExact && (RK == kPLMP || RK == kPPMP)) {
      // plmp_cuda.cu:113-120 (ppmp_cuda.cu:160-171): the squared velocities summed left to right
      const T v2 = (st[m].vx * st[m].vx + st[m].vy * st[m].vy) + st[m].vz * st[m].vz;
      st[m].p    = Fmax((s[m].p_or_E - T(0.5) * st[m].d * v2) * (gamma - T(1.0)), T(kTinyNumber));
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
  } else if constexpr (RK == kPLMP) {
    ReconPlmp<P>(st, L, R);
  } else if constexpr (RK == kPPMP) {
    ReconPpmp<P>(st, L, R);
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
    return RStateFromCons<P>(q.d, q.vx, q.vy, q.vz, q.p, gamma);  // q holds (d, momenta, E)
  } else if constexpr (P::kExact && (RK == kPLMP || RK == kPPMP)) {  // ppmp_cuda.cu:645-650 likewise
    const T E = q.p / (gamma - T(1.0)) + T(0.5) * q.d * ((q.vx * q.vx + q.vy * q.vy) + q.vz * q.vz);
    return RStateFromCons<P>(q.d, q.d * q.vx, q.d * q.vy, q.d * q.vz, E, gamma);
  } else {
    return RStateFromPrim<P, DE>(q, gamma, inv_gm1, ge);
  }
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
  static constexpr int kMinWaves = P::kId == PolicyMixed::kId ? 1 : 2;
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

// LDS of one stage-2 tile.
template <class P, int RK, int TX, int TY, bool DE = false>
struct Stage2Smem {
  using T = typename P::T;
  using S = Stage2Shape<P, RK, TX, TY>;
  WCell<T, DE> plane[S::PY][S::PX];
  Prim<T> qL[S::NQ];
  Prim<T> qR[S::NQ];
  GeIfaces<T, S::NQ, DE> qg;
  T fl[kNumCons<DE>][S::NF];
  double red[TX * TY / 64];
};

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
    if constexpr (DE) c.g = W[5 * nc + id];
    return c;
  };

  // z window of this lane's column: win[m] = W(c - r + m) for the cell c whose z
  // interfaces are reconstructed next.  Prologue: cells k0-1 and k0 give the face k0-1/2.
  WCell<T, DE> win[NS];
#pragma unroll
  for (int m = 0; m < NS; m++) win[m] = loadW(ic, jc, k0 - 1 - r + m);
  Prim<T> qLz, qRz;
  T gLz = T(0), gRz = T(0);
  ReconDir<P, RK, 2, DE>(win, gamma, pf, qLz, qRz, &gLz, &gRz);  // cell k0-1
  Prim<T> qL_lo = qLz;                                       // left state of face k0-1/2
  T gL_lo       = gLz;
#pragma unroll
  for (int m = 0; m < NS - 1; m++) win[m] = win[m + 1];
  win[NS - 1] = loadW(ic, jc, k0 + r);
  ReconDir<P, RK, 2, DE>(win, gamma, pf, qLz, qRz, &gLz, &gRz);  // cell k0
  T fz_lo[NC];                                               // flux through face k-1/2 of the current cell
  FaceFlux<P, 2, DE>(IfaceState<P, RK, DE>(qL_lo, gamma, inv_gm1, gL_lo), IfaceState<P, RK, DE>(qRz, gamma, inv_gm1, gRz),
                     gamma, fz_lo);
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
    FaceFlux<P, 2, DE>(IfaceState<P, RK, DE>(qL_lo, gamma, inv_gm1, gL_lo),
                       IfaceState<P, RK, DE>(qRz, gamma, inv_gm1, gRz), gamma, fz_hi);
    qL_lo = qLz;
    gL_lo = gLz;
    __syncthreads();
    // x reconstruction of cells -1..TX of every row
    for (int it = tid; it < S::NQX; it += NT) {
      int rj = it / (TX + 2), ri = it - rj * (TX + 2);  // plane column of cell = ri - 1 + H
      WCell<T, DE> s[NS];
#pragma unroll
      for (int m = 0; m < NS; m++) s[m] = plane[rj + H][ri - 1 + H - r + m];
      if constexpr (DE) {
        ReconDir<P, RK, 0, DE>(s, gamma, pf, qL[it], qR[it], &qgL[it], &qgR[it]);
      } else {
        ReconDir<P, RK, 0>(s, gamma, pf, qL[it], qR[it]);
      }
    }
    __syncthreads();
    for (int it = tid; it < S::NFX; it += NT) {
      int fj = it / (TX + 1), fi = it - fj * (TX + 1);  // between recon cells fi and fi+1 of row fj
      int ql = fj * (TX + 2) + fi;
      T g[NC];
      if constexpr (DE) {
        FaceFlux<P, 0, DE>(IfaceState<P, RK, DE>(qL[ql], gamma, inv_gm1, qgL[ql]),
                           IfaceState<P, RK, DE>(qR[ql + 1], gamma, inv_gm1, qgR[ql + 1]), gamma, g);
      } else {
        FaceFlux<P, 0, DE>(IfaceState<P, RK>(qL[ql], gamma, inv_gm1), IfaceState<P, RK>(qR[ql + 1], gamma, inv_gm1),
                           gamma, g);
      }
#pragma unroll
      for (int q = 0; q < NC; q++) fl[q][it] = g[q];
    }
    __syncthreads();
    double acc[NC];
    {
      int fa = ty * (TX + 1) + tx;
#pragma unroll
      for (int q = 0; q < NC; q++) acc[q] = dtodx * (double(fl[q][fa]) - double(fl[q][fa + 1]));
    }
    // y reconstruction of cells -1..TY of every column
    for (int it = tid; it < S::NQY; it += NT) {
      int rj = it / TX, ri = it - rj * TX;
      WCell<T, DE> s[NS];
#pragma unroll
      for (int m = 0; m < NS; m++) s[m] = plane[rj - 1 + H - r + m][ri + H];
      if constexpr (DE) {
        ReconDir<P, RK, 1, DE>(s, gamma, pf, qL[it], qR[it], &qgL[it], &qgR[it]);
      } else {
        ReconDir<P, RK, 1>(s, gamma, pf, qL[it], qR[it]);
      }
    }
    __syncthreads();
    for (int it = tid; it < S::NFY; it += NT) {
      int fj = it / TX, fi = it - fj * TX;  // between recon rows fj and fj+1 of column fi
      const int lq = fj * TX + fi, rq = (fj + 1) * TX + fi;
      T g[NC];
      if constexpr (DE) {
        FaceFlux<P, 1, DE>(IfaceState<P, RK, DE>(qL[lq], gamma, inv_gm1, qgL[lq]),
                           IfaceState<P, RK, DE>(qR[rq], gamma, inv_gm1, qgR[rq]), gamma, g);
      } else {
        FaceFlux<P, 1, DE>(IfaceState<P, RK>(qL[lq], gamma, inv_gm1), IfaceState<P, RK>(qR[rq], gamma, inv_gm1), gamma,
                           g);
      }
#pragma unroll
      for (int q = 0; q < NC; q++) fl[q][it] = g[q];
    }
    __syncthreads();
    if (own) {
      const size_t id = Lin(i, j, k, nx, ny);
      double un[NC];
      double ustart[4];  // density and momenta at the start of the step (static gravity)
      {
        int fa = ty * TX + tx;
        // U += ((dtodx dFx + dtody dFy) + dtodz dFz), hydro_cuda.cu:241-257; GasEnergy
        // first takes the pdV source (the separate Partial_Update kernel, VL_3D_cuda.cu:314-324)
#pragma unroll
        for (int q = 0; q < NC; q++) {
          double axy = acc[q] + dtody * (double(fl[q][fa]) - double(fl[q][fa + TX]));
          double u0  = U[q * nc + id];
          if constexpr (P::kGrav) {
            if (q < 4) ustart[q] = u0;
          }
          if (DE && q == kGasEnergy) u0 = u0 + a.pdv[id];
          un[q] = u0 + (axy + dtodz * (double(fz_lo[q]) - double(fz_hi[q])));
        }
      }
      if ((a.flags & kFlagDensityFloor) && un[0] < a.density_floor) {  // hydro_cuda.cu:279-299
        if (un[0] > 0) {
          double d0 = un[0];
          un[0]     = a.density_floor;
#pragma unroll
          for (int q = 1; q < NC; q++) un[q] *= (a.density_floor / d0);
        } else {
          // Average_Cell_Single_Field (hydro_cuda.cu:1155-1177) over the six
          // neighbours as they are in memory now; the reference reads them
          // while other threads update them as well.
          un[0] = (U[id - 1] + U[id + 1] + U[id - nx] + U[id + nx] + U[id - size_t(nx) * ny] +
                   U[id + size_t(nx) * ny]) /
                  6;
        }
      }
      if constexpr (P::kGrav) {  // STATIC_GRAV, hydro_cuda.cu:301-315
        double g[3];
        GravField3D(a.grav.id, a.grav.Pos(0, i, ng, a.dx), a.grav.Pos(1, j, ng, a.dy), a.grav.Pos(2, k, ng, a.dz), g[0],
                    g[1], g[2]);
        GravSource<3>(un, ustart, g, dt);
      }
      if constexpr (DE) {
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

}  // namespace chl
