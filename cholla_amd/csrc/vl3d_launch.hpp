// vl3d_launch.hpp -- host-side dispatch of the fused VL kernels over tile shapes.
#pragma once

#include <hip/hip_runtime.h>

#include "vl3d_kernels.hpp"

namespace chl
{

// Tile shapes (TX x TY lanes) compiled for both stages.  The autotuner picks a
// pair (stage-1 tile, stage-2 tile) and the z-chunk lengths per box size.
struct TileShape {
  int tx, ty;
};
constexpr int kNumTiles                = 5;
constexpr TileShape kTiles[kNumTiles] = {{32, 8}, {16, 16}, {64, 4}, {16, 8}, {32, 16}};

inline int CeilDiv(int a, int b) { return (a + b - 1) / b; }

// 1D launch over the tile grid (gx, gy, gz); padded to a multiple of 8 for the
// XCD-contiguous order.
inline unsigned TileGrid(StepArgs &b, int gx, int gy, int gz)
{
  b.gx = gx, b.gy = gy, b.gz = gz;
  const long T = long(gx) * gy * gz;
  return unsigned(b.xcd_order ? (T + 7) / 8 * 8 : T);
}

template <class P, int TX, int TY, bool DE = false>
hipError_t LaunchStage1T(const StepArgs &a, int R1, hipStream_t s)
{
  int ox = a.nx - 2 * a.ng + 2 * R1, oy = a.ny - 2 * a.ng + 2 * R1, oz = a.nz - 2 * a.ng + 2 * R1;
  StepArgs b     = a;
  unsigned grid  = TileGrid(b, CeilDiv(ox, TX), CeilDiv(oy, TY), CeilDiv(oz, a.kchunk));
  hipLaunchKernelGGL((Stage1Kernel<P, TX, TY, DE>), dim3(grid), dim3(TX * TY), 0, s, b, R1);
  return hipGetLastError();
}

template <class P, int RK, int TX, int TY, bool DE = false>
hipError_t LaunchStage2T(const StepArgs &a, hipStream_t s)
{
  int ox = a.nx - 2 * a.ng, oy = a.ny - 2 * a.ng, oz = a.nz - 2 * a.ng;
  StepArgs b    = a;
  unsigned grid = TileGrid(b, CeilDiv(ox, TX), CeilDiv(oy, TY), CeilDiv(oz, a.kchunk));
  hipLaunchKernelGGL((Stage2Kernel<P, RK, TX, TY, DE>), dim3(grid), dim3(TX * TY), 0, s, b);
  return hipGetLastError();
}

// Dual-energy, HLL, PCM and PLMP kernels are compiled for the first kNumDeTiles tile shapes only.
constexpr int kNumDeTiles = 2;

// fn(Q{}) with Q = P carrying the solver `riemann` selects (HLLC: P itself).
template <class P, class Fn>
hipError_t WithSolver(int riemann, Fn &&fn)
{
  if (riemann == kRiemannHll) return fn(HllPolicy<P>{});
  if (riemann == kRiemannRoe) return fn(RoePolicy<P>{});
  if (riemann == kRiemannExact) return fn(ExactPolicy<P>{});
  return fn(P{});
}

template <class P>
hipError_t LaunchStage1(const StepArgs &a, int R1, int tile, hipStream_t s)
{
  if (a.riemann != kRiemannHllc) {  // HLL, Roe, exact: the first two tile shapes
    if (tile > 1) return hipErrorInvalidValue;
    return WithSolver<P>(a.riemann, [&](auto q) {
      using Q = decltype(q);
      if (a.de) return tile == 0 ? LaunchStage1T<Q, 32, 8, true>(a, R1, s) : LaunchStage1T<Q, 16, 16, true>(a, R1, s);
      return tile == 0 ? LaunchStage1T<Q, 32, 8>(a, R1, s) : LaunchStage1T<Q, 16, 16>(a, R1, s);
    });
  }
  if (a.de) {
    switch (tile) {
      case 0:
        return LaunchStage1T<P, 32, 8, true>(a, R1, s);
      case 1:
        return LaunchStage1T<P, 16, 16, true>(a, R1, s);
    }
    return hipErrorInvalidValue;
  }
  switch (tile) {
    case 0:
      return LaunchStage1T<P, 32, 8>(a, R1, s);
    case 1:
      return LaunchStage1T<P, 16, 16>(a, R1, s);
    case 2:
      return LaunchStage1T<P, 64, 4>(a, R1, s);
    case 3:
      return LaunchStage1T<P, 16, 8>(a, R1, s);
    case 4:
      return LaunchStage1T<P, 32, 16>(a, R1, s);
  }
  return hipErrorInvalidValue;
}

template <class P, int RK>
hipError_t LaunchStage2R(const StepArgs &a, int tile, hipStream_t s)
{
  if (a.grav.id != 0) {  // static gravity: the first two tile shapes, every solver, with or without DE
    if (tile > 1) return hipErrorInvalidValue;
    return WithSolver<P>(a.riemann, [&](auto q) {
      using Q = GravPolicy<decltype(q)>;
      if (a.de) return tile == 0 ? LaunchStage2T<Q, RK, 32, 8, true>(a, s) : LaunchStage2T<Q, RK, 16, 16, true>(a, s);
      return tile == 0 ? LaunchStage2T<Q, RK, 32, 8>(a, s) : LaunchStage2T<Q, RK, 16, 16>(a, s);
    });
  }
  // PCM, PLMP, PPMP, HLL, Roe and the exact solver: the first two tile shapes, with or
  // without dual energy
  auto narrow = [&]() {
    if (tile > 1) return hipErrorInvalidValue;
    return WithSolver<P>(a.riemann, [&](auto q) {
      using Q = decltype(q);
      if (a.de) return tile == 0 ? LaunchStage2T<Q, RK, 32, 8, true>(a, s) : LaunchStage2T<Q, RK, 16, 16, true>(a, s);
      return tile == 0 ? LaunchStage2T<Q, RK, 32, 8>(a, s) : LaunchStage2T<Q, RK, 16, 16>(a, s);
    });
  };
  if constexpr (RK == kPCM || RK == kPLMP || RK == kPPMP) {
    return narrow();
  } else {
    if (a.riemann != kRiemannHllc) return narrow();
    if (a.de) {
      switch (tile) {
        case 0:
          return LaunchStage2T<P, RK, 32, 8, true>(a, s);
        case 1:
          return LaunchStage2T<P, RK, 16, 16, true>(a, s);
      }
      return hipErrorInvalidValue;
    }
    switch (tile) {
      case 0:
        return LaunchStage2T<P, RK, 32, 8>(a, s);
      case 1:
        return LaunchStage2T<P, RK, 16, 16>(a, s);
      case 2:
        return LaunchStage2T<P, RK, 64, 4>(a, s);
      case 3:
        return LaunchStage2T<P, RK, 16, 8>(a, s);
      case 4:
        return LaunchStage2T<P, RK, 32, 16>(a, s);
    }
    return hipErrorInvalidValue;
  }
}

template <class P>
hipError_t LaunchStage2(const StepArgs &a, int recon, int tile, hipStream_t s)
{
  if (recon == kPPMC) return LaunchStage2R<P, kPPMC>(a, tile, s);
  if (recon == kPLMC) return LaunchStage2R<P, kPLMC>(a, tile, s);
  if (recon == kPLMP) return LaunchStage2R<P, kPLMP>(a, tile, s);
  if (recon == kPPMP) return LaunchStage2R<P, kPPMP>(a, tile, s);
  if (recon == kPCM) return LaunchStage2R<P, kPCM>(a, tile, s);
  return hipErrorInvalidValue;
}

// The repairs after stage 2 (RepairKernel): one workgroup.
template <class P>
hipError_t LaunchRepair(const StepArgs &a, hipStream_t s)
{
  if (a.de) {
    hipLaunchKernelGGL((RepairKernel<P, true>), dim3(1), dim3(kRepairThreads), 0, s, a);
  } else {
    hipLaunchKernelGGL((RepairKernel<P, false>), dim3(1), dim3(kRepairThreads), 0, s, a);
  }
  return hipGetLastError();
}

// The side buffer of a density repair next to fused periodic images (hydro_exact.hip).
hipError_t LaunchRepairSide(const StepArgs &a, hipStream_t s);

// Explicitly instantiated per policy in hydro_exact.hip / hydro_fast.hip.
hipError_t LaunchStage1Policy(int policy, const StepArgs &a, int R1, int tile, hipStream_t s);
hipError_t LaunchStage2Policy(int policy, const StepArgs &a, int recon, int tile, hipStream_t s);
hipError_t LaunchStage1Exact(const StepArgs &a, int R1, int tile, hipStream_t s);
hipError_t LaunchStage2Exact(const StepArgs &a, int recon, int tile, hipStream_t s);
hipError_t LaunchStage1Fast(int policy, const StepArgs &a, int R1, int tile, hipStream_t s);
hipError_t LaunchStage2Fast(int policy, const StepArgs &a, int recon, int tile, hipStream_t s);
hipError_t LaunchRepairExact(const StepArgs &a, hipStream_t s);
hipError_t LaunchRepairFast(int policy, const StepArgs &a, hipStream_t s);

}  // namespace chl
