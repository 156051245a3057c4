// ensemble_launch.hpp -- host dispatch of the ensemble kernels over tile shapes.
#pragma once

#include "ensemble_kernels.hpp"
#include "vl3d_launch.hpp"

namespace chl
{

inline int PaddedItems(int n) { return (n + 7) / 8 * 8; }

template <class P, int TX, int TY, bool DE = false>
hipError_t EnsStage1T(const EnsembleArgs &e, hipStream_t s)
{
  if (e.n_items == 0) return hipSuccess;
  hipLaunchKernelGGL((EnsembleStage1Kernel<P, TX, TY, DE>), dim3(PaddedItems(e.n_items)), dim3(TX * TY), 0, s, e);
  return hipGetLastError();
}

template <class P, int RK, int TX, int TY, bool DE = false>
hipError_t EnsStage2T(const EnsembleArgs &e, hipStream_t s)
{
  if (e.n_items == 0) return hipSuccess;
  hipLaunchKernelGGL((EnsembleStage2Kernel<P, RK, TX, TY, DE>), dim3(PaddedItems(e.n_items)), dim3(TX * TY), 0, s,
                     e);
  return hipGetLastError();
}

template <int TX, int TY, int PH>
hipError_t EnsDeT(const EnsembleArgs &e, hipStream_t s)
{
  if (e.n_items == 0) return hipSuccess;
  hipLaunchKernelGGL((EnsembleDeKernel<TX, TY, PH>), dim3(PaddedItems(e.n_items)), dim3(TX * TY), 0, s, e);
  return hipGetLastError();
}

template <int PH>
hipError_t EnsDe(const EnsembleArgs &e, int tile, hipStream_t s)
{
  switch (tile) {
    case 0:
      return EnsDeT<32, 8, PH>(e, s);
    case 1:
      return EnsDeT<16, 16, PH>(e, s);
  }
  return hipErrorInvalidValue;
}

template <class P>
hipError_t EnsStage1(const EnsembleArgs &e, int tile, hipStream_t s)
{
  if (e.riemann != kRiemannHllc) {  // HLL, Roe, exact: the first two tile shapes
    if (tile > 1) return hipErrorInvalidValue;
    return WithSolver<P>(e.riemann, [&](auto q) {
      using Q = decltype(q);
      if (e.de) return tile == 0 ? EnsStage1T<Q, 32, 8, true>(e, s) : EnsStage1T<Q, 16, 16, true>(e, s);
      return tile == 0 ? EnsStage1T<Q, 32, 8>(e, s) : EnsStage1T<Q, 16, 16>(e, s);
    });
  }
  if (e.de) {
    switch (tile) {
      case 0:
        return EnsStage1T<P, 32, 8, true>(e, s);
      case 1:
        return EnsStage1T<P, 16, 16, true>(e, s);
    }
    return hipErrorInvalidValue;
  }
  switch (tile) {
    case 0:
      return EnsStage1T<P, 32, 8>(e, s);
    case 1:
      return EnsStage1T<P, 16, 16>(e, s);
    case 2:
      return EnsStage1T<P, 64, 4>(e, s);
    case 3:
      return EnsStage1T<P, 16, 8>(e, s);
    case 4:
      return EnsStage1T<P, 32, 16>(e, s);
  }
  return hipErrorInvalidValue;
}

template <class P, int RK>
hipError_t EnsStage2R(const EnsembleArgs &e, int tile, hipStream_t s)
{
  // PCM, PLMP, PPMP and the HLL, Roe and exact solvers: the first two tile shapes, with or
  // without dual energy
  if (RK == kPCM || RK == kPLMP || RK == kPPMP || e.riemann != kRiemannHllc) {
    if (tile > 1) return hipErrorInvalidValue;
    return WithSolver<P>(e.riemann, [&](auto q) {
      using Q = decltype(q);
      if (e.de) return tile == 0 ? EnsStage2T<Q, RK, 32, 8, true>(e, s) : EnsStage2T<Q, RK, 16, 16, true>(e, s);
      return tile == 0 ? EnsStage2T<Q, RK, 32, 8>(e, s) : EnsStage2T<Q, RK, 16, 16>(e, s);
    });
  }
  if constexpr (RK == kPCM || RK == kPLMP || RK == kPPMP) {
    return hipErrorInvalidValue;  // handled above
  } else if (e.de) {
    switch (tile) {
      case 0:
        return EnsStage2T<P, RK, 32, 8, true>(e, s);
      case 1:
        return EnsStage2T<P, RK, 16, 16, true>(e, s);
    }
    return hipErrorInvalidValue;
  } else {
    switch (tile) {
      case 0:
        return EnsStage2T<P, RK, 32, 8>(e, s);
      case 1:
        return EnsStage2T<P, RK, 16, 16>(e, s);
      case 2:
        return EnsStage2T<P, RK, 64, 4>(e, s);
      case 3:
        return EnsStage2T<P, RK, 16, 8>(e, s);
      case 4:
        return EnsStage2T<P, RK, 32, 16>(e, s);
    }
    return hipErrorInvalidValue;
  }
}

template <class P>
hipError_t EnsStage2(const EnsembleArgs &e, int recon, int tile, hipStream_t s)
{
  if (recon == kPPMC) return EnsStage2R<P, kPPMC>(e, tile, s);
  if (recon == kPLMC) return EnsStage2R<P, kPLMC>(e, tile, s);
  if (recon == kPLMP) return EnsStage2R<P, kPLMP>(e, tile, s);
  if (recon == kPPMP) return EnsStage2R<P, kPPMP>(e, tile, s);
  if (recon == kPCM) return EnsStage2R<P, kPCM>(e, tile, s);
  return hipErrorInvalidValue;
}

template <class P, int RK, int A, int B, bool DE = false>
struct PersistentPair {
  static constexpr int TX1 = kTiles[A].tx, TY1 = kTiles[A].ty, TX2 = kTiles[B].tx, TY2 = kTiles[B].ty;
  static constexpr int NT  = TX2 * TY2;
  static hipError_t Launch(const EnsembleArgs &e, int grid, hipStream_t s)
  {
    if (e.n_items == 0) return hipSuccess;
    hipLaunchKernelGGL((EnsemblePersistentKernel<P, RK, TX1, TY1, TX2, TY2, DE>), dim3(grid), dim3(NT), 0, s, e);
    return hipGetLastError();
  }
  static hipError_t Slots(int *per_cu)
  {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu,
                                                        EnsemblePersistentKernel<P, RK, TX1, TY1, TX2, TY2, DE>, NT, 0);
  }
};

// Pairs compiled with dual energy: both tiles among the dual-energy tile shapes.
constexpr bool DePair(int pair) { return kPairs[pair][0] < kNumDeTiles && kPairs[pair][1] < kNumDeTiles; }

template <class P, int RK>
hipError_t EnsPersistentR(const EnsembleArgs &e, int pair, int grid, int *per_cu, hipStream_t s)
{
  static_assert(kNumPairs == 4, "pair table and dispatch below go together");
  static_assert(!DePair(0) && !DePair(1) && DePair(2) && DePair(3), "dual-energy pairs below");
  if (e.de) {
    switch (pair) {
      case 2:
        return per_cu ? PersistentPair<P, RK, kPairs[2][0], kPairs[2][1], true>::Slots(per_cu)
                      : PersistentPair<P, RK, kPairs[2][0], kPairs[2][1], true>::Launch(e, grid, s);
      case 3:
        return per_cu ? PersistentPair<P, RK, kPairs[3][0], kPairs[3][1], true>::Slots(per_cu)
                      : PersistentPair<P, RK, kPairs[3][0], kPairs[3][1], true>::Launch(e, grid, s);
    }
    return hipErrorInvalidValue;
  }
  switch (pair) {
    case 0:
      return per_cu ? PersistentPair<P, RK, kPairs[0][0], kPairs[0][1]>::Slots(per_cu)
                    : PersistentPair<P, RK, kPairs[0][0], kPairs[0][1]>::Launch(e, grid, s);
    case 1:
      return per_cu ? PersistentPair<P, RK, kPairs[1][0], kPairs[1][1]>::Slots(per_cu)
                    : PersistentPair<P, RK, kPairs[1][0], kPairs[1][1]>::Launch(e, grid, s);
    case 2:
      return per_cu ? PersistentPair<P, RK, kPairs[2][0], kPairs[2][1]>::Slots(per_cu)
                    : PersistentPair<P, RK, kPairs[2][0], kPairs[2][1]>::Launch(e, grid, s);
    case 3:
      return per_cu ? PersistentPair<P, RK, kPairs[3][0], kPairs[3][1]>::Slots(per_cu)
                    : PersistentPair<P, RK, kPairs[3][0], kPairs[3][1]>::Launch(e, grid, s);
  }
  return hipErrorInvalidValue;
}

template <class P>
hipError_t EnsRepair(const EnsembleArgs &e, hipStream_t s)
{
  if (e.n_boxes == 0) return hipSuccess;
  if (e.de) {
    hipLaunchKernelGGL((EnsembleRepairKernel<P, true>), dim3(e.n_boxes), dim3(kRepairThreads), 0, s, e);
  } else {
    hipLaunchKernelGGL((EnsembleRepairKernel<P, false>), dim3(e.n_boxes), dim3(kRepairThreads), 0, s, e);
  }
  return hipGetLastError();
}

// per_cu != null: occupancy query instead of a launch.
template <class P>
hipError_t EnsPersistent(const EnsembleArgs &e, int recon, int pair, int grid, int *per_cu, hipStream_t s)
{
  if (recon == kPPMC) return EnsPersistentR<P, kPPMC>(e, pair, grid, per_cu, s);
  if (recon == kPLMC) return EnsPersistentR<P, kPLMC>(e, pair, grid, per_cu, s);
  return hipErrorInvalidValue;
}

}  // namespace chl
