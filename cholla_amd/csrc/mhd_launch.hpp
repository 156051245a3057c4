// mhd_launch.hpp -- host dispatch of the MHD step kernels.
#pragma once

#include <algorithm>

#include "mhd_kernels.hpp"
#include "support_kernels.hpp"

namespace chl
{

inline unsigned MhdBlocks(size_t n) { return unsigned((n + 255) / 256); }

// The three stage-2 flux launches of a reconstruction RK (1 PLMC, 2 PPMC) over the faces
// with every index in [lo, hi], in XCD-contiguous order: x faces in 64 x 4 face tiles, y and
// z faces in 64 x 4 columns that march kMarch faces along the direction.
template <class P, int RK>
void MhdFlux2Tiles(const mhd::MhdArgs &a, int lo, int hx, int hy, int hz, hipStream_t s)
{
  const unsigned ex = unsigned(hx - lo + 1), ey = unsigned(hy - lo + 1), ez = unsigned(hz - lo + 1);
  auto blocks = [](unsigned t) { return dim3((t + 7) / 8 * 8); };
  constexpr unsigned K = unsigned(mhd::kMarch);
  hipLaunchKernelGGL((mhd::MhdFlux2TileKernel<P, RK>), blocks((ex + 63) / 64 * ((ey + 3) / 4) * ez), dim3(256), 0,
                     s, a, lo, hx, hy, hz);
  hipLaunchKernelGGL((mhd::MhdFlux2MarchKernel<P, 1, RK>), blocks((ex + 63) / 64 * ((ez + 3) / 4) * ((ey + K - 1) / K)),
                     dim3(256), 0, s, a, lo, hx, hy, hz);
  hipLaunchKernelGGL((mhd::MhdFlux2MarchKernel<P, 2, RK>), blocks((ex + 63) / 64 * ((ey + 3) / 4) * ((ez + K - 1) / K)),
                     dim3(256), 0, s, a, lo, hx, hy, hz);
}

// CT edge fields fused with one stage's update (MhdCtUpdateKernel) over the cells [clo, ch*]
// and the edges [elo, eh*]: 32 x 8 tiles, z-chunks of 16 planes (on Orszag-Tang 256^3 these
// ran 2-5 % faster than 16 x 16 tiles or z-chunks of 8 and 32).
template <class P, bool FULL>
void MhdCtUpdate(const mhd::MhdArgs &a, int clo, int chx, int chy, int chz, int elo, int ehx, int ehy, int ehz,
                 hipStream_t s)
{
  constexpr int CTX = 32, CTY = 8, KCHUNK = 16;
  const unsigned tiles =
      unsigned(((chx - clo + CTX) / CTX) * ((chy - clo + CTY) / CTY) * ((chz - clo + KCHUNK) / KCHUNK));
  hipLaunchKernelGGL((mhd::MhdCtUpdateKernel<P, CTX, CTY, FULL>), dim3((tiles + 7) / 8 * 8), dim3(CTX * CTY), 0, s, a,
                     clo, chx, chy, chz, elo, ehx, ehy, ehz, KCHUNK);
}

// One VL step of MHD: stage-1 fluxes, edge fields + half update, stage-2 fluxes, edge
// fields + full update.  The ghost refresh is a separate launch; the inverse-dt reduction is
// folded into the stage-1 flux launch (kFlagFoldDt, device clock advanced right after it)
// or left to MhdCalcDtKernel.
template <class P>
hipError_t MhdStepT(const mhd::MhdArgs &a, hipStream_t s)
{
  const int nx = a.nx, ny = a.ny, nz = a.nz, ng = a.ng;
  // stage-1 faces i+1/2 of every cell with i <= n-2 (other indices over the whole grid)
  const mhd::MhdRegion m1 = mhd::MakeMhdSlab(0, nx - 1, ny - 1, nz - 1);
  if constexpr (P::kId == PolicyMixed::kId) {
    // z-march in 32 x 8 tiles of 8 planes (OT 256^3: 0.98 ms against 1.02 for the per-cell
    // launch; the FP64 forms run slower there, at 2 waves per SIMD against 3)
    constexpr int TX = 32, TY = 8, KZ = 8;
    const unsigned blocks = unsigned(((nx + TX - 1) / TX) * ((ny + TY - 1) / TY) * ((nz + KZ - 1) / KZ));
    hipLaunchKernelGGL((mhd::MhdFlux1MarchKernel<P, TX, TY>), dim3((blocks + 7) / 8 * 8), dim3(TX * TY), 0, s, a, KZ);
  } else {
    hipLaunchKernelGGL((mhd::MhdFlux1AllKernel<P>), dim3(mhd::MhdRegionBlocks(m1)), dim3(256), 0, s, a, m1);
  }
  if (a.flags & kFlagFoldDt) {
    const hipError_t e = LaunchDtAdvance(a.dti, a.clk_dt, a.clk_t, a.clk_tout, a.clk_cfl, a.clk_nstep, s, true);
    if (e != hipSuccess) return e;
  }
  // edges [2, n-1], half update of the cells [1, n-2]
  MhdCtUpdate<P, false>(a, 1, nx - 2, ny - 2, nz - 2, 2, nx - 1, ny - 1, nz - 1, s);
  {  // stage-2 faces with every index in [ng - 1, n - ng]
    const int lo = ng - 1, hx = nx - ng, hy = ny - ng, hz = nz - ng;
    if (a.recon == 0) {
      const mhd::MhdRegion m2 = mhd::MakeMhdSlab(lo, hx, hy, hz);
      hipLaunchKernelGGL((mhd::MhdFlux2PcmKernel<P>), dim3(mhd::MhdRegionBlocks(m2)), dim3(256), 0, s, a, m2);
    } else if (a.recon == 2) {
      MhdFlux2Tiles<P, 2>(a, lo, hx, hy, hz, s);
    } else {
      MhdFlux2Tiles<P, 1>(a, lo, hx, hy, hz, s);
    }
  }
  // edges [ng, n-ng], full update of the real cells and of the face fields just below them
  MhdCtUpdate<P, true>(a, ng, nx - ng - 1, ny - ng - 1, nz - ng - 1, ng, nx - ng, ny - ng, nz - ng, s);
  return hipGetLastError();
}

hipError_t LaunchMhdStepExact(const mhd::MhdArgs &a, hipStream_t s);

}  // namespace chl
