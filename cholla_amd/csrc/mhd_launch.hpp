// mhd_launch.hpp -- host dispatch of the MHD step kernels.
#pragma once

#include <algorithm>

#include "mhd_kernels.hpp"

namespace chl
{

inline unsigned MhdBlocks(size_t n) { return unsigned((n + 255) / 256); }

// The three stage-2 flux launches of a reconstruction RK (1 PLMC, 2 PPMC) over the faces
// with every index in [lo, hi]: 64 x 4 face tiles in XCD-contiguous order.
template <class P, int RK>
void MhdFlux2Tiles(const mhd::MhdArgs &a, int lo, int hx, int hy, int hz, hipStream_t s)
{
  const unsigned ex = unsigned(hx - lo + 1), ey = unsigned(hy - lo + 1), ez = unsigned(hz - lo + 1);
  auto blocks = [](unsigned t) { return dim3((t + 7) / 8 * 8); };
  hipLaunchKernelGGL((mhd::MhdFlux2TileKernel<P, 0, RK>), blocks((ex + 63) / 64 * ((ey + 3) / 4) * ez), dim3(256), 0,
                     s, a, lo, hx, hy, hz);
  hipLaunchKernelGGL((mhd::MhdFlux2TileKernel<P, 1, RK>), blocks((ex + 63) / 64 * ((ey + 3) / 4) * ez), dim3(256), 0,
                     s, a, lo, hx, hy, hz);
  hipLaunchKernelGGL((mhd::MhdFlux2TileKernel<P, 2, RK>), blocks((ex + 63) / 64 * ((ez + 3) / 4) * ey), dim3(256), 0,
                     s, a, lo, hx, hy, hz);
}

// One VL step of MHD: stage-1 fluxes, edge fields, half update, stage-2 fluxes, edge
// fields, full update.  The inverse-dt reduction and ghost refresh are separate launches.
template <class P>
hipError_t MhdStepT(const mhd::MhdArgs &a, hipStream_t s)
{
  using T     = typename P::T;
  const T *Fx = reinterpret_cast<const T *>(a.F[0]), *Fy = reinterpret_cast<const T *>(a.F[1]),
          *Fz = reinterpret_cast<const T *>(a.F[2]);
  const int nx = a.nx, ny = a.ny, nz = a.nz, ng = a.ng;
  constexpr int CTX = 32, CTY = 8, CTZ = 32;  // edge tile and z chunk of the CT launches
  auto ct = [&](const double *Us, int lo, int hx, int hy, int hz) {
    const unsigned tiles = unsigned(((hx - lo + CTX) / CTX) * ((hy - lo + CTY) / CTY) * ((hz - lo + CTZ) / CTZ));
    hipLaunchKernelGGL((mhd::MhdCtTileKernel<T, CTX, CTY>), dim3((tiles + 7) / 8 * 8), dim3(CTX * CTY), 0, s, Fx, Fy,
                       Fz, Us, a.E, nx, ny, nz, lo, hx, hy, hz, CTZ, a.dt_dev);
  };
  // stage-1 faces i+1/2 of every cell with i <= n-2 (other indices over the whole grid)
  const mhd::MhdRegion m1 = mhd::MakeMhdSlab(0, nx - 1, ny - 1, nz - 1);
  hipLaunchKernelGGL((mhd::MhdFlux1AllKernel<P>), dim3(mhd::MhdRegionBlocks(m1)), dim3(256), 0, s, a, m1);
  ct(a.U, 2, nx - 1, ny - 1, nz - 1);  // edges [2, n-1]
  const mhd::MhdRegion mh = mhd::MakeMhdSlab(1, nx - 2, ny - 2, nz - 2);
  hipLaunchKernelGGL((mhd::MhdHalfKernel<P>), dim3(mhd::MhdRegionBlocks(mh)), dim3(256), 0, s, a, mh);
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
  ct(a.Uh, ng, nx - ng, ny - ng, nz - ng);  // edges [ng, n-ng]
  const mhd::MhdRegion mf = mhd::MakeMhdSlab(ng - 1, nx - ng - 1, ny - ng - 1, nz - ng - 1);
  hipLaunchKernelGGL((mhd::MhdFullKernel<P>), dim3(mhd::MhdRegionBlocks(mf)), dim3(256), 0, s, a, mf);
  return hipGetLastError();
}

hipError_t LaunchMhdStepExact(const mhd::MhdArgs &a, hipStream_t s);

}  // namespace chl
