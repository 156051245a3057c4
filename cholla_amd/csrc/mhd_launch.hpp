// mhd_launch.hpp -- host dispatch of the MHD step kernels.
#pragma once

#include <algorithm>

#include "mhd_kernels.hpp"

namespace chl
{

inline unsigned MhdBlocks(size_t n) { return unsigned((n + 255) / 256); }

// One VL step of MHD: stage-1 fluxes, edge fields, half update, stage-2 fluxes, edge
// fields, full update.  The inverse-dt reduction and ghost refresh are separate launches.
template <class P>
hipError_t MhdStepT(const mhd::MhdArgs &a, hipStream_t s)
{
  using T          = typename P::T;
  const size_t n   = size_t(a.nx) * a.ny * a.nz;
  const unsigned g = MhdBlocks(n);
  const T *Fx = reinterpret_cast<const T *>(a.F[0]), *Fy = reinterpret_cast<const T *>(a.F[1]),
          *Fz = reinterpret_cast<const T *>(a.F[2]);
  const int ng = a.ng;
  hipLaunchKernelGGL((mhd::MhdFlux1Kernel<P, 0>), dim3(g), dim3(256), 0, s, a);
  hipLaunchKernelGGL((mhd::MhdFlux1Kernel<P, 1>), dim3(g), dim3(256), 0, s, a);
  hipLaunchKernelGGL((mhd::MhdFlux1Kernel<P, 2>), dim3(g), dim3(256), 0, s, a);
  hipLaunchKernelGGL(mhd::MhdCellEmfKernel<0>, dim3(g), dim3(256), 0, s, a.U, a.Ec, a.nx, a.ny, a.nz, 1, a.nx - 1,
                     a.ny - 1, a.nz - 1, a.dt_dev);
  hipLaunchKernelGGL((mhd::MhdEmfKernel<T>), dim3(g), dim3(256), 0, s, Fx, Fy, Fz, a.U, a.E, a.nx, a.ny, a.nz, 2,
                     a.nx - 1, a.ny - 1, a.nz - 1, a.dt_dev, a.Ec);
  hipLaunchKernelGGL((mhd::MhdHalfKernel<P>), dim3(g), dim3(256), 0, s, a);
  {  // stage-2 faces with every index in [ng - 1, n - ng]
    const int lo = ng - 1, hx = a.nx - ng, hy = a.ny - ng, hz = a.nz - ng;
    const unsigned ex = unsigned(hx - lo + 1), ey = unsigned(hy - lo + 1), ez = unsigned(hz - lo + 1);
    hipLaunchKernelGGL((mhd::MhdFlux2TileKernel<P, 0>), dim3((ex + 63) / 64, (ey + 3) / 4, ez), dim3(256), 0, s, a,
                       lo, hx, hy, hz);
    hipLaunchKernelGGL((mhd::MhdFlux2TileKernel<P, 1>), dim3((ex + 63) / 64, (ey + 3) / 4, ez), dim3(256), 0, s, a,
                       lo, hx, hy, hz);
    hipLaunchKernelGGL((mhd::MhdFlux2TileKernel<P, 2>), dim3((ex + 63) / 64, (ez + 3) / 4, ey), dim3(256), 0, s, a,
                       lo, hx, hy, hz);
  }
  hipLaunchKernelGGL(mhd::MhdCellEmfKernel<0>, dim3(g), dim3(256), 0, s, a.Uh, a.Ec, a.nx, a.ny, a.nz, ng - 1,
                     a.nx - ng, a.ny - ng, a.nz - ng, a.dt_dev);
  hipLaunchKernelGGL((mhd::MhdEmfKernel<T>), dim3(g), dim3(256), 0, s, Fx, Fy, Fz, a.Uh, a.E, a.nx, a.ny, a.nz, ng,
                     a.nx - ng, a.ny - ng, a.nz - ng, a.dt_dev, a.Ec);
  hipLaunchKernelGGL((mhd::MhdFullKernel<P>), dim3(g), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t LaunchMhdStepExact(const mhd::MhdArgs &a, hipStream_t s);

}  // namespace chl
