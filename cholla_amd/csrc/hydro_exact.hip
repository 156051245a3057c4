// hydro_exact.hip -- PolicyExact instantiations of the fused VL kernels.
// Built with -ffp-contract=off: every product and sum rounds as the reference
// source writes it, so results match a strict IEEE restatement bit for bit.
#include "vl3d_launch.hpp"

namespace chl
{
hipError_t LaunchStage1Exact(const StepArgs &a, int R1, int tile, hipStream_t s)
{
  return LaunchStage1<PolicyExact>(a, R1, tile, s);
}
hipError_t LaunchStage2Exact(const StepArgs &a, int recon, int tile, hipStream_t s)
{
  return LaunchStage2<PolicyExact>(a, recon, tile, s);
}
hipError_t LaunchRepairExact(const StepArgs &a, hipStream_t s) { return LaunchRepair<PolicyExact>(a, s); }

// The side buffer of a density repair next to fused periodic images: one lane per face cell.
__global__ void RepairSideKernel(const double *U, double *side, int nx, int ny, int nz, int ng)
{
  RepairSideCell(U, side, nx, ny, nz, ng, long(blockIdx.x) * blockDim.x + threadIdx.x);
}

hipError_t LaunchRepairSide(const StepArgs &a, hipStream_t s)
{
  const long Lx = a.nx - 2 * a.ng, Ly = a.ny - 2 * a.ng, Lz = a.nz - 2 * a.ng;
  const long n  = 2 * (Ly * Lz + Lx * Lz + Lx * Ly);
  hipLaunchKernelGGL(RepairSideKernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, a.U, a.rep.side, a.nx, a.ny,
                     a.nz, a.ng);
  return hipGetLastError();
}
}  // namespace chl
