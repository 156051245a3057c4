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

// The side buffer (RepairSideIndex): the depth-1 ghost layer's densities on the six faces,
// copied before stage 2 rewrites the ghosts.  One lane per face cell.
__global__ void RepairSideKernel(const double *U, double *side, int nx, int ny, int nz, int ng)
{
  const int Lx = nx - 2 * ng, Ly = ny - 2 * ng, Lz = nz - 2 * ng;
  const long fx = long(Ly) * Lz, fy = long(Lx) * Lz, fz = long(Lx) * Ly;
  long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
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

hipError_t LaunchRepairSide(const StepArgs &a, hipStream_t s)
{
  const long Lx = a.nx - 2 * a.ng, Ly = a.ny - 2 * a.ng, Lz = a.nz - 2 * a.ng;
  const long n  = 2 * (Ly * Lz + Lx * Lz + Lx * Ly);
  hipLaunchKernelGGL(RepairSideKernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, a.U, a.rep.side, a.nx, a.ny,
                     a.nz, a.ng);
  return hipGetLastError();
}
}  // namespace chl
