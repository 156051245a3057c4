// ensemble_exact.hip -- PolicyExact instantiations of the ensemble kernels
// (built with -ffp-contract=off, like hydro_exact.hip).
#include "ensemble_launch.hpp"

namespace chl
{
hipError_t LaunchEnsembleStage1Exact(const EnsembleArgs &e, int tile, hipStream_t s)
{
  return EnsStage1<PolicyExact>(e, tile, s);
}
hipError_t LaunchEnsembleStage2Exact(const EnsembleArgs &e, int recon, int tile, hipStream_t s)
{
  return EnsStage2<PolicyExact>(e, recon, tile, s);
}
hipError_t LaunchEnsemblePersistentExact(const EnsembleArgs &e, int recon, int pair, int grid, int *per_cu,
                                         hipStream_t s)
{
  return EnsPersistent<PolicyExact>(e, recon, pair, grid, per_cu, s);
}
hipError_t LaunchEnsembleRepairExact(const EnsembleArgs &e, hipStream_t s) { return EnsRepair<PolicyExact>(e, s); }

__global__ void EnsembleRepairSideKernel(EnsembleArgs e)
{
  const BoxDesc &d = e.box[blockIdx.y];
  if (!d.rep.side) return;
  RepairSideCell(d.U, d.rep.side, d.nx, d.ny, d.nz, e.ng, long(blockIdx.x) * blockDim.x + threadIdx.x);
}

// max_face_cells: the largest 2 (Ly Lz + Lx Lz + Lx Ly) over the boxes.
hipError_t LaunchEnsembleRepairSide(const EnsembleArgs &e, long max_face_cells, hipStream_t s)
{
  if (e.n_boxes == 0 || max_face_cells == 0) return hipSuccess;
  hipLaunchKernelGGL(EnsembleRepairSideKernel, dim3(unsigned((max_face_cells + 255) / 256), unsigned(e.n_boxes)),
                     dim3(256), 0, s, e);
  return hipGetLastError();
}

hipError_t LaunchEnsembleDe(int phase, const EnsembleArgs &e, int tile, hipStream_t s)
{
  if (phase == 0) return EnsDe<0>(e, tile, s);
  if (phase == 1) return EnsDe<1>(e, tile, s);
  return hipErrorInvalidValue;
}
}  // namespace chl
