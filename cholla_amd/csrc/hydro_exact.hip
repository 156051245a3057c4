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
}  // namespace chl
