// hydro_fast.hip -- PolicyFast (FP64) and PolicyMixed (FP32 reconstruction and
// Riemann solve, FP64 state/accumulation/pressure recovery/dt) instantiations.
// Built with -ffp-contract=on (FMA only within one expression).
#include "vl3d_launch.hpp"

namespace chl
{
hipError_t LaunchStage1Fast(int policy, const StepArgs &a, int R1, int tile, hipStream_t s)
{
  if (policy == PolicyMixed::kId) return LaunchStage1<PolicyMixed>(a, R1, tile, s);
  return LaunchStage1<PolicyFast>(a, R1, tile, s);
}
hipError_t LaunchStage2Fast(int policy, const StepArgs &a, int recon, int tile, hipStream_t s)
{
  if (policy == PolicyMixed::kId) return LaunchStage2<PolicyMixed>(a, recon, tile, s);
  return LaunchStage2<PolicyFast>(a, recon, tile, s);
}
hipError_t LaunchRepairFast(int policy, const StepArgs &a, hipStream_t s)
{
  if (policy == PolicyMixed::kId) return LaunchRepair<PolicyMixed>(a, s);
  return LaunchRepair<PolicyFast>(a, s);
}
}  // namespace chl
