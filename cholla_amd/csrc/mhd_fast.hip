// mhd_fast.hip -- PolicyFast (FP64) and PolicyMixed (FP32 reconstruction and HLLD, FP64
// state, accumulation, edge fields and dt) instantiations of the MHD step, and the
// policy dispatch.  Built with -ffp-contract=on (FMA only within one expression).
#include "mhd_launch.hpp"

namespace chl
{
hipError_t LaunchMhdStep(int policy, const mhd::MhdArgs &a, hipStream_t s)
{
  if (policy == PolicyExact::kId) return LaunchMhdStepExact(a, s);
  if (policy == PolicyMixed::kId) return MhdStepT<PolicyMixed>(a, s);
  return MhdStepT<PolicyFast>(a, s);
}
}  // namespace chl
