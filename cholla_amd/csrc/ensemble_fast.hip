// ensemble_fast.hip -- PolicyFast and PolicyMixed instantiations of the ensemble
// kernels (built with -ffp-contract=on, like hydro_fast.hip), and the policy dispatch.
#include "ensemble_launch.hpp"

namespace chl
{
hipError_t LaunchEnsembleStage1Exact(const EnsembleArgs &e, int tile, hipStream_t s);
hipError_t LaunchEnsembleStage2Exact(const EnsembleArgs &e, int recon, int tile, hipStream_t s);
hipError_t LaunchEnsemblePersistentExact(const EnsembleArgs &e, int recon, int pair, int grid, int *per_cu,
                                         hipStream_t s);

hipError_t LaunchEnsembleStage1(int policy, const EnsembleArgs &e, int tile, hipStream_t s)
{
  if (policy == PolicyExact::kId) return LaunchEnsembleStage1Exact(e, tile, s);
  if (policy == PolicyMixed::kId) return EnsStage1<PolicyMixed>(e, tile, s);
  return EnsStage1<PolicyFast>(e, tile, s);
}

hipError_t LaunchEnsembleStage2(int policy, const EnsembleArgs &e, int recon, int tile, hipStream_t s)
{
  if (policy == PolicyExact::kId) return LaunchEnsembleStage2Exact(e, recon, tile, s);
  if (policy == PolicyMixed::kId) return EnsStage2<PolicyMixed>(e, recon, tile, s);
  return EnsStage2<PolicyFast>(e, recon, tile, s);
}

hipError_t LaunchEnsemblePersistent(int policy, const EnsembleArgs &e, int recon, int pair, int grid, hipStream_t s)
{
  if (policy == PolicyExact::kId) return LaunchEnsemblePersistentExact(e, recon, pair, grid, nullptr, s);
  if (policy == PolicyMixed::kId) return EnsPersistent<PolicyMixed>(e, recon, pair, grid, nullptr, s);
  return EnsPersistent<PolicyFast>(e, recon, pair, grid, nullptr, s);
}

hipError_t LaunchEnsembleRepairExact(const EnsembleArgs &e, hipStream_t s);
hipError_t LaunchEnsembleRepair(int policy, const EnsembleArgs &e, hipStream_t s)
{
  if (policy == PolicyExact::kId) return LaunchEnsembleRepairExact(e, s);
  if (policy == PolicyMixed::kId) return EnsRepair<PolicyMixed>(e, s);
  return EnsRepair<PolicyFast>(e, s);
}

hipError_t EnsemblePersistentSlots(int policy, int recon, int pair, int de, int *per_cu)
{
  EnsembleArgs none{};
  none.de = de;
  if (policy == PolicyExact::kId) return LaunchEnsemblePersistentExact(none, recon, pair, 0, per_cu, nullptr);
  if (policy == PolicyMixed::kId) return EnsPersistent<PolicyMixed>(none, recon, pair, 0, per_cu, nullptr);
  return EnsPersistent<PolicyFast>(none, recon, pair, 0, per_cu, nullptr);
}
}  // namespace chl
