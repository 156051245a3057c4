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
hipError_t LaunchEnsembleDe(int phase, const EnsembleArgs &e, int tile, hipStream_t s)
{
  if (phase == 0) return EnsDe<0>(e, tile, s);
  if (phase == 1) return EnsDe<1>(e, tile, s);
  return hipErrorInvalidValue;
}
}  // namespace chl
