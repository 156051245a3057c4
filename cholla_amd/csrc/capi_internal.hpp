// capi_internal.hpp -- helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/cholla_amd.h"

namespace chl_capi
{
// Records msg as the text of chl_last_error() (per thread) and returns code.
int Fail(int code, const std::string &msg);

constexpr double kCfl = 0.3;  // C_cfl, src/grid/grid3D.cpp:157

inline int ReconRadius(int recon)
{
  if (recon == CHL_RECON_PPMC || recon == CHL_RECON_PPMP) return 2;
  return (recon == CHL_RECON_PLMC || recon == CHL_RECON_PLMP) ? 1 : 0;
}
}  // namespace chl_capi

#define CHL_HIP(expr)                                                                                          \
  do {                                                                                                         \
    hipError_t e_ = (expr);                                                                                    \
    if (e_ != hipSuccess) return chl_capi::Fail(CHL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CHL_TRY(expr)              \
  do {                             \
    int rc_ = (expr);              \
    if (rc_ != CHL_OK) return rc_; \
  } while (0)
