// noalloc.hip -- one VL step of a 3D box without allocating device memory.
//
// chl_hydro_vl_step_noalloc is chl_hydro_vl_step for callers that cannot allocate: code
// under HIP graph capture, serving loops, harnesses that forbid hipMalloc.  The integrator
// scratch that chl_hydro_create allocates (the half-step primitives W, dual energy's pdV
// sources, the device scalars and the repair marks) lives instead in `scratch`, a fixed
// device array of this translation unit: every GPU the library runs on holds one copy of
// kScratchBytes until the process ends.  The step runs the same kernels as
// chl_hydro_vl_step on the caller's stream and returns without synchronizing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "capi_internal.hpp"
#include "vl3d_launch.hpp"

namespace
{
constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
constexpr int kMaxDevices      = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Zero n bytes at p (n a multiple of 16, p 16-byte aligned): the repair counts and marks
// must start at zero, and the region may hold another box's data from an earlier call.
// A kernel, not hipMemsetAsync, so the step can be captured into a graph.
__global__ void ZeroFillKernel(uint4 *p, size_t n16)
{
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += size_t(gridDim.x) * blockDim.x)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
}

std::mutex g_mu;
void *g_scratch[kMaxDevices];  // device address of `scratch` on each device, once looked up

// The device address of this device's `scratch` (looked up on the device's first call:
// in host code the array's name is not a device address).
int ScratchAddress(void **out)
{
  int dev = 0;
  CHL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return chl_capi::Fail(CHL_ERR_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return chl_capi::Fail(CHL_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return CHL_OK;
}
}  // namespace

extern "C" size_t chl_noalloc_scratch_bytes(void) { return kScratchBytes; }

extern "C" int chl_hydro_vl_step_noalloc(const chl_hydro_config *cfg, double *d_conserved, double dx, double dy,
                                         double dz, double dt, double gamma, double density_floor, void *stream)
{
  if (!cfg || !d_conserved) return chl_capi::Fail(CHL_ERR_INVALID, "null argument");
  chl_hydro_config c = *cfg;
  CHL_TRY(chl_capi::PrepareConfig(c));
  if (c.ny == 1 || c.nz == 1) return chl_capi::Fail(CHL_ERR_UNSUPPORTED, "the allocation-free step is for 3D boxes");
  // the scratch layout: repair state | device scalars | W (+ pdV sources)
  size_t w_main      = 0;
  const size_t rep   = chl_capi::RepairBytes(c);  // a multiple of 256
  const size_t w     = chl_capi::HalfStepBytes(c, &w_main);
  const size_t bytes = rep + 256 + w;
  if (bytes > kScratchBytes)
    return chl_capi::Fail(CHL_ERR_INVALID, "box needs " + std::to_string(bytes) + " bytes of scratch; the allocation-free "
                                           "step holds " + std::to_string(kScratchBytes));
  void *base = nullptr;
  CHL_TRY(ScratchAddress(&base));
  unsigned char *b = static_cast<unsigned char *>(base);
  const hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(ZeroFillKernel, dim3(unsigned(std::min<size_t>((rep / 16 + 255) / 256, 1024))), dim3(256), 0, s,
                     reinterpret_cast<uint4 *>(b), rep / 16);
  CHL_HIP(hipGetLastError());
  chl_hydro h;  // a handle over the scratch, on the stack: nothing is allocated
  h.cfg     = c;
  h.n_cells = size_t(c.nx) * c.ny * c.nz;
  chl_capi::PlaceRepair(&h, b);
  h.scal    = b + rep;
  h.W       = b + rep + 256;
  h.W_bytes = w;
  h.pdv     = c.dual_energy ? reinterpret_cast<double *>(b + rep + 256 + w_main) : nullptr;
  h.stream  = s;
  return chl_capi::EnqueueVlStep(&h, d_conserved, dx, dy, dz, dt, gamma, density_floor);
}
