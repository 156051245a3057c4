// snapshot_kernels.hip -- real cells of one field between Cholla's grid layout and the
// snapshot layout of its HDF5 files.
//
// The grid is x-fastest with ghost cells, id = i + j*nx + k*nx*ny; a snapshot dataset of
// real cells is z-fastest, buf = k + j*nz_real + i*nz_real*ny_real (CopyReal3D_GPU_Kernel,
// src/io/io_gpu.cu:44-60; Fill_HDF5_Buffer_From_Grid_CPU and Fill_Grid_From_HDF5_Buffer,
// src/io/io.cpp:1232-1268 and 2450-2490).  Between the two the fastest index changes from x
// to z, so each workgroup moves a 32 (x) by 32 (z) tile of one y row through LDS: the grid
// side reads (or writes) rows of 32 consecutive x and the snapshot side writes (or reads)
// rows of 32 consecutive z, both as whole 256-byte lines.
#include <hip/hip_runtime.h>

#include <cstddef>

#include "snapshot_kernels.hpp"

namespace chl
{
namespace
{

constexpr int kT = 32;  // tile edge (x and z)
constexpr int kR = 8;   // rows of threads: 32 x 8 = 256 threads

// PACK: grid -> buf (converted to OUT); !PACK: buf -> grid.  (ox, oy, oz) is the grid index of
// the dataset's first element: n_ghost, or n_ghost - 1 along a magnetic field's own axis
// (its dataset holds one more face than there are real cells).
template <bool PACK, typename OUT>
__global__ void __launch_bounds__(kT *kR) SnapshotTileKernel(double *grid, OUT *buf, int nx, int ny, int nxr, int nyr,
                                                            int nzr, int ox, int oy, int oz)
{
  __shared__ double t[kT][kT + 1];
  const int i0 = blockIdx.x * kT, k0 = blockIdx.y * kT, j = blockIdx.z;
  const int tx = threadIdx.x % kT, ty = threadIdx.x / kT;
  const size_t plane = size_t(nx) * ny;
  auto gid = [&](int i, int k) { return size_t(i + ox) + size_t(j + oy) * nx + size_t(k + oz) * plane; };
  auto bid = [&](int i, int k) { return size_t(k) + size_t(j) * nzr + size_t(i) * nzr * nyr; };
  if constexpr (PACK) {
    for (int kk = ty; kk < kT; kk += kR) {  // grid rows: consecutive x
      const int i = i0 + tx, k = k0 + kk;
      if (i < nxr && k < nzr) t[kk][tx] = grid[gid(i, k)];
    }
    __syncthreads();
    for (int ii = ty; ii < kT; ii += kR) {  // snapshot rows: consecutive z
      const int i = i0 + ii, k = k0 + tx;
      if (i < nxr && k < nzr) buf[bid(i, k)] = OUT(t[tx][ii]);
    }
  } else {
    for (int ii = ty; ii < kT; ii += kR) {
      const int i = i0 + ii, k = k0 + tx;
      if (i < nxr && k < nzr) t[tx][ii] = double(buf[bid(i, k)]);
    }
    __syncthreads();
    for (int kk = ty; kk < kT; kk += kR) {
      const int i = i0 + tx, k = k0 + kk;
      if (i < nxr && k < nzr) grid[gid(i, k)] = t[kk][tx];
    }
  }
}

template <bool PACK, typename OUT>
hipError_t Launch(double *grid, OUT *buf, const SnapshotShape &s, hipStream_t st)
{
  dim3 g((s.nxr + kT - 1) / kT, (s.nzr + kT - 1) / kT, s.nyr);
  hipLaunchKernelGGL((SnapshotTileKernel<PACK, OUT>), g, dim3(kT * kR), 0, st, grid, buf, s.nx, s.ny, s.nxr, s.nyr,
                     s.nzr, s.ox, s.oy, s.oz);
  return hipGetLastError();
}

}  // namespace

hipError_t LaunchSnapshotPack(const double *grid, double *buf, const SnapshotShape &s, hipStream_t st)
{
  return Launch<true, double>(const_cast<double *>(grid), buf, s, st);
}

hipError_t LaunchSnapshotPack32(const double *grid, float *buf, const SnapshotShape &s, hipStream_t st)
{
  return Launch<true, float>(const_cast<double *>(grid), buf, s, st);
}

hipError_t LaunchSnapshotUnpack(const double *buf, double *grid, const SnapshotShape &s, hipStream_t st)
{
  return Launch<false, double>(grid, const_cast<double *>(buf), s, st);
}

}  // namespace chl
