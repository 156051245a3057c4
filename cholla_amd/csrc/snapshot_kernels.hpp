// snapshot_kernels.hpp -- launchers of the grid <-> snapshot-layout copies (snapshot_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace chl
{

// One field's real region: grid totals (nx, ny, nz), dataset extents (nxr, nyr, nzr) and the
// grid index (ox, oy, oz) of the dataset's first element.
struct SnapshotShape {
  int nx, ny, nz;
  int nxr, nyr, nzr;
  int ox, oy, oz;

  // every element the copy touches lies inside the grid, and the launch grid fits
  bool Valid() const
  {
    return nx > 0 && ny > 0 && nz > 0 && nxr > 0 && nyr > 0 && nzr > 0 && ox >= 0 && oy >= 0 && oz >= 0 &&
           ox + nxr <= nx && oy + nyr <= ny && oz + nzr <= nz && nyr <= 65535;
  }
};

hipError_t LaunchSnapshotPack(const double *grid, double *buf, const SnapshotShape &s, hipStream_t st);
hipError_t LaunchSnapshotPack32(const double *grid, float *buf, const SnapshotShape &s, hipStream_t st);
hipError_t LaunchSnapshotUnpack(const double *buf, double *grid, const SnapshotShape &s, hipStream_t st);

}  // namespace chl
