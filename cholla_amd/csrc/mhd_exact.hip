// mhd_exact.hip -- PolicyExact instantiations of the MHD step, the MHD inverse-dt and
// divergence reductions, and the reference-shaped MHD stage kernels used to check the
// device physics against the reference's unit vectors.  Built with -ffp-contract=off.
#include "mhd_launch.hpp"

namespace chl
{
namespace
{

using mhd::BlockMaxAtomic;

// Calc_dt_3D with MHD (src/hydro/hydro_cuda.cu:514-560) over the real cells, marching them
// (MakeMhdMarch; the maximum is independent of the order the cells are visited in)
__global__ void __launch_bounds__(256) MhdCalcDtKernel(const double *U, int nx, int ny, int nz, double dx,
                                                       double dy, double dz, double g, mhd::MhdRegion mr,
                                                       unsigned long long *dti)
{
  const size_t n = size_t(nx) * ny * nz;
  int x, y, z0, z1;
  double m = 0.0;
  if (mhd::MarchCell(mr, x, y, z0, z1)) {  // every lane reaches the workgroup reduction
    for (int z = z0; z < z1; z++) {
      const size_t id = size_t(x) + size_t(y) * nx + size_t(z) * size_t(nx) * ny;
      m               = Fmax(m, mhd::MhdInverseCrossingTime(U, id, x, y, z, n, nx, ny, dx, dy, dz, g));
    }
  }
  BlockMaxAtomic(m, dti);
}

// mhd::calculateMagneticDivergence (src/mhd/magnetic_divergence.cu:33-75)
__global__ void __launch_bounds__(256) MhdDivergenceKernel(const double *U, int nx, int ny, int nz, double dx,
                                                           double dy, double dz, unsigned long long *out)
{
  const size_t n  = size_t(nx) * ny * nz;
  const double *B = U + mhd::kBx * n;
  double m        = 0.0;
  for (size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x; id < n; id += size_t(blockDim.x) * gridDim.x) {
    const int x = int(id % nx), y = int((id / nx) % ny), z = int(id / (size_t(nx) * ny));
    if (x > 1 && y > 1 && z > 1) {
      const double div = ((B[id] - B[id - 1]) / dx) + ((B[n + id] - B[n + id - size_t(nx)]) / dy) +
                         ((B[2 * n + id] - B[2 * n + id - size_t(nx) * ny]) / dz);
      m = Fmax(m, Fabs(div));
    }
  }
  BlockMaxAtomic(m, out);
}

// Calculate_HLLD_Fluxes_CUDA over interface arrays (hlld_cuda.cu:31-164)
__global__ void __launch_bounds__(256) HlldGridKernel(const double *QL, const double *QR, const double *magX,
                                                      double *F, long n, double g, int dir)
{
  const long id = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= n) return;
  auto load = [&](const double *Q) {
    mhd::ConsM<double> c;
    c.d = Q[id];
    mhd::RotM(dir, Q[n + id], Q[2 * n + id], Q[3 * n + id], c.mx, c.my, c.mz);
    c.E = Q[4 * n + id], c.by = Q[5 * n + id], c.bz = Q[6 * n + id];
    return c;
  };
  mhd::ConsM<double> f;
  if (mhd::HlldM<PolicyExact>(load(QL), load(QR), magX[id], g, f)) mhd::StoreFluxM(F, size_t(id), size_t(n), dir, f);
}

// reconstruction::Write_Data with MHD (reconstruction.h:918-970) into interface arrays: the
// conserved state at index at, momenta in their global slots, the face field of at for the
// energy.
__device__ __forceinline__ void WriteGridState(const double *U, const mhd::PrimM<double> &s, double *Q, size_t at,
                                               size_t n, int dir, double g)
{
  mhd::ConsM<double> c = mhd::WriteDataM<PolicyExact>(s, U[(mhd::kBx + dir) * n + at], g);
  double gx, gy, gz;
  if (dir == 0) {
    gx = c.mx, gy = c.my, gz = c.mz;
  } else if (dir == 1) {
    gy = c.mx, gz = c.my, gx = c.mz;
  } else {
    gz = c.mx, gx = c.my, gy = c.mz;
  }
  Q[at] = c.d, Q[n + at] = gx, Q[2 * n + at] = gy, Q[3 * n + at] = gz, Q[4 * n + at] = c.E;
  Q[5 * n + at] = c.by, Q[6 * n + at] = c.bz;
}

// PLMC_cuda with MHD over a grid (plmc_cuda.cu:24-300): Q_L at i, Q_R at i-1
__global__ void __launch_bounds__(256) MhdPlmcGridKernel(const double *U, double *QL, double *QR, int nx, int ny,
                                                         int nz, double g, int dir)
{
  const size_t n  = size_t(nx) * ny * nz;
  const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const int x = int(id % nx), y = int((id / nx) % ny), z = int(id / (size_t(nx) * ny));
  // reconstruction::Thread_Guard<2> (reconstruction.h:95-119)
  bool out = x < 1 || x >= nx - 2;
  if (ny > 1) out = y < 1 || y >= ny - 2 || out;
  if (nz > 1) {
    out = z < 1 || z >= nz - 2 || out;
  } else {
    out = z >= nz || out;
  }
  if (out) return;
  mhd::PrimM<double> Lp, Rm;
  mhd::PlmcCellM<PolicyExact>(U, x, y, z, nx, ny, n, dir, g, Lp, Rm);
  const size_t step = dir == 0 ? 1 : (dir == 1 ? size_t(nx) : size_t(nx) * ny);
  WriteGridState(U, Lp, QL, id, n, dir, g);
  WriteGridState(U, Rm, QR, id - step, n, dir, g);
}

// PPMC_VL with MHD over a grid (ppmc_cuda.cu:540-695): Q_L at i, Q_R at i-1
__global__ void __launch_bounds__(256) MhdPpmcGridKernel(const double *U, double *QL, double *QR, int nx, int ny,
                                                         int nz, double g, int dir)
{
  const size_t n  = size_t(nx) * ny * nz;
  const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const int x = int(id % nx), y = int((id / nx) % ny), z = int(id / (size_t(nx) * ny));
  // reconstruction::Thread_Guard<3> (reconstruction.h:95-119)
  bool out = x < 2 || x >= nx - 3;
  if (ny > 1) out = y < 2 || y >= ny - 3 || out;
  if (nz > 1) {
    out = z < 2 || z >= nz - 3 || out;
  } else {
    out = z >= nz || out;
  }
  if (out) return;
  mhd::PrimM<double> Lp, Rm;
  mhd::PpmcCellM<PolicyExact>(U, x, y, z, nx, ny, n, dir, g, Lp, Rm);
  const size_t step = dir == 0 ? 1 : (dir == 1 ? size_t(nx) : size_t(nx) * ny);
  WriteGridState(U, Lp, QL, id, n, dir, g);
  WriteGridState(U, Rm, QR, id - step, n, dir, g);
}

// mhd::Update_Magnetic_Field_3D over a grid (magnetic_update.cu:21-81)
__global__ void __launch_bounds__(256) MagneticUpdateGridKernel(const double *src, double *dst, const double *E,
                                                                int nx, int ny, int nz, double dt, double dx,
                                                                double dy, double dz)
{
  const size_t n  = size_t(nx) * ny * nz;
  const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const int x = int(id % nx), y = int((id / nx) % ny), z = int(id / (size_t(nx) * ny));
  if (!(x > 0 && y > 0 && z > 0 && x < nx - 1 && y < ny - 1 && z < nz - 1)) return;
  const double dtodx = dt / dx, dtody = dt / dy, dtodz = dt / dz;
  const size_t sx = 1, sy = size_t(nx), sz = size_t(nx) * ny;
  const double ex1 = E[id + sy], ex2 = E[id + sz], ex3 = E[id + sy + sz];
  const double ey1 = E[n + id + sx], ey2 = E[n + id + sz], ey3 = E[n + id + sx + sz];
  const double ez1 = E[2 * n + id + sx], ez2 = E[2 * n + id + sy], ez3 = E[2 * n + id + sx + sy];
  dst[5 * n + id] = src[5 * n + id] + dtodz * (ey3 - ey1) + dtody * (ez1 - ez3);
  dst[6 * n + id] = src[6 * n + id] + dtodx * (ez3 - ez2) + dtodz * (ex1 - ex3);
  dst[7 * n + id] = src[7 * n + id] + dtody * (ex3 - ex2) + dtodx * (ey2 - ey3);
}

}  // namespace

hipError_t LaunchMhdStepExact(const mhd::MhdArgs &a, hipStream_t s) { return MhdStepT<PolicyExact>(a, s); }

hipError_t LaunchMhdCalcDt(const double *U, int nx, int ny, int nz, int ng, double dx, double dy, double dz,
                           double gamma, unsigned long long *dti, int num_cu, hipStream_t s)
{
  const int Lx = nx - 2 * ng, Ly = ny - 2 * ng, Lz = nz - 2 * ng;
  if (Lx <= 0 || Ly <= 0 || Lz <= 0) return hipSuccess;
  // about 8 workgroups per CU, each marching a chunk of whole z planes
  const int nxy    = (Lx * Ly + 255) / 256;
  const int zsplit = std::max(1, std::min(Lz, num_cu * 8 / nxy));
  const mhd::MhdRegion mr = mhd::MakeMhdMarch(ng, nx - ng - 1, ny - ng - 1, nz - ng - 1, (Lz + zsplit - 1) / zsplit);
  hipLaunchKernelGGL(MhdCalcDtKernel, dim3(mhd::MhdRegionBlocks(mr)), dim3(256), 0, s, U, nx, ny, nz, dx, dy, dz, gamma,
                     mr, dti);
  return hipGetLastError();
}

hipError_t LaunchMhdDivergence(const double *U, int nx, int ny, int nz, double dx, double dy, double dz,
                               unsigned long long *out, int num_cu, hipStream_t s)
{
  const size_t n = size_t(nx) * ny * nz;
  unsigned nb    = unsigned(std::min<size_t>(MhdBlocks(n), size_t(num_cu) * 8));
  hipLaunchKernelGGL(MhdDivergenceKernel, dim3(nb), dim3(256), 0, s, U, nx, ny, nz, dx, dy, dz, out);
  return hipGetLastError();
}

hipError_t LaunchHlldGrid(const double *QL, const double *QR, const double *magX, double *F, long n_cells,
                          double gamma, int dir, hipStream_t s)
{
  hipLaunchKernelGGL(HlldGridKernel, dim3(MhdBlocks(size_t(n_cells))), dim3(256), 0, s, QL, QR, magX, F, n_cells,
                     gamma, dir);
  return hipGetLastError();
}

hipError_t LaunchMhdPlmcGrid(const double *U, double *QL, double *QR, int nx, int ny, int nz, double gamma, int dir,
                             hipStream_t s)
{
  hipLaunchKernelGGL(MhdPlmcGridKernel, dim3(MhdBlocks(size_t(nx) * ny * nz)), dim3(256), 0, s, U, QL, QR, nx, ny,
                     nz, gamma, dir);
  return hipGetLastError();
}

hipError_t LaunchMhdPpmcGrid(const double *U, double *QL, double *QR, int nx, int ny, int nz, double gamma, int dir,
                             hipStream_t s)
{
  hipLaunchKernelGGL(MhdPpmcGridKernel, dim3(MhdBlocks(size_t(nx) * ny * nz)), dim3(256), 0, s, U, QL, QR, nx, ny,
                     nz, gamma, dir);
  return hipGetLastError();
}

hipError_t LaunchCtGrid(const double *Fx, const double *Fy, const double *Fz, const double *U, double *E, int nx,
                        int ny, int nz, hipStream_t s)
{
  hipLaunchKernelGGL((mhd::MhdEmfKernel<double>), dim3(MhdBlocks(size_t(nx) * ny * nz)), dim3(256), 0, s, Fx, Fy,
                     Fz, U, E, nx, ny, nz, 1, nx - 1, ny - 1, nz - 1);
  return hipGetLastError();
}

hipError_t LaunchMagneticUpdateGrid(const double *src, double *dst, const double *E, int nx, int ny, int nz,
                                    double dt, double dx, double dy, double dz, hipStream_t s)
{
  hipLaunchKernelGGL(MagneticUpdateGridKernel, dim3(MhdBlocks(size_t(nx) * ny * nz)), dim3(256), 0, s, src, dst, E,
                     nx, ny, nz, dt, dx, dy, dz);
  return hipGetLastError();
}

}  // namespace chl
