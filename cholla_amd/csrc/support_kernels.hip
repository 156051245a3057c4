// support_kernels.hip -- the rest of the hot path for gfx950, plus the
// reference-shaped stage kernels used to check the device physics against the
// reference's own unit-test vectors.
//
// Built with -ffp-contract=off (PolicyExact arithmetic).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "de_cells.hpp"
#include "ensemble_kernels.hpp"
#include "hydro_physics.hpp"
#include "support_kernels.hpp"
#include "vl3d_kernels.hpp"

namespace chl
{

// ============================================================================ ghost cells
// FindIndex, src/grid/cuda_boundaries.cu:240-302
__device__ __forceinline__ int FindIndexDev(int ig, int n, int flag, int face, int ng, double *a)
{
  if (face == 0) {
    switch (flag) {
      case 2:
        *a = -1.0;
        return 2 * ng - ig - 1;
      case 3:
        return ng;
      case 4:
        return -1;
      case 5:
        return ig;
      default:
        return ig + n - 2 * ng;
    }
  }
  switch (flag) {
    case 2:
      *a = -1.0;
      return 2 * (n - ng) - ig - 1;
    case 3:
      return n - ng - 1;
    case 4:
      return -1;
    case 5:
      return ig;
    default:
      return ig - n + 2 * ng;
  }
}

// One thread per ghost cell of all six faces.  Each ghost cell belongs to exactly
// one face slab of Set_Boundary_Extents (boundary_conditions.cpp:419-505): z slabs
// span everything, y slabs the real k range, x slabs the real j and k ranges.  The
// value written is what SetGhostCellsKernel (cuda_boundaries.cu:71-149) writes for
// that slab: the mapped real cell (all three indices mapped by SetBoundaryMapping,
// :161-238), momentum sign flips for reflective faces and the outflow diode of
// transmissive faces (hydro states only: the MHD build compiles the diode out,
// cuda_boundaries.cu:121-147, and fl.diode is 0 for it).  Sources are always real cells,
// so one launch reproduces the six sequential launches exactly.
__global__ void GhostFillKernel(double *c, int nx, int ny, int nz, int n_fields, int ng, GhostFlags fl, long n_x,
                                long n_y, long n_z)
{
  long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n_x + n_y + n_z) return;
  const int jl = ny > 1 ? ng : 0, ju = ny > 1 ? ny - ng : ny;
  const int kl = nz > 1 ? ng : 0, ku = nz > 1 ? nz - ng : nz;
  int i, j, k, dir;
  if (t < n_x) {  // x slabs: 2*ng x (ju-jl) x (ku-kl)
    int w = 2 * ng, hy = ju - jl;
    int ii = int(t % w);
    long r = t / w;
    j      = jl + int(r % hy);
    k      = kl + int(r / hy);
    i      = ii < ng ? ii : nx - 2 * ng + ii;
    dir    = ii < ng ? 0 : 1;
  } else if (t < n_x + n_y) {  // y slabs: nx x 2*ng x (ku-kl)
    long s = t - n_x;
    i      = int(s % nx);
    long r = s / nx;
    int jj = int(r % (2 * ng));
    k      = kl + int(r / (2 * ng));
    j      = jj < ng ? jj : ny - 2 * ng + jj;
    dir    = jj < ng ? 2 : 3;
  } else {  // z slabs: nx x ny x 2*ng
    long s = t - n_x - n_y;
    i      = int(s % nx);
    long r = s / nx;
    j      = int(r % ny);
    int kk = int(r / ny);
    k      = kk < ng ? kk : nz - 2 * ng + kk;
    dir    = kk < ng ? 4 : 5;
  }
  const int *flags = fl.f;
  if (flags[dir] == 4 || flags[dir] == 5) return;  // custom / MPI faces are set elsewhere
  double a[3] = {1., 1., 1.};
  long idx    = 0;
  int ir, jr, kr;
  if (nx > 1) {
    ir = i < ng ? FindIndexDev(i, nx, flags[0], 0, ng, &a[0])
                : (i >= nx - ng ? FindIndexDev(i, nx, flags[1], 1, ng, &a[0]) : i);
    if (ir < 0) return;
    idx += ir;
  }
  if (ny > 1) {
    jr = j < ng ? FindIndexDev(j, ny, flags[2], 0, ng, &a[1])
                : (j >= ny - ng ? FindIndexDev(j, ny, flags[3], 1, ng, &a[1]) : j);
    if (jr < 0) return;
    idx += long(nx) * jr;
  }
  if (nz > 1) {
    kr = k < ng ? FindIndexDev(k, nz, flags[4], 0, ng, &a[2])
                : (k >= nz - ng ? FindIndexDev(k, nz, flags[5], 1, ng, &a[2]) : k);
    if (kr < 0) return;
    idx += long(nx) * ny * kr;
  }
  const size_t nc = size_t(nx) * ny * nz;
  const size_t g  = Lin(i, j, k, nx, ny);
  for (int f = 0; f < n_fields; f++) c[g + f * nc] = c[idx + f * nc];
  if (flags[0] == 2 || flags[1] == 2) c[g + nc] *= a[0];
  if (flags[2] == 2 || flags[3] == 2) c[g + 2 * nc] *= a[1];
  if (flags[4] == 2 || flags[5] == 2) c[g + 3 * nc] *= a[2];
  if (fl.diode && flags[dir] == 3) {
    size_t mom = g + size_t(dir / 2 + 1) * nc;
    if (dir % 2 == 0) {
      if (c[mom] > 0.0) {
        c[g + 4 * nc] -= 0.5 * (c[mom] * c[mom]) / c[g];
        c[mom] = 0.0;
      }
    } else {
      if (c[mom] < 0.0) {
        c[g + 4 * nc] -= 0.5 * (c[mom] * c[mom]) / c[g];
        c[mom] = 0.0;
      }
    }
  }
}

// One face slab exactly as SetGhostCells(...) launches it (cuda_boundaries.cu:151-159).
__global__ void GhostSlabKernel(double *c, int nx, int ny, int nz, int n_fields, int ng, GhostFlags fl, int isize,
                                int jsize, int ksize, int imin, int jmin, int kmin, int dir)
{
  long id = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= long(isize) * jsize * ksize) return;
  int k = int(id / (long(isize) * jsize));
  int j = int((id - long(k) * isize * jsize) / isize);
  int i = int(id - long(k) * isize * jsize - long(j) * isize);
  i += imin, j += jmin, k += kmin;
  const int *flags = fl.f;
  double a[3]      = {1., 1., 1.};
  long idx         = 0;
  if (nx > 1) {
    int ir = i < ng ? FindIndexDev(i, nx, flags[0], 0, ng, &a[0])
                    : (i >= nx - ng ? FindIndexDev(i, nx, flags[1], 1, ng, &a[0]) : i);
    if (ir < 0) return;
    idx += ir;
  }
  if (ny > 1) {
    int jr = j < ng ? FindIndexDev(j, ny, flags[2], 0, ng, &a[1])
                    : (j >= ny - ng ? FindIndexDev(j, ny, flags[3], 1, ng, &a[1]) : j);
    if (jr < 0) return;
    idx += long(nx) * jr;
  }
  if (nz > 1) {
    int kr = k < ng ? FindIndexDev(k, nz, flags[4], 0, ng, &a[2])
                    : (k >= nz - ng ? FindIndexDev(k, nz, flags[5], 1, ng, &a[2]) : k);
    if (kr < 0) return;
    idx += long(nx) * ny * kr;
  }
  const size_t nc = size_t(nx) * ny * nz;
  const size_t g  = Lin(i, j, k, nx, ny);
  for (int f = 0; f < n_fields; f++) c[g + f * nc] = c[idx + f * nc];
  if (flags[0] == 2 || flags[1] == 2) c[g + nc] *= a[0];
  if (flags[2] == 2 || flags[3] == 2) c[g + 2 * nc] *= a[1];
  if (flags[4] == 2 || flags[5] == 2) c[g + 3 * nc] *= a[2];
  if (fl.diode && flags[dir] == 3) {
    size_t mom = g + size_t(dir / 2 + 1) * nc;
    if (dir % 2 == 0) {
      if (c[mom] > 0.0) {
        c[g + 4 * nc] -= 0.5 * (c[mom] * c[mom]) / c[g];
        c[mom] = 0.0;
      }
    } else {
      if (c[mom] < 0.0) {
        c[g + 4 * nc] -= 0.5 * (c[mom] * c[mom]) / c[g];
        c[mom] = 0.0;
      }
    }
  }
}

// ============================================================================ dt and floors
__device__ __forceinline__ void BlockMaxToGlobal(double v, unsigned long long *out)
{
  __shared__ double red[16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = Fmax(v, __shfl_xor(v, off, 64));
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int q = 1; q < int(blockDim.x >> 6); q++) m = Fmax(m, red[q]);
    atomicMax(out, (unsigned long long)__double_as_longlong(m));
  }
}

// Calc_dt_3D, src/hydro/hydro_cuda.cu:514-560 (grid-stride over all cells, real cells only)
__global__ void CalcDt3DKernel(const double *U, int nx, int ny, int nz, int ng, double dx, double dy, double dz,
                               double gamma, unsigned long long *dti)
{
  const size_t nc = size_t(nx) * ny * nz;
  double m        = 0.0;
  for (size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x; id < nc; id += size_t(blockDim.x) * gridDim.x) {
    int zid = int(id / (size_t(nx) * ny));
    int yid = int((id - size_t(zid) * nx * ny) / nx);
    int xid = int(id - size_t(zid) * nx * ny - size_t(yid) * nx);
    if (xid > ng - 1 && xid < nx - ng && yid > ng - 1 && yid < ny - ng && zid > ng - 1 && zid < nz - ng) {
      double d = U[id], dinv = 1.0 / d;
      double vx = U[nc + id] * dinv, vy = U[2 * nc + id] * dinv, vz = U[3 * nc + id] * dinv;
      m = Fmax(m, InverseCrossingTime(U[4 * nc + id], d, dinv, vx, vy, vz, dx, dy, dz, gamma));
    }
  }
  BlockMaxToGlobal(m, dti);
}

// Calc_dt_1D, src/hydro/hydro_cuda.cu:437-472
__global__ void CalcDt1DKernel(const double *U, int nx, int ng, double dx, double gamma, unsigned long long *dti)
{
  const size_t nc = size_t(nx);
  double m        = 0.0;
  for (size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x; id < nc; id += size_t(blockDim.x) * gridDim.x) {
    if (long(id) > long(ng) - 1 && long(id) < long(nx) - ng) {
      double d = U[id], dinv = 1.0 / d;
      double vx = U[nc + id] * dinv, vy = U[2 * nc + id] * dinv, vz = U[3 * nc + id] * dinv;
      double P  = (U[4 * nc + id] - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0);
      P         = Fmax(P, kTinyNumber);
      double cs = __builtin_sqrt(dinv * gamma * P);
      m         = Fmax(m, (Fabs(vx) + cs) / dx);
    }
  }
  BlockMaxToGlobal(m, dti);
}

// Temperature_Floor_Kernel, src/hydro/hydro_cuda.cu:1117-1153 (with the DE branch
// :1146-1151 when the state carries GasEnergy, n_fields == 6)
__global__ void TemperatureFloorKernel(double *U, int nx, int ny, int nz, int ng, double U_floor, int de)
{
  const size_t nc = size_t(nx) * ny * nz;
  size_t id       = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= nc) return;
  int zid = int(id / (size_t(nx) * ny));
  int yid = int((id - size_t(zid) * nx * ny) / nx);
  int xid = int(id - size_t(zid) * nx * ny - size_t(yid) * nx);
  if (xid > ng - 1 && xid < nx - ng && yid > ng - 1 && yid < ny - ng && zid > ng - 1 && zid < nz - ng) {
    double d = U[id], dinv = 1.0 / d;
    double vx = U[nc + id] * dinv, vy = U[2 * nc + id] * dinv, vz = U[3 * nc + id] * dinv;
    double E    = U[4 * nc + id];
    double Ekin = 0.5 * d * (vx * vx + vy * vy + vz * vz);
    double u    = (E - Ekin) / d;
    if (u < U_floor) U[4 * nc + id] = Ekin + d * U_floor;
    if (de) {
      u = U[kGasEnergy * nc + id] / d;
      if (u < U_floor) U[kGasEnergy * nc + id] = d * U_floor;
    }
  }
}

// ============================================================================ dual energy
// Select_Internal_Energy_3D over the real cells (DeSelectCell), then Sync_Energies_3D fused
// with the temperature floor, the inverse-dt reduction and the periodic ghost images
// (DeSyncCell); two launches because the selection reads the neighbours' energy, which
// the sync rewrites.
__global__ void DeSelectKernel(double *U, int nx, int ny, int nz, int ng)
{
  int i, j, k;
  if (!RealCell(long(blockIdx.x) * blockDim.x + threadIdx.x, nx, ny, nz, ng, i, j, k)) return;
  DeSelectCell(U, size_t(nx) * ny * nz, nx, ny, nz, ng, i, j, k);
}

__global__ void DeSyncKernel(double *U, int nx, int ny, int nz, int ng, int flags, double U_floor, double dx,
                             double dy, double dz, double gamma, unsigned long long *dti)
{
  int i, j, k;
  double m = 0.0;
  if (RealCell(long(blockIdx.x) * blockDim.x + threadIdx.x, nx, ny, nz, ng, i, j, k))
    m = DeSyncCell(U, size_t(nx) * ny * nz, nx, ny, nz, ng, i, j, k, flags, U_floor, dx, dy, dz, gamma);
  if (flags & kFlagReduceDt) BlockMaxToGlobal(m, dti);
}

// GasEnergy = E - kinetic energy of every cell (the DE state of a 5-field state).
__global__ void GasEnergyInitKernel(double *U, long nc)
{
  long id = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= nc) return;
  double d = U[id];
  double vx = U[nc + id] / d, vy = U[2 * nc + id] / d, vz = U[3 * nc + id] / d;
  U[kGasEnergy * nc + id] = U[4 * nc + id] - 0.5 * d * (vx * vx + vy * vy + vz * vz);
}

// Device-side time control for graph-captured stepping: dt = C_cfl / dti clamped
// to the output time as in main.cpp:254-261, and t advanced by dt.
// take: the accumulator is read and zeroed by one atomic exchange.  (A plain read followed by
// a store of 0 is not safe here: the compiler makes the read a scalar load, and the vector
// store may reach the cache before that load does.)
__global__ void DtAdvanceKernel(unsigned long long *dti_bits, double *dt_dev, double *t_dev, double tout,
                                double cfl, int *n_step, int take)
{
  const unsigned long long bits = take ? atomicExch(dti_bits, 0ull) : *dti_bits;
  double dti = __longlong_as_double((long long)bits);
  double t   = *t_dev;
  double dt  = cfl / dti;
  if (t + dt > tout) dt = tout - t;
  if (!(t < tout)) dt = 0.0;
  *dt_dev = dt;
  *t_dev  = t + dt;
  if (dt > 0.0) *n_step += 1;
}

// One thread per box: the same dt and clock arithmetic as DtAdvanceKernel, then the
// accumulators of the coming step are cleared.  Thread 0 also rewinds the per-XCD
// queues of the persistent launch.
__global__ void EnsembleClockKernel(BoxClock *clk, const BoxDesc *box, int n_boxes, double cfl, unsigned int *head)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b == 0 && head)
    for (int x = 0; x < kQueueXcds; x++) head[x * kHeadStride] = 0u;
  if (b >= n_boxes) return;
  BoxClock c       = clk[b];
  const double tout = box[b].tout;
  double dti       = __longlong_as_double((long long)c.dti);
  double dt        = cfl / dti;
  if (c.t + dt > tout) dt = tout - c.t;
  if (!(c.t < tout)) dt = 0.0;
  c.dt = dt;
  c.t  = c.t + dt;
  if (dt > 0.0) c.nstep += 1;
  c.dti   = 0ull;
  c.done1 = 0;
  c.done2 = 0;
  c.done3 = 0;
  clk[b]  = c;
}

// ============================================================================ turbulence IC
// Velocity field sum_m [a_m cos(2 pi k_m.x) + b_m sin(2 pi k_m.x)] * vscale at the cell
// centres x = (i - ng + 1/2) / n of a periodic unit box (positions as Grid3D::Get_Position,
// src/grid/grid3D.cpp:84-86); ghost cells take their periodic image's value.  Density and
// pressure uniform; energy as Calc_Energy_Primitive (src/utils/hydro_utilities.h:106-118).
// modes: 9 doubles per mode (kx, ky, kz, ax, ay, az, bx, by, bz).
__global__ void TurbulenceIcKernel(double *U, int n, int ng, int n_modes, const double *modes, double vscale,
                                   double rho, double p, double gamma)
{
  const int N     = n + 2 * ng;
  const size_t nc = size_t(N) * N * N;
  const size_t id = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= nc) return;
  const int i = int(id % N), j = int((id / N) % N), k = int(id / (size_t(N) * N));
  auto wrap   = [&](int a) { return a < ng ? a + n : (a >= N - ng ? a - n : a); };
  const double x = (wrap(i) - ng + 0.5) / n, y = (wrap(j) - ng + 0.5) / n, z = (wrap(k) - ng + 0.5) / n;
  double vx = 0, vy = 0, vz = 0;
  for (int m = 0; m < n_modes; m++) {
    const double *md = modes + 9 * m;
    double sn, cs;
    sincospi(2.0 * (md[0] * x + md[1] * y + md[2] * z), &sn, &cs);
    vx += md[3] * cs + md[6] * sn;
    vy += md[4] * cs + md[7] * sn;
    vz += md[5] * cs + md[8] * sn;
  }
  vx *= vscale, vy *= vscale, vz *= vscale;
  U[id]          = rho;
  U[nc + id]     = rho * vx;
  U[2 * nc + id] = rho * vy;
  U[3 * nc + id] = rho * vz;
  U[4 * nc + id] = (fmax(p, kTinyNumber) / (gamma - 1.)) + 0.5 * rho * (vx * vx + ((vy * vy) + (vz * vz)));
}

// ============================================================================ VL 1D and 2D: arithmetic forms
// The 1D and 2D integrators run in the three forms of the 3D one.  P is the form (with its
// Riemann solver): PolicyExact is the reference's order in FP64 (the oracle's bits);
// PolicyFast FP64 with the reciprocal-multiply and FMA arithmetic; PolicyMixed the
// reconstruction and Riemann fluxes in FP32, with the conserved state and its updates, the
// primitive (pressure) recovery and dt in FP64.  The mixed solvers take the split pressure
// of the 3D kernels: each face's reference p0 is the high FP32 part of its left cell's FP64
// pressure, both sides' momentum pressures are relative to it, and the normal-momentum flux
// comes back less p0, added back here in FP64 (the sum of two FP32 values, exact).

// The Riemann input of a cell's conserved values (the PCM states).  FP64 forms: the form's
// own recovery.  Mixed: the FP64 recovery, then the state in FP32; plo the low part of the
// FP64 pressure (p64 - p).
template <class P, bool DE>
__device__ __forceinline__ RState<typename P::T> SideFromCons(double d, double mx, double my, double mz, double E,
                                                              double gamma, double dge, double &plo)
{
  using T = typename P::T;
  if constexpr (sizeof(T) == sizeof(double)) {
    plo = 0.0;
    return RStateFromCons<P, DE>(d, mx, my, mz, E, gamma, dge);
  } else {
    const RState<double> s = RStateFromCons<PolicyFast, DE>(d, mx, my, mz, E, gamma, dge);
    RState<T> r{};
    r.d = T(s.d), r.mx = T(s.mx), r.my = T(s.my), r.mz = T(s.mz), r.E = T(s.E);
    r.vx = T(s.vx), r.vy = T(s.vy), r.vz = T(s.vz), r.p = T(s.p), r.inv_d = T(s.inv_d), r.dge = T(s.dge);
    plo = s.p - double(r.p);
    return r;
  }
}

// The flux of one face in FP64 from its two Riemann inputs (the split pressure's p0 added
// back to the normal-momentum flux); fge the GasEnergy flux under DE.
template <class P, bool DE>
__device__ __forceinline__ Cons<double> FluxFp64(const RState<typename P::T> &l, const RState<typename P::T> &r,
                                                 typename P::T gamma, double *fge)
{
  using T = typename P::T;
  T g     = T(0);
  const Cons<T> f = RiemannFlux<P, DE>(l, r, gamma, &g);
  Cons<double> o{double(f.d), double(f.mx), double(f.my), double(f.mz), double(f.E)};
  if constexpr (SplitPressure<P>) o.mx += double(l.p0);
  *fge = double(g);
  return o;
}

// The flux between two cells' PCM states a (left) and b (right), with the low parts of
// their FP64 pressures (the 3D kernels' PcmFaceFlux split).
template <class P, bool DE>
__device__ __forceinline__ Cons<double> PcmFlux(RState<typename P::T> a, double alo, RState<typename P::T> b,
                                                double blo, typename P::T gamma, double *fge)
{
  using T = typename P::T;
  if constexpr (SplitPressure<P>) {
    a.p0 = b.p0 = a.p;
    b.pd        = (b.p - a.p) + T(blo);
    a.pd        = T(alo);
  }
  return FluxFp64<P, DE>(a, b, gamma, fge);
}

// Stage 1 of VL_Algorithm_1D_CUDA (VL_1D_cuda.cu:64-91): PCM states, the Riemann
// solve of P (HLLC, or another solver under a SolverPolicy) on faces id-1/2 and id+1/2
// of every cell 1..n-2 and the dt/2 update; W holds the conserved half state.  With
// dual energy (DE) the GasEnergy half update carries the half-step pdV source of
// Update_Conserved_Variables_1D_half (hydro_cuda.cu:222-263): the pressure from the
// total energy, the neighbours' velocities by division.
template <class P, bool DE>
__global__ void Vl1dStage1Kernel(const double *U, double *W, int n, double dx, const double *dt_dev, double dt_host,
                                 double gamma, unsigned long long *dti_reset)
{
  using T = typename P::T;
  int id  = blockIdx.x * blockDim.x + threadIdx.x;
  if (dti_reset && id == 0) *dti_reset = 0ull;  // a later launch of this step reduces into it
  if (id < 1 || id >= n - 1) return;
  const double dt    = 0.5 * (dt_dev ? *dt_dev : dt_host);
  const double dtodx = dt / dx;
  double lo[3];
  auto side = [&](int c, double &plo) {
    return SideFromCons<P, DE>(U[c], U[n + c], U[2 * n + c], U[3 * n + c], U[4 * n + c], gamma,
                               DE ? U[5 * n + c] : 0.0, plo);
  };
  RState<T> a = side(id - 1, lo[0]), b = side(id, lo[1]), cc = side(id + 1, lo[2]);
  double gm = 0.0, gp = 0.0;
  Cons<double> fm = PcmFlux<P, DE>(a, lo[0], b, lo[1], T(gamma), &gm);
  Cons<double> fp = PcmFlux<P, DE>(b, lo[1], cc, lo[2], T(gamma), &gp);
  W[id]         = U[id] + dtodx * (fm.d - fp.d);
  W[n + id]     = U[n + id] + dtodx * (fm.mx - fp.mx);
  W[2 * n + id] = U[2 * n + id] + dtodx * (fm.my - fp.my);
  W[3 * n + id] = U[3 * n + id] + dtodx * (fm.mz - fp.mz);
  W[4 * n + id] = U[4 * n + id] + dtodx * (fm.E - fp.E);
  if constexpr (DE) {
    const double d = U[id], d_inv = 1.0 / d;
    const double vx = U[n + id] * d_inv, vy = U[2 * n + id] * d_inv, vz = U[3 * n + id] * d_inv;
    const double pr     = (U[4 * n + id] - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0);
    const double vx_imo = U[n + id - 1] / U[id - 1], vx_ipo = U[n + id + 1] / U[id + 1];
    W[5 * n + id]       = U[5 * n + id] + dtodx * (gm - gp) + 0.5 * pr * (dtodx * (vx_imo - vx_ipo));
  }
}

// The interface values of cell c from the conserved half state W (fields nf apart) along a
// line of stride st, whose normal and transverse momenta are the fields o1, o2, o3:
// primitives as Load_Data forms them (PLMC, PPMC) or as PLMP_cuda does
// (plmp_cuda.cu:104-120), in FP64, then the reconstruction in the form's precision; PCM
// passes the conserved values through in the exact form (pcm_cuda.cu:13-97) and the
// primitives in the others.  With dual energy (PLMC and PPMC only) the pressure is
// Load_Data's DE pressure and the specific GasEnergy is reconstructed beside
// (reconstruction.h:179-191, plmc_cuda.cu, ppmc_cuda.cu:668-672).  pc: the cell's own
// pressure in the form's precision (the mixed form's face reference).
template <class P, int RK, bool DE = false>
__device__ __forceinline__ void ReconLine(const double *W, long nf, long c, long st, int o1, int o2, int o3,
                                          double gamma, Prim<typename P::T> &L, Prim<typename P::T> &R,
                                          typename P::T &geL, typename P::T &geR, typename P::T &pc)
{
  using T          = typename P::T;
  constexpr int r  = ReconTraits<RK>::kRadius;
  constexpr int NS = 2 * r + 1;
  static_assert(!DE || RK != kPCM, "the PCM states carry the conserved GasEnergy (Vl2dPcmFluxKernel)");
  Prim<T> s[NS];
  Prim<double> w[NS];  // the FP64 primitives (the forms other than the exact PCM pass-through)
  T ge[NS];
#pragma unroll
  for (int m = 0; m < NS; m++) {
    const long q   = c + long(m - r) * st;
    const double d = W[q], mx = W[o1 * nf + q], my = W[o2 * nf + q], mz = W[o3 * nf + q], E = W[4 * nf + q];
    if constexpr (RK == kPCM && P::kExact) {
      s[m] = {d, mx, my, mz, E};
      continue;
    } else if constexpr (RK == kPLMP || RK == kPPMP) {
      const double vx = mx / d, vy = my / d, vz = mz / d;
      if constexpr (DE) {  // plmp_cuda.cu:116-121
        const double g = W[5 * nf + q], E_kin = 0.5 * d * (vx * vx + vy * vy + vz * vz);
        w[m]  = {d, vx, vy, vz, Fmax(PressureFromDE(E, E - E_kin, g, gamma), kTinyNumber)};
        ge[m] = T(g / d);
      } else {
        w[m] = {d, vx, vy, vz, Fmax((E - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0), kTinyNumber)};
      }
    } else if constexpr (DE) {
      const double vx = mx / d, vy = my / d, vz = mz / d, g = W[5 * nf + q];
      w[m]  = {d, vx, vy, vz, PressureFromDE(E, E - 0.5 * d * SquareMag(vx, vy, vz), g, gamma)};
      ge[m] = T(g / d);
    } else {
      w[m] = LoadPrim(d, mx, my, mz, E, gamma);
    }
    s[m] = {T(w[m].d), T(w[m].vx), T(w[m].vy), T(w[m].vz), T(w[m].p)};
  }
  pc = s[r].p;
  const T g = T(gamma);
  if constexpr (RK == kPCM) {
    L = R = s[0];
  } else if constexpr (RK == kPLMP) {
    ReconPlmp<P>(s, L, R);
    if constexpr (DE) ReconPlmpScalar<P>(ge, geL, geR);
  } else if constexpr (RK == kPPMP) {
    ReconPpmp<P>(s, L, R);
    if constexpr (DE) ReconPpmpScalar<P>(ge, geL, geR);
  } else if constexpr (RK == kPPMC && SplitPressure<P>) {
    // the mixed form's PPMC works on deviations from the centre cell (ReconPpmcCore), here
    // taken in FP64, with the centre's state for the eigenvectors; the centre is added back
    Prim<T> dv[NS];
#pragma unroll
    for (int m = 0; m < NS; m++)
      dv[m] = {T(w[m].d - w[r].d), T(w[m].vx - w[r].vx), T(w[m].vy - w[r].vy), T(w[m].vz - w[r].vz),
               T(w[m].p - w[r].p)};
    ReconPpmcCore<P>(dv, s[r], g, L, R);
    const Prim<T> &q = s[r];
    L = {L.d + q.d, L.vx + q.vx, L.vy + q.vy, L.vz + q.vz, L.p + q.p};
    R = {R.d + q.d, R.vx + q.vx, R.vy + q.vy, R.vz + q.vz, R.p + q.p};
    L.d = Fmax(L.d, T(kTinyNumber)), R.d = Fmax(R.d, T(kTinyNumber));
    L.p = Fmax(L.p, T(kTinyNumber)), R.p = Fmax(R.p, T(kTinyNumber));
    if constexpr (DE) ReconPpmcScalar<P>(ge, geL, geR);
  } else if constexpr (RK == kPPMC) {
    ReconPpmc<P>(s, g, L, R);
    if constexpr (DE) ReconPpmcScalar<P>(ge, geL, geR);
  } else {
    ReconPlmc<P>(s, g, L, R);
    if constexpr (DE) ReconPlmcScalar<P>(ge, geL, geR);
  }
}

// The flux of the face between a cell whose interface value there is L and the next cell,
// whose value there is R; p0 the left cell's own pressure (the mixed form's reference).
template <class P, int RK, bool DE>
__device__ __forceinline__ Cons<double> ReconFlux(const Prim<typename P::T> &L, const Prim<typename P::T> &R,
                                                  typename P::T geL, typename P::T geR, typename P::T p0,
                                                  typename P::T gamma, typename P::T inv_gm1, double *fge)
{
  using T = typename P::T;
  RState<T> l, r;
  if constexpr (SplitPressure<P>) {
    Prim<T> a = L, b = R;
    a.p       = Fmax(L.p, T(kTinyNumber));
    b.p       = Fmax(R.p, T(kTinyNumber));
    l         = RStateFromPrim<P, DE>(a, gamma, inv_gm1, geL);
    r         = RStateFromPrim<P, DE>(b, gamma, inv_gm1, geR);
    l.p0 = r.p0 = p0;
    l.pd        = Fmax(l.p - p0, -p0);
    r.pd        = Fmax(r.p - p0, -p0);
  } else {
    l = IfaceState<P, RK, DE>(L, gamma, inv_gm1, geL);
    r = IfaceState<P, RK, DE>(R, gamma, inv_gm1, geR);
  }
  return FluxFp64<P, DE>(l, r, gamma, fge);
}

// Stage 2 of VL_Algorithm_1D_CUDA (VL_1D_cuda.cu:95-137) for the real cells, with the
// next inverse dt (Calc_dt_1D) reduced in the epilogue (without dual energy; with it the
// energy sync that follows reduces it).  With DE the GasEnergy takes its flux difference
// on top of the full-step pdV source Vl1dPdvKernel added; PCM with DE passes the half state
// through with its conserved GasEnergy (PCM_Reconstruction_1D, pcm_cuda.cu:13-97), so its
// faces are stage 1's PCM faces on W.
template <class P, int RK, bool DE>
__global__ void Vl1dStage2Kernel(double *U, const double *W, int n, int ng, double dx, const double *dt_dev,
                                 double dt_host, double gamma, unsigned long long *dti, GravSpec grav)
{
  using T         = typename P::T;
  int id          = blockIdx.x * blockDim.x + threadIdx.x;
  double m        = 0.0;
  const double dt = dt_dev ? *dt_dev : dt_host;
  if (id >= ng && id < n - ng) {
    const T g = T(gamma), inv_gm1 = T(1.0 / (gamma - 1.0));
    double gm = 0.0, gp = 0.0;
    Cons<double> fm, fp;
    if constexpr (RK == kPCM && DE) {
      double lo[3];
      auto side = [&](int c, double &plo) {
        return SideFromCons<P, DE>(W[c], W[n + c], W[2 * n + c], W[3 * n + c], W[4 * n + c], gamma, W[5 * n + c],
                                   plo);
      };
      RState<T> a = side(id - 1, lo[0]), b = side(id, lo[1]), cc = side(id + 1, lo[2]);
      fm = PcmFlux<P, DE>(a, lo[0], b, lo[1], g, &gm);
      fp = PcmFlux<P, DE>(b, lo[1], cc, lo[2], g, &gp);
    } else {
      Prim<T> Lm, Rm, L0, R0, Lp, Rp;
      T gLm = 0, gRm = 0, gL0 = 0, gR0 = 0, gLp = 0, gRp = 0, pm, p0, pp;
      ReconLine<P, RK, DE>(W, n, id - 1, 1, 1, 2, 3, gamma, Lm, Rm, gLm, gRm, pm);
      ReconLine<P, RK, DE>(W, n, id, 1, 1, 2, 3, gamma, L0, R0, gL0, gR0, p0);
      ReconLine<P, RK, DE>(W, n, id + 1, 1, 1, 2, 3, gamma, Lp, Rp, gLp, gRp, pp);
      fm = ReconFlux<P, RK, DE>(Lm, R0, gLm, gR0, pm, g, inv_gm1, &gm);
      fp = ReconFlux<P, RK, DE>(L0, Rp, gL0, gRp, p0, g, inv_gm1, &gp);
    }
    const double dtodx = dt / dx;
    double un[5];
    un[0] = U[id] + dtodx * (fm.d - fp.d);
    un[1] = U[n + id] + dtodx * (fm.mx - fp.mx);
    un[2] = U[2 * n + id] + dtodx * (fm.my - fp.my);
    un[3] = U[3 * n + id] + dtodx * (fm.mz - fp.mz);
    un[4] = U[4 * n + id] + dtodx * (fm.E - fp.E);
    if (grav.id != 0) {  // STATIC_GRAV, hydro_cuda.cu:59-67
      const double start[4] = {U[id], U[n + id], U[2 * n + id], U[3 * n + id]};
      const double gf[1]    = {GravField1D(grav.id, grav.Pos(0, id, ng, dx))};
      GravSource<1>(un, start, gf, dt);
    }
    for (int q = 0; q < 5; q++) U[q * n + id] = un[q];
    if constexpr (DE) {
      U[5 * n + id] += dtodx * (gm - gp);
    } else {
      double d = un[0], dinv = 1.0 / d;
      double vx = un[1] * dinv, vy = un[2] * dinv, vz = un[3] * dinv;
      double pr = (un[4] - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0);
      pr        = Fmax(pr, kTinyNumber);
      double cs = __builtin_sqrt(dinv * gamma * pr);
      m         = (Fabs(vx) + cs) / dx;
    }
  }
  if (!DE && dti) BlockMaxToGlobal(m, dti);
}

// Partial_Update_Advected_Internal_Energy_1D, hydro_cuda.cu:663-700: the full-step pdV
// source from the state at the start of the step (each cell writes only its GasEnergy and
// reads its neighbours' density and momentum).
__global__ void Vl1dPdvKernel(double *U, int n, int ng, double dx, const double *dt_dev, double dt_host, double gamma)
{
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id < ng || id >= n - ng) return;
  const double dtodx = (dt_dev ? *dt_dev : dt_host) / dx;
  const double d = U[id], d_inv = 1.0 / d;
  const double vx = U[n + id] * d_inv, vy = U[2 * n + id] * d_inv, vz = U[3 * n + id] * d_inv;
  const double E = U[4 * n + id], GE = U[5 * n + id];
  const double E_kin  = 0.5 * d * (vx * vx + vy * vy + vz * vz);
  const double pr     = Fmax(PressureFromDE(E, E - E_kin, GE, gamma), kTinyNumber);
  const double vx_imo = U[n + id - 1] / U[id - 1], vx_ipo = U[n + id + 1] / U[id + 1];
  U[5 * n + id] += 0.5 * pr * (dtodx * (vx_imo - vx_ipo));
}

// Select_Internal_Energy_1D, hydro_cuda.cu:835-885: writes GasEnergy only; the neighbours'
// total energies (clamped to real cells) are read, not written.
__global__ void Vl1dSelectKernel(double *U, int n, int ng)
{
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id < ng || id >= n - ng) return;
  const int imo = max(id - 1, ng), ipo = min(id + 1, n - ng - 1);
  const double d = U[id], d_inv = 1.0 / d;
  const double vx = U[n + id] * d_inv, vy = U[2 * n + id] * d_inv, vz = U[3 * n + id] * d_inv;
  const double E = U[4 * n + id], U_adv = U[5 * n + id];
  const double U_tot = E - 0.5 * d * (vx * vx + vy * vy + vz * vz);
  double Emax        = Fmax(U[4 * n + imo], E);
  Emax               = Fmax(Emax, U[4 * n + ipo]);
  const double u     = ((U_tot / E > kDeEta1) || (U_tot / Emax > kDeEta2)) ? U_tot : U_adv;
  U[5 * n + id]      = Fmax(u, kTinyNumber);
}

// Sync_Energies_1D, hydro_cuda.cu:1018-1040, then the next inverse dt (Calc_dt_1D) of the
// synchronised cell.
__global__ void Vl1dSyncKernel(double *U, int n, int ng, double dx, double gamma, unsigned long long *dti)
{
  int id   = blockIdx.x * blockDim.x + threadIdx.x;
  double m = 0.0;
  if (id >= ng && id < n - ng) {
    const double d = U[id], d_inv = 1.0 / d;
    const double vx = U[n + id] * d_inv, vy = U[2 * n + id] * d_inv, vz = U[3 * n + id] * d_inv;
    const double E  = 0.5 * d * (vx * vx + vy * vy + vz * vz) + U[5 * n + id];
    U[4 * n + id]   = E;
    double pr       = (E - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0);
    pr              = Fmax(pr, kTinyNumber);
    m               = (Fabs(vx) + __builtin_sqrt(d_inv * gamma * pr)) / dx;
  }
  if (dti) BlockMaxToGlobal(m, dti);
}

// ============================================================================ VL 2D
// PCM_Reconstruction_2D with the Riemann solve of P (VL_2D_cuda.cu:49-71, and stage 2 of
// the PCM build): the x face i+1/2 and y face j+1/2 of every cell with a neighbour there,
// from the conserved states as they are.  Thread 0 clears the step's dt reduction.
template <class P, bool DE = false>
__global__ void Vl2dPcmFluxKernel(const double *U, double *Fx, double *Fy, int nx, int ny, double gamma,
                                  unsigned long long *dti_reset)
{
  using T       = typename P::T;
  const long nc = long(nx) * ny;
  const long id = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (dti_reset && id == 0) *dti_reset = 0ull;
  if (id >= nc) return;
  const int j = int(id / nx), i = int(id - long(j) * nx);
  auto state = [&](long c, int o1, int o2, int o3, double &plo) {
    return SideFromCons<P, DE>(U[c], U[o1 * nc + c], U[o2 * nc + c], U[o3 * nc + c], U[4 * nc + c], gamma,
                               DE ? U[5 * nc + c] : 0.0, plo);
  };
  auto solve = [&](double *F, long other, int o1, int o2, int o3) {
    double fge = 0.0, alo, blo;
    const RState<T> a = state(id, o1, o2, o3, alo), b = state(other, o1, o2, o3, blo);
    const Cons<double> f = PcmFlux<P, DE>(a, alo, b, blo, T(gamma), &fge);
    F[id] = f.d, F[o1 * nc + id] = f.mx, F[o2 * nc + id] = f.my, F[o3 * nc + id] = f.mz, F[4 * nc + id] = f.E;
    if constexpr (DE) F[5 * nc + id] = fge;
  };
  if (i < nx - 1) solve(Fx, id + 1, 1, 2, 3);
  if (j < ny - 1) solve(Fy, id + nx, 2, 3, 1);
}

// Stage 2 of the 2D integrator in the fast and mixed forms: each face's two interface values
// reconstructed from the half state along its line (ReconLine, as the 1D integrator does) and
// solved in one launch, without the exact form's interface arrays; the x faces i+1/2 with
// i in [ng-1, nx-ng-1] of the real rows and the y faces likewise, which the update reads.
template <class P, int RK, bool DE>
__global__ void Vl2dFlux2Kernel(const double *Uh, double *Fx, double *Fy, int nx, int ny, int ng, double gamma)
{
  using T       = typename P::T;
  const long nc = long(nx) * ny;
  const long id = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= nc) return;
  const int j = int(id / nx), i = int(id - long(j) * nx);
  const T g = T(gamma), inv_gm1 = T(1.0 / (gamma - 1.0));
  auto face = [&](double *F, long st, int o1, int o2, int o3) {
    Prim<T> La, Ra, Lb, Rb;
    T gLa = 0, gRa = 0, gLb = 0, gRb = 0, pa, pb;
    ReconLine<P, RK, DE>(Uh, nc, id, st, o1, o2, o3, gamma, La, Ra, gLa, gRa, pa);
    ReconLine<P, RK, DE>(Uh, nc, id + st, st, o1, o2, o3, gamma, Lb, Rb, gLb, gRb, pb);
    double fge          = 0.0;
    const Cons<double> f = ReconFlux<P, RK, DE>(La, Rb, gLa, gRb, pa, g, inv_gm1, &fge);
    F[id] = f.d, F[o1 * nc + id] = f.mx, F[o2 * nc + id] = f.my, F[o3 * nc + id] = f.mz, F[4 * nc + id] = f.E;
    if constexpr (DE) F[5 * nc + id] = fge;
  };
  if (j >= ng && j < ny - ng && i >= ng - 1 && i < nx - ng) face(Fx, 1, 1, 2, 3);
  if (i >= ng && i < nx - ng && j >= ng - 1 && j < ny - ng) face(Fy, nx, 2, 3, 1);
}

// Update_Conserved_Variables_2D_half, VL_2D_cuda.cu:183-256 (no DE): cells 1..n-2 along
// both axes, the y term added after the x term.
template <bool DE = false>
__global__ void Vl2dHalfKernel(const double *U, double *Uh, const double *Fx, const double *Fy, int nx, int ny,
                               double dx, double dy, const double *dt_dev, double dt_host, double gamma)
{
  const long nc = long(nx) * ny;
  const long id = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= nc) return;
  const int j = int(id / nx), i = int(id - long(j) * nx);
  if (i < 1 || i >= nx - 1 || j < 1 || j >= ny - 1) return;
  const double dt = 0.5 * (dt_dev ? *dt_dev : dt_host), dtodx = dt / dx, dtody = dt / dy;
  const long imo = id - 1, jmo = id - nx;
#pragma unroll
  for (int f = 0; f < 5 + int(DE); f++) {
    const long o = f * nc;
    Uh[o + id]   = U[o + id] + dtodx * (Fx[o + imo] - Fx[o + id]) + dtody * (Fy[o + jmo] - Fy[o + id]);
  }
  if constexpr (DE) {  // the half-step pdV source (VL_2D_cuda.cu:203-218, 248-252)
    const double d = U[id], d_inv = 1.0 / d;
    const double vx = U[nc + id] * d_inv, vy = U[2 * nc + id] * d_inv, vz = U[3 * nc + id] * d_inv;
    const double pr = (U[4 * nc + id] - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0);
    const long ipo = id + 1, jpo = id + nx, o = 5 * nc;
    const double vx_imo = U[nc + imo] / U[imo], vx_ipo = U[nc + ipo] / U[ipo];
    const double vy_jmo = U[2 * nc + jmo] / U[jmo], vy_jpo = U[2 * nc + jpo] / U[jpo];
    Uh[o + id] = U[o + id] + dtodx * (Fx[o + imo] - Fx[o + id]) + dtody * (Fy[o + jmo] - Fy[o + id]) +
                 0.5 * pr * (dtodx * (vx_imo - vx_ipo) + dtody * (vy_jmo - vy_jpo));
  }
}

// Partial_Update_Advected_Internal_Energy_2D (hydro_cuda.cu:702-750): the full-step pdV
// source from the start-of-step state (writes GasEnergy only).
__global__ void Vl2dPdvKernel(double *U, int nx, int ny, int ng, double dx, double dy, const double *dt_dev,
                              double dt_host, double gamma)
{
  const long nc = long(nx) * ny;
  const long id = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= nc) return;
  const int j = int(id / nx), i = int(id - long(j) * nx);
  if (i < ng || i >= nx - ng || j < ng || j >= ny - ng) return;
  const double dt = dt_dev ? *dt_dev : dt_host, dtodx = dt / dx, dtody = dt / dy;
  const double d = U[id], d_inv = 1.0 / d;
  const double vx = U[nc + id] * d_inv, vy = U[2 * nc + id] * d_inv, vz = U[3 * nc + id] * d_inv;
  const double E = U[4 * nc + id], GE = U[5 * nc + id];
  const double pr = Fmax(PressureFromDE(E, E - 0.5 * d * (vx * vx + vy * vy + vz * vz), GE, gamma), kTinyNumber);
  const long imo = id - 1, ipo = id + 1, jmo = id - nx, jpo = id + nx;
  const double vx_imo = U[nc + imo] / U[imo], vx_ipo = U[nc + ipo] / U[ipo];
  const double vy_jmo = U[2 * nc + jmo] / U[jmo], vy_jpo = U[2 * nc + jpo] / U[jpo];
  U[5 * nc + id] += 0.5 * pr * (dtodx * (vx_imo - vx_ipo) + dtody * (vy_jmo - vy_jpo));
}

// Select_Internal_Energy_2D (hydro_cuda.cu:887-935): GasEnergy only; neighbours clamped.
__global__ void Vl2dSelectKernel(double *U, int nx, int ny, int ng)
{
  const long nc = long(nx) * ny;
  const long id = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= nc) return;
  const int j = int(id / nx), i = int(id - long(j) * nx);
  if (i < ng || i >= nx - ng || j < ng || j >= ny - ng) return;
  const long imo = max(i - 1, ng) + long(j) * nx, ipo = min(i + 1, nx - ng - 1) + long(j) * nx;
  const long jmo = i + long(max(j - 1, ng)) * nx, jpo = i + long(min(j + 1, ny - ng - 1)) * nx;
  const double d = U[id], d_inv = 1.0 / d;
  const double vx = U[nc + id] * d_inv, vy = U[2 * nc + id] * d_inv, vz = U[3 * nc + id] * d_inv;
  const double E = U[4 * nc + id], U_adv = U[5 * nc + id];
  const double U_tot = E - 0.5 * d * (vx * vx + vy * vy + vz * vz);
  double Emax        = Fmax(U[4 * nc + imo], E);
  Emax               = Fmax(Emax, U[4 * nc + ipo]);
  Emax               = Fmax(Emax, U[4 * nc + jmo]);
  Emax               = Fmax(Emax, U[4 * nc + jpo]);
  const double u     = ((U_tot / E > kDeEta1) || (U_tot / Emax > kDeEta2)) ? U_tot : U_adv;
  U[5 * nc + id]     = Fmax(u, kTinyNumber);
}

// Sync_Energies_2D (hydro_cuda.cu:1042-1068), then Calc_dt_2D of the synchronised cell.
__global__ void Vl2dSyncKernel(double *U, int nx, int ny, int ng, double dx, double dy, double gamma,
                               unsigned long long *dti)
{
  const long nc = long(nx) * ny;
  const long id = long(blockIdx.x) * blockDim.x + threadIdx.x;
  double m      = 0.0;
  if (id < nc) {
    const int j = int(id / nx), i = int(id - long(j) * nx);
    if (i >= ng && i < nx - ng && j >= ng && j < ny - ng) {
      const double d = U[id], d_inv = 1.0 / d;
      const double vx = U[nc + id] * d_inv, vy = U[2 * nc + id] * d_inv, vz = U[3 * nc + id] * d_inv;
      const double E  = 0.5 * d * (vx * vx + vy * vy + vz * vz) + U[5 * nc + id];
      U[4 * nc + id]  = E;
      double pr       = Fmax((E - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0), 1.0e-20);
      const double cs = __builtin_sqrt(d_inv * gamma * pr);
      m               = Fmax((Fabs(vx) + cs) / dx, (Fabs(vy) + cs) / dy);
    }
  }
  if (dti) BlockMaxToGlobal(m, dti);
}

// Update_Conserved_Variables_2D (hydro_cuda.cu:84-160, no gravity, no DE) on the real
// cells, the x and y terms summed before they are added, then Calc_dt_2D
// (hydro_cuda.cu:474-510) of the updated cell into *dti.
template <bool DE = false>
__global__ void Vl2dUpdateKernel(double *U, const double *Fx, const double *Fy, int nx, int ny, int ng, double dx,
                                 double dy, const double *dt_dev, double dt_host, double gamma,
                                 unsigned long long *dti, GravSpec grav)
{
  const long nc = long(nx) * ny;
  const long id = long(blockIdx.x) * blockDim.x + threadIdx.x;
  double m      = 0.0;
  if (id < nc) {
    const int j = int(id / nx), i = int(id - long(j) * nx);
    if (i >= ng && i < nx - ng && j >= ng && j < ny - ng) {
      const double dt = dt_dev ? *dt_dev : dt_host, dtodx = dt / dx, dtody = dt / dy;
      const long imo = id - 1, jmo = id - nx;
      double un[5];
      const double start[4] = {U[id], U[nc + id], U[2 * nc + id], U[3 * nc + id]};
#pragma unroll
      for (int f = 0; f < 5; f++) {
        const long o = f * nc;
        un[f]        = U[o + id] + (dtodx * (Fx[o + imo] - Fx[o + id]) + dtody * (Fy[o + jmo] - Fy[o + id]));
      }
      if (grav.id != 0) {  // STATIC_GRAV, hydro_cuda.cu:142-153
        double g[2];
        GravField2D(grav.id, grav.Pos(0, i, ng, dx), grav.Pos(1, j, ng, dy), g[0], g[1]);
        GravSource<2>(un, start, g, dt);
      }
#pragma unroll
      for (int f = 0; f < 5; f++) U[f * nc + id] = un[f];
      if constexpr (DE) {  // GasEnergy: its fluxes on top of Vl2dPdvKernel's source; dt after the sync
        const long o = 5 * nc;
        U[o + id] += dtodx * (Fx[o + imo] - Fx[o + id]) + dtody * (Fy[o + jmo] - Fy[o + id]);
      }
      const double d = un[0], dinv = 1.0 / d;
      const double vx = un[1] * dinv, vy = un[2] * dinv, vz = un[3] * dinv;
      double pr       = (un[4] - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0);
      pr              = Fmax(pr, 1.0e-20);
      const double cs = __builtin_sqrt(dinv * gamma * pr);
      m               = Fmax((Fabs(vx) + cs) / dx, (Fabs(vy) + cs) / dy);
    }
  }
  if (!DE && dti) BlockMaxToGlobal(m, dti);
}

// Calc_dt_2D, hydro_cuda.cu:474-510
__global__ void CalcDt2DKernel(const double *U, int nx, int ny, int ng, double dx, double dy, double gamma,
                               unsigned long long *dti)
{
  const long nc = long(nx) * ny;
  double m      = 0.0;
  for (long id = long(blockIdx.x) * blockDim.x + threadIdx.x; id < nc; id += long(blockDim.x) * gridDim.x) {
    const int j = int(id / nx), i = int(id - long(j) * nx);
    if (i < ng || i >= nx - ng || j < ng || j >= ny - ng) continue;
    const double d = U[id], dinv = 1.0 / d;
    const double vx = U[nc + id] * dinv, vy = U[2 * nc + id] * dinv, vz = U[3 * nc + id] * dinv;
    double pr       = (U[4 * nc + id] - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0);
    pr              = Fmax(pr, 1.0e-20);
    const double cs = __builtin_sqrt(dinv * gamma * pr);
    m               = Fmax(m, Fmax((Fabs(vx) + cs) / dx, (Fabs(vy) + cs) / dy));
  }
  BlockMaxToGlobal(m, dti);
}

// ============================================================================ reference-shaped stage kernels
__device__ __forceinline__ bool ThreadGuard(int order, int nx, int ny, int nz, int xid, int yid, int zid)
{
  bool out = xid < order - 1 || xid >= nx - order;
  if (ny > 1) out = yid < order - 1 || yid >= ny - order || out;
  if (nz > 1) {
    out = zid < order - 1 || zid >= nz - order || out;
  } else {
    out = zid >= nz || out;
  }
  return out;
}

__device__ __forceinline__ void DirSlots(int dir, int &o1, int &o2, int &o3)
{
  o1 = dir == 0 ? 1 : (dir == 1 ? 2 : 3);
  o2 = dir == 0 ? 2 : (dir == 1 ? 3 : 1);
  o3 = dir == 0 ? 3 : (dir == 1 ? 1 : 2);
}

// PLMC_cuda / PPMC_VL / PLMP_cuda over a whole grid with the reference's thread guard and
// interface layout (plmc_cuda.cu:24-300, ppmc_cuda.cu:540-695): Q_L at cell i,
// Q_R at cell i-1 along dir, conserved variables in global momentum slots.
template <int RK, bool DE = false>
__global__ void ReconGridKernel(const double *U, double *QL, double *QR, int nx, int ny, int nz, double gamma, int dir)
{
  constexpr int r  = ReconTraits<RK>::kRadius;
  constexpr int NS = 2 * r + 1;
  const size_t nc  = size_t(nx) * ny * nz;
  long t           = long(blockIdx.x) * blockDim.x + threadIdx.x;
  int zid = int(t / (long(nx) * ny)), yid = int((t - long(zid) * nx * ny) / nx);
  int xid = int(t - long(zid) * nx * ny - long(yid) * nx);
  int o1, o2, o3;
  DirSlots(dir, o1, o2, o3);
  int ix = dir == 0, iy = dir == 1, iz = dir == 2;
  if constexpr (RK == kPLMP || RK == kPPMP) {  // plmp_cuda.cu:80-103, ppmp_cuda.cu:125-149: along dir only
    const int a = ix ? xid : (iy ? yid : zid), na = ix ? nx : (iy ? ny : nz);
    if (xid >= nx || yid >= ny || zid >= nz || a < r || a >= na - r - 1) return;
  } else {
    if (ThreadGuard(r + 1, nx, ny, nz, xid, yid, zid)) return;
  }
  Prim<double> s[NS];
  double ge[NS];
#pragma unroll
  for (int m = 0; m < NS; m++) {
    size_t id = Lin(xid + (m - r) * ix, yid + (m - r) * iy, zid + (m - r) * iz, nx, ny);
    if constexpr (RK == kPLMP || RK == kPPMP) {  // plmp_cuda.cu:104-132
      const double d = U[id], vx = U[o1 * nc + id] / d, vy = U[o2 * nc + id] / d, vz = U[o3 * nc + id] / d;
      if constexpr (DE) {  // plmp_cuda.cu:116-121: Get_Pressure_From_DE
        const double E = U[4 * nc + id], E_kin = 0.5 * d * (vx * vx + vy * vy + vz * vz), g = U[5 * nc + id];
        s[m]  = {d, vx, vy, vz, Fmax(PressureFromDE(E, E - E_kin, g, gamma), kTinyNumber)};
        ge[m] = g / d;
      } else {
        s[m] = {d, vx, vy, vz, Fmax((U[4 * nc + id] - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0),
                                    kTinyNumber)};
      }
    } else if constexpr (DE) {  // Load_Data's DE branch, reconstruction.h:179-191
      const double d = U[id], vx = U[o1 * nc + id] / d, vy = U[o2 * nc + id] / d, vz = U[o3 * nc + id] / d;
      const double E = U[4 * nc + id], g = U[5 * nc + id];
      s[m]  = {d, vx, vy, vz, PressureFromDE(E, E - 0.5 * d * SquareMag(vx, vy, vz), g, gamma)};
      ge[m] = g / d;
    } else {
      s[m] = LoadPrim(U[id], U[o1 * nc + id], U[o2 * nc + id], U[o3 * nc + id], U[4 * nc + id], gamma);
    }
  }
  Prim<double> L, R;
  double gL = 0.0, gR = 0.0;
  if constexpr (RK == kPPMC) {
    ReconPpmc<PolicyExact>(s, gamma, L, R);
    if constexpr (DE) ReconPpmcScalar<PolicyExact>(ge, gL, gR);
  } else if constexpr (RK == kPLMP) {
    ReconPlmp<PolicyExact>(s, L, R);
    if constexpr (DE) ReconPlmpScalar<PolicyExact>(ge, gL, gR);
  } else if constexpr (RK == kPPMP) {
    ReconPpmp<PolicyExact>(s, L, R);
    if constexpr (DE) ReconPpmpScalar<PolicyExact>(ge, gL, gR);
  } else {
    ReconPlmc<PolicyExact>(s, gamma, L, R);
    if constexpr (DE) ReconPlmcScalar<PolicyExact>(ge, gL, gR);
  }
  const double inv_gm1 = 1.0 / (gamma - 1.0);
  auto write = [&](const Prim<double> &q, double *Q, size_t id) {  // Write_Data, reconstruction.h:918-958
    Q[id]           = q.d;
    Q[o1 * nc + id] = q.d * q.vx;
    Q[o2 * nc + id] = q.d * q.vy;
    Q[o3 * nc + id] = q.d * q.vz;
    if constexpr (RK == kPLMP || RK == kPPMP) {  // plmp_cuda.cu:224-225, ppmp_cuda.cu:645-650
      Q[4 * nc + id] = q.p / (gamma - 1.0) + 0.5 * q.d * (q.vx * q.vx + q.vy * q.vy + q.vz * q.vz);
    } else {
      Q[4 * nc + id] = EnergyFromPrim<PolicyExact>(q, gamma, inv_gm1);
    }
  };
  write(L, QL, Lin(xid, yid, zid, nx, ny));
  write(R, QR, Lin(xid - ix, yid - iy, zid - iz, nx, ny));
  if constexpr (DE) {  // Write_Data, reconstruction.h:960-962 (plmp_cuda.cu:228-229 likewise)
    QL[5 * nc + Lin(xid, yid, zid, nx, ny)]                = L.d * gL;
    QR[5 * nc + Lin(xid - ix, yid - iy, zid - iz, nx, ny)] = R.d * gR;
  }
}

// Calculate_HLLC_Fluxes_CUDA over a whole grid (hllc_cuda.cu:19-307), or, under
// HllPolicy / RoePolicy / ExactPolicy, Calculate_HLL_Fluxes_CUDA (hll_cuda.cu:19-268) /
// Calculate_Roe_Fluxes_CUDA (roe_cuda.cu:19-400) / Calculate_Exact_Fluxes_CUDA (exact_cuda.cu:20-334)
template <class P, bool DE = false>
__global__ void RiemannGridKernel(const double *QL, const double *QR, double *F, int nx, int ny, int nz, double gamma,
                                  int dir)
{
  const size_t nc = size_t(nx) * ny * nz;
  size_t t        = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nc) return;
  int o1, o2, o3;
  DirSlots(dir, o1, o2, o3);
  RState<double> l = RStateFromCons<PolicyExact, DE>(QL[t], QL[o1 * nc + t], QL[o2 * nc + t], QL[o3 * nc + t],
                                                     QL[4 * nc + t], gamma, DE ? QL[5 * nc + t] : 0.0);
  RState<double> r = RStateFromCons<PolicyExact, DE>(QR[t], QR[o1 * nc + t], QR[o2 * nc + t], QR[o3 * nc + t],
                                                     QR[4 * nc + t], gamma, DE ? QR[5 * nc + t] : 0.0);
  double fge       = 0.0;
  Cons<double> f   = RiemannFlux<P, DE>(l, r, gamma, &fge);
  F[t]             = f.d;
  F[o1 * nc + t]   = f.mx;
  F[o2 * nc + t]   = f.my;
  F[o3 * nc + t]   = f.mz;
  F[4 * nc + t]    = f.E;
  if constexpr (DE) F[5 * nc + t] = fge;
}

// ============================================================================ launchers
static inline unsigned Blocks(long n, int tpb) { return unsigned((n + tpb - 1) / tpb); }

hipError_t LaunchGhostFill(double *U, int nx, int ny, int nz, int n_fields, int ng, const int flags[6],
                           hipStream_t s, bool mhd)
{
  GhostFlags fl;
  for (int q = 0; q < 6; q++) fl.f[q] = flags[q];
  fl.diode = mhd ? 0 : 1;
  if (nx <= 1) return hipSuccess;
  const int jl = ny > 1 ? ng : 0, ju = ny > 1 ? ny - ng : ny;
  const int kl = nz > 1 ? ng : 0, ku = nz > 1 ? nz - ng : nz;
  long n_x = long(2 * ng) * (ju - jl) * (ku - kl);
  long n_y = ny > 1 ? long(nx) * (2 * ng) * (ku - kl) : 0;
  long n_z = nz > 1 ? long(nx) * ny * (2 * ng) : 0;
  long tot = n_x + n_y + n_z;
  if (tot == 0) return hipSuccess;
  hipLaunchKernelGGL(GhostFillKernel, dim3(Blocks(tot, 256)), dim3(256), 0, s, U, nx, ny, nz, n_fields, ng, fl, n_x,
                     n_y, n_z);
  return hipGetLastError();
}

hipError_t LaunchGhostSlab(double *U, int nx, int ny, int nz, int n_fields, int ng, const int flags[6], int isize,
                           int jsize, int ksize, int imin, int jmin, int kmin, int dir, hipStream_t s, bool mhd)
{
  GhostFlags fl;
  for (int q = 0; q < 6; q++) fl.f[q] = flags[q];
  fl.diode = mhd ? 0 : 1;
  long tot = long(isize) * jsize * ksize;
  if (tot <= 0) return hipSuccess;
  hipLaunchKernelGGL(GhostSlabKernel, dim3(Blocks(tot, 256)), dim3(256), 0, s, U, nx, ny, nz, n_fields, ng, fl, isize,
                     jsize, ksize, imin, jmin, kmin, dir);
  return hipGetLastError();
}

hipError_t LaunchCalcDt(const double *U, int nx, int ny, int nz, int ng, double dx, double dy, double dz,
                        double gamma, unsigned long long *dti, int num_cu, hipStream_t s)
{
  // Calc_dt_GPU's dimension test (hydro_cuda.cu:572-592); every other shape than 1D and
  // 2D, including a single cell, takes the 3D kernel
  if (nx > 1 && ny == 1 && nz == 1) {
    hipLaunchKernelGGL(CalcDt1DKernel, dim3(1), dim3(256), 0, s, U, nx, ng, dx, gamma, dti);
  } else if (nx > 1 && ny > 1 && nz == 1) {
    long nc     = long(nx) * ny;
    unsigned nb = unsigned(std::min<long>(Blocks(nc, 256), long(num_cu) * 8));
    hipLaunchKernelGGL(CalcDt2DKernel, dim3(nb), dim3(256), 0, s, U, nx, ny, ng, dx, dy, gamma, dti);
  } else {
    long nc     = long(nx) * ny * nz;
    unsigned nb = unsigned(std::min<long>(Blocks(nc, 256), long(num_cu) * 8));
    hipLaunchKernelGGL(CalcDt3DKernel, dim3(nb), dim3(256), 0, s, U, nx, ny, nz, ng, dx, dy, dz, gamma, dti);
  }
  return hipGetLastError();
}

hipError_t LaunchTemperatureFloor(double *U, int nx, int ny, int nz, int ng, double U_floor, int n_fields,
                                  hipStream_t s)
{
  long nc = long(nx) * ny * nz;
  hipLaunchKernelGGL(TemperatureFloorKernel, dim3(Blocks(nc, 256)), dim3(256), 0, s, U, nx, ny, nz, ng, U_floor,
                     int(n_fields == 6));
  return hipGetLastError();
}

hipError_t LaunchGasEnergyInit(double *U, long n_cells, hipStream_t s)
{
  if (n_cells <= 0) return hipSuccess;
  hipLaunchKernelGGL(GasEnergyInitKernel, dim3(Blocks(n_cells, 256)), dim3(256), 0, s, U, n_cells);
  return hipGetLastError();
}

hipError_t LaunchDeSelect(double *U, int nx, int ny, int nz, int ng, hipStream_t s)
{
  long nr = long(nx - 2 * ng) * (ny - 2 * ng) * (nz - 2 * ng);
  if (nr <= 0) return hipSuccess;
  hipLaunchKernelGGL(DeSelectKernel, dim3(Blocks(nr, 256)), dim3(256), 0, s, U, nx, ny, nz, ng);
  return hipGetLastError();
}

hipError_t LaunchDeSync(double *U, int nx, int ny, int nz, int ng, int flags, double U_floor, double dx, double dy,
                        double dz, double gamma, unsigned long long *dti, hipStream_t s)
{
  long nr = long(nx - 2 * ng) * (ny - 2 * ng) * (nz - 2 * ng);
  if (nr <= 0) return hipSuccess;
  hipLaunchKernelGGL(DeSyncKernel, dim3(Blocks(nr, 256)), dim3(256), 0, s, U, nx, ny, nz, ng, flags, U_floor, dx, dy,
                     dz, gamma, dti);
  return hipGetLastError();
}

hipError_t LaunchDtAdvance(unsigned long long *dti, double *dt_dev, double *t_dev, double tout, double cfl,
                           int *n_step, hipStream_t s, bool take)
{
  hipLaunchKernelGGL(DtAdvanceKernel, dim3(1), dim3(1), 0, s, dti, dt_dev, t_dev, tout, cfl, n_step, int(take));
  return hipGetLastError();
}

hipError_t LaunchEnsembleClock(BoxClock *clk, const BoxDesc *box, int n_boxes, double cfl, unsigned int *head,
                               hipStream_t s)
{
  hipLaunchKernelGGL(EnsembleClockKernel, dim3(Blocks(std::max(n_boxes, 1), 256)), dim3(256), 0, s, clk, box, n_boxes,
                     cfl, head);
  return hipGetLastError();
}

hipError_t LaunchTurbulenceIc(double *U, int n, int ng, int n_modes, const double *modes, double vscale, double rho,
                              double p, double gamma, hipStream_t s)
{
  const long N = n + 2 * ng;
  hipLaunchKernelGGL(TurbulenceIcKernel, dim3(Blocks(N * N * N, 256)), dim3(256), 0, s, U, n, ng, n_modes, modes,
                     vscale, rho, p, gamma);
  return hipGetLastError();
}

// dti_reset: the inverse-dt accumulator, zeroed by stage 1's first thread (null: left alone);
// dti: the same accumulator when this launch sequence also reduces into it (null: a later
// launch does).  Nothing is cleared with a memset, so the sequence can be graph-captured.
template <class P, bool DE>
static hipError_t LaunchVl1dP(double *U, double *W, int n, int ng, int recon, double dx, const double *dt_dev,
                              double dt, double gamma, unsigned long long *dti, unsigned long long *dti_reset,
                              const GravSpec &grav, hipStream_t s)
{
  const unsigned nb = Blocks(n, 256);
  hipLaunchKernelGGL((Vl1dStage1Kernel<P, DE>), dim3(nb), dim3(256), 0, s, U, W, n, dx, dt_dev, dt, gamma, dti_reset);
  if constexpr (DE) {
    hipLaunchKernelGGL(Vl1dPdvKernel, dim3(nb), dim3(256), 0, s, U, n, ng, dx, dt_dev, dt, gamma);
    if (recon == kPPMC) {
      hipLaunchKernelGGL((Vl1dStage2Kernel<P, kPPMC, true>), dim3(nb), dim3(256), 0, s, U, W, n, ng, dx, dt_dev, dt,
                         gamma, nullptr, grav);
    } else if (recon == kPLMC) {
      hipLaunchKernelGGL((Vl1dStage2Kernel<P, kPLMC, true>), dim3(nb), dim3(256), 0, s, U, W, n, ng, dx, dt_dev, dt,
                         gamma, nullptr, grav);
    } else if (recon == kPLMP) {
      hipLaunchKernelGGL((Vl1dStage2Kernel<P, kPLMP, true>), dim3(nb), dim3(256), 0, s, U, W, n, ng, dx, dt_dev, dt,
                         gamma, nullptr, grav);
    } else if (recon == kPPMP) {
      hipLaunchKernelGGL((Vl1dStage2Kernel<P, kPPMP, true>), dim3(nb), dim3(256), 0, s, U, W, n, ng, dx, dt_dev, dt,
                         gamma, nullptr, grav);
    } else if (recon == kPCM) {
      hipLaunchKernelGGL((Vl1dStage2Kernel<P, kPCM, true>), dim3(nb), dim3(256), 0, s, U, W, n, ng, dx, dt_dev, dt,
                         gamma, nullptr, grav);
    } else {
      return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(Vl1dSelectKernel, dim3(nb), dim3(256), 0, s, U, n, ng);
    hipLaunchKernelGGL(Vl1dSyncKernel, dim3(nb), dim3(256), 0, s, U, n, ng, dx, gamma, dti);
    return hipGetLastError();
  } else {
    // one block when a dt reduction is fused so the block max is the grid max
    const unsigned nb2 = dti ? 1u : nb;
    const int tpb      = dti ? 1024 : 256;
    if (recon == kPPMC) {
      hipLaunchKernelGGL((Vl1dStage2Kernel<P, kPPMC, false>), dim3(nb2), dim3(tpb), 0, s, U, W, n, ng, dx, dt_dev, dt,
                         gamma, dti, grav);
    } else if (recon == kPLMC) {
      hipLaunchKernelGGL((Vl1dStage2Kernel<P, kPLMC, false>), dim3(nb2), dim3(tpb), 0, s, U, W, n, ng, dx, dt_dev, dt,
                         gamma, dti, grav);
    } else if (recon == kPLMP) {
      hipLaunchKernelGGL((Vl1dStage2Kernel<P, kPLMP, false>), dim3(nb2), dim3(tpb), 0, s, U, W, n, ng, dx, dt_dev, dt,
                         gamma, dti, grav);
    } else if (recon == kPPMP) {
      hipLaunchKernelGGL((Vl1dStage2Kernel<P, kPPMP, false>), dim3(nb2), dim3(tpb), 0, s, U, W, n, ng, dx, dt_dev, dt,
                         gamma, dti, grav);
    } else if (recon == kPCM) {
      hipLaunchKernelGGL((Vl1dStage2Kernel<P, kPCM, false>), dim3(nb2), dim3(tpb), 0, s, U, W, n, ng, dx, dt_dev, dt,
                         gamma, dti, grav);
    } else {
      return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
}

template <class B, bool DE>
static hipError_t LaunchVl1dS(int riemann, double *U, double *W, int n, int ng, int recon, double dx,
                              const double *dt_dev, double dt, double gamma, unsigned long long *dti,
                              unsigned long long *dti_reset, const GravSpec &grav, hipStream_t s)
{
  if (riemann == kRiemannHll) return LaunchVl1dP<HllPolicy<B>, DE>(U, W, n, ng, recon, dx, dt_dev, dt, gamma, dti, dti_reset, grav, s);
  if (riemann == kRiemannRoe) return LaunchVl1dP<RoePolicy<B>, DE>(U, W, n, ng, recon, dx, dt_dev, dt, gamma, dti, dti_reset, grav, s);
  if (riemann == kRiemannExact)
    return LaunchVl1dP<ExactPolicy<B>, DE>(U, W, n, ng, recon, dx, dt_dev, dt, gamma, dti, dti_reset, grav, s);
  if (riemann == kRiemannHllc) return LaunchVl1dP<B, DE>(U, W, n, ng, recon, dx, dt_dev, dt, gamma, dti, dti_reset, grav, s);
  return hipErrorInvalidValue;
}

hipError_t LaunchVl1d(double *U, double *W, int n, int ng, int recon, int riemann, bool de, int form, double dx,
                      const double *dt_dev, double dt, double gamma, unsigned long long *dti,
                      unsigned long long *dti_reset, const GravSpec &grav, hipStream_t s)
{
  if (dti && n > 1024 && !de) return hipErrorInvalidValue;
  if (recon != kPPMC && recon != kPLMC && recon != kPLMP && recon != kPPMP && recon != kPCM) return hipErrorInvalidValue;
  switch (form) {
    case PolicyExact::kId:
      return de ? LaunchVl1dS<PolicyExact, true>(riemann, U, W, n, ng, recon, dx, dt_dev, dt, gamma, dti, dti_reset, grav, s)
                : LaunchVl1dS<PolicyExact, false>(riemann, U, W, n, ng, recon, dx, dt_dev, dt, gamma, dti, dti_reset, grav, s);
    case PolicyFast::kId:
      return de ? LaunchVl1dS<PolicyFast, true>(riemann, U, W, n, ng, recon, dx, dt_dev, dt, gamma, dti, dti_reset, grav, s)
                : LaunchVl1dS<PolicyFast, false>(riemann, U, W, n, ng, recon, dx, dt_dev, dt, gamma, dti, dti_reset, grav, s);
    case PolicyMixed::kId:
      return de ? LaunchVl1dS<PolicyMixed, true>(riemann, U, W, n, ng, recon, dx, dt_dev, dt, gamma, dti, dti_reset, grav, s)
                : LaunchVl1dS<PolicyMixed, false>(riemann, U, W, n, ng, recon, dx, dt_dev, dt, gamma, dti, dti_reset, grav, s);
  }
  return hipErrorInvalidValue;
}

template <class P, bool DE>
static hipError_t LaunchVl2dP(double *U, double *scratch, int nx, int ny, int ng, int recon, double dx, double dy,
                              const double *dt_dev, double dt, double gamma, unsigned long long *dti,
                              const GravSpec &grav, hipStream_t s)
{
  const long nc = long(nx) * ny, fs = 6 * nc;  // 6 fields of room per array (GasEnergy last with DE)
  double *Uh = scratch, *QLx = scratch + fs, *QRx = scratch + 2 * fs, *QLy = scratch + 3 * fs;
  double *QRy = scratch + 4 * fs, *Fx = scratch + 5 * fs, *Fy = scratch + 6 * fs;
  const unsigned nb = Blocks(nc, 256), nb64 = Blocks(nc, 64);
  hipLaunchKernelGGL((Vl2dPcmFluxKernel<P, DE>), dim3(nb), dim3(256), 0, s, U, Fx, Fy, nx, ny, gamma, dti);
  hipLaunchKernelGGL((Vl2dHalfKernel<DE>), dim3(nb), dim3(256), 0, s, U, Uh, Fx, Fy, nx, ny, dx, dy, dt_dev, dt,
                     gamma);
  if (recon == kPCM) {
    hipLaunchKernelGGL((Vl2dPcmFluxKernel<P, DE>), dim3(nb), dim3(256), 0, s, Uh, Fx, Fy, nx, ny, gamma, nullptr);
  } else if constexpr (!P::kExact) {  // the fast and mixed forms: one launch, no interface arrays
    if (recon == kPPMC) {
      hipLaunchKernelGGL((Vl2dFlux2Kernel<P, kPPMC, DE>), dim3(nb), dim3(256), 0, s, Uh, Fx, Fy, nx, ny, ng, gamma);
    } else if (recon == kPLMC) {
      hipLaunchKernelGGL((Vl2dFlux2Kernel<P, kPLMC, DE>), dim3(nb), dim3(256), 0, s, Uh, Fx, Fy, nx, ny, ng, gamma);
    } else if (recon == kPLMP) {
      hipLaunchKernelGGL((Vl2dFlux2Kernel<P, kPLMP, DE>), dim3(nb), dim3(256), 0, s, Uh, Fx, Fy, nx, ny, ng, gamma);
    } else if (recon == kPPMP) {
      hipLaunchKernelGGL((Vl2dFlux2Kernel<P, kPPMP, DE>), dim3(nb), dim3(256), 0, s, Uh, Fx, Fy, nx, ny, ng, gamma);
    } else {
      return hipErrorInvalidValue;
    }
  } else {
    if constexpr (DE) {
      auto both = [&](auto rk) {
        constexpr int RK = decltype(rk)::value;
        hipLaunchKernelGGL((ReconGridKernel<RK, true>), dim3(nb64), dim3(64), 0, s, Uh, QLx, QRx, nx, ny, 1, gamma, 0);
        hipLaunchKernelGGL((ReconGridKernel<RK, true>), dim3(nb64), dim3(64), 0, s, Uh, QLy, QRy, nx, ny, 1, gamma, 1);
      };
      if (recon == kPPMC) {
        both(std::integral_constant<int, kPPMC>{});
      } else if (recon == kPLMC) {
        both(std::integral_constant<int, kPLMC>{});
      } else if (recon == kPLMP) {
        both(std::integral_constant<int, kPLMP>{});
      } else {
        both(std::integral_constant<int, kPPMP>{});
      }
    } else {
      hipError_t e = LaunchReconGrid(recon, Uh, QLx, QRx, nx, ny, 1, gamma, 0, s);
      if (e == hipSuccess) e = LaunchReconGrid(recon, Uh, QLy, QRy, nx, ny, 1, gamma, 1, s);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((RiemannGridKernel<P, DE>), dim3(nb64), dim3(64), 0, s, QLx, QRx, Fx, nx, ny, 1, gamma, 0);
    hipLaunchKernelGGL((RiemannGridKernel<P, DE>), dim3(nb64), dim3(64), 0, s, QLy, QRy, Fy, nx, ny, 1, gamma, 1);
  }
  if constexpr (DE) hipLaunchKernelGGL(Vl2dPdvKernel, dim3(nb), dim3(256), 0, s, U, nx, ny, ng, dx, dy, dt_dev, dt, gamma);
  hipLaunchKernelGGL((Vl2dUpdateKernel<DE>), dim3(nb), dim3(256), 0, s, U, Fx, Fy, nx, ny, ng, dx, dy, dt_dev, dt,
                     gamma, dti, grav);
  if constexpr (DE) {
    hipLaunchKernelGGL(Vl2dSelectKernel, dim3(nb), dim3(256), 0, s, U, nx, ny, ng);
    hipLaunchKernelGGL(Vl2dSyncKernel, dim3(nb), dim3(256), 0, s, U, nx, ny, ng, dx, dy, gamma, dti);
  }
  return hipGetLastError();
}

template <class B, bool DE>
static hipError_t LaunchVl2dS(int riemann, double *U, double *scratch, int nx, int ny, int ng, int recon, double dx,
                              double dy, const double *dt_dev, double dt, double gamma, unsigned long long *dti,
                              const GravSpec &grav, hipStream_t s)
{
  if (riemann == kRiemannHll)
    return LaunchVl2dP<HllPolicy<B>, DE>(U, scratch, nx, ny, ng, recon, dx, dy, dt_dev, dt, gamma, dti, grav, s);
  if (riemann == kRiemannRoe)
    return LaunchVl2dP<RoePolicy<B>, DE>(U, scratch, nx, ny, ng, recon, dx, dy, dt_dev, dt, gamma, dti, grav, s);
  if (riemann == kRiemannExact)
    return LaunchVl2dP<ExactPolicy<B>, DE>(U, scratch, nx, ny, ng, recon, dx, dy, dt_dev, dt, gamma, dti, grav, s);
  if (riemann == kRiemannHllc)
    return LaunchVl2dP<B, DE>(U, scratch, nx, ny, ng, recon, dx, dy, dt_dev, dt, gamma, dti, grav, s);
  return hipErrorInvalidValue;
}

// VL_Algorithm_2D_CUDA (VL_2D_cuda.cu:24-168) in the form `form'; scratch holds 7 x 6 x nx*ny
// doubles (half state, four interface arrays, two flux arrays, 6 fields of room each).
hipError_t LaunchVl2d(double *U, double *scratch, int nx, int ny, int ng, int recon, int riemann, bool de, int form,
                      double dx, double dy, const double *dt_dev, double dt, double gamma, unsigned long long *dti,
                      const GravSpec &grav, hipStream_t s)
{
  if (recon != kPPMC && recon != kPLMC && recon != kPLMP && recon != kPPMP && recon != kPCM) return hipErrorInvalidValue;
  switch (form) {
    case PolicyExact::kId:
      return de ? LaunchVl2dS<PolicyExact, true>(riemann, U, scratch, nx, ny, ng, recon, dx, dy, dt_dev, dt, gamma, dti, grav, s)
                : LaunchVl2dS<PolicyExact, false>(riemann, U, scratch, nx, ny, ng, recon, dx, dy, dt_dev, dt, gamma, dti, grav, s);
    case PolicyFast::kId:
      return de ? LaunchVl2dS<PolicyFast, true>(riemann, U, scratch, nx, ny, ng, recon, dx, dy, dt_dev, dt, gamma, dti, grav, s)
                : LaunchVl2dS<PolicyFast, false>(riemann, U, scratch, nx, ny, ng, recon, dx, dy, dt_dev, dt, gamma, dti, grav, s);
    case PolicyMixed::kId:
      return de ? LaunchVl2dS<PolicyMixed, true>(riemann, U, scratch, nx, ny, ng, recon, dx, dy, dt_dev, dt, gamma, dti, grav, s)
                : LaunchVl2dS<PolicyMixed, false>(riemann, U, scratch, nx, ny, ng, recon, dx, dy, dt_dev, dt, gamma, dti, grav, s);
  }
  return hipErrorInvalidValue;
}

hipError_t LaunchReconGrid(int recon, const double *U, double *QL, double *QR, int nx, int ny, int nz, double gamma,
                           int dir, hipStream_t s)
{
  long nc = long(nx) * ny * nz;
  if (recon == kPPMC) {
    hipLaunchKernelGGL(ReconGridKernel<kPPMC>, dim3(Blocks(nc, 64)), dim3(64), 0, s, U, QL, QR, nx, ny, nz, gamma,
                       dir);
  } else if (recon == kPLMC) {
    hipLaunchKernelGGL(ReconGridKernel<kPLMC>, dim3(Blocks(nc, 64)), dim3(64), 0, s, U, QL, QR, nx, ny, nz, gamma,
                       dir);
  } else if (recon == kPLMP) {
    hipLaunchKernelGGL(ReconGridKernel<kPLMP>, dim3(Blocks(nc, 64)), dim3(64), 0, s, U, QL, QR, nx, ny, nz, gamma,
                       dir);
  } else if (recon == kPPMP) {
    hipLaunchKernelGGL(ReconGridKernel<kPPMP>, dim3(Blocks(nc, 64)), dim3(64), 0, s, U, QL, QR, nx, ny, nz, gamma,
                       dir);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t LaunchRiemannGrid(int riemann, const double *QL, const double *QR, double *F, int nx, int ny, int nz,
                             double gamma, int dir, hipStream_t s)
{
  long nc = long(nx) * ny * nz;
  if (riemann == kRiemannHll) {
    hipLaunchKernelGGL(RiemannGridKernel<HllPolicy<PolicyExact>>, dim3(Blocks(nc, 64)), dim3(64), 0, s, QL, QR, F, nx,
                       ny, nz, gamma, dir);
  } else if (riemann == kRiemannRoe) {
    hipLaunchKernelGGL(RiemannGridKernel<RoePolicy<PolicyExact>>, dim3(Blocks(nc, 64)), dim3(64), 0, s, QL, QR, F, nx,
                       ny, nz, gamma, dir);
  } else if (riemann == kRiemannExact) {
    hipLaunchKernelGGL(RiemannGridKernel<ExactPolicy<PolicyExact>>, dim3(Blocks(nc, 64)), dim3(64), 0, s, QL, QR, F,
                       nx, ny, nz, gamma, dir);
  } else if (riemann == kRiemannHllc) {
    hipLaunchKernelGGL(RiemannGridKernel<PolicyExact>, dim3(Blocks(nc, 64)), dim3(64), 0, s, QL, QR, F, nx, ny, nz,
                       gamma, dir);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace chl
