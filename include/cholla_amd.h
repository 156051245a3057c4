/* cholla_amd.h -- C ABI of the MI355X (gfx950) hydro hot path.
 *
 * The library replaces the device side of Cholla's van Leer hydro step: the VL
 * integrator with PLMC/PPMC reconstruction and HLLC fluxes, the conserved update,
 * the temperature floor, the CFL inverse-timestep reduction and the ghost-cell fill.
 * Every entry point takes plain pointers and sizes; device pointers hold the
 * conserved state in Cholla's own layout: field-major doubles,
 *   U[f*n_cells + i + j*nx + k*nx*ny],  f = density, momentum_x, momentum_y,
 *   momentum_z, Energy  (src/grid/grid_enum.h:20-30, src/grid/grid3D.cpp:321-342),
 * with nx, ny, nz the total cell counts including n_ghost ghost cells per side.
 *
 * All calls are synchronous at return (the reference synchronises after every
 * stage, src/utils/gpu.hpp:139-155).  They return CHL_OK or an error code;
 * chl_last_error() gives the message.  The C++ adapter (cholla_dropin.hpp) turns
 * a non-zero code into the reference's abort.
 */
#ifndef CHOLLA_AMD_H
#define CHOLLA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes */
#define CHL_OK 0
#define CHL_ERR_INVALID 1     /* bad argument or unsupported configuration */
#define CHL_ERR_HIP 2         /* HIP runtime error (message in chl_last_error) */
#define CHL_ERR_NO_DEVICE 3   /* no HIP device visible */
#define CHL_ERR_UNSUPPORTED 4 /* option compiled out / not implemented */

/* ---- build options of the reference's make.type files, as run-time values */
#define CHL_RECON_PCM 0  /* PCM  */
#define CHL_RECON_PLMC 1 /* PLMC (builds/make.type.hydro:5) */
#define CHL_RECON_PPMC 2 /* PPMC */
#define CHL_RECON_PLMP 3 /* PLMP: limited in the primitive variables (src/reconstruction/plmp_cuda.cu) */
#define CHL_RECON_PPMP 4 /* PPMP: parabolic in the primitive variables (src/reconstruction/ppmp_cuda.cu) */
#define CHL_RIEMANN_HLLC 0 /* HLLC (builds/make.type.hydro) */
#define CHL_RIEMANN_HLL 1  /* HLL, Davis wave speeds (-DHLL, src/riemann_solvers/hll_cuda.cu) */
#define CHL_RIEMANN_ROE 2  /* Roe with the HLLE fallback (-DROE, src/riemann_solvers/roe_cuda.cu) */
#define CHL_RIEMANN_EXACT 3 /* exact solver, FP64 (-DEXACT, src/riemann_solvers/exact_cuda.cu) */
/* arithmetic forms of the same step */
#define CHL_PREC_FP64_EXACT 0 /* FP64, reference operation order, no FMA contraction */
#define CHL_PREC_FP64 1       /* FP64, shared reciprocals + FMA */
#define CHL_PREC_MIXED 2      /* FP32 reconstruction/Riemann, FP64 state, accumulation, pressure, dt */

/* boundary flags, Parameters::xl_bcnd..zu_bcnd (src/global/global.h:223-228) */
#define CHL_BC_PERIODIC 1
#define CHL_BC_REFLECTIVE 2
#define CHL_BC_TRANSMISSIVE 3

typedef struct chl_hydro_config {
  int nx, ny, nz;        /* total cells per axis incl. ghosts (Header::nx, ny, nz); 1D: ny = nz = 1,
                            2D: nz = 1 (every precision; the 1D half state is FP64 in every form) */
  int n_ghost;           /* Header::n_ghost: PLMC 3, PPMC 4 (src/grid/grid3D.cpp:53-76) */
  int n_fields;          /* 5 (hydro), or 6 with dual_energy (GasEnergy last); no scalars */
  int reconstruction;    /* CHL_RECON_* ; stage 1 is always PCM as in VL.  PCM, PLMP and PPMP: 3D
                            tile shapes 0 and 1 (n_ghost: PCM 2, PLMP 3, PPMP 4) */
  int riemann;           /* CHL_RIEMANN_HLLC, or _HLL / _ROE / _EXACT (3D tile shapes 0 and 1) */
  int precision;         /* CHL_PREC_* */
  int density_floor;     /* DENSITY_FLOOR build option, 0/1 */
  int temperature_floor; /* TEMPERATURE_FLOOR build option, 0/1 */
  int bc_flags[6];       /* xl, xu, yl, yu, zl, zu */
  int tile1, tile2;      /* tile shapes of the two fused stages, -1 = default */
  int kchunk1, kchunk2;  /* z cells per workgroup in the two stages, 0 = default */
  int tile_order;        /* 0: each XCD takes a contiguous run of tiles (default); 1: dispatch order */
  int dual_energy;       /* DE build option, 0/1: GasEnergy as field 5 (n_fields 6); every solver and
                            reconstruction in 1D, 2D and 3D (3D: tile shapes 0 and 1) */
  int positivity_fallback; /* 0/1, PPMC in 3D only; no counterpart in the reference: where a
                              cell's PPMC interface density or pressure comes out non-positive,
                              both of its interfaces in that sweep take the cell-centred values
                              (the reference floors each at TINY_NUMBER, ppmc_cuda.cu:681-684) */
} chl_hydro_config;

typedef struct chl_hydro chl_hydro; /* owns the integrator scratch of one box */

const char *chl_version(void);
const char *chl_last_error(void);
int chl_device_count(int *n);
int chl_device_name(char *buf, int len);

/* ---- device memory helpers for hosts without their own allocator */
int chl_malloc(void **ptr, size_t bytes);
int chl_free(void *ptr);
int chl_memcpy_htod(void *dst, const void *src, size_t bytes);
int chl_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int chl_memcpy_dtod(void *dst, const void *src, size_t bytes);
int chl_synchronize(void);

/* ---- integrator handle.  Replaces the static scratch that VL_Algorithm_3D_CUDA
 * allocates on first call and Free_Memory_VL_3D releases
 * (src/integrators/VL_3D_cuda.cu:58-133, :362-377; src/global/global_cuda.h:18-32). */
int chl_hydro_create(const chl_hydro_config *cfg, chl_hydro **out);
int chl_hydro_destroy(chl_hydro *h);

/* One VL step of the state d_conserved by dt (VL_Algorithm_3D_CUDA,
 * src/integrators/VL_3D_cuda.h:9-11; VL_Algorithm_1D_CUDA, VL_1D_cuda.cu:30 when
 * ny == nz == 1).  Real cells are updated in place; ghost cells are left to
 * chl_set_boundaries, as in the reference. */
int chl_hydro_vl_step(chl_hydro *h, double *d_conserved, double dx, double dy, double dz, double dt, double gamma,
                      double density_floor);

/* chl_hydro_vl_step without a handle and without allocating device memory.  No counterpart
 * in the reference: it is for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  cfg describes a 3D box as for
 * chl_hydro_create; the step runs the same kernels as chl_hydro_vl_step, with its scratch
 * (half-step state, repair marks) in a fixed device array of the library of
 * chl_noalloc_scratch_bytes() bytes on each GPU, held until the process ends; a box that
 * needs more is refused.  The launches go on `stream` (a hipStream_t; NULL = the default
 * stream) and the call returns without synchronizing.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU race. */
int chl_hydro_vl_step_noalloc(const chl_hydro_config *cfg, double *d_conserved, double dx, double dy, double dz,
                              double dt, double gamma, double density_floor, void *stream);
size_t chl_noalloc_scratch_bytes(void);

/* Grid3D::Update_Hydro_Grid (src/grid/grid3D.cpp:474-585) fused: VL step, temperature
 * floor (U_floor as grid3D.cpp:497), and the next inverse timestep (Calc_dt_GPU)
 * returned in *dti_out.  With fill_ghosts != 0 the ghost cells are also refreshed,
 * as the following Set_Boundary_Conditions_Grid would (main.cpp:287-290). */
int chl_hydro_update_grid(chl_hydro *h, double *d_conserved, double dx, double dy, double dz, double dt,
                          double gamma, double density_floor, double U_floor, int fill_ghosts, double *dti_out);

/* The reference driver loop (src/main.cpp:247-290) with dt kept on the device:
 * for up to max_steps steps while t < tout: dt = 0.3/dti clamped to tout, VL step,
 * floors, next dti, t += dt, ghost fill.  The state must hold valid ghost cells on
 * entry.  With use_graph != 0 one step is captured once in a hipGraph and replayed.
 * Returns the final time and the number of steps taken. */
int chl_hydro_run(chl_hydro *h, double *d_conserved, double dx, double dy, double dz, double gamma,
                  double density_floor, double U_floor, double t0, double tout, int max_steps, int use_graph,
                  double *t_out, int *steps_out);

/* Run n_steps device-timed steps (as chl_hydro_run with tout = +inf) and report the
 * elapsed time measured with HIP events on the library's stream. */
int chl_hydro_time_steps(chl_hydro *h, double *d_conserved, double dx, double dy, double dz, double gamma,
                         int n_steps, int use_graph, float *ms_out, double *dt_last_out);

/* Average duration of each fused kernel over n_steps steps, timed with HIP events
 * around each launch on the library's stream (per-kernel roofline figures). */
int chl_hydro_kernel_times(chl_hydro *h, double *d_conserved, double dx, double dy, double dz, double gamma,
                           int n_steps, float *ms_stage1, float *ms_stage2, float *ms_other);

/* Pick the stage tile shapes and z-chunks for this box by timing candidates on a
 * scratch copy of d_conserved (the state itself is not modified). */
int chl_hydro_autotune(chl_hydro *h, const double *d_conserved, double dx, double dy, double dz, double gamma,
                       int steps_per_candidate, int *tile1, int *tile2, int *kchunk1, int *kchunk2);
int chl_hydro_get_config(const chl_hydro *h, chl_hydro_config *out);

/* ---- individual stages, same semantics and arguments as the reference functions */
/* Calc_dt_GPU (src/hydro/hydro_cuda.h:69-70): max over real cells of the inverse
 * crossing time; dt = C_cfl / *dti_out. */
int chl_calc_dt(const double *dev_conserved, int nx, int ny, int nz, int n_ghost, int n_fields, double dx, double dy,
                double dz, double gamma, double *dti_out);
/* Apply_Temperature_Floor (src/hydro/hydro_cuda.h:87) */
int chl_apply_temperature_floor(double *dev_conserved, int nx, int ny, int nz, int n_ghost, int n_fields,
                                double U_floor);
/* SetGhostCells (src/grid/cuda_boundaries.h:15-16): one face slab */
int chl_set_ghost_cells(double *c_head, int nx, int ny, int nz, int n_fields, int n_cells, int n_ghost,
                        const int flags[6], int isize, int jsize, int ksize, int imin, int jmin, int kmin, int dir);
/* Set_Boundary_Conditions (single process, src/grid/boundary_conditions.cpp:96-111):
 * all faces in one launch */
int chl_set_boundaries(double *dev_conserved, int nx, int ny, int nz, int n_fields, int n_ghost, const int flags[6]);

/* STATIC_GRAV build option (src/gravity/static_grav.h; the source of hydro_cuda.cu:59-67,
 * 142-153, 301-315): the analytic field Parameters::custom_grav selects (0 none; 1D and 3D:
 * 1 NFW halo + Miyamoto-Nagai disk; 2D: 1 Gresho, 2 Rayleigh-Taylor, 3 Keplerian, 4 Kuzmin
 * disk + NFW halo) and the grid's offsets and lower bounds, as VL_Algorithm_*_CUDA receive
 * them (x_off.., xbound..).  Applies to every later step of the handle.  3D: every solver and
 * reconstruction, with or without dual energy, stage-2 tile shapes 0 and 1. */
int chl_hydro_set_static_gravity(chl_hydro *h, int custom_grav, int x_off, int y_off, int z_off, double xbound,
                                 double ybound, double zbound);

/* ---- reference-shaped stage kernels (PolicyExact), for checking the device
 * physics against the reference's unit-test vectors */
/* PLMC_cuda / PPMC_VL (plmc_cuda.cu:24, ppmc_cuda.cu:540) on device arrays */
int chl_reconstruct(int reconstruction, const double *dev_conserved, double *dev_bounds_L, double *dev_bounds_R,
                    int nx, int ny, int nz, double gamma, int dir);
/* Calculate_HLLC_Fluxes_CUDA (hllc_cuda.cu:19) on device arrays */
int chl_hllc_fluxes(const double *dev_bounds_L, const double *dev_bounds_R, double *dev_flux, int nx, int ny, int nz,
                    double gamma, int dir);
/* Calculate_HLL_Fluxes_CUDA (hll_cuda.cu:19) on device arrays */
int chl_hll_fluxes(const double *dev_bounds_L, const double *dev_bounds_R, double *dev_flux, int nx, int ny, int nz,
                   double gamma, int dir);
/* Calculate_Roe_Fluxes_CUDA (roe_cuda.cu:19) on device arrays */
int chl_roe_fluxes(const double *dev_bounds_L, const double *dev_bounds_R, double *dev_flux, int nx, int ny, int nz,
                   double gamma, int dir);
/* Calculate_Exact_Fluxes_CUDA (exact_cuda.cu:20) on device arrays */
int chl_exact_fluxes(const double *dev_bounds_L, const double *dev_bounds_R, double *dev_flux, int nx, int ny, int nz,
                     double gamma, int dir);

/* ---- snapshot layout of one field (the HDF5 datasets of src/io/io.cpp).  The grid is
 * x-fastest with ghost cells; a dataset holds the real cells z-fastest,
 * buf[k + j*nz_real + i*nz_real*ny_real] (1D: buf[i]).  nx_real, ny_real, nz_real are the
 * dataset extents; mhd_direction 0/1/2 marks a face-centred magnetic field whose dataset has
 * one more face along that axis and starts at grid index n_ghost - 1 there, -1 a
 * cell-centred field.  3D and 1D grids (2D: CHL_ERR_UNSUPPORTED). */
/* CopyReal3D_GPU_Kernel (double destination), src/io/io_gpu.cu:44-60 (and the 1D branch of
 * Fill_HDF5_Buffer_From_Grid_CPU, io.cpp:1258-1264) */
int chl_pack_real_cells(const double *d_field, double *d_buf, int nx, int ny, int nz, int n_ghost, int nx_real,
                        int ny_real, int nz_real, int mhd_direction);
/* CopyReal3D_GPU_Kernel (float destination), io_gpu.cu:64-83, for the float32 snapshots of
 * Output_Float32 (io.cpp:267-378) */
int chl_pack_real_cells_f32(const double *d_field, float *d_buf, int nx, int ny, int nz, int n_ghost, int nx_real,
                            int ny_real, int nz_real, int mhd_direction);
/* Fill_Grid_From_HDF5_Buffer (io.cpp:2450-2490) and the magnetic-field loops of
 * Read_Grid_HDF5 (io.cpp:2566-2700), on the device */
int chl_unpack_real_cells(const double *d_buf, double *d_field, int nx, int ny, int nz, int n_ghost, int nx_real,
                          int ny_real, int nz_real, int mhd_direction);

/* ---- ensembles of independent periodic boxes.  The reference runs one box per
 * process (src/main.cpp:247-379); this advances many boxes of different sizes
 * together, each with its own CFL dt and clock, with the step of every box fused
 * into one launch per stage ("split") or one launch per step ("persistent"). */
#define CHL_LAUNCH_SPLIT 0
#define CHL_LAUNCH_PERSISTENT 1

typedef struct chl_ensemble_config {
  int n_ghost;           /* as chl_hydro_config; every box shares it */
  int n_fields;          /* 5, or 6 with dual_energy */
  int reconstruction;    /* CHL_RECON_PLMC or CHL_RECON_PPMC */
  int riemann;           /* CHL_RIEMANN_HLLC */
  int precision;         /* CHL_PREC_* */
  int density_floor;     /* 0/1 */
  int temperature_floor; /* 0/1 */
  int launch;            /* CHL_LAUNCH_SPLIT or CHL_LAUNCH_PERSISTENT */
  int tile1, tile2;      /* stage tile shapes, -1 = default (the persistent launch
                            takes the pairs listed by chl_ensemble_persistent_pairs) */
  int kchunk1, kchunk2;  /* z cells per work item, 0 = default */
  int dual_energy;       /* DE build option, 0/1: GasEnergy as field 5 of every box (n_fields 6) */
  int positivity_fallback; /* 0/1, PPMC only: as chl_hydro_config.positivity_fallback */
} chl_ensemble_config;

typedef struct chl_ensemble chl_ensemble; /* owns the scratch and work lists of the boxes */

/* dims: 3*n_boxes totals (nx, ny, nz incl. ghosts) per box; every box is periodic
 * on all faces and needs at least n_ghost real cells per axis. */
int chl_ensemble_create(const chl_ensemble_config *cfg, int n_boxes, const int *dims, chl_ensemble **out);
int chl_ensemble_destroy(chl_ensemble *e);
int chl_ensemble_get_config(const chl_ensemble *e, chl_ensemble_config *out);
/* The (tile1, tile2) pairs compiled for the persistent launch: pairs[2*i], pairs[2*i+1]. */
int chl_ensemble_persistent_pairs(int *pairs, int max_pairs, int *n_pairs);
/* Borrow the boxes' device states (one pointer per box, Cholla layout) and their
 * cell widths (dxyz: dx, dy, dz per box). */
int chl_ensemble_bind(chl_ensemble *e, double *const *d_states, const double *dxyz);
/* Start the clocks: t0 and tout per box (null: 0 and +inf), initial inverse dt
 * from each box's state (Calc_dt_GPU).  States must hold valid ghost cells. */
int chl_ensemble_start(chl_ensemble *e, double gamma, double density_floor, double U_floor, const double *t0,
                       const double *tout);
/* Advance every box n_steps steps (a box at its tout stays put), one hipGraph
 * replay per step with use_graph != 0; *ms_out (may be null) is the HIP-event
 * time of the n_steps steps on the library's stream. */
int chl_ensemble_advance(chl_ensemble *e, int n_steps, int use_graph, float *ms_out);
/* The same steps as a loop over boxes, each box launched on its own (dt kernel,
 * stage 1, stage 2 per box per step): the baseline the fused launch replaces. */
int chl_ensemble_advance_looped(chl_ensemble *e, int n_steps, float *ms_out);
/* Per-box clocks (each array may be null). */
int chl_ensemble_clocks(chl_ensemble *e, double *t, double *dt, int *n_steps);
/* Pick launch form, tiles and z-chunks by timing candidates on a scratch copy of
 * the bound states (states and clocks are left as they were).  Needs a started ensemble. */
int chl_ensemble_autotune(chl_ensemble *e, int steps_per_candidate, chl_ensemble_config *chosen);

/* ---- ideal MHD (the reference's MHD build: -DMHD -DPLMC -DHLLD -DVL, builds/make.type.mhd).
 * The state holds 8 fields: density, momentum x/y/z, Energy, then the face-centred
 * magnetic_x/y/z (magnetic_?[i] on the face i+1/2 of cell i; src/grid/grid_enum.h:60-66). */
typedef struct chl_mhd_config {
  int nx, ny, nz;        /* totals incl. ghosts */
  int n_ghost;           /* >= the reconstruction's ghosts + 1 for the face fields: PCM 3, PLMC 4,
                            PPMC 5 (src/grid/grid3D.cpp:53-76) */
  int precision;         /* CHL_PREC_*; the mixed form runs the reconstruction and HLLD in FP32 */
  int density_floor;     /* 0/1 */
  int temperature_floor; /* 0/1 */
  int bc_flags[6];
  int reconstruction;    /* stage 2: CHL_RECON_PCM, CHL_RECON_PLMC or CHL_RECON_PPMC
                            (src/utils/error_handling.cpp:98-101) */
} chl_mhd_config;

typedef struct chl_mhd chl_mhd; /* owns the MHD integrator scratch of one box */

int chl_mhd_create(const chl_mhd_config *cfg, chl_mhd **out);
int chl_mhd_destroy(chl_mhd *h);
/* VL_Algorithm_3D_CUDA with MHD (src/integrators/VL_3D_cuda.cu:38-360): one step by dt;
 * real cells and their face fields are updated in place, ghosts are left to the caller. */
int chl_mhd_vl_step(chl_mhd *h, double *d_conserved, double dx, double dy, double dz, double dt, double gamma,
                    double density_floor);
/* The driver loop (src/main.cpp:247-290) for the MHD build with dt on the device, as
 * chl_hydro_run: dt = 0.3/dti clamped to tout, step, temperature floor, next dti, ghosts. */
int chl_mhd_run(chl_mhd *h, double *d_conserved, double dx, double dy, double dz, double gamma,
                double density_floor, double U_floor, double t0, double tout, int max_steps, int use_graph,
                double *t_out, int *steps_out);
int chl_mhd_time_steps(chl_mhd *h, double *d_conserved, double dx, double dy, double dz, double gamma, int n_steps,
                       int use_graph, float *ms_out, double *dt_last_out);
/* Calc_dt_GPU with MHD (src/hydro/hydro_cuda.cu:514-597, mhdInverseCrossingTime) */
int chl_mhd_calc_dt(const double *dev_conserved, int nx, int ny, int nz, int n_ghost, double dx, double dy,
                    double dz, double gamma, double *dti_out);
/* SetGhostCells and Set_Boundary_Conditions as the MHD build compiles them: the same slab
 * fills on 8-field states without the transmissive outflow diode, which
 * src/grid/cuda_boundaries.cu:121-147 leaves out under MHD (face fields copy like the rest). */
int chl_mhd_set_ghost_cells(double *c_head, int nx, int ny, int nz, int n_fields, int n_cells, int n_ghost,
                            const int flags[6], int isize, int jsize, int ksize, int imin, int jmin, int kmin, int dir);
int chl_mhd_set_boundaries(double *dev_conserved, int nx, int ny, int nz, int n_ghost, const int flags[6]);
/* mhd::checkMagneticDivergence (src/mhd/magnetic_divergence.cu:79): max |div B| */
int chl_mhd_divergence(const double *dev_conserved, int nx, int ny, int nz, double dx, double dy, double dz,
                       double *max_divergence);
/* reference-shaped MHD stage kernels (FP64, reference operation order) on device arrays */
/* mhd::Calculate_HLLD_Fluxes_CUDA (src/riemann_solvers/hlld_cuda.cu:31) */
int chl_hlld_fluxes(const double *dev_bounds_L, const double *dev_bounds_R, const double *dev_magnetic_face,
                    double *dev_flux, long n_cells, double gamma, int dir);
/* PLMC_cuda with MHD (src/reconstruction/plmc_cuda.cu:24) */
int chl_mhd_reconstruct_plmc(const double *dev_conserved, double *dev_bounds_L, double *dev_bounds_R, int nx, int ny,
                             int nz, double gamma, int dir);
/* PPMC_VL with MHD (src/reconstruction/ppmc_cuda.cu:540) */
int chl_mhd_reconstruct_ppmc(const double *dev_conserved, double *dev_bounds_L, double *dev_bounds_R, int nx, int ny,
                             int nz, double gamma, int dir);
/* mhd::Calculate_CT_Electric_Fields (src/mhd/ct_electric_fields.cu:20) */
int chl_ct_electric_fields(const double *fluxX, const double *fluxY, const double *fluxZ,
                           const double *dev_conserved, double *ctElectricFields, int nx, int ny, int nz);
/* mhd::Update_Magnetic_Field_3D (src/mhd/magnetic_update.cu:21) */
int chl_update_magnetic_field(const double *sourceGrid, double *destinationGrid, const double *ctElectricFields,
                              int nx, int ny, int nz, double dt, double dx, double dy, double dz);

/* ---- problem setup on the host (not timed): initial states of the benchmark boxes */
/* Kelvin-Helmholtz box: n^3 real cells + ng ghosts, unit box, periodic, central slab
 * |y-0.5|<0.25 rho=2 vx=+0.5, outside rho=1 vx=-0.5, p=2.5, vy = 0.01 sin(4 pi x)
 * + 1e-4 u, u ~ U(-1,1) from std::mt19937_64(seed) drawn k, j, i with i fastest.
 * host_U must hold 5*(n+2ng)^3 doubles; ghost cells are filled periodically. */
int chl_ic_kelvin_helmholtz(double *host_U, int n, int ng, double gamma, uint64_t seed);

/* Ensemble layout: per box an edge drawn uniformly from {32, 48, 64, 96, 128}
 * (std::mt19937_64(seed)() % 5) and an rms Mach number from U[0.5, 3]
 * (0.5 + 2.5 * u, u = 53 high bits of the next draw / 2^53), box by box. */
int chl_ensemble_layout(uint64_t seed, int n_boxes, int *edges, double *mach);
/* Modes of the turbulent velocity field: every integer wavevector k with
 * 1 <= |k| <= 4 in the half space (kx > 0, or kx == 0 and ky > 0, or kx == ky == 0
 * and kz > 0), in the order kx, ky, kz ascending (kz fastest).  Per mode two
 * 3-vectors a, b of independent normals (Box-Muller on std::mt19937_64(seed),
 * drawn ax, ay, az, bx, by, bz) with standard deviation |k|^-2, so the per-mode
 * power is proportional to k^-4 and the shell-summed spectrum E(k) (about 4 pi k^2
 * modes per unit shell) to k^-2, then projected solenoidal (a -= k (k.a)/|k|^2,
 * likewise b).  modes: 9 doubles per mode (k, a, b). */
int chl_turbulence_modes(uint64_t seed, double *modes, int max_modes, int *n_modes);
/* Periodic turbulence box on the device: n^3 real cells + ng ghosts, unit box,
 * rho and p uniform, v = s * sum_m [a_m cos(2 pi k_m.x) + b_m sin(2 pi k_m.x)] at
 * the cell centres with s set so the rms |v| over the box is mach * sqrt(gamma p / rho);
 * ghost cells hold their periodic images.  d_U: 5*(n+2ng)^3 doubles on the device. */
int chl_ic_turbulence(double *d_U, int n, int ng, double gamma, double rho, double p, double mach, uint64_t seed);
/* GasEnergy of every cell (ghosts included) set to the internal energy E - kinetic of the
 * state's first five fields: the field the DE build adds, which the reference's initial
 * conditions fill with P/(gamma-1) (e.g. src/grid/initial_conditions.cpp:228-229).
 * d_U: 6 fields. */
int chl_init_gas_energy(double *d_U, int nx, int ny, int nz);
/* mhd::utils::Init_Magnetic_Field_With_Vector_Potential (src/utils/mhd_utilities.cu:22-44):
 * face fields = differences of the edge potential A (3 x n_cells, x/y/z components) for
 * every cell with i, j, k >= 1.  host_U: 8 fields. */
int chl_magnetic_field_from_vector_potential(double *host_U, int nx, int ny, int nz, double dx, double dy, double dz,
                                             const double *A);
/* Orszag-Tang vortex (Grid3D::Orszag_Tang_Vortex, src/grid/initial_conditions.cpp:1926-1984) on a
 * periodic unit box of n^3 real cells + ng ghosts, with a z-dependent divergence-free
 * perturbation through the same potential: A_x, A_y += amplitude / (2 pi) *
 * sum_{m=1..4} a_m / m cos(2 pi m z + phi_m), a ~ U(-1,1), phi ~ U(0, 2 pi) from
 * std::mt19937_64(seed) drawn a_x, phi_x, a_y, phi_y per m.  host_U: 8 x (n+2ng)^3. */
int chl_ic_orszag_tang(double *host_U, int n, int ng, double gamma, uint64_t seed, double amplitude);

#ifdef __cplusplus
}
#endif

#endif /* CHOLLA_AMD_H */
