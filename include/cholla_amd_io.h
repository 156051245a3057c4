/* cholla_amd_io.h -- snapshots in Cholla's HDF5 format (libcholla_amd_io.so).
 *
 * Replaces the HDF5 branch of the reference's Output_Data / Output_Float32 and
 * Read_Grid (src/io/io.cpp:198-378, 2255-2316) for the hydro, dual-energy and MHD states
 * this library steps: one file per call, the header attributes of Write_Header_HDF5
 * (io.cpp:605-735) and the real cells of every conserved field as big-endian datasets in
 * the reference's z-fastest order (io.cpp:1232-1268, io_gpu.cu:44-60).  The state stays on
 * the device: each field is packed there (chl_pack_real_cells) and crosses PCIe once.
 * Depends on libcholla_amd.so and the HDF5 C library. */
#ifndef CHOLLA_AMD_IO_H
#define CHOLLA_AMD_IO_H

#ifdef __cplusplus
extern "C" {
#endif

#define CHL_ERR_IO 5 /* file or HDF5 error (message in chl_io_last_error) */

typedef struct chl_io_header {
  int nx, ny, nz;      /* grid totals incl. ghosts (Header::nx, ny, nz); ny = nz = 1 for 1D */
  int n_ghost;         /* Header::n_ghost */
  int n_fields;        /* 5; 6 with dual_energy (GasEnergy last); 8 with mhd */
  int dual_energy;     /* DE build: /GasEnergy is written and read */
  int mhd;             /* MHD build: /magnetic_x,y,z (one more face along their own axis) */
  int float32;         /* write: datasets as big-endian float32 (Output_Float32, hydro fields only) */
  int n_step;          /* Header::n_step (read back by chl_io_read_hdf5) */
  double gamma;        /* gama (read back) */
  double t, dt;        /* Header::t (read back), Header::dt */
  double bounds[3];    /* lower corner of the domain (xbound, ybound, zbound) */
  double domain[3];    /* domain lengths (xdglobal, ydglobal, zdglobal) */
  double dx[3];        /* cell widths */
  const char *macro_flags; /* "Macro Flags" attribute (NULL: empty) */
} chl_io_header;

const char *chl_io_last_error(void);

/* Output_Data with HDF5 (io.cpp:198-245): H5Fcreate(path, truncate), Write_Header_HDF5
 * (io.cpp:605-735) and Write_Grid_HDF5 (io.cpp:1350-1497) with every field written
 * (Output_Complete_Data).  With h->float32, Output_Float32 (io.cpp:267-378) instead: the
 * hydro fields (and GasEnergy) as float32, 3D only, no magnetic fields.  d_U: the device
 * state in the library's layout. */
int chl_io_write_hdf5(const char *path, const double *d_U, const chl_io_header *h);

/* Read_Grid with HDF5 (io.cpp:2255-2316) and Read_Grid_HDF5 (io.cpp:2508-2700): reads
 * gamma, t and n_step into h and the real cells of every field into d_U (ghost cells are
 * left as they are: the caller sets boundaries next, as main.cpp does).  The file's
 * datasets must have the extents h describes (checked; the reference assumes them). */
int chl_io_read_hdf5(const char *path, double *d_U, chl_io_header *h);

#ifdef __cplusplus
}
#endif

#endif /* CHOLLA_AMD_IO_H */
