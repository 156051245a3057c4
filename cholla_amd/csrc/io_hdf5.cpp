// io_hdf5.cpp -- libcholla_amd_io.so: snapshots of device states in Cholla's HDF5 format
// (declared in include/cholla_amd_io.h).  Host code over the HDF5 C library and the C ABI
// of libcholla_amd.so: each field is packed into the dataset order on the device, copied
// to the host once and written; reads run the same way backwards.
#include <hdf5.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cholla_amd.h"
#include "../../include/cholla_amd_io.h"

namespace
{

thread_local std::string g_err;

int Fail(int code, const std::string &msg)
{
  g_err = msg;
  return code;
}

// Unit constants Write_Header_HDF5 records (src/global/global.h:40-48).
constexpr double kTimeUnit     = 3.15569e10;
constexpr double kLengthUnit   = 3.08567758e21;
constexpr double kMassUnit     = 1.98847e33;
constexpr double kDensityUnit  = kMassUnit / (kLengthUnit * kLengthUnit * kLengthUnit);
constexpr double kVelocityUnit = kLengthUnit / kTimeUnit;
constexpr double kEnergyUnit   = kDensityUnit * kVelocityUnit * kVelocityUnit;

// Silences HDF5's own error-stack printing for the duration of one entry point (errors are
// returned as codes with a message instead) and restores the caller's handler.
class QuietHdf5
{
 public:
  QuietHdf5() { H5Eget_auto2(H5E_DEFAULT, &fn_, &data_), H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
  ~QuietHdf5() { H5Eset_auto2(H5E_DEFAULT, fn_, data_); }

 private:
  H5E_auto2_t fn_ = nullptr;
  void *data_     = nullptr;
};

// An HDF5 identifier closed on scope exit.
class Id
{
 public:
  Id(hid_t id, herr_t (*close)(hid_t)) : id_(id), close_(close) {}
  ~Id()
  {
    if (id_ >= 0) close_(id_);
  }
  Id(const Id &)            = delete;
  Id &operator=(const Id &) = delete;
  hid_t get() const { return id_; }
  bool ok() const { return id_ >= 0; }

 private:
  hid_t id_;
  herr_t (*close_)(hid_t);
};

// A device buffer freed on scope exit.
class DeviceBuffer
{
 public:
  ~DeviceBuffer()
  {
    if (p_) chl_free(p_);
  }
  int Allocate(size_t bytes) { return chl_malloc(&p_, bytes); }
  void *get() const { return p_; }

 private:
  void *p_ = nullptr;
};

int WriteAttr(hid_t loc, const char *name, hid_t file_type, hid_t mem_type, const void *data, hsize_t n)
{
  Id space(H5Screate_simple(1, &n, nullptr), H5Sclose);
  if (!space.ok()) return Fail(CHL_ERR_IO, std::string("attribute space for ") + name);
  Id attr(H5Acreate2(loc, name, file_type, space.get(), H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (!attr.ok() || H5Awrite(attr.get(), mem_type, data) < 0)
    return Fail(CHL_ERR_IO, std::string("writing attribute ") + name);
  return CHL_OK;
}

int WriteString(hid_t loc, const char *name, const char *value)
{
  Id type(H5Tcopy(H5T_C_S1), H5Tclose);
  if (!type.ok() || H5Tset_size(type.get(), H5T_VARIABLE) < 0) return Fail(CHL_ERR_IO, "string type");
  const char *v = value ? value : "";
  return WriteAttr(loc, name, type.get(), type.get(), &v, 1);
}

int ReadAttr(hid_t loc, const char *name, hid_t mem_type, void *out)
{
  Id attr(H5Aopen(loc, name, H5P_DEFAULT), H5Aclose);
  if (!attr.ok() || H5Aread(attr.get(), mem_type, out) < 0)
    return Fail(CHL_ERR_IO, std::string("reading attribute ") + name);
  return CHL_OK;
}

#define IO_TRY(expr)              \
  do {                            \
    int rc_ = (expr);             \
    if (rc_ != CHL_OK) return rc_; \
  } while (0)

// One dataset: its name, the field it holds and, for a face-centred magnetic field, its axis.
struct Field {
  const char *name;
  int index;
  int mhd_direction;
};

// The fields of a state (grid_enum.h: hydro, then GasEnergy with DE, or the three face
// fields with MHD), in Write_Grid_HDF5's order (io.cpp:1404-1485).
int Fields(const chl_io_header &h, std::vector<Field> &out)
{
  out = {{"/density", 0, -1}, {"/momentum_x", 1, -1}, {"/momentum_y", 2, -1}, {"/momentum_z", 3, -1},
         {"/Energy", 4, -1}};
  if (h.dual_energy && h.mhd) return Fail(CHL_ERR_UNSUPPORTED, "dual energy with MHD is not built");
  if (h.dual_energy) out.push_back({"/GasEnergy", 5, -1});
  if (h.mhd) out.insert(out.end(), {{"/magnetic_x", 5, 0}, {"/magnetic_y", 6, 1}, {"/magnetic_z", 7, 2}});
  const int nf = 5 + (h.dual_energy ? 1 : 0) + (h.mhd ? 3 : 0);
  if (h.n_fields != nf) return Fail(CHL_ERR_INVALID, "n_fields does not match dual_energy / mhd");
  return CHL_OK;
}

struct Geometry {
  bool one_d;
  int nxr, nyr, nzr;
  size_t n_cells, max_elems;
};

int Check(const chl_io_header &h, Geometry &g)
{
  g.one_d = h.ny == 1 && h.nz == 1;
  if (!g.one_d && (h.ny == 1 || h.nz == 1)) return Fail(CHL_ERR_UNSUPPORTED, "2D grids are not built");
  if (g.one_d && h.mhd) return Fail(CHL_ERR_UNSUPPORTED, "MHD is 3D only");
  if (h.n_ghost < 1 || h.nx <= 2 * h.n_ghost || (!g.one_d && (h.ny <= 2 * h.n_ghost || h.nz <= 2 * h.n_ghost)))
    return Fail(CHL_ERR_INVALID, "grid smaller than its ghost layers");
  g.nxr       = h.nx - 2 * h.n_ghost;
  g.nyr       = g.one_d ? 1 : h.ny - 2 * h.n_ghost;
  g.nzr       = g.one_d ? 1 : h.nz - 2 * h.n_ghost;
  g.n_cells   = size_t(h.nx) * h.ny * h.nz;
  g.max_elems = size_t(g.nxr + (h.mhd ? 1 : 0)) * (g.nyr + (h.mhd ? 1 : 0)) * (g.nzr + (h.mhd ? 1 : 0));
  return CHL_OK;
}

// Dataset extents of one field (rank 1 in 1D, else 3: nx_real, ny_real, nz_real with one
// more face along a magnetic field's axis).
int Extents(const Geometry &g, const Field &f, hsize_t dims[3])
{
  dims[0] = hsize_t(g.nxr + (f.mhd_direction == 0));
  dims[1] = hsize_t(g.nyr + (f.mhd_direction == 1));
  dims[2] = hsize_t(g.nzr + (f.mhd_direction == 2));
  return g.one_d ? 1 : 3;
}

int WriteHeader(hid_t file, const chl_io_header &h, const Geometry &g)
{
  const hid_t F64 = H5T_IEEE_F64BE, I32 = H5T_STD_I32BE, D = H5T_NATIVE_DOUBLE, I = H5T_NATIVE_INT;
  IO_TRY(WriteAttr(file, "gamma", F64, D, &h.gamma, 1));
  IO_TRY(WriteString(file, "Git Commit Hash", chl_version()));
  IO_TRY(WriteString(file, "Macro Flags", h.macro_flags));
  IO_TRY(WriteString(file, "cholla", ""));
  IO_TRY(WriteAttr(file, "t", F64, D, &h.t, 1));
  IO_TRY(WriteAttr(file, "dt", F64, D, &h.dt, 1));
  IO_TRY(WriteAttr(file, "n_step", I32, I, &h.n_step, 1));
  IO_TRY(WriteAttr(file, "n_fields", I32, I, &h.n_fields, 1));
  const double units[6] = {kTimeUnit, kLengthUnit, kMassUnit, kVelocityUnit, kDensityUnit, kEnergyUnit};
  const char *unit_names[6] = {"time_unit", "length_unit", "mass_unit", "velocity_unit", "density_unit",
                               "energy_unit"};
  for (int q = 0; q < 6; q++) IO_TRY(WriteAttr(file, unit_names[q], F64, D, &units[q], 1));
  if (h.mhd) {
    const double b_unit = std::sqrt(kMassUnit / kLengthUnit) / kTimeUnit;  // MAGNETIC_FIELD_UNIT
    IO_TRY(WriteAttr(file, "magnetic_field_unit", F64, D, &b_unit, 1));
  }
  const int dims[3] = {g.nxr, g.one_d ? 1 : g.nyr, g.one_d ? 1 : g.nzr};
  IO_TRY(WriteAttr(file, "dims", I32, I, dims, 3));
  IO_TRY(WriteAttr(file, "bounds", F64, D, h.bounds, 3));
  IO_TRY(WriteAttr(file, "domain", F64, D, h.domain, 3));
  IO_TRY(WriteAttr(file, "dx", F64, D, h.dx, 3));
  return CHL_OK;
}

}  // namespace

extern "C" {

const char *chl_io_last_error(void) { return g_err.c_str(); }

int chl_io_write_hdf5(const char *path, const double *d_U, const chl_io_header *hp)
{
  if (!path || !d_U || !hp) return Fail(CHL_ERR_INVALID, "null argument");
  const chl_io_header &h = *hp;
  Geometry g;
  IO_TRY(Check(h, g));
  std::vector<Field> fields;
  IO_TRY(Fields(h, fields));
  const bool f32 = h.float32 != 0;
  if (f32 && (g.one_d || h.mhd))
    return Fail(CHL_ERR_UNSUPPORTED, "float32 snapshots are 3D hydro only (Output_Float32, io.cpp:299)");
  if (f32) fields.resize(h.dual_energy ? 6 : 5);
  const size_t elem = f32 ? sizeof(float) : sizeof(double);
  DeviceBuffer dbuf;
  if (dbuf.Allocate(g.max_elems * elem) != CHL_OK) return Fail(CHL_ERR_IO, std::string("device buffer: ") + chl_last_error());
  std::vector<unsigned char> hbuf(g.max_elems * elem);

  QuietHdf5 quiet;
  Id file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return Fail(CHL_ERR_IO, std::string("cannot create ") + path);
  IO_TRY(WriteHeader(file.get(), h, g));
  for (const Field &f : fields) {
    hsize_t dims[3];
    const int rank    = Extents(g, f, dims);
    const size_t n    = size_t(dims[0]) * (rank == 3 ? size_t(dims[1]) * dims[2] : 1);
    const double *src = d_U + size_t(f.index) * g.n_cells;
    int rc            = f32 ? chl_pack_real_cells_f32(src, static_cast<float *>(dbuf.get()), h.nx, h.ny, h.nz,
                                                      h.n_ghost, int(dims[0]), int(dims[1]), int(dims[2]), f.mhd_direction)
                            : chl_pack_real_cells(src, static_cast<double *>(dbuf.get()), h.nx, h.ny, h.nz, h.n_ghost,
                                                  int(dims[0]), int(dims[1]), int(dims[2]), f.mhd_direction);
    if (rc != CHL_OK) return Fail(rc, std::string("packing ") + f.name + ": " + chl_last_error());
    if (chl_memcpy_dtoh(hbuf.data(), dbuf.get(), n * elem) != CHL_OK)
      return Fail(CHL_ERR_IO, std::string("device to host copy: ") + chl_last_error());
    Id space(H5Screate_simple(rank, dims, nullptr), H5Sclose);
    Id dset(H5Dcreate2(file.get(), f.name, f32 ? H5T_IEEE_F32BE : H5T_IEEE_F64BE, space.get(), H5P_DEFAULT,
                       H5P_DEFAULT, H5P_DEFAULT),
            H5Dclose);
    if (!space.ok() || !dset.ok() ||
        H5Dwrite(dset.get(), f32 ? H5T_NATIVE_FLOAT : H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, hbuf.data()) <
            0)
      return Fail(CHL_ERR_IO, std::string("writing dataset ") + f.name);
  }
  if (H5Fflush(file.get(), H5F_SCOPE_LOCAL) < 0) return Fail(CHL_ERR_IO, std::string("flushing ") + path);
  return CHL_OK;
}

int chl_io_read_hdf5(const char *path, double *d_U, chl_io_header *hp)
{
  if (!path || !d_U || !hp) return Fail(CHL_ERR_INVALID, "null argument");
  chl_io_header &h = *hp;
  Geometry g;
  IO_TRY(Check(h, g));
  std::vector<Field> fields;
  IO_TRY(Fields(h, fields));
  DeviceBuffer dbuf;
  if (dbuf.Allocate(g.max_elems * sizeof(double)) != CHL_OK)
    return Fail(CHL_ERR_IO, std::string("device buffer: ") + chl_last_error());
  std::vector<double> hbuf(g.max_elems);

  QuietHdf5 quiet;
  Id file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (!file.ok()) return Fail(CHL_ERR_IO, std::string("cannot open ") + path);
  double gamma = 0, t = 0;
  int n_step   = 0;
  IO_TRY(ReadAttr(file.get(), "gamma", H5T_NATIVE_DOUBLE, &gamma));
  IO_TRY(ReadAttr(file.get(), "t", H5T_NATIVE_DOUBLE, &t));
  IO_TRY(ReadAttr(file.get(), "n_step", H5T_NATIVE_INT, &n_step));
  for (const Field &f : fields) {
    hsize_t want[3], got[3] = {0, 0, 0};
    const int rank = Extents(g, f, want);
    Id dset(H5Dopen2(file.get(), f.name, H5P_DEFAULT), H5Dclose);
    if (!dset.ok()) return Fail(CHL_ERR_IO, std::string("no dataset ") + f.name + " in " + path);
    Id space(H5Dget_space(dset.get()), H5Sclose);
    if (!space.ok() || H5Sget_simple_extent_ndims(space.get()) != rank ||
        H5Sget_simple_extent_dims(space.get(), got, nullptr) != rank || std::memcmp(got, want, rank * sizeof(hsize_t)))
      return Fail(CHL_ERR_IO, std::string("dataset ") + f.name + " does not have the grid's extents");
    const size_t n = size_t(want[0]) * (rank == 3 ? size_t(want[1]) * want[2] : 1);
    if (H5Dread(dset.get(), H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, hbuf.data()) < 0)
      return Fail(CHL_ERR_IO, std::string("reading dataset ") + f.name);
    if (chl_memcpy_htod(dbuf.get(), hbuf.data(), n * sizeof(double)) != CHL_OK)
      return Fail(CHL_ERR_IO, std::string("host to device copy: ") + chl_last_error());
    int rc = chl_unpack_real_cells(static_cast<const double *>(dbuf.get()), d_U + size_t(f.index) * g.n_cells, h.nx,
                                   h.ny, h.nz, h.n_ghost, int(want[0]), int(want[1]), int(want[2]), f.mhd_direction);
    if (rc != CHL_OK) return Fail(rc, std::string("unpacking ") + f.name + ": " + chl_last_error());
  }
  h.gamma = gamma, h.t = t, h.n_step = n_step;
  return CHL_OK;
}

}  // extern "C"
