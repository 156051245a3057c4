"""Cholla-format HDF5 snapshots of device states: ctypes binding of libcholla_amd_io.so
(include/cholla_amd_io.h).  The library replaces the HDF5 branch of the reference's
Output_Data / Output_Float32 and Read_Grid (src/io/io.cpp:198-378, 2255-2316, 2508-2700).
"""
import ctypes
import os

from . import HERE, ChollaError, DeviceArray, lib as core_lib

IO_LIB_PATH = os.path.join(HERE, "libcholla_amd_io.so")
EXPORTS = ["chl_io_last_error", "chl_io_write_hdf5", "chl_io_read_hdf5"]
CHL_ERR_IO = 5


class IoHeader(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("n_ghost", ctypes.c_int),
        ("n_fields", ctypes.c_int), ("dual_energy", ctypes.c_int), ("mhd", ctypes.c_int), ("float32", ctypes.c_int),
        ("n_step", ctypes.c_int), ("gamma", ctypes.c_double), ("t", ctypes.c_double), ("dt", ctypes.c_double),
        ("bounds", ctypes.c_double * 3), ("domain", ctypes.c_double * 3), ("dx", ctypes.c_double * 3),
        ("macro_flags", ctypes.c_char_p),
    ]


_io = None


def io_lib():
    """Load the snapshot library (raises if it was not built: it needs the HDF5 C library)."""
    global _io
    if _io is not None:
        return _io
    core_lib()  # the snapshot library resolves its core symbols against the loaded core library
    if not os.path.exists(IO_LIB_PATH):
        raise ChollaError(f"{IO_LIB_PATH} is missing: build with `python -m cholla_amd.build` where HDF5 is installed")
    L = ctypes.CDLL(IO_LIB_PATH)
    L.chl_io_last_error.restype = ctypes.c_char_p
    L.chl_io_last_error.argtypes = []
    for name in ("chl_io_write_hdf5", "chl_io_read_hdf5"):
        fn = getattr(L, name)
        fn.restype = ctypes.c_int
        fn.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.POINTER(IoHeader)]
    _io = L
    return L


def _check(rc):
    if rc != 0:
        raise ChollaError(f"cholla_amd snapshot error {rc}: {io_lib().chl_io_last_error().decode()}")


def header(nx, ny, nz, n_ghost, gamma=5.0 / 3.0, t=0.0, dt=0.0, n_step=0, dx=(1.0, 1.0, 1.0), bounds=(0.0, 0.0, 0.0),
           domain=None, dual_energy=False, mhd=False, float32=False, macro_flags=""):
    h = IoHeader()
    h.nx, h.ny, h.nz, h.n_ghost = nx, ny, nz, n_ghost
    h.dual_energy, h.mhd, h.float32 = int(dual_energy), int(mhd), int(float32)
    h.n_fields = 5 + int(dual_energy) + 3 * int(mhd)
    h.n_step, h.gamma, h.t, h.dt = n_step, gamma, t, dt
    reals = (nx - 2 * n_ghost, max(ny - 2 * n_ghost, 1), max(nz - 2 * n_ghost, 1))
    dom = domain if domain is not None else tuple(r * d for r, d in zip(reals, dx))
    for q in range(3):
        h.dx[q], h.bounds[q], h.domain[q] = dx[q], bounds[q], dom[q]
    h.macro_flags = macro_flags.encode()
    return h


def write_hdf5(path, U: DeviceArray, h: IoHeader):
    """Write the device state U as one snapshot file (chl_io_write_hdf5)."""
    _check(io_lib().chl_io_write_hdf5(os.fsencode(path), U.ptr, ctypes.byref(h)))


def read_hdf5(path, U: DeviceArray, h: IoHeader):
    """Read a snapshot's real cells into U; h.gamma, h.t and h.n_step take the file's values."""
    _check(io_lib().chl_io_read_hdf5(os.fsencode(path), U.ptr, ctypes.byref(h)))
    return {"gamma": h.gamma, "t": h.t, "n_step": h.n_step}
