"""ctypes binding of the CPU oracle (libcholla_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's cpu_baseline leg.  The product package (cholla_amd)
never imports this module.
"""
import contextlib
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcholla_oracle.so")

PCM, PLMC, PPMC, PLMP, PPMP = 0, 1, 2, 3, 4

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None


def build():
    """Compile the oracle with its Makefile (g++ -ffp-contract=off, OpenMP)."""
    import subprocess

    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        d, i, l = ctypes.c_double, ctypes.c_int, ctypes.c_long
        sig = {
            "orc_ppm_single_variable": (None, [d, d, d, d, d, _dp, _dp]),
            "orc_thread_guard": (i, [i, i, i, i, i, i, i]),
            "orc_load_data": (None, [_dp, i, i, i, i, i, i, i, d, _dp]),
            "orc_compute_slope": (None, [_dp, _dp, d, _dp]),
            "orc_van_leer_slope": (None, [_dp, _dp, _dp]),
            "orc_interface_linear": (None, [_dp, _dp, d, _dp]),
            "orc_monotonize_char_return_prim": (None, [_dp, _dp, _dp, _dp, _dp, d, d, _dp]),
            "orc_plm_limit_interfaces": (None, [_dp, _dp, _dp, _dp, _dp]),
            "orc_write_data": (None, [_dp, _dp, l, l, i, d]),
            "orc_hydro_inverse_crossing_time": (d, [d, d, d, d, d, d, d, d, d, d]),
            "orc_fill_uniform": (None, [ctypes.c_uint64, d, d, l, _dp]),
            "orc_pcm_3d": (None, [_dp] * 7 + [i, i, i, i]),
            "orc_pcm_1d": (None, [_dp, _dp, _dp, i, i]),
            "orc_plmc": (None, [_dp, _dp, _dp, i, i, i, d, i]),
            "orc_plmp": (None, [_dp, _dp, _dp, i, i, i, d, i]),
            "orc_ppmp": (None, [_dp, _dp, _dp, i, i, i, d, i]),
            "orc_roe": (None, [_dp, _dp, _dp, i, i, i, d, i]),
            "orc_exact": (None, [_dp, _dp, _dp, i, i, i, d, i]),
            "orc_set_static_gravity": (None, [i, i, i, i, d, d, d]),
            "orc_static_gravity_field": (None, [i, d, d, d, _dp]),
            "orc_ppmc_vl": (None, [_dp, _dp, _dp, i, i, i, d, i]),
            "orc_hllc": (None, [_dp, _dp, _dp, i, i, i, d, i]),
            "orc_update_half_3d": (None, [_dp] * 5 + [i, i, i, d, d, d, d, d, i, i, d]),
            "orc_partial_update_ge_3d": (None, [_dp, i, i, i, i, d, d, d, d, d]),
            "orc_select_internal_energy_3d": (None, [_dp, i, i, i, i]),
            "orc_sync_energies_3d": (None, [_dp, i, i, i, i]),
            "orc_set_dual_energy": (None, [i]),
            "orc_get_dual_energy": (i, []),
            "orc_set_positivity_fallback": (None, [i]),
            "orc_set_riemann": (None, [i]),
            "orc_hll": (None, [_dp, _dp, _dp, i, i, i, d, i]),
            "orc_pressure_from_de": (d, [d, d, d, d]),
            "orc_update_3d": (None, [_dp] * 4 + [i, i, i, i, d, d, d, d, i, i, d, d]),
            "orc_set_crash_repair": (None, [i]),
            "orc_set_flux_rounding_model": (None, [d, i]),
            "orc_temperature_floor": (None, [_dp, i, i, i, i, d]),
            "orc_calc_dti": (d, [_dp, i, i, i, i, d, d, d, d]),
            "orc_set_ghost_cells": (None, [_dp, i, i, i, i, i, _ip, i, i, i, i, i, i, i]),
            "orc_set_boundaries": (None, [_dp, i, i, i, i, i, _ip]),
            "orc_set_boundaries_mhd": (None, [_dp, i, i, i, i, i, _ip]),
            "orc_set_ghost_cells_mhd": (None, [_dp, i, i, i, i, i, _ip, i, i, i, i, i, i, i]),
            "orc_vl3d": (None, [_dp, i, i, i, i, d, d, d, d, d, i, i, i, d]),
            "orc_vl1d": (None, [_dp, i, i, d, d, d, i, i]),
            "orc_run": (i, [_dp, i, i, i, i, i, d, d, d, d, i, _ip, _dp, d, i, i, d, i, d, _dp]),
            "orc_num_threads": (i, []),
            # MHD (mhd_oracle.cpp)
            "orc_mhd_fast_speed": (d, [d] * 6),
            "orc_mhd_slow_speed": (d, [d] * 6),
            "orc_mhd_alfven_speed": (d, [d, d]),
            "orc_mhd_total_pressure": (d, [d] * 4),
            "orc_mhd_pressure_prim": (d, [d] * 9),
            "orc_mhd_energy_prim": (d, [d] * 9),
            "orc_mhd_cell_centered_b": (None, [_dp, i, i, i, i, i, i, _dp]),
            "orc_mhd_load_data": (None, [_dp, i, i, i, i, i, i, i, d, _dp]),
            "orc_mhd_eigenvectors": (None, [_dp, d, d, d, _dp]),
            "orc_mhd_prim_to_char": (None, [_dp, _dp, d, d, d, _dp]),
            "orc_mhd_char_to_prim": (None, [_dp, _dp, d, d, d, _dp]),
            "orc_mhd_inverse_crossing_time": (d, [d] * 13),
            "orc_mhd_pcm_3d": (None, [_dp] * 7 + [i, i, i]),
            "orc_mhd_plmc": (None, [_dp, _dp, _dp, i, i, i, d, i]),
            "orc_mhd_ppmc": (None, [_dp, _dp, _dp, i, i, i, d, i]),
            "orc_mhd_set_reconstruction": (None, [i]),
            "orc_mhd_get_reconstruction": (i, []),
            "orc_hlld": (None, [_dp, _dp, _dp, _dp, l, d, i]),
            "orc_ct_slope": (d, [_dp, _dp, d, i, i, i, i, i, i, i, i, i, i, l, l, l]),
            "orc_ct_electric_fields": (None, [_dp] * 5 + [i, i, i, i]),
            "orc_update_magnetic": (None, [_dp, _dp, _dp, i, i, i, d, d, d, d]),
            "orc_mhd_calc_dti": (d, [_dp, i, i, i, i, d, d, d, d]),
            "orc_mhd_divergence": (d, [_dp, i, i, i, d, d, d]),
            "orc_mhd_vl3d": (None, [_dp, i, i, i, i, d, d, d, d, d, i, d]),
            "orc_mhd_run": (i, [_dp, i, i, i, i, d, d, d, d, _ip, _dp, d, i, i, d, i, d, _dp]),
            "orc_fill_uniform32": (None, [ctypes.c_uint32, d, d, l, _dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    """Pointer to a C-contiguous float64 array (no copy)."""
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def uniform(seed, a, b, n):
    out = np.empty(n, dtype=np.float64)
    lib().orc_fill_uniform(seed, a, b, n, ptr(out))
    return out


def calc_dti(U, nx, ny, nz, n_ghost, dx, dy, dz, gamma):
    return lib().orc_calc_dti(ptr(U), nx, ny, nz, n_ghost, dx, dy, dz, gamma)


def set_boundaries(U, nx, ny, nz, n_fields, n_ghost, flags):
    f, fp = iptr(flags)
    lib().orc_set_boundaries(ptr(U), nx, ny, nz, n_fields, n_ghost, fp)


def mhd_set_boundaries(U, nx, ny, nz, n_ghost, flags):
    """The MHD build's boundary fill (8 fields, no outflow diode)."""
    f, fp = iptr(flags)
    lib().orc_set_boundaries_mhd(ptr(U), nx, ny, nz, 8, n_ghost, fp)


@contextlib.contextmanager
def positivity_fallback(on=True):
    """The PPMC positivity fallback (an option of this implementation, not of the
    reference) for the calls inside the block."""
    lib().orc_set_positivity_fallback(int(on))
    try:
        yield
    finally:
        lib().orc_set_positivity_fallback(0)


HLLC, HLL, ROE, EXACT = 0, 1, 2, 3


@contextlib.contextmanager
def riemann(solver=HLLC):
    """The Riemann solver build option (HLLC, HLL, ROE or EXACT) for the calls inside the block."""
    lib().orc_set_riemann(int(solver))
    try:
        yield
    finally:
        lib().orc_set_riemann(HLLC)


@contextlib.contextmanager
def crash_repair(on=True):
    """A build without both floors for orc_vl3d calls inside the block: cells that come out of
    the update with a negative or NaN density or energy get Average_Cell_All_Fields."""
    lib().orc_set_crash_repair(int(on))
    try:
        yield
    finally:
        lib().orc_set_crash_repair(0)


MHD_PCM, MHD_PLMC, MHD_PPMC = 0, 1, 2


@contextlib.contextmanager
def mhd_reconstruction(recon):
    """The MHD build's stage-2 reconstruction (MHD_PCM, MHD_PLMC or MHD_PPMC) for the MHD
    step calls inside the block."""
    prev = lib().orc_mhd_get_reconstruction()
    lib().orc_mhd_set_reconstruction(int(recon))
    try:
        yield
    finally:
        lib().orc_mhd_set_reconstruction(prev)


@contextlib.contextmanager
def dual_energy(on=True):
    """The reference's DE build option for the calls inside the block (GasEnergy = field 5)."""
    prev = lib().orc_get_dual_energy()
    lib().orc_set_dual_energy(int(on))
    try:
        yield
    finally:
        lib().orc_set_dual_energy(prev)


def pressure_from_de(E, U_total, U_advected, gamma):
    return lib().orc_pressure_from_de(E, U_total, U_advected, gamma)


@contextlib.contextmanager
def static_gravity(custom_grav, offsets=(0, 0, 0), bounds=(0.0, 0.0, 0.0)):
    """The STATIC_GRAV build option (Parameters::custom_grav and the grid's offsets and lower
    bounds) for the calls inside the block."""
    lib().orc_set_static_gravity(int(custom_grav), *[int(o) for o in offsets], *[float(b) for b in bounds])
    try:
        yield
    finally:
        lib().orc_set_static_gravity(0, 0, 0, 0, 0.0, 0.0, 0.0)


def gravity_field(dims, x, y=0.0, z=0.0):
    g = np.zeros(3)
    lib().orc_static_gravity_field(int(dims), float(x), float(y), float(z), ptr(g))
    return g


def run(U, nx, ny, nz, n_ghost, n_fields, dx, dy, dz, gamma, recon, flags, t, tout, max_steps,
        density_floor_on=True, density_floor=0.0, temperature_floor_on=True, U_floor=0.0, dt_log=None,
        dual_energy_on=False, positivity_fallback_on=False, solver=HLLC):
    """main.cpp's hydro loop; U (with ghost cells already set) is advanced in place.
    dual_energy_on: the DE build (3D only, n_fields 6 with GasEnergy last).

    Returns (steps, t)."""
    f, fp = iptr(flags)
    tt = ctypes.c_double(t)
    log = dt_log if dt_log is not None else None
    with dual_energy(dual_energy_on), positivity_fallback(positivity_fallback_on), riemann(solver):
        n = lib().orc_run(ptr(U), nx, ny, nz, n_ghost, n_fields, dx, dy, dz, gamma, recon, fp,
                          ctypes.byref(tt), tout, max_steps, int(density_floor_on), density_floor,
                          int(temperature_floor_on), U_floor, ptr(log) if log is not None else None)
    if n < 0:
        raise ValueError("dual energy needs a 3D grid with n_fields == 6")
    return n, tt.value


def uniform32(seed, a, b, n):
    """std::uniform_real_distribution<double>(a, b) draws of std::mt19937(seed)."""
    out = np.empty(n, dtype=np.float64)
    lib().orc_fill_uniform32(seed, a, b, n, ptr(out))
    return out


def mhd_run(U, nx, ny, nz, n_ghost, dx, dy, dz, gamma, flags, t, tout, max_steps, density_floor_on=True,
            density_floor=0.0, temperature_floor_on=True, U_floor=0.0, dt_log=None):
    """main.cpp's loop for the MHD build (PLMC + HLLD + CT); U (8 fields, ghosts set) advances in place.

    Returns (steps, t)."""
    f, fp = iptr(flags)
    tt = ctypes.c_double(t)
    n = lib().orc_mhd_run(ptr(U), nx, ny, nz, n_ghost, dx, dy, dz, gamma, fp, ctypes.byref(tt), tout, max_steps,
                          int(density_floor_on), density_floor, int(temperature_floor_on), U_floor,
                          ptr(dt_log) if dt_log is not None else None)
    return n, tt.value


def mhd_calc_dti(U, nx, ny, nz, n_ghost, dx, dy, dz, gamma):
    return lib().orc_mhd_calc_dti(ptr(U), nx, ny, nz, n_ghost, dx, dy, dz, gamma)


def mhd_divergence(U, nx, ny, nz, dx, dy, dz):
    return lib().orc_mhd_divergence(ptr(U), nx, ny, nz, dx, dy, dz)
