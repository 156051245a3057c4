"""cholla_amd -- Python host binding of libcholla_amd.so (the MI355X hydro hot path).

The product is the C ABI in include/cholla_amd.h; this module binds it with
ctypes so tests, the smoke check and the benchmark can drive it.  Device memory
is allocated through the library itself (no torch types cross the boundary).
Every call raises ChollaError on a non-zero status.  There is no CPU fallback:
without the built library or a HIP device the calls fail loudly.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcholla_amd.so")

CHL_OK = 0
RECON_PCM, RECON_PLMC, RECON_PPMC, RECON_PLMP, RECON_PPMP = 0, 1, 2, 3, 4
RIEMANN_HLLC = 0
RIEMANN_HLL = 1
RIEMANN_ROE = 2
RIEMANN_EXACT = 3
PREC_FP64_EXACT, PREC_FP64, PREC_MIXED = 0, 1, 2
PRECISIONS = {"fp64_exact": PREC_FP64_EXACT, "fp64": PREC_FP64, "mixed": PREC_MIXED}
BC_PERIODIC, BC_REFLECTIVE, BC_TRANSMISSIVE = 1, 2, 3

# every symbol include/cholla_amd.h declares
EXPORTS = [
    "chl_version", "chl_last_error", "chl_device_count", "chl_device_name", "chl_malloc", "chl_free",
    "chl_memcpy_htod", "chl_memcpy_dtoh", "chl_memcpy_dtod", "chl_synchronize", "chl_hydro_create",
    "chl_hydro_destroy", "chl_hydro_vl_step", "chl_hydro_update_grid", "chl_hydro_run", "chl_hydro_time_steps",
    "chl_hydro_kernel_times", "chl_hydro_autotune", "chl_hydro_get_config", "chl_calc_dt",
    "chl_apply_temperature_floor", "chl_set_ghost_cells", "chl_set_boundaries", "chl_reconstruct",
    "chl_hllc_fluxes", "chl_hll_fluxes", "chl_roe_fluxes", "chl_exact_fluxes", "chl_hydro_set_static_gravity", "chl_ic_kelvin_helmholtz", "chl_ensemble_create", "chl_ensemble_destroy",
    "chl_ensemble_get_config", "chl_ensemble_persistent_pairs", "chl_ensemble_bind", "chl_ensemble_start",
    "chl_ensemble_advance", "chl_ensemble_advance_looped", "chl_ensemble_clocks", "chl_ensemble_autotune",
    "chl_ensemble_layout", "chl_turbulence_modes", "chl_ic_turbulence", "chl_init_gas_energy", "chl_mhd_create",
    "chl_mhd_destroy",
    "chl_mhd_vl_step", "chl_mhd_run", "chl_mhd_time_steps", "chl_mhd_calc_dt", "chl_mhd_divergence",
    "chl_mhd_set_ghost_cells", "chl_mhd_set_boundaries",
    "chl_hlld_fluxes", "chl_mhd_reconstruct_plmc", "chl_mhd_reconstruct_ppmc", "chl_ct_electric_fields", "chl_update_magnetic_field",
    "chl_magnetic_field_from_vector_potential", "chl_ic_orszag_tang", "chl_pack_real_cells",
    "chl_pack_real_cells_f32", "chl_unpack_real_cells", "chl_hydro_vl_step_noalloc", "chl_noalloc_scratch_bytes",
]
LAUNCH_SPLIT, LAUNCH_PERSISTENT = 0, 1


class ChollaError(RuntimeError):
    pass


class HydroConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("n_ghost", ctypes.c_int),
        ("n_fields", ctypes.c_int), ("reconstruction", ctypes.c_int), ("riemann", ctypes.c_int),
        ("precision", ctypes.c_int), ("density_floor", ctypes.c_int), ("temperature_floor", ctypes.c_int),
        ("bc_flags", ctypes.c_int * 6), ("tile1", ctypes.c_int), ("tile2", ctypes.c_int),
        ("kchunk1", ctypes.c_int), ("kchunk2", ctypes.c_int), ("tile_order", ctypes.c_int),
        ("dual_energy", ctypes.c_int), ("positivity_fallback", ctypes.c_int),
    ]


class EnsembleConfig(ctypes.Structure):
    _fields_ = [
        ("n_ghost", ctypes.c_int), ("n_fields", ctypes.c_int), ("reconstruction", ctypes.c_int),
        ("riemann", ctypes.c_int), ("precision", ctypes.c_int), ("density_floor", ctypes.c_int),
        ("temperature_floor", ctypes.c_int), ("launch", ctypes.c_int), ("tile1", ctypes.c_int),
        ("tile2", ctypes.c_int), ("kchunk1", ctypes.c_int), ("kchunk2", ctypes.c_int),
        ("dual_energy", ctypes.c_int), ("positivity_fallback", ctypes.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MhdConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("n_ghost", ctypes.c_int),
        ("precision", ctypes.c_int), ("density_floor", ctypes.c_int), ("temperature_floor", ctypes.c_int),
        ("bc_flags", ctypes.c_int * 6), ("reconstruction", ctypes.c_int),
    ]


_lib = None
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_d = ctypes.c_double
_i = ctypes.c_int


def lib():
    """Load the built library (raises if it was never built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ChollaError(f"{LIB_PATH} is missing: build it with `python -m cholla_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "chl_version": (ctypes.c_char_p, []),
        "chl_last_error": (ctypes.c_char_p, []),
        "chl_device_count": (_i, [_ip]),
        "chl_device_name": (_i, [ctypes.c_char_p, _i]),
        "chl_malloc": (_i, [ctypes.POINTER(_vp), ctypes.c_size_t]),
        "chl_free": (_i, [_vp]),
        "chl_memcpy_htod": (_i, [_vp, _vp, ctypes.c_size_t]),
        "chl_memcpy_dtoh": (_i, [_vp, _vp, ctypes.c_size_t]),
        "chl_memcpy_dtod": (_i, [_vp, _vp, ctypes.c_size_t]),
        "chl_synchronize": (_i, []),
        "chl_hydro_create": (_i, [ctypes.POINTER(HydroConfig), ctypes.POINTER(_vp)]),
        "chl_hydro_destroy": (_i, [_vp]),
        "chl_hydro_vl_step": (_i, [_vp, _vp, _d, _d, _d, _d, _d, _d]),
        "chl_hydro_vl_step_noalloc": (_i, [ctypes.POINTER(HydroConfig), _vp, _d, _d, _d, _d, _d, _d, _vp]),
        "chl_noalloc_scratch_bytes": (ctypes.c_size_t, []),
        "chl_hydro_update_grid": (_i, [_vp, _vp, _d, _d, _d, _d, _d, _d, _d, _i, _dp]),
        "chl_hydro_run": (_i, [_vp, _vp, _d, _d, _d, _d, _d, _d, _d, _d, _i, _i, _dp, _ip]),
        "chl_hydro_time_steps": (_i, [_vp, _vp, _d, _d, _d, _d, _i, _i, ctypes.POINTER(ctypes.c_float), _dp]),
        "chl_hydro_kernel_times": (_i, [_vp, _vp, _d, _d, _d, _d, _i] + [ctypes.POINTER(ctypes.c_float)] * 3),
        "chl_hydro_autotune": (_i, [_vp, _vp, _d, _d, _d, _d, _i, _ip, _ip, _ip, _ip]),
        "chl_hydro_get_config": (_i, [_vp, ctypes.POINTER(HydroConfig)]),
        "chl_calc_dt": (_i, [_vp, _i, _i, _i, _i, _i, _d, _d, _d, _d, _dp]),
        "chl_apply_temperature_floor": (_i, [_vp, _i, _i, _i, _i, _i, _d]),
        "chl_set_ghost_cells": (_i, [_vp, _i, _i, _i, _i, _i, _i, _ip, _i, _i, _i, _i, _i, _i, _i]),
        "chl_set_boundaries": (_i, [_vp, _i, _i, _i, _i, _i, _ip]),
        "chl_mhd_set_ghost_cells": (_i, [_vp, _i, _i, _i, _i, _i, _i, _ip, _i, _i, _i, _i, _i, _i, _i]),
        "chl_mhd_set_boundaries": (_i, [_vp, _i, _i, _i, _i, _ip]),
        "chl_reconstruct": (_i, [_i, _vp, _vp, _vp, _i, _i, _i, _d, _i]),
        "chl_hllc_fluxes": (_i, [_vp, _vp, _vp, _i, _i, _i, _d, _i]),
        "chl_hll_fluxes": (_i, [_vp, _vp, _vp, _i, _i, _i, _d, _i]),
        "chl_roe_fluxes": (_i, [_vp, _vp, _vp, _i, _i, _i, _d, _i]),
        "chl_exact_fluxes": (_i, [_vp, _vp, _vp, _i, _i, _i, _d, _i]),
        "chl_hydro_set_static_gravity": (_i, [_vp, _i, _i, _i, _i, _d, _d, _d]),
        "chl_ic_kelvin_helmholtz": (_i, [_dp, _i, _i, _d, ctypes.c_uint64]),
        "chl_ensemble_create": (_i, [ctypes.POINTER(EnsembleConfig), _i, _ip, ctypes.POINTER(_vp)]),
        "chl_ensemble_destroy": (_i, [_vp]),
        "chl_ensemble_get_config": (_i, [_vp, ctypes.POINTER(EnsembleConfig)]),
        "chl_ensemble_persistent_pairs": (_i, [_ip, _i, _ip]),
        "chl_ensemble_bind": (_i, [_vp, ctypes.POINTER(_vp), _dp]),
        "chl_ensemble_start": (_i, [_vp, _d, _d, _d, _dp, _dp]),
        "chl_ensemble_advance": (_i, [_vp, _i, _i, ctypes.POINTER(ctypes.c_float)]),
        "chl_ensemble_advance_looped": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_float)]),
        "chl_ensemble_clocks": (_i, [_vp, _dp, _dp, _ip]),
        "chl_ensemble_autotune": (_i, [_vp, _i, ctypes.POINTER(EnsembleConfig)]),
        "chl_ensemble_layout": (_i, [ctypes.c_uint64, _i, _ip, _dp]),
        "chl_turbulence_modes": (_i, [ctypes.c_uint64, _dp, _i, _ip]),
        "chl_ic_turbulence": (_i, [_vp, _i, _i, _d, _d, _d, _d, ctypes.c_uint64]),
        "chl_init_gas_energy": (_i, [_vp, _i, _i, _i]),
        "chl_mhd_create": (_i, [ctypes.POINTER(MhdConfig), ctypes.POINTER(_vp)]),
        "chl_mhd_destroy": (_i, [_vp]),
        "chl_mhd_vl_step": (_i, [_vp, _vp, _d, _d, _d, _d, _d, _d]),
        "chl_mhd_run": (_i, [_vp, _vp, _d, _d, _d, _d, _d, _d, _d, _d, _i, _i, _dp, _ip]),
        "chl_mhd_time_steps": (_i, [_vp, _vp, _d, _d, _d, _d, _i, _i, ctypes.POINTER(ctypes.c_float), _dp]),
        "chl_mhd_calc_dt": (_i, [_vp, _i, _i, _i, _i, _d, _d, _d, _d, _dp]),
        "chl_mhd_divergence": (_i, [_vp, _i, _i, _i, _d, _d, _d, _dp]),
        "chl_hlld_fluxes": (_i, [_vp, _vp, _vp, _vp, ctypes.c_long, _d, _i]),
        "chl_mhd_reconstruct_plmc": (_i, [_vp, _vp, _vp, _i, _i, _i, _d, _i]),
        "chl_mhd_reconstruct_ppmc": (_i, [_vp, _vp, _vp, _i, _i, _i, _d, _i]),
        "chl_ct_electric_fields": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i]),
        "chl_update_magnetic_field": (_i, [_vp, _vp, _vp, _i, _i, _i, _d, _d, _d, _d]),
        "chl_magnetic_field_from_vector_potential": (_i, [_dp, _i, _i, _i, _d, _d, _d, _dp]),
        "chl_ic_orszag_tang": (_i, [_dp, _i, _i, _d, ctypes.c_uint64, _d]),
        "chl_pack_real_cells": (_i, [_vp, _vp] + [_i] * 8),
        "chl_pack_real_cells_f32": (_i, [_vp, _vp] + [_i] * 8),
        "chl_unpack_real_cells": (_i, [_vp, _vp] + [_i] * 8),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != CHL_OK:
        raise ChollaError(f"cholla_amd error {rc}: {lib().chl_last_error().decode()}")


def device_count():
    n = ctypes.c_int(0)
    lib().chl_device_count(ctypes.byref(n))
    return n.value


def device_name():
    buf = ctypes.create_string_buffer(256)
    check(lib().chl_device_name(buf, 256))
    return buf.value.decode()


class DeviceArray:
    """A float64 device buffer owned through the library."""

    def __init__(self, n, fill=None):
        self.n = int(n)
        self.nbytes = self.n * 8
        p = _vp()
        check(lib().chl_malloc(ctypes.byref(p), max(self.nbytes, 8)))
        self.ptr = p
        if fill is not None:
            self.upload(fill)

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        d = cls(a.size)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        assert a.size == self.n
        check(lib().chl_memcpy_htod(self.ptr, a.ctypes.data_as(_vp), self.nbytes))

    def download(self):
        out = np.empty(self.n, dtype=np.float64)
        check(lib().chl_memcpy_dtoh(out.ctypes.data_as(_vp), self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr is not None:
            lib().chl_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Hydro:
    """Integrator handle for one box (chl_hydro_* entry points)."""

    def __init__(self, nx, ny, nz, n_ghost, reconstruction=RECON_PPMC, precision=PREC_FP64_EXACT,
                 bc=(1, 1, 1, 1, 1, 1), density_floor=True, temperature_floor=True, tile1=-1, tile2=-1,
                 kchunk1=0, kchunk2=0, tile_order=0, dual_energy=False, positivity_fallback=False,
                 riemann=RIEMANN_HLLC):
        cfg = HydroConfig()
        cfg.nx, cfg.ny, cfg.nz, cfg.n_ghost = nx, ny, nz, n_ghost
        cfg.n_fields = 6 if dual_energy else 5
        cfg.dual_energy = int(dual_energy)
        cfg.positivity_fallback = int(positivity_fallback)
        cfg.reconstruction, cfg.riemann, cfg.precision = reconstruction, riemann, precision
        cfg.density_floor, cfg.temperature_floor = int(density_floor), int(temperature_floor)
        for q in range(6):
            cfg.bc_flags[q] = bc[q]
        cfg.tile1, cfg.tile2, cfg.kchunk1, cfg.kchunk2 = tile1, tile2, kchunk1, kchunk2
        cfg.tile_order = tile_order
        h = _vp()
        check(lib().chl_hydro_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.dims = (nx, ny, nz)
        self.n_ghost = n_ghost
        self.n_fields = cfg.n_fields

    def config(self):
        c = HydroConfig()
        check(lib().chl_hydro_get_config(self.h, ctypes.byref(c)))
        return c

    def set_static_gravity(self, custom_grav, offsets=(0, 0, 0), bounds=(0.0, 0.0, 0.0)):
        """STATIC_GRAV: the analytic field custom_grav and the grid's offsets / lower bounds."""
        check(lib().chl_hydro_set_static_gravity(self.h, int(custom_grav), *[int(o) for o in offsets],
                                                 *[float(b) for b in bounds]))

    def vl_step(self, U, dx, dy, dz, dt, gamma, density_floor=0.0):
        check(lib().chl_hydro_vl_step(self.h, U.ptr, dx, dy, dz, dt, gamma, density_floor))

    def update_grid(self, U, dx, dy, dz, dt, gamma, density_floor=0.0, U_floor=0.0, fill_ghosts=True):
        dti = ctypes.c_double()
        check(lib().chl_hydro_update_grid(self.h, U.ptr, dx, dy, dz, dt, gamma, density_floor, U_floor,
                                          int(fill_ghosts), ctypes.byref(dti)))
        return dti.value

    def run(self, U, dx, dy, dz, gamma, t0, tout, max_steps, use_graph=True, density_floor=0.0, U_floor=0.0):
        t = ctypes.c_double()
        n = ctypes.c_int()
        check(lib().chl_hydro_run(self.h, U.ptr, dx, dy, dz, gamma, density_floor, U_floor, t0, tout, max_steps,
                                  int(use_graph), ctypes.byref(t), ctypes.byref(n)))
        return n.value, t.value

    def time_steps(self, U, dx, dy, dz, gamma, n_steps, use_graph=True):
        ms = ctypes.c_float()
        dt = ctypes.c_double()
        check(lib().chl_hydro_time_steps(self.h, U.ptr, dx, dy, dz, gamma, n_steps, int(use_graph), ctypes.byref(ms),
                                         ctypes.byref(dt)))
        return ms.value, dt.value

    def kernel_times(self, U, dx, dy, dz, gamma, n_steps):
        a, b, c = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        check(lib().chl_hydro_kernel_times(self.h, U.ptr, dx, dy, dz, gamma, n_steps, ctypes.byref(a),
                                           ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def autotune(self, U, dx, dy, dz, gamma, steps=2):
        out = [ctypes.c_int() for _ in range(4)]
        check(lib().chl_hydro_autotune(self.h, U.ptr, dx, dy, dz, gamma, steps, *[ctypes.byref(o) for o in out]))
        return tuple(o.value for o in out)

    def close(self):
        if getattr(self, "h", None):
            lib().chl_hydro_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def calc_dti(U, nx, ny, nz, n_ghost, dx, dy, dz, gamma):
    out = ctypes.c_double()
    check(lib().chl_calc_dt(U.ptr, nx, ny, nz, n_ghost, 5, dx, dy, dz, gamma, ctypes.byref(out)))
    return out.value


def set_boundaries(U, nx, ny, nz, n_ghost, flags, n_fields=5):
    f = (ctypes.c_int * 6)(*flags)
    check(lib().chl_set_boundaries(U.ptr, nx, ny, nz, n_fields, n_ghost, f))


def set_ghost_cells(U, nx, ny, nz, n_ghost, flags, isize, jsize, ksize, imin, jmin, kmin, direction, n_fields=5):
    f = (ctypes.c_int * 6)(*flags)
    check(lib().chl_set_ghost_cells(U.ptr, nx, ny, nz, n_fields, nx * ny * nz, n_ghost, f, isize, jsize, ksize, imin,
                                    jmin, kmin, direction))


def mhd_set_boundaries(U, nx, ny, nz, n_ghost, flags):
    """The MHD build's boundary fill of an 8-field state (no outflow diode)."""
    f = (ctypes.c_int * 6)(*flags)
    check(lib().chl_mhd_set_boundaries(U.ptr, nx, ny, nz, n_ghost, f))


def mhd_set_ghost_cells(U, nx, ny, nz, n_ghost, flags, isize, jsize, ksize, imin, jmin, kmin, direction):
    f = (ctypes.c_int * 6)(*flags)
    check(lib().chl_mhd_set_ghost_cells(U.ptr, nx, ny, nz, 8, nx * ny * nz, n_ghost, f, isize, jsize, ksize, imin,
                                        jmin, kmin, direction))


def apply_temperature_floor(U, nx, ny, nz, n_ghost, U_floor):
    check(lib().chl_apply_temperature_floor(U.ptr, nx, ny, nz, n_ghost, 5, U_floor))


def reconstruct(recon, U, QL, QR, nx, ny, nz, gamma, direction):
    check(lib().chl_reconstruct(recon, U.ptr, QL.ptr, QR.ptr, nx, ny, nz, gamma, direction))


def hllc_fluxes(QL, QR, F, nx, ny, nz, gamma, direction):
    check(lib().chl_hllc_fluxes(QL.ptr, QR.ptr, F.ptr, nx, ny, nz, gamma, direction))


def hll_fluxes(QL, QR, F, nx, ny, nz, gamma, direction):
    check(lib().chl_hll_fluxes(QL.ptr, QR.ptr, F.ptr, nx, ny, nz, gamma, direction))


def roe_fluxes(QL, QR, F, nx, ny, nz, gamma, direction):
    check(lib().chl_roe_fluxes(QL.ptr, QR.ptr, F.ptr, nx, ny, nz, gamma, direction))


def exact_fluxes(QL, QR, F, nx, ny, nz, gamma, direction):
    check(lib().chl_exact_fluxes(QL.ptr, QR.ptr, F.ptr, nx, ny, nz, gamma, direction))


def ic_kelvin_helmholtz(n, ng, gamma=5.0 / 3.0, seed=2024):
    """Host-built KH state (5 x (n+2ng)^3, ghosts filled periodically)."""
    N = n + 2 * ng
    U = np.empty(5 * N ** 3, dtype=np.float64)
    check(lib().chl_ic_kelvin_helmholtz(U.ctypes.data_as(_dp), n, ng, gamma, seed))
    return U


# ---------------------------------------------------------------- ensembles of boxes

class Ensemble:
    """Independent periodic boxes advanced together (chl_ensemble_* entry points).

    dims: list of (nx, ny, nz) totals including ghosts.  States are DeviceArrays
    owned by the caller and bound with bind()."""

    def __init__(self, dims, n_ghost, reconstruction=RECON_PPMC, precision=PREC_MIXED, launch=LAUNCH_SPLIT,
                 density_floor=True, temperature_floor=True, tile1=-1, tile2=-1, kchunk1=0, kchunk2=0,
                 dual_energy=False, positivity_fallback=False, riemann=RIEMANN_HLLC):
        cfg = EnsembleConfig()
        cfg.n_ghost, cfg.reconstruction, cfg.riemann = n_ghost, reconstruction, riemann
        cfg.n_fields, cfg.dual_energy = (6 if dual_energy else 5), int(dual_energy)
        cfg.positivity_fallback = int(positivity_fallback)
        cfg.precision, cfg.density_floor, cfg.temperature_floor = precision, int(density_floor), int(temperature_floor)
        cfg.launch, cfg.tile1, cfg.tile2, cfg.kchunk1, cfg.kchunk2 = launch, tile1, tile2, kchunk1, kchunk2
        d = np.ascontiguousarray(np.asarray(dims, dtype=np.int32).reshape(-1, 3))
        self.dims = [tuple(int(v) for v in row) for row in d]
        self.n = len(self.dims)
        h = _vp()
        check(lib().chl_ensemble_create(ctypes.byref(cfg), self.n, d.ctypes.data_as(_ip), ctypes.byref(h)))
        self.h = h
        self.n_ghost = n_ghost
        self._states = None

    def config(self):
        c = EnsembleConfig()
        check(lib().chl_ensemble_get_config(self.h, ctypes.byref(c)))
        return c

    def bind(self, states, dxyz):
        self._states = list(states)  # keep the buffers alive while bound
        ptrs = (_vp * self.n)(*[s.ptr for s in self._states])
        dx = np.ascontiguousarray(np.asarray(dxyz, dtype=np.float64).reshape(-1))
        assert dx.size == 3 * self.n
        check(lib().chl_ensemble_bind(self.h, ptrs, dx.ctypes.data_as(_dp)))

    def start(self, gamma, density_floor=0.0, U_floor=0.0, t0=None, tout=None):
        arr = lambda v: None if v is None else np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))
        t0a, touta = arr(t0), arr(tout)
        check(lib().chl_ensemble_start(self.h, gamma, density_floor, U_floor,
                                       None if t0a is None else t0a.ctypes.data_as(_dp),
                                       None if touta is None else touta.ctypes.data_as(_dp)))

    def advance(self, n_steps, use_graph=True):
        ms = ctypes.c_float()
        check(lib().chl_ensemble_advance(self.h, n_steps, int(use_graph), ctypes.byref(ms)))
        return ms.value

    def advance_looped(self, n_steps):
        ms = ctypes.c_float()
        check(lib().chl_ensemble_advance_looped(self.h, n_steps, ctypes.byref(ms)))
        return ms.value

    def clocks(self):
        t = np.empty(self.n)
        dt = np.empty(self.n)
        ns = np.empty(self.n, dtype=np.int32)
        check(lib().chl_ensemble_clocks(self.h, t.ctypes.data_as(_dp), dt.ctypes.data_as(_dp),
                                        ns.ctypes.data_as(_ip)))
        return t, dt, ns

    def autotune(self, steps=1):
        c = EnsembleConfig()
        check(lib().chl_ensemble_autotune(self.h, steps, ctypes.byref(c)))
        return c

    def close(self):
        if getattr(self, "h", None):
            lib().chl_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def persistent_pairs():
    n = ctypes.c_int()
    buf = (ctypes.c_int * 64)()
    check(lib().chl_ensemble_persistent_pairs(buf, 32, ctypes.byref(n)))
    return [(buf[2 * i], buf[2 * i + 1]) for i in range(n.value)]


def ensemble_layout(seed, n_boxes):
    """Edges and rms Mach numbers of the ensemble boxes (chl_ensemble_layout)."""
    edges = np.empty(n_boxes, dtype=np.int32)
    mach = np.empty(n_boxes)
    check(lib().chl_ensemble_layout(seed, n_boxes, edges.ctypes.data_as(_ip), mach.ctypes.data_as(_dp)))
    return edges, mach


def turbulence_modes(seed):
    """(m, 9) array: k, a, b of every mode of the turbulent velocity field."""
    n = ctypes.c_int()
    check(lib().chl_turbulence_modes(seed, None, 0, ctypes.byref(n)))
    out = np.empty((n.value, 9))
    check(lib().chl_turbulence_modes(seed, out.ctypes.data_as(_dp), n.value, ctypes.byref(n)))
    return out


def ic_turbulence(U, n, ng, gamma=5.0 / 3.0, rho=1.0, p=1.0, mach=1.0, seed=1):
    """Turbulence box on the device (U: DeviceArray of 5*(n+2ng)^3, or 6*(n+2ng)^3 for a
    dual-energy state, whose GasEnergy is then set too)."""
    N = n + 2 * ng
    check(lib().chl_ic_turbulence(U.ptr, n, ng, gamma, rho, p, mach, seed))
    if U.n == 6 * N ** 3:
        init_gas_energy(U, N, N, N)


def init_gas_energy(U, nx, ny, nz):
    """GasEnergy (field 5) = E - kinetic energy of every cell (chl_init_gas_energy)."""
    check(lib().chl_init_gas_energy(U.ptr, nx, ny, nz))


# ---------------------------------------------------------------- ideal MHD

class Mhd:
    """MHD integrator handle for one box (chl_mhd_* entry points); states hold 8 fields."""

    def __init__(self, nx, ny, nz, n_ghost=4, precision=PREC_FP64_EXACT, bc=(1, 1, 1, 1, 1, 1), density_floor=True,
                 temperature_floor=True, reconstruction=None):
        cfg = MhdConfig()
        cfg.nx, cfg.ny, cfg.nz, cfg.n_ghost, cfg.precision = nx, ny, nz, n_ghost, precision
        cfg.reconstruction = RECON_PLMC if reconstruction is None else reconstruction
        cfg.density_floor, cfg.temperature_floor = int(density_floor), int(temperature_floor)
        for q in range(6):
            cfg.bc_flags[q] = bc[q]
        h = _vp()
        check(lib().chl_mhd_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.dims = (nx, ny, nz)
        self.n_ghost = n_ghost

    def vl_step(self, U, dx, dy, dz, dt, gamma, density_floor=0.0):
        check(lib().chl_mhd_vl_step(self.h, U.ptr, dx, dy, dz, dt, gamma, density_floor))

    def run(self, U, dx, dy, dz, gamma, t0, tout, max_steps, use_graph=True, density_floor=0.0, U_floor=0.0):
        t = ctypes.c_double()
        n = ctypes.c_int()
        check(lib().chl_mhd_run(self.h, U.ptr, dx, dy, dz, gamma, density_floor, U_floor, t0, tout, max_steps,
                                int(use_graph), ctypes.byref(t), ctypes.byref(n)))
        return n.value, t.value

    def time_steps(self, U, dx, dy, dz, gamma, n_steps, use_graph=True):
        ms = ctypes.c_float()
        dt = ctypes.c_double()
        check(lib().chl_mhd_time_steps(self.h, U.ptr, dx, dy, dz, gamma, n_steps, int(use_graph), ctypes.byref(ms),
                                       ctypes.byref(dt)))
        return ms.value, dt.value

    def close(self):
        if getattr(self, "h", None):
            lib().chl_mhd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mhd_calc_dti(U, nx, ny, nz, n_ghost, dx, dy, dz, gamma):
    out = ctypes.c_double()
    check(lib().chl_mhd_calc_dt(U.ptr, nx, ny, nz, n_ghost, dx, dy, dz, gamma, ctypes.byref(out)))
    return out.value


def mhd_divergence(U, nx, ny, nz, dx, dy, dz):
    out = ctypes.c_double()
    check(lib().chl_mhd_divergence(U.ptr, nx, ny, nz, dx, dy, dz, ctypes.byref(out)))
    return out.value


def hlld_fluxes(QL, QR, magX, F, n_cells, gamma, direction):
    check(lib().chl_hlld_fluxes(QL.ptr, QR.ptr, magX.ptr, F.ptr, n_cells, gamma, direction))


def mhd_reconstruct_plmc(U, QL, QR, nx, ny, nz, gamma, direction):
    check(lib().chl_mhd_reconstruct_plmc(U.ptr, QL.ptr, QR.ptr, nx, ny, nz, gamma, direction))


def mhd_reconstruct_ppmc(U, QL, QR, nx, ny, nz, gamma, direction):
    check(lib().chl_mhd_reconstruct_ppmc(U.ptr, QL.ptr, QR.ptr, nx, ny, nz, gamma, direction))


def ct_electric_fields(Fx, Fy, Fz, U, E, nx, ny, nz):
    check(lib().chl_ct_electric_fields(Fx.ptr, Fy.ptr, Fz.ptr, U.ptr, E.ptr, nx, ny, nz))


def update_magnetic_field(src, dst, E, nx, ny, nz, dt, dx, dy, dz):
    check(lib().chl_update_magnetic_field(src.ptr, dst.ptr, E.ptr, nx, ny, nz, dt, dx, dy, dz))


def magnetic_field_from_vector_potential(U, nx, ny, nz, dx, dy, dz, A):
    """Face fields of the host state U (8 fields, modified in place) from the edge potential A."""
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(-1)
    assert U.dtype == np.float64 and U.flags["C_CONTIGUOUS"]
    check(lib().chl_magnetic_field_from_vector_potential(U.ctypes.data_as(_dp), nx, ny, nz, dx, dy, dz,
                                                         A.ctypes.data_as(_dp)))


def ic_orszag_tang(n, ng=4, gamma=5.0 / 3.0, seed=2024, amplitude=1e-3):
    """Host-built Orszag-Tang state (8 x (n+2ng)^3, ghosts filled periodically)."""
    N = n + 2 * ng
    U = np.empty(8 * N ** 3, dtype=np.float64)
    check(lib().chl_ic_orszag_tang(U.ctypes.data_as(_dp), n, ng, gamma, seed, amplitude))
    return U
