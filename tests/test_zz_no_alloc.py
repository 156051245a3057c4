"""The allocation-free VL step (chl_hydro_vl_step_noalloc) against chl_hydro_vl_step.

The allocation-free entry point runs the same kernels with the integrator scratch in a
fixed device array of the library, so its results equal the handle-based step bit for bit.
Two calls in a row, the second on a smaller box, check that the shared scratch (whose
repair marks must start at zero) carries nothing from one call to the next; the rough box
makes the step mark and repair non-positive densities.
"""
import ctypes
import sys

import numpy as np
import pytest

import cholla_amd as chl
from helpers import rough_box
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def L():
    if chl.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need a real MI355X")
    return chl.lib()


def config(N, ng, recon, precision, dual_energy=False):
    c = chl.HydroConfig()
    c.nx = c.ny = c.nz = N
    c.n_ghost = ng
    c.n_fields = 6 if dual_energy else 5
    c.reconstruction, c.riemann, c.precision = recon, chl.RIEMANN_HLLC, precision
    c.density_floor, c.temperature_floor = 1, 1
    for q in range(6):
        c.bc_flags[q] = 1
    c.tile1 = c.tile2 = -1
    c.dual_energy = int(dual_energy)
    return c


def step_both(L, U, N, ng, recon, precision, dt_scale=1.0):
    h = 1.0 / (N - 2 * ng)
    dt = dt_scale * 0.3 / orc.calc_dti(U, N, N, N, ng, h, h, h, G)
    A = chl.DeviceArray.from_host(U)
    B = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, recon, precision)
    hy.vl_step(A, h, h, h, dt, G, density_floor=0.0)
    hy.close()
    c = config(N, ng, recon, precision)
    rc = L.chl_hydro_vl_step_noalloc(ctypes.byref(c), B.ptr, h, h, h, dt, G, 0.0, None)
    chl.check(rc)
    chl.check(L.chl_synchronize())
    return A.download(), B.download()


@pytest.mark.parametrize("precision", [chl.PREC_FP64_EXACT, chl.PREC_FP64, chl.PREC_MIXED])
def test_noalloc_equals_the_handle_step(L, precision):
    # first call: a 24^3 rough box (repairs fire); second call: a smaller 12^3 box
    for n, seed in ((24, 0), (12, 1)):
        ng = 4
        U, N = rough_box(n, ng, seed=seed)
        say("noalloc precision", precision, "n", n)
        A, B = step_both(L, U, N, ng, chl.RECON_PPMC, precision)
        np.testing.assert_array_equal(B, A)
        assert (B.reshape(5, N, N, N)[0, ng:-ng, ng:-ng, ng:-ng] > 0).all()
    # and the exact form equals the oracle
    if precision == chl.PREC_FP64_EXACT:
        n, ng = 12, 4
        U, N = rough_box(n, ng, seed=1)
        h = 1.0 / n
        dt = 0.3 / orc.calc_dti(U, N, N, N, ng, h, h, h, G)
        Uo = U.copy()
        orc.lib().orc_vl3d(orc.ptr(Uo), N, N, N, ng, h, h, h, dt, G, 5, orc.PPMC, 1, 0.0)
        _, B = step_both(L, U, N, ng, chl.RECON_PPMC, precision)
        np.testing.assert_array_equal(B, Uo)


def test_noalloc_rejects_a_box_beyond_its_scratch(L):
    cap = L.chl_noalloc_scratch_bytes()
    N = int(round((cap / 40) ** (1 / 3))) + 8  # W alone (5 FP64 fields) exceeds the scratch
    c = config(N, 4, chl.RECON_PPMC, chl.PREC_FP64_EXACT)
    D = chl.DeviceArray(8)
    rc = L.chl_hydro_vl_step_noalloc(ctypes.byref(c), D.ptr, 0.1, 0.1, 0.1, 1e-3, G, 0.0, None)
    assert rc != chl.CHL_OK and b"scratch" in L.chl_last_error()
