"""GPU parity of the floors and repairs of the full update, bit for bit against the oracle.

The reference repairs a non-positive density with the mean of the six neighbours
(Average_Cell_Single_Field, src/hydro/hydro_cuda.cu:293-297) and, in a build without both
floors, a negative or NaN density or energy with the healthy 3x3x3 mean
(Average_Cell_All_Fields, :372-382).  Here stage 2 marks such cells and RepairKernel repairs
them in cell order after the update (the oracle's serial order, tests/test_oracle_repair.py).
The rough box (tests/helpers.py) leaves a few non-positive densities -- one next to a ghost
cell -- and dozens of negative energies after one step at CFL 0.3.
"""
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


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    if chl.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need a real MI355X")


def oracle_update_grid(U, N, ng, h, dt, density_floor_on, temperature_floor_on, U_floor, n_fields=5, de=False):
    """Update_Hydro_Grid's order: the VL step (with its floor and repairs), the temperature
    floor, the next inverse dt, then the ghost fill (main.cpp)."""
    with orc.crash_repair(not (density_floor_on and temperature_floor_on)), orc.dual_energy(de):
        orc.lib().orc_vl3d(orc.ptr(U), N, N, N, ng, h, h, h, dt, G, n_fields, orc.PPMC, int(density_floor_on), 0.0)
        if temperature_floor_on:
            orc.lib().orc_temperature_floor(orc.ptr(U), N, N, N, ng, U_floor)
        dti = orc.calc_dti(U, N, N, N, ng, h, h, h, G)
        orc.set_boundaries(U, N, N, N, n_fields, ng, [1] * 6)
    return dti


def test_density_repair_single_step():
    n, ng = 16, 4
    U, N = rough_box(n, ng, seed=0)
    h = 1.0 / n
    dt = 0.3 / orc.calc_dti(U, N, N, N, ng, h, h, h, G)
    say("density repair vl_step, rough box 16^3 seed 0")
    Uo = U.copy()
    orc.lib().orc_vl3d(orc.ptr(Uo), N, N, N, ng, h, h, h, dt, G, 5, orc.PPMC, 1, 0.0)
    D = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, chl.PREC_FP64_EXACT)
    hy.vl_step(D, h, h, h, dt, G, density_floor=0.0)
    A = D.download()
    np.testing.assert_array_equal(A, Uo)
    assert (A.reshape(5, N, N, N)[0, ng:-ng, ng:-ng, ng:-ng] > 0).all()
    # a second step on the same handle starts with no marks left
    hy.vl_step(D, h, h, h, 0.5 * dt, G, density_floor=0.0)
    orc.lib().orc_vl3d(orc.ptr(Uo), N, N, N, ng, h, h, h, 0.5 * dt, G, 5, orc.PPMC, 1, 0.0)
    np.testing.assert_array_equal(D.download(), Uo)
    hy.close()


@pytest.mark.parametrize("floors", [(True, True), (True, False), (False, True), (False, False)])
def test_update_grid_with_repairs(floors):
    """Density floor and temperature floor build options: with both, the fused path (periodic
    images, temperature floor and dt in stage 2's epilogue; a repaired cell next to a ghost reads
    the side buffer); without both, the crash repair and the unfused floor, dt and fill."""
    dfl, tfl = floors
    n, ng = 16, 4
    U, N = rough_box(n, ng, seed=0)
    h = 1.0 / n
    dt = 0.3 / orc.calc_dti(U, N, N, N, ng, h, h, h, G)
    U_floor = 0.05
    say("update_grid density_floor", dfl, "temperature_floor", tfl)
    Uo = U.copy()
    dti_o = oracle_update_grid(Uo, N, ng, h, dt, dfl, tfl, U_floor)
    D = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, chl.PREC_FP64_EXACT, density_floor=dfl, temperature_floor=tfl)
    dti = hy.update_grid(D, h, h, h, dt, G, density_floor=0.0, U_floor=U_floor)
    A = D.download()
    np.testing.assert_array_equal(A, Uo)
    assert dti == dti_o
    hy.close()


def test_dual_energy_density_repair():
    n, ng = 16, 4
    U5, N = rough_box(n, ng, seed=0)
    h = 1.0 / n
    V = U5.reshape(5, -1)
    d, E = V[0], V[4]
    ke = 0.5 * (V[1] ** 2 + V[2] ** 2 + V[3] ** 2) / d
    U = np.concatenate([U5, E - ke]).copy()  # GasEnergy = internal energy density
    dt = 0.3 / orc.calc_dti(U5, N, N, N, ng, h, h, h, G)
    say("dual energy density repair")
    Uo = U.copy()
    dti_o = oracle_update_grid(Uo, N, ng, h, dt, True, True, 0.0, n_fields=6, de=True)
    D = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, chl.PREC_FP64_EXACT, dual_energy=True, tile1=0, tile2=0)
    dti = hy.update_grid(D, h, h, h, dt, G, density_floor=0.0, U_floor=0.0)
    np.testing.assert_array_equal(D.download(), Uo)
    assert dti == dti_o
    hy.close()


@pytest.mark.parametrize("precision", [chl.PREC_FP64, chl.PREC_MIXED])
def test_repair_in_the_fast_forms(precision):
    """The FP64 and mixed forms repair through the same kernel: every density stays positive
    and each repaired cell holds the mean of its neighbours as the form computed them."""
    n, ng = 16, 4
    U, N = rough_box(n, ng, seed=0)
    h = 1.0 / n
    dt = 0.3 / orc.calc_dti(U, N, N, N, ng, h, h, h, G)
    D = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, precision, density_floor=True, temperature_floor=True)
    hy.vl_step(D, h, h, h, dt, G, density_floor=0.0)
    A = D.download().reshape(5, N, N, N)
    assert (A[0, ng:-ng, ng:-ng, ng:-ng] > 0).all() and np.isfinite(A).all()
    hy.close()


def test_positive_density_floor_fires_in_both_stages():
    """KH 24^3 (density 1 or 2) with a density floor of 1.5: every light cell is floored in the
    half step and in the full update, 4 device-timed steps, bit for bit."""
    n, ng = 24, 4
    N = n + 2 * ng
    h = 1.0 / n
    U = chl.ic_kelvin_helmholtz(n, ng, G, seed=5)
    Uo = U.copy()
    steps_o, t_o = orc.run(Uo, N, N, N, ng, 5, h, h, h, G, orc.PPMC, [1] * 6, 0.0, 1e30, 4, density_floor=1.5)
    D = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, chl.PREC_FP64_EXACT)
    steps, t = hy.run(D, h, h, h, G, 0.0, 1e30, 4, use_graph=True, density_floor=1.5)
    A = D.download()
    assert steps == steps_o and t == t_o
    np.testing.assert_array_equal(A, Uo)
    assert A.reshape(5, N, N, N)[0].min() >= 1.5
    hy.close()


def test_temperature_floor_in_the_fused_epilogue():
    """U_floor above the internal energy of the dense layer: the stage-2 epilogue's floor
    fires on every dense cell, 4 device-timed steps in a hipGraph, bit for bit."""
    n, ng = 24, 4
    N = n + 2 * ng
    h = 1.0 / n
    U = chl.ic_kelvin_helmholtz(n, ng, G, seed=6)
    Uo = U.copy()
    steps_o, t_o = orc.run(Uo, N, N, N, ng, 5, h, h, h, G, orc.PPMC, [1] * 6, 0.0, 1e30, 4, U_floor=3.0)
    D = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, chl.PREC_FP64_EXACT)
    steps, t = hy.run(D, h, h, h, G, 0.0, 1e30, 4, use_graph=True, U_floor=3.0)
    A = D.download()
    assert steps == steps_o and t == t_o
    np.testing.assert_array_equal(A, Uo)
    V = A.reshape(5, N, N, N)[:, ng:-ng, ng:-ng, ng:-ng]
    eint = (V[4] - 0.5 * (V[1] ** 2 + V[2] ** 2 + V[3] ** 2) / V[0]) / V[0]
    assert eint.min() >= 3.0 * (1 - 1e-12)
    hy.close()


@pytest.mark.parametrize("launch,looped", [(chl.LAUNCH_SPLIT, False), (chl.LAUNCH_PERSISTENT, False),
                                           (chl.LAUNCH_SPLIT, True)])
def test_ensemble_repairs_each_box(launch, looped):
    """Two rough boxes of different sizes in one ensemble (split, persistent and per-box loop
    launches): one step equals the oracle's on each box, repairs included."""
    ng = 4
    boxes = [rough_box(16, ng, seed=0), rough_box(12, ng, seed=2)]
    states = [U for U, _ in boxes]
    dims = [(N, N, N) for _, N in boxes]
    dxyz = [(1.0 / (N - 2 * ng),) * 3 for _, N in boxes]
    say("ensemble repairs, launch", launch, "looped", looped)
    ens = chl.Ensemble(dims, ng, chl.RECON_PPMC, chl.PREC_FP64_EXACT, launch, tile1=0, tile2=0)
    dev = [chl.DeviceArray.from_host(U) for U in states]
    ens.bind(dev, np.array(dxyz).reshape(-1))
    ens.start(G)
    if looped:
        ens.advance_looped(1)
    else:
        ens.advance(1)
    for b, U in enumerate(states):
        N = dims[b][0]
        Uo = U.copy()
        orc.run(Uo, N, N, N, ng, 5, *dxyz[b], G, orc.PPMC, [1] * 6, 0.0, 1e30, 1)
        np.testing.assert_array_equal(dev[b].download(), Uo)
    ens.close()


def rough_mhd_box(n, ng, seed):
    """The rough hydro box with a uniform field (no divergence) and its magnetic energy."""
    U5, N = rough_box(n, ng, seed=seed, mach=4.0, d_lo=1e-3)
    V = U5.reshape(5, -1)
    B = np.array([0.1, 0.05, 0.02])
    return np.concatenate([V[:4].reshape(-1), V[4] + 0.5 * (B ** 2).sum(),
                           np.full(N ** 3, B[0]), np.full(N ** 3, B[1]), np.full(N ** 3, B[2])]).copy(), N


@pytest.mark.parametrize("floors", [(True, True), (False, False)])
def test_mhd_repairs(floors):
    """An MHD step with six times the CFL dt on the rough box leaves five non-positive
    densities and negative energies; the density repair (both floors) and the crash repair
    (neither floor) match the oracle bit for bit."""
    dfl, tfl = floors
    n, ng = 16, 4
    U, N = rough_mhd_box(n, ng, seed=0)
    h = 1.0 / n
    dt = 6.0 * 0.3 / orc.mhd_calc_dti(U, N, N, N, ng, h, h, h, G)
    say("MHD repairs, floors", floors)
    Uo = U.copy()
    with orc.crash_repair(not (dfl and tfl)):
        orc.lib().orc_mhd_vl3d(orc.ptr(Uo), N, N, N, ng, h, h, h, dt, G, int(dfl), 0.0)
    D = chl.DeviceArray.from_host(U)
    hm = chl.Mhd(N, N, N, ng, chl.PREC_FP64_EXACT, density_floor=dfl, temperature_floor=tfl)
    hm.vl_step(D, h, h, h, dt, G, density_floor=0.0)
    A = D.download()
    # ghost faces of the result are the boundary fill's business on both sides (the oracle's
    # unfused field update also writes ghost faces, from this step's wild ghost-zone fluxes)
    orc.set_boundaries(A, N, N, N, 8, ng, [1] * 6)
    orc.set_boundaries(Uo, N, N, N, 8, ng, [1] * 6)
    a = Uo.reshape(8, N, N, N)[:, ng:-ng, ng:-ng, ng:-ng]
    assert (a[0] > 0).all() and np.isfinite(a).all()
    np.testing.assert_array_equal(A, Uo)
    hm.close()
