"""GPU parity of the HIP hot path against the CPU oracle and the reference's own vectors.

All calls go through the C ABI (libcholla_amd.so).  Bars:
  * PolicyExact (CHL_PREC_FP64_EXACT): bit-identical to the oracle; the test
    asserts per-field relative L1 <= 1e-12 and reports the max |diff| (expected 0);
  * CHL_PREC_FP64: per-field relative L1 <= 1e-12, except the
    Kelvin-Helmholtz box's m_z, a field ~5e-6 of the momentum on which the fast form's FP64
    rounding alone leaves ~2e-11 (KH_FP64_MISS; the exact form is the parity form);
  * CHL_PREC_MIXED: per-field relative L1 <= 1e-5 against the FP64 oracle on every field
    where the form meets it; the Kelvin-Helmholtz box's m_y and m_z do not, and their tests
    assert the measured miss (KH_MIXED_MISS) instead of hiding it.
"""
import json
import os
import sys

import numpy as np
import pytest

import cholla_amd as chl
from helpers import (KH_FP64_MISS, check_fp64, check_mixed, exact_riemann_cell_average, interior, linear_wave_ic,
                     rel_l1, sod_ic, strict_l1)
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_unit_vectors.json")))
TOL = {chl.PREC_FP64_EXACT: 1e-12, chl.PREC_FP64: 1e-12, chl.PREC_MIXED: 1e-5}


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    if chl.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need a real MI355X")
    say("device:", chl.device_name())


def nearly_equal(a, b):
    return abs(a - b) <= 1e-14 or abs(a - b) <= 4 * np.spacing(max(abs(a), abs(b)))


def field_l1(A, B, nf, nz, ny, nx, ng):
    """Strict per-field relative L1 over real cells, and max |diff|."""
    a = interior(A, nf, nz, ny, nx, ng)
    b = interior(B, nf, nz, ny, nx, ng)
    return [rel_l1(a[f], b[f]) for f in range(nf)], float(np.abs(a - b).max())


# The mixed form on the Kelvin-Helmholtz box (24^3 to 64^3, up to 50 steps): m_y and m_z
# miss the 1e-5 bar (m_z is ~5e-6 of the momentum, driven by dp/dz of the seeded noise
# alone); bounds a few times the measured values.
KH_MIXED_MISS = {"momentum_y": 1e-4, "momentum_z": 1e-1}


# ----------------------------------------------------------------------------- stage kernels vs reference vectors
def test_plmc_kernel_reference_vectors():
    g = GOLD["plmc_hydro"]
    n = g["nx"] * g["ny"] * g["nz"] * g["n_fields"]
    grid = orc.uniform(g["seed"], *g["dist"], 2 + n)[2:].copy()
    U = chl.DeviceArray.from_host(grid)
    for d in range(3):
        say("PLMC reference vectors dir", d)
        QL, QR = chl.DeviceArray(n, np.zeros(n)), chl.DeviceArray(n, np.zeros(n))
        nx, ny, nz = g["rotated_dims"][d]
        chl.reconstruct(chl.RECON_PLMC, U, QL, QR, nx, ny, nz, g["gamma"], d)
        for arr, fid in ((QL.download(), g["left"][d]), (QR.download(), g["right"][d])):
            for i, v in enumerate(arr):
                assert nearly_equal(fid.get(str(i), 0.0), v), (d, i, fid.get(str(i), 0.0), v)


def test_ppmc_kernel_reference_vectors():
    g = GOLD["ppmc_vl_hydro"]
    nc = g["nx"] * g["ny"] * g["nz"]
    grid = orc.uniform(g["seed"], *g["dist"], nc * 5)
    U = chl.DeviceArray.from_host(grid)
    for d in range(3):
        say("PPMC reference vectors dir", d)
        QL, QR = chl.DeviceArray(5 * nc, np.zeros(5 * nc)), chl.DeviceArray(5 * nc, np.zeros(5 * nc))
        chl.reconstruct(chl.RECON_PPMC, U, QL, QR, 6, 6, 6, g["gamma"], d)
        L, R = QL.download(), QR.download()
        for i in range(g["checked_length"]):
            assert nearly_equal(g["left"][d].get(str(i), 0.0), L[i]), (d, i)
            if i != 54:
                assert nearly_equal(g["right"][d].get(str(i), 0.0), R[i]), (d, i)
        assert nearly_equal(g["right"][d]["54"], L[4 * nc + 86])


@pytest.mark.parametrize("recon", [orc.PLMC, orc.PPMC])
def test_recon_and_hllc_kernels_bitwise_vs_oracle(recon):
    rng = np.random.default_rng(5)
    nx, ny, nz = 9, 8, 7
    nc = nx * ny * nz
    d = rng.uniform(0.2, 3, nc)
    v = rng.uniform(-2, 2, (3, nc))
    p = rng.uniform(0.1, 4, nc)
    g = 5 / 3
    U = np.concatenate([d, d * v[0], d * v[1], d * v[2], p / (g - 1) + 0.5 * d * (v ** 2).sum(0)])
    Ud = chl.DeviceArray.from_host(U)
    for dr in range(3):
        say("recon", recon, "dir", dr)
        QLo, QRo = np.zeros(5 * nc), np.zeros(5 * nc)
        fn = orc.lib().orc_ppmc_vl if recon == orc.PPMC else orc.lib().orc_plmc
        fn(orc.ptr(U), orc.ptr(QLo), orc.ptr(QRo), nx, ny, nz, g, dr)
        QL, QR = chl.DeviceArray(5 * nc, np.zeros(5 * nc)), chl.DeviceArray(5 * nc, np.zeros(5 * nc))
        chl.reconstruct(recon, Ud, QL, QR, nx, ny, nz, g, dr)
        np.testing.assert_array_equal(QL.download(), QLo)
        np.testing.assert_array_equal(QR.download(), QRo)
        # HLLC on valid interface pairs built from random states (every cell valid)
        Fo = np.zeros(5 * nc)
        L2 = U.copy()
        R2 = np.roll(U.reshape(5, nc), 1, axis=1).reshape(-1).copy()
        orc.lib().orc_hllc(orc.ptr(L2), orc.ptr(R2), orc.ptr(Fo), nx, ny, nz, g, dr)
        F = chl.DeviceArray(5 * nc, np.zeros(5 * nc))
        chl.hllc_fluxes(chl.DeviceArray.from_host(L2), chl.DeviceArray.from_host(R2), F, nx, ny, nz, g, dr)
        np.testing.assert_array_equal(F.download(), Fo)


# ----------------------------------------------------------------------------- boundaries, dt, floors
@pytest.mark.parametrize("flags", [[1] * 6, [3] * 6, [2] * 6, [1, 1, 2, 2, 3, 3], [3, 2, 1, 1, 2, 3]])
def test_ghost_fill_bitwise(flags):
    say("ghost fill", flags)
    rng = np.random.default_rng(9)
    ng, nx, ny, nz = 4, 13, 12, 12  # every axis holds >= n_ghost real cells, so ghosts map to real cells
    U = rng.uniform(0.5, 1.5, 5 * nx * ny * nz)
    Uo = U.copy()
    orc.set_boundaries(Uo, nx, ny, nz, 5, ng, flags)
    Ud = chl.DeviceArray.from_host(U)
    chl.set_boundaries(Ud, nx, ny, nz, ng, flags)
    np.testing.assert_array_equal(Ud.download(), Uo)
    # the single-slab entry point (SetGhostCells) composes to the same result
    Ud2 = chl.DeviceArray.from_host(U)
    jl, ju, kl, ku = ng, ny - ng, ng, nz - ng
    ext = [(0, ng, jl, ju, kl, ku), (nx - ng, nx, jl, ju, kl, ku), (0, nx, 0, ng, kl, ku), (0, nx, ny - ng, ny, kl, ku),
           (0, nx, 0, ny, 0, ng), (0, nx, 0, ny, nz - ng, nz)]
    for dr, (i0, i1, j0, j1, k0, k1) in enumerate(ext):
        chl.set_ghost_cells(Ud2, nx, ny, nz, ng, flags, i1 - i0, j1 - j0, k1 - k0, i0, j0, k0, dr)
    np.testing.assert_array_equal(Ud2.download(), Uo)


def test_calc_dt_and_temperature_floor_bitwise():
    g = GOLD["calc_dt_3d_one_cell"]
    say("calc_dt reference vector")
    U1 = chl.DeviceArray.from_host(np.array(g["conserved"]))
    assert nearly_equal(g["expected_dti"], chl.calc_dti(U1, 1, 1, 1, 0, 1.0, 1.0, 1.0, g["gamma"]))
    rng = np.random.default_rng(4)
    ng, nx, ny, nz = 3, 20, 17, 15
    nc = nx * ny * nz
    d = rng.uniform(0.2, 3, nc)
    v = rng.uniform(-2, 2, (3, nc))
    p = rng.uniform(-0.2, 4, nc)  # some negative pressures exercise the floor
    gam = 1.4
    U = np.concatenate([d, d * v[0], d * v[1], d * v[2], p / (gam - 1) + 0.5 * d * (v ** 2).sum(0)])
    Ud = chl.DeviceArray.from_host(U)
    say("calc_dt random box")
    assert chl.calc_dti(Ud, nx, ny, nz, ng, 0.1, 0.2, 0.3, gam) == orc.calc_dti(U, nx, ny, nz, ng, 0.1, 0.2, 0.3, gam)
    say("temperature floor")
    Uo = U.copy()
    orc.lib().orc_temperature_floor(orc.ptr(Uo), nx, ny, nz, ng, 0.05)
    chl.apply_temperature_floor(Ud, nx, ny, nz, ng, 0.05)
    np.testing.assert_array_equal(Ud.download(), Uo)


# ----------------------------------------------------------------------------- integrated steps
def _random_periodic_box(n, ng, seed=11, g=5 / 3, mach=0.8):
    N = n + 2 * ng
    rng = np.random.default_rng(seed)
    x = (np.arange(N) - ng + 0.5) / n
    d = 1.0 + 0.3 * np.sin(2 * np.pi * x)[None, None, :] * np.cos(2 * np.pi * x)[None, :, None]
    d = np.broadcast_to(d, (N, N, N)).copy() + 0.05 * rng.uniform(-1, 1, (N, N, N))
    v = mach * rng.uniform(-1, 1, (3, N, N, N))
    p = 1.0 + 0.1 * rng.uniform(-1, 1, (N, N, N))
    U = np.stack([d, d * v[0], d * v[1], d * v[2], p / (g - 1) + 0.5 * d * (v ** 2).sum(0)]).reshape(-1).copy()
    orc.set_boundaries(U, N, N, N, 5, ng, [1] * 6)
    return U, N, g


@pytest.mark.parametrize("precision", [chl.PREC_FP64_EXACT, chl.PREC_FP64, chl.PREC_MIXED])
@pytest.mark.parametrize("recon,ng", [(chl.RECON_PLMC, 3), (chl.RECON_PPMC, 4)])
def test_single_vl_step(recon, ng, precision):
    n = 20
    U, N, g = _random_periodic_box(n, ng)
    h = 1.0 / n
    dt = 0.3 / orc.calc_dti(U, N, N, N, ng, h, h, h, g)
    say("vl_step recon", recon, "precision", precision, "N", N)
    Uo = U.copy()
    orc.lib().orc_vl3d(orc.ptr(Uo), N, N, N, ng, h, h, h, dt, g, 5, recon, 1, 0.0)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, recon, precision)
    hy.vl_step(Ud, h, h, h, dt, g)
    A = Ud.download()
    l1, mx = field_l1(A, Uo, 5, N, N, N, ng)
    say("  rel L1", ["%.3e" % x for x in l1], "max|diff| %.3e" % mx)
    assert max(l1) <= TOL[precision], l1
    if precision == chl.PREC_FP64_EXACT:
        assert mx == 0.0, f"exact form differs from the oracle by {mx}"


@pytest.mark.parametrize("tile1,tile2", [(0, 0), (1, 1), (2, 2), (3, 3), (0, 3)])
def test_tiles_and_chunks_identical(tile1, tile2):
    n, ng = 24, 4
    U, N, g = _random_periodic_box(n, ng, seed=3)
    h = 1.0 / n
    say("tiles", tile1, tile2)
    outs = []
    for k1, k2 in ((0, 0), (7, 5)):
        Ud = chl.DeviceArray.from_host(U)
        hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, chl.PREC_FP64_EXACT, tile1=tile1, tile2=tile2, kchunk1=k1,
                       kchunk2=k2)
        hy.run(Ud, h, h, h, g, 0.0, 1.0, 3, use_graph=False)
        outs.append(Ud.download())
    Uo = U.copy()
    orc.run(Uo, N, N, N, ng, 5, h, h, h, g, orc.PPMC, [1] * 6, 0.0, 1.0, 3)
    np.testing.assert_array_equal(outs[0], Uo)
    np.testing.assert_array_equal(outs[1], Uo)


@pytest.mark.parametrize("precision", [chl.PREC_FP64_EXACT, chl.PREC_FP64, chl.PREC_MIXED])
def test_kelvin_helmholtz_multi_step(precision):
    """KH box (the benchmark's initial state) at 32^3, 20 steps, device-timed loop in a hipGraph."""
    n, ng, g = 32, 4, 5 / 3
    N = n + 2 * ng
    U = chl.ic_kelvin_helmholtz(n, ng, g, seed=7)
    h = 1.0 / n
    Uo = U.copy()
    dts = np.zeros(20)
    steps_o, to = orc.run(Uo, N, N, N, ng, 5, h, h, h, g, orc.PPMC, [1] * 6, 0.0, 1e30, 20, dt_log=dts)
    say("KH 32^3 x20 precision", precision)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, precision)
    steps, t = hy.run(Ud, h, h, h, g, 0.0, 1e30, 20, use_graph=True)
    assert steps == steps_o == 20
    assert abs(t - to) <= TOL[precision] * to
    A = Ud.download()
    l1, mx = field_l1(A, Uo, 5, N, N, N, ng)
    say("  rel L1", ["%.3e" % x for x in l1], "max|diff| %.3e" % mx)
    if precision == chl.PREC_FP64_EXACT:
        assert mx == 0.0
    elif precision == chl.PREC_MIXED:
        check_mixed(l1, KH_MIXED_MISS)
    else:
        check_fp64(l1, KH_FP64_MISS)
    # ghost cells after the fused periodic fill equal a fresh fill of the real cells
    B = A.copy()
    orc.set_boundaries(B, N, N, N, 5, ng, [1] * 6)
    np.testing.assert_array_equal(A, B)


@pytest.mark.xfail(strict=True, reason="measured miss: the mixed form's m_y and m_z on the KH box are above the "
                                      "1e-5 bar (FP32 flux rounding; tests/test_mixed_error_model.py)")
def test_kelvin_helmholtz_mixed_north_star_bar():
    """The north star's bar as stated, per-field relative L1 <= 1e-5 on all five fields, KH
    64^3 x 50 steps against the FP64 oracle.  Density, m_x and energy meet it; the test is
    expected to fail on m_y and m_z and reports their values."""
    n, ng, g = 64, 4, 5 / 3
    N = n + 2 * ng
    U = chl.ic_kelvin_helmholtz(n, ng, g, seed=7)
    h = 1.0 / n
    Uo = U.copy()
    orc.run(Uo, N, N, N, ng, 5, h, h, h, g, orc.PPMC, [1] * 6, 0.0, 1e30, 50)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, chl.PREC_MIXED)
    hy.run(Ud, h, h, h, g, 0.0, 1e30, 50, use_graph=True)
    l1 = strict_l1(Ud.download(), Uo, 5, N, N, N, ng)
    say("KH 64^3 x50 mixed strict L1", ["%.3e" % x for x in l1])
    hy.close()
    check_mixed(l1, {"momentum_y": 1e-4, "momentum_z": 1e-1})  # the measured miss, asserted
    check_mixed(l1)  # the bar itself


def test_graph_replay_matches_eager():
    n, ng, g = 16, 4, 5 / 3
    U, N, _ = _random_periodic_box(n, ng, seed=21)
    h = 1.0 / n
    res = []
    for use_graph in (False, True, True):
        say("graph replay", use_graph)
        Ud = chl.DeviceArray.from_host(U)
        hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, chl.PREC_FP64)
        hy.run(Ud, h, h, h, g, 0.0, 1e30, 7, use_graph=use_graph)
        res.append(Ud.download())
    np.testing.assert_array_equal(res[0], res[1])
    np.testing.assert_array_equal(res[0], res[2])


@pytest.mark.parametrize("recon", [chl.RECON_PLMC, chl.RECON_PPMC])
def test_linear_wave_system(recon):
    """hydro_system_tests.cpp:196-301 on the GPU: exact step count and L1 vs the initial state."""
    gw = GOLD["linear_waves_3d"]
    c = gw["cases"]["sound_right"]
    ng = 3 if recon == chl.RECON_PLMC else 4
    U, (nx, ny, nz), (dx, dy, dz) = linear_wave_ic(gw["nx"], gw["ny"], gw["nz"], ng, gw["xlen"], gw["ylen"], gw["zlen"],
                                                    gw["gamma"], gw["rho"], c["vx"], gw["P"], gw["A"], c["eig"])
    orc.set_boundaries(U, nx, ny, nz, 5, ng, [1] * 6)
    U0 = interior(U, 5, nz, ny, nx, ng).copy()
    say("linear wave recon", recon)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(nx, ny, nz, ng, recon, chl.PREC_FP64_EXACT)
    steps, t = hy.run(Ud, dx, dy, dz, gw["gamma"], 0.0, 1.0, 100000, use_graph=True)
    assert steps == c["steps"] and t == 1.0
    allowed = 2 * gw["allowed_L1"]["PLMC" if recon == chl.RECON_PLMC else "PPMC"]
    U1 = interior(Ud.download(), 5, nz, ny, nx, ng)
    l1 = [np.abs(U1[f] - U0[f]).mean() for f in range(5)]
    assert max(l1) < allowed, l1


@pytest.mark.parametrize("use_graph", [False, True])
def test_sod_1d_large_unfused_dt(use_graph):
    """A 1D tube above 1024 cells without dual energy: the next inverse dt is its own grid-wide
    reduction after the step, whose accumulator stage 1 zeroes in-kernel (no memset in the
    captured step).  2000 cells, PPMC + HLLC, to t = 0.05: bit-identical to the oracle."""
    U, nx, dx = sod_ic(n_real=2000, ng=4)
    ng = 4
    orc.set_boundaries(U, nx, 1, 1, 5, ng, [3, 3, 0, 0, 0, 0])
    Uo = U.copy()
    steps_o, to = orc.run(Uo, nx, 1, 1, ng, 5, dx, dx, dx, 1.4, orc.PPMC, [3, 3, 0, 0, 0, 0], 0.0, 0.05, 100000)
    say("Sod 1D, 2000 cells, graph", use_graph)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(nx, 1, 1, ng, chl.RECON_PPMC, chl.PREC_FP64_EXACT, bc=[3, 3, 1, 1, 1, 1])
    steps, t = hy.run(Ud, dx, dx, dx, 1.4, 0.0, 0.05, 100000, use_graph=use_graph)
    assert steps == steps_o and t == to
    A = Ud.download().reshape(5, nx)
    np.testing.assert_array_equal(A[:, ng:nx - ng], Uo.reshape(5, nx)[:, ng:nx - ng])
    hy.close()


def test_sod_1d():
    U, nx, dx = sod_ic()
    ng = 3
    flags = [3, 3, 1, 1, 1, 1]
    orc.set_boundaries(U, nx, 1, 1, 5, ng, [3, 3, 0, 0, 0, 0])
    Uo = U.copy()
    steps_o, to = orc.run(Uo, nx, 1, 1, ng, 5, dx, dx, dx, 1.4, orc.PLMC, [3, 3, 0, 0, 0, 0], 0.0, 0.2, 100000)
    say("Sod 1D")
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(nx, 1, 1, ng, chl.RECON_PLMC, chl.PREC_FP64_EXACT, bc=flags)
    steps, t = hy.run(Ud, dx, dx, dx, 1.4, 0.0, 0.2, 100000, use_graph=False)
    assert steps == steps_o and t == to
    A = Ud.download().reshape(5, nx)
    np.testing.assert_array_equal(A[:, ng:nx - ng], Uo.reshape(5, nx)[:, ng:nx - ng])
    ex = exact_riemann_cell_average(1.0, 0.0, 1.0, 0.1, 0.0, 0.1, 1.4, np.arange(101) * dx, 0.5, 0.2)
    l1 = np.abs(A[0, ng:nx - ng] - ex[0]).mean()
    say("  Sod density L1 vs exact", l1)
    assert l1 < 6e-3
