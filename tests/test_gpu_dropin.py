"""The reference-signature entry points (include/cholla_dropin.hpp) driven the way Cholla's
main loop and Grid3D drive them (cholla_amd/dropin_driver_*), checked against the oracle.

The drivers are plain C++ programs compiled against the adapter with the reference's
make.type macros (-DPLMC or -DPPMC, -DDENSITY_FLOOR, -DTEMPERATURE_FLOOR); they use the
host-side dt sequence (Calc_dt_GPU -> set_dt -> clamp) of src/main.cpp, so their result
must be bit-identical to the oracle's driver loop."""
import json
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import cholla_amd as chl
from helpers import interior, linear_wave_ic, sod_ic
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_unit_vectors.json")))
HDR = "<12i8d"


def run_driver(name, U, dims, ng, dxyz, gamma, t, tout, max_steps, flags, tmp_path, U_floor=0.0, dfloor=0.0, nf=5):
    nx, ny, nz = dims
    fin, fout = os.path.join(tmp_path, "in.bin"), os.path.join(tmp_path, "out.bin")
    with open(fin, "wb") as f:
        f.write(struct.pack(HDR, nx, ny, nz, ng, nf, max_steps, *flags, *dxyz, gamma, t, tout, U_floor, dfloor))
        f.write(np.ascontiguousarray(U, dtype="<f8").tobytes())
    exe = os.path.join(os.path.dirname(chl.LIB_PATH), name)
    print("[gpu-case] driver", name, dims, flush=True)
    r = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(fout, "rb").read()
    hsz = struct.calcsize(HDR)
    n = nf * nx * ny * nz
    A = np.frombuffer(raw[hsz:hsz + 8 * n], dtype="<f8").copy()
    ns, = struct.unpack("<i", raw[hsz + 8 * n:hsz + 8 * n + 4])
    tt, = struct.unpack("<d", raw[hsz + 8 * n + 4:hsz + 8 * n + 12])
    return A, ns, tt


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    if chl.device_count() < 1:
        pytest.fail("no HIP device visible")


def test_dropin_linear_wave_ppmc(tmp_path):
    gw = GOLD["linear_waves_3d"]
    c = gw["cases"]["sound_right"]
    ng = 4
    U, dims, dxyz = linear_wave_ic(gw["nx"], gw["ny"], gw["nz"], ng, gw["xlen"], gw["ylen"], gw["zlen"], gw["gamma"],
                                   gw["rho"], c["vx"], gw["P"], gw["A"], c["eig"])
    A, ns, t = run_driver("dropin_driver_ppmc", U, dims, ng, dxyz, gw["gamma"], 0.0, 1.0, 100000, [1] * 6, tmp_path)
    Uo = U.copy()
    orc.set_boundaries(Uo, *dims, 5, ng, [1] * 6)
    steps_o, to = orc.run(Uo, *dims, ng, 5, *dxyz, gw["gamma"], orc.PPMC, [1] * 6, 0.0, 1.0, 100000)
    assert ns == steps_o == c["steps"] and t == to
    np.testing.assert_array_equal(A, Uo)


def test_dropin_kelvin_helmholtz_ppmc(tmp_path):
    n, ng, g = 24, 4, 5 / 3
    N = n + 2 * ng
    U = chl.ic_kelvin_helmholtz(n, ng, g, seed=11)
    A, ns, t = run_driver("dropin_driver_ppmc", U, (N, N, N), ng, (1 / n,) * 3, g, 0.0, 1e30, 12, [1] * 6, tmp_path)
    Uo = U.copy()
    steps_o, to = orc.run(Uo, N, N, N, ng, 5, 1 / n, 1 / n, 1 / n, g, orc.PPMC, [1] * 6, 0.0, 1e30, 12)
    assert ns == steps_o == 12 and t == to
    np.testing.assert_array_equal(A, Uo)


def test_dropin_sod_plmc(tmp_path):
    U, nx, dx = sod_ic()
    ng = 3
    flags = [3, 3, 0, 0, 0, 0]
    A, ns, t = run_driver("dropin_driver_plmc", U, (nx, 1, 1), ng, (dx, dx, dx), 1.4, 0.0, 0.2, 100000, flags,
                          tmp_path)
    Uo = U.copy()
    orc.set_boundaries(Uo, nx, 1, 1, 5, ng, flags)
    steps_o, to = orc.run(Uo, nx, 1, 1, ng, 5, dx, dx, dx, 1.4, orc.PLMC, flags, 0.0, 0.2, 100000)
    assert ns == steps_o and t == to == 0.2
    np.testing.assert_array_equal(A, Uo)


def test_dropin_mixed_boundaries_plmc(tmp_path):
    """Reflective + transmissive + periodic faces through SetGhostCells slab by slab."""
    n, ng, g = 16, 3, 1.4
    N = n + 2 * ng
    rng = np.random.default_rng(2)
    d = 1 + 0.2 * rng.uniform(-1, 1, (N, N, N))
    v = 0.3 * rng.uniform(-1, 1, (3, N, N, N))
    p = 1 + 0.2 * rng.uniform(-1, 1, (N, N, N))
    U = np.stack([d, d * v[0], d * v[1], d * v[2], p / (g - 1) + 0.5 * d * (v ** 2).sum(0)]).reshape(-1).copy()
    flags = [2, 3, 1, 1, 3, 2]
    A, ns, t = run_driver("dropin_driver_plmc", U, (N, N, N), ng, (1 / n,) * 3, g, 0.0, 1e30, 8, flags, tmp_path)
    Uo = U.copy()
    orc.set_boundaries(Uo, N, N, N, 5, ng, flags)
    steps_o, to = orc.run(Uo, N, N, N, ng, 5, 1 / n, 1 / n, 1 / n, g, orc.PLMC, flags, 0.0, 1e30, 8)
    assert ns == steps_o == 8 and t == to
    np.testing.assert_array_equal(A, Uo)


@pytest.mark.parametrize("driver,recon,ng", [("dropin_driver_mhd", orc.MHD_PLMC, 4),
                                              ("dropin_driver_mhd_ppmc", orc.MHD_PPMC, 5),
                                              ("dropin_driver_mhd_pcm", orc.MHD_PCM, 3)])
def test_dropin_mhd_orszag_tang(tmp_path, driver, recon, ng):
    """The MHD builds of the adapter (-DMHD with -DPLMC, -DPPMC or -DPCM, each with the
    reference's ghost count) driven like Cholla's loop: bit-identical to the MHD oracle's
    driver loop on an Orszag-Tang box."""
    n, g = 16, 5 / 3
    N = n + 2 * ng
    h = 1.0 / n
    U = chl.ic_orszag_tang(n, ng, g, seed=9)
    A, ns, t = run_driver(driver, U, (N, N, N), ng, (h, h, h), g, 0.0, 1e30, 6, [1] * 6, tmp_path, nf=8)
    Uo = U.copy()
    with orc.mhd_reconstruction(recon):
        n_o, t_o = orc.mhd_run(Uo, N, N, N, ng, h, h, h, g, [1] * 6, 0.0, 1e30, 6)
    assert ns == n_o == 6 and t == t_o
    np.testing.assert_array_equal(A, Uo)


def test_dropin_dual_energy_ppmc(tmp_path):
    """The DE build of the adapter (-DPPMC -DDE): six fields with GasEnergy last, driven like
    Cholla's loop (integrator with internal-energy selection and sync, temperature floor of
    E and GasEnergy, dt), bit-identical to the oracle's DE driver loop."""
    from test_gpu_de import slab_box

    n, ng, g = 16, 4, 5 / 3
    U, dims = slab_box(n, ng, seed=21)
    N = dims[0]
    h = 1.0 / n
    A, ns, t = run_driver("dropin_driver_ppmc_de", U, dims, ng, (h, h, h), g, 0.0, 1e30, 8, [1] * 6, tmp_path,
                          U_floor=1.5, nf=6)
    Uo = U.copy()
    steps_o, to = orc.run(Uo, N, N, N, ng, 6, h, h, h, g, orc.PPMC, [1] * 6, 0.0, 1e30, 8, U_floor=1.5,
                          dual_energy_on=True)
    assert ns == steps_o == 8 and t == to
    np.testing.assert_array_equal(A, Uo)


def test_dropin_static_gravity_plmc(tmp_path):
    """The -DSTATIC_GRAV build of the adapter with Parameters::custom_grav = 1 (NFW halo +
    Miyamoto-Nagai disk) on a 3D box, driven like Cholla's loop: the adapter leaves the stage-2
    tile to the library, which must take one with a gravity kernel.  The field calls log, pow and
    sqrt, whose last bit differs between the host's and the device's math libraries, so the bar
    is a per-field relative L1 of 1e-12 against the oracle's driver loop."""
    from test_gpu_gravity import G, disk_box, rel_l1

    n, ng = 16, 3
    U, (N, _, _) = disk_box((n, n, n), ng, vamp=1e-2)
    orc.set_boundaries(U, N, N, N, 5, ng, [1] * 6)
    h = 0.6
    exe = os.path.join(os.path.dirname(chl.LIB_PATH), "dropin_driver_plmc_grav")
    fin, fout = os.path.join(tmp_path, "in.bin"), os.path.join(tmp_path, "out.bin")
    with open(fin, "wb") as f:
        f.write(struct.pack(HDR, N, N, N, ng, 5, 5, *([1] * 6), h, h, h, G, 0.0, 1e30, 0.0, 0.0))
        f.write(np.ascontiguousarray(U, dtype="<f8").tobytes())
    print("[gpu-case] driver dropin_driver_plmc_grav", N, flush=True)
    r = subprocess.run([exe, fin, fout, "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(fout, "rb").read()
    hsz = struct.calcsize(HDR)
    A = np.frombuffer(raw[hsz:hsz + 8 * U.size], dtype="<f8").copy()
    ns, = struct.unpack("<i", raw[hsz + 8 * U.size:hsz + 8 * U.size + 4])
    Uo = U.copy()
    with orc.static_gravity(1, (0, 0, 0), (0.0, 0.0, 0.0)):
        steps_o, _ = orc.run(Uo, N, N, N, ng, 5, h, h, h, G, orc.PLMC, [1] * 6, 0.0, 1e30, 5)
    assert ns == steps_o == 5
    l1 = rel_l1(interior(A, 5, N, N, N, ng), interior(Uo, 5, N, N, N, ng))
    assert max(l1) <= 1e-12, l1
    # the field changed the state: the same run without gravity differs by far more
    Un = U.copy()
    orc.run(Un, N, N, N, ng, 5, h, h, h, G, orc.PLMC, [1] * 6, 0.0, 1e30, 5)
    assert max(rel_l1(interior(Un, 5, N, N, N, ng), interior(Uo, 5, N, N, N, ng))) > 1e-9


def test_dropin_mhd_checks_divergence(tmp_path):
    """The MHD driver calls mhd::checkMagneticDivergence(G) after every step (main.cpp:375-378),
    which prints the maximum |div B| and would exit above 1e-9; on an Orszag-Tang box it prints
    one round-off value per step."""
    n, ng, g = 16, 4, 5 / 3
    N = n + 2 * ng
    h = 1.0 / n
    U = chl.ic_orszag_tang(n, ng, g, seed=9)
    fin, fout = os.path.join(tmp_path, "in.bin"), os.path.join(tmp_path, "out.bin")
    with open(fin, "wb") as f:
        f.write(struct.pack(HDR, N, N, N, ng, 8, 3, *([1] * 6), h, h, h, g, 0.0, 1e30, 0.0, 0.0))
        f.write(np.ascontiguousarray(U, dtype="<f8").tobytes())
    exe = os.path.join(os.path.dirname(chl.LIB_PATH), "dropin_driver_mhd")
    r = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("Global maximum magnetic divergence")]
    assert len(lines) == 3, r.stdout
    assert all(float(l.split("=")[1]) < 1e-10 for l in lines), lines
