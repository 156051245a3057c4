"""GPU parity of the HLL, Roe and exact Riemann solver options (the reference's -DHLL, -DROE
and -DEXACT builds: Calculate_HLL_Fluxes_CUDA, src/riemann_solvers/hll_cuda.cu:19-268,
Calculate_Roe_Fluxes_CUDA, src/riemann_solvers/roe_cuda.cu:19-400, and
Calculate_Exact_Fluxes_CUDA, src/riemann_solvers/exact_cuda.cu:20-334;
chl_hydro_config.riemann = CHL_RIEMANN_HLL / _ROE / _EXACT) against the oracle's solvers
(tests/test_oracle_hll.py, test_oracle_roe.py and test_oracle_exact.py pin them).  The
exact solver runs in FP64 in every form; its FP32 powers are the correctly rounded float
of the double pow on both sides.

Bars, as for HLLC:
  * the component entries chl_hll_fluxes / chl_roe_fluxes are bit-identical to the oracle
    on random interface pairs in all three directions (for Roe, pairs far enough apart
    that the HLLE fallback fires on some faces);
  * PolicyExact: a VL step and a graph-captured KH run with PLMC and PPMC are
    bit-identical to the oracle with the solver (and differ from the HLLC run, so the
    switch is live); PolicyFast and the mixed form within per-field relative L1 1e-12
    and strict L1 1e-5 (the KH box's m_y, m_z: the measured miss);
  * 1D Sod with PLMC bit-identical to the oracle's 1D driver.  The reference's VL_1D
    integrator has no HLL branch (VL_1D_cuda.cu:69-124 names Exact, Roe and HLLC only);
    the 1D HLL form follows the oracle's VL_1D with the solver swapped.
"""
import sys

import numpy as np
import pytest

import cholla_amd as chl
from helpers import sod_ic
from oracle import oracle as orc
from helpers import KH_FP64_MISS, check_fp64, check_mixed
from test_gpu_parity import KH_MIXED_MISS, TOL, _random_periodic_box, field_l1

pytestmark = pytest.mark.gpu
# (C-ABI solver, oracle solver, component entry, oracle component)
SOLVERS = {"hll": (chl.RIEMANN_HLL, orc.HLL, chl.hll_fluxes, "orc_hll"),
           "roe": (chl.RIEMANN_ROE, orc.ROE, chl.roe_fluxes, "orc_roe"),
           "exact": (chl.RIEMANN_EXACT, orc.EXACT, chl.exact_fluxes, "orc_exact")}


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    if chl.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need a real MI355X")


@pytest.mark.parametrize("name", list(SOLVERS))
def test_component_bitwise_vs_oracle(name):
    solver, _, entry, orc_entry = SOLVERS[name]
    rng = np.random.default_rng(11)
    nx, ny, nz = 9, 8, 7
    nc = nx * ny * nz
    g = 5 / 3
    d = rng.uniform(0.2, 3, nc)
    v = rng.uniform(-3, 3, (3, nc))  # subsonic and supersonic faces; receding pairs for Roe's HLLE
    p = rng.uniform(0.1, 4, nc)
    U = np.concatenate([d, d * v[0], d * v[1], d * v[2], p / (g - 1) + 0.5 * d * (v ** 2).sum(0)])
    L2 = U.copy()
    R2 = np.roll(U.reshape(5, nc), 1, axis=1).reshape(-1).copy()
    for dr in range(3):
        say(name, "component dir", dr)
        Fo = np.zeros(5 * nc)
        getattr(orc.lib(), orc_entry)(orc.ptr(L2), orc.ptr(R2), orc.ptr(Fo), nx, ny, nz, g, dr)
        F = chl.DeviceArray(5 * nc, np.zeros(5 * nc))
        entry(chl.DeviceArray.from_host(L2), chl.DeviceArray.from_host(R2), F, nx, ny, nz, g, dr)
        np.testing.assert_array_equal(F.download(), Fo)


@pytest.mark.parametrize("name", list(SOLVERS))
@pytest.mark.parametrize("precision", [chl.PREC_FP64_EXACT, chl.PREC_FP64, chl.PREC_MIXED])
@pytest.mark.parametrize("recon,ng", [(chl.RECON_PLMC, 3), (chl.RECON_PPMC, 4)])
def test_single_vl_step(recon, ng, precision, name):
    solver, orc_solver = SOLVERS[name][:2]
    n = 20
    U, N, g = _random_periodic_box(n, ng, seed=9)
    h = 1.0 / n
    dt = 0.3 / orc.calc_dti(U, N, N, N, ng, h, h, h, g)
    say(name, "vl_step recon", recon, "precision", precision, "N", N)
    Uo, Uc = U.copy(), U.copy()
    with orc.riemann(orc_solver):
        orc.lib().orc_vl3d(orc.ptr(Uo), N, N, N, ng, h, h, h, dt, g, 5, recon, 1, 0.0)
    orc.lib().orc_vl3d(orc.ptr(Uc), N, N, N, ng, h, h, h, dt, g, 5, recon, 1, 0.0)
    assert not np.array_equal(Uo, Uc)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, recon, precision, riemann=solver)
    hy.vl_step(Ud, h, h, h, dt, g)
    A = Ud.download()
    hy.close()
    l1, mx = field_l1(A, Uo, 5, N, N, N, ng)
    say("  rel L1", ["%.3e" % x for x in l1], "max|diff| %.3e" % mx)
    assert max(l1) <= TOL[precision], l1
    if precision == chl.PREC_FP64_EXACT:
        assert mx == 0.0, f"exact form differs from the oracle by {mx}"


@pytest.mark.parametrize("name", list(SOLVERS))
@pytest.mark.parametrize("tile", [0, 1])
@pytest.mark.parametrize("precision", [chl.PREC_FP64_EXACT, chl.PREC_FP64, chl.PREC_MIXED])
def test_kelvin_helmholtz(precision, tile, name):
    solver, orc_solver = SOLVERS[name][:2]
    n, ng, g = 24, 4, 5 / 3
    N = n + 2 * ng
    U = chl.ic_kelvin_helmholtz(n, ng, g, seed=5)
    h = 1.0 / n
    Uo = U.copy()
    steps_o, to = orc.run(Uo, N, N, N, ng, 5, h, h, h, g, orc.PPMC, [1] * 6, 0.0, 1e30, 6, solver=orc_solver)
    say(name, "KH 24^3 x6 precision", precision, "tile", tile)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, precision, riemann=solver, tile1=tile, tile2=tile)
    steps, t = hy.run(Ud, h, h, h, g, 0.0, 1e30, 6, use_graph=True)
    hy.close()
    assert steps == steps_o == 6
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


@pytest.mark.parametrize("name", list(SOLVERS))
def test_sod_1d(name):
    solver, orc_solver = SOLVERS[name][:2]
    U, nx, dx = sod_ic()
    ng = 3
    orc.set_boundaries(U, nx, 1, 1, 5, ng, [3, 3, 0, 0, 0, 0])
    Uo = U.copy()
    steps_o, to = orc.run(Uo, nx, 1, 1, ng, 5, dx, dx, dx, 1.4, orc.PLMC, [3, 3, 0, 0, 0, 0], 0.0, 0.2, 100000,
                          solver=orc_solver)
    say("Sod 1D", name)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(nx, 1, 1, ng, chl.RECON_PLMC, chl.PREC_FP64_EXACT, bc=[3, 3, 1, 1, 1, 1], riemann=solver)
    steps, t = hy.run(Ud, dx, dx, dx, 1.4, 0.0, 0.2, 100000, use_graph=False)
    hy.close()
    assert steps == steps_o and t == to
    A = Ud.download().reshape(5, nx)
    np.testing.assert_array_equal(A[:, ng:nx - ng], Uo.reshape(5, nx)[:, ng:nx - ng])


@pytest.mark.parametrize("name", list(SOLVERS))
def test_option_limits(name):
    solver = SOLVERS[name][0]
    # dual energy is built for every solver and reconstruction on tile shapes 0 and 1
    chl.Hydro(24, 24, 24, 4, chl.RECON_PPMC, chl.PREC_FP64, riemann=solver, dual_energy=True, tile1=1,
              tile2=1).close()
    chl.Hydro(24, 24, 24, 4, chl.RECON_PLMP, chl.PREC_FP64, riemann=solver, dual_energy=True).close()
    with pytest.raises(chl.ChollaError, match="tile shapes 0 and 1"):
        chl.Hydro(24, 24, 24, 4, chl.RECON_PLMP, chl.PREC_FP64, riemann=solver, dual_energy=True, tile2=3)
    with pytest.raises(chl.ChollaError, match="tile shapes 0 and 1"):
        chl.Hydro(24, 24, 24, 4, chl.RECON_PPMC, chl.PREC_FP64, riemann=solver, tile2=3)
