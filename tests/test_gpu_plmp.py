"""GPU parity of the PLMP, PPMP and PCM reconstruction options (the reference's -DPLMP, -DPPMP, -DPCM
builds: PLMP_cuda, src/reconstruction/plmp_cuda.cu:20-365, and PCM_Reconstruction_3D/1D,
src/reconstruction/pcm_cuda.cu, in the second VL stage; chl_hydro_config.reconstruction =
CHL_RECON_PLMP / CHL_RECON_PCM) against the oracle (tests/test_oracle_plmp.py pins PLMP).

Bars, as for PLMC and PPMC:
  * PolicyExact: single VL steps, graph-captured KH runs in both tile shapes, PLMP with
    HLL, and 1D Sod are bit-identical to the oracle; PolicyFast within per-field relative
    L1 1e-12, the mixed form within strict L1 1e-5 (the KH box's m_y, m_z: the measured miss);
  * the options are refused with dual energy and for tile shapes the build leaves out.
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
NG = {chl.RECON_PLMP: 3, chl.RECON_PCM: 2, chl.RECON_PPMP: 4}


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    if chl.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need a real MI355X")


@pytest.mark.parametrize("precision", [chl.PREC_FP64_EXACT, chl.PREC_FP64, chl.PREC_MIXED])
@pytest.mark.parametrize("recon", [chl.RECON_PLMP, chl.RECON_PCM, chl.RECON_PPMP])
def test_single_vl_step(recon, precision):
    n, ng = 20, NG[recon]
    U, N, g = _random_periodic_box(n, ng, seed=13)
    h = 1.0 / n
    dt = 0.3 / orc.calc_dti(U, N, N, N, ng, h, h, h, g)
    say("vl_step recon", recon, "precision", precision, "N", N)
    Uo = U.copy()
    orc.lib().orc_vl3d(orc.ptr(Uo), N, N, N, ng, h, h, h, dt, g, 5, recon, 1, 0.0)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, recon, precision)
    hy.vl_step(Ud, h, h, h, dt, g)
    A = Ud.download()
    hy.close()
    l1, mx = field_l1(A, Uo, 5, N, N, N, ng)
    say("  rel L1", ["%.3e" % x for x in l1], "max|diff| %.3e" % mx)
    assert max(l1) <= TOL[precision], l1
    if precision == chl.PREC_FP64_EXACT:
        assert mx == 0.0, f"exact form differs from the oracle by {mx}"


@pytest.mark.parametrize("tile", [0, 1])
@pytest.mark.parametrize("precision", [chl.PREC_FP64_EXACT, chl.PREC_FP64, chl.PREC_MIXED])
@pytest.mark.parametrize("recon", [chl.RECON_PLMP, chl.RECON_PCM, chl.RECON_PPMP])
def test_kelvin_helmholtz(recon, precision, tile):
    n, ng, g = 24, NG[recon], 5 / 3
    N = n + 2 * ng
    U = chl.ic_kelvin_helmholtz(n, ng, g, seed=3)
    h = 1.0 / n
    Uo = U.copy()
    steps_o, to = orc.run(Uo, N, N, N, ng, 5, h, h, h, g, recon, [1] * 6, 0.0, 1e30, 6)
    say("KH 24^3 x6 recon", recon, "precision", precision, "tile", tile)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, recon, precision, tile1=tile, tile2=tile)
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


def test_plmp_with_hll_exact():
    n, ng, g = 16, 3, 5 / 3
    U, N, _ = _random_periodic_box(n, ng, seed=17)
    h = 1.0 / n
    Uo = U.copy()
    orc.run(Uo, N, N, N, ng, 5, h, h, h, g, orc.PLMP, [1] * 6, 0.0, 1e30, 4, solver=orc.HLL)
    say("PLMP + HLL exact")
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, chl.RECON_PLMP, chl.PREC_FP64_EXACT, riemann=chl.RIEMANN_HLL)
    hy.run(Ud, h, h, h, g, 0.0, 1e30, 4, use_graph=False)
    hy.close()
    np.testing.assert_array_equal(Ud.download(), Uo)


@pytest.mark.parametrize("recon", [chl.RECON_PLMP, chl.RECON_PCM, chl.RECON_PLMC, chl.RECON_PPMC, chl.RECON_PPMP])
def test_sod_1d(recon):
    ng = 4 if recon in (chl.RECON_PPMC, chl.RECON_PPMP) else 3
    U, nx, dx = sod_ic(ng=ng)
    orc.set_boundaries(U, nx, 1, 1, 5, ng, [3, 3, 0, 0, 0, 0])
    Uo = U.copy()
    steps_o, to = orc.run(Uo, nx, 1, 1, ng, 5, dx, dx, dx, 1.4, recon, [3, 3, 0, 0, 0, 0], 0.0, 0.2, 100000)
    say("Sod 1D recon", recon)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(nx, 1, 1, ng, recon, chl.PREC_FP64_EXACT, bc=[3, 3, 1, 1, 1, 1])
    steps, t = hy.run(Ud, dx, dx, dx, 1.4, 0.0, 0.2, 100000, use_graph=False)
    hy.close()
    assert steps == steps_o and t == to
    A = Ud.download().reshape(5, nx)
    np.testing.assert_array_equal(A[:, ng:nx - ng], Uo.reshape(5, nx)[:, ng:nx - ng])


def test_option_limits():
    with pytest.raises(chl.ChollaError, match="tile shapes 0 and 1"):
        chl.Hydro(24, 24, 24, 3, chl.RECON_PLMP, chl.PREC_FP64, tile2=3)
    with pytest.raises(chl.ChollaError, match="n_ghost"):
        chl.Hydro(24, 24, 24, 1, chl.RECON_PCM, chl.PREC_FP64)
