"""GPU parity of the 2D integrator (VL_Algorithm_2D_CUDA, src/integrators/VL_2D_cuda.cu; 2D
boxes: nz == 1 < ny) against the oracle's 2D driver (tests/test_oracle_2d.py pins it):
bit-identical in the FP64 reference-order form for every reconstruction and Riemann solver,
eager and graph-captured, on square and non-square cells, with periodic and transmissive
boundaries."""
import sys

import numpy as np
import pytest

import cholla_amd as chl
from oracle import oracle as orc
from test_oracle_2d import NG, kh_2d

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0
RECON = {orc.PLMC: chl.RECON_PLMC, orc.PPMC: chl.RECON_PPMC, orc.PLMP: chl.RECON_PLMP, orc.PCM: chl.RECON_PCM,
         orc.PPMP: chl.RECON_PPMP}
SOLVER = {orc.HLLC: chl.RIEMANN_HLLC, orc.HLL: chl.RIEMANN_HLL, orc.ROE: chl.RIEMANN_ROE, orc.EXACT: chl.RIEMANN_EXACT}


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    if chl.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need a real MI355X")


def run_both(U, N, ng, recon, solver, h, hy, steps, flags, use_graph):
    Uo = U.copy()
    s_o, t_o = orc.run(Uo, N, N, 1, ng, 5, h, hy, h, G, recon, flags[:4] + [0, 0], 0.0, 1e30, steps, solver=solver)
    Ud = chl.DeviceArray.from_host(U)
    hd = chl.Hydro(N, N, 1, ng, RECON[recon], chl.PREC_FP64_EXACT, bc=flags[:4] + [1, 1], riemann=SOLVER[solver])
    s, t = hd.run(Ud, h, hy, h, G, 0.0, 1e30, steps, use_graph=use_graph)
    hd.close()
    assert s == s_o == steps and t == t_o
    return Ud.download(), Uo


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("aspect", [1.0, 0.5])
@pytest.mark.parametrize("recon", [orc.PLMC, orc.PPMC, orc.PLMP, orc.PCM, orc.PPMP])
def test_kh_2d_bitwise(recon, aspect, use_graph):
    ng, n = NG[recon], 32
    U, N = kh_2d(n, ng, seed=6)
    say("2D KH recon", recon, "aspect", aspect, "graph", use_graph)
    A, B = run_both(U, N, ng, recon, orc.HLLC, 1.0 / n, aspect / n, 10, [1, 1, 1, 1], use_graph)
    np.testing.assert_array_equal(A, B)


@pytest.mark.parametrize("solver", [orc.HLL, orc.ROE, orc.EXACT])
def test_kh_2d_solvers_bitwise(solver):
    ng, n = 3, 24
    U, N = kh_2d(n, ng, seed=8)
    say("2D KH solver", solver)
    A, B = run_both(U, N, ng, orc.PLMC, solver, 1.0 / n, 1.0 / n, 6, [1, 1, 1, 1], False)
    np.testing.assert_array_equal(A, B)


def test_kh_2d_transmissive_reflective_bitwise():
    ng, n = 4, 24
    U, N = kh_2d(n, ng, seed=9)
    orc.set_boundaries(U, N, N, 1, 5, ng, [3, 3, 2, 2, 0, 0])
    say("2D KH flags 3,3,2,2")
    A, B = run_both(U, N, ng, orc.PPMC, orc.HLLC, 1.0 / n, 1.0 / n, 8, [3, 3, 2, 2], True)
    np.testing.assert_array_equal(A, B)


def test_2d_limits():
    with pytest.raises(chl.ChollaError, match="2D boxes need ny"):
        chl.Hydro(40, 8, 1, 4, chl.RECON_PPMC, chl.PREC_MIXED)


@pytest.mark.parametrize("solver", [orc.HLL, orc.ROE, orc.EXACT])
def test_dual_energy_2d_solvers_bitwise(solver):
    """2D DE with the HLL, Roe and exact solvers: bit-identical to the oracle's loop."""
    from test_oracle_de_2d import warm_2d
    ng = NG[orc.PPMC]
    U, N = warm_2d(16, ng)
    h = 1.0 / 32
    Uo = U.copy()
    s_o, t_o = orc.run(Uo, N, N, 1, ng, 6, h, h, h, G, orc.PPMC, [1, 1, 1, 1, 0, 0], 0.0, 1e30, 10,
                       dual_energy_on=True, solver=solver)
    say("2D DE solver", solver)
    Ud = chl.DeviceArray.from_host(U)
    hd = chl.Hydro(N, N, 1, ng, chl.RECON_PPMC, chl.PREC_FP64_EXACT, bc=[1] * 6, dual_energy=True,
                   riemann=SOLVER[solver])
    s, t = hd.run(Ud, h, h, h, G, 0.0, 1e30, 10)
    hd.close()
    assert s == s_o == 10 and t == t_o
    np.testing.assert_array_equal(Ud.download(), Uo)


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("recon", [orc.PLMC, orc.PPMC, orc.PCM, orc.PLMP, orc.PPMP])
@pytest.mark.parametrize("case", ["warm", "cold_x", "cold_y"])
def test_dual_energy_2d_bitwise(case, recon, use_graph):
    """2D DE (VL_2D_cuda.cu:140-152, 203-252): a warm flow and cold compressions along each axis."""
    from test_oracle_de_2d import warm_2d
    from test_oracle_de_1d import state_1d
    ng = NG[recon]
    if case == "warm":
        U, N = warm_2d(16, ng)
        nx = ny = N
    else:
        n = 32
        x = (np.arange(n) + 0.5) / n
        U1, nx = state_1d(np.ones(n), 3.0 + 0.5 * np.sin(2 * np.pi * x), 0.2 * np.cos(2 * np.pi * x), 0 * x,
                          1e-4 * np.ones(n), ng)
        ny = 3 * ng
        U = np.repeat(U1.reshape(6, 1, nx), ny, axis=1)
        if case == "cold_y":
            U = U.transpose(0, 2, 1)[[0, 2, 1, 3, 4, 5]]
            nx, ny = ny, nx
        U = U.reshape(-1).copy()
        orc.set_boundaries(U, nx, ny, 1, 6, ng, [1, 1, 1, 1, 0, 0])
    h = 1.0 / 32
    Uo = U.copy()
    s_o, t_o = orc.run(Uo, nx, ny, 1, ng, 6, h, h, h, G, recon, [1, 1, 1, 1, 0, 0], 0.0, 1e30, 10,
                       dual_energy_on=True)
    say("2D DE", case, "recon", recon, "graph", use_graph)
    Ud = chl.DeviceArray.from_host(U)
    hd = chl.Hydro(nx, ny, 1, ng, RECON[recon], chl.PREC_FP64_EXACT, bc=[1] * 6, dual_energy=True)
    s, t = hd.run(Ud, h, h, h, G, 0.0, 1e30, 10, use_graph=use_graph)
    hd.close()
    assert s == s_o == 10 and t == t_o
    np.testing.assert_array_equal(Ud.download(), Uo)
