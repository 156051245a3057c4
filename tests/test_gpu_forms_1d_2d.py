"""The 1D and 2D integrators (VL_Algorithm_1D_CUDA, VL_1D_cuda.cu:30-146; VL_Algorithm_2D_CUDA,
VL_2D_cuda.cu:24-256) in the 3D integrator's other two arithmetic forms, against the oracle's
FP64 reference order on the same inputs:

* CHL_PREC_FP64 (reciprocal-multiply and FMA arithmetic, FP64 throughout): strict per-field
  relative L1 at the FP64 form's bar, 1e-12;
* CHL_PREC_MIXED (reconstruction and Riemann fluxes in FP32 with the split pressure, the state,
  its updates, the primitive recovery and dt in FP64): strict per-field L1 at the mixed bar,
  1e-5, except the fields each case lists with its measured miss (bounds a few times the
  measured values).

Cases: the Sod tube of the 1D configuration (PLMC + HLLC to t = 0.2) with every reconstruction
and every solver, dual energy with every reconstruction; the 2D Kelvin-Helmholtz box of the 2D parity
tests (32^2, 10 steps) with every reconstruction and solver, dual energy with PCM, PLMC and PPMC.
The exact form of both integrators is bit-identical to the oracle (test_gpu_parity.py,
test_gpu_2d.py, test_gpu_de.py)."""
import sys

import numpy as np
import pytest

import cholla_amd as chl
from helpers import exact_riemann_cell_average, rel_l1, sod_ic
from oracle import oracle as orc
from test_oracle_2d import kh_2d

pytestmark = pytest.mark.gpu
G2 = 5.0 / 3.0
RECON = {orc.PLMC: chl.RECON_PLMC, orc.PPMC: chl.RECON_PPMC, orc.PLMP: chl.RECON_PLMP, orc.PCM: chl.RECON_PCM,
         orc.PPMP: chl.RECON_PPMP}
SOLVER = {orc.HLLC: chl.RIEMANN_HLLC, orc.HLL: chl.RIEMANN_HLL, orc.ROE: chl.RIEMANN_ROE, orc.EXACT: chl.RIEMANN_EXACT}
NG1 = {orc.PLMC: 3, orc.PPMC: 4, orc.PLMP: 3, orc.PCM: 3, orc.PPMP: 4}
NG2 = {orc.PLMC: 3, orc.PPMC: 4, orc.PLMP: 3, orc.PCM: 2, orc.PPMP: 4}
NAMES = ["density", "momentum_x", "momentum_y", "momentum_z", "Energy", "GasEnergy"]
FP64_BAR, MIXED_BAR = 1e-12, 1e-5


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    if chl.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need a real MI355X")


def field_l1(A, B, nf, shape, ng):
    a = A.reshape((nf,) + shape)
    b = B.reshape((nf,) + shape)
    sl = tuple(slice(ng, s - ng) if s > 1 else slice(None) for s in shape)
    return [rel_l1(a[(f,) + sl], b[(f,) + sl]) for f in range(nf)]


def check(l1, precision, miss=None):
    miss = miss or {}
    bar = FP64_BAR if precision == chl.PREC_FP64 else MIXED_BAR
    bad = [(NAMES[f], v, miss.get(NAMES[f], bar)) for f, v in enumerate(l1) if not v <= miss.get(NAMES[f], bar)]
    assert not bad, f"L1 over its bound: {bad}"


def sod_run(recon, solver, precision, de=False):
    ng = NG1[recon]
    U, nx, dx = sod_ic(ng=ng)
    nf = 5
    if de:
        U5 = U.reshape(5, nx)
        ge = U5[4] - 0.5 * (U5[1] ** 2) / np.where(U5[0] > 0, U5[0], 1.0)
        U = np.concatenate([U5, ge[None]]).reshape(-1).copy()
        nf = 6
    orc.set_boundaries(U, nx, 1, 1, nf, ng, [3, 3, 0, 0, 0, 0])
    Uo = U.copy()
    s_o, t_o = orc.run(Uo, nx, 1, 1, ng, nf, dx, dx, dx, 1.4, recon, [3, 3, 0, 0, 0, 0], 0.0, 0.2, 100000,
                       dual_energy_on=de, solver=solver)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(nx, 1, 1, ng, RECON[recon], precision, bc=[3, 3, 1, 1, 1, 1], riemann=SOLVER[solver],
                   dual_energy=de)
    s, t = hy.run(Ud, dx, dx, dx, 1.4, 0.0, 0.2, 100000, use_graph=True)
    hy.close()
    assert t == pytest.approx(t_o, rel=1e-12) and t == pytest.approx(0.2, rel=1e-12)
    A = Ud.download()
    return field_l1(A, Uo, nf, (nx,), ng), A.reshape(nf, nx), ng, dx


# Measured (MI355X): the mixed form meets 1e-5 on every field of every case here (Sod: 2e-8
# to 6.4e-7; m_y and m_z stay exactly 0 and compare 0 with 0).  The fast FP64 form meets 1e-12
# except PPMP on the Sod tube, whose m_x reads 1.5e-12: PPMP's discrete steepening and
# flattening switches take the other branch in a few cells on a last-bit difference.
SOD_MIXED_MISS = {}
SOD_FP64_MISS = {orc.PPMP: {"momentum_x": 5e-12}}


@pytest.mark.parametrize("precision", [chl.PREC_FP64, chl.PREC_MIXED])
@pytest.mark.parametrize("recon", [orc.PLMC, orc.PPMC, orc.PLMP, orc.PCM, orc.PPMP])
def test_sod_forms_reconstructions(recon, precision):
    say("Sod 1D form", precision, "recon", recon)
    l1, A, ng, dx = sod_run(recon, orc.HLLC, precision)
    say("  L1", ["%.2e" % v for v in l1])
    check(l1, precision, SOD_MIXED_MISS if precision == chl.PREC_MIXED else SOD_FP64_MISS.get(recon))
    if recon == orc.PLMC:  # the exact Riemann solution still holds the form to the scheme's error
        ex = exact_riemann_cell_average(1.0, 0.0, 1.0, 0.1, 0.0, 0.1, 1.4, np.arange(101) * dx, 0.5, 0.2)
        assert np.abs(A[0, ng:-ng] - ex[0]).mean() < 6e-3


@pytest.mark.parametrize("precision", [chl.PREC_FP64, chl.PREC_MIXED])
@pytest.mark.parametrize("solver", [orc.HLL, orc.ROE, orc.EXACT])
def test_sod_forms_solvers(solver, precision):
    say("Sod 1D form", precision, "solver", solver)
    l1, _, _, _ = sod_run(orc.PLMC, solver, precision)
    say("  L1", ["%.2e" % v for v in l1])
    check(l1, precision, SOD_MIXED_MISS if precision == chl.PREC_MIXED else None)


@pytest.mark.parametrize("precision", [chl.PREC_FP64, chl.PREC_MIXED])
@pytest.mark.parametrize("recon", [orc.PLMC, orc.PPMC, orc.PLMP, orc.PPMP, orc.PCM])
def test_sod_forms_dual_energy(recon, precision):
    say("Sod 1D DE form", precision, "recon", recon)
    l1, _, _, _ = sod_run(recon, orc.HLLC, precision, de=True)
    say("  L1", ["%.2e" % v for v in l1])
    check(l1, precision, SOD_MIXED_MISS if precision == chl.PREC_MIXED else SOD_FP64_MISS.get(recon))


def kh_run(recon, solver, precision, de=False, steps=10, n=32):
    ng = NG2[recon]
    U, N = kh_2d(n, ng, seed=6)
    nf = 5
    h = 1.0 / n
    if de:
        U5 = U.reshape(5, N * N)
        ge = U5[4] - 0.5 * (U5[1] ** 2 + U5[2] ** 2 + U5[3] ** 2) / U5[0]
        U = np.concatenate([U5, ge[None]]).reshape(-1).copy()
        nf = 6
    Uo = U.copy()
    s_o, t_o = orc.run(Uo, N, N, 1, ng, nf, h, h, h, G2, recon, [1, 1, 1, 1, 0, 0], 0.0, 1e30, steps,
                       dual_energy_on=de, solver=solver)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, 1, ng, RECON[recon], precision, bc=[1, 1, 1, 1, 1, 1], riemann=SOLVER[solver],
                   dual_energy=de)
    s, t = hy.run(Ud, h, h, h, G2, 0.0, 1e30, steps, use_graph=True)
    hy.close()
    assert s == s_o == steps
    return field_l1(Ud.download(), Uo, nf, (N, N), ng)


# Measured: the mixed form's largest field on the 2D KH box is m_y (the 0.01 perturbation),
# 1.2e-6 to 5.9e-6; the others 1e-8 to 1e-7.  The fast FP64 form: 1e-16 to 2.5e-14.
KH2_MIXED_MISS = {}


@pytest.mark.parametrize("precision", [chl.PREC_FP64, chl.PREC_MIXED])
@pytest.mark.parametrize("recon", [orc.PLMC, orc.PPMC, orc.PLMP, orc.PCM, orc.PPMP])
def test_kh_2d_forms_reconstructions(recon, precision):
    say("2D KH form", precision, "recon", recon)
    l1 = kh_run(recon, orc.HLLC, precision)
    say("  L1", ["%.2e" % v for v in l1])
    check(l1, precision, KH2_MIXED_MISS if precision == chl.PREC_MIXED else None)


@pytest.mark.parametrize("precision", [chl.PREC_FP64, chl.PREC_MIXED])
@pytest.mark.parametrize("solver", [orc.HLL, orc.ROE, orc.EXACT])
def test_kh_2d_forms_solvers(solver, precision):
    say("2D KH form", precision, "solver", solver)
    l1 = kh_run(orc.PLMC, solver, precision)
    say("  L1", ["%.2e" % v for v in l1])
    check(l1, precision, KH2_MIXED_MISS if precision == chl.PREC_MIXED else None)


@pytest.mark.parametrize("precision", [chl.PREC_FP64, chl.PREC_MIXED])
@pytest.mark.parametrize("recon", [orc.PCM, orc.PLMC, orc.PPMC, orc.PLMP, orc.PPMP])
def test_kh_2d_forms_dual_energy(recon, precision):
    say("2D KH DE form", precision, "recon", recon)
    l1 = kh_run(recon, orc.HLLC, precision, de=True)
    say("  L1", ["%.2e" % v for v in l1])
    check(l1, precision, KH2_MIXED_MISS if precision == chl.PREC_MIXED else None)
