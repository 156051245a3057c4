"""GPU parity of the static-gravity option (STATIC_GRAV; chl_hydro_set_static_gravity) against
the oracle (tests/test_oracle_gravity.py pins it): 3D in the fused stage 2 (every form), 2D
and 1D in their integrators.  The fields use log, pow, atan2, cos and sin, whose last bit
differs between the host's libm and the device's math library, so the bar is a per-field
relative L1 of 1e-12 in the FP64 forms (1e-5 mixed); everything else in the step is
bit-identical.  Offsets and bounds are non-trivial so the cell positions are checked too."""
import sys

import numpy as np
import pytest

import cholla_amd as chl
from oracle import oracle as orc
from helpers import check_mixed, strict_l1
from test_gpu_parity import TOL

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    if chl.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need a real MI355X")


def disk_box(shape, ng, seed=3, p0=1e-3, vamp=1e-3):
    """A gas at pressure ~p0 moving at ~vamp (the MW fields are in kpc / kyr)."""
    rng = np.random.default_rng(seed)
    N = tuple(n + 2 * ng for n in shape)
    d = 1.0 + 0.1 * rng.uniform(-1, 1, N)
    v = vamp * rng.uniform(-1, 1, (3,) + N)
    p = p0 * (1.0 + 0.1 * rng.uniform(-1, 1, N))
    U = np.stack([d, d * v[0], d * v[1], d * v[2], p / (G - 1) + 0.5 * d * (v ** 2).sum(0)]).reshape(-1).copy()
    return U, N


def rel_l1(A, B, nf=5):
    A, B = A.reshape(nf, -1), B.reshape(nf, -1)
    return [np.abs(A[f] - B[f]).sum() / max(np.abs(B[f]).sum(), 1e-300) for f in range(nf)]


def run_3d(U, N, ng, h, precision, grav, off, bnd):
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, precision, tile2=1)
    hy.set_static_gravity(grav, off, bnd)
    s, t = hy.run(Ud, h, h, h, G, 0.0, 1e30, 5, use_graph=True)
    hy.close()
    return Ud.download(), s, t


def oracle_3d(U, N, ng, h, grav, off, bnd):
    Uo = U.copy()
    with orc.static_gravity(grav, off, bnd):
        s, t = orc.run(Uo, N, N, N, ng, 5, h, h, h, G, orc.PPMC, [1] * 6, 0.0, 1e30, 5)
    return Uo, s, t


@pytest.mark.parametrize("precision", [chl.PREC_FP64_EXACT, chl.PREC_FP64, chl.PREC_MIXED])
def test_3d_static_gravity(precision):
    """A Mach ~0.3 flow (velocities ~1e-2 kpc/kyr, sound speed 0.03), which the ~1e-7 kpc/kyr^2
    field changes measurably over five steps."""
    n, ng = 16, 4
    U, (N, _, _) = disk_box((n, n, n), ng, vamp=1e-2)
    orc.set_boundaries(U, N, N, N, 5, ng, [1] * 6)
    h, off, bnd = 0.6, (3, -2, 5), (-7.0, 2.0, -3.0)
    say("3D static gravity precision", precision)
    Uo, s_o, t_o = oracle_3d(U, N, ng, h, 1, off, bnd)
    A, s, t = run_3d(U, N, ng, h, precision, 1, off, bnd)
    assert s == s_o == 5 and abs(t - t_o) <= TOL[precision] * t_o
    l1 = strict_l1(A, Uo, 5, N, N, N, ng)
    say("  strict L1", ["%.2e" % x for x in l1])
    if precision == chl.PREC_MIXED:
        check_mixed(l1)
    else:
        assert max(l1) <= TOL[precision], l1
    Uz, _, _ = oracle_3d(U, N, ng, h, 0, off, bnd)
    assert max(strict_l1(Uz, Uo, 5, N, N, N, ng)) > 1e-6  # the source is live


@pytest.mark.parametrize("precision", [chl.PREC_FP64_EXACT, chl.PREC_FP64, chl.PREC_MIXED])
def test_3d_static_gravity_dual_energy(precision):
    """STATIC_GRAV with DE in 3D: the source on the DE build's update (GasEnergy untouched by
    it, hydro_cuda.cu:301-315) against the oracle's DE loop with the same field."""
    n, ng = 16, 4
    U5, (N, _, _) = disk_box((n, n, n), ng, vamp=1e-2)
    U5 = U5.reshape(5, -1)
    d, ek = U5[0], 0.5 * (U5[1] ** 2 + U5[2] ** 2 + U5[3] ** 2) / U5[0]
    U = np.concatenate([U5, (U5[4] - ek)[None]]).reshape(-1).copy()  # GasEnergy = internal energy
    orc.set_boundaries(U, N, N, N, 6, ng, [1] * 6)
    h, off, bnd = 0.6, (3, -2, 5), (-7.0, 2.0, -3.0)
    say("3D static gravity + DE precision", precision)
    Uo = U.copy()
    with orc.static_gravity(1, off, bnd):
        s_o, t_o = orc.run(Uo, N, N, N, ng, 6, h, h, h, G, orc.PPMC, [1] * 6, 0.0, 1e30, 5, dual_energy_on=True)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, precision, tile2=1, dual_energy=True)
    hy.set_static_gravity(1, off, bnd)
    s, t = hy.run(Ud, h, h, h, G, 0.0, 1e30, 5, use_graph=True)
    hy.close()
    A = Ud.download()
    assert s == s_o == 5 and abs(t - t_o) <= TOL[precision] * t_o
    l1 = strict_l1(A, Uo, 6, N, N, N, ng)
    say("  strict L1", ["%.2e" % x for x in l1])
    if precision == chl.PREC_MIXED:
        check_mixed(l1, names=["density", "momentum_x", "momentum_y", "momentum_z", "Energy", "GasEnergy"])
    else:
        assert max(l1) <= TOL[precision], l1


SOLVERS = {chl.RIEMANN_HLL: orc.HLL, chl.RIEMANN_ROE: orc.ROE, chl.RIEMANN_EXACT: orc.EXACT}
RECONS = {chl.RECON_PCM: (orc.PCM, 2), chl.RECON_PLMP: (orc.PLMP, 3), chl.RECON_PPMP: (orc.PPMP, 4),
          chl.RECON_PLMC: (orc.PLMC, 3), chl.RECON_PPMC: (orc.PPMC, 4)}


@pytest.mark.parametrize("de", [False, True])
@pytest.mark.parametrize("riemann", [chl.RIEMANN_HLL, chl.RIEMANN_ROE, chl.RIEMANN_EXACT])
def test_3d_static_gravity_solvers(riemann, de):
    """STATIC_GRAV with the HLL, Roe and exact solvers (the source is the full update's, the
    solver only the faces'), with and without DE, exact form, PPMC: the oracle's loop with the
    same solver and field, at the static-gravity bar."""
    n, ng = 12, 4
    U, (N, _, _) = disk_box((n, n, n), ng, vamp=1e-2, seed=5)
    nf = 5
    if de:
        U5 = U.reshape(5, -1)
        U = np.concatenate([U5, (U5[4] - 0.5 * (U5[1] ** 2 + U5[2] ** 2 + U5[3] ** 2) / U5[0])[None]])
        U = U.reshape(-1).copy()
        nf = 6
    orc.set_boundaries(U, N, N, N, nf, ng, [1] * 6)
    h, off, bnd = 0.6, (3, -2, 5), (-7.0, 2.0, -3.0)
    say("3D static gravity solver", riemann, "DE", de)
    Uo = U.copy()
    with orc.static_gravity(1, off, bnd):
        s_o, t_o = orc.run(Uo, N, N, N, ng, nf, h, h, h, G, orc.PPMC, [1] * 6, 0.0, 1e30, 5, dual_energy_on=de,
                           solver=SOLVERS[riemann])
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, chl.PREC_FP64_EXACT, riemann=riemann, dual_energy=de)
    hy.set_static_gravity(1, off, bnd)
    s, t = hy.run(Ud, h, h, h, G, 0.0, 1e30, 5, use_graph=True)
    hy.close()
    assert s == s_o == 5 and abs(t - t_o) <= TOL[chl.PREC_FP64_EXACT] * t_o
    l1 = strict_l1(Ud.download(), Uo, nf, N, N, N, ng)
    say("  strict L1", ["%.2e" % x for x in l1])
    assert max(l1) <= TOL[chl.PREC_FP64], l1


@pytest.mark.parametrize("recon", [chl.RECON_PCM, chl.RECON_PLMP, chl.RECON_PPMP, chl.RECON_PLMC])
def test_3d_static_gravity_reconstructions_dual_energy(recon):
    """STATIC_GRAV with DE under the other reconstructions, HLLC, exact form, against the
    oracle's DE loop with the same field."""
    n = 12
    orc_recon, ng = RECONS[recon]
    U5, (N, _, _) = disk_box((n, n, n), ng, vamp=1e-2, seed=6)
    U5 = U5.reshape(5, -1)
    U = np.concatenate([U5, (U5[4] - 0.5 * (U5[1] ** 2 + U5[2] ** 2 + U5[3] ** 2) / U5[0])[None]]).reshape(-1).copy()
    orc.set_boundaries(U, N, N, N, 6, ng, [1] * 6)
    h, off, bnd = 0.6, (3, -2, 5), (-7.0, 2.0, -3.0)
    say("3D static gravity + DE recon", recon)
    Uo = U.copy()
    with orc.static_gravity(1, off, bnd):
        s_o, t_o = orc.run(Uo, N, N, N, ng, 6, h, h, h, G, orc_recon, [1] * 6, 0.0, 1e30, 5, dual_energy_on=True)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, recon, chl.PREC_FP64_EXACT, dual_energy=True)
    hy.set_static_gravity(1, off, bnd)
    s, t = hy.run(Ud, h, h, h, G, 0.0, 1e30, 5, use_graph=True)
    hy.close()
    assert s == s_o == 5 and abs(t - t_o) <= TOL[chl.PREC_FP64_EXACT] * t_o
    l1 = strict_l1(Ud.download(), Uo, 6, N, N, N, ng)
    say("  strict L1", ["%.2e" % x for x in l1])
    assert max(l1) <= TOL[chl.PREC_FP64], l1


@pytest.mark.parametrize("grav", [1, 2, 3, 4])
def test_2d_static_gravity(grav):
    n, ng = 24, 3
    U, (N, _) = disk_box((n, n), ng, seed=grav, p0=1.0)  # a sound speed ~1: dt g stays small
    U = U.reshape(5, N, N)
    U[0] += 1.0  # keep the Keplerian and RT cases well away from the density floor
    U = U.reshape(-1).copy()
    orc.set_boundaries(U, N, N, 1, 5, ng, [1, 1, 1, 1, 0, 0])
    h = {1: 0.05, 2: 0.05, 3: 1.0, 4: 0.6}[grav]
    off, bnd = (1, -3, 0), (-0.6, -0.4, 0.0)
    if grav == 3:  # a patch 1e8 cm from a solar mass: g ~ 0.5
        bnd = (1.0e8, 1.0e8, 0.0)
    Uo = U.copy()
    with orc.static_gravity(grav, off, bnd):
        s_o, t_o = orc.run(Uo, N, N, 1, ng, 5, h, h, h, G, orc.PLMC, [1, 1, 1, 1, 0, 0], 0.0, 1e30, 6)
    say("2D static gravity", grav)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, 1, ng, chl.RECON_PLMC, chl.PREC_FP64_EXACT, bc=[1] * 6)
    hy.set_static_gravity(grav, off, bnd)
    s, t = hy.run(Ud, h, h, h, G, 0.0, 1e30, 6, use_graph=False)
    hy.close()
    assert s == s_o == 6
    l1 = rel_l1(Ud.download(), Uo)
    assert max(l1) <= 1e-12, l1


def test_1d_static_gravity():
    n, ng = 64, 3
    U, (N,) = disk_box((n,), ng, seed=11)
    orc.set_boundaries(U, N, 1, 1, 5, ng, [1, 1, 0, 0, 0, 0])
    h, off, bnd = 0.1, (4, 0, 0), (-3.5, 0.0, 0.0)
    Uo = U.copy()
    with orc.static_gravity(1, off, bnd):
        s_o, t_o = orc.run(Uo, N, 1, 1, ng, 5, h, h, h, G, orc.PLMC, [1, 1, 0, 0, 0, 0], 0.0, 1e30, 8)
    say("1D static gravity")
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, 1, 1, ng, chl.RECON_PLMC, chl.PREC_FP64_EXACT, bc=[1] * 6)
    hy.set_static_gravity(1, off, bnd)
    s, t = hy.run(Ud, h, h, h, G, 0.0, 1e30, 8, use_graph=False)
    hy.close()
    assert s == s_o == 8
    assert max(rel_l1(Ud.download(), Uo)) <= 1e-12


def test_static_gravity_limits():
    hy = chl.Hydro(24, 24, 24, 4, chl.RECON_PPMC, chl.PREC_FP64, tile2=3)  # an explicit tile without gravity
    with pytest.raises(chl.ChollaError, match="tile shapes 0 and 1"):
        hy.set_static_gravity(1)
    with pytest.raises(chl.ChollaError, match="out of range"):
        hy.set_static_gravity(2)
    hy.close()
    # a stage-2 tile left to the library moves to one with a gravity kernel
    hy = chl.Hydro(24, 24, 24, 4, chl.RECON_PPMC, chl.PREC_FP64)
    assert hy.config().tile2 == 3
    hy.set_static_gravity(1)
    assert hy.config().tile2 == 1
    hy.close()
