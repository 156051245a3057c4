"""Pins the oracle's PLMP and PPMP reconstructions under the dual-energy build option
(plmp_cuda.cu:55-68, 116-131, 204-206, 228-229, 309-311; ppmp_cuda.cu:84-90, 163-174, 353-361,
640-683): the stencil pressure from Get_Pressure_From_DE with the reconstruction's own kinetic
energy, and the specific GasEnergy dge / d reconstructed as each primitive is, written back as
d * ge.

The reference holds no DE vectors for these reconstructions.  The checks, each chosen so that a
plausible slip fails it:
  * warm gas (E - E_kin far above eta_1 E): the DE pressure is the total-energy pressure, so the
    five hydro fields of both interface arrays equal the non-DE build's bit for bit (a DE
    pressure taken from the wrong field, or a kinetic energy in another order, fails);
  * a uniform specific GasEnergy reaches every interface as d_iface * ge;
  * a linear specific GasEnergy is reproduced on the interfaces, ge_i +- slope / 2 (PLMP's
    limited slope and PPMP's parabola of a line), times the interface density: a GasEnergy
    limited as a conserved instead of a specific variable, or written to the wrong face, fails;
  * cold gas (E - E_kin below eta_1 E): the interface pressure follows the GasEnergy,
    (gamma - 1) dge, not the total energy;
  * the 1D VL driver with PLMP and PPMP and DE converges on Sod like the non-DE build.
"""
import numpy as np
import pytest

from helpers import exact_riemann_cell_average, sod_ic
from oracle import oracle as orc

G = 1.4
KGE = 5


def state(d, vx, vy, vz, p, ge_spec, g=G):
    """Conserved 6-field state; E from p, dge = d * ge_spec (independent of p)."""
    E = p / (g - 1) + 0.5 * d * (vx * vx + vy * vy + vz * vz)
    return np.concatenate([d, d * vx, d * vy, d * vz, E, d * ge_spec])


def sweep(fn, U, n, de, g=G):
    nf = 6 if de else 5
    QL, QR = np.zeros(nf * n), np.zeros(nf * n)
    with orc.dual_energy(de):
        getattr(orc.lib(), fn)(orc.ptr(U), orc.ptr(QL), orc.ptr(QR), n, 1, 1, g, 0)
    return QL, QR


RANGE = {"orc_plmp": (1, -2), "orc_ppmp": (2, -3)}  # cells whose two interfaces the sweep writes


@pytest.mark.parametrize("fn", ["orc_plmp", "orc_ppmp"])
def test_warm_gas_hydro_fields_match_the_non_de_build(fn):
    rng = np.random.default_rng(3)
    n = 24
    d = 1.0 + 0.3 * rng.random(n)
    vx, vy, vz = (0.2 * rng.standard_normal(n) for _ in range(3))
    p = 1.0 + 0.5 * rng.random(n)
    U6 = state(d, vx, vy, vz, p, p / ((G - 1) * d))
    L6, R6 = sweep(fn, U6, n, True)
    L5, R5 = sweep(fn, U6[:5 * n].copy(), n, False)
    np.testing.assert_array_equal(L6[:5 * n], L5)
    np.testing.assert_array_equal(R6[:5 * n], R5)


@pytest.mark.parametrize("fn", ["orc_plmp", "orc_ppmp"])
def test_uniform_and_linear_gas_energy(fn):
    n = 16
    x = np.arange(n, dtype=float)
    d = 1.0 + 0.04 * x  # linear density too: the interface density multiplies ge
    one = np.ones(n)
    a, b = RANGE[fn]
    for ge, slope in ((0.7 * one, 0.0), (0.5 + 0.03 * x, 0.03)):
        U = state(d, 0.1 * one, 0.0 * one, 0.0 * one, 2.0 * one, ge)
        QL, QR = sweep(fn, U, n, True)
        for i in range(a, n + b):
            dl, dr = QL[i], QR[i - 1]  # left state of face i+1/2 at i; right state of face i-1/2 at i-1
            assert abs(dl - (d[i] + 0.02)) < 1e-12 and abs(dr - (d[i] - 0.02)) < 1e-12
            assert abs(QL[KGE * n + i] / dl - (ge[i] + 0.5 * slope)) < 1e-12
            assert abs(QR[KGE * n + i - 1] / dr - (ge[i] - 0.5 * slope)) < 1e-12


@pytest.mark.parametrize("fn", ["orc_plmp", "orc_ppmp"])
def test_cold_gas_pressure_follows_the_gas_energy(fn):
    n = 16
    one = np.ones(n)
    d, ge = 1.2 * one, 0.05 * one
    # kinetic energy 0.5 * 1.2 * 9 = 5.4 and an internal energy of 1e-6 in E: below eta_1 E
    U = state(d, 3.0 * one, 0.0 * one, 0.0 * one, 1e-6 * (G - 1) * one, ge)
    QL, QR = sweep(fn, U, n, True)
    a, b = RANGE[fn]
    for i in range(a, n + b):
        dl, mx, E = QL[i], QL[n + i], QL[4 * n + i]
        p_iface = (E - 0.5 * mx * mx / dl) * (G - 1)
        assert abs(p_iface - (G - 1) * 1.2 * 0.05) < 1e-12 * 5.4 / (G - 1)  # not the 1e-6 of E
        assert abs(QL[KGE * n + i] - 1.2 * 0.05) < 1e-15


@pytest.mark.parametrize("recon", [orc.PLMP, orc.PPMP])
def test_sod_with_dual_energy_converges(recon):
    ng = 4
    U5, nx, dx = sod_ic(ng=ng)
    U5 = U5.reshape(5, nx)
    U = np.concatenate([U5, U5[4:5]]).reshape(-1).copy()  # at rest: GasEnergy = E
    orc.set_boundaries(U, nx, 1, 1, 6, ng, [3, 3, 0, 0, 0, 0])
    steps, t = orc.run(U, nx, 1, 1, ng, 6, dx, dx, dx, G, recon, [3, 3, 0, 0, 0, 0], 0.0, 0.2, 100000,
                       dual_energy_on=True)
    assert t == pytest.approx(0.2, rel=1e-12) and steps > 10
    ex = exact_riemann_cell_average(1.0, 0.0, 1.0, 0.1, 0.0, 0.1, G, np.arange(101) * dx, 0.5, 0.2)
    A = U.reshape(6, nx)
    assert np.abs(A[0, ng:nx - ng] - ex[0]).mean() < 8e-3
    assert np.all(A[5, ng:nx - ng] > 0)
