"""Pins the 1D dual-energy branches of the oracle (the DE build of VL_Algorithm_1D_CUDA:
the half-step pdV term of Update_Conserved_Variables_1D_half, hydro_cuda.cu:222-263;
Partial_Update_Advected_Internal_Energy_1D, Select_Internal_Energy_1D and Sync_Energies_1D,
hydro_cuda.cu:663-700, 835-885, 1018-1040; VL_1D_cuda.cu:126-142).  The reference holds no
1D DE vectors; the checks, as for 3D (tests/test_oracle_de.py):
  * a uniform supersonic cold flow is a fixed point, bit for bit, with the advected
    internal energy kept;
  * a warm flow (the total-energy pressure selected everywhere) gives the non-DE hydro
    fields to 1e-12, with GasEnergy = E - kinetic after every step; Sod likewise;
  * a cold smooth isentropic compression keeps (gamma-1) GasEnergy / rho^gamma constant,
    converging at first order (a wrong factor or sign on either pdV term moves it by tens
    of percent).
"""
import numpy as np
import pytest

from oracle import oracle as orc
from helpers import sod_ic

G = 5.0 / 3.0


def state_1d(d, vx, vy, vz, p, ng, ge=None):
    n = d.size
    nx = n + 2 * ng
    U = np.zeros((6, nx))
    r = slice(ng, nx - ng)
    U[0, r], U[1, r], U[2, r], U[3, r] = d, d * vx, d * vy, d * vz
    U[5, r] = p / (G - 1) if ge is None else ge
    dinv = 1.0 / U[0, r]
    ux, uy, uz = U[1, r] * dinv, U[2, r] * dinv, U[3, r] * dinv
    U[4, r] = 0.5 * U[0, r] * (ux * ux + uy * uy + uz * uz) + U[5, r]
    U = U.reshape(-1).copy()
    orc.set_boundaries(U, nx, 1, 1, 6, ng, [1, 1, 0, 0, 0, 0])
    return U, nx


def run(U, nx, ng, dx, steps, recon, de=True, n_fields=6, tout=1e30, flags=(1, 1, 0, 0, 0, 0)):
    return orc.run(U, nx, 1, 1, ng, n_fields, dx, dx, dx, G, recon, list(flags), 0.0, tout, steps,
                   dual_energy_on=de)


@pytest.mark.parametrize("recon,ng", [(orc.PLMC, 3), (orc.PPMC, 4)])
def test_uniform_supersonic_flow_is_a_fixed_point(recon, ng):
    n = 16
    one = np.ones(n)
    U, nx = state_1d(one, 3.0 * one, -2.0 * one, 1.5 * one, 1e-4 * one, ng)
    U0 = U.copy()
    steps, t = run(U, nx, ng, 1.0 / n, 6, recon)
    assert steps == 6 and t > 0
    np.testing.assert_array_equal(U, U0)


@pytest.mark.parametrize("recon,ng", [(orc.PLMC, 3), (orc.PPMC, 4)])
def test_warm_flow_matches_non_de_build(recon, ng):
    n = 48
    rng = np.random.default_rng(9)
    x = (np.arange(n) + 0.5) / n
    d = 1.0 + 0.2 * np.sin(2 * np.pi * x) + 0.02 * rng.uniform(-1, 1, n)
    vx = 0.5 * np.sin(4 * np.pi * x)
    vy, vz = 0.3 * np.cos(2 * np.pi * x), 0.1 * np.ones(n)
    p = 1.0 + 0.1 * rng.uniform(-1, 1, n)
    U6, nx = state_1d(d, vx, vy, vz, p, ng)
    U5 = U6.reshape(6, -1)[:5].reshape(-1).copy()
    for _ in range(6):  # GasEnergy = E - kinetic after every step
        assert run(U6, nx, ng, 1.0 / n, 1, recon)[0] == 1
        a = U6.reshape(6, nx)[:, ng:-ng]
        ke = 0.5 * a[0] * ((a[1] / a[0]) ** 2 + (a[2] / a[0]) ** 2 + (a[3] / a[0]) ** 2)
        assert np.abs(a[5] - (a[4] - ke)).max() <= 1e-13 * np.abs(a[4]).max()
    assert run(U5, nx, ng, 1.0 / n, 6, recon, de=False, n_fields=5)[0] == 6
    a, b = U6.reshape(6, nx)[:, ng:-ng], U5.reshape(5, nx)[:, ng:-ng]
    for f in range(5):
        rel = np.abs(a[f] - b[f]).sum() / np.abs(b[f]).sum()
        assert rel <= 1e-12, (f, rel)


def test_sod_with_dual_energy_matches_plain():
    U5, nx, dx = sod_ic()
    ng = 3
    U6 = np.concatenate([U5, np.zeros(nx)])
    U6[5 * nx:] = U6[4 * nx:5 * nx]  # at rest: GasEnergy = E
    flags = (3, 3, 0, 0, 0, 0)
    orc.set_boundaries(U5, nx, 1, 1, 5, ng, list(flags))
    orc.set_boundaries(U6, nx, 1, 1, 6, ng, list(flags))
    s5 = orc.run(U5, nx, 1, 1, ng, 5, dx, dx, dx, 1.4, orc.PLMC, list(flags), 0.0, 0.2, 100000)
    s6 = orc.run(U6, nx, 1, 1, ng, 6, dx, dx, dx, 1.4, orc.PLMC, list(flags), 0.0, 0.2, 100000, dual_energy_on=True)
    assert s5 == s6
    a, b = U6.reshape(6, nx)[:, ng:-ng], U5.reshape(5, nx)[:, ng:-ng]
    for f in range(5):
        assert np.abs(a[f] - b[f]).sum() <= 1e-12 * np.abs(b[f]).sum() + 1e-300, f


def _entropy_error(n, t_end=0.1):
    ng = 4
    x = (np.arange(n) + 0.5) / n
    d = np.ones(n)
    vx = 3.0 + 0.5 * np.sin(2 * np.pi * x)
    p0 = 1e-4
    U, nx = state_1d(d, vx, 0 * d, 0 * d, p0 * d, ng)
    steps, t = run(U, nx, ng, 1.0 / n, 100000, orc.PPMC, tout=t_end)
    assert abs(t - t_end) < 1e-12
    a = U.reshape(6, nx)[:, ng:-ng]
    assert a[0].max() - a[0].min() > 0.5  # the flow compressed
    s = (G - 1) * a[5] / a[0] ** G
    return np.abs(s / p0 - 1).max()


def test_cold_compression_keeps_entropy_first_order():
    e32, e64 = _entropy_error(32), _entropy_error(64)
    # 3.9e-3 at 64 cells; without the half-step pdV term of the half update the error
    # doubles (7.3e-3), with the full-step term's sign flipped it grows tenfold
    assert e64 < 5e-3, e64
    assert e32 / e64 > 1.7, (e32, e64)
