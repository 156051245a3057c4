"""Pins the dual-energy (DE build option) branches of the CPU oracle.

The reference holds two known-answer tests for this path: Get_Pressure_From_DE
(src/utils/hydro_utilities_tests.cpp:225-236, both branches of the eta_1 test) and
Load_Data under DE (src/reconstruction/reconstruction_tests.cu:205-248).  The rest
is fixed by what the scheme must do, each chosen so that a plausible slip fails:
  * a uniform supersonic flow is a fixed point bit for bit, with the advected
    internal energy kept (a wrong branch in Select_Internal_Energy or a GasEnergy
    flux / pdV term that does not vanish for uniform flow moves it);
  * for a warm subsonic flow (thermal fraction far above eta_1) the DE build gives
    the non-DE hydro fields to round-off and GasEnergy = E - kinetic after every
    step (a wrong GasEnergy slot, flux or sync breaks one of the two);
  * for a cold, smooth, isentropic supersonic compression the advected internal
    energy keeps p / rho^gamma constant to truncation error, converging at first
    order (the pdV source uses the pressure and velocities at the start of the step,
    hydro_cuda.cu:754-833, so it is first order in time) while the density changes
    by ~70 % (the pdV source with a wrong factor or sign changes the entropy by tens
    of percent);
  * Select / Sync / temperature-floor known answers on hand-built cells.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from test_oracle import nearly_equal

G = 5.0 / 3.0


def idx(i, j, k, nx, ny):
    return i + j * nx + k * nx * ny


# ----------------------------------------------------------------------------- reference vectors
def test_pressure_from_de_reference_vectors():
    """hydro_utilities_tests.cpp:225-236 (TestParams E, U_total, U_advected, gamma 5/3)."""
    E = [20.9342082433e-90, 20.9976906577e10, 20.9487120853e300]
    U_total = [2.389074039e-10, 4.890374019e2, 6.8731436293e100]
    U_adv = [1.3847303413e-10, 1.0756968986e2, 1.0882403847e100]
    fid = [1.5927160260000002e-10, 71.713126573333341, 7.2549358980000001e+99]
    for e, ut, ua, f in zip(E, U_total, U_adv, fid):
        assert nearly_equal(f, orc.pressure_from_de(e, ut, ua, G)), (e, ut, ua, f)


def test_load_data_de_reference_vector():
    """reconstruction_tests.cu:205-248 with DE: iota grid of 6 fields, Energy x 500."""
    nx = ny = nz = 3
    n = nx * ny * nz
    U = np.arange(6 * n, dtype=np.float64)
    U[4 * n:5 * n] *= 5.0e2
    out = np.zeros(5)
    with orc.dual_energy():
        orc.lib().orc_load_data(orc.ptr(U), 1, 1, 1, nx, ny, nz, 0, G, orc.ptr(out))
    fid = [13, 3.0769230769230771, 5.1538461538461542, 7.2307692307692308, 39950.641025641031]
    for f, v in zip(fid, out):
        assert nearly_equal(f, v), (fid, out)


# ----------------------------------------------------------------------------- helpers
def periodic_state(nxr, nyr, nzr, ng, d, vx, vy, vz, p, ge=None):
    """6-field periodic state from cell-centred arrays (z, y, x) of the real cells;
    Energy as Sync_Energies_3D forms it, GasEnergy = p/(gamma-1) unless given."""
    nx, ny, nz = nxr + 2 * ng, nyr + 2 * ng, nzr + 2 * ng
    U = np.zeros((6, nz, ny, nx))
    r = (slice(ng, nz - ng), slice(ng, ny - ng), slice(ng, nx - ng))
    U[(0,) + r] = d
    U[(1,) + r] = d * vx
    U[(2,) + r] = d * vy
    U[(3,) + r] = d * vz
    g_e = p / (G - 1) if ge is None else ge
    U[(5,) + r] = g_e
    dd = U[(0,) + r]
    dinv = 1.0 / dd
    ux, uy, uz = U[(1,) + r] * dinv, U[(2,) + r] * dinv, U[(3,) + r] * dinv
    U[(4,) + r] = 0.5 * dd * (ux * ux + uy * uy + uz * uz) + U[(5,) + r]
    U = U.reshape(-1).copy()
    orc.set_boundaries(U, nx, ny, nz, 6, ng, [1] * 6)
    return U, (nx, ny, nz)


def run_de(U, dims, ng, dx, steps, recon=orc.PPMC, de=True, n_fields=6):
    nx, ny, nz = dims
    return orc.run(U, nx, ny, nz, ng, n_fields, dx, dx, dx, G, recon, [1] * 6, 0.0, 1e30, steps,
                   dual_energy_on=de)


# ----------------------------------------------------------------------------- invariants
@pytest.mark.parametrize("recon", [orc.PLMC, orc.PPMC])
def test_uniform_supersonic_flow_is_a_fixed_point(recon):
    ng = 3 if recon == orc.PLMC else 4
    n = 8
    one = np.ones((n, n, n))
    p = 1e-4 * one  # sound speed 0.013, Mach ~ 300: the advected internal energy is the one used
    U, dims = periodic_state(n, n, n, ng, one, 3.0 * one, -2.0 * one, 1.5 * one, p)
    U0 = U.copy()
    nx, ny, nz = dims
    # U_total = E - kinetic differs from the stored GasEnergy in its last bits
    r = U0.reshape(6, nz, ny, nx)[:, ng:-ng, ng:-ng, ng:-ng]
    ut = r[4] - 0.5 * r[0] * ((r[1] / r[0]) ** 2 + (r[2] / r[0]) ** 2 + (r[3] / r[0]) ** 2)
    assert (np.abs(ut / r[4]) < 1e-3).all() and not np.array_equal(ut, r[5])
    steps, t = run_de(U, dims, ng, 1.0 / n, 6, recon)
    assert steps == 6 and t > 0
    np.testing.assert_array_equal(U, U0)


@pytest.mark.parametrize("recon", [orc.PLMC, orc.PPMC])
def test_warm_flow_matches_non_de_build(recon):
    ng = 3 if recon == orc.PLMC else 4
    n = 12
    rng = np.random.default_rng(5)
    z, y, x = np.meshgrid(*(((np.arange(n) + 0.5) / n),) * 3, indexing="ij")
    d = 1.0 + 0.2 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + 0.02 * rng.uniform(-1, 1, (n, n, n))
    vx = 0.5 * np.sin(2 * np.pi * z) + 0.05 * rng.uniform(-1, 1, (n, n, n))
    vy = 0.4 * np.cos(2 * np.pi * x)
    vz = 0.3 * np.sin(2 * np.pi * y)
    p = 1.0 + 0.1 * rng.uniform(-1, 1, (n, n, n))
    U6, dims = periodic_state(n, n, n, ng, d, vx, vy, vz, p)
    nx, ny, nz = dims
    U5 = U6.reshape(6, -1)[:5].reshape(-1).copy()
    steps = 8
    assert run_de(U6, dims, ng, 1.0 / n, steps, recon)[0] == steps
    assert run_de(U5, dims, ng, 1.0 / n, steps, recon, de=False, n_fields=5)[0] == steps
    a = U6.reshape(6, nz, ny, nx)[:, ng:-ng, ng:-ng, ng:-ng]
    b = U5.reshape(5, nz, ny, nx)[:, ng:-ng, ng:-ng, ng:-ng]
    for f in range(5):
        rel = np.abs(a[f] - b[f]).sum() / np.abs(b[f]).sum()
        assert rel <= 1e-12, (f, rel)
    ke = 0.5 * a[0] * ((a[1] / a[0]) ** 2 + (a[2] / a[0]) ** 2 + (a[3] / a[0]) ** 2)
    assert np.abs(a[5] - (a[4] - ke)).max() <= 1e-13 * np.abs(a[4]).max()


def _entropy_error(n, t_end=0.1):
    """max |s/s0 - 1|, s = (gamma-1) GasEnergy / rho^gamma, of a cold isentropic
    compression vx = 3 + 0.5 sin(2 pi x), rho = 1, p = 1e-4 (Mach > 100, so the
    advected internal energy is selected everywhere), periodic, PPMC."""
    ng, ny = 4, 4
    x = (np.arange(n) + 0.5) / n
    shape = (ny, ny, n)
    d = np.ones(shape)
    vx = np.broadcast_to(3.0 + 0.5 * np.sin(2 * np.pi * x), shape)
    p0 = 1e-4
    U, dims = periodic_state(n, ny, ny, ng, d, vx, 0 * d, 0 * d, p0 * d)
    nx, nyt, nzt = dims
    with orc.dual_energy():
        steps, t = orc.run(U, nx, nyt, nzt, ng, 6, 1.0 / n, 1.0 / ny, 1.0 / ny, G, orc.PPMC, [1] * 6, 0.0, t_end,
                           100000, dual_energy_on=True)
    assert abs(t - t_end) < 1e-12
    a = U.reshape(6, nzt, nyt, nx)[:, ng:-ng, ng:-ng, ng:-ng]
    s = (G - 1) * a[5] / a[0] ** G
    # the flow compressed: the density spread is ~70 %
    assert a[0].max() - a[0].min() > 0.5
    return np.abs(s / p0 - 1).max()


def test_cold_compression_keeps_entropy_first_order():
    e32, e64 = _entropy_error(32), _entropy_error(64)
    assert e64 < 1e-2, e64
    assert e32 / e64 > 1.7, (e32, e64)


# ----------------------------------------------------------------------------- known answers
def test_select_and_sync_known_answers():
    ng, nr = 1, 3
    n = nr + 2 * ng
    U = np.zeros((6, n, n, n))
    d, v = 2.0, 1.0  # kinetic energy density 0.5 * 2 * 3 = 3
    U[0] = d
    U[1] = U[2] = U[3] = d * v
    U[4] = 3.0 + 1.0  # U_total = 1: thermal fraction 0.25 > eta_1
    U[5] = 0.7       # advected value, ignored where U_total is selected
    # cell (1,1,1): thermal fraction 2e-4 < eta_1 (and 1.5e-4 of the neighbours' maximum
    # energy < eta_2) -> the advected 0.7 is kept and E is synced to 3.7
    U[4, 1, 1, 1] = 3.0 / (1 - 2e-4)
    # cell (2,2,2): negative advected value with a cold cell -> floored at TINY_NUMBER
    U[4, 2, 2, 2] = 3.0 * (1 + 1e-5)
    U[5, 2, 2, 2] = -0.5
    flat = U.reshape(-1).copy()
    with orc.dual_energy():
        orc.lib().orc_select_internal_energy_3d(orc.ptr(flat), n, n, n, ng)
        sel = flat.reshape(6, n, n, n)[5].copy()
        orc.lib().orc_sync_energies_3d(orc.ptr(flat), n, n, n, ng)
    out = flat.reshape(6, n, n, n)
    r = (slice(ng, n - ng),) * 3
    ut = U[4] - 0.5 * d * 3 * v * v
    want = ut.copy()
    want[1, 1, 1] = 0.7
    want[2, 2, 2] = 1e-20
    np.testing.assert_array_equal(sel[r], want[r])
    np.testing.assert_array_equal(out[4][r], 0.5 * d * 3 * v * v + want[r])
    # ghost cells are untouched
    assert (out[5][0] == U[5][0]).all() and (out[4][0] == U[4][0]).all()


def test_temperature_floor_de_known_answer():
    ng, nr = 1, 2
    n = nr + 2 * ng
    U = np.zeros((6, n, n, n))
    U[0] = 2.0
    U[4] = 10.0
    U[5] = 0.01  # specific 0.005 < floor 0.1 -> GasEnergy = d * floor = 0.2
    U[5, 1, 1, 2] = 1.0  # specific 0.5 > floor: kept
    flat = U.reshape(-1).copy()
    with orc.dual_energy():
        orc.lib().orc_temperature_floor(orc.ptr(flat), n, n, n, ng, 0.1)
    out = flat.reshape(6, n, n, n)
    r = (slice(ng, n - ng),) * 3
    want = np.full((nr, nr, nr), 0.2)
    want[0, 0, 1] = 1.0
    np.testing.assert_array_equal(out[5][r], want)
    np.testing.assert_array_equal(out[4][r], U[4][r])  # E - Ekin = 10 / 2 = 5 per mass > floor
