"""Pins of the MHD oracle (oracle/mhd_oracle.cpp) before it is trusted as the checker.

Each oracle function is checked against the reference's own unit-test vectors
(tests/golden/mhd_unit_vectors.json, tests/golden/hlld_vectors.json) or against what
the mathematics fixes:
  orc_update_magnetic      magnetic_update_tests.cu fiducials
  orc_ct_electric_fields   ct_electric_fields_tests.cu fiducials (three upwind cases)
  orc_ct_slope             the 24 _ctSlope fiducials
  orc_mhd_plmc             tMHDPlmcReconstructor fiducials (all three directions)
  orc_hlld                 72 whole-solver vectors x 3 directions (Brio & Wu, Dai & Woodward,
                           Ryu & Jones 4d, Einfeldt, constant, degenerate states)
  utilities                mhd_utilities_tests.cu fiducials
  orc_mhd_pcm_3d / run     a uniform state stays exactly uniform; mass is conserved; a
                           divergence-free face field stays divergence-free to round-off
  orc_mhd_run              the linear-wave system tests (exact step counts, L1 bounds)
"""
import json
import math
import os

import numpy as np
import pytest

from helpers import interior, mhd_linear_wave_ic, vector_potential_b
from oracle import oracle as orc

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "mhd_unit_vectors.json")))
HLLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "hlld_vectors.json")))
L = orc.lib()
P = orc.ptr


def near(a, b, ulps=4, absolute=1e-14):
    a, b = float(a), float(b)
    if a == b or abs(a - b) <= absolute:
        return True
    return abs(a - b) <= ulps * np.spacing(max(abs(a), abs(b)))


def test_magnetic_update_vectors():
    g = GOLD["magnetic_update"]
    n = g["nx"] ** 3
    src = np.arange(8 * n, dtype=float)
    E = np.arange(3 * n, dtype=float) + src[-1] + 1
    dst = np.full(8 * n, g["default"])
    L.orc_update_magnetic(P(src), P(dst), P(E), g["nx"], g["nx"], g["nx"], g["dt"], g["dx"], g["dx"], g["dx"])
    for i, v in enumerate(dst):
        assert near(v, g["expect"].get(str(i), g["default"])), (i, v)


@pytest.mark.parametrize("case", ["positive", "negative", "zero"])
def test_ct_electric_field_vectors(case):
    g = GOLD["ct_electric_fields"]
    n = g["nx"] ** 3
    fx = np.arange(7 * n, dtype=float)
    fy = np.arange(7 * n, dtype=float) + fx[-1] + 1
    fz = np.arange(7 * n, dtype=float) + fy[-1] + 1
    grid = np.arange(8 * n, dtype=float) + fz[-1] + 1
    for f in (fx, fy, fz):
        if case == "negative":
            f[:n] *= -1
        if case == "zero":
            f[:n] = 0
    E = np.full(3 * n, g["default"])
    L.orc_ct_electric_fields(P(fx), P(fy), P(fz), P(grid), P(E), 2, 2, 2, 8)
    want = g["cases"][case]["expect"]
    for i, v in enumerate(E):
        assert near(v, want.get(str(i), g["default"])), (i, v)


def test_ct_slope_vectors():
    g = GOLD["ct_slope"]
    nx = g["nx"]
    n = nx ** 3
    draws = orc.uniform32(g["seed_mt19937"], *g["dist"], 16 * n)
    conserved, flux = draws[:8 * n].copy(), draws[8 * n:].copy()
    for args, want in zip(g["calls"], g["expect"]):
        got = L.orc_ct_slope(P(flux), P(conserved), float(args[0]), *args[1:], 2, 2, 2, nx, nx, n, 8 * n, 8 * n)
        assert near(got, want), (args, got, want)


def test_plmc_mhd_vectors():
    g = GOLD["plmc_mhd"]
    nx = g["nx"]
    n = nx ** 3
    grid = orc.uniform(g["seed"], *g["dist"], 2 + 8 * n)[2:].copy()
    for d in range(3):
        QL, QR = np.zeros(7 * n), np.zeros(7 * n)
        L.orc_mhd_plmc(P(grid), P(QL), P(QR), nx, nx, nx, g["gamma"], d)
        for arr, fid in ((QL, g["left"][d]), (QR, g["right"][d])):
            for i, v in enumerate(arr):
                assert near(v, fid.get(str(i), 0.0)), (d, i, v)


def test_ppmc_mhd_vectors():
    """PPMC_VL with MHD against tALLPpmcVLReconstructor's MHD fiducials (one reconstructed
    cell of a 6^3 random grid, all seven interface fields, left and right, three directions)."""
    g = GOLD["ppmc_mhd"]
    nx = g["nx"]
    n = nx ** 3
    grid = orc.uniform(g["seed"], *g["dist"], 8 * n).copy()
    for d in range(3):
        QL, QR = np.zeros(7 * n), np.zeros(7 * n)
        L.orc_mhd_ppmc(P(grid), P(QL), P(QR), nx, nx, nx, g["gamma"], d)
        for arr, fid in ((QL, g["left"][d]), (QR, g["right"][d])):
            for i, v in enumerate(arr):
                assert near(v, fid.get(str(i), 0.0)), (d, i, v)


def _mhd_state(nx, d, vx, vy, vz, p, b, gamma=5.0 / 3.0):
    """8-field MHD state on nx^3 from per-cell primitive arrays (face fields = the cell's)."""
    n = nx ** 3
    U = np.zeros((8, n))
    bx, by, bz = b
    U[0] = d
    U[1], U[2], U[3] = d * vx, d * vy, d * vz
    U[5], U[6], U[7] = bx, by, bz
    U[4] = p / (gamma - 1) + 0.5 * d * (vx * vx + vy * vy + vz * vz) + 0.5 * (bx * bx + by * by + bz * bz)
    return U.ravel().copy()


@pytest.mark.parametrize("d", [0, 1, 2])
def test_ppmc_mhd_closed_forms(d):
    """PPMC-MHD pinned by what the mathematics fixes: a uniform state reproduces itself at both
    interfaces, and a density profile linear along the sweep (velocity, pressure and field
    uniform: only the entropy wave carries it) gives the linear interpolant d_i +- slope/2,
    which PPM's limiter leaves alone on smooth monotone data."""
    nx, g = 8, 5.0 / 3.0
    n = nx ** 3
    idx = np.indices((nx, nx, nx))  # (z, y, x)
    coord = idx[2 - d].ravel().astype(float)
    for dens, slope in ((np.full(n, 1.3), 0.0), (1.0 + 0.05 * coord, 0.05)):
        U = _mhd_state(nx, dens, 0.2, -0.1, 0.3, 0.9, (0.4, 0.7, -0.5), g)
        QL, QR = np.zeros(7 * n), np.zeros(7 * n)
        L.orc_mhd_ppmc(P(U), P(QL), P(QR), nx, nx, nx, g, d)
        QL, QR = QL.reshape(7, n), QR.reshape(7, n)
        step = (1, nx, nx * nx)[d]
        for c in range(n):
            z, y, x = c // (nx * nx), (c // nx) % nx, c % nx
            if min(x, y, z) < 2 or max(x, y, z) >= nx - 3:
                continue
            assert abs(QL[0, c] - (dens[c] + 0.5 * slope)) < 1e-13, (d, c, QL[0, c])
            assert abs(QR[0, c - step] - (dens[c] - 0.5 * slope)) < 1e-13, (d, c, QR[0, c - step])
            if slope == 0.0:  # every conserved slot of a uniform state comes back
                for f in range(1, 5):
                    assert abs(QL[f, c] - U.reshape(8, n)[f, c]) < 1e-12, (d, c, f)


def hlld_flux(left, right, gamma, d, solver=None):
    """The reference test's Compute_Fluxes (hlld_cuda_tests.cu:58-140): rotate the sweep-frame
    momenta into the global slots of direction d, pass Bx as the face field, map back."""
    def rot(s):
        m = list(s[1:4])
        if d:
            k = 3 - d
            m = m[k:] + m[:k]
        return np.array([s[0]] + m + [s[4], s[6], s[7]])
    QL, QR, bx, F = rot(left), rot(right), np.array([left[5]]), np.zeros(7)
    (solver or (lambda a, b, c, f: L.orc_hlld(P(a), P(b), P(c), P(f), 1, gamma, d)))(QL, QR, bx, F)
    f = list(F)
    f.insert(5, 0.0)
    m = f[1:4]
    f[1:4] = m[d:] + m[:d]
    return f


def hlld_close(a, b):
    # the reference's check: 2.7e-15 absolute or 7 ulp (hlld_cuda_tests.cu:195-205)
    if abs(a - b) <= 2.7e-15:
        return True
    if np.sign(a) != np.sign(b):
        return False
    return abs(int(np.float64(a).view(np.int64)) - int(np.float64(b).view(np.int64))) <= 7


def test_hlld_reference_vectors():
    bad = []
    for c in HLLD["cases"]:
        for d in range(3):
            f = hlld_flux(c["left"], c["right"], c["gamma"], d)
            for q, (a, b) in enumerate(zip(c["flux"], f)):
                if not hlld_close(a, b):
                    bad.append((c["test"], c["states"], d, q, a, b))
    assert len(HLLD["cases"]) >= 70
    assert not bad, bad[:5]


def test_mhd_utilities_vectors():
    g = GOLD["mhd_utilities"]
    gam = g["gamma"]
    for i in range(3):
        bx, by, bz = g["magneticX"][i], g["magneticY"][i], g["magneticZ"][i]
        d, p = g["density"][i], g["pressureGas"][i]
        assert near(L.orc_mhd_total_pressure(p, bx, by, bz), g["total_pressure"][i])
        assert L.orc_mhd_total_pressure(g["total_pressure_negative_multiplier"][i] * p, bx, by, bz) == (
            1e-20 if i else L.orc_mhd_total_pressure(p, bx, by, bz))
        c = g["fast_speed_coef"][i]
        assert near(L.orc_mhd_fast_speed(c * d, c * p, c * bx, c * by, c * bz, gam), g["fast_speed"][i])
        assert near(L.orc_mhd_fast_speed(-c * d, c * p, c * bx, c * by, c * bz, gam),
                    g["fast_speed_negative_density"][i])
        assert near(L.orc_mhd_alfven_speed(bx, d), g["alfven_speed"][i])
        assert near(L.orc_mhd_alfven_speed(bx, -d), g["alfven_speed_negative_density"][i])
        if i >= g["slow_speed_first_index"]:
            c = g["slow_speed_coef"]
            assert near(L.orc_mhd_slow_speed(d * c, p * c, bx * c, by * c, bz * c, gam), g["slow_speed"][i])
            assert near(L.orc_mhd_slow_speed(-d * c, p * c, bx * c, by * c, bz * c, gam),
                        g["slow_speed_negative_density"][i])
        # magnetic energy = E - (E with the field removed): pressure_prim differs by (gamma-1) * B^2/2
        if i == 1:
            e_with = L.orc_mhd_energy_prim(1.0, 1.0, 0, 0, 0, gam, bx, by, bz)
            e_none = L.orc_mhd_energy_prim(1.0, 1.0, 0, 0, 0, gam, 0, 0, 0)
            assert near(e_with - e_none, g["magnetic_energy"][i], ulps=64)
    # cell-centred field: the mean of the two faces
    c = GOLD["cell_centered_b"]
    grid = np.arange(8 * c["n_cells"], dtype=float)
    out = np.zeros(3)
    nx = c["nx"]
    # the reference test uses n_cells = 125 with nx = ny = 3; orc_mhd_cell_centered_b takes nz so that
    # nx*ny*nz = n_cells: 125 / 9 is not whole, so the mean is checked directly
    i = c["xid"]
    idx = i + i * nx + i * nx * nx
    n = c["n_cells"]
    want = [0.5 * (grid[5 * n + idx] + grid[5 * n + idx - 1]), 0.5 * (grid[6 * n + idx] + grid[6 * n + idx - nx]),
            0.5 * (grid[7 * n + idx] + grid[7 * n + idx - nx * nx])]
    assert want == c["expect"]


def test_vector_potential_vector():
    g = GOLD["vector_potential"]
    n = g["n"]
    A = np.arange(3 * n ** 3, dtype=float).reshape(3, n, n, n)
    B = vector_potential_b(A, n, n, n, g["d"], g["d"], g["d"])
    U = np.full((8, n, n, n), -999.0)
    U[5:8][:, 1:, 1:, 1:] = B[:, 1:, 1:, 1:]
    flat = U.reshape(-1)
    for i, v in enumerate(flat):
        assert near(v, g["expect"].get(str(i), -999.0)), (i, v)


# ----------------------------------------------------------------------------- invariants of the full step
def random_mhd_box(n, ng, seed, g=5 / 3):
    """Periodic box with smooth + noisy hydro fields and a face field from a periodic vector potential."""
    N = n + 2 * ng
    rng = np.random.default_rng(seed)
    x = (np.arange(N) - ng + 0.5) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    d = 1.0 + 0.2 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) + 0.02 * rng.uniform(-1, 1, (N, N, N))
    v = 0.5 * rng.uniform(-1, 1, (3, N, N, N))
    p = 1.0 + 0.1 * rng.uniform(-1, 1, (N, N, N))
    A = np.stack([0.05 * np.cos(2 * np.pi * (Y + Z)), 0.05 * np.sin(2 * np.pi * (X + Z)),
                  0.3 / (2 * np.pi) * (np.cos(2 * np.pi * X) + np.sin(2 * np.pi * Y))])
    h = 1.0 / n
    B = vector_potential_b(A, N, N, N, h, h, h)
    U = np.zeros((8, N, N, N))
    U[5:] = B
    U[0], U[1], U[2], U[3] = d, d * v[0], d * v[1], d * v[2]
    # cell-centred field for the energy (cellCenteredMagneticFields)
    bc = np.stack([0.5 * (B[0] + np.roll(B[0], 1, 2)), 0.5 * (B[1] + np.roll(B[1], 1, 1)),
                   0.5 * (B[2] + np.roll(B[2], 1, 0))])
    U[4] = p / (g - 1) + 0.5 * d * (v ** 2).sum(0) + 0.5 * (bc ** 2).sum(0)
    U = U.reshape(-1).copy()
    orc.set_boundaries(U, N, N, N, 8, ng, [1] * 6)
    return U, N, h


def test_uniform_state_is_exactly_steady():
    n, ng, g = 8, 4, 5 / 3
    N = n + 2 * ng
    U = np.zeros((8, N, N, N))
    U[0], U[1], U[2], U[3] = 1.3, 1.3 * 0.2, -1.3 * 0.1, 1.3 * 0.05
    U[5], U[6], U[7] = 0.7, -0.4, 0.2
    U[4] = 2.0 / (g - 1) + 0.5 * 1.3 * (0.2 ** 2 + 0.1 ** 2 + 0.05 ** 2) + 0.5 * (0.49 + 0.16 + 0.04)
    U = U.reshape(-1).copy()
    U0 = U.copy()
    steps, t = orc.mhd_run(U, N, N, N, ng, 1 / n, 1 / n, 1 / n, g, [1] * 6, 0.0, 1e30, 3)
    assert steps == 3 and t > 0
    np.testing.assert_array_equal(U, U0)


def test_divergence_and_mass_after_steps():
    n, ng, g = 16, 4, 5 / 3
    U, N, h = random_mhd_box(n, ng, seed=3)
    div0 = orc.mhd_divergence(U, N, N, N, h, h, h)
    bmax = np.abs(U.reshape(8, N, N, N)[5:]).max()
    assert div0 < 1e-12 * bmax / h
    mass0 = interior(U, 8, N, N, N, ng)[0].sum()
    steps, _ = orc.mhd_run(U, N, N, N, ng, h, h, h, g, [1] * 6, 0.0, 1e30, 5)
    assert steps == 5
    assert np.isfinite(U).all()
    div = orc.mhd_divergence(U, N, N, N, h, h, h)
    assert div < 1e-11 * bmax / h, div
    mass = interior(U, 8, N, N, N, ng)[0].sum()
    assert abs(mass - mass0) <= 1e-13 * mass0


# ----------------------------------------------------------------------------- linear waves (mhd_system_tests.cpp)
WAVES = {
    # wave speed, step count along x, right eigenvector (rho, mx, my, mz, E, Bx, By, Bz), PLMC L1 bound
    "fast_right": (2.0, 214, [x / (2 * math.sqrt(5)) for x in (2, 4, -2, 0, 9, 0, 4, 0)], 6.5e-8),
    "alfven_right": (1.0, 427, [0, 0, 0, -1, 0, 0, 0, 1], 3.0e-8),
}


@pytest.mark.slow
@pytest.mark.parametrize("case", sorted(WAVES))
def test_mhd_linear_wave_system(case):
    """mhd_system_tests.cpp:125-160, 277-311 along x (pitch = yaw = 0, domain 0.5, N = 32, PLMC):
    one wave crossing of the doubled domain; exact step count and the per-field L1 bound."""
    speed, steps_want, eig, allowed = WAVES[case]
    g, ng, N = 5 / 3, 4, 32
    U, (nx, ny, nz), (dx, dy, dz) = mhd_linear_wave_ic(2 * N, N, N, ng, 1.0, 0.5, 0.5, g, 1.0, 0.0, 1 / g,
                                                       (1.0, 1.5, 0.0), 1e-6, eig)
    orc.set_boundaries(U, nx, ny, nz, 8, ng, [1] * 6)
    U0 = interior(U, 8, nz, ny, nx, ng).copy()
    steps, t = orc.mhd_run(U, nx, ny, nz, ng, dx, dy, dz, g, [1] * 6, 0.0, 1.0 / speed, 100000)
    assert steps == steps_want
    U1 = interior(U, 8, nz, ny, nx, ng)
    l1 = [np.abs(U1[f] - U0[f]).mean() for f in range(8)]
    assert max(l1) < allowed and math.sqrt(sum(x * x for x in l1)) < allowed, l1
    assert np.abs(U1 - U0).max() < allowed
