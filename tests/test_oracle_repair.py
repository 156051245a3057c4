"""Pins the oracle's floors and repairs of the full update (orc_update_3d).

Reference: Update_Conserved_Variables_3D, src/hydro/hydro_cuda.cu:279-299 (density floor,
Average_Cell_Single_Field), :372-382 (the crash repair of a build without both floors,
Average_Cell_All_Fields, :1179-1268).  The reference runs the averages inside the update
kernel while neighbouring threads update too; the oracle runs them after the sweep, serially
in cell order, so a repaired cell sees its updated neighbours and the repairs of cells with a
lower index.  The checks below drive the update with zero fluxes, so the state after the
sweep is the planted one, and compare with closed forms written out here.
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle as orc
from helpers import rough_box

G = 5.0 / 3.0


def smooth_box(n, ng, seed):
    N = n + 2 * ng
    rng = np.random.default_rng(seed)
    d = 1.0 + 0.5 * rng.uniform(0, 1, (N, N, N))
    v = 0.3 * rng.uniform(-1, 1, (3, N, N, N))
    p = 1.0 + 0.2 * rng.uniform(0, 1, (N, N, N))
    U = np.stack([d, d * v[0], d * v[1], d * v[2], p / (G - 1) + 0.5 * d * (v ** 2).sum(0)])
    return U.reshape(-1).copy(), N


def zero_flux_update(U, N, ng, density_floor_on, density_floor, crash=False):
    F = np.zeros_like(U)
    with orc.crash_repair(crash):
        orc.lib().orc_update_3d(orc.ptr(U), orc.ptr(F), orc.ptr(F), orc.ptr(F), N, N, N, ng, 0.1, 0.1, 0.1, 0.01, 5,
                                int(density_floor_on), density_floor, G)


def six_mean(V, k, j, i):
    # hydro_cuda.cu:1165-1176: left, right, down, up, bottom, top
    return (V[k, j, i - 1] + V[k, j, i + 1] + V[k, j - 1, i] + V[k, j + 1, i] + V[k - 1, j, i] + V[k + 1, j, i]) / 6


def test_negative_density_takes_the_six_neighbour_mean():
    n, ng = 8, 4
    U, N = smooth_box(n, ng, seed=1)
    V = U.reshape(5, N, N, N)
    V[0, 6, 7, 5] = -0.25  # interior cell
    want = six_mean(V[0], 6, 7, 5)
    before = V.copy()
    zero_flux_update(U, N, ng, True, 0.0)
    assert V[0, 6, 7, 5] == want
    # nothing else changes, and the momenta / energy of the repaired cell are left as they were
    mask = np.ones(V.shape, bool)
    mask[0, 6, 7, 5] = False
    np.testing.assert_array_equal(V[mask], before[mask])


def test_repairs_run_in_cell_order():
    """Two neighbouring negative cells: the one with the lower index is averaged first (with
    its neighbour still negative), the second sees the first's repaired density."""
    n, ng = 8, 4
    U, N = smooth_box(n, ng, seed=2)
    V = U.reshape(5, N, N, N)
    V[0, 7, 6, 6] = -0.5
    V[0, 7, 6, 7] = -0.75  # i + 1: a higher index
    W = V[0].copy()
    W[7, 6, 6] = six_mean(W, 7, 6, 6)
    W[7, 6, 7] = six_mean(W, 7, 6, 7)
    assert W[7, 6, 7] != six_mean(V[0], 7, 6, 7)  # the order is visible
    zero_flux_update(U, N, ng, True, 0.0)
    assert V[0, 7, 6, 6] == W[7, 6, 6] and V[0, 7, 6, 7] == W[7, 6, 7]


def test_boundary_cell_reads_its_ghost_neighbour():
    n, ng = 8, 4
    U, N = smooth_box(n, ng, seed=3)
    V = U.reshape(5, N, N, N)
    V[0, ng, ng + 2, ng] = -1.0  # first real cell in x and z: two ghost neighbours
    want = six_mean(V[0], ng, ng + 2, ng)
    zero_flux_update(U, N, ng, True, 0.0)
    assert V[0, ng, ng + 2, ng] == want


def test_positive_density_below_the_floor_is_scaled():
    n, ng = 8, 4
    U, N = smooth_box(n, ng, seed=4)
    V = U.reshape(5, N, N, N)
    V[0, 5, 5, 9] = 0.02
    cell = V[:, 5, 5, 9].copy()
    zero_flux_update(U, N, ng, True, 0.1)
    # hydro_cuda.cu:281-292: density to the floor, every other field times floor / d0
    assert V[0, 5, 5, 9] == 0.1
    for f in range(1, 5):
        assert V[f, 5, 5, 9] == cell[f] * (0.1 / 0.02)
    # without the floor option nothing happens
    U2, _ = smooth_box(n, ng, seed=4)
    V2 = U2.reshape(5, N, N, N)
    V2[0, 5, 5, 9] = -0.02
    zero_flux_update(U2, N, ng, False, 0.1)
    assert V2[0, 5, 5, 9] == -0.02


def all_fields_mean(V, k, j, i):
    """Average_Cell_All_Fields written out (hydro_cuda.cu:1207-1262)."""
    sums = np.zeros(5)
    N_ = 0
    for kk in (k - 1, k, k + 1):
        for jj in (j - 1, j, j + 1):
            for ii in (i - 1, i, i + 1):
                d, mx, my, mz, E = V[:, kk, jj, ii]
                P = (E - (0.5 / d) * (mx * mx + my * my + mz * mz)) * (G - 1.0)
                if d > 0.0 and P > 0.0:
                    sums += [d, mx, my, mz, P / (G - 1.0)]
                    N_ += 1
    d_av, P_av = sums[0] / N_, sums[4] / N_
    vx, vy, vz = sums[1] / sums[0], sums[2] / sums[0], sums[3] / sums[0]
    return np.array([d_av, d_av * vx, d_av * vy, d_av * vz, P_av / (G - 1.0) + 0.5 * d_av * (vx * vx + vy * vy + vz * vz)])


@pytest.mark.parametrize("what", ["negative_energy", "nan_density", "negative_density"])
def test_crash_repair_averages_the_healthy_block(what):
    n, ng = 8, 4
    U, N = smooth_box(n, ng, seed=5)
    V = U.reshape(5, N, N, N)
    k, j, i = 7, 8, 6
    if what == "negative_energy":
        V[4, k, j, i] = -0.3
    elif what == "nan_density":
        V[0, k, j, i] = np.nan
    else:
        V[0, k, j, i] = -0.3
    # a neighbour with a non-positive pressure is left out of the mean
    V[4, k + 1, j, i - 1] = 0.5 * (V[1, k + 1, j, i - 1] ** 2 + V[2, k + 1, j, i - 1] ** 2 +
                                   V[3, k + 1, j, i - 1] ** 2) / V[0, k + 1, j, i - 1] * 0.99
    want = all_fields_mean(V, k, j, i)
    zero_flux_update(U, N, ng, False, 0.0, crash=True)
    np.testing.assert_array_equal(V[:, k, j, i], want)


def test_crash_repair_is_off_with_both_floors():
    n, ng = 8, 4
    U, N = smooth_box(n, ng, seed=6)
    V = U.reshape(5, N, N, N)
    V[4, 6, 6, 6] = -0.3
    zero_flux_update(U, N, ng, True, 0.0, crash=False)
    assert V[4, 6, 6, 6] == -0.3


def test_a_rough_step_triggers_both_repairs():
    """The box the GPU repair tests use: one step at CFL 0.3 leaves non-positive densities
    (density repair) and negative energies (crash repair) without any NaN."""
    n, ng = 16, 4
    U, N = rough_box(n, ng, seed=0)
    h = 1.0 / n
    dt = 0.3 / orc.calc_dti(U, N, N, N, ng, h, h, h, G)
    raw = U.copy()
    orc.lib().orc_vl3d(orc.ptr(raw), N, N, N, ng, h, h, h, dt, G, 5, orc.PPMC, 0, 0.0)
    a = raw.reshape(5, N, N, N)[:, ng:-ng, ng:-ng, ng:-ng]
    neg = np.argwhere(a[0] <= 0)
    assert len(neg) >= 1 and int((a[4] < 0).sum()) >= 5 and np.isfinite(a).all()
    assert any(((p == 0) | (p == n - 1)).any() for p in neg)  # one next to a ghost cell
    fixed = U.copy()
    orc.lib().orc_vl3d(orc.ptr(fixed), N, N, N, ng, h, h, h, dt, G, 5, orc.PPMC, 1, 0.0)
    b = fixed.reshape(5, N, N, N)[:, ng:-ng, ng:-ng, ng:-ng]
    assert (b[0] > 0).all()
    crash = U.copy()
    with orc.crash_repair(True):
        orc.lib().orc_vl3d(orc.ptr(crash), N, N, N, ng, h, h, h, dt, G, 5, orc.PPMC, 0, 0.0)
    c = crash.reshape(5, N, N, N)[:, ng:-ng, ng:-ng, ng:-ng]
    assert (c[0] > 0).all() and (c[4] > 0).all()
