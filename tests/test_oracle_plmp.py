"""Pins the oracle's PLMP reconstruction (the -DPLMP build, PLMP_cuda in the VL build,
src/reconstruction/plmp_cuda.cu:20-365) and the VL integrator with it.

The reference holds no PLMP test vectors, so it is pinned by what the reconstruction
must do, each check chosen so that a plausible slip fails it:
  * a uniform state is reproduced on every interface;
  * a linear profile in each primitive variable is reproduced exactly (the limited slope
    equals the common one-sided slope; a dropped 0.5 or a swapped side fails);
  * for monotone data the reference's chain of minima reduces to the harmonic mean
    dL dR / (dL + dR) (the Van Leer limiter), checked against that closed form;
  * a local extremum gets a zero slope, and every interface lies between the cell and
    its neighbours (TVD);
  * interfaces land on the right faces: the right state of face i-1/2 at i-1 in Q_R and
    the left state of face i+1/2 at i in Q_L;
  * a y or z sweep of a rotated state equals the x sweep;
  * Sod with PLMP converges to the exact Riemann solution (3D driver on a slab of 3 real
    cells in y and z, which stays one-dimensional bit for bit, and the 1D driver).
"""
import numpy as np
import pytest

from oracle import oracle as orc
from helpers import exact_riemann_cell_average, sod_ic

G = 1.4


def cons_from_prim(d, vx, vy, vz, p, g=G):
    return np.concatenate([d, d * vx, d * vy, d * vz, p / (g - 1) + 0.5 * d * (vx * vx + vy * vy + vz * vz)])


def prim_from_cons(Q, n, g=G):
    d, mx, my, mz, E = (Q[f * n:(f + 1) * n] for f in range(5))
    d = np.where(d == 0, 1.0, d)  # slots the sweep leaves unset
    vx, vy, vz = mx / d, my / d, mz / d
    return d, vx, vy, vz, (E - 0.5 * d * (vx * vx + vy * vy + vz * vz)) * (g - 1)


def plmp_x(prims, g=G):
    n = prims[0].size
    U = cons_from_prim(*prims, g)
    QL, QR = np.zeros(5 * n), np.zeros(5 * n)
    orc.lib().orc_plmp(orc.ptr(U), orc.ptr(QL), orc.ptr(QR), n, 1, 1, g, 0)
    return QL, QR, n


def test_uniform_and_linear_profiles():
    n = 12
    one = np.ones(n)
    QL, QR, _ = plmp_x((1.3 * one, 0.2 * one, -0.1 * one, 0.4 * one, 2.0 * one))
    U = cons_from_prim(1.3 * one, 0.2 * one, -0.1 * one, 0.4 * one, 2.0 * one)
    for f in range(5):
        np.testing.assert_allclose(QL[f * n + 1:f * n + n - 2], U[f * n + 1:f * n + n - 2], rtol=1e-15)
        np.testing.assert_allclose(QR[f * n:f * n + n - 3], U[f * n:f * n + n - 3], rtol=1e-15)
    x = np.arange(n, dtype=float)
    prims = (1.0 + 0.05 * x, 0.3 - 0.02 * x, 0.1 + 0.01 * x, -0.2 + 0.03 * x, 1.0 + 0.04 * x)
    QL, QR, _ = plmp_x(prims)
    left = prim_from_cons(QL, n)   # left state of face i+1/2, stored at i
    right = prim_from_cons(QR, n)  # right state of face i-1/2, stored at i-1
    slopes = (0.05, -0.02, 0.01, 0.03, 0.04)
    for v in range(5):
        for i in range(1, n - 2):
            assert abs(left[v][i] - (prims[v][i] + 0.5 * slopes[v])) < 1e-12
            assert abs(right[v][i - 1] - (prims[v][i] - 0.5 * slopes[v])) < 1e-12


@pytest.mark.parametrize("seed", range(3))
def test_harmonic_mean_limiter_extrema_and_tvd(seed):
    rng = np.random.default_rng(seed)
    n = 40
    # monotone density (increasing), pressure (decreasing), velocities arbitrary
    d = np.cumsum(rng.uniform(0.01, 0.5, n)) + 0.5
    p = 10.0 - np.cumsum(rng.uniform(0.01, 0.2, n))
    vx, vy, vz = rng.uniform(-1, 1, (3, n))
    QL, QR, _ = plmp_x((d, vx, vy, vz, p))
    left, right = prim_from_cons(QL, n), prim_from_cons(QR, n)
    for q, v in ((d, 0), (p, 4)):
        for i in range(1, n - 2):
            dl, dr = q[i] - q[i - 1], q[i + 1] - q[i]
            hm = dl * dr / (dl + dr)
            assert abs(left[v][i] - (q[i] + hm)) < 1e-12 * abs(q[i])
            assert abs(right[v][i - 1] - (q[i] - hm)) < 1e-12 * abs(q[i])
    # arbitrary data: TVD bounds and flat extrema, on the density and x velocity
    d = rng.uniform(0.5, 2, n)
    QL, QR, _ = plmp_x((d, vx, vy, vz, p))
    left, right = prim_from_cons(QL, n), prim_from_cons(QR, n)
    for q, v in ((d, 0), (vx, 1)):
        for i in range(1, n - 2):
            lo, hi = min(q[i - 1:i + 2]), max(q[i - 1:i + 2])
            for val in (left[v][i], right[v][i - 1]):
                assert lo - 1e-12 <= val <= hi + 1e-12
            if (q[i] - q[i - 1]) * (q[i + 1] - q[i]) <= 0:
                assert abs(left[v][i] - q[i]) < 1e-12 and abs(right[v][i - 1] - q[i]) < 1e-12


def test_direction_permutation():
    rng = np.random.default_rng(7)
    n = 10
    d, p = rng.uniform(0.5, 2, n), rng.uniform(0.5, 2, n)
    vx, vy, vz = rng.uniform(-1, 1, (3, n))
    QLx, QRx, _ = plmp_x((d, vx, vy, vz, p))
    Ux = cons_from_prim(d, vx, vy, vz, p).reshape(5, n)
    for dr, shape in ((1, (1, n, 1)), (2, (1, 1, n))):
        # the sweep momentum goes in the axis's slot; o1, o2, o3 rotate as in DirOrder
        order = [0, 1 + dr, 1 + (dr + 1) % 3, 1 + (dr + 2) % 3, 4]
        U = np.zeros((5, n))
        for f, g in enumerate(order):
            U[g] = Ux[f]
        U = U.reshape(-1).copy()
        QL, QR = np.zeros(5 * n), np.zeros(5 * n)
        orc.lib().orc_plmp(orc.ptr(U), orc.ptr(QL), orc.ptr(QR), shape[0], shape[1], shape[2], G, dr)
        for f, g in enumerate(order):
            np.testing.assert_array_equal(QL[g * n + 1:g * n + n - 2], QLx[f * n + 1:f * n + n - 2])
            np.testing.assert_array_equal(QR[g * n:g * n + n - 3], QRx[f * n:f * n + n - 3])


def sod_l1(n, recon, three_d):
    U, nx, dx = sod_ic(n_real=n)
    ng = 3
    if three_d:
        # a slab ng real cells thick in y and z (the periodic fill needs that many), periodic there
        N = 3 * ng
        V = np.repeat(U.reshape(5, 1, 1, nx), N, axis=1).repeat(N, axis=2).reshape(-1).copy()
        flags = [3, 3, 1, 1, 1, 1]
        orc.set_boundaries(V, nx, N, N, 5, ng, flags)
        steps, t = orc.run(V, nx, N, N, ng, 5, dx, dx, dx, G, recon, flags, 0.0, 0.2, 100000)
        A = V.reshape(5, N, N, nx)[:, ng, ng, ng:nx - ng]
        R = V.reshape(5, N, N, nx)[:, ng:N - ng, ng:N - ng, ng:nx - ng]
        assert np.array_equal(R, np.broadcast_to(A[:, None, None, :], R.shape))  # stays one-dimensional
        assert not R[2:4].any()
    else:
        orc.set_boundaries(U, nx, 1, 1, 5, ng, [3, 3, 0, 0, 0, 0])
        steps, t = orc.run(U, nx, 1, 1, ng, 5, dx, dx, dx, G, recon, [3, 3, 0, 0, 0, 0], 0.0, 0.2, 100000)
        A = U.reshape(5, nx)[:, ng:nx - ng]
    assert abs(t - 0.2) < 1e-15
    if recon != orc.PCM:  # PCM's smeared rarefaction head reaches the outflow boundary
        assert abs(A[0].sum() * dx - 0.55) < 1e-13
    ex = exact_riemann_cell_average(1.0, 0.0, 1.0, 0.1, 0.0, 0.1, G, np.arange(n + 1) * dx, 0.5, 0.2)
    return np.abs(A[0] - ex[0]).mean()


@pytest.mark.parametrize("three_d", [False, True])
def test_sod_plmp_converges(three_d):
    e100, e200 = sod_l1(100, orc.PLMP, three_d), sod_l1(200, orc.PLMP, three_d)
    assert e100 < 1e-2 and e200 < 0.7 * e100, (e100, e200)
    if not three_d:  # a second-order scheme beats first-order PCM on the same grid
        assert e100 < 0.8 * sod_l1(100, orc.PCM, three_d)
