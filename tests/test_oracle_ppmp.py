"""Pins the oracle's PPMP reconstruction (the -DPPMP build in VL: PPMP_cuda,
src/reconstruction/ppmp_cuda.cu:28-755, with STEEPENING and FLATTENING undefined as there).
No reference vectors exist; the checks:
  * uniform states and linear profiles are reproduced exactly on the right faces (for linear
    data every limited slope is the common one and the parabola is the line);
  * every interface lies between the cell and its neighbour on that side, and a local
    extremum is flattened to the cell value;
  * the parabola is monotone inside the cell: (q_R - q_L)(q_i - (q_L + q_R)/2) stays within
    (q_R - q_L)^2 / 6, Colella & Woodward's condition;
  * direction permutation; Sod converges, sharper than PLMP at the contact.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from helpers import exact_riemann_cell_average, sod_ic
from test_oracle_plmp import cons_from_prim, prim_from_cons

G = 1.4


def ppmp_x(prims):
    n = prims[0].size
    U = cons_from_prim(*prims, G)
    QL, QR = np.zeros(5 * n), np.zeros(5 * n)
    orc.lib().orc_ppmp(orc.ptr(U), orc.ptr(QL), orc.ptr(QR), n, 1, 1, G, 0)
    return prim_from_cons(QL, n, G), prim_from_cons(QR, n, G), n


def test_uniform_and_linear():
    n = 14
    x = np.arange(n, dtype=float)
    prims = (1.0 + 0.05 * x, 0.3 - 0.02 * x, 0.1 + 0.01 * x, -0.2 + 0.03 * x, 1.0 + 0.04 * x)
    left, right, _ = ppmp_x(prims)
    slopes = (0.05, -0.02, 0.01, 0.03, 0.04)
    for v in range(5):
        for i in range(2, n - 3):
            assert abs(left[v][i] - (prims[v][i] + 0.5 * slopes[v])) < 1e-12
            assert abs(right[v][i - 1] - (prims[v][i] - 0.5 * slopes[v])) < 1e-12


@pytest.mark.parametrize("seed", range(4))
def test_bounds_extrema_and_monotone_parabola(seed):
    rng = np.random.default_rng(seed)
    n = 60
    d = rng.uniform(0.5, 2, n)
    vx, vy, vz = rng.uniform(-1, 1, (3, n))
    p = rng.uniform(0.5, 2, n)
    left, right, _ = ppmp_x((d, vx, vy, vz, p))
    for q, v in ((d, 0), (vx, 1), (p, 4)):
        for i in range(2, n - 3):
            qR, qL = left[v][i], right[v][i - 1]  # the cell's right and left face values
            assert min(q[i], q[i - 1]) - 1e-12 <= qL <= max(q[i], q[i - 1]) + 1e-12
            assert min(q[i], q[i + 1]) - 1e-12 <= qR <= max(q[i], q[i + 1]) + 1e-12
            if (q[i] - q[i - 1]) * (q[i + 1] - q[i]) <= 0:
                assert abs(qL - q[i]) < 1e-12 and abs(qR - q[i]) < 1e-12
            dq = qR - qL
            assert abs(dq * (q[i] - 0.5 * (qL + qR))) <= dq * dq / 6 + 1e-12


def test_direction_permutation():
    rng = np.random.default_rng(7)
    n = 12
    d, p = rng.uniform(0.5, 2, n), rng.uniform(0.5, 2, n)
    vx, vy, vz = rng.uniform(-1, 1, (3, n))
    Ux = cons_from_prim(d, vx, vy, vz, p, G).reshape(5, n)
    QLx, QRx = np.zeros(5 * n), np.zeros(5 * n)
    orc.lib().orc_ppmp(orc.ptr(Ux.reshape(-1).copy()), orc.ptr(QLx), orc.ptr(QRx), n, 1, 1, G, 0)
    Uc = Ux.reshape(-1).copy()
    orc.lib().orc_ppmp(orc.ptr(Uc), orc.ptr(QLx), orc.ptr(QRx), n, 1, 1, G, 0)
    for dr, shape in ((1, (1, n, 1)), (2, (1, 1, n))):
        order = [0, 1 + dr, 1 + (dr + 1) % 3, 1 + (dr + 2) % 3, 4]
        U = np.zeros((5, n))
        for f, g in enumerate(order):
            U[g] = Ux[f]
        U = U.reshape(-1).copy()
        QL, QR = np.zeros(5 * n), np.zeros(5 * n)
        orc.lib().orc_ppmp(orc.ptr(U), orc.ptr(QL), orc.ptr(QR), shape[0], shape[1], shape[2], G, dr)
        for f, g in enumerate(order):
            np.testing.assert_array_equal(QL[g * n + 2:g * n + n - 3], QLx[f * n + 2:f * n + n - 3])
            np.testing.assert_array_equal(QR[g * n + 1:g * n + n - 4], QRx[f * n + 1:f * n + n - 4])


def sod_l1(n, recon):
    U, nx, dx = sod_ic(n_real=n, ng=4)
    orc.set_boundaries(U, nx, 1, 1, 5, 4, [3, 3, 0, 0, 0, 0])
    steps, t = orc.run(U, nx, 1, 1, 4, 5, dx, dx, dx, G, recon, [3, 3, 0, 0, 0, 0], 0.0, 0.2, 100000)
    assert abs(t - 0.2) < 1e-15
    V = U.reshape(5, nx)[:, 4:nx - 4]
    assert abs(V[0].sum() * dx - 0.55) < 1e-13
    ex = exact_riemann_cell_average(1.0, 0.0, 1.0, 0.1, 0.0, 0.1, G, np.arange(n + 1) * dx, 0.5, 0.2)
    return np.abs(V[0] - ex[0]).mean()


def test_sod_ppmp_converges_and_beats_plmp():
    e100, e200 = sod_l1(100, orc.PPMP), sod_l1(200, orc.PPMP)
    assert e100 < 1e-2 and e200 < 0.7 * e100, (e100, e200)
    assert e100 < sod_l1(100, orc.PLMP)

