"""Pins the oracle's HLL solver (the -DHLL build option, Calculate_HLL_Fluxes_CUDA,
src/riemann_solvers/hll_cuda.cu:19-268).

The reference holds no unit-test vectors for HLL, so it is pinned by what the solver
must do, each check chosen so that a plausible slip fails it:
  * consistency: F(U, U) is the physical flux of U, subsonic and supersonic;
  * upwinding: a supersonic left state gives the left physical flux bit for bit (a wrong
    wave-speed sign or a swapped side fails);
  * a closed form: for a subsonic face the flux equals the textbook HLL combination
    (S_R F_L - S_L F_R + S_L S_R (U_R - U_L)) / (S_R - S_L) with the Davis speeds
    S_L = min(u_L - c_L, u_R - c_R), S_R = max(u_L + c_L, u_R + c_R), evaluated here in
    numpy (a Roe-averaged speed, as HLLC uses, or a dropped dissipation term fails);
  * mirror symmetry and direction permutation;
  * Sod with PLMC + HLL converges to the exact Riemann solution, and HLL smears the
    contact more than HLLC on the same grid (HLL has no contact wave).
"""
import numpy as np
import pytest

from oracle import oracle as orc
from helpers import exact_riemann_cell_average, sod_ic

G = 5.0 / 3.0


def _hll_single(left, right, gamma, direction=0):
    QL, QR, F = np.array(left, float), np.array(right, float), np.zeros(5)
    orc.lib().orc_hll(orc.ptr(QL), orc.ptr(QR), orc.ptr(F), 1, 1, 1, gamma, direction)
    return F


def _cons(d, vx, vy, vz, p, g):
    return [d, d * vx, d * vy, d * vz, p / (g - 1) + 0.5 * d * (vx * vx + vy * vy + vz * vz)]


def _phys_flux_x(d, vx, vy, vz, p, g):
    E = p / (g - 1) + 0.5 * d * (vx * vx + vy * vy + vz * vz)
    return np.array([d * vx, d * vx * vx + p, d * vx * vy, d * vx * vz, (E + p) * vx])


@pytest.mark.parametrize("seed", range(4))
def test_hll_consistency_upwinding_closed_form(seed):
    rng = np.random.default_rng(seed)
    for _ in range(20):
        d, p = rng.uniform(0.1, 3), rng.uniform(0.1, 3)
        v = rng.uniform(-4, 4, 3)
        F = _hll_single(_cons(d, *v, p, G), _cons(d, *v, p, G), G)
        assert np.allclose(F, _phys_flux_x(d, *v, p, G), rtol=1e-12, atol=1e-12)
    FL = _phys_flux_x(1.0, 5.0, 0.1, -0.2, 1.0, G)
    F = _hll_single(_cons(1.0, 5.0, 0.1, -0.2, 1.0, G), _cons(0.9, 4.8, 0.0, 0.3, 1.1, G), G)
    np.testing.assert_array_equal(F, FL)
    FR = _phys_flux_x(0.9, -4.8, 0.0, 0.3, 1.1, G)
    F = _hll_single(_cons(1.0, -5.0, 0.1, -0.2, 1.0, G), _cons(0.9, -4.8, 0.0, 0.3, 1.1, G), G)
    np.testing.assert_array_equal(F, FR)
    for _ in range(20):
        a = (rng.uniform(0.1, 3), *rng.uniform(-1, 1, 3), rng.uniform(0.1, 3))
        b = (rng.uniform(0.1, 3), *rng.uniform(-1, 1, 3), rng.uniform(0.1, 3))
        UL, UR = np.array(_cons(*a, G)), np.array(_cons(*b, G))
        cl, cr = np.sqrt(G * a[4] / a[0]), np.sqrt(G * b[4] / b[0])
        SL, SR = min(a[1] - cl, b[1] - cr), max(a[1] + cl, b[1] + cr)
        assert SL < 0 < SR
        want = (SR * _phys_flux_x(*a, G) - SL * _phys_flux_x(*b, G) + SL * SR * (UR - UL)) / (SR - SL)
        np.testing.assert_allclose(_hll_single(UL, UR, G), want, rtol=1e-12, atol=1e-13)
        # mirror: swap sides and flip the normal velocity
        am, bm = (a[0], -a[1], a[2], a[3], a[4]), (b[0], -b[1], b[2], b[3], b[4])
        Fm = _hll_single(_cons(*bm, G), _cons(*am, G), G)
        np.testing.assert_allclose(_hll_single(UL, UR, G) * [-1, 1, -1, -1, -1], Fm, rtol=1e-10, atol=1e-12)
        # the z sweep of a rotated state equals the x sweep, rotated back
        Fz = _hll_single(_cons(a[0], a[2], a[3], a[1], a[4], G), _cons(b[0], b[2], b[3], b[1], b[4], G), G, 2)
        np.testing.assert_allclose([Fz[0], Fz[3], Fz[1], Fz[2], Fz[4]], _hll_single(UL, UR, G), rtol=1e-14,
                                   atol=1e-14)


def sod_l1(n, solver):
    U, nx, dx = sod_ic(n_real=n)
    ng = 3
    orc.set_boundaries(U, nx, 1, 1, 5, ng, [3, 3, 0, 0, 0, 0])
    steps, t = orc.run(U, nx, 1, 1, ng, 5, dx, dx, dx, 1.4, orc.PLMC, [3, 3, 0, 0, 0, 0], 0.0, 0.2, 100000,
                       solver=solver)
    assert abs(t - 0.2) < 1e-15
    edges = np.arange(n + 1) * dx
    ex = exact_riemann_cell_average(1.0, 0.0, 1.0, 0.1, 0.0, 0.1, 1.4, edges, 0.5, 0.2)
    V = U.reshape(5, nx)[:, ng:nx - ng]
    assert abs(V[0].sum() * dx - 0.55) < 1e-13
    return np.abs(V[0] - ex[0]).mean()


def test_sod_hll_converges_and_is_more_diffusive_than_hllc():
    e100, e200 = sod_l1(100, orc.HLL), sod_l1(200, orc.HLL)
    assert e100 < 1e-2 and e200 < 0.7 * e100, (e100, e200)
    assert e100 > sod_l1(100, orc.HLLC)
