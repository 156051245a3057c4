"""Pins the oracle's exact Riemann solver (the -DEXACT build, Calculate_Exact_Fluxes_CUDA,
src/riemann_solvers/exact_cuda.cu:20-334).

The reference holds no test vectors for it.  It is pinned against an independent exact
solver (tests/helpers.py, Toro ch. 4, double precision, converged to round-off): the
Godunov flux of the solution sampled on the interface, for Toro's five test problems
(Table 4.1) shifted through a range of frame velocities so that the interface falls in
each region in turn (data states, both fans, both star states, behind both shocks), with
transverse velocities upwinded by the sampled normal velocity.  The reference evaluates
its rarefaction powers with powf (FP32) and stops Newton at a relative change of 1e-6,
so the bar is a relative 2e-6 (a wrong branch, a swapped side or a sign slip is off by
far more).  Sod with PLMC and the exact solver converges.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from helpers import exact_riemann, exact_riemann_cell_average, sod_ic

G = 1.4
TORO = [  # (rho, u, p) left, right (Toro, Riemann Solvers, Table 4.1)
    ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1)),
    ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4)),
    ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01)),
    ((1.0, 0.0, 0.01), (1.0, 0.0, 100.0)),
    ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950)),
]


def cons(d, vx, vy, vz, p, g=G):
    return [d, d * vx, d * vy, d * vz, p / (g - 1) + 0.5 * d * (vx * vx + vy * vy + vz * vz)]


def orc_flux(a, b, direction=0):
    QL, QR, F = np.array(cons(*a)), np.array(cons(*b)), np.zeros(5)
    orc.lib().orc_exact(orc.ptr(QL), orc.ptr(QR), orc.ptr(F), 1, 1, 1, G, direction)
    return F


def godunov(a, b):
    """Flux of the independent exact solution at x/t = 0."""
    rho, u, p = (x[0] for x in exact_riemann(a[0], a[1], a[4], b[0], b[1], b[4], G, np.array([0.0])))
    vy, vz = (a[2], a[3]) if u >= 0 else (b[2], b[3])
    E = p / (G - 1) + 0.5 * rho * (u * u + vy * vy + vz * vz)
    return np.array([rho * u, rho * u * u + p, rho * u * vy, rho * u * vz, (E + p) * u])


@pytest.mark.parametrize("case", range(len(TORO)))
def test_godunov_flux_matches_independent_solver(case):
    (rl, ul, pl), (rr, ur, pr) = TORO[case]
    cl, cr = np.sqrt(G * pl / rl), np.sqrt(G * pr / rr)
    scale = max(abs(ul), abs(ur), cl, cr)
    for shift in np.linspace(-2.5, 2.5, 41) * scale:  # frame velocities sweeping every wave past x = 0
        a, b = (rl, ul + shift, 0.3, -0.2, pl), (rr, ur + shift, -0.5, 0.1, pr)
        F, want = orc_flux(a, b), godunov(a, b)
        ref = np.abs(want).max()
        np.testing.assert_allclose(F, want, rtol=2e-6, atol=2e-6 * ref, err_msg=f"case {case} shift {shift}")


def test_direction_permutation():
    rng = np.random.default_rng(3)
    for _ in range(30):
        a = (rng.uniform(0.2, 3), *rng.uniform(-2, 2, 3), rng.uniform(0.1, 3))
        b = (rng.uniform(0.2, 3), *rng.uniform(-2, 2, 3), rng.uniform(0.1, 3))
        F = orc_flux(a, b)
        Fy = orc_flux((a[0], a[3], a[1], a[2], a[4]), (b[0], b[3], b[1], b[2], b[4]), direction=1)
        # to round-off: the total energy of the permuted input sums the squares in another order
        np.testing.assert_allclose([Fy[0], Fy[2], Fy[3], Fy[1], Fy[4]], F, rtol=1e-13, atol=1e-13)


def test_sod_exact_solver_converges():
    errs = []
    for n in (100, 200):
        U, nx, dx = sod_ic(n_real=n)
        ng = 3
        orc.set_boundaries(U, nx, 1, 1, 5, ng, [3, 3, 0, 0, 0, 0])
        steps, t = orc.run(U, nx, 1, 1, ng, 5, dx, dx, dx, G, orc.PLMC, [3, 3, 0, 0, 0, 0], 0.0, 0.2, 100000,
                           solver=orc.EXACT)
        assert abs(t - 0.2) < 1e-15
        V = U.reshape(5, nx)[:, ng:nx - ng]
        assert abs(V[0].sum() * dx - 0.55) < 1e-13
        ex = exact_riemann_cell_average(1.0, 0.0, 1.0, 0.1, 0.0, 0.1, G, np.arange(n + 1) * dx, 0.5, 0.2)
        errs.append(np.abs(V[0] - ex[0]).mean())
    assert errs[0] < 1e-2 and errs[1] < 0.7 * errs[0], errs
