"""Pins the oracle's PPMC positivity fallback (an option of this implementation with no
counterpart in the reference, so there are no reference vectors for it).

PPMC_VL floors an interface density and pressure at TINY_NUMBER each on its own
(ppmc_cuda.cu:681-684).  Where the characteristic projection drives an interface
density below zero while the pressure stays finite, the floored state has a sound
speed of order sqrt(p / 1e-20), and the HLLC flux through that face injects energy
without bound.  With the fallback on, a cell whose interface density or pressure comes
out non-positive uses its cell-centred values on both of its interfaces in that sweep.

The checks, each failing under a plausible slip (one side replaced, the wrong cell's
values, a replacement that fires on healthy cells or misses failing ones):
  * on a rough state, every cell's two interfaces are either those of the plain PPMC
    or both equal to the cell's own conserved values, and exactly the cells whose
    plain interfaces were floored are replaced;
  * on a smooth flow the option changes no bit of a full run;
  * on a 32^3 rms-Mach-2.8 turbulence box the plain scheme's dt collapses by orders of
    magnitude within 40 steps (an interface pair floored to 1e-20 density next to a
    finite pressure), while with the fallback dt stays within a factor two of the
    initial one and the box's mass is conserved to round-off.
"""
import numpy as np

from oracle import oracle as orc
from test_oracle_de import periodic_state

G = 5.0 / 3.0
TINY = 1e-20


def rough_state(n, ng, seed, sd=0.3, mach=1.0, de=False):
    """Periodic box of log-normal density and pressure (sigma sd) and Gaussian
    velocities of rms mach per component: rough enough for PPMC to produce a few
    dozen non-positive interface densities or pressures.  5 fields, or 6 with the
    GasEnergy of the DE build (de=True)."""
    rng = np.random.default_rng(seed)
    d = np.exp(sd * rng.standard_normal((n, n, n)))
    p = np.exp(sd * rng.standard_normal((n, n, n)))
    v = mach * rng.standard_normal((3, n, n, n))
    U, dims = periodic_state(n, n, n, ng, d, v[0], v[1], v[2], p)
    return (U if de else U.reshape(6, -1)[:5].reshape(-1).copy()), dims


def ppmc(U, N, direction, fallback):
    nc = N ** 3
    QL, QR = np.zeros(5 * nc), np.zeros(5 * nc)
    with orc.positivity_fallback(fallback):
        orc.lib().orc_ppmc_vl(orc.ptr(U), orc.ptr(QL), orc.ptr(QR), N, N, N, G, direction)
    return QL.reshape(5, N, N, N), QR.reshape(5, N, N, N)


def floored(Q):
    """Interfaces whose density or pressure sits at the TINY_NUMBER floor."""
    d = Q[0]
    ke = 0.5 * (Q[1] ** 2 + Q[2] ** 2 + Q[3] ** 2) / d
    p = (Q[4] - ke) * (G - 1)
    return (d <= TINY) | (p <= 1e-12 * np.abs(Q[4]))


def test_fallback_replaces_exactly_the_failing_cells():
    n, ng = 12, 4
    U, (N, _, _) = rough_state(n, ng, seed=1)
    C = U.reshape(5, N, N, N)
    inner = slice(3, N - 3)  # PPMC's thread guard (ppmc_cuda.cu:556-560)
    fired_total = 0
    for direction in range(3):
        QL0, QR0 = ppmc(U, N, direction, False)
        QL1, QR1 = ppmc(U, N, direction, True)
        # cell i writes QL[i] (its face i+1/2) and QR[i - e] (its face i-1/2)
        sh = [0, 0, 0]
        sh[2 - direction] = 1  # array axes are (z, y, x)
        back = lambda Q: np.roll(Q, shift=sh, axis=(1, 2, 3))  # back(Q)[i] = Q[i - e]
        L0, L1, R0, R1 = QL0, QL1, back(QR0), back(QR1)
        cells = (slice(None), inner, inner, inner)
        L0, L1, R0, R1, Cc = L0[cells], L1[cells], R0[cells], R1[cells], C[cells]
        changed = (L0 != L1).any(0) | (R0 != R1).any(0)
        failing = floored(L0) | floored(R0)
        assert np.array_equal(changed, failing), direction
        fired = np.nonzero(changed)
        fired_total += len(fired[0])
        for Q in (L1, R1):
            got, want = Q[(slice(None),) + fired], Cc[(slice(None),) + fired]
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)
    assert fired_total >= 10


def test_fallback_is_a_no_op_on_a_smooth_flow():
    n, ng = 16, 4
    z, y, x = np.meshgrid(*(((np.arange(n) + 0.5) / n),) * 3, indexing="ij")
    d = 1 + 0.2 * np.sin(2 * np.pi * (x + y))
    vx = 0.5 * np.cos(2 * np.pi * z)
    vy = 0.3 * np.sin(2 * np.pi * x)
    vz = 0.2 * np.cos(2 * np.pi * y)
    p = 1 + 0.1 * np.cos(2 * np.pi * (y - z))
    U, (N, _, _) = periodic_state(n, n, n, ng, d, vx, vy, vz, p)
    U = U.reshape(6, -1)[:5].reshape(-1).copy()
    A, B = U.copy(), U.copy()
    h = 1.0 / n
    orc.run(A, N, N, N, ng, 5, h, h, h, G, orc.PPMC, [1] * 6, 0.0, 1e30, 8)
    orc.run(B, N, N, N, ng, 5, h, h, h, G, orc.PPMC, [1] * 6, 0.0, 1e30, 8, positivity_fallback_on=True)
    np.testing.assert_array_equal(A, B)


def turbulence_box(box, ng=4, seed=2024):
    """Box `box` of the seeded turbulence ensemble layout, built on the host from the
    library's mode table (the same field ic_turbulence writes)."""
    import cholla_amd as chl

    edges, mach = chl.ensemble_layout(seed, box + 1)
    n = int(edges[box])
    modes = chl.turbulence_modes(seed * 1000 + box)
    x = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    v = np.zeros((3, n, n, n))
    for m in modes:
        ph = 2 * np.pi * (m[0] * X + m[1] * Y + m[2] * Z)
        c, s = np.cos(ph), np.sin(ph)
        for q in range(3):
            v[q] += m[3 + q] * c + m[6 + q] * s
    v *= mach[box] * np.sqrt(G) / np.sqrt((v ** 2).sum(0).mean())
    one = np.ones((n, n, n))
    U, dims = periodic_state(n, n, n, ng, one, v[0], v[1], v[2], one)
    return U.reshape(6, -1)[:5].reshape(-1).copy(), dims, n, float(mach[box])


def test_fallback_keeps_a_mach3_box_steady():
    U0, (N, _, _), n, mach = turbulence_box(284)
    assert n == 32 and mach > 2.7
    h, ng, steps = 1.0 / n, 4, 40
    mass0 = U0.reshape(5, N, N, N)[0, ng:-ng, ng:-ng, ng:-ng].sum()
    out = {}
    for on in (False, True):
        U = U0.copy()
        dts = np.zeros(steps)
        orc.run(U, N, N, N, ng, 5, h, h, h, G, orc.PPMC, [1] * 6, 0.0, 1e30, steps, U_floor=1e-3, dt_log=dts,
                positivity_fallback_on=on)
        out[on] = (U, dts)
    assert out[False][1].min() < 1e-2 * out[False][1][0]
    U, dts = out[True]
    assert dts.min() > 0.5 * dts[0]
    A = U.reshape(5, N, N, N)[:, ng:-ng, ng:-ng, ng:-ng]
    assert np.isfinite(A).all() and A[0].min() > 0
    assert abs(A[0].sum() - mass0) <= 1e-12 * mass0
