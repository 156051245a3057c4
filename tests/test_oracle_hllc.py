"""Pins of the oracle's HLLC (orc_hllc, restating Calculate_HLLC_Fluxes_CUDA,
src/riemann_solvers/hllc_cuda.cu:19-307) in its subsonic star region, which the reference's
own HLLC test (hllc_cuda_tests.cu, disabled at :22) leaves open.

1. Closed forms.  HLLC resolves an isolated contact and shear wave exactly:
   * a stationary contact (u = 0 on both sides, equal pressures, any densities and transverse
     velocities) has flux (0, p, 0, 0, 0);
   * a moving contact with shear (equal normal velocity u and pressure, different densities
     and transverse velocities, |u| subsonic) has the upwind side's physical flux.
   HLL (and any solver that smears the contact) fails both.
2. An independent HLLC written here from Toro, "Riemann Solvers and Numerical Methods for
   Fluid Dynamics", 3rd ed., section 10.4 (eqs. 10.37-10.39, the F*_K = F_K + S_K (U*_K - U_K)
   form), with the reference's wave-speed estimates (Roe-averaged u +- a against the two sides'
   u -+ c, hllc_cuda.cu:141-175).  The reference writes the flux in the symmetric form of Batten
   et al. (hllc_cuda.cu:242-305), which equals Toro's in exact arithmetic; on random subsonic
   pairs in all three sweep directions the two agree to 1e-12 of the flux scale.

Planting a slip in the oracle's star region (the sign of p_L - p_R in S_M, the wave speed S_L
for S_M in the star density, or the p* S_M term of the star energy) fails these tests; see
the oracle pins listed in the design notes.
"""
import numpy as np
import pytest

from oracle import oracle as orc

G = 5.0 / 3.0


def _hllc(left, right, gamma, dir=0):
    """The oracle's HLLC on one face (1x1x1 grid); states and flux in global component order."""
    QL, QR, F = np.array(left, float), np.array(right, float), np.zeros(5)
    orc.lib().orc_hllc(orc.ptr(QL), orc.ptr(QR), orc.ptr(F), 1, 1, 1, gamma, dir)
    return F


def _cons(d, u, v, w, p, g):
    return [d, d * u, d * v, d * w, p / (g - 1) + 0.5 * d * (u * u + v * v + w * w)]


def _phys(d, u, v, w, p, g):
    E = p / (g - 1) + 0.5 * d * (u * u + v * v + w * w)
    return np.array([d * u, d * u * u + p, d * u * v, d * u * w, (E + p) * u])


def toro_hllc(l, r, g):
    """Independent HLLC flux (Toro section 10.4) for primitive states l, r = (d, u, v, w, p),
    u normal.  Wave speeds as the reference estimates them: S_L = min(u~ - a~, u_L - c_L),
    S_R = max(u~ + a~, u_R + c_R) with Roe averages weighted by sqrt(d)."""
    dl, ul, vl, wl, pl = l
    dr, ur, vr, wr, pr = r
    El = pl / (g - 1) + 0.5 * dl * (ul * ul + vl * vl + wl * wl)
    Er = pr / (g - 1) + 0.5 * dr * (ur * ur + vr * vr + wr * wr)
    cl, cr = np.sqrt(g * pl / dl), np.sqrt(g * pr / dr)
    sl, sr = np.sqrt(dl), np.sqrt(dr)
    avg = lambda a, b: (sl * a + sr * b) / (sl + sr)
    u_, v_, w_ = avg(ul, ur), avg(vl, vr), avg(wl, wr)
    H_ = avg((El + pl) / dl, (Er + pr) / dr)
    a_ = np.sqrt((g - 1) * (H_ - 0.5 * (u_ * u_ + v_ * v_ + w_ * w_)))
    SL, SR = min(u_ - a_, ul - cl), max(u_ + a_, ur + cr)
    # Toro (10.37)
    Ss = (pr - pl + dl * ul * (SL - ul) - dr * ur * (SR - ur)) / (dl * (SL - ul) - dr * (SR - ur))
    UL, UR = np.array(_cons(dl, ul, vl, wl, pl, g)), np.array(_cons(dr, ur, vr, wr, pr, g))
    FL, FR = _phys(dl, ul, vl, wl, pl, g), _phys(dr, ur, vr, wr, pr, g)

    def star(d, u, v, w, p, E, S):  # Toro (10.39)
        f = d * (S - u) / (S - Ss)
        return f * np.array([1.0, Ss, v, w, E / d + (Ss - u) * (Ss + p / (d * (S - u)))])

    if SL >= 0:
        return FL, Ss
    if SR <= 0:
        return FR, Ss
    if Ss >= 0:
        return FL + SL * (star(dl, ul, vl, wl, pl, El, SL) - UL), Ss
    return FR + SR * (star(dr, ur, vr, wr, pr, Er, SR) - UR), Ss


def _to_global(dir, q):
    """Rotated (normal, t1, t2) vector components -> global (x, y, z) slots for sweep dir."""
    n, t1, t2 = q
    return [(n, t1, t2), (t2, n, t1), (t1, t2, n)][dir]


def _face(l, r, dir):
    """The oracle's flux for rotated primitive states l, r, returned in rotated order."""
    gl = _cons(l[0], *_to_global(dir, l[1:4]), l[4], G)
    gr = _cons(r[0], *_to_global(dir, r[1:4]), r[4], G)
    F = _hllc(gl, gr, G, dir)
    m = F[1:4]
    rot = [(m[0], m[1], m[2]), (m[1], m[2], m[0]), (m[2], m[0], m[1])][dir]
    return np.array([F[0], *rot, F[4]])


@pytest.mark.parametrize("dir", [0, 1, 2])
def test_stationary_contact_flux_is_pressure_only(dir):
    rng = np.random.default_rng(10 + dir)
    for _ in range(50):
        p = rng.uniform(0.1, 5)
        l = (rng.uniform(0.1, 5), 0.0, *rng.uniform(-2, 2, 2), p)
        r = (rng.uniform(0.1, 5), 0.0, *rng.uniform(-2, 2, 2), p)
        F = _face(l, r, dir)
        assert np.allclose(F, [0, p, 0, 0, 0], rtol=0, atol=1e-14 * (p + 1)), (l, r, F)


@pytest.mark.parametrize("dir", [0, 1, 2])
def test_moving_contact_with_shear_is_upwinded_exactly(dir):
    rng = np.random.default_rng(20 + dir)
    n_left = n_right = 0
    for _ in range(100):
        p = rng.uniform(0.2, 5)
        dl, dr = rng.uniform(0.2, 5, 2)
        cmin = np.sqrt(G * p / max(dl, dr))
        u = rng.uniform(-0.9, 0.9) * cmin  # subsonic on both sides
        l = (dl, u, *rng.uniform(-2, 2, 2), p)
        r = (dr, u, *rng.uniform(-2, 2, 2), p)
        F = _face(l, r, dir)
        up = l if u > 0 else r
        n_left += u > 0
        n_right += u < 0
        Fup = _phys(*up, G)
        scale = np.abs(Fup).max() + p
        assert np.allclose(F, Fup, rtol=0, atol=1e-13 * scale), (l, r, F, Fup)
    assert n_left > 20 and n_right > 20


def test_hll_would_fail_the_contact_closed_form():
    """The closed form separates HLLC from HLL: the oracle's HLL smears the same contact."""
    l, r = (2.0, 0.3, 0.5, -0.2, 1.0), (0.5, 0.3, -0.4, 0.7, 1.0)
    FL = _phys(*l, G)
    gl, gr = _cons(*l, G), _cons(*r, G)
    QL, QR, F = np.array(gl), np.array(gr), np.zeros(5)
    orc.lib().orc_hll(orc.ptr(QL), orc.ptr(QR), orc.ptr(F), 1, 1, 1, G, 0)
    assert np.abs(F - FL).max() > 1e-3
    assert np.allclose(_hllc(gl, gr, G, 0), FL, rtol=0, atol=1e-13 * 10)


@pytest.mark.parametrize("dir", [0, 1, 2])
def test_random_subsonic_pairs_match_independent_hllc(dir):
    rng = np.random.default_rng(30 + dir)
    signs = {1: 0, -1: 0}
    for _ in range(300):
        dl, dr = rng.uniform(0.1, 4, 2)
        pl, pr = rng.uniform(0.1, 4, 2)
        c = np.sqrt(G * min(pl / dl, pr / dr))
        l = (dl, rng.uniform(-0.8, 0.8) * c, *rng.uniform(-2, 2, 2), pl)
        r = (dr, rng.uniform(-0.8, 0.8) * c, *rng.uniform(-2, 2, 2), pr)
        F = _face(l, r, dir)
        T, Ss = toro_hllc(l, r, G)
        signs[1 if Ss >= 0 else -1] += 1
        scale = np.abs(T).max() + max(pl, pr)
        assert np.allclose(F, T, rtol=0, atol=1e-12 * scale), (l, r, F, T)
    assert signs[1] > 50 and signs[-1] > 50  # both star states exercised
