"""Pins the oracle's static gravity (the STATIC_GRAV build: calc_g_1D/2D/3D,
src/gravity/static_grav.h, and the time-centred source of the full updates,
hydro_cuda.cu:59-67, 142-153, 301-315).  The reference holds no vectors for it:
  * every field is checked against the potential it derives from, differentiated
    numerically here (NFW halo + Miyamoto-Nagai disk in 3D and along z in 1D, Kuzmin disk +
    NFW halo and a point mass in 2D) or against the rotation it balances (Gresho: v_phi^2 / r),
    so a slip in an exponent, a sign or a component fails;
  * the source on a uniform gas at rest: after one step each cell's momentum is exactly
    dt g(x) d at the cell's own position (offsets and bounds included), the energy
    unchanged (the start and end velocities are zero), which fixes the factor, the sign,
    the time centring and the cell positions.
"""
import numpy as np
import pytest

from oracle import oracle as orc

GN = 4.49451e-18
MSUN = 1.98847e33


def phi_mw(R, z):
    M_vir, M_d, R_d, z_d, R_vir, c = 1.0e12, 6.5e10, 3.5, 0.7, 261.0, 20.0
    M_h, R_h = M_vir - M_d, R_vir / c
    r = np.sqrt(R * R + z * z)
    halo = -GN * M_h / (np.log(1 + c) - c / (1 + c)) * np.log(1 + r / R_h) / r
    disk = -GN * M_d / np.sqrt(R * R + (R_d + np.sqrt(z * z + z_d * z_d)) ** 2)
    return halo + disk


def phi_kuzmin_nfw(r):
    M_d, M_h, R_d, R_s, c = 6.5e10, 1.0e12 - 6.5e10, 3.5, 261.0 / 20, 20.0
    return -GN * M_d / np.sqrt(r * r + R_d * R_d) - GN * M_h / (np.log(1 + c) - c / (1 + c)) * np.log(1 + r / R_s) / r


def grad(f, p, h):
    p = np.array(p, float)
    out = []
    for a in range(p.size):
        e = np.zeros_like(p)
        e[a] = h
        out.append(-(f(*(p + e)) - f(*(p - e))) / (2 * h))
    return np.array(out)


POINTS3 = [(8.0, 0.5, 0.3), (-3.0, 4.0, -1.2), (0.7, -0.2, 2.5), (15.0, -9.0, 0.05)]


@pytest.mark.parametrize("p", POINTS3)
def test_3d_field_is_minus_grad_potential(p):
    with orc.static_gravity(1):
        g = orc.gravity_field(3, *p)
    want = grad(lambda x, y, z: phi_mw(np.hypot(x, y), z), p, 1e-5)
    np.testing.assert_allclose(g, want, rtol=1e-6, atol=1e-9 * np.abs(want).max())


@pytest.mark.parametrize("z", [-2.0, -0.3, 0.1, 1.5, 6.0])
def test_1d_field_is_vertical_gradient_at_fixed_radius(z):
    with orc.static_gravity(1):
        g = orc.gravity_field(1, z)
    want = grad(lambda zz: phi_mw(13.9211647546, zz), [z], 1e-5)[0]
    assert abs(g[0] - want) <= 1e-6 * abs(want)


@pytest.mark.parametrize("p", [(0.05, 0.1), (-0.25, 0.1), (0.3, -0.2), (0.5, 0.3), (-3.0, 7.0)])
def test_2d_fields(p):
    x, y = p
    r, rhat = np.hypot(x, y), np.array([x, y]) / np.hypot(x, y)
    with orc.static_gravity(1):  # Gresho: -v_phi^2 / r rhat
        v = 5 * r if r < 0.2 else (2 - 5 * r if r < 0.4 else 0.0)
        np.testing.assert_allclose(orc.gravity_field(2, x, y)[:2], -v * v / r * rhat, rtol=1e-12, atol=1e-14)
    with orc.static_gravity(2):
        np.testing.assert_array_equal(orc.gravity_field(2, x, y)[:2], [0.0, -1.0])
    with orc.static_gravity(3):
        np.testing.assert_allclose(orc.gravity_field(2, x, y)[:2], grad(lambda a, b: -GN * MSUN / np.hypot(a, b), p,
                                                                        1e-6 * r), rtol=1e-6)
    with orc.static_gravity(4):
        np.testing.assert_allclose(orc.gravity_field(2, x, y)[:2],
                                   grad(lambda a, b: phi_kuzmin_nfw(np.hypot(a, b)), p, 1e-5 * r), rtol=1e-6)


def uniform_at_rest(shape, ng, d=1.3, p=2.0, g=5.0 / 3.0):
    U = np.zeros((5,) + shape)
    U[0], U[4] = d, p / (g - 1)
    return U.reshape(-1).copy()


@pytest.mark.parametrize("dims", [1, 2, 3])
def test_source_on_uniform_gas_at_rest(dims):
    ng, n, G = 3, 8, 5.0 / 3.0
    N = n + 2 * ng
    shape = {1: (N,), 2: (N, N), 3: (N, N, N)}[dims]
    nx, ny, nz = N, N if dims > 1 else 1, N if dims > 2 else 1
    grav = {1: 1, 2: 4, 3: 1}[dims]
    off, bnd = (2, 5, 1), (-4.0, 3.0, -0.5)
    h = 0.7
    U = uniform_at_rest(shape, ng)
    flags = [1, 1, 1 if dims > 1 else 0, 1 if dims > 1 else 0, 1 if dims > 2 else 0, 1 if dims > 2 else 0]
    with orc.static_gravity(grav, off, bnd):
        steps, t = orc.run(U, nx, ny, nz, ng, 5, h, h, h, G, orc.PLMC, flags, 0.0, 1e30, 1)
        assert steps == 1
        V = U.reshape((5,) + shape)
        real = tuple(slice(ng, N - ng) for _ in shape)
        idx = np.arange(ng, N - ng)
        pos = [(off[a] + idx - ng + 0.5) * h + bnd[a] for a in range(dims)]
        for cell in np.ndindex(*(n,) * dims):
            q = tuple(ng + c for c in reversed(cell)) if dims > 1 else (ng + cell[0],)
            xyz = [pos[a][cell[a]] for a in range(dims)] + [0.0] * (3 - dims)
            g = orc.gravity_field(dims, *xyz)
            for a in range(dims):
                want = t * g[a] * 1.3  # dt g (d + d_n) / 2 with d_n = d
                assert abs(V[(1 + a,) + q] - want) <= 1e-15 * abs(want) + 1e-300
        np.testing.assert_array_equal(V[(4,) + real], 2.0 / (G - 1))
        np.testing.assert_array_equal(V[(0,) + real], 1.3)
