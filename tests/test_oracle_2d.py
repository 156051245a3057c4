"""Pins the oracle's 2D integrator (VL_Algorithm_2D_CUDA, src/integrators/VL_2D_cuda.cu:24-256,
with Update_Conserved_Variables_2D and Calc_dt_2D, hydro_cuda.cu:84-160, 474-510).  The
reference holds no 2D vectors; the 2D step differs from the pinned 3D and 1D steps only in
what a missing dimension removes, so it is pinned against them:
  * a z-uniform 3D box advanced by the 3D driver and the same data as a 2D box advanced by
    the 2D driver agree to round-off (the z fluxes cancel exactly; the 3D crossing time
    sums its terms in another order, so dt may differ in its last bits);
  * a y-uniform 2D Sod tube equals the 1D driver's result to round-off, and the same tube
    along y equals it along x bit for bit (direction permutation);
  * mass is conserved to round-off on a periodic 2D box.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from helpers import sod_ic

G = 5.0 / 3.0
NG = {orc.PLMC: 3, orc.PPMC: 4, orc.PLMP: 3, orc.PCM: 2, orc.PPMP: 4}


def kh_2d(n, ng, seed=4):
    rng = np.random.default_rng(seed)
    N = n + 2 * ng
    x = (np.arange(N) - ng + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="xy")
    inner = np.abs(Y - 0.5) < 0.25
    d = np.where(inner, 2.0, 1.0)
    vx = np.where(inner, 0.5, -0.5)
    vy = 0.01 * np.sin(4 * np.pi * X) + 1e-4 * rng.uniform(-1, 1, X.shape)
    p = 2.5 * np.ones_like(d)
    vz = 0.05 * np.cos(2 * np.pi * Y)
    U = np.stack([d, d * vx, d * vy, d * vz, p / (G - 1) + 0.5 * d * (vx * vx + vy * vy + vz * vz)]).reshape(-1)
    U = U.copy()
    orc.set_boundaries(U, N, N, 1, 5, ng, [1, 1, 1, 1, 0, 0])
    return U, N


@pytest.mark.parametrize("aspect", [1.0, 0.5])
@pytest.mark.parametrize("recon", [orc.PLMC, orc.PPMC, orc.PLMP, orc.PCM])
def test_2d_matches_z_uniform_3d(recon, aspect):
    """aspect = dy / dx: non-square cells catch a dx used in place of dy."""
    ng, n = NG[recon], 16
    U2, N = kh_2d(n, ng)
    Nz = 3 * ng
    U3 = np.repeat(U2.reshape(5, 1, N, N), Nz, axis=1).reshape(-1).copy()
    h = 1.0 / n
    hy = aspect * h
    s2, t2 = orc.run(U2, N, N, 1, ng, 5, h, hy, h, G, recon, [1, 1, 1, 1, 0, 0], 0.0, 1e30, 8)
    s3, t3 = orc.run(U3, N, N, Nz, ng, 5, h, hy, h, G, recon, [1] * 6, 0.0, 1e30, 8)
    assert s2 == s3 == 8 and abs(t2 - t3) <= 1e-14 * t3
    a = U2.reshape(5, N, N)[:, ng:-ng, ng:-ng]
    b = U3.reshape(5, Nz, N, N)[:, ng:-ng, ng:-ng, ng:-ng]
    for k in range(b.shape[1]):
        np.testing.assert_allclose(b[:, k], a, rtol=1e-12, atol=1e-13)
    mass = a[0].sum() * h * hy
    U0, _ = kh_2d(n, ng)
    assert abs(mass - U0.reshape(5, N, N)[0, ng:-ng, ng:-ng].sum() * h * hy) < 1e-13


@pytest.mark.parametrize("recon", [orc.PLMC, orc.PPMC, orc.PLMP])
def test_sod_2d_matches_1d_and_is_direction_symmetric(recon):
    ng = NG[recon]
    U1, nx, dx = sod_ic(ng=ng, gamma=1.4)
    ny = 2 * ng + 3
    # x-tube: the 1D state copied along y (periodic there)
    Ux = np.repeat(U1.reshape(5, 1, nx), ny, axis=1).reshape(-1).copy()
    fx = [3, 3, 1, 1, 0, 0]
    orc.set_boundaries(Ux, nx, ny, 1, 5, ng, fx)
    sx = orc.run(Ux, nx, ny, 1, ng, 5, dx, dx, dx, 1.4, recon, fx, 0.0, 0.2, 100000)
    orc.set_boundaries(U1, nx, 1, 1, 5, ng, [3, 3, 0, 0, 0, 0])
    s1 = orc.run(U1, nx, 1, 1, ng, 5, dx, dx, dx, 1.4, recon, [3, 3, 0, 0, 0, 0], 0.0, 0.2, 100000)
    assert sx == s1
    A = Ux.reshape(5, ny, nx)[:, ng:-ng, ng:-ng]
    B = U1.reshape(5, nx)[:, ng:-ng]
    for j in range(A.shape[1]):
        np.testing.assert_allclose(A[:, j], B, rtol=1e-13, atol=1e-14)
    # y-tube: transpose the tube, swap the momentum slots
    U1b, _, _ = sod_ic(ng=ng, gamma=1.4)
    Uy = np.repeat(U1b.reshape(5, nx, 1), ny, axis=2)[[0, 2, 1, 3, 4]].reshape(-1).copy()
    fy = [1, 1, 3, 3, 0, 0]
    orc.set_boundaries(Uy, ny, nx, 1, 5, ng, fy)
    sy = orc.run(Uy, ny, nx, 1, ng, 5, dx, dx, dx, 1.4, recon, fy, 0.0, 0.2, 100000)
    assert sy == sx
    C = Uy.reshape(5, nx, ny)[[0, 2, 1, 3, 4]][:, ng:-ng, ng:-ng]
    np.testing.assert_array_equal(C.transpose(0, 2, 1), A)
