"""Pins the 2D dual-energy branches of the oracle (the DE build of VL_Algorithm_2D_CUDA:
VL_2D_cuda.cu:140-152, 203-252; Partial_Update_Advected_Internal_Energy_2D,
Select_Internal_Energy_2D, Sync_Energies_2D, hydro_cuda.cu:702-750, 887-935, 1042-1068)
against the pinned 1D and 3D DE drivers:
  * a cold supersonic compression along x and along y (the advected internal energy
    selected) equals the 1D DE driver's result, both pdV terms included;
  * a warm 2D flow gives the non-DE 2D fields to 1e-12 with GasEnergy = E - kinetic, and a
    z-uniform 3D DE box gives the same as the 2D DE box.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from test_oracle_de_1d import state_1d

G = 5.0 / 3.0


def run(U, nx, ny, nz, ng, h, steps, recon, de, nf, flags, tout=1e30):
    return orc.run(U, nx, ny, nz, ng, nf, h, h, h, G, recon, flags, 0.0, tout, steps, dual_energy_on=de)


@pytest.mark.parametrize("axis", ["x", "y"])
@pytest.mark.parametrize("recon,ng", [(orc.PLMC, 3), (orc.PPMC, 4)])
def test_cold_compression_2d_equals_1d(recon, ng, axis):
    """Along x (uniform in y) and along y (uniform in x, the momenta swapped): the y-direction
    pdV terms are only exercised by the second."""
    n = 32
    x = (np.arange(n) + 0.5) / n
    d, vx, p = np.ones(n), 3.0 + 0.5 * np.sin(2 * np.pi * x), 1e-4 * np.ones(n)
    U1, nx = state_1d(d, vx, 0.2 * np.cos(2 * np.pi * x), 0 * d, p, ng)
    nt = 3 * ng
    U2 = np.repeat(U1.reshape(6, 1, nx), nt, axis=1)  # (field, y, x)
    dims = (nx, nt)
    if axis == "y":
        U2 = U2.transpose(0, 2, 1)[[0, 2, 1, 3, 4, 5]]
        dims = (nt, nx)
    U2 = U2.reshape(-1).copy()
    f1, f2 = [1, 1, 0, 0, 0, 0], [1, 1, 1, 1, 0, 0]
    orc.set_boundaries(U2, dims[0], dims[1], 1, 6, ng, f2)
    s1 = run(U1, nx, 1, 1, ng, 1.0 / n, 40, recon, True, 6, f1)
    s2 = run(U2, dims[0], dims[1], 1, ng, 1.0 / n, 40, recon, True, 6, f2)
    assert s1 == s2
    A = U2.reshape(6, dims[1], dims[0])
    if axis == "y":
        A = A.transpose(0, 2, 1)[[0, 2, 1, 3, 4, 5]]
    A = A[:, ng:-ng, ng:-ng]
    B = U1.reshape(6, nx)[:, ng:-ng]
    assert A[0].max() - A[0].min() > 0.3  # compressed
    for j in range(A.shape[1]):
        np.testing.assert_allclose(A[:, j], B, rtol=1e-12, atol=1e-14)


def warm_2d(n, ng, seed=2):
    rng = np.random.default_rng(seed)
    N = n + 2 * ng
    x = (np.arange(N) - ng + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="xy")
    d = 1.0 + 0.2 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    vx, vy, vz = 0.4 * np.sin(2 * np.pi * Y), 0.3 * np.cos(2 * np.pi * X), 0.1 * np.ones_like(X)
    p = 1.0 + 0.1 * rng.uniform(-1, 1, X.shape)
    ge = p / (G - 1)
    E = 0.5 * d * (vx * vx + vy * vy + vz * vz) + ge
    U = np.stack([d, d * vx, d * vy, d * vz, E, ge]).reshape(-1).copy()
    orc.set_boundaries(U, N, N, 1, 6, ng, [1, 1, 1, 1, 0, 0])
    return U, N


def test_warm_2d_matches_non_de_and_3d():
    ng, n = 4, 16
    U6, N = warm_2d(n, ng)
    U5 = U6.reshape(6, -1)[:5].reshape(-1).copy()
    Nz = 3 * ng
    U3 = np.repeat(U6.reshape(6, 1, N, N), Nz, axis=1).reshape(-1).copy()
    f2 = [1, 1, 1, 1, 0, 0]
    assert run(U6, N, N, 1, ng, 1.0 / n, 8, orc.PPMC, True, 6, f2)[0] == 8
    assert run(U5, N, N, 1, ng, 1.0 / n, 8, orc.PPMC, False, 5, f2)[0] == 8
    assert run(U3, N, N, Nz, ng, 1.0 / n, 8, orc.PPMC, True, 6, [1] * 6)[0] == 8
    a = U6.reshape(6, N, N)[:, ng:-ng, ng:-ng]
    b = U5.reshape(5, N, N)[:, ng:-ng, ng:-ng]
    for f in range(5):
        assert np.abs(a[f] - b[f]).sum() <= 1e-12 * np.abs(b[f]).sum(), f
    ke = 0.5 * a[0] * ((a[1] / a[0]) ** 2 + (a[2] / a[0]) ** 2 + (a[3] / a[0]) ** 2)
    assert np.abs(a[5] - (a[4] - ke)).max() <= 1e-13 * np.abs(a[4]).max()
    c = U3.reshape(6, Nz, N, N)[:, ng:-ng, ng:-ng, ng:-ng]
    for k in range(c.shape[1]):
        np.testing.assert_allclose(c[:, k], a, rtol=1e-12, atol=1e-13)
