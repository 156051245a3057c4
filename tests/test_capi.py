"""CPU-side checks of the C-ABI library: it loads, exports every symbol that
include/cholla_amd.h declares, refuses to run without a device (no silent
fallback), and its host-only problem setup is correct."""
import math
import os
import re
import subprocess

import numpy as np
import sys
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pytest

import cholla_amd as chl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cholla_amd.h")


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not os.path.exists(chl.LIB_PATH):
        from cholla_amd import build

        build.build(verbose=False)


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(chl_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(chl.EXPORTS)


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", chl.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(l.split()[-1] for l in out.stdout.splitlines() if l.strip())
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    L = chl.lib()  # binds every declared symbol with its signature
    assert L.chl_version().decode().startswith("cholla_amd")


def test_no_device_fails_loudly():
    if chl.device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(chl.ChollaError):
        chl.Hydro(16, 16, 16, 4)


def test_kh_initial_state_host():
    n, ng = 16, 4
    N = n + 2 * ng
    U = chl.ic_kelvin_helmholtz(n, ng, 5 / 3, seed=1).reshape(5, N, N, N)
    d = U[0, ng:-ng, ng:-ng, ng:-ng]
    # half the slab is dense: mean density 1.5, mean x-momentum 2*0.5*0.5 + 1*(-0.5)*0.5 = 0.25
    assert abs(d.mean() - 1.5) < 1e-14
    assert abs(U[1, ng:-ng, ng:-ng, ng:-ng].mean() - 0.25) < 1e-14
    assert np.all(U[3] == 0.0)
    p = (U[4] - 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0]) * (2 / 3)
    assert np.allclose(p, 2.5, rtol=0, atol=1e-12)
    # periodic ghosts
    assert np.array_equal(U[:, :ng], U[:, n:n + ng])
    assert np.array_equal(U[:, :, -ng:], U[:, :, ng:2 * ng])
    # same seed, same state; different seed, different noise
    assert np.array_equal(U.reshape(-1), chl.ic_kelvin_helmholtz(n, ng, 5 / 3, seed=1))
    assert not np.array_equal(U.reshape(-1), chl.ic_kelvin_helmholtz(n, ng, 5 / 3, seed=2))


# ----------------------------------------------------------------------------- ensemble setup (host side)
def test_ensemble_no_device_fails_loudly():
    if chl.device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(chl.ChollaError):
        chl.Ensemble([(40, 40, 40)], 4)
    with pytest.raises(chl.ChollaError):
        chl.check(chl.lib().chl_ic_turbulence(None, 16, 4, 5 / 3, 1.0, 1.0, 1.0, 1))


def test_ensemble_layout_draws():
    edges, mach = chl.ensemble_layout(2024, 384)
    assert set(np.unique(edges)) <= {32, 48, 64, 96, 128}
    assert len(np.unique(edges)) == 5
    assert mach.min() >= 0.5 and mach.max() <= 3.0
    # uniform over the five edges and over [0.5, 3]: loose 5-sigma bounds for 384 draws
    counts = np.array([(edges == e).sum() for e in (32, 48, 64, 96, 128)])
    assert np.all(np.abs(counts - 384 / 5) < 5 * np.sqrt(384 * 0.2 * 0.8)), counts
    assert abs(mach.mean() - 1.75) < 5 * 2.5 / np.sqrt(12 * 384)
    e2, m2 = chl.ensemble_layout(2024, 384)
    assert np.array_equal(edges, e2) and np.array_equal(mach, m2)
    e3, _ = chl.ensemble_layout(2025, 384)
    assert not np.array_equal(edges, e3)
    # the first draw is the engine's first output mod 5 (std::mt19937_64, seed 5489 -> 14514284786278117030)
    e0, _ = chl.ensemble_layout(5489, 1)
    assert e0[0] == [32, 48, 64, 96, 128][14514284786278117030 % 5]


def test_turbulence_modes():
    m = chl.turbulence_modes(7)
    k = m[:, :3]
    # every half-space integer vector with 1 <= |k|^2 <= 16, each once, in kx, ky, kz order
    want = [(kx, ky, kz) for kx in range(0, 5) for ky in range(-4, 5) for kz in range(-4, 5)
            if 1 <= kx * kx + ky * ky + kz * kz <= 16 and (kx > 0 or (kx == 0 and (ky > 0 or (ky == 0 and kz > 0))))]
    assert [tuple(int(v) for v in row) for row in k] == want
    full = sum(1 for kx in range(-4, 5) for ky in range(-4, 5) for kz in range(-4, 5)
               if 1 <= kx * kx + ky * ky + kz * kz <= 16)
    assert 2 * len(want) == full
    k2 = (k ** 2).sum(1)
    a, b = m[:, 3:6], m[:, 6:9]
    # solenoidal: k . a = k . b = 0
    assert np.abs((k * a).sum(1)).max() < 1e-14 and np.abs((k * b).sum(1)).max() < 1e-14
    # per-mode power ~ k^-4 (shell spectrum E(k) ~ k^-2): after the projection two of
    # three components remain, so E[|a|^2] = 2 / |k|^4; |a|^2 |k|^4 / 2 averages to 1
    # (chi-square with 2 d.o.f. / 2)
    s = np.concatenate([(a ** 2).sum(1) * k2 ** 2 / 2, (b ** 2).sum(1) * k2 ** 2 / 2])
    assert abs(s.mean() - 1.0) < 5 / np.sqrt(len(s)), s.mean()
    # and the power of a shell follows the count of its modes times |k|^-4: the shell
    # 1.5 <= |k| < 2.5 against 3.5 <= |k| <= 4 (expected ratio ~7; a |k|^-1 amplitude
    # would give ~2), within sampling noise (~0.2 relative for these mode counts)
    kk = np.sqrt(k2)
    pw = (a ** 2).sum(1) + (b ** 2).sum(1)
    s1, s2 = (kk >= 1.5) & (kk < 2.5), (kk >= 3.5) & (kk <= 4.0)
    expect = (kk[s1] ** -4.0).sum() / (kk[s2] ** -4.0).sum()
    ratio = pw[s1].sum() / pw[s2].sum()
    assert abs(ratio / expect - 1) < 0.5, (ratio, expect)
    assert not np.array_equal(m, chl.turbulence_modes(8))
    assert np.array_equal(m, chl.turbulence_modes(7))


# ----------------------------------------------------------------------------- MHD setup (host side)
def test_vector_potential_matches_reference_vector():
    """mhd_utilities_tests.cu tMHDInitMagneticFieldWithVectorPotential on the product's host routine."""
    import json

    g = json.load(open(os.path.join(ROOT, "tests", "golden", "mhd_unit_vectors.json")))["vector_potential"]
    n = g["n"]
    U = np.full(8 * n ** 3, -999.0)
    A = np.arange(3 * n ** 3, dtype=float)
    chl.magnetic_field_from_vector_potential(U, n, n, n, g["d"], g["d"], g["d"], A)
    for i, v in enumerate(U):
        assert v == g["expect"].get(str(i), -999.0), (i, v)


def test_orszag_tang_initial_state():
    from helpers import vector_potential_b

    n, ng, g = 16, 4, 5 / 3
    N = n + 2 * ng
    U = chl.ic_orszag_tang(n, ng, g, seed=5, amplitude=1e-3).reshape(8, N, N, N)
    h = 1.0 / n
    # divergence of the face field vanishes to round-off on every cell with a full stencil
    B = U[5:8]
    div = (B[0, 1:, 1:, 1:] - B[0, 1:, 1:, :-1]) / h + (B[1, 1:, 1:, 1:] - B[1, 1:, :-1, 1:]) / h + \
          (B[2, 1:, 1:, 1:] - B[2, :-1, 1:, 1:]) / h
    assert np.abs(div).max() < 1e-11
    # background field: B = (-sin 2 pi y, sin 4 pi x, 0) / sqrt(4 pi) up to the discrete curl and the perturbation
    x = (np.arange(N) - ng + 0.5) / n
    B0 = 1 / math.sqrt(4 * math.pi)
    rho = 25 / (36 * math.pi)
    assert np.all(U[0, ng:-ng, ng:-ng, ng:-ng] == rho)
    vx = U[1] / U[0]
    assert np.allclose(vx[ng:-ng, ng:-ng, ng:-ng], np.sin(2 * math.pi * x)[None, ng:-ng, None], atol=1e-14)
    by = U[6, ng:-ng, ng:-ng, ng:-ng]
    assert np.abs(by.mean(axis=(0, 1)) - B0 * np.sin(4 * math.pi * (x[ng:-ng] - 0.5 / n))).max() < 0.05 * B0
    assert np.abs(U[5, ng:-ng, ng:-ng, ng:-ng].mean()) < 1e-3 and np.abs(U[7]).max() < 1e-12
    # the perturbation is there (Bx varies along z) and is of the requested size
    dbx = U[5, ng:-ng, ng, ng] - U[5, ng:-ng, ng, ng].mean()
    assert 1e-5 < np.abs(dbx).max() < 1e-2
    # periodic ghosts, all fields
    assert np.array_equal(U[:, :ng], U[:, n:n + ng])
    assert np.array_equal(U[:, :, :, -ng:], U[:, :, :, ng:2 * ng])
    # pressure recovered from the energy with the cell-centred field is the uniform background
    bc = np.stack([0.5 * (B[0] + np.roll(B[0], 1, 2)), 0.5 * (B[1] + np.roll(B[1], 1, 1)),
                   0.5 * (B[2] + np.roll(B[2], 1, 0))])
    p = (U[4] - 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / U[0] - 0.5 * (bc ** 2).sum(0)) * (g - 1)
    assert np.allclose(p[ng:-ng, ng:-ng, ng:-ng], 5 / (12 * math.pi), rtol=1e-12)
    assert np.array_equal(U.reshape(-1), chl.ic_orszag_tang(n, ng, g, seed=5, amplitude=1e-3))
