"""CPU-side checks of the C-ABI library: it loads, exports every symbol that
include/cholla_amd.h declares, refuses to run without a device (no silent
fallback), and its host-only problem setup is correct."""
import os
import re
import subprocess

import numpy as np
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
