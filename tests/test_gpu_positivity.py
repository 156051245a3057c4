"""GPU parity of the PPMC positivity fallback (chl_hydro_config.positivity_fallback,
chl_ensemble_config.positivity_fallback; an option of this implementation, off by
default) against the oracle's fallback (tests/test_oracle_positivity.py pins it).

Bars:
  * PolicyExact: bit-identical to the oracle's driver loop with the fallback on a rough
    box where it fires (and the result differs from the plain PPMC run), single box and
    ensemble; PolicyFast: per-field relative L1 <= 1e-12;
  * on the benchmark's hardest turbulence boxes (32^3 and 48^3 at rms Mach 2.5-3), which
    the plain scheme drives to dt collapse and non-finite states within 20-110 steps, the
    fallback keeps every box finite with a steady dt through 120 steps in the FP64 and
    mixed forms and both launch forms, mass conserved to round-off.
"""
import sys

import numpy as np
import pytest

import cholla_amd as chl
from helpers import interior
from oracle import oracle as orc
from test_oracle_positivity import rough_state

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    if chl.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need a real MI355X")


def oracle_run(U, N, n, steps, pf):
    A = U.copy()
    h = 1.0 / n
    orc.run(A, N, N, N, 4, 5, h, h, h, G, orc.PPMC, [1] * 6, 0.0, 1e30, steps, positivity_fallback_on=pf)
    return A


@pytest.mark.parametrize("precision", [chl.PREC_FP64_EXACT, chl.PREC_FP64])
def test_fallback_single_box_matches_oracle(precision):
    n, steps = 16, 6
    U, (N, _, _) = rough_state(n, 4, seed=4)
    ref = oracle_run(U, N, n, steps, True)
    assert not np.array_equal(ref, oracle_run(U, N, n, steps, False))  # the fallback fired
    say("positivity fallback single box precision", precision)
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, 4, chl.RECON_PPMC, precision, positivity_fallback=True)
    h = 1.0 / n
    ns, t = hy.run(Ud, h, h, h, G, 0.0, 1e30, steps)
    A = Ud.download()
    hy.close()
    assert ns == steps
    if precision == chl.PREC_FP64_EXACT:
        np.testing.assert_array_equal(A, ref)
    else:
        a, b = interior(A, 5, N, N, N, 4), interior(ref, 5, N, N, N, 4)
        l1 = [np.abs(a[f] - b[f]).sum() / np.abs(b[f]).sum() for f in range(5)]
        assert max(l1) <= 1e-12, l1


def test_fallback_dual_energy_exact_matches_oracle():
    """The fallback under the DE build also replaces the cell's GasEnergy interfaces."""
    n, steps = 16, 6
    U, (N, _, _) = rough_state(n, 4, seed=6, de=True)
    h = 1.0 / n
    refs = {}
    for pf in (True, False):
        A = U.copy()
        orc.run(A, N, N, N, 4, 6, h, h, h, G, orc.PPMC, [1] * 6, 0.0, 1e30, steps, dual_energy_on=True,
                positivity_fallback_on=pf)
        refs[pf] = A
    assert not np.array_equal(refs[True], refs[False])
    say("positivity fallback dual energy exact")
    Ud = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, 4, chl.RECON_PPMC, chl.PREC_FP64_EXACT, dual_energy=True, positivity_fallback=True)
    ns, t = hy.run(Ud, h, h, h, G, 0.0, 1e30, steps)
    A = Ud.download()
    hy.close()
    assert ns == steps
    np.testing.assert_array_equal(A, refs[True])


def test_fallback_ensemble_exact_matches_oracle():
    boxes = [(12, 3), (20, 5)]
    inputs, refs, dims = [], [], []
    for n, seed in boxes:
        U, (N, _, _) = rough_state(n, 4, seed=seed)
        refs.append(oracle_run(U, N, n, 5, True))
        inputs.append(U)
        dims.append((N, N, N))
    for launch in (chl.LAUNCH_SPLIT, chl.LAUNCH_PERSISTENT):
        say("positivity fallback ensemble exact launch", launch)
        dev = [chl.DeviceArray.from_host(U) for U in inputs]
        ens = chl.Ensemble(dims, 4, chl.RECON_PPMC, chl.PREC_FP64_EXACT, launch, positivity_fallback=True)
        ens.bind(dev, np.array([[1.0 / n] * 3 for n, _ in boxes]).reshape(-1))
        ens.start(G)
        ens.advance(5)
        for q, d in enumerate(dev):
            np.testing.assert_array_equal(d.download(), refs[q])
        ens.close()


def test_fallback_requires_ppmc():
    with pytest.raises(chl.ChollaError):
        chl.Hydro(24, 24, 24, 3, chl.RECON_PLMC, chl.PREC_FP64, positivity_fallback=True)
    with pytest.raises(chl.ChollaError):
        chl.Hydro(100, 1, 1, 4, chl.RECON_PPMC, chl.PREC_FP64_EXACT, positivity_fallback=True)
    with pytest.raises(chl.ChollaError):
        chl.Ensemble([(40, 40, 40)], 3, chl.RECON_PLMC, chl.PREC_FP64, chl.LAUNCH_SPLIT, positivity_fallback=True)


@pytest.mark.parametrize("precision,launch", [(chl.PREC_FP64, chl.LAUNCH_SPLIT), (chl.PREC_MIXED, chl.LAUNCH_SPLIT),
                                              (chl.PREC_MIXED, chl.LAUNCH_PERSISTENT)])
def test_fallback_keeps_hard_turbulence_boxes_finite(precision, launch):
    edges, mach = chl.ensemble_layout(2024, 384)
    pick = [11, 76, 154, 158, 200, 210, 239, 258, 284, 345, 372]
    ng, steps = 4, 120
    dims = [(int(edges[b]) + 2 * ng,) * 3 for b in pick]
    dev = []
    for q, b in enumerate(pick):
        U = chl.DeviceArray(5 * dims[q][0] ** 3)
        chl.ic_turbulence(U, int(edges[b]), ng, G, 1.0, 1.0, float(mach[b]), 2024 * 1000 + b)
        dev.append(U)
    mass0 = [interior(d.download(), 5, *dims[q], ng)[0].sum() for q, d in enumerate(dev)]
    say("positivity fallback turbulence boxes precision", precision, "launch", launch)
    ens = chl.Ensemble(dims, ng, chl.RECON_PPMC, precision, launch, positivity_fallback=True)
    ens.bind(dev, np.array([[1.0 / edges[b]] * 3 for b in pick]).reshape(-1))
    ens.start(G, U_floor=1e-3)
    ens.advance(1)
    _, dt1, _ = ens.clocks()
    ens.advance(steps - 1)
    t, dt, ns = ens.clocks()
    assert (ns == steps).all()
    assert (dt > 0.3 * dt1).all(), list(zip(pick, dt / dt1))
    for q, d in enumerate(dev):
        a = interior(d.download(), 5, *dims[q], ng)
        assert np.isfinite(a).all() and a[0].min() > 0, pick[q]
        assert abs(a[0].sum() - mass0[q]) <= 1e-12 * mass0[q], (pick[q], a[0].sum(), mass0[q])
    ens.close()
