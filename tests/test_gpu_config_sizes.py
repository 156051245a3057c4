"""The exact form against the oracle at (or near) the benchmark configurations' own sizes.

The parity files check every stage on small boxes; these runs repeat the integrated path at
the sizes the configs state, where a 32-bit index, a tile-edge case or a z-chunk boundary that
small boxes never reach would show:
  * KH 128^3 x 100 steps and KH 256^3 x 10 steps (config 2's box), exact form, bit for bit;
    the fast FP64 form on the 256^3 box within its bounds;
  * Orszag-Tang 64^3 x 50 steps (config 5's setup), exact form, bit for bit;
  * a turbulence ensemble with a 128^3 box among smaller ones (config 4's box sizes and
    initial conditions), exact form, 20 steps, bit for bit per box.
Each keeps the oracle's share to well under a minute on the GPU host's cores.
"""
import sys

import numpy as np
import pytest

import cholla_amd as chl
from helpers import KH_FP64_MISS, check_fp64, strict_l1
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    if chl.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need a real MI355X")


@pytest.mark.parametrize("n,steps", [(128, 100), (256, 10)])
def test_kelvin_helmholtz_exact_at_config_size(n, steps):
    ng = 4
    N = n + 2 * ng
    h = 1.0 / n
    U = chl.ic_kelvin_helmholtz(n, ng, G, seed=2024)
    say(f"KH {n}^3 x{steps} exact vs oracle")
    Uo = U.copy()
    steps_o, t_o = orc.run(Uo, N, N, N, ng, 5, h, h, h, G, orc.PPMC, [1] * 6, 0.0, 1e30, steps)
    D = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, chl.PREC_FP64_EXACT, tile1=1, tile2=4, kchunk1=16, kchunk2=128)
    s, t = hy.run(D, h, h, h, G, 0.0, 1e30, steps, use_graph=True)
    hy.close()
    assert s == steps_o == steps and t == t_o
    A = D.download()
    assert np.array_equal(A, Uo), f"max|diff| {np.abs(A - Uo).max()}"
    if n == 256:
        D.upload(U)
        hf = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, chl.PREC_FP64)
        hf.run(D, h, h, h, G, 0.0, 1e30, steps, use_graph=True)
        hf.close()
        l1 = strict_l1(D.download(), Uo, 5, N, N, N, ng)
        say("  fp64 form strict L1", ["%.2e" % v for v in l1])
        check_fp64(l1, KH_FP64_MISS)
    D.free()


def test_orszag_tang_exact_at_64():
    n, ng = 64, 4
    N = n + 2 * ng
    h = 1.0 / n
    U = chl.ic_orszag_tang(n, ng, G, seed=2024, amplitude=1e-3)
    say("Orszag-Tang 64^3 x50 exact vs oracle")
    Uo = U.copy()
    steps_o, t_o = orc.mhd_run(Uo, N, N, N, ng, h, h, h, G, [1] * 6, 0.0, 1e30, 50)
    D = chl.DeviceArray.from_host(U)
    hm = chl.Mhd(N, N, N, ng, chl.PREC_FP64_EXACT)
    s, t = hm.run(D, h, h, h, G, 0.0, 1e30, 50, use_graph=True)
    hm.close()
    assert s == steps_o == 50 and t == t_o
    A = D.download()
    assert np.array_equal(A, Uo), f"max|diff| {np.abs(A - Uo).max()}"
    div = chl.mhd_divergence(D, N, N, N, h, h, h)
    assert div < 1e-11 * np.abs(A.reshape(8, N, N, N)[5:]).max() / h


@pytest.mark.parametrize("launch", [chl.LAUNCH_SPLIT, chl.LAUNCH_PERSISTENT])
def test_turbulence_ensemble_with_a_128_box(launch):
    ng, steps = 4, 20
    edges, machs = [128, 48, 32, 96], [1.5, 3.0, 0.5, 2.2]
    dims = [(e + 2 * ng,) * 3 for e in edges]
    dxyz = [(1.0 / e,) * 3 for e in edges]
    dev = []
    for b, (e, m) in enumerate(zip(edges, machs)):
        D = chl.DeviceArray(5 * (e + 2 * ng) ** 3)
        chl.ic_turbulence(D, e, ng, G, 1.0, 1.0, m, seed=40 + b)
        dev.append(D)
    states = [D.download() for D in dev]
    say("turbulence ensemble", edges, "launch", launch)
    ens = chl.Ensemble(dims, ng, chl.RECON_PPMC, chl.PREC_FP64_EXACT, launch, tile1=0, tile2=0, kchunk1=32,
                       kchunk2=32, positivity_fallback=True)
    ens.bind(dev, np.array(dxyz).reshape(-1))
    ens.start(G)
    ens.advance(steps)
    t, _, ns = ens.clocks()
    ens.close()
    for b, U in enumerate(states):
        N = dims[b][0]
        Uo = U.copy()
        n_o, t_o = orc.run(Uo, N, N, N, ng, 5, *dxyz[b], G, orc.PPMC, [1] * 6, 0.0, 1e30, steps,
                           positivity_fallback_on=True)
        assert ns[b] == n_o == steps and t[b] == t_o
        A = dev[b].download()
        assert np.array_equal(A, Uo), (b, float(np.abs(A - Uo).max()))
