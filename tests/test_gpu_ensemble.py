"""GPU parity of the ensemble integrator (chl_ensemble_*) and of the turbulence IC.

Bars:
  * PolicyExact ensemble, split and persistent launch: every box bit-identical to the
    oracle running that box alone (main.cpp's loop with periodic boundaries), and the
    box clocks equal to the oracle's times;
  * any policy: the split launch, the persistent launch and the per-box launch loop
    give bit-identical states and clocks (the same tile arithmetic in each);
  * mixed ensemble: per-field relative L1 <= 1e-5 against the FP64 oracle;
  * turbulence IC: the numpy evaluation of the same mode sum to 1e-12, rms Mach number
    as requested to 1e-12, ghost cells exact periodic images.
"""
import sys

import numpy as np
import pytest

import cholla_amd as chl
from helpers import check_mixed, interior, rel_l1
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    if chl.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need a real MI355X")


def periodic_box(nxr, nyr, nzr, ng, seed, g=5 / 3, mach=0.8):
    """Smooth + noisy periodic state of nxr x nyr x nzr real cells, ghosts filled."""
    nx, ny, nz = nxr + 2 * ng, nyr + 2 * ng, nzr + 2 * ng
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) - ng + 0.5) / nxr
    y = (np.arange(ny) - ng + 0.5) / nyr
    d = 1.0 + 0.3 * np.sin(2 * np.pi * x)[None, None, :] * np.cos(2 * np.pi * y)[None, :, None]
    d = np.broadcast_to(d, (nz, ny, nx)).copy() + 0.05 * rng.uniform(-1, 1, (nz, ny, nx))
    v = mach * rng.uniform(-1, 1, (3, nz, ny, nx))
    p = 1.0 + 0.1 * rng.uniform(-1, 1, (nz, ny, nx))
    U = np.stack([d, d * v[0], d * v[1], d * v[2], p / (g - 1) + 0.5 * d * (v ** 2).sum(0)]).reshape(-1).copy()
    orc.set_boundaries(U, nx, ny, nz, 5, ng, [1] * 6)
    return U, (nx, ny, nz), (1.0 / nxr, 1.0 / nyr, 1.0 / nzr)


REAL = [(12, 12, 12), (20, 8, 14), (8, 16, 10), (16, 12, 24), (6, 9, 7)]


def make_boxes(ng, seed0=100):
    boxes = [periodic_box(*r, ng, seed0 + b) for b, r in enumerate(REAL)]
    return [b[0] for b in boxes], [b[1] for b in boxes], [b[2] for b in boxes]


def run_ensemble(states, dims, dxyz, ng, steps, precision, launch, recon=chl.RECON_PPMC, looped=False,
                 use_graph=True, tout=None, **kw):
    ens = chl.Ensemble(dims, ng, recon, precision, launch, **kw)
    dev = [chl.DeviceArray.from_host(U) for U in states]
    ens.bind(dev, np.array(dxyz).reshape(-1))
    ens.start(5 / 3, tout=tout)
    if looped:
        ens.advance_looped(steps)
    else:
        ens.advance(steps, use_graph=use_graph)
    t, dt, ns = ens.clocks()
    out = [d.download() for d in dev]
    ens.close()
    return out, t, ns


@pytest.mark.parametrize("launch", [chl.LAUNCH_SPLIT, chl.LAUNCH_PERSISTENT])
@pytest.mark.parametrize("recon,ng", [(chl.RECON_PPMC, 4), (chl.RECON_PLMC, 3)])
def test_exact_ensemble_matches_oracle_per_box(launch, recon, ng):
    steps = 6
    states, dims, dxyz = make_boxes(ng)
    say("exact ensemble vs oracle, launch", launch, "recon", recon, "boxes", dims)
    out, t, ns = run_ensemble(states, dims, dxyz, ng, steps, chl.PREC_FP64_EXACT, launch, recon)
    for b, U in enumerate(states):
        Uo = U.copy()
        nx, ny, nz = dims[b]
        n_o, t_o = orc.run(Uo, nx, ny, nz, ng, 5, *dxyz[b], 5 / 3, recon, [1] * 6, 0.0, 1e30, steps)
        assert ns[b] == n_o == steps
        assert t[b] == t_o, (b, t[b], t_o)
        diff = float(np.abs(out[b] - Uo).max())
        say("  box", b, dims[b], "max|diff| %.3e" % diff)
        np.testing.assert_array_equal(out[b], Uo)


ENS_OPTIONS = [  # (reconstruction, riemann, n_ghost, oracle reconstruction, oracle solver)
    (chl.RECON_PPMC, chl.RIEMANN_HLL, 4, orc.PPMC, orc.HLL),
    (chl.RECON_PPMC, chl.RIEMANN_ROE, 4, orc.PPMC, orc.ROE),
    (chl.RECON_PLMC, chl.RIEMANN_EXACT, 3, orc.PLMC, orc.EXACT),
    (chl.RECON_PCM, chl.RIEMANN_HLLC, 2, orc.PCM, orc.HLLC),
    (chl.RECON_PLMP, chl.RIEMANN_HLLC, 3, orc.PLMP, orc.HLLC),
    (chl.RECON_PPMP, chl.RIEMANN_HLL, 4, orc.PPMP, orc.HLL),
]


@pytest.mark.parametrize("looped", [False, True])
@pytest.mark.parametrize("opt", range(len(ENS_OPTIONS)))
def test_exact_ensemble_options_match_oracle_per_box(opt, looped):
    """The make.type options inside the ensemble (split launch and the per-box loop): the HLL,
    Roe and exact solvers and the PCM, PLMP and PPMP reconstructions, each box bit-identical to
    the oracle's loop with the same options."""
    recon, riemann, ng, orc_recon, orc_solver = ENS_OPTIONS[opt]
    steps = 5
    states, dims, dxyz = make_boxes(ng, seed0=300 + opt)
    say("exact ensemble options", ENS_OPTIONS[opt][:3], "looped", looped)
    out, t, ns = run_ensemble(states, dims, dxyz, ng, steps, chl.PREC_FP64_EXACT, chl.LAUNCH_SPLIT, recon,
                              looped=looped, riemann=riemann)
    for b, U in enumerate(states):
        Uo = U.copy()
        nx, ny, nz = dims[b]
        n_o, t_o = orc.run(Uo, nx, ny, nz, ng, 5, *dxyz[b], 5 / 3, orc_recon, [1] * 6, 0.0, 1e30, steps,
                           solver=orc_solver)
        assert ns[b] == n_o == steps and t[b] == t_o, (b, t[b], t_o)
        np.testing.assert_array_equal(out[b], Uo)


def test_ensemble_option_limits():
    dims = [(20, 20, 20)]
    with pytest.raises(chl.ChollaError, match="split launch"):
        chl.Ensemble(dims, 4, chl.RECON_PPMC, chl.PREC_FP64, chl.LAUNCH_PERSISTENT, riemann=chl.RIEMANN_ROE,
                     tile1=0, tile2=0)
    with pytest.raises(chl.ChollaError, match="tile shapes 0 and 1"):
        chl.Ensemble(dims, 4, chl.RECON_PPMP, chl.PREC_FP64, chl.LAUNCH_SPLIT, tile1=2, tile2=1)


@pytest.mark.parametrize("recon,orc_recon,ng", [(chl.RECON_PCM, orc.PCM, 2), (chl.RECON_PLMP, orc.PLMP, 3),
                                               (chl.RECON_PPMP, orc.PPMP, 4)])
def test_pcm_plmp_ppmp_de_ensemble_matches_oracle_per_box(recon, orc_recon, ng):
    """PCM, PLMP and PPMP with dual energy inside the ensemble (split launch), each box
    bit-identical to the oracle's DE loop."""
    from test_oracle_de import periodic_state
    steps = 5
    states, dims, dxyz = [], [], []
    for b, r in enumerate(REAL[:3]):
        rng = np.random.default_rng(500 + b)
        shape = (r[2], r[1], r[0])
        d = 1 + 0.1 * rng.uniform(-1, 1, shape)
        v = 0.5 * rng.uniform(-1, 1, (3,) + shape)
        p = 1 + 0.1 * rng.uniform(-1, 1, shape)
        U, dm = periodic_state(r[0], r[1], r[2], ng, d, v[0], v[1], v[2], p)
        states.append(U), dims.append(dm), dxyz.append((1.0 / r[0], 1.0 / r[1], 1.0 / r[2]))
    say("DE ensemble recon", recon)
    out, t, ns = run_ensemble(states, dims, dxyz, ng, steps, chl.PREC_FP64_EXACT, chl.LAUNCH_SPLIT, recon,
                              dual_energy=True)
    for b, U in enumerate(states):
        Uo = U.copy()
        nx, ny, nz = dims[b]
        n_o, t_o = orc.run(Uo, nx, ny, nz, ng, 6, *dxyz[b], 5 / 3, orc_recon, [1] * 6, 0.0, 1e30, steps,
                           dual_energy_on=True)
        assert ns[b] == n_o == steps and t[b] == t_o
        np.testing.assert_array_equal(out[b], Uo)
    # a tile left to the library lands on a shape the option has kernels for
    ens = chl.Ensemble(dims, 4, chl.RECON_PPMC, chl.PREC_FP64, chl.LAUNCH_SPLIT, riemann=chl.RIEMANN_HLL)
    assert ens.config().tile1 <= 1 and ens.config().tile2 <= 1
    ens.close()


@pytest.mark.parametrize("precision", [chl.PREC_FP64, chl.PREC_MIXED])
def test_launch_forms_identical(precision):
    """Split, persistent (every compiled tile pair) and the per-box loop agree bit for bit."""
    ng, steps = 4, 5
    states, dims, dxyz = make_boxes(ng, seed0=300)
    say("launch forms, precision", precision)
    ref, t_ref, n_ref = run_ensemble(states, dims, dxyz, ng, steps, precision, chl.LAUNCH_SPLIT)
    runs = [("looped", dict(launch=chl.LAUNCH_SPLIT, looped=True)),
            ("split-eager", dict(launch=chl.LAUNCH_SPLIT, use_graph=False))]
    for t1, t2 in chl.persistent_pairs():
        for kc in (8, 32):
            runs.append((f"persistent {t1},{t2} kc{kc}",
                         dict(launch=chl.LAUNCH_PERSISTENT, tile1=t1, tile2=t2, kchunk1=kc, kchunk2=kc)))
    for name, kw in runs:
        say("  form", name)
        launch = kw.pop("launch")
        out, t, ns = run_ensemble(states, dims, dxyz, ng, steps, precision, launch, **kw)
        np.testing.assert_array_equal(ns, n_ref)
        np.testing.assert_array_equal(t, t_ref)
        for b in range(len(states)):
            np.testing.assert_array_equal(out[b], ref[b])


def test_split_tiles_and_chunks_identical():
    ng, steps = 4, 3
    states, dims, dxyz = make_boxes(ng, seed0=500)
    ref, t_ref, _ = run_ensemble(states, dims, dxyz, ng, steps, chl.PREC_FP64_EXACT, chl.LAUNCH_SPLIT)
    for t1, t2, k1, k2 in [(0, 3, 8, 64), (4, 4, 16, 16), (3, 2, 64, 8), (1, 0, 32, 32)]:
        say("split tiles", t1, t2, k1, k2)
        out, t, _ = run_ensemble(states, dims, dxyz, ng, steps, chl.PREC_FP64_EXACT, chl.LAUNCH_SPLIT, tile1=t1,
                                 tile2=t2, kchunk1=k1, kchunk2=k2)
        np.testing.assert_array_equal(t, t_ref)
        for b in range(len(states)):
            np.testing.assert_array_equal(out[b], ref[b])


def test_mixed_ensemble_l1_vs_oracle():
    ng, steps = 4, 10
    states, dims, dxyz = make_boxes(ng, seed0=700)
    say("mixed ensemble L1 vs FP64 oracle")
    out, t, ns = run_ensemble(states, dims, dxyz, ng, steps, chl.PREC_MIXED, chl.LAUNCH_PERSISTENT)
    for b, U in enumerate(states):
        Uo = U.copy()
        nx, ny, nz = dims[b]
        n_o, t_o = orc.run(Uo, nx, ny, nz, ng, 5, *dxyz[b], 5 / 3, orc.PPMC, [1] * 6, 0.0, 1e30, steps)
        assert ns[b] == n_o
        assert abs(t[b] - t_o) <= 1e-5 * t_o
        a, o = interior(out[b], 5, nz, ny, nx, ng), interior(Uo, 5, nz, ny, nx, ng)
        l1 = [rel_l1(a[f], o[f]) for f in range(5)]
        say("  box", b, "strict L1", ["%.2e" % v for v in l1])
        check_mixed(l1)


def test_per_box_output_time():
    """A box stops at its own tout (dt clamped to land on it), the others go on."""
    ng, steps = 4, 8
    states, dims, dxyz = make_boxes(ng, seed0=900)
    _, t_free, _ = run_ensemble(states, dims, dxyz, ng, 3, chl.PREC_FP64_EXACT, chl.LAUNCH_SPLIT)
    tout = np.full(len(states), 1e30)
    tout[1] = 0.5 * t_free[1]  # reached within the first steps
    tout[3] = t_free[3]
    out, t, ns = run_ensemble(states, dims, dxyz, ng, steps, chl.PREC_FP64_EXACT, chl.LAUNCH_PERSISTENT, tout=tout)
    assert t[1] == tout[1] and t[3] == tout[3]
    for b in (1, 3):
        Uo = states[b].copy()
        nx, ny, nz = dims[b]
        n_o, t_o = orc.run(Uo, nx, ny, nz, ng, 5, *dxyz[b], 5 / 3, orc.PPMC, [1] * 6, 0.0, tout[b], steps)
        assert ns[b] == n_o and t[b] == t_o
        np.testing.assert_array_equal(out[b], Uo)
    assert all(ns[b] == steps for b in (0, 2, 4))


def test_autotune_leaves_state_and_clocks():
    ng = 4
    states, dims, dxyz = make_boxes(ng, seed0=1100)
    ens = chl.Ensemble(dims, ng, chl.RECON_PPMC, chl.PREC_MIXED, chl.LAUNCH_SPLIT)
    dev = [chl.DeviceArray.from_host(U) for U in states]
    ens.bind(dev, np.array(dxyz).reshape(-1))
    ens.start(5 / 3)
    ens.advance(2)
    before = [d.download() for d in dev]
    clk_before = ens.clocks()
    say("ensemble autotune")
    cfg = ens.autotune(steps=1)
    say("  chose", cfg.as_dict())
    for b in range(len(states)):
        np.testing.assert_array_equal(dev[b].download(), before[b])
    for x, y in zip(ens.clocks(), clk_before):
        np.testing.assert_array_equal(x, y)
    ens.advance(2)
    _, _, ns = ens.clocks()
    assert (ns == 4).all()
    ens.close()


def test_turbulence_ic_matches_numpy():
    n, ng, g, rho, p, mach, seed = 16, 4, 5 / 3, 1.0, 1.0, 1.7, 42
    N = n + 2 * ng
    U = chl.DeviceArray(5 * N ** 3)
    say("turbulence IC n", n)
    chl.ic_turbulence(U, n, ng, g, rho, p, mach, seed)
    A = U.download().reshape(5, N, N, N)
    modes = chl.turbulence_modes(seed)
    x = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    v = np.zeros((3, n, n, n))
    for m in modes:
        ph = 2 * np.pi * (m[0] * X + m[1] * Y + m[2] * Z)
        c, s = np.cos(ph), np.sin(ph)
        for q in range(3):
            v[q] += m[3 + q] * c + m[6 + q] * s
    rms = np.sqrt((v ** 2).sum(0).mean())
    v *= mach * np.sqrt(g * p / rho) / rms
    real = (slice(ng, N - ng),) * 3
    got = A[(slice(None),) + real]
    assert np.all(got[0] == rho)
    for q in range(3):
        np.testing.assert_allclose(got[1 + q], rho * v[q], rtol=0, atol=1e-12 * np.abs(v).max())
    vg = got[1:4] / rho
    assert abs(np.sqrt((vg ** 2).sum(0).mean()) / np.sqrt(g * p / rho) - mach) <= 1e-12 * mach
    E = p / (g - 1) + 0.5 * rho * (vg ** 2).sum(0)
    np.testing.assert_allclose(got[4], E, rtol=1e-13)
    B = A.reshape(-1).copy()
    orc.set_boundaries(B, N, N, N, 5, ng, [1] * 6)
    np.testing.assert_array_equal(A.reshape(-1), B)


def test_turbulence_ensemble_step():
    """A few boxes of the benchmark layout (edges from the seeded draw, capped at 48 to stay
    small): every box advances, stays finite and conserves mass to round-off."""
    edges, mach = chl.ensemble_layout(7, 6)
    edges = np.minimum(edges, 48)
    ng = 4
    dims = [(int(e) + 2 * ng,) * 3 for e in edges]
    dev = []
    for b, e in enumerate(edges):
        U = chl.DeviceArray(5 * (int(e) + 2 * ng) ** 3)
        chl.ic_turbulence(U, int(e), ng, 5 / 3, 1.0, 1.0, float(mach[b]), 1000 + b)
        dev.append(U)
    mass0 = [interior(d.download(), 5, *dims[b], ng)[0].sum() for b, d in enumerate(dev)]
    ens = chl.Ensemble(dims, ng, chl.RECON_PPMC, chl.PREC_MIXED, chl.LAUNCH_PERSISTENT)
    ens.bind(dev, np.array([[1.0 / e] * 3 for e in edges]).reshape(-1))
    ens.start(5 / 3)
    say("turbulence ensemble edges", list(edges), "mach", ["%.2f" % m for m in mach])
    ens.advance(10)
    t, dt, ns = ens.clocks()
    assert (ns == 10).all() and (t > 0).all()
    for b, d in enumerate(dev):
        A = d.download()
        assert np.isfinite(A).all()
        m = interior(A, 5, *dims[b], ng)[0].sum()
        assert abs(m - mass0[b]) <= 1e-12 * mass0[b], (b, m, mass0[b])
    ens.close()


# ----------------------------------------------------------------------------- dual energy
def de_boxes(ng, seed0=500):
    """Uneven 6-field boxes: uniform density and pressure with a hypersonic slab
    (|y - 0.5| < 0.25 at vx = 100, thermal fraction ~3e-4 < DE_ETA_1) in a slow medium."""
    from test_oracle_de import periodic_state

    states, dims, dxyz = [], [], []
    for b, (nxr, nyr, nzr) in enumerate(REAL):
        rng = np.random.default_rng(seed0 + b)
        z, y, x = np.meshgrid((np.arange(nzr) + 0.5) / nzr, (np.arange(nyr) + 0.5) / nyr, (np.arange(nxr) + 0.5) / nxr,
                              indexing="ij")
        sh = (nzr, nyr, nxr)
        slab = np.abs(y - 0.5) < 0.25
        d = 1 + 0.05 * rng.uniform(-1, 1, sh)
        vx = np.where(slab, 100.0, -0.3) + 0.02 * rng.uniform(-1, 1, sh)
        vy = 0.05 * np.sin(4 * np.pi * x) + 0.02 * rng.uniform(-1, 1, sh)
        vz = 0.02 * rng.uniform(-1, 1, sh)
        p = 1 + 0.05 * rng.uniform(-1, 1, sh)
        U, dm = periodic_state(nxr, nyr, nzr, ng, d, vx, vy, vz, p)
        states.append(U)
        dims.append(dm)
        dxyz.append((1.0 / nxr, 1.0 / nyr, 1.0 / nzr))
    return states, dims, dxyz


@pytest.mark.parametrize("launch", [chl.LAUNCH_SPLIT, chl.LAUNCH_PERSISTENT])
def test_de_exact_ensemble_matches_oracle_per_box(launch):
    ng, steps = 4, 6
    states, dims, dxyz = de_boxes(ng)
    say("DE exact ensemble vs oracle, launch", launch)
    out, t, ns = run_ensemble(states, dims, dxyz, ng, steps, chl.PREC_FP64_EXACT, launch, dual_energy=True)
    for b, U in enumerate(states):
        Uo = U.copy()
        nx, ny, nz = dims[b]
        n_o, t_o = orc.run(Uo, nx, ny, nz, ng, 6, *dxyz[b], 5 / 3, orc.PPMC, [1] * 6, 0.0, 1e30, steps,
                           dual_energy_on=True)
        assert ns[b] == n_o == steps
        assert t[b] == t_o, (b, t[b], t_o)
        np.testing.assert_array_equal(out[b], Uo)


@pytest.mark.parametrize("precision", [chl.PREC_FP64, chl.PREC_MIXED])
def test_de_launch_forms_identical(precision):
    """With dual energy: split, persistent (the dual-energy tile pairs) and the per-box loop
    agree bit for bit."""
    ng, steps = 4, 5
    states, dims, dxyz = de_boxes(ng, seed0=700)
    say("DE launch forms, precision", precision)
    ref, t_ref, n_ref = run_ensemble(states, dims, dxyz, ng, steps, precision, chl.LAUNCH_SPLIT, dual_energy=True)
    runs = [("looped", dict(launch=chl.LAUNCH_SPLIT, looped=True)),
            ("split-eager", dict(launch=chl.LAUNCH_SPLIT, use_graph=False)),
            ("split 0,0", dict(launch=chl.LAUNCH_SPLIT, tile1=0, tile2=0))]
    for t1, t2 in ((0, 0), (1, 1)):
        for kc in (8, 32):
            runs.append((f"persistent {t1},{t2} kc{kc}",
                         dict(launch=chl.LAUNCH_PERSISTENT, tile1=t1, tile2=t2, kchunk1=kc, kchunk2=kc)))
    for name, kw in runs:
        say("  form", name)
        launch = kw.pop("launch")
        out, t, ns = run_ensemble(states, dims, dxyz, ng, steps, precision, launch, dual_energy=True, **kw)
        np.testing.assert_array_equal(ns, n_ref)
        np.testing.assert_array_equal(t, t_ref)
        for b in range(len(states)):
            np.testing.assert_array_equal(out[b], ref[b])


@pytest.mark.parametrize("precision", [chl.PREC_FP64, chl.PREC_MIXED])
def test_de_turbulence_ensemble_stays_finite(precision):
    """The benchmark's hardest boxes (32^3 at rms Mach 2.5-2.9) with dual energy, the
    PPMC positivity fallback and the benchmark's temperature floor, persistent launch:
    finite through 120 steps with a steady dt, mass conserved to round-off.  (Without
    the fallback box 258 collapses its dt within ~45 steps in the FP64 oracle as well:
    a floored 1e-20 interface density next to a finite pressure makes an HLLC flux that
    injects energy, which the explicit pdV source then amplifies.)"""
    edges, mach = chl.ensemble_layout(2024, 384)
    pick = [372, 76, 36, 200, 258, 284]
    ng, steps = 4, 120
    dims = [(int(edges[b]) + 2 * ng,) * 3 for b in pick]
    dev = []
    for q, b in enumerate(pick):
        U = chl.DeviceArray(6 * dims[q][0] ** 3)
        chl.ic_turbulence(U, int(edges[b]), ng, 5 / 3, 1.0, 1.0, float(mach[b]), 2024 * 1000 + b)
        dev.append(U)
    mass0 = [interior(d.download(), 6, *dims[q], ng)[0].sum() for q, d in enumerate(dev)]
    say("DE turbulence ensemble precision", precision, "steps", steps)
    ens = chl.Ensemble(dims, ng, chl.RECON_PPMC, precision, chl.LAUNCH_PERSISTENT, dual_energy=True,
                       positivity_fallback=True)
    ens.bind(dev, np.array([[1.0 / edges[b]] * 3 for b in pick]).reshape(-1))
    ens.start(5 / 3, U_floor=1e-3)
    ens.advance(1)
    _, dt1, _ = ens.clocks()
    ens.advance(steps - 1)
    t, dt, ns = ens.clocks()
    assert (ns == steps).all()
    assert (dt > 0.3 * dt1).all(), list(zip(pick, dt / dt1))
    for q, d in enumerate(dev):
        A = d.download()
        assert np.isfinite(A).all(), pick[q]
        a = interior(A, 6, *dims[q], ng)
        assert a[0].min() > 0 and a[5].min() > 0, pick[q]
        assert abs(a[0].sum() - mass0[q]) <= 1e-12 * mass0[q], (pick[q], a[0].sum(), mass0[q])
    ens.close()
