#!/usr/bin/env python3
"""Benchmark of the hydro hot path on one MI355X: cell-updates/s of the VL
PPMC+HLLC step on the 512^3 Kelvin-Helmholtz box, mixed FP32/FP64 form against
the FP64 forms, with the per-field L1 of the mixed form against FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--forms mixed,fp64,fp64_exact]
    python bench.py --workload ensemble [--boxes 384] [--steps 100]   # batched-ensemble line
    python bench.py --workload mhd [--n 256] [--steps 150]            # ideal-MHD line

One step = one full VL step (stage 1 + stage 2 with fused floors, dt reduction
and periodic ghost fill, plus the device-side dt update), graph-captured, state
resident in HBM.  Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
BYTES_PER_CELL_UPDATE = 80  # FP64 state (5 fields) read once + written once
L1_BAR = 1e-5               # the mixed form's per-field relative L1 bar against FP64
# HBM bytes per launch of the dominant kernel measured with rocprofv3 --pmc FETCH_SIZE /
# WRITE_SIZE (scaled by a plain copy of known size); written next to its summary.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "stage2_kh512_hbm_traffic.json")


def measured_traffic(form, tile, n):
    """The profiled HBM bytes per launch of stage 2 for this form, stage-2 tile and z-chunk
    (tile = [tile1, tile2, kchunk1, kchunk2]; stage 2's traffic depends on its own launch
    shape only) and box, if the committed profile covers that configuration (else None)."""
    try:
        with open(TRAFFIC_FILE) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None
    for rec in prof.get("records", []):
        rt = list(rec.get("tile", []))
        if rec.get("form") == form and rec.get("n") == n and tile and len(rt) == 4 and \
                (rt[1], rt[3]) == (tile[1], tile[3]):
            return rec
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=512, help="real cells per axis")
    p.add_argument("--forms", default="mixed,fp64,fp64_exact")
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--no-autotune", action="store_true")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--kernel-steps", type=int, default=10)
    p.add_argument("--tiles", default="", help="fixed tile1,tile2,kchunk1,kchunk2 (skips the autotuner)")
    p.add_argument("--workload", default="kh", choices=["kh", "ensemble", "mhd", "snapshot"])
    p.add_argument("--boxes", type=int, default=384, help="ensemble: number of boxes")
    p.add_argument("--looped-steps", type=int, default=5, help="ensemble: steps of the per-box loop baseline")
    p.add_argument("--u-floor", type=float, default=1e-3,
                   help="ensemble: temperature floor as specific internal energy (initial value 1.5)")
    p.add_argument("--dual-energy", action="store_true", help="ensemble: run with the DE build option")
    p.add_argument("--no-positivity-fallback", action="store_true",
                   help="ensemble: plain PPMC (the hardest Mach-3 boxes then collapse their dt and go non-finite)")
    p.add_argument("--launch", default="auto", choices=["auto", "split", "persistent"],
                   help="ensemble: launch form (auto = autotuned)")
    p.add_argument("--calibrate-copy", type=float, default=0.0,
                   help="only run a plain device copy of this many GiB (HBM counter calibration) and exit")
    return p.parse_args()


def calibrate_copy(gib):
    """A plain copy kernel of known size, for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE."""
    import torch

    n = int(gib * (1 << 30)) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copied {n * 8} bytes x3")


def physical_cores():
    """Physical cores of the host as lscpu counts them (unique core, socket pairs), or None."""
    import subprocess

    try:
        out = subprocess.run(["lscpu", "-p=Core,Socket"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    pairs = {tuple(l.split(",")) for l in out.splitlines() if l and not l.startswith("#")}
    return len(pairs) or None


def cpu_baseline(samples=((512, 2), (256, 10))):
    """The oracle (the C++ restatement of the reference's functions, OpenMP over the host's
    threads) on bounded samples of the same workload: KH 512^3 x 2 steps (the value) and
    256^3 x 10 steps, each after one untimed step (allocation and page-in)."""
    from oracle import oracle as orc
    import cholla_amd as chl

    ng, g = 4, 5.0 / 3.0
    threads = orc.lib().orc_num_threads()
    runs = []
    for n_sample, steps in samples:
        N = n_sample + 2 * ng
        U = chl.ic_kelvin_helmholtz(n_sample, ng, g, seed=7)
        h = 1.0 / n_sample
        orc.run(U, N, N, N, ng, 5, h, h, h, g, orc.PPMC, [1] * 6, 0.0, 1e30, 1)
        t0 = time.perf_counter()
        n_done, _ = orc.run(U, N, N, N, ng, 5, h, h, h, g, orc.PPMC, [1] * 6, 0.0, 1e30, steps)
        sec = time.perf_counter() - t0
        runs.append((n_sample, n_done, sec, n_sample ** 3 * n_done / sec))
        log(f"cpu baseline {n_sample}^3 x {n_done}: {sec:.1f} s, {runs[-1][3] / 1e6:.2f} M cell-updates/s")
        del U
    n0, s0, t0_, r0 = runs[0]
    return {"value": r0, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "lscpu_physical_cores": physical_cores(),
            "sample": "; ".join(f"KH {n}^3 PPMC+HLLC FP64, {k} steps in {t:.1f} s = {r / 1e6:.2f} M cell-updates/s"
                                for n, k, t, r in runs) + f" (oracle, OpenMP, {threads} threads)",
            "samples": [{"n": n, "steps": k, "seconds": t, "cell_updates_per_s": r} for n, k, t, r in runs]}


def bench_ensemble(args):
    """Batched ensemble: boxes of edges from {32..128} and rms Mach from U[0.5, 3] (seeded),
    periodic turbulence, PPMC+HLLC with the dual-energy option, per-box CFL dt; the fused
    launch (autotuned form and tiles, one graph replay per step) against a loop of per-box
    launches, and the mixed form's per-field L1 against the FP64 form after the same steps.

    The boxes run the reference's default hydro build options (density and temperature
    floors, no dual energy) with the PPMC positivity fallback on.  Without the fallback
    the reference algorithm itself loses the hardest boxes (32^3 at rms Mach 2.5-3): PPMC
    drives a pair of interface densities below zero next to a finite pressure, the
    TINY_NUMBER floors turn that into a sound speed of ~1e9, and the HLLC flux injects
    energy that collapses the box's dt by 1e-10 and then leaves it non-finite (the FP64
    oracle shows it on box 284 at step 22).  --no-positivity-fallback runs the plain
    scheme and reports the boxes it loses."""
    import cholla_amd as chl

    ng, g = 4, 5.0 / 3.0
    de = args.dual_energy
    pf = not args.no_positivity_fallback
    nf = 6 if de else 5
    steps = args.steps if args.steps != 200 else 100
    edges, mach = chl.ensemble_layout(args.seed, args.boxes)
    dims = [(int(e) + 2 * ng,) * 3 for e in edges]
    dxyz = np.array([[1.0 / e] * 3 for e in edges]).reshape(-1)
    cells = int(sum(int(e) ** 3 for e in edges))
    t0 = time.perf_counter()
    init = []
    for b, e in enumerate(edges):
        U = chl.DeviceArray(nf * (int(e) + 2 * ng) ** 3)
        chl.ic_turbulence(U, int(e), ng, g, 1.0, 1.0, float(mach[b]), args.seed * 1000 + b)
        init.append(U)
    log(f"ensemble: {args.boxes} boxes, {cells / 1e6:.1f} M real cells, ICs on device in "
        f"{time.perf_counter() - t0:.1f} s; edge counts "
        + str({int(v): int((edges == v).sum()) for v in np.unique(edges)}))

    def copies():
        out = []
        for U in init:
            c = chl.DeviceArray(U.n)
            chl.check(chl.lib().chl_memcpy_dtod(c.ptr, U.ptr, U.nbytes))
            out.append(c)
        return out

    launch = {"auto": chl.LAUNCH_SPLIT, "split": chl.LAUNCH_SPLIT, "persistent": chl.LAUNCH_PERSISTENT}[args.launch]
    tiles = [int(x) for x in args.tiles.split(",")] if args.tiles else [-1, -1, 0, 0]
    res, finals = {}, {}
    for form in [f.strip() for f in args.forms.split(",") if f.strip() in ("mixed", "fp64")]:
        prec = chl.PRECISIONS[form]
        states = copies()
        ens = chl.Ensemble(dims, ng, chl.RECON_PPMC, prec, launch, tile1=tiles[0], tile2=tiles[1],
                           kchunk1=tiles[2], kchunk2=tiles[3], dual_energy=de, positivity_fallback=pf)
        ens.bind(states, dxyz)
        ens.start(g, U_floor=args.u_floor)
        if args.launch == "auto" and not args.tiles and not args.no_autotune:
            t1 = time.perf_counter()
            cfg = ens.autotune(steps=2)
            log(f"[{form}] ensemble autotune: {cfg.as_dict()} ({time.perf_counter() - t1:.1f} s)")
        cfg = ens.config().as_dict()
        ens.advance(args.warmup)
        chl.check(chl.lib().chl_synchronize())
        w0 = time.perf_counter()
        ms = ens.advance(steps)
        wall = time.perf_counter() - w0
        t_box, dt_box, ns = ens.clocks()
        rate = cells * steps / max(ms * 1e-3, wall)
        r = {"cell_updates_per_s": rate, "ms_per_step": max(ms, 1e3 * wall) / steps, "config": cfg,
             "steps_per_box": [int(ns.min()), int(ns.max())]}
        if form == "mixed" or "mixed" not in args.forms:
            # the per-box loop baseline on the same kernels, continuing from the same state
            lms = ens.advance_looped(args.looped_steps)
            r["looped_cell_updates_per_s"] = cells * args.looped_steps / (lms * 1e-3)
            r["looped_ms_per_step"] = lms / args.looped_steps
            r["fused_vs_looped"] = rate / r["looped_cell_updates_per_s"]
            # and the same number of fused steps again (the L1 comparison below needs both
            # forms at equal step counts: the other forms take 2 x looped_steps fused steps)
            ens.advance(args.looped_steps)
        else:
            ens.advance(2 * args.looped_steps)
        log(f"[{form}] ensemble {rate / 1e9:.3f} G cell-updates/s ({r['ms_per_step']:.2f} ms/step)"
            + (f", looped {r['looped_cell_updates_per_s'] / 1e9:.3f} G ({r['looped_ms_per_step']:.2f} ms/step)"
               if "looped_cell_updates_per_s" in r else ""))
        # a box whose state stopped being finite invalidates the form's measurement
        r["non_finite_boxes"] = [b for b, U in enumerate(states) if not np.isfinite(U.download()).all()]
        if r["non_finite_boxes"]:
            log(f"[{form}] WARNING: {len(r['non_finite_boxes'])} boxes not finite after the run: "
                f"{r['non_finite_boxes'][:12]}")
        res[form] = r
        finals[form] = (states, ens)
    if args.launch == "auto" and "mixed" in res and res["mixed"]["config"]["launch"] != chl.LAUNCH_PERSISTENT:
        # the persistent launch (one kernel per step over a work queue), timed beside the
        # autotuner's pick on a fresh copy of the initial boxes
        states = copies()
        sc = res["mixed"]["config"]  # the split pick's tiles and z-chunks, where that pair is compiled
        pair = [sc["tile1"], sc["tile2"]] in [list(p) for p in chl.persistent_pairs()]
        kw = dict(tile1=sc["tile1"], tile2=sc["tile2"], kchunk1=sc["kchunk1"], kchunk2=sc["kchunk2"]) if pair else {}
        ens_p = chl.Ensemble(dims, ng, chl.RECON_PPMC, chl.PREC_MIXED, chl.LAUNCH_PERSISTENT, dual_energy=de,
                             positivity_fallback=pf, **kw)
        ens_p.bind(states, dxyz)
        ens_p.start(g, U_floor=args.u_floor)
        ens_p.advance(args.warmup)
        chl.check(chl.lib().chl_synchronize())
        pst = max(steps // 5, 10)
        w0 = time.perf_counter()
        ms = ens_p.advance(pst)
        wall = time.perf_counter() - w0
        prate = cells * pst / max(ms * 1e-3, wall)
        res["mixed"]["persistent"] = {"cell_updates_per_s": prate, "ms_per_step": max(ms, 1e3 * wall) / pst,
                                      "steps": pst, "config": ens_p.config().as_dict()}
        log(f"[mixed] persistent launch {prate / 1e9:.3f} G cell-updates/s ({max(ms, 1e3 * wall) / pst:.2f} ms/step)")
        ens_p.close()
        for U in states:
            U.free()
    l1 = None
    if "mixed" in finals and "fp64" in finals:
        names = ["density", "momentum_x", "momentum_y", "momentum_z", "Energy", "GasEnergy"][:nf]
        num, den = np.zeros(nf), np.zeros(nf)
        skip = set(res["mixed"]["non_finite_boxes"]) | set(res["fp64"]["non_finite_boxes"])
        for b, e in enumerate(edges):
            if b in skip:
                continue
            n = int(e)
            a = finals["mixed"][0][b].download().reshape(nf, n + 2 * ng, n + 2 * ng, n + 2 * ng)[:, ng:-ng, ng:-ng, ng:-ng]
            o = finals["fp64"][0][b].download().reshape(nf, n + 2 * ng, n + 2 * ng, n + 2 * ng)[:, ng:-ng, ng:-ng, ng:-ng]
            num += np.abs(a - o).reshape(nf, -1).sum(1)
            den += np.abs(o).reshape(nf, -1).sum(1)
        l1 = {"vs": "fp64", "boxes_compared": len(edges) - len(skip), "bar": L1_BAR,
              "strict": {names[f]: float(num[f] / den[f]) for f in range(nf)}}
        l1["fields_over_bar"] = [k for k, v in l1["strict"].items() if not v <= L1_BAR]
        log("ensemble mixed vs fp64 L1:", json.dumps(l1))
    # the line's value: the mixed form if every box stayed finite, else the FP64 form
    valid = [f for f in ("mixed", "fp64") if f in res and not res[f]["non_finite_boxes"]]
    head = valid[0] if valid else ("mixed" if "mixed" in res else next(iter(res)))
    r = res[head]
    out = {
        "metric": f"aggregate cell-updates/s, batched ensemble of {args.boxes} periodic turbulence boxes, "
                  "PPMC+HLLC" + (" with dual energy" if de else "") + ", per-box CFL dt",
        "value": r["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if head == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": f"ensemble {args.boxes} boxes, edges U{{32,48,64,96,128}}, Mach U[0.5,3], "
                               f"solenoidal modes 1<=|k|<=4 with E(k)~k^-2, seed {args.seed}, form={head}, "
                               f"dual energy {'on' if de else 'off'}, PPMC positivity fallback "
                               f"{'on' if pf else 'off'}, temperature floor U={args.u_floor:g} "
                               "(initial 1.5)",
                   "cells": cells, "n_ghost": ng, "gamma": g, "n_fields": nf},
        "step_roofline_frac": r["cell_updates_per_s"] * 16 * nf / 1e9 / HBM_PEAK_GBS,
        "forms": res, "l1_mixed_vs_fp64": l1,
    }
    print(json.dumps(out), flush=True)


def bench_mhd(args):
    """Ideal MHD, Orszag-Tang vortex with the seeded z perturbation (1e-3), n^3 periodic,
    PLMC + HLLD + CT: cell-updates/s of each form, the mixed form's L1 against FP64 and
    max |div B| after the timed steps."""
    import cholla_amd as chl

    n = args.n if args.n != 512 else 256
    steps = args.steps if args.steps != 200 else 150
    ng, g = 4, 5.0 / 3.0
    N = n + 2 * ng
    h = 1.0 / n
    cells = n ** 3
    t0 = time.perf_counter()
    U0 = chl.ic_orszag_tang(n, ng, g, seed=args.seed, amplitude=1e-3)
    log(f"Orszag-Tang {n}^3 initial state built on host in {time.perf_counter() - t0:.1f} s")
    D0 = chl.DeviceArray.from_host(U0)
    del U0
    res, finals = {}, {}
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    for form in forms:
        U = chl.DeviceArray(D0.n)
        chl.check(chl.lib().chl_memcpy_dtod(U.ptr, D0.ptr, D0.nbytes))
        hm = chl.Mhd(N, N, N, ng, chl.PRECISIONS[form])
        hm.time_steps(U, h, h, h, g, args.warmup, use_graph=True)
        chl.check(chl.lib().chl_synchronize())
        w0 = time.perf_counter()
        ms, dt_last = hm.time_steps(U, h, h, h, g, steps, use_graph=True)
        chl.check(chl.lib().chl_synchronize())
        wall = time.perf_counter() - w0
        rate = cells * steps / max(ms * 1e-3, wall)
        div = chl.mhd_divergence(U, N, N, N, h, h, h)
        res[form] = {"cell_updates_per_s": rate, "ms_per_step": max(ms, 1e3 * wall) / steps, "max_div_B": div,
                     "dt_last": dt_last}
        log(f"[mhd {form}] {rate / 1e9:.3f} G cell-updates/s, {res[form]['ms_per_step']:.2f} ms/step, "
            f"max|div B| {div:.2e}")
        finals[form] = U
        hm.close()
    l1 = None
    ref_form = "fp64" if "fp64" in finals else ("fp64_exact" if "fp64_exact" in finals else None)
    if "mixed" in finals and ref_form:
        a = finals["mixed"].download().reshape(8, N, N, N)[:, ng:-ng, ng:-ng, ng:-ng]
        b = finals[ref_form].download().reshape(8, N, N, N)[:, ng:-ng, ng:-ng, ng:-ng]
        names = ["density", "momentum_x", "momentum_y", "momentum_z", "Energy", "magnetic_x", "magnetic_y",
                 "magnetic_z"]
        l1 = {"vs": ref_form, "steps": args.warmup + steps, "bar": L1_BAR, "strict": {}}
        for f in range(8):
            l1["strict"][names[f]] = float(np.abs(a[f] - b[f]).sum() / np.abs(b[f]).sum())
        l1["fields_over_bar"] = [k for k, v in l1["strict"].items() if not v <= L1_BAR]
        del a, b
        log("mhd mixed vs", ref_form, "L1:", json.dumps(l1))
    head = "mixed" if "mixed" in res else forms[0]
    r = res[head]
    out = {
        "metric": f"cell-updates/s, ideal MHD Orszag-Tang {n}^3 PLMC+HLLD+CT",
        "value": r["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if head == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": f"Orszag-Tang vortex {n}^3 periodic + 1e-3 z perturbation, PLMC+HLLD+CT, VL, "
                               f"CFL 0.3, form={head}", "cells": cells, "n_ghost": ng, "gamma": g, "seed": args.seed},
        "step_roofline_frac": r["cell_updates_per_s"] * 128 / 1e9 / HBM_PEAK_GBS,
        "forms": res, "l1_mixed_vs_fp64": l1,
    }
    if not args.no_cpu:
        try:
            from oracle import oracle as orc

            ns = 64
            Nn = ns + 2 * ng
            Us = chl.ic_orszag_tang(ns, ng, g, seed=args.seed, amplitude=1e-3)
            orc.mhd_run(Us, Nn, Nn, Nn, ng, 1 / ns, 1 / ns, 1 / ns, g, [1] * 6, 0.0, 1e30, 1)
            c0 = time.perf_counter()
            k, _ = orc.mhd_run(Us, Nn, Nn, Nn, ng, 1 / ns, 1 / ns, 1 / ns, g, [1] * 6, 0.0, 1e30, 4)
            sec = time.perf_counter() - c0
            out["cpu_baseline"] = {"value": ns ** 3 * k / sec, "unit": "cell-updates/s",
                                   "cores": orc.lib().orc_num_threads(), "kind": "port",
                                   "sample": f"Orszag-Tang {ns}^3 PLMC+HLLD+CT FP64, {k} steps, {sec:.2f} s "
                                             "(oracle, OpenMP)"}
        except Exception as e:
            log("cpu baseline failed:", e)
            out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def bench_snapshot(args):
    """Snapshot I/O of a KH state: the device transposes between the grid layout and the
    HDF5 dataset order (8 B read + 8 B written per real cell, HBM-bound) at n^3, and a whole
    HDF5 snapshot written and read back (device -> file -> device, the file under $TMPDIR)
    at min(n, 256)^3, checked to restore every real cell."""
    import tempfile

    import cholla_amd as chl
    from cholla_amd import snapshot as snap

    n, ng, g = args.n, 4, 5.0 / 3.0
    N = n + 2 * ng
    U = chl.ic_kelvin_helmholtz(n, ng, g, seed=args.seed)
    D = chl.DeviceArray.from_host(U)
    del U
    buf = chl.DeviceArray(n ** 3)
    L = chl.lib()
    res = {}
    for name, fn, a, b in (("pack", L.chl_pack_real_cells, D.ptr, buf.ptr),
                           ("unpack", L.chl_unpack_real_cells, buf.ptr, D.ptr)):
        chl.check(fn(a, b, N, N, N, ng, n, n, n, -1))
        reps = max(args.kernel_steps, 5)
        t0 = time.perf_counter()
        for _ in range(reps):
            chl.check(fn(a, b, N, N, N, ng, n, n, n, -1))
        sec = (time.perf_counter() - t0) / reps
        res[name] = {"ms": sec * 1e3, "GB_per_s": 16 * n ** 3 / sec / 1e9}
        log(f"{name} {n}^3 field: {sec * 1e3:.3f} ms, {16 * n ** 3 / sec / 1e9:.0f} GB/s (16 B per real cell)")
    buf.free()
    m = min(n, 256)
    M = m + 2 * ng
    V = chl.ic_kelvin_helmholtz(m, ng, g, seed=args.seed)
    S = chl.DeviceArray.from_host(V)
    R = chl.DeviceArray(V.size, fill=np.zeros(V.size))
    h = snap.header(M, M, M, ng, gamma=g, dx=(1.0 / m,) * 3)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "0.h5")
        t0 = time.perf_counter()
        snap.write_hdf5(path, S, h)
        t_w = time.perf_counter() - t0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        snap.read_hdf5(path, R, snap.header(M, M, M, ng))
        t_r = time.perf_counter() - t0
    ok = bool(np.array_equal(R.download().reshape(5, M, M, M)[:, ng:-ng, ng:-ng, ng:-ng],
                             V.reshape(5, M, M, M)[:, ng:-ng, ng:-ng, ng:-ng]))
    log(f"snapshot {m}^3: {size / 1e9:.2f} GB file, write {t_w:.2f} s ({size / t_w / 1e9:.2f} GB/s), "
        f"read {t_r:.2f} s ({size / t_r / 1e9:.2f} GB/s), round trip exact: {ok}")
    out = {
        "metric": f"snapshot layout transpose GB/s ({n}^3 field, 16 B per real cell)",
        "value": res["pack"]["GB_per_s"], "unit": "GB/s", "n_gpus": 1, "steps": max(args.kernel_steps, 5),
        "warmup": 1, "ms_per_step": res["pack"]["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"KH {n}^3 state: pack/unpack of one field; HDF5 snapshot of the {m}^3 state"},
        "roofline": {"bound": "hbm", "achieved": res["pack"]["GB_per_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": res["pack"]["GB_per_s"] / HBM_PEAK_GBS, "traffic": None},
        "kernels": res,
        "file": {"n": m, "bytes": size, "write_s": t_w, "read_s": t_r, "round_trip_exact": ok},
    }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.calibrate_copy > 0:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU layout is a design note, "
              "not a measured configuration.", file=sys.stderr)
        sys.exit(2)
    import cholla_amd as chl

    if args.workload in ("ensemble", "mhd", "snapshot"):
        if chl.device_count() < 1:
            log("no HIP device visible")
            sys.exit(1)
        log("device:", chl.device_name())
        {"ensemble": bench_ensemble, "mhd": bench_mhd, "snapshot": bench_snapshot}[args.workload](args)
        return

    if chl.device_count() < 1:
        log("no HIP device visible")
        sys.exit(1)
    log("device:", chl.device_name())
    n, ng, g = args.n, 4, 5.0 / 3.0
    N = n + 2 * ng
    h = 1.0 / n
    cells = n ** 3
    t0 = time.perf_counter()
    U0 = chl.ic_kelvin_helmholtz(n, ng, g, seed=args.seed)
    log(f"KH {n}^3 initial state built on host in {time.perf_counter() - t0:.1f} s")
    D0 = chl.DeviceArray.from_host(U0)
    del U0
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    res, finals = {}, {}
    for form in forms:
        prec = chl.PRECISIONS[form]
        U = chl.DeviceArray(D0.n)
        chl.check(chl.lib().chl_memcpy_dtod(U.ptr, D0.ptr, D0.nbytes))
        tiles = [int(x) for x in args.tiles.split(",")] if args.tiles else [-1, -1, 0, 0]
        hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, prec, tile1=tiles[0], tile2=tiles[1], kchunk1=tiles[2],
                       kchunk2=tiles[3])
        tune = tuple(tiles) if args.tiles else None
        if not args.no_autotune and not args.tiles:
            t1 = time.perf_counter()
            tune = hy.autotune(U, h, h, h, g, steps=2)
            log(f"[{form}] autotune tile1,tile2,kchunk1,kchunk2 = {tune} ({time.perf_counter() - t1:.1f} s)")
        # warm-up steps advance the state (like the reference's run), then K timed steps continue it
        hy.time_steps(U, h, h, h, g, args.warmup, use_graph=True)
        chl.check(chl.lib().chl_synchronize())
        w0 = time.perf_counter()
        ms, dt_last = hy.time_steps(U, h, h, h, g, args.steps, use_graph=True)
        chl.check(chl.lib().chl_synchronize())
        wall = time.perf_counter() - w0
        s1, s2, so = hy.kernel_times(U, h, h, h, g, args.kernel_steps)
        rate_ev = cells * args.steps / (ms * 1e-3)
        rate_wall = cells * args.steps / wall
        res[form] = {
            "cell_updates_per_s": min(rate_ev, rate_wall), "ms_per_step": max(ms / args.steps, 1e3 * wall / args.steps),
            "event_ms_per_step": ms / args.steps, "wall_ms_per_step": 1e3 * wall / args.steps,
            "stage1_ms": s1, "stage2_ms": s2, "dt_kernel_ms": so, "tile": tune, "dt_last": dt_last,
        }
        log(f"[{form}] {res[form]['cell_updates_per_s'] / 1e9:.3f} G cell-updates/s, "
            f"{res[form]['ms_per_step']:.2f} ms/step (stage1 {s1:.2f} ms, stage2 {s2:.2f} ms)")
        finals[form] = U
        hy.close()
    # mixed vs FP64: per-field relative L1 after the same number of steps from the same state
    l1 = None
    ref_form = "fp64" if "fp64" in finals else ("fp64_exact" if "fp64_exact" in finals else None)
    if "mixed" in finals and ref_form:
        a = finals["mixed"].download().reshape(5, N, N, N)[:, ng:-ng, ng:-ng, ng:-ng]
        b = finals[ref_form].download().reshape(5, N, N, N)[:, ng:-ng, ng:-ng, ng:-ng]
        names = ["density", "momentum_x", "momentum_y", "momentum_z", "Energy"]
        l1 = {"vs": ref_form, "steps": args.warmup + args.steps, "bar": L1_BAR, "strict": {}}
        for f in range(5):
            l1["strict"][names[f]] = float(np.abs(a[f] - b[f]).sum() / np.abs(b[f]).sum())
        l1["fields_over_bar"] = [k for k, v in l1["strict"].items() if not v <= L1_BAR]
        del a, b
        log("mixed vs", ref_form, "L1:", json.dumps(l1))
    head = "mixed" if "mixed" in res else forms[0]
    r = res[head]
    # roofline of the dominant kernel (stage 2 completes the cell update): algorithmic
    # 80 B per cell-update x n^3 cells per launch / the launch's average duration
    dom_ms = max(r["stage1_ms"], r["stage2_ms"])
    dom = "stage2" if r["stage2_ms"] >= r["stage1_ms"] else "stage1"
    achieved = BYTES_PER_CELL_UPDATE * cells / (r["stage2_ms"] * 1e-3) / 1e9
    trec = measured_traffic(head, r["tile"], n)
    out = {
        "metric": "cell-updates/s at 512^3 PPMC+HLLC, mixed-FP32 vs FP64, at stated L1 error",
        "value": r["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"Kelvin-Helmholtz {n}^3 periodic, PPMC+HLLC, VL, CFL 0.3, form={head}",
                   "cells": cells, "n_ghost": ng, "gamma": g, "seed": args.seed},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": trec["bytes_per_launch"] if trec else None,
                     "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT) if trec else None, "kernel": "Stage2Kernel",
                     "dominant": dom, "dominant_ms": dom_ms,
                     "algorithmic_bytes_per_launch": BYTES_PER_CELL_UPDATE * cells},
        "step_roofline_frac": r["cell_updates_per_s"] * BYTES_PER_CELL_UPDATE / 1e9 / HBM_PEAK_GBS,
        "forms": res,
        "l1_mixed_vs_fp64": l1,
    }
    if "fp64" in res and "mixed" in res:
        out["speedup_mixed_vs_fp64"] = res["mixed"]["cell_updates_per_s"] / res["fp64"]["cell_updates_per_s"]
    if "fp64_exact" in res and "mixed" in res:
        out["speedup_mixed_vs_fp64_exact"] = res["mixed"]["cell_updates_per_s"] / res["fp64_exact"]["cell_updates_per_s"]
    if not args.no_cpu:
        try:
            cb = cpu_baseline()
            log("cpu baseline:", json.dumps(cb))
            out["cpu_baseline"] = cb
        except Exception as e:  # the CPU leg must not hide the GPU measurement
            log("cpu baseline failed:", e)
            out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
