#!/usr/bin/env python3
"""Transcribe the HLLD solver test vectors of the reference into a JSON fixture.

Reads (as text) src/riemann_solvers/hlld_cuda_tests.cu of cholla-hydro/cholla and
writes tests/golden/hlld_vectors.json: for every whole-solver check of the test
fixture tMHDCalculateHLLDFluxesCUDA, the left and right states (conserved, in the
order density, momentum x/y/z, energy, Bx, By, Bz as that test builds them with
its Primitive_2_Conserved helper: momentum = density * velocity, energy =
Calc_Energy_Primitive with the magnetic energy), the adiabatic index and the
fiducial flux (density, momentum x/y/z, energy, Bx, By, Bz).  Scalar and
dual-energy fluxes are not carried (those build options are off here).

Run once in a container that holds the reference:
    python tests/golden/extract_hlld_vectors.py /root/reference
The tests read only the JSON it writes.
"""
import json
import math
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = ["BrioAndWuShockTube", "DaiAndWoodwardShockTube", "RyuAndJones4dShockTube", "EinfeldtStrongRarefaction",
         "ConstantStatesExpectCorrectFlux", "DegenerateState"]


def cexpr(text, env):
    t = " ".join(text.split()).replace("std::sqrt", "math.sqrt").replace("M_PI", "math.pi").replace("std::pow", "math.pow")
    t = re.sub(r"(\d)\.(?!\d)", r"\1.0", t)  # C "2." literals
    t = re.sub(r"(\d[\d.]*)[eE]([+-]?\d+)", r"\1e\2", t)
    return float(eval(t, {"math": math}, env))


def energy(p, d, vx, vy, vz, g, bx, by, bz):
    # hydro_utilities::Calc_Energy_Primitive with MHD (src/utils/hydro_utilities.h:106-118)
    e = (max(p, 1e-20) / (g - 1.)) + 0.5 * d * (vx * vx + ((vy * vy) + (vz * vz)))
    e += 0.5 * (bx * bx + ((by * by) + (bz * bz)))
    return e


def prim_to_cons(v, g):
    d, vx, vy, vz, p, bx, by, bz = v
    return [d, vx * d, vy * d, vz * d, energy(p, d, vx, vy, vz, g, bx, by, bz), bx, by, bz]


def split_top(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return out


def main(ref):
    src = open(os.path.join(ref, "src", "riemann_solvers", "hlld_cuda_tests.cu")).read()
    src = re.sub(r"//[^\n]*", "", src)
    starts = [m.start() for m in re.finditer(r"TEST_F\(tMHDCalculateHLLDFluxesCUDA,", src)]
    cases, skipped = [], 0
    for si, st in enumerate(starts):
        body = src[st: starts[si + 1] if si + 1 < len(starts) else len(src)]
        name = re.match(r"TEST_F\(tMHDCalculateHLLDFluxesCUDA,\s*(\w+)", body).group(1)
        if not any(name.startswith(t) for t in TESTS):
            continue
        env = {}
        for m in re.finditer(r"Real const (\w+)\s*=\s*([^;]+);", body):
            try:
                env[m.group(1)] = cexpr(m.group(2), env)
            except Exception:
                pass
        g = env["gamma"]
        states = {}
        for m in re.finditer(r"(\w+)\s*=\s*Primitive_2_Conserved\(\s*\{([^}]*)\}\s*,\s*gamma\s*,\s*primitiveScalar\)",
                             body):
            states[m.group(1)] = prim_to_cons([cexpr(x, env) for x in split_top(m.group(2))], g)
        # each check: fiducialFlux{...} ... Compute_Fluxes(A, B, gamma, direction)
        for m in re.finditer(r"fiducialFlux\s*\{([^}]*)\};(.*?)Compute_Fluxes\((\w+),\s*(\w+),\s*gamma,\s*direction\)",
                             body, flags=re.S):
            fid = [cexpr(x, env) for x in split_top(m.group(1))]
            between, a, b = m.group(2), m.group(3), m.group(4)
            local = dict(states)
            ok = True
            for dm in re.finditer(r"std::vector<Real>\s+(?!const)([^;]+);", between):
                for part in split_top(dm.group(1)):
                    lhs, rhs = [x.strip() for x in part.split("=")]
                    if rhs in local:
                        local[lhs] = list(local[rhs])
                    else:
                        ok = False
            for am in re.finditer(r"(\w+)\[(\d+)\]\s*=\s*-(\w+)\[(\d+)\];", between):
                if am.group(1) != am.group(3) or am.group(2) != am.group(4):
                    ok = False
                    continue
                local[am.group(1)][int(am.group(2))] = -local[am.group(1)][int(am.group(2))]
            if re.search(r"\b\w+\[\d+\]\s*=(?!=)", re.sub(r"(\w+)\[(\d+)\]\s*=\s*-\1\[\2\];", "", between)):
                ok = False
            if not ok or a not in local or b not in local or len(fid) != 8:
                skipped += 1
                continue
            cases.append({"test": name, "gamma": g, "left": local[a], "right": local[b], "flux": fid,
                          "states": [a, b]})
    out = {"_about": "Transcribed from src/riemann_solvers/hlld_cuda_tests.cu (tMHDCalculateHLLDFluxesCUDA) by "
                     "tests/golden/extract_hlld_vectors.py; states are conserved (d, mx, my, mz, E, Bx, By, Bz) in "
                     "the sweep frame, flux likewise; the reference checks each in all three directions with "
                     "absolute tolerance 2.7e-15 or 7 ulp (hlld_cuda_tests.cu:195-205).",
           "cases": cases}
    with open(os.path.join(HERE, "hlld_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"{len(cases)} cases written, {skipped} skipped")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
