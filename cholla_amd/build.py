"""Build libcholla_amd.so (gfx950) in-tree with hipcc.

    python -m cholla_amd.build        # or cholla_amd.build.build()

Every HIP translation unit is compiled for gfx950 with packed-FP32 instructions
disabled.  The FP64 reference-order kernels (hydro_exact.hip, support_kernels.hip,
ensemble_exact.hip, mhd_exact.hip) are built with -ffp-contract=off.  The fast FP64
and mixed-precision kernels (hydro_fast.hip, ensemble_fast.hip, mhd_fast.hip) are
built with -ffp-contract=on: a product and a sum are fused only where they sit in
one source expression, so a tile body rounds the same way in every kernel it is
inlined into (the single-box, ensemble-split, ensemble-persistent and per-box loop
launches give identical bits, and a face flux that two neighbouring tiles both
compute is the same number in both, which keeps the update conservative).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libcholla_amd.so")
IO_LIB = os.path.join(HERE, "libcholla_amd_io.so")
BUILD_DIR = os.path.join(ROOT, "_build")

ARCH = os.environ.get("CHL_OFFLOAD_ARCH", "gfx950")
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Wno-unused-variable", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}",
]
UNITS = [
    ("hydro_exact.hip", ["-ffp-contract=off"]),
    ("support_kernels.hip", ["-ffp-contract=off"]),
    ("snapshot_kernels.hip", ["-ffp-contract=off"]),
    ("hydro_fast.hip", ["-ffp-contract=on"]),
    ("ensemble_exact.hip", ["-ffp-contract=off"]),
    ("ensemble_fast.hip", ["-ffp-contract=on"]),
    ("mhd_exact.hip", ["-ffp-contract=off"]),
    ("mhd_fast.hip", ["-ffp-contract=on"]),
    ("capi.cpp", ["-ffp-contract=off"]),
    ("noalloc.hip", ["-ffp-contract=off"]),
    ("ensemble.cpp", ["-ffp-contract=off"]),
    ("mhd.cpp", ["-ffp-contract=off"]),
]
# Host programs built on the reference-signature adapter (cholla_dropin.cpp), one per
# reconstruction build option, as Cholla's make.type files select them.  They need no HIP
# headers: plain g++ against the C ABI.
DRIVERS = {
    "dropin_driver_plmc": ["-DPLMC", "-DDENSITY_FLOOR", "-DTEMPERATURE_FLOOR"],
    "dropin_driver_ppmc": ["-DPPMC", "-DDENSITY_FLOOR", "-DTEMPERATURE_FLOOR"],
    "dropin_driver_mhd": ["-DMHD", "-DPLMC", "-DHLLD", "-DDENSITY_FLOOR", "-DTEMPERATURE_FLOOR"],
    "dropin_driver_mhd_ppmc": ["-DMHD", "-DPPMC", "-DHLLD", "-DDENSITY_FLOOR", "-DTEMPERATURE_FLOOR"],
    "dropin_driver_mhd_pcm": ["-DMHD", "-DPCM", "-DHLLD", "-DDENSITY_FLOOR", "-DTEMPERATURE_FLOOR"],
    "dropin_driver_ppmc_de": ["-DPPMC", "-DDE", "-DDENSITY_FLOOR", "-DTEMPERATURE_FLOOR"],
    "dropin_driver_plmc_grav": ["-DPLMC", "-DSTATIC_GRAV", "-DDENSITY_FLOOR", "-DTEMPERATURE_FLOOR"],
}


def hipcc():
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found")
    return p


def _compile(src, extra):
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    cmd = [hipcc(), *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature and says so; drop that line
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{err}")
    return obj


def build(verbose=True):
    os.makedirs(BUILD_DIR, exist_ok=True)
    units = [(s, e) for s, e in UNITS if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(len(units), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda u: _compile(*u), units))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, LIB)
    for name, defs in DRIVERS.items():
        out = os.path.join(HERE, name)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", *defs, f"-I{os.path.join(ROOT, 'include')}",
               os.path.join(CSRC, "cholla_dropin.cpp"), os.path.join(CSRC, "dropin_driver.cpp"),
               f"-L{HERE}", "-lcholla_amd", "-Wl,-rpath,$ORIGIN", "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stderr}")
    io = build_io()
    if verbose:
        print(f"built {LIB}, {', '.join(DRIVERS)}" + (f" and {IO_LIB}" if io else " (no HDF5: snapshot library skipped)"))
    return LIB


def hdf5_prefix():
    """Install prefix of the HDF5 C library (headers and libhdf5.so), or None."""
    for pre in [os.environ.get("CHL_HDF5_PREFIX", ""), "/opt/conda", "/usr"]:
        if pre and os.path.exists(os.path.join(pre, "include", "hdf5.h")) and \
                os.path.exists(os.path.join(pre, "lib", "libhdf5.so")):
            return pre
    return None


def build_io():
    """libcholla_amd_io.so: Cholla-format HDF5 snapshots over the core library's C ABI
    (host code only: g++ against the HDF5 C library).  The HDF5 prefix goes in as a
    RUNPATH (direct dependencies only) and the C++ runtime is linked statically, so the
    prefix's own, older libstdc++ is never what the HIP runtime resolves to."""
    pre = hdf5_prefix()
    if pre is None:
        return None
    tmp = IO_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-fPIC", "-shared", f"-I{os.path.join(ROOT, 'include')}",
           f"-I{os.path.join(pre, 'include')}", os.path.join(CSRC, "io_hdf5.cpp"), "-o", tmp,
           f"-L{HERE}", "-lcholla_amd", f"-L{os.path.join(pre, 'lib')}", "-lhdf5",
           "-static-libstdc++", "-static-libgcc", f"-Wl,--enable-new-dtags,-rpath,$ORIGIN:{os.path.join(pre, 'lib')}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"snapshot library build failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, IO_LIB)
    return IO_LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
