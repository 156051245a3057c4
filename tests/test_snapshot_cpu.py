"""CPU-side checks of the snapshot library (libcholla_amd_io.so, include/cholla_amd_io.h) and
the core library's snapshot-layout entry points: the library loads, exports what its header
declares, and refuses unsupported or inconsistent requests before touching a device."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import cholla_amd as chl
from cholla_amd import snapshot as snap

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cholla_amd_io.h")


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not os.path.exists(snap.IO_LIB_PATH):
        pytest.fail("libcholla_amd_io.so is not built (python -m cholla_amd.build; needs the HDF5 C library)")


def declared():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(chl_[a-z0-9_]+)\s*\(", text)))


def test_header_binding_and_exports_agree():
    assert declared() == sorted(snap.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", snap.IO_LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(l.split()[-1] for l in out.stdout.splitlines() if l.strip())
    assert set(declared()) <= exported


def test_header_struct_layout_matches_c():
    # the C struct: 9 ints, 3 doubles, 3 x double[3], one pointer (natural alignment)
    assert ctypes.sizeof(snap.IoHeader) == 9 * 4 + 4 + 3 * 8 + 9 * 8 + 8
    assert snap.IoHeader.gamma.offset == 40 and snap.IoHeader.macro_flags.offset == 136


class Fake:
    ptr = ctypes.c_void_p(16)  # never dereferenced: every case below fails its checks first


@pytest.mark.parametrize("kw,msg", [
    (dict(nx=20, ny=20, nz=1, n_ghost=4), "2D"),
    (dict(nx=20, ny=20, nz=20, n_ghost=4, dual_energy=True, mhd=True), "dual energy with MHD"),
    (dict(nx=106, ny=1, nz=1, n_ghost=3, mhd=True), "MHD is 3D only"),
    (dict(nx=8, ny=8, nz=8, n_ghost=4), "ghost"),
    (dict(nx=106, ny=1, nz=1, n_ghost=3, float32=True), "float32"),
])
def test_write_refuses_before_device_use(kw, msg, tmp_path):
    h = snap.header(**kw)
    with pytest.raises(chl.ChollaError, match=msg):
        snap.write_hdf5(str(tmp_path / "x.h5"), Fake(), h)
    assert not (tmp_path / "x.h5").exists()


def test_inconsistent_field_count_is_refused(tmp_path):
    h = snap.header(20, 20, 20, 4)
    h.n_fields = 6
    with pytest.raises(chl.ChollaError, match="n_fields"):
        snap.read_hdf5(str(tmp_path / "missing.h5"), Fake(), h)


def test_core_layout_entry_points_check_regions():
    L = chl.lib()
    p = ctypes.c_void_p(16)
    # 2D grids, bad direction, a region past the grid: refused before any launch
    assert L.chl_pack_real_cells(p, p, 20, 20, 1, 4, 12, 12, 1, -1) == 4
    assert L.chl_pack_real_cells(p, p, 20, 20, 20, 4, 12, 12, 12, 3) == 1
    assert L.chl_unpack_real_cells(p, p, 20, 20, 20, 4, 17, 12, 12, -1) == 1
    assert L.chl_pack_real_cells_f32(p, p, 106, 1, 1, 3, 100, 1, 1, 0) == 1
    assert L.chl_pack_real_cells(None, p, 20, 20, 20, 4, 12, 12, 12, -1) == 1
