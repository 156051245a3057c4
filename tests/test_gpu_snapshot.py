"""Snapshots in Cholla's HDF5 format (libcholla_amd_io.so) on the GPU.

The reference holds no snapshot files for this path (its test-data submodule is not
vendored), so the format is pinned against the reference's writer and reader code, checked
from outside with the HDF5 tools' own reader (h5dump), not with this library:
  * the header attributes of Write_Header_HDF5 (io.cpp:605-735) with their file types
    (F64BE / I32BE / variable-length strings) and values;
  * each dataset big-endian F64 of extent (nx_real, ny_real, nz_real) holding
    U[f, k+ng, j+ng, i+ng] at [i][j][k] (io.cpp:1232-1268, io_gpu.cu:44-60), magnetic
    fields with one more face along their own axis starting at grid index ng-1 (io.cpp:
    1472-1480, 2566-2700), 1D datasets of extent (nx_real), float32 files as F32BE;
and by behaviour: a snapshot read back restores every real cell bit for bit and leaves
ghost cells alone, and a run restarted from a snapshot (read, boundaries, continue) is
bit-identical to the uninterrupted run.
"""
import os
import re
import shutil
import subprocess
import sys

import numpy as np
import pytest

import cholla_amd as chl
from cholla_amd import snapshot as snap
from helpers import interior

pytestmark = pytest.mark.gpu
G = 5.0 / 3.0
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _require_gpu():
    if chl.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need a real MI355X")
    if not os.path.exists(snap.IO_LIB_PATH):
        pytest.fail("libcholla_amd_io.so is not built")


def need_h5dump():
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump (the HDF5 tools) is not installed: the format checks need an independent reader")


def h5_header(path):
    """{name: (kind, datatype, extents)} of the root group's attributes and datasets."""
    out = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    items = {}
    for m in re.finditer(r'(ATTRIBUTE|DATASET) "([^"]+)" \{\s*DATATYPE\s+(H5T_STRING|\S+).*?DATASPACE\s+SIMPLE \{ \( ([\d, ]+) \)',
                         out, re.S):
        items[m.group(2)] = (m.group(1), m.group(3), tuple(int(x) for x in m.group(4).split(",")))
    return items


def h5_attr(path, name):
    """An attribute's values as h5dump prints them (17 significant digits)."""
    out = subprocess.run([H5DUMP, "-m", "%.17g", "-a", name, path], capture_output=True, text=True,
                         check=True).stdout
    body = out[out.index("DATA {"):]
    vals = ",".join(re.findall(r"\(\d+\):\s*(.*)", body))
    return [float(v) for v in vals.split(",") if v.strip()]


def h5_data(path, name, tmp, dtype="<f8"):
    binf = os.path.join(tmp, name.strip("/") + ".bin")
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", binf, path], capture_output=True, check=True)
    return np.fromfile(binf, dtype=dtype)


def kh_state(n, ng=4, seed=11):
    U = chl.ic_kelvin_helmholtz(n, ng, G, seed=seed)
    return U, n + 2 * ng


def test_hydro_snapshot_format_and_round_trip(tmp_path):
    need_h5dump()
    n, ng = 12, 4
    U, N = kh_state(n, ng)
    # a non-cubic box: 12 x 10 x 8 real cells
    A = U.reshape(5, N, N, N)[:, : 8 + 2 * ng, : 10 + 2 * ng, :].copy()
    nx, ny, nz = N, 10 + 2 * ng, 8 + 2 * ng
    A[:, :, :, :] += np.arange(nx)[None, None, None, :] * 1e-3  # make x distinguishable
    D = chl.DeviceArray.from_host(A)
    path = str(tmp_path / "3.h5")
    h = snap.header(nx, ny, nz, ng, gamma=G, t=0.125, dt=1e-3, n_step=7, dx=(1 / 12, 1 / 10, 1 / 8),
                    bounds=(-0.5, 0.0, 0.25), macro_flags="-DPPMC -DHLLC -DVL")
    say("snapshot hydro write", nx, ny, nz)
    snap.write_hdf5(path, D, h)
    items = h5_header(path)
    for name in ("gamma", "t", "dt", "time_unit", "length_unit", "mass_unit", "velocity_unit", "density_unit",
                 "energy_unit", "bounds", "domain", "dx"):
        assert items[name][:2] == ("ATTRIBUTE", "H5T_IEEE_F64BE"), name
    for name in ("n_step", "n_fields", "dims"):
        assert items[name][:2] == ("ATTRIBUTE", "H5T_STD_I32BE"), name
    for name in ("Git Commit Hash", "Macro Flags", "cholla"):
        assert items[name][:2] == ("ATTRIBUTE", "H5T_STRING"), name
    assert "magnetic_field_unit" not in items and "GasEnergy" not in items
    assert h5_attr(path, "gamma") == pytest.approx([G], rel=1e-12)
    assert h5_attr(path, "t") == pytest.approx([0.125]) and h5_attr(path, "n_step") == [7]
    assert h5_attr(path, "n_fields") == [5] and h5_attr(path, "dims") == [12, 10, 8]
    assert h5_attr(path, "bounds") == pytest.approx([-0.5, 0.0, 0.25])
    assert h5_attr(path, "domain") == pytest.approx([1.0, 1.0, 1.0], rel=1e-12)
    assert h5_attr(path, "length_unit") == pytest.approx([3.08567758e21], rel=1e-9)
    real = A[:, ng:-ng, ng:-ng, ng:-ng]  # (f, k, j, i)
    for f, name in enumerate(["density", "momentum_x", "momentum_y", "momentum_z", "Energy"]):
        assert items[name] == ("DATASET", "H5T_IEEE_F64BE", (12, 10, 8)), name
        got = h5_data(path, "/" + name, str(tmp_path)).reshape(12, 10, 8)
        np.testing.assert_array_equal(got, real[f].transpose(2, 1, 0))
    # read back into a state whose ghost cells hold a marker
    B = chl.DeviceArray.from_host(np.full(A.size, -7.0))
    h2 = snap.header(nx, ny, nz, ng)
    info = snap.read_hdf5(path, B, h2)
    assert info == {"gamma": G, "t": 0.125, "n_step": 7}
    R = B.download().reshape(A.shape)
    np.testing.assert_array_equal(R[:, ng:-ng, ng:-ng, ng:-ng], real)
    mask = np.ones(A.shape[1:], bool)
    mask[ng:-ng, ng:-ng, ng:-ng] = False
    assert (R[:, mask] == -7.0).all()


def test_restart_from_snapshot_is_bit_identical(tmp_path):
    n, ng = 16, 4
    U, N = kh_state(n, ng, seed=5)
    h = 1.0 / n
    say("snapshot restart")
    cont = chl.DeviceArray.from_host(U)
    hy = chl.Hydro(N, N, N, ng, chl.RECON_PPMC, chl.PREC_FP64_EXACT)
    n1, t1 = hy.run(cont, h, h, h, G, 0.0, 1e30, 6)
    first = chl.DeviceArray.from_host(U)
    n0, t0 = hy.run(first, h, h, h, G, 0.0, 1e30, 3)
    path = str(tmp_path / "restart.h5")
    snap.write_hdf5(path, first, snap.header(N, N, N, ng, gamma=G, t=t0, n_step=n0, dx=(h, h, h)))
    again = chl.DeviceArray(U.size, fill=np.zeros(U.size))
    info = snap.read_hdf5(path, again, snap.header(N, N, N, ng))
    chl.set_boundaries(again, N, N, N, ng, [chl.BC_PERIODIC] * 6)
    n2, t2 = hy.run(again, h, h, h, G, info["t"], 1e30, 3)
    hy.close()
    assert n0 + n2 == n1 == 6 and t2 == t1
    np.testing.assert_array_equal(again.download(), cont.download())


def test_dual_energy_and_float32_snapshots(tmp_path):
    need_h5dump()
    n, ng = 12, 4
    U, N = kh_state(n, ng, seed=2)
    A = np.concatenate([U.reshape(5, -1), (U.reshape(5, -1)[4] * 0.25)[None]]).reshape(-1)
    D = chl.DeviceArray.from_host(A)
    p64, p32 = str(tmp_path / "de.h5"), str(tmp_path / "de.h5.float32")
    say("snapshot DE and float32")
    snap.write_hdf5(p64, D, snap.header(N, N, N, ng, gamma=G, dual_energy=True))
    snap.write_hdf5(p32, D, snap.header(N, N, N, ng, gamma=G, dual_energy=True, float32=True))
    real = A.reshape(6, N, N, N)[:, ng:-ng, ng:-ng, ng:-ng]
    it64, it32 = h5_header(p64), h5_header(p32)
    names = ["density", "momentum_x", "momentum_y", "momentum_z", "Energy", "GasEnergy"]
    assert h5_attr(p64, "n_fields") == [6]
    for f, name in enumerate(names):
        assert it64[name] == ("DATASET", "H5T_IEEE_F64BE", (n, n, n))
        assert it32[name] == ("DATASET", "H5T_IEEE_F32BE", (n, n, n))
        got = h5_data(p32, "/" + name, str(tmp_path), "<f4").reshape(n, n, n)
        np.testing.assert_array_equal(got, real[f].transpose(2, 1, 0).astype(np.float32))
    B = chl.DeviceArray(A.size, fill=np.zeros(A.size))
    snap.read_hdf5(p64, B, snap.header(N, N, N, ng, dual_energy=True))
    np.testing.assert_array_equal(B.download().reshape(6, N, N, N)[:, ng:-ng, ng:-ng, ng:-ng], real)


def test_mhd_snapshot_face_fields(tmp_path):
    need_h5dump()
    n, ng = 12, 4
    N = n + 2 * ng
    U = chl.ic_orszag_tang(n, ng, G, seed=3, amplitude=1e-3)
    D = chl.DeviceArray.from_host(U)
    path = str(tmp_path / "ot.h5")
    say("snapshot MHD")
    snap.write_hdf5(path, D, snap.header(N, N, N, ng, gamma=G, mhd=True))
    items = h5_header(path)
    assert items["magnetic_field_unit"][:2] == ("ATTRIBUTE", "H5T_IEEE_F64BE")
    assert h5_attr(path, "n_fields") == [8]
    C = U.reshape(8, N, N, N)
    for f, (name, ext, sl) in enumerate([
            ("magnetic_x", (n + 1, n, n), (slice(ng, N - ng), slice(ng, N - ng), slice(ng - 1, N - ng))),
            ("magnetic_y", (n, n + 1, n), (slice(ng, N - ng), slice(ng - 1, N - ng), slice(ng, N - ng))),
            ("magnetic_z", (n, n, n + 1), (slice(ng - 1, N - ng), slice(ng, N - ng), slice(ng, N - ng)))]):
        assert items[name] == ("DATASET", "H5T_IEEE_F64BE", ext), name
        got = h5_data(path, "/" + name, str(tmp_path)).reshape(ext)
        np.testing.assert_array_equal(got, C[5 + f][sl].transpose(2, 1, 0))
    B = chl.DeviceArray(U.size, fill=np.zeros(U.size))
    snap.read_hdf5(path, B, snap.header(N, N, N, ng, mhd=True))
    R = B.download().reshape(8, N, N, N)
    np.testing.assert_array_equal(R[:5, ng:-ng, ng:-ng, ng:-ng], C[:5, ng:-ng, ng:-ng, ng:-ng])
    np.testing.assert_array_equal(R[5, ng:-ng, ng:-ng, ng - 1:-ng], C[5, ng:-ng, ng:-ng, ng - 1:-ng])
    np.testing.assert_array_equal(R[7, ng - 1:-ng, ng:-ng, ng:-ng], C[7, ng - 1:-ng, ng:-ng, ng:-ng])


def test_one_d_snapshot_and_extent_check(tmp_path):
    need_h5dump()
    nx, ng = 100, 3
    N = nx + 2 * ng
    A = np.random.default_rng(4).uniform(0.5, 1.5, (5, N))
    D = chl.DeviceArray.from_host(A)
    path = str(tmp_path / "sod.h5")
    say("snapshot 1D")
    snap.write_hdf5(path, D, snap.header(N, 1, 1, ng, gamma=1.4, dx=(0.01, 1, 1)))
    items = h5_header(path)
    assert items["density"] == ("DATASET", "H5T_IEEE_F64BE", (nx,))
    assert h5_attr(path, "dims") == [nx, 1, 1]
    np.testing.assert_array_equal(h5_data(path, "/Energy", str(tmp_path)), A[4, ng:-ng])
    B = chl.DeviceArray(A.size, fill=np.zeros(A.size))
    snap.read_hdf5(path, B, snap.header(N, 1, 1, ng))
    np.testing.assert_array_equal(B.download().reshape(5, N)[:, ng:-ng], A[:, ng:-ng])
    # a snapshot of another size is refused, not read past
    C = chl.DeviceArray(5 * (N + 10), fill=np.zeros(5 * (N + 10)))
    with pytest.raises(chl.ChollaError, match="extents"):
        snap.read_hdf5(path, C, snap.header(N + 10, 1, 1, ng))
