"""CPU tests of the NetCDF-4 field files (lib/libpdfs_nc.so over HDF5): the
layout the reference's MPI build writes and reads (mpi_openacc/src/
ncwrap.c:121-225), checked by round trips and by HDF5's own h5dump."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import pdfs_amd

pytestmark = pytest.mark.skipif(not os.path.exists(pdfs_amd.NC_LIB_PATH),
                                reason="lib/libpdfs_nc.so not built (no HDF5)")
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def test_library_exports():
    L = C.CDLL(pdfs_amd.NC_LIB_PATH)
    for name in pdfs_amd.NC_EXPORTED:
        assert hasattr(L, name), name


def test_round_trip_bitwise(tmp_path):
    rng = np.random.default_rng(5)
    Z, Y, X = 6, 9, 11
    fl = rng.integers(-1, 2, size=(Z, Y, X)).astype(np.int32)
    d = rng.uniform(-3, 900, size=(Z, Y, X))
    d[0, 0, 0], d[1, 2, 3], d[2, 3, 4] = np.inf, np.nan, 0.0
    p = tmp_path / "f.nc"
    assert pdfs_amd.nc_write(p, fl, d, (0.5, 0.25, 2.0)) == 0
    rc, v = pdfs_amd.nc_read(p)
    assert rc == 0 and v["h"] == (0.5, 0.25, 2.0)
    assert np.array_equal(v["flags"], fl)
    assert np.array_equal(v["distance"].view(np.uint64), d.view(np.uint64))


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="no h5dump")
def test_netcdf4_layout_seen_by_h5dump(tmp_path):
    """The objects ncPutDim creates (ncwrap.c:156-225), as HDF5 itself
    reports them: x, y, z are float dimension scales with values delta * k,
    _Netcdf4Dimid 0, 1, 2 and a double 'delta'; Flags (32-bit int) and
    Distance (64-bit float) are (z, y, x) with DIMENSION_LIST -> z, y, x."""
    Z, Y, X = 3, 4, 5
    p = tmp_path / "f.nc"
    assert pdfs_amd.nc_write(p, np.zeros((Z, Y, X), np.int32), np.ones((Z, Y, X)),
                             (0.5, 0.25, 2.0)) == 0
    full = subprocess.run([H5DUMP, "-A", str(p)], capture_output=True, text=True,
                          timeout=60).stdout
    blocks = dict(re.findall(r'DATASET "(\w+)" \{(.*?)\n   \}', full, re.S))
    assert set(blocks) == {"x", "y", "z", "Flags", "Distance"}
    for i, (ax, n, dl) in enumerate((("x", X, 0.5), ("y", Y, 0.25), ("z", Z, 2.0))):
        b = blocks[ax]
        assert "H5T_IEEE_F32LE" in b and f"SIMPLE {{ ( {n} ) / ( {n} ) }}" in b
        assert '"DIMENSION_SCALE"' in b and f'(0): "{ax}"' in b
        m = re.search(r'ATTRIBUTE "_Netcdf4Dimid" \{.*?SCALAR.*?\(0\): (\d+)', b, re.S)
        assert m and int(m.group(1)) == i
        m = re.search(r'ATTRIBUTE "delta" \{.*?H5T_IEEE_F64LE.*?\(0\): ([-\d.e]+)', b, re.S)
        assert m and float(m.group(1)) == dl
    for var, t in (("Flags", "H5T_STD_I32LE"), ("Distance", "H5T_IEEE_F64LE")):
        b = blocks[var]
        assert t in b and f"SIMPLE {{ ( {Z}, {Y}, {X} ) / ( {Z}, {Y}, {X} ) }}" in b
        refs = re.findall(r"DATASET \d+ /(\w)", b)
        assert refs == ["z", "y", "x"], refs
    vals = subprocess.run([H5DUMP, "-d", "y", "-y", "-w", "0", str(p)], capture_output=True,
                          text=True, timeout=60).stdout
    data = re.search(r"DATA \{\s*(.*?)\s*\}", vals, re.S).group(1)
    assert [float(v) for v in data.replace(",", " ").split()] == [0.25 * k for k in range(Y)]


def test_read_errors(tmp_path):
    assert pdfs_amd.nc_read(tmp_path / "missing.nc")[0] == -1
    p = tmp_path / "text.nc"
    p.write_text("not an HDF5 file\n")
    assert pdfs_amd.nc_read(p)[0] == -1
    # an HDF5 file without the variables: made through HDF5's own C API
    h5 = C.CDLL(os.path.join(os.path.dirname(os.path.realpath(H5DUMP)), "..", "lib",
                             "libhdf5.so"))
    h5.H5open()
    h5.H5Fcreate.restype = C.c_int64
    h5.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, C.c_int64, C.c_int64]
    h5.H5Fclose.argtypes = [C.c_int64]
    e = tmp_path / "empty.nc"
    f = h5.H5Fcreate(str(e).encode(), 2, 0, 0)  # H5F_ACC_TRUNC, default property lists
    assert f >= 0 and h5.H5Fclose(f) >= 0
    assert pdfs_amd.nc_read(e)[0] == -2
