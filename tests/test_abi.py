"""CPU checks of the C-ABI libraries: they load, and export every function
that include/*.h declares (include/pdfs_nc.h: lib/libpdfs_nc.so; the other
headers: lib/libpdfs.so). No compute calls (no GPU here)."""
import ctypes
import os
import re
import subprocess

import pdfs_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


NC_HEADER = "pdfs_nc.h"


def declared_functions(nc=False):
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h") or (fn == NC_HEADER) != nc:
            continue
        src = open(os.path.join(ROOT, "include", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_loads_and_exports_declared_symbols():
    L = pdfs_amd.lib()
    names = declared_functions()
    assert "run_fsm" in names and "runFSM" in names and "pdfs_solve" in names
    out = subprocess.run(["nm", "-D", "--defined-only", pdfs_amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(L, n)


def test_netcdf_library_exports_declared_symbols():
    if not os.path.exists(pdfs_amd.NC_LIB_PATH):
        import pytest
        pytest.skip("lib/libpdfs_nc.so not built (no HDF5)")
    names = declared_functions(nc=True)
    assert names == sorted(pdfs_amd.NC_EXPORTED)
    out = subprocess.run(["nm", "-D", "--defined-only", pdfs_amd.NC_LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(names) <= exported


def test_phi_struct_layouts_match_reference():
    # single_gpu_cuda/src/phi3D.h:45-55 (64 bytes) and mpi_openacc/src/phi3D.h:45-54 (56 bytes)
    assert ctypes.sizeof(pdfs_amd.Phi) == 64
    assert pdfs_amd.Phi.distance.offset == 48 and pdfs_amd.Phi.location.offset == 40
    assert ctypes.sizeof(pdfs_amd.PhiACC) == 56
    assert pdfs_amd.PhiACC.distance.offset == 40
