"""CPU tests of the PDFS CLI's argument and input errors, which end before any
GPU work: the single-GPU build's usage and read errors (single_gpu_cuda/src/
main.c:56-80) and the MPI build's argument checks (mpi_openacc/src/main.c:
82-104, 172-225: usage on a bad or missing argument, exit status 0)."""
import os
import subprocess

import numpy as np
import pytest

import pdfs_amd
from vti_util import write_input_vti

pytestmark = pytest.mark.skipif(not os.path.exists(pdfs_amd.CLI_PATH),
                                reason="bin/PDFS not built")


def run(*args):
    return subprocess.run([pdfs_amd.CLI_PATH, *map(str, args)], capture_output=True, text=True,
                          timeout=60)


def test_vti_usage_and_read_errors(tmp_path):
    r = run()
    assert r.returncode == 1 and r.stdout.startswith("Usage: ") and "<file.vti> <prefix>" in r.stdout
    r = run(tmp_path / "only.vti")
    assert r.returncode == 1 and "Usage: " in r.stdout
    r = run(tmp_path / "missing.vti", tmp_path / "o")
    assert r.returncode == 1
    assert r.stdout == f'Unable to read file "{tmp_path / "missing.vti"}"!\n'
    bad = tmp_path / "bad.vti"
    bad.write_text("<?xml version=\"1.0\"?>\nnot a VTI header\n")
    r = run(bad, tmp_path / "o")
    assert r.returncode == 1 and "Malformed VTI file" in r.stdout
    # an extension flag it does not know
    good = tmp_path / "good.vti"
    write_input_vti(good, np.zeros((2, 3, 4), np.int32), np.ones((2, 3, 4)))
    r = run(good, tmp_path / "o", "--bogus")
    assert r.returncode == 1 and "[--passes N | --converge] [--fp32]" in r.stdout
    assert not any(p.name.startswith("o_") for p in tmp_path.iterdir())


def test_netcdf_argument_checks(tmp_path):
    usage = "Usage: "
    prefix = tmp_path / "o"
    for args in (["-i", tmp_path / "f.nc"],                   # no -p
                 ["-p", prefix],                              # no -i
                 ["-i", tmp_path / "f.nc", "-p", prefix, "--nx", "0"],
                 ["-i", tmp_path / "f.nc", "-p", prefix, "--ny", "2"]):
        r = run(*args)
        assert r.returncode == 0 and usage in r.stdout, (args, r.stdout)
        assert "Input:" not in r.stdout
    r = run("-i", tmp_path / "f.nc", "-p", prefix, "--nz", "3")
    assert "Total processes must equal total blocks in the decmposition!" in r.stdout
    assert not os.path.exists(str(prefix) + "_soln.nc")


@pytest.mark.skipif(not os.path.exists(pdfs_amd.NC_LIB_PATH), reason="no lib/libpdfs_nc.so")
def test_netcdf_input_errors(tmp_path):
    prefix = tmp_path / "o"
    r = run("-i", tmp_path / "missing.nc", "-p", prefix)
    assert r.returncode == 2 and "cannot open the NetCDF-4 file" in r.stdout
    assert " Input: " in r.stdout and "nx: 1, ny: 1, nz: 1" in r.stdout
    txt = tmp_path / "text.nc"
    txt.write_text("not HDF5\n")
    r = run("-i", txt, "-p", prefix)
    assert r.returncode == 2 and "cannot open the NetCDF-4 file" in r.stdout
    assert not os.path.exists(str(prefix) + "_soln.nc")
