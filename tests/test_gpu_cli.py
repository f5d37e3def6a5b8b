"""GPU tests of the CLI path: pre/post kernels and calc_dist_field against
the oracle pipeline (bit-exact), and the PDFS binary end to end."""
import os
import subprocess

import numpy as np
import pytest

import oracle_ctypes as O
import pdfs_amd
from vti_util import random_cli_input, reference_vti_text, write_input_vti

pytestmark = pytest.mark.gpu


def _padded(loc, dist):
    Z, Y, X = loc.shape
    L = np.full((Z + 2, Y + 2, X + 2), -1, np.int32)
    D = O.padded(X, Y, Z)
    L[1:-1, 1:-1, 1:-1] = loc
    D[1:-1, 1:-1, 1:-1] = dist
    return L, D


def oracle_pipeline(L, D, passes=1, h=(1.0, 1.0, 1.0)):
    D = D.copy()
    X, Y, Z = O.dims(D)
    O.lib().orc_update_distance(L, D, D.size)
    for _ in range(passes):
        for sn in range(1, 9):
            O.sweep_ref64(D, sn, h)
    O.lib().orc_set_negative_inside(L, D, D.size)
    O.lib().orc_adjust_boundary(D, X, Y, Z)
    return D


@pytest.mark.parametrize("op", [0, 1, 2])
def test_prepost_kernels_bitwise(op):
    loc, dist = random_cli_input(9, 7, 5, 10 + op)
    dist[1, 2, 3] = np.nan
    L, D = _padded(loc, dist)
    want = D.copy()
    X, Y, Z = O.dims(D)
    if op == 0:
        O.lib().orc_update_distance(L, want, want.size)
    elif op == 1:
        O.lib().orc_set_negative_inside(L, want, want.size)
    else:
        O.lib().orc_adjust_boundary(want, X, Y, Z)
    got = D.copy()
    pdfs_amd.prepost(op, L, got)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.5, 0.75, 1.25)])
@pytest.mark.parametrize("shape", [(13, 11, 9), (24, 16, 10)])
def test_calc_dist_field_matches_oracle(h, shape):
    """The CLI pipeline on the GPU against the oracle pipeline, bit for bit.
    24 x 16 tiles the FP64 path's 8 x 8 columns exactly, so its pass runs as
    one overlapped launch."""
    loc, dist = random_cli_input(*shape, 21)
    L, D = _padded(loc, dist)
    want = oracle_pipeline(L, D, 1, h)
    got = D.copy()
    pdfs_amd.calc_dist_field(L, got, h, passes=1)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_cli_end_to_end(tmp_path):
    assert os.path.exists(pdfs_amd.CLI_PATH), "build the PDFS binary"
    loc, dist = random_cli_input(10, 8, 6, 33)
    h = (1.0, 1.0, 1.0)
    inp = str(tmp_path / "field.vti")
    write_input_vti(inp, loc, dist, h)
    r = subprocess.run([pdfs_amd.CLI_PATH, inp, str(tmp_path / "run")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Parallel FSM time:" in r.stdout and "VTI file created:" in r.stdout
    assert "x: 10\tdx: 1.000000" in r.stdout
    L, D = _padded(loc, dist)
    want = reference_vti_text(oracle_pipeline(L, D, 1, h), h)
    got = open(str(tmp_path / "run_distfield.vti")).read()
    assert got == want


REF_HOST = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                        "_ref", "PDFS_ref_host")


@pytest.mark.skipif(not os.path.exists(REF_HOST), reason="reference host program not built")
@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.5, 0.75, 1.25)])
@pytest.mark.parametrize("shape", [(12, 9, 7), (16, 16, 8)])
def test_reference_host_program_linked_to_libpdfs(tmp_path, h, shape):
    """The drop-in as the reference builds it: single_gpu_cuda/src/{main,
    phi3D,vti_parser,file_writer}.c compiled unmodified (oracle/Makefile
    `dropin`) and linked against lib/libpdfs.so in place of the nvcc object
    (single_gpu_cuda/Makefile:19-29). Its run_fsm call (phi3D.c:119) goes
    through this build's C-ABI: the output file equals bin/PDFS's and the
    oracle pipeline's byte for byte, and stdout has the reference's lines."""
    loc, dist = random_cli_input(*shape, 47)
    inp = str(tmp_path / "field.vti")
    write_input_vti(inp, loc, dist, h)
    r = subprocess.run([REF_HOST, inp, str(tmp_path / "ref")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Parallel FSM time:" in r.stdout and f"x: {shape[0]}\tdx:" in r.stdout
    q = subprocess.run([pdfs_amd.CLI_PATH, inp, str(tmp_path / "ours")], capture_output=True,
                       text=True, timeout=120)
    assert q.returncode == 0, q.stdout + q.stderr
    L, D = _padded(loc, dist)
    want = reference_vti_text(oracle_pipeline(L, D, 1, h), h)
    got_ref = open(str(tmp_path / "ref_distfield.vti")).read()
    got_ours = open(str(tmp_path / "ours_distfield.vti")).read()
    assert got_ref == want
    assert got_ours == got_ref


@pytest.mark.skipif(not os.path.exists(pdfs_amd.NC_LIB_PATH), reason="no lib/libpdfs_nc.so")
@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.5, 0.75, 1.25)])
def test_cli_netcdf_end_to_end(tmp_path, h):
    """PDFS -i <file.nc> -p <prefix>, the MPI build on one process
    (mpi_openacc/src/main.c:71-145): the output <prefix>_soln.nc holds the
    padded Flags (input inside, -1 on the border, process_file.c:204-333)
    and the padded Distance of the one-pass pipeline, bit-exact against the
    oracle pipeline; stdout has the reference's lines."""
    assert os.path.exists(pdfs_amd.CLI_PATH), "build the PDFS binary"
    loc, dist = random_cli_input(12, 9, 7, 34)
    dist[2, 3, 4] = np.nan
    inp = tmp_path / "field.nc"
    assert pdfs_amd.nc_write(inp, loc, dist, h) == 0
    r = subprocess.run([pdfs_amd.CLI_PATH, "-i", str(inp), "-p", str(tmp_path / "run")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = str(tmp_path / "run_soln.nc")
    for line in (f" Input: {inp}", f"Output: {out}", "nx: 1, ny: 1, nz: 1",
                 "BOX_X: 14, BOX_Y: 11, BOX_Z: 9", "Parallel MPI/OpenACC Block FSM time:"):
        assert line in r.stdout, (line, r.stdout)
    rc, v = pdfs_amd.nc_read(out)
    assert rc == 0 and v["h"] == h
    L, D = _padded(loc, dist)
    want = oracle_pipeline(L, D, 1, h)
    assert np.array_equal(v["flags"], L)
    assert np.array_equal(v["distance"].view(np.uint64), want.view(np.uint64))


@pytest.mark.skipif(not os.path.exists(pdfs_amd.NC_LIB_PATH), reason="no lib/libpdfs_nc.so")
def test_cli_netcdf_options(tmp_path):
    """Decompositions other than 1x1x1 need as many processes as blocks
    (main.c:212-215), so on one process they are refused with the
    reference's message and usage; so are missing -i / -p. The FP32 and
    until-converged extensions converge to the FP64 solution."""
    loc, dist = random_cli_input(10, 8, 6, 35)
    inp = tmp_path / "f.nc"
    assert pdfs_amd.nc_write(inp, loc, dist) == 0
    prefix = str(tmp_path / "o")
    r = subprocess.run([pdfs_amd.CLI_PATH, "-i", str(inp), "-p", prefix, "--nx", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert "Total processes must equal total blocks in the decmposition!" in r.stdout
    assert "Usage:" in r.stdout and not os.path.exists(prefix + "_soln.nc")
    r = subprocess.run([pdfs_amd.CLI_PATH, "-i", str(inp)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0 and "Usage:" in r.stdout
    outs = {}
    for extra in (["--converge"], ["--converge", "--fp32"]):
        r = subprocess.run([pdfs_amd.CLI_PATH, "-i", str(inp), "-p", prefix] + extra,
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        outs[len(extra)] = pdfs_amd.nc_read(prefix + "_soln.nc")[1]["distance"]
    L, D = _padded(loc, dist)
    D64 = D.copy()
    X, Y, Z = O.dims(D64)
    O.lib().orc_update_distance(L, D64, D64.size)
    O.solve_ref64(D64, tol=1e-12)  # the CLI's until-converged tolerance
    O.lib().orc_set_negative_inside(L, D64, D64.size)
    O.lib().orc_adjust_boundary(D64, X, Y, Z)
    assert np.array_equal(outs[1].view(np.uint64), D64.view(np.uint64))
    assert np.abs(outs[2] - D64).max() <= 1e-4
