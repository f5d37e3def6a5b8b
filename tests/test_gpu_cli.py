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
def test_calc_dist_field_matches_oracle(h):
    loc, dist = random_cli_input(13, 11, 9, 21)
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
