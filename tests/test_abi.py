"""CPU checks of the C-ABI libraries: they load, and export every function
that include/*.h declares (include/pdfs_nc.h: lib/libpdfs_nc.so; the other
headers: lib/libpdfs.so). No compute calls (no GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

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


def test_entry_points_reject_bad_arguments_before_touching_the_gpu():
    """Argument checks come first in every status-returning entry point:
    PDFS_EINVAL (-1) and a message naming the call, with no context, no
    device memory and no GPU. Spacings must be finite and positive (the
    batch, single-field, FP64 and judge paths alike)."""
    import math

    import numpy as np
    C = ctypes
    L = pdfs_amd.lib()
    st = pdfs_amd.Stats()
    opts = pdfs_amd.Opts()
    L.pdfs_default_opts(C.byref(opts))

    def err():
        return L.pdfs_last_error().decode()

    a = np.zeros((6, 6, 6))
    for h in ((0.0, 1.0, 1.0), (1.0, -2.0, 1.0), (1.0, 1.0, math.nan), (math.inf, 1.0, 1.0)):
        assert L.pdfs_solve(a.ctypes.data, 4, 4, 4, *h, C.byref(opts), C.byref(st)) == -1, h
        assert "pdfs_solve" in err()
    assert L.pdfs_solve(None, 4, 4, 4, 1.0, 1.0, 1.0, C.byref(opts), C.byref(st)) == -1
    assert L.pdfs_solve(a.ctypes.data, 0, 4, 4, 1.0, 1.0, 1.0, C.byref(opts), C.byref(st)) == -1
    calls = [
        ("pdfs_solve_f32_dev", lambda: L.pdfs_solve_f32_dev(None, None, None, 4, 4, 4, 1.0, 1.0,
                                                            1.0, C.byref(opts), C.byref(st))),
        ("pdfs_sweep_f32_dev", lambda: L.pdfs_sweep_f32_dev(None, None, 4, 4, 4, 1.0, 1.0, 1.0,
                                                            1, -1, None)),
        ("pdfs_solve_ref64_dev", lambda: L.pdfs_solve_ref64_dev(None, None, 4, 4, 4, 1.0, 1.0,
                                                                1.0, 1, 0.0, C.byref(st))),
        ("pdfs_solve_speed64_dev", lambda: L.pdfs_solve_speed64_dev(
            None, None, None, 4, 4, 4, 1.0, 1.0, 1.0, 1, 0.0, C.byref(st))),
        ("pdfs_judge_f32_dev", lambda: L.pdfs_judge_f32_dev(None, None, None, 4, 4, 4, 1.0, 1.0,
                                                            1.0, None, None)),
        ("pdfs_solve_batch_f32_dev", lambda: L.pdfs_solve_batch_f32_dev(
            None, 1, None, None, None, 1.0, 1.0, 1.0, C.byref(opts), C.byref(st), None, None)),
        ("generate", lambda: L.pdfs_generate_dev(None, pdfs_amd.GEN_POINT, 1, 4, 4, 4, 1.0, None,
                                                 None)),
        ("pdfs_warm_start_f32_dev", lambda: L.pdfs_warm_start_f32_dev(None, None, None, 4, 4, 4,
                                                                      1.0)),
        ("pdfs_solve_warm_f32_dev", lambda: L.pdfs_solve_warm_f32_dev(
            None, None, None, None, 4, 4, 4, 1.0, 1.0, 1.0, C.byref(opts), C.byref(st))),
        ("pdfs_solve_speed_f32_dev", lambda: L.pdfs_solve_speed_f32_dev(
            None, None, None, None, 4, 4, 4, 1.0, 1.0, 1.0, C.byref(opts), C.byref(st))),
        ("pdfs_solve_batch_speed_f32_dev", lambda: L.pdfs_solve_batch_speed_f32_dev(
            None, 1, None, None, None, None, 1.0, 1.0, 1.0, C.byref(opts), C.byref(st), None,
            None)),
        ("pdfs_judge_speed_f32_dev", lambda: L.pdfs_judge_speed_f32_dev(
            None, None, None, None, 4, 4, 4, 1.0, 1.0, 1.0, None, None)),
    ]
    for name, call in calls:
        assert call() == -1, name
        assert name in err(), (name, err())


REF_HOST = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle",
                        "_ref", "PDFS_ref_host")


@pytest.mark.skipif(not os.path.exists(REF_HOST), reason="reference host program not built")
def test_reference_host_program_links_run_fsm_from_libpdfs():
    """The reference's single-GPU host C, compiled unmodified, takes run_fsm
    from lib/libpdfs.so: the symbol is undefined in the program and the
    library resolves (the GPU test runs it)."""
    import subprocess
    nm = subprocess.run(["nm", "-D", REF_HOST], capture_output=True, text=True, check=True).stdout
    assert any(l.split()[-1] == "run_fsm" and " U " in f" {l} " for l in nm.splitlines())
    ldd = subprocess.run(["ldd", REF_HOST], capture_output=True, text=True, check=True).stdout
    line = next(l for l in ldd.splitlines() if "libpdfs.so" in l)
    assert "not found" not in line and os.path.samefile(line.split("=>")[1].split()[0],
                                                        pdfs_amd.LIB_PATH)


def test_opts_stats_layouts_and_enums_match_the_header(tmp_path):
    """The ctypes mirrors of pdfs_opts and pdfs_stats have the header's size
    and field offsets, and the enum values the module exports equal the
    header's (a C program built against include/pdfs.h prints them)."""
    import ctypes as C
    import re
    import subprocess

    opts = pdfs_amd.Opts
    stats = pdfs_amd.Stats
    header = open(os.path.join(ROOT, "include", "pdfs.h")).read()
    # the module keeps each name, or drops its PDFS_ prefix (GEN_*, TILE_AUTO)
    pyname = {}
    for e in sorted(set(re.findall(r"\b(PDFS_[A-Z0-9_]+)\s*=", header))):
        for n in (e, e[len("PDFS_"):]):
            if hasattr(pdfs_amd, n):
                pyname[e] = n
                break
    enums = sorted(pyname)
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pdfs.h"', "int main(void) {"]
    lines.append('  printf("opts %zu\\n", sizeof(pdfs_opts));')
    for name, _ in opts._fields_:
        lines.append(f'  printf("opts.{name} %zu\\n", offsetof(pdfs_opts, {name}));')
    lines.append('  printf("stats %zu\\n", sizeof(pdfs_stats));')
    for name, _ in stats._fields_:
        lines.append(f'  printf("stats.{name} %zu\\n", offsetof(pdfs_stats, {name}));')
    for e in enums:
        lines.append(f'  printf("{e} %d\\n", (int){e});')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    str(src)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                 text=True).stdout.splitlines())
    assert int(got["opts"]) == C.sizeof(opts)
    assert int(got["stats"]) == C.sizeof(stats)
    for name, _ in opts._fields_:
        assert int(got[f"opts.{name}"]) == getattr(opts, name).offset, name
    for name, _ in stats._fields_:
        assert int(got[f"stats.{name}"]) == getattr(stats, name).offset, name
    assert "PDFS_OVERLAP_AUTO" in enums and "PDFS_OV_AUTO_FIELDS" in enums
    assert "PDFS_TILE_AUTO" in enums and "PDFS_GEN_SPHERES" in enums
    for e in enums:
        assert int(got[e]) == getattr(pdfs_amd, pyname[e]), e
