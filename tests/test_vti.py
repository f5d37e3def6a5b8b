"""CPU tests of the VTI reader/writer exported by libpdfs.so (no GPU)."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

import pdfs_amd
from vti_util import random_cli_input, reference_vti_text, write_input_vti


class Vti(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int), ("dx", C.c_double),
                ("dy", C.c_double), ("dz", C.c_double), ("location", C.POINTER(C.c_int)),
                ("distance", C.POINTER(C.c_double))]


def _lib():
    L = pdfs_amd.lib()
    L.pdfs_vti_read.argtypes = [C.c_char_p, C.POINTER(Vti)]
    L.pdfs_vti_read.restype = C.c_int
    L.pdfs_vti_free.argtypes = [C.POINTER(Vti)]
    L.pdfs_vti_write.argtypes = [C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                 C.c_double, C.c_double, C.c_double]
    L.pdfs_vti_write.restype = C.c_int
    return L


def read(path):
    L = _lib()
    v = Vti()
    rc = L.pdfs_vti_read(path.encode(), C.byref(v))
    if rc:
        return rc, None
    X, Y, Z = v.x, v.y, v.z
    n = (X + 2) * (Y + 2) * (Z + 2)
    loc = np.ctypeslib.as_array(v.location, (n,)).copy().reshape(Z + 2, Y + 2, X + 2)
    dist = np.ctypeslib.as_array(v.distance, (n,)).copy().reshape(Z + 2, Y + 2, X + 2)
    out = dict(dims=(X, Y, Z), h=(v.dx, v.dy, v.dz), loc=loc, dist=dist)
    L.pdfs_vti_free(C.byref(v))
    return 0, out


def test_reader_matches_layout(tmp_path):
    loc, dist = random_cli_input(7, 5, 4, 1)
    dist[0, 0, 0] = np.nan
    p = str(tmp_path / "in.vti")
    write_input_vti(p, loc, dist, h=(0.5, 0.25, 2.0))
    rc, v = read(p)
    assert rc == 0 and v["dims"] == (7, 5, 4) and v["h"] == (0.5, 0.25, 2.0)
    L, D = v["loc"], v["dist"]
    assert np.array_equal(L[1:-1, 1:-1, 1:-1], loc)
    ghost = np.ones(L.shape, bool)
    ghost[1:-1, 1:-1, 1:-1] = False
    assert np.all(L[ghost] == -1) and np.all(np.isinf(D[ghost]))
    inner = D[1:-1, 1:-1, 1:-1]
    assert np.isnan(inner[0, 0, 0])
    assert np.array_equal(inner.reshape(-1)[1:], dist.reshape(-1)[1:])


def test_reader_errors(tmp_path):
    assert read(str(tmp_path / "missing.vti"))[0] == -1
    p = tmp_path / "nohdr.vti"
    p.write_text("hello\n")
    assert read(str(p))[0] == -2
    loc, dist = random_cli_input(4, 3, 2, 2)
    p = str(tmp_path / "short.vti")
    write_input_vti(p, loc, dist)
    txt = open(p).read()
    open(p, "w").write(txt[: len(txt) // 2])
    assert read(p)[0] == -4


def test_writer_byte_for_byte(tmp_path):
    rng = np.random.default_rng(3)
    a = rng.uniform(-3, 900, size=(5, 4, 6))
    a[0, 0, 0] = -1.0
    p = str(tmp_path / "out.vti")
    h = (1.0, 0.5, 0.125)
    rc = _lib().pdfs_vti_write(p.encode(), a.ctypes.data, 6, 4, 5, *h)
    assert rc == 0
    assert open(p).read() == reference_vti_text(a, h)


def test_writer_output_is_not_rereadable_like_the_reference(tmp_path):
    """The writer puts Origin between WholeExtent and Spacing
    (file_writer.c:160-161); the reader's sscanf (vti_parser.c:92-94) cannot
    parse that, in the reference as here."""
    a = np.zeros((3, 3, 3))
    p = str(tmp_path / "o.vti")
    _lib().pdfs_vti_write(p.encode(), a.ctypes.data, 3, 3, 3, 1.0, 1.0, 1.0)
    assert read(p)[0] == -2


def _check_read(v, loc, dist):
    L, D = v["loc"], v["dist"]
    assert np.array_equal(L[1:-1, 1:-1, 1:-1], loc)
    assert np.array_equal(D[1:-1, 1:-1, 1:-1], dist)
    ghost = np.ones(L.shape, bool)
    ghost[1:-1, 1:-1, 1:-1] = False
    assert np.all(L[ghost] == -1) and np.all(np.isinf(D[ghost]))


def test_reader_parallel_path_large_file(tmp_path):
    """A file of several MiB is parsed in parallel chunks; the values land
    where the serial parse puts them."""
    loc, dist = random_cli_input(96, 64, 40, 11)
    p = str(tmp_path / "big.vti")
    write_input_vti(p, loc, dist, h=(0.5, 0.5, 0.5))
    assert os.path.getsize(p) > 2 << 20
    rc, v = read(p)
    assert rc == 0 and v["dims"] == (96, 64, 40)
    _check_read(v, loc, dist)


def test_reader_parallel_split_token_reads_like_serial(tmp_path):
    """'1.25-0.5' is two numbers to the serial strtod loop (as to the
    reference's fscanf); a chunked parse that sees one token must give the
    same values."""
    loc, dist = random_cli_input(96, 64, 40, 12)
    t = 70001
    dist.reshape(-1)[t] = 1.25
    dist.reshape(-1)[t + 1] = -0.5
    p = str(tmp_path / "split.vti")
    write_input_vti(p, loc, dist)
    txt = open(p).read()
    assert txt.count("1.25 -0.5") == 1
    open(p, "w").write(txt.replace("1.25 -0.5", "1.25-0.5"))
    rc, v = read(p)
    assert rc == 0
    _check_read(v, loc, dist)


def test_reader_parallel_errors(tmp_path):
    loc, dist = random_cli_input(96, 64, 40, 13)
    p = str(tmp_path / "trunc.vti")
    write_input_vti(p, loc, dist)
    txt = open(p).read()
    open(p, "w").write(txt[: int(len(txt) * 0.9)])
    assert read(p)[0] == -4
    t = txt.index("distance")
    k = txt.index(" ", t + 200000)
    open(p, "w").write(txt[:k] + " abc" + txt[k:])
    assert read(p)[0] == -4


def test_reader_rejects_extents_that_cannot_be_held(tmp_path):
    """Header extents of 2^30 or more per axis, or whose padded product
    passes 2^36, are a malformed header (-2), not an allocation of a
    wrapped-around size."""
    loc, dist = random_cli_input(4, 3, 2, 2)
    p = str(tmp_path / "ok.vti")
    write_input_vti(p, loc, dist)
    txt = open(p).read()
    for ext in ("0 1073741824 0 2 0 1", "0 3000000 0 3000000 0 3000000", "0 -5 0 2 0 1",
                "0 nan 0 2 0 1"):
        q = str(tmp_path / "bad.vti")
        open(q, "w").write(txt.replace("WholeExtent=\"0 3 0 2 0 1\"", f"WholeExtent=\"{ext}\"", 1))
        assert read(q)[0] == -2, ext


def _asan_gcc_works(tmp_path):
    src = tmp_path / "t.c"
    src.write_text("int main(void) { return 0; }\n")
    r = subprocess.run(["gcc", "-fsanitize=address,undefined", str(src), "-o", str(tmp_path / "t")],
                       capture_output=True, timeout=120)
    return r.returncode == 0


def test_reader_survives_mutated_files_under_sanitizers(tmp_path):
    """The reader (serial and parallel parse paths) on mutated input files,
    built with AddressSanitizer and UndefinedBehaviorSanitizer (leak checks
    on): every file ends in a documented return code with no memory error,
    undefined behaviour or leak, and every accepted file has in-range extents
    and the ghost layer (tests/vti_mutate.c)."""
    if shutil.which("gcc") is None or not _asan_gcc_works(tmp_path):
        pytest.skip("gcc with AddressSanitizer is not available")
    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(here)
    exe = tmp_path / "vti_mutate"
    subprocess.run(["gcc", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-fopenmp",
                    "-I" + os.path.join(root, "include"), os.path.join(here, "vti_mutate.c"),
                    os.path.join(root, "pdfs_amd", "csrc", "pdfs_vti.c"), "-lm", "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    rng = np.random.default_rng(11)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", OMP_NUM_THREADS="4")
    # small files parse serially; the large one (over 1 MiB of numbers) in
    # parallel chunks
    for name, (Z, Y, X), iters in (("small", (3, 4, 5), 1500), ("mid", (7, 9, 11), 1500),
                                   ("large", (40, 60, 60), 60)):
        seed = tmp_path / f"seed_{name}.vti"
        write_input_vti(seed, rng.integers(-1, 2, size=(Z, Y, X)),
                        rng.uniform(0.0, 50.0, size=(Z, Y, X)), (0.5, 1.0, 2.0))
        r = subprocess.run([str(exe), str(seed), str(iters), str(tmp_path / "cur.vti")], env=env,
                           capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, (name, r.stdout[-2000:], r.stderr[-4000:])
        counts = dict(zip(r.stdout.split()[0::2], map(int, r.stdout.split()[1::2])))
        assert sum(counts.values()) == iters and counts["ok"] > 0 and counts["data"] > 0, counts
