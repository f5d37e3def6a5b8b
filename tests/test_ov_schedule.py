"""CPU checks of the overlapped-pass ticket order (pdfsi_ov_order, the host
function that orders the tickets of k_sweep_col<..., true>), and of the
argument checks of its launcher.

The kernel's waits are restated here from its source: a ticket (slot o,
field f, virtual tile a, b, physical column p) waits for the edge entries of
its upwind tiles (a - 1, b) and (a, b - 1) in slot o, for the done flags of
slot o - 1 at p and at its physical neighbours p -/+ nTy (x) and p -/+ 1 (y),
and for every ticket of slot o - 2 (its edge region is reused). The
persistent grid cannot deadlock when every waited-on ticket comes earlier in
the order, and the sweep gives the bits of consecutive orderings when the
waits cover the planes a column reads (its own and its four neighbours').
"""
import ctypes as C
import random

import numpy as np
import pytest

import pdfs_amd


def dirs(sn):
    desc = sn in (2, 5, 7, 8)
    return int(desc ^ (sn in (4, 8))), int(desc ^ (sn in (2, 6)))


def order(tiles, active, TX, TY, sn0, sn1, grid, Zs):
    """tiles: [(nTx, nTy)] per field; active: field indices whose columns
    are in the plan (all their tiles, in the virtual ready-key order)."""
    nf = len(tiles)
    cols = []
    for f in active:
        nTx, nTy = tiles[f]
        for a in range(nTx):
            for b in range(nTy):
                cols.append((a * (TX + 6) + b * (TY + 6), f, a, b))
    cols.sort()
    flat = [v for _, f, a, b in cols for v in (f, a, b)]
    nc = len(cols)
    ns = sn1 - sn0 + 1
    out = (C.c_int * (5 * max(1, ns * nc)))()
    I = C.c_int * max(1, nf)
    L = pdfs_amd.lib()
    n = L.pdfsi_ov_order(nf, I(*[t[0] for t in tiles]), I(*[t[1] for t in tiles]), I(*Zs), nc,
                         (C.c_int * max(1, len(flat)))(*flat), TX, TY, sn0, sn1, grid, out)
    return n, np.array(out[:5 * max(0, n)], dtype=np.int64).reshape(-1, 5), nc


def check(tiles, active, TX, TY, sn0, sn1, grid, Zs):
    n, t, nc = order(tiles, active, TX, TY, sn0, sn1, grid, Zs)
    ns = sn1 - sn0 + 1
    assert n == ns * nc
    pbase = np.concatenate([[0], np.cumsum([x * y for x, y in tiles])])
    pos_v = {}   # (o, f, a, b) -> position
    pos_p = {}   # (o, p) -> position
    for i, (o, f, a, b, p) in enumerate(t):
        nTx, nTy = tiles[f]
        fx, fy = dirs(sn0 + o)
        pa, pb = (nTx - 1 - a if fx else a), (nTy - 1 - b if fy else b)
        assert p == pbase[f] + pa * nTy + pb
        assert (o, f, a, b) not in pos_v
        pos_v[(o, f, a, b)] = i
        pos_p[(o, p)] = i
    assert len(pos_v) == ns * nc
    last_of_slot = {o: max(i for (oo, *_), i in pos_v.items() if oo == o) for o in range(ns)}
    deps = []
    for i, (o, f, a, b, p) in enumerate(t):
        nTx, nTy = tiles[f]
        fx, fy = dirs(sn0 + o)
        pa, pb = (nTx - 1 - a if fx else a), (nTy - 1 - b if fy else b)
        d = []
        if a > 0:
            d.append(pos_v[(o, f, a - 1, b)])
        if b > 0:
            d.append(pos_v[(o, f, a, b - 1)])
        if o > 0:
            q = [p]
            if pa > 0:
                q.append(p - nTy)
            if pa + 1 < nTx:
                q.append(p + nTy)
            if pb > 0:
                q.append(p - 1)
            if pb + 1 < nTy:
                q.append(p + 1)
            for qq in q:
                j = pos_p[(o - 1, qq)]
                # the waited flag belongs to the same field and a tile at or
                # next to this one
                oo, f2, a2, b2, p2 = t[j]
                assert f2 == f and p2 == qq
                d.append(j)
        if o > 1:
            d.append(last_of_slot[o - 2])
        assert all(j < i for j in d), (i, d)
        deps.append(d)
    # worker-pool emulation: `grid` workers take tickets in order; a ticket
    # ends after its dependencies end (random durations); every ticket ends
    rng = random.Random(7)
    free = [0.0] * min(grid, len(t))
    end = [0.0] * len(t)
    for i in range(len(t)):
        w = min(range(len(free)), key=lambda k: free[k])
        st = max([free[w]] + [end[j] for j in deps[i]])
        end[i] = st + rng.uniform(0.5, 1.5)
        free[w] = end[i]
    return t


@pytest.mark.parametrize("seg", [(1, 8), (1, 4), (5, 8)])
def test_single_512_field(seg):
    # 512^3 with 32 x 16 tiles: 16 x 32 columns, 256 concurrent
    check([(16, 32)], [0], 32, 16, seg[0], seg[1], 256, [512])


@pytest.mark.parametrize("grid", [1, 7, 256, 5000])
def test_batch_and_grids(grid):
    tiles = [(16, 32)] * 3 + [(2, 3), (1, 1), (1, 5), (4, 1)]
    check(tiles, list(range(len(tiles))), 32, 16, 1, 8, grid, [512, 256, 128, 64, 32, 100, 33])


def test_active_subset_and_small_tiles():
    tiles = [(3, 4), (5, 5), (1, 2), (8, 8)]
    check(tiles, [1, 3], 16, 16, 1, 8, 13, [40, 70, 20, 96])
    check(tiles, [0, 2], 8, 8, 5, 8, 3, [17, 5, 9, 64])


def test_order_follows_the_schedule():
    """Tickets of slot 1 start before slot 0 has finished (the overlap), and
    the first ticket of each later slot follows the last of two slots
    before."""
    t = check([(16, 32)], [0], 32, 16, 1, 8, 256, [512])
    first = {o: int(np.argmax(t[:, 0] == o)) for o in range(8)}
    last = {o: int(len(t) - 1 - np.argmax(t[::-1, 0] == o)) for o in range(8)}
    assert first[1] < last[0]
    for o in range(2, 8):
        assert first[o] > last[o - 2]


class _Batch(C.Structure):
    # pdfsk_batch (pdfs_amd/csrc/pdfs_kernels.h)
    _fields_ = [("fields", C.c_void_p), ("cols", C.c_void_p), ("state", C.c_void_p),
                ("changed", C.c_void_p), ("streak", C.c_void_p), ("nfields", C.c_int),
                ("maxdiff", C.c_void_p), ("maxdiff_last", C.c_void_p), ("err", C.c_void_p),
                ("edge_x", C.c_void_p), ("edge_y", C.c_void_p), ("edge_x_bytes", C.c_size_t),
                ("edge_y_bytes", C.c_size_t)]


def test_overlapped_launcher_rejects_bad_arguments():
    """pdfsk_sweep_batch_ov refuses, before any HIP call, orderings outside
    1..8, an empty or missing ticket table, missing flag or counter arrays,
    a tile the mode cannot run, and an FP64 or fixed-point mode without its
    change maxima. (Accepted arguments are not launched here: no GPU.)"""
    L = pdfs_amd.lib()
    f = L.pdfsk_sweep_batch_ov
    vp = C.c_void_p
    f.argtypes = [C.POINTER(_Batch), C.c_int, C.c_int, vp, C.c_int, C.c_int, vp, vp, vp,
                  C.c_size_t, C.c_size_t, C.c_int, vp, vp, C.c_int, C.c_int, vp]
    f.restype = C.c_int
    b = _Batch()
    b.nfields = 1
    b.maxdiff = 4096
    dummy = 8192

    def call(sn0=1, sn1=8, otab=dummy, nt=8, done=dummy, cnt=dummy, chg=dummy, mode=0,
             tile=pdfs_amd.TILE_32x16, grid=1):
        return f(C.byref(b), sn0, sn1, otab, nt, 1, done, cnt, chg, 0, 0, mode, None, None,
                 tile, grid, None)

    assert call(sn0=0) == -1 and call(sn1=9) == -1 and call(sn0=5, sn1=4) == -1
    assert call(otab=None) == -1 and call(nt=0) == -1 and call(grid=0) == -1
    assert call(done=None) == -1 and call(cnt=None) == -1 and call(chg=None) == -1
    assert call(mode=5, tile=pdfs_amd.TILE_32x16) == -1  # the speed sweep has no 32x16 variant
    b.maxdiff = None
    assert call(mode=7) == -1 and call(mode=2, tile=pdfs_amd.TILE_8x8) == -1
