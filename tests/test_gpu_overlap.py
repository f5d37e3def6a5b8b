"""GPU tests of the overlapped passes (pdfs_opts.overlap: a pass, or half
pass, of orderings in one persistent launch, through the C-ABI): the same
bits and pass counts as one launch per ordering and as the oracle, for the
fixed-point and FP32 storage, single fields and batches whose extents are
multiples of the column tile, plain, certified and graph-replayed."""
import numpy as np
import pytest
import torch

import oracle_ctypes as O
import pdfs_amd

pytestmark = pytest.mark.gpu
ALWAYS, OFF = pdfs_amd.PDFS_OVERLAP_ALWAYS, pdfs_amd.PDFS_OVERLAP_OFF


def _start(X, Y, Z, seed, kind="spheres"):
    if kind == "spheres":
        a = O.gen_spheres(X, Y, Z, seed, n=6, r=(2.0, 0.15 * min(X, Y, Z) + 2.0))
    else:
        a = O.random_field(X, Y, Z, seed)
    return np.ascontiguousarray(O.interior(a)).astype(np.float32)


def _dev(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


@pytest.mark.parametrize("tile,shape", [(pdfs_amd.TILE_32x16, (64, 64, 40)),
                                        (pdfs_amd.TILE_32x16, (96, 48, 33)),
                                        (pdfs_amd.TILE_16x16, (48, 80, 57)),
                                        (pdfs_amd.TILE_16x8, (32, 40, 20))])
@pytest.mark.parametrize("storage", [pdfs_amd.PDFS_STORE_FIXED, pdfs_amd.PDFS_STORE_F32])
def test_overlap_equals_per_ordering_launches(ctx, dev, tile, shape, storage):
    X, Y, Z = shape
    start = _start(X, Y, Z, 31 + X)
    a, b = _dev(start, dev), _dev(start, dev)
    sa = ctx.solve_f32(a, None, judge=0, tile=tile, storage=storage, polish=0, overlap=ALWAYS)
    sb = ctx.solve_f32(b, None, judge=0, tile=tile, storage=storage, polish=0, overlap=OFF)
    print(f"{shape} tile {tile} storage {storage}: passes {sa['passes']} / {sb['passes']}, "
          f"launches {sa['sweep_launches']} / {sb['sweep_launches']}", flush=True)
    assert sa["passes"] == sb["passes"]
    assert sa["sweep_launches"] < sb["sweep_launches"]
    assert torch.equal(a.view(torch.int32), b.view(torch.int32))
    if storage == pdfs_amd.PDFS_STORE_FIXED:
        want, p_or, _, _ = O.solve_fixed(start)
        assert sa["passes"] == p_or
        assert np.array_equal(a.cpu().numpy().view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("certify", [0, 1])
def test_overlap_batch_equals_per_ordering(ctx, dev, graph, certify):
    shapes = [(64, 64, 40), (96, 64, 24), (32, 32, 60), (64, 32, 17)]
    starts = [_start(*s, 50 + i, "random" if i % 2 else "spheres") for i, s in enumerate(shapes)]
    fa = [_dev(s, dev) for s in starts]
    fb = [_dev(s, dev) for s in starts]
    inits = [_dev(s, dev) for s in starts]
    kw = dict(tile=pdfs_amd.TILE_32x16, graph=graph, certify=certify)
    sa, pa, _ = ctx.solve_batch_f32(fa, inits, overlap=ALWAYS, **kw)
    sb, pb, _ = ctx.solve_batch_f32(fb, inits, overlap=OFF, **kw)
    print(f"batch graph {graph} certify {certify}: passes {pa} / {pb}, launches "
          f"{sa['sweep_launches']} / {sb['sweep_launches']}", flush=True)
    assert pa == pb and sa["judge_flagged"] == 0
    for x, y in zip(fa, fb):
        assert torch.equal(x.view(torch.int32), y.view(torch.int32))


def test_overlap_repeated_solves_and_unaligned_fallback(ctx, dev):
    """Several overlapped solves in a row on one context (epochs, done flags
    and counters carry over) give the same bits each time; a field whose
    extents are not multiples of the tile runs one launch per ordering."""
    start = _start(64, 48, 30, 9)
    ref = None
    for r in range(4):
        t = _dev(start, dev)
        st = ctx.solve_f32(t, None, judge=0, tile=pdfs_amd.TILE_16x16, overlap=ALWAYS)
        if ref is None:
            ref = t.clone()
        assert torch.equal(t.view(torch.int32), ref.view(torch.int32)), r
    odd = _start(50, 37, 20, 3)
    t = _dev(odd, dev)
    st = ctx.solve_f32(t, None, judge=0, tile=pdfs_amd.TILE_16x16, overlap=ALWAYS)
    assert st["sweep_launches"] >= 8 * (st["passes"] - 1)
    want, _, _, _ = O.solve_fixed(odd)
    assert np.array_equal(t.cpu().numpy().view(np.uint32), want.view(np.uint32))


def test_auto_takes_overlap_for_one_field_not_for_many(ctx, dev):
    """The default (PDFS_OVERLAP_AUTO) overlaps the passes of one field whose
    ordering is dominated by its critical path, and keeps one launch per
    ordering for a batch of more fields than PDFS_OV_AUTO_FIELDS."""
    start = _start(64, 64, 40, 5)
    t = _dev(start, dev)
    st = ctx.solve_f32(t, None, judge=0, tile=pdfs_amd.TILE_32x16)
    assert st["sweep_launches"] * 4 <= st["orderings"]
    want, _, _, _ = O.solve_fixed(start)
    assert np.array_equal(t.cpu().numpy().view(np.uint32), want.view(np.uint32))
    n = pdfs_amd.PDFS_OV_AUTO_FIELDS + 1
    fs = [_dev(_start(64, 64, 40, 6 + i), dev) for i in range(n)]
    st3, _, _ = ctx.solve_batch_f32(fs, [f.clone() for f in fs], tile=pdfs_amd.TILE_32x16)
    assert st3["sweep_launches"] >= st3["orderings"]


def test_ticket_tables_follow_the_shapes(ctx, dev):
    """Alternating shapes on one context (the ticket table is kept while the
    shapes repeat and rebuilt when they change): every solve gives the bits
    of one launch per ordering."""
    shapes = [(64, 64, 40), (64, 64, 24), (64, 64, 40), (32, 96, 40), (64, 64, 40)]
    for i, s in enumerate(shapes):
        start = _start(*s, 70 + i)
        a, b = _dev(start, dev), _dev(start, dev)
        sa = ctx.solve_f32(a, None, judge=0, tile=pdfs_amd.TILE_16x16, overlap=ALWAYS)
        sb = ctx.solve_f32(b, None, judge=0, tile=pdfs_amd.TILE_16x16, overlap=OFF)
        assert sa["passes"] == sb["passes"], (i, s)
        assert torch.equal(a.view(torch.int32), b.view(torch.int32)), (i, s)


@pytest.mark.parametrize("certify", [0, 1])
def test_overlap_keeps_the_unchanged_ordering_streak(ctx, dev, certify):
    """A point source in a corner is exact after the pass's first ordering,
    so the orderings after it change nothing: one launch per ordering counts
    that streak and ends the field by the plain stop (its last pass counted,
    not launched) before the judge would run. An overlapped launch leaves
    the same streak, so both give the same passes and bits, alone and in a
    batch beside a field that needs more passes."""
    X, Y, Z = 32, 32, 32
    corner = np.full((Z, Y, X), np.inf, dtype=np.float32)
    corner[0, 0, 0] = 0.0
    other = _start(64, 64, 40, 12)
    res = {}
    for ov in (ALWAYS, OFF):
        t = _dev(corner, dev)
        s1 = ctx.solve_f32(t, _dev(corner, dev), certify=certify, tile=pdfs_amd.TILE_16x16,
                           overlap=ov)
        fs = [_dev(corner, dev), _dev(other, dev)]
        sb, pb, _ = ctx.solve_batch_f32(fs, [f.clone() for f in fs], certify=certify,
                                        tile=pdfs_amd.TILE_16x16, overlap=ov)
        res[ov] = (s1["passes"], t, pb, fs, s1["sweep_launches"], sb["sweep_launches"])
    a, b = res[ALWAYS], res[OFF]
    print(f"certify {certify}: single passes {a[0]} / {b[0]}, batch {a[2]} / {b[2]}, "
          f"launches {a[4]} / {b[4]}, {a[5]} / {b[5]}", flush=True)
    assert a[0] == b[0] and a[2] == b[2]
    assert a[4] < b[4] and a[5] < b[5]
    assert torch.equal(a[1].view(torch.int32), b[1].view(torch.int32))
    for x, y in zip(a[3], b[3]):
        assert torch.equal(x.view(torch.int32), y.view(torch.int32))


def _set64(ctx, v):
    assert pdfs_amd.lib().pdfsi_set_overlap64(ctx.c, v) == 0


@pytest.mark.parametrize("shape,h", [((64, 48, 40), (1.0, 1.0, 1.0)),
                                     ((48, 64, 24), (1.0, 0.7, 1.3))])
def test_overlap_fp64_reference_arithmetic(ctx, dev, shape, h):
    """The FP64 reference-arithmetic solve (the path behind run_fsm) with
    overlapped passes: after one and two passes and converged, the bits of
    the oracle's sweeps, and of one launch per ordering, in fewer launches."""
    X, Y, Z = shape
    a = O.gen_spheres(X, Y, Z, seed=X + Z, n=5, r=(2.0, 6.0))
    try:
        for passes in (1, 2, 0):
            want = a.copy()
            if passes:
                for _ in range(passes):
                    for sn in range(1, 9):
                        O.sweep_ref64(want, sn, h)
            else:
                p_or, _ = O.solve_ref64(want, h, tol=1e-12)
            got = {}
            for ov in (ALWAYS, OFF):
                _set64(ctx, ov)
                t = torch.from_numpy(a.copy()).to(dev)
                st = ctx.solve_ref64(t, passes=passes, tol=1e-12, h=h)
                got[ov] = (t.cpu().numpy(), st)
            print(f"{shape} h={h} passes={passes}: launches {got[ALWAYS][1]['sweep_launches']}"
                  f" / {got[OFF][1]['sweep_launches']}", flush=True)
            assert got[ALWAYS][1]["sweep_launches"] < got[OFF][1]["sweep_launches"]
            assert got[ALWAYS][1]["passes"] == got[OFF][1]["passes"]
            if not passes:
                assert got[ALWAYS][1]["passes"] == p_or
            for ov in (ALWAYS, OFF):
                assert np.array_equal(got[ov][0].view(np.uint64), want.view(np.uint64)), (passes, ov)
    finally:
        _set64(ctx, pdfs_amd.PDFS_OVERLAP_AUTO)


def test_overlap_speed_path(ctx, dev):
    """The FP32 speed sweep (a slowness per point) with overlapped passes,
    then its FP64 speed polish: the bits and passes of one launch per
    ordering."""
    X, Y, Z = 64, 48, 40
    start = _start(X, Y, Z, 21)
    z = np.arange(Z, dtype=np.float32)[:, None, None]
    x = np.arange(X, dtype=np.float32)[None, None, :]
    slow = (1.0 + 0.5 * np.sin(z / 7.0) + 0.25 * np.cos(x / 11.0)).astype(np.float32)
    slow = np.broadcast_to(slow, (Z, Y, X)).copy()
    s = _dev(slow, dev)
    res = {}
    for ov in (ALWAYS, OFF):
        t = _dev(start, dev)
        st = ctx.solve_speed_f32(t, s, init=_dev(start, dev), tile=pdfs_amd.TILE_32x8, overlap=ov)
        res[ov] = (t, st)
    a, b = res[ALWAYS], res[OFF]
    print(f"speed: passes {a[1]['passes']} / {b[1]['passes']}, launches "
          f"{a[1]['sweep_launches']} / {b[1]['sweep_launches']}", flush=True)
    assert a[1]["passes"] == b[1]["passes"] and a[1]["polish_passes"] == b[1]["polish_passes"]
    assert a[1]["sweep_launches"] < b[1]["sweep_launches"]
    assert torch.equal(a[0].view(torch.int32), b[0].view(torch.int32))


@pytest.mark.parametrize("shape,tile", [((32, 16, 40), pdfs_amd.TILE_32x16),
                                        ((16, 80, 24), pdfs_amd.TILE_16x16),
                                        ((128, 8, 12), pdfs_amd.TILE_16x8),
                                        ((64, 32, 1), pdfs_amd.TILE_32x16),
                                        ((48, 48, 2), pdfs_amd.TILE_16x16)])
def test_overlap_degenerate_shapes(ctx, dev, shape, tile):
    """One column per ordering, one-tile-wide strips, a single plane and two
    planes: overlapped passes give the bits and passes of one launch per
    ordering and of the oracle."""
    X, Y, Z = shape
    start = _start(X, Y, Z, 3 * X + Z, "random")
    a, b = _dev(start, dev), _dev(start, dev)
    sa = ctx.solve_f32(a, None, judge=0, tile=tile, overlap=ALWAYS)
    sb = ctx.solve_f32(b, None, judge=0, tile=tile, overlap=OFF)
    print(f"{shape}: passes {sa['passes']} / {sb['passes']}, launches {sa['sweep_launches']} / "
          f"{sb['sweep_launches']}", flush=True)
    assert sa["passes"] == sb["passes"]
    assert sa["sweep_launches"] < sb["sweep_launches"]
    assert torch.equal(a.view(torch.int32), b.view(torch.int32))
    want, p_or, _, _ = O.solve_fixed(start)
    assert sa["passes"] == p_or
    assert np.array_equal(a.cpu().numpy().view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (1.0, 0.8, 1.25)])
def test_overlap_fp64_polish(ctx, dev, h):
    """FP32 storage then the FP64 polish (modes 3/4), both overlapped: the
    polish stops on its last ordering's largest change, which the overlapped
    launch keeps apart. Same bits and passes as one launch per ordering, and
    (isotropic) as the oracle's FP32 solve and polish."""
    X, Y, Z = 64, 48, 40
    padded = O.gen_spheres(X, Y, Z, seed=77, n=5, r=(2.0, 6.0)).astype(np.float32).astype(np.float64)
    start = np.ascontiguousarray(O.interior(padded)).astype(np.float32)
    res = {}
    for ov in (ALWAYS, OFF):
        t = _dev(start, dev)
        st = ctx.solve_f32(t, _dev(start, dev), h=h, judge=0, storage=pdfs_amd.PDFS_STORE_F32,
                           tile=pdfs_amd.TILE_16x16, overlap=ov)
        res[ov] = (t, st)
    a, b = res[ALWAYS], res[OFF]
    print(f"polish h={h}: passes {a[1]['passes']} / {b[1]['passes']}, polish passes "
          f"{a[1]['polish_passes']} / {b[1]['polish_passes']}", flush=True)
    assert a[1]["polish_passes"] >= 1
    assert a[1]["passes"] == b[1]["passes"] and a[1]["polish_passes"] == b[1]["polish_passes"]
    assert a[1]["sweep_launches"] < b[1]["sweep_launches"]
    assert torch.equal(a[0].view(torch.int32), b[0].view(torch.int32))
    if h == (1.0, 1.0, 1.0):
        want, p32, ppol = O.solve_f32_polished(padded)
        assert ppol == a[1]["polish_passes"]
        assert np.array_equal(a[0].cpu().numpy().view(np.uint32), O.interior(want).view(np.uint32))
