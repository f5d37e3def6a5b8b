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
    sa = ctx.solve_f32(a, None, judge=0, tile=tile, storage=storage, polish=0, overlap=1)
    sb = ctx.solve_f32(b, None, judge=0, tile=tile, storage=storage, polish=0, overlap=0)
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
    sa, pa, _ = ctx.solve_batch_f32(fa, inits, overlap=1, **kw)
    sb, pb, _ = ctx.solve_batch_f32(fb, inits, overlap=0, **kw)
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
        st = ctx.solve_f32(t, None, judge=0, tile=pdfs_amd.TILE_16x16, overlap=1)
        if ref is None:
            ref = t.clone()
        assert torch.equal(t.view(torch.int32), ref.view(torch.int32)), r
    odd = _start(50, 37, 20, 3)
    t = _dev(odd, dev)
    st = ctx.solve_f32(t, None, judge=0, tile=pdfs_amd.TILE_16x16, overlap=1)
    assert st["sweep_launches"] >= 8 * (st["passes"] - 1)
    want, _, _, _ = O.solve_fixed(odd)
    assert np.array_equal(t.cpu().numpy().view(np.uint32), want.view(np.uint32))
