"""GPU test of the entry point that allocates nothing
(pdfs_sweep_f32_noalloc): through the C-ABI it gives the bits and the
change count of pdfs_sweep_f32_dev (the context entry point with allocated
scratch) ordering by ordering, on two fields in a row, the second smaller;
sizes beyond its scratch are rejected before any GPU work."""
import numpy as np
import pytest
import torch

import oracle_ctypes as O
import pdfs_amd

pytestmark = pytest.mark.gpu


def _dense(a, dev):
    return torch.from_numpy(np.ascontiguousarray(O.interior(a)).astype(np.float32)).to(dev)


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.5, 0.7, 1.3)])
def test_noalloc_sweep_equals_context_sweep(ctx, dev, h):
    for shape, seed in (((61, 47, 38), 11), ((23, 9, 14), 12)):
        X, Y, Z = shape
        a = O.random_field(X, Y, Z, seed=seed)
        want = _dense(a, dev)
        got = want.clone()
        res = torch.zeros(2, dtype=torch.int64, device=dev)
        for sn in range(1, 9):
            print(f"noalloc {shape} h={h} sn={sn}", flush=True)
            ch = ctx.sweep_f32(want, sn, h)
            before = int(res[0])
            pdfs_amd.sweep_f32_noalloc(got, sn, res, h)
            torch.cuda.current_stream().synchronize()
            assert int(res[1]) == 0
            assert int(res[0]) - before == ch, (shape, sn)
            assert torch.equal(got.view(torch.int32), want.view(torch.int32)), (shape, sn)


def test_noalloc_rejects_fields_beyond_its_scratch(dev):
    t = torch.zeros((4, 4, 4), dtype=torch.float32, device=dev)
    res = torch.zeros(2, dtype=torch.int64, device=dev)
    big = torch.zeros((600, 512, 512), dtype=torch.float32, device=dev)
    with pytest.raises(pdfs_amd.PdfsError) as e:
        pdfs_amd.sweep_f32_noalloc(big, 1, res)
    assert e.value.code == pdfs_amd.PDFS_EINVAL
    with pytest.raises(pdfs_amd.PdfsError):
        pdfs_amd.sweep_f32_noalloc(t, 9, res)
