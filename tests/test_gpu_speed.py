"""GPU tests of spatially varying speed on the FP32 column pipeline
(pdfs_solve_speed_f32_dev, pdfs_solve_batch_speed_f32_dev,
pdfs_judge_speed_f32_dev): the Eikonal |grad u| = s with a dense FP32
slowness s = 1/F per point, through the C-ABI.

Bars:
* FP32 sweeps: bit-exact against the oracle (orc_solve_f32_speed) on every
  column variant that runs the mode, and bit-exact against the unit-speed
  solve for s = 1 and the doubled-spacing solve for s = 2;
* FP32 + FP64 polish: bit-exact against the oracle's FP32 speed solve
  followed by orc_polish64_speed;
* at the terrain configuration's full size (256 x 256 x 128) the polished
  result lies within 1e-4 h of the converged FP64 speed solution (reference
  arithmetic, pdfs_solve_speed64_dev, itself bit-exact against the oracle).
"""
import numpy as np
import pytest
import torch

import oracle_ctypes as O
import pdfs_amd

pytestmark = pytest.mark.gpu

# these tests check the FP32-storage path (and its polish) against the FP32 oracle
S32 = pdfs_amd.PDFS_STORE_F32

F32_VS_F64_TOL = 1e-4
SPACINGS = [(1.0, 1.0, 1.0), (0.5, 0.7, 1.3)]
SPEED_TILES = tuple(t for t in pdfs_amd.TILES if t != pdfs_amd.TILE_32x16)


def _dense(a, dev):
    return torch.from_numpy(np.ascontiguousarray(O.interior(a)).astype(np.float32)).to(dev)


def _bits32(x):
    return np.ascontiguousarray(x).view(np.uint32)


def _slowness(X, Y, Z, seed, lo=0.5, hi=2.0):
    """Padded FP32 slowness: layers in z times smooth lateral waves, in
    [lo, hi] (ghost values 1, never read)."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(Z + 2.0), np.arange(Y + 2.0), np.arange(X + 2.0),
                          indexing="ij")
    f = np.zeros(z.shape)
    for _ in range(3):
        k = rng.uniform(-0.3, 0.3, size=3)
        f += np.sin(k[0] * x + k[1] * y + k[2] * z + rng.uniform(0, 6.3))
    f += 0.8 * np.floor(z / max(2, Z // 5)) % 2
    f = (f - f.min()) / (f.max() - f.min())
    s = (lo + (hi - lo) * f).astype(np.float32)
    s[0], s[-1], s[:, 0], s[:, -1], s[:, :, 0], s[:, :, -1] = 1, 1, 1, 1, 1, 1
    return s


@pytest.mark.parametrize("tile", SPEED_TILES)
@pytest.mark.parametrize("h", SPACINGS)
def test_speed_f32_bitwise_vs_oracle(ctx, dev, tile, h):
    """FP32 speed sweeps to the fixed point on each column variant: the
    oracle's bits and pass count."""
    X, Y, Z = 41, 30, 19
    a = O.gen_spheres(X, Y, Z, seed=60 + tile, n=4, r=(1.5, 4.0))
    s = _slowness(X, Y, Z, seed=tile)
    want = a.astype(np.float32)
    p_o, _ = O.solve_f32_speed(want, s, h)
    t = _dense(a, dev)
    print(f"speed f32 tile={tile} h={h}", flush=True)
    st = ctx.solve_speed_f32(t, _dense(s, dev), h=h, tile=tile, polish=0)
    assert st["tile"] == tile
    assert st["passes"] == p_o, (st["passes"], p_o)
    assert np.array_equal(_bits32(t.cpu().numpy()), _bits32(O.interior(want)))


@pytest.mark.parametrize("h", SPACINGS)
def test_speed_unit_and_doubled_slowness(ctx, dev, h):
    """s = 1 is the unit-speed solve bit for bit; s = 2 is the unit-speed
    solve with doubled spacings bit for bit (both FP32, polish off)."""
    X, Y, Z = 37, 26, 22
    a = O.gen_terrain(X, Y, Z, seed=5, amp=(1.0, 8.0), sig=(2.0, 9.0))
    for sc in (1.0, 2.0):
        t1, t2 = _dense(a, dev), _dense(a, dev)
        print(f"speed s={sc} h={h}", flush=True)
        ctx.solve_speed_f32(t1, torch.full_like(t1, sc), h=h, polish=0)
        ctx.solve_f32(t2, h=tuple(sc * v for v in h), polish=0, storage=S32)
        assert torch.equal(t1.view(torch.int32), t2.view(torch.int32)), sc


def test_speed_32x16_request_falls_back(ctx, dev):
    """The 32 x 16 variant cannot hold the slowness delay line: asking for it
    (or for online tuning) runs the default variant, with the same bits."""
    X, Y, Z = 70, 50, 12
    a = O.gen_spheres(X, Y, Z, seed=2, n=3, r=(2.0, 4.0))
    s = _dense(_slowness(X, Y, Z, seed=4), dev)
    t1, t2, t3 = _dense(a, dev), _dense(a, dev), _dense(a, dev)
    print("speed tile 32x16 request", flush=True)
    st = ctx.solve_speed_f32(t1, s, tile=pdfs_amd.TILE_32x16, polish=0)
    assert st["tile"] != pdfs_amd.TILE_32x16
    st = ctx.solve_speed_f32(t3, s, tile=pdfs_amd.TILE_AUTO, polish=0)
    assert st["tile"] != pdfs_amd.TILE_32x16
    ctx.solve_speed_f32(t2, s, tile=pdfs_amd.TILE_16x16, polish=0)
    assert torch.equal(t1.view(torch.int32), t2.view(torch.int32))
    assert torch.equal(t3.view(torch.int32), t2.view(torch.int32))


@pytest.mark.parametrize("shape", [(5, 4, 3), (23, 17, 11), (40, 33, 27)])
@pytest.mark.parametrize("h", SPACINGS)
def test_speed_polish_bitwise_vs_oracle(ctx, dev, shape, h):
    """The default speed solve with an init (FP32 passes, then the FP64
    speed polish) equals the oracle's FP32 speed solve followed by
    orc_polish64_speed bit for bit, with the same pass counts; the speed
    judge passes it."""
    X, Y, Z = shape
    a = O.gen_spheres(X, Y, Z, seed=90 + X, n=4, r=(1.5, 4.0))
    s = _slowness(X, Y, Z, seed=X)
    want, p32, pp = O.solve_f32_speed_polished(a, s, h=h)
    t = _dense(a, dev)
    print(f"speed polish {shape} h={h}", flush=True)
    st = ctx.solve_speed_f32(t, _dense(s, dev), init=t.clone(), h=h)
    assert st["passes"] == p32 and st["polish_passes"] == pp, (st["passes"], p32,
                                                               st["polish_passes"], pp)
    assert np.array_equal(_bits32(t.cpu().numpy()), _bits32(O.interior(want)))
    assert st["judge_flagged"] == 0


def test_speed_judge_counts_match_oracle(ctx, dev):
    """The GPU speed judge flags exactly the points the oracle's speed judge
    flags, on a converged field with planted defects."""
    X, Y, Z = 45, 38, 21
    h = (0.6, 0.8, 1.1)
    a = O.gen_spheres(X, Y, Z, seed=13, n=4, r=(1.5, 4.0))
    s = _slowness(X, Y, Z, seed=13)
    v = a.astype(np.float32)
    O.solve_f32_speed(v, s, h)
    rng = np.random.default_rng(1)
    for _ in range(40):
        k, j, i = (int(rng.integers(1, n + 1)) for n in (X, Y, Z))
        v[i, j, k] -= np.float32(rng.uniform(1e-4, 1e-1))
    init = a.astype(np.float32)
    want = O.judge_f32_speed(v, init, s, h)
    assert want >= 40
    print("speed judge", flush=True)
    got = ctx.judge_speed_f32(_dense(v, dev), _dense(init, dev), _dense(s, dev), h=h)
    assert got == want
    assert ctx.judge_f32(_dense(v, dev), _dense(init, dev), h=h) != want


def test_speed_batch_equals_single(ctx, dev):
    """Uneven fields with their own slownesses in one batch (plain and with
    the polish) give each field the bits of its own solve."""
    rng = np.random.default_rng(31)
    fields, slows = [], []
    for i in range(6):
        X, Y, Z = (int(v) for v in rng.integers(3, 40, 3))
        fields.append(O.gen_spheres(X, Y, Z, seed=500 + i, n=3, r=(1.5, 5.0)))
        slows.append(_slowness(X, Y, Z, seed=500 + i, lo=0.3, hi=3.0))
    for polish in (0, 1):
        singles = []
        for a, s in zip(fields, slows):
            t = _dense(a, dev)
            ctx.solve_speed_f32(t, _dense(s, dev), init=t.clone() if polish else None,
                                polish=polish)
            singles.append(t.cpu())
        ts = [_dense(a, dev) for a in fields]
        print(f"speed batch polish={polish}", flush=True)
        st, passes, _ = ctx.solve_batch_speed_f32(
            ts, [_dense(s, dev) for s in slows],
            inits=[t.clone() for t in ts] if polish else None, polish=polish)
        for i, (t, w) in enumerate(zip(ts, singles)):
            assert torch.equal(t.cpu().view(torch.int32), w.view(torch.int32)), (polish, i)


def test_speed_layered_medium_exact_on_gpu(ctx, dev):
    """A source plane z = 1 in a z-layered medium (anisotropic spacings):
    the FP32 result is exactly the FP32 recursion u_k = u_{k-1} + fl(dz s_k)."""
    X, Y, Z = 33, 21, 40
    h = (0.75, 1.0, 0.5)
    a = np.full((Z, Y, X), 1e6, np.float32)
    a[0] = 0.0
    layers = np.random.default_rng(7).uniform(0.25, 4.0, size=Z).astype(np.float32)
    s = np.broadcast_to(layers[:, None, None], (Z, Y, X)).copy()
    t = torch.from_numpy(a).to(dev)
    print("speed layered", flush=True)
    ctx.solve_speed_f32(t, torch.from_numpy(s).to(dev), h=h, polish=0)
    got = t.cpu().numpy()
    u = np.float32(0.0)
    for k in range(Z):
        if k:
            u = np.float32(u + np.float32(np.float32(h[2]) * layers[k]))
        assert np.all(got[k] == u), k


def test_speed_terrain_full_size_within_tolerance(ctx, dev):
    """Terrain 256 x 256 x 128 (h = 1) in a layered, laterally varying medium
    (s in [0.5, 2]): the polished FP32 speed solve is within 1e-4 h of the
    converged FP64 speed solution; the certified stop too."""
    X, Y, Z = 256, 256, 128
    a = O.gen_terrain(X, Y, Z, seed=3)
    s = _slowness(X, Y, Z, seed=3)
    pad = torch.from_numpy(a.astype(np.float32).astype(np.float64)).to(dev)
    print("speed terrain FP64 reference solve", flush=True)
    ctx.solve_speed64(pad, torch.from_numpy(s.astype(np.float64)).to(dev), passes=0, tol=1e-12)
    ref = pad[1:-1, 1:-1, 1:-1].contiguous()
    sd = _dense(s, dev)
    errs = {}
    for name, kw in (("raw", dict(polish=0)), ("plain", {}), ("certified", dict(certify=1))):
        t = _dense(a, dev)
        print(f"speed terrain {name}", flush=True)
        st = ctx.solve_speed_f32(t, sd, init=t.clone(), **kw)
        fin = torch.isfinite(ref)
        errs[name] = float((t.double() - ref)[fin].abs().max())
        if name != "raw":
            assert st["judge_flagged"] == 0, name
    print("speed terrain max errors (h):", errs, flush=True)
    assert errs["plain"] <= F32_VS_F64_TOL and errs["certified"] <= F32_VS_F64_TOL


def test_speed_graph_replay_and_certified_batch(ctx, dev):
    """Passes replayed from a captured hipGraph give the eager speed solve's
    bits, on every replay (several solves in a row); the certified batch
    stop leaves every field within the speed judge's tolerance (judge
    flags 0) and no later than the plain stop."""
    X, Y, Z = 48, 40, 30
    a = O.gen_spheres(X, Y, Z, seed=71, n=4, r=(1.5, 4.0))
    s = _dense(_slowness(X, Y, Z, seed=71), dev)
    want = _dense(a, dev)
    ctx.solve_speed_f32(want, s, init=want.clone(), graph=0)
    for rep in range(3):
        t = _dense(a, dev)
        print(f"speed graph replay {rep}", flush=True)
        ctx.solve_speed_f32(t, s, init=t.clone(), graph=1)
        assert torch.equal(t.view(torch.int32), want.view(torch.int32)), rep
    rng = np.random.default_rng(9)
    fields, slows = [], []
    for i in range(5):
        X, Y, Z = (int(v) for v in rng.integers(8, 48, 3))
        fields.append(O.gen_terrain(X, Y, Z, seed=900 + i, amp=(1.0, 6.0), sig=(2.0, 9.0)))
        slows.append(_dense(_slowness(X, Y, Z, seed=900 + i), dev))
    ts = [_dense(a, dev) for a in fields]
    print("speed certified batch", flush=True)
    stc, pc, _ = ctx.solve_batch_speed_f32(ts, slows, inits=[t.clone() for t in ts], certify=1)
    assert stc["judge_flagged"] == 0
    for t, sl, a in zip(ts, slows, fields):
        assert ctx.judge_speed_f32(t, _dense(a, dev), sl) == 0
    tp = [_dense(a, dev) for a in fields]
    stp, pp, _ = ctx.solve_batch_speed_f32(tp, slows, inits=[t.clone() for t in tp])
    assert all(c <= p for c, p in zip(pc, pp)), (pc, pp)
