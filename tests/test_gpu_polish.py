"""GPU tests of the FP64 column pipeline: the reference-arithmetic solve
(run_fsm, runFSM, pdfs_solve F64) and the FP64 polish that ends every
converged FP32 solve, through the C-ABI.

Bars:
* reference arithmetic: bit-exact against the oracle (which is bit-exact
  against the reference's own compiled core);
* polish: bit-exact against the oracle's FP32 solve + FP64 polish
  (orc_polish64) at sizes the oracle finishes in seconds;
* at the configurations' full sizes (terrain 256x256x128, spheres 512^3, the
  512-field batch): the plain, certified and warm-started FP32 results lie
  within F32_VS_F64_TOL of the converged FP64 reference-arithmetic solution
  computed on the GPU (itself bit-exact against the oracle at small sizes).
"""
import os

import numpy as np
import pytest
import torch

import oracle_ctypes as O
import pdfs_amd
from pdfs_amd import workloads

# these tests check the FP32-storage path (and its polish) against the FP32 oracle
S32 = pdfs_amd.PDFS_STORE_F32

pytestmark = pytest.mark.gpu

# max |FP32 result - FP64 solution| in units of h
F32_VS_F64_TOL = 1e-4

SPACINGS = [(1.0, 1.0, 1.0), (0.5, 0.7, 1.3)]
TILES64 = (pdfs_amd.TILE_16x16, pdfs_amd.TILE_16x8, pdfs_amd.TILE_16x16_1CU,
           pdfs_amd.TILE_16x8_2CU)


def _dense(a, dev):
    return torch.from_numpy(np.ascontiguousarray(O.interior(a)).astype(np.float32)).to(dev)


def _bits32(x):
    return np.ascontiguousarray(x).view(np.uint32)


@pytest.mark.parametrize("tile", TILES64)
@pytest.mark.parametrize("h", SPACINGS)
def test_ref64_column_tiles_bitwise(tile, h):
    """pdfs_solve in FP64 reference arithmetic on each FP64 column tile:
    one pass and converged, bit-exact against the oracle (and its pass
    count)."""
    a = O.gen_spheres(37, 29, 23, seed=17, n=5, r=(2.0, 5.0))
    want1 = a.copy()
    for sn in range(1, 9):
        O.sweep_ref64(want1, sn, h)
    got1 = a.copy()
    print(f"ref64 column tile={tile} h={h} one pass", flush=True)
    pdfs_amd.solve(got1, h=h, precision=pdfs_amd.PDFS_F64_REF, fixed_passes=1, tile64=tile, storage=S32)
    assert np.array_equal(want1.view(np.uint64), got1.view(np.uint64))
    want = a.copy()
    p_o, _ = O.solve_ref64(want, h=h, tol=1e-12)
    got = a.copy()
    print(f"ref64 column tile={tile} h={h} converged", flush=True)
    st = pdfs_amd.solve(got, h=h, precision=pdfs_amd.PDFS_F64_REF, tol=1e-12, tile64=tile, storage=S32)
    assert st["passes"] == p_o
    assert np.array_equal(want.view(np.uint64), got.view(np.uint64))


def test_ref64_ghosts_not_inf_take_the_level_path():
    """A padded field whose ghost layer is not all +inf (the column pipeline
    keeps ghosts implicit at +inf) runs on the per-level kernel, which reads
    the ghost cells as the reference does: bit-exact per ordering."""
    X, Y, Z = 21, 18, 14
    for sn in range(1, 9):
        a = O.random_field(X, Y, Z, seed=300 + sn)
        a[0, 3:9, 4:12] = 2.5          # z = 0 ghost plane
        a[5:9, :, -1] = 0.75           # x = X + 1 ghost plane
        b = a.copy()
        print(f"ref64 finite ghosts sn={sn}", flush=True)
        O.sweep_ref64(a, sn)
        pdfs_amd.runFSM(b, sn)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), sn


@pytest.mark.parametrize("shape", [(5, 4, 3), (17, 13, 9), (33, 40, 20), (48, 17, 31)])
@pytest.mark.parametrize("h", SPACINGS)
def test_polish_bitwise_vs_oracle(ctx, dev, shape, h):
    """The default FP32 solve (FP32 passes, then the FP64 polish) equals the
    oracle's FP32 solve followed by orc_polish64 bit for bit, with the same
    FP32 and polish pass counts."""
    X, Y, Z = shape
    a = O.gen_spheres(X, Y, Z, seed=40 + X, n=4, r=(1.5, 4.0))
    want, p32, pp = O.solve_f32_polished(a, h=h)
    t = _dense(a, dev)
    print(f"polish {shape} h={h}", flush=True)
    st = ctx.solve_f32(t, t.clone(), h=h, storage=S32)
    assert st["passes"] == p32 and st["polish_passes"] == pp, (st["passes"], p32,
                                                               st["polish_passes"], pp)
    assert np.array_equal(_bits32(t.cpu().numpy()), _bits32(O.interior(want)))
    assert st["judge_flagged"] == 0


@pytest.mark.parametrize("tile", TILES64)
def test_polish_tiles_and_batch_equal_single(ctx, dev, tile):
    """Uneven fields polished in one batch, on each FP64 tile, give each
    field the bits of its own polished solve."""
    rng = np.random.default_rng(77 + tile)
    fields = []
    for i in range(7):
        X, Y, Z = (int(v) for v in rng.integers(3, 45, 3))
        fields.append(O.gen_spheres(X, Y, Z, seed=800 + i, n=3, r=(1.5, 5.0)) if i % 2 else
                      O.gen_terrain(X, Y, Z, seed=800 + i, amp=(1.0, 6.0), sig=(2.0, 9.0)))
    wants = []
    for a in fields:
        w, _, _ = O.solve_f32_polished(a)
        wants.append(O.interior(w))
    ts = [_dense(a, dev) for a in fields]
    print(f"polish batch tile64={tile}", flush=True)
    st, passes, _ = ctx.solve_batch_f32(ts, [t.clone() for t in ts], tile64=tile, storage=S32)
    for i, (t, w) in enumerate(zip(ts, wants)):
        assert np.array_equal(_bits32(t.cpu().numpy()), _bits32(w)), i
    assert st["judge_flagged"] == 0 and st["polish_passes"] >= 1


def test_polish_raises_values_below_the_solution(ctx, dev):
    """Unlike a min(u, G(u)) sweep, the polish is the fixed-point map
    u = min(init, G(u)) itself: a region planted 0.5 h below the FP64
    solution comes back to it (within 1e-9 h before the FP32 rounding)."""
    a = O.gen_spheres(40, 36, 28, seed=5, n=4, r=(2.0, 6.0))
    u64 = a.copy()
    O.solve_d64(u64, tol=0.0)
    bad = u64.astype(np.float32)
    bad[8:14, 10:16, 12:20] -= 0.5
    t = torch.from_numpy(np.ascontiguousarray(O.interior(bad))).to(dev)
    init = _dense(a, dev)
    # FP32 min(u, G) passes cannot raise the planted region (they lower the
    # points downwind of it instead); the polish lifts all of them back
    st = ctx.solve_f32(t, init, storage=S32)
    got = t.cpu().numpy().astype(np.float64)
    err = np.abs(got - O.interior(u64)).max()
    print(f"planted region: polish passes {st['polish_passes']} max err {err:.3e}", flush=True)
    assert err <= 2e-5  # the FP32 rounding of values below 64 h


def _f64_solution(ctx, dev, kind, seed, X, Y, Z, nobj, p, q):
    pad = torch.empty((Z + 2, Y + 2, X + 2), dtype=torch.float64, device=dev)
    init = torch.empty((Z, Y, X), dtype=torch.float32, device=dev)
    ctx.generate_ex(kind, seed, X, Y, Z, 1.0, nobj, p, q, dense=init, padded=pad)
    st = ctx.solve_ref64(pad, passes=0, tol=1e-12)
    return init, pad[1:-1, 1:-1, 1:-1].contiguous(), st


def _err(f, u64):
    e = (f.double() - u64).abs()
    return float(e.max()), float(e.pow(2).mean().sqrt())


def _modes_within_tolerance(ctx, dev, init, u64, label):
    out = {}
    for mode in ("raw", "plain", "certified", "warm"):
        f = init.clone()
        print(f"{label}: {mode}", flush=True)
        if mode == "raw":
            st = ctx.solve_f32(f, init, polish=0, storage=S32)
        elif mode == "plain":
            st = ctx.solve_f32(f, init, storage=S32)
        elif mode == "certified":
            st = ctx.solve_f32(f, init, certify=1, storage=S32)
        else:
            if not os.path.exists(pdfs_amd.WEIGHTS_PATH):
                continue
            net = pdfs_amd.Net(ctx, path=pdfs_amd.WEIGHTS_PATH)
            try:
                st = ctx.solve_warm_f32(net, f, init, storage=S32)
            finally:
                net.close()
            assert st["fallbacks"] == 0
        emax, erms = _err(f, u64)
        out[mode] = (emax, erms, st["passes"], st["polish_passes"])
        print(f"{label} {mode}: max {emax:.3e} h rms {erms:.3e} h passes {st['passes']} "
              f"polish {st['polish_passes']} (max change {st['polish_maxdiff']:.2e})", flush=True)
        assert st["judge_flagged"] == 0
        del f
    for mode in ("plain", "certified", "warm"):
        if mode in out:
            assert out[mode][0] <= F32_VS_F64_TOL, (mode, out[mode])
    return out


def test_c1_terrain_modes_within_tolerance(ctx, dev):
    """c1: terrain 256x256x128 (8 hills, amplitude U(5,40) h, width
    U(10,60) h, seed 777): the FP32 fixed point alone misses the bound
    (about 1.0e-4 h); plain, certified and warm-started solves, each ending in
    the FP64 polish, are within 1e-4 h of the converged FP64 solution."""
    init, u64, st64 = _f64_solution(ctx, dev, pdfs_amd.GEN_TERRAIN, 777, 256, 256, 128, 8,
                                    (5.0, 40.0), (10.0, 60.0))
    print(f"c1 FP64 reference arithmetic: {st64['passes']} passes", flush=True)
    _modes_within_tolerance(ctx, dev, init, u64, "c1")


def test_c2_spheres512_modes_within_tolerance(ctx, dev):
    """c2: 64 spheres in 512^3 (radius U(4,32) h, seed 2024): the FP32 fixed
    point alone is 2.8e-4 h off; plain, certified and warm-started solves,
    each ending in the FP64 polish, are within 1e-4 h of the converged FP64
    solution."""
    init, u64, st64 = _f64_solution(ctx, dev, pdfs_amd.GEN_SPHERES, 2024, 512, 512, 512, 64,
                                    (4.0, 32.0), (0.0, 0.0))
    print(f"c2 FP64 reference arithmetic: {st64['passes']} passes", flush=True)
    out = _modes_within_tolerance(ctx, dev, init, u64, "c2")
    assert out["raw"][0] > F32_VS_F64_TOL  # the polish is what brings it inside
    del init, u64
    torch.cuda.empty_cache()


def test_c3_terrain_warm_start_bound_and_tolerance(ctx, dev):
    """c3 evaluation leg on a fresh 256x256x128 terrain (seed 9090): the
    committed network's warm start lies above the converged solution
    everywhere, the warm-started solve certifies without a fallback, and the
    result is within 1e-4 h of the FP64 solution."""
    if not os.path.exists(pdfs_amd.WEIGHTS_PATH):
        pytest.skip("no trained weights")
    init, u64, _ = _f64_solution(ctx, dev, pdfs_amd.GEN_TERRAIN, 9090, 256, 256, 128, 8,
                                 (5.0, 40.0), (10.0, 60.0))
    net = pdfs_amd.Net(ctx, path=pdfs_amd.WEIGHTS_PATH)
    try:
        w = torch.empty_like(init)
        net.warm_start(init, w)
        below = int((w.double() < u64 - 1e-3).sum())
        f = init.clone()
        st = ctx.solve_warm_f32(net, f, init, storage=S32)
    finally:
        net.close()
    emax, erms = _err(f, u64)
    print(f"c3 terrain: below {below}, fallbacks {st['fallbacks']}, max {emax:.3e} h, "
          f"passes {st['passes']} polish {st['polish_passes']}", flush=True)
    assert below == 0
    assert st["fallbacks"] == 0 and st["judge_flagged"] == 0
    assert emax <= F32_VS_F64_TOL


@pytest.mark.parametrize("mode", ["plain", "certified", "warm"])
def test_c4_batch_equals_single_solves(ctx, dev, mode):
    """c4: the 512 uneven fields (workloads.c4_specs) solved in shared
    persistent launches. Every field equals its own single solve in the same
    mode bit for bit; the judge flags nothing; fallbacks are counted; a
    sample of fields equals the oracle's FP32 solve + polish bit for bit."""
    specs = workloads.c4_specs()
    inits = workloads.generate_c4(ctx, specs)
    fields = [t.clone() for t in inits]
    net = None
    if mode == "warm":
        if not os.path.exists(pdfs_amd.WEIGHTS_PATH):
            pytest.skip("no trained weights")
        net = pdfs_amd.Net(ctx, path=pdfs_amd.WEIGHTS_PATH)
    try:
        print(f"c4 {mode}: batch", flush=True)
        if mode == "warm":
            st, passes, fb = ctx.solve_batch_warm_f32(net, fields, inits, graph=1,
                                                      tile=pdfs_amd.TILE_AUTO, storage=S32)
        else:
            st, passes, _ = ctx.solve_batch_f32(fields, inits, graph=1, tile=pdfs_amd.TILE_AUTO,
                                                certify=int(mode == "certified"), storage=S32)
            fb = [0] * len(fields)
        print(f"c4 {mode}: passes max {max(passes)}, fallbacks {sum(fb)}, polish passes "
              f"{st['polish_passes']}", flush=True)
        assert st["judge_flagged"] == 0 and st["polish_skipped"] == 0
        for i, (t, f) in enumerate(zip(inits, fields)):
            g = t.clone()
            if mode == "warm":
                s1 = ctx.solve_warm_f32(net, g, t, storage=S32)
                assert s1["fallbacks"] == fb[i], i
            else:
                s1 = ctx.solve_f32(g, t, certify=int(mode == "certified"), storage=S32)
            assert s1["passes"] == passes[i], (i, s1["passes"], passes[i])
            assert torch.equal(g.view(torch.int32), f.view(torch.int32)), (i, specs[i][:2])
    finally:
        if net is not None:
            net.close()
    if mode == "plain":
        small = [i for i, sp in enumerate(specs) if np.prod(sp[1]) <= 1_500_000][:6]
        for i in small:
            kind, (X, Y, Z) = specs[i][:2]
            a = np.full((Z + 2, Y + 2, X + 2), np.inf)
            a[1:-1, 1:-1, 1:-1] = inits[i].double().cpu().numpy()
            want, _, _ = O.solve_f32_polished(a)
            print(f"c4 field {i} {X}x{Y}x{Z} vs oracle", flush=True)
            assert np.array_equal(_bits32(fields[i].cpu().numpy()), _bits32(O.interior(want))), i
    del inits, fields
    torch.cuda.empty_cache()
