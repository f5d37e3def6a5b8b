"""GPU tests of the fixed-point storage of the FP32 path (pdfs_opts.storage =
PDFS_STORE_FIXED, the default for converged solves with equal spacings and
unit speed), through the C-ABI.

Bars:
* bit-exact against the oracle's fixed-point pipeline (orc_fx_encode,
  orc_sweep_fx / orc_solve_fx, orc_fx_decode) per pass and converged, on
  ragged, tiny and cubic shapes, two spacings and every tile variant;
* the fixed-point judge's counts and flag bits equal orc_judge_fx's;
* batches: every field equals its own single solve bit for bit;
* at the configurations' full sizes (terrain 256x256x128, spheres 512^3, the
  512-field batch): plain, certified and warm-started results lie within
  F32_VS_F64_TOL of the converged FP64 reference-arithmetic solution computed
  on the GPU (bit-exact against the oracle at small sizes), with no polish.
"""
import os

import numpy as np
import pytest
import torch

import oracle_ctypes as O
import pdfs_amd
from pdfs_amd import workloads

pytestmark = pytest.mark.gpu

FIXED = pdfs_amd.PDFS_STORE_FIXED
# max |result - FP64 solution| in units of h
F32_VS_F64_TOL = 1e-4
SHAPES = [(5, 4, 3), (17, 13, 9), (33, 40, 20), (48, 17, 31), (64, 64, 64), (100, 37, 21)]


def _bits32(x):
    return np.ascontiguousarray(x).view(np.uint32)


def _start(kind, X, Y, Z, seed, h=1.0):
    if kind == "spheres":
        a = O.gen_spheres(X, Y, Z, seed, h=h, n=6, r=(1.5, 0.2 * min(X, Y, Z) + 2.0))
    else:
        a = O.random_field(X, Y, Z, seed)
        a[1:-1, 1:-1, 1:-1] *= h
    return np.ascontiguousarray(O.interior(a)).astype(np.float32)


def _dev(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("h", [1.0, 0.37])
@pytest.mark.parametrize("kind", ["spheres", "random"])
def test_fixed_converged_bitwise_vs_oracle(ctx, dev, shape, h, kind):
    X, Y, Z = shape
    start = _start(kind, X, Y, Z, 11 + X, h)
    want, p_or, (k, base, d), _ = O.solve_fixed(start, h=h)
    t = _dev(start, dev)
    init = t.clone()
    st = ctx.solve_f32(t, init, h=(h, h, h))
    print(f"{kind} {shape} h={h}: passes gpu {st['passes']} oracle {p_or}, k={k}", flush=True)
    assert st["storage"] == FIXED and st["polish_passes"] == 0
    assert st["passes"] == p_or
    assert st["judge_flagged"] == 0
    assert np.array_equal(_bits32(t.cpu().numpy()), _bits32(want))


@pytest.mark.parametrize("stop", [0, 1, 4, 50])
def test_fixed_stop_rule_matches_oracle(ctx, dev, stop):
    """The plain stop (a pass whose largest drop is at most stop_quanta
    quanta; 0: a pass that changes nothing) ends on the oracle's pass, with
    its bits, on fields that need several passes."""
    for shape, kind in (((33, 40, 20), "spheres"), ((64, 64, 64), "random"),
                        ((90, 20, 70), "spheres")):
        start = _start(kind, *shape, 3 + stop)
        want, p_or, _, _ = O.solve_fixed(start, stop_quanta=stop)
        t = _dev(start, dev)
        st = ctx.solve_f32(t, None, judge=0, stop_quanta=stop)
        print(f"stop {stop} {kind} {shape}: passes gpu {st['passes']} oracle {p_or}", flush=True)
        assert st["passes"] == p_or, (shape, st["passes"], p_or)
        assert np.array_equal(_bits32(t.cpu().numpy()), _bits32(want)), shape


@pytest.mark.parametrize("tile", pdfs_amd.TILES)
def test_fixed_per_pass_bitwise_every_tile(ctx, dev, tile):
    """One and two passes (capped solves) on a ragged random field equal the
    oracle's orderings 1..8 (1..16) on the encoded field, on every tile."""
    X, Y, Z = 48, 17, 31
    start = _start("random", X, Y, Z, 5)
    k, base, d = O.fx_params(start)
    q = O.fx_padded(O.fx_encode(start, base, d))
    for npass in (1, 2):
        for sn in range(1, 9):
            O.sweep_fx(q, sn, k)
        want = O.fx_decode(O.interior(q).copy(), base, d)
        t = _dev(start, dev)
        st = ctx.solve_f32(t, None, judge=0, tile=tile, max_passes=npass, allow_notconv=True)
        assert st["storage"] == FIXED and st["passes"] == npass
        assert np.array_equal(_bits32(t.cpu().numpy()), _bits32(want)), (tile, npass)


@pytest.mark.parametrize("graph", [0, 1])
def test_fixed_batch_equals_single_and_oracle(ctx, dev, graph):
    shapes = [(40, 33, 29), (17, 13, 9), (64, 48, 40), (5, 4, 3), (96, 72, 24)]
    starts = [_start("spheres" if i % 2 else "random", *s, 100 + i) for i, s in enumerate(shapes)]
    fields = [_dev(s, dev) for s in starts]
    inits = [f.clone() for f in fields]
    st, passes, _ = ctx.solve_batch_f32(fields, inits, graph=graph, tile=pdfs_amd.TILE_AUTO)
    assert st["storage"] == FIXED and st["judge_flagged"] == 0
    for i, s in enumerate(starts):
        want, p_or, _, _ = O.solve_fixed(s)
        assert passes[i] == p_or, i
        assert np.array_equal(_bits32(fields[i].cpu().numpy()), _bits32(want)), i
        g = _dev(s, dev)
        s1 = ctx.solve_f32(g, inits[i])
        assert s1["passes"] == passes[i]
        assert torch.equal(g.view(torch.int32), fields[i].view(torch.int32)), i


def test_fixed_certified_stops_early_within_tolerance(ctx, dev):
    """The certified stop (fixed-point judge, tolerance judge_tol h + 4
    quanta) needs no more passes than the plain stop and stays within the
    bound of the converged FP64 solution."""
    X, Y, Z = 120, 96, 80
    a = O.gen_spheres(X, Y, Z, 7, n=8, r=(2.0, 12.0))
    ref = a.copy()
    O.solve_d64(ref, tol=0.0)
    start = np.ascontiguousarray(O.interior(a)).astype(np.float32)
    t = _dev(start, dev)
    init = t.clone()
    sp = ctx.solve_f32(t.clone(), init)
    sc = ctx.solve_f32(t, init, certify=1)
    err = np.abs(t.cpu().numpy().astype(np.float64) - O.interior(ref)).max()
    print(f"certified: passes {sc['passes']} (plain {sp['passes']}) orderings {sc['orderings']} "
          f"(plain {sp['orderings']}) max err {err:.3e} h", flush=True)
    assert sc["judge_flagged"] == 0 and sc["storage"] == FIXED
    assert sc["passes"] <= sp["passes"] and sc["orderings"] <= sp["orderings"]
    assert err <= F32_VS_F64_TOL


@pytest.mark.parametrize("shape", [(33, 40, 20), (70, 9, 65)])
def test_fixed_judge_counts_and_flags_equal_oracle(ctx, dev, shape):
    """Planted defects (points lowered or raised by 3 to 300 quanta beyond
    the tolerance, at faces, z-tile edges and corners, and +inf points) are
    flagged exactly where orc_judge_fx flags them; the converged field
    passes."""
    X, Y, Z = shape
    start = _start("spheres", X, Y, Z, 21)
    _, _, (k, base, d), q = O.solve_fixed(start)
    init_p = np.full((Z + 2, Y + 2, X + 2), np.inf, np.float32)
    init_p[1:-1, 1:-1, 1:-1] = start
    tol = 2e-6 + 4 * d
    init_t = _dev(start, dev)
    qi = O.interior(q).copy()
    assert ctx.judge_fixed(_dev(qi.view(np.int32), dev), init_t, 1.0, base, d, tol) == 0
    assert O.judge_fx(q, init_p, 1.0, base, d, tol) == 0
    rng = np.random.default_rng(3)
    bad = qi.copy()
    pts = [(0, 0, 0), (Z - 1, Y - 1, X - 1), (31 % Z, 5 % Y, 7 % X), (32 % Z, 5 % Y, 7 % X),
           (Z // 2, 0, X // 2), (Z // 2, Y // 2, 0)]
    pts += [tuple(int(rng.integers(0, n)) for n in (Z, Y, X)) for _ in range(20)]
    for i, (z, y, x) in enumerate(pts):
        step = int(tol / d) + 3 + 13 * i
        v = int(bad[z, y, x])
        if v == O.FX_INF:
            continue
        bad[z, y, x] = v - step if (i % 2 == 0 and v > step) else v + step
    bad[Z // 3, Y // 3, X // 3] = O.FX_INF
    qb = O.fx_padded(bad)
    n_or, fl_or = O.judge_fx(qb, init_p, 1.0, base, d, tol, want_flags=True)
    nwords = (X * Y * Z + 31) // 32
    flags = torch.zeros(nwords, dtype=torch.int32, device=dev)
    n_gpu = ctx.judge_fixed(_dev(bad.view(np.int32), dev), init_t, 1.0, base, d, tol, flags=flags)
    bits = np.unpackbits(flags.cpu().numpy().view(np.uint8), bitorder="little")[:X * Y * Z]
    print(f"judge: gpu {n_gpu} oracle {n_or}", flush=True)
    assert n_gpu == n_or and n_or > 0
    assert np.array_equal(bits.reshape(Z, Y, X), O.interior(fl_or))


def test_fixed_storage_fallbacks(ctx, dev):
    """A start field holding -inf is swept in FP32 storage (its bits are the
    FP32-storage solve's); anisotropic spacing, a fixed pass count and a
    speed field also keep FP32 storage; NaN start values act as +inf."""
    X, Y, Z = 20, 18, 16
    start = _start("spheres", X, Y, Z, 4)
    s2 = start.copy()
    s2[3, 4, 5] = -np.inf
    t = _dev(s2, dev)
    st = ctx.solve_f32(t, None, judge=0)
    u = _dev(s2, dev)
    ctx.solve_f32(u, None, judge=0, storage=pdfs_amd.PDFS_STORE_F32, polish=0)
    assert st["storage"] == pdfs_amd.PDFS_STORE_F32
    assert torch.equal(t.view(torch.int32), u.view(torch.int32))
    t = _dev(start, dev)
    assert ctx.solve_f32(t, t.clone(), h=(1.0, 1.0, 1.5))["storage"] == pdfs_amd.PDFS_STORE_F32
    t = _dev(start, dev)
    assert ctx.solve_f32(t, None, judge=0, fixed_passes=1)["storage"] == pdfs_amd.PDFS_STORE_F32
    s3 = start.copy()
    s4 = start.copy()
    far = np.argwhere(start > 1e5)[0]
    s3[tuple(far)] = np.nan
    s4[tuple(far)] = np.inf
    a, b = _dev(s3, dev), _dev(s4, dev)
    ctx.solve_f32(a, None, judge=0)
    ctx.solve_f32(b, None, judge=0)
    assert torch.equal(a.view(torch.int32), b.view(torch.int32))


# ---- full-size configurations ------------------------------------------------

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
    for mode in ("plain", "certified", "warm"):
        f = init.clone()
        print(f"{label}: {mode}", flush=True)
        if mode == "plain":
            st = ctx.solve_f32(f, init)
        elif mode == "certified":
            st = ctx.solve_f32(f, init, certify=1)
        else:
            if not os.path.exists(pdfs_amd.WEIGHTS_PATH):
                continue
            net = pdfs_amd.Net(ctx, path=pdfs_amd.WEIGHTS_PATH)
            try:
                st = ctx.solve_warm_f32(net, f, init)
            finally:
                net.close()
            assert st["fallbacks"] == 0
        emax, erms = _err(f, u64)
        out[mode] = (emax, erms, st["passes"], st["orderings"])
        print(f"{label} {mode}: max {emax:.3e} h rms {erms:.3e} h passes {st['passes']} "
              f"orderings {st['orderings']}", flush=True)
        assert st["judge_flagged"] == 0 and st["storage"] == FIXED
        assert st["polish_passes"] == 0
        assert emax <= F32_VS_F64_TOL, (mode, emax)
        del f
    return out


def test_c1_terrain_fixed_within_tolerance(ctx, dev):
    """c1: terrain 256x256x128 (8 hills, amplitude U(5,40) h, width
    U(10,60) h, seed 777), fixed-point storage, no polish."""
    init, u64, st64 = _f64_solution(ctx, dev, pdfs_amd.GEN_TERRAIN, 777, 256, 256, 128, 8,
                                    (5.0, 40.0), (10.0, 60.0))
    _modes_within_tolerance(ctx, dev, init, u64, "c1")


def test_c2_spheres512_fixed_within_tolerance(ctx, dev):
    """c2: 64 spheres in 512^3 (radius U(4,32) h, seed 2024), fixed-point
    storage, no polish: plain, certified and warm-started."""
    init, u64, st64 = _f64_solution(ctx, dev, pdfs_amd.GEN_SPHERES, 2024, 512, 512, 512, 64,
                                    (4.0, 32.0), (0.0, 0.0))
    _modes_within_tolerance(ctx, dev, init, u64, "c2")
    del init, u64
    torch.cuda.empty_cache()


def test_c3_terrain_warm_fixed_within_tolerance(ctx, dev):
    """c3 evaluation leg on a fresh 256x256x128 terrain (seed 9090), fixed
    point: the warm-started solve certifies without a fallback within the
    bound."""
    if not os.path.exists(pdfs_amd.WEIGHTS_PATH):
        pytest.skip("no trained weights")
    init, u64, _ = _f64_solution(ctx, dev, pdfs_amd.GEN_TERRAIN, 9090, 256, 256, 128, 8,
                                 (5.0, 40.0), (10.0, 60.0))
    net = pdfs_amd.Net(ctx, path=pdfs_amd.WEIGHTS_PATH)
    try:
        f = init.clone()
        st = ctx.solve_warm_f32(net, f, init)
    finally:
        net.close()
    emax, _ = _err(f, u64)
    print(f"c3 terrain fixed: fallbacks {st['fallbacks']}, max {emax:.3e} h, "
          f"passes {st['passes']}", flush=True)
    assert st["fallbacks"] == 0 and st["judge_flagged"] == 0 and st["storage"] == FIXED
    assert emax <= F32_VS_F64_TOL


@pytest.mark.parametrize("mode", ["plain", "certified", "warm"])
def test_c4_fixed_batch_equals_single_solves(ctx, dev, mode):
    """c4: the 512 uneven fields in shared persistent launches, fixed-point
    storage. Every field equals its own single solve in the same mode bit for
    bit; the judge flags nothing; a sample equals the oracle's fixed-point
    solve bit for bit."""
    specs = workloads.c4_specs()
    inits = workloads.generate_c4(ctx, specs)
    fields = [t.clone() for t in inits]
    net = None
    if mode == "warm":
        if not os.path.exists(pdfs_amd.WEIGHTS_PATH):
            pytest.skip("no trained weights")
        net = pdfs_amd.Net(ctx, path=pdfs_amd.WEIGHTS_PATH)
    try:
        print(f"c4 fixed {mode}: batch", flush=True)
        if mode == "warm":
            st, passes, fb = ctx.solve_batch_warm_f32(net, fields, inits, graph=1,
                                                      tile=pdfs_amd.TILE_AUTO)
        else:
            st, passes, _ = ctx.solve_batch_f32(fields, inits, graph=1, tile=pdfs_amd.TILE_AUTO,
                                                certify=int(mode == "certified"))
            fb = [0] * len(fields)
        print(f"c4 fixed {mode}: passes max {max(passes)}, fallbacks {sum(fb)}", flush=True)
        assert st["judge_flagged"] == 0 and st["storage"] == FIXED
        for i, (t, f) in enumerate(zip(inits, fields)):
            g = t.clone()
            if mode == "warm":
                s1 = ctx.solve_warm_f32(net, g, t)
                assert s1["fallbacks"] == fb[i], i
            else:
                s1 = ctx.solve_f32(g, t, certify=int(mode == "certified"))
            assert s1["passes"] == passes[i], (i, s1["passes"], passes[i])
            assert torch.equal(g.view(torch.int32), f.view(torch.int32)), (i, specs[i][:2])
    finally:
        if net is not None:
            net.close()
    if mode == "plain":
        small = [i for i, sp in enumerate(specs) if np.prod(sp[1]) <= 1_500_000][:6]
        for i in small:
            kind, (X, Y, Z) = specs[i][:2]
            want, p_or, _, _ = O.solve_fixed(inits[i].cpu().numpy())
            print(f"c4 field {i} {X}x{Y}x{Z} vs oracle", flush=True)
            assert p_or == passes[i], i
            assert np.array_equal(_bits32(fields[i].cpu().numpy()), _bits32(want)), i
    del inits, fields
    torch.cuda.empty_cache()


@pytest.mark.parametrize("offset", [1, 2, 3])
def test_unaligned_field_pointer(ctx, dev, offset):
    """A field that starts 4, 8 or 12 bytes into its allocation (the
    conversions then take their one-value-per-lane path for all of it) gives
    the bits of the same field 16-byte aligned, and of the oracle."""
    X, Y, Z = 37, 29, 23
    start = _start("spheres", X, Y, Z, 90 + offset)
    n = X * Y * Z
    buf = torch.empty(n + 4, dtype=torch.float32, device=dev)
    t = buf[offset:offset + n].view(Z, Y, X)
    t.copy_(torch.from_numpy(start))
    assert t.data_ptr() % 16 != 0
    a = _dev(start, dev)
    sa = ctx.solve_f32(t, None, judge=0)
    sb = ctx.solve_f32(a, None, judge=0)
    assert sa["storage"] == FIXED and sa["passes"] == sb["passes"]
    assert torch.equal(t.view(torch.int32), a.view(torch.int32))
    want, _, _, _ = O.solve_fixed(start)
    assert np.array_equal(t.cpu().numpy().view(np.uint32), want.view(np.uint32))
