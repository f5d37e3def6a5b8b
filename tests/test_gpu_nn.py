"""GPU checks of the warm-start network (BF16 MFMA kernels) against its numpy
restatement, and of the warm-started solve against the FSM oracle.

Tolerances: activations are stored in BF16 (8 significant bits) after every
layer on both sides; the GPU accumulates in FP32 in another order, so a
stored value can round to the neighbouring BF16 number and the difference
propagates. Forward: max |gpu - ref| <= 3e-2 * max |ref|. Gradients (the
GPU also rounds the ReLU-masked output gradients to BF16): per layer,
relative L2 error <= 6e-2 and cosine >= 0.995.
"""
import os
import sys

import numpy as np
import pytest
import torch

import oracle_ctypes as O
import pdfs_amd

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import nn_ref as R  # noqa: E402

# these tests check the FP32-storage path (and its polish) against the FP32 oracle
S32 = pdfs_amd.PDFS_STORE_F32

pytestmark = pytest.mark.gpu

FWD_TOL = 3e-2
GRAD_REL = 6e-2
GRAD_COS = 0.995


def _bf16_bits(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _mask_tensor(m, dev):
    """[nb, Dz, Dy, Dx] float mask -> BF16 bits [nb, Dz, Dy, Dx, 16] (channel 0)."""
    x = np.zeros(m.shape + (16,), np.float32)
    x[..., 0] = m
    bits = _bf16_bits(x)
    return torch.from_numpy(bits.view(np.int16)).to(dev), R.bf16_bits_to_f64(bits)


@pytest.fixture(scope="module")
def net(ctx):
    n = pdfs_amd.Net(ctx, seed=11)
    yield n
    n.close()


def _random_params(seed):
    rng = np.random.default_rng(seed)
    woff, boff, n = R.offsets()
    p = np.zeros(n, np.float32)
    for i, (ci, co) in enumerate(zip(R.CIN, R.COUT)):
        p[woff[i]:woff[i] + 27 * ci * co] = rng.normal(0, np.sqrt(2.0 / (27 * ci)), 27 * ci * co)
        p[boff[i]:boff[i] + co] = rng.normal(0, 0.1, co)
    return p


@pytest.mark.parametrize("shape", [(1, 8, 8, 8), (2, 16, 8, 24)])
def test_forward_matches_numpy(net, dev, shape):
    p = _random_params(3 + shape[0])
    net.set_params(p)
    rng = np.random.default_rng(7)
    m = (rng.random(shape) < 0.3) * rng.random(shape)
    xt, xr = _mask_tensor(m, dev)
    pred = torch.empty(shape + (16,), dtype=torch.float32, device=dev)
    net.forward(xt, pred)
    got = pred.cpu().numpy().astype(np.float64)
    want, _ = R.forward(p, xr, bf16=True)
    err = np.abs(got - want).max()
    scale = np.abs(want).max()
    print(f"forward {shape}: max err {err:.3e} of max {scale:.3e}")
    assert err <= FWD_TOL * scale
    assert np.corrcoef(got[..., 0].ravel(), want[..., 0].ravel())[0, 1] > 0.999


def test_gradients_match_numpy(net, dev):
    shape = (2, 8, 8, 16)
    p = _random_params(21)
    net.set_params(p)
    rng = np.random.default_rng(8)
    m = (rng.random(shape) < 0.3) * rng.random(shape)
    xt, xr = _mask_tensor(m, dev)
    t = rng.normal(size=shape).astype(np.float32)
    loss = net.train_step(xt, torch.from_numpy(t).to(dev), lr=0.0)
    g = net.params(which=1).astype(np.float64)
    out, cache = R.forward(p, xr, bf16=True)
    L, g0 = R.mse(out, t.astype(np.float64))
    gw = R.backward(cache, g0)
    assert abs(loss - L) <= 1e-2 * abs(L)
    woff, boff, _ = R.offsets()
    for i, name in enumerate(R.NAMES):
        for lo, n in ((woff[i], 27 * R.CIN[i] * R.COUT[i]), (boff[i], R.COUT[i])):
            a, b = g[lo:lo + n], gw[lo:lo + n]
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            cos = a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
            print(f"{name}[{n}]: rel {rel:.3e} cos {cos:.6f}")
            assert rel <= GRAD_REL and cos >= GRAD_COS, (name, n, rel, cos)


def test_adam_reduces_loss(ctx, dev):
    n = pdfs_amd.Net(ctx, seed=5)
    shape = (2, 8, 8, 8)
    rng = np.random.default_rng(9)
    m = (rng.random(shape) < 0.2).astype(np.float64)
    xt, _ = _mask_tensor(m, dev)
    t = torch.from_numpy(rng.random(shape).astype(np.float32)).to(dev)
    l0 = n.train_step(xt, t, lr=0.0)
    for _ in range(60):
        n.train_step(xt, t, lr=2e-3)
    l1 = n.train_step(xt, t, lr=0.0)
    print(f"loss {l0:.4f} -> {l1:.4f}")
    assert l1 < 0.5 * l0
    n.close()


def test_features_and_warm_start_match_numpy(net, dev):
    X, Y, Z = 45, 38, 29
    a = O.gen_spheres(X, Y, Z, seed=77, n=4, r=(2.0, 5.0))
    init = np.ascontiguousarray(O.interior(a)).astype(np.float32)
    ti = torch.from_numpy(init).to(dev)
    Dz, Dy, Dx = pdfs_amd.net_grid(X, Y, Z)
    assert (Dz, Dy, Dx) == (8, 16, 16)
    mask = torch.empty((1, Dz, Dy, Dx, 16), dtype=torch.int16, device=dev)
    net.features(ti, mask)
    mb = mask.cpu().numpy().view(np.uint16)
    want = R.mask_down4(init, Dz, Dy, Dx)
    assert np.array_equal(R.bf16_bits_to_f64(mb[0, ..., 0]), R.bf16_round(want))
    assert not mb[..., 1:].any()
    net.set_params(_random_params(31))
    net.set_calibration(1.0 / 32, 1.25, 2.0)
    pred = torch.empty((1, Dz, Dy, Dx, 16), dtype=torch.float32, device=dev)
    net.forward(mask, pred)
    w = torch.empty_like(ti)
    net.warm_start(ti, w, hmax=1.5)
    want_w = R.warm_up4(pred.cpu().numpy()[0, ..., 0], 1.5 * 32, 1.25, 2.0 * 1.5, init)
    assert np.array_equal(w.cpu().numpy().view(np.uint32), want_w.view(np.uint32))


def test_warm_start_batch_matches_per_field(net, dev):
    """The uneven-batch warm start (all fields in one launch per layer) gives
    every field exactly the per-field warm start: fields of different coarse
    grids, including a single point and fields thinner than one 4-cell block."""
    net.set_params(_random_params(53))
    net.set_calibration(1.0 / 32, 1.25, 2.0)
    shapes = [(45, 38, 29), (1, 1, 1), (70, 17, 33), (64, 8, 9), (3, 100, 5), (96, 96, 40)]
    inits, ws_one = [], []
    for i, (X, Y, Z) in enumerate(shapes):
        a = O.gen_spheres(X, Y, Z, seed=300 + i, n=3, r=(1.0, 4.0))
        init = torch.from_numpy(np.ascontiguousarray(O.interior(a)).astype(np.float32)).to(dev)
        w = torch.empty_like(init)
        net.warm_start(init, w, hmax=1.25)
        inits.append(init)
        ws_one.append(w)
    ws_batch = [torch.full_like(t, -1.0) for t in inits]
    net.warm_start_batch(inits, ws_batch, hmax=1.25)
    for i, (a, b) in enumerate(zip(ws_one, ws_batch)):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32)), f"field {i} {shapes[i]}"


def _spheres(seed, X=64, Y=56, Z=48):
    return O.gen_spheres(X, Y, Z, seed=seed, n=5, r=(3.0, 8.0))


def test_warm_solve_good_bound_certifies_without_fallback(ctx, net, dev):
    """A loose upper bound (large inflation) always certifies without a
    fallback; the result is within the FP32 tolerance of the FP64 oracle."""
    a = _spheres(90)
    ref64 = a.copy()
    O.solve_d64(ref64, tol=0.0)
    init = torch.from_numpy(np.ascontiguousarray(O.interior(a)).astype(np.float32)).to(dev)
    net.set_params(_random_params(41))
    net.set_calibration(1.0 / 32, 0.0, 1000.0)   # w = min(init, 1000 h): above everything
    f = init.clone()
    st = ctx.solve_warm_f32(net, f, init, storage=S32)
    err = np.abs(f.cpu().numpy().astype(np.float64) - O.interior(ref64)).max()
    assert st["fallbacks"] == 0 and st["judge_flagged"] == 0
    assert err <= 1e-4


def test_warm_solve_bad_bound_falls_back(ctx, net, dev):
    """An underestimating warm start (everything 0.5 h) leaves a wrong field;
    the judge flags it and the solver falls back to the plain solve, whose
    result equals the oracle's FP32 solve + FP64 polish bit for bit."""
    a = _spheres(91)
    ref32, _, _ = O.solve_f32_polished(a)
    init = torch.from_numpy(np.ascontiguousarray(O.interior(a)).astype(np.float32)).to(dev)
    net.set_calibration(1.0 / 32, 0.0, 0.5)
    f = init.clone()
    st = ctx.solve_warm_f32(net, f, init, storage=S32)
    assert st["fallbacks"] == 1 and st["judge_flagged"] == 0
    assert np.array_equal(f.cpu().numpy().view(np.uint32), O.interior(ref32).view(np.uint32))


@pytest.mark.skipif(not os.path.exists(pdfs_amd.WEIGHTS_PATH), reason="no trained weights")
def test_trained_weights_warm_start_is_an_upper_bound(ctx, dev):
    """The committed network with its calibration, on fresh fields of each
    generator kind (the point source is outside its training set): the
    warm-started solve ends certified and within 1e-4 h of the plain solve,
    by itself or through the fallback; the points where the warm start lies
    below the solution are reported."""
    n = pdfs_amd.Net(ctx, path=pdfs_amd.WEIGHTS_PATH)
    X = Y = Z = 96
    for kind in (pdfs_amd.GEN_POINT, pdfs_amd.GEN_TERRAIN, pdfs_amd.GEN_SPHERES):
        init = torch.empty((Z, Y, X), dtype=torch.float32, device=dev)
        ctx.generate(kind, 7000 + kind, X, Y, Z, 1.0, dense=init)
        plain = init.clone()
        sp = ctx.solve_f32(plain, init, storage=S32)
        f = init.clone()
        st = ctx.solve_warm_f32(n, f, init, storage=S32)
        w = torch.empty_like(init)
        n.warm_start(init, w)
        below = int((w < plain).sum())
        print(f"kind {kind}: plain {sp['passes']} passes, warm {st['passes']} "
              f"(fallbacks {st['fallbacks']}), points below the solution {below}")
        assert st["judge_flagged"] == 0
        assert (f - plain).abs().max().item() <= 1e-4
    n.close()


def test_warm_start_cut_keeps_init_beyond(ctx, net, dev, tmp_path):
    """Cells whose raw prediction reaches the cutoff keep their init, on the
    dense and the batched warm start; elsewhere the warm start is unchanged.
    The cutoff travels with the calibration in the weights file."""
    p = _random_params(43)
    p[R.offsets()[1][-1]] += 2.0   # head bias of channel 0: predictions around 64 h
    net.set_params(p)
    net.set_calibration(1.0 / 32, 1.0, 0.0)
    net.set_cut(0.0)
    shapes = [(40, 48, 56), (33, 17, 25)]
    inits = []
    for i, (Z, Y, X) in enumerate(shapes):
        t = torch.empty((Z, Y, X), dtype=torch.float32, device=dev)
        ctx.generate_ex(pdfs_amd.GEN_SPHERES, 700 + i, X, Y, Z, 1.0, 3, (2.0, 5.0), (0.0, 0.0),
                        dense=t)
        inits.append(t)
    raw = [torch.empty_like(t) for t in inits]
    for t, r in zip(inits, raw):
        net.warm_start(t, r)
    unk = inits[0] >= 1e6
    cut = float(raw[0][unk].median())
    assert 0.0 < cut < 1e6
    net.set_cut(cut)
    assert net.cut == pytest.approx(cut)
    want = [torch.where(r >= cut, t, r) for t, r in zip(inits, raw)]
    assert bool((want[0] == inits[0])[unk].any()) and bool((want[0] != inits[0])[unk].any())
    got = torch.empty_like(inits[0])
    net.warm_start(inits[0], got)
    assert torch.equal(got, want[0])
    outs = [torch.empty_like(t) for t in inits]
    net.warm_start_batch(inits, outs)
    for o, w in zip(outs, want):
        assert torch.equal(o, w)
    p = str(tmp_path / "w.bin")
    net.save(p)
    other = pdfs_amd.Net(ctx, seed=5, path=p)
    try:
        assert other.cut == pytest.approx(cut)
        assert other.calibration == pytest.approx(net.calibration)
    finally:
        other.close()
    net.set_cut(0.0)


@pytest.mark.skipif(not os.path.exists(pdfs_amd.WEIGHTS_PATH), reason="no trained weights")
def test_trained_warm_start_bounds_far_corners_at_512(ctx, dev):
    """512^3 sphere fields whose far corners lie beyond the distances the
    network resolves (without the cutoff, its prediction levelled off up to
    31 h below the solution there, and the judge sent the field back to the
    plain solve): with the committed calibration and cutoff the warm start
    lies above the solution everywhere, and the warm-started solve certifies
    without a fallback."""
    n = pdfs_amd.Net(ctx, path=pdfs_amd.WEIGHTS_PATH)
    S = 512
    try:
        assert n.cut > 0
        for seed in (2026, 2032):
            init = torch.empty((S, S, S), dtype=torch.float32, device=dev)
            ctx.generate(pdfs_amd.GEN_SPHERES, seed, S, S, S, 1.0, dense=init)
            u = init.clone()
            ctx.solve_f32(u, None, judge=0, storage=S32)
            w = torch.empty_like(init)
            n.warm_start(init, w)
            assert int((w < u).sum()) == 0
            f = init.clone()
            st = ctx.solve_warm_f32(n, f, init, storage=S32)
            assert st["fallbacks"] == 0 and st["judge_flagged"] == 0
            del init, u, w, f
    finally:
        n.close()
    torch.cuda.empty_cache()


@pytest.mark.skipif(not os.path.exists(pdfs_amd.WEIGHTS_PATH), reason="no trained weights")
def test_trained_network_gain_over_untrained_and_plain(ctx, dev):
    """What the trained network gains over the method without it, measured on
    fresh 128^3 fields of the training generators (terrain, spheres):

    * its raw prediction (before the upper-bound inflation) is far closer to
      the converged field than an untrained network's, and its calibrated
      warm start lies above the solution everywhere (no fallback);
    * the calibrated warm start is a far tighter upper bound than the plain
      init (1e6 on unknown cells) on the cells within its cutoff, which are
      nearly all of them;
    * the warm-started certified solve needs no more passes than the plain
      solve. (On these distance fields a warm start cannot cut the passes of
      the certified solve below the plain certified count: even u* + 0.5 h
      certifies in the same number of passes; only the exact field
      certifies after one.)"""
    trained = pdfs_amd.Net(ctx, path=pdfs_amd.WEIGHTS_PATH)
    untrained = pdfs_amd.Net(ctx, seed=123)
    scale, a, b = trained.calibration
    cut = trained.cut
    S = 128
    try:
        for kind, nobj, p, q, seed in ((pdfs_amd.GEN_TERRAIN, 8, (5.0, 40.0), (10.0, 60.0), 31337),
                                       (pdfs_amd.GEN_SPHERES, 4, (4.0, 32.0), (0.0, 0.0), 31338)):
            init = torch.empty((S, S, S), dtype=torch.float32, device=dev)
            ctx.generate_ex(kind, seed, S, S, S, 1.0, nobj, p, q, dense=init)
            u = init.clone()
            sp = ctx.solve_f32(u, init, storage=S32)
            unk = init >= 1e6
            err = {}
            for name, net in (("trained", trained), ("untrained", untrained)):
                # a = 1, b = 0, no cutoff: the warm start is min(init, d) = d on
                # unknown cells
                net.set_calibration(scale, 1.0, 0.0)
                net.set_cut(0.0)
                d = torch.empty_like(init)
                net.warm_start(init, d)
                err[name] = float((d - u)[unk].abs().mean())
            trained.set_calibration(scale, a, b)
            trained.set_cut(cut)
            w = torch.empty_like(init)
            trained.warm_start(init, w)
            slack = (w - u)[unk]
            # unknown cells the network covers (the rest, beyond the cutoff,
            # keep their init)
            cov = unk & (w < init)
            beyond = 1.0 - float(cov.sum()) / float(unk.sum())
            f = init.clone()
            sw = ctx.solve_warm_f32(trained, f, init, storage=S32)
            print(f"kind {kind}: mean |d - u| trained {err['trained']:.2f} h, untrained "
                  f"{err['untrained']:.2f} h; warm start slack mean {float((w - u)[cov].mean()):.1f} h "
                  f"on the covered cells, {beyond:.4f} beyond the cutoff "
                  f"(below: {int((slack < 0).sum())}), init slack "
                  f"{float((init - u)[unk].mean()):.3g} h; passes plain {sp['passes']} warm "
                  f"{sw['passes']} fallbacks {sw['fallbacks']}")
            assert err["trained"] < 0.25 * err["untrained"]
            assert int((slack < 0).sum()) == 0 and sw["fallbacks"] == 0
            assert float((w - u)[cov].mean()) < 1e-4 * float((init - u)[unk].mean())
            assert beyond < 0.05
            assert sw["passes"] <= sp["passes"] and sw["judge_flagged"] == 0
            assert (f - u).abs().max().item() <= 1e-4
    finally:
        trained.close()
        untrained.close()
