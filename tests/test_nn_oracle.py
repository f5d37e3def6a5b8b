"""CPU pinning of the warm-start network's numpy restatement (oracle/nn_ref.py),
the checker of the GPU kernels: convolution against a direct loop, every
gradient against central finite differences of the float64 network, pooling
and upsampling against their adjoints, the features and the upsampling
against brute force and closed forms."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import nn_ref as R  # noqa: E402


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    woff, boff, n = R.offsets()
    p = np.zeros(n)
    for i, (ci, co) in enumerate(zip(R.CIN, R.COUT)):
        p[woff[i]:woff[i] + 27 * ci * co] = rng.normal(0, scale * np.sqrt(2.0 / (27 * ci)), 27 * ci * co)
        p[boff[i]:boff[i] + co] = rng.normal(0, 0.05, co)
    return p


def test_conv3_matches_direct_loop():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 3, 4, 5, 3))
    W = rng.normal(size=(27, 2, 3))
    b = rng.normal(size=2)
    got = R.conv3(x, W, b)
    want = np.zeros((2, 3, 4, 5, 2))
    for n in range(2):
        for z in range(3):
            for y in range(4):
                for xx in range(5):
                    acc = b.copy()
                    for dz in range(3):
                        for dy in range(3):
                            for dx in range(3):
                                zz, yy, xq = z + dz - 1, y + dy - 1, xx + dx - 1
                                if 0 <= zz < 3 and 0 <= yy < 4 and 0 <= xq < 5:
                                    acc += W[(dz * 3 + dy) * 3 + dx] @ x[n, zz, yy, xq]
                    want[n, z, y, xx] = acc
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_conv3_adjoints():
    """<conv(x), g> = <x, dgrad(g)> = <W, wgrad(x, g)>"""
    rng = np.random.default_rng(2)
    x = rng.normal(size=(1, 4, 3, 5, 4))
    g = rng.normal(size=(1, 4, 3, 5, 6))
    W = rng.normal(size=(27, 6, 4))
    lhs = np.sum(R.conv3(x, W) * g)
    assert np.isclose(lhs, np.sum(x * R.conv3_dgrad(g, W)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(W * R.conv3_wgrad(x, g)), rtol=1e-12)


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2, 4, 6, 8, 3))
    g = rng.normal(size=(2, 2, 3, 4, 3))
    assert np.isclose(np.sum(R.pool2(x) * g), np.sum(x * R.pool2_bwd(g)), rtol=1e-12)
    assert np.isclose(np.sum(R.up2(g) * x), np.sum(g * R.up2_bwd(x)), rtol=1e-12)
    # pool of a constant is the constant; up then pool is the identity
    assert np.allclose(R.pool2(np.full_like(x, 3.5)), 3.5)
    assert np.allclose(R.pool2(R.up2(g)), g)


def test_network_gradient_matches_finite_differences():
    """Full float64 network: the packed gradient's directional derivative and
    single coordinates of every layer against central differences."""
    rng = np.random.default_rng(4)
    p = _params(5)
    x = np.zeros((1, 8, 8, 8, 16))
    x[..., 0] = rng.random((1, 8, 8, 8))
    t = rng.normal(size=(1, 8, 8, 8))

    def loss(q):
        out, _ = R.forward(q, x, bf16=False)
        return R.mse(out, t)[0]

    out, cache = R.forward(p, x, bf16=False)
    L, g0 = R.mse(out, t)
    g = R.backward(cache, g0)
    eps = 1e-6
    d = rng.normal(size=p.size)
    fd = (loss(p + eps * d) - loss(p - eps * d)) / (2 * eps)
    assert np.isclose(fd, g @ d, rtol=1e-5, atol=1e-9), (fd, g @ d)
    woff, boff, _ = R.offsets()
    for i in range(len(R.CIN)):
        for k in (woff[i] + 7 * i % (27 * R.CIN[i] * R.COUT[i]), boff[i]):
            e = np.zeros_like(p)
            e[k] = 1.0
            fdk = (loss(p + 1e-5 * e) - loss(p - 1e-5 * e)) / 2e-5
            assert np.isclose(fdk, g[k], rtol=1e-4, atol=1e-8), (R.NAMES[i], k, fdk, g[k])


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.0078125, 1.00390625, 1.01171875, -2.5, 3.0e-3], np.float32)
    got = R.bf16_round(vals)
    # 7 stored mantissa bits: 1 + 2^-7 is exact; 1 + 2^-8 is a tie -> even
    # (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6)
    assert got[0] == 1.0 and got[1] == 1.0078125 and got[2] == 1.0 and got[3] == 1.015625
    assert got[4] == -2.5
    assert abs(got[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_mask_down4_brute_force():
    rng = np.random.default_rng(6)
    init = np.where(rng.random((9, 7, 13)) < 0.2, 0.0, 1e6)
    m = R.mask_down4(init, 8, 8, 8)
    for z in range(8):
        for y in range(8):
            for x in range(8):
                blk = init[4 * z:4 * z + 4, 4 * y:4 * y + 4, 4 * x:4 * x + 4]
                assert m[z, y, x] == np.sum(blk < 1e6) / 64.0


def test_warm_up4_reproduces_linear_fields_and_clamps():
    """Trilinear interpolation is exact for affine functions of the coarse
    coordinates (away from the clamped far edge), matches the samples at
    coarse points, and the init caps the result."""
    Dz, Dy, Dx = 8, 8, 8
    zc, yc, xc = np.meshgrid(np.arange(Dz), np.arange(Dy), np.arange(Dx), indexing="ij")
    pred = (1.0 + 0.5 * zc + 0.25 * yc + 2.0 * xc).astype(np.float32)
    Z, Y, X = 29, 29, 29
    init = np.full((Z, Y, X), 1e6, np.float32)
    w = R.warm_up4(pred, 1.0, 1.0, 0.0, init)
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    want = 1.0 + 0.5 * z / 4 + 0.25 * y / 4 + 2.0 * x / 4
    assert np.allclose(w, want, rtol=0, atol=1e-4)
    assert np.array_equal(w[::4, ::4, ::4], pred[: (Z + 3) // 4, : (Y + 3) // 4, : (X + 3) // 4])
    init[3, 4, 5] = 0.0
    w2 = R.warm_up4(pred, 2.0, 1.1, 3.0, init)
    assert w2[3, 4, 5] == 0.0
    assert np.isclose(w2[0, 0, 0], 1.1 * 2.0 * 1.0 + 3.0)
