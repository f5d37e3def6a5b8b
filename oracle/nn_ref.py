"""CPU restatement (numpy) of the warm-start network, as test infrastructure:
the checker of the HIP kernels' forward pass and gradients. It has no
counterpart in the reference (the learned warm start is new).

Layout is that of the library: channels-last activations [nb, z, y, x, C];
weights [27, Cout, Cin] with tap = (dz*3 + dy)*3 + dx (input offset
dz-1, dy-1, dx-1), parameters packed as all layers' weights, then all
layers' biases, in layer order.

`forward(..., bf16=True)` rounds activations to BF16 where the GPU stores
them (every ReLU conv output and every pooling output); the arithmetic is
float64. `backward` is exact float64 backpropagation through those stored
activations.
"""
import numpy as np

CIN = [16, 16, 16, 32, 32, 64, 64, 64, 128, 96, 48, 16]
COUT = [16, 16, 32, 32, 64, 64, 64, 64, 64, 32, 16, 16]
NAMES = ["e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "d2", "d1", "d0", "head"]


def offsets():
    woff, boff, off = [], [], 0
    for ci, co in zip(CIN, COUT):
        woff.append(off)
        off += 27 * ci * co
    for co in COUT:
        boff.append(off)
        off += co
    return woff, boff, off


def unpack(p):
    woff, boff, _ = offsets()
    W = [p[woff[i]:woff[i] + 27 * CIN[i] * COUT[i]].reshape(27, COUT[i], CIN[i]).astype(np.float64)
         for i in range(len(CIN))]
    B = [p[boff[i]:boff[i] + COUT[i]].astype(np.float64) for i in range(len(CIN))]
    return W, B


def pack(gW, gB):
    woff, boff, n = offsets()
    out = np.zeros(n, np.float64)
    for i in range(len(CIN)):
        out[woff[i]:woff[i] + gW[i].size] = gW[i].ravel()
        out[boff[i]:boff[i] + gB[i].size] = gB[i]
    return out


def bf16_round(x):
    """Round to nearest even BF16 (finite values), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits_to_f64(h):
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _shifted(x, dz, dy, dx):
    """x[:, z+dz, y+dy, x+dx, :] with zeros outside."""
    nb, Z, Y, X, C = x.shape
    out = np.zeros_like(x)
    zs, ys, xs = slice(max(0, -dz), Z - max(0, dz)), slice(max(0, -dy), Y - max(0, dy)), \
        slice(max(0, -dx), X - max(0, dx))
    zd, yd, xd = slice(max(0, dz), Z - max(0, -dz)), slice(max(0, dy), Y - max(0, -dy)), \
        slice(max(0, dx), X - max(0, -dx))
    out[:, zs, ys, xs] = x[:, zd, yd, xd]
    return out


def conv3(x, W, b=None):
    """out[p, co] = sum_tap sum_ci x[p + shift(tap), ci] W[tap, co, ci] (+ b)."""
    out = 0.0
    for tap in range(27):
        dz, dy, dx = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
        out = out + _shifted(x, dz, dy, dx) @ W[tap].T
    if b is not None:
        out = out + b
    return out


def conv3_wgrad(x, g):
    """dW[tap, co, ci] = sum_p g[p, co] x[p + shift(tap), ci]."""
    C_in, C_out = x.shape[-1], g.shape[-1]
    dW = np.zeros((27, C_out, C_in))
    g2 = g.reshape(-1, C_out)
    for tap in range(27):
        dz, dy, dx = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
        dW[tap] = g2.T @ _shifted(x, dz, dy, dx).reshape(-1, C_in)
    return dW


def conv3_dgrad(g, W):
    """dx[q, ci] = sum_tap sum_co g[q - shift(tap), co] W[tap, co, ci]."""
    out = 0.0
    for tap in range(27):
        dz, dy, dx = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
        out = out + _shifted(g, -dz, -dy, -dx) @ W[tap]
    return out


def pool2(x):
    nb, Z, Y, X, C = x.shape
    return x.reshape(nb, Z // 2, 2, Y // 2, 2, X // 2, 2, C).mean(axis=(2, 4, 6))


def pool2_bwd(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 1), 2, 2), 2, 3) / 8.0


def up2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up2_bwd(g):
    nb, Z, Y, X, C = g.shape
    return g.reshape(nb, Z // 2, 2, Y // 2, 2, X // 2, 2, C).sum(axis=(2, 4, 6))


def forward(p, x, bf16=True):
    """x: [nb, Dz, Dy, Dx, 16] (float). Returns (out [.., 16] float64, cache)."""
    W, B = unpack(p)
    if bf16:
        W = [bf16_round(w) for w in W]
    r = bf16_round if bf16 else (lambda v: v)
    relu = lambda v: np.maximum(v, 0.0)
    a = {}
    a["x"] = x.astype(np.float64)
    a["a0"] = r(relu(conv3(a["x"], W[0], B[0])))
    a["s0"] = r(relu(conv3(a["a0"], W[1], B[1])))
    a["p0"] = r(pool2(a["s0"]))
    a["a1"] = r(relu(conv3(a["p0"], W[2], B[2])))
    a["s1"] = r(relu(conv3(a["a1"], W[3], B[3])))
    a["p1"] = r(pool2(a["s1"]))
    a["a2"] = r(relu(conv3(a["p1"], W[4], B[4])))
    a["s2"] = r(relu(conv3(a["a2"], W[5], B[5])))
    a["p2"] = r(pool2(a["s2"]))
    a["a3"] = r(relu(conv3(a["p2"], W[6], B[6])))
    a["b3"] = r(relu(conv3(a["a3"], W[7], B[7])))
    a["c2"] = np.concatenate([up2(a["b3"]), a["s2"]], axis=-1)
    a["d2"] = r(relu(conv3(a["c2"], W[8], B[8])))
    a["c1"] = np.concatenate([up2(a["d2"]), a["s1"]], axis=-1)
    a["d1"] = r(relu(conv3(a["c1"], W[9], B[9])))
    a["c0"] = np.concatenate([up2(a["d1"]), a["s0"]], axis=-1)
    a["d0"] = r(relu(conv3(a["c0"], W[10], B[10])))
    out = conv3(a["d0"], W[11], B[11])
    return out, (W, a)


def mse(out, t):
    d = out[..., 0] - t
    return float((d * d).mean()), 2.0 * d / d.size


def backward(cache, g0):
    """g0: dL/d out[..., 0] ([nb, Dz, Dy, Dx]). Returns the packed gradient."""
    W, a = cache
    gW, gB = [None] * 12, [None] * 12
    mask = lambda g, y: g * (y > 0)

    def lay(i, xin, gpre, need_dx=True):
        gW[i] = conv3_wgrad(xin, gpre)
        gB[i] = gpre.reshape(-1, gpre.shape[-1]).sum(0)
        return conv3_dgrad(gpre, W[i]) if need_dx else None

    gout = np.zeros(g0.shape + (16,))
    gout[..., 0] = g0
    gd0 = lay(11, a["d0"], gout)
    gc0 = lay(10, a["c0"], mask(gd0, a["d0"]))
    gd1, gs0 = up2_bwd(gc0[..., :32]), gc0[..., 32:]
    gc1 = lay(9, a["c1"], mask(gd1, a["d1"]))
    gd2, gs1 = up2_bwd(gc1[..., :64]), gc1[..., 64:]
    gc2 = lay(8, a["c2"], mask(gd2, a["d2"]))
    gb3, gs2 = up2_bwd(gc2[..., :64]), gc2[..., 64:]
    ga3 = lay(7, a["a3"], mask(gb3, a["b3"]))
    gp2 = lay(6, a["p2"], mask(ga3, a["a3"]))
    gs2 = gs2 + pool2_bwd(gp2)
    ga2 = lay(5, a["a2"], mask(gs2, a["s2"]))
    gp1 = lay(4, a["p1"], mask(ga2, a["a2"]))
    gs1 = gs1 + pool2_bwd(gp1)
    ga1 = lay(3, a["a1"], mask(gs1, a["s1"]))
    gp0 = lay(2, a["p0"], mask(ga1, a["a1"]))
    gs0 = gs0 + pool2_bwd(gp0)
    ga0 = lay(1, a["a0"], mask(gs0, a["s0"]))
    lay(0, a["x"], mask(ga0, a["a0"]), need_dx=False)
    return pack(gW, gB)


def mask_down4(init, Dz, Dy, Dx, far=1e6):
    """Fraction of known cells (init < far) per 4x4x4 block, [Dz, Dy, Dx]."""
    Z, Y, X = init.shape
    pad = np.zeros((4 * Dz, 4 * Dy, 4 * Dx), np.float64)
    pad[:Z, :Y, :X] = init < far
    return pad.reshape(Dz, 4, Dy, 4, Dx, 4).sum(axis=(1, 3, 5)) / 64.0


def warm_up4(pred0, scale_inv, a, b, init):
    """Trilinear upsampling (fine point i at coarse coordinate i/4, clamped)
    of pred0 [Dz, Dy, Dx], w = min(init, a * max(d, 0) + b), float32 math."""
    Z, Y, X = init.shape
    Dz, Dy, Dx = pred0.shape
    f32 = np.float32
    P = pred0.astype(f32)

    def axis(n, D):
        f = (np.arange(n, dtype=f32) * f32(0.25)).astype(f32)
        i0 = np.minimum(f.astype(np.int64), D - 1)
        i1 = np.minimum(i0 + 1, D - 1)
        return i0, i1, (f - i0.astype(f32)).astype(f32)

    z0, z1, tz = axis(Z, Dz)
    y0, y1, ty = axis(Y, Dy)
    x0, x1, tx = axis(X, Dx)
    one = f32(1)
    tx_, ty_, tz_ = tx[None, None, :], ty[None, :, None], tz[:, None, None]

    def g(zi, yi, xi):
        return P[zi[:, None, None], yi[None, :, None], xi[None, None, :]]

    c00 = g(z0, y0, x0) * (one - tx_) + g(z0, y0, x1) * tx_
    c01 = g(z0, y1, x0) * (one - tx_) + g(z0, y1, x1) * tx_
    c10 = g(z1, y0, x0) * (one - tx_) + g(z1, y0, x1) * tx_
    c11 = g(z1, y1, x0) * (one - tx_) + g(z1, y1, x1) * tx_
    c0 = c00 * (one - ty_) + c01 * ty_
    c1 = c10 * (one - ty_) + c11 * ty_
    d = (c0 * (one - tz_) + c1 * tz_) * f32(scale_inv)
    up = f32(a) * np.maximum(d, f32(0)) + f32(b)
    return np.minimum(init.astype(f32), up.astype(f32))
