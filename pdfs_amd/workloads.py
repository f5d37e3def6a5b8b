"""Synthetic workloads of the benchmark configurations.

c0  point source 64^3 (FP64 reference arithmetic).
c1  terrain 256x256x128: 8 hills, amplitude U(5,40) h, width U(10,60) h.
c2  multi-sphere 512^3: 64 spheres, radius U(4,32) h.
c4  batch of 512 fields, each axis uniform over {32, 48, ..., 192}, an equal
    mix of terrain, spheres and point sources. Object parameters scale with
    the field: terrain amplitude x Z/128 and width x min(X,Y)/256 (the c1
    proportions); spheres: 8 per field, radius U(max(1.5, 4s), max(3, 32s)) h
    with s = (XYZ / 512^3)^(1/3) (the c2 proportions by volume).
"""
import numpy as np

from . import GEN_POINT, GEN_SPHERES, GEN_TERRAIN

C4_SIZES = list(range(32, 193, 16))


def c4_specs(seed=4, nfields=512):
    """[(kind, (X, Y, Z), field_seed, nobj, p_range, q_range)] for the batch."""
    rng = np.random.default_rng(seed)
    out = []
    kinds = (GEN_TERRAIN, GEN_SPHERES, GEN_POINT)
    for i in range(nfields):
        X, Y, Z = (int(v) for v in rng.choice(C4_SIZES, 3))
        kind = kinds[i % 3]
        fseed = int(seed) * 1000003 + i
        if kind == GEN_TERRAIN:
            a, w = Z / 128.0, min(X, Y) / 256.0
            out.append((kind, (X, Y, Z), fseed, 8, (5.0 * a, 40.0 * a), (10.0 * w, 60.0 * w)))
        elif kind == GEN_SPHERES:
            s = (X * Y * Z / 512.0 ** 3) ** (1.0 / 3.0)
            out.append((kind, (X, Y, Z), fseed, 8, (max(1.5, 4 * s), max(3.0, 32 * s)),
                        (0.0, 0.0)))
        else:
            out.append((kind, (X, Y, Z), fseed, 0, (0.0, 0.0), (0.0, 0.0)))
    return out


def generate_c4(ctx, specs, device="cuda:0"):
    """Dense FP32 device fields for c4 specs."""
    import torch
    fields = []
    for kind, (X, Y, Z), fseed, nobj, p, q in specs:
        t = torch.empty((Z, Y, X), dtype=torch.float32, device=device)
        ctx.generate_ex(kind, fseed, X, Y, Z, 1.0, nobj, p, q, dense=t)
        fields.append(t)
    return fields
