"""Train the warm-start network on FSM-solved synthetic fields (GPU).

Data: `--fields` fields of 128^3 (h = 1), alternating the terrain generator
(8 hills, amplitude U(5,40) h, width U(10,60) h) and the sphere generator
(radius U(4,32) h; 1-8 spheres per field, the 512^3 workload's density of
64 spheres being 1 per 128^3). Each field is solved by the batched FP32 FSM until a pass changes
nothing; the network input is the source mask at 1/4 resolution, the target
the solution at the coarse points scaled by 1/32.

Training: Adam, batch 8, cosine learning-rate decay, random axis flips (the
mapping from sources to distances commutes with them).

Calibration: the network resolves distances up to about its receptive
field; beyond a cutoff its prediction levels off below the solution, so
cells whose raw prediction d reaches the cutoff keep their init. On held-out
fields, for each cutoff and each inflation a in a small set, b = max over the
unknown cells inside the cutoff of (u - a d) + margin, so that
w = min(init, a d + b) lies above the solution on every held-out cell; the
(cutoff, a, b) leaving the least work to the sweeps is kept (a cell inside
counts its slack a d + b - u, a cell beyond its distance u).

Calibration fields also include larger grids (256x256x128 terrain, 512^3
spheres by default), whose distances exceed the 128^3 training fields'.

Usage: python -m pdfs_amd.train_warm --out weights.bin [--fields 2000]
       [--steps 3000] [--seed 1]
       python -m pdfs_amd.train_warm --out weights.bin --recalibrate old.bin
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np


def log(*a):
    print(*a, flush=True)


def draw_spec(rng, i, S):
    if i % 2 == 0:
        return (1, 8, (5.0, 40.0), (10.0, 60.0))
    return (2, int(rng.integers(1, 9)), (4.0, 32.0), (0.0, 0.0))


CAL_A = (1.0, 1.1, 1.2, 1.35, 1.5, 1.75, 2.0)
# cutoff candidates in h (0: none): beyond the distances the network resolves
# (its receptive field spans about 150 cells) the warm start keeps the init
CAL_CUTS = (48.0, 64.0, 80.0, 96.0, 112.0, 128.0, 160.0, 0.0)


def make_set(ctx, net, n, seed, S=128, chunk=125, dev="cuda:0"):
    """masks BF16 [n, S/4, S/4, S/4, 16] (int16 storage), targets [n, S/4, S/4, S/4]."""
    import torch

    import pdfs_amd
    D = S // 4
    masks = torch.empty((n, D, D, D, 16), dtype=torch.int16, device=dev)
    targets = torch.empty((n, D, D, D), dtype=torch.float32, device=dev)
    rng = np.random.default_rng(seed)
    passes = []
    for c0 in range(0, n, chunk):
        idx = list(range(c0, min(n, c0 + chunk)))
        inits = []
        for i in idx:
            kind, nobj, p, q = draw_spec(rng, i, S)
            t = torch.empty((S, S, S), dtype=torch.float32, device=dev)
            ctx.generate_ex(kind, seed * 1000003 + i, S, S, S, 1.0, nobj, p, q, dense=t)
            inits.append(t)
        fields = [t.clone() for t in inits]
        _, ps, _ = ctx.solve_batch_f32(fields, None, judge=0, graph=1, tile=pdfs_amd.TILE_AUTO)
        passes += ps
        for j, i in enumerate(idx):
            net.features(inits[j], masks[i:i + 1])
            net.target(fields[j], targets[i:i + 1])
        del inits, fields
    return masks, targets, passes


def raw_prediction(ctx, net, init):
    """Upsampled prediction d (cells) without inflation: the warm start with
    a = 1, b = 0 is min(init, d), which is d on every unknown cell (init
    1e6); the network's input is the mask of this init."""
    import torch
    scale, a, b = net.calibration
    net.set_calibration(scale, 1.0, 0.0)
    d = torch.empty_like(init)
    net.warm_start(init, d)
    net.set_calibration(scale, a, b)
    return d


LARGE_CALIB = (  # (kind, X, Y, Z, nobj, p, q)
    (1, 256, 256, 128, 8, (5.0, 40.0), (10.0, 60.0)),
    (1, 256, 256, 128, 8, (5.0, 40.0), (10.0, 60.0)),
    (2, 512, 512, 512, 64, (4.0, 32.0), (0.0, 0.0)),
    (2, 512, 512, 512, 64, (4.0, 32.0), (0.0, 0.0)),
)


def calibrate(ctx, net, n, seed, S=128, dev="cuda:0", margin=1.0, large=LARGE_CALIB):
    import torch

    import pdfs_amd
    rng = np.random.default_rng(seed)
    stats = []

    def add(init):
        u = init.clone()
        ctx.solve_f32(u, None, judge=0)
        d = raw_prediction(ctx, net, init).clamp(min=0)
        unk = init >= 1.0e6
        # a subsample of the unknown cells keeps the host memory small
        idx = torch.nonzero(unk.flatten()).flatten()
        if idx.numel() > 4_000_000:
            idx = idx[:: idx.numel() // 4_000_000 + 1]
        uu, dd = u.flatten()[idx].double(), d.flatten()[idx].double()
        # the worst cell inside each cutoff over all unknowns is kept exactly
        worst = {}
        for cut in CAL_CUTS:
            m = unk & (d < cut) if cut > 0 else unk
            um, dm = u[m], d[m]
            for a in CAL_A:
                worst[(a, cut)] = float((um - a * dm).max()) if um.numel() else -1e30
        stats.append((uu, dd, worst))

    for i in range(n):
        kind, nobj, p, q = draw_spec(rng, i, S)
        init = torch.empty((S, S, S), dtype=torch.float32, device=dev)
        ctx.generate_ex(kind, seed * 1000003 + i, S, S, S, 1.0, nobj, p, q, dense=init)
        add(init)
    for j, (kind, X, Y, Z, nobj, p, q) in enumerate(large):
        init = torch.empty((Z, Y, X), dtype=torch.float32, device=dev)
        ctx.generate_ex(kind, seed * 7919 + j, X, Y, Z, 1.0, nobj, p, q, dense=init)
        add(init)
        del init
    # per cutoff, the (a, b) that bound every unknown cell inside it with the
    # least slack; across cutoffs, the least work left to the sweeps: a cell
    # inside counts its slack, a cell beyond (kept at its init) its distance
    cands = []
    for cut in CAL_CUTS:
        lim = cut if cut > 0 else float("inf")
        best_a = None
        for a in CAL_A:
            worst = max(w[(a, cut)] for _, _, w in stats)
            b = max(0.0, worst + margin)
            cost = float(np.mean([float(torch.where(d < lim, a * d + b - u, u).mean())
                                  for u, d, _ in stats]))
            beyond = float(np.mean([float((d >= lim).double().mean()) for _, d, _ in stats]))
            c = {"cut": cut, "a": a, "b": b, "cost": cost, "beyond": beyond}
            if best_a is None or cost < best_a["cost"]:
                best_a = c
        cands.append(best_a)
        log(f"[calib] cut={cut:.0f} a={best_a['a']:.2f} b={best_a['b']:.2f} "
            f"cost {best_a['cost']:.2f} h, {best_a['beyond']:.3f} of the cells beyond")
    best = min(cands, key=lambda c: c["cost"])
    return best, cands


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--fields", type=int, default=2000)
    ap.add_argument("--val", type=int, default=96)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--recalibrate", default=None,
                    help="load these weights, only recompute the calibration, save to --out")
    args = ap.parse_args()

    import torch

    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import pdfs_amd

    dev = "cuda:0"
    ctx = pdfs_amd.Context()
    net = pdfs_amd.Net(ctx, seed=args.seed)
    log(f"[train] network: {net.num_params} parameters")
    if args.recalibrate:
        net.load(args.recalibrate)
        scale = net.calibration[0]
        best, cands = calibrate(ctx, net, args.val, args.seed + 104729)
        net.set_calibration(scale, best["a"], best["b"])
        net.set_cut(best["cut"])
        net.save(args.out)
        jp = os.path.splitext(args.recalibrate)[0] + ".json"
        summary = json.load(open(jp)) if os.path.exists(jp) else {}
        summary["calibration"] = {"scale": scale, "a": best["a"], "b": best["b"], "cut": best["cut"],
                                  "candidates": cands,
                                  "large_fields": [list(map(str, c)) for c in LARGE_CALIB]}
        with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
            json.dump(summary, f, indent=1)
        log(json.dumps(summary["calibration"]))
        net.close()
        ctx.close()
        return
    t0 = time.perf_counter()
    masks, targets, passes = make_set(ctx, net, args.fields, args.seed)
    torch.cuda.synchronize()
    tgen = time.perf_counter() - t0
    log(f"[train] {args.fields} fields of 128^3 solved in {tgen:.1f} s "
        f"(passes min {min(passes)} max {max(passes)})")
    g = torch.Generator(device="cpu").manual_seed(args.seed)
    losses = []
    t0 = time.perf_counter()
    for step in range(args.steps):
        idx = torch.randint(0, args.fields, (args.batch,), generator=g)
        m, t = masks[idx.to(dev)], targets[idx.to(dev)]
        flips = [d for d in (1, 2, 3) if torch.rand(1, generator=g).item() < 0.5]
        if flips:
            m, t = m.flip(flips).contiguous(), t.flip(flips).contiguous()
        lr = args.lr * 0.5 * (1.0 + math.cos(math.pi * step / args.steps))
        loss = net.train_step(m, t, lr)
        losses.append(loss)
        if step % 100 == 0 or step == args.steps - 1:
            log(f"[train] step {step} loss {np.mean(losses[-100:]):.5f} lr {lr:.2e} "
                f"{time.perf_counter() - t0:.1f} s")
    ttrain = time.perf_counter() - t0
    del masks, targets
    # held-out error of the prediction at the coarse points, in cells
    vm, vt, _ = make_set(ctx, net, args.val, args.seed + 7919)
    pred = torch.empty(tuple(vm.shape[:4]) + (16,), dtype=torch.float32, device=dev)
    err = []
    scale = net.calibration[0]
    for i in range(0, args.val, 8):
        j = min(args.val, i + 8)
        net.forward(vm[i:j].contiguous(), pred[i:j])
        err.append(((pred[i:j, ..., 0] - vt[i:j]) / scale).abs().flatten())
    err = torch.cat(err)
    log(f"[train] held-out |pred - u| in cells: mean {err.mean().item():.3f} "
        f"p99 {err.quantile(0.99).item():.3f} max {err.max().item():.3f}")
    best, cands = calibrate(ctx, net, args.val, args.seed + 104729)
    net.set_calibration(scale, best["a"], best["b"])
    net.set_cut(best["cut"])
    net.save(args.out)
    summary = {"parameters": net.num_params, "fields": args.fields, "field_size": 128,
               "steps": args.steps, "batch": args.batch, "lr": args.lr, "seed": args.seed,
               "seconds_data": round(tgen, 1), "seconds_train": round(ttrain, 1),
               "loss_first100": float(np.mean(losses[:100])),
               "loss_last100": float(np.mean(losses[-100:])),
               "heldout_abs_err_cells": {"mean": err.mean().item(),
                                         "p99": err.quantile(0.99).item(),
                                         "max": err.max().item()},
               "calibration": {"scale": scale, "a": best["a"], "b": best["b"],
                               "cut": best["cut"], "candidates": cands}}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(summary, f, indent=1)
    log(json.dumps(summary))
    net.close()
    ctx.close()


if __name__ == "__main__":
    main()
