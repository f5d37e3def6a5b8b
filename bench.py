"""Benchmark: converged fields/s of the FP32 FSM on one MI355X, plus the
sweep kernel's achieved HBM bandwidth against the 8 TB/s peak.

Workload (the `value`): multi-sphere distance field, 512^3, h = 1, 64 seeded
spheres of radius U(4,32) (interior 0, shell within 1.5h exact distance,
rest 1e6), generated on the GPU. A step = reset the device field from the
resident init (device copy) + one converged plain FSM solve (the default:
passes of the 8 orderings over the field held in 32-bit fixed point, FP32
local solves, until a pass changes no value; within 1e-4 h of the FP64
solution with no polish) + the FP64 residual judge.

Also reported (in "extra"): FP32 storage (its fixed point alone, and with
the FP64 polish that brings it within the bound); the certified stop (passes
until the FP64 judge passes) and the learned warm start on the same field and
on a fresh 256x256x128 terrain field; a batch of 512^3 sphere fields (this field and
the next seeds) in shared persistent launches, plain, certified and
warm-started, against the single-field plain rate; the 512-field uneven batch
solved in shared persistent launches (plain, certified, warm-started) against
each field in its own launches; FP32 vs FP64 error; the reference CPU core's
rate.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 512]
       [--no-oracle] [--no-cpu] [--batch N] [--batch512 K] [--no-warm]
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--tile", type=int, default=-1)
    ap.add_argument("--no-oracle", action="store_true", help="skip the FP64 reference-arithmetic check")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU reference baseline")
    ap.add_argument("--cpu-orderings", type=int, default=3)
    ap.add_argument("--batch", type=int, default=512, help="fields of the c4 batch leg (0: skip)")
    ap.add_argument("--batch512", type=int, default=8,
                    help="fields of the same-size batch leg (0: skip)")
    ap.add_argument("--no-warm", action="store_true", help="skip the certified / warm-start legs")
    ap.add_argument("--no-speed", action="store_true",
                    help="skip the spatially varying speed leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run plain 1 GiB copies (for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--profile-ref64", action="store_true",
                    help="only run the FP64 reference-arithmetic solve of the --size^3 sphere field "
                         "(the path behind run_fsm), twice, for a rocprofv3 run; prints its statistics")
    ap.add_argument("--profile-batch", choices=("512", "c4"),
                    help="only run one batch's plain solves (8 x 512^3 spheres, or the c4 batch of "
                         "--batch fields) for a rocprofv3 run; prints their statistics")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.profile_batch:
        profile_batch(args)
        return
    if args.profile_ref64:
        profile_ref64(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the single-field solve does not shard across GPUs in this build.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import pdfs_amd

    dev = torch.device("cuda:0")
    n = args.size
    X = Y = Z = n
    N = X * Y * Z
    ctx = pdfs_amd.Context()
    init = torch.empty((Z, Y, X), dtype=torch.float32, device=dev)
    log(f"[bench] generating spheres {n}^3 seed={args.seed}")
    ctx.generate(pdfs_amd.GEN_SPHERES, args.seed, X, Y, Z, 1.0, dense=init)
    field = torch.empty_like(init)

    def step(profile=0):
        field.copy_(init)
        return ctx.solve_f32(field, init, tile=args.tile, profile=profile)

    for w in range(args.warmup):
        st = step()
        log(f"[bench] warmup {w}: passes={st['passes']} ms_total={st['ms_total']:.2f} "
            f"flagged={st['judge_flagged']}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for k in range(args.steps):
        stats.append(step())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    fields_per_s = args.steps / elapsed
    st = stats[-1]
    log(f"[bench] timed {args.steps} steps: {ms_per_step:.2f} ms/step passes={st['passes']}")

    # rooflines of the two kernels of a solve, from per-launch HIP events on
    # the solver's stream in one profiled solve.
    # Column sweep (fixed point or FP32 storage): the solver counts its
    # launches (an overlapped launch sweeps a pass, or half a pass, of
    # orderings; otherwise one launch per ordering). Algorithmic bytes: 8 B
    # per executed point-update (32-bit read + write), averaged over the
    # same launches as the time (and as rocprofv3's per-dispatch average).
    # FP64 polish sweep: 20 B per point-update (FP64 read + write, FP32 init
    # read), one launch per polish ordering.
    pst = step(profile=1)
    launches = sweep_launches(pst)
    avg_launch_ms = pst["ms_sweep_kernels"] / launches
    algo_bytes = 8.0 * pst["point_updates"] / launches
    achieved = algo_bytes / (avg_launch_ms * 1e-3) / 1e9
    sweep_gbs_total = 8.0 * N * pst["orderings"] / (pst["ms_sweeps"] * 1e-3) / 1e9
    pl = max(1, pst["polish_orderings"])
    pol_ms = pst["ms_polish_kernels"] / pl
    pol_bytes = 20.0 * pst["polish_point_updates"] / pl
    pol_gbs = pol_bytes / (pol_ms * 1e-3) / 1e9 if pol_ms > 0 else 0.0
    log(f"[bench] FP32 sweep avg {avg_launch_ms:.3f} ms/launch -> {achieved:.0f} GB/s "
        f"({achieved / HBM_PEAK_GBS:.1%} of peak); FP64 polish avg {pol_ms:.3f} ms/launch x "
        f"{pl} -> {pol_gbs:.0f} GB/s; judge {pst['ms_judge']:.3f} ms")
    result = field.clone()

    extra = {"passes": st["passes"], "orderings": st["orderings"],
             "storage": "fixed point" if st["storage"] == pdfs_amd.PDFS_STORE_FIXED else "fp32",
             "ms_convert": round(pst["ms_convert"], 3),
             "polish_passes": st["polish_passes"], "polish_orderings": st["polish_orderings"],
             "ms_polish": round(pst["ms_polish"], 3), "ms_fp32_sweeps": round(pst["ms_sweeps"], 3),
             "judge_flagged": int(st["judge_flagged"]), "ms_judge": round(pst["ms_judge"], 4),
             # the FP64 judge reads the FP32 field and its init: 8 B per point
             # (flag writes only for failing points), one run per plain solve
             "judge_gbs": round(8.0 * N / (pst["ms_judge"] * 1e-3) / 1e9, 1)
             if pst["judge_runs"] == 1 and pst["ms_judge"] > 0 else None,
             "sweep_pass_gbs": round(sweep_gbs_total, 1)}
    # FP64 reference-arithmetic solve of the same field on the GPU (the
    # column pipeline behind run_fsm, bit-identical to the reference core at
    # small sizes): the FP64 solution the FP32 results are measured against
    ref_passes = None
    u64 = None
    if not args.no_oracle:
        padded = torch.empty((Z + 2, Y + 2, X + 2), dtype=torch.float64, device=dev)
        ctx.generate(pdfs_amd.GEN_SPHERES, args.seed, X, Y, Z, 1.0, padded=padded)
        torch.cuda.synchronize()
        t = time.perf_counter()
        rst = ctx.solve_ref64(padded, passes=0, tol=1e-12)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t
        log(f"[bench] FP64 reference-arithmetic solve: passes={rst['passes']} {t64:.3f} s")
        ref_passes = rst["passes"]
        u64 = padded[1:-1, 1:-1, 1:-1].contiguous()
        del padded
        err = (result.double() - u64).abs()
        extra["f32_vs_f64_max_err"] = float(err.max())
        extra["f32_vs_f64_rms_err"] = float(err.pow(2).mean().sqrt())
        extra["within_1e-4h"] = bool(extra["f32_vs_f64_max_err"] <= 1e-4)
        extra["ref64_passes"] = ref_passes
        extra["ref64_seconds"] = round(t64, 4)
        extra["ref64_ms_per_ordering"] = round(rst["ms_sweeps"] / max(1, rst["orderings"]), 4)
        log(f"[bench] result vs FP64: max {extra['f32_vs_f64_max_err']:.3e} "
            f"rms {extra['f32_vs_f64_rms_err']:.3e}")
        del err
    net = None
    if not args.no_warm:
        net = pdfs_amd.Net(ctx, path=pdfs_amd.WEIGHTS_PATH) if os.path.exists(pdfs_amd.WEIGHTS_PATH) else None
        extra["modes_512"] = solve_modes(ctx, net, init, field, args.tile, ms_per_step, u64)
        del u64
        u64 = None
        torch.cuda.empty_cache()
        tinit = torch.empty((128, 256, 256), dtype=torch.float32, device=dev)
        tpad = torch.empty((130, 258, 258), dtype=torch.float64, device=dev)
        ctx.generate_ex(pdfs_amd.GEN_TERRAIN, 777, 256, 256, 128, 1.0, 8, (5.0, 40.0), (10.0, 60.0),
                        dense=tinit, padded=tpad)
        ctx.solve_ref64(tpad, passes=0, tol=1e-12)
        tu64 = tpad[1:-1, 1:-1, 1:-1].contiguous()
        del tpad
        extra["modes_terrain_256x256x128"] = solve_modes(ctx, net, tinit, torch.empty_like(tinit),
                                                         args.tile, None, tu64)
        del tinit, tu64
    del u64

    if not args.no_speed:
        extra["speed_512"] = speed_leg(ctx, init, field, args)
    if not args.no_oracle:
        extra["parity"] = parity_leg(ctx, args)
    if args.batch512 > 0:
        extra["batch_512"] = same_size_batch_leg(ctx, args, init, net, fields_per_s)
    if args.batch > 0:
        extra["batch"] = batch_leg(ctx, args.batch, net)
        if not args.no_cpu:
            extra["batch"]["cpu_reference"] = cpu_batch_baseline(ctx, args.batch, extra["batch"])
    if net is not None:
        net.close()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ctx, args, ref_passes, X, Y, Z)

    out = {
        "metric": "converged fields/sec + GB/s vs HBM peak at 512^3 FP32; warm-start vs plain FSM",
        "value": round(fields_per_s, 4),
        "unit": "fields/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (GPU generator, seeded)",
        "config": {"workload": f"plain FSM, multi-sphere distance {n}^3, h=1, 64 spheres r~U(4,32), "
                               "FP32 local solves on the field held in 32-bit fixed point until a "
                               "pass changes nothing, + FP64 judge",
                   "grid": [X, Y, Z], "seed": args.seed, "solver": "plain", "tile": st["tile"],
                   "storage": extra["storage"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(
                         "fixed" if st["storage"] == pdfs_amd.PDFS_STORE_FIXED else "sweep"),
                     "kernel": ("k_sweep_col (fixed point, mode 7)"
                                if st["storage"] == pdfs_amd.PDFS_STORE_FIXED
                                else "k_sweep_col (FP32, mode 0)"),
                     "avg_launch_ms": round(avg_launch_ms, 4),
                     "algorithmic_bytes_per_launch": round(algo_bytes),
                     "launches": launches, "orderings_swept": pst["orderings"]},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    # the other kernel of a solve with FP32 storage: the FP64 polish sweep
    # (mode 3); the fixed-point default runs no polish
    if pst["polish_orderings"] > 0:
        out["extra"]["roofline_polish"] = {
            "bound": "hbm", "achieved": round(pol_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(pol_gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic("polish"),
            "kernel": "k_sweep_col (FP64 polish, mode 3)",
            "avg_launch_ms": round(pol_ms, 4), "algorithmic_bytes_per_launch": round(pol_bytes),
            "launches": pst["polish_orderings"], "ms_total": round(pst["ms_polish_kernels"], 3)}
        # the dominant kernel by time carries the top-level roofline
        if pst["ms_polish_kernels"] > pst["ms_sweep_kernels"]:
            out["roofline"], out["extra"]["roofline_polish"] = (
                out["extra"]["roofline_polish"], out["roofline"])
    print(json.dumps(out), flush=True)
    ctx.close()


def copy_calibration():
    """Plain copies of known size for the HBM counter calibration: 1 GiB read
    + 1 GiB written by torch's copy (runtime blit kernel) and by an
    element-wise kernel (x + 0), three times each."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    for _ in range(3):
        torch.add(a, 0.0, out=b)
    torch.cuda.synchronize()
    print(f"copy calibration: {n * 4} bytes read and written per kernel", flush=True)


def profile_ref64(args):
    """The FP64 reference-arithmetic solve (column pipeline, overlapped passes
    where the field tiles exactly) of the sphere field, until a pass changes
    nothing by more than 1e-12, twice; one JSON line with the second solve's
    passes, launches and times. Algorithmic bytes: 16 B per point-update."""
    import torch

    import pdfs_amd
    n = args.size
    ctx = pdfs_amd.Context()
    padded = torch.empty((n + 2, n + 2, n + 2), dtype=torch.float64, device="cuda:0")
    ctx.generate(pdfs_amd.GEN_SPHERES, args.seed, n, n, n, 1.0, padded=padded)
    work = torch.empty_like(padded)
    for _ in range(2):
        work.copy_(padded)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = ctx.solve_ref64(work, passes=0, tol=1e-12)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    launches = max(1, st["sweep_launches"])
    out = {"workload": f"FP64 reference arithmetic, spheres {n}^3, seed {args.seed}",
           "seconds": round(dt, 4), "passes": st["passes"], "orderings": st["orderings"],
           "sweep_launches": launches, "ms_sweeps": round(st["ms_sweeps"], 3),
           "algorithmic_bytes_per_launch": round(16.0 * st["point_updates"] / launches),
           "gbs_over_ms_sweeps": round(16.0 * st["point_updates"] / (st["ms_sweeps"] * 1e-3) / 1e9, 1)}
    print(json.dumps(out), flush=True)
    ctx.close()


def profile_batch(args):
    """The plain batched solve of the 8 x 512^3 batch (seeds seed .. seed + 7)
    or of the c4 batch, twice: the first tunes the tile (its first pass tries
    each variant), the second runs on the tuned tile with per-launch HIP
    events. Prints the second solve's statistics as one JSON line: sweep and
    polish launches, their average times and algorithmic GB/s (8 B per
    executed FP32 point-update, 20 B per polish point-update)."""
    import torch

    import pdfs_amd
    from pdfs_amd.workloads import c4_specs, generate_c4

    ctx = pdfs_amd.Context()
    dev = torch.device("cuda:0")
    if args.profile_batch == "512":
        inits = []
        for i in range(8):
            t = torch.empty((512, 512, 512), dtype=torch.float32, device=dev)
            ctx.generate(pdfs_amd.GEN_SPHERES, args.seed + i, 512, 512, 512, 1.0, dense=t)
            inits.append(t)
        name = "8 x 512^3 spheres"
    else:
        inits = generate_c4(ctx, c4_specs(seed=4, nfields=args.batch))
        name = f"c4, {args.batch} uneven fields"
    fields = [t.clone() for t in inits]
    ctx.solve_batch_f32(fields, inits, tile=pdfs_amd.TILE_AUTO, judge=0)
    for f, t in zip(fields, inits):
        f.copy_(t)
    st, passes, _ = ctx.solve_batch_f32(fields, inits, tile=pdfs_amd.TILE_AUTO, judge=0, profile=1)
    launches = sweep_launches(st)
    sweep_ms = st["ms_sweep_kernels"] / max(1, launches)
    pol_n = max(1, st["polish_orderings"])
    pol_ms = st["ms_polish_kernels"] / pol_n
    out = {"batch": name, "points": sum(t.numel() for t in inits), "tile": st["tile"],
           "passes_max": max(passes), "orderings_swept": st["orderings"], "sweep_launches": launches,
           "sweep_avg_launch_ms": round(sweep_ms, 4),
           "sweep_gbs": round(8.0 * st["point_updates"] / (st["ms_sweep_kernels"] * 1e-3) / 1e9, 1),
           "polish_launches": st["polish_orderings"], "polish_avg_launch_ms": round(pol_ms, 4),
           "polish_gbs": round(20.0 * st["polish_point_updates"] /
                               (st["ms_polish_kernels"] * 1e-3) / 1e9, 1)
           if st["ms_polish_kernels"] > 0 else None,
           "ms_sweeps": round(st["ms_sweeps"], 3), "ms_polish": round(st["ms_polish"], 3)}
    print(json.dumps(out), flush=True)
    ctx.close()


def sweep_launches(st):
    """Launches of the column sweep (fixed point or FP32 storage) in a solve,
    as the solver counts them (the FP64 polish's launches are not among
    them)."""
    return max(1, st["sweep_launches"])


def measured_traffic(kernel="sweep"):
    """HBM bytes per launch on this workload from the committed rocprofv3 PMC
    summaries (FETCH_SIZE + WRITE_SIZE, each scaled by the counter's reading
    for a plain copy of known size): the fixed-point sweep (kernel="fixed"),
    the FP32-storage sweep ("sweep") or the FP64 polish ("polish"); or None."""
    name = "k_sweep_col_fixed_spheres512.json" if kernel == "fixed" else "k_sweep_col_spheres512.json"
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    if kernel == "polish":
        return d.get("polish", {}).get("hbm_bytes_per_launch")
    return d.get("hbm_bytes_per_launch")


# warm-start network convolutions: (Cin, Cout, level); a level-l grid has 1/8^l
# of the coarse points (the coarse grid is 1/4 of the field per axis)
NET_CONVS = ((16, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2),
             (64, 64, 3), (64, 64, 3), (128, 64, 2), (96, 32, 1), (48, 16, 0), (16, 16, 0))


def network_gflop(X, Y, Z):
    """Forward FLOPs of the warm-start network on a field: sum over its 3x3x3
    convolutions of 2 * 27 * Cin * Cout * points."""
    import pdfs_amd
    Dz, Dy, Dx = pdfs_amd.net_grid(X, Y, Z)
    v0 = Dz * Dy * Dx
    return sum(2.0 * 27 * ci * co * v0 / 8 ** lvl for ci, co, lvl in NET_CONVS) / 1e9


def solve_modes(ctx, net, init, field, tile, plain_ms, u64=None):
    """Plain, certified and warm-started solves of one field: passes, time,
    judge result, fallbacks (best of 3 for the timed modes), and with the
    FP64 solution u64 each mode's max / RMS error against it (in h)."""
    import torch

    import pdfs_amd
    out = {}

    def err(d):
        if u64 is not None:
            e = (field.double() - u64).abs()
            d["max_err"] = float(e.max())
            d["rms_err"] = float(e.pow(2).mean().sqrt())
            d["within_1e-4h"] = bool(d["max_err"] <= 1e-4)
        return d

    def timed(fn, reps=3):
        best, st = None, None
        for _ in range(reps):
            field.copy_(init)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            s = fn()
            torch.cuda.synchronize()
            dt = 1000.0 * (time.perf_counter() - t0)
            if best is None or dt < best:
                best, st = dt, s
        return best, st

    if plain_ms is None:
        plain_ms, sp = timed(lambda: ctx.solve_f32(field, init, tile=tile))
    else:
        field.copy_(init)
        sp = ctx.solve_f32(field, init, tile=tile)
    S32 = pdfs_amd.PDFS_STORE_F32
    out["plain"] = err({"passes": sp["passes"], "orderings": sp["orderings"],
                        "ms": round(plain_ms, 3), "ms_convert": round(sp["ms_convert"], 3),
                        "sweep_launches": sp["sweep_launches"],
                        "judge_flagged": int(sp["judge_flagged"])})
    # the same solve with one launch per ordering (no overlapped passes)
    oms, so = timed(lambda: ctx.solve_f32(field, init, tile=tile,
                                          overlap=pdfs_amd.PDFS_OVERLAP_OFF))
    out["plain_one_launch_per_ordering"] = err(
        {"passes": so["passes"], "ms": round(oms, 3), "sweep_launches": so["sweep_launches"],
         "judge_flagged": int(so["judge_flagged"]),
         "overlapped_speedup": round(oms / plain_ms, 3)})
    pms, sq = timed(lambda: ctx.solve_f32(field, init, tile=tile, storage=S32))
    out["plain_fp32_storage_polished"] = err(
        {"passes": sq["passes"], "orderings": sq["orderings"], "ms": round(pms, 3),
         "polish_passes": sq["polish_passes"], "ms_polish": round(sq["ms_polish"], 3),
         "judge_flagged": int(sq["judge_flagged"]),
         "note": "FP32 storage, then the FP64 polish"})
    rms, sr = timed(lambda: ctx.solve_f32(field, init, tile=tile, storage=S32, polish=0))
    out["plain_fp32_only"] = err({"passes": sr["passes"], "ms": round(rms, 3),
                                  "judge_flagged": int(sr["judge_flagged"]),
                                  "note": "the FP32-storage fixed point without the polish"})
    cms, sc = timed(lambda: ctx.solve_f32(field, init, tile=tile, certify=1))
    out["certified"] = err({"passes": sc["passes"], "orderings": sc["orderings"],
                            "ms": round(cms, 3),
                            "judge_runs": sc["judge_runs"], "polish_passes": sc["polish_passes"],
                            "judge_flagged": int(sc["judge_flagged"]),
                            "speedup_vs_plain": round(plain_ms / cms, 3)})
    if net is not None:
        wms, sw = timed(lambda: ctx.solve_warm_f32(net, field, init, tile=tile))
        Z, Y, X = init.shape
        gflop = network_gflop(X, Y, Z)
        out["warm"] = err({"passes": sw["passes"], "ms": round(wms, 3),
                           "ms_network": round(sw["ms_warm"], 3),
                           "network_gflop": round(gflop, 1),
                           "network_tflops": round(gflop / sw["ms_warm"], 2) if sw["ms_warm"] > 0 else None,
                           "polish_passes": sw["polish_passes"],
                           "fallbacks": sw["fallbacks"], "judge_flagged": int(sw["judge_flagged"]),
                           "speedup_vs_plain": round(plain_ms / wms, 3)})
    log(f"[bench] modes {tuple(init.shape)}: {out}")
    return out


def slowness_field(Z, Y, X, dev):
    """Slowness (1 / speed) in [0.5, 2] for the speed leg: four layers in z,
    modulated by two smooth lateral waves (float32 [Z, Y, X])."""
    import torch
    z = torch.arange(Z, device=dev, dtype=torch.float32).view(Z, 1, 1)
    y = torch.arange(Y, device=dev, dtype=torch.float32).view(1, Y, 1)
    x = torch.arange(X, device=dev, dtype=torch.float32).view(1, 1, X)
    layer = torch.floor(z * 4.0 / Z)
    wave = torch.sin(0.031 * x + 0.017 * y + 0.4) + torch.cos(0.023 * y - 0.011 * z)
    f = layer / 3.0 * 0.6 + (wave + 2.0) / 4.0 * 0.4
    return (0.5 + 1.5 * f).contiguous()


def speed_leg(ctx, init, field, args):
    """Spatially varying speed at the headline size: the multi-sphere field
    with a layered, laterally varying slowness in [0.5, 2]. The default
    solve (FP32 speed sweeps to the fixed point, FP64 speed polish, judge)
    timed best of 3; the FP32 fixed point alone; and both against the
    converged FP64 speed solution (reference arithmetic with the point's
    spacings scaled by its slowness, pdfs_solve_speed64_dev)."""
    import torch
    Z, Y, X = init.shape
    dev = init.device
    s = slowness_field(Z, Y, X, dev)
    out = {"slowness": "[0.5, 2]: 4 z-layers x smooth lateral waves"}

    def timed(fn, reps=3):
        best, st = None, None
        for _ in range(reps):
            field.copy_(init)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            dt = 1000.0 * (time.perf_counter() - t0)
            if best is None or dt < best:
                best, st = dt, r
        return best, st

    ms, st = timed(lambda: ctx.solve_speed_f32(field, s, init=init, tile=args.tile, profile=0))
    res = field.clone()
    ms_raw, sr = timed(lambda: ctx.solve_speed_f32(field, s, init=init, tile=args.tile, polish=0),
                       reps=1)
    raw = field.clone()
    field.copy_(init)
    pst = ctx.solve_speed_f32(field, s, init=init, tile=args.tile, profile=1)
    launches = sweep_launches(pst)
    out.update({"ms": round(ms, 3), "fields_per_s": round(1000.0 / ms, 3), "passes": st["passes"],
                "orderings": st["orderings"], "polish_passes": st["polish_passes"],
                "judge_flagged": int(st["judge_flagged"]), "tile": st["tile"],
                "ms_fp32_only": round(ms_raw, 3),
                "fp32_ms_per_launch": round(pst["ms_sweep_kernels"] / launches, 4),
                "polish_ms_per_launch": round(pst["ms_polish_kernels"]
                                              / max(1, pst["polish_orderings"]), 4)})
    if not args.no_oracle:
        pad = torch.full((Z + 2, Y + 2, X + 2), float("inf"), dtype=torch.float64, device=dev)
        pad[1:-1, 1:-1, 1:-1] = init.double()
        sp = torch.ones_like(pad)
        sp[1:-1, 1:-1, 1:-1] = s.double()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r64 = ctx.solve_speed64(pad, sp, passes=0, tol=1e-12)
        torch.cuda.synchronize()
        out["speed64_seconds"] = round(time.perf_counter() - t0, 3)
        out["speed64_passes"] = r64["passes"]
        u64 = pad[1:-1, 1:-1, 1:-1]
        for name, v in (("", res), ("fp32_only_", raw)):
            e = (v.double() - u64).abs()
            out[name + "max_err"] = float(e.max())
            out[name + "rms_err"] = float(e.pow(2).mean().sqrt())
        out["within_1e-4h"] = bool(out["max_err"] <= 1e-4)
        del pad, sp, u64
    del s, res, raw
    torch.cuda.empty_cache()
    log(f"[bench] speed {tuple(init.shape)}: {out}")
    return out


def parity_leg(ctx, args):
    """GPU against the reference on the small configurations: c0 (64^3
    point source, seed 12345, FP64 to max |delta| <= 1e-12) on the GPU's
    reference-arithmetic path and on the reference core (oracle/_ref, when
    present): passes, time, SHA-256 of the padded field, bitwise equality;
    and the converged FP32 field's max / RMS error against the converged FP64
    field on c0 and on the c1 terrain (256x256x128, 8 hills)."""
    import hashlib

    import numpy as np
    import torch

    import pdfs_amd
    out = {}

    def f32_err(kind, seed, X, Y, Z, nobj=0, p=(0.0, 0.0), q=(0.0, 0.0)):
        pad = torch.empty((Z + 2, Y + 2, X + 2), dtype=torch.float64, device="cuda:0")
        dense = torch.empty((Z, Y, X), dtype=torch.float32, device="cuda:0")
        ctx.generate_ex(kind, seed, X, Y, Z, 1.0, nobj, p, q, dense=dense, padded=pad)
        t = time.perf_counter()
        r64 = ctx.solve_ref64(pad, passes=0, tol=1e-12)
        t64 = time.perf_counter() - t
        init = dense.clone()
        S32 = pdfs_amd.PDFS_STORE_F32
        s32 = ctx.solve_f32(dense, init, judge=0, polish=0, storage=S32)
        err = (dense.double() - pad[1:-1, 1:-1, 1:-1]).abs()
        dense.copy_(init)
        ctx.solve_f32(dense, init, judge=0, storage=S32)
        perr = float((dense.double() - pad[1:-1, 1:-1, 1:-1]).abs().max())
        dense.copy_(init)
        ctx.solve_f32(dense, init, judge=0)
        ferr = (dense.double() - pad[1:-1, 1:-1, 1:-1]).abs()
        perr = {"f32_polished_vs_f64_max_err": perr,
                "fixed_vs_f64_max_err": float(ferr.max()),
                "fixed_vs_f64_rms_err": float(ferr.pow(2).mean().sqrt())}
        return pad, r64, t64, s32, float(err.max()), float(err.pow(2).mean().sqrt()), perr

    X = Y = Z = 64
    pad, r64, t64, s32, emax, erms, perr = f32_err(pdfs_amd.GEN_POINT, 12345, X, Y, Z)
    host = pad.cpu().numpy()
    c0 = {"grid": [X, Y, Z], "seed": 12345, "gpu_f64_passes": r64["passes"],
          "gpu_f64_seconds": round(t64, 4), "sha256": hashlib.sha256(host.tobytes()).hexdigest(),
          "f32_passes": s32["passes"], "f32_only_vs_f64_max_err": emax,
          "f32_only_vs_f64_rms_err": erms, **perr}
    so = os.path.join(ROOT, "oracle", "_ref", "libpdfs_ref.so")
    if os.path.exists(so):
        ref = torch.empty_like(pad)
        ctx.generate(pdfs_amd.GEN_POINT, 12345, X, Y, Z, 1.0, padded=ref)
        a = ref.cpu().numpy()
        passes, sec = _ref_solve_worker((so, X, Y, Z, a))
        c0.update({"cpu_reference_passes": passes, "cpu_reference_seconds": round(sec, 4),
                   "cpu_reference_sha256": hashlib.sha256(a.tobytes()).hexdigest(),
                   "bitwise_equal": bool(np.array_equal(a.view(np.uint64), host.view(np.uint64)))})
    out["c0"] = c0
    X, Y, Z = 256, 256, 128
    _, r64, t64, s32, emax, erms, perr = f32_err(pdfs_amd.GEN_TERRAIN, 777, X, Y, Z, 8,
                                                 (5.0, 40.0), (10.0, 60.0))
    out["c1"] = {"grid": [X, Y, Z], "seed": 777, "gpu_f64_passes": r64["passes"],
                 "gpu_f64_seconds": round(t64, 3), "f32_passes": s32["passes"],
                 "f32_only_vs_f64_max_err": emax, "f32_only_vs_f64_rms_err": erms, **perr}
    if os.path.exists(so) and not args.no_cpu:
        # the reference core on one host core, two passes of the same init
        # (the converged solve takes about a minute there), against two
        # passes of the GPU's reference-arithmetic path: bitwise, and the CPU
        # time per pass scaled to the passes the GPU's FP64 solve converged in
        ref = torch.empty((Z + 2, Y + 2, X + 2), dtype=torch.float64, device="cuda:0")
        ctx.generate_ex(pdfs_amd.GEN_TERRAIN, 777, X, Y, Z, 1.0, 8, (5.0, 40.0), (10.0, 60.0),
                        padded=ref)
        a = ref.cpu().numpy()
        ctx.solve_ref64(ref, passes=2)
        passes, sec = _ref_solve_worker((so, X, Y, Z, a, -1.0, 2))
        out["c1"].update({"cpu_reference_passes_run": passes,
                          "cpu_reference_seconds_per_pass": round(sec / passes, 3),
                          "cpu_reference_converged_seconds_scaled": round(
                              sec / passes * r64["passes"], 2),
                          "bitwise_equal_after_2_passes": bool(np.array_equal(
                              a.view(np.uint64), ref.cpu().numpy().view(np.uint64)))})
        del ref, a
    log(f"[bench] parity: {out}")
    torch.cuda.empty_cache()
    return out


def same_size_batch_leg(ctx, args, init, net, single_fields_per_s):
    """K fields of the value's workload (seeds seed, seed + 1, ...) solved
    together: each ordering is one persistent launch over all K fields
    (graph-replayed passes, tile tuned online and cached), plain, certified
    and warm-started. Fields/s against the single-field plain rate; the sweep
    kernel's algorithmic bandwidth over the plain batch's executed
    point-updates."""
    import torch

    import pdfs_amd
    K = args.batch512
    Z, Y, X = init.shape
    inits = [init]
    for i in range(1, K):
        t = torch.empty_like(init)
        ctx.generate(pdfs_amd.GEN_SPHERES, args.seed + i, X, Y, Z, 1.0, dense=t)
        inits.append(t)
    fields = [torch.empty_like(t) for t in inits]
    log(f"[bench] {K} x {X}x{Y}x{Z} batch")

    def reset():
        for f, t in zip(fields, inits):
            f.copy_(t)
        torch.cuda.synchronize()

    def timed(fn):
        reset()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, r

    auto = pdfs_amd.TILE_AUTO
    # plain: FP32 passes until a pass changes nothing, then the FP64 polish
    # (the inits are passed for it), no final judge
    timed(lambda: ctx.solve_batch_f32(fields, inits, tile=auto, graph=1, judge=0))  # tune
    tp, (sp, pp, _) = timed(lambda: ctx.solve_batch_f32(fields, inits, tile=auto, graph=1, judge=0))
    _, (sk, _, _) = timed(lambda: ctx.solve_batch_f32(fields, inits, tile=auto, profile=1, judge=0))
    gbs = 8.0 * sk["point_updates"] / (sk["ms_sweep_kernels"] * 1e-3) / 1e9
    pgbs = (20.0 * sk["polish_point_updates"] / (sk["ms_polish_kernels"] * 1e-3) / 1e9
            if sk["ms_polish_kernels"] > 0 else None)
    out = {"fields": K, "grid": [X, Y, Z], "tile": sp["tile"],
           "plain": {"seconds": round(tp, 4), "fields_per_s": round(K / tp, 2),
                     "passes_max": max(pp), "polish_passes": sp["polish_passes"],
                     "ms_polish": round(sp["ms_polish"], 3),
                     "speedup_vs_single_plain": round(K / tp / single_fields_per_s, 3),
                     "sweep_gbs": round(gbs, 1), "sweep_frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
                     "polish_gbs": round(pgbs, 1) if pgbs else None}}
    tc, (sc, pc, _) = timed(lambda: ctx.solve_batch_f32(fields, inits, tile=auto, graph=1, certify=1))
    out["certified"] = {"seconds": round(tc, 4), "fields_per_s": round(K / tc, 2),
                        "passes_max": max(pc), "judge_flagged": int(sc["judge_flagged"]),
                        "ms_polish": round(sc["ms_polish"], 3),
                        "speedup_vs_single_plain": round(K / tc / single_fields_per_s, 3)}
    if net is not None:
        tw, (sw, pw, fb) = timed(lambda: ctx.solve_batch_warm_f32(net, fields, inits, tile=auto,
                                                                 graph=1))
        out["warm"] = {"seconds": round(tw, 4), "fields_per_s": round(K / tw, 2),
                       "passes_max": max(pw), "fallbacks": int(sum(fb)),
                       "ms_network": round(sw["ms_warm"], 3), "judge_flagged": int(sw["judge_flagged"]),
                       "ms_polish": round(sw["ms_polish"], 3),
                       "speedup_vs_single_plain": round(K / tw / single_fields_per_s, 3)}
    log(f"[bench] {K} x 512^3 batch: {out}")
    del inits, fields
    torch.cuda.empty_cache()
    return out


def batch_leg(ctx, nfields, net=None):
    """c4: uneven fields in one persistent launch per ordering (graph-replayed
    passes, tile tuned online), against each field solved by its own
    launches, with per-field completion times (ms since the batch start);
    the same batch with certified stop and with the learned warm start."""
    import numpy as np
    import torch

    import pdfs_amd
    from pdfs_amd.workloads import c4_specs, generate_c4

    specs = c4_specs(seed=4, nfields=nfields)
    inits = generate_c4(ctx, specs)
    npts = sum(t.numel() for t in inits)
    log(f"[bench] c4 batch: {nfields} fields, {npts / 1e6:.1f} M points")
    # warm the tuner / graph once, then time (plain stop, then the FP64
    # polish: the inits are passed for it; no final judge)
    fields = [t.clone() for t in inits]
    ctx.solve_batch_f32(fields, inits, tile=pdfs_amd.TILE_AUTO, graph=1, judge=0)
    fields = [t.clone() for t in inits]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st, passes, ms = ctx.solve_batch_f32(fields, inits, tile=pdfs_amd.TILE_AUTO, graph=1, judge=0)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    batched = [f.clone() for f in fields]
    # each field in its own launches, one after another
    fields = [t.clone() for t in inits]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = []
    for f, i0 in zip(fields, inits):
        ctx.solve_f32(f, i0, judge=0)
        done.append(1000.0 * (time.perf_counter() - t0))
    ts = time.perf_counter() - t0
    same = all(torch.equal(a, b) for a, b in zip(batched, fields))
    out = {"fields": nfields, "points": npts, "seconds": round(tb, 4),
           "fields_per_s": round(nfields / tb, 2), "passes_max": max(passes),
           "polish_passes": st["polish_passes"], "ms_polish": round(st["ms_polish"], 3),
           "p50_p99_note": "per-field completion of the FP32 passes; the FP64 polish then runs "
                           "for all fields together (ms_polish)",
           "p50_ms": round(float(np.percentile(ms, 50)), 3),
           "p99_ms": round(float(np.percentile(ms, 99)), 3), "tile": st["tile"],
           "per_field_launches": {"seconds": round(ts, 4), "fields_per_s": round(nfields / ts, 2),
                                  "p50_ms": round(float(np.percentile(done, 50)), 3),
                                  "p99_ms": round(float(np.percentile(done, 99)), 3)},
           "speedup": round(ts / tb, 2), "bitwise_equal_to_per_field": bool(same)}
    # certified stop in the batch (per-field FP64 judge from pass 1)
    fields = [t.clone() for t in inits]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stc, pc, _ = ctx.solve_batch_f32(fields, inits, tile=pdfs_amd.TILE_AUTO, graph=1, certify=1)
    torch.cuda.synchronize()
    tc = time.perf_counter() - t0
    out["certified"] = {"seconds": round(tc, 4), "fields_per_s": round(nfields / tc, 2),
                        "passes_max": max(pc), "judge_flagged": int(stc["judge_flagged"]),
                        "speedup_vs_per_field_plain": round(ts / tc, 2)}
    if net is not None:
        fields = [t.clone() for t in inits]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stw, pw, fb = ctx.solve_batch_warm_f32(net, fields, inits, tile=pdfs_amd.TILE_AUTO, graph=1)
        torch.cuda.synchronize()
        tw = time.perf_counter() - t0
        out["warm"] = {"seconds": round(tw, 4), "fields_per_s": round(nfields / tw, 2),
                       "ms_network": round(stw["ms_warm"], 3), "fallbacks": int(sum(fb)),
                       "passes_max": max(pw), "judge_flagged": int(stw["judge_flagged"]),
                       "speedup_vs_per_field_plain": round(ts / tw, 2)}
    log(f"[bench] c4 certified / warm: {out.get('certified')} {out.get('warm')}")
    log(f"[bench] c4 batched {tb * 1e3:.1f} ms ({out['fields_per_s']} fields/s, p50 "
        f"{out['p50_ms']} p99 {out['p99_ms']} ms) vs per-field {ts * 1e3:.1f} ms; equal={same}")
    return out


def _ref_solve_worker(job):
    """One field on the reference core (oracle/_ref), in place: passes until
    one has max |delta| <= tol (default 1e-12, at most 100 passes), or
    exactly max_passes passes with tol < 0. Returns (passes, sweep seconds).
    Runs in a worker process for the c4 CPU sample."""
    import ctypes as C

    import numpy as np
    so, X, Y, Z, dist = job[:5]
    tol, max_passes = job[5:] if len(job) > 5 else (1e-12, 100)
    R = C.CDLL(so)
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    R.ref_solve.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                            C.c_double, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    R.ref_solve.restype = C.c_int
    sec, md = C.c_double(0), C.c_double(0)
    passes = R.ref_solve(dist, X, Y, Z, 1.0, 1.0, 1.0, tol, max_passes, C.byref(sec),
                         C.byref(md))
    return passes, sec.value


def cpu_batch_baseline(ctx, nfields, gpu, sample=48):
    """The reference FSM core (oracle/_ref) on a sample of the c4 batch: the
    first `sample` fields, each converged (max |delta| <= 1e-12) in its own
    worker process, one process per host core (at most 16, the box's CPU
    share for one GPU). fields/s = sample / wall time of the pool."""
    import multiprocessing as mp

    import torch

    import pdfs_amd
    from pdfs_amd.workloads import c4_specs

    so = os.path.join(ROOT, "oracle", "_ref", "libpdfs_ref.so")
    if not os.path.exists(so):
        return None
    specs = c4_specs(seed=4, nfields=nfields)[:sample]
    jobs = []
    for kind, (X, Y, Z), fseed, nobj, p, q in specs:
        t = torch.empty((Z + 2, Y + 2, X + 2), dtype=torch.float64, device="cuda:0")
        ctx.generate_ex(kind, fseed, X, Y, Z, 1.0, nobj, p, q, padded=t)
        jobs.append((so, X, Y, Z, t.cpu().numpy()))
        del t
    procs = max(1, min(16, os.cpu_count() or 1, len(jobs)))
    log(f"[bench] CPU reference core on {len(jobs)} c4 fields, {procs} processes ...")
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(procs) as pool:
        res = pool.map(_ref_solve_worker, jobs, chunksize=1)
    wall = time.perf_counter() - t0
    fps = len(jobs) / wall
    log(f"[bench] CPU c4 sample: {wall:.1f} s -> {fps:.2f} fields/s")
    return {"value": round(fps, 3), "unit": "fields/s", "cores": procs, "kind": "reference",
            "sample": f"first {len(jobs)} of the {nfields} c4 fields, each converged to max |delta| "
                      f"<= 1e-12 by the reference runFSM (mpi_openacc/src/phi3D_fsm.c, gcc -O3) "
                      f"in one of {procs} worker processes",
            "wall_seconds": round(wall, 2), "sweep_seconds": round(sum(r[1] for r in res), 2),
            "passes_max": max(r[0] for r in res),
            "gpu_batched_vs_cpu": round(gpu["fields_per_s"] / fps, 1)}


def cpu_baseline(ctx, args, ref_passes, X, Y, Z):
    """Reference FSM core (mpi_openacc/src/phi3D_fsm.c compiled from source,
    oracle/_ref) on one host core: a bounded sample of orderings over the
    same 512^3 field, scaled to converged fields/s with the reference
    arithmetic's own pass count (measured on the GPU's bitwise-equal FP64
    path)."""
    import ctypes as C

    import numpy as np
    import torch

    import pdfs_amd

    so = os.path.join(ROOT, "oracle", "_ref", "libpdfs_ref.so")
    if not os.path.exists(so):
        log("[bench] no oracle/_ref build: cpu_baseline null")
        return None
    R = C.CDLL(so)
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    R.ref_time_orderings.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.c_int]
    R.ref_time_orderings.restype = C.c_double
    padded = torch.empty((Z + 2, Y + 2, X + 2), dtype=torch.float64, device="cuda:0")
    ctx.generate(pdfs_amd.GEN_SPHERES, args.seed, X, Y, Z, 1.0, padded=padded)
    host = padded.cpu().numpy()
    del padded
    k = args.cpu_orderings
    log(f"[bench] CPU reference core: {k} orderings of {X}x{Y}x{Z} on 1 core ...")
    sec = R.ref_time_orderings(host, X, Y, Z, 1.0, 1.0, 1.0, k)
    per_ordering = sec / k
    passes = ref_passes if ref_passes else 2
    fields_per_s = 1.0 / (per_ordering * 8 * passes)
    log(f"[bench] CPU: {per_ordering:.2f} s/ordering -> {fields_per_s:.3e} fields/s "
        f"({passes} passes)")
    return {"value": fields_per_s, "unit": "fields/s", "cores": 1, "kind": "reference",
            "sample": f"{k} orderings (sn=1..{k}) of the same {X}^3 field, reference runFSM "
                      f"(mpi_openacc/src/phi3D_fsm.c, gcc -O3, serial), {sec:.1f} s; scaled by "
                      f"8 orderings x {passes} passes (FP64 reference-arithmetic pass count "
                      f"to 1e-12)",
            "seconds_per_ordering": per_ordering}


if __name__ == "__main__":
    main()
