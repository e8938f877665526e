"""Render-pass benchmark on one MI355X.

Workload: the hyperbolic cell scene, 640x480, 64 samples per pixel, 8 light
hops, 192 primitives (64 cells x (1 horosphere + 2 geodesic planes)). A step
is one full 64-spp render of the frame (19.66 M paths) with the pass-batching
kernel (`render_passes`, one launch per step). Inputs (scene, canvas, seeds)
are resident in HBM before the timed region. The plain per-pass kernel
(`render`, one launch + sync per pass, as the reference runs) is timed beside
it on the same frame.

Prints one JSON line. `roofline` prices the dominant kernel's algorithmic
FP32 flops (counted by the CPU oracle's counting scalar on a sample of the
same paths) against the MI355X FP32 vector peak. `cpu_baseline` times the CPU
oracle (a restatement of the reference kernel) on a crop of the same frame on
the host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpaths/s + time-to-RMSE 0.01 vs FP64 oracle: chosen form vs plain kernel"
FP32_VECTOR_PEAK = 157.3  # TFLOP/s, MI355X published FP32 vector (packed / dual-issue)
FP32_NONPACKED_PEAK = 78.6  # TFLOP/s without packed FP32 (this project compiles packed FP32 off)
# HBM bytes per render_passes launch on this workload, from the FETCH_SIZE /
# WRITE_SIZE counter runs scaled by the calibration copy
# (profiles/render_passes_hyperbolic_cells_strict.md): 6 615 KiB + 6 000 KiB.
MEASURED_TRAFFIC_BYTES = (6615 + 6000) * 1024


def count_flops_per_path(scene_spec, w, h, seeds, stride=16):
    """Algorithmic flops per path, counted by the oracle on every stride-th
    pixel of every stride-th row (first pass of each)."""
    import oracle

    total, paths = 0, 0
    for y in range(0, h, stride):
        c, s, f = oracle.render(scene_spec, w, h, 1, seeds, precision=oracle.COUNT, crop=(0, y, w, 1), threads=0)
        total += f
        paths += w
    return total / paths, paths


def cpu_baseline(scene_spec, w, h, seeds, budget_s=12.0):
    """Oracle FP32 on rows of the frame, all passes of one sample each, until
    about `budget_s` of wall time."""
    import oracle

    threads = min(os.cpu_count() or 1, 16)
    paths, t0 = 0, time.perf_counter()
    y = 0
    rows = 4
    while time.perf_counter() - t0 < budget_s and y < h:
        oracle.render(scene_spec, w, h, 1, seeds, crop=(0, y, w, min(rows, h - y)), threads=threads)
        paths += w * min(rows, h - y)
        y += rows
    dt = time.perf_counter() - t0
    return {"value": paths / dt / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (FP32, OpenMP) on {paths} paths: first {y} rows x 1 spp of the same frame, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="time only the chosen form")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run the extract kernel over 4096x4096 float4 (256 MiB read + 256 MiB written): "
                         "the known-size copy that scales FETCH_SIZE / WRITE_SIZE counter readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the render pass shards by "
              "pixel rows, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from hypertrace_amd import scenes
    from hypertrace_amd.processing import FAST_FP, Canvas, Context, EntityBuffer, Render, seed_fill
    from hypertrace_amd.types import X32_F32

    wl = scenes.hyperbolic_cells()
    W, H, SPP = wl.width, wl.height, wl.spp
    paths_per_step = W * H * SPP
    spec = wl.scene.spec_text(X32_F32)
    seeds0 = seed_fill(W * H, wl.seed_mode, wl.seed)

    with Context(0, X32_F32) as ctx:
        buf = EntityBuffer(ctx, wl.scene)

        def time_chosen(render, steps, warmup):
            canvas = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
            for _ in range(warmup):
                render.render_passes(buf, canvas, SPP)
            ctx.sync()
            # timed region: K steps, HIP events on the launch stream
            t0 = time.perf_counter()
            ctx.timer_begin()
            for _ in range(steps):
                render.render_passes(buf, canvas, SPP)
            gpu_ms = ctx.timer_end()
            ctx.sync()
            wall = time.perf_counter() - t0
            return wall, gpu_ms

        render = Render(ctx, wl.scene)  # default program: strict IEEE FP32
        wall, gpu_ms = time_chosen(render, args.steps, args.warmup)
        ms_per_step = wall * 1e3 / args.steps
        kernel_ms = gpu_ms / args.steps
        mpaths = paths_per_step * args.steps / wall / 1e6

        fast = None
        if not args.no_plain:
            render_fast = Render(ctx, wl.scene, fp_options=FAST_FP)
            fwall, fgpu = time_chosen(render_fast, args.steps, args.warmup)
            fast = {"kernel": "render_passes, fast FP32 (FMA contraction, reciprocal division, device libm)",
                    "value": paths_per_step * args.steps / fwall / 1e6, "unit": "Mpaths/s",
                    "kernel_ms": fgpu / args.steps}

        if args.calibrate:
            from hypertrace_amd.processing import Extractor, ImageBuffer
            big = Canvas(ctx, (4096, 4096))
            big.passes = 1
            Extractor(ctx).process(big, ImageBuffer(ctx, (4096, 4096)))
            ctx.sync()

        plain = None
        if not args.no_plain:
            canvas_p = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
            render.render(buf, canvas_p)  # warm
            ctx.sync()
            t0 = time.perf_counter()
            ctx.timer_begin()
            for _ in range(SPP):
                render.render(buf, canvas_p)
            plain_gpu_ms = ctx.timer_end()
            plain_wall = time.perf_counter() - t0
            plain = {"kernel": "render (one launch + sync per pass, as the reference runs)",
                     "value": W * H * SPP / plain_wall / 1e6, "unit": "Mpaths/s",
                     "ms_per_frame": plain_wall * 1e3, "gpu_ms_per_frame": plain_gpu_ms}

    flops_per_path, sample_paths = count_flops_per_path(spec, W, H, seeds0)
    flops_per_launch = flops_per_path * paths_per_step
    achieved = flops_per_launch / (kernel_ms * 1e-3) / 1e12
    roofline = {"bound": "valu", "achieved": achieved, "peak": FP32_VECTOR_PEAK, "unit": "TFLOP/s",
                "frac": achieved / FP32_VECTOR_PEAK, "frac_nonpacked": achieved / FP32_NONPACKED_PEAK,
                "kernel": "render_passes", "kernel_ms": kernel_ms,
                "flops_per_path": flops_per_path, "flops_sample_paths": sample_paths,
                "traffic": MEASURED_TRAFFIC_BYTES,
                "algorithmic_bytes": 40 * W * H + len(wl.scene.to_bytes(X32_F32))}

    out = {
        "metric": METRIC,
        "value": mpaths,
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "hyperbolic_cells", "width": W, "height": H, "spp": SPP, "hops": 8,
                   "primitives": 192, "kernel": "render_passes"},
        "roofline": roofline,
        "plain": plain,
        "fast_fp": fast,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(spec, W, H, seeds0)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
