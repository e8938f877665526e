"""Render-pass benchmark on one MI355X.

Workload: the hyperbolic cell scene, 640x480, 64 samples per pixel, 8 light
hops, 192 primitives (64 cells x (1 horosphere + 2 geodesic planes)). A step
is one full 64-spp render of the frame (19.66 M paths) with the chosen form
(`render_passes`: one launch per frame, in-thread path regeneration, culled
and compacted nearest-hit scan, strict IEEE FP32 — bit-identical to the plain
kernel). Inputs (scene, canvas, seeds) are resident in HBM before the timed
region. Beside it:

* plain: the per-pass kernel (`render`, one launch + sync per pass, as the
  reference runs it) with the element-by-element scan, same strict FP32;
* time-to-RMSE: wall time until the image reaches RMSE <= 0.01 against an
  FP64 4096-spp oracle image rendered on the GPU from independent seeds (RMSE
  over all pixels x RGB of the extracted, gamma-1/2.2, clamped [0, 1] image),
  chosen form vs plain kernel. Both forms produce bit-identical canvases for
  the same number of passes, so the spp needed is found once and each form is
  then timed over exactly that many passes;
* roofline: the dominant kernel's algorithmic flops (counted by the CPU
  oracle's counting scalars on a sample of the same paths, for the algorithm
  the kernel runs: FP32 exact tests of the cull's candidates plus the FP64
  cull tests) per second of its HIP-event time, against the MI355X FP32
  vector peak; the plain algorithm's count is reported beside it;
* cpu_baseline: the CPU oracle (a restatement of the reference kernel) on
  rows of the same frame, on the host's cores.

The whole result (tuner tables, RMSE curves, every leg) goes to a JSON
file (--detail; by default bench_detail.json in the run's output folder);
the last line of stdout is a summary of at most 2 KB: the headline, its
roofline, the CPU baseline, the time-to-RMSE figures and one number per
leg. Progress lines go to stderr.
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
# HBM bytes per render_passes launch on this workload for the tuner's picks
# (scalar bounds, groups of 4 with the group queue, hit order, sparse level
# 1 at 32 or 16 rays, the pixel queue in the golden-ratio order with or
# without the probe's ordering, helper lanes, wave priority), from the
# FETCH_SIZE / WRITE_SIZE counter runs scaled by the calibration copy
# (FETCH x 2.00, WRITE x 1.00; profiles/render_passes_hyperbolic_cells_culled.md),
# against 12.3 MB algorithmic: the excess is the programs' 92-96 B per lane
# of register spills (24-25 MB of scratch over the resident lanes, more
# than the L2 holds). `traffic` is null for a pick whose counters were not read.
_PICK = ("-ffp-contract=off -DHT_STRICT_FP=1 -DHT_CULL_UNROLL=4 -DHT_WAVES_PER_EU=4 -DHT_SCALAR_BOUNDS=1 "
         "-DHT_CULL_GROUPS=4 -DHT_PIXEL_QUEUE=1 -DHT_WAVE_PRIO=1 -DHT_GROUP_QUEUE=1 -DHT_HIT_ORDER=1 -DHT_HELPERS=1 "
         "-DHT_SPARSE_L1={} -DHT_PIXEL_SPREAD=1")
_PROBED = " -DHT_PIXEL_CLASSES=1"
MEASURED_TRAFFIC_BYTES = {_PICK.format(32): (2 * 35282 + 581161) * 1024,
                          _PICK.format(16): (2 * 35960 + 572511) * 1024,
                          _PICK.format(32) + _PROBED: (2 * 40926 + 652623) * 1024,
                          _PICK.format(16) + _PROBED: (2 * 41423 + 645645) * 1024}
NO_CULL = "-ffp-contract=off -DHT_STRICT_FP=1 -DHT_NO_CULL=1"
ORACLE_SEED_XOR = 0x0BAD5EED


def count_mode(knobs):
    """The oracle's counting mode for the scan a tuner setting runs: 1 = the
    flat culled scan; g = two-level over groups of g (the group tests, then
    the leaf tests of the groups each ray passes); 1000 + g = the same with
    the group queue (the pair's cull state rebuilt per passing group); 2000
    + g = the group queue with the hit order (crossing intervals of the
    passing pairs, exact tests only of the pairs the interval filter keeps)."""
    g = int(knobs.get("groups") or 0)
    if not g or not knobs.get("scalar_bounds"):
        return 1  # groups over LDS bounds build their tree per block; counted as the flat scan's tests
    if knobs.get("group_queue"):
        return (2000 if knobs.get("hit_order") else 1000) + g
    return g


def count_flops_per_path(scene_spec, w, h, seeds, stride=16, mode=1):
    """Flops per path counted by the oracle's counting scalars on every pixel
    of every stride-th row (first pass of each): the algorithm the chosen
    program's scan runs (`mode`, see count_mode: FP32 exact tests of the
    candidates + the FP64 cull tests) and the plain reference algorithm
    (every element's exact test)."""
    import oracle

    culled32 = culled64 = plain32 = 0
    paths = 0
    for y in range(0, h, stride):
        _, _, c32, c64 = oracle.count_flops(scene_spec, w, h, 1, seeds, mode, crop=(0, y, w, 1))
        _, _, p32, _ = oracle.count_flops(scene_spec, w, h, 1, seeds, False, crop=(0, y, w, 1))
        culled32 += c32
        culled64 += c64
        plain32 += p32
        paths += w
    return {"fp32": culled32 / paths, "fp64": culled64 / paths, "plain_fp32": plain32 / paths, "paths": paths}


def host_threads():
    """Host CPUs this process may run on (its affinity mask), with what the
    machine reports beside it."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return usable, {"os_cpu_count": os.cpu_count(), "affinity": usable,
                    "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(scene_spec, w, h, seeds, budget_s=12.0):
    """Oracle FP32 on bands of rows of the frame, one sample per pixel per
    pass, pass after pass (seeds advancing as on the GPU), until about
    `budget_s` of wall time, on every host CPU the process may use."""
    import oracle

    threads, host = host_threads()
    rows = 16
    canvas = np.zeros((h, w, 4), np.float32)
    seeds = np.array(seeds, dtype=np.uint32).reshape(h, w)
    paths, t0, bands, y = 0, time.perf_counter(), 0, 0
    while time.perf_counter() - t0 < budget_s:
        n = min(rows, h - y)
        canvas, seeds, _ = oracle.render(scene_spec, w, h, 1, seeds, canvas, crop=(0, y, w, n), threads=threads)
        paths += w * n
        bands += 1
        y = (y + n) % h
    dt = time.perf_counter() - t0
    return {"value": paths / dt / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (FP32, OpenMP, {threads} threads) on {paths} paths of the same frame: "
                      f"{bands} bands of {rows} rows x 1 spp, cycling over the frame, {dt:.1f} s",
            "host": host}


SUMMARY_MAX_BYTES = 2048
DETAIL_PATH = os.path.join(ROOT, "gpurun_out", "bench_detail.json")


def _sig(x, digits=5):
    """Floats rounded to `digits` significant digits (recursively)."""
    if isinstance(x, float):
        return float(f"{x:.{digits}g}")
    if isinstance(x, dict):
        return {k: _sig(v, digits) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_sig(v, digits) for v in x]
    return x


def _pick(d, keys):
    d = d or {}
    return {k: d[k] for k in keys if k in d}


def summary(out, detail_path=None):
    """The one-line result: the headline metric, its roofline, the CPU
    baseline, the time-to-RMSE figures and one headline number per leg. The
    tuner tables and RMSE curves stay in the detail file."""
    s = _pick(out, ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                    "scaling", "vs_baseline", "dtype", "data"))
    s["config"] = _pick(out.get("config"), ("workload", "width", "height", "spp", "hops", "primitives", "kernel"))
    s["roofline"] = _pick(out.get("roofline"), ("bound", "achieved", "peak", "unit", "frac", "frac_nonpacked",
                                                "kernel", "kernel_ms", "flops_per_path", "traffic",
                                                "algorithmic_bytes"))
    if "cpu_baseline" in out:
        cb = out["cpu_baseline"]
        s["cpu_baseline"] = _pick(cb, ("value", "unit", "cores", "kind", "sample"))
        if "sample" in s["cpu_baseline"]:
            s["cpu_baseline"]["sample"] = s["cpu_baseline"]["sample"][:200]
    if "time_to_rmse" in out:
        s["time_to_rmse"] = _pick(out["time_to_rmse"], ("target", "spp", "chosen_s", "plain_s", "speedup"))
    legs = {}
    for name in ("plain", "dropin_per_pass", "fast_fp"):
        if name in out:
            legs[name] = out[name].get("value")
    eu = out.get("euclidean_320x240")
    if eu:
        legs["euclidean_320x240"] = ({"chosen": eu["chosen"]["value"], "plain": eu["plain"]["value"]}
                                     if "chosen" in eu else {"error": str(eu.get("error"))[:120]})
    rc = out.get("radiance_cache_1080p")
    if rc:
        legs["radiance_cache_1080p"] = (
            {"bit_exact": _pick(rc.get("bit_exact_forms"), ("spp", "chosen_s", "plain_s", "speedup")),
             "cached": _pick(rc.get("cached"), ("reached", "spp", "seconds", "speedup_vs_chosen"))}
            if "cached" in rc else {"error": str(rc.get("error"))[:120]})
    bt = out.get("batch_32_jobs")
    if bt:
        legs["batch_32_jobs"] = (
            {"graph": (bt.get("graph_replay") or {}).get("value"), "eager": bt["eager"]["value"],
             "one_job_per_launch": bt["one_job_per_launch"]["value"],
             "bit_identical": bt.get("bit_identical_to_single_jobs"),
             "small_jobs_graph_speedup": (bt.get("small_jobs") or {}).get("graph_speedup")}
            if "eager" in bt else {"error": str(bt.get("error"))[:120]})
    wf = out.get("wavefront_cells")
    if wf:
        legs["wavefront_cells"] = ({"ratio_to_megakernel": wf["wavefront_over_megakernel"],
                                    "bit_identical": wf["bit_identical"]}
                                   if "wavefront" in wf else {"error": str(wf.get("error"))[:120]})
    sd = out.get("sdf_1080p")
    if sd:
        legs["sdf_1080p"] = sd["value"] if "value" in sd else {"error": str(sd.get("error"))[:120]}
    s["legs"] = legs
    if detail_path:
        s["detail"] = os.path.relpath(detail_path, ROOT)
    s = _sig(s)
    # the line must stay small: drop legs, then the CPU sample text, if ever needed
    for key in ("legs", "cpu_baseline"):
        if len(json.dumps(s)) <= SUMMARY_MAX_BYTES:
            break
        if key == "legs":
            s.pop("legs", None)
        elif "cpu_baseline" in s:
            s["cpu_baseline"].pop("sample", None)
    return s


def emit(out, detail_path=DETAIL_PATH, stream=None):
    """Writes the whole result to `detail_path` and prints the summary as
    the last line of stdout."""
    stream = stream or sys.stdout
    if detail_path:
        try:
            os.makedirs(os.path.dirname(detail_path), exist_ok=True)
            with open(detail_path, "w") as f:
                json.dump(out, f, indent=1)
        except OSError as e:
            print(f"bench: detail file not written: {e}", file=sys.stderr)
            detail_path = None
    stream.write(json.dumps(summary(out, detail_path)) + "\n")
    stream.flush()


def progress(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def gamma_image(canvas):
    """The image RMSE is taken on: extract, gamma 1/2.2, clamp to [0, 1]; RGB."""
    c = canvas.astype(np.float64)
    img = c[..., :3] / c[..., 3:4]
    return np.clip(np.power(np.maximum(img, 0.0), 1 / 2.2), 0.0, 1.0)


def rmse(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)))


def fp64_reference(wl, oracle_spp):
    """FP64 image of the workload at oracle_spp from two independent halves:
    (gamma image, its own noise RMSE, seconds). The halves differ by twice
    the variance of their mean."""
    from hypertrace_amd.processing import Canvas, Context, EntityBuffer, Render
    from hypertrace_amd.types import X32_F64

    W, H = wl.width, wl.height
    t_or = time.perf_counter()
    half = max(1, oracle_spp // 2)
    halves = []
    with Context(0, X32_F64) as ctx64:
        r64 = Render(ctx64, wl.scene)
        b64 = EntityBuffer(ctx64, wl.scene)
        for k in range(2):
            c = Canvas(ctx64, (W, H), 1, wl.seed ^ ORACLE_SEED_XOR ^ (k + 1))
            done = 0
            while done < half:
                n = min(256, half - done)
                r64.render_passes(b64, c, n)
                done += n
            halves.append(c.load_image())
    ref = gamma_image(halves[0] + halves[1])
    noise = rmse(gamma_image(halves[0]), gamma_image(halves[1])) / 2.0
    return ref, noise, time.perf_counter() - t_or, 2 * half


def time_to_rmse(ctx, wl, buf, chosen, plain_render, args, reference=None):
    W, H = wl.width, wl.height
    from hypertrace_amd.processing import Canvas

    ref, oracle_noise, oracle_s, oracle_spp = reference if reference is not None else fp64_reference(
        wl, args.oracle_spp)
    half = oracle_spp // 2

    # spp at which the FP32 image first reaches the target: steps of 16
    # passes, then pass by pass inside the last step.
    c = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
    spp, step, found = 0, 16, None
    curve = []
    while spp < 4096:
        chosen.render_passes(buf, c, step)
        spp += step
        e = rmse(gamma_image(c.load_image()), ref)
        curve.append((spp, round(e, 6)))
        if e <= args.target_rmse:
            found = spp
            break
    if found is None:
        return {"reached": False, "curve": curve, "oracle_spp": 2 * half, "oracle_noise_rmse": oracle_noise}
    c = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
    start = found - step
    if start > 0:
        chosen.render_passes(buf, c, start)
    spp = start
    while spp < found:
        chosen.render_passes(buf, c, 1)
        spp += 1
        if rmse(gamma_image(c.load_image()), ref) <= args.target_rmse:
            break
    need = spp

    c = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
    ctx.sync()
    t0 = time.perf_counter()
    chosen.render_passes(buf, c, need)
    ctx.sync()
    t_chosen = time.perf_counter() - t0
    res = {"reached": True, "target": args.target_rmse, "spp": need, "chosen_s": t_chosen,
           "final_rmse": rmse(gamma_image(c.load_image()), ref), "oracle_spp": 2 * half,
           "oracle_noise_rmse": oracle_noise, "oracle_render_s": oracle_s, "curve": curve,
           "rmse_definition": "all pixels x RGB of the extracted, gamma-1/2.2, clamped [0,1] image"}
    if plain_render is not None:
        c = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(need):
            plain_render.render(buf, c)
        t_plain = time.perf_counter() - t0
        res["plain_s"] = t_plain
        res["speedup"] = t_plain / t_chosen
    return res


def same_seed_fp64_error(wl, spp, image32, fp_options=None):
    """The FP32 image against the FP64 program's image of the same paths
    (same seeds, same spp; C4 images): precision error, with the branch flips
    it causes, next to the FP64 image's own Monte Carlo noise at that spp
    (from a second FP64 image on independent seeds)."""
    from hypertrace_amd.processing import Canvas, Context, EntityBuffer, Render
    from hypertrace_amd.types import X32_F64

    W, H = wl.width, wl.height
    with Context(0, X32_F64) as c64:
        r64 = Render(c64, wl.scene, fp_options=fp_options)
        b64 = EntityBuffer(c64, wl.scene)
        imgs = []
        for seed in (wl.seed, wl.seed ^ ORACLE_SEED_XOR):
            c = Canvas(c64, (W, H), wl.seed_mode, seed)
            r64.render_passes(b64, c, spp)
            imgs.append(gamma_image(c.load_image()))
    g32 = gamma_image(image32)
    return {"spp": spp, "rmse_fp32_vs_fp64_same_seeds": rmse(g32, imgs[0]),
            "max_abs": float(np.max(np.abs(g32 - imgs[0]))),
            "pixels_within_1e-3": float(np.mean(np.all(np.abs(g32 - imgs[0]) <= 1e-3, axis=-1))),
            "fp64_noise_rmse_at_spp": rmse(imgs[0], imgs[1]) / np.sqrt(2.0)}


def euclidean_leg(args):
    """The Euclidean sphere scene, 320x240, 16 passes of 1 spp, depth 4, 8
    spheres + 1 emitter, gradient sky (the reference's CPU-runnable case):
    one 16-pass frame of the chosen form (one render_passes launch) and of
    the plain kernel (one launch + sync per pass, as the reference's loop),
    time to RMSE 0.01 against the FP64 4096-spp image, and the FP32 image's
    error against the FP64 program on the same seeds."""
    from hypertrace_amd import scenes
    from hypertrace_amd.processing import Canvas, Context, EntityBuffer, Render
    from hypertrace_amd.types import X32_F32

    from hypertrace_amd.processing import program_options

    wl = scenes.euclidean_spheres()
    W, H, SPP = wl.width, wl.height, wl.spp
    reference = fp64_reference(wl, args.oracle_spp)
    with Context(0, X32_F32) as ctx:
        buf = EntityBuffer(ctx, wl.scene)
        chosen = Render(ctx, wl.scene)
        plain = Render(ctx, wl.scene, fp_options=NO_CULL)
        # the same forms with the FP32 hit window 1e-5 instead of the
        # reference's 1e-6 (the reference's window leaves FP32 images 0.01
        # RMSE from FP64 ones on this scene)
        chosen_w = Render(ctx, wl.scene, fp_options=program_options(eps32=1e-5))
        plain_w = Render(ctx, wl.scene, fp_options=program_options(eps32=1e-5, no_cull=True))
        c = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
        chosen.render_passes(buf, c, SPP)  # warm
        plain.render(buf, c)
        times = []
        for _ in range(3):
            c = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
            ctx.sync()
            t0 = time.perf_counter()
            chosen.render_passes(buf, c, SPP)
            ctx.sync()
            times.append(time.perf_counter() - t0)
        t_chosen = sorted(times)[1]
        image32 = c.load_image()
        c = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(SPP):
            plain.render(buf, c)
        ctx.sync()
        t_plain = time.perf_counter() - t0
        ttr = time_to_rmse(ctx, wl, buf, chosen, plain, args, reference)
        ttr_w = time_to_rmse(ctx, wl, buf, chosen_w, plain_w, args, reference)
        cw = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
        chosen_w.render_passes(buf, cw, SPP)
        image32_w = cw.load_image()
    paths = W * H * SPP
    keys = ("reached", "spp", "chosen_s", "plain_s", "speedup", "final_rmse", "oracle_noise_rmse")

    def summary(t):
        out = {k: t.get(k) for k in keys}
        if not t.get("reached"):
            out["rmse_at_max_spp"] = t["curve"][-1]  # (spp, RMSE): the FP32 image's distance floor
        return out

    return {"workload": "euclidean_spheres 320x240, 16 x 1 spp, depth 4, 8 spheres + 1 emitter (radiance 5), "
                        "gradient sky, strict FP32",
            "chosen": {"value": paths / t_chosen / 1e6, "unit": "Mpaths/s", "ms_per_frame": t_chosen * 1e3,
                       "kernel": "render_passes, culled scan, one launch per frame"},
            "plain": {"value": paths / t_plain / 1e6, "unit": "Mpaths/s", "ms_per_frame": t_plain * 1e3,
                      "kernel": "render, element-by-element scan, one launch + sync per pass"},
            "time_to_rmse": summary(ttr),
            "fp64_error": same_seed_fp64_error(wl, SPP, image32),
            "fp32_hit_window_1e-5": {
                "options": "-DHT_EPS32=1e-05f (not the reference's results: a wider self-intersection window)",
                "time_to_rmse": summary(ttr_w),
                "fp64_error": same_seed_fp64_error(wl, SPP, image32_w)}}


def cache_leg(args):
    """The radiance-cache method on the 1920x1080 hyperbolic cell scene with
    256 primitives: time until RMSE <= target against an FP64 4096-spp image,
    next to the chosen bit-exact form and the plain per-pass kernel. The
    cached passes are timed with their training (4 Adam steps of 16384
    suffixes per pass); image readbacks for the RMSE checks are not."""
    from hypertrace_amd import scenes
    from hypertrace_amd.maps import compose_on_device
    from hypertrace_amd.processing import Canvas, Context, EntityBuffer, Render
    from hypertrace_amd.radiance_cache import CachedRender, RadianceCache, init_params
    from hypertrace_amd.types import X32_F32

    wl = scenes.hyperbolic_cells(1920, 1080, 1, primitives=256)
    W, H = wl.width, wl.height
    reference = fp64_reference(wl, args.oracle_spp)
    ref = reference[0]
    with Context(0, X32_F32) as ctx:
        buf = EntityBuffer(ctx, wl.scene)
        compose_on_device(ctx, buf, wl.scene, wl.word_maps)
        from hypertrace_amd.tuning import tune_render
        # 32-spp frames: the timed launch runs 150+ passes, over which each
        # lane's work evens out, so very short frames overrate the pixel queue
        tuning = tune_render(ctx, wl.scene, W, H, 32, wl.seed_mode, wl.seed)
        chosen = Render(ctx, wl.scene, fp_options=tuning["best"]["options"])
        plain = None if args.no_plain else Render(ctx, wl.scene, fp_options=NO_CULL)
        base = time_to_rmse(ctx, wl, buf, chosen, plain, args, reference)
        cache = RadianceCache(ctx, W * H, 65536, 16384)
        # the cached kernel runs the same scan: build it with the tuner's best
        # options among those that apply to it (the pixel queue only changes
        # render_passes; the cached kernel runs one pass per launch)
        cached_knobs = min((t for t in tuning["table"] if not t["pixel_queue"]), key=lambda t: t["ms"])
        cr = CachedRender(ctx, wl.scene, cache, query_hop=2, train_every=16, steps_per_pass=4, batch=16384,
                          lr=1e-3, fp_options=cached_knobs["options"])
        canvas = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
        t, spp, curve, reached, check = 0.0, 0, [], None, 4
        while spp < args.cache_max_spp:
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(check):
                cr.render_pass(buf, canvas)
            ctx.sync()
            t += time.perf_counter() - t0
            spp += check
            e = rmse(gamma_image(canvas.load_image()), ref)
            curve.append((spp, round(e, 6)))
            if e <= args.target_rmse:
                reached = spp
                break
        queries, records = cache.counts()
        # the same method with the paths ending into the cache at their 2nd
        # and 4th hit: RMSE curves over a fixed number of passes, for the
        # bias / variance split of each
        hops = {}
        for qh in (1, 3):
            crh = CachedRender(ctx, wl.scene, cache, query_hop=qh, train_every=16, steps_per_pass=4,
                               batch=16384, lr=1e-3, fp_options=cached_knobs["options"])
            cache.set_params(init_params(0))  # a fresh network for each query hop
            ch = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
            th, sh, cv = 0.0, 0, []
            while sh < args.cache_split_spp:
                ctx.sync()
                t0 = time.perf_counter()
                for _ in range(8):
                    crh.render_pass(buf, ch)
                ctx.sync()
                th += time.perf_counter() - t0
                sh += 8
                cv.append((sh, round(rmse(gamma_image(ch.load_image()), ref), 6)))
            fit = bias_variance(cv)
            hops[f"query_hop_{qh}"] = {"ms_per_pass": th * 1e3 / sh, "curve": cv[::4] + cv[-1:], "fit": fit,
                                       "spp_for_target": spp_for(fit, args.target_rmse)}
    chosen_fit = bias_variance(base.get("curve") or [])
    cached_fit = bias_variance(curve)
    split = {"chosen": {"fit": chosen_fit, "spp_for_target": spp_for(chosen_fit, args.target_rmse),
                        "ms_per_pass": (base["chosen_s"] * 1e3 / base["spp"]) if base.get("spp") else None},
             "query_hop_2": {"fit": cached_fit, "spp_for_target": spp_for(cached_fit, args.target_rmse),
                             "ms_per_pass": t * 1e3 / max(spp, 1)}}
    split.update(hops)
    out = {"workload": "hyperbolic_cells 1920x1080, 256 primitives, 8 hops", "target": args.target_rmse,
           "oracle_spp": reference[3], "oracle_noise_rmse": reference[1], "oracle_render_s": reference[2],
           "bit_exact_forms": {k: base.get(k) for k in ("reached", "spp", "chosen_s", "plain_s", "speedup")},
           "chosen_program": tuning["best"], "cached_program": cached_knobs, "render_tuning": tuning,
           "cached": {"reached": reached is not None, "spp": reached if reached is not None else spp,
                      "seconds": t, "ms_per_pass": t * 1e3 / max(spp, 1), "curve": curve[::4] + curve[-1:],
                      "queries_last_pass": queries, "training_records_last_pass": records,
                      "method": "paths end into the MLP cache at their third hit; 1/16 of pixels per pass "
                                "run unbiased to full depth and train it (4 Adam steps x 16384 per pass)"},
           "bias_variance": split}
    if reached is not None and base.get("plain_s"):
        out["cached"]["speedup_vs_plain"] = base["plain_s"] / t
    if reached is not None and base.get("chosen_s"):
        out["cached"]["speedup_vs_chosen"] = base["chosen_s"] / t
    return out


def bias_variance(curve, min_spp=8):
    """Least-squares fit of RMSE(spp)^2 = bias^2 + var / spp to the points
    of an RMSE curve with spp >= min_spp: the error floor no number of
    samples removes (the method's bias plus the reference image's own
    noise) and the per-sample variance. Returns the two terms and the spp
    the fit needs for a target (None when the floor is above it)."""
    pts = [(float(n), float(e)) for n, e in curve if n >= min_spp]
    if len(pts) < 2:
        return None
    a = np.array([[1.0, 1.0 / n] for n, _ in pts])
    y = np.array([e * e for _, e in pts])
    (b2, v), *_ = np.linalg.lstsq(a, y, rcond=None)
    b2, v = max(float(b2), 0.0), max(float(v), 0.0)
    return {"floor_rmse": b2 ** 0.5, "var_per_spp": v, "points": len(pts)}


def spp_for(fit, target):
    if fit is None or fit["floor_rmse"] >= target:
        return None
    return fit["var_per_spp"] / (target * target - fit["floor_rmse"] ** 2)


def fixed_knobs(text, base=None):
    """A render-tuner result for the setting "UNROLL,WAVES[,sb][,gN][,q][,p][,gq][,ho][,hl][,sN][,spN][,pc]" (no timing)."""
    from hypertrace_amd.processing import STRICT_FP
    from hypertrace_amd.tuning import options
    knobs = text.split(",")
    unroll, waves = int(knobs[0]), int(knobs[1])
    flags = knobs[2:]
    sb = "sb" in flags
    groups = next((int(k[1:]) for k in flags if k.startswith("g") and k[1:].isdigit()), 0)
    queue = "q" in flags
    prio = "p" in flags
    gq = "gq" in flags
    ho = "ho" in flags
    hl = "hl" in flags
    sparse = next((int(k[1:]) for k in flags if k.startswith("s") and k[1:].isdigit()), 0)
    spread = next((int(k[2:]) for k in flags if k.startswith("sp") and k[2:].isdigit()), 0)
    classes = "pc" in flags
    return {"best": {"unroll": unroll, "waves_per_eu": waves, "scalar_bounds": sb, "groups": groups,
                     "pixel_queue": queue, "wave_prio": prio, "group_queue": gq, "hit_order": ho, "helpers": hl,
                     "sparse": sparse, "spread": spread, "classes": classes,
                     "options": options(unroll, waves, base or STRICT_FP, scalar_bounds=sb, groups=groups,
                                        pixel_queue=queue, wave_prio=prio, group_queue=gq, hit_order=ho,
                                        helpers=hl, sparse=sparse, spread=spread, classes=classes)},
            "table": [], "fixed": True}


def sdf_leg(args):
    """The sphere-traced scene at 1920x1080: 1024 SDF shapes, 256 spp, depth
    16, Russian roulette from the 3rd hit, strict FP32: Mpaths/s of one frame
    (one render_passes launch, inputs resident), and the FP64 reference image
    at the same spp from independent seeds: RMSE between them next to the
    FP64 image's own noise (its two halves)."""
    from hypertrace_amd import scenes
    from hypertrace_amd.processing import Canvas, Context, EntityBuffer, Render, program_options
    from hypertrace_amd.types import X32_F32, X32_F64

    wl = scenes.sdf_shapes()
    W, H, SPP = wl.width, wl.height, args.sdf_spp
    opts = program_options(rr_depth=3)
    with Context(0, X32_F32) as ctx:
        # the render tuner over 4-spp frames (all settings render the same images)
        from hypertrace_amd.tuning import tune_render
        if args.sdf_knobs:
            tuning = fixed_knobs(args.sdf_knobs, opts)
        else:
            tuning = tune_render(ctx, wl.scene, W, H, 4, wl.seed_mode, wl.seed, base=opts)
        r = Render(ctx, wl.scene, fp_options=tuning["best"]["options"])
        b = EntityBuffer(ctx, wl.scene)
        c = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
        r.render_passes(b, c, 1)
        c = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.timer_begin()
        r.render_passes(b, c, SPP)
        gpu_ms = ctx.timer_end()
        wall = time.perf_counter() - t0
        img32 = gamma_image(c.load_image())
        c16 = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
        r.render_passes(b, c16, 16)
        image16 = c16.load_image()
    same_seeds = same_seed_fp64_error(wl, 16, image16, fp_options=opts)
    halves = []
    t64 = time.perf_counter()
    with Context(0, X32_F64) as c64:
        r64 = Render(c64, wl.scene, fp_options=opts)
        b64 = EntityBuffer(c64, wl.scene)
        for k in range(2):
            cc = Canvas(c64, (W, H), 1, wl.seed ^ ORACLE_SEED_XOR ^ (k + 1))
            r64.render_passes(b64, cc, SPP // 2)
            halves.append(cc.load_image())
    t64 = time.perf_counter() - t64
    ref = gamma_image(halves[0] + halves[1])
    return {"workload": f"sdf_shapes 1920x1080, 1024 shapes (sphere:box:torus 1:1:1), {SPP} spp, depth 16, "
                        "Russian roulette from the 3rd hit", "dtype": "f32",
            "value": W * H * SPP / wall / 1e6, "unit": "Mpaths/s", "ms_per_frame": wall * 1e3,
            "gpu_ms_per_frame": gpu_ms, "render_tuning": tuning,
            "fp64_reference": {"spp": 2 * (SPP // 2), "render_s": t64,
                               "rmse_fp32_vs_fp64": rmse(img32, ref),
                               "fp64_noise_rmse": rmse(gamma_image(halves[0]), gamma_image(halves[1])) / 2.0,
                               "note": "independent seeds: the RMSE is MC noise of both images plus any FP32 bias"},
            "fp64_error": same_seeds}


def wavefront_leg(args, options=None):
    """The cell frame (640x480, 64 spp, 192 primitives) through the wavefront
    form (ht_render_wavefront: SoA path state in HBM, extend / material-sorted
    shade kernels) next to the megakernel's render_passes on the same
    program options (the main frame's tuned ones): time per frame, and
    whether canvas and seeds agree bit for bit. The wavefront call is
    synchronous (it reads the live-path count back every few iterations), so
    its time is wall-clock."""
    import numpy as np

    from hypertrace_amd import scenes
    from hypertrace_amd.processing import Canvas, Context, EntityBuffer, Render
    from hypertrace_amd.types import X32_F32

    wl = scenes.hyperbolic_cells()
    W, H, SPP = wl.width, wl.height, wl.spp
    with Context(0, X32_F32) as ctx:
        buf = EntityBuffer(ctx, wl.scene)
        mega = Render(ctx, wl.scene, fp_options=options)
        wave = Render(ctx, wl.scene, fp_options=options, wavefront=True)
        results = {}
        for name, run in (("megakernel", lambda c: mega.render_passes(buf, c, SPP)),
                          ("wavefront", lambda c: wave.render_wavefront(buf, c, SPP))):
            c = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
            run(c)  # warm-up, and the frame compared below
            ctx.sync()
            img, seeds = c.load_image().copy(), c.load_seeds().copy()
            times = []
            for _ in range(args.wavefront_steps):
                c2 = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
                ctx.sync()
                t0 = time.perf_counter()
                run(c2)
                ctx.sync()
                times.append(time.perf_counter() - t0)
            results[name] = {"ms_per_frame": min(times) * 1e3, "value": W * H * SPP / min(times) / 1e6,
                             "unit": "Mpaths/s", "img": img, "seeds": seeds}
        same = bool(np.array_equal(results["megakernel"]["img"].view(np.uint32),
                                   results["wavefront"]["img"].view(np.uint32)) and
                    np.array_equal(results["megakernel"]["seeds"], results["wavefront"]["seeds"]))
    for r in results.values():
        del r["img"], r["seeds"]
    return {"workload": f"hyperbolic_cells {W}x{H}, {SPP} spp, 192 primitives (strict FP32, culled scan)",
            "options": options, "megakernel": results["megakernel"], "wavefront": results["wavefront"],
            "bit_identical": same,
            "wavefront_over_megakernel": results["wavefront"]["ms_per_frame"] / results["megakernel"]["ms_per_frame"]}


def batch_leg(args):
    """32 uneven jobs (half Euclidean, half hyperbolic; 128x128..1024x768
    log-uniform, spp U{1..64}, depth U{4..12}, 8..256 primitives of the
    diffuse / mirror / glass 1.5 mix) in one persistent launch per frame:
    the autotuner's table (block size x tile size x grid), then frames of the
    best candidate eager and as hipGraph replays (canvas clears, launch and
    the post pass of every job), next to the same kernel run one job per
    launch."""
    from hypertrace_amd import scenes
    from hypertrace_amd.batch import Batch, BatchProgram, autotune
    from hypertrace_amd.processing import Context
    from hypertrace_amd.types import X32_F32

    specs = scenes.batch_jobs()
    paths = sum(s.width * s.height * s.spp for s in specs)
    with Context(0, X32_F32) as ctx:
        if args.batch_program:
            # a fixed program (NAME,BLOCK,TILE,GRID; NAME a program option set of batch.py), no tuning
            from hypertrace_amd import batch as batch_mod
            name, block, tile, grid_ = args.batch_program.split(",")
            best = {"block": int(block), "tile": int(tile), "grid": int(grid_), "options": getattr(batch_mod, name)}
            tune = {"best": best, "table": [dict(best, ms=0.0)], "fixed": True}
        else:
            tune = autotune(ctx, specs, reps=1, progress=lambda m: print(m, file=sys.stderr, flush=True))
            best = tune["best"]
        bprog = BatchProgram(ctx, best["block"], fp_options=best.get("options"))
        batch = Batch(ctx, bprog, specs, best["tile"], with_rgba8=True)
        grid = best["grid"]

        def timed(fn, n=3):
            fn()
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(n):
                fn()
            ctx.sync()
            return (time.perf_counter() - t0) / n

        eager_s = timed(lambda: batch.render(grid, clear=True, post=True))
        graph_s = None
        if not args.no_graph:
            batch.capture(grid, clear=True, post=True)
            graph_s = timed(batch.replay)
        # one job per launch with the best block-tiled candidate (a wave-queue
        # program's per-wave tiles leave most of the device idle on one job)
        solo = min((r for r in tune["table"] if "HT_BATCH_WAVE_QUEUE" not in (r.get("options") or "")),
                   key=lambda r: r["ms"], default=None)
        if solo is None:  # a fixed wave-queue program: block tiles of the same scan for the solo launches
            from hypertrace_amd.batch import BLOCK_HIT_ORDER
            solo = {"block": 128, "tile": 128, "grid": 2048, "options": BLOCK_HIT_ORDER}
        sprog = BatchProgram(ctx, solo["block"], fp_options=solo.get("options"))
        singles = [Batch(ctx, sprog, [s], solo["tile"]) for s in specs]

        def one_by_one():
            for b in singles:
                b.render(solo["grid"], clear=True)

        single_s = timed(one_by_one, 2)
        # one frame of the tuned batch program and of every job alone, both
        # from fresh seeds: canvases and seeds must agree bit for bit
        for b in singles:
            b.close()
        fresh = Batch(ctx, bprog, specs, best["tile"])
        fresh.render(grid, clear=True)
        same = True
        for s, cb in zip(specs, fresh.canvases):
            one = Batch(ctx, sprog, [s], solo["tile"])
            one.render(solo["grid"], clear=True)
            c1 = one.canvases[0]
            same = same and bool(np.array_equal(cb.load_image().view(np.uint32), c1.load_image().view(np.uint32))
                                 and np.array_equal(cb.load_seeds(), c1.load_seeds()))
            one.close()
        fresh.close()
        # many tiny jobs (256 thumbnails of 16x16..32x32, 1 spp, 8..16
        # primitives) with the same program: the frame's 2 x 256 + 3 launches
        # (a canvas clear and a post pass per job, the counter reset, the
        # bounds fill, the persistent kernel) take the host longer to enqueue
        # than the device to run, which is where one hipGraph replay pays
        small = scenes.batch_jobs(256, (16, 16), (32, 32), (1, 1), (4, 6), (8, 16))
        small_paths = sum(s.width * s.height * s.spp for s in small)
        sbatch = Batch(ctx, bprog, small, best["tile"], with_rgba8=True)
        small_eager_s = timed(lambda: sbatch.render(grid, clear=True, post=True), 20)
        small_graph_s = None
        if not args.no_graph:
            sbatch.capture(grid, clear=True, post=True)
            small_graph_s = timed(sbatch.replay, 20)
        sbatch.close()
    small_out = {"workload": "256 jobs, 16x16..32x32, 1 spp, depth 4..6, 8..16 primitives, post pass per job",
                 "paths_per_frame": small_paths,
                 "eager": {"ms_per_frame": small_eager_s * 1e3, "value": small_paths / small_eager_s / 1e6,
                           "unit": "Mpaths/s"},
                 "graph_replay": None if small_graph_s is None else
                 {"ms_per_frame": small_graph_s * 1e3, "value": small_paths / small_graph_s / 1e6, "unit": "Mpaths/s"},
                 "graph_speedup": None if small_graph_s is None else small_eager_s / small_graph_s}
    return {"workload": "32 jobs, half Euclidean / half hyperbolic, 128x128..1024x768, spp 1..64, depth 4..12, "
                        "8..256 primitives, diffuse/mirror/glass 0.6/0.25/0.15",
            "paths_per_frame": paths, "autotuned": best,
            "graph_replay": None if graph_s is None else
            {"ms_per_frame": graph_s * 1e3, "value": paths / graph_s / 1e6, "unit": "Mpaths/s"},
            "eager": {"ms_per_frame": eager_s * 1e3, "value": paths / eager_s / 1e6, "unit": "Mpaths/s"},
            "one_job_per_launch": {"ms_per_frame": single_s * 1e3, "value": paths / single_s / 1e6,
                                   "unit": "Mpaths/s", "program": solo},
            "bit_identical_to_single_jobs": same,
            "small_jobs": small_out,
            "tuner_table": tune["table"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="time only the chosen form")
    ap.add_argument("--no-cache", action="store_true", help="skip the 1080p radiance-cache leg")
    ap.add_argument("--cache-max-spp", type=int, default=1024)
    ap.add_argument("--cache-split-spp", type=int, default=256,
                    help="passes of the query-hop 1 and 3 runs of the radiance-cache leg")
    ap.add_argument("--no-sdf", action="store_true", help="skip the 1080p SDF leg")
    ap.add_argument("--no-euclidean", action="store_true", help="skip the 320x240 Euclidean leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 32-job batched leg")
    ap.add_argument("--no-wavefront", action="store_true", help="skip the wavefront-form leg")
    ap.add_argument("--batch-program", default=None, metavar="NAME,BLOCK,TILE,GRID",
                    help="batch leg: this program (NAME = an option set of hypertrace_amd/batch.py, e.g. "
                         "WAVE_QUEUE_HIT_ORDER) instead of the autotuner's pick")
    ap.add_argument("--no-graph", action="store_true",
                    help="batch leg: eager frames only, no hipGraph capture (for profilers that do not run graphs)")
    ap.add_argument("--wavefront-steps", type=int, default=2)
    ap.add_argument("--sdf-spp", type=int, default=256)
    ap.add_argument("--no-rmse", action="store_true", help="skip the time-to-RMSE measurement")
    ap.add_argument("--oracle-spp", type=int, default=4096)
    ap.add_argument("--target-rmse", type=float, default=0.01)
    ap.add_argument("--render-knobs", default=None, metavar="UNROLL,WAVES[,sb][,gN][,q][,p][,gq][,ho][,hl][,sN][,spN][,pc]",
                    help="use this cull unroll / waves-per-SIMD (/ scalar-bounds) setting for the main frame "
                         "instead of tuning (keeps a profiled trace to one program)")
    ap.add_argument("--sdf-knobs", default=None, metavar="UNROLL,WAVES[,sb][,gN][,q][,p][,gq][,ho][,hl][,sN][,spN][,pc]",
                    help="the same for the SDF leg's program")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run the extract kernel over 4096x4096 float4 (256 MiB read + 256 MiB written): "
                         "the known-size copy that scales FETCH_SIZE / WRITE_SIZE counter readings")
    ap.add_argument("--detail", default=DETAIL_PATH,
                    help="file for the whole result (tuner tables, RMSE curves); the last stdout line is a summary")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the render pass would shard by "
              "pixel rows, one row band per GPU, with per-pixel results unchanged)", file=sys.stderr)
        sys.exit(2)

    from hypertrace_amd import scenes
    from hypertrace_amd.processing import FAST_FP, Canvas, Context, EntityBuffer, Render, seed_fill
    from hypertrace_amd.types import X32_F32

    wl = scenes.hyperbolic_cells()
    W, H, SPP = wl.width, wl.height, wl.spp
    paths_per_step = W * H * SPP
    spec = wl.scene.spec_text(X32_F32)
    seeds0 = seed_fill(W * H, wl.seed_mode, wl.seed)
    extra = {}

    with Context(0, X32_F32) as ctx:
        buf = EntityBuffer(ctx, wl.scene)
        # The 192 cell maps are composed on the GPU in FP64 (det = 1
        # reprojection after every product) straight into the scene buffer;
        # they must equal the host composition bit for bit.
        from hypertrace_amd.maps import compose_on_device
        compose_on_device(ctx, buf, wl.scene, wl.word_maps)
        maps_ok = buf.read().tobytes() == wl.scene.to_bytes(X32_F32)
        if not maps_ok:
            raise RuntimeError("GPU-composed scene maps differ from the host composition")

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
            return time.perf_counter() - t0, gpu_ms

        # chosen program: strict IEEE FP32, culled scan, with the unroll /
        # waves-per-SIMD setting the tuner picks for this frame (all settings
        # render bit-identical images)
        from hypertrace_amd.tuning import tune_render
        if args.render_knobs:
            tuning = fixed_knobs(args.render_knobs)
        else:
            tuning = tune_render(ctx, wl.scene, W, H, SPP, wl.seed_mode, wl.seed)
        extra["render_tuning"] = tuning
        chosen = Render(ctx, wl.scene, fp_options=tuning["best"]["options"])
        wall, gpu_ms = time_chosen(chosen, args.steps, args.warmup)
        ms_per_step = wall * 1e3 / args.steps
        kernel_ms = gpu_ms / args.steps
        mpaths = paths_per_step * args.steps / wall / 1e6

        if args.calibrate:
            from hypertrace_amd.processing import Extractor, ImageBuffer
            big = Canvas(ctx, (4096, 4096))
            big.passes = 1
            Extractor(ctx).process(big, ImageBuffer(ctx, (4096, 4096)))
            ctx.sync()

        plain_render = None
        if not args.no_plain:
            plain_render = Render(ctx, wl.scene, fp_options=NO_CULL)
            canvas_p = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
            plain_render.render(buf, canvas_p)  # warm
            ctx.sync()
            t0 = time.perf_counter()
            ctx.timer_begin()
            for _ in range(SPP):
                plain_render.render(buf, canvas_p)
            plain_gpu_ms = ctx.timer_end()
            plain_wall = time.perf_counter() - t0
            extra["plain"] = {"kernel": "render: one launch + sync per pass, element-by-element scan, strict FP32",
                              "value": W * H * SPP / plain_wall / 1e6, "unit": "Mpaths/s",
                              "ms_per_frame": plain_wall * 1e3, "gpu_ms_per_frame": plain_gpu_ms}
            # the reference's loop unchanged (one synchronous render per pass)
            # over the chosen program: what a drop-in caller gets
            canvas_d = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
            chosen.render(buf, canvas_d)  # warm
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(SPP):
                chosen.render(buf, canvas_d)
            dropin_wall = time.perf_counter() - t0
            extra["dropin_per_pass"] = {"kernel": "render: one launch + sync per pass on the chosen program "
                                                  "(the reference's loop, culled scan)",
                                        "value": W * H * SPP / dropin_wall / 1e6, "unit": "Mpaths/s",
                                        "ms_per_frame": dropin_wall * 1e3}
            fast = Render(ctx, wl.scene, fp_options=FAST_FP)
            fwall, fgpu = time_chosen(fast, args.steps, args.warmup)
            extra["fast_fp"] = {"kernel": "render_passes, culled scan, fast FP32 (FMA contraction, reciprocal "
                                          "division, device libm; not bit-identical)",
                                "value": paths_per_step * args.steps / fwall / 1e6, "unit": "Mpaths/s",
                                "kernel_ms": fgpu / args.steps}

        if not args.no_rmse:
            extra["time_to_rmse"] = time_to_rmse(ctx, wl, buf, chosen, plain_render, args)
            frame = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
            chosen.render_passes(buf, frame, SPP)
            extra["fp64_error"] = same_seed_fp64_error(wl, SPP, frame.load_image())

    fl = count_flops_per_path(spec, W, H, seeds0, mode=count_mode(tuning["best"]))
    flops_per_path = fl["fp32"] + fl["fp64"]
    achieved = flops_per_path * paths_per_step / (kernel_ms * 1e-3) / 1e12
    roofline = {"bound": "valu", "achieved": achieved, "peak": FP32_VECTOR_PEAK, "unit": "TFLOP/s",
                "frac": achieved / FP32_VECTOR_PEAK, "frac_nonpacked": achieved / FP32_NONPACKED_PEAK,
                "kernel": "render_passes", "kernel_ms": kernel_ms,
                "flops_per_path": flops_per_path, "fp32_flops_per_path": fl["fp32"],
                "fp64_flops_per_path": fl["fp64"], "flops_sample_paths": fl["paths"],
                "flops_note": "algorithmic flops of the algorithm the kernel runs (FP32 exact tests of the "
                              "candidates, FP64 cull tests of the chosen scan; FP64 and non-packed FP32 issue at the "
                              "same rate), counted by the CPU oracle on the same paths",
                "flops_count_mode": count_mode(tuning["best"]),
                "plain_flops_per_path": fl["plain_fp32"],
                "plain_equivalent_tflops": fl["plain_fp32"] * paths_per_step / (kernel_ms * 1e-3) / 1e12,
                "traffic": MEASURED_TRAFFIC_BYTES.get(tuning["best"]["options"]),
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
                   "primitives": 192, "kernel": "render_passes",
                   "maps": "192 Moebius words composed on the GPU in FP64 with det=1 reprojection, bit-identical "
                           "to the host composition"},
        "roofline": roofline,
    }
    out.update(extra)
    progress(f"headline {mpaths:.1f} Mpaths/s, {ms_per_step:.2f} ms per frame, roofline frac {roofline['frac']:.4f}")
    # The legs after the headline; a failing leg is reported, the headline stays valid.
    legs = [("euclidean_320x240", not args.no_euclidean and not args.no_rmse, lambda: euclidean_leg(args)),
            ("radiance_cache_1080p", not args.no_cache and not args.no_rmse, lambda: cache_leg(args)),
            ("batch_32_jobs", not args.no_batch, lambda: batch_leg(args)),
            ("wavefront_cells", not args.no_wavefront,
             lambda: wavefront_leg(args, out["render_tuning"]["best"]["options"])),
            ("sdf_1080p", not args.no_sdf, lambda: sdf_leg(args))]
    for name, wanted, run in legs:
        if not wanted:
            continue
        t0 = time.perf_counter()
        try:
            out[name] = run()
        except Exception as e:
            out[name] = {"error": f"{type(e).__name__}: {e}"}
        progress(f"{name} done in {time.perf_counter() - t0:.1f} s")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(spec, W, H, seeds0)
        progress(f"cpu baseline {out['cpu_baseline']['value']:.3f} Mpaths/s on {out['cpu_baseline']['cores']} threads")
    emit(out, args.detail)


if __name__ == "__main__":
    main()
