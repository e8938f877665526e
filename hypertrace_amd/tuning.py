"""Autotuner of the render program's code-generation knobs for a workload.

Four compile-time knobs of the scene program: HT_CULL_UNROLL (cull tests
per step of the nearest-hit scan) and HT_WAVES_PER_EU (minimum waves per
SIMD the register allocation must allow; above 4 the compiler spills a
little) trade instruction-level parallelism against registers and so
against resident waves; HT_SCALAR_BOUNDS moves the element bounds from LDS
(FP32, widened per lane) to a global FP64 array read through the scalar
cache, which frees the LDS that limits occupancy on scenes with many
elements; HT_CULL_GROUPS scans groups of Morton-ordered elements first and
skips a group's elements when no lane of the wave passes its sphere (with
scalar bounds the bounds_fill kernel writes the tree's leaves and group
spheres to the global array, so the render kernels hold no tree in LDS);
HT_PIXEL_QUEUE launches only as many blocks as the device holds at once and
has each lane take its next pixel from a global counter when its pixel's
samples are done, which removes the tail of waves whose slowest pixel holds
the other 63 lanes (frames with more pixels than resident lanes gain; a
frame with fewer than twice as many pixels as resident lanes loses);
HT_WAVE_PRIO steps each wave's issue priority down as its slowest lane
completes each quarter of the passes, so the waves with the most work left
(the frame's critical path) issue first (with the pixel queue: once the
queue is empty, by the work each wave's current pixels have left);
HT_GROUP_QUEUE compacts the two-level scan at both levels: the (group,
ray) pairs that pass a group sphere are queued and tested on full lanes
against the group's leaves, so each ray pays for the groups it passes
rather than for the groups any ray of its wave passes; HT_HIT_ORDER (with
the group queue) gives each passing pair the interval of hyperbolic
distance at which its geodesic crosses the element's bound spheres, drops
the pairs whose geodesic never enters the sphere ahead of the start, and
skips those that lie beyond the ray's nearest certain hit, so most rays
run one or two exact tests instead of five; HT_HELPERS keeps the lanes
whose pixels are done (with the pixel queue: the lanes without a pixel
once the queue is empty) in the loop as helpers that take other rays' pair
tests, so a wave's last rays are not tested on nearly empty lanes;
HT_SPARSE_L1 = R (with the group queue) runs the group tests of a scan
with at most R rays over its (group, ray) pairs, one pair per lane, instead
of every group in every lane; HT_PIXEL_SPREAD hands the pixels to the
lanes (or out of the pixel queue) in an order that spreads each wave's 64
pixels over the frame (1: the golden-ratio order t * a mod n; B = 2..64:
one run of 64 / B pixels from each of B bands), so that a cluster of
expensive pixels is spread over many waves whose other lanes finish early
and help (with the pixel queue and the golden-ratio order the 640x480 cell
frame takes 16.8 instead of 20.6 ms); HT_PIXEL_CLASSES (with the queue and
the golden-ratio order) probes one camera ray per pixel before the frame
and hands out the pixels whose probe hits the scene first, the sky pixels
after them (3 % on the 640x480 cell frame). Every
setting renders bit-identical images (only the schedule differs), so the
tuner simply times one frame of the caller's workload per candidate on the
GPU and keeps the fastest. Small frames (few waves per SIMD) prefer fewer
registers less than large ones.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Optional, Sequence, Tuple

from .processing import STRICT_FP, Canvas, Context, EntityBuffer, Render

# (cull unroll, waves per SIMD, scalar bounds, group size or 0, pixel queue, wave priority, group queue)
CANDIDATES: Tuple[Tuple, ...] = (
    (4, 4, False, 0, False), (4, 5, False, 0, False), (2, 5, False, 0, False), (4, 6, False, 0, False),
    (4, 4, True, 0, False), (4, 5, True, 0, False), (4, 6, True, 0, False), (4, 5, False, 4, False),
    (4, 4, False, 0, True), (4, 4, True, 0, True), (4, 6, True, 0, True),
    (4, 4, True, 4, False), (4, 6, True, 4, False), (4, 6, True, 8, False), (4, 6, True, 4, True),
    (4, 4, True, 4, True),
    (4, 4, True, 0, False, True), (4, 5, True, 0, False, True), (4, 6, True, 4, False, True),
    (4, 4, True, 0, True, True), (4, 4, True, 4, True, True),
    (4, 4, True, 4, False, True, True), (4, 4, True, 8, False, True, True), (4, 6, True, 4, False, True, True),
    (4, 4, True, 4, True, True, True), (4, 6, True, 4, True, True, True),
    (4, 3, True, 4, False, True, True), (4, 3, True, 4, True, True, True),
    (4, 4, True, 4, False, True, True, True), (4, 3, True, 4, False, True, True, True),
    (4, 4, True, 4, True, True, True, True), (4, 3, True, 4, True, True, True, True),
    (4, 4, True, 4, False, True, True, True, True), (4, 3, True, 4, False, True, True, True, True),
    (4, 4, True, 4, False, True, True, False, True), (4, 4, True, 0, False, True, False, False, True),
    (4, 4, True, 4, False, True, True, True, True, 32), (4, 4, True, 4, False, True, True, True, True, 16),
    (4, 3, True, 4, False, True, True, True, True, 32),
    (4, 4, True, 4, True, True, True, True, True), (4, 4, True, 4, True, True, True, True, True, 16),
    (4, 4, True, 4, True, True, True, False, True), (4, 4, True, 4, True, True, True, False, True, 16),
    (4, 4, True, 4, True, True, True, True, True, 32, 1), (4, 4, True, 4, True, True, True, True, True, 16, 1),
    (4, 3, True, 4, True, True, True, True, True, 32, 1), (4, 4, True, 4, True, True, True, True, True, 32, 64),
    (4, 4, True, 4, True, True, True, True, True, 32, 1, True), (4, 4, True, 4, True, True, True, True, True, 16, 1, True),
    (4, 3, True, 4, True, True, True, True, True, 32, 1, True),
    (4, 4, True, 4, True, True, True, False, True, 16, 1), (4, 4, True, 4, True, True, True, False, True, 16, 1, True))


def options(unroll: int, waves: int, base: str = STRICT_FP, scalar_bounds: bool = False, groups: int = 0,
            pixel_queue: bool = False, wave_prio: bool = False, group_queue: bool = False,
            hit_order: bool = False, helpers: bool = False, sparse: int = 0, spread: int = 0,
            classes: bool = False) -> str:
    opts = f"{base} -DHT_CULL_UNROLL={int(unroll)} -DHT_WAVES_PER_EU={int(waves)}"
    if scalar_bounds:
        opts += " -DHT_SCALAR_BOUNDS=1"
    if groups:
        opts += f" -DHT_CULL_GROUPS={int(groups)}"
    if pixel_queue:
        opts += " -DHT_PIXEL_QUEUE=1"
    if wave_prio:
        opts += " -DHT_WAVE_PRIO=1"
    if group_queue:
        opts += " -DHT_GROUP_QUEUE=1"
    if hit_order:
        opts += " -DHT_HIT_ORDER=1"
    if helpers:
        opts += " -DHT_HELPERS=1"
    if sparse:
        opts += f" -DHT_SPARSE_L1={int(sparse)}"
    if spread:
        opts += f" -DHT_PIXEL_SPREAD={int(spread)}"
    if classes:
        opts += " -DHT_PIXEL_CLASSES=1"
    return opts


def _knobs(candidate: Sequence) -> Tuple[int, int, bool, int, bool, bool, bool, bool, bool, int, int, bool]:
    unroll, waves = int(candidate[0]), int(candidate[1])
    sb = bool(candidate[2]) if len(candidate) > 2 else False
    groups = int(candidate[3]) if len(candidate) > 3 else 0
    queue = bool(candidate[4]) if len(candidate) > 4 else False
    prio = bool(candidate[5]) if len(candidate) > 5 else False
    gq = bool(candidate[6]) if len(candidate) > 6 else False
    ho = bool(candidate[7]) if len(candidate) > 7 else False
    hl = bool(candidate[8]) if len(candidate) > 8 else False
    sparse = int(candidate[9]) if len(candidate) > 9 else 0
    spread = int(candidate[10]) if len(candidate) > 10 else 0
    classes = bool(candidate[11]) if len(candidate) > 11 else False
    return unroll, waves, sb, groups, queue, prio, gq, ho, hl, sparse, spread, classes


def candidate_options(candidate: Sequence, base: str = STRICT_FP) -> str:
    """Program options of one tuner candidate tuple."""
    return options(*_knobs(candidate)[:2], base, *_knobs(candidate)[2:])


# A candidate that spills registers is picked only when it beats the best
# spill-free candidate by more than this fraction of that candidate's time.
SPILL_TOLERANCE = 0.015


def pick(table: Sequence[Dict], tolerance: float = SPILL_TOLERANCE) -> Dict:
    """The fastest candidate whose render_passes kernel spills no registers
    to scratch (scratch bytes 0), unless a spilling one is faster by more
    than `tolerance` (then the spills are out of its hot loop, or cheaper
    than the occupancy a spill-free register budget costs); the fastest of
    all when every candidate spills."""
    fastest = min(table, key=lambda t: t["ms"])
    clean = [t for t in table if not t.get("scratch")]
    if clean:
        best_clean = min(clean, key=lambda t: t["ms"])
        if fastest["ms"] * (1.0 + tolerance) >= best_clean["ms"]:
            return best_clean
    return fastest


def tune_render(ctx: Context, scene, width: int, height: int, spp: int, seed_mode: int = 1, seed: int = 0,
                candidates: Sequence[Sequence] = CANDIDATES, reps: int = 3, base: str = STRICT_FP,
                cache_path: Optional[str] = None, spill_tolerance: float = SPILL_TOLERANCE) -> Dict:
    """Times `spp` passes of render_passes (HIP events, median of `reps`) for
    each (unroll, waves[, scalar bounds[, groups[, pixel queue[, wave priority[, group queue]]]]]) candidate,
    with the kernel's registers and scratch bytes per lane; returns {"best":
    {...}, "table": [...], "fastest_any": {...}}: the best is the fastest
    candidate that spills no registers unless a spilling one is faster by
    more than `spill_tolerance` (see `pick`)."""
    key = f"gfx950:{scene.cpp()}:{width}x{height}:{spp}:{base}:spill{spill_tolerance}"
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            known = json.load(f)
        if key in known:
            return known[key]
    buf = EntityBuffer(ctx, scene)
    table = []
    for cand in candidates:
        unroll, waves, sb, groups, queue, prio, gq, ho, hl, sparse, spread, classes = _knobs(cand)
        opts = options(unroll, waves, base, sb, groups, queue, prio, gq, ho, hl, sparse, spread, classes)
        r = Render(ctx, scene, fp_options=opts)
        c = Canvas(ctx, (width, height), seed_mode, seed)
        r.render_passes(buf, c, 1)
        ctx.sync()
        times = []
        for _ in range(reps):
            ctx.timer_begin()
            r.render_passes(buf, c, spp)
            times.append(ctx.timer_end())
        res = r.program.kernel_info("render_passes")
        table.append({"unroll": unroll, "waves_per_eu": waves, "scalar_bounds": sb, "groups": groups,
                      "pixel_queue": queue, "wave_prio": prio, "group_queue": gq, "hit_order": ho, "helpers": hl,
                      "sparse": sparse, "spread": spread, "classes": classes, "ms": sorted(times)[len(times) // 2],
                      "options": opts,
                      "vgprs": res["vgprs"], "scratch": res["scratch"]})
    best = pick(table, spill_tolerance)
    result = {"best": best, "table": table, "fastest_any": min(table, key=lambda t: t["ms"]),
              "spill_tolerance": spill_tolerance}
    if cache_path:
        known = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                known = json.load(f)
        known[key] = result
        os.makedirs(os.path.dirname(cache_path), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(known, f, indent=1)
    return result
