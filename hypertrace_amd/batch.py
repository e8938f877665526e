"""Batched render of many jobs in one persistent launch (host side of
`ht_batch_*` in include/hypertrace_hip.h), its hipGraph frame, and the
autotuner that picks the block size and tile size of the batch program.

A batch program is compiled for two scene types (Euclidean and hyperbolic
jobs); block size (64..1024 threads) and tile size (pixels per work item)
are compile-time, so each candidate of the tuner is its own program (hipRTC,
cached on disk like every program). The tuner times whole frames of the
caller's job set on the GPU and keeps the fastest candidate; the result is
stored in a JSON file keyed by device and job-set signature.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
import time
from ctypes import Structure, c_uint32, c_void_p
from typing import Dict, List, Optional, Sequence, Tuple

from . import codegen, native
from .processing import Buffer, Canvas, Context, EntityBuffer, ImageBuffer, Program, precompile, scene_type_key
from .scenes import BatchJobSpec, batch_euclidean_scene, batch_hyperbolic_scene
from .types import Config

import numpy as np


class _Job(Structure):
    _fields_ = [("type", c_uint32), ("w", c_uint32), ("h", c_uint32), ("spp", c_uint32), ("depth", c_uint32),
                ("scene", c_void_p), ("canvas", c_void_p), ("seeds", c_void_p), ("rgba8", c_void_p)]


def _lib():
    return native.lib()


def prototype_scenes():
    """Scene instances of the two batch types (only their types matter)."""
    rng = np.random.default_rng(0)
    return batch_euclidean_scene(rng, 8), batch_hyperbolic_scene(rng, 9)


def batch_root(config: Config, block: int, max_elements: int = 256) -> str:
    s0, s1 = prototype_scenes()
    return codegen.batch_source(s0, s1, config, block, max_elements)


class BatchProgram:
    def __init__(self, ctx: Context, block: int = 256, max_elements: int = 256, fp_options: Optional[str] = None):
        self.block = block
        self.program = Program(ctx, None, name="batch.hip", fp_options=fp_options,
                               source=batch_root(ctx.config, block, max_elements))


def precompile_batch(config: Config, block: int, fp_options: Optional[str] = None) -> bool:
    return precompile(None, config, name="batch.hip", fp_options=fp_options, source=batch_root(config, block))


class Batch:
    """The device-side job set of one frame: buffers of every job and the
    job table."""

    def __init__(self, ctx: Context, bprog: BatchProgram, specs: Sequence[BatchJobSpec], tile: Optional[int] = None,
                 with_rgba8: bool = False):
        self.ctx = ctx
        self.bprog = bprog
        self.tile = tile if tile is not None else bprog.block
        self.specs = list(specs)
        # each job's scene must be of its type's scene type (Scene0 / Scene1
        # of the batch program), or the kernel would read it with another layout
        types = [scene_type_key(p, ctx.config) for p in prototype_scenes()]
        for k, s in enumerate(self.specs):
            if s.type not in (0, 1):
                raise native.HtError(-3, f"job {k}: type {s.type} is neither 0 nor 1")
            got = scene_type_key(s.scene, ctx.config)
            if got != types[s.type]:
                raise native.HtError(-3, f"job {k}: scene of type {got} is not the batch program's type {s.type}")
        self.scene_bufs = [EntityBuffer(ctx, s.scene) for s in self.specs]
        self.canvases = [Canvas(ctx, (s.width, s.height), 1, s.seed) for s in self.specs]
        self.rgba8 = [ImageBuffer(ctx, (s.width, s.height), 4, np.uint8) if with_rgba8 else None for s in self.specs]
        arr = (_Job * len(self.specs))()
        for k, (s, b, c, r) in enumerate(zip(self.specs, self.scene_bufs, self.canvases, self.rgba8)):
            arr[k] = _Job(s.type, s.width, s.height, s.spp, s.depth, b.handle, c.image.handle, c.seeds.handle,
                          r.handle if r is not None else None)
        h = c_void_p()
        native.check(_lib().ht_batch_create(bprog.program.handle, len(self.specs), ctypes.cast(arr, c_void_p),
                                            self.tile, ctypes.byref(h)))
        self.handle = h
        self.tiles = _lib().ht_batch_tiles(h)
        self.paths = sum(s.width * s.height * s.spp for s in self.specs)

    def close(self):
        if self.handle:
            _lib().ht_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, grid: int, clear: bool = True, gamma: float = 1 / 2.2, post: bool = False) -> None:
        native.check(_lib().ht_batch_render(self.handle, grid, int(clear), gamma, int(post)))
        for c, s in zip(self.canvases, self.specs):
            if clear:
                c.passes = 0
            c.add_pass(s.spp)

    def capture(self, grid: int, clear: bool = True, gamma: float = 1 / 2.2, post: bool = False) -> None:
        native.check(_lib().ht_batch_capture(self.handle, grid, int(clear), gamma, int(post)))
        self._graph_clear = clear

    def replay(self) -> None:
        native.check(_lib().ht_batch_replay(self.handle))
        for c, s in zip(self.canvases, self.specs):
            if self._graph_clear:
                c.passes = 0
            c.add_pass(s.spp)


def _signature(specs: Sequence[BatchJobSpec]) -> str:
    h = hashlib.sha1()
    for s in specs:
        h.update(f"{s.type},{s.width},{s.height},{s.spp},{s.depth},{len(s.scene.object.items)};".encode())
    return h.hexdigest()[:16]


BLOCKS = (64, 128, 256, 512, 1024)
TILE_MULTS = (1, 2, 4)
CU_COUNT = 256
# Program options tried per block size: the default (bounds staged in LDS
# per tile) and, from 128 threads up, the per-job FP64 bounds read through
# the scalar cache (same images; without bounds in LDS the register cap of
# 4 waves per SIMD sets the occupancy).
SCALAR_BOUNDS = "-ffp-contract=off -DHT_STRICT_FP=1 -DHT_SCALAR_BOUNDS=1"
# Scalar-cache bounds, and each wave takes its own tiles and hands the
# tile's pixels to its lanes as they finish (same images).
WAVE_QUEUE = SCALAR_BOUNDS + " -DHT_BATCH_WAVE_QUEUE=1"
# Wave priority (HT_WAVE_PRIO): block tiles step each wave's issue priority
# down as its pixels' samples complete; the wave queue raises it on the
# frame's last tiles by the work the wave's tile has left. Either way the
# waves that will finish last issue first.
SCALAR_BOUNDS_PRIO = SCALAR_BOUNDS + " -DHT_WAVE_PRIO=1"
WAVE_QUEUE_PRIO = WAVE_QUEUE + " -DHT_WAVE_PRIO=1"
# The wave queue with each job's bounds scanned in two levels (groups of 4
# Morton-ordered elements; the per-job trees are built by batch_bounds_fill).
WAVE_QUEUE_GROUPS = WAVE_QUEUE_PRIO + " -DHT_CULL_GROUPS=4"
# The wave queue with the scene programs' two-level scan compacted at both
# levels (group queue; each job's leaves read from its own FP64 array), the
# hit order, helper lanes (lanes without a pixel once the tile's pixels are
# taken join the scans) and sparse level 1.
WAVE_QUEUE_HIT_ORDER = WAVE_QUEUE_GROUPS + " -DHT_GROUP_QUEUE=1 -DHT_HIT_ORDER=1 -DHT_HELPERS=1 -DHT_SPARSE_L1=32"
# Block tiles with the same scan (static helper lanes).
BLOCK_HIT_ORDER = SCALAR_BOUNDS_PRIO + " -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1 -DHT_HIT_ORDER=1 -DHT_HELPERS=1" \
                                       " -DHT_SPARSE_L1=32"
# The same two at a register cap of 3 waves per SIMD (170 VGPRs), where
# the 4-wave cap (128) spills.
WAVE_QUEUE_HIT_ORDER_W3 = WAVE_QUEUE_HIT_ORDER + " -DHT_WAVES_PER_EU=3"
BLOCK_HIT_ORDER_W3 = BLOCK_HIT_ORDER + " -DHT_WAVES_PER_EU=3"
OPTIONS = ((None, BLOCKS), (SCALAR_BOUNDS, (128, 256, 512, 1024)), (WAVE_QUEUE, (64, 128, 256)),
           (WAVE_QUEUE_PRIO, (64, 128, 256)), (WAVE_QUEUE_GROUPS, (64, 128, 256)), (SCALAR_BOUNDS_PRIO, (128, 256)),
           (WAVE_QUEUE_HIT_ORDER, (64, 128, 256)), (BLOCK_HIT_ORDER, (128, 256)),
           (WAVE_QUEUE_HIT_ORDER_W3, (64, 128, 256)), (BLOCK_HIT_ORDER_W3, (128, 256)))


def tile_unit(opts: Optional[str], block: int) -> int:
    """Tiles are multiples of the block, or of one wave (64 pixels) in
    wave-queue programs, which take tiles per wave."""
    return 64 if opts and "HT_BATCH_WAVE_QUEUE" in opts else block


def autotune(ctx: Context, specs: Sequence[BatchJobSpec], blocks=BLOCKS, tile_mults=TILE_MULTS,
             grid_waves_per_cu=(8, 16), reps: int = 2, cache_path: Optional[str] = None,
             options=OPTIONS, progress=None) -> Dict:
    """Times one frame of `specs` for every (program options, block, tile,
    grid) candidate and returns the fastest with the whole table. grid =
    blocks such that each CU gets the given number of waves; `options` pairs
    program options with the block sizes they are tried at; `progress`, if
    given, is called with a line of text after each program's candidates."""
    key = f"gfx950:{_signature(specs)}"
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            known = json.load(f)
        if key in known:
            return known[key]
    table = []
    candidates = [(opts, block) for opts, opt_blocks in options for block in blocks if block in opt_blocks]
    for opts, block in candidates:
        bprog = BatchProgram(ctx, block, fp_options=opts)
        res = bprog.program.kernel_info("render_batch")
        for m in tile_mults:
            tile = tile_unit(opts, block) * m
            batch = Batch(ctx, bprog, specs, tile)
            for wpc in grid_waves_per_cu:
                grid = max(1, CU_COUNT * wpc * 64 // block)
                batch.render(grid)  # warm
                ctx.sync()
                times = []
                for _ in range(reps):
                    t0 = time.perf_counter()
                    batch.render(grid)
                    ctx.sync()
                    times.append(time.perf_counter() - t0)
                table.append({"block": block, "tile": tile, "grid": grid, "options": opts, "ms": 1e3 * min(times),
                              "vgprs": res["vgprs"], "scratch": res["scratch"]})
            batch.close()
        if progress is not None:
            progress(f"batch autotune: block {block} [{opts}]: best "
                     f"{min(r['ms'] for r in table[-len(tile_mults) * len(grid_waves_per_cu):]):.1f} ms "
                     f"({res['vgprs']} VGPRs, {res['scratch']} B scratch)")
    best = min(table, key=lambda r: r["ms"])
    result = {"best": best, "table": table, "paths": sum(s.width * s.height * s.spp for s in specs)}
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
