"""Learned radiance cache: host side of the cached render method.

`RadianceCache` owns the MLP (FP32 master parameters, Adam state, BF16
weight images), its query and training queues, and the training buffers on
the GPU (`ht_cache_*` in include/hypertrace_hip.h). `CachedRender` runs one
cached pass per call: the scene program's `render_cached` kernel, the MLP
inference that resolves the queried pixels, and a few online training steps
on the suffixes the pass recorded.

The method (no reference counterpart): a path ends into the cache at its
`query_hop`-th hit (2: after two bounces), where the network predicts the
radiance the rest of the path would gather, times the path's throughput.
One pixel in `train_every` per pass (16) continues unbiased to full depth;
those paths stay in the image and their suffixes train the network.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import numpy as np

from . import native
from .objects import SceneImpl
from .processing import Buffer, Canvas, Context, EntityBuffer, Program, check_scene_type

N_PARAMS = 26115
N_FEATURES = 12
HIDDEN = 64
IN = 144


def param_layout():
    """(name, offset, shape) of the FP32 master parameters."""
    out, off = [], 0
    shapes = [("W0", (HIDDEN, IN)), ("b0", (HIDDEN,))]
    for l in range(1, 5):
        shapes += [(f"W{l}", (HIDDEN, HIDDEN)), (f"b{l}", (HIDDEN,))]
    shapes += [("W5", (3, HIDDEN)), ("b5", (3,))]
    for name, shape in shapes:
        out.append((name, off, shape))
        off += int(np.prod(shape))
    assert off == N_PARAMS
    return out


def unpack_params(flat: np.ndarray) -> dict:
    return {name: flat[off:off + int(np.prod(shape))].reshape(shape) for name, off, shape in param_layout()}


def pack_params(p: dict) -> np.ndarray:
    flat = np.zeros(N_PARAMS, np.float32)
    for name, off, shape in param_layout():
        flat[off:off + int(np.prod(shape))] = np.asarray(p[name], np.float32).reshape(-1)
    return flat


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights for the ReLU layers, zero biases."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, _, shape in param_layout():
        if name.startswith("W"):
            bound = np.sqrt(6.0 / shape[1])
            p[name] = rng.uniform(-bound, bound, shape)
        else:
            p[name] = np.zeros(shape)
    return pack_params(p)


class RadianceCache:
    def __init__(self, ctx: Context, query_capacity: int, train_capacity: int = 65536, max_batch: int = 16384,
                 seed: int = 0):
        self.ctx = ctx
        self.max_batch = max_batch
        h = ctypes.c_void_p()
        native.check(native.lib().ht_cache_create(ctx.handle, query_capacity, train_capacity, max_batch,
                                                  ctypes.byref(h)))
        self.handle = h
        self.set_params(init_params(seed))

    def close(self) -> None:
        if self.handle:
            native.lib().ht_cache_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, flat: np.ndarray) -> None:
        flat = np.ascontiguousarray(flat, np.float32)
        native.check(native.lib().ht_cache_set_params(self.handle, flat.ctypes.data_as(ctypes.c_void_p), flat.size))

    def params(self) -> np.ndarray:
        out = np.zeros(N_PARAMS, np.float32)
        native.check(native.lib().ht_cache_get_params(self.handle, out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out

    def grads(self) -> np.ndarray:
        out = np.zeros(N_PARAMS, np.float32)
        native.check(native.lib().ht_cache_get_grads(self.handle, out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out

    def counts(self) -> Tuple[int, int]:
        q, r = ctypes.c_uint32(), ctypes.c_uint32()
        native.check(native.lib().ht_cache_counts(self.handle, ctypes.byref(q), ctypes.byref(r)))
        return q.value, r.value

    # explicit-data entry points (tests, offline use)
    def forward(self, features: np.ndarray) -> np.ndarray:
        """features [n, 12] -> predictions [n, 3]."""
        f = np.ascontiguousarray(np.asarray(features, np.float32).T)
        n = f.shape[1]
        fb = Buffer(self.ctx, f.nbytes)
        fb.write(f)
        pb = Buffer(self.ctx, 12 * n)
        native.check(native.lib().ht_mlp_forward(self.handle, fb.handle, n, pb.handle))
        return pb.read(np.float32, 3 * n).reshape(3, n).T.copy()

    def train_step(self, features: np.ndarray, targets: np.ndarray, lr: float = 1e-3, apply: bool = True) -> float:
        f = np.ascontiguousarray(np.asarray(features, np.float32).T)
        t = np.ascontiguousarray(np.asarray(targets, np.float32).T)
        n = f.shape[1]
        fb = Buffer(self.ctx, f.nbytes)
        fb.write(f)
        tb = Buffer(self.ctx, t.nbytes)
        tb.write(t)
        loss = ctypes.c_float()
        native.check(native.lib().ht_mlp_train_step(self.handle, fb.handle, tb.handle, n, lr, int(apply),
                                                    ctypes.byref(loss)))
        return float(loss.value)


class CachedRender:
    """Cached render method for a scene type whose program has the
    `render_cached` kernel (a vector of mapped covered objects)."""

    def __init__(self, ctx: Context, scene: SceneImpl, cache: RadianceCache, query_hop: int = 2,
                 train_every: int = 16, steps_per_pass: int = 4, batch: int = 16384, lr: float = 1e-3,
                 fp_options: Optional[str] = None):
        self.ctx = ctx
        self.cache = cache
        self.program = Program(ctx, scene, fp_options=fp_options)
        self.query_hop = query_hop
        self.train_every = train_every
        self.steps_per_pass = steps_per_pass
        self.batch = batch
        self.lr = lr
        self.passes_done = 0

    def render_pass(self, scene_buffer: EntityBuffer, canvas: Canvas, train: bool = True) -> None:
        check_scene_type(self.program.scene_type, scene_buffer)
        lib = native.lib()
        native.check(lib.ht_render_cached(self.program.handle, self.cache.handle, canvas.width, canvas.height,
                                          scene_buffer.handle, canvas.image.handle, canvas.seeds.handle,
                                          self.passes_done, self.train_every, self.query_hop))
        canvas.add_pass()
        self.passes_done += 1
        if train and self.steps_per_pass > 0:
            native.check(lib.ht_cache_train(self.cache.handle, self.steps_per_pass, self.batch, self.lr))
