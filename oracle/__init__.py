"""CPU oracle of the render pass — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
this package, and only as the checker. The product path never does.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_size_t, c_uint8, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

F32, F64, COUNT = 0, 1, 2

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        l = ctypes.CDLL(LIB_PATH)
        l.oracle_last_error.restype = c_char_p
        l.oracle_render.restype = c_int
        l.oracle_render.argtypes = [c_char_p, c_int, c_uint32, c_uint32, c_uint32, c_void_p, c_void_p, c_uint32,
                                    c_uint32, c_uint32, c_uint32, c_int, POINTER(c_uint64)]
        l.oracle_count_flops.restype = c_int
        l.oracle_count_flops.argtypes = [c_char_p, c_int, c_uint32, c_uint32, c_uint32, c_void_p, c_void_p,
                                         c_uint32, c_uint32, c_uint32, c_uint32, c_int, POINTER(c_uint64),
                                         POINTER(c_uint64)]
        l.oracle_lcg_next.restype = c_uint32
        l.oracle_lcg_next.argtypes = [POINTER(c_uint32)]
        l.oracle_uniform_f32.restype = c_float
        l.oracle_uniform_f32.argtypes = [POINTER(c_uint32)]
        l.oracle_hemisphere_cosine_f64.argtypes = [POINTER(c_uint32), c_void_p]
        l.oracle_hash.restype = c_uint32
        l.oracle_hash.argtypes = [c_void_p, c_int]
        for n in ("oracle_q_mul", "oracle_q_div", "oracle_c22_dot", "oracle_mo_apply", "oracle_rot_apply"):
            getattr(l, n).argtypes = [c_void_p, c_void_p, c_void_p]
        l.oracle_mo_deriv.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p]
        l.oracle_hy_distance.restype = c_double
        l.oracle_hy_distance.argtypes = [c_void_p, c_void_p]
        l.oracle_look_at.argtypes = [c_void_p, c_void_p]
        l.oracle_shape_detect.restype = c_double
        l.oracle_shape_detect.argtypes = [c_int, c_double, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
        l.oracle_sdf_detect.restype = c_double
        l.oracle_sdf_detect.argtypes = [c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]
        l.oracle_post.argtypes = [c_void_p, c_size_t, c_uint32, c_float, c_void_p, c_void_p]
        l.oracle_material_interact.restype = c_int
        l.oracle_material_interact.argtypes = [c_char_p, c_void_p, c_void_p, c_void_p, c_void_p, POINTER(c_uint32),
                                               c_void_p, c_void_p, c_void_p]
        l.oracle_background.restype = c_int
        l.oracle_background.argtypes = [c_char_p, c_void_p, c_void_p, c_void_p]
        l.oracle_tiled_select.restype = c_int
        l.oracle_tiled_select.argtypes = [c_int, c_int, c_double, c_double, c_int, c_void_p]
        _lib = l
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(c_void_p)


def render(spec: str, w: int, h: int, passes: int, seeds: np.ndarray, canvas: np.ndarray = None,
           precision: int = F32, crop=None, threads: int = 0):
    """Returns (canvas float32[h, w, 4], seeds uint32[h, w], flops). Only the
    crop (x0, y0, cw, ch) is rendered; other pixels keep their inputs."""
    seeds = np.array(seeds, dtype=np.uint32).reshape(h, w).copy()
    canvas = np.zeros((h, w, 4), np.float32) if canvas is None else np.array(canvas, np.float32).reshape(h, w, 4).copy()
    x0, y0, cw, ch = crop if crop is not None else (0, 0, w, h)
    flops = c_uint64(0)
    rc = lib().oracle_render(spec.encode(), precision, w, h, passes, _p(canvas), _p(seeds), x0, y0, cw, ch, threads,
                             ctypes.byref(flops))
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return canvas, seeds, int(flops.value)


def count_flops(spec: str, w: int, h: int, passes: int, seeds: np.ndarray, culled, crop=None,
                threads: int = 0):
    """FP32 render of the crop on the counting scalars. Returns (canvas,
    seeds, fp32 flops, fp64 flops); culled: 0 / False = the reference's
    plain scan, 1 / True = the flat culled scan, g = the two-level scan over
    groups of g, 1000 + g = the same with the group queue. culled=True counts the culled scan's
    algorithm (FP64 cull tests, exact tests of the candidates only)."""
    seeds = np.array(seeds, dtype=np.uint32).reshape(h, w).copy()
    canvas = np.zeros((h, w, 4), np.float32)
    x0, y0, cw, ch = crop if crop is not None else (0, 0, w, h)
    f32, f64 = c_uint64(0), c_uint64(0)
    rc = lib().oracle_count_flops(spec.encode(), int(culled), w, h, passes, _p(canvas), _p(seeds), x0, y0, cw, ch,
                                  threads, ctypes.byref(f32), ctypes.byref(f64))
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return canvas, seeds, int(f32.value), int(f64.value)


def vec_call(name: str, *arrays, out_len: int = 4):
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
    out = np.zeros(out_len, np.float64)
    getattr(lib(), name)(*[_p(a) for a in args], _p(out))
    return out


def hy_distance(a, b) -> float:
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return float(lib().oracle_hy_distance(_p(a), _p(b)))


def shape_detect(kind: int, start, direction, repeat: bool = False, param: float = 0.0):
    s = np.zeros(4, np.float64)
    d = np.zeros(4, np.float64)
    s[: len(start)] = start
    d[: len(direction)] = direction
    os_, od, on = np.zeros(4), np.zeros(4), np.zeros(4)
    dist = lib().oracle_shape_detect(kind, param, int(repeat), _p(s), _p(d), _p(os_), _p(od), _p(on))
    return float(dist), os_, od, on


def sdf_detect(kind: int, params, start, direction, repeat: bool = False):
    """Sphere-traced shape in FP64 (0 sphere r, 1 box h, 2 torus R r):
    (distance or -1, hit point, normal)."""
    pr = np.zeros(3, np.float64)
    pr[: len(params)] = params
    s = np.asarray(start, np.float64).copy()
    d = np.asarray(direction, np.float64).copy()
    os_, on = np.zeros(3), np.zeros(3)
    dist = lib().oracle_sdf_detect(kind, _p(pr), int(repeat), _p(s), _p(d), _p(os_), _p(on))
    return float(dist), os_, on


def material_interact(spec: str, normal, direction, intensity=(1.0, 1.0, 1.0), emission=(0.0, 0.0, 0.0),
                      rng_state: int = 1):
    """One material interaction (FP64 geometry, FP32 colours). Returns
    (goes_on, direction, intensity, emission, rng_state after)."""
    n = np.ascontiguousarray(normal, np.float64)
    d = np.ascontiguousarray(direction, np.float64)
    i = np.ascontiguousarray(intensity, np.float32)
    e = np.ascontiguousarray(emission, np.float32)
    st = c_uint32(rng_state)
    od, oi, oe = np.zeros(3), np.zeros(3, np.float32), np.zeros(3, np.float32)
    r = lib().oracle_material_interact(spec.encode(), _p(n), _p(d), _p(i), _p(e), ctypes.byref(st), _p(od), _p(oi),
                                       _p(oe))
    if r < 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return bool(r), od, oi, oe, int(st.value)


def background(spec: str, direction, intensity=(1.0, 1.0, 1.0)):
    """Emission of a Euclidean background for an escaping path."""
    d = np.ascontiguousarray(direction, np.float64)
    i = np.ascontiguousarray(intensity, np.float32)
    out = np.zeros(3, np.float32)
    if lib().oracle_background(spec.encode(), _p(d), _p(i), _p(out)) != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return out


def tiled_select(obj: int, tiling: int, cell: float, border: float, n_materials: int, p) -> int:
    """Material index a tiled object picks at p (-1: border). obj 0 = tiled
    plane, 1 = tiled horosphere."""
    q = np.zeros(4, np.float64)
    q[: len(p)] = p
    r = lib().oracle_tiled_select(obj, tiling, cell, border, n_materials, _p(q))
    if r < -1:
        raise ValueError("bad tiled-object argument")
    return r


def post(canvas: np.ndarray, passes: int, gamma: float):
    canvas = np.ascontiguousarray(canvas, np.float32)
    n = canvas.size // 4
    image = np.zeros(canvas.shape, np.float32)
    rgba = np.zeros(canvas.shape, np.uint8)
    lib().oracle_post(_p(canvas), n, passes, gamma, _p(image), _p(rgba))
    return image, rgba
