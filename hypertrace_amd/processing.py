"""Host runtime over the C ABI, mirroring the reference `processing` crate.

Same names, argument meaning and error behaviour as the reference types:
`Context` (processing/src/context.rs:3-16), `EntityBuffer` (buffer.rs:8-34),
`Render` (render.rs:8-119), `Canvas` / `ImageBuffer` / `Extractor` /
`Packer` (canvas.rs:11-221), `GammaFilter` (filter/gamma.rs:3-58) and
`Pipeline` (pipeline.rs:8-67). Errors surface as `native.HtError`.
"""
from __future__ import annotations

import ctypes
from ctypes import c_float, c_void_p
from typing import Optional, Tuple

import numpy as np

from . import codegen, native
from .objects import SceneImpl
from .types import Config, X32_F32

REFERENCE_SEED = 0xDEADBEEFDEADBEEF  # processing/src/canvas.rs:75


class Context:
    """One device and one in-order stream, plus the layout config."""

    def __init__(self, device: int = 0, config: Config = X32_F32):
        self.config = config
        self._lib = native.lib()
        h = c_void_p()
        native.check(self._lib.ht_ctx_create(device, ctypes.byref(h)))
        self.handle = h

    def sync(self) -> None:
        native.check(self._lib.ht_ctx_sync(self.handle))

    def close(self) -> None:
        if self.handle:
            self._lib.ht_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def timer_begin(self) -> None:
        native.check(self._lib.ht_timer_begin(self.handle))

    def timer_end(self) -> float:
        ms = c_float()
        native.check(self._lib.ht_timer_end(self.handle, ctypes.byref(ms)))
        return float(ms.value)


class Buffer:
    """Device buffer owned by the caller."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        h = c_void_p()
        native.check(native.lib().ht_buf_alloc(ctx.handle, self.nbytes, ctypes.byref(h)))
        self.handle = h

    def write(self, data) -> None:
        if isinstance(data, (bytes, bytearray)):
            arr = np.frombuffer(bytes(data), dtype=np.uint8)
        else:
            arr = np.ascontiguousarray(data)
        native.check(native.lib().ht_buf_write(self.handle, arr.ctypes.data_as(c_void_p), arr.nbytes))

    def read(self, dtype=np.uint8, count: Optional[int] = None) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else count
        out = np.empty(n, dtype=dt)
        native.check(native.lib().ht_buf_read(self.handle, out.ctypes.data_as(c_void_p), out.nbytes))
        return out

    def zero(self) -> None:
        native.check(native.lib().ht_buf_zero(self.handle))

    def free(self) -> None:
        if self.handle:
            native.check(native.lib().ht_buf_free(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class EntityBuffer(Buffer):
    """The scene serialised with the context's config and copied to the
    device (processing/src/buffer.rs:14-29)."""

    def __init__(self, ctx: Context, scene: SceneImpl):
        data = scene.to_bytes(ctx.config)
        super().__init__(ctx, max(len(data), 1))
        self.write(data)
        self.scene_bytes = data
        # The reference's EntityBuffer<S> fixes the scene type at compile
        # time; here the type travels with the bytes and is checked before a
        # program reads them (a buffer of another type would be read with
        # the wrong layout, and its lengths could send reads past its end).
        self.scene_type = scene_type_key(scene, ctx.config)


def scene_type_key(scene: SceneImpl, config: Config) -> str:
    """The scene type (its device type expression) and the layout config:
    what a program and a scene buffer must agree on."""
    return f"{scene.cpp()}|x{config.address_width}|f{64 if config.double_support else 32}"


def check_scene_type(expected: Optional[str], buffer: "EntityBuffer") -> None:
    """HtError (Error::Other) if `buffer` holds a scene of another type than
    the program was built for."""
    got = getattr(buffer, "scene_type", None)
    if expected is not None and got != expected:
        raise native.HtError(-3, f"scene buffer of type {got} given to a program built for {expected}")


class ImageBuffer(Buffer):
    """w*h pixels of N channels of `dtype` (canvas.rs:125-189)."""

    def __init__(self, ctx: Context, shape: Tuple[int, int], channels: int = 4, dtype=np.float32):
        self.shape = (int(shape[0]), int(shape[1]))
        self.channels = channels
        self.dtype = np.dtype(dtype)
        super().__init__(ctx, self.shape[0] * self.shape[1] * channels * self.dtype.itemsize)

    @property
    def width(self) -> int:
        return self.shape[0]

    @property
    def height(self) -> int:
        return self.shape[1]

    def load(self) -> np.ndarray:
        return self.read(self.dtype).reshape(self.height, self.width, self.channels)

    def store(self, image: np.ndarray) -> None:
        image = np.ascontiguousarray(image, dtype=self.dtype)
        if image.size != self.width * self.height * self.channels:
            raise native.HtError(-3, f"image of {image.size} values for a {self.shape} x {self.channels} buffer")
        self.write(image)


class Canvas:
    """Progressive float4 accumulation image plus the per-pixel seed buffer
    (canvas.rs:11-118). `clean()` zeroes the image and the pass count but
    keeps the seeds, as the reference does (canvas.rs:102-106)."""

    def __init__(self, ctx: Context, shape: Tuple[int, int], seed_mode: int = 0, seed: int = REFERENCE_SEED):
        self.ctx = ctx
        self.shape = (int(shape[0]), int(shape[1]))
        self.image = ImageBuffer(ctx, self.shape, 4, np.float32)
        self.image.zero()
        self.seeds = ImageBuffer(ctx, self.shape, 1, np.uint32)
        native.check(native.lib().ht_seed_init(self.seeds.handle, self.shape[0], self.shape[1], seed_mode, seed))
        self.passes = 0

    @property
    def width(self) -> int:
        return self.shape[0]

    @property
    def height(self) -> int:
        return self.shape[1]

    def add_pass(self, n: int = 1) -> None:
        self.passes += n

    def clean(self) -> None:
        self.image.zero()
        self.ctx.sync()
        self.passes = 0

    def load_image(self) -> np.ndarray:
        return self.image.load()

    def load_seeds(self) -> np.ndarray:
        return self.seeds.load()[..., 0]

    def store_seeds(self, seeds: np.ndarray) -> None:
        self.seeds.store(np.asarray(seeds, dtype=np.uint32).reshape(self.height, self.width, 1))


def seed_fill(n: int, mode: int = 0, seed: int = REFERENCE_SEED) -> np.ndarray:
    """Host-side seed generation (same modes as Canvas)."""
    out = np.empty(n, dtype=np.uint32)
    native.check(native.lib().ht_seed_fill(out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n, mode, seed))
    return out


class Program:
    """A compiled render program for one scene type."""

    def __init__(self, ctx: Context, scene: Optional[SceneImpl], name: str = "render.hip",
                 fp_options: Optional[str] = None, source: Optional[str] = None):
        """`source` replaces the generated root source (batch programs)."""
        self.ctx = ctx
        self.fp_options = fp_options
        self.source = source if source is not None else codegen.root_source(scene, ctx.config)
        hdrs = codegen.program_headers(scene)
        names = native.cstr_array(list(hdrs.keys()))
        contents = native.cstr_array(list(hdrs.values()))
        h = c_void_p()
        lib = native.lib()
        native.check(lib.ht_prog_build_ex(ctx.handle, name.encode(), self.source.encode(), len(hdrs), names,
                                          contents, 1 if ctx.config.double_support else 0,
                                          ctx.config.address_width,
                                          fp_options.encode() if fp_options is not None else None, ctypes.byref(h)))
        self.cache_hit = bool(lib.ht_prog_cache_hit())
        self.handle = h
        self.scene_type = scene_type_key(scene, ctx.config) if scene is not None else None

    def kernel_info(self, kernel: str = "render_passes") -> dict:
        """Vector registers and scratch bytes per lane (register spills) and
        static LDS bytes per block of one of the program's kernels."""
        v, sc, lds = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        native.check(native.lib().ht_prog_kernel_info(self.handle, kernel.encode(), ctypes.byref(v),
                                                      ctypes.byref(sc), ctypes.byref(lds)))
        return {"vgprs": v.value, "scratch": sc.value, "lds": lds.value}

    def close(self) -> None:
        if self.handle:
            native.lib().ht_prog_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def program_options(fast: bool = False, rr_depth: Optional[int] = None, no_cull: bool = False,
                    scalar_bounds: bool = False, hy_distance_fp64: bool = False,
                    eps32: Optional[float] = None) -> str:
    """Program build options: strict (default) or fast FP32; Russian roulette
    from the rr_depth-th hit (a different method: it draws one more uniform
    per hit from there on); the element-by-element scan (comparisons); the
    culled scan's bounds in a global FP64 array read through the scalar
    cache instead of LDS (same images; frees LDS on scenes with many
    elements); the hyperbolic distance in FP64 with a cancellation-free
    acosh (a precision method: not bit-identical to the reference); the FP32
    hit window `eps32` instead of the reference's 1e-6 (a precision method:
    FP32 hit points at world scale ~10 round by ~1e-6, so the reference's
    window lets FP32 images drift from FP64 ones; 1e-5 removes that)."""
    opts = FAST_FP if fast else STRICT_FP
    if rr_depth is not None:
        opts += f" -DHT_RR_DEPTH={int(rr_depth)}"
    if no_cull:
        opts += " -DHT_NO_CULL=1"
    if scalar_bounds:
        opts += " -DHT_SCALAR_BOUNDS=1"
    if hy_distance_fp64:
        opts += " -DHT_HY_DIST_FP64=1"
    if eps32 is not None:
        opts += f" -DHT_EPS32={float(eps32)!r}f"
    return opts


# Program FP modes (see include/hypertrace_hip.h, ht_prog_build_ex).
STRICT_FP = "-ffp-contract=off -DHT_STRICT_FP=1"  # the default
FAST_FP = "-fno-hip-fp32-correctly-rounded-divide-sqrt"
NO_CULL_FP = STRICT_FP + " -DHT_NO_CULL=1"  # strict, element-by-element scan


def precompile(scene: Optional[SceneImpl], config: Config, name: str = "render.hip",
               fp_options: Optional[str] = None, source: Optional[str] = None) -> bool:
    """Compile a scene's program (or the given root source) into the
    code-object cache (no GPU needed). Returns True if it was already cached."""
    source = source if source is not None else codegen.root_source(scene, config)
    hdrs = codegen.program_headers(scene)
    lib = native.lib()
    native.check(lib.ht_prog_precompile(name.encode(), source.encode(), len(hdrs),
                                        native.cstr_array(list(hdrs.keys())),
                                        native.cstr_array(list(hdrs.values())),
                                        1 if config.double_support else 0, config.address_width,
                                        fp_options.encode() if fp_options is not None else None))
    return bool(lib.ht_prog_cache_hit())


class Render:
    """Render program of a scene type (render.rs:8-119). `render` runs one
    synchronous pass; `render_passes` queues several passes in one launch;
    with `wavefront`, `render_wavefront` computes the same passes as queued
    extend / shade kernels (ht_render_wavefront)."""

    def __init__(self, ctx: Context, scene: SceneImpl, fp_options: Optional[str] = None, wavefront: bool = False):
        self.ctx = ctx
        source = codegen.root_source(scene, ctx.config, wavefront=True) if wavefront else None
        self.program = Program(ctx, scene, fp_options=fp_options, source=source)

    def render(self, scene_buffer: EntityBuffer, canvas: Canvas) -> None:
        check_scene_type(self.program.scene_type, scene_buffer)
        native.check(native.lib().ht_render(self.program.handle, canvas.width, canvas.height, scene_buffer.handle,
                                            canvas.image.handle, canvas.seeds.handle))
        canvas.add_pass()

    def render_passes(self, scene_buffer: EntityBuffer, canvas: Canvas, passes: int) -> None:
        check_scene_type(self.program.scene_type, scene_buffer)
        native.check(native.lib().ht_render_passes(self.program.handle, canvas.width, canvas.height,
                                                   scene_buffer.handle, canvas.image.handle, canvas.seeds.handle,
                                                   passes))
        canvas.add_pass(passes)

    def render_wavefront(self, scene_buffer: EntityBuffer, canvas: Canvas, passes: int) -> None:
        check_scene_type(self.program.scene_type, scene_buffer)
        native.check(native.lib().ht_render_wavefront(self.program.handle, canvas.width, canvas.height,
                                                      scene_buffer.handle, canvas.image.handle,
                                                      canvas.seeds.handle, passes))
        canvas.add_pass(passes)


class _TypeOnly:
    def __init__(self, key: str):
        self.scene_type = key


class Extractor:
    """image = canvas / passes (canvas.rs:120-168)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def process(self, canvas: Canvas, image: ImageBuffer) -> None:
        if canvas.shape != image.shape:
            raise native.HtError(-3, "shape mismatch")
        native.check(native.lib().ht_extract(self.ctx.handle, image.width, image.height, canvas.passes,
                                             canvas.image.handle, image.handle))


class GammaFilter:
    """out = pow(max(in, 0), gamma) (filter/gamma.rs:3-58)."""

    def __init__(self, ctx: Context, gamma: float):
        self.ctx = ctx
        self.gamma = float(gamma)

    def process(self, inp: ImageBuffer, out: ImageBuffer) -> None:
        if inp.shape != out.shape:
            raise native.HtError(-3, "shape mismatch")
        native.check(native.lib().ht_gamma(self.ctx.handle, inp.width, inp.height, self.gamma, inp.handle,
                                           out.handle))


class Packer:
    """RGBA8 = trunc(255 * clamp(x, 0, 1)) (canvas.rs:170-221)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def process(self, image: ImageBuffer, packed: ImageBuffer) -> None:
        if image.shape != packed.shape:
            raise native.HtError(-3, "shape mismatch")
        native.check(native.lib().ht_pack(self.ctx.handle, image.width, image.height, image.handle, packed.handle))


class Pipeline:
    """Scene buffer + render + canvas + extract/gamma/pack (pipeline.rs:8-67).
    `get_image` runs the three post passes in order like the reference;
    `get_image(fused=True)` runs them as one kernel with identical bytes."""

    def __init__(self, ctx: Context, size: Tuple[int, int], scene: SceneImpl, gamma: float = 1.0 / 2.2,
                 seed_mode: int = 0, seed: int = REFERENCE_SEED):
        self.ctx = ctx
        self.scene_buffer = EntityBuffer(ctx, scene)
        self.render_ = Render(ctx, scene)
        self.extractor = Extractor(ctx)
        self.filter = GammaFilter(ctx, gamma)
        self.packer = Packer(ctx)
        self.canvas = Canvas(ctx, size, seed_mode, seed)
        self.image_buf_src = ImageBuffer(ctx, size)
        self.image_buf_dst = ImageBuffer(ctx, size)
        self.image_buf_packed = ImageBuffer(ctx, size, 4, np.uint8)

    def set_scene(self, scene: SceneImpl) -> None:
        """Re-serialises and uploads the scene (pipeline.rs:44-47); the scene
        must be of the type the pipeline's program was built for."""
        check_scene_type(self.render_.program.scene_type, _TypeOnly(scene_type_key(scene, self.ctx.config)))
        self.scene_buffer = EntityBuffer(self.ctx, scene)

    def render(self) -> None:
        self.render_.render(self.scene_buffer, self.canvas)

    def render_passes(self, passes: int) -> None:
        self.render_.render_passes(self.scene_buffer, self.canvas, passes)

    def get_image(self, fused: bool = False) -> np.ndarray:
        if fused:
            native.check(native.lib().ht_post(self.ctx.handle, self.canvas.width, self.canvas.height,
                                              self.canvas.passes, self.filter.gamma, self.canvas.image.handle,
                                              self.image_buf_packed.handle))
        else:
            self.extractor.process(self.canvas, self.image_buf_src)
            self.filter.process(self.image_buf_src, self.image_buf_dst)
            # The reference filter sequence swaps src/dst after each filter
            # (filter/sequence.rs:12-19), so the packer reads the filtered image.
            self.image_buf_src, self.image_buf_dst = self.image_buf_dst, self.image_buf_src
            self.packer.process(self.image_buf_src, self.image_buf_packed)
        return self.image_buf_packed.load()

    def clean(self) -> None:
        self.canvas.clean()

    def save_image(self, path: str, fused: bool = True) -> np.ndarray:
        """Post passes, then the RGB of the RGBA8 image to a binary PPM (P6)
        file at `path` (for visual regression); returns the RGBA8 image."""
        rgba = self.get_image(fused=fused)
        write_ppm(path, rgba)
        return rgba


def write_ppm(path: str, rgba8: np.ndarray) -> None:
    """Binary PPM (P6, maxval 255) of an [h, w, 4] uint8 image, rows top
    to bottom as the canvas stores them (pixel x + w * y)."""
    img = np.ascontiguousarray(np.asarray(rgba8, np.uint8)[..., :3])
    h, w = img.shape[:2]
    with open(path, "wb") as f:
        f.write(f"P6\n{w} {h}\n255\n".encode())
        f.write(img.tobytes())


def read_ppm(path: str) -> np.ndarray:
    """[h, w, 3] uint8 image of a binary PPM written by write_ppm."""
    with open(path, "rb") as f:
        data = f.read()
    parts, pos = [], 0
    while len(parts) < 4:
        while data[pos:pos + 1].isspace():
            pos += 1
        end = pos
        while not data[end:end + 1].isspace():
            end += 1
        parts.append(data[pos:end])
        pos = end
    pos += 1
    if parts[0] != b"P6" or int(parts[3]) != 255:
        raise ValueError("not a binary 8-bit PPM")
    w, h = int(parts[1]), int(parts[2])
    return np.frombuffer(data[pos:pos + 3 * w * h], np.uint8).reshape(h, w, 3)
