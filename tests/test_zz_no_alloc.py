"""ht_render_wavefront_noalloc through the C ABI: the same canvas and seeds as
ht_render_wavefront (which keeps a workspace it allocates), on two calls in a
row with the second frame smaller, and a clear refusal of a frame whose
queues exceed the fixed workspace."""
import numpy as np
import pytest

from hypertrace_amd import native, scenes
from hypertrace_amd.native import HtError
from hypertrace_amd.processing import Canvas, Context, EntityBuffer, Render
from hypertrace_amd.types import X32_F32

pytestmark = pytest.mark.gpu


def _noalloc(render, buf, canvas, passes):
    native.check(native.lib().ht_render_wavefront_noalloc(render.program.handle, canvas.width, canvas.height,
                                                          buf.handle, canvas.image.handle, canvas.seeds.handle,
                                                          passes))
    canvas.add_pass(passes)


def _alloc(render, buf, canvas, passes):
    native.check(native.lib().ht_render_wavefront(render.program.handle, canvas.width, canvas.height, buf.handle,
                                                  canvas.image.handle, canvas.seeds.handle, passes))
    canvas.add_pass(passes)


def test_noalloc_equals_wavefront_on_two_calls():
    with Context(0, X32_F32) as ctx:
        first = scenes.hyperbolic_cells(160, 120, 1)
        render = Render(ctx, first.scene, wavefront=True)
        for (w, h, passes) in ((160, 120, 3), (96, 72, 2)):  # the second call smaller
            wl = scenes.hyperbolic_cells(w, h, 1)
            buf = EntityBuffer(ctx, wl.scene)
            print(f"case noalloc {w}x{h} passes {passes}", flush=True)
            a = Canvas(ctx, (w, h), wl.seed_mode, wl.seed)
            b = Canvas(ctx, (w, h), wl.seed_mode, wl.seed)
            _alloc(render, buf, a, passes)
            _noalloc(render, buf, b, passes)
            np.testing.assert_array_equal(b.load_image().view(np.uint32), a.load_image().view(np.uint32))
            np.testing.assert_array_equal(b.load_seeds(), a.load_seeds())
            assert np.all(b.load_image()[..., 3] == passes)


def test_noalloc_refuses_frames_beyond_its_workspace():
    with Context(0, X32_F32) as ctx:
        wl = scenes.hyperbolic_cells(640, 480, 1)  # ~550 B of queues per pixel: above 64 MiB
        render = Render(ctx, wl.scene, wavefront=True)
        buf = EntityBuffer(ctx, wl.scene)
        c = Canvas(ctx, (640, 480), wl.seed_mode, wl.seed)
        with pytest.raises(HtError, match="fixed workspace"):
            _noalloc(render, buf, c, 1)
        assert np.all(c.load_image() == 0.0)  # nothing ran
