"""Runtime object lifetimes through the C ABI: buffers and programs retain
their context (as OpenCL memory objects retain theirs), so releasing them
after the context is destroyed is valid and leaves no error behind for the
next call."""
import numpy as np
import pytest

from hypertrace_amd import scenes
from hypertrace_amd.processing import Canvas, Context, EntityBuffer, Extractor, ImageBuffer, Render
from hypertrace_amd.types import X32_F32

pytestmark = pytest.mark.gpu


def test_release_after_context_destroy():
    wl = scenes.euclidean_spheres(32, 24, 1)
    ctx = Context(0, X32_F32)
    buf = EntityBuffer(ctx, wl.scene)
    render = Render(ctx, wl.scene)
    canvas = Canvas(ctx, (32, 24), wl.seed_mode, wl.seed)
    render.render(buf, canvas)
    ctx.close()
    # the context's owner is gone; the objects still hold it
    want = canvas.load_image()
    del render  # releases its program
    buf.free()
    canvas.image.free()
    canvas.seeds.free()
    # a new context launches cleanly: no stale error from the releases above
    with Context(0, X32_F32) as ctx2:
        c2 = Canvas(ctx2, (32, 24))
        c2.passes = 1
        img = ImageBuffer(ctx2, (32, 24))
        Extractor(ctx2).process(c2, img)
        ctx2.sync()
    assert want.shape == (24, 32, 4) and np.all(want[..., 3] == 1.0)
