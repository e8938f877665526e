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


def test_scalar_bounds_program_shared_by_threads():
    """A scalar-bounds program keeps one bounds array, filled before each
    launch: two threads rendering different scenes of the same type through
    one program must each get the images of rendering alone."""
    import threading

    from hypertrace_amd.processing import program_options

    W, H, N = 64, 48, 24
    wls = [scenes.hyperbolic_cells(W, H, 1, primitives=192), scenes.hyperbolic_cells(W, H, 1, primitives=150)]
    with Context(0, X32_F32) as ctx:
        render = Render(ctx, wls[0].scene, fp_options=program_options(scalar_bounds=True))
        bufs = [EntityBuffer(ctx, wl.scene) for wl in wls]
        alone = []
        for wl, buf in zip(wls, bufs):
            c = Canvas(ctx, (W, H), wl.seed_mode, wl.seed)
            for _ in range(N):
                render.render_passes(buf, c, 1)
            alone.append(c.load_image().copy())
        shared = [Canvas(ctx, (W, H), wl.seed_mode, wl.seed) for wl in wls]
        errors = []

        def work(k):
            try:
                for _ in range(N):
                    render.render_passes(bufs[k], shared[k], 1)
            except Exception as e:  # reported below
                errors.append(e)

        threads = [threading.Thread(target=work, args=(k,)) for k in range(2)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        assert not errors
        for k in range(2):
            np.testing.assert_array_equal(shared[k].load_image().view(np.uint32), alone[k].view(np.uint32))
