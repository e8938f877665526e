"""Plugin entities on the GPU: a scene whose view, background, shapes,
materials and one object kind are plugins written against the reference's
device interface (tests/plugin_scenes.py) renders canvases and seeds bit for
bit equal to its built-in twin, in FP32 and FP64 programs, one pass per call
and many passes per launch; the built-in twin itself meets the oracle's
1e-4 / 99.9 % bar on a pass, so the plugin path is tied to the oracle."""
import numpy as np
import pytest

import oracle
import plugin_scenes as ps
from hypertrace_amd.processing import Canvas, Context, EntityBuffer, Render
from hypertrace_amd.types import X32_F32, X32_F64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[X32_F32, X32_F64], ids=["f32", "f64"])
def ctx(request):
    c = Context(0, request.param)
    yield c
    c.close()


def _render(ctx, scene, w, h, passes, batched):
    r = Render(ctx, scene)
    buf = EntityBuffer(ctx, scene)
    canvas = Canvas(ctx, (w, h), 1, 0x5EED)
    seeds0 = canvas.load_seeds().copy()
    if batched:
        r.render_passes(buf, canvas, passes)
    else:
        for _ in range(passes):
            r.render(buf, canvas)
    return seeds0, canvas.load_image(), canvas.load_seeds()


@pytest.mark.parametrize("frame", ps.FRAMES, ids=[f"{w}x{h}x{p}" for w, h, p in ps.FRAMES])
@pytest.mark.parametrize("batched", [False, True], ids=["per_pass", "passes"])
def test_plugin_scene_equals_builtin_twin(ctx, frame, batched):
    w, h, passes = frame
    built, plug = ps.scene_pair()
    print(f"case plugins f64={ctx.config.double_support} {w}x{h}x{passes} batched={batched}", flush=True)
    _, img_b, seeds_b = _render(ctx, built, w, h, passes, batched)
    _, img_p, seeds_p = _render(ctx, plug, w, h, passes, batched)
    assert np.all(np.isfinite(img_b))
    assert np.all(img_b[..., 3] == passes)
    lit = (img_b[..., :3] > 0).any(axis=-1).mean()
    assert lit > 0.5, f"only {lit:.2f} of pixels gathered light"
    assert np.array_equal(img_p.view(np.uint32), img_b.view(np.uint32))
    assert np.array_equal(seeds_p, seeds_b)


def test_builtin_twin_meets_the_oracle(ctx):
    w, h, _ = ps.FRAMES[0]
    built, _ = ps.scene_pair()
    print(f"case plugin twin vs oracle f64={ctx.config.double_support}", flush=True)
    seeds0, got, got_seeds = _render(ctx, built, w, h, 1, False)
    precision = oracle.F64 if ctx.config.double_support else oracle.F32
    ref, ref_seeds, _ = oracle.render(built.spec_text(ctx.config), w, h, 1, seeds0, precision=precision)
    ok = np.all(np.abs(got - ref) <= 1e-4 * np.abs(ref) + 1e-6, axis=-1)
    assert ok.mean() >= 0.999, f"only {ok.mean():.5f} of pixels within 1e-4"
    assert (got_seeds[ok] == ref_seeds[ok]).mean() >= 0.999
