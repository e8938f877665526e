"""The wavefront form (ht_render_wavefront: SoA path state in device memory,
extend / material-sorted shade kernels, regeneration) against the
megakernel's render_passes on the same scene, canvas and seeds: both must
leave the canvas and the seeds bit-identical, across scene families,
program options, image sizes that are not multiples of a wave, and a
second call that continues the same canvas."""
import numpy as np
import pytest

from hypertrace_amd import scenes
from hypertrace_amd.processing import Canvas, Context, EntityBuffer, Render, program_options
from hypertrace_amd.types import X32_F32

pytestmark = pytest.mark.gpu


def _batch_scene(kind):
    rng = np.random.default_rng(11)
    return scenes.batch_euclidean_scene(rng, 40) if kind == 0 else scenes.batch_hyperbolic_scene(rng, 45)


def _wl(wl):
    return wl.scene, wl.width, wl.height, wl.seed_mode, wl.seed


CASES = {
    # name: (() -> (scene, width, height, seed mode, seed), program options)
    "eu_ball": (lambda: _wl(scenes.euclidean_spheres(96, 72)), None),
    "eu_example": (lambda: _wl(scenes.reference_euclidean_example(77, 51)), None),
    "hy_example": (lambda: _wl(scenes.reference_hyperbolic_example(77, 51)), None),
    "hy_cells": (lambda: _wl(scenes.hyperbolic_cells(96, 72)), None),
    "hy_cells_scalar_bounds": (lambda: _wl(scenes.hyperbolic_cells(96, 72)), program_options(scalar_bounds=True)),
    "sdf_rr": (lambda: _wl(scenes.sdf_shapes(96, 72, 1, shapes=48, extent=4.0, camera_z=12.0)),
               program_options(rr_depth=3)),
    "batch_eu_glass": (lambda: (_batch_scene(0), 101, 67, 1, 21), None),
    "batch_hy_glass": (lambda: (_batch_scene(1), 101, 67, 1, 22), None),
}


@pytest.mark.parametrize("name", list(CASES))
def test_wavefront_identical_to_render_passes(ctx, name):
    make, opts = CASES[name]
    scene, w, h, mode, seed = make()
    print(f"case {name} {w}x{h} opts={opts}", flush=True)
    buf = EntityBuffer(ctx, scene)
    mega = Render(ctx, scene, fp_options=opts)
    wave = Render(ctx, scene, fp_options=opts, wavefront=True)
    out = []
    for run in (lambda c, k: mega.render_passes(buf, c, k), lambda c, k: wave.render_wavefront(buf, c, k)):
        c = Canvas(ctx, (w, h), mode, seed)
        run(c, 5)
        run(c, 3)  # continues the same canvas and seeds (and reuses the workspace)
        ctx.sync()
        out.append((c.load_image().copy(), c.load_seeds().copy()))
    np.testing.assert_array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert np.all(out[1][0][..., 3] == 8.0)


def test_wavefront_needs_its_kernels(ctx):
    from hypertrace_amd.native import HtError

    wl = scenes.euclidean_spheres(16, 8, 1)
    buf = EntityBuffer(ctx, wl.scene)
    c = Canvas(ctx, (16, 8), wl.seed_mode, wl.seed)
    with pytest.raises(HtError, match="without the wavefront kernels"):
        Render(ctx, wl.scene).render_wavefront(buf, c, 1)


@pytest.fixture(scope="module")
def ctx():
    with Context(0, X32_F32) as c:
        yield c



def test_wavefront_fp64_program():
    """FP64 programs (DOUBLE_PRECISION: real = double, FP64 hit keys) through
    the wavefront form give the same canvas and seeds as render_passes."""
    from hypertrace_amd.types import X32_F64

    wl = scenes.hyperbolic_cells(80, 60)
    with Context(0, X32_F64) as c64:
        buf = EntityBuffer(c64, wl.scene)
        mega = Render(c64, wl.scene)
        wave = Render(c64, wl.scene, wavefront=True)
        out = []
        for run in (lambda c: mega.render_passes(buf, c, 4), lambda c: wave.render_wavefront(buf, c, 4)):
            c = Canvas(c64, (80, 60), wl.seed_mode, wl.seed)
            run(c)
            c64.sync()
            out.append((c.load_image().copy(), c.load_seeds().copy()))
    np.testing.assert_array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    np.testing.assert_array_equal(out[0][1], out[1][1])
