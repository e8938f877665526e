"""The batched persistent render on the GPU: every job's canvas and seeds
bit-identical to rendering that job alone (its own program with HOPS =
depth), for several block and tile sizes; hipGraph replays identical to
eager frames; argument checks."""
import numpy as np
import pytest

from hypertrace_amd import scenes
import dataclasses

from hypertrace_amd.batch import (BLOCK_HIT_ORDER, SCALAR_BOUNDS, SCALAR_BOUNDS_PRIO, WAVE_QUEUE, WAVE_QUEUE_HIT_ORDER,
                                  WAVE_QUEUE_PRIO, Batch, BatchProgram, autotune)
from hypertrace_amd.native import HtError
from hypertrace_amd.objects import SceneImpl
from hypertrace_amd.processing import Canvas, Context, EntityBuffer, Render
from hypertrace_amd.types import X32_F32

pytestmark = pytest.mark.gpu

GROUPS4 = " -DHT_CULL_GROUPS=4"
SMALL = dict(n=8, min_size=(24, 16), max_size=(96, 72), spp_range=(1, 4), prim_range=(8, 64), depths=(4, 7, 12))


@pytest.fixture(scope="module")
def ctx():
    c = Context(0, X32_F32)
    yield c
    c.close()


def alone(ctx, spec, frames=1):
    """The job rendered by itself: its scene with HOPS = depth, render_passes."""
    s = spec.scene
    scene = SceneImpl(s.geo, s.view, s.object, s.background, spec.depth)
    r = Render(ctx, scene)
    buf = EntityBuffer(ctx, scene)
    c = Canvas(ctx, (spec.width, spec.height), 1, spec.seed)
    for _ in range(frames):
        r.render_passes(buf, c, spec.spp)
    return c.load_image(), c.load_seeds()


# (block, tile / block, program options) of the identity test
IDENTITY_CASES = [(256, 1, None), (64, 4, None), (1024, 1, None), (256, 1, SCALAR_BOUNDS), (1024, 2, SCALAR_BOUNDS),
                  (128, 1, WAVE_QUEUE), (256, 4, WAVE_QUEUE), (64, 1, WAVE_QUEUE), (256, 0.25, WAVE_QUEUE),
                  (256, 1, SCALAR_BOUNDS + GROUPS4), (128, 0.5, WAVE_QUEUE + GROUPS4),
                  (64, 2, WAVE_QUEUE + " -DHT_CULL_GROUPS=8"), (128, 1, SCALAR_BOUNDS + " -DHT_WAVE_PRIO=1"),
                  (128, 1, WAVE_QUEUE + " -DHT_WAVE_PRIO=1"), (64, 2, WAVE_QUEUE + GROUPS4 + " -DHT_WAVE_PRIO=1"),
                  # the group queue with the hit order (each job's leaves from its own array), helper lanes and
                  # sparse level 1, over wave tiles and block tiles; the group queue alone
                  (64, 1, WAVE_QUEUE_HIT_ORDER), (128, 2, WAVE_QUEUE_HIT_ORDER), (256, 0.25, WAVE_QUEUE_HIT_ORDER),
                  (128, 1, BLOCK_HIT_ORDER), (256, 1, WAVE_QUEUE + GROUPS4 + " -DHT_GROUP_QUEUE=1")]
GRAPH_OPTS = [None, SCALAR_BOUNDS, WAVE_QUEUE, WAVE_QUEUE + GROUPS4, WAVE_QUEUE_HIT_ORDER]  # 256-thread programs


@pytest.mark.parametrize("block,mult,opts", IDENTITY_CASES)
def test_batch_jobs_identical_to_single_renders(ctx, block, mult, opts):
    """Every job's canvas and seeds equal rendering it alone, with the bounds
    staged in LDS per tile or (scalar bounds) read from the job's own FP64
    array, and with each wave taking its own tiles and handing pixels to its
    lanes as they finish (wave queue)."""
    specs = scenes.batch_jobs(**SMALL)
    tile = int(block * mult)  # wave-queue tiles are per wave: a multiple of 64, below the block too
    print(f"case block {block} tile {tile} opts {opts}", flush=True)
    bprog = BatchProgram(ctx, block, fp_options=opts)
    batch = Batch(ctx, bprog, specs, tile)
    batch.render(grid=7, clear=False)
    ctx.sync()
    for k, spec in enumerate(specs):
        img, seeds = alone(ctx, spec)
        got = batch.canvases[k]
        np.testing.assert_array_equal(got.load_image().view(np.uint32), img.view(np.uint32), err_msg=f"job {k}")
        np.testing.assert_array_equal(got.load_seeds(), seeds, err_msg=f"job {k}")


@pytest.mark.parametrize("opts", GRAPH_OPTS)
def test_graph_replays_match_eager_frames(ctx, opts):
    specs = scenes.batch_jobs(**SMALL)
    bprog = BatchProgram(ctx, 256, fp_options=opts)
    eager = Batch(ctx, bprog, specs, with_rgba8=True)
    graph = Batch(ctx, bprog, specs, with_rgba8=True)
    graph.capture(grid=16, clear=True, post=True)
    other = scenes.euclidean_spheres(32, 24, 1)
    r_other = Render(ctx, other.scene)
    b_other = EntityBuffer(ctx, other.scene)
    c_other = Canvas(ctx, (32, 24), 1, 5)
    for frame in range(3):
        eager.render(grid=16, clear=True, post=True)
        graph.replay()
        r_other.render_passes(b_other, c_other, 2)  # unrelated eager work between replays
        ctx.sync()
        for k in range(len(specs)):
            a, b = eager.canvases[k].load_image(), graph.canvases[k].load_image()
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=f"frame {frame} job {k}")
            np.testing.assert_array_equal(eager.rgba8[k].load(), graph.rgba8[k].load())
            assert np.all(a[..., 3] == specs[k].spp)  # each frame starts from a cleared canvas


@pytest.mark.parametrize("opts", [None, WAVE_QUEUE])
def test_batch_jobs_with_zero_spp(ctx, opts):
    """Jobs with no samples leave their canvas and seeds as they were; the
    other jobs are unaffected."""
    specs = scenes.batch_jobs(**SMALL)
    specs = [dataclasses.replace(s, spp=0) if k in (0, 3) else s for k, s in enumerate(specs)]
    print(f"case zero spp opts {opts}", flush=True)
    bprog = BatchProgram(ctx, 256, fp_options=opts)
    batch = Batch(ctx, bprog, specs, 256)
    seeds0 = [c.load_seeds().copy() for c in batch.canvases]
    batch.render(grid=5, clear=False)
    ctx.sync()
    for k, spec in enumerate(specs):
        got = batch.canvases[k]
        if spec.spp == 0:
            assert np.all(got.load_image() == 0.0)
            np.testing.assert_array_equal(got.load_seeds(), seeds0[k])
        else:
            img, seeds = alone(ctx, spec)
            np.testing.assert_array_equal(got.load_image().view(np.uint32), img.view(np.uint32), err_msg=f"job {k}")
            np.testing.assert_array_equal(got.load_seeds(), seeds, err_msg=f"job {k}")


# Job sets at the edges of the culled scan: more elements than the batch
# program's capacity (256: the unculled scan), across the tree's 16-element
# threshold, and up to the capacity; (primitive range, program options, block).
EDGE_SETS = [((257, 300), None, 256), ((257, 300), WAVE_QUEUE + GROUPS4, 128), ((12, 20), WAVE_QUEUE + GROUPS4, 128),
             ((250, 256), SCALAR_BOUNDS + GROUPS4, 256), ((12, 20), WAVE_QUEUE_HIT_ORDER, 128),
             ((250, 256), WAVE_QUEUE_HIT_ORDER, 64), ((257, 300), WAVE_QUEUE_HIT_ORDER, 128)]


def edge_specs(prims):
    return scenes.batch_jobs(n=4, min_size=(24, 16), max_size=(48, 32), spp_range=(1, 2), prim_range=prims,
                             depths=(4, 7))


@pytest.mark.parametrize("prims,opts,block", EDGE_SETS)
def test_batch_element_counts_at_the_edges(ctx, prims, opts, block):
    specs = edge_specs(prims)
    print(f"case primitives {prims} opts {opts} block {block}", flush=True)
    bprog = BatchProgram(ctx, block, fp_options=opts)
    batch = Batch(ctx, bprog, specs, block)
    batch.render(grid=6, clear=False)
    ctx.sync()
    for k, spec in enumerate(specs):
        img, seeds = alone(ctx, spec)
        np.testing.assert_array_equal(batch.canvases[k].load_image().view(np.uint32), img.view(np.uint32),
                                      err_msg=f"job {k}")
        np.testing.assert_array_equal(batch.canvases[k].load_seeds(), seeds, err_msg=f"job {k}")


def test_batch_argument_checks(ctx):
    specs = scenes.batch_jobs(**SMALL)[:2]
    bprog = BatchProgram(ctx, 256)
    with pytest.raises(HtError, match="tile"):
        Batch(ctx, bprog, specs, 128)  # not a multiple of the block
    with pytest.raises(HtError, match="multiple of 64"):
        Batch(ctx, BatchProgram(ctx, 128, fp_options=WAVE_QUEUE), specs, 96)  # per-wave tiles: multiples of 64
    wl = scenes.euclidean_spheres(16, 16, 1)
    scene_prog = Render(ctx, wl.scene).program

    class NotBatch:
        program = scene_prog
        block = 256

    with pytest.raises(HtError, match="not a batch program"):
        Batch(ctx, NotBatch(), specs)


# ---- config 3 at its workload: 32 jobs, half Euclidean / half hyperbolic,
# 128x128..1024x768, spp 1..64, depth 4..12, 8..256 primitives
# The candidates the test's autotune picks from (the bench tunes over all).
CFG3_OPTIONS = ((WAVE_QUEUE_PRIO, (128,)), (SCALAR_BOUNDS_PRIO, (256,)), (WAVE_QUEUE_HIT_ORDER, (128,)))


def test_config3_32_jobs_identical_to_single_renders(ctx):
    """The 32-job frame of the batched workload in one persistent launch of
    an autotuned batch program: every job's canvas and seeds equal
    rendering that job alone with its own program."""
    specs = scenes.batch_jobs()
    assert len(specs) == 32
    tune = autotune(ctx, specs, blocks=(128, 256), tile_mults=(1,), grid_waves_per_cu=(8,), reps=1,
                    options=CFG3_OPTIONS)
    best = tune["best"]
    print(f"case config 3: {len(specs)} jobs, {tune['paths']} paths, best {best}", flush=True)
    batches = []
    # the tuner's pick, and the hit-order wave-queue program whatever it picks
    for block, opts, tile, grid in ((best["block"], best["options"], best["tile"], best["grid"]),
                                    (128, WAVE_QUEUE_HIT_ORDER, 128, 2048)):
        bprog = BatchProgram(ctx, block, fp_options=opts)
        batch = Batch(ctx, bprog, specs, tile)
        batch.render(grid, clear=True)
        batches.append(batch)
    ctx.sync()
    for k, spec in enumerate(specs):
        img, seeds = alone(ctx, spec)
        assert np.all(img[..., 3] == spec.spp)
        for batch in batches:
            got = batch.canvases[k]
            np.testing.assert_array_equal(got.load_image().view(np.uint32), img.view(np.uint32), err_msg=f"job {k}")
            np.testing.assert_array_equal(got.load_seeds(), seeds, err_msg=f"job {k}")


@pytest.mark.parametrize("opts", [WAVE_QUEUE_PRIO, WAVE_QUEUE_HIT_ORDER])
@pytest.mark.parametrize("geo_type", [0, 1])
def test_config3_one_pass_crop_vs_oracle(ctx, geo_type, opts):
    """One pass of the first Euclidean (0) or hyperbolic (1) job of the
    32-job set, rendered by a batch program, against the CPU oracle on a
    crop: 1e-4 relative on 99.9 % of pixels, seeds equal."""
    import oracle

    spec = next(s for s in scenes.batch_jobs() if s.type == geo_type)
    one = dataclasses.replace(spec, spp=1)
    print(f"case config 3 crop: type {geo_type}, {one.width}x{one.height}, depth {one.depth}, {opts}", flush=True)
    bprog = BatchProgram(ctx, 128, fp_options=opts)
    batch = Batch(ctx, bprog, [one], 128)
    seeds0 = batch.canvases[0].load_seeds().copy()
    batch.render(grid=64, clear=True)
    ctx.sync()
    got, got_seeds = batch.canvases[0].load_image(), batch.canvases[0].load_seeds()
    s = one.scene
    scene = SceneImpl(s.geo, s.view, s.object, s.background, one.depth)
    cw, ch = min(64, one.width), min(48, one.height)
    x0, y0 = (one.width - cw) // 2, (one.height - ch) // 2
    ref, ref_seeds, _ = oracle.render(scene.spec_text(X32_F32), one.width, one.height, 1, seeds0,
                                      crop=(x0, y0, cw, ch))
    sl = (slice(y0, y0 + ch), slice(x0, x0 + cw))
    ok = np.all(np.abs(got[sl] - ref[sl]) <= 1e-4 * np.abs(ref[sl]) + 1e-6, axis=-1)
    print(f"pixels within 1e-4: {ok.mean():.5f}, seeds equal {np.mean(got_seeds[sl] == ref_seeds[sl]):.5f}")
    assert ok.mean() >= 0.999
    assert np.mean(got_seeds[sl] == ref_seeds[sl]) >= 0.999
