"""GPU parity of the render pass against the CPU oracle, through the C ABI.

Bars (per pixel, per channel): |gpu - oracle| <= 1e-4 |oracle| + 1e-6 on at
least 99.9 % of pixels for a single-sample pass with the identical seeded RNG
stream (grazing branch flips are the allowed remainder); bit-exact for the
RNG state where the path took the same branches, for the pass-batching
kernel versus repeated single passes, for the post passes and the seeds.
"""
import numpy as np
import pytest

import oracle
from hypertrace_amd import scenes
from hypertrace_amd.processing import (NO_CULL_FP, STRICT_FP, Canvas, Context, EntityBuffer, Extractor, GammaFilter, ImageBuffer, Packer,
                                       Pipeline, Render, program_options, seed_fill)
from hypertrace_amd.types import X32_F32, X32_F64

pytestmark = pytest.mark.gpu

REL = 1e-4
ABS = 1e-6


@pytest.fixture(scope="module")
def ctx():
    c = Context(0, X32_F32)
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx64():
    c = Context(0, X32_F64)
    yield c
    c.close()


def within(got, ref):
    return np.all(np.abs(got - ref) <= REL * np.abs(ref) + ABS, axis=-1)


def one_pass(ctx, wl, passes=1, fp_options=None):
    render = Render(ctx, wl.scene, fp_options=fp_options)
    buf = EntityBuffer(ctx, wl.scene)
    canvas = Canvas(ctx, (wl.width, wl.height), wl.seed_mode, wl.seed)
    seeds0 = canvas.load_seeds().copy()
    for _ in range(passes):
        render.render(buf, canvas)
    return seeds0, canvas.load_image(), canvas.load_seeds()


WORKLOADS = {
    "eu_ball": lambda: scenes.euclidean_spheres(96, 72),
    "eu_unit": lambda: scenes.euclidean_spheres(96, 72, unit_spheres=True),
    "hy_cells": lambda: scenes.hyperbolic_cells(96, 72),
    "eu_example": lambda: scenes.reference_euclidean_example(96, 72),
    "hy_example": lambda: scenes.reference_hyperbolic_example(96, 72),
    "sdf": lambda: scenes.sdf_shapes(96, 72, 1, shapes=48, extent=4.0, camera_z=12.0),
    "sdf_1024": lambda: scenes.sdf_shapes(320, 240, 1),
}


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_single_pass_parity_f32(ctx, name):
    wl = WORKLOADS[name]()
    print(f"case {name} {wl.width}x{wl.height}", flush=True)
    seeds0, got, got_seeds = one_pass(ctx, wl)
    ref, ref_seeds, _ = oracle.render(wl.scene.spec_text(X32_F32), wl.width, wl.height, 1, seeds0)
    ok = within(got, ref)
    assert ok.mean() >= 0.999, f"{name}: only {ok.mean():.5f} of pixels within 1e-4"
    # Alpha counts passes exactly.
    assert np.all(got[..., 3] == 1.0)
    # Where the radiance agrees the path made the same choices: same RNG state.
    assert (got_seeds[ok] == ref_seeds[ok]).mean() >= 0.999


@pytest.mark.parametrize("name", ["eu_unit", "hy_cells", "eu_example", "hy_example", "sdf"])
def test_single_pass_parity_f64(ctx64, name):
    wl = WORKLOADS[name]()
    print(f"case {name} f64", flush=True)
    seeds0, got, got_seeds = one_pass(ctx64, wl)
    ref, ref_seeds, _ = oracle.render(wl.scene.spec_text(X32_F64), wl.width, wl.height, 1, seeds0,
                                      precision=oracle.F64)
    ok = within(got, ref)
    assert ok.mean() >= 0.999, f"{name}: only {ok.mean():.5f} of pixels within 1e-4"


HY_DIST_FP64_CASES = ["hy_cells", "hy_example"]


@pytest.mark.parametrize("name", HY_DIST_FP64_CASES)
def test_hy_distance_fp64_option(ctx, name):
    """Program option HT_HY_DIST_FP64 (the hyperbolic distance in FP64 with a
    cancellation-free acosh; a precision method, not the drop-in). Against
    the FP64 oracle with the same seeds its single-sample radiance agrees at
    least as often as the reference's FP32 formula does, and it stays with
    the FP32 oracle on nearly every pixel (the distance only decides near
    ties and short hops)."""
    from hypertrace_amd.processing import program_options

    wl = WORKLOADS[name]()
    print(f"case {name} hy_distance_fp64", flush=True)
    seeds0, base, _ = one_pass(ctx, wl)
    _, opt, _ = one_pass(ctx, wl, fp_options=program_options(hy_distance_fp64=True))
    ref32, _, _ = oracle.render(wl.scene.spec_text(X32_F32), wl.width, wl.height, 1, seeds0)
    ref64, _, _ = oracle.render(wl.scene.spec_text(X32_F64), wl.width, wl.height, 1, seeds0, precision=oracle.F64)
    base64, opt64, opt32 = within(base, ref64).mean(), within(opt, ref64).mean(), within(opt, ref32).mean()
    print(f"within 1e-4 of FP64: reference formula {base64:.5f}, FP64 distance {opt64:.5f}; "
          f"FP64 distance within 1e-4 of the FP32 oracle {opt32:.5f}")
    assert opt64 >= base64 - 0.002
    assert opt32 >= 0.98


@pytest.mark.parametrize("name", ["eu_ball", "hy_cells"])
def test_render_passes_identical_to_repeated_render(ctx, name):
    wl = WORKLOADS[name]()
    print(f"case {name} passes", flush=True)
    k = 5
    _, img_a, seeds_a = one_pass(ctx, wl, passes=k)
    render = Render(ctx, wl.scene)
    buf = EntityBuffer(ctx, wl.scene)
    canvas = Canvas(ctx, (wl.width, wl.height), wl.seed_mode, wl.seed)
    render.render_passes(buf, canvas, k)
    img_b, seeds_b = canvas.load_image(), canvas.load_seeds()
    assert canvas.passes == k
    np.testing.assert_array_equal(img_a.view(np.uint32), img_b.view(np.uint32))
    np.testing.assert_array_equal(seeds_a, seeds_b)


@pytest.mark.parametrize("name", ["sdf", "eu_ball"])
def test_russian_roulette_parity(ctx, name):
    """Program option HT_RR_DEPTH=3 against the oracle's `rr 3` scene prefix:
    same survival draws and throughput rescaling, per pixel."""
    wl = WORKLOADS[name]()
    print(f"case {name} rr", flush=True)
    render = Render(ctx, wl.scene, fp_options=program_options(rr_depth=3))
    buf = EntityBuffer(ctx, wl.scene)
    canvas = Canvas(ctx, (wl.width, wl.height), wl.seed_mode, wl.seed)
    seeds0 = canvas.load_seeds().copy()
    render.render_passes(buf, canvas, 2)
    got = canvas.load_image()
    ref, _, _ = oracle.render("rr 3 " + wl.scene.spec_text(X32_F32), wl.width, wl.height, 2, seeds0)
    ok = within(got, ref)
    assert ok.mean() >= 0.995, f"{name}: only {ok.mean():.5f} of pixels within 1e-4"


def test_multi_pass_parity_hyperbolic(ctx):
    wl = scenes.hyperbolic_cells(64, 48)
    seeds0 = seed_fill(wl.width * wl.height, wl.seed_mode, wl.seed)
    render = Render(ctx, wl.scene)
    buf = EntityBuffer(ctx, wl.scene)
    canvas = Canvas(ctx, (wl.width, wl.height), wl.seed_mode, wl.seed)
    render.render_passes(buf, canvas, 8)
    got = canvas.load_image()
    ref, _, _ = oracle.render(wl.scene.spec_text(X32_F32), wl.width, wl.height, 8, seeds0)
    # Once a path flips, later samples of that pixel use a shifted RNG
    # stream, so the multi-pass bar is looser than the single-pass one.
    ok = within(got, ref)
    assert ok.mean() >= 0.99


def test_seed_init_matches_host(ctx):
    for mode, seed in ((0, 0xDEADBEEFDEADBEEF), (1, 0x48595054 ^ 2), (2, 0xDEADBEEFDEADBEEF)):
        c = Canvas(ctx, (37, 11), mode, seed)
        np.testing.assert_array_equal(c.load_seeds().reshape(-1), seed_fill(37 * 11, mode, seed))


def test_post_passes_bitexact(ctx):
    rng = np.random.default_rng(5)
    w, h, passes = 53, 17, 7
    canvas_host = (rng.uniform(-0.2, 1.6, (h, w, 4)) * passes).astype(np.float32)
    canvas_host[..., 3] = passes
    canvas_host[0, 0, 0] = np.nan
    canvas_host[0, 1, 1] = np.inf
    c = Canvas(ctx, (w, h))
    c.image.store(canvas_host)
    c.passes = passes
    img = ImageBuffer(ctx, (w, h))
    out = ImageBuffer(ctx, (w, h))
    packed = ImageBuffer(ctx, (w, h), 4, np.uint8)
    Extractor(ctx).process(c, img)
    GammaFilter(ctx, 1 / 2.2).process(img, out)
    Packer(ctx).process(out, packed)
    extracted, gammaed, got = img.load(), out.load(), packed.load()
    # extract: IEEE division, bit-exact
    np.testing.assert_array_equal(extracted.view(np.uint32), (canvas_host / np.float32(passes)).view(np.uint32))
    # gamma: pow within 2 ulp of the CPU oracle's
    ref_img, _ = oracle.post(canvas_host, passes, np.float32(1 / 2.2))
    finite = np.isfinite(ref_img)
    assert np.all(np.abs(gammaed[finite] - ref_img[finite]) <= 2 * np.spacing(np.abs(ref_img[finite])))
    # pack: bit-exact on the same float input (trunc(255 * clamp(x, 0, 1)), NaN -> 0)
    ref_pack = (np.float32(255.0) * np.fmin(np.fmax(gammaed, np.float32(0)), np.float32(1))).astype(np.uint8)
    np.testing.assert_array_equal(got, ref_pack)


def test_pipeline_fused_post_equals_separate(ctx):
    wl = scenes.euclidean_spheres(64, 48)
    p = Pipeline(ctx, (wl.width, wl.height), wl.scene, seed_mode=wl.seed_mode, seed=wl.seed)
    for _ in range(3):
        p.render()
    a = p.get_image().copy()
    b = p.get_image(fused=True).copy()
    np.testing.assert_array_equal(a, b)
    assert np.all(a[..., 3] == 255)


NO_CULL = NO_CULL_FP


# Program options of the culled scan's variants (all must render the same
# images): the default flat scan over LDS bounds, the sphere tree, a render
# tuner setting that spills, and the scalar-cache bounds.
CULLED_VARIANTS = {
    "flat": None,
    "tree": STRICT_FP + " -DHT_CULL_TREE=1",
    "unroll2_waves5": STRICT_FP + " -DHT_CULL_UNROLL=2 -DHT_WAVES_PER_EU=5",
    "scalar_bounds": STRICT_FP + " -DHT_SCALAR_BOUNDS=1",
    "scalar_bounds_u2_w5": STRICT_FP + " -DHT_CULL_UNROLL=2 -DHT_WAVES_PER_EU=5 -DHT_SCALAR_BOUNDS=1",
    "groups4": STRICT_FP + " -DHT_CULL_GROUPS=4",
    "groups8_w5": STRICT_FP + " -DHT_CULL_GROUPS=8 -DHT_WAVES_PER_EU=5",
    "pixel_queue": STRICT_FP + " -DHT_PIXEL_QUEUE=1",
    "pixel_queue_sb_w6": STRICT_FP + " -DHT_PIXEL_QUEUE=1 -DHT_SCALAR_BOUNDS=1 -DHT_WAVES_PER_EU=6",
    "sb_groups4": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4",
    "sb_groups8_w6_q": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=8 -DHT_WAVES_PER_EU=6 -DHT_PIXEL_QUEUE=1",
    "sb_prio_w5": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_WAVES_PER_EU=5 -DHT_WAVE_PRIO=1",
    "prio": STRICT_FP + " -DHT_WAVE_PRIO=1",
    "sb_q_prio": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_PIXEL_QUEUE=1 -DHT_WAVE_PRIO=1",
    "sb_g4_q_prio": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_PIXEL_QUEUE=1 -DHT_WAVE_PRIO=1",
    "gq4_prio": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1 -DHT_WAVE_PRIO=1",
    "gq8": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=8 -DHT_GROUP_QUEUE=1",
    "gq4_q_prio_w6": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1 -DHT_PIXEL_QUEUE=1"
                     " -DHT_WAVE_PRIO=1 -DHT_WAVES_PER_EU=6",
    "gq16_u2": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=16 -DHT_GROUP_QUEUE=1 -DHT_CULL_UNROLL=2",
    "gq4_ho_prio": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1 -DHT_WAVE_PRIO=1"
                   " -DHT_HIT_ORDER=1",
    "gq4_ho_w3": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1 -DHT_WAVE_PRIO=1"
                 " -DHT_HIT_ORDER=1 -DHT_WAVES_PER_EU=3",
    "gq4_ho_q_prio": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1 -DHT_PIXEL_QUEUE=1"
                     " -DHT_WAVE_PRIO=1 -DHT_HIT_ORDER=1",
    "gq2_ho_u2": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=2 -DHT_GROUP_QUEUE=1 -DHT_CULL_UNROLL=2"
                 " -DHT_HIT_ORDER=1",
    "gq4_ho_hl": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1 -DHT_WAVE_PRIO=1"
                 " -DHT_HIT_ORDER=1 -DHT_HELPERS=1",
    "gq4_hl_w3": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1 -DHT_HELPERS=1"
                 " -DHT_WAVES_PER_EU=3",
    "sb_hl": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_HELPERS=1 -DHT_WAVE_PRIO=1",
    "hl": STRICT_FP + " -DHT_HELPERS=1",
    # level 1 over (group, ray) pairs in waves with few rays (HT_SPARSE_L1):
    # with helper lanes (all 64 lanes in the scan), with the pixel queue
    # (only the lanes holding a pixel), and always on (64)
    "gq4_ho_hl_s32": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1 -DHT_WAVE_PRIO=1"
                     " -DHT_HIT_ORDER=1 -DHT_HELPERS=1 -DHT_SPARSE_L1=32",
    "gq4_ho_q_s16": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1 -DHT_PIXEL_QUEUE=1"
                    " -DHT_WAVE_PRIO=1 -DHT_HIT_ORDER=1 -DHT_SPARSE_L1=16",
    "gq4_s64": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1 -DHT_SPARSE_L1=64",
    "gq8_s8_u2": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=8 -DHT_GROUP_QUEUE=1 -DHT_CULL_UNROLL=2"
                 " -DHT_SPARSE_L1=8",
    # the pixel queue with helper lanes (the lanes without a pixel once the
    # queue is empty join the scans)
    "gq4_ho_q_hl_s16": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1 -DHT_PIXEL_QUEUE=1"
                       " -DHT_WAVE_PRIO=1 -DHT_HIT_ORDER=1 -DHT_HELPERS=1 -DHT_SPARSE_L1=16",
    "gq4_q_hl": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1 -DHT_PIXEL_QUEUE=1"
                " -DHT_HELPERS=1",
    "q_hl": STRICT_FP + " -DHT_PIXEL_QUEUE=1 -DHT_HELPERS=1",
    # pixels handed out in a spread order (HT_PIXEL_SPREAD), statically or
    # from the pixel queue: 1 = the golden-ratio order, B = 2..64 = one run
    # of 64 / B pixels from each of B bands per wave
    "gq4_ho_q_hl_s32_sp1": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1"
                           " -DHT_PIXEL_QUEUE=1 -DHT_WAVE_PRIO=1 -DHT_HIT_ORDER=1 -DHT_HELPERS=1"
                           " -DHT_SPARSE_L1=32 -DHT_PIXEL_SPREAD=1",
    "q_hl_sp1": STRICT_FP + " -DHT_PIXEL_QUEUE=1 -DHT_HELPERS=1 -DHT_PIXEL_SPREAD=1",
    "sp1_w3": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_PIXEL_SPREAD=1 -DHT_WAVES_PER_EU=3",
    # the queue hands out the pixels whose probe ray hits the scene first
    # (HT_PIXEL_CLASSES, with the golden-ratio order)
    "gq4_ho_q_hl_s32_sp1_pc": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1"
                              " -DHT_PIXEL_QUEUE=1 -DHT_WAVE_PRIO=1 -DHT_HIT_ORDER=1 -DHT_HELPERS=1"
                              " -DHT_SPARSE_L1=32 -DHT_PIXEL_SPREAD=1 -DHT_PIXEL_CLASSES=1",
    "q_sp1_pc": STRICT_FP + " -DHT_PIXEL_QUEUE=1 -DHT_PIXEL_SPREAD=1 -DHT_PIXEL_CLASSES=1",
    "gq4_ho_q_hl_s32_sp64": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1"
                            " -DHT_PIXEL_QUEUE=1 -DHT_WAVE_PRIO=1 -DHT_HIT_ORDER=1 -DHT_HELPERS=1"
                            " -DHT_SPARSE_L1=32 -DHT_PIXEL_SPREAD=64",
    "gq4_ho_hl_s32_sp8_w3": STRICT_FP + " -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1"
                            " -DHT_WAVE_PRIO=1 -DHT_HIT_ORDER=1 -DHT_HELPERS=1 -DHT_SPARSE_L1=32"
                            " -DHT_PIXEL_SPREAD=8 -DHT_WAVES_PER_EU=3",
    "q_hl_sp4": STRICT_FP + " -DHT_PIXEL_QUEUE=1 -DHT_HELPERS=1 -DHT_PIXEL_SPREAD=4",
    "sp64": STRICT_FP + " -DHT_PIXEL_SPREAD=64",
}


@pytest.mark.parametrize("name,full,variant", [
    ("eu_ball", False, "flat"), ("eu_unit", False, "flat"), ("hy_cells", False, "flat"), ("hy_cells", True, "flat"),
    ("eu_example", False, "flat"), ("hy_example", False, "flat"), ("sdf", False, "flat"),
    ("sdf_1024", False, "flat"), ("hy_cells", False, "tree"), ("sdf_1024", False, "tree"),
    ("hy_cells", True, "unroll2_waves5"), ("eu_ball", False, "scalar_bounds"), ("hy_cells", True, "scalar_bounds"),
    ("hy_example", False, "scalar_bounds"), ("sdf_1024", False, "scalar_bounds"),
    ("hy_cells", False, "scalar_bounds_u2_w5"), ("hy_cells", True, "groups4"), ("sdf_1024", False, "groups4"),
    ("hy_cells", False, "groups8_w5"), ("hy_cells", True, "pixel_queue"), ("eu_ball", False, "pixel_queue"),
    ("sdf_1024", False, "pixel_queue_sb_w6"), ("hy_example", False, "pixel_queue_sb_w6"),
    ("hy_cells", True, "sb_groups4"), ("sdf_1024", False, "sb_groups4"), ("eu_ball", False, "sb_groups4"),
    ("hy_cells", False, "sb_groups8_w6_q"), ("sdf_1024", False, "sb_groups8_w6_q"),
    ("hy_example", False, "sb_groups8_w6_q"), ("hy_cells", True, "sb_prio_w5"), ("sdf_1024", False, "sb_prio_w5"),
    ("eu_ball", False, "prio"), ("hy_example", False, "prio"), ("hy_cells", True, "sb_q_prio"),
    ("sdf_1024", False, "sb_g4_q_prio"), ("hy_cells", False, "sb_g4_q_prio"),
    ("hy_cells", True, "gq4_prio"), ("hy_cells", False, "gq8"), ("sdf_1024", False, "gq4_prio"),
    ("sdf_1024", False, "gq8"), ("eu_ball", False, "gq4_prio"), ("hy_example", False, "gq8"),
    ("eu_example", False, "gq4_prio"), ("hy_cells", True, "gq4_q_prio_w6"), ("sdf_1024", False, "gq4_q_prio_w6"),
    ("hy_cells", False, "gq16_u2"), ("sdf", False, "gq8"),
    ("hy_cells", True, "gq4_ho_prio"), ("hy_cells", False, "gq4_ho_w3"), ("hy_example", False, "gq4_ho_prio"),
    ("eu_ball", False, "gq4_ho_prio"), ("sdf_1024", False, "gq4_ho_w3"), ("hy_cells", True, "gq4_ho_q_prio"),
    ("eu_example", False, "gq4_ho_w3"), ("hy_cells", False, "gq2_ho_u2"),
    ("hy_cells", True, "gq4_ho_hl"), ("hy_cells", False, "gq4_hl_w3"), ("hy_example", False, "gq4_ho_hl"),
    ("sdf_1024", False, "gq4_ho_hl"), ("eu_ball", False, "sb_hl"), ("eu_example", False, "hl"),
    ("sdf", False, "hl"), ("eu_unit", False, "gq4_ho_hl"),
    ("hy_cells", True, "gq4_ho_hl_s32"), ("hy_cells", True, "gq4_ho_q_s16"), ("hy_cells", False, "gq4_s64"),
    ("sdf_1024", False, "gq4_s64"), ("sdf_1024", False, "gq4_ho_hl_s32"), ("eu_ball", False, "gq4_ho_q_s16"),
    ("hy_example", False, "gq4_ho_q_s16"), ("eu_example", False, "gq8_s8_u2"), ("hy_cells", False, "gq8_s8_u2"),
    ("hy_cells", True, "gq4_ho_q_hl_s16"), ("sdf_1024", False, "gq4_q_hl"), ("eu_ball", False, "gq4_ho_q_hl_s16"),
    ("hy_example", False, "gq4_q_hl"), ("sdf", False, "q_hl"), ("eu_example", False, "q_hl"),
    ("hy_cells", True, "gq4_ho_q_hl_s32_sp64"), ("hy_cells", False, "gq4_ho_hl_s32_sp8_w3"),
    ("sdf_1024", False, "gq4_ho_q_hl_s32_sp64"), ("eu_ball", False, "q_hl_sp4"), ("hy_example", False, "sp64"),
    ("hy_cells", True, "gq4_ho_q_hl_s32_sp1"), ("sdf_1024", False, "gq4_ho_q_hl_s32_sp1"),
    ("eu_example", False, "q_hl_sp1"), ("hy_cells", False, "sp1_w3"),
    ("hy_cells", True, "gq4_ho_q_hl_s32_sp1_pc"), ("sdf_1024", False, "gq4_ho_q_hl_s32_sp1_pc"),
    ("hy_example", False, "q_sp1_pc"), ("eu_ball", False, "q_sp1_pc")])
def test_culled_scan_identical_to_plain_scan(ctx, name, full, variant):
    """The culled, compacted nearest-hit scan (flat over LDS or scalar-cache
    bounds, or walking the sphere tree with -DHT_CULL_TREE) must give
    bit-identical canvas and seeds to the element-by-element scan."""
    wl = scenes.hyperbolic_cells() if full else WORKLOADS[name]()
    print(f"case {name} full={full} variant={variant}", flush=True)
    buf = EntityBuffer(ctx, wl.scene)
    out = []
    culled = CULLED_VARIANTS[variant]
    for fp in (culled, NO_CULL):
        r = Render(ctx, wl.scene, fp_options=fp)
        c = Canvas(ctx, (wl.width, wl.height), wl.seed_mode, wl.seed)
        r.render_passes(buf, c, 4)
        r.render(buf, c)
        out.append((c.load_image(), c.load_seeds()))
    np.testing.assert_array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    np.testing.assert_array_equal(out[0][1], out[1][1])


HELPER_RR_OPTIONS = {"hl": " -DHT_HELPERS=1",
                     "q_gq_hl_s16": " -DHT_PIXEL_QUEUE=1 -DHT_SCALAR_BOUNDS=1 -DHT_CULL_GROUPS=4 -DHT_GROUP_QUEUE=1"
                                    " -DHT_HELPERS=1 -DHT_SPARSE_L1=16"}


@pytest.mark.parametrize("helpers", list(HELPER_RR_OPTIONS))
@pytest.mark.parametrize("name", ["sdf_1024", "eu_ball"])
def test_helpers_with_russian_roulette(ctx, name, helpers):
    """Helper lanes (HT_HELPERS; also with the pixel queue, whose lanes
    without a pixel help) with Russian roulette, where a path can end before
    its scan: the same canvas and seeds as the program without helpers."""
    wl = WORKLOADS[name]()
    print(f"case helpers + roulette {name} {helpers}", flush=True)
    buf = EntityBuffer(ctx, wl.scene)
    out = []
    for fp in (program_options(rr_depth=3) + HELPER_RR_OPTIONS[helpers], program_options(rr_depth=3)):
        r = Render(ctx, wl.scene, fp_options=fp)
        c = Canvas(ctx, (wl.width, wl.height), wl.seed_mode, wl.seed)
        r.render_passes(buf, c, 6)
        out.append((c.load_image(), c.load_seeds()))
    np.testing.assert_array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    np.testing.assert_array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("variant", ["gq4_ho_prio", "gq4_ho_w3", "gq4_ho_hl", "gq4_ho_hl_s32", "gq4_ho_q_s16",
                                     "gq4_ho_q_hl_s16", "gq4_ho_q_hl_s32_sp64", "gq4_ho_q_hl_s32_sp1",
                                     "gq4_ho_q_hl_s32_sp1_pc"])
@pytest.mark.parametrize("frame", ["640x480", "1920x1080"])
def test_hit_order_identical_on_full_frames(ctx, frame, variant):
    """The hit-order scan (each passing pair's crossing interval: pairs that
    never enter the element's sphere ahead of the start are dropped, pairs
    beyond the ray's nearest certain hit skipped) on the benchmark frames:
    the 192-primitive cell scene at 640x480 over 16 passes and the
    256-primitive scene at 1920x1080 over 3 passes, canvas and seeds equal
    to the element-by-element scan's."""
    if frame == "640x480":
        wl, passes = scenes.hyperbolic_cells(), 16
    else:
        wl, passes = scenes.hyperbolic_cells(1920, 1080, 1, primitives=256), 3
    print(f"case hit order {variant} {frame} {passes} passes", flush=True)
    buf = EntityBuffer(ctx, wl.scene)
    out = []
    for fp in (CULLED_VARIANTS[variant], NO_CULL):
        r = Render(ctx, wl.scene, fp_options=fp)
        c = Canvas(ctx, (wl.width, wl.height), wl.seed_mode, wl.seed)
        r.render_passes(buf, c, passes)
        out.append((c.load_image(), c.load_seeds()))
    np.testing.assert_array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    np.testing.assert_array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("variant", ["pixel_queue", "sb_q_prio", "gq4_q_prio_w6", "gq4_ho_q_prio", "gq4_ho_hl",
                                     "gq4_ho_q_s16", "gq4_ho_q_hl_s16", "q_hl", "gq4_ho_q_hl_s32_sp64", "sp64",
                                     "q_hl_sp4", "gq4_ho_hl_s32_sp8_w3", "gq4_ho_q_hl_s32_sp1", "q_hl_sp1",
                                     "sp1_w3", "gq4_ho_q_hl_s32_sp1_pc", "q_sp1_pc"])
@pytest.mark.parametrize("size", [(1, 1), (7, 3), (65, 1), (17, 19), (257, 5)])
def test_pixel_queue_ragged_frames(ctx, size, variant):
    """The pixel-queue and spread-order programs on frames smaller than a
    block or with a partial last wave (the spread order's padding items take
    no pixel), launched twice (the counter is reset before every launch):
    the same canvas and seeds as the one-thread-per-pixel kernel."""
    w, h = size
    wl = scenes.euclidean_spheres(w, h)
    print(f"case {variant} {w}x{h}", flush=True)
    buf = EntityBuffer(ctx, wl.scene)
    out = []
    for fp in (CULLED_VARIANTS[variant], None):
        r = Render(ctx, wl.scene, fp_options=fp)
        c = Canvas(ctx, (w, h), wl.seed_mode, wl.seed)
        r.render_passes(buf, c, 3)
        r.render_passes(buf, c, 2)
        out.append((c.load_image(), c.load_seeds()))
    np.testing.assert_array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert np.all(out[0][0][..., 3] == 5.0)


def test_fp32_hit_window_option(ctx, ctx64):
    """-DHT_EPS32=1e-5f: the culled scan equals the element-by-element scan
    under the same window (bit for bit), and the FP32 image of the Euclidean
    sphere scene moves to the FP64 one. With the reference's window (1e-6)
    FP32 hit points at world scale ~8 round by about the window's size, and
    the converged FP32 image sits ~0.01 RMSE from the FP64 image (measured on
    this scene with the CPU oracle too); FP64 decides."""
    wl = scenes.euclidean_spheres(96, 72)
    print("case fp32 hit window 1e-5, euclidean spheres 96x72", flush=True)
    spp = 256
    buf = EntityBuffer(ctx, wl.scene)
    imgs = {}
    for name, fp in (("ref", None), ("wide", program_options(eps32=1e-5)),
                     ("wide_plain", program_options(eps32=1e-5, no_cull=True))):
        r = Render(ctx, wl.scene, fp_options=fp)
        c = Canvas(ctx, (wl.width, wl.height), wl.seed_mode, wl.seed)
        r.render_passes(buf, c, spp)
        imgs[name] = (c.load_image(), c.load_seeds())
    np.testing.assert_array_equal(imgs["wide"][0].view(np.uint32), imgs["wide_plain"][0].view(np.uint32))
    np.testing.assert_array_equal(imgs["wide"][1], imgs["wide_plain"][1])
    r64 = Render(ctx64, wl.scene)
    b64 = EntityBuffer(ctx64, wl.scene)
    c64 = Canvas(ctx64, (wl.width, wl.height), wl.seed_mode, wl.seed)
    r64.render_passes(b64, c64, spp)
    ref64 = c64.load_image()

    def img(c):
        c = c.astype(np.float64)
        return np.clip(np.maximum(c[..., :3] / c[..., 3:4], 0.0) ** (1 / 2.2), 0.0, 1.0)

    def dist(a):
        return float(np.sqrt(np.mean((img(a) - img(ref64)) ** 2)))

    d_ref, d_wide = dist(imgs["ref"][0]), dist(imgs["wide"][0])
    print(f"rmse to fp64 (same seeds, {spp} spp): reference window {d_ref:.5f}, 1e-5 window {d_wide:.5f}", flush=True)
    assert d_wide < 0.25 * d_ref


# A cell scene whose program has bounds capacity 64, and the variants run on it.
CAPACITY_BASE = lambda: scenes.hyperbolic_cells(64, 48, primitives=48)  # noqa: E731
CAPACITY_VARIANTS = ["flat", "scalar_bounds", "sb_groups4", "groups4", "pixel_queue", "sb_g4_q_prio", "gq4_ho_prio",
                     "gq4_ho_hl", "gq4_ho_hl_s32", "gq4_s64", "gq4_ho_q_hl_s16"]


@pytest.mark.parametrize("variant", CAPACITY_VARIANTS)
def test_element_counts_across_the_bounds_capacity(ctx, variant):
    """A program built for 48 primitives (bounds capacity 64) renders cell
    scenes of the same type with 0, 3, 12, 16, 48, 64 and 120 primitives: an
    empty vector, scenes below and at the tree's 16-element threshold (flat
    layout, then the tree with groups), and more elements than the capacity
    (the unculled scan). Each equals the element-by-element scan bit for
    bit."""
    from hypertrace_amd.objects import ObjectVec, SceneImpl

    base = CAPACITY_BASE().scene
    culled = Render(ctx, base, fp_options=CULLED_VARIANTS[variant])
    plain = Render(ctx, base, fp_options=NO_CULL)
    for n in (0, 3, 12, 16, 48, 64, 120):
        print(f"case {variant}: {n} primitives", flush=True)
        if n == 0:
            scene = SceneImpl(base.geo, base.view, ObjectVec([], proto=base.object.items[0]), base.background,
                              base.hops)
        else:
            scene = scenes.hyperbolic_cells(64, 48, primitives=n).scene
        buf = EntityBuffer(ctx, scene)
        out = []
        for r in (culled, plain):
            c = Canvas(ctx, (64, 48), 1, 11)
            r.render_passes(buf, c, 3)
            r.render(buf, c)
            out.append((c.load_image(), c.load_seeds()))
        np.testing.assert_array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32), err_msg=f"{n} primitives")
        np.testing.assert_array_equal(out[0][1], out[1][1], err_msg=f"{n} primitives")
        assert np.all(out[0][0][..., 3] == 4.0)
