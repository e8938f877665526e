"""On-disk image output for visual regression: binary PPM round trip (CPU)
and the pipeline's saved frame equal to its RGBA8 image (GPU)."""
import numpy as np
import pytest

from hypertrace_amd.processing import read_ppm, write_ppm


def test_ppm_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (7, 11, 4), dtype=np.uint8)
    p = tmp_path / "x.ppm"
    write_ppm(str(p), img)
    data = p.read_bytes()
    assert data.startswith(b"P6\n11 7\n255\n") and len(data) == len(b"P6\n11 7\n255\n") + 7 * 11 * 3
    assert np.array_equal(read_ppm(str(p)), img[..., :3])


@pytest.mark.gpu
def test_pipeline_saves_its_frame(tmp_path):
    from hypertrace_amd import scenes
    from hypertrace_amd.processing import Context, Pipeline
    from hypertrace_amd.types import X32_F32

    wl = scenes.euclidean_spheres(40, 30, 1)
    with Context(0, X32_F32) as ctx:
        pipe = Pipeline(ctx, (40, 30), wl.scene, seed_mode=wl.seed_mode, seed=wl.seed)
        pipe.render_passes(4)
        rgba = pipe.save_image(str(tmp_path / "f.ppm"))
    assert rgba.shape == (30, 40, 4) and np.all(rgba[..., 3] == 255)
    assert np.array_equal(read_ppm(str(tmp_path / "f.ppm")), rgba[..., :3])
