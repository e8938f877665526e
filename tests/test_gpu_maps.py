"""GPU FP64 composition of the scene's Moebius words (ht_mo_compose) against
its host mirror: bit-identical scene bytes in FP32 and FP64 programs, and the
argument checks of the entry point."""
import numpy as np
import pytest

from hypertrace_amd import scenes
from hypertrace_amd.maps import WordMaps, compose_on_device
from hypertrace_amd.native import HtError
from hypertrace_amd.objects import field_offsets
from hypertrace_amd.processing import Context, EntityBuffer
from hypertrace_amd.types import X32_F32, X32_F64

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg", [X32_F32, X32_F64], ids=["f32", "f64"])
def test_gpu_composition_bit_identical_to_host(cfg):
    wl = scenes.hyperbolic_cells()
    want = wl.scene.to_bytes(cfg)
    with Context(0, cfg) as ctx:
        buf = EntityBuffer(ctx, wl.scene)
        # poison the maps so that only the GPU composition can restore them
        data = bytearray(want)
        for o in field_offsets(wl.scene, cfg, wl.word_maps.targets):
            data[o:o + 8 * cfg.real] = b"\xff" * (8 * cfg.real)
        buf.write(np.frombuffer(bytes(data), np.uint8))
        compose_on_device(ctx, buf, wl.scene, wl.word_maps)
        got = buf.read().tobytes()
    assert got == want


def test_gpu_composition_rejects_bad_arguments():
    wl = scenes.hyperbolic_cells(cells=4)
    wm = wl.word_maps
    with Context(0, X32_F32) as ctx:
        buf = EntityBuffer(ctx, wl.scene)
        bad_letter = WordMaps(wm.gens, [list(w) for w in wm.words], wm.posts, wm.targets)
        bad_letter.words[0] = [len(wm.gens)]
        with pytest.raises(HtError, match="letter"):
            compose_on_device(ctx, buf, wl.scene, bad_letter)
        too_long = WordMaps(wm.gens, [list(w) for w in wm.words], wm.posts, wm.targets)
        too_long.words[1] = [0] * 17
        with pytest.raises(ValueError):
            compose_on_device(ctx, buf, wl.scene, too_long)
        # a scene whose buffer is too small for the recorded offsets
        small = EntityBuffer(ctx, scenes.hyperbolic_cells(cells=1).scene)
        with pytest.raises(HtError, match="outside"):
            compose_on_device(ctx, small, wl.scene, wm)
        buf_ok = buf.read().tobytes()
    assert buf_ok == wl.scene.to_bytes(X32_F32)
