"""Scene-type checks on the host (no GPU): a program reads a scene buffer
with its own type's layout, so a buffer of another scene type is refused
before any launch (the reference's generic EntityBuffer<S> / Render<S> fix
the type at compile time, processing/src/buffer.rs:8-34, render.rs:8-119)."""
import pytest

from hypertrace_amd import scenes
from hypertrace_amd.native import HtError
from hypertrace_amd.processing import Pipeline, check_scene_type, scene_type_key
from hypertrace_amd.types import X32_F32, X32_F64


class _Buf:
    def __init__(self, key):
        self.scene_type = key


def test_type_key_follows_the_type_not_the_values():
    a = scenes.hyperbolic_cells(64, 48).scene
    b = scenes.hyperbolic_cells(64, 48, 1, primitives=96).scene  # same type, other values
    e = scenes.euclidean_spheres(64, 48).scene
    assert scene_type_key(a, X32_F32) == scene_type_key(b, X32_F32)
    assert scene_type_key(a, X32_F32) != scene_type_key(e, X32_F32)
    assert scene_type_key(a, X32_F32) != scene_type_key(a, X32_F64)  # another byte layout


def test_mismatch_is_refused():
    a = scene_type_key(scenes.hyperbolic_cells(64, 48).scene, X32_F32)
    e = scene_type_key(scenes.euclidean_spheres(64, 48).scene, X32_F32)
    check_scene_type(a, _Buf(a))
    check_scene_type(None, _Buf(e))  # batch programs check per job instead
    with pytest.raises(HtError) as err:
        check_scene_type(a, _Buf(e))
    assert err.value.code == -3
    with pytest.raises(HtError):
        check_scene_type(a, object())  # a buffer without a type


def test_pipeline_set_scene_refuses_another_type():
    class _Prog:
        scene_type = scene_type_key(scenes.hyperbolic_cells(64, 48).scene, X32_F32)

    class _Render:
        program = _Prog()

    class _Ctx:
        config = X32_F32

    p = Pipeline.__new__(Pipeline)  # no device: set_scene must refuse before any upload
    p.ctx, p.render_ = _Ctx(), _Render()
    with pytest.raises(HtError):
        p.set_scene(scenes.euclidean_spheres(64, 48).scene)
