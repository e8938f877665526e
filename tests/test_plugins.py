"""The device plugin seam on the CPU: entities written against the
reference's device interface (free functions named <prefix>_detect /
_interact / _sample over C structs, declared by define_*_interface;
kernel/ocl/{shape,object,material,background,view}/interface.hh) compile into
a render program through hipRTC, their bytes equal the built-in twins', and
mistakes in a plugin (a C struct whose layout disagrees with the declared
fields, a function whose signature is not the interface's) stop the build
with a message naming them. Rendering is compared on the GPU
(test_gpu_plugins.py)."""
import pytest

import plugin_scenes as ps
from hypertrace_amd import codegen, native
from hypertrace_amd.native import HtError
from hypertrace_amd.objects import Covered, Plugin, PointView, SceneImpl, Sphere, plugin_field
from hypertrace_amd.processing import precompile, scene_type_key
from hypertrace_amd.types import X32_F32, X32_F64, Config

X64_F32 = Config(64, False)


@pytest.mark.parametrize("cfg", [X32_F32, X32_F64, X64_F32], ids=["x32f32", "x32f64", "x64f32"])
def test_plugin_scene_bytes_equal_builtin_twin(cfg):
    built, plug = ps.scene_pair()
    assert plug.to_bytes(cfg) == built.to_bytes(cfg)


def test_plugin_root_source_and_headers():
    built, plug = ps.scene_pair()
    src = codegen.root_source(plug, X32_F32)
    assert src.index("#include <ht_plugin/begin.hh>") < src.index("#include <user/ref_sphere.hh>")
    assert src.index("#include <ht_plugin/end.hh>") < src.index("typedef ")
    # each plugin type once in the header map and the include list
    prefixes = [p.prefix for p in codegen.plugins(plug)]
    assert sorted(prefixes) == sorted(
        ["ref_pinhole", "ref_sphere", "ref_diffuse", "ref_tinted", "ref_mirror", "ref_glow", "ref_ball", "ref_sky"])
    hdrs = codegen.program_headers(plug)
    for name in ps.SOURCES:
        assert hdrs[name] == ps.SOURCES[name]
    for name in ("material/interface.hh", "shape/interface.hh", "render/light/eu.hh", "transform/rotation.hh"):
        assert name in hdrs
    # a scene without plugins gets the device library only
    assert set(codegen.program_headers(built)) == set(codegen.device_headers())
    assert "ht_plugin" not in codegen.root_source(built, X32_F32)
    # the layout of every plugin's C struct is checked against its fields
    assert "static_assert(ht_plugin::ref_glow::SIZE == 16 && ht_plugin::ref_glow::ALIGN == 16" in src
    assert "static_assert(ht_plugin::ref_pinhole::SIZE == 4 && ht_plugin::ref_pinhole::ALIGN == 4" in src


def test_plugin_scene_type_differs_from_builtin():
    built, plug = ps.scene_pair()
    assert scene_type_key(built, X32_F32) != scene_type_key(plug, X32_F32)


@pytest.mark.parametrize("cfg", [X32_F32, X32_F64], ids=["f32", "f64"])
def test_plugin_program_builds(cfg):
    native.lib()
    _, plug = ps.scene_pair()
    precompile(plug, cfg)


def _one_material_scene(material):
    return SceneImpl("Eu", PointView(0.5), Covered(Sphere(), material), ps.ref_sky((0.1, 0.2, 0.3)), 3)


def test_plugin_layout_mismatch_is_a_build_error():
    # The struct holds a float3 (16 B) but the entity declares a real (4 B).
    bad = Plugin("material", "ref_glow", "user/ref_materials.hh", ps.MAT_SRC, [plugin_field("real", 1.0)])
    with pytest.raises(HtError, match="layout of ht_plugin::ref_glow"):
        precompile(_one_material_scene(bad), X32_F32)


def test_plugin_signature_mismatch_is_a_build_error():
    src = """#pragma once
#include <material/interface.hh>
typedef void Wrong;
define_material_interface(Wrong, wrong)
bool wrong_interact(__global const Wrong *self, Context *context, real3 normal, LightLocal *light) {
    return true;
}
"""
    bad = Plugin("material", "wrong", "user/wrong.hh", {"user/wrong.hh": src})
    with pytest.raises(HtError, match="wrong_interact"):
        precompile(_one_material_scene(bad), X32_F32)


def test_plugin_argument_checks():
    with pytest.raises(ValueError):
        Plugin("texture", "x", "a.hh", {"a.hh": ""})
    with pytest.raises(ValueError):
        Plugin("material", "not an identifier", "a.hh", {"a.hh": ""})
    with pytest.raises(ValueError):
        Plugin("material", "x", "a.hh", {"b.hh": ""})
    with pytest.raises(ValueError):
        plugin_field("double9", 0.0)
    with pytest.raises(NotImplementedError):
        ps.ref_mirror().spec("Eu", X32_F32)


def test_plugin_source_name_clash_is_refused():
    a = Plugin("material", "a_mat", "user/x.hh", {"user/x.hh": "// one"})
    b = Plugin("material", "b_mat", "user/x.hh", {"user/x.hh": "// two"})
    scene = SceneImpl("Eu", PointView(0.5), Covered(Sphere(), a), ps.ref_sky((0, 0, 0)), 1)
    scene2 = SceneImpl("Eu", PointView(0.5), Covered(Sphere(), b), ps.ref_sky((0, 0, 0)), 1)
    with pytest.raises(ValueError, match="clashes"):
        codegen.program_headers(scene, scene2)
