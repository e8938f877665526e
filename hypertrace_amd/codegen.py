"""Scene-to-kernel codegen: the root source and header map of one render
program.

The reference built an OpenCL program per scene type from `kernel::SOURCE`
plus generated headers and a root file that typedef'd the scene and included
render/render.cl (processing/src/render.rs:14-54). Here the root source
typedefs the scene's C++ template type, checks its byte layout against the
host serializer with static_asserts, and instantiates the render kernels;
the header map carries the device library.
"""
from __future__ import annotations

import glob
import os
from typing import Dict, List, Tuple

from .objects import (Covered, Entity, Mapped, ObjectVec, Plugin, SceneImpl, _ChoiceBase, _Component, _Tiled,
                      _VecBase)
from .types import Config

DEVICE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "device")

# Include names of the reference's device library (kernel/ocl/...) that a
# plugin's sources may include. The plugin region (ht_plugin/begin.hh)
# already provides their vocabulary, so each maps to an empty header.
REFERENCE_INCLUDES = (
    "macros.hh", "types.hh", "random.hh", "hash.hh",
    "algebra/real.hh", "algebra/vector.hh", "algebra/color.hh", "algebra/complex.hh", "algebra/quaternion.hh",
    "render/context.hh", "render/light/base.hh", "render/light/eu.hh", "render/light/hy.hh",
    "geometry/euclidean.hh", "geometry/hyperbolic.hh", "geometry/eu.hh", "geometry/hy.hh",
    "geometry/ray_eu.hh", "geometry/ray_hy.hh", "transform/rotation.hh",
    "shape/interface.hh", "object/interface.hh", "material/interface.hh", "background/interface.hh",
    "view/interface.hh",
)


def _read_dir(sub: str) -> Dict[str, str]:
    out = {}
    for path in sorted(glob.glob(os.path.join(DEVICE_DIR, sub, "*.hh"))):
        name = os.path.relpath(path, DEVICE_DIR)
        with open(path) as f:
            out[name] = f.read()
    return out


def device_headers() -> Dict[str, str]:
    """Device library files keyed by include name ("ht/base.hh", ...)."""
    return _read_dir("ht")


def plugins(*scenes: Entity) -> List[Plugin]:
    """The plugin entities of the scenes, each type once, in tree order."""
    out: List[Plugin] = []
    seen = set()
    stack: List[Entity] = list(reversed(scenes))
    while stack:
        e = stack.pop()
        if isinstance(e, Plugin) and e.prefix not in seen:
            seen.add(e.prefix)
            out.append(e)
        stack.extend(reversed(_children(e)))
    return out


def program_headers(*scenes: Entity) -> Dict[str, str]:
    """Header map of a program over `scenes`: the device library, plus, when
    a scene holds plugin entities, the plugin region, the reference's include
    names it stands in for and the plugins' own sources."""
    out = device_headers()
    found = plugins(*[s for s in scenes if s is not None])
    if not found:
        return out
    out.update(_read_dir("ht_plugin"))
    for name in REFERENCE_INCLUDES:
        out[name] = "#pragma once\n// Provided by the plugin region (ht_plugin/begin.hh).\n"
    for p in found:
        for name, text in p.sources.items():
            if name in out and out[name] != text:
                raise ValueError(f"plugin {p.prefix}: source {name!r} clashes with another file of that name")
            out[name] = text
    return out


def _plugin_includes(*scenes: Entity) -> List[str]:
    found = plugins(*scenes)
    if not found:
        return []
    return (["#include <ht_plugin/begin.hh>"] + [f"#include <{h}>" for h in dict.fromkeys(p.header for p in found)]
            + ["#include <ht_plugin/end.hh>"])


def _children(e: Entity) -> List[Entity]:
    if isinstance(e, SceneImpl):
        return [e.view, e.background, e.object]
    if isinstance(e, _VecBase):
        return [e.proto]
    if isinstance(e, _ChoiceBase):
        return list(e.variants)
    if isinstance(e, Mapped):
        return [e.map, e.inner]
    if isinstance(e, _Tiled):
        return list(e.materials) + [e.border_material]
    if isinstance(e, Plugin):
        return []
    kids = []
    for name in ("inner", "shape", "material"):
        if hasattr(e, name) and isinstance(getattr(e, name), Entity):
            kids.append(getattr(e, name))
    if hasattr(e, "components"):
        kids += [c.material for c in e.components]
    return kids


def layout_asserts(scene: SceneImpl, cfg: Config) -> List[str]:
    """static_asserts pinning every entity type's SIZE/ALIGN to the host
    serializer's numbers."""
    seen: Dict[str, Tuple[int, int]] = {}
    stack: List[Entity] = [scene]
    while stack:
        e = stack.pop()
        t = e.cpp(scene.geo)
        if t not in seen:
            seen[t] = (e.static_size(cfg), e.align(cfg))
        stack.extend(_children(e))
    lines = []
    for t, (size, align) in sorted(seen.items()):
        lines.append(f"static_assert({t}::SIZE == {size} && {t}::ALIGN == {align}, \"layout of {t}\");")
    return lines


def bounds_cap(scene: SceneImpl) -> int:
    """LDS capacity (elements) of the culled scan: the next power of two
    holding the scene's top-level vector, at least 64. A later scene of the
    same type with more elements falls back to the plain scan."""
    n = len(scene.object.items) if isinstance(scene.object, _VecBase) else 0
    cap = 64
    while cap < n:
        cap *= 2
    return min(cap, 4096)


def cache_supported(scene: SceneImpl) -> bool:
    """The radiance-cache kernel needs a top-level vector of mapped covered
    objects (it reads the hit's normal and material through them)."""
    obj = scene.object
    return isinstance(obj, ObjectVec) and len(obj.items) > 0 and all(
        isinstance(e, Mapped) and isinstance(e.inner, Covered) for e in obj.items)


def root_source(scene: SceneImpl, cfg: Config, wavefront: bool = False) -> str:
    """Root source of a scene's render program: the render kernels, the
    radiance-cache kernel where the scene supports it, and with `wavefront`
    the kernels of the wavefront form (ht_render_wavefront)."""
    cached = cache_supported(scene)
    lines = [
        "// Generated render program for one scene type.",
        f"#define HT_BOUNDS_CAP {bounds_cap(scene)}",
        "#include <ht/cache.hh>" if cached else "#include <ht/render.hh>",
    ]
    if wavefront:
        lines.append("#include <ht/wavefront.hh>")
    lines += _plugin_includes(scene)
    lines.append(f"typedef {scene.cpp()} Scene;")
    lines += layout_asserts(scene, cfg)
    lines.append("HT_DEFINE_RENDER_KERNELS(Scene)")
    if cached:
        lines.append("HT_DEFINE_CACHED_KERNEL(Scene)")
    if wavefront:
        lines.append("HT_DEFINE_WAVEFRONT_KERNELS(Scene)")
    return "\n".join(lines) + "\n"


def batch_source(scene0: SceneImpl, scene1: SceneImpl, cfg: Config, block: int = 256,
                 max_elements: int = 256) -> str:
    """Root source of a batch program: one persistent kernel (`render_batch`)
    over jobs of two scene types; `block` threads per block (64..1024),
    culled-scan bounds for up to `max_elements` top-level objects per scene
    (larger scenes fall back to the plain scan)."""
    if block % 64 or not 64 <= block <= 1024:
        raise ValueError("block must be 64..1024 in steps of 64")
    cap = 64
    while cap < max_elements:
        cap *= 2
    lines = [
        "// Generated batch program: two scene types, one persistent kernel.",
        f"#define HT_BOUNDS_CAP {min(cap, 4096)}",
        f"#define HT_MAX_WAVES {block // 64}",
        f"#define HT_BATCH_BLOCK {block}",
        "#include <ht/batch.hh>",
    ]
    lines += _plugin_includes(scene0, scene1)
    lines += [
        f"typedef {scene0.cpp()} Scene0;",
        f"typedef {scene1.cpp()} Scene1;",
    ]
    lines += layout_asserts(scene0, cfg)
    lines += layout_asserts(scene1, cfg)
    lines.append("HT_DEFINE_BATCH_KERNEL(Scene0, Scene1)")
    return "\n".join(dict.fromkeys(lines)) + "\n"
