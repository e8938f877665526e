"""Scene entity vocabulary (host side).

Python mirror of the reference's Rust scene types (objects/src/**, the map and
geometry entities of types/src/map.rs and types/src/geometry.rs). Each entity
knows three things:

* `cpp(geo)`  the C++ type expression of the device template that renders it
  (hypertrace_amd/csrc/device/ht/entities.hh); this is the scene-to-kernel
  codegen of the reference (`*_source`, e.g. objects/src/scene.rs:49-67),
  which emitted OpenCL macro instantiations;
* `align(cfg)` / `size(cfg)` / `store(cfg, w)`  its bytes in the scene stream,
  following the reference serializer and derive rules
  (type-macros/src/entity.rs:7-361, types/src/containers/vector.rs:19-58,
  types/src/primitive/*.rs);
* `spec(geo, cfg)`  a token description for the CPU oracle.

Field order of every struct is the reference's declaration order, because it
fixes the byte layout.
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

from .types import Config, Writer, aligned_add, lcm, upper_multiple

EU = "Eu"
HY = "Hy"


def _h32(v: float) -> str:
    return float(np.float32(v)).hex()


def _hr(v: float, cfg: Config) -> str:
    return float(v).hex() if cfg.double_support else float(np.float32(v)).hex()


# --------------------------------------------------------------------- leaves

class _Leaf:
    dyn = False

    def size(self, cfg: Config) -> int:
        return self.static_size(cfg)


class _Real(_Leaf):
    def __init__(self, v):
        self.v = float(v)

    def align(self, cfg):
        return cfg.real

    def static_size(self, cfg):
        return cfg.real

    def store(self, cfg, w):
        w.real(self.v)


class _F32(_Leaf):
    def __init__(self, v):
        self.v = float(v)

    def align(self, cfg):
        return 4

    def static_size(self, cfg):
        return 4

    def store(self, cfg, w):
        w.f32(self.v)


class _Color(_Leaf):
    """Vector<f32, 3>: three floats plus one pad slot, align 16
    (types/src/primitive/vector.rs:68-112)."""

    def __init__(self, rgb):
        self.v = [float(x) for x in rgb]
        assert len(self.v) == 3

    def align(self, cfg):
        return 16

    def static_size(self, cfg):
        return 16

    def store(self, cfg, w):
        for x in self.v:
            w.f32(x)
        w.skip(4)


class _Real3(_Leaf):
    """Vector<f64, 3> stored as `real3`: three reals plus one pad slot."""

    def __init__(self, v):
        self.v = [float(x) for x in v]
        assert len(self.v) == 3

    def align(self, cfg):
        return 4 * cfg.real

    def static_size(self, cfg):
        return 4 * cfg.real

    def store(self, cfg, w):
        for x in self.v:
            w.real(x)
        w.skip(cfg.real)


class _RealN(_Leaf):
    """VectorN (N = 2, 4, 8, 16) of reals: size = align = N reals."""

    def __init__(self, v):
        self.v = [float(x) for x in v]

    def align(self, cfg):
        return len(self.v) * cfg.real

    def static_size(self, cfg):
        return len(self.v) * cfg.real

    def store(self, cfg, w):
        for x in self.v:
            w.real(x)


def _struct_align(cfg, fields) -> int:
    a = 1
    for f in fields:
        a = lcm(a, f.align(cfg))
    return a


def _struct_size(cfg, fields, dynamic: bool) -> int:
    acc = 0
    for f in fields:
        acc = aligned_add(acc, f.size(cfg) if dynamic else f.static_size(cfg), f.align(cfg))
    return upper_multiple(acc, _struct_align(cfg, fields))


def _struct_store(cfg, w, fields, align) -> None:
    for f in fields:
        w.align(f.align(cfg))
        if w.marks is not None:
            w.marks.setdefault(id(f), w.position)
        f.store(cfg, w)
    w.align(align)


def field_offsets(root: "Entity", cfg: Config, targets) -> List[int]:
    """Byte offsets, in root's serialized form, of the given entity instances
    (each stored as a field of some struct, e.g. the map of a Mapped)."""
    w = Writer(cfg)
    w.marks = {}
    root.store(cfg, w)
    return [w.marks[id(t)] for t in targets]


# --------------------------------------------------------------------- entity base

class Entity:
    """Base of scene entities. Subclasses with fixed fields return them from
    `fields()`; containers override the layout methods."""

    dyn = False

    def fields(self) -> list:
        return []

    def align(self, cfg: Config) -> int:
        return _struct_align(cfg, self.fields())

    def static_size(self, cfg: Config) -> int:
        return _struct_size(cfg, self.fields(), False)

    def size(self, cfg: Config) -> int:
        return _struct_size(cfg, self.fields(), True)

    def store(self, cfg: Config, w: Writer) -> None:
        _struct_store(cfg, w, self.fields(), self.align(cfg))

    def cpp(self, geo: str) -> str:
        raise NotImplementedError

    def spec(self, geo: str, cfg: Config) -> List[str]:
        raise NotImplementedError

    def to_bytes(self, cfg: Config) -> bytes:
        w = Writer(cfg)
        self.store(cfg, w)
        out = w.bytes()
        if len(out) != self.size(cfg):
            raise AssertionError(f"{type(self).__name__}: stored {len(out)} bytes, layout says {self.size(cfg)}")
        return out


class _Zst(Entity):
    """Unit struct: size 0, align 1, no bytes."""

    def align(self, cfg):
        return 1

    def static_size(self, cfg):
        return 0

    def size(self, cfg):
        return 0

    def store(self, cfg, w):
        pass


# --------------------------------------------------------------------- maps

class Shift3(Entity):
    """Euclidean translation (types/src/map.rs:54-70, transform/shift.hh:27-29)."""

    geo = EU

    def __init__(self, v):
        self.v = [float(x) for x in v]

    def fields(self):
        return [_Real3(self.v)]

    def cpp(self, geo):
        return "ht::Shift3"

    def spec(self, geo, cfg):
        return ["shift"] + [_hr(x, cfg) for x in self.v]


class Rotation3(Entity):
    """Rotation by the unit quaternion (w, x, y, z) (types/src/map.rs:92-108)."""

    geo = EU

    def __init__(self, q):
        self.q = [float(x) for x in q]

    @staticmethod
    def from_axis(axis, angle):
        s = np.sin(angle / 2)
        return Rotation3([np.cos(angle / 2), s * axis[0], s * axis[1], s * axis[2]])

    def fields(self):
        return [_RealN(self.q)]

    def cpp(self, geo):
        return "ht::Rotation3"

    def spec(self, geo, cfg):
        return ["rot"] + [_hr(x, cfg) for x in self.q]


class Homogenous3(Entity):
    """Chain{outer: Shift3, inner: Rotation3} (types/src/map.rs:165-259)."""

    geo = EU

    def __init__(self, shift=(0, 0, 0), rot=(1, 0, 0, 0)):
        self.shift = shift if isinstance(shift, Shift3) else Shift3(shift)
        self.rot = rot if isinstance(rot, Rotation3) else Rotation3(rot)

    def fields(self):
        return [self.shift, self.rot]

    def cpp(self, geo):
        return "ht::Homogenous3"

    def spec(self, geo, cfg):
        return ["hom"] + self.shift.spec(geo, cfg)[1:] + self.rot.spec(geo, cfg)[1:]


class Moebius(Entity):
    """Moebius transform [[a, b], [c, d]] of the upper half-space, stored as
    comp2x2 = real8 (types/src/map.rs:147-162, transform/moebius.hh:8-10)."""

    geo = HY

    def __init__(self, m):
        m = np.asarray(m, dtype=np.complex128).reshape(2, 2)
        self.m = m

    @staticmethod
    def identity():
        return Moebius([[1, 0], [0, 1]])

    def chain(self, other: "Moebius") -> "Moebius":
        return Moebius(self.m @ other.m)

    def normalized(self) -> "Moebius":
        det = self.m[0, 0] * self.m[1, 1] - self.m[0, 1] * self.m[1, 0]
        return Moebius(self.m / np.sqrt(det))

    def reals(self):
        a, b, c, d = self.m.reshape(4)
        return [a.real, a.imag, b.real, b.imag, c.real, c.imag, d.real, d.imag]

    def fields(self):
        return [_RealN(self.reals())]

    def cpp(self, geo):
        return "ht::Moebius"

    def spec(self, geo, cfg):
        return ["mo"] + [_hr(x, cfg) for x in self.reals()]


# Hyperbolic isometry generators, geometry/hyperbolic.cc:35-62 (FP64 on host).
def hy_zshift(l):
    e = np.exp(l / 2)
    return Moebius([[e, 0], [0, 1 / e]])


def hy_xshift(l):
    c, s = np.cosh(l / 2), np.sinh(l / 2)
    return Moebius([[c, s], [s, c]])


def hy_yshift(l):
    c, s = np.cosh(l / 2), np.sinh(l / 2)
    return Moebius([[c, 1j * s], [-1j * s, c]])


def hy_zrotate(phi):
    c, s = np.cos(phi / 2), np.sin(phi / 2)
    return Moebius([[c + 1j * s, 0], [0, c - 1j * s]])


def hy_xrotate(theta):
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return Moebius([[c, 1j * s], [1j * s, c]])


def hy_yrotate(theta):
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return Moebius([[c, -s], [s, c]])


# --------------------------------------------------------------------- shapes

class Sphere(_Zst):
    """Unit sphere (objects/src/shape/sphere.rs, shape/eu/sphere.cc)."""

    def cpp(self, geo):
        assert geo == EU
        return "ht::Sphere"

    def spec(self, geo, cfg):
        return ["sphere"]


class Ball(Entity):
    """Sphere of radius r (not in the reference; keeps world distances exact
    for varied radii)."""

    def __init__(self, r):
        self.r = float(r)

    def fields(self):
        return [_Real(self.r)]

    def cpp(self, geo):
        assert geo == EU
        return "ht::Ball"

    def spec(self, geo, cfg):
        return ["ball", _hr(self.r, cfg)]


class SdfSphere(Entity):
    """Sphere-traced sphere of radius r (not in the reference: its shapes are
    analytic; device ht/sdf.hh)."""

    def __init__(self, r=1.0):
        self.r = float(r)

    def fields(self):
        return [_Real(self.r)]

    def cpp(self, geo):
        assert geo == EU
        return "ht::SdfSphere"

    def spec(self, geo, cfg):
        return ["sdfsphere", _hr(self.r, cfg)]


class SdfBox(Entity):
    """Sphere-traced box with half extents h (real3)."""

    def __init__(self, h=(1.0, 1.0, 1.0)):
        self.h = [float(x) for x in h]

    def fields(self):
        return [_Real3(self.h)]

    def cpp(self, geo):
        assert geo == EU
        return "ht::SdfBox"

    def spec(self, geo, cfg):
        return ["sdfbox"] + [_hr(x, cfg) for x in self.h]


class SdfTorus(Entity):
    """Sphere-traced torus around the local y axis: major radius R, minor r."""

    def __init__(self, big=1.0, small=0.25):
        self.big, self.small = float(big), float(small)

    def fields(self):
        return [_Real(self.big), _Real(self.small)]

    def cpp(self, geo):
        assert geo == EU
        return "ht::SdfTorus"

    def spec(self, geo, cfg):
        return ["sdftorus", _hr(self.big, cfg), _hr(self.small, cfg)]


class Cube(_Zst):
    """Unit cube [-1, 1]^3 (shape/eu/cube.cc)."""

    def cpp(self, geo):
        assert geo == EU
        return "ht::Cube"

    def spec(self, geo, cfg):
        return ["cube"]


class Plane(_Zst):
    """Plane z = 0 (Euclidean) or the unit hemisphere (hyperbolic)."""

    def cpp(self, geo):
        return f"ht::Plane<ht::{geo}>"

    def spec(self, geo, cfg):
        return ["plane"]


class Horosphere(_Zst):
    """Horosphere z = 1 (shape/hy/horosphere.cc)."""

    def cpp(self, geo):
        assert geo == HY
        return "ht::Horosphere"

    def spec(self, geo, cfg):
        return ["horosphere"]


# --------------------------------------------------------------------- containers

class _VecBase(Entity):
    dyn = True

    def __init__(self, items: Sequence[Entity], proto: Entity = None):
        self.items = list(items)
        self.proto = proto if proto is not None else (self.items[0] if self.items else None)
        if self.proto is None:
            raise ValueError("an empty vector needs a prototype element")

    def _check(self, geo):
        t = self.proto.cpp(geo)
        for it in self.items:
            if it.cpp(geo) != t:
                raise TypeError(f"vector elements differ in type: {it.cpp(geo)} vs {t}")
        return t

    def align(self, cfg):
        return lcm(cfg.usize, self.proto.align(cfg))

    def static_size(self, cfg):
        return upper_multiple(cfg.usize, self.proto.align(cfg))

    def size(self, cfg):
        return upper_multiple(self.static_size(cfg) + len(self.items) * self.proto.static_size(cfg), self.align(cfg))

    def store(self, cfg, w):
        w.usize(len(self.items))
        w.align(self.proto.align(cfg))
        for it in self.items:
            if it.dyn:
                raise TypeError("vector elements must be statically sized")
            it.store(cfg, w)
        w.align(self.align(cfg))


class ShapeVec(_VecBase):
    """Nearest hit over shapes (shape/vector.inl)."""

    def cpp(self, geo):
        return f"ht::ShapeVec<ht::{geo}, {self._check(geo)}>"

    def spec(self, geo, cfg):
        out = ["shapevec", str(len(self.items))]
        for it in self.items:
            out += it.spec(geo, cfg)
        return out


class ObjectVec(_VecBase):
    """Nearest hit over objects (object/vector.inl, objects/src/object/vector.rs)."""

    def cpp(self, geo):
        return f"ht::ObjectVec<ht::{geo}, {self._check(geo)}>"

    def spec(self, geo, cfg):
        out = ["objvec", str(len(self.items))]
        for it in self.items:
            out += it.spec(geo, cfg)
        return out


class _ChoiceBase(Entity):
    """Rust enum with single-field variants: {usize index; union {...}}
    (type-macros/src/entity.rs:239-275, :315-361)."""

    def __init__(self, variants: Sequence[Entity], index: int):
        self.variants = list(variants)
        if not 0 <= index < len(self.variants):
            raise ValueError("choice index out of range")
        self.index = index

    def align(self, cfg):
        a = cfg.usize
        for v in self.variants:
            a = lcm(a, v.align(cfg))
        return a

    def static_size(self, cfg):
        m = 0
        for v in self.variants:
            m = max(m, aligned_add(cfg.usize, v.static_size(cfg), v.align(cfg)))
        return upper_multiple(m, self.align(cfg))

    def size(self, cfg):
        return self.static_size(cfg)

    def store(self, cfg, w):
        p0 = w.position
        w.usize(self.index)
        w.align(self.align(cfg))
        v = self.variants[self.index]
        w.align(v.align(cfg))
        v.store(cfg, w)
        written = w.position - p0
        ms = self.static_size(cfg)
        if written < ms:
            w.skip(ms - written)
        w.align(self.align(cfg))

    def _args(self, geo):
        return ", ".join(v.cpp(geo) for v in self.variants)


class ShapeChoice(_ChoiceBase):
    """`shape_choice!` enum (objects/src/shape/choice.rs:1-57)."""

    def cpp(self, geo):
        return f"ht::ShapeChoice<ht::{geo}, {self._args(geo)}>"

    def spec(self, geo, cfg):
        return self.variants[self.index].spec(geo, cfg)


class ObjectChoice(_ChoiceBase):
    """`object_choice!` enum (objects/src/object/choice.rs:1-78)."""

    def cpp(self, geo):
        return f"ht::ObjectChoice<ht::{geo}, {self._args(geo)}>"

    def spec(self, geo, cfg):
        return self.variants[self.index].spec(geo, cfg)


# --------------------------------------------------------------------- mapped

class Mapped(Entity):
    """Mapped {map, inner} (objects/src/mapped.rs:5-11): a shape, object or
    view placed by an isometry."""

    def __init__(self, inner: Entity, map: Entity):
        self.inner = inner
        self.map = map
        self.dyn = inner.dyn

    def fields(self):
        return [self.map, self.inner]

    def cpp(self, geo):
        return f"ht::Mapped<ht::{geo}, {self.inner.cpp(geo)}, {self.map.cpp(geo)}>"

    def spec(self, geo, cfg):
        return ["mapped"] + self.map.spec(geo, cfg) + self.inner.spec(geo, cfg)


# --------------------------------------------------------------------- materials

class Lambertian(_Zst):
    def cpp(self, geo):
        return "ht::Lambertian"

    def spec(self, geo, cfg):
        return ["lambertian"]


class Specular(_Zst):
    def cpp(self, geo):
        return "ht::Specular"

    def spec(self, geo, cfg):
        return ["specular"]


class Transparent(_Zst):
    def cpp(self, geo):
        return "ht::Transparent"

    def spec(self, geo, cfg):
        return ["transparent"]


class Absorbing(_Zst):
    def cpp(self, geo):
        return "ht::Absorbing"

    def spec(self, geo, cfg):
        return ["absorbing"]


class Refractive(Entity):
    """Refractive {index} (objects/src/material/refractive.rs)."""

    def __init__(self, index):
        self.index = float(index)

    def fields(self):
        return [_Real(self.index)]

    def cpp(self, geo):
        return "ht::Refractive"

    def spec(self, geo, cfg):
        return ["refractive", _hr(self.index, cfg)]


class Colored(Entity):
    """Colored {color, inner} (objects/src/material/modifier.rs:13-18)."""

    def __init__(self, inner: Entity, color):
        self.inner = inner
        self.color = [float(x) for x in color]

    def fields(self):
        return [_Color(self.color), self.inner]

    def cpp(self, geo):
        return f"ht::Colored<{self.inner.cpp(geo)}>"

    def spec(self, geo, cfg):
        return ["colored"] + [_h32(x) for x in self.color] + self.inner.spec(geo, cfg)


class Emissive(Entity):
    """Emissive {emission, inner} (objects/src/material/modifier.rs:47-52)."""

    def __init__(self, inner: Entity, emission):
        self.inner = inner
        self.emission = [float(x) for x in emission]

    def fields(self):
        return [_Color(self.emission), self.inner]

    def cpp(self, geo):
        return f"ht::Emissive<{self.inner.cpp(geo)}>"

    def spec(self, geo, cfg):
        return ["emissive"] + [_h32(x) for x in self.emission] + self.inner.spec(geo, cfg)


class _Component(Entity):
    def __init__(self, material: Entity, portion: float):
        self.material = material
        self.portion = float(portion)

    def fields(self):
        return [self.material, _Real(self.portion)]


class Mixture(Entity):
    """`mixture!` struct of Component{material, portion}
    (objects/src/material/mixture.rs:4-79)."""

    def __init__(self, components: Sequence[Tuple[Entity, float]]):
        self.components = [_Component(m, p) for m, p in components]

    def fields(self):
        return self.components

    def cpp(self, geo):
        return "ht::Mixture<" + ", ".join(c.material.cpp(geo) for c in self.components) + ">"

    def spec(self, geo, cfg):
        out = ["mixture", str(len(self.components))]
        for c in self.components:
            out += [_hr(c.portion, cfg)] + c.material.spec(geo, cfg)
        return out


# --------------------------------------------------------------------- objects

class Covered(Entity):
    """Covered {material, shape} (objects/src/object/covered.rs:11-23)."""

    def __init__(self, shape: Entity, material: Entity):
        self.shape = shape
        self.material = material

    def fields(self):
        return [self.material, self.shape]

    def cpp(self, geo):
        return f"ht::Covered<ht::{geo}, {self.shape.cpp(geo)}, {self.material.cpp(geo)}>"

    def spec(self, geo, cfg):
        return ["covered"] + self.shape.spec(geo, cfg) + self.material.spec(geo, cfg)


class _Array(Entity):
    """[T; N]: elements back to back, align of T (types/src/containers/array.rs)."""

    def __init__(self, items):
        self.items = list(items)

    def align(self, cfg):
        return self.items[0].align(cfg)

    def static_size(self, cfg):
        return len(self.items) * self.items[0].static_size(cfg)

    def size(self, cfg):
        return self.static_size(cfg)

    def store(self, cfg, w):
        for it in self.items:
            it.store(cfg, w)


PLANE_TILINGS = {"uniform": 0, "pentagonal": 1, "pentastar": 2}
HOROSPHERE_TILINGS = {"uniform": 0, "square": 1, "hexagonal": 2}


class _Tiled(Entity):
    """{materials: [M; N], border_material: M, cell_size, border_width}
    (objects/src/object/tiled_plane.rs:35-42, tiled_horosphere.rs:35-42)."""

    TILINGS: dict = {}
    CPP = ""
    TOKEN = ""

    def __init__(self, materials, cell_size, border_width, border_material, tiling):
        self.materials = list(materials)
        if tiling not in self.TILINGS:
            raise ValueError(f"unknown tiling {tiling!r}")
        self.tiling = tiling
        self.cell_size = float(cell_size)
        self.border_width = float(border_width)
        self.border_material = border_material
        t = border_material.cpp(HY)
        for m in self.materials:
            if m.cpp(HY) != t:
                raise TypeError("tile materials must have one type")

    def fields(self):
        return [_Array(self.materials), self.border_material, _Real(self.cell_size), _Real(self.border_width)]

    def cpp(self, geo):
        assert geo == HY
        return (f"ht::{self.CPP}<{self.border_material.cpp(geo)}, {self.TILINGS[self.tiling]}, "
                f"{len(self.materials)}>")

    def spec(self, geo, cfg):
        out = [self.TOKEN, str(self.TILINGS[self.tiling]), str(len(self.materials)), _hr(self.cell_size, cfg),
               _hr(self.border_width, cfg)]
        for m in self.materials:
            out += m.spec(geo, cfg)
        return out + self.border_material.spec(geo, cfg)


class TiledPlane(_Tiled):
    """Geodesic plane tiled uniformly or by the {5,4} pentagonal /
    pentastar tiling (object/hy/tiled_plane/impl.inl)."""

    TILINGS = PLANE_TILINGS
    CPP = "TiledPlane"
    TOKEN = "tiledplane"


class TiledHorosphere(_Tiled):
    """Horosphere tiled uniformly, by squares or by hexagons
    (object/hy/tiled_horosphere/impl.inl)."""

    TILINGS = HOROSPHERE_TILINGS
    CPP = "TiledHorosphere"
    TOKEN = "tiledhoro"


# --------------------------------------------------------------------- views

class PointView(Entity):
    """Pinhole view {fov} (objects/src/view/point.rs:10-43); `fov` is the
    tangent of half the vertical field of view."""

    def __init__(self, fov: float):
        self.fov = float(fov)

    def fields(self):
        return [_Real(self.fov)]

    def cpp(self, geo):
        return f"ht::PointView<ht::{geo}>"

    def spec(self, geo, cfg):
        return ["pointview", _hr(self.fov, cfg)]


# --------------------------------------------------------------------- backgrounds

class ConstBg(Entity):
    """Constant colour {color} (objects/src/background/basic.rs:11-21)."""

    def __init__(self, color):
        self.color = [float(x) for x in color]

    def fields(self):
        return [_Color(self.color)]

    def cpp(self, geo):
        return "ht::ConstBg"

    def spec(self, geo, cfg):
        return ["constbg"] + [_h32(x) for x in self.color]


class GradBg(Entity):
    """Gradient {direction, colors[2], power} (objects/src/background/basic.rs:52-69)."""

    def __init__(self, direction, colors, power):
        self.direction = [float(x) for x in direction]
        self.colors = [[float(x) for x in c] for c in colors]
        self.power = float(power)

    def fields(self):
        return [_Real3(self.direction), _Color(self.colors[0]), _Color(self.colors[1]), _F32(self.power)]

    def cpp(self, geo):
        assert geo == EU, "GradBg is Euclidean only"
        return "ht::GradBg"

    def spec(self, geo, cfg):
        return (["gradbg"] + [_hr(x, cfg) for x in self.direction] + [_h32(x) for x in self.colors[0]]
                + [_h32(x) for x in self.colors[1]] + [_h32(self.power)])


# --------------------------------------------------------------------- plugins

PLUGIN_KINDS = ("shape", "object", "material", "background", "view")

_FIELD_KINDS = {
    "real": _Real,      # real (f32 or f64 with the program's precision)
    "float": _F32,      # f32
    "float3": _Color,   # Vector<f32, 3>: 16 B
    "real3": _Real3,    # Vector<real, 3>: 4 reals
    "real2": _RealN,    # Vector<real, N>, N = 2, 4, 8, 16: N reals
    "real4": _RealN,
}


def plugin_field(kind: str, value):
    """A field of a plugin entity's C struct, in declaration order: `kind`
    is its device type name ("real", "float", "float3", "real3", "real2",
    "real4"), as the reference's derive names the field types
    (types/src/primitive/vector.rs:16-19, scalar.rs)."""
    if kind not in _FIELD_KINDS:
        raise ValueError(f"plugin field kind must be one of {sorted(_FIELD_KINDS)}")
    return _FIELD_KINDS[kind](value)


class Plugin(Entity):
    """An entity whose device code is a plugin written against the
    reference's device interface: a C struct of its fields and free
    functions named `<prefix>_detect` / `<prefix>_interact` /
    `<prefix>_sample`, declared by `define_<kind>_interface(Self, prefix,
    ...)` in the header `header` (the reference's user-entity seam,
    kernel/ocl/{shape,object,material,background,view}/interface.hh).
    `sources` maps include names to file texts (the header, the files it
    includes); `fields` are `plugin_field`s in declaration order, which fix
    the bytes as the reference's derive does (type-macros/src/entity.rs).
    The device type is the adaptor `ht_plugin::<prefix>` the interface macro
    defines. The oracle has no counterpart: a plugin runs on the device only."""

    def __init__(self, kind: str, prefix: str, header: str, sources, fields: Sequence = ()):
        if kind not in PLUGIN_KINDS:
            raise ValueError(f"plugin kind must be one of {PLUGIN_KINDS}")
        if not prefix.isidentifier():
            raise ValueError("plugin prefix must be a C identifier")
        self.kind = kind
        self.prefix = prefix
        self.header = header
        self.sources = dict(sources)
        if header not in self.sources:
            raise ValueError(f"plugin sources lack the header {header!r}")
        self._fields = list(fields)

    def fields(self):
        return self._fields

    def cpp(self, geo):
        return f"ht_plugin::{self.prefix}"

    def spec(self, geo, cfg):
        raise NotImplementedError("plugin entities have no oracle counterpart")


# --------------------------------------------------------------------- scene

class SceneImpl(Entity):
    """SceneImpl {view, background, object} with `hops` light hops
    (objects/src/scene.rs:20-42)."""

    def __init__(self, geo: str, view: Entity, object: Entity, background: Entity, hops: int):
        if geo not in (EU, HY):
            raise ValueError("geometry must be 'Eu' or 'Hy'")
        self.geo = geo
        self.view = view
        self.object = object
        self.background = background
        self.hops = int(hops)
        self.dyn = object.dyn

    def fields(self):
        return [self.view, self.background, self.object]

    def cpp(self, geo=None):
        g = self.geo
        return (f"ht::SceneImpl<ht::{g}, {self.view.cpp(g)}, {self.object.cpp(g)}, "
                f"{self.background.cpp(g)}, {self.hops}>")

    def spec(self, geo=None, cfg: Config = None):
        g = self.geo
        return (["scene", g.lower(), str(self.hops)] + self.view.spec(g, cfg) + self.background.spec(g, cfg)
                + self.object.spec(g, cfg))

    def spec_text(self, cfg: Config) -> str:
        return " ".join(self.spec(self.geo, cfg))
