"""The reference's own layout known answers, restated on the host layout
code that every scene buffer goes through (hypertrace_amd/objects.py,
types.py): objects/src/macros/derive_tests.rs (derived structs, tuples and
enums, X64 with f64) and processing/src/tests/ocl.rs (OpenCL struct sizes
of the tiled-horosphere shape, zero-size types)."""
from hypertrace_amd.objects import (Colored, Entity, Lambertian, Mixture, TiledHorosphere, _ChoiceBase,
                                    _Component, _Zst)
from hypertrace_amd.types import X32_F32, Config

X64_F64 = Config(64, True)


class Prim(Entity):
    """A primitive of fixed size and alignment (u8, u16, i32, f32, float4 ...)."""

    def __init__(self, size, align=None):
        self._size, self._align = size, align if align is not None else size

    def align(self, cfg):
        return self._align

    def static_size(self, cfg):
        return self._size

    def size(self, cfg):
        return self._size


class Struct(Entity):
    def __init__(self, *fields):
        self._fields = list(fields)

    def fields(self):
        return self._fields


class Enum(_ChoiceBase):
    """Single-entity variants (the derive's struct-like variants are their
    fields laid out after the index, the same as one nested struct here)."""


def u8():
    return Prim(1)


def test_derive_unit():
    # derive_tests.rs:13-20: Unit0 -> align 1, size 0
    z = _Zst()
    assert z.align(X64_F64) == 1 and z.static_size(X64_F64) == 0


def test_derive_struct_and_tuple():
    # derive_tests.rs:22-44: {u8, i32, (), [u8; 3]} -> align 4, size 12
    s = Struct(u8(), Prim(4), _Zst(), Prim(3, 1))
    assert s.align(X64_F64) == 4
    assert s.static_size(X64_F64) == 12


def test_derive_enum():
    # derive_tests.rs:46-56: enum {A {u8, f32}, B((), [u16; 3]), C} -> align 8, size 16 (X64 index)
    e = Enum([Struct(u8(), Prim(4)), Struct(_Zst(), Prim(6, 2)), _Zst()], 0)
    assert e.align(X64_F64) == 8
    assert e.static_size(X64_F64) == 16


def test_ocl_zero_size_types():
    # ocl.rs zst: struct {} and struct { struct {} } have size 0
    assert _Zst().static_size(X32_F32) == 0
    assert Struct(_Zst()).static_size(X32_F32) == 0


def test_ocl_tiled_horosphere_sizes():
    # ocl.rs tiled_horosphere_size: Material {float4; float; float} = 32,
    # Object {Material[3]; Material; float; float} = 144, {uint; Material; Object} = 192
    material = Struct(Prim(16), Prim(4), Prim(4))
    assert material.align(X32_F32) == 16 and material.static_size(X32_F32) == 32
    obj = Struct(Prim(96, 16), material, Prim(4), Prim(4))
    assert obj.static_size(X32_F32) == 4 * 2 * 4 * 4 + 4 * 4
    choice = Struct(Prim(4), material, obj)
    assert choice.static_size(X32_F32) == 4 * 4 + 2 * 4 * 4 + 4 * 2 * 4 * 4 + 4 * 4


def test_tiled_horosphere_entity_matches_the_ocl_layout():
    """The real entity with a 32-byte material ({color3; inner}, a mixture
    component {Colored, portion}) has the sizes the OpenCL test asserts."""
    m = Mixture([(Colored(Lambertian(), (1, 1, 1)), 1.0)])
    comp = _Component(Colored(Lambertian(), (1, 1, 1)), 1.0)
    assert comp.static_size(X32_F32) == 32 and comp.align(X32_F32) == 16
    assert m.static_size(X32_F32) == 32
    t = TiledHorosphere([m] * 3, 1.0, 0.02, m, "hexagonal")
    assert t.static_size(X32_F32) == 144
    assert len(t.to_bytes(X32_F32)) == 144

