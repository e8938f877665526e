"""Consistency of the Rust `backend_hip` crate (rust/processing_hip) with the
C ABI and the Python codegen. No Rust toolchain is installed here, so these
are textual checks: every function of include/hypertrace_hip.h is declared in
ffi.rs with the same number of parameters, the reference include names match
the codegen's, every device type expression the crate emits is one the
Python entity vocabulary emits, and the sources' brackets balance."""
import os
import re

import pytest

from hypertrace_amd import codegen, objects

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CRATE = os.path.join(ROOT, "rust", "processing_hip")


def _read(*parts):
    with open(os.path.join(*parts)) as f:
        return f.read()


def _strip_c_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def _arity(params: str) -> int:
    params = params.strip()
    if params in ("", "void"):
        return 0
    return params.count(",") + 1


def header_functions():
    text = _strip_c_comments(_read(ROOT, "include", "hypertrace_hip.h"))
    out = {}
    for m in re.finditer(r"\b(ht_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text):
        out[m.group(1)] = _arity(m.group(2))
    return out


def rust_functions():
    text = _read(CRATE, "src", "ffi.rs")
    block = text[text.index('extern "C" {'):]
    out = {}
    for m in re.finditer(r"pub fn (ht_[a-z0-9_]+)\s*\(([^)]*)\)", block):
        params = re.sub(r"\s+", " ", m.group(2)).strip().rstrip(",")
        out[m.group(1)] = 0 if not params else params.count(":")
    return out


def test_ffi_declares_every_c_function_with_its_arity():
    c = header_functions()
    r = rust_functions()
    assert len(c) >= 45
    assert set(c) == set(r), f"only in C: {set(c) - set(r)}, only in Rust: {set(r) - set(c)}"
    for name in c:
        assert c[name] == r[name], f"{name}: {c[name]} parameters in C, {r[name]} in Rust"


def test_batch_job_struct_matches():
    c = _strip_c_comments(_read(ROOT, "include", "hypertrace_hip.h"))
    body = re.search(r"typedef struct ht_batch_job \{(.*?)\} ht_batch_job;", c, re.S).group(1)
    c_fields = re.findall(r"\*?(\w+)\s*[,;]", body)
    rs = _read(CRATE, "src", "ffi.rs")
    r_body = re.search(r"pub struct ht_batch_job \{(.*?)\}", rs, re.S).group(1)
    r_fields = [f.rstrip("_") for f in re.findall(r"pub (\w+):", r_body)]
    assert c_fields == r_fields


def test_error_codes_match():
    c = _read(ROOT, "include", "hypertrace_hip.h")
    rs = _read(CRATE, "src", "ffi.rs")
    for name in ("HT_OK", "HT_ERR_IO", "HT_ERR_BACKEND", "HT_ERR_OTHER"):
        cv = int(re.search(rf"#define {name} \(?(-?\d+)\)?", c).group(1))
        rv = int(re.search(rf"pub const {name}: c_int = (-?\d+);", rs).group(1))
        assert cv == rv, name


def test_reference_include_list_matches_codegen():
    rs = _read(CRATE, "src", "device.rs")
    start = rs.index("pub const REFERENCE_INCLUDES")
    block = rs[start:rs.index("];", start)]
    assert tuple(re.findall(r'"([^"]+)"', block)) == codegen.REFERENCE_INCLUDES


def _python_types():
    eu, hy = objects.EU, objects.HY
    out = set()
    leaves = [objects.Sphere(), objects.Cube(), objects.Horosphere(), objects.Plane(), objects.Lambertian(),
              objects.Specular(), objects.Transparent(), objects.Absorbing(), objects.Refractive(1.5),
              objects.Shift3((0, 0, 0)), objects.Rotation3((1, 0, 0, 0)), objects.Homogenous3(),
              objects.Moebius.identity(), objects.PointView(1.0), objects.ConstBg((0, 0, 0)),
              objects.GradBg((0, 1, 0), [(0, 0, 0), (1, 1, 1)], 1.0)]
    for e in leaves:
        for g in (eu, hy):
            try:
                out.add(e.cpp(g))
            except AssertionError:
                pass
    return out


def test_leaf_type_expressions_are_the_python_vocabulary():
    rs = _read(CRATE, "src", "source.rs")
    names = re.findall(r'^leaf!\(.*, \w+, "([^"]+)"\);$', rs, re.M)
    assert len(names) == 13
    py = _python_types()
    for n in names:
        assert n in py, n
    # the templated ones, in the form the Python vocabulary writes them
    for pattern, example in [
        (r'"ht::Plane<ht::\{\}>"', objects.Plane().cpp("Hy")),
        (r'"ht::PointView<ht::\{\}>"', objects.PointView(1.0).cpp("Eu")),
        (r'"ht::Covered<ht::\{\}, \{\}, \{\}>"', objects.Covered(objects.Sphere(), objects.Lambertian()).cpp("Eu")),
        (r'"ht::Mapped<ht::\{\}, \{\}, \{\}>"',
         objects.Mapped(objects.Sphere(), objects.Homogenous3()).cpp("Eu")),
        (r'"ht::Colored<\{\}>"', objects.Colored(objects.Lambertian(), (1, 1, 1)).cpp("Eu")),
        (r'"ht::Emissive<\{\}>"', objects.Emissive(objects.Lambertian(), (1, 1, 1)).cpp("Eu")),
        (r'"ht::ObjectVec<ht::\{\}, \{\}>"', None),
        (r'"ht::ShapeVec<ht::\{\}, \{\}>"', None),
        (r'"ht::SceneImpl<ht::\{\}, \{\}, \{\}, \{\}, \{\}>"', None),
        (r'"ht::TiledPlane<\{\}, \{\}, \{\}>"', None),
        (r'"ht::TiledHorosphere<\{\}, \{\}, \{\}>"', None),
        (r'"ht::ObjectChoice<ht::\{\}, \{\}>"', None),
        (r'"ht::ShapeChoice<ht::\{\}, \{\}>"', None),
        (r'"ht::Mixture<\{\}>"', None),
        (r'"ht_plugin::\{\}"', None),
    ]:
        assert re.search(pattern, rs), pattern
        if example is not None:
            head = re.search(r'"(ht::\w+)<', pattern.replace("\\", "")).group(1)
            assert example.startswith(head + "<")


def test_tiling_indices_match_python():
    rs = _read(CRATE, "src", "source.rs")
    for trait, table in (("HipPlaneTiling", objects.PLANE_TILINGS), ("HipHorosphereTiling", objects.HOROSPHERE_TILINGS)):
        for name, idx in table.items():
            rust_name = name.capitalize()
            m = re.search(rf"impl {trait} for {rust_name} \{{\s*const INDEX: u32 = (\d+);", rs)
            assert m and int(m.group(1)) == idx, (trait, name)


@pytest.mark.parametrize("path", sorted(
    os.path.relpath(os.path.join(d, f), CRATE) for d, _, fs in os.walk(CRATE) for f in fs if f.endswith(".rs")))
def test_rust_sources_balance_brackets(path):
    text = _read(CRATE, path)
    # drop comments, string and char literals (lifetimes such as 'a stay)
    text = re.sub(r'//[^\n]*|"(?:\\.|[^"\\])*"|\'(?:\\.|[^\\\'])\'', " ", text)
    stack = []
    pairs = {")": "(", "]": "[", "}": "{"}
    for ch in text:
        if ch in "([{":
            stack.append(ch)
        elif ch in ")]}":
            assert stack and stack[-1] == pairs[ch], f"{path}: unbalanced {ch}"
            stack.pop()
    assert not stack, f"{path}: unclosed {stack[-1]}"
