"""Host side of the FP64 Moebius-word composition (hypertrace_amd/maps.py):
the host mirror against numpy's complex arithmetic and the det = 1 property,
the principal square root on every half-plane, and the byte offsets the GPU
writes to."""
import numpy as np

from hypertrace_amd import scenes
from hypertrace_amd.maps import _csqrt, compose_word
from hypertrace_amd.objects import Moebius, field_offsets
from hypertrace_amd.types import X32_F32, X32_F64


def _det(m: np.ndarray) -> complex:
    return m[0, 0] * m[1, 1] - m[0, 1] * m[1, 0]


def test_principal_sqrt_all_half_planes():
    rng = np.random.default_rng(7)
    zs = list(rng.normal(size=(200, 2)) * np.exp(rng.uniform(-20, 20, (200, 1))))
    # |z| within 1e-150..1e150 (x*x + y*y must neither underflow nor overflow)
    zs += [(-4.0, 0.0), (-4.0, -0.0), (4.0, 0.0), (0.0, 9.0), (0.0, -9.0), (-1e-100, 1e-100)]
    for x, y in zs:
        r = complex(*_csqrt((float(x), float(y))))
        z = complex(x, y)
        assert r.real >= 0.0
        assert abs(r * r - z) <= 4e-16 * abs(z), (z, r)
        # principal branch, as numpy's
        if not (x < 0 and y == 0):
            assert abs(r - np.sqrt(z)) <= 4e-16 * abs(r), (z, r, np.sqrt(z))


def test_compose_word_matches_complex_products_and_has_det_one():
    wl = scenes.hyperbolic_cells(cells=64)
    wm = wl.word_maps
    assert len(wm.targets) == 192
    worst = 0.0
    for letters, post, m in zip(wm.words, wm.posts, wm.targets):
        ref = np.eye(2, dtype=np.complex128)
        for k in letters:
            ref = ref @ wm.gens[k].m
        ref = ref @ post.m
        ref = ref / np.sqrt(_det(ref))
        # the same map up to the sign of the square root
        s = 1.0 if np.abs(ref - m.m).max() < np.abs(ref + m.m).max() else -1.0
        worst = max(worst, np.abs(s * ref - m.m).max() / np.abs(ref).max())
        assert abs(_det(m.m) - 1.0) < 1e-14
    assert worst < 1e-13, worst


def test_fp64_reprojection_beats_fp32_composition():
    """Why the words are composed in FP64: narrowed to FP32 once, a map's det
    is 1 to within FP32 rounding (relative to |M|^2) whatever the word length;
    the same products done in FP32 drift off det = 1 as the word grows (the
    kernel's inverse assumes det = 1, transform/moebius.cc:66-72)."""
    gens = list(scenes.CELL_GENERATORS)
    rng = np.random.default_rng(3)
    drift64, drift32 = 0.0, 0.0
    for _ in range(200):
        letters = [int(k) for k in rng.integers(0, 4, 16)]
        m = compose_word(gens, letters, Moebius.identity())
        f = m.m.astype(np.complex64).astype(np.complex128)
        drift64 = max(drift64, abs(_det(f) - 1.0) / np.abs(f).max() ** 2)
        acc = np.eye(2, dtype=np.complex64)
        for k in letters:
            acc = (acc @ gens[k].m.astype(np.complex64)).astype(np.complex64)
        a = acc.astype(np.complex128)
        drift32 = max(drift32, abs(_det(a) - 1.0) / np.abs(a).max() ** 2)
    assert drift64 < 3e-7, drift64
    assert drift32 > 4 * drift64, (drift32, drift64)


def test_map_offsets_point_at_the_serialized_maps():
    wl = scenes.hyperbolic_cells(cells=16)
    for cfg, dt in ((X32_F32, np.float32), (X32_F64, np.float64)):
        data = wl.scene.to_bytes(cfg)
        offs = field_offsets(wl.scene, cfg, wl.word_maps.targets)
        assert len(set(offs)) == len(offs)
        for o, m in zip(offs, wl.word_maps.targets):
            assert o % (8 * cfg.real) == 0
            got = np.frombuffer(data, dt, count=8, offset=o)
            assert np.array_equal(got, np.array(m.reals(), dt))


def test_compose_word_single_letter_and_identity_post():
    g = Moebius([[1.5 + 0.5j, 2.0], [0.25j, 1.0 - 0.5j]])
    m = compose_word([g], [0], Moebius.identity())
    ref = g.m / np.sqrt(_det(g.m))
    assert np.allclose(m.m, ref, rtol=0, atol=1e-15) or np.allclose(m.m, -ref, rtol=0, atol=1e-15)
