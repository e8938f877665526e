"""Moebius maps of a scene composed from words over generators, in FP64 with
det = 1 reprojection after every product.

The reference composes its hyperbolic isometries on the host in f64 and
narrows them to the device `real` when it serializes the scene
(types/src/primitive/scalar.rs:198-207); its kernel normalises a matrix by
dividing by sqrt(det) (c22_normalize, kernel/ocl/algebra/matrix_complex.cc:50-52).
Here the same composition runs on the GPU (`ht_mo_compose`), writing each
element's map straight into the scene buffer. `compose_word` is its host
mirror: the same IEEE double operations in the same order, so both give the
same bits.

    M = N(...N(N(I G[w0]) G[w1]) ... P),   N(m) = m / sqrt(det m)

The square root is the principal one, evaluated algebraically with the
cancellation-free branch for each half-plane (not the polar form), so only
+, -, *, / and sqrt appear.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from .objects import Moebius, field_offsets
from .types import Config

WORD_MAX = 16  # HT_MO_WORD_MAX


def _cmul(a, b):
    return (a[0] * b[0] - a[1] * b[1], a[0] * b[1] + a[1] * b[0])


def _cadd(a, b):
    return (a[0] + b[0], a[1] + b[1])


def _csub(a, b):
    return (a[0] - b[0], a[1] - b[1])


def _csqrt(z):
    r = math.sqrt(z[0] * z[0] + z[1] * z[1])
    if z[0] >= 0.0:
        re = math.sqrt((r + z[0]) * 0.5)
        return (re, z[1] / (re + re))
    im = math.sqrt((r - z[0]) * 0.5)
    if z[1] < 0.0:
        im = -im
    return (z[1] / (im + im), im)


def _mul(p, q):
    a, b, c, d = p
    e, f, g, h = q
    return (_cadd(_cmul(a, e), _cmul(b, g)), _cadd(_cmul(a, f), _cmul(b, h)),
            _cadd(_cmul(c, e), _cmul(d, g)), _cadd(_cmul(c, f), _cmul(d, h)))


def _renorm(m):
    a, b, c, d = m
    s = _csqrt(_csub(_cmul(a, d), _cmul(b, c)))
    s2 = s[0] * s[0] + s[1] * s[1]
    inv = (s[0] / s2, -s[1] / s2)
    return (_cmul(a, inv), _cmul(b, inv), _cmul(c, inv), _cmul(d, inv))


def _pairs(m: Moebius):
    r = [float(x) for x in m.reals()]
    return ((r[0], r[1]), (r[2], r[3]), (r[4], r[5]), (r[6], r[7]))


def compose_word(gens: Sequence[Moebius], letters: Sequence[int], post: Moebius) -> Moebius:
    """Host mirror of the GPU composition: the same bits as ht_mo_compose."""
    m = ((1.0, 0.0), (0.0, 0.0), (0.0, 0.0), (1.0, 0.0))
    for k in letters:
        m = _renorm(_mul(m, _pairs(gens[k])))
    m = _renorm(_mul(m, _pairs(post)))
    return Moebius([[complex(*m[0]), complex(*m[1])], [complex(*m[2]), complex(*m[3])]])


@dataclass
class WordMaps:
    """Maps of a scene given as words: element i's map object `targets[i]`
    (a Moebius instance inside the scene) equals compose_word(gens,
    words[i], posts[i])."""
    gens: List[Moebius]
    words: List[List[int]]
    posts: List[Moebius]
    targets: List[Moebius]

    def arrays(self, scene, cfg: Config):
        n = len(self.targets)
        offsets = np.array(field_offsets(scene, cfg, self.targets), np.uint32)
        letters = np.zeros((n, WORD_MAX), np.uint8)
        lengths = np.zeros(n, np.uint8)
        for i, w in enumerate(self.words):
            if len(w) > WORD_MAX:
                raise ValueError(f"word {i} longer than {WORD_MAX}")
            letters[i, :len(w)] = w
            lengths[i] = len(w)
        gens = np.array([m.reals() for m in self.gens], np.float64).reshape(-1, 8)
        posts = np.array([m.reals() for m in self.posts], np.float64).reshape(-1, 8)
        return offsets, letters, lengths, gens, posts


def compose_on_device(ctx, buffer, scene, word_maps: WordMaps) -> None:
    """Overwrite the maps of `buffer` (an EntityBuffer holding `scene`) with
    their GPU FP64 composition."""
    from . import native

    offsets, letters, lengths, gens, posts = word_maps.arrays(scene, ctx.config)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    native.check(native.lib().ht_mo_compose(ctx.handle, buffer.handle, ctx.config.real, len(offsets), p(offsets),
                                            p(letters), p(lengths), p(gens), len(gens), p(posts)))
