"""Synthetic benchmark scenes.

All scenes are generated from a master seed (0x48595054, "HYPT"); workload k
uses `MASTER_SEED ^ k`. Generation runs in FP64 on the host, as the
reference composes its maps in FP64 before narrowing them to the device's
`real` (types/src/primitive/scalar.rs:198-207).

euclidean_spheres   320x240, 16 passes x 1 spp, 4 hops: 8 diffuse spheres
                    (centres U[-4,4]^3, radii U[0.2,1], albedo U[0.2,0.9]^3)
                    plus one emissive sphere (radiance 5), gradient sky, pinhole
                    at (0,0,9) looking down -z with a 60 degree vertical FOV.
hyperbolic_cells    640x480, 64 spp, 8 hops: 64 cells of 1 horosphere + 2
                    geodesic planes in the upper half-space, each cell placed by
                    a word of 1..6 letters over 4 fixed isometries; diffuse
                    albedo U[0.3,0.9]^3 mixed with a mirror at probability 0.3;
                    constant sky.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

from .objects import (
    EU,
    HY,
    Ball,
    Colored,
    ConstBg,
    Covered,
    Emissive,
    GradBg,
    Homogenous3,
    Horosphere,
    Lambertian,
    Mapped,
    Mixture,
    Moebius,
    ObjectVec,
    Plane,
    PointView,
    SceneImpl,
    Shift3,
    ShapeChoice,
    Specular,
    Sphere,
    hy_xshift,
    hy_yshift,
    hy_zrotate,
    hy_zshift,
)

MASTER_SEED = 0x48595054


@dataclass
class Workload:
    name: str
    scene: SceneImpl
    width: int
    height: int
    spp: int
    seed_mode: int
    seed: int


def _rng(k: int) -> np.random.Generator:
    return np.random.default_rng(MASTER_SEED ^ k)


def euclidean_spheres(width: int = 320, height: int = 240, spp: int = 16, unit_spheres: bool = False) -> Workload:
    """The Euclidean sphere scene. `unit_spheres=True` gives the variant with
    unit spheres placed by Homogenous3 maps (the reference's own entity
    types); the default uses `Ball` for the varied radii."""
    rng = _rng(1)
    objs = []
    for i in range(9):
        centre = rng.uniform(-4.0, 4.0, 3)
        radius = rng.uniform(0.2, 1.0)
        albedo = rng.uniform(0.2, 0.9, 3)
        emission = [5.0, 5.0, 5.0] if i == 8 else [0.0, 0.0, 0.0]
        material = Emissive(Colored(Lambertian(), albedo), emission)
        if unit_spheres:
            objs.append(Mapped(Covered(Sphere(), material), Homogenous3(centre, (1, 0, 0, 0))))
        else:
            objs.append(Mapped(Covered(Ball(radius), material), Shift3(centre)))
    fov = math.tan(math.radians(30.0))
    view = Mapped(PointView(fov), Homogenous3((0.0, 0.0, 9.0), (1, 0, 0, 0)))
    bg = GradBg((0.0, 1.0, 0.0), [(0.8, 0.9, 1.0), (0.05, 0.05, 0.08)], 1.0)
    scene = SceneImpl(EU, view, ObjectVec(objs), bg, 4)
    return Workload("euclidean_spheres", scene, width, height, spp, 1, MASTER_SEED ^ 1)


def _translate(x: complex) -> Moebius:
    return Moebius([[1, x], [0, 1]])


def _inversion() -> Moebius:
    # p -> -p^-1: the horosphere z = 1 maps to the sphere of diameter 1
    # tangent to the boundary at the origin.
    return Moebius([[0, 1], [-1, 0]])


def _scale(s: float) -> Moebius:
    return hy_zshift(math.log(s))


CELL_GENERATORS = (
    hy_xshift(0.9),
    hy_yshift(0.9),
    hy_xshift(-0.7),
    hy_zrotate(2.0 * math.pi / 5.0),
)


def hyperbolic_cells(width: int = 640, height: int = 480, spp: int = 64, cells: int = 64,
                     hops: int = 8) -> Workload:
    """The hyperbolic cell scene; every primitive's map is the FP64 product
    word * base, normalised to det 1."""
    rng = _rng(2)
    choice = [Plane(), Horosphere()]
    objs = []
    for _ in range(cells):
        word = Moebius.identity()
        for _ in range(int(rng.integers(1, 7))):
            word = word.chain(CELL_GENERATORS[int(rng.integers(0, 4))])
        # Cell layout below the camera: a horosphere of diameter dh at the
        # cell centre and two half-sphere planes beside it.
        dh = rng.uniform(0.15, 0.35)
        locals_ = [
            (1, _translate(0.0).chain(_scale(dh)).chain(_inversion())),
            (0, _translate(complex(*rng.uniform(-0.5, 0.5, 2))).chain(_scale(rng.uniform(0.1, 0.3)))),
            (0, _translate(complex(*rng.uniform(-0.5, 0.5, 2))).chain(_scale(rng.uniform(0.1, 0.3)))),
        ]
        for kind, local in locals_:
            m = word.chain(local).normalized()
            albedo = rng.uniform(0.3, 0.9, 3)
            material = Mixture([(Colored(Lambertian(), albedo), 0.7), (Specular(), 0.3)])
            objs.append(Mapped(Covered(ShapeChoice(choice, kind), material), m))
    view = Mapped(PointView(1.0), hy_zshift(math.log(2.0)))
    bg = ConstBg((1.0, 1.0, 1.0))
    scene = SceneImpl(HY, view, ObjectVec(objs), bg, hops)
    return Workload("hyperbolic_cells", scene, width, height, spp, 1, MASTER_SEED ^ 2)
