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
from typing import List, Optional, Tuple

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
    # Scene maps given as words over generators (GPU-composable), if any.
    word_maps: Optional["WordMaps"] = None


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
                     hops: int = 8, primitives: Optional[int] = None) -> Workload:
    """The hyperbolic cell scene; every primitive's map is the FP64 product
    word * base with det = 1 reprojection after each factor
    (maps.compose_word, bit-identical to the GPU composition). `primitives`
    keeps only the first that many (the cells needed are generated)."""
    if primitives is not None:
        cells = (primitives + 2) // 3
    from .maps import WordMaps, compose_word

    rng = _rng(2)
    choice = [Plane(), Horosphere()]
    objs = []
    wm = WordMaps(list(CELL_GENERATORS), [], [], [])
    for _ in range(cells):
        letters = [int(rng.integers(0, 4)) for _ in range(int(rng.integers(1, 7)))]
        # Cell layout below the camera: a horosphere of diameter dh at the
        # cell centre and two half-sphere planes beside it.
        dh = rng.uniform(0.15, 0.35)
        locals_ = [
            (1, _translate(0.0).chain(_scale(dh)).chain(_inversion())),
            (0, _translate(complex(*rng.uniform(-0.5, 0.5, 2))).chain(_scale(rng.uniform(0.1, 0.3)))),
            (0, _translate(complex(*rng.uniform(-0.5, 0.5, 2))).chain(_scale(rng.uniform(0.1, 0.3)))),
        ]
        for kind, local in locals_:
            m = compose_word(CELL_GENERATORS, letters, local)
            wm.words.append(letters)
            wm.posts.append(local)
            wm.targets.append(m)
            albedo = rng.uniform(0.3, 0.9, 3)
            material = Mixture([(Colored(Lambertian(), albedo), 0.7), (Specular(), 0.3)])
            objs.append(Mapped(Covered(ShapeChoice(choice, kind), material), m))
    if primitives is not None:
        objs = objs[:primitives]
        wm = WordMaps(wm.gens, wm.words[:primitives], wm.posts[:primitives], wm.targets[:primitives])
    view = Mapped(PointView(1.0), hy_zshift(math.log(2.0)))
    bg = ConstBg((1.0, 1.0, 1.0))
    scene = SceneImpl(HY, view, ObjectVec(objs), bg, hops)
    return Workload("hyperbolic_cells", scene, width, height, spp, 1, MASTER_SEED ^ 2, wm)


def sdf_shapes(width: int = 1920, height: int = 1080, spp: int = 256, shapes: int = 1024, hops: int = 16,
               extent: float = 10.0, camera_z: float = 26.0) -> Workload:
    """The sphere-traced scene: `shapes` SDF primitives (sphere : box : torus =
    1 : 1 : 1 by index mod 3) with sizes s ~ U[0.1, 1] (sphere radius s; box
    half extents s * U[0.5, 1]^3; torus R = s, r = s * U[0.2, 0.4]), uniformly
    random rotations, centres U[-extent, extent]^3; materials diffuse (albedo
    U[0.3, 0.9]^3) / mirror / glass (IOR 1.5) with probabilities 0.6 / 0.25 /
    0.15 as one-hot mixtures, 2 % of the shapes emissive (radiance 4); gradient
    sky; pinhole at (0, 0, camera_z) looking down -z, 40 degree vertical FOV.
    Russian roulette from the 3rd hit is a program option
    (processing.program_options(rr_depth=3))."""
    from .objects import Refractive, SdfBox, SdfSphere, SdfTorus

    rng = _rng(4)
    objs = []
    for i in range(shapes):
        kind = i % 3
        size = rng.uniform(0.1, 1.0)
        variants = [SdfSphere(1.0), SdfBox((1.0, 1.0, 1.0)), SdfTorus(1.0, 0.25)]
        if kind == 0:
            variants[0] = SdfSphere(size)
        elif kind == 1:
            variants[1] = SdfBox(size * rng.uniform(0.5, 1.0, 3))
        else:
            variants[2] = SdfTorus(size, size * rng.uniform(0.2, 0.4))
        q = rng.normal(size=4)
        q = q / np.linalg.norm(q)
        centre = rng.uniform(-extent, extent, 3)
        u = rng.uniform()
        onehot = (1.0, 0.0, 0.0) if u < 0.6 else (0.0, 1.0, 0.0) if u < 0.85 else (0.0, 0.0, 1.0)
        albedo = rng.uniform(0.3, 0.9, 3)
        emission = (4.0, 4.0, 4.0) if rng.uniform() < 0.02 else (0.0, 0.0, 0.0)
        material = Emissive(Mixture([(Colored(Lambertian(), albedo), onehot[0]), (Specular(), onehot[1]),
                                     (Refractive(1.5), onehot[2])]), emission)
        objs.append(Mapped(Covered(ShapeChoice(variants, kind), material), Homogenous3(centre, q)))
    fov = math.tan(math.radians(20.0))
    view = Mapped(PointView(fov), Homogenous3((0.0, 0.0, camera_z), (1, 0, 0, 0)))
    bg = GradBg((0.0, 1.0, 0.0), [(1.0, 1.0, 1.0), (0.1, 0.1, 0.12)], 1.0)
    scene = SceneImpl(EU, view, ObjectVec(objs), bg, hops)
    return Workload("sdf_shapes", scene, width, height, spp, 1, MASTER_SEED ^ 4)


def _unpack_color(rgb: int):
    """0xRRGGBB -> linear colour, as main/examples/hy.rs:27-35 (gamma 2.2)."""
    return [(((rgb >> s) & 0xFF) / 255.0) ** 2.2 for s in (16, 8, 0)]


def reference_euclidean_example(width: int = 800, height: int = 600, spp: int = 1) -> Workload:
    """The reference's Euclidean example scene (main/examples/eu.rs:30-136):
    a sphere (refractive / specular mixture), a cube and a plane, gradient sky,
    camera at its initial controller position (0, 0.5, 2), 4 hops."""
    from .objects import Cube, Refractive

    choice = [Plane(), Sphere(), Cube()]

    def mixture(d, s, r):
        return Mixture([(Colored(Lambertian(), d[0]), d[1]), (Specular(), s), (Colored(Refractive(r[0]), r[1]), r[2])])

    objs = [
        Mapped(Covered(ShapeChoice(choice, 1), mixture(([1.0, 0.2, 0.2], 0.0), 0.1, (1.2, [1.0, 1.0, 0.2], 0.9))),
               Shift3((0.0, 1.0, 0.0))),
        Mapped(Covered(ShapeChoice(choice, 2), mixture(([0.2, 0.8, 0.8], 1.0), 0.0, (1.0, [1.0, 1.0, 1.0], 0.0))),
               Shift3((0.0, -1.0, 0.0))),
        Mapped(Covered(ShapeChoice(choice, 0), mixture(([1.0, 1.0, 1.0], 0.9), 0.1, (1.0, [1.0, 1.0, 1.0], 0.0))),
               Shift3((0.0, 0.0, -1.0))),
    ]
    view = Mapped(PointView(1.0), Homogenous3((0.0, 0.5, 2.0), (1, 0, 0, 0)))
    bg = GradBg((0.0, 1.0, 0.0), [(1.0, 1.0, 1.0), (0.0, 0.0, 0.0)], 2.4)
    scene = SceneImpl(EU, view, ObjectVec(objs), bg, 4)
    return Workload("reference_euclidean_example", scene, width, height, spp, 0, 0xDEADBEEFDEADBEEF)


def reference_hyperbolic_example(width: int = 800, height: int = 600, spp: int = 1) -> Workload:
    """The reference's hyperbolic example scene (main/examples/hy.rs:37-183):
    two tiled horospheres (hexagonal, square) and two tiled planes
    (pentastar, pentagonal) with emissive borders, constant sky, camera at its
    initial controller map rotate_z(5pi/6) rotate_x(2pi/5) shift_z(2), 3 hops.
    The reference builds its maps with vecmat/ccgeom; here they are built in
    FP64 from the generator formulas of geometry/hyperbolic.cc:35-62."""
    from .objects import ObjectChoice, TiledHorosphere, TiledPlane, Transparent, hy_xrotate

    def make_material(diffuse, specularity, transparency, emission=None):
        mix = Mixture([(Colored(Lambertian(), diffuse), 1.0 - specularity - transparency), (Specular(), specularity),
                       (Transparent(), transparency)])
        return Emissive(mix, emission if emission is not None else [0.0, 0.0, 0.0])

    border = make_material(_unpack_color(0xE4E4E4), 0.0, 0.0, [1.0, 1.0, 1.0])
    proto_m = make_material([0.5, 0.5, 0.5], 0.0, 0.0)
    protos = [TiledPlane([proto_m] * 2, 0.0, 0.0, proto_m, "pentastar"),
              TiledPlane([proto_m] * 2, 0.0, 0.0, proto_m, "pentagonal"),
              TiledHorosphere([proto_m] * 3, 1.0, 0.0, proto_m, "hexagonal"),
              TiledHorosphere([proto_m] * 4, 1.0, 0.0, proto_m, "square")]

    def choice(index, value):
        variants = list(protos)
        variants[index] = value
        return ObjectChoice(variants, index)

    objs = [
        Mapped(choice(2, TiledHorosphere([make_material(_unpack_color(c), 0.1, 0.1) for c in
                                          (0xFE0000, 0xFFAA01, 0x35ADAE)], 0.5, 0.02, border, "hexagonal")),
               Moebius.identity()),
        Mapped(choice(3, TiledHorosphere([make_material(_unpack_color(c), 0.1, 0.1) for c in
                                          (0xFE7401, 0xFE0000, 0xFFAA01, 0xFED601)], 0.5, 0.02, border, "square")),
               Moebius([[1.0, math.sqrt(2.0)], [0.0, 1.0]]).chain(hy_xrotate(math.pi))),
        Mapped(choice(0, TiledPlane([make_material(_unpack_color(c), 0.1, 0.0) for c in (0xFE7401, 0x35ADAE)],
                                    float("nan"), 0.01, border, "pentastar")),
               Moebius.identity()),
        Mapped(choice(1, TiledPlane([make_material(_unpack_color(c), 0.1, 0.0) for c in (0xFE0000, 0xFED601)],
                                    float("nan"), 0.02, border, "pentagonal")),
               Moebius([[1.0, 2.0j], [0.0, 1.0]])),
    ]
    cam = hy_zrotate(5.0 * math.pi / 6.0).chain(hy_xrotate(2.0 * math.pi / 5.0)).chain(hy_zshift(2.0))
    view = Mapped(PointView(1.0), cam)
    scene = SceneImpl(HY, view, ObjectVec(objs), ConstBg(_unpack_color(0xEEEEEE)), 3)
    return Workload("reference_hyperbolic_example", scene, width, height, spp, 0, 0xDEADBEEFDEADBEEF)


# ------------------------------------------------------------ batched jobs

def _mix(rng):
    """One-hot mixture: diffuse (albedo U[0.3, 0.9]^3) / mirror / glass
    (IOR 1.5) with probabilities 0.6 / 0.25 / 0.15."""
    from .objects import Refractive

    u = rng.uniform()
    onehot = (1.0, 0.0, 0.0) if u < 0.6 else (0.0, 1.0, 0.0) if u < 0.85 else (0.0, 0.0, 1.0)
    albedo = rng.uniform(0.3, 0.9, 3)
    return Mixture([(Colored(Lambertian(), albedo), onehot[0]), (Specular(), onehot[1]), (Refractive(1.5), onehot[2])])


BATCH_MAX_HOPS = 12


def batch_euclidean_scene(rng, prims: int, hops: int = BATCH_MAX_HOPS) -> SceneImpl:
    """Batch type 0: balls (radius U[0.2, 0.8], centres U[-5, 5]^3) under a
    gradient sky, pinhole at (0, 0, 14), 60 degree vertical FOV."""
    objs = [Mapped(Covered(Ball(rng.uniform(0.2, 0.8)), _mix(rng)), Shift3(rng.uniform(-5.0, 5.0, 3)))
            for _ in range(prims)]
    view = Mapped(PointView(math.tan(math.radians(30.0))), Homogenous3((0.0, 0.0, 14.0), (1, 0, 0, 0)))
    bg = GradBg((0.0, 1.0, 0.0), [(1.0, 1.0, 1.0), (0.08, 0.08, 0.1)], 1.0)
    return SceneImpl(EU, view, ObjectVec(objs), bg, hops)


def batch_hyperbolic_scene(rng, prims: int, hops: int = BATCH_MAX_HOPS) -> SceneImpl:
    """Batch type 1: cells of 1 horosphere + 2 planes as in the cell scene,
    each primitive's map a word of 1..6 generators (FP64, reprojected) times
    its local map; constant sky; camera at zshift(ln 2)."""
    from .maps import compose_word

    choice = [Plane(), Horosphere()]
    objs = []
    letters, dh = [], 0.25
    for i in range(prims):
        if i % 3 == 0:
            letters = [int(rng.integers(0, 4)) for _ in range(int(rng.integers(1, 7)))]
            dh = rng.uniform(0.15, 0.35)
            kind, local = 1, _translate(0.0).chain(_scale(dh)).chain(_inversion())
        else:
            kind = 0
            local = _translate(complex(*rng.uniform(-0.5, 0.5, 2))).chain(_scale(rng.uniform(0.1, 0.3)))
        m = compose_word(CELL_GENERATORS, letters, local)
        objs.append(Mapped(Covered(ShapeChoice(choice, kind), _mix(rng)), m))
    view = Mapped(PointView(1.0), hy_zshift(math.log(2.0)))
    return SceneImpl(HY, view, ObjectVec(objs), ConstBg((1.0, 1.0, 1.0)), hops)


@dataclass
class BatchJobSpec:
    type: int  # 0 Euclidean, 1 hyperbolic
    scene: SceneImpl
    width: int
    height: int
    spp: int
    depth: int
    seed: int


def batch_jobs(n: int = 32, min_size=(128, 128), max_size=(1024, 768), spp_range=(1, 64), depth_range=(4, 12),
               prim_range=(8, 256), depths=None) -> List[BatchJobSpec]:
    """The batched workload: n independent jobs, half Euclidean, half
    hyperbolic (alternating), width and height log-uniform between min_size
    and max_size, spp U{spp_range}, depth U{depth_range} (or drawn from
    `depths`), U{prim_range} primitives with the diffuse / mirror / glass mix.
    Seeds: per-pixel hash seeding with MASTER_SEED ^ 3 ^ (job << 8)."""
    rng = _rng(3)
    jobs = []
    for j in range(n):
        w = int(round(math.exp(rng.uniform(math.log(min_size[0]), math.log(max_size[0])))))
        h = int(round(math.exp(rng.uniform(math.log(min_size[1]), math.log(max_size[1])))))
        spp = int(rng.integers(spp_range[0], spp_range[1] + 1))
        depth = int(rng.integers(depth_range[0], depth_range[1] + 1)) if depths is None else int(
            depths[int(rng.integers(0, len(depths)))])
        prims = int(rng.integers(prim_range[0], prim_range[1] + 1))
        t = j % 2
        scene = batch_euclidean_scene(rng, prims) if t == 0 else batch_hyperbolic_scene(rng, prims)
        jobs.append(BatchJobSpec(t, scene, w, h, spp, depth, MASTER_SEED ^ 3 ^ (j << 8)))
    return jobs
