"""Full-path pins of the CPU oracle, where no reference test exercises the
composed path (the reference instantiates scene.inl on the host but never
calls `_sample`): closed-form cavities, brute-force nearest hit and isometry
invariance."""
import math

import numpy as np

import oracle
from hypertrace_amd.objects import (EU, HY, Absorbing, Ball, Colored, ConstBg, Covered, Emissive, GradBg, Sphere,
                                    Homogenous3, Horosphere, Lambertian, Mapped, Mixture, Moebius, ObjectVec, Plane,
                                    PointView, SceneImpl, ShapeChoice, Shift3, Specular, hy_xshift, hy_yshift,
                                    hy_zrotate, hy_zshift)
from hypertrace_amd.processing import seed_fill
from hypertrace_amd.types import X32_F32


def render(scene, w, h, passes=1, seed=123, precision=oracle.F32):
    seeds = seed_fill(w * h, 1, seed)
    c, s, _ = oracle.render(scene.spec_text(X32_F32), w, h, passes, seeds, precision=precision)
    return c, seeds


def series(e, rho, hops, start):
    """Emission accumulated in FP32 the way the path does it: at each hit
    emission += I * e with I the product of albedos so far."""
    acc = np.zeros(3, np.float32)
    i = np.ones(3, np.float32)
    e = np.float32(e)
    rho = np.asarray(rho, np.float32)
    for k in range(hops):
        if start == "emit_first":
            acc = acc + i * e
            i = i * rho
        else:
            i = i * rho
            acc = acc + i * e
    return acc


def test_euclidean_cavity_closed_form():
    # Camera inside the unit sphere: every bounce hits it again, so each path's
    # radiance is the FP32 series exactly. (A large sphere does not work in
    # FP32: the self-hit test's rounding exceeds EPS = 1e-6 and paths leak out,
    # in the reference's algorithm as here.)
    rho = (0.5, 0.7, 0.9)
    for precision in (oracle.F32, oracle.F64):
        for order in ("emit_first", "color_first"):
            mat = (Emissive(Colored(Lambertian(), rho), (2.0, 2.0, 2.0)) if order == "emit_first"
                   else Colored(Emissive(Lambertian(), (2.0, 2.0, 2.0)), rho))
            obj = ObjectVec([Mapped(Covered(Sphere(), mat), Homogenous3((0.3, -0.2, 0.1)))])
            scene = SceneImpl(EU, Mapped(PointView(1.0), Homogenous3()), obj,
                              GradBg((0, 1, 0), [(9, 9, 9), (9, 9, 9)], 1), 5)
            c, _ = render(scene, 24, 16, precision=precision)
            expect = series(2.0, rho, 5, order)
            ok = np.all(np.abs(c[..., :3] - expect) <= 2e-6 * np.abs(expect), axis=-1)
            assert ok.mean() >= 0.99, (precision, order, ok.mean())
            assert np.all(c[..., 3] == 1)


def test_hyperbolic_horoball_cavity_closed_form():
    rho = (0.6, 0.8, 0.3)
    mat = Emissive(Colored(Lambertian(), rho), (1.5, 1.5, 1.5))
    choice = ShapeChoice([Plane(), Horosphere()], 1)
    obj = ObjectVec([Mapped(Covered(choice, mat), Moebius.identity())])
    # camera at (0, 0, 2), inside the horoball z > 1
    scene = SceneImpl(HY, Mapped(PointView(1.0), hy_zshift(math.log(2.0))), obj, ConstBg((9, 9, 9)), 6)
    c, _ = render(scene, 24, 16)
    expect = series(1.5, rho, 6, "emit_first")
    ok = np.all(np.abs(c[..., :3] - expect) <= 2e-6 * np.abs(expect), axis=-1)
    assert ok.mean() >= 0.99


def camera_rays(seeds, w, h, fov):
    """Camera ray directions from the seeds: two LCG draws per pixel (x then
    y jitter), pixel mapping of render.cl:12-18, view/point/impl.inl:18-25."""
    dirs = np.zeros((h, w, 3))
    for y in range(h):
        for x in range(w):
            s = int(seeds[x + w * y])
            s = (1103515245 * s + 12345) & 0xFFFFFFFF
            ux = s / 2 ** 32
            s = (1103515245 * s + 12345) & 0xFFFFFFFF
            uy = s / 2 ** 32
            px = ((2 * x + 1) - w) / h + (2 / h) * (ux - 0.5)
            py = (h - (2 * y + 1)) / h + (2 / h) * (uy - 0.5)
            d = np.array([px, py, -1 / fov])
            dirs[y, x] = d / np.linalg.norm(d)
    return dirs


def test_euclidean_nearest_hit_brute_force():
    rng = np.random.default_rng(21)
    n = 12
    centres = rng.uniform(-3, 3, (n, 3))
    radii = rng.uniform(0.3, 1.2, n)
    items = []
    for i in range(n):
        emit = ((i + 1) / n, 1.0 - (i + 1) / (2 * n), 0.25)
        items.append(Mapped(Covered(Ball(radii[i]), Emissive(Absorbing(), emit)), Shift3(centres[i])))
    cam = np.array([0.0, 0.0, 10.0])
    fov = 0.6
    scene = SceneImpl(EU, Mapped(PointView(fov), Homogenous3(cam)), ObjectVec(items),
                      GradBg((0, 1, 0), [(0, 0, 1), (0, 0, 1)], 1), 3)
    w, h = 48, 36
    c, seeds = render(scene, w, h)
    dirs = camera_rays(seeds, w, h, fov)
    agree = 0
    for y in range(h):
        for x in range(w):
            d = dirs[y, x]
            best, best_t = -1, np.inf
            for i in range(n):
                oc = cam - centres[i]
                b = -np.dot(d, oc)
                disc = radii[i] ** 2 - np.sum((oc + b * d) ** 2)
                if disc < 0:
                    continue
                t = b - math.sqrt(disc)
                if t < 0:
                    t = b + math.sqrt(disc)
                if 0 < t < best_t:
                    best, best_t = i, t
            want = np.array([0, 0, 1.0]) if best < 0 else np.array([(best + 1) / n, 1 - (best + 1) / (2 * n), 0.25])
            agree += np.allclose(c[y, x, :3], want, atol=1e-6)
    assert agree / (w * h) >= 0.995
    # the image is not trivial: several objects and the sky are visible
    assert len(np.unique(np.round(c[..., 0], 4))) >= 4


def test_hyperbolic_isometry_invariance():
    rng = np.random.default_rng(22)
    gens = [hy_xshift(0.7), hy_yshift(-0.5), hy_zrotate(1.1)]
    items = []
    for i in range(10):
        m = Moebius.identity()
        for _ in range(3):
            m = m.chain(gens[int(rng.integers(0, 3))])
        m = m.chain(Moebius([[1, complex(*rng.uniform(-0.6, 0.6, 2))], [0, 1]])).chain(hy_zshift(math.log(rng.uniform(0.2, 0.5))))
        kind = int(rng.integers(0, 2))
        mat = Mixture([(Colored(Lambertian(), rng.uniform(0.3, 0.9, 3)), 0.7), (Specular(), 0.3)])
        items.append((kind, mat, m.normalized()))
    g = hy_xshift(0.4).chain(hy_zrotate(0.3)).chain(hy_yshift(0.2)).normalized()
    cam = hy_zshift(math.log(2.0))

    def scene(pre):
        objs = [Mapped(Covered(ShapeChoice([Plane(), Horosphere()], k), mat), pre.chain(m).normalized())
                for k, mat, m in items]
        return SceneImpl(HY, Mapped(PointView(1.0), pre.chain(cam).normalized()), ObjectVec(objs), ConstBg((1, 1, 1)), 4)

    a, _ = render(scene(Moebius.identity()), 40, 30, precision=oracle.F64)
    b, _ = render(scene(g), 40, 30, precision=oracle.F64)
    ok = np.all(np.abs(a - b) <= 1e-6 * np.abs(a) + 1e-7, axis=-1)
    assert ok.mean() >= 0.99
    assert (a[..., 0] < 0.999).mean() > 0.2  # objects are in view


def test_culled_flop_count_keeps_the_render():
    """The culled-scan flop count (the roofline's numerator) runs the FP64
    cull before the exact tests; its render must equal the plain count's, hit
    for hit, and it must count fewer FP32 flops plus some FP64 ones, in both
    geometries."""
    from hypertrace_amd import scenes

    for wl in (scenes.hyperbolic_cells(64, 48, 1, cells=16), scenes.euclidean_spheres(64, 48, 1)):
        spec = wl.scene.spec_text(X32_F32)
        seeds = seed_fill(64 * 48, wl.seed_mode, wl.seed)
        c0, s0, p32, p64 = oracle.count_flops(spec, 64, 48, 2, seeds, False)
        c1, s1, c32, c64 = oracle.count_flops(spec, 64, 48, 2, seeds, True)
        assert np.array_equal(c0, c1) and np.array_equal(s0, s1), wl.name
        assert p64 == 0 and c64 > 0 and c32 < p32, (wl.name, p32, c32, c64)
        # the two-level scan (groups of g; 1000 + g with the group queue):
        # the same render, the same exact tests, fewer cull flops than the
        # flat scan where the scene has the tree's 16 elements
        for mode in (4, 8, 1004, 1008):
            cg, sg, g32, g64 = oracle.count_flops(spec, 64, 48, 2, seeds, mode)
            assert np.array_equal(c0, cg) and np.array_equal(s0, sg), (wl.name, mode)
            assert g32 == c32, (wl.name, mode)
            if len(wl.scene.object.items) >= 16:
                assert 0 < g64 < c64, (wl.name, mode, g64, c64)
            else:
                assert g64 == c64, (wl.name, mode)


def test_hit_order_count_keeps_the_render():
    """The hit-order scan's count (2000 + g: each passing pair's crossing
    interval, exact tests only of the pairs its filter keeps) renders exactly
    what the group-queue count and the plain scan render, on the cell scene
    (192 primitives, the tree's groups), the 256-primitive cell scene and the
    reference's hyperbolic example (fewer elements than the tree's threshold:
    the flat scan, no interval), and it runs fewer exact tests (FP32 flops)
    on the cell scenes."""
    from hypertrace_amd import scenes

    for wl, passes in ((scenes.hyperbolic_cells(160, 120), 3),
                       (scenes.hyperbolic_cells(192, 108, 1, primitives=256), 2),
                       (scenes.reference_hyperbolic_example(64, 48), 2),
                       (scenes.euclidean_spheres(64, 48, 1), 2)):
        w, h = wl.width, wl.height
        spec = wl.scene.spec_text(X32_F32)
        seeds = seed_fill(w * h, wl.seed_mode, wl.seed)
        c0, s0, _, _ = oracle.count_flops(spec, w, h, passes, seeds, False)
        cq, sq, q32, q64 = oracle.count_flops(spec, w, h, passes, seeds, 1004)
        ch, sh, h32, h64 = oracle.count_flops(spec, w, h, passes, seeds, 2004)
        assert np.array_equal(c0.view(np.uint32), ch.view(np.uint32)) and np.array_equal(s0, sh), wl.name
        assert np.array_equal(cq.view(np.uint32), ch.view(np.uint32)), wl.name
        if wl.scene.geo == "Hy" and len(wl.scene.object.items) >= 16:
            assert h32 < 0.7 * q32 and h64 > q64, (wl.name, q32, h32, q64, h64)
        else:
            assert (h32, h64) == (q32, q64), wl.name
