"""Pins of the CPU oracle's materials, mixture selection, backgrounds and
tiled objects (no GPU). Each check is the reference's own unit test of that
function restated (material/specular.cc:21-45, transparent.cc:20-43,
absorbing.cc:20-42, lambertian.cc:28-69, material/test/modifier.cpp:49-106)
or a property the mathematics fixes (Snell's law and total internal
reflection for refractive.cc:3-22, which the reference does not test; the
interval a uniform falls in for mixture/impl.inl:15-27; the end points and
the half-angle form of the gradient sky, gradient.cc:3-13; the symmetries,
tile size and border width of the {5,4} tilings and the lattices of the
horosphere tilings, tiled_plane/impl.inl:47-136, tiled_horosphere/impl.inl:52-98),
chosen so that a dropped term, a wrong sign or a swapped operand fails."""
import numpy as np
import pytest

import oracle

RNG = np.random.default_rng(0xABCD)
ATTEMPTS = 1024  # the reference's TEST_ATTEMPTS (algebra/real.hh:33)


def unit(rng, n=None):
    v = rng.normal(size=(3,) if n is None else (n, 3))
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def incoming(rng):
    """A unit normal and a unit direction arriving against it (dir . n <= 0),
    as the reference's material tests draw them."""
    n, d = unit(rng), unit(rng)
    if d @ n > 0:
        d = -d
    return n, d


def lcg(s):
    return (1103515245 * s + 12345) & 0xFFFFFFFF


# ------------------------------------------------------------ specular, transparent, absorbing, modifiers
def test_specular_rule():
    for _ in range(ATTEMPTS):
        n, d = incoming(RNG)
        on, od, oi, oe, st = oracle.material_interact("specular", n, d, rng_state=77)
        assert on and st == 77  # goes on, draws nothing
        assert np.all(oe == 0) and np.all(oi == 1)
        assert abs(np.linalg.norm(od) - 1) < 1e-12
        # specular.cc:42-43: normalize(out - dir) == normal
        diff = od - d
        np.testing.assert_allclose(diff / np.linalg.norm(diff), n, atol=1e-9)
        # mirror law: the tangential part is kept, the normal part flips
        np.testing.assert_allclose(od - (od @ n) * n, d - (d @ n) * n, atol=1e-12)
        assert abs(od @ n + d @ n) < 1e-12


def test_transparent_and_absorbing_rules():
    for _ in range(ATTEMPTS // 4):
        n, d = incoming(RNG)
        on, od, _, oe, st = oracle.material_interact("transparent", n, d, rng_state=5)
        assert on and st == 5 and np.all(oe == 0)
        np.testing.assert_array_equal(od, d)
        on, _, _, oe, st = oracle.material_interact("absorbing", n, d, rng_state=5)
        assert not on and st == 5 and np.all(oe == 0)


def test_colored_and_emissive_modifiers():
    """modifier.cpp:49-106: Colored multiplies the intensity, Emissive adds
    intensity x emission to the emission; direction and the other quantity
    unchanged. Nested, each sees the intensity its outer modifiers left."""
    for _ in range(ATTEMPTS // 4):
        n, d = incoming(RNG)
        color = RNG.uniform(0, 1, 3).astype(np.float32)
        mat = RNG.uniform(0, 1, 3).astype(np.float32)
        base = RNG.uniform(0, 1, 3).astype(np.float32)
        spec = "colored %r %r %r transparent" % tuple(float(x) for x in mat)
        on, od, oi, oe, _ = oracle.material_interact(spec, n, d, intensity=color)
        assert on and np.all(oe == 0)
        np.testing.assert_array_equal(od, d)
        np.testing.assert_allclose(oi, color * mat, rtol=1e-6)
        spec = "emissive %r %r %r transparent" % tuple(float(x) for x in mat)
        on, od, oi, oe, _ = oracle.material_interact(spec, n, d, intensity=color, emission=base)
        assert on
        np.testing.assert_array_equal(od, d)
        np.testing.assert_array_equal(oi, color)
        np.testing.assert_allclose(oe, base + color * mat, rtol=1e-6)
    # order: emission sees the intensity before (outer emissive) or after
    # (inner emissive) the colour
    c, e, i0 = np.float32([0.5, 0.25, 1.0]), np.float32([2.0, 4.0, 8.0]), np.float32([1.0, 1.0, 0.5])
    n, d = incoming(RNG)
    _, _, oi, oe, _ = oracle.material_interact("emissive 2 4 8 colored 0.5 0.25 1 transparent", n, d, intensity=i0)
    np.testing.assert_allclose(oe, i0 * e)
    np.testing.assert_allclose(oi, i0 * c)
    _, _, oi, oe, _ = oracle.material_interact("colored 0.5 0.25 1 emissive 2 4 8 transparent", n, d, intensity=i0)
    np.testing.assert_allclose(oe, i0 * c * e)


# ------------------------------------------------------------ refractive: Snell's law, total internal reflection
def _angle_sin(v, n):
    return np.linalg.norm(np.cross(v, n)) / np.linalg.norm(v)


@pytest.mark.parametrize("index", [1.5, 1.33, 2.4])
def test_refractive_snell_and_total_internal_reflection(index):
    spec = f"refractive {index!r}"
    rng = np.random.default_rng(int(index * 100))  # independent of test order
    entering = leaving = tir = 0
    for _ in range(ATTEMPTS):
        n, d = unit(rng), unit(rng)
        cos_i = d @ n
        if abs(cos_i) < 1e-3:
            continue  # grazing: the reference's EPS window decides
        on, od, oi, oe, st = oracle.material_interact(spec, n, d, rng_state=9)
        assert on and st == 9 and np.all(oe == 0) and np.all(oi == 1)
        assert abs(np.linalg.norm(od) - 1) < 1e-9
        # out lies in the plane of d and n
        assert abs(np.dot(np.cross(d, n), od)) < 1e-9
        sin_i = _angle_sin(d, n)
        # d against the normal: entering the medium (n1 / n2 = 1 / index);
        # along it: leaving (n1 / n2 = index)
        ratio = 1.0 / index if cos_i < 0 else index
        if ratio * sin_i < 1.0 - 1e-6:
            # Snell: n1 sin(i) = n2 sin(t), and the path crosses the surface
            assert abs(_angle_sin(od, n) - ratio * sin_i) < 1e-9
            assert np.sign(od @ n) == np.sign(cos_i)
            # the tangential direction is kept
            t_in, t_out = d - cos_i * n, od - (od @ n) * n
            if np.linalg.norm(t_in) > 1e-6:
                assert np.dot(t_in, t_out) > 0
            entering += cos_i < 0
            leaving += cos_i > 0
        elif ratio * sin_i > 1.0 + 1e-6:
            # total internal reflection: the mirror direction
            assert cos_i > 0  # only when leaving the denser medium
            np.testing.assert_allclose(od, d - 2 * cos_i * n, atol=1e-12)
            tir += 1
    # leaving directions refract only inside the critical cone, a fraction
    # 1 - cos(asin(1 / index)) of the hemisphere (about 9 % at index 2.4)
    expect_leaving = ATTEMPTS / 2 * (1 - np.sqrt(1 - 1 / index**2))
    assert entering > 100 and leaving > expect_leaving / 2 and tir > 50


def test_refractive_normal_incidence_passes_straight():
    n = np.array([0.0, 0.0, 1.0])
    for d in (np.array([0.0, 0.0, -1.0]), np.array([0.0, 0.0, 1.0])):
        _, od, _, _, _ = oracle.material_interact("refractive 1.5", n, d)
        np.testing.assert_allclose(od, d, atol=1e-12)


# ------------------------------------------------------------ mixture: which component one uniform picks
def test_mixture_picks_the_interval_of_the_uniform():
    """One uniform u = state' / 2^32 per interaction (random.cc:8-14); the
    component whose interval [p0 + .. + p(k-1), p0 + .. + pk) contains u is
    taken, in declaration order (mixture/impl.inl:15-27); beyond the last
    interval none is (the path ends)."""
    comps = [(0.2, (1, 0, 0)), (0.5, (0, 1, 0)), (0.25, (0, 0, 1))]  # sum 0.95
    spec = "mixture 3 " + " ".join(f"{p!r} emissive {c[0]} {c[1]} {c[2]} absorbing" for p, c in comps)
    edges = np.cumsum([p for p, _ in comps])
    seen = set()
    n, d = incoming(RNG)
    for s in list(RNG.integers(0, 2 ** 32, 3000)) + [0, 0xFFFFFFFF, 858993459]:
        s = int(s)
        u = lcg(s) / 2.0 ** 32
        if np.min(np.abs(edges - u)) < 1e-9:
            continue
        k = int(np.searchsorted(edges, u, side="right"))
        on, _, _, oe, st = oracle.material_interact(spec, n, d, rng_state=s)
        assert st == lcg(s)  # exactly one draw
        assert not on  # every component absorbs
        want = np.zeros(3, np.float32) if k == len(comps) else np.float32(comps[k][1])
        np.testing.assert_array_equal(oe, want)
        seen.add(k)
    assert seen == {0, 1, 2, 3}


def test_mixture_selection_order_matters():
    # same portions, components swapped: the picks swap with them
    n, d = incoming(RNG)
    a = "mixture 2 0.5 emissive 1 0 0 absorbing 0.5 emissive 0 1 0 absorbing"
    b = "mixture 2 0.5 emissive 0 1 0 absorbing 0.5 emissive 1 0 0 absorbing"
    for s in range(200):
        ea = oracle.material_interact(a, n, d, rng_state=s)[3]
        eb = oracle.material_interact(b, n, d, rng_state=s)[3]
        np.testing.assert_array_equal(ea, eb[[1, 0, 2]])


# ------------------------------------------------------------ lambertian: cosine distribution
def _frame(n):
    a = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
    t = np.cross(n, a)
    t /= np.linalg.norm(t)
    return t, np.cross(n, t), n


def test_lambertian_cosine_distribution():
    """lambertian.cc:28-69: outgoing directions are unit, on the side of the
    normal that faces the incoming light (the normal is flipped for light
    arriving along it), and cosine-distributed: their projection onto the
    tangent plane is uniform on the unit disk (P x R equal-area cells within
    the reference's grid bar, and a chi-square bar; mean within 2 x 3 /
    sqrt(ATTEMPTS))."""
    P, R = 16, 16
    N = P * R * ATTEMPTS
    rng = np.random.default_rng(0xDEAD)
    normal = unit(rng)
    t, b, _ = _frame(normal)
    state = 0xDEADBEEF
    xy = np.zeros((N, 2))
    for i in range(N):
        d = unit(rng)
        flip = d @ normal > 0  # half the rays arrive along the normal
        ok, od, _, oe, state2 = oracle.material_interact("lambertian", normal, d, rng_state=state)
        assert lcg(lcg(state)) == state2  # two draws
        state = state2
        n_eff = -normal if flip else normal
        assert ok and np.all(oe == 0)
        assert abs(np.linalg.norm(od) - 1) < 1e-9
        assert od @ n_eff > 0
        # local frame with z along the effective normal (flip keeps handedness irrelevant for the disk)
        xy[i] = (od @ t, od @ b) if not flip else (od @ t, -(od @ b))
    conf = 3 / np.sqrt(ATTEMPTS)
    assert np.all(np.abs(xy.mean(0)) < 2 * conf)
    r = np.sqrt(np.sum(xy ** 2, 1))
    sector = np.minimum(((np.arctan2(xy[:, 1], xy[:, 0]) % (2 * np.pi)) / (2 * np.pi) * P).astype(int), P - 1)
    # the reference's DiskGrid check (random.cc:101-125, random.hh:48-64):
    # rings of equal radial step, each cell's share of N within CONF / size
    # (relative) of its area share
    ring = np.minimum((r * R).astype(int), R - 1)
    counts = np.bincount(ring * P + sector, minlength=P * R)
    size = np.repeat(((np.arange(R) + 1.0) ** 2 - np.arange(R) ** 2) / R ** 2 / P, P)
    total = counts / size / N
    assert np.all(np.abs(total - 1) <= conf / size)
    # and a calibrated one: chi-square over P x R equal-area cells (255
    # degrees of freedom: mean 255, sd 22.6) below mean + 5 sd; a uniform
    # instead of a cosine hemisphere gives ~1.8e5
    ring = np.minimum((r ** 2 * R).astype(int), R - 1)
    counts = np.bincount(ring * P + sector, minlength=P * R)
    expect = N / (P * R)
    chi2 = float(np.sum((counts - expect) ** 2 / expect))
    dof = P * R - 1
    assert chi2 < dof + 5 * np.sqrt(2 * dof), chi2


# ------------------------------------------------------------ backgrounds
def test_gradient_sky():
    """gradient.cc:3-13 with c0 != c1 and power != 1: along `direction` the
    sky is c0, against it c1, and in between the blend weight of c0 is
    cos(theta / 2)^(2 power) for the angle theta to `direction`."""
    c0, c1, p = np.float32([1.0, 0.5, 0.0]), np.float32([0.0, 0.2, 1.0]), 2.5
    intensity = np.float32([0.5, 1.0, 2.0])
    spec = "gradbg 0 1 0 1 0.5 0 0 0.2 1 2.5"
    np.testing.assert_allclose(oracle.background(spec, [0, 1, 0], intensity), c0 * intensity, rtol=1e-6)
    np.testing.assert_allclose(oracle.background(spec, [0, -1, 0], intensity), c1 * intensity, rtol=1e-6)
    for theta in np.linspace(0.1, 3.0, 12):
        d = np.array([np.sin(theta), np.cos(theta), 0.0])
        w = np.cos(theta / 2) ** (2 * p)
        want = (w * c0 + (1 - w) * c1) * intensity
        np.testing.assert_allclose(oracle.background(spec, d, intensity), want, rtol=2e-6, atol=1e-7)


def test_constant_sky():
    np.testing.assert_allclose(oracle.background("constbg 0.25 0.5 2", [0.3, -0.4, 0.866], [2, 1, 0.5]),
                               [0.5, 0.5, 1.0], rtol=1e-7)


# ------------------------------------------------------------ tiled plane: the {5,4} tilings
def _c22(m):
    return np.array([m[0, 0].real, m[0, 0].imag, m[0, 1].real, m[0, 1].imag, m[1, 0].real, m[1, 0].imag,
                     m[1, 1].real, m[1, 1].imag])


def _xshift(l):
    c, s = np.cosh(l / 2), np.sinh(l / 2)
    return np.array([[c, s], [s, c]], complex)


def _zrot(phi):
    return np.array([[np.exp(0.5j * phi), 0], [0, np.exp(-0.5j * phi)]], complex)


def _apply(m, p):
    return oracle.vec_call("oracle_mo_apply", _c22(m), p)


QJ = np.array([0.0, 0.0, 1.0, 0.0])
# a vertex of the {5,4} tiling sits at QJ with edges along the x and y axes;
# a tile centre lies on the diagonal at the circumradius R of the regular
# right-angled pentagon, cosh R = cot(pi/5) cot(pi/4); its inradius r has
# cosh r = cos(pi/4) / sin(pi/5)
R_CIRC = np.arccosh(1 / np.tan(np.pi / 5))
R_IN = np.arccosh(np.cos(np.pi / 4) / np.sin(np.pi / 5))
TO_CENTRE = _zrot(np.pi / 4) @ _xshift(R_CIRC)


def _plane_points(n, rng):
    """Points of the unit hemisphere (the plane's surface) with Klein
    coordinates uniform in the disk of radius 0.95."""
    r = 0.95 * np.sqrt(rng.uniform(0, 1, n))
    t = rng.uniform(0, 2 * np.pi, n)
    x, y = r * np.cos(t), r * np.sin(t)
    return np.stack([x, y, np.sqrt(1 - x * x - y * y), np.zeros(n)], 1)


@pytest.mark.parametrize("tiling", [1, 2])
def test_tiled_plane_symmetries(tiling):
    """A 2-colouring of the {5,4} tiling: reflection in an edge line (x = 0,
    y = 0) and the quarter turn about the vertex swap the colours; the 72
    degree turn about a tile centre keeps them; all three keep the border
    (its width is measured in hyperbolic distance)."""
    rng = np.random.default_rng(tiling)
    pts = _plane_points(1500, rng)

    def sel(p):
        return oracle.tiled_select(0, tiling, 1.0, 0.05, 2, p)

    m = np.array([sel(p) for p in pts])
    assert 0.05 < np.mean(m == -1) < 0.7 and np.mean(m == 0) > 0.1 and np.mean(m == 1) > 0.1
    turn = TO_CENTRE @ _zrot(2 * np.pi / 5) @ np.linalg.inv(TO_CENTRE)
    cases = [(lambda p: p * np.array([-1, 1, 1, 1]), True), (lambda p: p * np.array([1, -1, 1, 1]), True),
             (lambda p: _apply(_zrot(np.pi / 2), p), True), (lambda p: _apply(turn, p), False)]
    for f, swaps in cases:
        mi = np.array([sel(f(p)) for p in pts])
        np.testing.assert_array_equal(mi == -1, m == -1)
        inner = m >= 0
        if swaps:
            np.testing.assert_array_equal(mi[inner], 1 - m[inner])
        else:
            np.testing.assert_array_equal(mi[inner], m[inner])
    assert sel(_apply(TO_CENTRE, QJ)) >= 0  # a tile centre is inside its tile


@pytest.mark.parametrize("border", [0.02, 0.05, 0.1])
def test_tiled_plane_tile_size_and_border_width(border):
    """From a tile centre toward an edge midpoint the border begins where the
    hyperbolic distance to the edge (at the inradius) is w, sinh w = border x
    cosh(inradius): a band of constant hyperbolic width."""
    def at(t):
        p = _apply(TO_CENTRE @ _zrot(4 * np.pi / 5) @ _xshift(t), QJ)
        return oracle.tiled_select(0, 1, 1.0, border, 2, p)

    lo, hi = 0.0, R_IN
    assert at(lo) >= 0 and at(hi) == -1
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        lo, hi = (lo, mid) if at(mid) == -1 else (mid, hi)
    assert abs(lo - (R_IN - np.arcsinh(border * np.cosh(R_IN)))) < 1e-9


# ------------------------------------------------------------ tiled horosphere: square and hexagonal lattices
def test_tiled_horosphere_square_cells():
    """Square cells of side `cell` on the horosphere z = 1 (flat there):
    the 2 x 2 block of cells around the origin takes 4 distinct materials
    (cell (0, 0) the first), the pattern repeats every 2 cells in x and y
    across negative coordinates too, neighbours differ with 2 materials,
    and the border is the band within `border` of the grid lines."""
    cell, border = 0.8, 0.06
    sel = lambda x, y, n=4: oracle.tiled_select(1, 1, cell, border, n, [x, y, 1.0, 0.0])
    block = {(i, j): sel((i + 0.5) * cell, (j + 0.5) * cell) for i in (0, 1) for j in (0, 1)}
    assert block[(0, 0)] == 0 and sorted(block.values()) == [0, 1, 2, 3]
    for i in range(-4, 4):
        for j in range(-4, 4):
            x, y = (i + 0.5) * cell, (j + 0.5) * cell
            assert sel(x, y) == block[(i % 2, j % 2)]
            assert sel(x, y, 2) != sel(x + cell, y, 2) and sel(x, y, 2) != sel(x, y + cell, 2)
    rng = np.random.default_rng(3)
    for x, y in rng.uniform(-5, 5, (3000, 2)):
        gx = abs(x - cell * np.round(x / cell))
        gy = abs(y - cell * np.round(y / cell))
        near = min(gx, gy)
        if abs(near - border) < 1e-9:
            continue
        assert (sel(x, y) == -1) == (near < border)


def test_tiled_horosphere_hexagonal_cells():
    """Hexagonal cells: the cell regions' centres form a triangular lattice
    of spacing 1.5 x cell (6 neighbours each), 3 materials colour it
    properly (neighbours differ), and the midpoint of two neighbouring
    centres lies on the border."""
    from scipy import ndimage

    cell, border, n = 1.0, 0.05, 3
    xs = np.linspace(-3 * cell, 3 * cell, 481)
    grid = np.array([[oracle.tiled_select(1, 2, cell, border, n, [x, y, 1.0, 0.0]) for x in xs] for y in xs])
    cents, mats = [], []
    for k in range(n):
        lab, count = ndimage.label(grid == k)
        for j in range(1, count + 1):
            ys, xi = np.nonzero(lab == j)
            if ys.min() == 0 or xi.min() == 0 or ys.max() == len(xs) - 1 or xi.max() == len(xs) - 1:
                continue  # cut by the window
            cents.append((xs[xi].mean(), xs[ys].mean()))
            mats.append(k)
    c = np.array(cents)
    assert len(c) >= 10
    dist = np.linalg.norm(c[:, None] - c[None], axis=-1)
    np.fill_diagonal(dist, np.inf)
    spacing = 1.5 * cell
    assert np.all(np.abs(dist.min(1) - spacing) < 0.01 * cell)
    centre = np.argmin(np.linalg.norm(c, axis=1))
    nb = np.nonzero(np.abs(dist[centre] - spacing) < 0.02 * cell)[0]
    assert len(nb) == 6
    for i in range(len(c)):
        for j in np.nonzero(np.abs(dist[i] - spacing) < 0.02 * cell)[0]:
            assert mats[i] != mats[j]
            mid = 0.5 * (c[i] + c[j])
            assert oracle.tiled_select(1, 2, cell, border, n, [mid[0], mid[1], 1.0, 0.0]) == -1
