// Culled, wavefront-compacted nearest-hit scan over a vector of mapped
// objects, with results identical to the plain scan (object/vector.inl:24-62).
//
// 1. Bounds. At kernel start every block computes, for each element, a
//    Euclidean sphere that contains the world image of its primitive and
//    stages it in LDS (FP32, 32 B per element; with HT_SCALAR_BOUNDS a
//    bounds_fill kernel writes them once per launch as FP64 into a global
//    array read through the scalar cache): a unit sphere / ball / cube mapped
//    by an isometry, or, in the upper half-space, the image of the unit
//    hemisphere (a hemisphere over the image of the unit circle) or of the
//    horosphere z = 1 (a horoball) under the element's Moebius map. Images
//    that pass through infinity (a plane) get no bound.
// 2. Cull. Per ray and element, an FP64 test against the geodesic: the
//    straight ray in Euclidean space; in the half-space the semicircle in the
//    vertical plane of the ray, restricted to the part ahead of the start.
//    The sphere is inflated by a margin far above the rounding error of the
//    element's own exact test, so a culled element is one whose exact test
//    returns a miss; the exact test is skipped only where its result cannot
//    matter.
// 3. Compaction. The surviving (ray, element) pairs of a wavefront are
//    appended to a per-wave LDS ring with a ballot and a lane prefix count,
//    and after each scan step processed in batches of one pair per active
//    lane: each lane runs the exact test of one pair for another lane's ray
//    and folds the distance into that lane's slot with a 64-bit LDS atomic
//    min on (ordered distance, element index),
//    which selects the first element with the smallest distance, as the
//    reference loop does. Each lane then reruns its winner's exact test to
//    get the hit state, and maps it back to the parent frame once.
#pragma once
#include "entities.hh"
#include "sdf.hh"
#include "tiled.hh"

#ifndef HT_BOUNDS_CAP
#define HT_BOUNDS_CAP 256
#endif

namespace ht {

static constexpr int BOUNDS_CAP = HT_BOUNDS_CAP;
#ifndef HT_CULL_UNROLL
#define HT_CULL_UNROLL 4
#endif
static constexpr int CULL_UNROLL = HT_CULL_UNROLL;
// The FP64 cull tests may contract a * b + c into FMAs (the programs build
// with -ffp-contract=off for the exact FP32 tests): a cull test only has to
// be conservative, its bounds carry a relative margin of 1e-3 against
// roundings of 1e-16, and which elements it passes does not change any
// result (the exact test decides). Measured on MI355X: 4 % faster on the
// cell frame, 22 % on the 1024-shape SDF scene; images bit-identical.
#define HT_CULL_FP _Pragma("clang fp contract(fast)")
static constexpr int WAVE = 64;
// Pair-queue ring of the culled scan: holds the pairs left after a batch
// (< 64) plus the appends between two drains (<= 64 per element: the
// CULL_UNROLL elements of a flat step, or the leaves of a group node with
// HT_CULL_GROUPS). Every index formed is below twice its size, so one
// conditional subtraction wraps it.
#if defined(HT_CULL_GROUPS) && HT_CULL_GROUPS > HT_CULL_UNROLL
#define HT_RING_STEP HT_CULL_GROUPS
#else
#define HT_RING_STEP HT_CULL_UNROLL
#endif
static constexpr uint RING_STEP = HT_RING_STEP;
#ifdef HT_HIT_ORDER
// With HT_HIT_ORDER the exact tests wait in the ring until it may not take
// another level-2 batch (one leaf per group leaf and lane), so that most
// scans filter all their candidates against the ray's final certain hit.
static constexpr uint QUEUE_RING = 512;
static_assert(RING_STEP <= 4, "the hit-order ring holds one level-2 batch of groups of up to 4 leaves");
static constexpr uint HIT_ORDER_DRAIN = QUEUE_RING - WAVE * RING_STEP;
#else
static constexpr uint QUEUE_RING = WAVE * (RING_STEP + 1);
#endif
HD uint ring_wrap(uint i) { return i >= QUEUE_RING ? i - QUEUE_RING : i; }
// Program option HT_GROUP_QUEUE (with HT_SCALAR_BOUNDS and HT_CULL_GROUPS):
// the two-level scan compacts at both levels. The group spheres are tested
// by every lane (wave-uniform scalar loads); the passing (group, ray) pairs
// go to a second per-wave ring, drained in batches of one pair per active
// lane: each lane tests the pair's group leaves (staged in LDS per block)
// for the pair's ray and pushes the passing (element, ray) pairs to the
// exact tests' ring. A ray then pays for the groups it passes, not for the
// groups any ray of its wave passes. The group ring holds the pairs left
// after a batch (< 64) plus one level-1 step's appends (<= 64 per group).
static constexpr uint GROUP_RING = WAVE * (HT_CULL_UNROLL + 1);
HD uint gring_wrap(uint i) { return i >= GROUP_RING ? i - GROUP_RING : i; }
#if defined(HT_GROUP_QUEUE) && !(defined(HT_SCALAR_BOUNDS) && defined(HT_CULL_GROUPS))
#error "HT_GROUP_QUEUE needs HT_SCALAR_BOUNDS and HT_CULL_GROUPS"
#endif
#if defined(HT_SPARSE_L1) && !defined(HT_GROUP_QUEUE)
#error "HT_SPARSE_L1 is an option of the group queue (HT_GROUP_QUEUE)"
#endif
// Program option HT_HIT_ORDER (with HT_GROUP_QUEUE): before the exact tests,
// each passing (element, ray) pair gets the interval of the ray's
// hyperbolic distance at which its geodesic crosses the element's outer
// sphere, and a "certain hit" bound where the geodesic crosses the inner
// sphere (the element's surface lies between the two spheres, so the
// exact test must hit it no later). A pair whose geodesic never crosses
// the outer sphere ahead of the start is dropped; a pair whose entry lies
// at least HIT_ORDER_MARGIN beyond the ray's nearest certain hit is skipped.
// Neither can be the reference loop's nearest hit: the exact tests' FP32
// distances of these scenes stay within 1e-3 of the geometric ones, far
// inside the margin. The winner, its distance and its hit state are those
// of the plain scan (the exact test decides every hit that is kept).
#if defined(HT_HIT_ORDER) && !defined(HT_GROUP_QUEUE)
#error "HT_HIT_ORDER needs HT_GROUP_QUEUE"
#endif
#ifdef HT_HIT_ORDER
static_assert(HT_BOUNDS_CAP <= 1024, "hit-order ring entries hold a 10-bit element index");
// exp(0.02): a skipped pair's entry is at least 0.02 farther (hyperbolic
// distance) than the nearest certain hit; rounded up for the float product.
#define HIT_ORDER_MARGIN 1.0203f
#endif
// Pair-ring entry: (element << 6) | ray lane; with HT_HIT_ORDER the upper
// 16 bits hold the pair's lb (the upper half of its float bits).
#ifdef HT_HIT_ORDER
#define HIT_LB_ONE 0x3F80u     // 1.0f: no interval
#define HIT_LB_PENDING 0u      // the interval is not computed yet
#define HIT_LB_FIELD(lb16) ((uint)(lb16) << 16)
#define HIT_ELEM(e) (((e) >> 6) & 0x3FFu)
#define HIT_ORDER_LEAF(slot, el) (slot)
#else
#define HIT_LB_ONE 0u
#define HIT_LB_PENDING 0u
#define HIT_LB_FIELD(lb16) 0u
#define HIT_ELEM(e) ((e) >> 6)
#define HIT_ORDER_LEAF(slot, el) (el)
#endif
#ifndef HT_MAX_WAVES
#define HT_MAX_WAVES 4  // blocks of up to 256 threads
#endif
static constexpr int MAX_WAVES = HT_MAX_WAVES;

// World-space sphere of an element: r >= the outer radius that contains the
// primitive's image with margin (r < 0: no bound, always a candidate);
// ri <= an inner radius that the image contains with margin (used only for
// "the geodesic stays inside the sphere", where shrinking is the
// conservative direction); horo = 1: a horoball, tangent to the boundary
// below its centre (2: a tree node whose subtree holds one). FP32, 32 B:
// measured on MI355X, FP64 storage with the squares precomputed (64 B, four
// LDS reads per element) is slower (39.4 vs 38.0 ms per cell frame).
struct Bound {
    float cx, cy, cz, r;
    float ri, horo, pad1, pad2;
};
HD double b_cx(const Bound &b) { return (double)b.cx; }
HD double b_cy(const Bound &b) { return (double)b.cy; }
HD double b_cz(const Bound &b) { return (double)b.cz; }
HD double b_r(const Bound &b) { return (double)b.r; }
HD double b_r2(const Bound &b) { return (double)b.r * (double)b.r; }
HD double b_ri2(const Bound &b) { return (double)b.ri * (double)b.ri; }
HD double b_cz2(const Bound &b) { return (double)b.cz * (double)b.cz; }
HD float b_horo(const Bound &b) { return b.horo; }

// The same bound widened to FP64 with the squares precomputed (the values
// the accessors above compute), 64 B: the layout of the scalar-load scan
// (program option HT_SCALAR_BOUNDS), where a kernel fills one array of them
// in global memory per launch and the scan reads element e's bound at a
// wave-uniform index, so the loads go through the scalar cache into SGPRs
// once per wave instead of as LDS reads and conversions in every lane.
struct BoundS {
    double cx, cy, r2, ri2, cz2, cz, r;
    float horo, pad;
};
static_assert(sizeof(BoundS) == 64, "BoundS layout");
HD double b_cx(const BoundS &b) { return b.cx; }
HD double b_cy(const BoundS &b) { return b.cy; }
HD double b_cz(const BoundS &b) { return b.cz; }
HD double b_r(const BoundS &b) { return b.r; }
HD double b_r2(const BoundS &b) { return b.r2; }
HD double b_ri2(const BoundS &b) { return b.ri2; }
HD double b_cz2(const BoundS &b) { return b.cz2; }
HD float b_horo(const BoundS &b) { return b.horo; }
HD BoundS widen(const Bound &b) {
    BoundS w;
    w.cx = b_cx(b);
    w.cy = b_cy(b);
    w.r2 = b_r2(b);
    w.ri2 = b_ri2(b);
    w.cz2 = b_cz2(b);
    w.cz = b_cz(b);
    w.r = b_r(b);
    w.horo = b.horo;
    w.pad = 0.0f;
    return w;
}
#ifdef HT_SCALAR_BOUNDS
// The scan reads the bounds through a constant-address-space pointer (the
// same memory; no kernel that reads them writes them), which lets
// wave-uniform loads go through the scalar cache into SGPRs.
typedef __attribute__((address_space(4))) const BoundS ConstBoundS;
#ifdef HT_BATCH_BLOCK
// Batch programs: one array per job (filled by batch_bounds_fill before the
// frame); each wave publishes the address of its current job's array in its
// own LDS slot before the tile's paths run, and the scan reads it as a
// wave-uniform (SGPR) base.
HD unsigned long long &batch_bounds_slot() {
    __shared__ unsigned long long slot[HT_MAX_WAVES];
    return slot[threadIdx.x / 64];
}
HD const ConstBoundS *scalar_bounds_base() {
    unsigned long long v = batch_bounds_slot();
    uint lo = __builtin_amdgcn_readfirstlane((uint)v);
    uint hi = __builtin_amdgcn_readfirstlane((uint)(v >> 32));
    return (const ConstBoundS *)(((unsigned long long)hi << 32) | lo);
}
// The same array for loads at per-lane indices (vector loads).
HD const BoundS *vector_bounds_base() {
    unsigned long long v = batch_bounds_slot();
    uint lo = __builtin_amdgcn_readfirstlane((uint)v);
    uint hi = __builtin_amdgcn_readfirstlane((uint)(v >> 32));
    return (const BoundS *)(((unsigned long long)hi << 32) | lo);
}
#else
// Scene programs: the top-level vector's bounds, written by the program's
// bounds_fill kernel before each render launch (launches of one program on
// one stream are ordered, so one array per loaded program).
// With HT_CULL_GROUPS the array holds the tree's leaves in Morton order
// followed by the group nodes (fill_scalar_block).
#ifdef HT_CULL_GROUPS
__device__ BoundS ht_scalar_bounds[HT_BOUNDS_CAP + HT_BOUNDS_CAP / HT_CULL_GROUPS];
#else
__device__ BoundS ht_scalar_bounds[HT_BOUNDS_CAP];
#endif
HD const ConstBoundS *scalar_bounds_base() { return (const ConstBoundS *)ht_scalar_bounds; }
HD const BoundS *vector_bounds_base() { return ht_scalar_bounds; }
#endif
// One 64-byte scalar load per bound (s_load_dwordx16), whatever fields the
// test uses.
typedef unsigned int u32x16 __attribute__((ext_vector_type(16)));
HD BoundS load_bound(const ConstBoundS *p) {
    u32x16 v = *reinterpret_cast<__attribute__((address_space(4))) const u32x16 *>(p);
    BoundS b;
    __builtin_memcpy(&b, &v, sizeof(b));
    return b;
}
#endif

// Relative cull margin for an element whose exact test runs through a map
// with squared Frobenius norm n2: generous against the FP32 rounding of the
// mapped ray (relative error grows with the map's conditioning).
HD double cull_margin(double n2) { return 1e-3 + 1e-4 * n2 * n2; }

// ------------------------------------------------------------ local kinds
// Primitive kind inside a Covered object: 0 = none (no bound), 1 = unit
// sphere, 2 = ball (radius in the data), 3 = unit cube, 4 = unit hemisphere
// (hyperbolic plane), 5 = horosphere z = 1.
template <class S> struct PrimKind { static HD int kind(const u8 *) { return 0; } static HD double radius(const u8 *) { return 0.0; } };
template <> struct PrimKind<Sphere> { static HD int kind(const u8 *) { return 1; } static HD double radius(const u8 *) { return 1.0; } };
template <> struct PrimKind<Ball> { static HD int kind(const u8 *) { return 2; } static HD double radius(const u8 *s) { return (double)ld<real>(s); } };
// Sphere-traced shapes: the march's bounding sphere (bound + 2 eps) plus 2 eps.
template <class S> struct SdfKind {
    static HD int kind(const u8 *) { return 2; }
    static HD double radius(const u8 *s) { return (double)S::bound(s) + 4.0 * (double)SDF_EPS; }
};
template <> struct PrimKind<SdfSphere> : SdfKind<SdfSphere> {};
template <> struct PrimKind<SdfBox> : SdfKind<SdfBox> {};
template <> struct PrimKind<SdfTorus> : SdfKind<SdfTorus> {};
template <> struct PrimKind<Cube> { static HD int kind(const u8 *) { return 3; } static HD double radius(const u8 *) { return 1.7320508075688772; } };
template <> struct PrimKind<Plane<Hy>> { static HD int kind(const u8 *) { return 4; } static HD double radius(const u8 *) { return 1.0; } };
template <> struct PrimKind<Horosphere> { static HD int kind(const u8 *) { return 5; } static HD double radius(const u8 *) { return 1.0; } };

template <class G, class... Vs> struct PrimKind<ShapeChoice<G, Vs...>> {
    typedef ChoiceLayout<Vs...> C;
    template <usize I, class V, class... Rest> static HD int kind_at(usize idx, const u8 *d) {
        if (idx == I) return PrimKind<V>::kind(d);
        if constexpr (sizeof...(Rest) > 0) return kind_at<I + 1, Rest...>(idx, d); else return 0;
    }
    template <usize I, class V, class... Rest> static HD double radius_at(usize idx, const u8 *d) {
        if (idx == I) return PrimKind<V>::radius(d);
        if constexpr (sizeof...(Rest) > 0) return radius_at<I + 1, Rest...>(idx, d); else return 0.0;
    }
    static HD int kind(const u8 *s) { return kind_at<0, Vs...>(C::index(s), C::data(s)); }
    static HD double radius(const u8 *s) { return radius_at<0, Vs...>(C::index(s), C::data(s)); }
};

// Primitive of an object entity, read from its bytes.
template <class TI> struct ObjPrim {
    static constexpr bool OK = false;
    static HD int kind(const u8 *) { return 0; }
    static HD double radius(const u8 *) { return 0.0; }
};
template <class G, class S, class Mat> struct ObjPrim<Covered<G, S, Mat>> {
    static constexpr bool OK = true;
    typedef Covered<G, S, Mat> C;
    static HD int kind(const u8 *o) { return PrimKind<S>::kind(o + C::L::offset(1)); }
    static HD double radius(const u8 *o) { return PrimKind<S>::radius(o + C::L::offset(1)); }
};
template <class Mat, int TL, int N> struct ObjPrim<TiledPlane<Mat, TL, N>> {
    static constexpr bool OK = true;
    static HD int kind(const u8 *) { return 4; }
    static HD double radius(const u8 *) { return 1.0; }
};
template <class Mat, int TL, int N> struct ObjPrim<TiledHorosphere<Mat, TL, N>> {
    static constexpr bool OK = true;
    static HD int kind(const u8 *) { return 5; }
    static HD double radius(const u8 *) { return 1.0; }
};

// Elements the culled scan handles: a mapped object over one primitive.
template <class T> struct Cullable { static constexpr bool value = false; };
template <class G, class TI, class M> struct Cullable<Mapped<G, TI, M>> {
    static constexpr bool value = ObjPrim<TI>::OK;
};

// ------------------------------------------------------------ world bounds
template <class B> struct MakeBound;
template <> struct MakeBound<Bound> {
    static HD Bound none() { return {0.f, 0.f, 0.f, -1.f, -1.f, 0.f, 0.f, 0.f}; }
    // Outer radius inflated, inner radius shrunk, in FP64, each rounded
    // outwards (resp. inwards) to FP32 together with the centre's rounding.
    static HD Bound make(double cx, double cy, double cz, double r, double margin) {
        float fx = (float)cx, fy = (float)cy, fz = (float)cz;
        double err = fabs_((double)fx - cx) + fabs_((double)fy - cy) + fabs_((double)fz - cz);
        double ro = r * (1.0 + margin) + err + 1e-30;
        float fo = (float)ro;
        if ((double)fo < ro) fo = __uint_as_float(__float_as_uint(fo) + 1u);  // next float up (fo > 0)
        double rin = r * (1.0 - margin) - err;
        float fi = rin > 0.0 ? (float)rin : 0.0f;
        if ((double)fi > rin && fi > 0.0f) fi = __uint_as_float(__float_as_uint(fi) - 1u);  // next float down
        return {fx, fy, fz, fo, fi, 0.f, 0.f, 0.f};
    }
    static HD Bound node(float4 q, bool horo) { return {q.x, q.y, q.z, q.w, 0.f, horo ? 2.f : 0.f, 0.f, 0.f}; }
};
HD Bound no_bound() { return MakeBound<Bound>::none(); }
HD Bound make_bound(double cx, double cy, double cz, double r, double margin) {
    return MakeBound<Bound>::make(cx, cy, cz, r, margin);
}

// The element sphere as FP32 (x, y, z, r), rounded outwards (tree leaves).
HD float4 leaf_sphere(const Bound &b) {
    float fx = (float)b.cx, fy = (float)b.cy, fz = (float)b.cz;
    double err = fabs_((double)fx - (double)b.cx) + fabs_((double)fy - (double)b.cy) + fabs_((double)fz - (double)b.cz);
    double ro = (double)b.r + err;
    float fr = (float)ro;
    if ((double)fr < ro) fr = __uint_as_float(__float_as_uint(fr) + 1u);
    return make_float4(fx, fy, fz, fr);
}

template <class M> struct WorldBound;

template <> struct WorldBound<Moebius> {
    static HD Bound of(const Moebius &m, int kind, double) {
        double ar = m.v.a.x, ai = m.v.a.y, br = m.v.b.x, bi = m.v.b.y;
        double cr = m.v.c.x, ci = m.v.c.y, dr = m.v.d.x, di = m.v.d.y;
        double n2 = ar * ar + ai * ai + br * br + bi * bi + cr * cr + ci * ci + dr * dr + di * di;
        double margin = cull_margin(n2);
        if (kind == 5) {
            // horosphere z = 1 -> horoball tangent at a/c, diameter 1/|c|^2
            double c2 = cr * cr + ci * ci;
            if (!(c2 > 1e-12 * n2)) return no_bound();
            double px = (ar * cr + ai * ci) / c2, py = (ai * cr - ar * ci) / c2;
            double rad = 0.5 / c2;
            Bound b = make_bound(px, py, rad, rad, margin);
            b.horo = 1.f;
            return b;
        }
        if (kind == 4) {
            // unit hemisphere -> hemisphere over f(unit circle), f(w) = (aw+b)/(cw+d):
            // centre (b conj(d) - a conj(c)) / (|d|^2 - |c|^2), radius 1 / ||d|^2 - |c|^2|
            double den = (dr * dr + di * di) - (cr * cr + ci * ci);
            if (!(fabs_(den) > 1e-9 * n2)) return no_bound();
            double nx = (br * dr + bi * di) - (ar * cr + ai * ci);
            double ny = (bi * dr - br * di) - (ai * cr - ar * ci);
            return make_bound(nx / den, ny / den, 0.0, 1.0 / fabs_(den), margin);
        }
        return no_bound();
    }
};

template <> struct WorldBound<Homogenous3> {
    static HD Bound of(const Homogenous3 &m, int kind, double radius) {
        if (kind < 1 || kind > 3) return no_bound();
        double cx = m.outer.v.x, cy = m.outer.v.y, cz = m.outer.v.z;  // rotation fixes the origin
        double dist = fabs_(cx) + fabs_(cy) + fabs_(cz);
        return make_bound(cx, cy, cz, radius + 1e-4 * dist, 1e-4);
    }
};

template <> struct WorldBound<Shift3> {
    static HD Bound of(const Shift3 &m, int kind, double radius) {
        if (kind < 1 || kind > 3) return no_bound();
        double cx = m.v.x, cy = m.v.y, cz = m.v.z;
        double dist = fabs_(cx) + fabs_(cy) + fabs_(cz);
        return make_bound(cx, cy, cz, radius + 1e-4 * dist, 1e-4);
    }
};

template <class T, class M> HD Bound element_bound(const u8 *elem) {
    typedef typename T::Inner TI;
    const u8 *inner = T::inner(elem);
    return WorldBound<M>::of(T::map(elem), ObjPrim<TI>::kind(inner), ObjPrim<TI>::radius(inner));
}

// ------------------------------------------------------------ ray cull tests
#ifdef HT_HIT_ORDER
// Crossing interval of a (ray, element) pair; see CullRayHy::interval.
struct HitIv {
    float lb, ub;
    bool drop, certain;
};
HD float float_down(double x) {
    float f = (float)x;
    if ((double)f > x) f = __uint_as_float(__float_as_uint(f) - 1u);  // x >= 1 here
    return f;
}
HD float float_up(double x) {
    float f = (float)x;
    if ((double)f < x) f = __uint_as_float(__float_as_uint(f) + 1u);
    return f;
}
#endif

// Euclidean ray p + t d, t >= 0.
struct CullRayEu {
    static constexpr bool HYPERBOLIC = false;
    double px, py, pz, dx, dy, dz;
    bool valid;
    static HD CullRayEu make(const Ray<Eu> &r) {
        return {(double)r.start.x, (double)r.start.y, (double)r.start.z,
                (double)r.direction.x, (double)r.direction.y, (double)r.direction.z, true};
    }
    // Branch-free: every term is evaluated, so several elements' tests can
    // be interleaved; the decision is the same boolean function.
    template <bool NODE = false, class B> HD bool candidate(const B &b) const {
        HT_CULL_FP
        double sx = b_cx(b) - px, sy = b_cy(b) - py, sz = b_cz(b) - pz;
        double r = b_r(b);
        double t = sx * dx + sy * dy + sz * dz;
        double qx = sx - t * dx, qy = sy - t * dy, qz = sz - t * dz;
        bool hit = !(t < -r) && qx * qx + qy * qy + qz * qz <= b_r2(b);  // not wholly behind, and within r
        return r < 0 || !valid || hit;
    }
};

// Hyperbolic geodesic from p with tangent d: the circle centred on the
// boundary plane in the vertical plane through p along u = d.xy / |d.xy|,
// centre at p.xy + gam u (gam = p.z d.z / |d.xy|), radius
// rg = p.z |d| / |d.xy|; the part ahead of p is the part with horizontal
// coordinate >= 0 along u. A sphere (centre c, radius r) is culled when the
// plane misses it, its cut (centre at sig along u and height c.z, radius
// rp) lies wholly behind p, the two circles are disjoint, or one lies
// strictly inside the other. The tests are written multiplied through by
// m = |d.xy|^2 (and m^2), with S = sig |d.xy| = (c - p).xy . d.xy,
// D = delta |d.xy| = (c - p).xy x d.xy, RP = rp^2 m = r^2 m - D^2 and
// O = (d^2 - rg^2 - rp^2) m = S^2 - 2 p.z d.z S + (c.z^2 - p.z^2) m - RP
// (d: the distance between the circle centres), so the cull state needs no
// square root and no division: disjoint <=> O > 0 and O^2 > 4 p.z^2 |d|^2 RP;
// the cut inside the geodesic's circle <=> p.z^2 |d|^2 > RP, O < 0 and the
// same square test; the geodesic's circle inside the inner sphere's cut
// <=> RPI > p.z^2 |d|^2, I = (ri^2 - r^2) m - O > 0 and
// I^2 > 4 p.z^2 |d|^2 RPI (RPI = ri^2 m - D^2).
struct CullRayHy {
    static constexpr bool HYPERBOLIC = true;
    double px, py, pz, dx, dy, dz, dn2, m, pz2, pz2n, four_pz2n, two_pzdz;
    bool valid;
    static HD CullRayHy make(const Ray<Hy> &r) {
        HT_CULL_FP
        CullRayHy c;
        double dx = r.direction.x, dy = r.direction.y, dz = r.direction.z;
        double pz = r.start.z;
        double m = dx * dx + dy * dy;
        double dn2 = m + dz * dz;
        c.valid = m > 1e-8 * dn2 && pz > 0.0;  // |d.xy| > 1e-4 |d|
        c.px = r.start.x;
        c.py = r.start.y;
        c.pz = pz;
        c.dx = dx;
        c.dy = dy;
        c.dz = dz;
        c.dn2 = dn2;
        c.m = m;
        c.pz2 = pz * pz;
        c.pz2n = c.pz2 * dn2;
        c.four_pz2n = 4.0 * c.pz2n;
        c.two_pzdz = 2.0 * pz * dz;
        return c;
    }
    HD bool toward_tangent_point(double tx, double ty) const;
    // Branch-free (the same boolean function as the sequence of tests it
    // describes), so several elements' tests can be interleaved. NODE: a
    // tree node's sphere (inner radius 0, so the last test never culls; a
    // node over a horoball keeps the "behind" case).
    template <bool NODE = false, class B> HD bool candidate(const B &b) const {
        HT_CULL_FP
        double rx = b_cx(b) - px, ry = b_cy(b) - py;
        double S = rx * dx + ry * dy;   // along the plane, from p (x |d.xy|)
        double D = -rx * dy + ry * dx;  // distance from the vertical plane (x |d.xy|)
        double d2 = D * D;
        double rp = b_r2(b) * m - d2;   // squared radius of the plane's cut of the outer sphere (x m)
        double s2 = S * S;
        // 1. the plane misses the outer sphere
        bool cull = rp < 0.0;
        // 2. the cut circle lies wholly behind p (horizontal coordinate < 0);
        //    not applied to a horoball when the ray is near the geodesic
        //    through its tangent point T: there the element's local ray is
        //    near vertical, and the horosphere test (shape/hy/horosphere.cc:16-24)
        //    applies its EPS window before dividing by |d.xy|^2, so it accepts
        //    crossings behind the start.
        bool behind = S < 0.0 && s2 > rp;
        // horo = 2 marks a tree node whose subtree holds a horoball: no "behind" cull
        float horo = b_horo(b);
        if constexpr (NODE) {
            behind = behind && !(horo > 0);
        } else {
            if (behind && horo > 0) behind = horo < 1.5f && !toward_tangent_point(b_cx(b), b_cy(b));
        }
        cull = cull || behind;
        // 3. disjoint circles; 4. the cut circle strictly inside the geodesic circle
        double o = s2 - two_pzdz * S + (b_cz2(b) - pz2) * m - rp;
        double o2 = o * o, f = four_pz2n * rp;
        cull = cull || (o > 0.0 && o2 > f);
        cull = cull || (pz2n > rp && o < 0.0 && o2 > f);
        // 5. the geodesic circle strictly inside the inner sphere's cut
        if constexpr (!NODE) {
            double ri2 = b_ri2(b);
            double rpi = ri2 * m - d2;
            double in2 = (ri2 - b_r2(b)) * m - o;
            cull = cull || (rpi > pz2n && in2 > 0.0 && in2 * in2 > four_pz2n * rpi);
        }
        return b_r(b) < 0 || !valid || !cull;
    }
};

// True if the ray direction is within ~6 degrees of the tangent (either
// sense) of the geodesic from p to the boundary point T = (tx, ty, 0): the
// angle to the local vertical of a horoball tangent at T (maps are conformal).
HD bool CullRayHy::toward_tangent_point(double tx, double ty) const {
    HT_CULL_FP
    double vx = tx - px, vy = ty - py;
    double v2 = vx * vx + vy * vy, h = pz;
    // tangent at p of the circle through p and T orthogonal to the boundary,
    // scaled by |v|: (h v, (|v|^2 - h^2) / 2); d is not normalised, so both
    // sides carry |d|^2
    double tz = 0.5 * (v2 - h * h);
    double n = h * (dx * vx + dy * vy) + dz * tz;
    double den2 = (h * h * v2 + tz * tz) * dn2;
    return den2 - n * n < 1e-2 * den2;
}

#ifdef HT_HIT_ORDER
// Crossing interval of a geodesic ray with an element's bound spheres.
//
// Parametrise the ray's semicircle by the angle theta seen from its centre,
// measured from the end E the ray moves to; the start is at theta_p and
// along the ray t = tan(theta / 2) falls from tp = tan(theta_p / 2) to 0.
// The hyperbolic distance from the start is ln(tp / t) (arc length along a
// geodesic semicircle), so with w = tp / t = e^D the crossings of a sphere
// with the semicircle are the roots of gam w^2 - 2 bet w + alp = 0, where
// in units of p.z (S, Z: the sphere centre's coordinates along the ray's
// horizontal direction and in height, P: the radius of the sphere's cut by
// the ray's vertical plane):
//   alp = (1 + S tp)^2 + (Z tp)^2 - P^2 tp^2   (the power of the far end W)
//   gam = (tp - S)^2 + Z^2 - P^2               (the power of E)
//   bet = (tp^2 + 1) Z
// and the start (w = 1) is inside the sphere iff S^2 + (1 - Z)^2 < P^2.
// With q = bet + sqrt(bet^2 - alp gam) (bet >= 0) the roots are q / gam and
// alp / q; the first crossing ahead of the start is alp / q.
//   drop:    the start is outside the outer sphere and no crossing of it
//            lies ahead (no root, or both roots below 1);
//   lb:      e^D of the first crossing of the outer sphere ahead (1 when
//            the start is inside it), rounded down;
//   certain: the start is outside the outer sphere and the geodesic
//            crosses the inner sphere ahead, at ub = e^D rounded up: the
//            element's surface lies between the spheres, so its first
//            crossing ahead is no farther than ub.
// Exempt pairs (the element the ray leaves, whose test applies the repeat
// rule, and a horoball whose test may accept crossings behind the start)
// get lb = 1 and no certain bound.
// Reciprocal and square root in FP64 from the hardware approximations and
// one Newton step (relative error far below the interval's 1e-6 slack).
HD double rcp_nr(double x) {
    HT_CULL_FP
    double r = __builtin_amdgcn_rcp(x);
    return r + r * (1.0 - x * r);
}
HD double sqrt_nr(double x) {
    HT_CULL_FP
    if (!(x > 0.0)) return 0.0;
    double y = __builtin_amdgcn_rsq(x);
    y = y * (1.5 - 0.5 * x * y * y);
    return x * y;
}
// Per-ray constants of hit_interval, one float4 per ray: the horizontal unit
// direction (ux, uy), tp and 1 / p.z (tp = 0: no interval for this ray, as
// where the cull state is not valid; straight rays have none).
HD float4 hit_interval_consts(const Ray<Eu> &) { return make_float4(0.f, 0.f, 0.f, 0.f); }
HD float4 hit_interval_consts(const Ray<Hy> &r) {
    HT_CULL_FP
    double dx = r.direction.x, dy = r.direction.y, dz = r.direction.z, pz = r.start.z;
    double m = dx * dx + dy * dy, n = m + dz * dz;
    if (!(m > 1e-8 * n && pz > 0.0)) return make_float4(0.f, 0.f, 0.f, 0.f);
    double sm = sqrt_(m), dn = sqrt_(n), inv = 1.0 / sm;
    // (|d| + d.z) / |d.xy|, written without cancellation for d.z < 0
    double tp = dz >= 0.0 ? (dn + dz) * inv : sm / (dn - dz);
    return make_float4((float)(dx * inv), (float)(dy * inv), (float)tp, (float)(1.0 / pz));
}
// True if the direction d at p is within ~6 degrees of the tangent (either
// sense) of the geodesic from p to the boundary point (tx, ty, 0) (as
// CullRayHy::toward_tangent_point).
HD bool toward_tangent(double px, double py, double pz, double dx, double dy, double dz, double tx, double ty) {
    HT_CULL_FP
    double vx = tx - px, vy = ty - py;
    double v2 = vx * vx + vy * vy, h = pz;
    double tz = 0.5 * (v2 - h * h);
    double n = h * (dx * vx + dy * vy) + dz * tz;
    double den2 = (h * h * v2 + tz * tz) * (dx * dx + dy * dy + dz * dz);
    return den2 - n * n < 1e-2 * den2;
}
// The interval of a pair (see above): ray start (px, py, pz) with the
// constants k of hit_interval_consts; exempt pairs get lb = 1.
template <class B> HD HitIv hit_interval(const B &b, double px, double py, double pz, float4 k, bool exempt) {
    HT_CULL_FP
    HitIv iv = {1.0f, 0.0f, false, false};
    double tp = k.z, ipz = k.w;
    if (!(tp > 0.0) || exempt || b_r(b) < 0) return iv;
    double ux = k.x, uy = k.y;
    double rx = b_cx(b) - px, ry = b_cy(b) - py;
    double sig = rx * ux + ry * uy, delta = -rx * uy + ry * ux;
    double d2p = delta * delta;
    double S = sig * ipz, Z = b_cz(b) * ipz, ipz2 = ipz * ipz;
    double po2 = (b_r2(b) - d2p) * ipz2;
    double zm = 1.0 - Z;
    (void)pz;
    if (S * S + zm * zm < po2) return iv;  // the start is inside the outer sphere
    if (!(po2 > 0.0)) {
        iv.drop = true;  // the ray's vertical plane misses the sphere
        return iv;
    }
    double tp2 = tp * tp;
    double u = 1.0 + S * tp, v = tp - S, zt = Z * tp;
    double a0 = u * u + zt * zt, g0 = v * v + Z * Z, bet = (tp2 + 1.0) * Z;
    // inner sphere first: a certain pair is never dropped
    double pi2 = (b_ri2(b) - d2p) * ipz2;
    if (pi2 > 0.0) {
        double ai = a0 - pi2 * tp2, gi = g0 - pi2;
        double di = bet * bet - ai * gi;
        if (di >= 0.0) {
            double qi = bet + sqrt_nr(di);
            if (qi > 0.0 && (gi <= 0.0 || qi >= gi * (1.0 + 1e-6))) {
                double ub = ai * rcp_nr(qi) * (1.0 + 1e-9);
                if (ub >= 1.0 && ub < 1e30) {
                    iv.certain = true;
                    iv.ub = float_up(ub);
                }
            }
        }
    }
    double ao = a0 - po2 * tp2, go = g0 - po2;
    double disc = bet * bet - ao * go;
    double q = bet + sqrt_nr(disc);
    if (!iv.certain && (disc < 0.0 || (go > 0.0 && q < go * (1.0 - 1e-6)))) {
        iv.drop = true;
        return iv;
    }
    double w = ao * rcp_nr(q);
    iv.lb = w > 1.0 ? float_down(w * (1.0 - 1e-9)) : 1.0f;
    return iv;
}
// The interval of a leaf's pair with a ray: a horoball whose test may accept
// crossings behind the start (the ray near the geodesic through its tangent
// point) is exempt as well; straight rays have no interval.
template <class B> HD HitIv hit_order_interval(const B &b, const Ray<Hy> &r, float4 k, bool exempt) {
    if (!exempt && b_horo(b) > 0.f)
        exempt = toward_tangent(r.start.x, r.start.y, r.start.z, r.direction.x, r.direction.y, r.direction.z, b_cx(b),
                                b_cy(b));
    return hit_interval(b, r.start.x, r.start.y, r.start.z, k, exempt);
}
template <class B> HD HitIv hit_order_interval(const B &, const Ray<Eu> &, float4, bool) {
    return {1.0f, 0.0f, false, false};
}
#endif

template <class G> struct CullRayOf;
template <> struct CullRayOf<Eu> { typedef CullRayEu type; };
template <> struct CullRayOf<Hy> { typedef CullRayHy type; };

// ------------------------------------------------------------ LDS staging
template <class G> struct WaveScratch {
    Ray<G> ray[WAVE];
    uint orig_hasher[WAVE];
    uint prev_hash[WAVE];
    unsigned long long best[WAVE];
    uint queue[QUEUE_RING];
#ifdef HT_GROUP_QUEUE
    uint gqueue[GROUP_RING];  // (group << 6) | ray lane
#endif
#ifdef HT_HIT_ORDER
    uint tw[WAVE];       // float bits of the ray's nearest certain hit e^D (+inf: none yet)
    float4 ivc[WAVE];     // the ray's interval constants (hit_interval_consts)
#endif
#ifdef HT_SPARSE_L1
    unsigned char rlane[WAVE];  // lane of the wave's j-th ray (sparse level 1)
#endif
};

template <class Tag> HD Bound *bounds_lds() {
    __shared__ Bound b[BOUNDS_CAP];
    return b;
}
#ifdef HT_BATCH_BLOCK
// Batch programs inline both geometries' scans, but a wave scans one job at
// a time and every scan rewrites the scratch it reads: the two geometries
// share one LDS buffer.
static constexpr size_t WAVE_SCRATCH_BYTES =
    ((sizeof(WaveScratch<Eu>) > sizeof(WaveScratch<Hy>) ? sizeof(WaveScratch<Eu>) : sizeof(WaveScratch<Hy>)) + 15) /
    16 * 16;
HD unsigned char *wave_scratch_raw() {
    __shared__ __attribute__((aligned(16))) unsigned char raw[MAX_WAVES * WAVE_SCRATCH_BYTES];
    return raw;
}
// Wave w's slot, WAVE_SCRATCH_BYTES apart whatever the geometry.
template <class G> HD WaveScratch<G> *wave_scratch_lds(uint w) {
    static_assert(alignof(WaveScratch<G>) <= 16 && WAVE_SCRATCH_BYTES % 16 == 0, "scratch alignment");
    return reinterpret_cast<WaveScratch<G> *>(wave_scratch_raw() + w * WAVE_SCRATCH_BYTES);
}
#else
template <class G> HD WaveScratch<G> *wave_scratch_lds(uint w) {
    __shared__ WaveScratch<G> s[MAX_WAVES];
    return s + w;
}
#endif

// ------------------------------------------------------------ bounding-sphere tree
// Implicit complete binary tree over the elements in Morton order of their
// bound centres: P = BOUNDS_CAP leaves (leaf j = node P + j -> element
// order[j]), internal node k (1 <= k < P) has children 2k and 2k + 1, root 1.
// Internal nodes hold a sphere containing both children's spheres, inflated
// by a relative 1e-6 and rounded outwards, so a node test is weaker than the
// tests of all its leaves and culling a node never drops a candidate. Node
// radius -1: the subtree holds an element without a bound (always descend);
// -2: no elements. Built per block in LDS by TopScan::prepare.
// Program option HT_CULL_TREE: measured on MI355X the tree walk is slower
// than the flat four-wide scan for these scenes (cell scene 640x480: 47.1 ms
// vs 38.0 ms per 64-spp frame; 1024 SDF shapes: equal), because each step
// is a dependent chain (node load -> FP64 test -> next node) at low
// occupancy, and nodes holding a horoball cannot use the "behind" cull. The
// flat scan is the default.
#if defined(HT_CULL_TREE) && defined(HT_SCALAR_BOUNDS)
#error "HT_CULL_TREE walks the LDS bounds; HT_SCALAR_BOUNDS has none"
#endif
// Program option HT_CULL_GROUPS = g (2..16, a power of two): the same tree,
// scanned in two levels with every lane of a wave in step: the nodes that
// cover g consecutive leaves are tested first, and a node's leaves only when
// some lane's ray passes it (a wave-uniform skip). With HT_SCALAR_BOUNDS the
// bounds_fill kernel builds the tree in its own LDS once per launch and
// writes the leaves (in Morton order, each with its element index) and the
// group nodes to the global FP64 array, which the scan reads through the
// scalar cache: no LDS in the render kernels (batch programs: one such
// array per job, filled by batch_bounds_fill, one block per job).
#ifdef HT_CULL_GROUPS
static constexpr uint CULL_GROUP = HT_CULL_GROUPS;
static_assert(CULL_GROUP >= 2 && CULL_GROUP <= 16 && (CULL_GROUP & (CULL_GROUP - 1)) == 0, "group size");
static_assert(CULL_GROUP <= RING_STEP, "the pair ring holds one group's leaves");
#endif
static_assert((uint)CULL_UNROLL <= RING_STEP, "the pair ring holds one flat step");
#if defined(HT_CULL_TREE) || defined(HT_CULL_GROUPS)
static constexpr bool TREE_OK = BOUNDS_CAP >= 64 && BOUNDS_CAP <= 1024;
#else
static constexpr bool TREE_OK = false;
#endif
static constexpr uint TREE_MIN_ELEMENTS = 16;  // below this the flat scan is cheaper
static constexpr unsigned short NO_ELEMENT = 0xFFFF;

struct TreeLds {
    float4 node[BOUNDS_CAP];            // also the sort's key array (8 B x P) before the nodes are built
    unsigned char horo[BOUNDS_CAP];     // 1: the subtree holds a horoball
    unsigned short order[BOUNDS_CAP];   // leaf j -> element index, NO_ELEMENT for padding
    uint lo[3], hi[3];                  // scene box of the bounded centres (ordered float keys)
};
template <class Tag> HD TreeLds *tree_lds() {
    __shared__ TreeLds t;
    return &t;
}

HD uint expand10(uint v) {
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
HD uint float_key(float f);
HD float key_float(uint k);

// Sphere containing spheres a and b (x, y, z, r >= 0), in FP64, rounded out.
HD float4 merge_spheres(float4 a, float4 b) {
    double dx = (double)b.x - a.x, dy = (double)b.y - a.y, dz = (double)b.z - a.z;
    double d = sqrt_(dx * dx + dy * dy + dz * dz);
    double cx, cy, cz, r;
    if (d + b.w <= a.w) {
        cx = a.x, cy = a.y, cz = a.z, r = a.w;
    } else if (d + a.w <= b.w) {
        cx = b.x, cy = b.y, cz = b.z, r = b.w;
    } else {
        r = 0.5 * (d + (double)a.w + (double)b.w);
        double t = (r - a.w) / d;
        cx = a.x + dx * t, cy = a.y + dy * t, cz = a.z + dz * t;
    }
    float fx = (float)cx, fy = (float)cy, fz = (float)cz;
    double err = fabs_((double)fx - cx) + fabs_((double)fy - cy) + fabs_((double)fz - cz);
    double ro = r * (1.0 + 1e-6) + err + 1e-30;
    float fr = (float)ro;
    if ((double)fr < ro) fr = __uint_as_float(__float_as_uint(fr) + 1u);
    return make_float4(fx, fy, fz, fr);
}

HD void wave_sync() {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

HD unsigned long long lane_mask_below() {
    uint lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Monotonic unsigned key of a float.
// Key of (distance, element index) whose unsigned order is the reference
// loop's choice: smaller distance first, then the lower index. FP32: the
// ordered distance bits above a 32-bit index, exact. FP64: the ordered
// distance bits with the low 12 mantissa bits replaced by the index (up to
// 4096 elements), so distances within 2^-40 relative tie and go to the lower
// index; FP64 programs serve as the accuracy oracle, not as a bit-exact
// drop-in.
#ifdef HT_DOUBLE
HD unsigned long long hit_key(double d, usize i) {
    unsigned long long u = (unsigned long long)__double_as_longlong(d);
    unsigned long long k = (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
    return (k & ~0xFFFull) | (unsigned long long)(i & 0xFFF);
}
HD usize key_index(unsigned long long k) { return (usize)(k & 0xFFFull); }
#else
HD unsigned long long hit_key(float d, usize i);
HD usize key_index(unsigned long long k) { return (usize)(k & 0xFFFFFFFFull); }
#endif
HD uint order_key(float f) {
    uint u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
HD uint float_key(float f) { return order_key(f); }
HD float key_float(uint k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k); }

#ifndef HT_DOUBLE
HD unsigned long long hit_key(float d, usize i) {
    return ((unsigned long long)order_key(d) << 32) | (unsigned long long)i;
}
#endif

// ------------------------------------------------------------ the scan
// Top-level object of a scene: the culled scan for a vector of cullable
// elements in FP32 programs, the element-by-element scan otherwise.
template <class T> struct TopScan {
    static HD void prepare(const u8 *) {}
    static HD void fill_scalar(const u8 *, BoundS *, uint, uint) {}
    static HD void fill_scalar_block(const u8 *, BoundS *) {}
    // has_ray = false: a lane without a ray of its own (nothing to do here)
    template <class G>
    static HD real detect(const u8 *self, Context *ctx, typename T::Cache *cache, Light<G> *light,
                          bool has_ray = true) {
        if (!has_ray) return (real)-1;
        return T::detect(self, ctx, cache, light);
    }
};


template <class G, class E> struct TopScan<ObjectVec<G, E>> {
    typedef ObjectVec<G, E> OV;
    typedef typename OV::V V;
#if defined(HT_NO_CULL)
    static constexpr bool CULL = false;  // the plain scan on request (for comparisons)
#elif defined(HT_DOUBLE)
    static constexpr bool CULL = Cullable<E>::value && (BOUNDS_CAP <= 4096);
#else
    static constexpr bool CULL = Cullable<E>::value;
#endif

    // All threads of the block, before any of them diverges; the caller
    // synchronises the block afterwards.
    static HD void prepare(const u8 *self) {
#if defined(HT_GROUP_QUEUE) && !defined(HT_BATCH_BLOCK)
        // the tree's leaves (Morton order, element index in pad1) from the
        // global array bounds_fill wrote, narrowed to FP32: their values are
        // FP32 numbers widened (the squares FP64-exact), so the narrowed
        // bound decides exactly as the FP64 one
        if constexpr (CULL && TREE_OK) {
            usize n = V::len(self);
            if (n > (usize)BOUNDS_CAP || n < TREE_MIN_ELEMENTS) return;
            Bound *b = bounds_lds<OV>();
            for (uint j = threadIdx.x; j < (uint)BOUNDS_CAP; j += blockDim.x) {
                BoundS q = ht_scalar_bounds[j];
                Bound x;
                x.cx = (float)q.cx;
                x.cy = (float)q.cy;
                x.cz = (float)q.cz;
                x.r = (float)q.r;
                x.ri = q.r < 0.0 ? -1.f : (float)sqrt_(q.ri2);
                x.horo = q.horo;
                x.pad1 = q.pad;
                x.pad2 = 0.f;
                b[j] = x;
            }
        } else {
            (void)self;
        }
#elif defined(HT_SCALAR_BOUNDS)
        (void)self;  // the bounds are in ht_scalar_bounds (fill_scalar)
#else
        if constexpr (CULL) {
            usize n = V::len(self);
            if (n > (usize)BOUNDS_CAP) return;
            Bound *b = bounds_lds<OV>();
            for (uint i = threadIdx.x; i < (uint)n; i += blockDim.x) {
                b[i] = element_bound<E, typename E::Map>(V::elem(self, i));
            }
            if constexpr (TREE_OK) {
                if (n >= TREE_MIN_ELEMENTS) {
                    __syncthreads();
                    build_tree((uint)n, b);
                }
            }
        }
#endif
    }

    // The bounds of the scalar-load scan into `out`, element i by thread
    // first + k * stride; a scene with more elements than BOUNDS_CAP is
    // scanned unculled.
    static HD void fill_scalar(const u8 *self, BoundS *out, uint first, uint stride) {
        if constexpr (CULL) {
            usize n = V::len(self);
            if (n > (usize)BOUNDS_CAP) return;
            for (uint i = first; i < (uint)n; i += stride)
                out[i] = widen(element_bound<E, typename E::Map>(V::elem(self, i)));
        } else {
            (void)self, (void)out, (void)first, (void)stride;
        }
    }

#if defined(HT_SCALAR_BOUNDS) && defined(HT_CULL_GROUPS)
    // All threads of one block: the bounds of the two-level scalar-load scan.
    // With the tree (TREE_OK and at least TREE_MIN_ELEMENTS elements) `out`
    // gets leaf j of the tree at j (element index in `pad`, NO_ELEMENT for
    // padding) and group node FIRST + k at P + k; otherwise element i's bound
    // at i, as fill_scalar writes them (detect takes the same decision).
    static HD void fill_scalar_block(const u8 *self, BoundS *out) {
        if constexpr (CULL && TREE_OK) {
            usize n = V::len(self);
            if (n > (usize)BOUNDS_CAP) return;
            Bound *b = bounds_lds<OV>();
            for (uint i = threadIdx.x; i < (uint)n; i += blockDim.x) b[i] = element_bound<E, typename E::Map>(V::elem(self, i));
            __syncthreads();
            if (n < TREE_MIN_ELEMENTS) {
                for (uint i = threadIdx.x; i < (uint)n; i += blockDim.x) out[i] = widen(b[i]);
                return;
            }
            build_tree((uint)n, b);
            constexpr uint P = BOUNDS_CAP, FIRST = P / CULL_GROUP;
            const TreeLds *t = tree_lds<OV>();
            for (uint j = threadIdx.x; j < P; j += blockDim.x) {
                unsigned short e = t->order[j];
                BoundS w = widen(e == NO_ELEMENT ? no_bound() : b[e]);
                w.pad = __uint_as_float((uint)e);
                out[j] = w;
            }
            for (uint k = FIRST + threadIdx.x; k < 2 * FIRST; k += blockDim.x)
                out[P + (k - FIRST)] = widen(MakeBound<Bound>::node(t->node[k], t->horo[k] != 0));
        } else {
            fill_scalar(self, out, threadIdx.x, blockDim.x);
        }
    }
#endif

    // All threads of the block (see TreeLds).
    static HD void build_tree(uint n, const Bound *b) {
        constexpr uint P = BOUNDS_CAP;
        TreeLds *t = tree_lds<OV>();
        if (threadIdx.x < 3) {
            t->lo[threadIdx.x] = ~0u;
            t->hi[threadIdx.x] = 0u;
        }
        __syncthreads();
        for (uint i = threadIdx.x; i < n; i += blockDim.x) {
            if (b[i].r < 0) continue;
            float c[3] = {(float)b[i].cx, (float)b[i].cy, (float)b[i].cz};
            for (int a = 0; a < 3; ++a) {
                atomicMin(&t->lo[a], float_key(c[a]));
                atomicMax(&t->hi[a], float_key(c[a]));
            }
        }
        __syncthreads();
        // keys: Morton code of the centre in the box, element index; elements
        // without a bound after all others, padding last
        unsigned long long *key = reinterpret_cast<unsigned long long *>(t->node);
        for (uint i = threadIdx.x; i < P; i += blockDim.x) {
            unsigned long long k = ~0ull;
            if (i < n) {
                uint code = 0xFFFFFFFFu;
                if (b[i].r >= 0) {
                    float c[3] = {(float)b[i].cx, (float)b[i].cy, (float)b[i].cz};
                    code = 0;
                    for (int a = 0; a < 3; ++a) {
                        float lo = key_float(t->lo[a]), hi = key_float(t->hi[a]);
                        float u = hi > lo ? (c[a] - lo) / (hi - lo) : 0.5f;
                        uint q = (uint)fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
                        code |= expand10(q) << (2 - a);
                    }
                }
                k = ((unsigned long long)code << 32) | i;
            }
            key[i] = k;
        }
        __syncthreads();
        for (uint k = 2; k <= P; k <<= 1) {
            for (uint j = k >> 1; j > 0; j >>= 1) {
                for (uint i = threadIdx.x; i < P; i += blockDim.x) {
                    uint l = i ^ j;
                    if (l > i) {
                        unsigned long long x = key[i], y = key[l];
                        bool up = (i & k) == 0;
                        if ((x > y) == up) {
                            key[i] = y;
                            key[l] = x;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (uint i = threadIdx.x; i < P; i += blockDim.x) {
            uint low = (uint)(key[i] & 0xFFFFFFFFull);
            t->order[i] = low < n ? (unsigned short)low : NO_ELEMENT;
        }
        __syncthreads();
        // internal nodes, level by level from the leaves' parents to the root
        for (uint first = P / 2; first >= 1; first >>= 1) {
            for (uint k = first + threadIdx.x; k < 2 * first; k += blockDim.x) {
                float4 s[2];
                bool h[2];
                for (int c = 0; c < 2; ++c) {
                    uint ch = 2 * k + c;
                    if (ch >= P) {
                        unsigned short e = t->order[ch - P];
                        if (e == NO_ELEMENT) {
                            s[c] = make_float4(0.f, 0.f, 0.f, -2.f);
                            h[c] = false;
                        } else {
                            const Bound &bb = b[e];
                            s[c] = bb.r < 0 ? make_float4(0.f, 0.f, 0.f, -1.f) : leaf_sphere(bb);
                            h[c] = bb.horo > 0;
                        }
                    } else {
                        s[c] = t->node[ch];
                        h[c] = t->horo[ch] != 0;
                    }
                }
                float4 m;
                if (s[0].w == -2.f) {
                    m = s[1];
                } else if (s[1].w == -2.f) {
                    m = s[0];
                } else if (s[0].w < 0.f || s[1].w < 0.f) {
                    m = make_float4(0.f, 0.f, 0.f, -1.f);
                } else {
                    m = merge_spheres(s[0], s[1]);
                }
                t->node[k] = m;
                t->horo[k] = (h[0] || h[1]) ? 1 : 0;
            }
            __syncthreads();
        }
    }

    // Exact test of element i for a ray, with the hasher state the reference
    // loop would have at that element.
    static HD real element_test(const u8 *self, usize i, Ray<G> ray, uint orig_hasher, uint prev_hash,
                                typename E::Cache *cc, Ray<G> *local_out) {
        const u8 *elem = V::elem(self, i);
        Context c2;
        c2.rng.state = 0;
        c2.hasher.state = orig_hasher;
        hash_usize(&c2.hasher, i);
        c2.prev_hash = prev_hash;
        Light<G> l;
        l.ray = ray_map<G>(ray, E::map(elem).inverse());
        real dist = E::Inner::detect(E::inner(elem), &c2, cc, &l);
        *local_out = l.ray;
        return dist;
    }

    static HD void process(const u8 *self, WaveScratch<G> *ws, uint arank, uint k, uint head, bool slots = false) {
        if (arank < k) {
            uint e = ws->queue[ring_wrap(head + arank)];
            uint i = HIT_ELEM(e), L = e & 63u;
#ifdef HT_HIT_ORDER
            if (slots) i = leaf_element_at(i);  // the entry names the tree's leaf
#else
            (void)slots;
#endif
            typename E::Cache cc;
            Ray<G> local;
            real dist = element_test(self, i, ws->ray[L], ws->orig_hasher[L], ws->prev_hash[L], &cc, &local);
            if (dist > -EPS) {
                atomicMin(&ws->best[L], hit_key(dist, i));
            }
        }
    }

    // The sphere tree's leaves in walk order, each with its cull decision,
    // into `append` (the flat scan's pair queue). HT_CULL_TREE only.
    template <class CR, class Append> static HD void walk_tree(const CR &cr, Append &append) {
        const Bound *bounds = bounds_lds<OV>();
        // Stackless depth-first walk of the sphere tree, one node per
        // lane per step: descend to 2k when node k passes, else move
        // to the next subtree (drop the trailing 1 bits, then +1);
        // node 0 = done. Lanes walk different nodes on the same code.
        constexpr uint P = BOUNDS_CAP;
        const TreeLds *t = tree_lds<OV>();
        uint node = 1;
        for (;;) {
            bool live = node != 0;
            if (__ballot(live) == 0) break;
            uint elem = 0;
            bool cand = false;
            if (live) {
                Bound bb;
                if (node >= P) {
                    unsigned short e = t->order[node - P];
                    elem = e;
                    if (e != NO_ELEMENT) {
                        bb = bounds[e];
                    } else {
                        bb.r = -2.f;
                    }
                } else {
                    bb = MakeBound<Bound>::node(t->node[node], t->horo[node] != 0);
                }
                bool pass = bb.r != -2 && cr.candidate(bb);
                if (node >= P) {
                    cand = pass;
                }
                if (pass && node < P) {
                    node = 2 * node;
                } else {
                    node >>= __builtin_ctz(~node);
                    if (node) node += 1;
                }
            }
            append(elem, cand);
        }
    }

#if defined(HT_GROUP_QUEUE)
#ifdef HT_BATCH_BLOCK
    // Leaf slot i of the tree: batch programs read it from the job's global
    // array (the waves of a block serve different jobs, so there is no
    // per-block LDS copy) with vector loads; the FP64 values bounds_fill
    // wrote decide as their FP32 narrowing in the scene programs' LDS copy.
    static HD BoundS group_leaf(uint i) { return vector_bounds_base()[i]; }
    static HD uint leaf_element(const BoundS &b) { return __float_as_uint(b.pad); }
    static HD uint leaf_element_at(uint i) { return __float_as_uint(vector_bounds_base()[i].pad); }
#else
    static HD const Bound &group_leaf(uint i) { return bounds_lds<OV>()[i]; }
    static HD uint leaf_element(const Bound &b) { return __float_as_uint(b.pad1); }
    static HD uint leaf_element_at(uint i) { return __float_as_uint(bounds_lds<OV>()[i].pad1); }
#endif

    // The two-level scan compacted at both levels (HT_GROUP_QUEUE).
    // push_owned(element, candidate, ray lane) appends to the exact tests'
    // ring, drain() runs its full batches.
    template <class CR, class PushO, class Drain>
    static HD void walk_groups_queued(usize n, bool has_ray, const ConstBoundS *sb, WaveScratch<G> *ws, uint lane,
                                      unsigned long long below, uint nact, uint arank, PushO &push_owned,
                                      Drain &drain) {
        constexpr uint P = BOUNDS_CAP, GS = CULL_GROUP, FIRST = P / GS;
        static_assert(FIRST % CULL_UNROLL == 0, "level-1 steps cover the group nodes");
        uint gcount = 0, ghead = 0;
        auto gpush = [&](uint k, bool pass, uint owner) {
            unsigned long long m = __ballot(pass);
            if (m == 0) return;
            if (pass) ws->gqueue[gring_wrap(ghead + gcount + (uint)__popcll(m & below))] = (k << 6) | owner;
            gcount += (uint)__popcll(m);
        };
        // one batch of k <= nact (group, ray) pairs from the head of the ring
        auto gbatch = [&](uint k) {
            wave_sync();
            bool have = arank < k;
            uint gk = 0, L = 0;
            if (have) {
                uint e = ws->gqueue[gring_wrap(ghead + arank)];
                gk = e >> 6;
                L = e & 63u;
            }
            ghead = gring_wrap(ghead + k);
            gcount -= k;
            const CR c2 = CR::make(ws->ray[have ? L : lane]);
#pragma unroll
            for (uint j = 0; j < GS; ++j) {
                bool cand = false;
                uint el = 0;
                if (have) {
                    const auto &lb = group_leaf(gk * GS + j);
                    el = leaf_element(lb);
                    cand = el != NO_ELEMENT && c2.candidate(lb);
                }
                // with HT_HIT_ORDER the entry names the leaf (its interval is
                // computed when the ring is filtered, one pair per lane)
                push_owned(HIT_ORDER_LEAF(gk * GS + j, el), cand, L, HIT_LB_PENDING);
            }
            drain();
        };
        // the n elements fill the leaves [0, n) (bounded ones in Morton order,
        // then the unbounded): the groups after the last one holding an
        // element are padding only
        const uint groups = ((uint)n + GS - 1) / GS;
#ifdef HT_SPARSE_L1
        // A wave with few rays (program option HT_SPARSE_L1 = the most rays
        // this applies to) runs level 1 over its (group, ray) pairs, one
        // pair per active lane: groups * R pairs in ceil(groups * R / nact)
        // steps instead of every group in every lane. Ray-minor order: the
        // lanes of a step mostly read the same group's bound (vector loads
        // of the array the other form reads through the scalar cache).
        {
            const unsigned long long live = __ballot(has_ray);
            const uint R = (uint)__popcll(live);
            if (R <= (uint)HT_SPARSE_L1) {
                if (R == 0) return;
                if (has_ray) ws->rlane[__popcll(live & below)] = (unsigned char)lane;  // the wave's j-th ray
                wave_sync();
                const uint qs = nact / R, rs = nact % R;
                uint r = arank % R, g = arank / R;  // pair base + arank = g * R + r
                const BoundS *nodes = vector_bounds_base() + P;
                for (uint base = 0; base < groups * R; base += nact) {
                    const uint L = ws->rlane[r];
                    bool pass = false;
                    if (g < groups) {
                        const CR c1 = CR::make(ws->ray[L]);
                        const BoundS q = nodes[g];
                        pass = q.r != -2.0 && c1.template candidate<true>(q);  // -2: a node over padding only
                    }
                    gpush(g, pass, L);
                    while (gcount >= nact) gbatch(nact);
                    r += rs;
                    g += qs;
                    if (r >= R) {
                        r -= R;
                        g += 1;
                    }
                }
                if (gcount > 0) gbatch(gcount);
                return;
            }
        }
#endif
        for (uint k0 = 0; k0 < groups; k0 += CULL_UNROLL) {
            // the lane's own cull state, rebuilt from its ray in LDS at each
            // step (a few multiplications): it is then not live in registers
            // across the level-2 batches and the exact tests they drain
            const CR c1 = CR::make(ws->ray[lane]);
            bool pass[CULL_UNROLL];
#pragma unroll
            for (int u = 0; u < CULL_UNROLL; ++u) {
                BoundS q = load_bound(sb + P + k0 + u);
                pass[u] = has_ray && q.r != -2.0 && c1.template candidate<true>(q);  // -2: a node over padding only
            }
#pragma unroll
            for (int u = 0; u < CULL_UNROLL; ++u) gpush(k0 + u, pass[u], lane);
            while (gcount >= nact) gbatch(nact);
        }
        if (gcount > 0) gbatch(gcount);
    }
#endif

#if defined(HT_CULL_GROUPS) && defined(HT_SCALAR_BOUNDS)
    // walk_groups over the global array fill_scalar_block writes: every
    // node and leaf bound is one scalar load.
    template <class CR, class Push, class Drain>
    static HD void walk_groups_scalar(usize n, const CR &cr, const ConstBoundS *sb, Push &push, Drain &drain) {
        constexpr uint P = BOUNDS_CAP, GS = CULL_GROUP;
        const uint groups = ((uint)n + GS - 1) / GS;  // the later groups are padding only
        for (uint k = 0; k < groups; ++k) {
            BoundS q = load_bound(sb + P + k);
            if (q.r == -2.0) continue;  // no elements under this node
            bool pass = cr.template candidate<true>(q);
            if (__ballot(pass) == 0) continue;
            bool c[GS];
            uint el[GS];
#pragma unroll
            for (uint j = 0; j < GS; ++j) {
                BoundS lb = load_bound(sb + k * GS + j);
                el[j] = __float_as_uint(lb.pad);
                c[j] = el[j] != NO_ELEMENT && pass && cr.candidate(lb);
            }
#pragma unroll
            for (uint j = 0; j < GS; ++j) push(el[j], c[j]);
            drain();
        }
    }
#elif defined(HT_CULL_GROUPS)
    // The group nodes of the tree in order, each tested by every lane; a
    // node's CULL_GROUP leaves are tested (and their pairs queued) only when
    // some lane's ray passes the node. HT_CULL_GROUPS only.
    template <class CR, class Push, class Drain>
    static HD void walk_groups(const CR &cr, Push &push, Drain &drain) {
        constexpr uint P = BOUNDS_CAP, GS = CULL_GROUP, FIRST = P / GS;
        const Bound *bounds = bounds_lds<OV>();
        const TreeLds *t = tree_lds<OV>();
        for (uint k = FIRST; k < 2 * FIRST; ++k) {
            float4 q = t->node[k];
            if (q.w == -2.f) continue;  // no elements under this node
            bool pass = cr.template candidate<true>(MakeBound<Bound>::node(q, t->horo[k] != 0));
            if (__ballot(pass) == 0) continue;
            bool c[GS];
            unsigned short el[GS];
#pragma unroll
            for (uint j = 0; j < GS; ++j) {
                el[j] = t->order[k * GS - P + j];
                c[j] = el[j] != NO_ELEMENT && cr.candidate(bounds[el[j] != NO_ELEMENT ? el[j] : 0]) && pass;
            }
#pragma unroll
            for (uint j = 0; j < GS; ++j) push(el[j], c[j]);
            drain();
        }
    }
#endif

    // has_ray = false (program option HT_HELPERS): a lane whose pixel is
    // done joins the wave's scan without a ray of its own; it pushes no
    // pairs but takes other rays' pairs in the batches, so the last rays of
    // a wave get their pair tests on full lanes.
    static HD real detect(const u8 *self, Context *ctx, typename OV::Cache *cache, Light<G> *light,
                          bool has_ray = true) {
        if constexpr (!CULL) {
            if (!has_ray) return -R1;
            return OV::detect(self, ctx, cache, light);
        } else {
            usize n = V::len(self);
            if (n > (usize)BOUNDS_CAP) {
                if (!has_ray) return -R1;
                return OV::detect(self, ctx, cache, light);
            }
#ifdef HT_SCALAR_BOUNDS
            const ConstBoundS *sb = scalar_bounds_base();
            auto bound = [&](usize e) { return load_bound(sb + e); };
#else
            const Bound *lb = bounds_lds<OV>();
            auto bound = [&](usize e) -> const Bound & { return lb[e]; };
#endif
            WaveScratch<G> *ws = wave_scratch_lds<G>(threadIdx.x / WAVE);
            uint lane = __lane_id();
            uint orig = ctx->hasher.state;
            ws->ray[lane] = light->ray;
            ws->orig_hasher[lane] = orig;
            ws->prev_hash[lane] = ctx->prev_hash;
            ws->best[lane] = ~0ull;
#ifdef HT_HIT_ORDER
            ws->tw[lane] = 0x7F800000u;  // +inf
            ws->ivc[lane] = hit_interval_consts(light->ray);
#endif
            unsigned long long below = lane_mask_below();
            unsigned long long active = __ballot(1);
            uint nact = (uint)__popcll(active);
            uint arank = (uint)__popcll(active & below);
            typename CullRayOf<G>::type cr = CullRayOf<G>::type::make(light->ray);
            wave_sync();
            bool tree = false;
            if constexpr (TREE_OK) tree = n >= TREE_MIN_ELEMENTS;
            uint qcount = 0, qhead = 0;
            // Pairs in a ring: push adds element i for the lanes where it is
            // a candidate; drain runs the exact tests of full batches (one
            // pair per active lane) from the head, once per scan step, so the
            // exact test's code appears once in the loop.
            // (with HT_HIT_ORDER an entry also holds the upper 16 bits of
            // the pair's lb as a float: a lower bound)
            auto push_owned = [&](usize i, bool cand, uint owner, uint lb16) {
                unsigned long long m = __ballot(cand);
                if (m == 0) return;
                if (cand)
                    ws->queue[ring_wrap(qhead + qcount + (uint)__popcll(m & below))] =
                        HIT_LB_FIELD(lb16) | ((uint)i << 6) | owner;
                qcount += (uint)__popcll(m);
            };
            auto push = [&](usize i, bool cand) { push_owned(i, cand && has_ray, lane, HIT_LB_ONE); };
#ifdef HT_HIT_ORDER
            // Drops the ring's pairs whose entry lies beyond their ray's
            // nearest certain hit by the margin (stable in-place compaction,
            // one pair per active lane per step: a step writes no slot that a
            // later step reads).
            auto filter = [&]() {
                wave_sync();  // the pushes of the last batch
                if (tree) {
                    // the intervals of the pairs pushed since the last filter,
                    // one pair per active lane; a certain pair lowers its
                    // ray's bound, a dropped one gets lb = +inf
                    for (uint base = 0; base < qcount; base += nact) {
                        uint pos = base + arank;
                        if (pos < qcount) {
                            uint at = ring_wrap(qhead + pos);
                            uint e = ws->queue[at];
                            if ((e >> 16) == HIT_LB_PENDING) {
                                uint L = e & 63u;
                                const auto &b = group_leaf(HIT_ELEM(e));
                                Hasher hh = {ws->orig_hasher[L]};
                                hash_usize(&hh, leaf_element(b));
                                bool exempt = hh.state == ws->prev_hash[L];  // the repeat rule decides its test
                                HitIv iv = hit_order_interval(b, ws->ray[L], ws->ivc[L], exempt);
                                if (iv.certain) atomicMin(&ws->tw[L], __float_as_uint(iv.ub));
                                uint lb16 = iv.drop ? 0x7F80u : (__float_as_uint(iv.lb) >> 16);  // truncation rounds down
                                ws->queue[at] = (e & 0xFFFFu) | (lb16 << 16);
                            }
                        }
                    }
                    wave_sync();
                }
                uint out = 0;
                for (uint base = 0; base < qcount; base += nact) {
                    uint pos = base + arank;
                    bool keep = false;
                    uint e = 0;
                    if (pos < qcount) {
                        e = ws->queue[ring_wrap(qhead + pos)];
                        float lbf = __uint_as_float(e & 0xFFFF0000u);
                        float t = __uint_as_float(ws->tw[e & 63u]);
                        keep = !(lbf >= t * HIT_ORDER_MARGIN);
                    }
                    unsigned long long km = __ballot(keep);
                    wave_sync();
                    if (keep) ws->queue[ring_wrap(qhead + out + (uint)__popcll(km & below))] = e;
                    out += (uint)__popcll(km);
                    wave_sync();
                }
                qcount = out;
            };
#endif
            auto drain = [&]() {
#ifdef HT_HIT_ORDER
                // Wait while the ring can take another level-2 batch; then
                // filter it and test only as many pairs as make room for one:
                // the others wait for their ray's final certain bound.
                if (qcount <= HIT_ORDER_DRAIN) return;
                filter();
                while (qcount > HIT_ORDER_DRAIN) {
#else
                while (qcount >= nact) {
#endif
                    wave_sync();
                    process(self, ws, arank, nact, qhead, tree);
                    qhead = ring_wrap(qhead + nact);
                    qcount -= nact;
                    wave_sync();
                }
            };
            auto append = [&](usize i, bool cand) {
                push(i, cand);
                drain();
            };
            if (tree) {
#if defined(HT_GROUP_QUEUE)
                if constexpr (TREE_OK)
                    walk_groups_queued<typename CullRayOf<G>::type>(n, has_ray, sb, ws, lane, below, nact, arank,
                                                                   push_owned, drain);
#elif defined(HT_CULL_GROUPS) && defined(HT_SCALAR_BOUNDS)
                if constexpr (TREE_OK) walk_groups_scalar(n, cr, sb, push, drain);
#elif defined(HT_CULL_GROUPS)
                if constexpr (TREE_OK) walk_groups(cr, push, drain);
#else
                if constexpr (TREE_OK) walk_tree(cr, append);
#endif
            } else {
                // CULL_UNROLL independent cull tests per step (instruction-level
                // parallelism, traded against registers and so occupancy).
                usize e = 0;
                for (; e + CULL_UNROLL <= n; e += CULL_UNROLL) {
                    bool c[CULL_UNROLL];
#pragma unroll
                    for (int u = 0; u < CULL_UNROLL; ++u) c[u] = cr.candidate(bound(e + u));
#pragma unroll
                    for (int u = 0; u < CULL_UNROLL; ++u) push(e + u, c[u]);
                    drain();
                }
                for (; e < n; ++e) append(e, cr.candidate(bound(e)));
            }
#ifdef HT_HIT_ORDER
            if (qcount > 0) {
                filter();
                while (qcount >= nact) {
                    wave_sync();
                    process(self, ws, arank, nact, qhead, tree);
                    qhead = ring_wrap(qhead + nact);
                    qcount -= nact;
                }
            }
#endif
            if (qcount > 0) {
                wave_sync();
                process(self, ws, arank, qcount, qhead, tree);
            }
            wave_sync();
            if (!has_ray) return -R1;
            unsigned long long best = ws->best[lane];
            if (best == ~0ull) {
                ctx->hasher.state = orig;
                return -R1;
            }
            usize i = key_index(best);
            typename E::Cache cc;
            Ray<G> local;
            real dist = element_test(self, i, light->ray, orig, ctx->prev_hash, &cc, &local);
            ctx->hasher.state = orig;
            hash_usize(&ctx->hasher, i);
            cache->inner = cc;
            cache->index = i;
            light->ray = ray_map<G>(local, E::map(V::elem(self, i)));
            return dist;
        }
    }
};

}  // namespace ht
