// CPU oracle of hypertrace's per-pixel Monte Carlo render pass.
//
// TEST INFRASTRUCTURE ONLY. This is the checker: only tests/, the smoke check
// and bench.py's CPU-baseline leg may load it. The product path (the HIP
// library) never calls it.
//
// A restatement of the reference render path as a plain interpreter over a
// token description of the scene (emitted by hypertrace_amd/objects.py
// `spec`). Structure: virtual nodes built by a recursive-descent parser; the
// numerics follow the reference formulas in the reference's operation order,
// with the file:line cited at each function. Templated on the scalar:
//   <float, float>   the reference FP32 build
//   <double, float>  the reference DOUBLE_PRECISION build (geometry FP64,
//                    colours FP32, kernel/ocl/algebra/color.hh:6-7)
//   <Cnt, Cnt>       flop counting: add/sub/mul/div/sqrt = 1, each
//                    transcendental = 1, comparisons/negation/abs/min/max = 0
// Built with -ffp-contract=off so every operation rounds on its own; FP32
// transcendentals are evaluated in FP64 and rounded once (correctly rounded
// except in ~2^-28 of cases), so the oracle does not carry one libm's ulp
// errors. The reference leaves these to its OpenCL implementation's builtins.
//
// Which check pins each part: tests/test_oracle_pins.py (leaf functions
// against the reference's compiled leaf sources, closed forms and brute
// force), tests/test_oracle_materials.py (materials, mixture selection,
// backgrounds and the tiled objects against the reference's own unit-test
// rules and the tilings' geometry, through the oracle_material_interact /
// oracle_background / oracle_tiled_select entry points), tests/test_oracle_sdf.py
// (sphere tracing, roulette), tests/test_oracle_scenes.py (whole paths
// against closed forms). Every function here has such a check.
#include <atomic>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ------------------------------------------------------------ counting scalar
thread_local uint64_t g_flops = 0;

struct Cnt {
    float v;
    Cnt() : v(0) {}
    Cnt(float x) : v(x) {}
    Cnt(double x) : v((float)x) {}
    Cnt(int x) : v((float)x) {}
    Cnt(unsigned x) : v((float)x) {}
    explicit operator float() const { return v; }
};
inline Cnt operator+(Cnt a, Cnt b) { ++g_flops; return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { ++g_flops; return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { ++g_flops; return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { ++g_flops; return Cnt(a.v / b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline Cnt &operator+=(Cnt &a, Cnt b) { return a = a + b; }
inline Cnt &operator-=(Cnt &a, Cnt b) { return a = a - b; }
inline Cnt &operator*=(Cnt &a, Cnt b) { return a = a * b; }
inline Cnt &operator/=(Cnt &a, Cnt b) { return a = a / b; }
inline bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
inline bool operator<=(Cnt a, Cnt b) { return a.v <= b.v; }
inline bool operator>=(Cnt a, Cnt b) { return a.v >= b.v; }
inline bool operator==(Cnt a, Cnt b) { return a.v == b.v; }

inline float fsqrt(float x) { return std::sqrt(x); }
inline double fsqrt(double x) { return std::sqrt(x); }
inline Cnt fsqrt(Cnt x) { ++g_flops; return Cnt(std::sqrt(x.v)); }
inline float flog(float x) { return (float)std::log((double)x); }
inline double flog(double x) { return std::log(x); }
inline Cnt flog(Cnt x) { ++g_flops; return Cnt(std::log(x.v)); }
inline float fcos(float x) { return (float)std::cos((double)x); }
inline double fcos(double x) { return std::cos(x); }
inline Cnt fcos(Cnt x) { ++g_flops; return Cnt(std::cos(x.v)); }
inline float fsin(float x) { return (float)std::sin((double)x); }
inline double fsin(double x) { return std::sin(x); }
inline Cnt fsin(Cnt x) { ++g_flops; return Cnt(std::sin(x.v)); }
inline float fcosh(float x) { return (float)std::cosh((double)x); }
inline double fcosh(double x) { return std::cosh(x); }
inline Cnt fcosh(Cnt x) { ++g_flops; return Cnt((float)std::cosh((double)x.v)); }
inline float fsinh(float x) { return (float)std::sinh((double)x); }
inline double fsinh(double x) { return std::sinh(x); }
inline Cnt fsinh(Cnt x) { ++g_flops; return Cnt((float)std::sinh((double)x.v)); }
inline float ffloor(float x) { return std::floor(x); }
inline double ffloor(double x) { return std::floor(x); }
inline Cnt ffloor(Cnt x) { return Cnt(std::floor(x.v)); }
inline float fpow(float a, float b) { return (float)std::pow((double)a, (double)b); }
inline double fpow(double a, double b) { return std::pow(a, b); }
inline Cnt fpow(Cnt a, Cnt b) { ++g_flops; return Cnt(std::pow(a.v, b.v)); }
inline float fabs_(float x) { return std::fabs(x); }
inline double fabs_(double x) { return std::fabs(x); }
inline Cnt fabs_(Cnt x) { return Cnt(std::fabs(x.v)); }
inline float fmin_(float a, float b) { return std::fmin(a, b); }
inline double fmin_(double a, double b) { return std::fmin(a, b); }
inline Cnt fmin_(Cnt a, Cnt b) { return Cnt(std::fmin(a.v, b.v)); }
inline float fmax_(float a, float b) { return std::fmax(a, b); }
inline double fmax_(double a, double b) { return std::fmax(a, b); }
inline Cnt fmax_(Cnt a, Cnt b) { return Cnt(std::fmax(a.v, b.v)); }
template <class R> inline float tofloat(R x) { return (float)x; }

template <class R> struct Consts;
template <> struct Consts<float> {
    static float eps() { return 1e-6f; }
    static float pi() { return 3.14159265358979f; }
    static float fract_max() { return 0x1.fffffep-1f; }
};
template <> struct Consts<double> {
    static double eps() { return 1e-12; }
    static double pi() { return 3.141592653589793; }
    static double fract_max() { return 0x1.fffffffffffffp-1; }
};
template <> struct Consts<Cnt> {
    static Cnt eps() { return Cnt(1e-6f); }
    static Cnt pi() { return Cnt(3.14159265358979f); }
    static Cnt fract_max() { return Cnt(0x1.fffffep-1f); }
};

// ------------------------------------------------------------ vectors
template <class R> struct V2 { R x, y; };
template <class R> struct V3 { R x, y, z; };
template <class R> struct V4 { R x, y, z, w; };

template <class R> V2<R> operator+(V2<R> a, V2<R> b) { return {a.x + b.x, a.y + b.y}; }
template <class R> V2<R> operator-(V2<R> a, V2<R> b) { return {a.x - b.x, a.y - b.y}; }
template <class R> V2<R> operator*(V2<R> a, V2<R> b) { return {a.x * b.x, a.y * b.y}; }
template <class R> V2<R> operator*(V2<R> a, R s) { return {a.x * s, a.y * s}; }
template <class R> V2<R> operator/(V2<R> a, R s) { return {a.x / s, a.y / s}; }
template <class R> V2<R> operator-(V2<R> a) { return {-a.x, -a.y}; }
template <class R> R dot(V2<R> a, V2<R> b) { return a.x * b.x + a.y * b.y; }

template <class R> V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

template <class R> V4<R> operator+(V4<R> a, V4<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
template <class R> V4<R> operator-(V4<R> a, V4<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
template <class R> V4<R> operator*(V4<R> a, R s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
template <class R> V4<R> operator*(R s, V4<R> a) { return {s * a.x, s * a.y, s * a.z, s * a.w}; }
template <class R> V4<R> operator/(V4<R> a, R s) { return {a.x / s, a.y / s, a.z / s, a.w / s}; }
template <class R> V4<R> operator-(V4<R> a) { return {-a.x, -a.y, -a.z, -a.w}; }
template <class R> R dot(V4<R> a, V4<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

template <class V> auto length2(V a) { return dot(a, a); }
template <class V> auto length(V a) { return fsqrt(dot(a, a)); }
// OpenCL normalize(v) = v / length(v)
template <class V> V normalize(V a) { return a / length(a); }

// ------------------------------------------------------------ algebra
// complex.cc:4-30
template <class R> V2<R> c_conj(V2<R> a) { return {a.x, -a.y}; }
template <class R> V2<R> c_mul(V2<R> a, V2<R> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <class R> V2<R> c_inverse(V2<R> a) { return c_conj(a) / dot(a, a); }
// quaternion.cc:4-56; a quat is (x = real part, y, z, w)
template <class R> V4<R> q_conj(V4<R> a) { return {a.x, -a.y, -a.z, -a.w}; }
template <class R> V4<R> q_mul(V4<R> a, V4<R> b) {
    return {a.x * b.x - a.y * b.y - a.z * b.z - a.w * b.w,
            a.x * b.y + a.y * b.x + a.z * b.w - a.w * b.z,
            a.x * b.z + a.z * b.x - a.y * b.w + a.w * b.y,
            a.x * b.w + a.w * b.x + a.y * b.z - a.z * b.y};
}
template <class R> V4<R> cq_mul(V2<R> a, V4<R> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x, a.x * b.z - a.y * b.w, a.x * b.w + a.y * b.z};
}
template <class R> V4<R> q_inverse(V4<R> a) { return q_conj(a) / dot(a, a); }
template <class R> V4<R> q_div(V4<R> a, V4<R> b) { return q_mul(a, q_inverse(b)); }
template <class R> V4<R> q_of(V2<R> a, V2<R> b) { return {a.x, a.y, b.x, b.y}; }

// 2x2 complex matrix, matrix_complex.cc:38-65
template <class R> struct C22 { V2<R> a, b, c, d; };
template <class R> C22<R> c22_dot(C22<R> x, C22<R> y) {
    return {c_mul(x.a, y.a) + c_mul(x.b, y.c), c_mul(x.a, y.b) + c_mul(x.b, y.d),
            c_mul(x.c, y.a) + c_mul(x.d, y.c), c_mul(x.c, y.b) + c_mul(x.d, y.d)};
}

// ------------------------------------------------------------ RNG, hash
// random.cc:8-14
struct Rng {
    uint32_t state;
    uint32_t next() { return state = 1103515245u * state + 12345u; }
};
template <class R> R uniform(Rng &rng) {
    uint32_t s = rng.next();
    // int -> real rounds to nearest; dividing by 2^32 is exact.
    return R(double(s)) / R(4294967296.0);
}
// random.cc:53-59
template <class R> V3<R> hemisphere_cosine(Rng &rng) {
    R phi = (R(2) * Consts<R>::pi()) * uniform<R>(rng);
    R sqr_cos = uniform<R>(rng);
    R c = fsqrt(sqr_cos);
    R s = fsqrt(R(1) - sqr_cos);
    return {fcos(phi) * s, fsin(phi) * s, c};
}
// hash.cc:11-24
inline uint32_t hash_step(uint32_t h, uint32_t x) { return x + (h << 6) + (h << 16) - h; }

struct Ctx {
    Rng rng;
    uint32_t hasher;
    uint32_t prev_hash;
    bool repeat() const { return hasher == prev_hash; }  // context.cc:11-13
};

// ------------------------------------------------------------ geometry
template <class R> struct GeoEu {
    typedef V3<R> Pos;
    typedef V3<R> Dir;
    static Pos origin() { return {R(0), R(0), R(0)}; }
    static V3<R> to_local(Pos, Dir d) { return d; }
    static Dir from_local(Pos, V3<R> d) { return d; }
};
template <class R> struct GeoHy {
    typedef V4<R> Pos;
    typedef V4<R> Dir;
    static Pos origin() { return {R(0), R(0), R(1), R(0)}; }  // hyperbolic.cc:4-6
    static V3<R> to_local(Pos, Dir d) { return {d.x, d.y, d.z}; }
    static Dir from_local(Pos, V3<R> d) { return {d.x, d.y, d.z, R(0)}; }
};

// hyperbolic.cc:20-23
template <class R> R hy_distance(V4<R> a, V4<R> b) {
    R x = R(1) + length2(a - b) / (R(2) * a.z * b.z);
    return flog(x + fsqrt(x * x - R(1)));
}
// hyperbolic.cc:25-33
template <class R> V4<R> hy_dir_at(V4<R> p, V4<R> d, V4<R> h) {
    V2<R> pxy = {p.x, p.y}, hxy = {h.x, h.y}, dxy = {d.x, d.y};
    return {h.z / p.z * d.x, h.z / p.z * d.y, d.z - length(pxy - hxy) / p.z * length(dxy), R(0)};
}

template <class R, class G> struct Ray {
    typename G::Pos start;
    typename G::Dir dir;
};
template <class F> struct C3 { F x, y, z; };
template <class F> C3<F> operator+(C3<F> a, C3<F> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class F> C3<F> operator*(C3<F> a, C3<F> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class F> C3<F> operator*(F s, C3<F> a) { return {s * a.x, s * a.y, s * a.z}; }

template <class R, class F, class G> struct Light {
    C3<F> intensity;
    Ray<R, G> ray;
};
template <class R, class F> struct LightLocal {
    C3<F> intensity;
    V3<R> dir;
};

// ------------------------------------------------------------ maps
template <class R, class G> struct Map {
    virtual ~Map() {}
    virtual typename G::Pos apply_pos(typename G::Pos p) const = 0;
    virtual typename G::Dir apply_dir(typename G::Pos p, typename G::Dir d) const = 0;
    virtual std::unique_ptr<Map> inverse() const = 0;
};

// rotation.cc:64-66: (q (0,p) q*).yzw
template <class R> V3<R> rot_apply(V4<R> q, V3<R> p) {
    V4<R> r = q_mul(q_mul(q, V4<R>{R(0), p.x, p.y, p.z}), q_conj(q));
    return {r.y, r.z, r.w};
}

// Homogenous3 = outer Shift3 after inner Rotation3 (chain.inl:36-73).
// Shift3 and Rotation3 alone are the special cases with identity parts.
template <class R> struct HomMap : Map<R, GeoEu<R>> {
    V3<R> shift;
    V4<R> rot;
    bool has_shift, has_rot;
    HomMap(V3<R> s, V4<R> q, bool hs, bool hr) : shift(s), rot(q), has_shift(hs), has_rot(hr) {}
    V3<R> apply_pos(V3<R> p) const override {
        V3<R> r = has_rot ? rot_apply(rot, p) : p;
        return has_shift ? r + shift : r;
    }
    V3<R> apply_dir(V3<R>, V3<R> d) const override { return has_rot ? rot_apply(rot, d) : d; }
    std::unique_ptr<Map<R, GeoEu<R>>> inverse() const override {
        // inverse = {Shift(R^-1 (-v)), R^-1} (chain.inl:65-73, rotation.cc:122-124)
        V4<R> qi = q_conj(rot);
        V3<R> s = -shift;
        if (has_rot && has_shift) s = rot_apply(qi, s);
        return std::unique_ptr<Map<R, GeoEu<R>>>(new HomMap(s, qi, has_shift, has_rot));
    }
};

// Moebius on the quaternion half-space, moebius.cc:20-56
template <class R> struct MoMap : Map<R, GeoHy<R>> {
    C22<R> m;
    explicit MoMap(C22<R> v) : m(v) {}
    V4<R> apply_pos(V4<R> p) const override {
        V2<R> z = {R(0), R(0)};
        return q_div(cq_mul(m.a, p) + q_of(m.b, z), cq_mul(m.c, p) + q_of(m.d, z));
    }
    V4<R> deriv(V4<R> p, V4<R> v) const {
        V2<R> z = {R(0), R(0)};
        V4<R> u = cq_mul(m.a, p) + q_of(m.b, z);
        V4<R> d = cq_mul(m.c, p) + q_of(m.d, z);
        R d2 = dot(d, d);
        V4<R> g1 = q_div(cq_mul(m.a, v), d);
        V4<R> g21 = q_conj(cq_mul(m.c, v));
        V4<R> g22 = (R(2) * dot(d, cq_mul(m.c, v)) / d2) * q_conj(d);
        V4<R> g2 = q_mul(u, (g21 - g22) / d2);
        return g1 + g2;
    }
    V4<R> apply_dir(V4<R> p, V4<R> v) const override { return normalize(deriv(p, v)); }
    std::unique_ptr<Map<R, GeoHy<R>>> inverse() const override {
        C22<R> n = {m.d, -m.b, -m.c, m.a};  // c22_inverse_n, det = 1 assumed
        return std::unique_ptr<Map<R, GeoHy<R>>>(new MoMap(n));
    }
};

// ray_map.inl:15-21
template <class R, class G> Ray<R, G> ray_map(const Ray<R, G> &ray, const Map<R, G> &m) {
    Ray<R, G> r;
    r.start = m.apply_pos(ray.start);
    r.dir = normalize(m.apply_dir(ray.start, ray.dir));
    return r;
}

// ------------------------------------------------------------ shapes
template <class R, class F, class G> struct Shape {
    virtual ~Shape() {}
    // Returns the distance or a negative value; updates light.ray and normal.
    virtual R detect(Ctx &ctx, typename G::Dir &normal, Light<R, F, G> &light) const = 0;
};

template <class R, class F> struct SphereEu : Shape<R, F, GeoEu<R>> {
    R radius;
    bool unit;
    SphereEu(R r, bool u) : radius(r), unit(u) {}
    // shape/eu/sphere.cc:3-29 (unit) and the radius-r variant
    R detect(Ctx &ctx, V3<R> &normal, Light<R, F, GeoEu<R>> &light) const override {
        const R EPS = Consts<R>::eps();
        bool repeat = ctx.repeat();
        V3<R> pos = light.ray.start, dir = light.ray.dir;
        R b = -dot(dir, pos);
        V3<R> g = pos + dir * b;
        R d = (unit ? R(1) : radius * radius) - length2(g);
        if (d < R(0)) return R(-1);
        d = fsqrt(d);
        R e = b - d;
        R w = EPS * R(2 * (int)repeat - 1);
        if (e < w) {
            e = b + d;
            if (e < w) return R(-1);
        }
        if (unit) {
            V3<R> h = normalize(pos + dir * e);
            normal = h;
            light.ray.start = h;
        } else {
            V3<R> h = pos + dir * e;
            normal = normalize(h);
            light.ray.start = h;
        }
        return e;
    }
};

template <class R, class F> struct PlaneEu : Shape<R, F, GeoEu<R>> {
    // shape/eu/plane.cc:3-21
    R detect(Ctx &ctx, V3<R> &normal, Light<R, F, GeoEu<R>> &light) const override {
        if (ctx.repeat()) return R(-1);
        V3<R> pos = light.ray.start, dir = light.ray.dir;
        if (pos.z * dir.z > R(0)) return R(-1);
        R t = -pos.z / dir.z;
        light.ray.start = {pos.x + dir.x * t, pos.y + dir.y * t, R(0)};
        normal = {R(0), R(0), R(-1)};
        return t;
    }
};

template <class R> R sgn(R a) { return R((R(0) < a) - (a < R(0))); }

// Sphere-traced signed-distance shapes (not in the reference; the method of
// hypertrace_amd/csrc/device/ht/sdf.hh): kind 0 sphere {r}, 1 box {hx, hy, hz},
// 2 torus around y {R, r}. The march runs inside the bounding sphere of radius
// bound + 2 eps, stepping |f| until |f| < eps, at most 128 evaluations; a ray
// leaving the primitive it just hit first steps max(|f|, eps) until |f| >= 2 eps;
// the normal is the normalised tetrahedron gradient estimate (h = 4 eps).
template <class R> struct SdfConsts;
template <> struct SdfConsts<float> { static float eps() { return 1e-4f; } };
template <> struct SdfConsts<double> { static double eps() { return 1e-4; } };
template <> struct SdfConsts<Cnt> { static Cnt eps() { return Cnt(1e-4f); } };

template <class R, class F> struct SdfEu : Shape<R, F, GeoEu<R>> {
    int kind = 0;
    R a = R(0), b = R(0), c = R(0);
    R eval(V3<R> x) const {
        if (kind == 0) return length(x) - a;
        if (kind == 1) {
            V3<R> q = {fabs_(x.x) - a, fabs_(x.y) - b, fabs_(x.z) - c};
            V3<R> qp = {fmax_(q.x, R(0)), fmax_(q.y, R(0)), fmax_(q.z, R(0))};
            return length(qp) + fmin_(fmax_(q.x, fmax_(q.y, q.z)), R(0));
        }
        R qx = fsqrt(x.x * x.x + x.z * x.z) - a;
        return fsqrt(qx * qx + x.y * x.y) - b;
    }
    R bound() const {
        if (kind == 0) return a;
        if (kind == 1) return length(V3<R>{a, b, c});
        return a + b;
    }
    R detect(Ctx &ctx, V3<R> &normal, Light<R, F, GeoEu<R>> &light) const override {
        const R EPS = SdfConsts<R>::eps();
        bool repeat = ctx.repeat();
        V3<R> p = light.ray.start, d = light.ray.dir;
        R rb = bound() + R(2) * EPS;
        R bb = -dot(d, p);
        V3<R> g = p + d * bb;
        R disc = rb * rb - length2(g);
        if (disc < R(0)) return R(-1);
        R s = fsqrt(disc);
        R t1 = bb + s;
        if (t1 < R(0)) return R(-1);
        R t = fmax_(bb - s, R(0));
        bool leaving = repeat, hit = false;
        for (int k = 0; k < 128; ++k) {
            R av = fabs_(eval(p + d * t));
            if (leaving) {
                if (av >= R(2) * EPS) {
                    leaving = false;
                } else {
                    t += fmax_(av, EPS);
                    if (t > t1) break;
                    continue;
                }
            }
            if (av < EPS) {
                hit = true;
                break;
            }
            t += av;
            if (t > t1) break;
        }
        if (!hit) return R(-1);
        V3<R> x = p + d * t;
        R h = R(4) * EPS;
        R f0 = eval(x + V3<R>{h, -h, -h}), f1 = eval(x + V3<R>{-h, -h, h});
        R f2 = eval(x + V3<R>{-h, h, -h}), f3 = eval(x + V3<R>{h, h, h});
        V3<R> n = {f0 - f1 - f2 + f3, -f0 - f1 + f2 + f3, -f0 + f1 - f2 + f3};
        normal = normalize(n);
        light.ray.start = x;
        return t;
    }
};

template <class R, class F> struct CubeEu : Shape<R, F, GeoEu<R>> {
    // shape/eu/cube.cc:4-21
    static R nearest(V3<R> n, V3<R> &normal) {
        bool xy = n.x > n.y, yz = n.y > n.z, xz = n.x > n.z;
        if (xy && xz) {
            normal.x = R(1);
            return n.x;
        } else if (yz) {
            normal.y = R(1);
            return n.y;
        }
        normal.z = R(1);
        return n.z;
    }
    // shape/eu/cube.cc:23-65
    R detect(Ctx &ctx, V3<R> &normal, Light<R, F, GeoEu<R>> &light) const override {
        const R EPS = Consts<R>::eps();
        bool repeat = ctx.repeat();
        V3<R> d = light.ray.dir, s = light.ray.start;
        V3<R> inv = {R(1) / d.x, R(1) / d.y, R(1) / d.z};
        V3<R> vmin = (V3<R>{R(-1), R(-1), R(-1)} - s) * inv;
        V3<R> vmax = (V3<R>{R(1), R(1), R(1)} - s) * inv;
        V3<R> nr = {fmin_(vmin.x, vmax.x), fmin_(vmin.y, vmax.y), fmin_(vmin.z, vmax.z)};
        V3<R> fr = {fmax_(vmin.x, vmax.x), fmax_(vmin.y, vmax.y), fmax_(vmin.z, vmax.z)};
        V3<R> sg = {sgn(d.x), sgn(d.y), sgn(d.z)};
        V3<R> nin = {R(0), R(0), R(0)};
        float din = tofloat(nearest(nr, nin));
        nin = nin * (-sg);
        V3<R> nout = {R(0), R(0), R(0)};
        float dout = -tofloat(nearest(-fr, nout));
        nout = nout * sg;
        if (din > dout) return R(-1);
        R dist = R(din);
        V3<R> nn = nin;
        R w = R(2) * EPS * R(2 * (int)repeat - 1);
        if (dist < w) {
            dist = R(dout);
            nn = nout;
            if (dist < w) return R(-1);
        }
        light.ray.start = light.ray.start + dist * light.ray.dir;
        normal = nn;
        return dist;
    }
};

template <class R, class F> struct PlaneHy : Shape<R, F, GeoHy<R>> {
    // shape/hy/plane.cc:4-37
    R detect(Ctx &ctx, V4<R> &normal, Light<R, F, GeoHy<R>> &light) const override {
        const R EPS = Consts<R>::eps();
        if (ctx.repeat()) return R(-1);
        V4<R> p = light.ray.start, d = light.ray.dir;
        R pd = dot(p, d);
        if (fabs_(pd) < EPS) return R(-1);
        R s = R(1) - length2(p);
        R t = s / (R(2) * pd);
        if (t < -EPS) return R(-1);
        V4<R> h = {p.x + d.x * t, p.y + d.y * t, R(0), R(0)};
        R pxy2 = h.x * h.x + h.y * h.y;
        if (pxy2 > R(1)) return R(-1);
        h.z = fsqrt(R(1) - pxy2);
        light.ray.start = h;
        light.ray.dir = hy_dir_at(p, d, h);
        normal = -h;
        return hy_distance(p, h);
    }
};

template <class R, class F> struct HorosphereHy : Shape<R, F, GeoHy<R>> {
    // shape/hy/horosphere.cc:4-34
    R detect(Ctx &ctx, V4<R> &normal, Light<R, F, GeoHy<R>> &light) const override {
        const R EPS = Consts<R>::eps();
        bool repeat = ctx.repeat();
        V4<R> p = light.ray.start, d = light.ray.dir;
        R dxy = length(V2<R>{d.x, d.y});
        if (p.z < dxy) return R(-1);
        R dt = fsqrt(p.z * p.z - dxy * dxy);
        R t = p.z * d.z - dt;
        R w = EPS * R(2 * (int)repeat - 1);
        if (t < w) {
            t += R(2) * dt;
            if (t < w) return R(-1);
        }
        t /= dxy * dxy;
        V4<R> h = {p.x + d.x * t, p.y + d.y * t, R(1), R(0)};
        light.ray.start = h;
        light.ray.dir = hy_dir_at(p, d, h);
        normal = {R(0), R(0), R(-1), R(0)};
        return hy_distance(p, h);
    }
};

// shape/mapped.inl:22-30
template <class R, class F, class G> struct MappedShape : Shape<R, F, G> {
    std::unique_ptr<Map<R, G>> map, inv;
    std::unique_ptr<Shape<R, F, G>> inner;
    R detect(Ctx &ctx, typename G::Dir &normal, Light<R, F, G> &light) const override {
        light.ray = ray_map(light.ray, *inv);
        R dist = inner->detect(ctx, normal, light);
        normal = map->apply_dir(light.ray.start, normal);  // apply_normal == apply_dir for isometries
        light.ray = ray_map(light.ray, *map);
        return dist;
    }
};

// shape/vector.inl:23-53
template <class R, class F, class G> struct VecShape : Shape<R, F, G> {
    std::vector<std::unique_ptr<Shape<R, F, G>>> items;
    R detect(Ctx &ctx, typename G::Dir &normal, Light<R, F, G> &light) const override {
        const R EPS = Consts<R>::eps();
        uint32_t orig = ctx.hasher;
        R min_dist = R(-1);
        typename G::Dir min_normal{};
        Light<R, F, G> min_light{};
        uint32_t min_hasher = 0;
        for (size_t i = 0; i < items.size(); ++i) {
            Light<R, F, G> cur = light;
            typename G::Dir cn{};
            ctx.hasher = hash_step(ctx.hasher, (uint32_t)i);
            R dist = items[i]->detect(ctx, cn, cur);
            if (dist > -EPS && (min_dist <= -EPS || dist < min_dist)) {
                min_dist = dist;
                min_normal = cn;
                min_light = cur;
                min_hasher = ctx.hasher;
            }
            ctx.hasher = orig;
        }
        if (min_dist > -EPS) {
            normal = min_normal;
            light = min_light;
            ctx.hasher = min_hasher;
            return min_dist;
        }
        return R(-1);
    }
};

// ------------------------------------------------------------ materials
template <class R, class F> struct Material {
    virtual ~Material() {}
    virtual bool interact(Ctx &ctx, V3<R> normal, LightLocal<R, F> &light, C3<F> &emission) const = 0;
};

// rotation.cc:103-120
template <class R> V4<R> look_at_cont(V3<R> d) {
    V3<R> z = {R(0), R(0), R(-1)};
    R c2 = dot(d, z);
    R c = fsqrt((R(1) + c2) / R(2));
    R sm = R(1) / (R(2) * c);
    V3<R> w = sm * cross(z, d);
    return {c, w.x, w.y, w.z};
}
template <class R> V4<R> look_at(V3<R> d) {
    if (d.z < R(0)) return look_at_cont(d);
    R half = Consts<R>::pi() / R(2);
    V4<R> flip = {fcos(half), fsin(half) * R(1), fsin(half) * R(0), fsin(half) * R(0)};
    return q_mul(look_at_cont(-d), flip);
}

template <class R, class F> struct Lambertian : Material<R, F> {
    // material/lambertian.cc:5-14
    bool interact(Ctx &ctx, V3<R> normal, LightLocal<R, F> &light, C3<F> &) const override {
        if (dot(light.dir, normal) > R(0)) normal = -normal;
        V3<R> r = hemisphere_cosine<R>(ctx.rng);
        light.dir = rot_apply(look_at(-normal), r);
        return true;
    }
};
template <class R, class F> struct Specular : Material<R, F> {
    // material/specular.cc:3-7
    bool interact(Ctx &, V3<R> normal, LightLocal<R, F> &light, C3<F> &) const override {
        light.dir = light.dir - (R(2) * dot(light.dir, normal)) * normal;
        return true;
    }
};
template <class R, class F> struct Refractive : Material<R, F> {
    R index;
    explicit Refractive(R k) : index(k) {}
    // material/refractive.cc:3-22
    bool interact(Ctx &, V3<R> normal, LightLocal<R, F> &light, C3<F> &) const override {
        const R EPS = Consts<R>::eps();
        V3<R> d = light.dir;
        R k = index;
        R a = dot(d, normal);
        if (a < -EPS) k = R(1) / k;
        R x = (a * a - R(1)) * (k * k) + R(1);
        if (x > EPS) {
            R b = sgn(a) * fsqrt(x);
            d = d * k + (b - a * k) * normal;
        } else {
            d = d - (R(2) * a) * normal;
        }
        light.dir = d;
        return true;
    }
};
template <class R, class F> struct Transparent : Material<R, F> {
    bool interact(Ctx &, V3<R>, LightLocal<R, F> &, C3<F> &) const override { return true; }
};
template <class R, class F> struct Absorbing : Material<R, F> {
    bool interact(Ctx &, V3<R>, LightLocal<R, F> &, C3<F> &) const override { return false; }
};
template <class R, class F> struct Colored : Material<R, F> {
    C3<F> color;
    std::unique_ptr<Material<R, F>> inner;
    // material/colored.inl:14-23
    bool interact(Ctx &ctx, V3<R> n, LightLocal<R, F> &light, C3<F> &e) const override {
        light.intensity = light.intensity * color;
        return inner->interact(ctx, n, light, e);
    }
};
template <class R, class F> struct Emissive : Material<R, F> {
    C3<F> emission;
    std::unique_ptr<Material<R, F>> inner;
    // material/emissive.inl:14-23
    bool interact(Ctx &ctx, V3<R> n, LightLocal<R, F> &light, C3<F> &e) const override {
        e = e + light.intensity * emission;
        return inner->interact(ctx, n, light, e);
    }
};
template <class R, class F> struct Mixture : Material<R, F> {
    std::vector<R> portions;
    std::vector<std::unique_ptr<Material<R, F>>> items;
    // material/mixture/impl.inl:15-27, mixture/macros.hh:3-13
    bool interact(Ctx &ctx, V3<R> n, LightLocal<R, F> &light, C3<F> &e) const override {
        R alpha = uniform<R>(ctx.rng);
        for (size_t i = 0; i < items.size(); ++i) {
            alpha -= portions[i];
            if (alpha < R(0)) return items[i]->interact(ctx, n, light, e);
        }
        return false;
    }
};

// ------------------------------------------------------------ objects
template <class R, class G> struct Cache {
    typename G::Dir normal;
    uint32_t idx[16];
};

template <class R, class F, class G> struct Object {
    virtual ~Object() {}
    virtual R detect(Ctx &ctx, Cache<R, G> &cache, Light<R, F, G> &light) const = 0;
    virtual bool interact(Ctx &ctx, Cache<R, G> &cache, Light<R, F, G> &light, C3<F> &emission) const = 0;
};

template <class R, class F> inline LightLocal<R, F> to_local(const Light<R, F, GeoEu<R>> &l) {
    return {l.intensity, l.ray.dir};
}
template <class R, class F> inline void from_local(Light<R, F, GeoEu<R>> &l, const LightLocal<R, F> &ll) {
    l.ray.dir = ll.dir;
    l.intensity = ll.intensity;
}
template <class R, class F> inline LightLocal<R, F> to_local(const Light<R, F, GeoHy<R>> &l) {
    return {l.intensity, V3<R>{l.ray.dir.x, l.ray.dir.y, l.ray.dir.z}};
}
template <class R, class F> inline void from_local(Light<R, F, GeoHy<R>> &l, const LightLocal<R, F> &ll) {
    l.ray.dir = {ll.dir.x, ll.dir.y, ll.dir.z, R(0)};
    l.intensity = ll.intensity;
}

// object/covered.inl:23-56
template <class R, class F, class G> struct Covered : Object<R, F, G> {
    std::unique_ptr<Shape<R, F, G>> shape;
    std::unique_ptr<Material<R, F>> material;
    R detect(Ctx &ctx, Cache<R, G> &cache, Light<R, F, G> &light) const override {
        return shape->detect(ctx, cache.normal, light);
    }
    bool interact(Ctx &ctx, Cache<R, G> &cache, Light<R, F, G> &light, C3<F> &e) const override {
        LightLocal<R, F> ll = to_local(light);
        bool ret = material->interact(ctx, G::to_local(light.ray.start, cache.normal), ll, e);
        from_local(light, ll);
        return ret;
    }
};

// Hyperbolic generators in R, geometry/hyperbolic.cc:40-53 (c*C1 = (c*1, c*0)).
template <class R> C22<R> hy_xshift_m(R l) {
    R l2 = l / R(2);
    R c = fcosh(l2), s = fsinh(l2);
    return {{c * R(1), c * R(0)}, {s * R(1), s * R(0)}, {s * R(1), s * R(0)}, {c * R(1), c * R(0)}};
}
template <class R> C22<R> hy_zrotate_m(R phi) {
    R c = fcos(phi / R(2)), s = fsin(phi / R(2));
    return {{c, s}, {R(0), R(0)}, {R(0), R(0)}, {c, -s}};
}
template <class R> R fract_ocl(R x, R &ip) {  // OpenCL fract: kept below 1
    ip = ffloor(x);
    return fmin_(x - ip, Consts<R>::fract_max());
}
inline int rem_i(int a, int b) { return ((a % b) + b) % b; }

// Tiled plane: object/hy/tiled_plane/impl.inl:22-148.
template <class R, class F> struct TiledPlaneObj : Object<R, F, GeoHy<R>> {
    int tiling = 0;
    R cell = R(0), border = R(0);
    std::vector<std::unique_ptr<Material<R, F>>> mats;
    std::unique_ptr<Material<R, F>> border_mat;
    R detect(Ctx &ctx, Cache<R, GeoHy<R>> &cache, Light<R, F, GeoHy<R>> &light) const override {
        return PlaneHy<R, F>().detect(ctx, cache.normal, light);
    }
    const Material<R, F> *select(V4<R> p) const {
        if (tiling != 1 && tiling != 2) return mats[0].get();
        const R PI = Consts<R>::pi();
        bool w = false;
        R Q = fsqrt(fcos(PI / R(4) + PI / R(5)) / fcos(PI / R(4) - PI / R(5)));
        R T = fsqrt(fcos(PI / R(4) + PI / R(5)) * fcos(PI / R(4) - PI / R(5)));
        R S = (fcos(PI / R(4)) - fsin(PI / R(5))) / T;
        R L = T / fcos(PI / R(4));
        R K = L * (R(2) * fcos(PI / R(5)) - R(1) / fcos(PI / R(5)));
        Q = flog((R(1) + Q) / (R(1) - Q));
        S = flog((R(1) + S) / (R(1) - S));
        if (p.x < R(0)) {
            p.x = -p.x;
            w = !w;
        }
        if (p.y < R(0)) {
            p.y = -p.y;
            w = !w;
        }
        p = MoMap<R>(c22_dot(hy_xshift_m(-Q), hy_zrotate_m(-PI / R(4)))).apply_pos(p);
        bool e = false;
        for (int j = 0; j < 5; ++j) {
            bool a[3] = {false, false, false};
            for (int i = 0; i < 3 - (int)e; ++i) {
                R o = R(2) * PI * R(i - 1) / R(5);
                a[i] = (fcos(o) * p.x + fsin(o) * p.y) < L;
            }
            a[2] = a[2] || e;
            int s = (int)a[0] + (int)a[1] + (int)a[2];
            if (s == 3) break;
            if (s == 2) {
                int i = (int)(!a[1]) + 2 * (int)(!a[2]);
                R o = R(2) * PI * R(i - 1) / R(5);
                p = MoMap<R>(c22_dot(hy_zrotate_m(-PI / R(5)), c22_dot(hy_xshift_m(R(-2) * S), hy_zrotate_m(-o))))
                        .apply_pos(p);
                e = true;
                w = !w;
            } else {
                int i = (int)a[0];
                R o = PI * R(2 * i - 1) / R(5);
                p = MoMap<R>(c22_dot(hy_xshift_m(R(-2) * Q), hy_zrotate_m(-o))).apply_pos(p);
                e = false;
            }
        }
        bool bh = false;
        for (int i = 0; i < 5; ++i) {
            R o = R(2) * PI * R(i) / R(5);
            R dd = fcos(o) * p.x + fsin(o) * p.y;
            bh = bh || (dd > (L - border * p.z));
            if (tiling == 2) {
                R ps = K + dd;
                w = w ^ (ps < R(0));
                bh = bh || (fabs_(ps) < border * p.z);
            }
        }
        if (!bh) return (w || mats.size() < 2) ? mats[0].get() : mats[1].get();
        return border_mat.get();
    }
    bool interact(Ctx &ctx, Cache<R, GeoHy<R>> &cache, Light<R, F, GeoHy<R>> &light, C3<F> &e) const override {
        const Material<R, F> *m = select(light.ray.start);
        LightLocal<R, F> ll = to_local(light);
        bool ret = m->interact(ctx, GeoHy<R>::to_local(light.ray.start, cache.normal), ll, e);
        from_local(light, ll);
        return ret;
    }
};

// Tiled horosphere: object/hy/tiled_horosphere/impl.inl:29-110.
template <class R, class F> struct TiledHoroObj : Object<R, F, GeoHy<R>> {
    int tiling = 0;
    R cell = R(0), border = R(0);
    std::vector<std::unique_ptr<Material<R, F>>> mats;
    std::unique_ptr<Material<R, F>> border_mat;
    R detect(Ctx &ctx, Cache<R, GeoHy<R>> &cache, Light<R, F, GeoHy<R>> &light) const override {
        return HorosphereHy<R, F>().detect(ctx, cache.normal, light);
    }
    const Material<R, F> *select(V4<R> p) const {
        int n = (int)mats.size();
        int material_no = 0;
        bool is_border = false;
        R br = border / cell;
        if (tiling == 1) {
            R kx, ky;
            R fx = fract_ocl(p.x / cell, kx), fy = fract_ocl(p.y / cell, ky);
            if (fx < br || fx > R(1) - br || fy < br || fy > R(1) - br) {
                is_border = true;
            } else {
                int hx = rem_i((int)tofloat(kx), 2), hy = rem_i((int)tofloat(ky), 2);
                material_no = 3 * hy - 2 * hx * hy + hx;
            }
        } else if (tiling == 2) {
            R s3 = fsqrt(R(3));
            R bxx = R(2) / s3, bxy = R(0), byx = R(-1) / s3, byy = R(1);
            R s = cell * s3 / R(2);
            R hxv = (bxx * p.x + bxy * p.y) / s, hyv = (byx * p.x + byy * p.y) / s;
            int hx = (int)tofloat(ffloor((ffloor(hxv) - ffloor(hyv)) / R(3)));
            int hy = (int)tofloat(ffloor((ffloor(hxv + hyv) - R(hx)) / R(2)));
            R sx = R(hx) * R(2) + R(hy) * R(1), sy = R(hx) * R(-1) + R(hy) * R(1);
            hxv = hxv - sx;
            hyv = hyv - sy;
            if (fabs_(hxv - R(1)) > R(1) - br || fabs_(hyv) > R(1) - br || fabs_(hxv + hyv - R(1)) > R(1) - br) {
                is_border = true;
            } else {
                material_no = 2 * hx + hy;
            }
        }
        if (is_border) return border_mat.get();
        return mats[rem_i(material_no, n)].get();
    }
    bool interact(Ctx &ctx, Cache<R, GeoHy<R>> &cache, Light<R, F, GeoHy<R>> &light, C3<F> &e) const override {
        const Material<R, F> *m = select(light.ray.start);
        LightLocal<R, F> ll = to_local(light);
        bool ret = m->interact(ctx, GeoHy<R>::to_local(light.ray.start, cache.normal), ll, e);
        from_local(light, ll);
        return ret;
    }
};

// object/mapped.inl:24-51
template <class R, class F, class G> struct MappedObject : Object<R, F, G> {
    std::unique_ptr<Map<R, G>> map, inv;
    std::unique_ptr<Object<R, F, G>> inner;
    R detect(Ctx &ctx, Cache<R, G> &cache, Light<R, F, G> &light) const override {
        light.ray = ray_map(light.ray, *inv);
        R dist = inner->detect(ctx, cache, light);
        light.ray = ray_map(light.ray, *map);
        return dist;
    }
    bool interact(Ctx &ctx, Cache<R, G> &cache, Light<R, F, G> &light, C3<F> &e) const override {
        light.ray = ray_map(light.ray, *inv);
        bool ret = inner->interact(ctx, cache, light, e);
        light.ray = ray_map(light.ray, *map);
        return ret;
    }
};

// ------------------------------------------------------------ culled-scan flop counting
// The HIP kernel's nearest-hit scan over a top-level vector of mapped
// primitives tests each element's world-space bounding sphere against the
// ray's geodesic in FP64 and runs the exact test only for candidates; each
// ray then reruns its winner's exact test and maps that one hit back to the
// world frame (hypertrace_amd/csrc/device/ht/scan.hh). To count the flops of
// that algorithm, its bounds and cull tests are restated here on a counting
// double (FP64 flops, g_flops64); the exact tests run on the counting float
// as everywhere else. Only the flop count uses this path; it selects the same
// hit as the reference loop below.
thread_local uint64_t g_flops64 = 0;
bool g_count_culled = false;
// The two-level scan's count (program options HT_SCALAR_BOUNDS with
// HT_CULL_GROUPS = g_count_group): the group spheres' tests, then the leaf
// tests of the groups the ray passes; with g_count_queue (HT_GROUP_QUEUE)
// also the rebuild of the ray's cull state for each (group, ray) pair.
int g_count_group = 0;
bool g_count_queue = false;
// HT_HIT_ORDER (with the group queue): each passing (element, ray) pair's
// crossing interval; exact tests only for the pairs the interval keeps.
bool g_count_hit_order = false;

// Crossing interval of a pair (scan.hh CullRayHy::interval): drop = no
// crossing of the outer sphere ahead of the start; lb = e^D of the first
// crossing ahead (1 if the start is inside the outer sphere), rounded down;
// certain: the geodesic crosses the inner sphere ahead, no farther than ub.
struct HitIv {
    float lb = 1.0f, ub = 0.0f;
    bool drop = false, certain = false;
};

struct CntD {
    double v;
    CntD() : v(0) {}
    CntD(double x) : v(x) {}
};
inline CntD operator+(CntD a, CntD b) { ++g_flops64; return CntD(a.v + b.v); }
inline CntD operator-(CntD a, CntD b) { ++g_flops64; return CntD(a.v - b.v); }
inline CntD operator*(CntD a, CntD b) { ++g_flops64; return CntD(a.v * b.v); }
inline CntD operator/(CntD a, CntD b) { ++g_flops64; return CntD(a.v / b.v); }
inline CntD sqrtd(CntD a) { ++g_flops64; return CntD(std::sqrt(a.v)); }

struct BoundF {
    float cx = 0, cy = 0, cz = 0, r = -1, ri = -1;
    bool horo = false;
    bool horo_node = false;  // a tree node over a horoball: no "behind" cull
};

inline float next_up(float f) { uint32_t u; std::memcpy(&u, &f, 4); ++u; std::memcpy(&f, &u, 4); return f; }
inline float next_down(float f) { uint32_t u; std::memcpy(&u, &f, 4); --u; std::memcpy(&f, &u, 4); return f; }

// Sphere (cx, cy, cz, r) rounded to FP32: outer radius inflated by the margin
// and the centre's rounding and rounded up, inner radius shrunk and rounded down.
inline BoundF make_bound(double cx, double cy, double cz, double r, double margin) {
    BoundF b;
    b.cx = (float)cx;
    b.cy = (float)cy;
    b.cz = (float)cz;
    double err = std::fabs((double)b.cx - cx) + std::fabs((double)b.cy - cy) + std::fabs((double)b.cz - cz);
    double ro = r * (1.0 + margin) + err + 1e-30;
    b.r = (float)ro;
    if ((double)b.r < ro) b.r = next_up(b.r);
    double rin = r * (1.0 - margin) - err;
    b.ri = rin > 0.0 ? (float)rin : 0.0f;
    if ((double)b.ri > rin && b.ri > 0.0f) b.ri = next_down(b.ri);
    return b;
}

// Bound of a unit hemisphere (kind 4) or of the horosphere z = 1 (kind 5)
// under the Moebius map [[a, b], [c, d]]: the hemisphere over the image of
// the unit circle, centre (b conj(d) - a conj(c)) / (|d|^2 - |c|^2), radius
// 1 / ||d|^2 - |c|^2|; the horoball tangent to the boundary at a / c with
// diameter 1 / |c|^2. Margin 1e-3 + 1e-4 n2^2 for the squared norm n2.
template <class R> BoundF mo_bound(const C22<R> &m, int kind) {
    double ar = tofloat(m.a.x), ai = tofloat(m.a.y), br = tofloat(m.b.x), bi = tofloat(m.b.y);
    double cr = tofloat(m.c.x), ci = tofloat(m.c.y), dr = tofloat(m.d.x), di = tofloat(m.d.y);
    double n2 = ar * ar + ai * ai + br * br + bi * bi + cr * cr + ci * ci + dr * dr + di * di;
    double margin = 1e-3 + 1e-4 * n2 * n2;
    if (kind == 5) {
        double c2 = cr * cr + ci * ci;
        if (!(c2 > 1e-12 * n2)) return BoundF();
        double rad = 0.5 / c2;
        BoundF b = make_bound((ar * cr + ai * ci) / c2, (ai * cr - ar * ci) / c2, rad, rad, margin);
        b.horo = true;
        return b;
    }
    if (kind == 4) {
        double den = (dr * dr + di * di) - (cr * cr + ci * ci);
        if (!(std::fabs(den) > 1e-9 * n2)) return BoundF();
        double nx = (br * dr + bi * di) - (ar * cr + ai * ci);
        double ny = (bi * dr - br * di) - (ai * cr - ar * ci);
        return make_bound(nx / den, ny / den, 0.0, 1.0 / std::fabs(den), margin);
    }
    return BoundF();
}

// Straight Euclidean ray p + t d, t >= 0, against a sphere.
struct CullRayEu {
    static constexpr bool HYPERBOLIC = false;
    CntD px, py, pz, dx, dy, dz;
    CntD tp, ipz, iux, iuy;  // unused: straight rays have no crossing interval
    template <class R> static CullRayEu make(const V3<R> &p, const V3<R> &d) {
        return {CntD(tofloat(p.x)), CntD(tofloat(p.y)), CntD(tofloat(p.z)),
                CntD(tofloat(d.x)), CntD(tofloat(d.y)), CntD(tofloat(d.z))};
    }
    // branch-free, as the kernel evaluates it: every term is computed
    bool candidate(const BoundF &b, bool = false) const {
        CntD sx = CntD(b.cx) - px, sy = CntD(b.cy) - py, sz = CntD(b.cz) - pz;
        CntD r = b.r;
        CntD t = sx * dx + sy * dy + sz * dz;
        CntD qx = sx - t * dx, qy = sy - t * dy, qz = sz - t * dz;
        bool hit = !(t.v < -r.v) && (qx * qx + qy * qy + qz * qz).v <= (r * r).v;
        return b.r < 0.f || hit;
    }
    // straight rays: no crossing interval
    template <class R> void interval_consts(const V3<R> &, const V3<R> &) {}
    HitIv interval(const BoundF &, bool) const { return HitIv(); }
};

// Hyperbolic geodesic from p with tangent d: the semicircle in the vertical
// plane through p along u = d.xy / |d.xy|, centre p.xy + gam u, radius rg;
// the part ahead of p. A sphere is culled when the plane misses it, its cut
// lies wholly behind p, the circles are disjoint, or one lies strictly inside
// the other; horoballs near the geodesic through their tangent point keep
// the behind-the-start case (the horosphere test accepts crossings behind
// the start for near-vertical local rays, shape/hy/horosphere.cc:16-24).
struct CullRayHy {
    static constexpr bool HYPERBOLIC = true;
    CntD px, py, pz, dx, dy, dz, dn2, m, pz2, pz2n, four_pz2n, two_pzdz;
    bool valid = false;
    // no square root and no division, as the kernel builds it: the tests
    // are multiplied through by m = |d.xy|^2 (see candidate)
    template <class R> static CullRayHy make(const V4<R> &p, const V4<R> &d) {
        CullRayHy c;
        CntD x = tofloat(d.x), y = tofloat(d.y), z = tofloat(d.z);
        CntD pz = tofloat(p.z);
        CntD m = x * x + y * y;
        CntD dn2 = m + z * z;
        c.valid = m.v > (CntD(1e-8) * dn2).v && pz.v > 0.0;
        c.px = tofloat(p.x);
        c.py = tofloat(p.y);
        c.pz = pz;
        c.dx = x;
        c.dy = y;
        c.dz = z;
        c.dn2 = dn2;
        c.m = m;
        c.pz2 = pz * pz;
        c.pz2n = c.pz2 * dn2;
        c.four_pz2n = CntD(4.0) * c.pz2n;
        c.two_pzdz = CntD(2.0) * pz * z;
        return c;
    }
    // Crossing interval (scan.hh CullRayHy::interval, which derives it):
    // with w = e^D along the ray, a sphere's crossings with the geodesic are
    // the roots of gam w^2 - 2 bet w + alp = 0.
    CntD tp, ipz, iux, iuy;  // the interval's per-ray constants (the kernel's hit_interval_consts)
    template <class R> void interval_consts(const V4<R> &p, const V4<R> &d) {
        CntD x = tofloat(d.x), y = tofloat(d.y), z = tofloat(d.z), pzz = tofloat(p.z);
        CntD m = x * x + y * y, n = m + z * z;
        bool ok = m.v > 1e-8 * n.v && pzz.v > 0.0;
        CntD sm = sqrtd(ok ? m : CntD(1.0)), dn = sqrtd(ok ? n : CntD(1.0)), inv = CntD(1.0) / sm;
        CntD t = z.v >= 0.0 ? (dn + z) * inv : sm / (dn - z);
        tp = ok ? t : CntD(0.0);
        ipz = ok ? CntD(1.0) / pzz : CntD(0.0);
        iux = x * inv;
        iuy = y * inv;
    }
    static float down(double x) {
        float f = (float)x;
        if ((double)f > x) f = next_down(f);
        return f;
    }
    static float up(double x) {
        float f = (float)x;
        if ((double)f < x) f = next_up(f);
        return f;
    }
    HitIv interval(const BoundF &b, bool exempt) const {
        HitIv iv;
        if (!(tp.v > 0.0) || exempt || b.r < 0.f) return iv;
        CntD rx = CntD(b.cx) - px, ry = CntD(b.cy) - py;
        CntD sig = rx * iux + ry * iuy, delta = CntD(0.0) - rx * iuy + ry * iux;
        CntD d2p = delta * delta;
        CntD S = sig * ipz, Z = CntD(b.cz) * ipz, ipz2 = ipz * ipz;
        CntD po2 = (CntD(b.r) * CntD(b.r) - d2p) * ipz2;
        CntD zm = CntD(1.0) - Z;
        if ((S * S + zm * zm).v < po2.v) return iv;
        if (!(po2.v > 0.0)) {
            iv.drop = true;
            return iv;
        }
        CntD tp2 = tp * tp;
        CntD u = CntD(1.0) + S * tp, v = tp - S, zt = Z * tp;
        CntD a0 = u * u + zt * zt, g0 = v * v + Z * Z, bet = (tp2 + CntD(1.0)) * Z;
        CntD pi2 = (CntD(b.ri) * CntD(b.ri) - d2p) * ipz2;
        if (pi2.v > 0.0) {
            CntD ai = a0 - pi2 * tp2, gi = g0 - pi2;
            CntD di = bet * bet - ai * gi;
            if (di.v >= 0.0) {
                CntD qi = bet + sqrtd(di);
                if (qi.v > 0.0 && (gi.v <= 0.0 || qi.v >= gi.v * (1.0 + 1e-6))) {
                    double ub = (ai / qi).v * (1.0 + 1e-9);
                    if (ub >= 1.0 && ub < 1e30) {
                        iv.certain = true;
                        iv.ub = up(ub);
                    }
                }
            }
        }
        CntD ao = a0 - po2 * tp2, go = g0 - po2;
        CntD disc = bet * bet - ao * go;
        CntD q = bet + sqrtd(disc.v > 0.0 ? disc : CntD(0.0));
        if (!iv.certain && (disc.v < 0.0 || (go.v > 0.0 && q.v < go.v * (1.0 - 1e-6)))) {
            iv.drop = true;
            return iv;
        }
        double w = (ao / q).v;
        iv.lb = w > 1.0 ? down(w * (1.0 - 1e-9)) : 1.0f;
        return iv;
    }
    bool toward_tangent_point(double tx, double ty) const {
        CntD vx = CntD(tx) - px, vy = CntD(ty) - py;
        CntD v2 = vx * vx + vy * vy, h = pz;
        CntD tz = CntD(0.5) * (v2 - h * h);
        CntD n = h * (dx * vx + dy * vy) + dz * tz;
        CntD den2 = (h * h * v2 + tz * tz) * dn2;
        return (den2 - n * n).v < (CntD(1e-2) * den2).v;
    }
    // branch-free, as the kernel evaluates it: every term is computed; the
    // tangent-point test only for a horoball whose cut lies behind p
    // branch-free, as the kernel evaluates it: every term is computed; the
    // tangent-point test only for a horoball whose cut lies behind p; a tree
    // node (node = true) skips the inner-sphere test, which cannot cull it
    bool candidate(const BoundF &b, bool node = false) const {
        CntD rx = CntD(b.cx) - px, ry = CntD(b.cy) - py;
        CntD S = rx * dx + ry * dy;
        CntD D = CntD(0.0) - rx * dy + ry * dx;
        CntD d2 = D * D;
        CntD rp = CntD(b.r) * CntD(b.r) * m - d2;
        CntD s2 = S * S;
        bool cull = rp.v < 0.0;
        bool behind = S.v < 0.0 && s2.v > rp.v;
        if (behind && b.horo_node) behind = false;
        else if (behind && b.horo) behind = !toward_tangent_point(b.cx, b.cy);
        cull = cull || behind;
        CntD o = s2 - two_pzdz * S + (CntD(b.cz) * CntD(b.cz) - pz2) * m - rp;
        CntD o2 = o * o, f = four_pz2n * rp;
        cull = cull || (o.v > 0.0 && o2.v > f.v);
        cull = cull || (pz2n.v > rp.v && o.v < 0.0 && o2.v > f.v);
        if (!node) {
            CntD ri2 = CntD(b.ri) * CntD(b.ri);
            CntD rpi = ri2 * m - d2;
            CntD in2 = (ri2 - CntD(b.r) * CntD(b.r)) * m - o;
            cull = cull || (rpi.v > pz2n.v && in2.v > 0.0 && (in2 * in2).v > (four_pz2n * rpi).v);
        }
        return b.r < 0.f || !valid || !cull;
    }
};

template <class R> inline CullRayEu cull_ray(const Ray<R, GeoEu<R>> &r) { return CullRayEu::make(r.start, r.dir); }
template <class R> inline CullRayHy cull_ray(const Ray<R, GeoHy<R>> &r) { return CullRayHy::make(r.start, r.dir); }

// World bound of one vector element; false if the element is not a mapped
// single primitive (then the kernel scans the whole vector plainly).
template <class R, class F> bool element_bound(const Object<R, F, GeoHy<R>> *o, BoundF &b);
template <class R, class F> bool element_bound(const Object<R, F, GeoEu<R>> *o, BoundF &b);

// object/vector.inl:24-74
template <class R, class F, class G> struct VecObject : Object<R, F, G> {
    std::vector<std::unique_ptr<Object<R, F, G>>> items;
    int level = 0;
    std::vector<BoundF> bounds;
    bool cullable = false;

    // the group tree of the device's bounds_fill (scan.hh build_tree /
    // fill_scalar_block): leaves in Morton order of the FP32 bound centres
    // (element index order among equal codes, unbounded elements after all
    // others, padding last) over P = the program's bounds capacity, nodes
    // merged bottom-up in FP64 and rounded outwards to FP32
    size_t cap = 0;
    std::vector<uint32_t> order;   // leaf j -> element index, 0xFFFF = padding
    std::vector<BoundF> group;     // the group nodes covering g_count_group leaves each
    std::vector<bool> group_empty; // a node over padding only

    static uint32_t expand10(uint32_t v) {
        v &= 1023u;
        v = (v | (v << 16)) & 0x030000FFu;
        v = (v | (v << 8)) & 0x0300F00Fu;
        v = (v | (v << 4)) & 0x030C30C3u;
        v = (v | (v << 2)) & 0x09249249u;
        return v;
    }
    static V4<float> merge_spheres(V4<float> a, V4<float> b) {
        double dx = (double)b.x - a.x, dy = (double)b.y - a.y, dz = (double)b.z - a.z;
        double d = std::sqrt(dx * dx + dy * dy + dz * dz);
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
        double err = std::fabs((double)fx - cx) + std::fabs((double)fy - cy) + std::fabs((double)fz - cz);
        double ro = r * (1.0 + 1e-6) + err + 1e-30;
        float fr = (float)ro;
        if ((double)fr < ro) fr = next_up(fr);
        return {fx, fy, fz, fr};
    }
    void build_groups(int g) {
        size_t n = items.size();
        cap = 64;
        while (cap < n) cap *= 2;
        order.clear();
        group.clear();
        group_empty.clear();
        if (g <= 0 || cap > 1024 || n < 16) return;  // the device's flat scan
        const size_t P = cap;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (auto &b : bounds) {
            if (b.r < 0) continue;
            float c[3] = {b.cx, b.cy, b.cz};
            for (int a = 0; a < 3; ++a) lo[a] = std::fmin(lo[a], c[a]), hi[a] = std::fmax(hi[a], c[a]);
        }
        std::vector<uint64_t> key(P, ~0ull);
        for (size_t i = 0; i < n; ++i) {
            uint32_t code = 0xFFFFFFFFu;
            if (bounds[i].r >= 0) {
                float c[3] = {bounds[i].cx, bounds[i].cy, bounds[i].cz};
                code = 0;
                for (int a = 0; a < 3; ++a) {
                    float u = hi[a] > lo[a] ? (c[a] - lo[a]) / (hi[a] - lo[a]) : 0.5f;
                    uint32_t q = (uint32_t)std::fmin(std::fmax(u * 1024.0f, 0.0f), 1023.0f);
                    code |= expand10(q) << (2 - a);
                }
            }
            key[i] = ((uint64_t)code << 32) | i;
        }
        std::sort(key.begin(), key.end());
        order.resize(P);
        for (size_t j = 0; j < P; ++j) {
            uint32_t low = (uint32_t)(key[j] & 0xFFFFFFFFull);
            order[j] = low < n ? low : 0xFFFFu;
        }
        // nodes k = 1 .. P - 1 of the implicit tree, children 2k, 2k + 1
        std::vector<V4<float>> node(P);
        std::vector<char> nh(P, 0);
        for (size_t first = P / 2; first >= 1; first >>= 1) {
            for (size_t k = first; k < 2 * first; ++k) {
                V4<float> sp[2];
                bool h[2];
                for (int c = 0; c < 2; ++c) {
                    size_t ch = 2 * k + c;
                    if (ch >= P) {
                        uint32_t e = order[ch - P];
                        if (e == 0xFFFFu) {
                            sp[c] = {0.f, 0.f, 0.f, -2.f};
                            h[c] = false;
                        } else {
                            const BoundF &bb = bounds[e];
                            sp[c] = bb.r < 0 ? V4<float>{0.f, 0.f, 0.f, -1.f} : V4<float>{bb.cx, bb.cy, bb.cz, bb.r};
                            h[c] = bb.horo;
                        }
                    } else {
                        sp[c] = node[ch];
                        h[c] = nh[ch] != 0;
                    }
                }
                V4<float> m;
                if (sp[0].w == -2.f) m = sp[1];
                else if (sp[1].w == -2.f) m = sp[0];
                else if (sp[0].w < 0.f || sp[1].w < 0.f) m = {0.f, 0.f, 0.f, -1.f};
                else m = merge_spheres(sp[0], sp[1]);
                node[k] = m;
                nh[k] = (h[0] || h[1]) ? 1 : 0;
            }
            if (first == 1) break;
        }
        size_t FIRST = P / (size_t)g;
        for (size_t k = FIRST; k < 2 * FIRST; ++k) {
            BoundF q;
            q.cx = node[k].x, q.cy = node[k].y, q.cz = node[k].z, q.r = node[k].w;
            q.ri = 0.f;
            q.horo_node = nh[k] != 0;
            group.push_back(q);
            group_empty.push_back(node[k].w == -2.f);
        }
    }

    void prepare_cull() {
        bounds.assign(items.size(), BoundF());
        cullable = !items.empty() && items.size() <= 4096;
        for (size_t i = 0; i < items.size() && cullable; ++i) cullable = element_bound(items[i].get(), bounds[i]);
        if (cullable && g_count_group > 0) build_groups(g_count_group);
    }

    // the candidate elements of a ray, in the order the scan finds them
    template <class CR, class Ray_> void candidates(const CR &cr, const Ray_ &ray, std::vector<uint32_t> &out,
                                                    std::vector<HitIv> *ivs = nullptr, uint32_t orig = 0,
                                                    uint32_t prev = 0) const {
        out.clear();
        if (ivs) ivs->clear();
        if (group.empty()) {
            for (size_t i = 0; i < items.size(); ++i)
                if (cr.candidate(bounds[i])) out.push_back((uint32_t)i);
            return;
        }
        const size_t g = (size_t)g_count_group;
        for (size_t k = 0; k < group.size(); ++k) {
            if (group_empty[k] || !cr.candidate(group[k], true)) continue;
            CR c2 = cr;
            if (g_count_queue) c2 = cull_ray(ray);  // the pair's lane rebuilds the ray's cull state
            if (ivs) c2.tp = cr.tp, c2.ipz = cr.ipz, c2.iux = cr.iux, c2.iuy = cr.iuy;  // read from the ray's slot
            for (size_t j = 0; j < g; ++j) {
                uint32_t e = order[k * g + j];
                if (e != 0xFFFFu && c2.candidate(bounds[e])) {
                    if (ivs) {
                        // the ray leaving this element: its test applies the repeat rule
                        bool exempt = hash_step(orig, e) == prev;
                        if constexpr (CR::HYPERBOLIC) {
                            if (!exempt && bounds[e].horo) exempt = c2.toward_tangent_point(bounds[e].cx, bounds[e].cy);
                        }
                        HitIv iv = c2.interval(bounds[e], exempt);
                        if (iv.drop) continue;
                        ivs->push_back(iv);
                    }
                    out.push_back(e);
                }
            }
        }
    }

    // The culled scan's work: cull tests, exact tests of the candidates
    // without the map back, then the winner's exact test and its map back.
    R detect_culled(Ctx &ctx, Cache<R, G> &cache, Light<R, F, G> &light) const {
        const R EPS = Consts<R>::eps();
        uint32_t orig = ctx.hasher;
        auto cr = cull_ray(light.ray);
        R min_dist = R(-1);
        uint32_t min_index = 0;
        std::vector<uint32_t> cand;
        if (g_count_hit_order && !group.empty()) {
            // the per-ray interval constants, once per scan
            cr.interval_consts(light.ray.start, light.ray.dir);
            std::vector<HitIv> ivs;
            candidates(cr, light.ray, cand, &ivs, orig, ctx.prev_hash);
            // the ring's filter against the ray's nearest certain hit (the
            // entry keeps the upper 16 bits of lb: rounded down)
            float t = INFINITY;
            for (const HitIv &iv : ivs)
                if (iv.certain) t = std::min(t, iv.ub);
            std::vector<uint32_t> kept;
            for (size_t k = 0; k < cand.size(); ++k) {
                uint32_t bits;
                std::memcpy(&bits, &ivs[k].lb, 4);
                bits &= 0xFFFF0000u;
                float lbf;
                std::memcpy(&lbf, &bits, 4);
                if (!(lbf >= t * 1.0203f)) kept.push_back(cand[k]);
            }
            cand.swap(kept);
        } else {
            candidates(cr, light.ray, cand);
        }
        std::sort(cand.begin(), cand.end());  // the reference loop's order decides ties
        for (uint32_t i : cand) {
            auto mo = static_cast<const MappedObject<R, F, G> *>(items[i].get());
            Light<R, F, G> cur = light;
            Cache<R, G> cc = cache;
            ctx.hasher = hash_step(orig, (uint32_t)i);
            cur.ray = ray_map(cur.ray, *mo->inv);
            R dist = mo->inner->detect(ctx, cc, cur);
            ctx.hasher = orig;
            if (dist > -EPS && (min_dist <= -EPS || dist < min_dist)) {
                min_dist = dist;
                min_index = (uint32_t)i;
            }
        }
        if (!(min_dist > -EPS)) return R(-1);
        auto mo = static_cast<const MappedObject<R, F, G> *>(items[min_index].get());
        Cache<R, G> cc = cache;
        ctx.hasher = hash_step(orig, min_index);
        light.ray = ray_map(light.ray, *mo->inv);
        R dist = mo->inner->detect(ctx, cc, light);
        light.ray = ray_map(light.ray, *mo->map);
        cache = cc;
        cache.idx[level] = min_index;
        return dist;
    }

    R detect(Ctx &ctx, Cache<R, G> &cache, Light<R, F, G> &light) const override {
        if (g_count_culled && cullable) return detect_culled(ctx, cache, light);
        const R EPS = Consts<R>::eps();
        uint32_t orig = ctx.hasher;
        R min_dist = R(-1);
        uint32_t min_index = 0;
        Light<R, F, G> min_light{};
        Cache<R, G> min_cache{};
        uint32_t min_hasher = 0;
        for (size_t i = 0; i < items.size(); ++i) {
            Light<R, F, G> cur = light;
            Cache<R, G> cc = cache;
            ctx.hasher = hash_step(ctx.hasher, (uint32_t)i);
            R dist = items[i]->detect(ctx, cc, cur);
            if (dist > -EPS && (min_dist <= -EPS || dist < min_dist)) {
                min_dist = dist;
                min_index = (uint32_t)i;
                min_light = cur;
                min_cache = cc;
                min_hasher = ctx.hasher;
            }
            ctx.hasher = orig;
        }
        if (min_dist > -EPS) {
            cache = min_cache;
            cache.idx[level] = min_index;
            light = min_light;
            ctx.hasher = min_hasher;
            return min_dist;
        }
        return R(-1);
    }
    bool interact(Ctx &ctx, Cache<R, G> &cache, Light<R, F, G> &light, C3<F> &e) const override {
        return items[cache.idx[level]]->interact(ctx, cache, light, e);
    }
};

template <class R, class F> bool element_bound(const Object<R, F, GeoHy<R>> *o, BoundF &b) {
    auto mo = dynamic_cast<const MappedObject<R, F, GeoHy<R>> *>(o);
    if (!mo) return false;
    auto mm = dynamic_cast<const MoMap<R> *>(mo->map.get());
    if (!mm) return false;
    const Object<R, F, GeoHy<R>> *in = mo->inner.get();
    int kind = 0;
    if (auto cov = dynamic_cast<const Covered<R, F, GeoHy<R>> *>(in)) {
        const Shape<R, F, GeoHy<R>> *s = cov->shape.get();
        if (dynamic_cast<const PlaneHy<R, F> *>(s)) kind = 4;
        else if (dynamic_cast<const HorosphereHy<R, F> *>(s)) kind = 5;
    } else if (dynamic_cast<const TiledPlaneObj<R, F> *>(in)) {
        kind = 4;
    } else if (dynamic_cast<const TiledHoroObj<R, F> *>(in)) {
        kind = 5;
    } else {
        return false;
    }
    b = mo_bound(mm->m, kind);
    return true;
}

// Isometric Euclidean maps fix no origin but move it by the shift: a sphere
// of the primitive's radius (unit sphere 1, ball r, cube sqrt 3) about the
// shift, inflated by 1e-4 of the shift's L1 norm and by 1e-4 relative.
template <class R, class F> bool element_bound(const Object<R, F, GeoEu<R>> *o, BoundF &b) {
    auto mo = dynamic_cast<const MappedObject<R, F, GeoEu<R>> *>(o);
    if (!mo) return false;
    auto hm = dynamic_cast<const HomMap<R> *>(mo->map.get());
    auto cov = dynamic_cast<const Covered<R, F, GeoEu<R>> *>(mo->inner.get());
    if (!hm || !cov) return false;
    const Shape<R, F, GeoEu<R>> *s = cov->shape.get();
    double radius = -1.0;
    if (auto sp = dynamic_cast<const SphereEu<R, F> *>(s)) radius = sp->unit ? 1.0 : (double)tofloat(sp->radius);
    else if (dynamic_cast<const CubeEu<R, F> *>(s)) radius = 1.7320508075688772;
    if (radius < 0.0) {
        b = BoundF();
        return true;
    }
    double cx = hm->has_shift ? tofloat(hm->shift.x) : 0.0;
    double cy = hm->has_shift ? tofloat(hm->shift.y) : 0.0;
    double cz = hm->has_shift ? tofloat(hm->shift.z) : 0.0;
    double dist = std::fabs(cx) + std::fabs(cy) + std::fabs(cz);
    b = make_bound(cx, cy, cz, radius + 1e-4 * dist, 1e-4);
    return true;
}

// ------------------------------------------------------------ views, backgrounds
template <class R, class G> struct View {
    virtual ~View() {}
    virtual Ray<R, G> sample(Ctx &ctx, V2<R> pos, V2<R> size) const = 0;
};
// view/point/impl.inl:18-25 (jitter x then y)
template <class R, class G> struct PointView : View<R, G> {
    R fov;
    Ray<R, G> sample(Ctx &ctx, V2<R> pos, V2<R> size) const override {
        R ux = uniform<R>(ctx.rng);
        R uy = uniform<R>(ctx.rng);
        V2<R> sp = pos + size * V2<R>{ux - R(0.5f), uy - R(0.5f)};
        Ray<R, G> r;
        r.start = G::origin();
        r.dir = G::from_local(r.start, normalize(V3<R>{sp.x, sp.y, R(-1) / fov}));
        return r;
    }
};
// view/mapped.inl:20-26
template <class R, class G> struct MappedView : View<R, G> {
    std::unique_ptr<Map<R, G>> map;
    std::unique_ptr<View<R, G>> inner;
    Ray<R, G> sample(Ctx &ctx, V2<R> pos, V2<R> size) const override {
        Ray<R, G> r = inner->sample(ctx, pos, size);
        r.dir = map->apply_dir(r.start, r.dir);
        r.start = map->apply_pos(r.start);
        return r;
    }
};

template <class R, class F, class G> struct Background {
    virtual ~Background() {}
    virtual void interact(const Light<R, F, G> &light, C3<F> &e) const = 0;
};
// background/constant.inl:18-21
template <class R, class F, class G> struct ConstBg : Background<R, F, G> {
    C3<F> color;
    void interact(const Light<R, F, G> &light, C3<F> &e) const override { e = e + color * light.intensity; }
};
// background/gradient.cc:3-13
template <class R, class F> struct GradBg : Background<R, F, GeoEu<R>> {
    V3<R> dir;
    C3<F> c0, c1;
    F power;
    void interact(const Light<R, F, GeoEu<R>> &light, C3<F> &e) const override {
        F t = F(0.5f) * (F(tofloat(dot(light.ray.dir, dir))) + F(1.0f));
        F factor = fpow(fmin_(fmax_(t, F(0.0f)), F(1.0f)), power);
        C3<F> c = factor * c0 + (F(1.0f) - factor) * c1;
        e = e + c * light.intensity;
    }
};

// ------------------------------------------------------------ scene + parser
template <class R, class F, class G> struct Scene {
    int hops = 0;
    int rr_depth = -1;  // Russian roulette from this hit on (-1: none)
    std::unique_ptr<View<R, G>> view;
    std::unique_ptr<Background<R, F, G>> bg;
    std::unique_ptr<Object<R, F, G>> object;

    // render/scene.inl:16-44
    C3<F> sample(Ctx &ctx, V2<R> pos, V2<R> size) const {
        Light<R, F, G> light;
        light.intensity = {F(1.0f), F(1.0f), F(1.0f)};
        light.ray = view->sample(ctx, pos, size);
        C3<F> e = {F(0.0f), F(0.0f), F(0.0f)};
        ctx.prev_hash = 0xFFFFFFFFu;
        for (int i = 0; i < hops; ++i) {
            if (rr_depth >= 0 && i >= rr_depth) {
                // survive with q = min(max(throughput), 0.95), throughput /= q
                C3<F> I = light.intensity;
                F q = fmin_(fmax_(fmax_(I.x, I.y), I.z), F(0.95f));
                R u = uniform<R>(ctx.rng);
                if (u >= R(tofloat(q))) break;
                light.intensity = {I.x / q, I.y / q, I.z / q};
            }
            ctx.hasher = 0;
            Cache<R, G> cache{};
            R dist = object->detect(ctx, cache, light);
            if (dist > R(-0.5f)) {
                if (!object->interact(ctx, cache, light, e)) break;
                ctx.prev_hash = ctx.hasher;
            } else {
                bg->interact(light, e);
                break;
            }
        }
        return e;
    }
};

struct Tokens {
    std::vector<std::string> t;
    size_t i = 0;
    explicit Tokens(const char *s) {
        std::istringstream in(s);
        std::string w;
        while (in >> w) t.push_back(w);
    }
    const std::string &next() {
        if (i >= t.size()) throw std::runtime_error("scene spec ended early");
        return t[i++];
    }
    double num() { return std::strtod(next().c_str(), nullptr); }
    long integer() { return std::strtol(next().c_str(), nullptr, 10); }
    void expect(const char *w) {
        const std::string &s = next();
        if (s != w) throw std::runtime_error("scene spec: expected " + std::string(w) + ", got " + s);
    }
};

template <class R, class F, class G> struct Parser;
template <class R, class F> struct ShapeParse;

template <class R, class F> struct GeoParse;
template <class R, class F> struct GeoParse {
    static std::unique_ptr<Map<R, GeoEu<R>>> map_eu(Tokens &tk) {
        std::string k = tk.next();
        if (k == "shift") {
            V3<R> v = {R(tk.num()), R(tk.num()), R(tk.num())};
            return std::unique_ptr<Map<R, GeoEu<R>>>(new HomMap<R>(v, V4<R>{R(1), R(0), R(0), R(0)}, true, false));
        }
        if (k == "rot") {
            V4<R> q = {R(tk.num()), R(tk.num()), R(tk.num()), R(tk.num())};
            return std::unique_ptr<Map<R, GeoEu<R>>>(new HomMap<R>(V3<R>{R(0), R(0), R(0)}, q, false, true));
        }
        if (k == "hom") {
            V3<R> v = {R(tk.num()), R(tk.num()), R(tk.num())};
            V4<R> q = {R(tk.num()), R(tk.num()), R(tk.num()), R(tk.num())};
            return std::unique_ptr<Map<R, GeoEu<R>>>(new HomMap<R>(v, q, true, true));
        }
        throw std::runtime_error("unknown Euclidean map " + k);
    }
    static std::unique_ptr<Map<R, GeoHy<R>>> map_hy(Tokens &tk) {
        tk.expect("mo");
        C22<R> m;
        m.a = {R(tk.num()), R(tk.num())};
        m.b = {R(tk.num()), R(tk.num())};
        m.c = {R(tk.num()), R(tk.num())};
        m.d = {R(tk.num()), R(tk.num())};
        return std::unique_ptr<Map<R, GeoHy<R>>>(new MoMap<R>(m));
    }
};

template <class R, class F, class G> struct Parser {
    Tokens &tk;
    bool hy;
    explicit Parser(Tokens &t, bool h) : tk(t), hy(h) {}

    std::unique_ptr<Map<R, G>> map();

    C3<F> color() { return {F(tk.num()), F(tk.num()), F(tk.num())}; }

    std::unique_ptr<Shape<R, F, G>> shape();
    std::unique_ptr<Material<R, F>> material() {
        std::string k = tk.next();
        if (k == "lambertian") return std::unique_ptr<Material<R, F>>(new Lambertian<R, F>());
        if (k == "specular") return std::unique_ptr<Material<R, F>>(new Specular<R, F>());
        if (k == "transparent") return std::unique_ptr<Material<R, F>>(new Transparent<R, F>());
        if (k == "absorbing") return std::unique_ptr<Material<R, F>>(new Absorbing<R, F>());
        if (k == "refractive") return std::unique_ptr<Material<R, F>>(new Refractive<R, F>(R(tk.num())));
        if (k == "colored") {
            auto m = new Colored<R, F>();
            m->color = color();
            m->inner = material();
            return std::unique_ptr<Material<R, F>>(m);
        }
        if (k == "emissive") {
            auto m = new Emissive<R, F>();
            m->emission = color();
            m->inner = material();
            return std::unique_ptr<Material<R, F>>(m);
        }
        if (k == "mixture") {
            auto m = new Mixture<R, F>();
            long n = tk.integer();
            for (long i = 0; i < n; ++i) {
                m->portions.push_back(R(tk.num()));
                m->items.push_back(material());
            }
            return std::unique_ptr<Material<R, F>>(m);
        }
        throw std::runtime_error("unknown material " + k);
    }
    std::unique_ptr<Object<R, F, G>> object(int level) {
        std::string k = tk.next();
        if (k == "covered") {
            auto o = new Covered<R, F, G>();
            o->shape = shape();
            o->material = material();
            return std::unique_ptr<Object<R, F, G>>(o);
        }
        if (k == "mapped") {
            auto o = new MappedObject<R, F, G>();
            o->map = map();
            o->inv = o->map->inverse();
            o->inner = object(level);
            return std::unique_ptr<Object<R, F, G>>(o);
        }
        if (k == "tiledplane" || k == "tiledhoro") {
            auto o = ShapeParse<R, F>::tiled(*this, k);
            if (!o) throw std::runtime_error(k + " needs the hyperbolic geometry");
            return o;
        }
        if (k == "objvec") {
            auto o = new VecObject<R, F, G>();
            o->level = level;
            if (level >= 16) throw std::runtime_error("object vectors nested too deep");
            long n = tk.integer();
            for (long i = 0; i < n; ++i) o->items.push_back(object(level + 1));
            return std::unique_ptr<Object<R, F, G>>(o);
        }
        throw std::runtime_error("unknown object " + k);
    }
    std::unique_ptr<View<R, G>> view() {
        std::string k = tk.next();
        if (k == "pointview") {
            auto v = new PointView<R, G>();
            v->fov = R(tk.num());
            return std::unique_ptr<View<R, G>>(v);
        }
        if (k == "mapped") {
            auto v = new MappedView<R, G>();
            v->map = map();
            v->inner = view();
            return std::unique_ptr<View<R, G>>(v);
        }
        throw std::runtime_error("unknown view " + k);
    }
    std::unique_ptr<Background<R, F, G>> background();
};

template <class R, class F, class G> std::unique_ptr<Shape<R, F, G>> shape_common(Parser<R, F, G> &p, const std::string &k);

template <class R, class F> struct ShapeParse {
    static std::unique_ptr<Shape<R, F, GeoEu<R>>> make(Parser<R, F, GeoEu<R>> &p, const std::string &k) {
        if (k == "sphere") return std::unique_ptr<Shape<R, F, GeoEu<R>>>(new SphereEu<R, F>(R(1), true));
        if (k == "ball") return std::unique_ptr<Shape<R, F, GeoEu<R>>>(new SphereEu<R, F>(R(p.tk.num()), false));
        if (k == "cube") return std::unique_ptr<Shape<R, F, GeoEu<R>>>(new CubeEu<R, F>());
        if (k == "sdfsphere" || k == "sdfbox" || k == "sdftorus") {
            auto sh = new SdfEu<R, F>();
            sh->kind = k == "sdfsphere" ? 0 : k == "sdfbox" ? 1 : 2;
            sh->a = R(p.tk.num());
            if (sh->kind >= 1) sh->b = R(p.tk.num());
            if (sh->kind == 1) sh->c = R(p.tk.num());
            return std::unique_ptr<Shape<R, F, GeoEu<R>>>(sh);
        }
        if (k == "plane") return std::unique_ptr<Shape<R, F, GeoEu<R>>>(new PlaneEu<R, F>());
        return nullptr;
    }
    static std::unique_ptr<Shape<R, F, GeoHy<R>>> make(Parser<R, F, GeoHy<R>> &p, const std::string &k) {
        (void)p;
        if (k == "plane") return std::unique_ptr<Shape<R, F, GeoHy<R>>>(new PlaneHy<R, F>());
        if (k == "horosphere") return std::unique_ptr<Shape<R, F, GeoHy<R>>>(new HorosphereHy<R, F>());
        return nullptr;
    }
    template <class T> static std::unique_ptr<Object<R, F, GeoHy<R>>> tiled_fill(Parser<R, F, GeoHy<R>> &p, T *o) {
        o->tiling = (int)p.tk.integer();
        long n = p.tk.integer();
        o->cell = R(p.tk.num());
        o->border = R(p.tk.num());
        for (long i = 0; i < n; ++i) o->mats.push_back(p.material());
        o->border_mat = p.material();
        return std::unique_ptr<Object<R, F, GeoHy<R>>>(o);
    }
    static std::unique_ptr<Object<R, F, GeoHy<R>>> tiled(Parser<R, F, GeoHy<R>> &p, const std::string &k) {
        if (k == "tiledplane") return tiled_fill(p, new TiledPlaneObj<R, F>());
        return tiled_fill(p, new TiledHoroObj<R, F>());
    }
    static std::unique_ptr<Object<R, F, GeoEu<R>>> tiled(Parser<R, F, GeoEu<R>> &, const std::string &) {
        return nullptr;
    }
    static std::unique_ptr<Map<R, GeoEu<R>>> map(Parser<R, F, GeoEu<R>> &p) { return GeoParse<R, F>::map_eu(p.tk); }
    static std::unique_ptr<Map<R, GeoHy<R>>> map(Parser<R, F, GeoHy<R>> &p) { return GeoParse<R, F>::map_hy(p.tk); }
    static std::unique_ptr<Background<R, F, GeoEu<R>>> bg(Parser<R, F, GeoEu<R>> &p, const std::string &k) {
        if (k == "gradbg") {
            auto b = new GradBg<R, F>();
            b->dir = {R(p.tk.num()), R(p.tk.num()), R(p.tk.num())};
            b->c0 = p.color();
            b->c1 = p.color();
            b->power = F(p.tk.num());
            return std::unique_ptr<Background<R, F, GeoEu<R>>>(b);
        }
        return nullptr;
    }
    static std::unique_ptr<Background<R, F, GeoHy<R>>> bg(Parser<R, F, GeoHy<R>> &, const std::string &) {
        return nullptr;
    }
};

template <class R, class F, class G> std::unique_ptr<Map<R, G>> Parser<R, F, G>::map() {
    return ShapeParse<R, F>::map(*this);
}

template <class R, class F, class G> std::unique_ptr<Shape<R, F, G>> Parser<R, F, G>::shape() {
    std::string k = tk.next();
    if (k == "mapped") {
        auto s = new MappedShape<R, F, G>();
        s->map = map();
        s->inv = s->map->inverse();
        s->inner = shape();
        return std::unique_ptr<Shape<R, F, G>>(s);
    }
    if (k == "shapevec") {
        auto s = new VecShape<R, F, G>();
        long n = tk.integer();
        for (long i = 0; i < n; ++i) s->items.push_back(shape());
        return std::unique_ptr<Shape<R, F, G>>(s);
    }
    auto s = ShapeParse<R, F>::make(*this, k);
    if (!s) throw std::runtime_error("unknown shape " + k);
    return s;
}

template <class R, class F, class G> std::unique_ptr<Background<R, F, G>> Parser<R, F, G>::background() {
    std::string k = tk.next();
    if (k == "constbg") {
        auto b = new ConstBg<R, F, G>();
        b->color = color();
        return std::unique_ptr<Background<R, F, G>>(b);
    }
    auto b = ShapeParse<R, F>::bg(*this, k);
    if (!b) throw std::runtime_error("unknown background " + k);
    return b;
}

template <class R, class F, class G> std::unique_ptr<Scene<R, F, G>> parse_scene(Tokens &tk, int rr_depth) {
    auto s = std::unique_ptr<Scene<R, F, G>>(new Scene<R, F, G>());
    s->rr_depth = rr_depth;
    s->hops = (int)tk.integer();
    Parser<R, F, G> p(tk, std::is_same<G, GeoHy<R>>::value);
    s->view = p.view();
    s->bg = p.background();
    s->object = p.object(0);
    if (auto v = dynamic_cast<VecObject<R, F, G> *>(s->object.get())) v->prepare_cull();
    return s;
}

// render.cl:12-26 for each pixel of the crop; `passes` samples per pixel.
template <class R, class F, class G>
uint64_t render_impl(const Scene<R, F, G> &scene, uint32_t w, uint32_t h, uint32_t passes, float *canvas,
                     uint32_t *seeds, uint32_t x0, uint32_t y0, uint32_t cw, uint32_t ch, int threads,
                     uint64_t *flops64 = nullptr) {
    std::atomic<uint64_t> total{0}, total64{0};
    long n = (long)cw * (long)ch;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
#endif
    for (long k = 0; k < n; ++k) {
        g_flops = 0;
        g_flops64 = 0;
        uint32_t x = x0 + (uint32_t)(k % cw), y = y0 + (uint32_t)(k / cw);
        uint32_t idx = x + w * y;
        V2<R> pos = V2<R>{R(double(2 * x + 1)) - R(double(w)), R(double(h)) - R(double(2 * y + 1))} / R(double(h));
        R sz = R(2) / R(double(h));
        V2<R> size = {sz, sz};
        Ctx ctx;
        ctx.rng.state = seeds[idx];
        ctx.hasher = 0;
        ctx.prev_hash = 0xFFFFFFFFu;
        float *c = canvas + 4 * (size_t)idx;
        for (uint32_t p = 0; p < passes; ++p) {
            C3<F> e = scene.sample(ctx, pos, size);
            c[0] += tofloat(e.x);
            c[1] += tofloat(e.y);
            c[2] += tofloat(e.z);
            c[3] += 1.0f;
        }
        seeds[idx] = ctx.rng.state;
        total += g_flops;
        total64 += g_flops64;
    }
    if (flops64) *flops64 = total64.load();
    return total.load();
}

template <class R, class F>
uint64_t render_any(const char *spec, uint32_t w, uint32_t h, uint32_t passes, float *canvas, uint32_t *seeds,
                    uint32_t x0, uint32_t y0, uint32_t cw, uint32_t ch, int threads, uint64_t *flops64 = nullptr) {
    // optional prefix "rr <depth>": Russian roulette from that hit on
    Tokens tk(spec);
    int rr = -1;
    std::string first = tk.next();
    if (first == "rr") {
        rr = (int)tk.integer();
        first = tk.next();
    }
    if (first != "scene") throw std::runtime_error("spec must start with 'scene'");
    std::string geo = tk.next();
    if (geo == "eu") {
        auto s = parse_scene<R, F, GeoEu<R>>(tk, rr);
        return render_impl(*s, w, h, passes, canvas, seeds, x0, y0, cw, ch, threads, flops64);
    }
    if (geo == "hy") {
        auto s = parse_scene<R, F, GeoHy<R>>(tk, rr);
        return render_impl(*s, w, h, passes, canvas, seeds, x0, y0, cw, ch, threads, flops64);
    }
    throw std::runtime_error("unknown geometry " + geo);
}

thread_local std::string g_err;

}  // namespace orc

using namespace orc;

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

// Renders `passes` passes over the crop [x0, x0+cw) x [y0, y0+ch) of a w x h
// image, accumulating into canvas (float4 per pixel) and advancing seeds.
// precision: 0 = FP32, 1 = FP64 geometry with FP32 colours, 2 = FP32 with
// flop counting (the count is returned in *flops).
int oracle_render(const char *spec, int precision, uint32_t w, uint32_t h, uint32_t passes, float *canvas,
                  uint32_t *seeds, uint32_t x0, uint32_t y0, uint32_t cw, uint32_t ch, int threads,
                  uint64_t *flops) {
    try {
        if (x0 + cw > w || y0 + ch > h) throw std::runtime_error("crop outside the image");
        uint64_t f = 0;
        if (precision == 0) {
            render_any<float, float>(spec, w, h, passes, canvas, seeds, x0, y0, cw, ch, threads);
        } else if (precision == 1) {
            render_any<double, float>(spec, w, h, passes, canvas, seeds, x0, y0, cw, ch, threads);
        } else if (precision == 2) {
            f = render_any<Cnt, Cnt>(spec, w, h, passes, canvas, seeds, x0, y0, cw, ch, threads);
        } else {
            throw std::runtime_error("unknown precision");
        }
        if (flops) *flops = f;
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return -1;
    }
}

// Flop count of the FP32 render over a crop: *flops32 = FP32 flops, *flops64
// = FP64 flops. culled = 0: the reference algorithm (every element's exact
// test, each mapped back); culled = 1: the culled scan's algorithm (FP64
// cull tests, exact tests of candidates, one map back per hit); culled = g
// (2..16): the two-level scan over groups of g (group-sphere tests, then
// the leaf tests of the passing groups); culled = 1000 + g: the same with
// the group queue (the ray's cull state rebuilt per passing group); culled
// = 2000 + g: the group queue with the hit order (each passing pair's
// crossing interval; exact tests only of the pairs its filter keeps).
int oracle_count_flops(const char *spec, int culled, uint32_t w, uint32_t h, uint32_t passes, float *canvas,
                       uint32_t *seeds, uint32_t x0, uint32_t y0, uint32_t cw, uint32_t ch, int threads,
                       uint64_t *flops32, uint64_t *flops64) {
    try {
        if (x0 + cw > w || y0 + ch > h) throw std::runtime_error("crop outside the image");
        g_count_culled = culled != 0;
        g_count_queue = culled >= 1000;
        g_count_hit_order = culled >= 2000;
        g_count_group = culled >= 2 ? culled % 1000 : 0;
        uint64_t f64 = 0;
        uint64_t f32 = render_any<Cnt, Cnt>(spec, w, h, passes, canvas, seeds, x0, y0, cw, ch, threads, &f64);
        g_count_culled = false;
        g_count_group = 0;
        g_count_queue = false;
        g_count_hit_order = false;
        if (flops32) *flops32 = f32;
        if (flops64) *flops64 = f64;
        return 0;
    } catch (const std::exception &e) {
        g_count_culled = false;
        g_count_group = 0;
        g_count_queue = false;
        g_count_hit_order = false;
        g_err = e.what();
        return -1;
    }
}

// ---- leaf functions, for pinning against closed forms and known answers.

uint32_t oracle_lcg_next(uint32_t *state) {
    Rng r{*state};
    uint32_t v = r.next();
    *state = r.state;
    return v;
}
float oracle_uniform_f32(uint32_t *state) {
    Rng r{*state};
    float u = uniform<float>(r);
    *state = r.state;
    return u;
}
void oracle_hemisphere_cosine_f64(uint32_t *state, double *out3) {
    Rng r{*state};
    V3<double> v = hemisphere_cosine<double>(r);
    *state = r.state;
    out3[0] = v.x;
    out3[1] = v.y;
    out3[2] = v.z;
}
uint32_t oracle_hash(const uint32_t *xs, int n) {
    uint32_t h = 0;
    for (int i = 0; i < n; ++i) h = hash_step(h, xs[i]);
    return h;
}
// quaternions as (x = real, y, z, w)
void oracle_q_mul(const double *a, const double *b, double *out) {
    V4<double> r = q_mul(V4<double>{a[0], a[1], a[2], a[3]}, V4<double>{b[0], b[1], b[2], b[3]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
void oracle_q_div(const double *a, const double *b, double *out) {
    V4<double> r = q_div(V4<double>{a[0], a[1], a[2], a[3]}, V4<double>{b[0], b[1], b[2], b[3]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
static C22<double> c22_of(const double *m) {
    return {{m[0], m[1]}, {m[2], m[3]}, {m[4], m[5]}, {m[6], m[7]}};
}
void oracle_c22_dot(const double *a, const double *b, double *out) {
    C22<double> r = c22_dot(c22_of(a), c22_of(b));
    double v[8] = {r.a.x, r.a.y, r.b.x, r.b.y, r.c.x, r.c.y, r.d.x, r.d.y};
    std::memcpy(out, v, sizeof v);
}
void oracle_mo_apply(const double *m, const double *p, double *out) {
    V4<double> r = MoMap<double>(c22_of(m)).apply_pos(V4<double>{p[0], p[1], p[2], p[3]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
void oracle_mo_deriv(const double *m, const double *p, const double *v, double *out) {
    V4<double> r = MoMap<double>(c22_of(m)).deriv(V4<double>{p[0], p[1], p[2], p[3]}, V4<double>{v[0], v[1], v[2], v[3]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
double oracle_hy_distance(const double *a, const double *b) {
    return hy_distance(V4<double>{a[0], a[1], a[2], a[3]}, V4<double>{b[0], b[1], b[2], b[3]});
}
void oracle_look_at(const double *d, double *q) {
    V4<double> r = look_at(V3<double>{d[0], d[1], d[2]});
    q[0] = r.x; q[1] = r.y; q[2] = r.z; q[3] = r.w;
}
void oracle_rot_apply(const double *q, const double *p, double *out) {
    V3<double> r = rot_apply(V4<double>{q[0], q[1], q[2], q[3]}, V3<double>{p[0], p[1], p[2]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// One shape test in FP64 on a ray (start, dir as 4 values; Euclidean shapes
// use the first 3). kind: 0 sphere, 1 cube, 2 eu plane, 3 hy plane,
// 4 horosphere, 5 ball (param = radius). Returns the distance (negative =
// miss); writes the new start, direction and normal.
double oracle_shape_detect(int kind, double param, int repeat, const double *start, const double *dir,
                           double *out_start, double *out_dir, double *out_normal) {
    Ctx ctx{};
    ctx.hasher = 7;
    ctx.prev_hash = repeat ? 7 : 8;
    if (kind == 3 || kind == 4) {
        Light<double, float, GeoHy<double>> l{};
        l.ray.start = {start[0], start[1], start[2], start[3]};
        l.ray.dir = {dir[0], dir[1], dir[2], dir[3]};
        V4<double> n{};
        double d = kind == 3 ? PlaneHy<double, float>().detect(ctx, n, l) : HorosphereHy<double, float>().detect(ctx, n, l);
        double s[4] = {l.ray.start.x, l.ray.start.y, l.ray.start.z, l.ray.start.w};
        double v[4] = {l.ray.dir.x, l.ray.dir.y, l.ray.dir.z, l.ray.dir.w};
        double nn[4] = {n.x, n.y, n.z, n.w};
        std::memcpy(out_start, s, sizeof s);
        std::memcpy(out_dir, v, sizeof v);
        std::memcpy(out_normal, nn, sizeof nn);
        return d;
    }
    Light<double, float, GeoEu<double>> l{};
    l.ray.start = {start[0], start[1], start[2]};
    l.ray.dir = {dir[0], dir[1], dir[2]};
    V3<double> n{};
    double d;
    if (kind == 0) d = SphereEu<double, float>(1.0, true).detect(ctx, n, l);
    else if (kind == 5) d = SphereEu<double, float>(param, false).detect(ctx, n, l);
    else if (kind == 1) d = CubeEu<double, float>().detect(ctx, n, l);
    else d = PlaneEu<double, float>().detect(ctx, n, l);
    double s[4] = {l.ray.start.x, l.ray.start.y, l.ray.start.z, 0};
    double v[4] = {l.ray.dir.x, l.ray.dir.y, l.ray.dir.z, 0};
    double nn[4] = {n.x, n.y, n.z, 0};
    std::memcpy(out_start, s, sizeof s);
    std::memcpy(out_dir, v, sizeof v);
    std::memcpy(out_normal, nn, sizeof nn);
    return d;
}

// Sphere-traced shape in FP64: kind 0 sphere (p0), 1 box (p0, p1, p2), 2 torus (p0, p1).
double oracle_sdf_detect(int kind, const double *params, int repeat, const double *start, const double *dir,
                         double *out_start, double *out_normal) {
    Ctx ctx{};
    ctx.hasher = 7;
    ctx.prev_hash = repeat ? 7 : 8;
    SdfEu<double, float> sh;
    sh.kind = kind;
    sh.a = params[0];
    sh.b = params[1];
    sh.c = params[2];
    Light<double, float, GeoEu<double>> l{};
    l.ray.start = {start[0], start[1], start[2]};
    l.ray.dir = {dir[0], dir[1], dir[2]};
    V3<double> n{};
    double d = sh.detect(ctx, n, l);
    out_start[0] = l.ray.start.x;
    out_start[1] = l.ray.start.y;
    out_start[2] = l.ray.start.z;
    out_normal[0] = n.x;
    out_normal[1] = n.y;
    out_normal[2] = n.z;
    return d;
}

// One material interaction in FP64 geometry with FP32 colours (the
// reference's DOUBLE_PRECISION build): `spec` is a material in the scene
// spec's grammar ("specular", "mixture 2 0.3 absorbing 0.7 lambertian", ...);
// the light arrives along `dir` with `intensity` at a surface of local
// normal `normal`, emission starts at `emission_in`. Writes the outgoing
// direction, intensity and emission, advances *rng_state by the draws the
// material takes, and returns 1 if the path goes on, 0 if it ends, -1 on a
// parse error.
int oracle_material_interact(const char *spec, const double *normal, const double *dir, const float *intensity,
                             const float *emission_in, uint32_t *rng_state, double *out_dir, float *out_intensity,
                             float *out_emission) {
    try {
        Tokens tk(spec);
        Parser<double, float, GeoEu<double>> p(tk, false);
        auto m = p.material();
        Ctx ctx{};
        ctx.rng.state = *rng_state;
        LightLocal<double, float> ll;
        ll.intensity = {intensity[0], intensity[1], intensity[2]};
        ll.dir = {dir[0], dir[1], dir[2]};
        C3<float> e = {emission_in[0], emission_in[1], emission_in[2]};
        bool ret = m->interact(ctx, V3<double>{normal[0], normal[1], normal[2]}, ll, e);
        *rng_state = ctx.rng.state;
        out_dir[0] = ll.dir.x;
        out_dir[1] = ll.dir.y;
        out_dir[2] = ll.dir.z;
        out_intensity[0] = ll.intensity.x;
        out_intensity[1] = ll.intensity.y;
        out_intensity[2] = ll.intensity.z;
        out_emission[0] = e.x;
        out_emission[1] = e.y;
        out_emission[2] = e.z;
        return ret ? 1 : 0;
    } catch (const std::exception &ex) {
        g_err = ex.what();
        return -1;
    }
}

// Emission a Euclidean background adds for a path escaping along `dir` with
// `intensity` (FP64 geometry, FP32 colours); `spec` in the scene spec's
// grammar ("gradbg dx dy dz r0 g0 b0 r1 g1 b1 power", "constbg r g b").
int oracle_background(const char *spec, const double *dir, const float *intensity, float *out_emission) {
    try {
        Tokens tk(spec);
        Parser<double, float, GeoEu<double>> p(tk, false);
        auto bg = p.background();
        Light<double, float, GeoEu<double>> l{};
        l.ray.dir = {dir[0], dir[1], dir[2]};
        l.intensity = {intensity[0], intensity[1], intensity[2]};
        C3<float> e = {0.0f, 0.0f, 0.0f};
        bg->interact(l, e);
        out_emission[0] = e.x;
        out_emission[1] = e.y;
        out_emission[2] = e.z;
        return 0;
    } catch (const std::exception &ex) {
        g_err = ex.what();
        return -1;
    }
}

// Material a tiled object picks at the point p of its surface (FP64):
// object 0 = tiled plane (tiling 1 pentagonal, 2 pentastar), 1 = tiled
// horosphere (tiling 1 square, 2 hexagonal); returns the material's index
// 0..n_materials-1, -1 for the border material, -2 on a bad argument.
int oracle_tiled_select(int object, int tiling, double cell, double border, int n_materials, const double *p) {
    if (n_materials < 1 || n_materials > 64) return -2;
    auto fill = [&](auto *o) {
        o->tiling = tiling;
        o->cell = cell;
        o->border = border;
        for (int i = 0; i < n_materials; ++i) o->mats.emplace_back(new Transparent<double, float>());
        o->border_mat.reset(new Transparent<double, float>());
        const Material<double, float> *m = o->select(V4<double>{p[0], p[1], p[2], p[3]});
        if (m == o->border_mat.get()) return -1;
        for (int i = 0; i < n_materials; ++i)
            if (m == o->mats[i].get()) return i;
        return -2;
    };
    if (object == 0) {
        TiledPlaneObj<double, float> o;
        return fill(&o);
    }
    if (object == 1) {
        TiledHoroObj<double, float> o;
        return fill(&o);
    }
    return -2;
}

// Post passes on the CPU: extract -> gamma -> pack (processing/src/canvas.rs:126-135,
// filter/gamma.rs:10-19, canvas.rs:176-186).
void oracle_post(const float *canvas, size_t n, uint32_t passes, float gamma, float *image, uint8_t *rgba8) {
    for (size_t i = 0; i < 4 * n; ++i) {
        float e = canvas[i] / (float)passes;
        float g = std::pow(std::fmax(e, 0.0f), gamma);
        if (image) image[i] = g;
        rgba8[i] = (uint8_t)(255.0f * std::fmin(std::fmax(g, 0.0f), 1.0f));
    }
}

}  // extern "C"
