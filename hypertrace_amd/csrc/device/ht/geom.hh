// Algebra (complex, quaternion, 2x2 complex), isometries (Shift3, Rotation3,
// Homogenous3, Moebius), Euclidean and hyperbolic geometry, rays and lights.
//
// Each function follows the reference's formula and operation order so that
// FP32 results stay within rounding of the reference kernel:
//   complex      kernel/ocl/algebra/complex.cc:4-30
//   quaternion   kernel/ocl/algebra/quaternion.cc:4-56
//   comp2x2      kernel/ocl/algebra/matrix_complex.cc:38-65
//   Moebius      kernel/ocl/transform/moebius.cc:3-60
//   Shift3       kernel/ocl/transform/shift.cc:34-59
//   Rotation3    kernel/ocl/transform/rotation.cc:51-127
//   Homogenous3  kernel/ocl/transform/chain.inl:28-73 with homogenous.hh
//   geometry     kernel/ocl/geometry/euclidean.cc, hyperbolic.cc
//   rays         kernel/ocl/geometry/ray.inl, ray_map.inl
//   lights       kernel/ocl/render/light/{base.hh,eu.cc,hy.cc}
#pragma once
#include "base.hh"

namespace ht {

// ---------------------------------------------------------------- complex
typedef real2 comp;
HD comp c_conj(comp a) { return {a.x, -a.y}; }
HD real c_abs2(comp a) { return dot(a, a); }
HD comp c_mul(comp a, comp b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
HD comp c_inverse(comp a) { return c_conj(a) / c_abs2(a); }
HD comp c_div(comp a, comp b) { return c_mul(a, c_inverse(b)); }

// ---------------------------------------------------------------- quaternion
typedef real4 quat;
HD quat q_new(comp a, comp b) { return {a.x, a.y, b.x, b.y}; }
HD quat q_new3(real3 v, real w) { return {v.x, v.y, v.z, w}; }
HD quat q_conj(quat a) { return {a.x, -a.y, -a.z, -a.w}; }
HD real q_abs2(quat a) { return dot(a, a); }
HD quat q_mul(quat a, quat b) {
    return {a.x * b.x - a.y * b.y - a.z * b.z - a.w * b.w,
            a.x * b.y + a.y * b.x + a.z * b.w - a.w * b.z,
            a.x * b.z + a.z * b.x - a.y * b.w + a.w * b.y,
            a.x * b.w + a.w * b.x + a.y * b.z - a.z * b.y};
}
HD quat cq_mul(comp a, quat b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x, a.x * b.z - a.y * b.w, a.x * b.w + a.y * b.z};
}
HD quat q_inverse(quat a) { return q_conj(a) / q_abs2(a); }
HD quat q_div(quat a, quat b) { return q_mul(a, q_inverse(b)); }

// ---------------------------------------------------------------- comp2x2
// Stored as real8 {a, b, c, d} = (s01, s23, s45, s67).
struct c22 { comp a, b, c, d; };
HD c22 ld_c22(const u8 *p) {
    const real *r = reinterpret_cast<const real *>(p);
    return {{r[0], r[1]}, {r[2], r[3]}, {r[4], r[5]}, {r[6], r[7]}};
}
HD c22 c22_dot(c22 x, c22 y) {
    return {c_mul(x.a, y.a) + c_mul(x.b, y.c), c_mul(x.a, y.b) + c_mul(x.b, y.d),
            c_mul(x.c, y.a) + c_mul(x.d, y.c), c_mul(x.c, y.b) + c_mul(x.d, y.d)};
}
HD c22 c22_inverse_n(c22 m) { return {m.d, -m.b, -m.c, m.a}; }

// ---------------------------------------------------------------- maps
// A map type M provides: LAYOUT (Leaf/Fields), load(p), inverse(m),
// apply_pos(m, p), apply_dir(m, p, d), apply_normal(m, p, d), chain(a, b).

// Shift3 {real3 v}
struct Shift3 {
    typedef LReal3 L;
    static constexpr size_t_ SIZE = L::SIZE, ALIGN = L::ALIGN;
    static constexpr bool DYN = false;
    real3 v;
    static HD Shift3 load(const u8 *p) { return {ld_real3(p)}; }
    static HD Shift3 identity() { return {r3(R0, R0, R0)}; }
    HD real3 apply_pos(real3 p) const { return p + v; }
    HD real3 apply_dir(real3, real3 d) const { return d; }
    HD real3 apply_normal(real3, real3 d) const { return d; }
    HD Shift3 inverse() const { return {-v}; }
    static HD Shift3 chain(Shift3 a, Shift3 b) { return {a.v + b.v}; }
};

// Rotation3 {quat v}: p -> (v * (0,p) * conj(v)).yzw
struct Rotation3 {
    typedef LReal4 L;
    static constexpr size_t_ SIZE = L::SIZE, ALIGN = L::ALIGN;
    static constexpr bool DYN = false;
    quat v;
    static HD Rotation3 load(const u8 *p) { return {ld_real4(p)}; }
    static HD Rotation3 identity() { return {r4(R1, R0, R0, R0)}; }
    HD real3 apply_pos(real3 p) const {
        quat r = q_mul(q_mul(v, r4(R0, p.x, p.y, p.z)), q_conj(v));
        return {r.y, r.z, r.w};
    }
    HD real3 apply_dir(real3, real3 d) const { return apply_pos(d); }
    HD real3 apply_normal(real3 p, real3 d) const { return apply_dir(p, d); }
    HD Rotation3 inverse() const { return {q_conj(v)}; }
    static HD Rotation3 chain(Rotation3 a, Rotation3 b) { return {q_mul(a.v, b.v)}; }
    static HD Rotation3 from_axis(real3 axis, real angle) {
        real s = sin_(angle / 2);
        return {r4(cos_(angle / 2), s * axis.x, s * axis.y, s * axis.z)};
    }
    // rotation.cc:103-120
    static HD Rotation3 look_at_cont(real3 d) {
        real3 z = r3(R0, R0, -R1);
        real c2 = dot(d, z);
        real c = sqrt_((R1 + c2) / R2);
        real sm = R1 / (R2 * c);
        real3 w = sm * cross(z, d);
        return {r4(c, w.x, w.y, w.z)};
    }
    static HD Rotation3 look_at(real3 d) {
        if (d.z < R0) {
            return look_at_cont(d);
        } else {
            return chain(look_at_cont(-d), from_axis(r3(R1, R0, R0), PI));
        }
    }
};

// Homogenous3 = chain{outer: Shift3, inner: Rotation3}
struct Homogenous3 {
    typedef Fields<Shift3, Rotation3> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;
    Shift3 outer;
    Rotation3 inner;
    static HD Homogenous3 load(const u8 *p) {
        return {Shift3::load(p + L::offset(0)), Rotation3::load(p + L::offset(1))};
    }
    HD real3 apply_pos(real3 p) const { return outer.apply_pos(inner.apply_pos(p)); }
    HD real3 apply_dir(real3 p, real3 d) const {
        return outer.apply_dir(inner.apply_pos(p), inner.apply_dir(p, d));
    }
    HD real3 apply_normal(real3 p, real3 d) const {
        return outer.apply_normal(inner.apply_pos(p), inner.apply_normal(p, d));
    }
    // chain.inl:65-73 with rot3_shf3_reorder (rotation.cc:122-124)
    HD Homogenous3 inverse() const {
        Homogenous3 s = {outer.inverse(), inner.inverse()};
        s.outer.v = s.inner.apply_pos(s.outer.v);
        return s;
    }
};

// Moebius {comp2x2 v} acting on the quaternion upper half-space.
struct Moebius {
    typedef LComp2x2 L;
    static constexpr size_t_ SIZE = L::SIZE, ALIGN = L::ALIGN;
    static constexpr bool DYN = false;
    c22 v;
    static HD Moebius load(const u8 *p) { return {ld_c22(p)}; }
    static HD Moebius identity() {
        return {{r2(R1, R0), r2(R0, R0), r2(R0, R0), r2(R1, R0)}};
    }
    // moebius.cc:20-25
    HD quat apply_pos(quat p) const {
        return q_div(cq_mul(v.a, p) + q_new(v.b, r2(R0, R0)), cq_mul(v.c, p) + q_new(v.d, r2(R0, R0)));
    }
    // moebius.cc:33-42
    HD quat deriv(quat p, quat w) const {
        quat u = cq_mul(v.a, p) + q_new(v.b, r2(R0, R0));
        quat d = cq_mul(v.c, p) + q_new(v.d, r2(R0, R0));
        real d2 = q_abs2(d);
        quat g1 = q_div(cq_mul(v.a, w), d);
        quat cw = cq_mul(v.c, w);
        quat g21 = q_conj(cw);
        quat g22 = (2 * dot(d, cw) / d2) * q_conj(d);
        quat g2 = q_mul(u, (g21 - g22) / d2);
        return g1 + g2;
    }
    HD quat apply_dir(quat p, quat w) const { return normalize(deriv(p, w)); }
    HD quat apply_normal(quat p, quat w) const { return apply_dir(p, w); }
    // moebius.cc:54-56: adjugate, assumes det = 1
    HD Moebius inverse() const { return {c22_inverse_n(v)}; }
    static HD Moebius chain(Moebius a, Moebius b) { return {c22_dot(a.v, b.v)}; }
};

// ---------------------------------------------------------------- geometry
struct LightBase {
    color3 intensity;
    bool diffuse;
};
struct LightLocal {
    LightBase base;
    real3 direction;
};

// Euclidean: Pos = Dir = real3, Map = Homogenous3 (euclidean.cc:5-81)
struct Eu {
    typedef real3 Pos;
    typedef real3 Dir;
    typedef LReal3 LDir;
    static HD Pos origin() { return r3(R0, R0, R0); }
    static HD real3 dir_to_local(Pos, Dir d) { return d; }
    static HD Dir dir_from_local(Pos, real3 d) { return d; }
    static HD Dir dir_at(Pos, Dir d, Pos) { return d; }
};

// Hyperbolic upper half-space: Pos = Dir = quat with w = 0 (hyperbolic.cc)
struct Hy {
    typedef quat Pos;
    typedef quat Dir;
    typedef LReal4 LDir;
    static HD Pos origin() { return r4(R0, R0, R1, R0); }
    static HD real3 dir_to_local(Pos, Dir d) { return xyz(d); }
    static HD Dir dir_from_local(Pos, real3 d) { return q_new3(d, R0); }
    // hyperbolic.cc:20-23. Program option HT_HY_DIST_FP64 (a precision
    // method, not the drop-in: results differ from the reference in the last
    // bits): the same distance in FP64 through the cancellation-free form
    // acosh(1 + u) = log1p(u + sqrt(u (u + 2))), u = |a - b|^2 / (2 a.z b.z),
    // which keeps its relative accuracy both for nearby points (u -> 0,
    // where x * x - 1 cancels) and near the boundary (a.z b.z -> 0).
    static HD real distance(quat a, quat b) {
#ifdef HT_HY_DIST_FP64
        double dx = (double)a.x - (double)b.x, dy = (double)a.y - (double)b.y, dz = (double)a.z - (double)b.z;
        double u = (dx * dx + dy * dy + dz * dz) / (2.0 * (double)a.z * (double)b.z);
        return (real)::log1p(u + ::sqrt(u * (u + 2.0)));
#else
        real x = 1 + length2(a - b) / (2 * a.z * b.z);
        return log_(x + sqrt_(x * x - 1));
#endif
    }
    // hyperbolic.cc:25-33
    static HD Dir dir_at(quat p, quat d, quat h) {
        return {h.z / p.z * d.x, h.z / p.z * d.y, d.z - length(xy(p) - xy(h)) / p.z * length(xy(d)), R0};
    }
};

template <class G> struct Ray {
    typename G::Pos start;
    typename G::Dir direction;
};

// ray_map.inl:15-21: the direction is renormalised after the map.
template <class G, class M> HD Ray<G> ray_map(Ray<G> ray, const M &m) {
    Ray<G> r;
    r.start = m.apply_pos(ray.start);
    r.direction = normalize(m.apply_dir(ray.start, ray.direction));
    return r;
}

template <class G> struct Light {
    LightBase base;
    Ray<G> ray;
};

// light/eu.cc, light/hy.cc
HD LightLocal light_to_local(const Light<Eu> *l) { return {l->base, l->ray.direction}; }
HD void light_update_local(Light<Eu> *l, LightLocal ll) {
    l->ray.direction = ll.direction;
    l->base = ll.base;
}
HD LightLocal light_to_local(const Light<Hy> *l) { return {l->base, xyz(l->ray.direction)}; }
HD void light_update_local(Light<Hy> *l, LightLocal ll) {
    l->ray.direction = q_new3(ll.direction, R0);
    l->base = ll.base;
}

}  // namespace ht
