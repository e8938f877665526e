// Sphere-traced signed-distance shapes (new: the reference's shape set is
// analytic only, shape/eu/*.cc). In the shape's local frame:
//   SdfSphere {r}          |x| - r
//   SdfBox    {h (real3)}  |max(q, 0)| + min(max(q.x, q.y, q.z), 0),  q = |x| - h
//   SdfTorus  {R, r}       |(|x.xz| - R, x.y)| - r        (around the y axis)
// placed by isometric maps (Mapped, Homogenous3), so local distances are world
// distances. The march runs inside the shape's bounding sphere (radius + 2 eps):
// from its entry (or the ray start), step |f| until |f| < eps (hit) or the
// exit; at most SDF_STEPS evaluations. Marching on |f| also finds the exit of
// a ray inside the shape (glass). A ray leaving the primitive it just hit
// (the reference's repeat rule, render/context.cc:3-13) first steps
// max(|f|, eps) until |f| >= 2 eps, so it does not re-hit its start point.
// Normal: the tetrahedron estimate of the gradient (4 evaluations, h = 4 eps),
// normalised; outward.
#pragma once
#include "entities.hh"

namespace ht {

static constexpr int SDF_STEPS = 128;
#ifdef HT_DOUBLE
static constexpr real SDF_EPS = 1e-4;
#else
static constexpr real SDF_EPS = 1e-4f;
#endif

HD real sdf_sphere(real3 x, real r) { return length(x) - r; }
HD real sdf_box(real3 x, real3 h) {
    real3 q = r3(fabs_(x.x) - h.x, fabs_(x.y) - h.y, fabs_(x.z) - h.z);
    real3 qp = r3(fmax_(q.x, R0), fmax_(q.y, R0), fmax_(q.z, R0));
    return length(qp) + fmin_(fmax_(q.x, fmax_(q.y, q.z)), R0);
}
HD real sdf_torus(real3 x, real big, real small) {
    real qx = sqrt_(x.x * x.x + x.z * x.z) - big;
    return sqrt_(qx * qx + x.y * x.y) - small;
}

template <class F>
HD real sdf_trace(const F &f, real bound, Context *context, real3 *normal, Light<Eu> *light) {
    bool repeat = context_is_repeat(context);
    real3 p = light->ray.start, d = light->ray.direction;
    real rb = bound + 2 * SDF_EPS;
    real b = -dot(d, p);
    real3 g = p + d * b;
    real disc = rb * rb - length2(g);
    if (disc < R0) return -R1;
    real s = sqrt_(disc);
    real t1 = b + s;
    if (t1 < R0) return -R1;
    real t = fmax_(b - s, R0);
    bool leaving = repeat;
    bool hit = false;
    for (int k = 0; k < SDF_STEPS; ++k) {
        real a = fabs_(f(p + d * t));
        if (leaving) {
            if (a >= 2 * SDF_EPS) {
                leaving = false;
            } else {
                t += fmax_(a, SDF_EPS);
                if (t > t1) break;
                continue;
            }
        }
        if (a < SDF_EPS) {
            hit = true;
            break;
        }
        t += a;
        if (t > t1) break;
    }
    if (!hit) return -R1;
    real3 x = p + d * t;
    real h = 4 * SDF_EPS;
    real f0 = f(x + r3(h, -h, -h)), f1 = f(x + r3(-h, -h, h)), f2 = f(x + r3(-h, h, -h)), f3 = f(x + r3(h, h, h));
    real3 n = r3(f0 - f1 - f2 + f3, -f0 - f1 + f2 + f3, -f0 + f1 - f2 + f3);
    *normal = normalize(n);
    light->ray.start = x;
    return t;
}

struct SdfSphere {
    typedef Fields<LReal> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;
    static HD real bound(const u8 *self) { return ld<real>(self); }
    static HD real shape_detect(const u8 *self, Context *context, real3 *normal, Light<Eu> *light) {
        real r = ld<real>(self);
        return sdf_trace([r](real3 x) { return sdf_sphere(x, r); }, r, context, normal, light);
    }
};

struct SdfBox {
    typedef Fields<LReal3> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;
    static HD real bound(const u8 *self) { return length(ld_real3(self)); }
    static HD real shape_detect(const u8 *self, Context *context, real3 *normal, Light<Eu> *light) {
        real3 h = ld_real3(self);
        return sdf_trace([h](real3 x) { return sdf_box(x, h); }, length(h), context, normal, light);
    }
};

struct SdfTorus {
    typedef Fields<LReal, LReal> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;
    static HD real bound(const u8 *self) { return ld<real>(self + L::offset(0)) + ld<real>(self + L::offset(1)); }
    static HD real shape_detect(const u8 *self, Context *context, real3 *normal, Light<Eu> *light) {
        real big = ld<real>(self + L::offset(0)), small = ld<real>(self + L::offset(1));
        return sdf_trace([big, small](real3 x) { return sdf_torus(x, big, small); }, big + small, context, normal,
                         light);
    }
};

}  // namespace ht
