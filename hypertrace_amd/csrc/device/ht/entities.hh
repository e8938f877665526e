// Scene entity vocabulary on the device: shapes, materials, objects, views,
// backgrounds and the scene itself, as C++ templates over the scene byte
// stream. Every entity type T carries its byte layout (T::SIZE, T::ALIGN,
// T::DYN; for dynamically sized types SIZE is the minimum size) and static
// device functions that take `const u8 *self`, the address of its bytes.
// Template names mirror the reference's Rust entity types (objects/src/**), so
// that a scene type written in Rust maps one-to-one to a type expression here.
//
// Device interfaces replaced (same meaning and argument order):
//   shape      real shape_detect(self, ctx, Dir *normal, Light *)   shape/interface.hh:6-15
//   object     real detect(self, ctx, Cache *, Light *)              object/interface.hh:4-22
//              bool interact(self, ctx, Cache *, Light *, color3 *)
//   material   bool interact(self, ctx, real3 normal, LightLocal *, color3 *)  material/interface.hh:5-12
//   background void bg_interact(self, ctx, const Light *, color3 *)  background/interface.hh:7-16
//   view       Ray sample(self, ctx, real2 pos, real2 size)           view/interface.hh:5-14
#pragma once
#include "geom.hh"

namespace ht {

// ================================================================ shapes

// Unit sphere at the origin, shape/eu/sphere.cc:3-29.
struct Sphere {
    static constexpr size_t_ SIZE = 0, ALIGN = 1;
    static constexpr bool DYN = false;
    static HD real shape_detect(const u8 *, Context *context, real3 *normal, Light<Eu> *light) {
        bool repeat = context_is_repeat(context);
        real3 pos = light->ray.start;
        real3 dir = light->ray.direction;
        real b = -dot(dir, pos);
        real3 g = pos + dir * b;
        real d = R1 - length2(g);
        if (d < R0) return -R1;
        d = sqrt_(d);
        real e = b - d;
        real w = EPS * (2 * (int)repeat - 1);
        if (e < w) {
            e = b + d;
            if (e < w) return -R1;
        }
        real3 h = normalize(pos + dir * e);
        *normal = h;
        light->ray.start = h;
        return e;
    }
};

// Sphere of radius r at the origin. Not in the reference (whose only sphere is
// the unit sphere under isometric maps); it keeps world-space distances exact
// for scenes with varied radii. Same EPS/repeat rule as Sphere.
struct Ball {
    typedef Fields<LReal> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;
    static HD real shape_detect(const u8 *self, Context *context, real3 *normal, Light<Eu> *light) {
        real r = ld<real>(self);
        bool repeat = context_is_repeat(context);
        real3 pos = light->ray.start;
        real3 dir = light->ray.direction;
        real b = -dot(dir, pos);
        real3 g = pos + dir * b;
        real d = r * r - length2(g);
        if (d < R0) return -R1;
        d = sqrt_(d);
        real e = b - d;
        real w = EPS * (2 * (int)repeat - 1);
        if (e < w) {
            e = b + d;
            if (e < w) return -R1;
        }
        real3 h = pos + dir * e;
        *normal = normalize(h);
        light->ray.start = h;
        return e;
    }
};

// Unit cube [-1,1]^3, shape/eu/cube.cc:4-65.
struct Cube {
    static constexpr size_t_ SIZE = 0, ALIGN = 1;
    static constexpr bool DYN = false;
    static HD real nearest(real3 n, real3 *normal) {
        bool xy = n.x > n.y;
        bool yz = n.y > n.z;
        bool xz = n.x > n.z;
        real dist;
        if (xy && xz) {
            dist = n.x;
            normal->x = R1;
        } else if (yz) {
            dist = n.y;
            normal->y = R1;
        } else {
            dist = n.z;
            normal->z = R1;
        }
        return dist;
    }
    static HD real shape_detect(const u8 *, Context *context, real3 *normal, Light<Eu> *light) {
        bool repeat = context_is_repeat(context);
        Ray<Eu> *ray = &light->ray;
        real3 inv_dir = r3(R1 / ray->direction.x, R1 / ray->direction.y, R1 / ray->direction.z);
        real3 vmin = (r3(-R1, -R1, -R1) - ray->start) * inv_dir;
        real3 vmax = (r3(R1, R1, R1) - ray->start) * inv_dir;
        real3 nr = fmin_(vmin, vmax);
        real3 fr = fmax_(vmin, vmax);
        real3 norm_in = r3(R0, R0, R0);
        float dist_in = (float)nearest(nr, &norm_in);
        norm_in *= -sign_(ray->direction);
        real3 norm_out = r3(R0, R0, R0);
        float dist_out = -(float)nearest(-fr, &norm_out);
        norm_out *= sign_(ray->direction);
        if (dist_in > dist_out) return -R1;
        real dist = dist_in;
        real3 norm = norm_in;
        real w = 2 * EPS * (2 * (int)repeat - 1);
        if (dist < w) {
            dist = dist_out;
            norm = norm_out;
            if (dist < w) return -R1;
        }
        ray->start += dist * ray->direction;
        *normal = norm;
        return dist;
    }
};

template <class G> struct Plane;

// Plane z = 0 with normal (0,0,-1), shape/eu/plane.cc:3-21.
template <> struct Plane<Eu> {
    static constexpr size_t_ SIZE = 0, ALIGN = 1;
    static constexpr bool DYN = false;
    static HD real shape_detect(const u8 *, Context *context, real3 *normal, Light<Eu> *light) {
        if (context_is_repeat(context)) return -R1;
        real3 pos = light->ray.start;
        real3 dir = light->ray.direction;
        if (pos.z * dir.z > R0) return -R1;
        real t = -pos.z / dir.z;
        light->ray.start = r3(pos.x + dir.x * t, pos.y + dir.y * t, R0);
        *normal = r3(R0, R0, -R1);
        return t;
    }
};

// Geodesic plane = unit hemisphere |p| = 1, shape/hy/plane.cc:4-37.
template <> struct Plane<Hy> {
    static constexpr size_t_ SIZE = 0, ALIGN = 1;
    static constexpr bool DYN = false;
    static HD real shape_detect(const u8 *, Context *context, quat *normal, Light<Hy> *light) {
        if (context_is_repeat(context)) return -R1;
        quat p = light->ray.start, d = light->ray.direction;
        real pd = dot(p, d);
        if (fabs_(pd) < EPS) return -R1;
        real s = R1 - length2(p);
        real t = s / (2 * pd);
        if (t < -EPS) return -R1;
        quat h = r4(p.x + d.x * t, p.y + d.y * t, R0, R0);
        real pxy2 = length2(xy(h));
        if (pxy2 > R1) return -R1;
        h.z = sqrt_(R1 - pxy2);
        light->ray.start = h;
        light->ray.direction = Hy::dir_at(p, d, h);
        *normal = -h;
        return Hy::distance(p, h);
    }
};

// Horosphere z = 1 with normal (0,0,-1,0), shape/hy/horosphere.cc:4-34.
struct Horosphere {
    static constexpr size_t_ SIZE = 0, ALIGN = 1;
    static constexpr bool DYN = false;
    static HD real shape_detect(const u8 *, Context *context, quat *normal, Light<Hy> *light) {
        bool repeat = context_is_repeat(context);
        quat p = light->ray.start, d = light->ray.direction;
        real dxy = length(xy(d));
        if (p.z < dxy) return -R1;
        real dt = sqrt_(p.z * p.z - dxy * dxy);
        real t = p.z * d.z - dt;
        real w = EPS * (2 * (int)repeat - 1);
        if (t < w) {
            t += 2 * dt;
            if (t < w) return -R1;
        }
        t /= dxy * dxy;
        quat h = r4(p.x + d.x * t, p.y + d.y * t, R1, R0);
        light->ray.start = h;
        light->ray.direction = Hy::dir_at(p, d, h);
        *normal = r4(R0, R0, -R1, R0);
        return Hy::distance(p, h);
    }
};

// ================================================================ materials

// Cosine-weighted diffuse, material/lambertian.cc:5-14.
struct Lambertian {
    static constexpr size_t_ SIZE = 0, ALIGN = 1;
    static constexpr bool DYN = false;
    static HD bool interact(const u8 *, Context *context, real3 normal, LightLocal *light, color3 *) {
        if (dot(light->direction, normal) > 0) normal = -normal;
        real3 rnd = random_hemisphere_cosine(&context->rng);
        Rotation3 rot = Rotation3::look_at(-normal);
        light->direction = rot.apply_dir(r3(R0, R0, R0), rnd);
        return true;
    }
};

// Mirror, material/specular.cc:3-7.
struct Specular {
    static constexpr size_t_ SIZE = 0, ALIGN = 1;
    static constexpr bool DYN = false;
    static HD bool interact(const u8 *, Context *, real3 normal, LightLocal *light, color3 *) {
        light->direction -= (2 * dot(light->direction, normal)) * normal;
        return true;
    }
};

// Snell refraction or total internal reflection, material/refractive.cc:3-22.
struct Refractive {
    typedef Fields<LReal> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;
    static HD bool interact(const u8 *self, Context *, real3 normal, LightLocal *light, color3 *) {
        real3 d = light->direction;
        real k = ld<real>(self);
        real a = dot(d, normal);
        if (a < -EPS) k = R1 / k;
        real x = (a * a - R1) * (k * k) + R1;
        if (x > EPS) {
            real b = sign_(a) * sqrt_(x);
            d = d * k + (b - a * k) * normal;
        } else {
            d -= (2 * a) * normal;
        }
        light->direction = d;
        return true;
    }
};

// material/transparent.cc:3-6, material/absorbing.cc:3-6.
struct Transparent {
    static constexpr size_t_ SIZE = 0, ALIGN = 1;
    static constexpr bool DYN = false;
    static HD bool interact(const u8 *, Context *, real3, LightLocal *, color3 *) { return true; }
};
struct Absorbing {
    static constexpr size_t_ SIZE = 0, ALIGN = 1;
    static constexpr bool DYN = false;
    static HD bool interact(const u8 *, Context *, real3, LightLocal *, color3 *) { return false; }
};

// Colored {color, inner}: intensity *= color, material/colored.inl:14-23.
template <class M> struct Colored {
    typedef Fields<LColor3, M> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;
    static HD bool interact(const u8 *self, Context *ctx, real3 normal, LightLocal *light, color3 *emission) {
        light->base.intensity *= ld_color3(self + L::offset(0));
        return M::interact(self + L::offset(1), ctx, normal, light, emission);
    }
};

// Emissive {emission, inner}: emission += intensity * e, material/emissive.inl:14-23.
template <class M> struct Emissive {
    typedef Fields<LColor3, M> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;
    static HD bool interact(const u8 *self, Context *ctx, real3 normal, LightLocal *light, color3 *emission) {
        *emission += light->base.intensity * ld_color3(self + L::offset(0));
        return M::interact(self + L::offset(1), ctx, normal, light, emission);
    }
};

// Mixture component {material, portion}, objects/src/material/mixture.rs:4-9.
template <class M> struct Component {
    typedef Fields<M, LReal> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;
};

// Stochastic mixture: one uniform, portions subtracted in declaration order,
// material/mixture/impl.inl:15-27 with mixture/macros.hh:3-13.
template <class... Ms> struct Mixture {
    typedef Fields<Component<Ms>...> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;

    template <size_t_ I, class M, class... Rest>
    static HD bool pick(const u8 *self, real alpha, Context *ctx, real3 normal, LightLocal *light, color3 *emission) {
        const u8 *comp = self + L::offset(I);
        alpha -= ld<real>(comp + Component<M>::L::offset(1));
        if (alpha < R0) {
            return M::interact(comp + Component<M>::L::offset(0), ctx, normal, light, emission);
        }
        if constexpr (sizeof...(Rest) > 0) {
            return pick<I + 1, Rest...>(self, alpha, ctx, normal, light, emission);
        } else {
            return false;
        }
    }
    static HD bool interact(const u8 *self, Context *ctx, real3 normal, LightLocal *light, color3 *emission) {
        real alpha = random_uniform(&ctx->rng);
        return pick<0, Ms...>(self, alpha, ctx, normal, light, emission);
    }
};

// ================================================================ containers

// Vec<T> {usize size; T elements[]}, types/src/containers/vector.rs:19-58,
// kernel/ocl/container/vector.inl:8-49.
template <class T> struct VecLayout {
    static constexpr size_t_ ALIGN = lcm_(sizeof(usize), T::ALIGN);
    static constexpr size_t_ OFF = up(sizeof(usize), T::ALIGN);
    static constexpr size_t_ SIZE = up(OFF, ALIGN);  // minimum size
    static constexpr size_t_ STRIDE = T::SIZE;
    static HD usize len(const u8 *self) { return ld<usize>(self); }
    static HD const u8 *elem(const u8 *self, usize i) { return self + OFF + (size_t_)i * STRIDE; }
};

// Enum {usize index; union {variants}} with single-field variants,
// type-macros/src/entity.rs:239-275, type-macros/src/source.rs:108-179.
template <class... Vs> struct ChoiceLayout {
    static constexpr size_t_ ALIGN_ = lcm_(sizeof(usize), Fields<Vs...>::align());
    static constexpr size_t_ OFF = up(sizeof(usize), ALIGN_);
    static constexpr size_t_ maxsize() {
        size_t_ s[] = {Vs::SIZE..., 0};
        size_t_ m = 0;
        for (size_t_ i = 0; i < sizeof...(Vs); ++i) m = s[i] > m ? s[i] : m;
        return m;
    }
    static constexpr size_t_ vsize() {
        // max over variants of aligned_add(sizeof(usize), variant size, variant align)
        size_t_ s[] = {Vs::SIZE..., 0};
        size_t_ a[] = {Vs::ALIGN..., 1};
        size_t_ m = sizeof(usize);
        for (size_t_ i = 0; i < sizeof...(Vs); ++i) {
            size_t_ v = up(sizeof(usize), a[i]) + s[i];
            m = v > m ? v : m;
        }
        return m;
    }
    static constexpr size_t_ SIZE = up(vsize(), ALIGN_);
    static constexpr size_t_ ALIGN = ALIGN_;
    static HD usize index(const u8 *self) { return ld<usize>(self); }
    static HD const u8 *data(const u8 *self) { return self + OFF; }
};

// ================================================================ mapped

// Mapped {map, inner}, objects/src/mapped.rs:5-11. Acts as a shape
// (shape/mapped.inl:22-30), an object (object/mapped.inl:24-51) or a view
// (view/mapped.inl:20-26) depending on where it is used.
// Cache type of an object entity; views and shapes have none.
template <class T, class = void> struct CacheOf { typedef int type; };
template <class T> struct CacheOf<T, void_t<typename T::Cache>> { typedef typename T::Cache type; };

template <class G, class T, class M> struct Mapped {
    typedef Fields<M, T> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = T::DYN;
    typedef typename CacheOf<T>::type Cache;
    typedef G Geo;
    typedef M Map;
    typedef T Inner;
    static HD M map(const u8 *self) { return M::load(self + L::offset(0)); }
    static HD const u8 *inner(const u8 *self) { return self + L::offset(1); }

    // shape
    static HD real shape_detect(const u8 *self, Context *ctx, typename G::Dir *normal, Light<G> *light) {
        M m = map(self);
        M inv = m.inverse();
        light->ray = ray_map<G>(light->ray, inv);
        real dist = T::shape_detect(inner(self), ctx, normal, light);
        *normal = m.apply_normal(light->ray.start, *normal);
        light->ray = ray_map<G>(light->ray, m);
        return dist;
    }
    // object
    static HD real detect(const u8 *self, Context *ctx, Cache *cache, Light<G> *light) {
        M m = map(self);
        M inv = m.inverse();
        light->ray = ray_map<G>(light->ray, inv);
        real dist = T::detect(inner(self), ctx, cache, light);
        light->ray = ray_map<G>(light->ray, m);
        return dist;
    }
    static HD bool interact(const u8 *self, Context *ctx, Cache *cache, Light<G> *light, color3 *emission) {
        M m = map(self);
        M inv = m.inverse();
        light->ray = ray_map<G>(light->ray, inv);
        bool ret = T::interact(inner(self), ctx, cache, light, emission);
        light->ray = ray_map<G>(light->ray, m);
        return ret;
    }
    // view
    static HD Ray<G> sample(const u8 *self, Context *ctx, real2 pos, real2 size) {
        M m = map(self);
        Ray<G> ray = T::sample(inner(self), ctx, pos, size);
        ray.direction = m.apply_dir(ray.start, ray.direction);
        ray.start = m.apply_pos(ray.start);
        return ray;
    }
};

template <class T> struct IsMapped { static constexpr bool value = false; };
template <class G, class T, class M> struct IsMapped<Mapped<G, T, M>> { static constexpr bool value = true; };

// ================================================================ shape containers

// Nearest hit over a vector of shapes, shape/vector.inl:23-53.
template <class G, class T> struct ShapeVec {
    typedef VecLayout<T> V;
    static constexpr size_t_ SIZE = V::SIZE, ALIGN = V::ALIGN;
    static constexpr bool DYN = true;
    static HD real shape_detect(const u8 *self, Context *ctx, typename G::Dir *normal, Light<G> *light) {
        Hasher orig_hasher = ctx->hasher;
        real min_dist = -R1;
        typename G::Dir min_normal;
        Light<G> min_light;
        Hasher min_hasher;
        usize n = V::len(self);
        for (usize index = 0; index < n; ++index) {
            Light<G> cur_light = *light;
            typename G::Dir cur_normal;
            hash_usize(&ctx->hasher, index);
            real dist = T::shape_detect(V::elem(self, index), ctx, &cur_normal, &cur_light);
            if (dist > -EPS && (min_dist <= -EPS || dist < min_dist)) {
                min_dist = dist;
                min_normal = cur_normal;
                min_light = cur_light;
                min_hasher = ctx->hasher;
            }
            ctx->hasher = orig_hasher;
        }
        if (min_dist > -EPS) {
            *normal = min_normal;
            *light = min_light;
            ctx->hasher = min_hasher;
            return min_dist;
        }
        return -R1;
    }
};

// Shape choice, shape/choice/impl.inl:19-32 with shape/choice/macros.hh:3-10.
template <class G, class... Vs> struct ShapeChoice {
    typedef ChoiceLayout<Vs...> C;
    static constexpr size_t_ SIZE = C::SIZE, ALIGN = C::ALIGN;
    static constexpr bool DYN = false;
    template <usize I, class V, class... Rest>
    static HD real pick(usize index, const u8 *data, Context *ctx, typename G::Dir *normal, Light<G> *light) {
        if (index == I) return V::shape_detect(data, ctx, normal, light);
        if constexpr (sizeof...(Rest) > 0) {
            return pick<I + 1, Rest...>(index, data, ctx, normal, light);
        } else {
            return -R1;
        }
    }
    static HD real shape_detect(const u8 *self, Context *ctx, typename G::Dir *normal, Light<G> *light) {
        return pick<0, Vs...>(C::index(self), C::data(self), ctx, normal, light);
    }
};

// ================================================================ objects

// Covered {material, shape}, object/covered.inl:23-56, objects/src/object/covered.rs:11-23.
template <class G> struct CoveredCache { typename G::Dir normal; };

template <class G, class S, class Mat> struct Covered {
    typedef Fields<Mat, S> L;
    typedef S ShapeT;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;
    typedef CoveredCache<G> Cache;
    static HD real detect(const u8 *self, Context *ctx, Cache *cache, Light<G> *light) {
        return S::shape_detect(self + L::offset(1), ctx, &cache->normal, light);
    }
    static HD bool interact(const u8 *self, Context *ctx, Cache *cache, Light<G> *light, color3 *emission) {
        LightLocal ll = light_to_local(light);
        bool ret = Mat::interact(self + L::offset(0), ctx, G::dir_to_local(light->ray.start, cache->normal), &ll, emission);
        light_update_local(light, ll);
        return ret;
    }
};

// Nearest hit over a vector of objects, object/vector.inl:24-74.
template <class T> struct ObjectVecCache {
    usize index;
    typename T::Cache inner;
};

template <class G, class T> struct ObjectVec {
    typedef VecLayout<T> V;
    static constexpr size_t_ SIZE = V::SIZE, ALIGN = V::ALIGN;
    static constexpr bool DYN = true;
    typedef ObjectVecCache<T> Cache;

    static HD real detect(const u8 *self, Context *ctx, Cache *cache, Light<G> *light) {
        if constexpr (IsMapped<T>::value) {
            return detect_mapped(self, ctx, cache, light);
        } else {
            return detect_plain(self, ctx, cache, light);
        }
    }

    static HD real detect_plain(const u8 *self, Context *ctx, Cache *cache, Light<G> *light) {
        Hasher orig_hasher = ctx->hasher;
        real min_dist = -R1;
        usize min_index = 0;
        Light<G> min_light;
        typename T::Cache min_cache;
        Hasher min_hasher;
        usize n = V::len(self);
        for (usize index = 0; index < n; ++index) {
            Light<G> cur_light = *light;
            typename T::Cache cur_cache;
            hash_usize(&ctx->hasher, index);
            real dist = T::detect(V::elem(self, index), ctx, &cur_cache, &cur_light);
            if (dist > -EPS && (min_dist <= -EPS || dist < min_dist)) {
                min_dist = dist;
                min_index = index;
                min_light = cur_light;
                min_cache = cur_cache;
                min_hasher = ctx->hasher;
            }
            ctx->hasher = orig_hasher;
        }
        if (min_dist > -EPS) {
            cache->inner = min_cache;
            cache->index = min_index;
            *light = min_light;
            ctx->hasher = min_hasher;
            return min_dist;
        }
        return -R1;
    }

    // Same result as detect_plain for Mapped elements. The reference maps
    // every candidate's ray back to the parent frame inside the loop
    // (object/mapped.inl:34); since that map is a pure function of the local
    // ray and the element's map, it is applied once, to the winner, after the
    // loop. The element maps are read with wavefront-uniform addresses, so
    // they come through the scalar cache.
    static HD real detect_mapped(const u8 *self, Context *ctx, Cache *cache, Light<G> *light) {
        typedef typename T::Inner TI;
        typedef typename T::Map TM;
        Hasher orig_hasher = ctx->hasher;
        real min_dist = -R1;
        usize min_index = 0;
        Ray<G> min_ray;
        typename T::Cache min_cache;
        Hasher min_hasher;
        usize n = V::len(self);
        for (usize index = 0; index < n; ++index) {
            const u8 *elem = V::elem(self, index);
            Light<G> cur_light = *light;
            typename T::Cache cur_cache;
            hash_usize(&ctx->hasher, index);
            TM inv = T::map(elem).inverse();
            cur_light.ray = ray_map<G>(cur_light.ray, inv);
            real dist = TI::detect(T::inner(elem), ctx, &cur_cache, &cur_light);
            if (dist > -EPS && (min_dist <= -EPS || dist < min_dist)) {
                min_dist = dist;
                min_index = index;
                min_ray = cur_light.ray;
                min_cache = cur_cache;
                min_hasher = ctx->hasher;
            }
            ctx->hasher = orig_hasher;
        }
        if (min_dist > -EPS) {
            cache->inner = min_cache;
            cache->index = min_index;
            light->ray = ray_map<G>(min_ray, T::map(V::elem(self, min_index)));
            ctx->hasher = min_hasher;
            return min_dist;
        }
        return -R1;
    }

    static HD bool interact(const u8 *self, Context *ctx, Cache *cache, Light<G> *light, color3 *emission) {
        return T::interact(V::elem(self, cache->index), ctx, &cache->inner, light, emission);
    }
};

// Object choice with a cache per variant, object/choice/impl.inl:22-51,
// object/choice/macros.hh:3-20, objects/src/object/choice.rs:1-78.
template <class... Vs> struct ChoiceCacheOf;
template <> struct ChoiceCacheOf<> {};
template <class V, class... Rest> struct ChoiceCacheOf<V, Rest...> {
    typename V::Cache head;
    ChoiceCacheOf<Rest...> tail;
};

template <class G, class... Vs> struct ObjectChoice {
    typedef ChoiceLayout<Vs...> C;
    static constexpr size_t_ SIZE = C::SIZE, ALIGN = C::ALIGN;
    static constexpr bool DYN = false;
    typedef ChoiceCacheOf<Vs...> Cache;

    template <usize I, class CC, class V, class... Rest>
    static HD real pick_detect(usize index, const u8 *data, Context *ctx, CC *cache, Light<G> *light) {
        if (index == I) return V::detect(data, ctx, &cache->head, light);
        if constexpr (sizeof...(Rest) > 0) {
            return pick_detect<I + 1, decltype(cache->tail), Rest...>(index, data, ctx, &cache->tail, light);
        } else {
            return -R1;
        }
    }
    template <usize I, class CC, class V, class... Rest>
    static HD bool pick_interact(usize index, const u8 *data, Context *ctx, CC *cache, Light<G> *light, color3 *emission) {
        if (index == I) return V::interact(data, ctx, &cache->head, light, emission);
        if constexpr (sizeof...(Rest) > 0) {
            return pick_interact<I + 1, decltype(cache->tail), Rest...>(index, data, ctx, &cache->tail, light, emission);
        } else {
            return false;
        }
    }
    static HD real detect(const u8 *self, Context *ctx, Cache *cache, Light<G> *light) {
        return pick_detect<0, Cache, Vs...>(C::index(self), C::data(self), ctx, cache, light);
    }
    static HD bool interact(const u8 *self, Context *ctx, Cache *cache, Light<G> *light, color3 *emission) {
        return pick_interact<0, Cache, Vs...>(C::index(self), C::data(self), ctx, cache, light, emission);
    }
};

// ================================================================ views

// Pinhole at the geometry's origin looking down -z, view/point/impl.inl:18-25.
// The pixel jitter draws x then y.
template <class G> struct PointView {
    typedef Fields<LReal> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;
    static HD Ray<G> sample(const u8 *self, Context *ctx, real2 pos, real2 size) {
        real fov = ld<real>(self);
        real ux = random_uniform(&ctx->rng);
        real uy = random_uniform(&ctx->rng);
        real2 sampled = pos + size * r2(ux - (real)0.5f, uy - (real)0.5f);
        Ray<G> ray;
        ray.start = G::origin();
        ray.direction = G::dir_from_local(ray.start, normalize(r3(sampled.x, sampled.y, -R1 / fov)));
        return ray;
    }
};

// ================================================================ backgrounds

// Constant colour, background/constant.inl:18-21.
struct ConstBg {
    typedef Fields<LColor3> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;
    template <class G> static HD void bg_interact(const u8 *self, Context *, const Light<G> *light, color3 *emission) {
        *emission += ld_color3(self) * light->base.intensity;
    }
};

// Vertical gradient (Euclidean only), background/gradient.cc:3-13.
struct GradBg {
    typedef Fields<LReal3, LColor3, LColor3, LFloat> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = false;
    static HD void bg_interact(const u8 *self, Context *, const Light<Eu> *light, color3 *emission) {
        real3 dir = ld_real3(self + L::offset(0));
        float t = 0.5f * ((float)dot(light->ray.direction, dir) + 1.0f);
        float factor = powf_(fminf(fmaxf(t, 0.0f), 1.0f), ld<float>(self + L::offset(3)));
        color3 color = factor * ld_color3(self + L::offset(1)) + (1.0f - factor) * ld_color3(self + L::offset(2));
        *emission += color * light->base.intensity;
    }
};

// ================================================================ scene

// SceneImpl {view, background, object} with H light hops,
// objects/src/scene.rs:20-67, render/scene.inl:16-44.
template <class G, class V, class T, class B, int H> struct SceneImpl {
    typedef Fields<V, B, T> L;
    static constexpr size_t_ SIZE = L::size(), ALIGN = L::align();
    static constexpr bool DYN = T::DYN;
    static constexpr int HOPS = H;
    typedef G Geo;
    typedef T Object;
    typedef V View;
    static HD const u8 *view(const u8 *s) { return s + L::offset(0); }
    static HD const u8 *background(const u8 *s) { return s + L::offset(1); }
    static HD const u8 *object(const u8 *s) { return s + L::offset(2); }

    static HD void bg(const u8 *self, Context *ctx, const Light<G> *light, color3 *emission) {
        B::bg_interact(background(self), ctx, light, emission);
    }
};

}  // namespace ht
