// Device plugin region, part 1 of 2 (ht_plugin/end.hh closes it).
//
// Entities written against the reference's OpenCL-C device interface compile
// here unchanged in form: a C struct of the entity's fields (the layout the
// reference's derive emits; zero-size entities are `void`), free functions
// with the interface's names and signatures, and the declaring macros
//   define_shape_interface(Self, self, Geo, geo)       shape/interface.hh:6-15
//     real self_detect(__global const Self *, Context *, GeoDir *normal, LightGeo *)
//   define_object_interface(Self, self, Cache, cache, Geo, geo)   object/interface.hh:4-22
//     real self_detect(__global const Self *, Context *, Cache *, LightGeo *)
//     bool self_interact(__global const Self *, Context *, Cache *, LightGeo *, color3 *emission)
//   define_material_interface(Self, self)              material/interface.hh:5-12
//     bool self_interact(__global const Self *, Context *, real3 normal, LightLocal *, float3 *emission)
//   define_background_interface(Self, self, Geo, geo)  background/interface.hh:7-16
//     void self_interact(__global const Self *, Context *, const LightGeo *, color3 *emission)
//   define_view_interface(Self, self, Geo, geo)        view/interface.hh:5-14
//     RayGeo self_sample(__global const Self *, Context *, real2 pixel_pos, real2 pixel_size)
// Each macro also defines the adaptor type `ht_plugin::self`, an entity of
// this library (SIZE / ALIGN from the C struct, the static interface
// functions of entities.hh) that converts the library's values to the
// reference's C types, calls the plugin's function and converts back; the
// scene's type expression names the plugin by that adaptor, e.g.
// ht::Covered<ht::Eu, ht_plugin::my_shape, ht_plugin::my_material>.
//
// Inside the region the reference's vocabulary is in the global namespace
// with the reference's layouts: real, real2/3/4 and float3 / double3 with
// OpenCL's 3-component size and alignment (those of 4 components), color3,
// quat, usize, Context, Rng, LightBase, LightLocal, LightEu / LightHy,
// RayEu / RayHy, EuPos / EuDir / HyPos / HyDir, Rotation3; the constants
// R0, R1, EPS, PI; make(T)(...) constructors; vector arithmetic, dot, length,
// length2, normalize, cross, distance; random_uniform, random_uniform2,
// random_hemisphere_cosine, rng_next; context_is_repeat; rot3_look_at,
// rot3_apply_pos / _dir / _normal; light_eu_to_local / light_hy_to_local
// and light_eu_update_local / light_hy_update_local; `__global` and the
// attribute macros expand to nothing. Every function defined in the region
// is compiled for the device (clang's force_cuda_host_device pragma), as the
// reference's OpenCL C functions were. Each bridged function computes with
// the library's own operation order, so a plugin that restates a built-in
// entity renders bit-identical images.
#pragma once

#include "../ht/entities.hh"

#ifdef HT_PLUGIN_REGION
#error "ht_plugin/begin.hh: plugin region already open"
#endif
#define HT_PLUGIN_REGION 1

namespace ht_cl {

// OpenCL C vectors: n = 3 takes the size and alignment of n = 4.
template <class T> struct alignas(2 * sizeof(T)) vec2 {
    T x, y;
    vec2() = default;
    HD explicit vec2(T s) : x(s), y(s) {}
    HD vec2(T a, T b) : x(a), y(b) {}
};
template <class T> struct alignas(4 * sizeof(T)) vec3 {
    T x, y, z;
    vec3() = default;
    HD explicit vec3(T s) : x(s), y(s), z(s) {}
    HD vec3(T a, T b, T c) : x(a), y(b), z(c) {}
};
template <class T> struct alignas(4 * sizeof(T)) vec4 {
    T x, y, z, w;
    vec4() = default;
    HD explicit vec4(T s) : x(s), y(s), z(s), w(s) {}
    HD vec4(T a, T b, T c, T d) : x(a), y(b), z(c), w(d) {}
};

// Scalars of vector-scalar operations convert to the component type (they
// take no part in deducing it), as OpenCL C converts them.
template <class T> struct scalar_of { typedef T type; };
#define S_ typename scalar_of<T>::type
#define HTP_OPS(V, EXPAND2, EXPAND1, NEG)                                                       \
    template <class T> HD V<T> operator+(V<T> a, V<T> b) { return EXPAND2(+); }                 \
    template <class T> HD V<T> operator-(V<T> a, V<T> b) { return EXPAND2(-); }                 \
    template <class T> HD V<T> operator*(V<T> a, V<T> b) { return EXPAND2(*); }                 \
    template <class T> HD V<T> operator/(V<T> a, V<T> b) { return EXPAND2(/); }                 \
    template <class T> HD V<T> operator*(V<T> a, S_ s) { return EXPAND1(a, *, s); }             \
    template <class T> HD V<T> operator*(S_ s, V<T> a) { return EXPAND1(a, *, s); }             \
    template <class T> HD V<T> operator/(V<T> a, S_ s) { return EXPAND1(a, /, s); }             \
    template <class T> HD V<T> operator-(V<T> a) { return NEG; }                                \
    template <class T> HD V<T> &operator+=(V<T> &a, V<T> b) { a = a + b; return a; }            \
    template <class T> HD V<T> &operator-=(V<T> &a, V<T> b) { a = a - b; return a; }            \
    template <class T> HD V<T> &operator*=(V<T> &a, V<T> b) { a = a * b; return a; }            \
    template <class T> HD V<T> &operator/=(V<T> &a, V<T> b) { a = a / b; return a; }            \
    template <class T> HD V<T> &operator*=(V<T> &a, S_ s) { a = a * s; return a; }              \
    template <class T> HD V<T> &operator/=(V<T> &a, S_ s) { a = a / s; return a; }

#define HTP_E2_2(OP) V2(a.x OP b.x, a.y OP b.y)
#define HTP_E1_2(A, OP, S) V2(A.x OP S, A.y OP S)
#define HTP_E2_3(OP) V3(a.x OP b.x, a.y OP b.y, a.z OP b.z)
#define HTP_E1_3(A, OP, S) V3(A.x OP S, A.y OP S, A.z OP S)
#define HTP_E2_4(OP) V4(a.x OP b.x, a.y OP b.y, a.z OP b.z, a.w OP b.w)
#define HTP_E1_4(A, OP, S) V4(A.x OP S, A.y OP S, A.z OP S, A.w OP S)
#define V2 vec2<T>
#define V3 vec3<T>
#define V4 vec4<T>
HTP_OPS(vec2, HTP_E2_2, HTP_E1_2, V2(-a.x, -a.y))
HTP_OPS(vec3, HTP_E2_3, HTP_E1_3, V3(-a.x, -a.y, -a.z))
HTP_OPS(vec4, HTP_E2_4, HTP_E1_4, V4(-a.x, -a.y, -a.z, -a.w))
#undef V2
#undef V3
#undef V4
#undef S_
#undef HTP_OPS
#undef HTP_E2_2
#undef HTP_E1_2
#undef HTP_E2_3
#undef HTP_E1_3
#undef HTP_E2_4
#undef HTP_E1_4

// Conversions between the library's values and the reference's C types.
HD ht::real2 to_ht(vec2<ht::real> a) { return {a.x, a.y}; }
HD ht::real3 to_ht(vec3<ht::real> a) { return {a.x, a.y, a.z}; }
HD ht::real4 to_ht(vec4<ht::real> a) { return {a.x, a.y, a.z, a.w}; }
HD ht::color3 to_color(vec3<float> a) { return {a.x, a.y, a.z}; }
HD vec2<ht::real> from_ht(ht::real2 a) { return {a.x, a.y}; }
HD vec3<ht::real> from_ht(ht::real3 a) { return {a.x, a.y, a.z}; }
HD vec4<ht::real> from_ht(ht::real4 a) { return {a.x, a.y, a.z, a.w}; }
// Colours convert by name: with FP32 geometry float3 and real3 are one type.
HD vec3<float> color_from_ht(ht::color3 a) { return {a.x, a.y, a.z}; }

// Size and alignment of a plugin's C struct; `void` is a zero-size entity.
template <class T> struct Layout {
    static constexpr ht::size_t_ SIZE = sizeof(T), ALIGN = alignof(T);
    static constexpr bool DYN = false;
};
template <> struct Layout<void> {
    static constexpr ht::size_t_ SIZE = 0, ALIGN = 1;
    static constexpr bool DYN = false;
};
// Storage the containers hold for a plugin object's cache type.
struct NoCache {};
template <class C> struct CacheStore { typedef C type; };
template <> struct CacheStore<void> { typedef NoCache type; };

template <class T> HD const T *self_of(const ht::u8 *p) { return reinterpret_cast<const T *>(p); }

}  // namespace ht_cl

// ------------------------------------------------------------ the reference's vocabulary
typedef ht::real real;
typedef ht::usize usize;
typedef unsigned int uint;
typedef ht_cl::vec2<ht::real> real2;
typedef ht_cl::vec3<ht::real> real3;
typedef ht_cl::vec4<ht::real> real4;
typedef ht_cl::vec3<float> htp_float3;
typedef ht_cl::vec3<double> htp_double3;
typedef htp_float3 color3;
typedef real4 quat;
typedef ht::Rotation3 Rotation3;
typedef real3 EuPos;
typedef real3 EuDir;
typedef quat HyPos;
typedef quat HyDir;

using ht::EPS;
using ht::PI;
using ht::R0;
using ht::R1;

// The path context (render/context.hh:7-11) as plugin types of their own,
// copied to and from the library's at each call, so that argument-dependent
// lookup never mixes the library's functions into the plugin's calls.
typedef struct Rng {
    uint state;
} Rng;
typedef struct Hasher {
    uint state;
} Hasher;
typedef struct Context {
    Rng rng;
    Hasher hasher;
    uint prev_hash;
} Context;

typedef struct LightBase {
    color3 intensity;
    bool diffuse;
} LightBase;
typedef struct LightLocal {
    LightBase base;
    real3 direction;
} LightLocal;
typedef struct RayEu {
    EuPos start;
    EuDir direction;
} RayEu;
typedef struct RayHy {
    HyPos start;
    HyDir direction;
} RayHy;
typedef struct LightEu {
    LightBase base;
    RayEu ray;
} LightEu;
typedef struct LightHy {
    LightBase base;
    RayHy ray;
} LightHy;

namespace ht_cl {
HD ::Rng from_ht(const ht::Rng &r) { return {r.state}; }
HD ht::Rng to_ht(const ::Rng &r) { return {r.state}; }
HD ::Context from_ht(const ht::Context &c) { return {{c.rng.state}, {c.hasher.state}, c.prev_hash}; }
HD ht::Context to_ht(const ::Context &c) { return {{c.rng.state}, {c.hasher.state}, c.prev_hash}; }
HD ::LightBase from_ht(const ht::LightBase &b) { return {color_from_ht(b.intensity), b.diffuse}; }
HD ht::LightBase to_ht(const ::LightBase &b) { return {to_color(b.intensity), b.diffuse}; }
HD ::LightLocal from_ht(const ht::LightLocal &l) { return {from_ht(l.base), from_ht(l.direction)}; }
HD ht::LightLocal to_ht(const ::LightLocal &l) { return {to_ht(l.base), to_ht(l.direction)}; }
HD ::RayEu from_ht(const ht::Ray<ht::Eu> &r) { return {from_ht(r.start), from_ht(r.direction)}; }
HD ht::Ray<ht::Eu> to_ht(const ::RayEu &r) { return {to_ht(r.start), to_ht(r.direction)}; }
HD ::RayHy from_ht(const ht::Ray<ht::Hy> &r) { return {from_ht(r.start), from_ht(r.direction)}; }
HD ht::Ray<ht::Hy> to_ht(const ::RayHy &r) { return {to_ht(r.start), to_ht(r.direction)}; }
HD ::LightEu from_ht(const ht::Light<ht::Eu> &l) { return {from_ht(l.base), from_ht(l.ray)}; }
HD ht::Light<ht::Eu> to_ht(const ::LightEu &l) { return {to_ht(l.base), to_ht(l.ray)}; }
HD ::LightHy from_ht(const ht::Light<ht::Hy> &l) { return {from_ht(l.base), from_ht(l.ray)}; }
HD ht::Light<ht::Hy> to_ht(const ::LightHy &l) { return {to_ht(l.base), to_ht(l.ray)}; }
// The reference's light type of a geometry.
template <class G> struct LightOf;
template <> struct LightOf<ht::Eu> { typedef ::LightEu type; typedef ::RayEu ray; };
template <> struct LightOf<ht::Hy> { typedef ::LightHy type; typedef ::RayHy ray; };
}  // namespace ht_cl

#pragma clang force_cuda_host_device begin

// The reference's spellings of address spaces and attributes (kernel/ocl/types.hh).
#define __global
#define __allow_unused_params__
#define __allow_multiple__ inline
// OpenCL-layout 3-component vectors under their OpenCL names (HIP's own
// float3 / double3 have the size of 3 components).
#define float3 htp_float3
#define double3 htp_double3
// make(T)(a, ...) builds a vector (one argument broadcasts), macros.hh.
#define make(T) T
// Token pasting of the reference's macro templates (macros.hh).
#define $2(a, b) a##b
#define $3(a, b, c) a##b##c

// ------------------------------------------------------------ bridged functions
inline real dot(real2 a, real2 b) { return ht::dot(ht_cl::to_ht(a), ht_cl::to_ht(b)); }
inline real dot(real3 a, real3 b) { return ht::dot(ht_cl::to_ht(a), ht_cl::to_ht(b)); }
inline real dot(real4 a, real4 b) { return ht::dot(ht_cl::to_ht(a), ht_cl::to_ht(b)); }
template <class V> inline real length2(V a) { return dot(a, a); }
template <class V> inline real length(V a) { return ht::length(ht_cl::to_ht(a)); }
template <class V> inline V normalize(V a) { return ht_cl::from_ht(ht::normalize(ht_cl::to_ht(a))); }
template <class V> inline real distance(V a, V b) { return length(a - b); }
inline real3 cross(real3 a, real3 b) { return ht_cl::from_ht(ht::cross(ht_cl::to_ht(a), ht_cl::to_ht(b))); }

inline uint rng_next(Rng *rng) {
    ht::Rng r = ht_cl::to_ht(*rng);
    uint v = ht::rng_next(&r);
    *rng = ht_cl::from_ht(r);
    return v;
}
inline real random_uniform(Rng *rng) {
    ht::Rng r = ht_cl::to_ht(*rng);
    real u = ht::random_uniform(&r);
    *rng = ht_cl::from_ht(r);
    return u;
}
inline real2 random_uniform2(Rng *rng) {
    real x = random_uniform(rng);
    real y = random_uniform(rng);
    return real2(x, y);
}
inline real3 random_hemisphere_cosine(Rng *rng) {
    ht::Rng r = ht_cl::to_ht(*rng);
    real3 v = ht_cl::from_ht(ht::random_hemisphere_cosine(&r));
    *rng = ht_cl::from_ht(r);
    return v;
}
inline bool context_is_repeat(const Context *context) {
    ht::Context c = ht_cl::to_ht(*context);
    return ht::context_is_repeat(&c);
}

inline Rotation3 rot3_look_at(real3 d) { return ht::Rotation3::look_at(ht_cl::to_ht(d)); }
inline real3 rot3_apply_pos(Rotation3 m, real3 p) { return ht_cl::from_ht(m.apply_pos(ht_cl::to_ht(p))); }
inline real3 rot3_apply_dir(Rotation3 m, real3 p, real3 d) {
    return ht_cl::from_ht(m.apply_dir(ht_cl::to_ht(p), ht_cl::to_ht(d)));
}
inline real3 rot3_apply_normal(Rotation3 m, real3 p, real3 d) {
    return ht_cl::from_ht(m.apply_normal(ht_cl::to_ht(p), ht_cl::to_ht(d)));
}

inline LightLocal light_eu_to_local(LightEu *light) {
    ht::Light<ht::Eu> l = ht_cl::to_ht(*light);
    return ht_cl::from_ht(ht::light_to_local(&l));
}
inline void light_eu_update_local(LightEu *light, LightLocal ll) {
    ht::Light<ht::Eu> l = ht_cl::to_ht(*light);
    ht::light_update_local(&l, ht_cl::to_ht(ll));
    *light = ht_cl::from_ht(l);
}
inline LightLocal light_hy_to_local(LightHy *light) {
    ht::Light<ht::Hy> l = ht_cl::to_ht(*light);
    return ht_cl::from_ht(ht::light_to_local(&l));
}
inline void light_hy_update_local(LightHy *light, LightLocal ll) {
    ht::Light<ht::Hy> l = ht_cl::to_ht(*light);
    ht::light_update_local(&l, ht_cl::to_ht(ll));
    *light = ht_cl::from_ht(l);
}

// ------------------------------------------------------------ interface macros
#define define_shape_interface(Self_, self_, Geo_, geo_)                                                     \
    real self_##_detect(__global const Self_ *self, Context *context, Geo_##Dir *normal, Light##Geo_ *light); \
    namespace ht_plugin {                                                                                    \
    struct self_ : ::ht_cl::Layout<Self_> {                                                                  \
        static HD ::ht::real shape_detect(const ::ht::u8 *p, ::ht::Context *ctx, ::ht::Geo_::Dir *normal,    \
                                          ::ht::Light<::ht::Geo_> *light) {                                  \
            ::Light##Geo_ l = ::ht_cl::from_ht(*light);                                                      \
            ::Geo_##Dir n = ::ht_cl::from_ht(*normal);                                                       \
            ::Context c = ::ht_cl::from_ht(*ctx);                                                            \
            ::ht::real t = ::self_##_detect(::ht_cl::self_of<Self_>(p), &c, &n, &l);                         \
            *ctx = ::ht_cl::to_ht(c);                                                                        \
            *light = ::ht_cl::to_ht(l);                                                                      \
            *normal = ::ht_cl::to_ht(n);                                                                     \
            return t;                                                                                        \
        }                                                                                                    \
    };                                                                                                       \
    }

#define define_object_interface(Self_, self_, Cache_, cache_, Geo_, geo_)                                    \
    real self_##_detect(__global const Self_ *self, Context *context, Cache_ *cache, Light##Geo_ *light);    \
    bool self_##_interact(__global const Self_ *self, Context *context, Cache_ *cache, Light##Geo_ *light,  \
                          color3 *emission);                                                                 \
    namespace ht_plugin {                                                                                    \
    struct self_ : ::ht_cl::Layout<Self_> {                                                                  \
        typedef typename ::ht_cl::CacheStore<Cache_>::type Cache;                                            \
        static HD ::ht::real detect(const ::ht::u8 *p, ::ht::Context *ctx, Cache *cache,                     \
                                    ::ht::Light<::ht::Geo_> *light) {                                        \
            ::Light##Geo_ l = ::ht_cl::from_ht(*light);                                                      \
            ::Context c = ::ht_cl::from_ht(*ctx);                                                            \
            ::ht::real t = ::self_##_detect(::ht_cl::self_of<Self_>(p), &c,                                  \
                                            static_cast<Cache_ *>(static_cast<void *>(cache)), &l);          \
            *ctx = ::ht_cl::to_ht(c);                                                                        \
            *light = ::ht_cl::to_ht(l);                                                                      \
            return t;                                                                                        \
        }                                                                                                    \
        static HD bool interact(const ::ht::u8 *p, ::ht::Context *ctx, Cache *cache,                         \
                                ::ht::Light<::ht::Geo_> *light, ::ht::color3 *emission) {                    \
            ::Light##Geo_ l = ::ht_cl::from_ht(*light);                                                      \
            ::color3 e = ::ht_cl::color_from_ht(*emission);                                                  \
            ::Context c = ::ht_cl::from_ht(*ctx);                                                            \
            bool r = ::self_##_interact(::ht_cl::self_of<Self_>(p), &c,                                      \
                                        static_cast<Cache_ *>(static_cast<void *>(cache)), &l, &e);          \
            *ctx = ::ht_cl::to_ht(c);                                                                        \
            *light = ::ht_cl::to_ht(l);                                                                      \
            *emission = ::ht_cl::to_color(e);                                                                \
            return r;                                                                                        \
        }                                                                                                    \
    };                                                                                                       \
    }

#define define_material_interface(Self_, self_)                                                              \
    bool self_##_interact(__global const Self_ *self, Context *context, real3 normal, LightLocal *light,     \
                          float3 *emission);                                                                 \
    namespace ht_plugin {                                                                                    \
    struct self_ : ::ht_cl::Layout<Self_> {                                                                  \
        static HD bool interact(const ::ht::u8 *p, ::ht::Context *ctx, ::ht::real3 normal,                   \
                                ::ht::LightLocal *light, ::ht::color3 *emission) {                           \
            ::LightLocal l = ::ht_cl::from_ht(*light);                                                       \
            ::color3 e = ::ht_cl::color_from_ht(*emission);                                                  \
            ::Context c = ::ht_cl::from_ht(*ctx);                                                            \
            bool r = ::self_##_interact(::ht_cl::self_of<Self_>(p), &c, ::ht_cl::from_ht(normal), &l, &e);   \
            *ctx = ::ht_cl::to_ht(c);                                                                        \
            *light = ::ht_cl::to_ht(l);                                                                      \
            *emission = ::ht_cl::to_color(e);                                                                \
            return r;                                                                                        \
        }                                                                                                    \
    };                                                                                                       \
    }

#define define_background_interface(Self_, self_, Geo_, geo_)                                                \
    void self_##_interact(__global const Self_ *self, Context *context, const Light##Geo_ *light,             \
                          color3 *emission);                                                                 \
    namespace ht_plugin {                                                                                    \
    struct self_ : ::ht_cl::Layout<Self_> {                                                                  \
        static HD void bg_interact(const ::ht::u8 *p, ::ht::Context *ctx, const ::ht::Light<::ht::Geo_> *light, \
                                   ::ht::color3 *emission) {                                                 \
            const ::Light##Geo_ l = ::ht_cl::from_ht(*light);                                                \
            ::color3 e = ::ht_cl::color_from_ht(*emission);                                                  \
            ::Context c = ::ht_cl::from_ht(*ctx);                                                            \
            ::self_##_interact(::ht_cl::self_of<Self_>(p), &c, &l, &e);                                      \
            *ctx = ::ht_cl::to_ht(c);                                                                        \
            *emission = ::ht_cl::to_color(e);                                                                \
        }                                                                                                    \
    };                                                                                                       \
    }

#define define_view_interface(Self_, self_, Geo_, geo_)                                                      \
    Ray##Geo_ self_##_sample(__global const Self_ *view, Context *context, real2 pixel_pos, real2 pixel_size); \
    namespace ht_plugin {                                                                                    \
    struct self_ : ::ht_cl::Layout<Self_> {                                                                  \
        static HD ::ht::Ray<::ht::Geo_> sample(const ::ht::u8 *p, ::ht::Context *ctx, ::ht::real2 pos,       \
                                               ::ht::real2 size) {                                           \
            ::Context c = ::ht_cl::from_ht(*ctx);                                                            \
            ::Ray##Geo_ ray = ::self_##_sample(::ht_cl::self_of<Self_>(p), &c, ::ht_cl::from_ht(pos),         \
                                               ::ht_cl::from_ht(size));                                      \
            *ctx = ::ht_cl::to_ht(c);                                                                        \
            return ::ht_cl::to_ht(ray);                                                                      \
        }                                                                                                    \
    };                                                                                                       \
    }
