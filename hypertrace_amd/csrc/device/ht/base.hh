// Device base layer: scalar types, small vectors, math wrappers, byte-layout
// arithmetic, the per-path LCG and the sdbm path hasher.
//
// Program-level switches (set by the program builder, like the reference's
// DOUBLE_PRECISION / ADDRESS_WIDTH flags, kernel/ocl/types.hh:3-58):
//   HT_DOUBLE          real = double (geometry in FP64; colours stay FP32)
//   HT_ADDRESS_WIDTH   32 or 64: width of `usize` in the scene byte stream
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#ifndef HT_ADDRESS_WIDTH
#define HT_ADDRESS_WIDTH 32
#endif

#define HD __device__ __forceinline__

namespace ht {

typedef unsigned int uint;
typedef unsigned char u8;
typedef unsigned long long u64;
typedef unsigned long size_t_;

#ifdef HT_DOUBLE
typedef double real;
#define HT_EPS 1e-12
#define HT_PI 3.141592653589793
#else
typedef float real;
#ifdef HT_EPS32
// Program option: a wider FP32 hit window than the reference's 1e-6
// (algebra/real.hh:13). FP32 hit points at world coordinates of order 10
// carry roundings of order 1e-6, the size of the reference's window, so its
// FP32 images differ from FP64 ones where rays leave a surface; a window of
// 1e-5 removes that difference. Not the reference's results.
#define HT_EPS HT_EPS32
#else
#define HT_EPS 1e-6f
#endif
#define HT_PI 3.14159265358979f
#endif

#if HT_ADDRESS_WIDTH == 64
typedef unsigned long long usize;
#elif HT_ADDRESS_WIDTH == 32
typedef unsigned int usize;
#else
#error "HT_ADDRESS_WIDTH must be 32 or 64"
#endif

static constexpr real R0 = real(0);
static constexpr real R1 = real(1);
static constexpr real R2 = real(2);
static constexpr real EPS = HT_EPS;
static constexpr real PI = HT_PI;

// ---------------------------------------------------------------- layout
// Byte layout of scene entities follows the reference serializer's rules
// (types/src/math.rs:6-16, type-macros/src/entity.rs:7-186): a field starts
// at the running end rounded up to the field's alignment, a struct's size is
// rounded up to the lcm of its field alignments, zero-size fields take no
// bytes.
constexpr size_t_ up(size_t_ x, size_t_ m) { return (x + m - 1) / m * m; }
constexpr size_t_ gcd_(size_t_ a, size_t_ b) { return b == 0 ? a : gcd_(b, a % b); }
constexpr size_t_ lcm_(size_t_ a, size_t_ b) { return a / gcd_(a, b) * b; }

template <class... F> struct Fields {
    static constexpr size_t_ N = sizeof...(F);
    static constexpr size_t_ offset(size_t_ k) {
        const size_t_ s[] = {F::SIZE..., 0};
        const size_t_ a[] = {F::ALIGN..., 1};
        size_t_ acc = 0;
        for (size_t_ i = 0; i < k; ++i) acc = up(acc, a[i]) + s[i];
        return up(acc, a[k]);
    }
    static constexpr size_t_ align() {
        const size_t_ a[] = {F::ALIGN..., 1};
        size_t_ r = 1;
        for (size_t_ i = 0; i < N; ++i) r = lcm_(r, a[i]);
        return r;
    }
    static constexpr size_t_ size() {
        const size_t_ s[] = {F::SIZE..., 0};
        const size_t_ a[] = {F::ALIGN..., 1};
        size_t_ acc = 0;
        for (size_t_ i = 0; i < N; ++i) acc = up(acc, a[i]) + s[i];
        return up(acc, align());
    }
};

template <class...> using void_t = void;

// Leaf layout descriptors for plain values stored in the byte stream.
template <size_t_ S, size_t_ A> struct Leaf {
    static constexpr size_t_ SIZE = S;
    static constexpr size_t_ ALIGN = A;
    static constexpr bool DYN = false;
};
typedef Leaf<sizeof(real), sizeof(real)> LReal;
typedef Leaf<4, 4> LFloat;
typedef Leaf<sizeof(usize), sizeof(usize)> LUsize;
typedef Leaf<4 * sizeof(real), 4 * sizeof(real)> LReal3;   // 3 slots + pad
typedef Leaf<4 * sizeof(real), 4 * sizeof(real)> LReal4;
typedef Leaf<16, 16> LColor3;                               // float3: 16 B
typedef Leaf<8 * sizeof(real), 8 * sizeof(real)> LComp2x2;  // real8
typedef Leaf<0, 1> LZero;

template <class T> HD T ld(const u8 *p) { return *reinterpret_cast<const T *>(p); }

// ---------------------------------------------------------------- vectors
struct real2 { real x, y; };
struct real3 { real x, y, z; };
struct real4 { real x, y, z, w; };
struct color3 { float x, y, z; };

HD real2 r2(real x, real y) { return {x, y}; }
HD real3 r3(real x, real y, real z) { return {x, y, z}; }
HD real4 r4(real x, real y, real z, real w) { return {x, y, z, w}; }
HD color3 c3(float x, float y, float z) { return {x, y, z}; }
HD color3 c3(float v) { return {v, v, v}; }

#define HT_VEC_OPS(V, S, EXPAND2, EXPAND1)                                       \
    HD V operator+(V a, V b) { return EXPAND2(+); }                              \
    HD V operator-(V a, V b) { return EXPAND2(-); }                              \
    HD V operator*(V a, V b) { return EXPAND2(*); }                              \
    HD V operator/(V a, V b) { return EXPAND2(/); }                              \
    HD V operator*(V a, S s) { return EXPAND1(a, *, s); }                        \
    HD V operator*(S s, V a) { return EXPAND1(a, *, s); }                        \
    HD V operator/(V a, S s) { return EXPAND1(a, /, s); }                        \
    HD V &operator+=(V &a, V b) { a = a + b; return a; }                         \
    HD V &operator-=(V &a, V b) { a = a - b; return a; }                         \
    HD V &operator*=(V &a, V b) { a = a * b; return a; }                         \
    HD V &operator*=(V &a, S s) { a = a * s; return a; }

#define HT_E2_2(OP) {a.x OP b.x, a.y OP b.y}
#define HT_E1_2(A, OP, S) {A.x OP S, A.y OP S}
#define HT_E2_3(OP) {a.x OP b.x, a.y OP b.y, a.z OP b.z}
#define HT_E1_3(A, OP, S) {A.x OP S, A.y OP S, A.z OP S}
#define HT_E2_4(OP) {a.x OP b.x, a.y OP b.y, a.z OP b.z, a.w OP b.w}
#define HT_E1_4(A, OP, S) {A.x OP S, A.y OP S, A.z OP S, A.w OP S}

HT_VEC_OPS(real2, real, HT_E2_2, HT_E1_2)
HT_VEC_OPS(real3, real, HT_E2_3, HT_E1_3)
HT_VEC_OPS(real4, real, HT_E2_4, HT_E1_4)
HT_VEC_OPS(color3, float, HT_E2_3, HT_E1_3)

HD real2 operator-(real2 a) { return {-a.x, -a.y}; }
HD real3 operator-(real3 a) { return {-a.x, -a.y, -a.z}; }
HD real4 operator-(real4 a) { return {-a.x, -a.y, -a.z, -a.w}; }

// ---------------------------------------------------------------- math
// Geometry runs in `real`; colours in float.
//
// Two FP32 code-generation modes, chosen at program build:
//  HT_STRICT_FP (the default program build): every operation rounds on its
//    own (no FMA contraction, IEEE division and square root, normalize as
//    v / length(v)), and FP32 transcendentals are evaluated in FP64 and
//    rounded once, which is correctly rounded except in ~2^-28 of cases. The
//    result no longer depends on one libm's ulp errors, so it agrees with any
//    other correctly rounded evaluation of the reference formulas.
//  fast: FMA contraction, hardware reciprocal-based division (within the
//    2.5 ulp OpenCL allows the reference), rsqrt-based normalize, the FP32
//    device libm.
HD float sqrt_(float x) { return __builtin_sqrtf(x); }
HD double sqrt_(double x) { return __builtin_sqrt(x); }
HD float rsqrt_(float x) { return __builtin_amdgcn_rsqf(x); }
HD double rsqrt_(double x) { return 1.0 / __builtin_sqrt(x); }
#ifdef HT_STRICT_FP
HD float log_(float x) { return (float)::log((double)x); }
HD float exp_(float x) { return (float)::exp((double)x); }
HD float cos_(float x) { return (float)::cos((double)x); }
HD float sin_(float x) { return (float)::sin((double)x); }
HD float cosh_(float x) { return (float)::cosh((double)x); }
HD float sinh_(float x) { return (float)::sinh((double)x); }
HD float atan2_(float y, float x) { return (float)::atan2((double)y, (double)x); }
HD float powf_(float a, float b) { return (float)::pow((double)a, (double)b); }
#else
HD float log_(float x) { return ::logf(x); }
HD float exp_(float x) { return ::expf(x); }
HD float cos_(float x) { return ::cosf(x); }
HD float sin_(float x) { return ::sinf(x); }
HD float cosh_(float x) { return ::coshf(x); }
HD float sinh_(float x) { return ::sinhf(x); }
HD float atan2_(float y, float x) { return ::atan2f(y, x); }
HD float powf_(float a, float b) { return ::powf(a, b); }
#endif
HD double log_(double x) { return ::log(x); }
HD double exp_(double x) { return ::exp(x); }
HD double cos_(double x) { return ::cos(x); }
HD double sin_(double x) { return ::sin(x); }
HD double cosh_(double x) { return ::cosh(x); }
HD double sinh_(double x) { return ::sinh(x); }
HD double atan2_(double y, double x) { return ::atan2(y, x); }
HD float fabs_(float x) { return __builtin_fabsf(x); }
HD double fabs_(double x) { return __builtin_fabs(x); }
HD float fmin_(float a, float b) { return __builtin_fminf(a, b); }
HD double fmin_(double a, double b) { return __builtin_fmin(a, b); }
HD float fmax_(float a, float b) { return __builtin_fmaxf(a, b); }
HD double fmax_(double a, double b) { return __builtin_fmax(a, b); }
HD float floor_(float x) { return __builtin_floorf(x); }
HD double floor_(double x) { return __builtin_floor(x); }
HD real sign_(real a) { return real((R0 < a) - (a < R0)); }

HD real dot(real2 a, real2 b) { return a.x * b.x + a.y * b.y; }
HD real dot(real3 a, real3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HD real dot(real4 a, real4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
template <class V> HD real length2(V a) { return dot(a, a); }
template <class V> HD real length(V a) { return sqrt_(dot(a, a)); }
#ifdef HT_STRICT_FP
// OpenCL normalize(v) = v / length(v), each operation rounded on its own.
template <class V> HD V normalize(V a) { return a / sqrt_(dot(a, a)); }
#else
template <class V> HD V normalize(V a) { return a * rsqrt_(dot(a, a)); }
#endif
HD real3 cross(real3 a, real3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
HD real3 fmin_(real3 a, real3 b) { return {fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
HD real3 fmax_(real3 a, real3 b) { return {fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }
HD real3 sign_(real3 a) { return {sign_(a.x), sign_(a.y), sign_(a.z)}; }
HD real2 xy(real3 a) { return {a.x, a.y}; }
HD real2 xy(real4 a) { return {a.x, a.y}; }
HD real3 xyz(real4 a) { return {a.x, a.y, a.z}; }

// Loads of vector values from the byte stream (real3 occupies 4 slots).
HD real3 ld_real3(const u8 *p) {
    const real *r = reinterpret_cast<const real *>(p);
    return {r[0], r[1], r[2]};
}
HD real4 ld_real4(const u8 *p) {
    const real *r = reinterpret_cast<const real *>(p);
    return {r[0], r[1], r[2], r[3]};
}
HD color3 ld_color3(const u8 *p) {
    const float *r = reinterpret_cast<const float *>(p);
    return {r[0], r[1], r[2]};
}

// ---------------------------------------------------------------- RNG
// Per-path LCG, reference kernel/ocl/random.cc:3-14: s <- 1103515245 s + 12345
// (mod 2^32); a uniform is the new state converted to real (round to nearest)
// times 2^-32 (exact, so it equals the reference's division by 2^32).
struct Rng { uint state; };
HD uint rng_next(Rng *rng) { return (rng->state = 1103515245u * rng->state + 12345u); }
HD real random_uniform(Rng *rng) {
    return (real)rng_next(rng) * (real)2.3283064365386963e-10;  // 2^-32
}
// random.cc:53-59: phi first, then cos^2(theta).
HD real3 random_hemisphere_cosine(Rng *rng) {
    real phi = (2 * PI) * random_uniform(rng);
    real sqr_cos_theta = random_uniform(rng);
    real cos_theta = sqrt_(sqr_cos_theta);
    real sin_theta = sqrt_(1 - sqr_cos_theta);
    return {cos_(phi) * sin_theta, sin_(phi) * sin_theta, cos_theta};
}

// ---------------------------------------------------------------- hash
// sdbm fold over the object-index path, reference kernel/ocl/hash.cc:5-24.
struct Hasher { uint state; };
HD Hasher hash_init() { return {0u}; }
HD void hash_uint(Hasher *h, uint x) { h->state = x + (h->state << 6) + (h->state << 16) - h->state; }
HD void hash_usize(Hasher *h, usize x) { hash_uint(h, (uint)x); }
HD uint hash_finish(const Hasher *h) { return h->state; }
HD uint hash_never() { return ~0u; }

// ---------------------------------------------------------------- context
// kernel/ocl/render/context.hh:7-11, context.cc:11-13.
struct Context {
    Rng rng;
    Hasher hasher;
    uint prev_hash;
};
HD bool context_is_repeat(const Context *c) { return hash_finish(&c->hasher) == c->prev_hash; }

}  // namespace ht
