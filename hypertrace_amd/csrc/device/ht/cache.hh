// Path tracing into a learned radiance cache (`render_cached`).
//
// A path runs the reference bounce loop (render/scene.inl:16-44) up to its
// `query_hop`-th hit (0 = the camera ray's first hit). There it either
//  * ends into the cache: the hit's 12 features and the path's throughput
//    are appended to a query queue; the cache's predicted outgoing radiance,
//    times that throughput, is added to the pixel by the MLP inference pass;
//  * or, for one pixel in `train_every` per pass (a hash of pixel and pass,
//    no RNG draw), continues unbiased to full depth. Its full-path radiance
//    goes to the canvas, and (features, radiance gathered after the hit /
//    throughput at the hit) is appended to the training queue.
// Paths that escape or are absorbed before the query hit end as usual.
//
// Features of a hit (all bounded): position (t/(1+|t|) of x, y and of z in
// Euclidean space or of ln z in the half-space), the ray's unit direction
// at the hit, the unit normal facing the ray, the material's mean albedo
// (mixture portions x colours).
//
// Queues are compacted per wavefront: one atomic per wave on the queue's
// counter (ballot + lane prefix count), SoA records.
#pragma once
#include "render.hh"

namespace ht {

static constexpr int CACHE_FEATURES = 12;

// ------------------------------------------------------------ mean albedo
template <class M> struct MatAlbedo {
    static HD color3 of(const u8 *) { return c3(1.0f); }
};
template <> struct MatAlbedo<Absorbing> {
    static HD color3 of(const u8 *) { return c3(0.0f); }
};
template <class M> struct MatAlbedo<Colored<M>> {
    typedef typename Colored<M>::L L;
    static HD color3 of(const u8 *s) { return ld_color3(s + L::offset(0)) * MatAlbedo<M>::of(s + L::offset(1)); }
};
template <class M> struct MatAlbedo<Emissive<M>> {
    typedef typename Emissive<M>::L L;
    static HD color3 of(const u8 *s) { return MatAlbedo<M>::of(s + L::offset(1)); }
};
template <class... Ms> struct MatAlbedo<Mixture<Ms...>> {
    typedef typename Mixture<Ms...>::L L;
    template <size_t_ I, class M, class... Rest> static HD color3 sum(const u8 *s) {
        const u8 *comp = s + L::offset(I);
        color3 a = MatAlbedo<M>::of(comp + Component<M>::L::offset(0)) *
                   (float)ld<real>(comp + Component<M>::L::offset(1));
        if constexpr (sizeof...(Rest) > 0) {
            return a + sum<I + 1, Rest...>(s);
        } else {
            return a;
        }
    }
    static HD color3 of(const u8 *s) { return sum<0, Ms...>(s); }
};

// ------------------------------------------------------------ hit normal and albedo
// World-space normal and mean albedo of the hit a detect left in `cache`.
template <class T> struct HitInfo {
    static constexpr bool OK = false;
};
template <class G, class S, class Mat> struct HitInfo<Covered<G, S, Mat>> {
    static constexpr bool OK = true;
    typedef Covered<G, S, Mat> C;
    static HD void get(const u8 *self, const typename C::Cache *cache, typename G::Pos, typename G::Dir *normal,
                       color3 *albedo) {
        *normal = cache->normal;
        *albedo = MatAlbedo<Mat>::of(self + C::L::offset(0));
    }
};
template <class G, class TI, class M> struct HitInfo<Mapped<G, TI, M>> {
    static constexpr bool OK = HitInfo<TI>::OK;
    typedef Mapped<G, TI, M> MP;
    static HD void get(const u8 *self, const typename MP::Cache *cache, typename G::Pos world,
                       typename G::Dir *normal, color3 *albedo) {
        M m = MP::map(self);
        typename G::Pos local = m.inverse().apply_pos(world);
        typename G::Dir n;
        HitInfo<TI>::get(MP::inner(self), cache, local, &n, albedo);
        *normal = m.apply_normal(local, n);
    }
};
template <class G, class E> struct HitInfo<ObjectVec<G, E>> {
    static constexpr bool OK = HitInfo<E>::OK;
    typedef ObjectVec<G, E> OV;
    static HD void get(const u8 *self, const typename OV::Cache *cache, typename G::Pos world,
                       typename G::Dir *normal, color3 *albedo) {
        HitInfo<E>::get(OV::V::elem(self, cache->index), &cache->inner, world, normal, albedo);
    }
};

// ------------------------------------------------------------ features
HD float squash(float t) { return t / (1.0f + fabs_(t)); }

template <class G> struct CacheGeo;
template <> struct CacheGeo<Hy> {
    static HD void pos(quat p, float *f) {
        f[0] = squash((float)p.x);
        f[1] = squash((float)p.y);
        f[2] = squash((float)log_(p.z));
    }
    static HD real3 dir3(quat d) { return xyz(d); }
};
template <> struct CacheGeo<Eu> {
    static HD void pos(real3 p, float *f) {
        f[0] = squash((float)p.x);
        f[1] = squash((float)p.y);
        f[2] = squash((float)p.z);
    }
    static HD real3 dir3(real3 d) { return d; }
};

template <class S>
HD void cache_features(const u8 *obj, const typename S::Object::Cache *cache, const Light<typename S::Geo> *light,
                       float *f) {
    typedef typename S::Geo G;
    typename G::Dir n;
    color3 albedo;
    HitInfo<typename S::Object>::get(obj, cache, light->ray.start, &n, &albedo);
    real3 d = CacheGeo<G>::dir3(light->ray.direction);
    real3 nn = normalize(CacheGeo<G>::dir3(n));
    if (dot(nn, d) > R0) nn = -nn;
    CacheGeo<G>::pos(light->ray.start, f);
    f[3] = (float)d.x;
    f[4] = (float)d.y;
    f[5] = (float)d.z;
    f[6] = (float)nn.x;
    f[7] = (float)nn.y;
    f[8] = (float)nn.z;
    f[9] = albedo.x;
    f[10] = albedo.y;
    f[11] = albedo.z;
}

HD uint mix32(uint x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Appends this lane's record (if `want`) to a queue: returns the slot, or
// ~0u when the lane has none or the queue is full. One atomic per wave.
HD uint queue_slot(bool want, uint *counter, uint cap) {
    unsigned long long m = __ballot(want);
    if (m == 0) return ~0u;
    uint leader = (uint)__ffsll((long long)m) - 1u;
    uint base = 0;
    if (__lane_id() == leader) base = atomicAdd(counter, (uint)__popcll(m));
    base = __shfl(base, (int)leader);
    if (!want) return ~0u;
    uint slot = base + (uint)__popcll(m & lane_mask_below());
    return slot < cap ? slot : ~0u;
}

template <class S>
HD void render_cached_impl(uint w, uint h, const u8 *__restrict__ scene, float4 *__restrict__ canvas,
                           uint *__restrict__ seeds, uint pass, uint *__restrict__ counters,
                           uint *__restrict__ q_pixel, float *__restrict__ q_feat, float *__restrict__ q_thr,
                           uint qcap, float *__restrict__ t_feat, float *__restrict__ t_target, uint tcap,
                           uint train_every, uint query_hop) {
    typedef typename S::Geo G;
    typedef typename S::Object T;
    static_assert(HitInfo<T>::OK, "the radiance cache needs a vector of mapped covered objects");
    TopScan<T>::prepare(S::object(scene));
    __syncthreads();
    uint idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= w * h) return;
    uint x = idx % w, y = idx / w;
    PixelSetup px = pixel_setup(x, y, w, h);
    Context ctx;
    ctx.rng.state = seeds[idx];
    ctx.hasher = hash_init();
    Light<G> light;
    color3 emission;
    start_path<S>(scene, &ctx, px, &light, &emission);
    bool train = train_every <= 1u || mix32(idx * 0x9E3779B9u ^ (pass + 1u) * 0x85EBCA6Bu) % train_every == 0u;
    bool query = false, record = false;
    float feat[CACHE_FEATURES];
    color3 thr_q = c3(0.0f), em_q = c3(0.0f);
    const u8 *obj = S::object(scene);
    for (int i = 0; i < S::HOPS; ++i) {
#ifdef HT_RR_DEPTH
        // the roulette of bounce() (render.hh), at the same point of the hop
        if (i >= HT_RR_DEPTH && russian_roulette(&ctx, &light.base)) break;
#endif
        ctx.hasher = hash_init();
        typename T::Cache cache;
        real dist = TopScan<T>::detect(obj, &ctx, &cache, &light);
        if (!(dist > (real)-0.5f)) {
            S::bg(scene, &ctx, &light, &emission);
            break;
        }
        if (i == (int)query_hop) {
            cache_features<S>(obj, &cache, &light, feat);
            thr_q = light.base.intensity;
            em_q = emission;
            if (!train) {
                query = true;
                break;
            }
            record = true;
        }
        if (!T::interact(obj, &ctx, &cache, &light, &emission)) break;
        ctx.prev_hash = hash_finish(&ctx.hasher);
    }
    float4 acc = canvas[idx];
    acc.x += emission.x;
    acc.y += emission.y;
    acc.z += emission.z;
    acc.w += 1.0f;
    canvas[idx] = acc;
    seeds[idx] = ctx.rng.state;

    uint qs = queue_slot(query, counters + 0, qcap);
    if (qs != ~0u) {
        q_pixel[qs] = idx;
        for (int f = 0; f < CACHE_FEATURES; ++f) q_feat[(size_t)f * qcap + qs] = feat[f];
        q_thr[qs] = thr_q.x;
        q_thr[(size_t)qcap + qs] = thr_q.y;
        q_thr[2 * (size_t)qcap + qs] = thr_q.z;
    }
    uint ts = queue_slot(record, counters + 1, tcap);
    if (ts != ~0u) {
        color3 after = emission - em_q;
        for (int f = 0; f < CACHE_FEATURES; ++f) t_feat[(size_t)f * tcap + ts] = feat[f];
        t_target[ts] = thr_q.x > 1e-20f ? after.x / thr_q.x : 0.0f;
        t_target[(size_t)tcap + ts] = thr_q.y > 1e-20f ? after.y / thr_q.y : 0.0f;
        t_target[2 * (size_t)tcap + ts] = thr_q.z > 1e-20f ? after.z / thr_q.z : 0.0f;
    }
}

}  // namespace ht

#define HT_DEFINE_CACHED_KERNEL(S)                                                                             \
    extern "C" __global__ void __launch_bounds__(256, HT_WAVES_PER_EU) render_cached(                          \
        unsigned w, unsigned h, const unsigned char *__restrict__ scene, float4 *__restrict__ canvas,          \
        unsigned *__restrict__ seeds, unsigned pass, unsigned *__restrict__ counters,                          \
        unsigned *__restrict__ q_pixel, float *__restrict__ q_feat, float *__restrict__ q_thr, unsigned qcap,  \
        float *__restrict__ t_feat, float *__restrict__ t_target, unsigned tcap, unsigned train_every,         \
        unsigned query_hop) {                                                                                  \
        ht::render_cached_impl<S>(w, h, scene, canvas, seeds, pass, counters, q_pixel, q_feat, q_thr, qcap,    \
                                  t_feat, t_target, tcap, train_every, query_hop);                             \
    }
