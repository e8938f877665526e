// Wavefront form of render_passes: the same per-pixel sample sequence, RNG
// stream and canvas additions (so the same canvas and seeds, bit for bit),
// split into kernels over queues of live paths instead of one thread
// looping over a pixel's samples.
//
// wf_generate  every pixel's first path (camera ray): extend queue 0 holds
//              all pixels.
// wf_extend    per queued path: Russian roulette (program option), the
//              nearest-hit scan, the sky on a miss. A hit moves the path,
//              with its hit record, into one of WF_KEYS shade queues by the
//              material component its next interaction will take (the
//              mixture's choice, read from a copy of the path's RNG, so the
//              interaction itself draws the same uniform later).
// wf_shade     the shade queues in key order (a wave runs one material): the
//              material interaction; a path that goes on moves to the next
//              extend queue, one that ends adds its radiance to its pixel and
//              starts the pixel's next sample (regeneration) until `passes`.
//
// SoA queues: a queue holds the records of its paths, word k of the record
// in slot i at base[k * cap + i] (cap = w * h; a pixel has one live path at
// a time, so no queue holds more). A kernel reads its input queue's slots in
// order and writes each surviving path's record into the slot the output
// queue assigns it; slots are handed out per wavefront (ballot + lane
// prefix count, one atomic per wave and queue), so a wave's 64 records land
// in consecutive slots and every load and store is coalesced. Extend-queue
// records: the ray part (RNG, hasher, light) and the accumulator part
// (emission, hop, samples done, pixel); shade-queue records add the hit.
//
// Counters (ctr): extend-queue sizes [WF_CTR_EXT + parity] and shade-queue
// sizes [WF_CTR_SH + parity * WF_KEYS + key], double-buffered by iteration
// parity. Iteration i's extend reads the extend size of parity i and its
// shade the shade sizes of parity i; both append to the extend queue of
// parity i + 1. wf_shade clears the two counters that iteration i + 1 will
// append to (neither is read or written by any other block of the running
// kernel), so the sequence needs no memset.
#pragma once
#include "render.hh"

namespace ht {

static constexpr uint WF_KEYS = 4;
static constexpr uint WF_CTR_EXT = 0, WF_CTR_SH = 2, WF_CTR_WORDS = 2 + 2 * WF_KEYS;

// Record i of a word-transposed SoA array of T.
template <class T> struct Soa {
    static constexpr uint WORDS = (uint)((sizeof(T) + 3) / 4);
    static HD T load(const uint *__restrict__ base, uint cap, uint i) {
        uint w[WORDS];
#pragma unroll
        for (uint k = 0; k < WORDS; ++k) w[k] = base[(size_t)k * cap + i];
        T v;
        __builtin_memcpy(&v, w, sizeof(T));
        return v;
    }
    static HD void store(uint *__restrict__ base, uint cap, uint i, const T &v) {
        uint w[WORDS];
        w[WORDS - 1] = 0u;
        __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
        for (uint k = 0; k < WORDS; ++k) base[(size_t)k * cap + i] = w[k];
    }
};

template <class S> struct WfRay {
    Context ctx;
    Light<typename S::Geo> light;
};
struct WfAcc {
    color3 emission;
    int hop;     // hits of the current path so far
    uint done;   // finished samples of the pixel
    uint pixel;  // index into canvas and seeds
};

// The records of one queue: ray part, accumulator part and (shade queues)
// the hit, one after the other, each `cap` records of its word count.
template <class S> struct WfQueue {
    typedef typename S::Object::Cache Hit;
    static constexpr uint RAY = Soa<WfRay<S>>::WORDS, ACC = Soa<WfAcc>::WORDS, HIT = Soa<Hit>::WORDS;
    static constexpr uint EXT_WORDS = RAY + ACC, SHADE_WORDS = RAY + ACC + HIT;
    static HD WfRay<S> ray(const uint *q, uint cap, uint i) { return Soa<WfRay<S>>::load(q, cap, i); }
    static HD WfAcc acc(const uint *q, uint cap, uint i) { return Soa<WfAcc>::load(q + (size_t)RAY * cap, cap, i); }
    static HD Hit hit(const uint *q, uint cap, uint i) {
        return Soa<Hit>::load(q + (size_t)EXT_WORDS * cap, cap, i);
    }
    static HD void put(uint *q, uint cap, uint i, const WfRay<S> &r, const WfAcc &a) {
        Soa<WfRay<S>>::store(q, cap, i, r);
        Soa<WfAcc>::store(q + (size_t)RAY * cap, cap, i, a);
    }
    static HD void put_hit(uint *q, uint cap, uint i, const Hit &h) {
        Soa<Hit>::store(q + (size_t)EXT_WORDS * cap, cap, i, h);
    }
};

// ------------------------------------------------------------ material keys
// Component a material's next interaction takes, without drawing: the
// mixture's uniform is drawn from a copy of the RNG (Mixture::interact,
// material/mixture/impl.inl:15-27, draws it first). Other materials: 0.
template <class M> struct MatKey {
    static HD uint key(const u8 *, Rng) { return 0u; }
};
template <class M> struct MatKey<Colored<M>> {
    static HD uint key(const u8 *s, Rng r) { return MatKey<M>::key(s + Colored<M>::L::offset(1), r); }
};
template <class M> struct MatKey<Emissive<M>> {
    static HD uint key(const u8 *s, Rng r) { return MatKey<M>::key(s + Emissive<M>::L::offset(1), r); }
};
template <class... Ms> struct MatKey<Mixture<Ms...>> {
    typedef typename Mixture<Ms...>::L L;
    template <size_t_ I, class M, class... Rest> static HD uint pick(const u8 *s, real alpha) {
        alpha -= ld<real>(s + L::offset(I) + Component<M>::L::offset(1));
        if (alpha < R0) return (uint)I;
        if constexpr (sizeof...(Rest) > 0) {
            return pick<I + 1, Rest...>(s, alpha);
        } else {
            return (uint)(I + 1);  // no component: the interaction absorbs
        }
    }
    static HD uint key(const u8 *s, Rng r) {
        real alpha = random_uniform(&r);
        return pick<0, Ms...>(s, alpha);
    }
};

// The material of the hit a detect left in `cache`.
template <class T> struct HitKey {
    template <class C> static HD uint key(const u8 *, const C *, Rng) { return 0u; }
};
template <class G, class S, class Mat> struct HitKey<Covered<G, S, Mat>> {
    template <class C> static HD uint key(const u8 *self, const C *, Rng r) {
        return MatKey<Mat>::key(self + Covered<G, S, Mat>::L::offset(0), r);
    }
};
template <class G, class TI, class M> struct HitKey<Mapped<G, TI, M>> {
    template <class C> static HD uint key(const u8 *self, const C *cache, Rng r) {
        return HitKey<TI>::key(Mapped<G, TI, M>::inner(self), cache, r);
    }
};
template <class G, class E> struct HitKey<ObjectVec<G, E>> {
    template <class C> static HD uint key(const u8 *self, const C *cache, Rng r) {
        return HitKey<E>::key(ObjectVec<G, E>::V::elem(self, cache->index), &cache->inner, r);
    }
};

// ------------------------------------------------------------ helpers
// Slot for this lane's record (if `want`) in a queue: one atomic per wave.
// Called by every lane of the wave (wave-uniform control flow). Returns ~0u
// when the lane has no record (or, which the one-path-per-pixel rule
// excludes, the queue is full).
HD uint wf_slot(bool want, uint *__restrict__ counter, uint cap) {
    unsigned long long m = __ballot(want);
    if (m == 0) return ~0u;
    uint leader = (uint)__ffsll((long long)m) - 1u;
    uint base = 0;
    if (__lane_id() == leader) base = atomicAdd(counter, (uint)__popcll(m));
    base = __shfl(base, (int)leader);
    uint slot = base + (uint)__popcll(m & lane_mask_below());
    return want && slot < cap ? slot : ~0u;
}

// The path has ended: its radiance joins its pixel (the canvas word holds
// the running sum, the order of render_pixel_passes' additions); the pixel's
// next sample starts (returns true) until `passes` are done, then its RNG
// state goes back to `seeds`.
template <class S>
HD bool wf_finish(uint w, uint h, const u8 *scene, float4 *__restrict__ canvas, uint *__restrict__ seeds,
                  uint passes, WfRay<S> *r, WfAcc *a) {
    uint p = a->pixel;
    float4 acc = canvas[p];
    acc.x += a->emission.x;
    acc.y += a->emission.y;
    acc.z += a->emission.z;
    acc.w += 1.0f;
    canvas[p] = acc;
    a->done += 1u;
    if (a->done < passes) {
        PixelSetup px = pixel_setup(p % w, p / w, w, h);
        start_path<S>(scene, &r->ctx, px, &r->light, &a->emission);
        a->hop = 0;
        return true;
    }
    seeds[p] = r->ctx.rng.state;
    return false;
}

// ------------------------------------------------------------ kernels
template <class S>
HD void wf_generate_impl(uint w, uint h, const u8 *__restrict__ scene, float4 *__restrict__ canvas,
                         uint *__restrict__ seeds, uint passes, uint *__restrict__ q0, uint *__restrict__ ctr) {
    uint n = w * h;
    uint idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx == 0) {
        for (uint c = 0; c < WF_CTR_WORDS; ++c) ctr[c] = 0u;
        ctr[WF_CTR_EXT] = (S::HOPS > 0 && passes > 0) ? n : 0u;
    }
    if (idx >= n || passes == 0) return;
    PixelSetup px = pixel_setup(idx % w, idx / w, w, h);
    WfRay<S> r;
    WfAcc a;
    r.ctx.rng.state = seeds[idx];
    r.ctx.hasher = hash_init();
    if constexpr (S::HOPS <= 0) {
        // no hits: every sample is the view's draw only (render_pixel_passes)
        float4 acc = canvas[idx];
        for (uint k = 0; k < passes; ++k) {
            start_path<S>(scene, &r.ctx, px, &r.light, &a.emission);
            acc.w += 1.0f;
        }
        canvas[idx] = acc;
        seeds[idx] = r.ctx.rng.state;
    } else {
        start_path<S>(scene, &r.ctx, px, &r.light, &a.emission);
        a.hop = 0;
        a.done = 0u;
        a.pixel = idx;
        WfQueue<S>::put(q0, n, idx, r, a);
    }
}

template <class S>
HD void wf_extend_impl(uint w, uint h, const u8 *__restrict__ scene, float4 *__restrict__ canvas,
                       uint *__restrict__ seeds, uint passes, const uint *__restrict__ q_in,
                       uint *__restrict__ q_next, uint *__restrict__ q_shade, uint *__restrict__ ctr, uint parity) {
    typedef typename S::Object T;
    typedef WfQueue<S> Q;
    const u8 *obj = S::object(scene);
    uint cap = w * h;
    uint n = min(ctr[WF_CTR_EXT + parity], cap);
    if (blockIdx.x * blockDim.x >= n) return;  // whole block: nothing queued for it
    TopScan<T>::prepare(obj);
    __syncthreads();
    uint stride = gridDim.x * blockDim.x;
    for (uint i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
        uint i = i0 + __lane_id();
        uint key = WF_KEYS;
        bool next = false;
        WfRay<S> r;
        WfAcc a;
        typename T::Cache cache;
        if (i < n) {
            r = Q::ray(q_in, cap, i);
            a = Q::acc(q_in, cap, i);
            bool end = false, miss = false;
#ifdef HT_RR_DEPTH
            if (a.hop >= HT_RR_DEPTH && russian_roulette(&r.ctx, &r.light.base)) end = true;
#endif
            if (!end) {
                r.ctx.hasher = hash_init();
                real dist = TopScan<T>::detect(obj, &r.ctx, &cache, &r.light);
                if (dist > (real)-0.5f) {
                    uint k = HitKey<T>::key(obj, &cache, r.ctx.rng);
                    key = k < WF_KEYS ? k : WF_KEYS - 1u;
                } else {
                    miss = end = true;
                }
            }
            if (end) {
                if (miss) S::bg(scene, &r.ctx, &r.light, &a.emission);
                next = wf_finish<S>(w, h, scene, canvas, seeds, passes, &r, &a);
            }
        }
        uint slot = wf_slot(next, ctr + WF_CTR_EXT + (parity ^ 1u), cap);
        if (slot != ~0u) Q::put(q_next, cap, slot, r, a);
#pragma unroll
        for (uint k = 0; k < WF_KEYS; ++k) {
            uint s = wf_slot(key == k, ctr + WF_CTR_SH + parity * WF_KEYS + k, cap);
            if (s != ~0u) {
                uint *q = q_shade + (size_t)k * Q::SHADE_WORDS * cap;
                Q::put(q, cap, s, r, a);
                Q::put_hit(q, cap, s, cache);
            }
        }
    }
}

template <class S>
HD void wf_shade_impl(uint w, uint h, const u8 *__restrict__ scene, float4 *__restrict__ canvas,
                      uint *__restrict__ seeds, uint passes, const uint *__restrict__ q_shade,
                      uint *__restrict__ q_next, uint *__restrict__ ctr, uint parity) {
    typedef typename S::Object T;
    typedef WfQueue<S> Q;
    const u8 *obj = S::object(scene);
    uint cap = w * h;
    uint counts[WF_KEYS], total = 0u;
#pragma unroll
    for (uint k = 0; k < WF_KEYS; ++k) {
        counts[k] = min(ctr[WF_CTR_SH + parity * WF_KEYS + k], cap);
        total += counts[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the counters iteration i + 1 appends to (see the header comment)
        ctr[WF_CTR_EXT + parity] = 0u;
#pragma unroll
        for (uint k = 0; k < WF_KEYS; ++k) ctr[WF_CTR_SH + (parity ^ 1u) * WF_KEYS + k] = 0u;
    }
    uint stride = gridDim.x * blockDim.x;
    for (uint i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < total; i0 += stride) {
        uint i = i0 + __lane_id();
        bool next = false;
        WfRay<S> r;
        WfAcc a;
        if (i < total) {
            // entry i of the key-ordered concatenation of the shade queues
            uint k = 0u, j = i;
            while (k + 1u < WF_KEYS && j >= counts[k]) {
                j -= counts[k];
                ++k;
            }
            const uint *q = q_shade + (size_t)k * Q::SHADE_WORDS * cap;
            r = Q::ray(q, cap, j);
            a = Q::acc(q, cap, j);
            typename T::Cache cache = Q::hit(q, cap, j);
            bool end = !T::interact(obj, &r.ctx, &cache, &r.light, &a.emission);
            if (!end) r.ctx.prev_hash = hash_finish(&r.ctx.hasher);
            a.hop += 1;
            if (a.hop >= S::HOPS) end = true;
            next = end ? wf_finish<S>(w, h, scene, canvas, seeds, passes, &r, &a) : true;
        }
        uint slot = wf_slot(next, ctr + WF_CTR_EXT + (parity ^ 1u), cap);
        if (slot != ~0u) Q::put(q_next, cap, slot, r, a);
    }
}

}  // namespace ht

// Kernel entry points of the wavefront form for the scene type `S`, and the
// queue sizes the runtime allocates by: {extend-record words, shade-record
// words, shade keys, counter words, HOPS}.
#define HT_DEFINE_WAVEFRONT_KERNELS(S)                                                                          \
    extern "C" __device__ unsigned ht_wf_layout[5] = {ht::WfQueue<S>::EXT_WORDS, ht::WfQueue<S>::SHADE_WORDS,   \
                                                      ht::WF_KEYS, ht::WF_CTR_WORDS, (unsigned)S::HOPS};        \
    extern "C" __global__ void __launch_bounds__(256) wf_generate(                                              \
        unsigned w, unsigned h, const unsigned char *__restrict__ scene, float4 *__restrict__ canvas,           \
        unsigned *__restrict__ seeds, unsigned passes, unsigned *__restrict__ q0, unsigned *__restrict__ ctr) { \
        ht::wf_generate_impl<S>(w, h, scene, canvas, seeds, passes, q0, ctr);                                   \
    }                                                                                                           \
    extern "C" __global__ void __launch_bounds__(256, HT_WAVES_PER_EU) wf_extend(                               \
        unsigned w, unsigned h, const unsigned char *__restrict__ scene, float4 *__restrict__ canvas,           \
        unsigned *__restrict__ seeds, unsigned passes, const unsigned *__restrict__ q_in,                       \
        unsigned *__restrict__ q_next, unsigned *__restrict__ q_shade, unsigned *__restrict__ ctr,              \
        unsigned parity) {                                                                                      \
        ht::wf_extend_impl<S>(w, h, scene, canvas, seeds, passes, q_in, q_next, q_shade, ctr, parity);          \
    }                                                                                                           \
    extern "C" __global__ void __launch_bounds__(256, HT_WAVES_PER_EU) wf_shade(                                \
        unsigned w, unsigned h, const unsigned char *__restrict__ scene, float4 *__restrict__ canvas,           \
        unsigned *__restrict__ seeds, unsigned passes, const unsigned *__restrict__ q_shade,                    \
        unsigned *__restrict__ q_next, unsigned *__restrict__ ctr, unsigned parity) {                           \
        ht::wf_shade_impl<S>(w, h, scene, canvas, seeds, passes, q_shade, q_next, ctr, parity);                 \
    }
