// Wavefront form of render_passes: the same per-pixel sample sequence, RNG
// stream and canvas additions (so the same canvas and seeds, bit for bit),
// split into kernels over queues of live paths instead of one thread
// looping over a pixel's samples.
//
// wf_generate  every pixel's first path (camera ray) into the SoA path
//              state; the extend queue of iteration 0 = all pixels.
// wf_extend    per queued path: Russian roulette (program option), the
//              nearest-hit scan, the sky on a miss. A hit's record goes to
//              the SoA hit store and the path to one of WF_KEYS shade queues
//              by the material component its next interaction will take
//              (the mixture's choice, read from a copy of the path's RNG, so
//              the interaction itself draws the same uniform later).
// wf_shade     the shade queues in key order (a wave runs one material): the
//              material interaction; a path that goes on re-enters the
//              extend queue, one that ends adds its radiance to the pixel and
//              starts the pixel's next sample (regeneration) until `passes`.
//
// Path state and hit records live in HBM as SoA: word k of path p's record
// at base[k * cap + p], cap = w * h, so every load and store of a wave is
// coalesced. The state is two records: the ray part (RNG, hasher, light),
// which the scan needs, and the accumulator part (emission, hop, samples
// done), which extend touches only when the path ends there. Queues are
// compacted per wavefront (ballot + lane prefix count, one atomic per wave
// and queue). A pixel has one live path at a time, so
// the queues never exceed cap and its canvas word has one writer.
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
    int hop;    // hits of the current path so far
    uint done;  // finished samples of the pixel
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
// Appends `value` for the lanes that `want`: one atomic per wave. Called by
// every lane of the wave (wave-uniform control flow). A queue holds at most
// one entry per pixel, so slots stay below `cap`; the bound is also checked.
HD void wf_push(bool want, uint *__restrict__ counter, uint *__restrict__ queue, uint value, uint cap) {
    unsigned long long m = __ballot(want);
    if (m == 0) return;
    uint leader = (uint)__ffsll((long long)m) - 1u;
    uint base = 0;
    if (__lane_id() == leader) base = atomicAdd(counter, (uint)__popcll(m));
    base = __shfl(base, (int)leader);
    uint slot = base + (uint)__popcll(m & lane_mask_below());
    if (want && slot < cap) queue[slot] = value;
}

// The path of pixel p has ended: its radiance joins the pixel (the canvas
// word holds the running sum, the order of render_pixel_passes' additions);
// the pixel's next sample starts (returns true) until `passes` are done,
// then its RNG state goes back to `seeds`.
template <class S>
HD bool wf_finish(uint w, uint h, uint p, const u8 *scene, float4 *__restrict__ canvas, uint *__restrict__ seeds,
                  uint passes, WfRay<S> *r, WfAcc *a) {
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
                         uint *__restrict__ seeds, uint passes, uint *__restrict__ rays, uint *__restrict__ accs,
                         uint *__restrict__ q0, uint *__restrict__ ctr) {
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
        Soa<WfRay<S>>::store(rays, n, idx, r);
        Soa<WfAcc>::store(accs, n, idx, a);
        q0[idx] = idx;
    }
}

template <class S>
HD void wf_extend_impl(uint w, uint h, const u8 *__restrict__ scene, float4 *__restrict__ canvas,
                       uint *__restrict__ seeds, uint passes, uint *__restrict__ rays, uint *__restrict__ accs,
                       uint *__restrict__ hits,
                       const uint *__restrict__ q_in, uint *__restrict__ q_next, uint *__restrict__ q_shade,
                       uint *__restrict__ ctr, uint parity) {
    typedef typename S::Object T;
    const u8 *obj = S::object(scene);
    uint cap = w * h;
    uint n = min(ctr[WF_CTR_EXT + parity], cap);
    if (blockIdx.x * blockDim.x >= n) return;  // whole block: nothing queued for it
    TopScan<T>::prepare(obj);
    __syncthreads();
    uint stride = gridDim.x * blockDim.x;
    for (uint i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
        uint i = i0 + __lane_id();
        uint p = 0u, key = WF_KEYS;
        bool next = false;
        if (i < n) p = q_in[i];
        if (i < n && p < cap) {
            WfRay<S> r = Soa<WfRay<S>>::load(rays, cap, p);
            bool end = false, miss = false;
#ifdef HT_RR_DEPTH
            if (Soa<WfAcc>::load(accs, cap, p).hop >= HT_RR_DEPTH && russian_roulette(&r.ctx, &r.light.base))
                end = true;
#endif
            if (!end) {
                r.ctx.hasher = hash_init();
                typename T::Cache cache;
                real dist = TopScan<T>::detect(obj, &r.ctx, &cache, &r.light);
                if (dist > (real)-0.5f) {
                    uint k = HitKey<T>::key(obj, &cache, r.ctx.rng);
                    key = k < WF_KEYS ? k : WF_KEYS - 1u;
                    Soa<typename T::Cache>::store(hits, cap, p, cache);
                } else {
                    miss = end = true;
                }
            }
            if (end) {
                WfAcc a = Soa<WfAcc>::load(accs, cap, p);
                if (miss) S::bg(scene, &r.ctx, &r.light, &a.emission);
                next = wf_finish<S>(w, h, p, scene, canvas, seeds, passes, &r, &a);
                Soa<WfAcc>::store(accs, cap, p, a);
            }
            Soa<WfRay<S>>::store(rays, cap, p, r);
        }
        wf_push(next, ctr + WF_CTR_EXT + (parity ^ 1u), q_next, p, cap);
#pragma unroll
        for (uint k = 0; k < WF_KEYS; ++k)
            wf_push(key == k, ctr + WF_CTR_SH + parity * WF_KEYS + k, q_shade + (size_t)k * cap, p, cap);
    }
}

template <class S>
HD void wf_shade_impl(uint w, uint h, const u8 *__restrict__ scene, float4 *__restrict__ canvas,
                      uint *__restrict__ seeds, uint passes, uint *__restrict__ rays, uint *__restrict__ accs,
                      const uint *__restrict__ hits, const uint *__restrict__ q_shade, uint *__restrict__ q_next,
                      uint *__restrict__ ctr, uint parity) {
    typedef typename S::Object T;
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
        uint p = 0u;
        bool next = false;
        if (i < total) {
            // entry i of the key-ordered concatenation of the shade queues
            uint k = 0u, j = i;
            while (k + 1u < WF_KEYS && j >= counts[k]) {
                j -= counts[k];
                ++k;
            }
            p = q_shade[(size_t)k * cap + j];
        }
        if (i < total && p < cap) {
            WfRay<S> r = Soa<WfRay<S>>::load(rays, cap, p);
            WfAcc a = Soa<WfAcc>::load(accs, cap, p);
            typename T::Cache cache = Soa<typename T::Cache>::load(hits, cap, p);
            bool end = !T::interact(obj, &r.ctx, &cache, &r.light, &a.emission);
            if (!end) r.ctx.prev_hash = hash_finish(&r.ctx.hasher);
            a.hop += 1;
            if (a.hop >= S::HOPS) end = true;
            next = end ? wf_finish<S>(w, h, p, scene, canvas, seeds, passes, &r, &a) : true;
            Soa<WfRay<S>>::store(rays, cap, p, r);
            Soa<WfAcc>::store(accs, cap, p, a);
        }
        wf_push(next, ctr + WF_CTR_EXT + (parity ^ 1u), q_next, p, cap);
    }
}

}  // namespace ht

// Kernel entry points of the wavefront form for the scene type `S`, and the
// record sizes the runtime allocates the SoA stores by: {ray words,
// accumulator words, hit words, shade keys, counter words, HOPS}.
#define HT_DEFINE_WAVEFRONT_KERNELS(S)                                                                          \
    extern "C" __device__ unsigned ht_wf_layout[6] = {                                                          \
        ht::Soa<ht::WfRay<S>>::WORDS, ht::Soa<ht::WfAcc>::WORDS, ht::Soa<typename S::Object::Cache>::WORDS,     \
        ht::WF_KEYS, ht::WF_CTR_WORDS, (unsigned)S::HOPS};                                                      \
    extern "C" __global__ void __launch_bounds__(256) wf_generate(                                              \
        unsigned w, unsigned h, const unsigned char *__restrict__ scene, float4 *__restrict__ canvas,           \
        unsigned *__restrict__ seeds, unsigned passes, unsigned *__restrict__ rays, unsigned *__restrict__ accs, \
        unsigned *__restrict__ q0, unsigned *__restrict__ ctr) {                                                \
        ht::wf_generate_impl<S>(w, h, scene, canvas, seeds, passes, rays, accs, q0, ctr);                       \
    }                                                                                                           \
    extern "C" __global__ void __launch_bounds__(256, HT_WAVES_PER_EU) wf_extend(                               \
        unsigned w, unsigned h, const unsigned char *__restrict__ scene, float4 *__restrict__ canvas,           \
        unsigned *__restrict__ seeds, unsigned passes, unsigned *__restrict__ rays, unsigned *__restrict__ accs, \
        unsigned *__restrict__ hits, const unsigned *__restrict__ q_in, unsigned *__restrict__ q_next,          \
        unsigned *__restrict__ q_shade, unsigned *__restrict__ ctr, unsigned parity) {                          \
        ht::wf_extend_impl<S>(w, h, scene, canvas, seeds, passes, rays, accs, hits, q_in, q_next, q_shade, ctr, \
                              parity);                                                                          \
    }                                                                                                           \
    extern "C" __global__ void __launch_bounds__(256, HT_WAVES_PER_EU) wf_shade(                                \
        unsigned w, unsigned h, const unsigned char *__restrict__ scene, float4 *__restrict__ canvas,           \
        unsigned *__restrict__ seeds, unsigned passes, unsigned *__restrict__ rays, unsigned *__restrict__ accs, \
        const unsigned *__restrict__ hits, const unsigned *__restrict__ q_shade, unsigned *__restrict__ q_next, \
        unsigned *__restrict__ ctr, unsigned parity) {                                                          \
        ht::wf_shade_impl<S>(w, h, scene, canvas, seeds, passes, rays, accs, hits, q_shade, q_next, ctr,        \
                             parity);                                                                           \
    }
