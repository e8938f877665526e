// Render kernels for one scene type.
//
// render       one pass = one path per pixel, accumulated into the float4
//              canvas; the drop-in for the reference kernel
//              `render(uint2 shape, Scene*, float4* canvas, uint* seeds)`
//              (kernel/ocl/render/render.cl:6-27).
// render_passes  `passes` consecutive passes in one launch. Each thread owns a
//              pixel and, when its path ends, starts the pixel's next sample
//              in place (path regeneration), so the wavefront keeps running the
//              nearest-hit scan on full lanes. Per pixel the sample sequence,
//              the RNG stream and the order of the canvas additions are the
//              same as `passes` calls of `render`, so the canvas and seeds come
//              out identical.
// Both run the bounce loop of render/scene.inl:16-44; the top-level object's
// nearest-hit scan goes through TopScan (scan.hh), which culls and compacts
// the exact tests where the scene allows it without changing any result.
#pragma once
#include "scan.hh"

namespace ht {

struct PixelSetup {
    real2 pos;
    real2 size;
};

// render.cl:12-18
HD PixelSetup pixel_setup(uint x, uint y, uint w, uint h) {
    PixelSetup p;
    p.pos = r2((real)(2 * x + 1) - (real)w, (real)h - (real)(2 * y + 1)) / (real)h;
    real s = (real)2 / (real)h;
    p.size = r2(s, s);
    return p;
}

// Russian roulette (program option HT_RR_DEPTH = d; not in the reference):
// from the d-th hit on (hop >= d), before the nearest-hit search, the path
// survives with probability q = min(max(throughput), 0.95) (one uniform) and
// its throughput is divided by q. Returns true when the path ends.
HD bool russian_roulette(Context *ctx, LightBase *base) {
    color3 I = base->intensity;
    float q = fminf(fmaxf(fmaxf(I.x, I.y), I.z), 0.95f);
    real u = random_uniform(&ctx->rng);
    if (u >= (real)q) return true;
    base->intensity = I / q;
    return false;
}

// One bounce of scene.inl:25-41 (`hop` = hits so far). Returns true when the
// path ends here.
template <class S>
HD bool bounce(const u8 *scene, Context *ctx, Light<typename S::Geo> *light, color3 *emission, int hop) {
    typedef typename S::Object T;
#ifdef HT_RR_DEPTH
    if (hop >= HT_RR_DEPTH && russian_roulette(ctx, &light->base)) return true;
#else
    (void)hop;
#endif
    const u8 *obj = S::object(scene);
    ctx->hasher = hash_init();
    typename T::Cache cache;
    real dist = TopScan<T>::detect(obj, ctx, &cache, light);
    if (dist > (real)-0.5f) {
        bool ret = T::interact(obj, ctx, &cache, light, emission);
        if (!ret) return true;
        ctx->prev_hash = hash_finish(&ctx->hasher);
        return false;
    }
    S::bg(scene, ctx, light, emission);
    return true;
}

template <class S>
HD void start_path(const u8 *scene, Context *ctx, const PixelSetup &px, Light<typename S::Geo> *light,
                   color3 *emission) {
    light->base.intensity = c3(1.0f);
    light->base.diffuse = false;
    light->ray = S::View::sample(S::view(scene), ctx, px.pos, px.size);
    *emission = c3(0.0f);
    ctx->prev_hash = hash_never();
}

template <class S>
HD void render_pixel_passes(uint w, uint h, uint idx, const u8 *__restrict__ scene, float4 *__restrict__ canvas,
                            uint *__restrict__ seeds, uint passes, int hops);

template <class S>
HD void render_pass_impl(uint w, uint h, const u8 *__restrict__ scene, float4 *__restrict__ canvas,
                         uint *__restrict__ seeds) {
    TopScan<typename S::Object>::prepare(S::object(scene));
    __syncthreads();
    uint idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= w * h) return;
#ifdef HT_HELPERS
    // one sample of render_pixel_passes: the same additions in the same
    // order, with the lanes whose path has ended helping the wave's scans
    render_pixel_passes<S>(w, h, idx, scene, canvas, seeds, 1u, S::HOPS);
#else
    uint x = idx % w, y = idx / w;
    PixelSetup px = pixel_setup(x, y, w, h);
    Context ctx;
    ctx.rng.state = seeds[idx];
    ctx.hasher = hash_init();
    Light<typename S::Geo> light;
    color3 emission;
    start_path<S>(scene, &ctx, px, &light, &emission);
    for (int i = 0; i < S::HOPS; ++i) {
        if (bounce<S>(scene, &ctx, &light, &emission, i)) break;
    }
    float4 acc = canvas[idx];
    acc.x += emission.x;
    acc.y += emission.y;
    acc.z += emission.z;
    acc.w += 1.0f;
    canvas[idx] = acc;
    seeds[idx] = ctx.rng.state;
#endif
}

// `passes` samples of pixel idx with in-thread path regeneration, each path
// at most `hops` hits (<= S::HOPS). The caller has prepared the scan.
template <class S>
HD void render_pixel_passes(uint w, uint h, uint idx, const u8 *__restrict__ scene, float4 *__restrict__ canvas,
                            uint *__restrict__ seeds, uint passes, int hops) {
    typedef typename S::Geo G;
    uint x = idx % w, y = idx / w;
    PixelSetup px = pixel_setup(x, y, w, h);
    Context ctx;
    ctx.rng.state = seeds[idx];
    ctx.hasher = hash_init();
    float4 acc = canvas[idx];

    Light<G> light;
    color3 emission = c3(0.0f);
    if (hops <= 0) {
        for (uint k = 0; k < passes; ++k) {
            start_path<S>(scene, &ctx, px, &light, &emission);
            acc.w += 1.0f;
        }
    } else {
        int hop = 0;
        uint done = 0;
        bool start = true;
#ifdef HT_WAVE_PRIO
        // Program option HT_WAVE_PRIO: the wave's issue priority steps down
        // from 3 as its slowest lane completes each quarter of the passes, so
        // the waves with the most work left (the frame's critical path) are
        // issued first on their SIMD. Scheduling only: same results.
        uint level = 3;
        __builtin_amdgcn_s_setprio(3);
#endif
#ifdef HT_HELPERS
        // Program option HT_HELPERS: a lane whose pixel has all its samples
        // stays in the loop and joins the scans of the wave's remaining rays
        // without a ray of its own (TopScan::detect with has_ray = false),
        // so their pair tests keep running on full lanes; the wave leaves
        // the loop when no lane has samples left. The bounce is split
        // around the scan so that every lane of the wave reaches it.
        typedef typename S::Object T;
        const u8 *obj = S::object(scene);
        bool alive = true;
        for (;;) {
            if (alive && start) {
                if (done == passes) {
                    alive = false;
                } else {
                    start_path<S>(scene, &ctx, px, &light, &emission);
                    hop = 0;
                    start = false;
                }
            }
            if (__ballot(alive) == 0) break;
            bool end = false;
#ifdef HT_RR_DEPTH
            if (alive && hop >= HT_RR_DEPTH && russian_roulette(&ctx, &light.base)) end = true;
#endif
            bool ray = alive && !end;
            if (ray) ctx.hasher = hash_init();
            typename T::Cache cache;
            real dist = TopScan<T>::detect(obj, &ctx, &cache, &light, ray);
            if (ray) {
                // the rest of bounce()
                if (dist > (real)-0.5f) {
                    if (!T::interact(obj, &ctx, &cache, &light, &emission)) end = true;
                    else ctx.prev_hash = hash_finish(&ctx.hasher);
                } else {
                    S::bg(scene, &ctx, &light, &emission);
                    end = true;
                }
            }
            if (!alive) continue;
            ++hop;
            if (hop >= hops) end = true;
            if (end) {
                acc.x += emission.x;
                acc.y += emission.y;
                acc.z += emission.z;
                acc.w += 1.0f;
                ++done;
                start = true;
            }
#ifdef HT_WAVE_PRIO
            if (level > 0 && __ballot(done < passes * (4 - level) / 4) == 0) {
                --level;
                if (level == 2) __builtin_amdgcn_s_setprio(2);
                else if (level == 1) __builtin_amdgcn_s_setprio(1);
                else __builtin_amdgcn_s_setprio(0);
            }
#endif
        }
#else
        for (;;) {
            if (start) {
                if (done == passes) break;
                start_path<S>(scene, &ctx, px, &light, &emission);
                hop = 0;
                start = false;
            }
            bool end = bounce<S>(scene, &ctx, &light, &emission, hop);
            ++hop;
            if (hop >= hops) end = true;
            if (end) {
                acc.x += emission.x;
                acc.y += emission.y;
                acc.z += emission.z;
                acc.w += 1.0f;
                ++done;
                start = true;
            }
#ifdef HT_WAVE_PRIO
            if (level > 0 && __ballot(done < passes * (4 - level) / 4) == 0) {
                --level;
                if (level == 2) __builtin_amdgcn_s_setprio(2);
                else if (level == 1) __builtin_amdgcn_s_setprio(1);
                else __builtin_amdgcn_s_setprio(0);
            }
#endif
        }
#endif
    }
    canvas[idx] = acc;
    seeds[idx] = ctx.rng.state;
}

// A lane's pixel in the queued forms (program options HT_PIXEL_QUEUE and
// HT_BATCH_WAVE_QUEUE): render_pixel_passes's state kept across the
// iterations of a loop in which each iteration advances every busy lane's
// path by one hit, so a lane can finish its pixel and take another while the
// others go on.
// The samples of a pixel run in order in one lane, so canvas and seeds are
// render_pixel_passes's bit for bit.
template <class S> struct QueuedPixel {
    typedef typename S::Geo G;
    Context ctx;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    Light<G> light;
    color3 emission = c3(0.0f);
    uint idx = 0, done = 0;
    int hop = 0;
    bool have = false, start = true;

    // Takes pixel i; a pixel with no bounce loop (hops <= 0) or no passes
    // is finished at once.
    HD void take(uint i, uint w, uint h, const u8 *__restrict__ scene, float4 *__restrict__ canvas,
                 uint *__restrict__ seeds, uint passes, int hops) {
        have = true;
        idx = i;
        ctx.rng.state = seeds[i];
        ctx.hasher = hash_init();
        acc = canvas[i];
        done = 0;
        start = true;
        if (hops <= 0 || passes == 0) {
            PixelSetup px = pixel_setup(i % w, i / w, w, h);
            for (uint k = 0; hops <= 0 && k < passes; ++k) {
                start_path<S>(scene, &ctx, px, &light, &emission);
                acc.w += 1.0f;
            }
            canvas[i] = acc;
            seeds[i] = ctx.rng.state;
            have = false;
        }
    }

    // One hit of the current sample (starting the sample first if needed);
    // after the pixel's last sample the pixel is written back.
    HD void step(uint w, uint h, const u8 *__restrict__ scene, float4 *__restrict__ canvas, uint *__restrict__ seeds,
                 uint passes, int hops) {
        if (start) {
            // the pixel position is recomputed per sample: fewer live registers
            start_path<S>(scene, &ctx, pixel_setup(idx % w, idx / w, w, h), &light, &emission);
            hop = 0;
            start = false;
        }
        bool end = bounce<S>(scene, &ctx, &light, &emission, hop);
        ++hop;
        if (hop >= hops) end = true;
        if (end) finish_sample(canvas, seeds, passes);
    }

    HD void finish_sample(float4 *__restrict__ canvas, uint *__restrict__ seeds, uint passes) {
        acc.x += emission.x;
        acc.y += emission.y;
        acc.z += emission.z;
        acc.w += 1.0f;
        ++done;
        start = true;
        if (done == passes) {
            canvas[idx] = acc;
            seeds[idx] = ctx.rng.state;
            have = false;
        }
    }

#ifdef HT_HELPERS
    // step() for every lane of the wave (program option HT_HELPERS): a lane
    // without a pixel joins the scan without a ray of its own and takes
    // other rays' pair tests. The bounce is split around the scan so that
    // every lane reaches it; a busy lane computes what step() computes.
    HD void step_all(uint w, uint h, const u8 *__restrict__ scene, float4 *__restrict__ canvas,
                     uint *__restrict__ seeds, uint passes, int hops) {
        typedef typename S::Object T;
        const bool busy = have;
        if (busy && start) {
            start_path<S>(scene, &ctx, pixel_setup(idx % w, idx / w, w, h), &light, &emission);
            hop = 0;
            start = false;
        }
        bool end = false;
#ifdef HT_RR_DEPTH
        if (busy && hop >= HT_RR_DEPTH && russian_roulette(&ctx, &light.base)) end = true;
#endif
        const bool ray = busy && !end;
        if (ray) ctx.hasher = hash_init();
        const u8 *obj = S::object(scene);
        typename T::Cache cache;
        real dist = TopScan<T>::detect(obj, &ctx, &cache, &light, ray);
        if (ray) {
            // the rest of bounce()
            if (dist > (real)-0.5f) {
                if (!T::interact(obj, &ctx, &cache, &light, &emission)) end = true;
                else ctx.prev_hash = hash_finish(&ctx.hasher);
            } else {
                S::bg(scene, &ctx, &light, &emission);
                end = true;
            }
        }
        if (!busy) return;
        ++hop;
        if (hop >= hops) end = true;
        if (end) finish_sample(canvas, seeds, passes);
    }
#endif
};

#ifdef HT_PIXEL_SPREAD
// Program option HT_PIXEL_SPREAD: the order in which render_passes hands
// the frame's n pixels to its lanes (statically, or out of the pixel
// queue). Work item t < extent takes pixel(t); items with pixel(t) >= n
// take none. Pixels are independent, so canvas and seeds are those of the
// identity order bit for bit; only which pixels share a wave changes. A
// cluster of expensive pixels (the scene's dense regions) is so spread over
// many waves instead of filling a few, each of which would hold the
// frame's end alone, and the lanes of a wave whose own pixels finish early
// help the scans of the rest (HT_HELPERS).
//   B = 2..64 (a power of two): bands. The pixels, padded to whole waves
//     (nw = ceil(n / 64)), are cut into B bands of nw * 64 / B; wave k takes
//     the k-th run of 64 / B consecutive pixels of every band (a bijection
//     of [0, nw * 64)).
//   1: golden-ratio order. pixel(t) = t * a mod n with a the first integer
//     >= 0.618 n coprime to n (a bijection of [0, n)): consecutive items lie
//     about 0.382 n apart, so a wave's 64 pixels sample the whole frame.
static_assert(HT_PIXEL_SPREAD == 1 ||
                  (HT_PIXEL_SPREAD >= 2 && HT_PIXEL_SPREAD <= 64 && (HT_PIXEL_SPREAD & (HT_PIXEL_SPREAD - 1)) == 0),
              "HT_PIXEL_SPREAD: 1 (golden-ratio order) or a power of two from 2 to 64 (bands)");
struct PixelOrder {
    uint n, extent, step;
    HD explicit PixelOrder(uint n_) : n(n_) {
#if HT_PIXEL_SPREAD == 1
        extent = n;
        step = (uint)((unsigned long long)n * 2654435769ull >> 32);  // floor(n * 0.6180339887)
        for (;; ++step) {
            uint a = step, b = n;
            while (b != 0u) {
                uint r = a % b;
                a = b;
                b = r;
            }
            if (a == 1u) break;  // gcd(step, n) = 1 (n = 1 keeps step 0: gcd(0, 1) = 1)
        }
#else
        extent = (n + 63u) / 64u * 64u;
        step = 0u;
#endif
    }
    HD uint pixel(uint t) const {
#if HT_PIXEL_SPREAD == 1
        return (uint)((unsigned long long)t * step % n);
#else
        const uint run = 64u / (uint)HT_PIXEL_SPREAD;
        const uint nw = extent / 64u;
        const uint wave = t / 64u, lane = t % 64u;
        return ((lane / run) * nw + wave) * run + lane % run;
#endif
    }
};
#endif

#ifdef HT_PIXEL_CLASSES
// Program option HT_PIXEL_CLASSES (with HT_PIXEL_QUEUE and HT_PIXEL_SPREAD =
// 1): before each render_passes launch the runtime launches
// pixel_classes(w, h, scene), which traces one probe camera ray per pixel
// (from a private RNG stream: the pixel's seed is neither read nor changed)
// and marks the pixels whose probe ray hits the scene. The pixel queue then
// hands out the marked pixels first and the others after them, each in the
// golden-ratio order: a pixel whose camera rays leave for the sky costs one
// scan per sample, the others several, so the long chains start first and
// the cheap pixels fill the frame's end (longest processing time first).
// Frames of more than HT_CLASS_CAP pixels get the golden-ratio order alone.
// Pixels stay independent: the same canvas and seeds.
#if !(defined(HT_PIXEL_QUEUE) && defined(HT_PIXEL_SPREAD) && HT_PIXEL_SPREAD == 1)
#error "HT_PIXEL_CLASSES needs HT_PIXEL_QUEUE and HT_PIXEL_SPREAD=1"
#endif
#define HT_CLASS_CAP (1u << 23)
__device__ uint ht_pixel_class[HT_CLASS_CAP / 32];  // bit p: pixel p's probe ray hit the scene

template <class S> HD void pixel_classes_impl(uint w, uint h, const u8 *__restrict__ scene) {
    typedef typename S::Object T;
    TopScan<T>::prepare(S::object(scene));
    __syncthreads();
    const uint n = w * h;
    const uint idx = blockIdx.x * blockDim.x + threadIdx.x;
    const bool have = idx < n;
    Context ctx;
    ctx.rng.state = idx * 0x9E3779B9u + 0x7F4A7C15u;  // private probe stream
    ctx.hasher = hash_init();
    Light<typename S::Geo> light;
    color3 emission;
    start_path<S>(scene, &ctx, pixel_setup(have ? idx % w : 0u, have ? idx / w : 0u, w, h), &light, &emission);
    ctx.hasher = hash_init();
    typename T::Cache cache;
    // every lane of the wave takes part in the scan; lanes past the frame
    // have no ray
    real dist = TopScan<T>::detect(S::object(scene), &ctx, &cache, &light, have);
    const unsigned long long m = __ballot(have && dist > (real)-0.5f);
    if (__lane_id() == 0 && idx < n) {  // the wave's 64 pixels start at idx, a multiple of 64
        ht_pixel_class[idx / 32u] = (uint)m;
        if (idx + 32u < n) ht_pixel_class[idx / 32u + 1u] = (uint)(m >> 32);
    }
}
HD bool pixel_class(uint p) { return (ht_pixel_class[p >> 5] >> (p & 31u)) & 1u; }
#endif

#ifdef HT_PIXEL_QUEUE
// Program option HT_PIXEL_QUEUE: render_passes as a persistent kernel (the
// runtime launches only the blocks that fit on the device at once, after
// pixel_queue_reset). A lane takes the next pixel from a global counter
// (one atomic per wave for all lanes that need one) and runs its `passes`
// samples with in-thread regeneration, exactly as render_pixel_passes does;
// when they are done it writes the pixel back and takes another. Lanes whose
// pixels need fewer hits no longer wait for the slowest lane of a static
// wave, which matters when the frame has many pixels per lane. Same canvas
// and seeds.
__device__ uint ht_pixel_next;

template <class S>
HD void render_passes_queued(uint w, uint h, const u8 *__restrict__ scene, float4 *__restrict__ canvas,
                             uint *__restrict__ seeds, uint passes) {
    const uint n = w * h;
#ifdef HT_PIXEL_SPREAD
    const PixelOrder order(n);
#ifdef HT_PIXEL_CLASSES
    // two rounds over the order: the marked pixels, then the others
    const bool classes = n <= HT_CLASS_CAP;
    const uint extent = classes ? 2u * order.extent : order.extent;
#else
    const uint extent = order.extent;  // the counter runs over the order's items
#endif
#else
    const uint extent = n;
#endif
    const unsigned long long below = lane_mask_below();
    QueuedPixel<S> p;
    bool drained = false;
#ifdef HT_WAVE_PRIO
    // Once the counter has passed the frame the waves still busy are the
    // frame's tail: each sets its issue priority by the work its slowest lane
    // has left (3: more than three quarters of the passes ... 0: less than a
    // quarter), so the longest remaining work issues first.
    uint level = 0;
#endif
    for (;;) {
#ifdef HT_WAVE_PRIO
        if (drained) {
            uint want = 0;
            if (__ballot(p.have && p.done < passes / 4) != 0) want = 3;
            else if (__ballot(p.have && p.done < passes / 2) != 0) want = 2;
            else if (__ballot(p.have && p.done < passes - passes / 4) != 0) want = 1;
            if (want != level) {
                level = want;
                if (want == 3) __builtin_amdgcn_s_setprio(3);
                else if (want == 2) __builtin_amdgcn_s_setprio(2);
                else if (want == 1) __builtin_amdgcn_s_setprio(1);
                else __builtin_amdgcn_s_setprio(0);
            }
        }
#endif
        if (!drained) {
            bool need = !p.have;
            unsigned long long m = __ballot(need);
            if (m != 0) {
                uint leader = (uint)__ffsll((long long)m) - 1u;
                uint base = 0;
                if (__lane_id() == leader) base = atomicAdd(&ht_pixel_next, (uint)__popcll(m));
                base = __shfl(base, (int)leader);
                if (base + (uint)__popcll(m) >= extent) drained = true;  // the counter only grows
                uint cand = base + (uint)__popcll(m & below);
#if defined(HT_PIXEL_CLASSES)
                if (need && cand < extent) {
                    const bool second = cand >= order.extent;
                    const uint q = order.pixel(second ? cand - order.extent : cand);
                    // a pixel of the other round: no pixel for this lane now (it asks again)
                    cand = (classes && pixel_class(q) == second) ? n : q;
                }
#elif defined(HT_PIXEL_SPREAD)
                if (need && cand < extent) cand = order.pixel(cand);
#endif
                if (need && cand < n) p.take(cand, w, h, scene, canvas, seeds, passes, S::HOPS);
            }
        }
        if (__ballot(p.have) == 0) {
            if (drained) break;
            continue;
        }
#ifdef HT_HELPERS
        p.step_all(w, h, scene, canvas, seeds, passes, S::HOPS);  // lanes without a pixel help the scan
#else
        if (p.have) p.step(w, h, scene, canvas, seeds, passes, S::HOPS);
#endif
    }
}
#endif

template <class S>
HD void render_passes_impl(uint w, uint h, const u8 *__restrict__ scene, float4 *__restrict__ canvas,
                           uint *__restrict__ seeds, uint passes) {
    TopScan<typename S::Object>::prepare(S::object(scene));
    __syncthreads();
#ifdef HT_PIXEL_QUEUE
    render_passes_queued<S>(w, h, scene, canvas, seeds, passes);
#else
    uint idx = blockIdx.x * blockDim.x + threadIdx.x;
#ifdef HT_PIXEL_SPREAD
    const PixelOrder order(w * h);
    if (idx >= order.extent) return;
    idx = order.pixel(idx);
#endif
    if (idx >= w * h) return;
    render_pixel_passes<S>(w, h, idx, scene, canvas, seeds, passes, S::HOPS);
#endif
}

}  // namespace ht

// Kernel entry points for the scene type `S`, with C linkage so the runtime
// finds them by name in the compiled code object. HT_WAVES_PER_EU: the
// minimum resident waves per SIMD the register allocation must allow (4:
// at most 128 VGPRs; measured best for 640x480 frames; 5 pays a small
// spill for more occupancy on larger frames; see tuning.py).
#ifndef HT_WAVES_PER_EU
#define HT_WAVES_PER_EU 4
#endif
// HT_SCALAR_BOUNDS programs also export bounds_fill(scene), which the
// runtime launches before each render / render_passes / render_cached
// launch of the program.
#if defined(HT_SCALAR_BOUNDS) && defined(HT_CULL_GROUPS)
// the tree is built by one block in its LDS (the launch's other blocks exit)
#define HT_DEFINE_BOUNDS_KERNEL(S)                                                                          \
    extern "C" __global__ void __launch_bounds__(256) bounds_fill(const unsigned char *__restrict__ scene) { \
        if (blockIdx.x == 0)                                                                                \
            ht::TopScan<typename S::Object>::fill_scalar_block(S::object(scene), ht::ht_scalar_bounds);     \
    }
#elif defined(HT_SCALAR_BOUNDS)
#define HT_DEFINE_BOUNDS_KERNEL(S)                                                                          \
    extern "C" __global__ void __launch_bounds__(256) bounds_fill(const unsigned char *__restrict__ scene) { \
        ht::TopScan<typename S::Object>::fill_scalar(S::object(scene), ht::ht_scalar_bounds,                \
                                                     blockIdx.x * blockDim.x + threadIdx.x,                 \
                                                     gridDim.x * blockDim.x);                               \
    }
#else
#define HT_DEFINE_BOUNDS_KERNEL(S)
#endif
#ifdef HT_PIXEL_QUEUE
#define HT_DEFINE_QUEUE_KERNEL                                                                        \
    extern "C" __global__ void pixel_queue_reset() {                                                  \
        if (threadIdx.x == 0) ht::ht_pixel_next = 0u;                                                 \
    }
#else
#define HT_DEFINE_QUEUE_KERNEL
#endif
#ifdef HT_PIXEL_CLASSES
#define HT_DEFINE_CLASS_KERNEL(S)                                                                     \
    extern "C" __global__ void __launch_bounds__(256, HT_WAVES_PER_EU)                                \
        pixel_classes(unsigned w, unsigned h, const unsigned char *__restrict__ scene) {              \
        ht::pixel_classes_impl<S>(w, h, scene);                                                       \
    }
#else
#define HT_DEFINE_CLASS_KERNEL(S)
#endif
#define HT_DEFINE_RENDER_KERNELS(S)                                                                   \
    HT_DEFINE_BOUNDS_KERNEL(S)                                                                        \
    HT_DEFINE_QUEUE_KERNEL                                                                            \
    HT_DEFINE_CLASS_KERNEL(S)                                                                         \
    extern "C" __global__ void __launch_bounds__(256, HT_WAVES_PER_EU) render(                        \
        unsigned w, unsigned h, const unsigned char *__restrict__ scene, float4 *__restrict__ canvas, \
        unsigned *__restrict__ seeds) {                                                               \
        ht::render_pass_impl<S>(w, h, scene, canvas, seeds);                                          \
    }                                                                                                 \
    extern "C" __global__ void __launch_bounds__(256, HT_WAVES_PER_EU) render_passes(                 \
        unsigned w, unsigned h, const unsigned char *__restrict__ scene, float4 *__restrict__ canvas, \
        unsigned *__restrict__ seeds, unsigned passes) {                                              \
        ht::render_passes_impl<S>(w, h, scene, canvas, seeds, passes);                                \
    }
