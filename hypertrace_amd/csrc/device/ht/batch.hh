// Batched render of many jobs (scenes of up to two types, views, image
// sizes, spp and depths) in one persistent launch.
//
// The frame's work is cut into tiles of `tile_px` pixels of one job (a
// multiple of the block size),
// numbered job after job (BatchJob::tile_begin). Each block takes the next
// tile from a global counter, stages that job's scene bounds for the culled
// scan (TopScan::prepare; with HT_SCALAR_BOUNDS it publishes the address of
// the job's FP64 bounds array, filled by batch_bounds_fill before the frame,
// instead), and renders the tile's pixels, `spp` passes each
// with in-thread path regeneration and at most `depth` hits per path; then it
// takes the next tile. Blocks that finish early take more tiles, so uneven
// jobs balance dynamically; the launch ends when the counter passes the last
// tile. Per pixel the work is exactly `render_passes` of that job's scene with
// HOPS = depth, so every job's canvas and seeds are bit-identical to
// rendering it alone.
//
// Program option HT_BATCH_WAVE_QUEUE (with HT_SCALAR_BOUNDS): each wave takes
// tiles from the counter on its own (no block barrier), and inside a tile
// each lane takes the tile's next pixel when its pixel's samples are done,
// so a wave waits for its slowest pixel once per tile instead of once for
// every pixel each lane renders in turn. Same canvas and seeds.
#pragma once
#include "render.hh"

#ifndef HT_BATCH_BLOCK
#define HT_BATCH_BLOCK 256
#endif

namespace ht {

static_assert(HT_BATCH_BLOCK % 64 == 0 && HT_BATCH_BLOCK <= 64 * HT_MAX_WAVES,
              "the scan's per-wave LDS scratch must cover every wave of a batch block");

// One job, written by the host (64 bytes; device pointers as integers;
// `bounds`: the job's BOUNDS_CAP FP64 bounds with HT_SCALAR_BOUNDS).
struct BatchJob {
    unsigned long long scene, canvas, seeds, bounds;
    uint type, w, h, spp, depth, tile_begin, tiles, pad0;
};
static_assert(sizeof(BatchJob) == 64, "BatchJob layout");

template <class S>
HD void batch_tile(const BatchJob &job, uint tile, uint tile_px) {
    const u8 *scene = reinterpret_cast<const u8 *>(job.scene);
#ifdef HT_SCALAR_BOUNDS
    if (threadIdx.x % 64 == 0) batch_bounds_slot() = job.bounds;
#else
    TopScan<typename S::Object>::prepare(S::object(scene));
#endif
    __syncthreads();
    uint npx = job.w * job.h;
    int hops = (int)job.depth < S::HOPS ? (int)job.depth : S::HOPS;
    for (uint k = threadIdx.x; k < tile_px; k += HT_BATCH_BLOCK) {
        uint idx = tile * tile_px + k;
        if (idx < npx)
            render_pixel_passes<S>(job.w, job.h, idx, scene, reinterpret_cast<float4 *>(job.canvas),
                                   reinterpret_cast<uint *>(job.seeds), job.spp, hops);
    }
    __syncthreads();  // the bounds are restaged for the next tile
}

// Job b's bounds (block b; any block size): HT_SCALAR_BOUNDS only.
template <class S0, class S1> HD void batch_bounds_fill_impl(const BatchJob *__restrict__ jobs, uint njobs) {
    if (blockIdx.x >= njobs) return;
    const BatchJob job = jobs[blockIdx.x];
    const u8 *scene = reinterpret_cast<const u8 *>(job.scene);
    BoundS *out = reinterpret_cast<BoundS *>(job.bounds);
#ifdef HT_CULL_GROUPS
    // the job's tree in this block's LDS, then its leaves and group nodes
    if (job.type == 0) {
        TopScan<typename S0::Object>::fill_scalar_block(S0::object(scene), out);
    } else {
        TopScan<typename S1::Object>::fill_scalar_block(S1::object(scene), out);
    }
#else
    if (job.type == 0) {
        TopScan<typename S0::Object>::fill_scalar(S0::object(scene), out, threadIdx.x, blockDim.x);
    } else {
        TopScan<typename S1::Object>::fill_scalar(S1::object(scene), out, threadIdx.x, blockDim.x);
    }
#endif
}

#ifdef HT_BATCH_WAVE_QUEUE
#ifndef HT_SCALAR_BOUNDS
#error "HT_BATCH_WAVE_QUEUE needs HT_SCALAR_BOUNDS (LDS bounds are staged per block)"
#endif
// The pixels [first, end) of one job by one wave: a lane whose pixel has its
// `spp` samples takes the next pixel of the range (one ballot for all lanes
// that need one); each pixel's samples run in order in one lane, as in
// render_pixel_passes.
template <class S>
HD void batch_wave_tile(const BatchJob &job, uint first, uint end, bool tail) {
    const u8 *scene = reinterpret_cast<const u8 *>(job.scene);
    float4 *canvas = reinterpret_cast<float4 *>(job.canvas);
    uint *seeds = reinterpret_cast<uint *>(job.seeds);
    const uint w = job.w, h = job.h, passes = job.spp;
    const int hops = (int)job.depth < S::HOPS ? (int)job.depth : S::HOPS;
    const unsigned long long below = lane_mask_below();
    uint cursor = first;  // the same in every lane
    QueuedPixel<S> p;
#ifdef HT_WAVE_PRIO
    // On the frame's last tiles (`tail`): the wave's issue priority follows
    // the work its tile has left (3 while pixels wait or its slowest lane has
    // more than three quarters of the passes left, down to 0), so the waves
    // that will finish last issue first. Back to 0 after the tile.
    uint level = 0;
#else
    (void)tail;
#endif
    for (;;) {
#ifdef HT_WAVE_PRIO
        if (tail) {
            uint want = 0;
            if (cursor < end || __ballot(p.have && p.done < passes / 4) != 0) want = 3;
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
        if (cursor < end) {
            bool need = !p.have;
            unsigned long long m = __ballot(need);
            uint cand = cursor + (uint)__popcll(m & below);
            if (need && cand < end) p.take(cand, w, h, scene, canvas, seeds, passes, hops);
            cursor += (uint)__popcll(m);
        }
        if (__ballot(p.have) == 0) {
            if (cursor >= end) break;
            continue;
        }
#ifdef HT_HELPERS
        // lanes without a pixel (the tile's pixels are all taken) help the
        // scans of the wave's remaining rays
        p.step_all(w, h, scene, canvas, seeds, passes, hops);
#else
        if (p.have) p.step(w, h, scene, canvas, seeds, passes, hops);
#endif
    }
#ifdef HT_WAVE_PRIO
    if (level != 0) __builtin_amdgcn_s_setprio(0);
#endif
}

template <class S0, class S1>
HD void render_batch_impl(const BatchJob *__restrict__ jobs, uint njobs, uint *__restrict__ counter,
                          uint total_tiles, uint tile_px) {
    for (;;) {
        uint t = 0;
        if (__lane_id() == 0) t = atomicAdd(counter, 1u);
        t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
        if (t >= total_tiles) return;
        uint j = 0;
        while (j + 1 < njobs && jobs[j + 1].tile_begin <= t) ++j;
        const BatchJob job = jobs[j];
        batch_bounds_slot() = job.bounds;  // this wave's slot; its own later reads see it
        uint npx = job.w * job.h;
        uint first = (t - job.tile_begin) * tile_px;
        uint end = first + (npx - first < tile_px ? npx - first : tile_px);
        // one of the last tiles: fewer tiles remain than waves run
        bool tail = t + gridDim.x * (blockDim.x / 64) >= total_tiles;
        if (job.type == 0) {
            batch_wave_tile<S0>(job, first, end, tail);
        } else {
            batch_wave_tile<S1>(job, first, end, tail);
        }
    }
}
#else
template <class S0, class S1>
HD void render_batch_impl(const BatchJob *__restrict__ jobs, uint njobs, uint *__restrict__ counter,
                          uint total_tiles, uint tile_px) {
    __shared__ uint next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1u);
        __syncthreads();
        uint t = next;
        __syncthreads();
        if (t >= total_tiles) return;
        uint j = 0;
        while (j + 1 < njobs && jobs[j + 1].tile_begin <= t) ++j;
        const BatchJob job = jobs[j];
        if (job.type == 0) {
            batch_tile<S0>(job, t - job.tile_begin, tile_px);
        } else {
            batch_tile<S1>(job, t - job.tile_begin, tile_px);
        }
    }
}
#endif

}  // namespace ht

// HT_SCALAR_BOUNDS batch programs also export batch_bounds_fill(jobs,
// njobs) (one block of 256 threads per job) and the number of 64-byte
// entries of a job's bounds array (with HT_CULL_GROUPS the tree's leaves
// and then its group nodes).
#ifdef HT_CULL_GROUPS
#define HT_BATCH_BOUND_ENTRIES (HT_BOUNDS_CAP + HT_BOUNDS_CAP / HT_CULL_GROUPS)
#else
#define HT_BATCH_BOUND_ENTRIES HT_BOUNDS_CAP
#endif
#ifdef HT_SCALAR_BOUNDS
#define HT_DEFINE_BATCH_BOUNDS_KERNEL(S0, S1)                                                            \
    extern "C" __device__ unsigned ht_batch_bounds_cap = HT_BATCH_BOUND_ENTRIES;                         \
    extern "C" __global__ void __launch_bounds__(256)                                                    \
        batch_bounds_fill(const ht::BatchJob *__restrict__ jobs, unsigned njobs) {                       \
        ht::batch_bounds_fill_impl<S0, S1>(jobs, njobs);                                                 \
    }
#else
#define HT_DEFINE_BATCH_BOUNDS_KERNEL(S0, S1)
#endif
// Wave-queue programs mark themselves: their tiles are per wave, so the
// runtime asks only for a multiple of 64 pixels.
#ifdef HT_BATCH_WAVE_QUEUE
#define HT_DEFINE_BATCH_TILE_MARK extern "C" __device__ unsigned ht_batch_wave_tiles = 1;
#else
#define HT_DEFINE_BATCH_TILE_MARK
#endif
#define HT_DEFINE_BATCH_KERNEL(S0, S1)                                                                   \
    HT_DEFINE_BATCH_BOUNDS_KERNEL(S0, S1)                                                                \
    HT_DEFINE_BATCH_TILE_MARK                                                                            \
    extern "C" __global__ void __launch_bounds__(HT_BATCH_BLOCK, HT_WAVES_PER_EU)                        \
        render_batch(const ht::BatchJob *__restrict__ jobs, unsigned njobs, unsigned *__restrict__ counter, \
                     unsigned total_tiles, unsigned tile_px) {                                           \
        ht::render_batch_impl<S0, S1>(jobs, njobs, counter, total_tiles, tile_px);                       \
    }
