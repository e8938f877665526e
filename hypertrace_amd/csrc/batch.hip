// Batched multi-job render on one MI355X: a job table, the persistent
// render_batch launch of a batch program (hypertrace_amd/csrc/device/ht/batch.hh),
// and a hipGraph of one frame (counter and canvas clears, the launch, the
// post pass of every job) that is captured once and replayed per frame.
// No reference counterpart: the reference renders one scene per program and
// per launch (processing/src/render.rs:102-118).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <algorithm>
#include <vector>

#include "internal.hh"

using ht_internal::ctx_release;
using ht_internal::ctx_retain;
using ht_internal::fail;

namespace {

// Mirror of ht::BatchJob in device/ht/batch.hh (64 bytes).
struct DevJob {
    uint64_t scene, canvas, seeds, bounds;
    uint32_t type, w, h, spp, depth, tile_begin, tiles, pad0;
};
static_assert(sizeof(DevJob) == 64, "BatchJob layout");

__global__ void k_batch_fill_u32(uint32_t *__restrict__ p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_batch_zero_f4(float4 *__restrict__ p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

}  // namespace

struct ht_batch {
    ht_prog *prog = nullptr;
    ht_ctx *ctx = nullptr;
    std::vector<ht_batch_job> jobs;
    uint32_t tile = 0, block = 0, total_tiles = 0;
    DevJob *dev_jobs = nullptr;
    uint32_t *counter = nullptr;
    void *bounds_mem = nullptr;  // scalar-bounds programs: every job's bounds array
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

int batch_check_job(const ht_batch_job &j, size_t i) {
    std::string who = "ht_batch_create job " + std::to_string(i);
    if (j.type > 1) return fail(HT_ERR_OTHER, who + ": type must be 0 or 1");
    if (!j.scene || !j.canvas || !j.seeds) return fail(HT_ERR_OTHER, who + ": null buffer");
    size_t px = (size_t)j.w * j.h;
    if (px == 0 || px > 0x80000000ull) return fail(HT_ERR_OTHER, who + ": image empty or larger than 2^31 pixels");
    if (j.canvas->size < 16 * px || j.seeds->size < 4 * px)
        return fail(HT_ERR_OTHER, who + ": canvas or seed buffer smaller than the image");
    if (j.scene->size == 0) return fail(HT_ERR_OTHER, who + ": empty scene buffer");
    if (j.rgba8 && j.rgba8->size < 4 * px) return fail(HT_ERR_OTHER, who + ": rgba8 buffer smaller than the image");
    return HT_OK;
}

// One frame on the context's stream: optional canvas clears, counter reset,
// the persistent launch, optional post passes.
int batch_frame(ht_batch *b, uint32_t grid, int clear, float gamma, int post) {
    hipStream_t st = b->ctx->stream;
    (void)hipGetLastError();
    if (clear) {
        for (const ht_batch_job &j : b->jobs) {
            uint32_t n = j.w * j.h;
            hipLaunchKernelGGL(k_batch_zero_f4, dim3((n + 255) / 256), dim3(256), 0, st, (float4 *)j.canvas->ptr, n);
        }
    }
    hipLaunchKernelGGL(k_batch_fill_u32, dim3(1), dim3(64), 0, st, b->counter, 1u, 0u);
    HIP_TRY(hipGetLastError());
    uint32_t njobs = (uint32_t)b->jobs.size();
    void *jp = b->dev_jobs, *cp = b->counter;
    if (b->prog->batch_bounds_fill) {
        // the scenes' bounds, per frame (a caller may rewrite scene buffers between frames)
        void *fill_args[] = {&jp, &njobs};
        HIP_TRY(hipModuleLaunchKernel(b->prog->batch_bounds_fill, njobs, 1, 1, 256, 1, 1, 0, st, fill_args,
                                      nullptr));
    }
    void *args[] = {&jp, &njobs, &cp, &b->total_tiles, &b->tile};
    HIP_TRY(hipModuleLaunchKernel(b->prog->render_batch, grid, 1, 1, b->block, 1, 1, 0, st, args, nullptr));
    if (post) {
        for (const ht_batch_job &j : b->jobs) {
            if (!j.rgba8) continue;
            int rc = ht_post(b->ctx, j.w, j.h, j.spp, gamma, j.canvas, j.rgba8);
            if (rc) return rc;
        }
    }
    return HT_OK;
}

}  // namespace

extern "C" int ht_batch_create(ht_prog *prog, uint32_t njobs, const ht_batch_job *jobs, uint32_t tile,
                               ht_batch **out) {
    if (!prog || !out || (njobs && !jobs)) return fail(HT_ERR_OTHER, "ht_batch_create: invalid argument");
    if (!prog->render_batch) return fail(HT_ERR_OTHER, "ht_batch_create: not a batch program (no render_batch)");
    if (njobs == 0) return fail(HT_ERR_OTHER, "ht_batch_create: no jobs");
    int block = 0;
    HIP_TRY(hipSetDevice(prog->ctx->device));
    HIP_TRY(hipFuncGetAttribute(&block, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, prog->render_batch));
    // tiles are per block, or per wave in wave-queue programs
    uint32_t unit = prog->batch_wave_tiles ? 64u : (uint32_t)block;
    if (tile == 0 || tile % unit)
        return fail(HT_ERR_OTHER, std::string("ht_batch_create: tile must be a positive multiple of ") +
                                      (prog->batch_wave_tiles ? "64 (a wave-queue program)"
                                                              : "the program's block (" + std::to_string(block) + ")"));
    for (uint32_t i = 0; i < njobs; ++i) {
        int rc = batch_check_job(jobs[i], i);
        if (rc) return rc;
        if (jobs[i].scene->ctx != prog->ctx || jobs[i].canvas->ctx != prog->ctx || jobs[i].seeds->ctx != prog->ctx)
            return fail(HT_ERR_OTHER, "ht_batch_create: buffers on another context");
    }
    // Tiles are numbered job after job with the jobs in descending order of
    // samples x depth per pixel (a bound on a pixel's work), so the longest
    // tiles are taken from the counter first and the frame does not end on
    // a heavy job's late tiles (longest processing time first). The device
    // table is in that order; each job's pixels are rendered as before, so
    // the order changes no result.
    std::vector<uint32_t> order(njobs);
    for (uint32_t i = 0; i < njobs; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) {
        return (uint64_t)jobs[a].spp * jobs[a].depth > (uint64_t)jobs[c].spp * jobs[c].depth;
    });
    std::vector<DevJob> table(njobs);
    uint32_t tiles = 0;
    for (uint32_t k = 0; k < njobs; ++k) {
        const uint32_t i = order[k];
        uint32_t px = jobs[i].w * jobs[i].h;
        DevJob &d = table[k];
        std::memset(&d, 0, sizeof d);
        d.scene = (uint64_t)(uintptr_t)jobs[i].scene->ptr;
        d.canvas = (uint64_t)(uintptr_t)jobs[i].canvas->ptr;
        d.seeds = (uint64_t)(uintptr_t)jobs[i].seeds->ptr;
        d.type = jobs[i].type;
        d.w = jobs[i].w;
        d.h = jobs[i].h;
        d.spp = jobs[i].spp;
        d.depth = jobs[i].depth;
        d.tile_begin = tiles;
        d.tiles = (px + tile - 1) / tile;
        tiles += d.tiles;
    }
    ht_batch *b = new ht_batch();
    b->prog = prog;
    b->ctx = prog->ctx;
    b->jobs.assign(jobs, jobs + njobs);
    b->tile = tile;
    b->block = (uint32_t)block;
    b->total_tiles = tiles;
    void *mem = nullptr;
    hipError_t e = hipMalloc(&mem, sizeof(DevJob) * njobs + 256);
    if (e != hipSuccess) {
        delete b;
        return fail(HT_ERR_BACKEND, std::string("ht_batch_create: hipMalloc: ") + hipGetErrorString(e));
    }
    if (prog->batch_bounds_fill) {
        size_t per_job = (size_t)prog->batch_bounds_cap * 64;
        e = hipMalloc(&b->bounds_mem, per_job * njobs);
        if (e != hipSuccess) {
            (void)hipFree(mem);
            delete b;
            return fail(HT_ERR_BACKEND, std::string("ht_batch_create: hipMalloc (bounds): ") + hipGetErrorString(e));
        }
        for (uint32_t i = 0; i < njobs; ++i)
            table[i].bounds = (uint64_t)(uintptr_t)((uint8_t *)b->bounds_mem + per_job * i);
    }
    b->dev_jobs = (DevJob *)mem;
    b->counter = (uint32_t *)((uint8_t *)mem + sizeof(DevJob) * njobs);
    e = hipMemcpyAsync(b->dev_jobs, table.data(), sizeof(DevJob) * njobs, hipMemcpyHostToDevice, b->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(mem);
        if (b->bounds_mem) (void)hipFree(b->bounds_mem);
        delete b;
        return fail(HT_ERR_BACKEND, std::string("ht_batch_create: ") + hipGetErrorString(e));
    }
    ctx_retain(b->ctx);
    *out = b;
    return HT_OK;
}

extern "C" void ht_batch_destroy(ht_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    (void)hipFree(b->dev_jobs);
    if (b->bounds_mem) (void)hipFree(b->bounds_mem);
    (void)hipGetLastError();
    ht_ctx *ctx = b->ctx;
    delete b;
    ctx_release(ctx);
}

extern "C" uint32_t ht_batch_tiles(const ht_batch *b) { return b ? b->total_tiles : 0; }
extern "C" uint32_t ht_batch_block(const ht_batch *b) { return b ? b->block : 0; }

extern "C" int ht_batch_render(ht_batch *b, uint32_t grid, int clear, float gamma, int post) {
    if (!b || grid == 0) return fail(HT_ERR_OTHER, "ht_batch_render: invalid argument");
    if (post && !clear) return fail(HT_ERR_OTHER, "ht_batch_render: post passes need per-frame canvases (clear)");
    HIP_TRY(hipSetDevice(b->ctx->device));
    return batch_frame(b, grid, clear, gamma, post);
}

extern "C" int ht_batch_capture(ht_batch *b, uint32_t grid, int clear, float gamma, int post) {
    if (!b || grid == 0) return fail(HT_ERR_OTHER, "ht_batch_capture: invalid argument");
    if (post && !clear) return fail(HT_ERR_OTHER, "ht_batch_capture: post passes need per-frame canvases (clear)");
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    HIP_TRY(hipStreamBeginCapture(b->ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_frame(b, grid, clear, gamma, post);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->ctx->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(HT_ERR_BACKEND, std::string("ht_batch_capture: ") + hipGetErrorString(e));
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return HT_OK;
}

extern "C" int ht_batch_replay(ht_batch *b) {
    if (!b || !b->exec) return fail(HT_ERR_OTHER, "ht_batch_replay: no captured frame");
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipGraphLaunch(b->exec, b->ctx->stream));
    return HT_OK;
}
