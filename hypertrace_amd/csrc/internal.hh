// Internal types of the C-ABI runtime shared by its source files
// (runtime.hip: contexts, buffers, programs, post passes; cache.hip: the
// radiance cache). Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>

#include "../../include/hypertrace_hip.h"

namespace ht_internal {

// Records the message for ht_last_error and returns the code.
int fail(int code, const std::string &msg);

}  // namespace ht_internal

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) {                                                                         \
            return ht_internal::fail(HT_ERR_BACKEND, std::string(#expr) + ": " + hipGetErrorString(e_)); \
        }                                                                                               \
    } while (0)

// A context is reference counted like an OpenCL cl_context: its owner holds
// one reference and every buffer, program and cache retains one, so
// destroying the context while they are alive (e.g. freed later by a garbage
// collector) keeps the stream valid until the last of them is released.
struct ht_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    std::atomic<int> refs{1};
};

struct ht_prog {
    ht_ctx *ctx = nullptr;
    hipModule_t module = nullptr;
    hipFunction_t render = nullptr;
    hipFunction_t render_passes = nullptr;
    hipFunction_t render_cached = nullptr;  // null unless the scene type supports the radiance cache
    hipFunction_t render_batch = nullptr;   // batch programs (several scene types, one persistent kernel)
    // scalar-bounds batch programs: fills every job's bounds array before a
    // frame; each array holds batch_bounds_cap bounds of 64 B
    hipFunction_t batch_bounds_fill = nullptr;
    unsigned batch_bounds_cap = 0;
    bool batch_wave_tiles = false;          // wave-queue batch program: tiles per wave (multiples of 64)
    hipFunction_t bounds_fill = nullptr;    // scalar-bounds programs: fills the bounds before each render launch
    // pixel-queue programs: resets the pixel counter before each
    // render_passes launch, which then runs as many blocks as fit at once
    hipFunction_t pixel_queue_reset = nullptr;
    // programs built with -DHT_PIXEL_CLASSES: the probe that orders the pixel
    // queue, launched before each render_passes launch
    hipFunction_t pixel_classes = nullptr;
    uint32_t resident_blocks = 0;
    // Wavefront form (programs whose root source adds the wavefront kernels):
    // kernels, queue sizes {extend-record words, shade-record words, keys,
    // counter words, HOPS} and the queue workspace, grown on demand.
    hipFunction_t wf_generate = nullptr, wf_extend = nullptr, wf_shade = nullptr;
    unsigned wf_layout[5] = {0, 0, 0, 0, 0};
    void *wf_mem = nullptr;
    size_t wf_bytes = 0;
    // Held from a scalar-bounds program's bounds_fill enqueue to its render
    // enqueue, so callers sharing the program from several threads cannot
    // interleave another scene's fill between the two (stream order does
    // the rest).
    std::mutex launch_mu;
};

struct ht_buf {
    ht_ctx *ctx = nullptr;
    void *ptr = nullptr;
    size_t size = 0;
};

namespace ht_internal {

inline void ctx_retain(ht_ctx *ctx) { ctx->refs.fetch_add(1); }

// Scalar-bounds programs: writes the scene's element bounds into the
// program's bounds array on the context's stream, ahead of a render launch
// (4 blocks; the kernel strides over any element count). No-op otherwise.
inline int fill_bounds(ht_prog *prog, void *scene_ptr) {
    if (!prog->bounds_fill) return HT_OK;
    void *args[] = {&scene_ptr};
    HIP_TRY(hipModuleLaunchKernel(prog->bounds_fill, 4, 1, 1, 256, 1, 1, 0, prog->ctx->stream, args, nullptr));
    return HT_OK;
}
inline void ctx_release(ht_ctx *ctx) {
    if (ctx->refs.fetch_sub(1) != 1) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipEventDestroy(ctx->t0);
    (void)hipEventDestroy(ctx->t1);
    (void)hipStreamDestroy(ctx->stream);
    (void)hipGetLastError();  // teardown errors are not reported to later calls
    delete ctx;
}

}  // namespace ht_internal
