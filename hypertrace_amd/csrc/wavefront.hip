// Host side of the wavefront render form (kernels in device/ht/wavefront.hh):
// workspace layout, the generate / extend / shade launch sequence, and the
// termination check on the live-path count.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "internal.hh"

using ht_internal::fail;

namespace {

// Iterations (extend + shade) queued between reads of the live-path count.
constexpr uint32_t WF_CHUNK = 8;
// Grid of the extend and shade kernels: they stride over their queue, so
// any grid covers it; this one fills the chip (256 CUs x 4 resident blocks
// x 2) without launching idle blocks for long queues.
constexpr uint32_t WF_MAX_BLOCKS = 2048;

int check_size(const char *what, const ht_buf *b, size_t need) {
    if (!b) return fail(HT_ERR_OTHER, std::string(what) + ": null buffer");
    if (b->size < need)
        return fail(HT_ERR_OTHER, std::string(what) + ": buffer of " + std::to_string(b->size) + " bytes, need " +
                                      std::to_string(need));
    return HT_OK;
}

// Checks the arguments and returns the workspace size in bytes (0 when the
// call has nothing to do: no passes or an empty image).
int wf_prepare(const char *who, ht_prog *prog, uint32_t w, uint32_t h, ht_buf *scene, ht_buf *canvas, ht_buf *seeds,
               uint32_t passes, size_t *bytes) {
    *bytes = 0;
    if (!prog || !scene) return fail(HT_ERR_OTHER, std::string(who) + ": invalid argument");
    if (!prog->wf_extend)
        return fail(HT_ERR_OTHER, std::string(who) + ": the program was built without the wavefront kernels");
    if (passes == 0 || w == 0 || h == 0) return HT_OK;
    if ((uint64_t)w * (uint64_t)h >= 0x80000000ull) return fail(HT_ERR_OTHER, std::string(who) + ": image too large");
    const uint32_t n = w * h;
    int rc;
    if ((rc = check_size((std::string(who) + " canvas").c_str(), canvas, (size_t)n * 16))) return rc;
    if ((rc = check_size((std::string(who) + " seeds").c_str(), seeds, (size_t)n * 4))) return rc;
    if (scene->size == 0) return fail(HT_ERR_OTHER, std::string(who) + ": empty scene buffer");
    const uint32_t ext_words = prog->wf_layout[0], shade_words = prog->wf_layout[1], keys = prog->wf_layout[2];
    const uint32_t ctr_words = prog->wf_layout[3];
    // workspace: 2 extend queues and `keys` shade queues of n SoA records
    // each, then the counters (cleared by wf_generate on every call)
    const size_t words = (size_t)n * (2 * (size_t)ext_words + (size_t)keys * shade_words) + ctr_words;
    *bytes = words * 4;
    return HT_OK;
}

// The launch sequence over a workspace of at least the wf_prepare size.
// Returns when every path has ended (the stream is then idle), or, for a
// program without hops, after wf_generate's launch.
int wf_run(const char *who, ht_prog *prog, uint32_t w, uint32_t h, ht_buf *scene, ht_buf *canvas, ht_buf *seeds,
           uint32_t passes, uint32_t *base) {
    const uint32_t n = w * h;
    const uint32_t ext_words = prog->wf_layout[0], keys = prog->wf_layout[2], hops = prog->wf_layout[4];
    const uint32_t shade_words = prog->wf_layout[1];
    ht_ctx *ctx = prog->ctx;
    uint32_t *q_ext[2] = {base, base + (size_t)n * ext_words};
    uint32_t *q_shade = q_ext[1] + (size_t)n * ext_words;
    uint32_t *ctr = q_shade + (size_t)n * keys * shade_words;

    void *sp = scene->ptr, *cp = canvas->ptr, *pp = seeds->ptr;
    int rc;
    if ((rc = ht_internal::fill_bounds(prog, sp))) return rc;
    {
        void *args[] = {&w, &h, &sp, &cp, &pp, &passes, &q_ext[0], &ctr};
        HIP_TRY(hipModuleLaunchKernel(prog->wf_generate, (n + 255) / 256, 1, 1, 256, 1, 1, 0, ctx->stream, args,
                                      nullptr));
    }
    if (hops == 0) return HT_OK;  // every sample ended in wf_generate
    const uint32_t blocks = (n + 255) / 256 < WF_MAX_BLOCKS ? (n + 255) / 256 : WF_MAX_BLOCKS;
    // each iteration advances every live path by one hit; a pixel needs at
    // most passes x hops of them
    const uint64_t max_iter = (uint64_t)passes * hops + WF_CHUNK;
    uint64_t iter = 0;
    for (;;) {
        for (uint32_t c = 0; c < WF_CHUNK; ++c, ++iter) {
            uint32_t parity = (uint32_t)(iter & 1);
            uint32_t *q_in = q_ext[parity], *q_next = q_ext[parity ^ 1];
            void *ext_args[] = {&w, &h, &sp, &cp, &pp, &passes, &q_in, &q_next, &q_shade, &ctr, &parity};
            HIP_TRY(hipModuleLaunchKernel(prog->wf_extend, blocks, 1, 1, 256, 1, 1, 0, ctx->stream, ext_args,
                                          nullptr));
            void *sh_args[] = {&w, &h, &sp, &cp, &pp, &passes, &q_shade, &q_next, &ctr, &parity};
            HIP_TRY(hipModuleLaunchKernel(prog->wf_shade, blocks, 1, 1, 256, 1, 1, 0, ctx->stream, sh_args,
                                          nullptr));
        }
        uint32_t live = 0;
        HIP_TRY(hipMemcpyAsync(&live, ctr + (iter & 1), 4, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (live == 0) return HT_OK;
        if (iter > max_iter)
            return fail(HT_ERR_OTHER, std::string(who) + ": paths still live after passes x hops iterations");
    }
}

// ------------------------------------------------------------ allocation-free form
// Workspace of ht_render_wavefront_noalloc: one array per GPU, part of the
// library's code object (so each GPU the library runs on holds it until the
// process ends). Host code uses the per-device address from
// hipGetSymbolAddress, never the array's name.
static __device__ __align__(256) uint32_t scratch[HT_WF_SCRATCH_BYTES / 4];

constexpr int WF_MAX_DEVICES = 64;
std::mutex g_scratch_mu;                 // guards g_scratch_ptr and serialises the calls
uint32_t *g_scratch_ptr[WF_MAX_DEVICES];  // per-device address of `scratch`, null until resolved

// The device's address of `scratch` (the device must be current).
int scratch_for(int device, uint32_t **out) {
    if (device < 0 || device >= WF_MAX_DEVICES)
        return fail(HT_ERR_OTHER, "ht_render_wavefront_noalloc: device index beyond the supported 64");
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(HT_ERR_BACKEND,
                        std::string("ht_render_wavefront_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<uint32_t *>(p);
    }
    *out = g_scratch_ptr[device];
    return HT_OK;
}

}  // namespace

extern "C" int ht_render_wavefront(ht_prog *prog, uint32_t w, uint32_t h, ht_buf *scene, ht_buf *canvas,
                                   ht_buf *seeds, uint32_t passes) {
    size_t bytes = 0;
    int rc = wf_prepare("ht_render_wavefront", prog, w, h, scene, canvas, seeds, passes, &bytes);
    if (rc || bytes == 0) return rc;
    ht_ctx *ctx = prog->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> lock(prog->launch_mu);
    if (prog->wf_bytes < bytes) {
        if (prog->wf_mem) {
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            HIP_TRY(hipFree(prog->wf_mem));
            prog->wf_mem = nullptr;
            prog->wf_bytes = 0;
        }
        HIP_TRY(hipMalloc(&prog->wf_mem, bytes));
        prog->wf_bytes = bytes;
    }
    return wf_run("ht_render_wavefront", prog, w, h, scene, canvas, seeds, passes, static_cast<uint32_t *>(prog->wf_mem));
}

extern "C" int ht_render_wavefront_noalloc(ht_prog *prog, uint32_t w, uint32_t h, ht_buf *scene, ht_buf *canvas,
                                           ht_buf *seeds, uint32_t passes) {
    size_t bytes = 0;
    int rc = wf_prepare("ht_render_wavefront_noalloc", prog, w, h, scene, canvas, seeds, passes, &bytes);
    if (rc || bytes == 0) return rc;
    if (bytes > (size_t)HT_WF_SCRATCH_BYTES)
        return fail(HT_ERR_OTHER, "ht_render_wavefront_noalloc: the frame needs " + std::to_string(bytes) +
                                      " bytes of queues, above the fixed workspace of " +
                                      std::to_string((size_t)HT_WF_SCRATCH_BYTES) +
                                      " bytes (use ht_render_wavefront)");
    ht_ctx *ctx = prog->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    // One call at a time per process: every call on a GPU shares its array.
    // The call returns with its stream idle (wf_run waits for the last
    // live-path count), so the next call starts on an unused workspace.
    std::lock_guard<std::mutex> glock(g_scratch_mu);
    uint32_t *ws = nullptr;
    if ((rc = scratch_for(ctx->device, &ws))) return rc;
    std::lock_guard<std::mutex> lock(prog->launch_mu);
    rc = wf_run("ht_render_wavefront_noalloc", prog, w, h, scene, canvas, seeds, passes, ws);
    // no launch of this call outlives the lock (a program without hops
    // returns from wf_run right after wf_generate's launch; a failed call
    // may leave launches queued)
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (rc) return rc;
    if (e != hipSuccess)
        return fail(HT_ERR_BACKEND, std::string("ht_render_wavefront_noalloc: ") + hipGetErrorString(e));
    return HT_OK;
}
