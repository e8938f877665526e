// Host side of the wavefront render form (kernels in device/ht/wavefront.hh):
// workspace layout, the generate / extend / shade launch sequence, and the
// termination check on the live-path count.
#include <hip/hip_runtime.h>

#include <cstdint>
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

}  // namespace

extern "C" int ht_render_wavefront(ht_prog *prog, uint32_t w, uint32_t h, ht_buf *scene, ht_buf *canvas,
                                   ht_buf *seeds, uint32_t passes) {
    if (!prog || !scene) return fail(HT_ERR_OTHER, "ht_render_wavefront: invalid argument");
    if (!prog->wf_extend)
        return fail(HT_ERR_OTHER, "ht_render_wavefront: the program was built without the wavefront kernels");
    if (passes == 0 || w == 0 || h == 0) return HT_OK;
    if ((uint64_t)w * (uint64_t)h >= 0x80000000ull) return fail(HT_ERR_OTHER, "ht_render_wavefront: image too large");
    const uint32_t n = w * h;
    int rc;
    if ((rc = check_size("ht_render_wavefront canvas", canvas, (size_t)n * 16))) return rc;
    if ((rc = check_size("ht_render_wavefront seeds", seeds, (size_t)n * 4))) return rc;
    if (scene->size == 0) return fail(HT_ERR_OTHER, "ht_render_wavefront: empty scene buffer");
    const uint32_t ext_words = prog->wf_layout[0], shade_words = prog->wf_layout[1], keys = prog->wf_layout[2];
    const uint32_t ctr_words = prog->wf_layout[3], hops = prog->wf_layout[4];
    // workspace: 2 extend queues and `keys` shade queues of n SoA records
    // each, then the counters
    const size_t words = (size_t)n * (2 * (size_t)ext_words + (size_t)keys * shade_words) + ctr_words;
    const size_t bytes = words * 4;
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
    uint32_t *base = static_cast<uint32_t *>(prog->wf_mem);
    uint32_t *q_ext[2] = {base, base + (size_t)n * ext_words};
    uint32_t *q_shade = q_ext[1] + (size_t)n * ext_words;
    uint32_t *ctr = q_shade + (size_t)n * keys * shade_words;

    void *sp = scene->ptr, *cp = canvas->ptr, *pp = seeds->ptr;
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
            return fail(HT_ERR_OTHER, "ht_render_wavefront: paths still live after passes x hops iterations");
    }
}
