/* hypertrace_hip.h — C ABI of the MI355X (gfx950) render pass.
 *
 * Drop-in for the OpenCL plumbing of the reference's `processing` crate:
 * every entry point names the reference code it replaces. All functions
 * return 0 on success and a negative code on failure; ht_last_error() then
 * holds a message (thread-local). Handles are opaque and owned by the caller.
 * One in-order HIP stream per context; calls on one context are not
 * reentrant.
 */
#ifndef HYPERTRACE_HIP_H
#define HYPERTRACE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ht_ctx ht_ctx;
typedef struct ht_prog ht_prog;
typedef struct ht_buf ht_buf;

/* Error codes, mapped onto processing::Error (processing/src/error.rs:3-9). */
#define HT_OK 0
#define HT_ERR_IO (-1)      /* Error::Io */
#define HT_ERR_BACKEND (-2) /* Error::Backend: HIP / hipRTC failure */
#define HT_ERR_OTHER (-3)   /* Error::Other: bad argument, size mismatch */

/* Last error message of the calling thread (never NULL). */
const char *ht_last_error(void);

/* Context: device + one in-order stream.
 * Replaces main/src/cli.rs:62-71 (platform/device/queue selection) and
 * processing/src/context.rs:3-16. */
int ht_ctx_create(int device_index, ht_ctx **out);
void ht_ctx_destroy(ht_ctx *ctx);
/* Block until all work queued on the context's stream is done. */
int ht_ctx_sync(ht_ctx *ctx);

/* Program build with hipRTC from a root source plus an in-memory header map
 * (names/contents), for one scene type. Replaces Render::source/new
 * (processing/src/render.rs:14-100), which assembled the OpenCL program from
 * kernel::SOURCE plus the generated scene headers and built it with
 * ocl::Program. double_precision mirrors DOUBLE_PRECISION, address_width
 * mirrors ADDRESS_WIDTH (32 or 64). Compiled code objects are cached on disk
 * under $HT_CACHE_DIR (default: the library's directory + "/../_cache"),
 * keyed by a hash of sources and options. If $HT_DUMP_SOURCE is set, the root
 * source is written to ./render.hip (render.rs:74-80 wrote ./render.ocl); if
 * $HT_LOG_COMPILES is set, every build that misses the cache says so on
 * stderr. */
int ht_prog_build(ht_ctx *ctx, const char *root_name, const char *root_src, int n_headers,
                  const char *const *names, const char *const *contents, int double_precision,
                  int address_width, ht_prog **out);
/* As ht_prog_build, with the floating-point code-generation options given
 * explicitly (space-separated compiler flags). NULL = the default,
 * "-ffp-contract=off -DHT_STRICT_FP=1": every operation rounded on its own,
 * IEEE division and square root, FP32 transcendentals correctly rounded.
 * "-fno-hip-fp32-correctly-rounded-divide-sqrt" is the fast mode: FMA
 * contraction, reciprocal division, the FP32 device libm. Further macros
 * select schedules that render the same images: -DHT_CULL_UNROLL=n (cull
 * tests per scan step, default 4), -DHT_WAVES_PER_EU=w (register cap,
 * default 4), -DHT_SCALAR_BOUNDS=1 (the scan's bounds in a global FP64
 * array read through the scalar cache, filled by the program's bounds_fill
 * kernel, which ht_render / ht_render_passes / ht_render_cached launch
 * first on the context's stream), -DHT_CULL_TREE=1 (sphere-tree walk),
 * -DHT_CULL_GROUPS=g (two-level scan over groups of g elements),
 * -DHT_PIXEL_QUEUE=1 (ht_render_passes launches only the blocks that fit
 * on the device at once, after the program's pixel_queue_reset kernel, and
 * lanes take pixels from a per-program counter; launches of one program
 * are ordered on its context's stream, so they never share it),
 * -DHT_WAVE_PRIO=1 (each wave's issue priority steps down as its pixels'
 * samples complete, so the waves with the most work left issue first),
 * -DHT_GROUP_QUEUE=1, -DHT_HIT_ORDER=1, -DHT_SPARSE_L1=r and -DHT_HELPERS=1
 * (the two-level scan compacted at both levels, its exact tests ordered
 * and skipped by crossing intervals, level 1 over (group, ray) pairs in
 * waves with at most r rays, lanes without a path joining the wave's
 * scans), -DHT_PIXEL_SPREAD=1 or 2..64 (the order in which render_passes
 * hands pixels to lanes: golden-ratio or banded) and -DHT_PIXEL_CLASSES=1
 * (with the pixel queue: ht_render_passes first launches the program's
 * pixel_classes probe, one camera ray per pixel from a private RNG stream,
 * and the queue hands out the pixels whose probe hits the scene first); and
 * methods that do not: -DHT_RR_DEPTH=d (Russian roulette),
 * -DHT_HY_DIST_FP64=1 (hyperbolic distance in FP64, cancellation-free),
 * -DHT_EPS32=<float> (FP32 programs: this hit window instead of the
 * reference's 1e-6, e.g. 1e-5f for scenes at world scale ~10, where FP32
 * roundings reach 1e-6). -DHT_NO_CULL=1 selects the element-by-element
 * scan (same images, for comparisons). No reference counterpart beyond
 * ocl::Program build options. */
int ht_prog_build_ex(ht_ctx *ctx, const char *root_name, const char *root_src, int n_headers,
                     const char *const *names, const char *const *contents, int double_precision,
                     int address_width, const char *fp_options, ht_prog **out);
void ht_prog_destroy(ht_prog *prog);
/* 1 if the last build on this thread was served from the cache. */
int ht_prog_cache_hit(void);
/* Resources of one kernel of a program (e.g. "render_passes"): vector
 * registers per lane, scratch bytes per lane (register spills) and static
 * LDS bytes per block; NULL outputs are skipped. No reference counterpart:
 * the render tuner reads it to keep programs that spill out of its picks. */
int ht_prog_kernel_info(ht_prog *prog, const char *kernel, int *vgprs, int *scratch_bytes, int *lds_bytes);
/* Compile only, into the code-object cache; needs no GPU. No reference
 * counterpart: lets a build step prepare programs ahead of time. */
int ht_prog_precompile(const char *root_name, const char *root_src, int n_headers,
                       const char *const *names, const char *const *contents, int double_precision,
                       int address_width, const char *fp_options);

/* Device buffers. Replace ocl::Buffer use in processing/src/buffer.rs:8-34
 * (EntityBuffer) and canvas.rs:125-189 (ImageBuffer). Writes copy the host
 * bytes; reads block until the data is on the host (canvas.rs:183-188). */
int ht_buf_alloc(ht_ctx *ctx, size_t bytes, ht_buf **out);
int ht_buf_free(ht_buf *buf);
size_t ht_buf_size(const ht_buf *buf);
int ht_buf_write(ht_buf *buf, const void *src, size_t bytes);
int ht_buf_read(ht_buf *buf, void *dst, size_t bytes);
int ht_buf_zero(ht_buf *buf);

/* One pass: one path per pixel, canvas[idx] += (radiance, 1), seeds advanced.
 * Synchronous like Render::render (processing/src/render.rs:102-118, the
 * queue.finish() at :115). canvas: float4[w*h]; seeds: uint32[w*h]. */
int ht_render(ht_prog *prog, uint32_t w, uint32_t h, ht_buf *scene, ht_buf *canvas_f32x4,
              ht_buf *seeds_u32);

/* `passes` consecutive passes in one launch with in-thread path
 * regeneration; canvas and seeds end up identical to `passes` calls of
 * ht_render. No reference counterpart (the reference runs one launch + one
 * finish per pass). Asynchronous: completes on the context's stream. */
int ht_render_passes(ht_prog *prog, uint32_t w, uint32_t h, ht_buf *scene, ht_buf *canvas_f32x4,
                     ht_buf *seeds_u32, uint32_t passes);

/* The same result as ht_render_passes (canvas and seeds bit for bit), as a
 * wavefront: paths live in SoA queues in device memory and advance one hit
 * per iteration through an extend kernel (nearest hit; hits sorted into
 * queues by the material component they take next) and a shade kernel
 * (material step in queue order, regeneration of the pixel's next sample).
 * The program's root source must add HT_DEFINE_WAVEFRONT_KERNELS(Scene)
 * (codegen.root_source(..., wavefront=True)); its workspace (2 extend and 4
 * shade queues of one record per pixel, ~550 B per pixel for the cell
 * scene) is kept by the program and grown on demand. Synchronous. No reference counterpart: the form the
 * megakernel is measured against. */
int ht_render_wavefront(ht_prog *prog, uint32_t w, uint32_t h, ht_buf *scene, ht_buf *canvas_f32x4,
                        ht_buf *seeds_u32, uint32_t passes);

/* ht_render_wavefront without any allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * queues live in one fixed array of HT_WF_SCRATCH_BYTES per GPU, part of the
 * library and held by each GPU it runs on until the process ends, instead
 * of the program's workspace; frames whose queues need more (about 550 B
 * per pixel for the cell scene: ~120 k pixels) are refused with
 * HT_ERR_OTHER. Every call on a GPU shares that GPU's array, so calls are
 * serialised within a process by a lock and each returns with its stream
 * idle; calls that overlapped on one GPU would race. Same canvas and seeds as
 * ht_render_wavefront. Synchronous. */
#define HT_WF_SCRATCH_BYTES (64u << 20)
int ht_render_wavefront_noalloc(ht_prog *prog, uint32_t w, uint32_t h, ht_buf *scene, ht_buf *canvas_f32x4,
                                ht_buf *seeds_u32, uint32_t passes);

/* Post passes (asynchronous until a read). image, out: float4[w*h];
 * rgba8: uchar4[w*h].
 * ht_extract: image = canvas / (float)passes   (processing/src/canvas.rs:126-135)
 * ht_gamma:   out = pow(max(in, 0), gamma)      (processing/src/filter/gamma.rs:10-19)
 * ht_pack:    rgba8 = (uchar)(255 * clamp(x, 0, 1)) truncated (canvas.rs:176-186)
 * ht_post:    the three fused in one kernel, bit-identical to running them
 *             in order (Pipeline::get_image, processing/src/pipeline.rs:53-62). */
int ht_extract(ht_ctx *ctx, uint32_t w, uint32_t h, uint32_t passes, ht_buf *canvas, ht_buf *image);
int ht_gamma(ht_ctx *ctx, uint32_t w, uint32_t h, float gamma, ht_buf *in, ht_buf *out);
int ht_pack(ht_ctx *ctx, uint32_t w, uint32_t h, ht_buf *image, ht_buf *rgba8);
int ht_post(ht_ctx *ctx, uint32_t w, uint32_t h, uint32_t passes, float gamma, ht_buf *canvas,
            ht_buf *rgba8);

/* Scene maps built on the GPU in FP64. Replaces the host-side composition of
 * hyperbolic isometries in f64 and their narrowing to the device `real`
 * (the vecmat/ccgeom products the examples chain, main/examples/hy.rs:37-183,
 * shipped through types/src/primitive/scalar.rs:198-207). For element i:
 *   M = N(...N(N(I * G[w0]) * G[w1]) ... * P_i),  N(m) = m / sqrt(det m)
 * (the det = 1 reprojection of c22_normalize, algebra/matrix_complex.cc:50-52)
 * over its word w = letters[i*HT_MO_WORD_MAX .. + lengths[i]) of generator
 * indices, all in FP64 with each operation rounded on its own, written as 8
 * reals (a, b, c, d as (re, im)) at byte offset offsets[i] of `scene`, in FP32
 * (real_size 4, round to nearest) or FP64 (real_size 8).
 * gens: n_gens x 8 doubles; posts: n x 8 doubles. Synchronous. */
#define HT_MO_WORD_MAX 16
int ht_mo_compose(ht_ctx *ctx, ht_buf *scene, int real_size, uint32_t n, const uint32_t *offsets,
                  const uint8_t *letters, const uint8_t *lengths, const double *gens, uint32_t n_gens,
                  const double *posts);

/* ---- Learned radiance cache. No reference counterpart (the reference has no
 * learned part); these extend the render pass with a cached method whose
 * error is measured against the FP64 program.
 *
 * MLP: 12 features (position, direction, normal, albedo), each encoded as
 * sin / cos of 2^b pi x for b < 6 (144 inputs), 5 hidden layers of 64 with
 * ReLU, 3 outputs (RGB radiance); BF16 weights and activations on MFMA with
 * FP32 accumulation; FP32 master parameters laid out as
 *   W0[64][144] b0[64] (W_l[64][64] b_l[64], l = 1..4) W5[3][64] b5[3].
 * Loss: relative L2, mean over samples x channels of (p - y)^2 / (p^2 + 0.01).
 * Adam: beta1 0.9, beta2 0.99, eps 1e-8, bias-corrected. */
#define HT_MLP_PARAMS 26115
#define HT_MLP_FEATURES 12
typedef struct ht_cache ht_cache;

/* Query queue of `query_capacity` entries (>= pixels per pass), training
 * queue of `train_capacity` records, training batches up to `max_batch`
 * samples (a multiple of 64). Parameters start at zero. */
int ht_cache_create(ht_ctx *ctx, uint32_t query_capacity, uint32_t train_capacity, uint32_t max_batch,
                    ht_cache **out);
void ht_cache_destroy(ht_cache *cache);
/* Master parameters (HT_MLP_PARAMS floats); setting them resets Adam. */
int ht_cache_set_params(ht_cache *cache, const float *params, size_t n);
int ht_cache_get_params(ht_cache *cache, float *params, size_t n);
/* Gradient of the last training batch (before any Adam update). */
int ht_cache_get_grads(ht_cache *cache, float *grads, size_t n);
/* Entries appended to the query and training queues by the last pass
 * (training records beyond the capacity are dropped, the count goes on). */
int ht_cache_counts(ht_cache *cache, uint32_t *queries, uint32_t *records);

/* Network on explicit data, synchronous. features: float [12][n] (feature-
 * major); pred, targets: float [3][n]. ht_mlp_train_step: one batch of n (a
 * multiple of 64, <= max_batch), forward + backward; apply = 1 also takes
 * the Adam step with learning rate lr; *loss (optional) = the batch loss. */
int ht_mlp_forward(ht_cache *cache, ht_buf *features, uint32_t n, ht_buf *pred);
int ht_mlp_train_step(ht_cache *cache, ht_buf *features, ht_buf *targets, uint32_t n, float lr, int apply,
                      float *loss);

/* One cached pass (asynchronous): the scene program's render_cached kernel
 * (paths end into the cache at their query_hop-th hit, 0 = the first; one
 * pixel in train_every, by a hash of pixel and pass_index, runs unbiased to
 * full depth and records a training sample), then the MLP inference that
 * adds throughput x prediction to each queried pixel. canvas gains one pass.
 * Needs a program whose scene is a vector of mapped covered objects. */
int ht_render_cached(ht_prog *prog, ht_cache *cache, uint32_t w, uint32_t h, ht_buf *scene, ht_buf *canvas,
                     ht_buf *seeds, uint32_t pass_index, uint32_t train_every, uint32_t query_hop);
/* `steps` Adam steps of `batch` samples on the training records of the last
 * pass (sample i of step k is record (k * batch + i) mod count). Asynchronous. */
int ht_cache_train(ht_cache *cache, uint32_t steps, uint32_t batch, float lr);

/* ---- Batched render: many jobs (scenes of the batch program's two types,
 * views, sizes, spp, depths) in one persistent launch. No reference
 * counterpart: the reference builds one program per scene type and renders
 * one scene per launch (processing/src/render.rs:14-118).
 *
 * A batch program is built with ht_prog_build_ex from a root source that
 * instantiates HT_DEFINE_BATCH_KERNEL(Scene0, Scene1) (codegen.batch_source);
 * its block size is the kernel's launch bound; `tile` (pixels per work item,
 * a multiple of the block size; of 64 for a program built with
 * -DHT_BATCH_WAVE_QUEUE=1, whose waves take tiles on their own and hand the
 * tile's pixels to their lanes as they finish) is chosen at
 * ht_batch_create. Per pixel the
 * work equals ht_render_passes of the job's scene with at most `depth` hits
 * per path, so canvases and seeds are bit-identical to rendering each job
 * alone. A batch program built with -DHT_SCALAR_BOUNDS=1 reads each job's
 * culling bounds from its own FP64 array (allocated by ht_batch_create,
 * 64 B x the program's bounds capacity per job, filled at the start of every
 * frame by the program's batch_bounds_fill kernel) instead of staging them
 * in LDS per tile; same images. The frame's tiles are handed out job after
 * job, the jobs in descending order of spp x depth (the heaviest tiles
 * first); the order changes no result. */
typedef struct ht_batch ht_batch;
typedef struct ht_batch_job {
    uint32_t type;             /* 0 or 1: Scene0 or Scene1 of the batch program */
    uint32_t w, h, spp, depth; /* image, passes per frame, hits per path (<= the type's HOPS) */
    ht_buf *scene, *canvas, *seeds;
    ht_buf *rgba8;             /* optional: post-pass output of each frame */
} ht_batch_job;
int ht_batch_create(ht_prog *batch_prog, uint32_t njobs, const ht_batch_job *jobs, uint32_t tile, ht_batch **out);
void ht_batch_destroy(ht_batch *batch);
uint32_t ht_batch_tiles(const ht_batch *batch);
uint32_t ht_batch_block(const ht_batch *batch);
/* One frame on the context's stream (asynchronous): clear = zero every
 * job's canvas first (a frame of `spp` passes; else the passes accumulate),
 * then the persistent launch over `grid` blocks; post = ht_post of each job
 * with an rgba8 buffer (needs clear). */
int ht_batch_render(ht_batch *batch, uint32_t grid, int clear, float gamma, int post);
/* The same frame captured into a hipGraph (clears by kernels, no memset
 * nodes), and its replay. */
int ht_batch_capture(ht_batch *batch, uint32_t grid, int clear, float gamma, int post);
int ht_batch_replay(ht_batch *batch);

/* Per-pixel seed buffer. Replaces Canvas::new's seed fill
 * (processing/src/canvas.rs:73-84).
 * mode 0: rand 0.8 SmallRng::seed_from_u64(seed).fill(&mut [u32]) on 64-bit
 *         targets: xoshiro256++ whose 32-byte seed is rand_core 0.6's
 *         default seed_from_u64 expansion (8 PCG32 words), since SmallRng
 *         does not override it; the reference uses seed = 0xdeadbeefdeadbeef.
 * mode 1: seeds[i] = ht_seed_hash(seed, i) (per-pixel hash seeding).
 * mode 2: as mode 0 with the generator seeded by SplitMix64, as
 *         rand_xoshiro's own Xoshiro256PlusPlus::seed_from_u64 does (what a
 *         SmallRng that forwarded seed_from_u64 to its generator would give). */
int ht_seed_init(ht_buf *seeds, uint32_t w, uint32_t h, int mode, uint64_t seed);
/* Host-side seed generation, same modes, into a caller array of n words. */
int ht_seed_fill(uint32_t *dst, size_t n, int mode, uint64_t seed);
uint32_t ht_seed_hash(uint64_t seed, uint32_t index);

/* Timing on the context's stream with HIP events: ht_timer_begin records a
 * start event; ht_timer_end records a stop event, waits for it and returns
 * the elapsed milliseconds. */
int ht_timer_begin(ht_ctx *ctx);
int ht_timer_end(ht_ctx *ctx, float *ms);

#ifdef __cplusplus
}
#endif

#endif /* HYPERTRACE_HIP_H */
