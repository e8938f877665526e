// C-ABI runtime of the render pass on one MI355X (gfx950): contexts, device
// buffers, hipRTC program build with an on-disk code-object cache, render
// launches, the post passes and the seed buffer. Declarations and the
// reference code each entry point replaces are in include/hypertrace_hip.h.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>
#include <utime.h>

#include <atomic>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hh"

using ht_internal::ctx_release;
using ht_internal::ctx_retain;
using ht_internal::fail;

namespace ht_internal {
thread_local std::string g_err = "";
int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
}  // namespace ht_internal

namespace {

thread_local int g_cache_hit = 0;

const char *kArch = "gfx950";
const char *kCacheTag = "ht-codeobj-v1";

uint64_t fnv1a(uint64_t h, const void *data, size_t n) {
    const unsigned char *p = static_cast<const unsigned char *>(data);
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}
uint64_t fnv1a_str(uint64_t h, const std::string &s) {
    uint64_t n = s.size();
    h = fnv1a(h, &n, sizeof(n));
    return fnv1a(h, s.data(), s.size());
}

std::string cache_dir() {
    const char *env = std::getenv("HT_CACHE_DIR");
    if (env && *env) return env;
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&ht_last_error), &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        size_t slash = p.rfind('/');
        std::string dir = slash == std::string::npos ? "." : p.substr(0, slash);
        return dir + "/../_cache";
    }
    return "./_cache";
}

bool read_file(const std::string &path, std::vector<char> *out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    f.seekg(0, std::ios::end);
    std::streamoff n = f.tellg();
    if (n <= 0) return false;
    f.seekg(0, std::ios::beg);
    out->resize(static_cast<size_t>(n));
    f.read(out->data(), n);
    return static_cast<bool>(f);
}

void write_file_atomic(const std::string &dir, const std::string &path, const std::vector<char> &data) {
    mkdir(dir.c_str(), 0755);
    std::string tmp = path + ".tmp" + std::to_string(getpid());
    {
        std::ofstream f(tmp, std::ios::binary);
        if (!f) return;
        f.write(data.data(), static_cast<std::streamsize>(data.size()));
        if (!f) return;
    }
    std::rename(tmp.c_str(), path.c_str());
}

}  // namespace

// ------------------------------------------------------------------ post kernels

// image = canvas / (float)passes, processing/src/canvas.rs:126-135.
__global__ void __launch_bounds__(256) k_extract(uint32_t n, uint32_t passes,
                                                 const float4 *__restrict__ canvas, float4 *__restrict__ image) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float p = (float)passes;
    float4 c = canvas[i];
    image[i] = make_float4(c.x / p, c.y / p, c.z / p, c.w / p);
}

__device__ __forceinline__ float gamma1(float x, float g) { return powf(fmaxf(x, 0.0f), g); }

// out = pow(max(in, 0), gamma), processing/src/filter/gamma.rs:10-19.
__global__ void __launch_bounds__(256) k_gamma(uint32_t n, float g, const float4 *__restrict__ in,
                                               float4 *__restrict__ out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 c = in[i];
    out[i] = make_float4(gamma1(c.x, g), gamma1(c.y, g), gamma1(c.z, g), gamma1(c.w, g));
}

// (uchar)(255 * clamp(x, 0, 1)), conversion toward zero; clamp = fmin(fmax(x,0),1)
// so NaN packs to 0. processing/src/canvas.rs:176-186.
__device__ __forceinline__ uint32_t pack1(float x) {
    return (uint32_t)(255.0f * fminf(fmaxf(x, 0.0f), 1.0f));
}
__device__ __forceinline__ uint32_t pack4(float4 c) {
    return pack1(c.x) | (pack1(c.y) << 8) | (pack1(c.z) << 16) | (pack1(c.w) << 24);
}

__global__ void __launch_bounds__(256) k_pack(uint32_t n, const float4 *__restrict__ image,
                                              uint32_t *__restrict__ rgba8) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rgba8[i] = pack4(image[i]);
}

// extract -> gamma -> pack in one pass: 16 B read + 4 B written per pixel.
__global__ void __launch_bounds__(256) k_post(uint32_t n, uint32_t passes, float g,
                                              const float4 *__restrict__ canvas, uint32_t *__restrict__ rgba8) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float p = (float)passes;
    float4 c = canvas[i];
    float4 e = make_float4(c.x / p, c.y / p, c.z / p, c.w / p);
    float4 v = make_float4(gamma1(e.x, g), gamma1(e.y, g), gamma1(e.z, g), gamma1(e.w, g));
    rgba8[i] = pack4(v);
}

// ------------------------------------------------------------------ Moebius words

// FP64 composition of Moebius maps with det = 1 reprojection after every
// product. Every operation rounds on its own (no contraction) and the square
// roots are IEEE, so the result is bit-identical to the same operations in
// any IEEE double arithmetic (the host mirror in hypertrace_amd/maps.py).
struct Cd {
    double x, y;
};
__device__ Cd cd_mul(Cd a, Cd b) {
#pragma clang fp contract(off)
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ Cd cd_add(Cd a, Cd b) {
#pragma clang fp contract(off)
    return {a.x + b.x, a.y + b.y};
}
__device__ Cd cd_sub(Cd a, Cd b) {
#pragma clang fp contract(off)
    return {a.x - b.x, a.y - b.y};
}
// Principal square root, the cancellation-free branch for each half-plane.
__device__ Cd cd_sqrt(Cd z) {
#pragma clang fp contract(off)
    double r = __builtin_sqrt(z.x * z.x + z.y * z.y);
    if (z.x >= 0.0) {
        double re = __builtin_sqrt((r + z.x) * 0.5);
        return {re, z.y / (re + re)};
    }
    double im = __builtin_sqrt((r - z.x) * 0.5);
    if (z.y < 0.0) im = -im;
    return {z.y / (im + im), im};
}
struct C22d {
    Cd a, b, c, d;
};
__device__ C22d c22d_mul(const C22d &p, const C22d &q) {
    return {cd_add(cd_mul(p.a, q.a), cd_mul(p.b, q.c)), cd_add(cd_mul(p.a, q.b), cd_mul(p.b, q.d)),
            cd_add(cd_mul(p.c, q.a), cd_mul(p.d, q.c)), cd_add(cd_mul(p.c, q.b), cd_mul(p.d, q.d))};
}
// m / sqrt(det m): det = 1 reprojection (c22_normalize, algebra/matrix_complex.cc:50-52,
// with an algebraic square root in place of the polar form).
__device__ C22d c22d_renorm(const C22d &m) {
#pragma clang fp contract(off)
    Cd s = cd_sqrt(cd_sub(cd_mul(m.a, m.d), cd_mul(m.b, m.c)));
    double s2 = s.x * s.x + s.y * s.y;
    Cd inv = {s.x / s2, -s.y / s2};
    return {cd_mul(m.a, inv), cd_mul(m.b, inv), cd_mul(m.c, inv), cd_mul(m.d, inv)};
}
__device__ C22d c22d_load(const double *v) { return {{v[0], v[1]}, {v[2], v[3]}, {v[4], v[5]}, {v[6], v[7]}}; }

// Element i: M = N(N(...N(N(I G[w0]) G[w1])...) P_i) over its word w (N =
// reprojection), stored as 8 reals (a, b, c, d as (re, im) pairs) at
// scene + offsets[i], in FP32 (round to nearest) or FP64.
__global__ void __launch_bounds__(64) k_mo_compose(uint8_t *__restrict__ scene, uint32_t n, int real_size,
                                                   const uint32_t *__restrict__ offsets,
                                                   const uint8_t *__restrict__ letters,
                                                   const uint8_t *__restrict__ lengths,
                                                   const double *__restrict__ gens,
                                                   const double *__restrict__ posts) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    C22d m = {{1.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {1.0, 0.0}};
    uint32_t len = lengths[i];
    for (uint32_t k = 0; k < len; ++k) {
        C22d g = c22d_load(gens + 8 * (size_t)letters[(size_t)i * HT_MO_WORD_MAX + k]);
        m = c22d_renorm(c22d_mul(m, g));
    }
    m = c22d_renorm(c22d_mul(m, c22d_load(posts + 8 * (size_t)i)));
    double v[8] = {m.a.x, m.a.y, m.b.x, m.b.y, m.c.x, m.c.y, m.d.x, m.d.y};
    uint8_t *dst = scene + offsets[i];
    if (real_size == 4) {
        float *f = reinterpret_cast<float *>(dst);
        for (int k = 0; k < 8; ++k) f[k] = (float)v[k];
    } else {
        double *d = reinterpret_cast<double *>(dst);
        for (int k = 0; k < 8; ++k) d[k] = v[k];
    }
}

// ------------------------------------------------------------------ seeds

namespace {

// SplitMix64 step: the seeding of rand_xoshiro's own
// Xoshiro256PlusPlus::seed_from_u64 (and of an all-zero from_seed).
uint64_t splitmix64(uint64_t *state) {
    *state += 0x9e3779b97f4a7c15ull;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// One word of rand_core 0.6's default SeedableRng::seed_from_u64: advance a
// 64-bit LCG state (PCG's multiplier and a fixed increment), then the PCG32
// XSH-RR output of the new state.
uint32_t pcg32_seed_word(uint64_t *state) {
    *state = *state * 6364136223846793005ull + 11634580027462260723ull;
    uint64_t s = *state;
    uint32_t xorshifted = (uint32_t)(((s >> 18) ^ s) >> 27);
    uint32_t rot = (uint32_t)(s >> 59);
    return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
}

inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Xoshiro256pp {
    uint64_t s[4];
    uint64_t next() {
        uint64_t result = rotl(s[0] + s[3], 23) + s[0];
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
};

}  // namespace

extern "C" uint32_t ht_seed_hash(uint64_t seed, uint32_t index) {
    // murmur3 fmix64 of (seed + golden * (index + 1)), high word.
    uint64_t z = seed + 0x9e3779b97f4a7c15ull * ((uint64_t)index + 1);
    z ^= z >> 33;
    z *= 0xff51afd7ed558ccdull;
    z ^= z >> 33;
    z *= 0xc4ceb9fe1a85ec53ull;
    z ^= z >> 33;
    return (uint32_t)(z >> 32);
}

extern "C" int ht_seed_fill(uint32_t *dst, size_t n, int mode, uint64_t seed) {
    if (!dst && n) return fail(HT_ERR_OTHER, "ht_seed_fill: null destination");
    if (mode == 0 || mode == 2) {
        Xoshiro256pp rng;
        uint64_t st = seed;
        if (mode == 0) {
            // SmallRng::seed_from_u64 of rand 0.8: SmallRng does not override
            // the trait's default, which fills the 32-byte seed with 8 words of
            // pcg32_seed_word (little-endian); Xoshiro256PlusPlus::from_seed
            // reads it as 4 little-endian u64.
            for (int i = 0; i < 4; ++i) {
                uint64_t lo = pcg32_seed_word(&st), hi = pcg32_seed_word(&st);
                rng.s[i] = lo | (hi << 32);
            }
        } else {
            // rand_xoshiro's Xoshiro256PlusPlus::seed_from_u64: SplitMix64
            for (int i = 0; i < 4; ++i) rng.s[i] = splitmix64(&st);
        }
        if (!(rng.s[0] | rng.s[1] | rng.s[2] | rng.s[3])) {
            // from_seed of an all-zero seed: seed_from_u64(0) of the generator
            st = 0;
            for (int i = 0; i < 4; ++i) rng.s[i] = splitmix64(&st);
        }
        // Rng::fill on [u32] fills the byte view: 8 bytes per next_u64
        // (little-endian: low word first); a trailing 4-byte remainder takes
        // next_u32 = the high word of next_u64.
        size_t i = 0;
        for (; i + 2 <= n; i += 2) {
            uint64_t v = rng.next();
            dst[i] = (uint32_t)v;
            dst[i + 1] = (uint32_t)(v >> 32);
        }
        if (i < n) dst[i] = (uint32_t)(rng.next() >> 32);
        return HT_OK;
    }
    if (mode == 1) {
        for (size_t i = 0; i < n; ++i) dst[i] = ht_seed_hash(seed, (uint32_t)i);
        return HT_OK;
    }
    return fail(HT_ERR_OTHER, "ht_seed_fill: unknown mode " + std::to_string(mode));
}

// ------------------------------------------------------------------ API

extern "C" const char *ht_last_error(void) { return ht_internal::g_err.c_str(); }

extern "C" int ht_ctx_create(int device_index, ht_ctx **out) {
    if (!out) return fail(HT_ERR_OTHER, "ht_ctx_create: null out");
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device_index < 0 || device_index >= count)
        return fail(HT_ERR_OTHER, "ht_ctx_create: device index " + std::to_string(device_index) + " out of range (" +
                                      std::to_string(count) + " devices)");
    HIP_TRY(hipSetDevice(device_index));
    ht_ctx *c = new ht_ctx();
    c->device = device_index;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(HT_ERR_BACKEND, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    if (hipEventCreate(&c->t0) != hipSuccess || hipEventCreate(&c->t1) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(HT_ERR_BACKEND, "hipEventCreate failed");
    }
    *out = c;
    return HT_OK;
}

extern "C" void ht_ctx_destroy(ht_ctx *ctx) {
    if (!ctx) return;
    ctx_release(ctx);
}

extern "C" int ht_ctx_sync(ht_ctx *ctx) {
    if (!ctx) return fail(HT_ERR_OTHER, "ht_ctx_sync: null ctx");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return HT_OK;
}

extern "C" int ht_prog_cache_hit(void) { return g_cache_hit; }

namespace {

// Compiles the program with hipRTC, or loads its code object from the cache.
// Floating-point code-generation options of a program build when the caller
// gives none: every operation rounded on its own (no FMA contraction, IEEE
// division and square root) and the device library's strict mode.
const char *kDefaultFpOptions = "-ffp-contract=off -DHT_STRICT_FP=1";

int get_code_object(const char *root_name, const char *root_src, int n_headers, const char *const *names,
                    const char *const *contents, int double_precision, int address_width, const char *fp_options,
                    std::vector<char> *code) {
    g_cache_hit = 0;
    if (!root_src || n_headers < 0 || (n_headers > 0 && (!names || !contents)))
        return fail(HT_ERR_OTHER, "program build: invalid argument");
    if (address_width != 32 && address_width != 64)
        return fail(HT_ERR_OTHER, "program build: address_width must be 32 or 64");
    std::string name = root_name ? root_name : "render.hip";
    std::vector<std::string> opts = {
        std::string("--offload-arch=") + kArch,
        "-O3",
        "-std=c++17",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-DHT_ADDRESS_WIDTH=" + std::to_string(address_width),
    };
    if (double_precision) opts.push_back("-DHT_DOUBLE=1");
    {
        std::istringstream fp(fp_options ? fp_options : kDefaultFpOptions);
        std::string o;
        while (fp >> o) opts.push_back(o);
    }

    if (const char *dump = std::getenv("HT_DUMP_SOURCE")) {
        if (*dump) {
            std::ofstream f("render.hip");
            f << root_src;
        }
    }

    uint64_t h = 1469598103934665603ull;
    h = fnv1a_str(h, kCacheTag);
    for (const auto &o : opts) h = fnv1a_str(h, o);
    h = fnv1a_str(h, name);
    h = fnv1a_str(h, root_src);
    for (int i = 0; i < n_headers; ++i) {
        h = fnv1a_str(h, names[i] ? names[i] : "");
        h = fnv1a_str(h, contents[i] ? contents[i] : "");
    }
    char key[32];
    std::snprintf(key, sizeof key, "%016llx", (unsigned long long)h);
    std::string dir = cache_dir();
    std::string path = dir + "/" + key + "." + kArch + ".hsaco";

    if (read_file(path, code)) {
        g_cache_hit = 1;
        // mark the entry as used (a build step prunes entries it did not use;
        // a read-only cache directory is fine)
        (void)utime(path.c_str(), nullptr);
        return HT_OK;
    }
    if (const char *log_compiles = std::getenv("HT_LOG_COMPILES")) {
        // diagnostics: which programs a run compiles instead of loading them from the cache
        if (*log_compiles) std::fprintf(stderr, "ht: hipRTC compile of %s (%s, options: %s)\n", name.c_str(), key,
                                        fp_options ? fp_options : kDefaultFpOptions);
    }
    hiprtcProgram prog;
    hiprtcResult r = hiprtcCreateProgram(&prog, root_src, name.c_str(), n_headers, contents, names);
    if (r != HIPRTC_SUCCESS) return fail(HT_ERR_BACKEND, std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r));
    std::vector<const char *> copts;
    for (const auto &o : opts) copts.push_back(o.c_str());
    r = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    std::string log(log_size, '\0');
    if (log_size) hiprtcGetProgramLog(prog, &log[0]);
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return fail(HT_ERR_BACKEND, std::string("hipRTC build of ") + name + " failed: " + hiprtcGetErrorString(r) + "\n" + log);
    }
    size_t code_size = 0;
    hiprtcGetCodeSize(prog, &code_size);
    code->resize(code_size);
    hiprtcGetCode(prog, code->data());
    hiprtcDestroyProgram(&prog);
    write_file_atomic(dir, path, *code);
    return HT_OK;
}

}  // namespace

extern "C" int ht_prog_build_ex(ht_ctx *ctx, const char *root_name, const char *root_src, int n_headers,
                                const char *const *names, const char *const *contents, int double_precision,
                                int address_width, const char *fp_options, ht_prog **out) {
    if (!ctx || !out) return fail(HT_ERR_OTHER, "ht_prog_build: invalid argument");
    std::vector<char> code;
    int rc = get_code_object(root_name, root_src, n_headers, names, contents, double_precision, address_width,
                             fp_options, &code);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    ht_prog *p = new ht_prog();
    p->ctx = ctx;
    hipError_t e = hipModuleLoadData(&p->module, code.data());
    if (e != hipSuccess) {
        delete p;
        return fail(HT_ERR_BACKEND, std::string("hipModuleLoadData: ") + hipGetErrorString(e));
    }
    // Scene programs have render/render_passes (and render_cached where the
    // radiance cache applies); batch programs have render_batch.
    auto find = [&](const char *name) -> hipFunction_t {
        hipFunction_t f = nullptr;
        if (hipModuleGetFunction(&f, p->module, name) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        return f;
    };
    p->render = find("render");
    p->render_passes = find("render_passes");
    p->render_cached = find("render_cached");
    p->render_batch = find("render_batch");
    p->bounds_fill = find("bounds_fill");
    p->pixel_queue_reset = find("pixel_queue_reset");
    p->pixel_classes = find("pixel_classes");
    if (p->pixel_queue_reset && p->render_passes) {
        int per_cu = 0, cus = 0;
        e = hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, p->render_passes, 256, 0);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
        if (e != hipSuccess || per_cu <= 0 || cus <= 0) {
            (void)hipModuleUnload(p->module);
            delete p;
            return fail(HT_ERR_BACKEND, "pixel-queue program: occupancy query failed");
        }
        p->resident_blocks = (uint32_t)(per_cu * cus);
    }
    p->batch_bounds_fill = find("batch_bounds_fill");
    if (p->batch_bounds_fill) {
        hipDeviceptr_t d = nullptr;
        size_t bytes = 0;
        e = hipModuleGetGlobal(&d, &bytes, p->module, "ht_batch_bounds_cap");
        if (e == hipSuccess && bytes == sizeof(unsigned)) e = hipMemcpyDtoH(&p->batch_bounds_cap, d, bytes);
        if (e != hipSuccess || bytes != sizeof(unsigned) || p->batch_bounds_cap == 0) {
            (void)hipModuleUnload(p->module);
            delete p;
            return fail(HT_ERR_BACKEND, "batch program has batch_bounds_fill but no readable ht_batch_bounds_cap");
        }
    }
    if (p->render_batch) {
        hipDeviceptr_t d = nullptr;
        size_t bytes = 0;
        p->batch_wave_tiles = hipModuleGetGlobal(&d, &bytes, p->module, "ht_batch_wave_tiles") == hipSuccess;
        (void)hipGetLastError();  // absent in programs without the option
    }
    p->wf_generate = find("wf_generate");
    p->wf_extend = find("wf_extend");
    p->wf_shade = find("wf_shade");
    if (p->wf_generate && p->wf_extend && p->wf_shade) {
        hipDeviceptr_t d = nullptr;
        size_t bytes = 0;
        e = hipModuleGetGlobal(&d, &bytes, p->module, "ht_wf_layout");
        if (e == hipSuccess && bytes == sizeof(p->wf_layout)) e = hipMemcpyDtoH(p->wf_layout, d, bytes);
        if (e != hipSuccess || bytes != sizeof(p->wf_layout)) {
            (void)hipModuleUnload(p->module);
            delete p;
            return fail(HT_ERR_BACKEND, "program has wavefront kernels but no readable ht_wf_layout");
        }
    } else {
        p->wf_generate = p->wf_extend = p->wf_shade = nullptr;
    }
    if (!(p->render && p->render_passes) && !p->render_batch) {
        (void)hipModuleUnload(p->module);
        delete p;
        return fail(HT_ERR_BACKEND, "program lacks the render/render_passes (or render_batch) kernels");
    }
    ctx_retain(ctx);
    *out = p;
    return HT_OK;
}

extern "C" int ht_prog_build(ht_ctx *ctx, const char *root_name, const char *root_src, int n_headers,
                             const char *const *names, const char *const *contents, int double_precision,
                             int address_width, ht_prog **out) {
    return ht_prog_build_ex(ctx, root_name, root_src, n_headers, names, contents, double_precision, address_width,
                            nullptr, out);
}

extern "C" int ht_prog_precompile(const char *root_name, const char *root_src, int n_headers,
                                  const char *const *names, const char *const *contents, int double_precision,
                                  int address_width, const char *fp_options) {
    std::vector<char> code;
    return get_code_object(root_name, root_src, n_headers, names, contents, double_precision, address_width,
                           fp_options, &code);
}

extern "C" int ht_prog_kernel_info(ht_prog *prog, const char *kernel, int *vgprs, int *scratch_bytes,
                                   int *lds_bytes) {
    if (!prog || !kernel) return fail(HT_ERR_OTHER, "ht_prog_kernel_info: invalid argument");
    HIP_TRY(hipSetDevice(prog->ctx->device));
    hipFunction_t f = nullptr;
    if (hipModuleGetFunction(&f, prog->module, kernel) != hipSuccess) {
        (void)hipGetLastError();
        return fail(HT_ERR_OTHER, std::string("ht_prog_kernel_info: the program has no kernel ") + kernel);
    }
    int v = 0, scratch = 0, lds = 0;
    HIP_TRY(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, f));
    HIP_TRY(hipFuncGetAttribute(&scratch, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, f));
    HIP_TRY(hipFuncGetAttribute(&lds, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, f));
    if (vgprs) *vgprs = v;
    if (scratch_bytes) *scratch_bytes = scratch;
    if (lds_bytes) *lds_bytes = lds;
    return HT_OK;
}

extern "C" void ht_prog_destroy(ht_prog *prog) {
    if (!prog) return;
    (void)hipSetDevice(prog->ctx->device);
    if (prog->wf_mem) {
        (void)hipStreamSynchronize(prog->ctx->stream);
        (void)hipFree(prog->wf_mem);
    }
    (void)hipModuleUnload(prog->module);
    (void)hipGetLastError();
    ht_ctx *ctx = prog->ctx;
    delete prog;
    ctx_release(ctx);
}

extern "C" int ht_buf_alloc(ht_ctx *ctx, size_t bytes, ht_buf **out) {
    if (!ctx || !out) return fail(HT_ERR_OTHER, "ht_buf_alloc: invalid argument");
    HIP_TRY(hipSetDevice(ctx->device));
    ht_buf *b = new ht_buf();
    b->ctx = ctx;
    b->size = bytes;
    if (bytes) {
        hipError_t e = hipMalloc(&b->ptr, bytes);
        if (e != hipSuccess) {
            delete b;
            return fail(HT_ERR_BACKEND, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
    }
    ctx_retain(ctx);
    *out = b;
    return HT_OK;
}

extern "C" int ht_buf_free(ht_buf *buf) {
    if (!buf) return HT_OK;
    ht_ctx *ctx = buf->ctx;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && buf->ptr) e = hipFree(buf->ptr);
    delete buf;
    ctx_release(ctx);
    if (e != hipSuccess) return fail(HT_ERR_BACKEND, std::string("ht_buf_free: ") + hipGetErrorString(e));
    return HT_OK;
}

extern "C" size_t ht_buf_size(const ht_buf *buf) { return buf ? buf->size : 0; }

extern "C" int ht_buf_write(ht_buf *buf, const void *src, size_t bytes) {
    if (!buf || (!src && bytes)) return fail(HT_ERR_OTHER, "ht_buf_write: invalid argument");
    if (bytes > buf->size) return fail(HT_ERR_OTHER, "ht_buf_write: " + std::to_string(bytes) + " bytes into a buffer of " + std::to_string(buf->size));
    if (!bytes) return HT_OK;
    HIP_TRY(hipSetDevice(buf->ctx->device));
    HIP_TRY(hipMemcpyAsync(buf->ptr, src, bytes, hipMemcpyHostToDevice, buf->ctx->stream));
    HIP_TRY(hipStreamSynchronize(buf->ctx->stream));
    return HT_OK;
}

extern "C" int ht_buf_read(ht_buf *buf, void *dst, size_t bytes) {
    if (!buf || (!dst && bytes)) return fail(HT_ERR_OTHER, "ht_buf_read: invalid argument");
    if (bytes > buf->size) return fail(HT_ERR_OTHER, "ht_buf_read: " + std::to_string(bytes) + " bytes from a buffer of " + std::to_string(buf->size));
    if (!bytes) return HT_OK;
    HIP_TRY(hipSetDevice(buf->ctx->device));
    HIP_TRY(hipMemcpyAsync(dst, buf->ptr, bytes, hipMemcpyDeviceToHost, buf->ctx->stream));
    HIP_TRY(hipStreamSynchronize(buf->ctx->stream));
    return HT_OK;
}

extern "C" int ht_buf_zero(ht_buf *buf) {
    if (!buf) return fail(HT_ERR_OTHER, "ht_buf_zero: null buffer");
    if (!buf->size) return HT_OK;
    HIP_TRY(hipSetDevice(buf->ctx->device));
    HIP_TRY(hipMemsetAsync(buf->ptr, 0, buf->size, buf->ctx->stream));
    return HT_OK;
}

namespace {

int check_image(const char *what, uint32_t w, uint32_t h, const ht_buf *b, size_t px_bytes) {
    if (!b) return fail(HT_ERR_OTHER, std::string(what) + ": null buffer");
    if ((uint64_t)w * (uint64_t)h > 0x80000000ull)
        return fail(HT_ERR_OTHER, std::string(what) + ": image larger than 2^31 pixels");
    size_t need = (size_t)w * (size_t)h * px_bytes;
    if (b->size < need)
        return fail(HT_ERR_OTHER, std::string(what) + ": buffer of " + std::to_string(b->size) + " bytes, need " + std::to_string(need));
    return HT_OK;
}

int launch_render(ht_prog *prog, hipFunction_t fn, uint32_t w, uint32_t h, ht_buf *scene, ht_buf *canvas,
                  ht_buf *seeds, uint32_t passes, bool with_passes) {
    if (!prog || !scene) return fail(HT_ERR_OTHER, "render: invalid argument");
    if (!fn) return fail(HT_ERR_OTHER, "render: the program has no per-scene render kernels (a batch program)");
    if (w == 0 || h == 0) return HT_OK;
    // at most 2^31 pixels (a 32 GiB canvas): pixel counters and the grid
    // size then stay far from 32-bit wrap-around
    if ((uint64_t)w * (uint64_t)h > 0x80000000ull) return fail(HT_ERR_OTHER, "render: image larger than 2^31 pixels");
    int rc;
    if ((rc = check_image("render canvas", w, h, canvas, 16))) return rc;
    if ((rc = check_image("render seeds", w, h, seeds, 4))) return rc;
    if (scene->size == 0) return fail(HT_ERR_OTHER, "render: empty scene buffer");
    ht_ctx *ctx = prog->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    uint32_t n = w * h;
    uint32_t block = 256;
    uint32_t grid = (n + block - 1) / block;
    void *sp = scene->ptr, *cp = canvas->ptr, *pp = seeds->ptr;
    std::lock_guard<std::mutex> lock(prog->launch_mu);
    if ((rc = ht_internal::fill_bounds(prog, sp))) return rc;
    if (with_passes && prog->pixel_classes && n <= (1u << 23)) {
        // the probe that marks the pixels the queue hands out first (the
        // kernel's HT_CLASS_CAP: larger frames are not probed and the
        // render kernel does not read the marks)
        void *args_probe[] = {&w, &h, &sp};
        HIP_TRY(hipModuleLaunchKernel(prog->pixel_classes, grid, 1, 1, block, 1, 1, 0, ctx->stream, args_probe,
                                      nullptr));
    }
    if (with_passes && prog->pixel_queue_reset) {
        // persistent: the blocks that fit at once pull pixels from the counter
        HIP_TRY(hipModuleLaunchKernel(prog->pixel_queue_reset, 1, 1, 1, 64, 1, 1, 0, ctx->stream, nullptr, nullptr));
        if (grid > prog->resident_blocks) grid = prog->resident_blocks;
    }
    void *args_pass[] = {&w, &h, &sp, &cp, &pp};
    void *args_passes[] = {&w, &h, &sp, &cp, &pp, &passes};
    HIP_TRY(hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, 0, ctx->stream,
                                  with_passes ? args_passes : args_pass, nullptr));
    return HT_OK;
}

}  // namespace

extern "C" int ht_render(ht_prog *prog, uint32_t w, uint32_t h, ht_buf *scene, ht_buf *canvas, ht_buf *seeds) {
    int rc = launch_render(prog, prog ? prog->render : nullptr, w, h, scene, canvas, seeds, 1, false);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(prog->ctx->stream));
    return HT_OK;
}

extern "C" int ht_render_passes(ht_prog *prog, uint32_t w, uint32_t h, ht_buf *scene, ht_buf *canvas, ht_buf *seeds,
                                uint32_t passes) {
    if (passes == 0) return HT_OK;
    return launch_render(prog, prog ? prog->render_passes : nullptr, w, h, scene, canvas, seeds, passes, true);
}

extern "C" int ht_extract(ht_ctx *ctx, uint32_t w, uint32_t h, uint32_t passes, ht_buf *canvas, ht_buf *image) {
    if (!ctx) return fail(HT_ERR_OTHER, "ht_extract: null ctx");
    int rc;
    if ((rc = check_image("ht_extract canvas", w, h, canvas, 16))) return rc;
    if ((rc = check_image("ht_extract image", w, h, image, 16))) return rc;
    uint32_t n = w * h;
    if (!n) return HT_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_extract, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, passes,
                       (const float4 *)canvas->ptr, (float4 *)image->ptr);
    HIP_TRY(hipGetLastError());
    return HT_OK;
}

extern "C" int ht_gamma(ht_ctx *ctx, uint32_t w, uint32_t h, float gamma, ht_buf *in, ht_buf *out) {
    if (!ctx) return fail(HT_ERR_OTHER, "ht_gamma: null ctx");
    int rc;
    if ((rc = check_image("ht_gamma in", w, h, in, 16))) return rc;
    if ((rc = check_image("ht_gamma out", w, h, out, 16))) return rc;
    uint32_t n = w * h;
    if (!n) return HT_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_gamma, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, gamma,
                       (const float4 *)in->ptr, (float4 *)out->ptr);
    HIP_TRY(hipGetLastError());
    return HT_OK;
}

extern "C" int ht_pack(ht_ctx *ctx, uint32_t w, uint32_t h, ht_buf *image, ht_buf *rgba8) {
    if (!ctx) return fail(HT_ERR_OTHER, "ht_pack: null ctx");
    int rc;
    if ((rc = check_image("ht_pack image", w, h, image, 16))) return rc;
    if ((rc = check_image("ht_pack rgba8", w, h, rgba8, 4))) return rc;
    uint32_t n = w * h;
    if (!n) return HT_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_pack, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, (const float4 *)image->ptr,
                       (uint32_t *)rgba8->ptr);
    HIP_TRY(hipGetLastError());
    return HT_OK;
}

extern "C" int ht_post(ht_ctx *ctx, uint32_t w, uint32_t h, uint32_t passes, float gamma, ht_buf *canvas,
                       ht_buf *rgba8) {
    if (!ctx) return fail(HT_ERR_OTHER, "ht_post: null ctx");
    int rc;
    if ((rc = check_image("ht_post canvas", w, h, canvas, 16))) return rc;
    if ((rc = check_image("ht_post rgba8", w, h, rgba8, 4))) return rc;
    uint32_t n = w * h;
    if (!n) return HT_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_post, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, passes, gamma,
                       (const float4 *)canvas->ptr, (uint32_t *)rgba8->ptr);
    HIP_TRY(hipGetLastError());
    return HT_OK;
}

extern "C" int ht_mo_compose(ht_ctx *ctx, ht_buf *scene, int real_size, uint32_t n, const uint32_t *offsets,
                             const uint8_t *letters, const uint8_t *lengths, const double *gens, uint32_t n_gens,
                             const double *posts) {
    if (!ctx || !scene) return fail(HT_ERR_OTHER, "ht_mo_compose: null ctx or scene");
    if (real_size != 4 && real_size != 8) return fail(HT_ERR_OTHER, "ht_mo_compose: real_size must be 4 or 8");
    if (!n) return HT_OK;
    if (!offsets || !letters || !lengths || !gens || !posts || !n_gens)
        return fail(HT_ERR_OTHER, "ht_mo_compose: null array");
    for (uint32_t i = 0; i < n; ++i) {
        if ((size_t)offsets[i] + 8 * (size_t)real_size > scene->size || offsets[i] % (8 * real_size))
            return fail(HT_ERR_OTHER, "ht_mo_compose: map offset " + std::to_string(offsets[i]) +
                                          " outside the scene buffer or misaligned");
        if (lengths[i] > HT_MO_WORD_MAX) return fail(HT_ERR_OTHER, "ht_mo_compose: word longer than HT_MO_WORD_MAX");
        for (uint32_t k = 0; k < lengths[i]; ++k)
            if (letters[(size_t)i * HT_MO_WORD_MAX + k] >= n_gens)
                return fail(HT_ERR_OTHER, "ht_mo_compose: letter outside the generator set");
    }
    // One staging allocation for the inputs: offsets | lengths | letters | gens | posts.
    size_t o_off = 0, o_len = o_off + 4 * (size_t)n, o_let = o_len + n;
    size_t o_gen = (o_let + (size_t)n * HT_MO_WORD_MAX + 7) & ~(size_t)7, o_post = o_gen + 64 * (size_t)n_gens;
    size_t total = o_post + 64 * (size_t)n;
    std::vector<uint8_t> host(total);
    std::memcpy(host.data() + o_off, offsets, 4 * (size_t)n);
    std::memcpy(host.data() + o_len, lengths, n);
    std::memcpy(host.data() + o_let, letters, (size_t)n * HT_MO_WORD_MAX);
    std::memcpy(host.data() + o_gen, gens, 64 * (size_t)n_gens);
    std::memcpy(host.data() + o_post, posts, 64 * (size_t)n);
    HIP_TRY(hipSetDevice(ctx->device));
    uint8_t *dev = nullptr;
    HIP_TRY(hipMalloc(&dev, total));
    hipError_t e = hipMemcpyAsync(dev, host.data(), total, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        (void)hipGetLastError();
        hipLaunchKernelGGL(k_mo_compose, dim3((n + 63) / 64), dim3(64), 0, ctx->stream, (uint8_t *)scene->ptr, n,
                           real_size, (const uint32_t *)(dev + o_off), dev + o_let, dev + o_len,
                           (const double *)(dev + o_gen), (const double *)(dev + o_post));
        e = hipGetLastError();
    }
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    hipError_t e3 = hipFree(dev);
    if (e != hipSuccess) return fail(HT_ERR_BACKEND, std::string("ht_mo_compose: ") + hipGetErrorString(e));
    if (e2 != hipSuccess) return fail(HT_ERR_BACKEND, std::string("ht_mo_compose: ") + hipGetErrorString(e2));
    if (e3 != hipSuccess) return fail(HT_ERR_BACKEND, std::string("ht_mo_compose: ") + hipGetErrorString(e3));
    return HT_OK;
}

extern "C" int ht_seed_init(ht_buf *seeds, uint32_t w, uint32_t h, int mode, uint64_t seed) {
    int rc;
    if ((rc = check_image("ht_seed_init", w, h, seeds, 4))) return rc;
    std::vector<uint32_t> host((size_t)w * h);
    if ((rc = ht_seed_fill(host.data(), host.size(), mode, seed))) return rc;
    return ht_buf_write(seeds, host.data(), host.size() * 4);
}

extern "C" int ht_timer_begin(ht_ctx *ctx) {
    if (!ctx) return fail(HT_ERR_OTHER, "ht_timer_begin: null ctx");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipEventRecord(ctx->t0, ctx->stream));
    return HT_OK;
}

extern "C" int ht_timer_end(ht_ctx *ctx, float *ms) {
    if (!ctx || !ms) return fail(HT_ERR_OTHER, "ht_timer_end: invalid argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipEventRecord(ctx->t1, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->t1));
    HIP_TRY(hipEventElapsedTime(ms, ctx->t0, ctx->t1));
    return HT_OK;
}
