// Learned radiance cache on one MI355X (gfx950): a 12-feature, 6-band
// frequency-encoded MLP (144 -> 5 x 64 ReLU -> 3), BF16 weights and
// activations on MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation,
// trained online with Adam on the tracer's own path suffixes. No reference
// counterpart: the reference has no learned part. The kernels:
//
// k_mlp_infer   encode + forward for the query queue; the epilogue adds
//               throughput x max(prediction, 0) to each query's pixel.
// k_mlp_train   encode + forward + relative-L2 loss gradient + backward for
//               a batch, storing each layer's input activations and output
//               gradients (BF16, feature-major) for the weight gradients.
// k_mlp_wgrad   dW = delta . act^T per layer as split-K MFMA GEMMs over
//               chunks of the batch (plus bias sums), one partial per chunk.
// k_mlp_adam    sums the partials in a fixed order, Adam, and rewrites the
//               BF16 weight images.
//
// Layout. The network runs transposed, Y^T = W X^T: weights are the MFMA's
// A operand (16 rows of W x 32 k) and 16 samples are the columns (B operand,
// 32 k x 16 samples). The accumulator tile of a layer gives lane (g, c) =
// (lane / 16, lane % 16) the outputs n = 16 t + 4 g + j (j < 4) of sample c,
// and the next layer's B operand wants, per lane, 8 consecutive k of the same
// sample c: taking the next layer's k in the order
//     kappa = 32 s + 8 g + j'  <->  feature phi(kappa) = 16 (2 s + j'/4) + 4 g + j'%4
// makes that operand exactly the values the lane already holds, so the
// activations of all six layers stay in registers. The weight images in LDS
// store W[n][phi(kappa)] (forward) and W[phi(kappa)][m] (backward, W^T) with
// rows padded by 8 BF16 to spread LDS banks.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hh"

using ht_internal::ctx_release;
using ht_internal::ctx_retain;
using ht_internal::fail;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NF = 12;       // features
constexpr int NBANDS = 6;    // frequency bands: sin, cos of 2^b pi x
constexpr int IN = 144;      // encoded inputs
constexpr int IN_PAD = 160;  // 5 k-steps of 32
constexpr int HID = 64;
constexpr int OUT = 3;
static_assert(IN == NF * 2 * NBANDS, "encoding width");

// FP32 master parameters: W0 [64][144], b0 [64], W1..W4 [64][64], b1..b4, W5 [3][64], b5 [3].
constexpr int P_W0 = 0, P_B0 = P_W0 + HID * IN;
constexpr int P_H = P_B0 + HID;  // first hidden layer's W
constexpr int P_HSTRIDE = HID * HID + HID;
constexpr int P_W5 = P_H + 4 * P_HSTRIDE, P_B5 = P_W5 + OUT * HID;
constexpr int NPARAM = P_B5 + OUT;
static_assert(NPARAM == HT_MLP_PARAMS, "parameter count");
__host__ __device__ constexpr int p_w(int l) { return l == 0 ? P_W0 : l == 5 ? P_W5 : P_H + (l - 1) * P_HSTRIDE; }
__host__ __device__ constexpr int p_b(int l) { return l == 0 ? P_B0 : l == 5 ? P_B5 : P_H + (l - 1) * P_HSTRIDE + HID * HID; }

// BF16 weight images (element offsets, padded row strides).
constexpr int S0 = IN_PAD + 8, SH = HID + 8, S5B = 32 + 8;
constexpr int I_F0 = 0;
constexpr int I_FH = I_F0 + HID * S0;  // F1..F4
constexpr int I_F5 = I_FH + 4 * HID * SH;
constexpr int I_FWD_END = I_F5 + 16 * SH;
constexpr int I_BH = I_FWD_END;  // B1..B4 (W^T)
constexpr int I_B5 = I_BH + 4 * HID * SH;
constexpr int I_END = I_B5 + HID * S5B;
__host__ __device__ constexpr int i_f(int l) { return l == 0 ? I_F0 : l == 5 ? I_F5 : I_FH + (l - 1) * HID * SH; }
__host__ __device__ constexpr int i_b(int l) { return l == 5 ? I_B5 : I_BH + (l - 1) * HID * SH; }

// Activation / gradient buffers for training (feature-major, n columns):
// act: A0 [160][n] (encoded input), A1..A5 [64][n] (layer inputs); del: D0..D4 [64][n], D5 [16][n].
__host__ __device__ constexpr size_t act_row(int l) { return l == 0 ? 0 : (size_t)IN_PAD + (size_t)(l - 1) * HID; }
constexpr int ACT_ROWS = IN_PAD + 5 * HID;
__host__ __device__ constexpr size_t del_row(int l) { return (size_t)l * HID; }
constexpr int DEL_ROWS = 5 * HID + 16;

constexpr float ADAM_B1 = 0.9f, ADAM_B2 = 0.99f, ADAM_EPS = 1e-8f;
constexpr float LOSS_EPS = 0.01f;  // relative L2: (p - y)^2 / (p^2 + eps)
constexpr int WGRAD_CHUNK = 128;   // samples per wgrad partial (16 384 samples: 128 blocks, 512 waves)

// kappa <-> feature (see the header comment); for K = 64 and K = 16 (t = 0 only).
__host__ __device__ inline int kappa_of(int f) {
    int t = f >> 4, g = (f >> 2) & 3, jj = f & 3;
    return 32 * (t >> 1) + 8 * g + 4 * (t & 1) + jj;
}

__device__ inline void wave_sync() {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Encoded input kappa of a sample's features x (LDS row): sin / cos of
// 2^b pi x_f with kappa = 12 f + 2 b + (0 sin, 1 cos); 0 for kappa >= 144.
// v_sin/v_cos take revolutions: sin(2^b pi x) = sin(2 pi * 2^(b-1) x).
__device__ inline float encode(const float *x, int kappa) {
    if (kappa >= IN) return 0.0f;
    int f = kappa / 12, r = kappa - 12 * f, b = r >> 1;
    float rev = x[f] * (float)(1 << b) * 0.5f;
    return (r & 1) ? __builtin_amdgcn_cosf(rev) : __builtin_amdgcn_sinf(rev);
}

// Copies `n16` 16-byte units from global to LDS with the whole block.
__device__ inline void load_image(__bf16 *lds, const __bf16 *img, int n16) {
    const uint4 *src = reinterpret_cast<const uint4 *>(img);
    uint4 *dst = reinterpret_cast<uint4 *>(lds);
    for (int i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
}

__device__ inline bf16x8 lds8(const __bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

// B operand of the next layer from this layer's accumulators (ReLU, BF16).
__device__ inline void relu_pack(const f32x4 acc[4], bf16x8 out[2]) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int jp = 0; jp < 8; ++jp) out[s][jp] = (__bf16)fmaxf(acc[2 * s + (jp >> 2)][jp & 3], 0.0f);
}

// One 64-out layer on the register-resident B operand (K = 64).
__device__ inline void layer64(const __bf16 *w, const float *bias, const bf16x8 bin[2], f32x4 acc[4], int g, int c) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[t][j] = bias[16 * t + 4 * g + j];
#pragma unroll
        for (int s = 0; s < 2; ++s) acc[t] = mfma(lds8(w + (16 * t + c) * SH + 32 * s + 8 * g), bin[s], acc[t]);
    }
}

// Forward through all layers for one 16-sample tile whose features are in
// LDS (x: [16][13]); returns the output tile (lane g = 0: RGB of sample c in
// [0..2]) and, if `keep`, each hidden layer's packed activations.
__device__ inline f32x4 forward_tile(const __bf16 *w, const float *__restrict__ params, const float *x, int g, int c,
                                     bf16x8 enc_out[5], bf16x8 keep[5][2]) {
    bf16x8 be[5];
#pragma unroll
    for (int s = 0; s < 5; ++s)
#pragma unroll
        for (int jp = 0; jp < 8; ++jp) be[s][jp] = (__bf16)encode(x + 13 * c, 32 * s + 8 * g + jp);
    if (enc_out) {
#pragma unroll
        for (int s = 0; s < 5; ++s) enc_out[s] = be[s];
    }
    f32x4 acc[4];
    const float *b0 = params + P_B0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[t][j] = b0[16 * t + 4 * g + j];
#pragma unroll
        for (int s = 0; s < 5; ++s) acc[t] = mfma(lds8(w + I_F0 + (16 * t + c) * S0 + 32 * s + 8 * g), be[s], acc[t]);
    }
    bf16x8 h[2];
#pragma unroll
    for (int l = 1; l <= 4; ++l) {
        relu_pack(acc, h);
        if (keep) {
            keep[l - 1][0] = h[0];
            keep[l - 1][1] = h[1];
        }
        layer64(w + i_f(l), params + p_b(l), h, acc, g, c);
    }
    relu_pack(acc, h);
    if (keep) {
        keep[4][0] = h[0];
        keep[4][1] = h[1];
    }
    f32x4 out;
    const float *b5 = params + P_B5;
#pragma unroll
    for (int j = 0; j < 4; ++j) out[j] = (4 * g + j < OUT) ? b5[4 * g + j] : 0.0f;
#pragma unroll
    for (int s = 0; s < 2; ++s) out = mfma(lds8(w + I_F5 + c * SH + 32 * s + 8 * g), h[s], out);
    return out;
}

// ------------------------------------------------------------ inference
// Persistent over the query queue: each wave takes 16-sample tiles.
__global__ void __launch_bounds__(256) k_mlp_infer(const __bf16 *__restrict__ image, const float *__restrict__ params,
                                                   const uint32_t *__restrict__ count_ptr, uint32_t count_const,
                                                   uint32_t cap, const float *__restrict__ feat,
                                                   const uint32_t *__restrict__ pixel, const float *__restrict__ thr,
                                                   float4 *__restrict__ canvas, float *__restrict__ pred) {
    __shared__ __attribute__((aligned(16))) __bf16 w[I_FWD_END];
    __shared__ float xs[4][16 * 13];
    load_image(w, image, I_FWD_END / 8);
    __syncthreads();
    uint32_t n = count_ptr ? min(*count_ptr, cap) : count_const;
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    uint32_t tiles = (n + 15) / 16;
    float *x = xs[wave];
    for (uint32_t tile = blockIdx.x * 4 + wave; tile < tiles; tile += gridDim.x * 4) {
        uint32_t base = tile * 16;
        for (int e = lane; e < 16 * NF; e += 64) {
            int si = e & 15, f = e >> 4;
            uint32_t s = base + si;
            x[13 * si + f] = s < n ? feat[(size_t)f * cap + s] : 0.0f;
        }
        wave_sync();
        f32x4 out = forward_tile(w, params, x, g, c, nullptr, nullptr);
        uint32_t s = base + c;
        if (g == 0 && s < n) {
            if (pred) {
                pred[s] = out[0];
                pred[(size_t)cap + s] = out[1];
                pred[2 * (size_t)cap + s] = out[2];
            }
            if (canvas) {
                uint32_t p = pixel[s];
                float4 a = canvas[p];
                a.x += thr[s] * fmaxf(out[0], 0.0f);
                a.y += thr[(size_t)cap + s] * fmaxf(out[1], 0.0f);
                a.z += thr[2 * (size_t)cap + s] * fmaxf(out[2], 0.0f);
                canvas[p] = a;
            }
        }
        wave_sync();
    }
}

// ------------------------------------------------------------ training: forward + backward
// Batch of n samples (n a multiple of 64): sample i is record (offset + i) mod
// count of the training queue (feat [12][ld], target [3][ld]). Each wave does
// one 16-sample tile.
__global__ void __launch_bounds__(256) k_mlp_train(const __bf16 *__restrict__ image, const float *__restrict__ params,
                                                   const uint32_t *__restrict__ count_ptr, uint32_t count_const,
                                                   uint32_t ld, uint32_t offset, uint32_t n,
                                                   const float *__restrict__ feat, const float *__restrict__ target,
                                                   __bf16 *__restrict__ act, __bf16 *__restrict__ del,
                                                   float *__restrict__ loss_sum) {
    __shared__ __attribute__((aligned(16))) __bf16 w[I_END];
    __shared__ float xs[4][16 * 13];
    __shared__ float ys[4][16 * 3];
    load_image(w, image, I_END / 8);
    __syncthreads();
    uint32_t count = count_ptr ? min(*count_ptr, ld) : count_const;
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    uint32_t tile = blockIdx.x * 4 + wave;
    if (tile * 16 >= n) return;
    uint32_t s0 = tile * 16;
    float *x = xs[wave], *y = ys[wave];
    for (int e = lane; e < 16 * NF; e += 64) {
        int si = e & 15, f = e >> 4;
        uint32_t r = count ? (offset + s0 + si) % count : 0;
        x[13 * si + f] = count ? feat[(size_t)f * ld + r] : 0.0f;
    }
    if (lane < 48) {
        int si = lane & 15, o = lane >> 4;
        uint32_t r = count ? (offset + s0 + si) % count : 0;
        y[3 * si + o] = count ? target[(size_t)o * ld + r] : 0.0f;
    }
    wave_sync();
    bf16x8 enc[5], h[5][2];
    f32x4 out = forward_tile(w, params, x, g, c, enc, h);
    uint32_t col = s0 + c;
    // layer inputs: A0 = encoding (kappa order), A_l = h[l-1] in feature order
#pragma unroll
    for (int s = 0; s < 5; ++s)
#pragma unroll
        for (int jp = 0; jp < 8; ++jp) act[(size_t)(32 * s + 8 * g + jp) * n + col] = enc[s][jp];
#pragma unroll
    for (int l = 1; l <= 5; ++l)
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int jp = 0; jp < 8; ++jp) {
                int f = 16 * (2 * s + (jp >> 2)) + 4 * g + (jp & 3);
                act[(act_row(l) + f) * n + col] = h[l - 1][s][jp];
            }
    // d loss / d out: relative L2 over n samples x 3 channels
    f32x4 d5;
    float lsum = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int o = 4 * g + j;
        float d = 0.0f;
        if (o < OUT && count) {
            float p = out[j], t = y[3 * c + o];
            float den = p * p + LOSS_EPS;
            d = 2.0f * (p - t) / (den * (float)n * (float)OUT);
            lsum += (p - t) * (p - t) / den;
        }
        d5[j] = d;
    }
    for (int m = 32; m >= 1; m >>= 1) lsum += __shfl_xor(lsum, m);
    if (lane == 0 && loss_sum) atomicAdd(loss_sum, lsum);
    bf16x8 bd[2];
#pragma unroll
    for (int jp = 0; jp < 8; ++jp) bd[0][jp] = (__bf16)(jp < 4 ? d5[jp] : 0.0f);
#pragma unroll
    for (int j = 0; j < 4; ++j) del[(del_row(5) + 4 * g + j) * n + col] = bd[0][j];
    // backward through layers 5..1: delta_{l-1} = (W_l^T delta_l) * relu'(h_{l-1})
    f32x4 acc[4];
#pragma unroll
    for (int l = 5; l >= 1; --l) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            if (l == 5) {
                acc[t] = mfma(lds8(w + I_B5 + (16 * t + c) * S5B + 8 * g), bd[0], acc[t]);
            } else {
#pragma unroll
                for (int s = 0; s < 2; ++s)
                    acc[t] = mfma(lds8(w + i_b(l) + (16 * t + c) * SH + 32 * s + 8 * g), bd[s], acc[t]);
            }
        }
        // mask by the ReLU of layer l-1's output (held packed in h[l-1])
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int jp = 0; jp < 8; ++jp) {
                int t = 2 * s + (jp >> 2), j = jp & 3;
                float v = (float)h[l - 1][s][jp] > 0.0f ? acc[t][j] : 0.0f;
                bd[s][jp] = (__bf16)v;
                del[(del_row(l - 1) + 16 * t + 4 * g + j) * n + col] = bd[s][jp];
            }
    }
}

// ------------------------------------------------------------ training: weight gradients
// Tiles (layer, 16-row block of outputs, 16-column block of inputs).
struct WTile {
    int l, tn, tk;
};
constexpr int NWTILES = 40 + 4 * 16 + 4;
__device__ inline WTile wtile(int i) {
    if (i < 40) return {0, i / 10, i % 10};
    i -= 40;
    if (i < 64) return {1 + i / 16, (i % 16) / 4, i % 4};
    i -= 64;
    return {5, 0, i};
}

// Block x = chunk of WGRAD_CHUNK samples; its WGRAD_WAVES waves loop over
// the tiles. partial[chunk][NPARAM]: every weight and bias written once per
// chunk.
constexpr int WGRAD_WAVES = 8;
__global__ void __launch_bounds__(64 * WGRAD_WAVES) k_mlp_wgrad(const __bf16 *__restrict__ act,
                                                                const __bf16 *__restrict__ del, uint32_t n,
                                                                float *__restrict__ partial) {
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    uint32_t chunk = blockIdx.x;
    uint32_t s_begin = chunk * WGRAD_CHUNK, s_end = min(s_begin + WGRAD_CHUNK, n);
    float *out = partial + (size_t)chunk * NPARAM;
    for (int ti = wave; ti < NWTILES; ti += WGRAD_WAVES) {
        WTile wt = wtile(ti);
        int l = wt.l;
        const __bf16 *dp = del + (del_row(l) + 16 * wt.tn + c) * (size_t)n;
        const __bf16 *ap = act + (act_row(l) + 16 * wt.tk + c) * (size_t)n;
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        float bsum = 0.0f;
        for (uint32_t s = s_begin; s < s_end; s += 32) {
            bf16x8 a = *reinterpret_cast<const bf16x8 *>(dp + s + 8 * g);
            bf16x8 b = *reinterpret_cast<const bf16x8 *>(ap + s + 8 * g);
            acc = mfma(a, b, acc);
            if (wt.tk == 0) {
#pragma unroll
                for (int j = 0; j < 8; ++j) bsum += (float)a[j];
            }
        }
        int nout = l == 5 ? OUT : HID, kin = l == 0 ? IN : HID;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int r = 16 * wt.tn + 4 * g + j, k = 16 * wt.tk + c;
            if (r < nout && k < kin) out[p_w(l) + r * kin + k] = acc[j];
        }
        if (wt.tk == 0) {
            bsum += __shfl_xor(bsum, 16);
            bsum += __shfl_xor(bsum, 32);
            int r = 16 * wt.tn + c;
            if (g == 0 && r < nout) out[p_b(l) + r] = bsum;
        }
    }
}

// ------------------------------------------------------------ parameters
// Writes master parameter i (a weight) into the BF16 images.
__device__ inline void pack_param(__bf16 *image, int i, float v) {
    __bf16 b = (__bf16)v;
    if (i < P_B0) {
        int r = i / IN, k = i % IN;
        image[I_F0 + r * S0 + k] = b;
        return;
    }
    for (int l = 1; l <= 4; ++l) {
        int o = i - p_w(l);
        if (o >= 0 && o < HID * HID) {
            int r = o / HID, k = o % HID;
            image[i_f(l) + r * SH + kappa_of(k)] = b;
            image[i_b(l) + k * SH + kappa_of(r)] = b;
            return;
        }
    }
    int o = i - P_W5;
    if (o >= 0 && o < OUT * HID) {
        int r = o / HID, k = o % HID;
        image[I_F5 + r * SH + kappa_of(k)] = b;
        image[I_B5 + k * S5B + kappa_of(r)] = b;
    }
}

__global__ void k_mlp_pack(const float *__restrict__ params, __bf16 *__restrict__ image) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < NPARAM) pack_param(image, i, params[i]);
}

// Sums the chunk partials, then (if apply) one Adam step; writes the
// gradient, the parameter and its BF16 image entries. A block holds
// ADAM_PARAMS parameters x ADAM_SUBS sub-sums: sub-sum s adds chunks
// s, s + ADAM_SUBS, ... in order, and the sub-sums are added in order of s,
// so the result is deterministic. Lanes of one sub read consecutive
// parameters (coalesced rows of the partial matrix).
constexpr int ADAM_SUBS = 8, ADAM_PARAMS = 256 / ADAM_SUBS;
__global__ void __launch_bounds__(256) k_mlp_adam(float *__restrict__ params, float *__restrict__ m1,
                                                  float *__restrict__ m2, const float *__restrict__ partial,
                                                  uint32_t nchunks, float lr, float c1, float c2, int apply,
                                                  float *__restrict__ grad, __bf16 *__restrict__ image) {
    __shared__ float sub_sum[ADAM_SUBS][ADAM_PARAMS];
    int col = threadIdx.x % ADAM_PARAMS, sub = threadIdx.x / ADAM_PARAMS;
    int i = blockIdx.x * ADAM_PARAMS + col;
    float part = 0.0f;
    if (i < NPARAM)
        for (uint32_t k = sub; k < nchunks; k += ADAM_SUBS) part += partial[(size_t)k * NPARAM + i];
    sub_sum[sub][col] = part;
    __syncthreads();
    if (sub != 0 || i >= NPARAM) return;
    float gsum = sub_sum[0][col];
#pragma unroll
    for (int s = 1; s < ADAM_SUBS; ++s) gsum += sub_sum[s][col];
    grad[i] = gsum;
    if (!apply) return;
    float a = ADAM_B1 * m1[i] + (1.0f - ADAM_B1) * gsum;
    float b = ADAM_B2 * m2[i] + (1.0f - ADAM_B2) * gsum * gsum;
    m1[i] = a;
    m2[i] = b;
    float p = params[i] - lr * (a * c1) / (sqrtf(b * c2) + ADAM_EPS);
    params[i] = p;
    pack_param(image, i, p);
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f32(float *__restrict__ p, uint32_t n, float v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

// ------------------------------------------------------------ the cache object

struct ht_cache {
    ht_ctx *ctx = nullptr;
    uint32_t qcap = 0, tcap = 0, max_batch = 0;
    uint32_t step = 0;
    // one allocation, carved
    void *mem = nullptr;
    float *params = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr, *partial = nullptr, *loss = nullptr;
    __bf16 *image = nullptr, *act = nullptr, *del = nullptr;
    uint32_t *counters = nullptr, *q_pixel = nullptr;
    float *q_feat = nullptr, *q_thr = nullptr, *t_feat = nullptr, *t_target = nullptr;
};

namespace {

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int launch_fill_f32(hipStream_t st, float *p, uint32_t n, float v) {
    if (!n) return HT_OK;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_fill_f32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
    HIP_TRY(hipGetLastError());
    return HT_OK;
}

int check_buf(const char *what, const ht_buf *b, size_t bytes) {
    if (!b) return fail(HT_ERR_OTHER, std::string(what) + ": null buffer");
    if (b->size < bytes)
        return fail(HT_ERR_OTHER, std::string(what) + ": buffer of " + std::to_string(b->size) + " bytes, needs " +
                                      std::to_string(bytes));
    return HT_OK;
}

// Forward + backward + weight gradients + reduction (+ Adam) for one batch.
int train_batch(ht_cache *c, const uint32_t *count_ptr, uint32_t count_const, uint32_t ld, uint32_t offset,
                uint32_t n, const float *feat, const float *target, float lr, int apply) {
    hipStream_t st = c->ctx->stream;
    int rc;
    if ((rc = launch_fill_f32(st, c->loss, 1, 0.0f))) return rc;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_mlp_train, dim3(n / 64), dim3(256), 0, st, c->image, c->params, count_ptr, count_const, ld,
                       offset, n, feat, target, c->act, c->del, c->loss);
    HIP_TRY(hipGetLastError());
    uint32_t nchunks = (n + WGRAD_CHUNK - 1) / WGRAD_CHUNK;
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(nchunks), dim3(64 * WGRAD_WAVES), 0, st, c->act, c->del, n, c->partial);
    HIP_TRY(hipGetLastError());
    float c1 = 1.0f, c2 = 1.0f;
    if (apply) {
        c->step += 1;
        c1 = (float)(1.0 / (1.0 - std::pow((double)ADAM_B1, (double)c->step)));
        c2 = (float)(1.0 / (1.0 - std::pow((double)ADAM_B2, (double)c->step)));
    }
    hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + ADAM_PARAMS - 1) / ADAM_PARAMS), dim3(256), 0, st, c->params,
                       c->m1, c->m2, c->partial, nchunks, lr, c1, c2, apply, c->grad, c->image);
    HIP_TRY(hipGetLastError());
    return HT_OK;
}

}  // namespace

extern "C" int ht_cache_create(ht_ctx *ctx, uint32_t query_capacity, uint32_t train_capacity, uint32_t max_batch,
                               ht_cache **out) {
    if (!ctx || !out) return fail(HT_ERR_OTHER, "ht_cache_create: invalid argument");
    if (max_batch == 0 || max_batch % 64)
        return fail(HT_ERR_OTHER, "ht_cache_create: max_batch must be a positive multiple of 64");
    ht_cache *c = new ht_cache();
    c->ctx = ctx;
    c->qcap = query_capacity;
    c->tcap = train_capacity;
    c->max_batch = max_batch;
    uint32_t nchunks = (max_batch + WGRAD_CHUNK - 1) / WGRAD_CHUNK;
    size_t sizes[] = {
        4 * (size_t)NPARAM, 4 * (size_t)NPARAM, 4 * (size_t)NPARAM, 4 * (size_t)NPARAM,  // params m1 m2 grad
        4 * (size_t)NPARAM * nchunks, 4,                                                 // partial loss
        2 * (size_t)I_END, 2 * (size_t)ACT_ROWS * max_batch, 2 * (size_t)DEL_ROWS * max_batch,
        4 * 4, 4 * (size_t)query_capacity,                                               // counters pixel
        4 * (size_t)NF * query_capacity, 4 * 3 * (size_t)query_capacity,                 // q_feat q_thr
        4 * (size_t)NF * train_capacity, 4 * 3 * (size_t)train_capacity};                // t_feat t_target
    size_t off[16], total = 0;
    for (size_t i = 0; i < sizeof(sizes) / sizeof(sizes[0]); ++i) {
        off[i] = total;
        total += align256(sizes[i]);
    }
    HIP_TRY(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(&c->mem, total);
    if (e != hipSuccess) {
        delete c;
        return fail(HT_ERR_BACKEND, std::string("ht_cache_create: hipMalloc: ") + hipGetErrorString(e));
    }
    uint8_t *b = static_cast<uint8_t *>(c->mem);
    c->params = (float *)(b + off[0]);
    c->m1 = (float *)(b + off[1]);
    c->m2 = (float *)(b + off[2]);
    c->grad = (float *)(b + off[3]);
    c->partial = (float *)(b + off[4]);
    c->loss = (float *)(b + off[5]);
    c->image = (__bf16 *)(b + off[6]);
    c->act = (__bf16 *)(b + off[7]);
    c->del = (__bf16 *)(b + off[8]);
    c->counters = (uint32_t *)(b + off[9]);
    c->q_pixel = (uint32_t *)(b + off[10]);
    c->q_feat = (float *)(b + off[11]);
    c->q_thr = (float *)(b + off[12]);
    c->t_feat = (float *)(b + off[13]);
    c->t_target = (float *)(b + off[14]);
    // everything starts at zero (image padding must stay zero)
    e = hipMemsetAsync(c->mem, 0, total, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(c->mem);
        delete c;
        return fail(HT_ERR_BACKEND, std::string("ht_cache_create: ") + hipGetErrorString(e));
    }
    ctx_retain(ctx);
    *out = c;
    return HT_OK;
}

extern "C" void ht_cache_destroy(ht_cache *c) {
    if (!c) return;
    (void)hipSetDevice(c->ctx->device);
    (void)hipStreamSynchronize(c->ctx->stream);
    (void)hipFree(c->mem);
    (void)hipGetLastError();
    ht_ctx *ctx = c->ctx;
    delete c;
    ctx_release(ctx);
}

extern "C" int ht_cache_set_params(ht_cache *c, const float *params, size_t n) {
    if (!c || !params) return fail(HT_ERR_OTHER, "ht_cache_set_params: invalid argument");
    if (n != (size_t)NPARAM) return fail(HT_ERR_OTHER, "ht_cache_set_params: expected HT_MLP_PARAMS values");
    HIP_TRY(hipSetDevice(c->ctx->device));
    hipStream_t st = c->ctx->stream;
    HIP_TRY(hipMemcpyAsync(c->params, params, 4 * (size_t)NPARAM, hipMemcpyHostToDevice, st));
    int rc;
    if ((rc = launch_fill_f32(st, c->m1, NPARAM, 0.0f))) return rc;
    if ((rc = launch_fill_f32(st, c->m2, NPARAM, 0.0f))) return rc;
    hipLaunchKernelGGL(k_mlp_pack, dim3((NPARAM + 255) / 256), dim3(256), 0, st, c->params, c->image);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    c->step = 0;
    return HT_OK;
}

extern "C" int ht_cache_get_params(ht_cache *c, float *params, size_t n) {
    if (!c || !params || n != (size_t)NPARAM) return fail(HT_ERR_OTHER, "ht_cache_get_params: invalid argument");
    HIP_TRY(hipSetDevice(c->ctx->device));
    HIP_TRY(hipMemcpyAsync(params, c->params, 4 * (size_t)NPARAM, hipMemcpyDeviceToHost, c->ctx->stream));
    HIP_TRY(hipStreamSynchronize(c->ctx->stream));
    return HT_OK;
}

extern "C" int ht_cache_get_grads(ht_cache *c, float *grads, size_t n) {
    if (!c || !grads || n != (size_t)NPARAM) return fail(HT_ERR_OTHER, "ht_cache_get_grads: invalid argument");
    HIP_TRY(hipSetDevice(c->ctx->device));
    HIP_TRY(hipMemcpyAsync(grads, c->grad, 4 * (size_t)NPARAM, hipMemcpyDeviceToHost, c->ctx->stream));
    HIP_TRY(hipStreamSynchronize(c->ctx->stream));
    return HT_OK;
}

extern "C" int ht_cache_counts(ht_cache *c, uint32_t *queries, uint32_t *records) {
    if (!c) return fail(HT_ERR_OTHER, "ht_cache_counts: null cache");
    uint32_t h[2];
    HIP_TRY(hipSetDevice(c->ctx->device));
    HIP_TRY(hipMemcpyAsync(h, c->counters, 8, hipMemcpyDeviceToHost, c->ctx->stream));
    HIP_TRY(hipStreamSynchronize(c->ctx->stream));
    if (queries) *queries = h[0];
    if (records) *records = h[1];
    return HT_OK;
}

extern "C" int ht_mlp_forward(ht_cache *c, ht_buf *features, uint32_t n, ht_buf *pred) {
    if (!c) return fail(HT_ERR_OTHER, "ht_mlp_forward: null cache");
    int rc;
    if ((rc = check_buf("ht_mlp_forward features", features, 4 * (size_t)NF * n))) return rc;
    if ((rc = check_buf("ht_mlp_forward pred", pred, 4 * 3 * (size_t)n))) return rc;
    if (!n) return HT_OK;
    HIP_TRY(hipSetDevice(c->ctx->device));
    uint32_t blocks = min((n + 63) / 64, 2048u);
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, c->ctx->stream, c->image, c->params, nullptr, n, n,
                       (const float *)features->ptr, nullptr, nullptr, nullptr, (float *)pred->ptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->ctx->stream));
    return HT_OK;
}

extern "C" int ht_mlp_train_step(ht_cache *c, ht_buf *features, ht_buf *targets, uint32_t n, float lr, int apply,
                                 float *loss) {
    if (!c) return fail(HT_ERR_OTHER, "ht_mlp_train_step: null cache");
    if (n == 0 || n % 64 || n > c->max_batch)
        return fail(HT_ERR_OTHER, "ht_mlp_train_step: n must be a positive multiple of 64 up to max_batch");
    int rc;
    if ((rc = check_buf("ht_mlp_train_step features", features, 4 * (size_t)NF * n))) return rc;
    if ((rc = check_buf("ht_mlp_train_step targets", targets, 4 * 3 * (size_t)n))) return rc;
    HIP_TRY(hipSetDevice(c->ctx->device));
    if ((rc = train_batch(c, nullptr, n, n, 0, n, (const float *)features->ptr, (const float *)targets->ptr, lr,
                          apply)))
        return rc;
    if (loss) {
        float l = 0.0f;
        HIP_TRY(hipMemcpyAsync(&l, c->loss, 4, hipMemcpyDeviceToHost, c->ctx->stream));
        HIP_TRY(hipStreamSynchronize(c->ctx->stream));
        *loss = l / (float)(n * OUT);
    } else {
        HIP_TRY(hipStreamSynchronize(c->ctx->stream));
    }
    return HT_OK;
}

extern "C" int ht_render_cached(ht_prog *prog, ht_cache *c, uint32_t w, uint32_t h, ht_buf *scene, ht_buf *canvas,
                                ht_buf *seeds, uint32_t pass_index, uint32_t train_every, uint32_t query_hop) {
    if (!prog || !c) return fail(HT_ERR_OTHER, "ht_render_cached: null program or cache");
    if (!prog->render_cached)
        return fail(HT_ERR_OTHER, "ht_render_cached: the program's scene type has no radiance-cache kernel "
                                  "(it needs a vector of mapped covered objects)");
    if (prog->ctx != c->ctx) return fail(HT_ERR_OTHER, "ht_render_cached: program and cache on different contexts");
    int rc;
    size_t npx = (size_t)w * h;
    if ((rc = check_buf("ht_render_cached scene", scene, 1))) return rc;
    if ((rc = check_buf("ht_render_cached canvas", canvas, 16 * npx))) return rc;
    if ((rc = check_buf("ht_render_cached seeds", seeds, 4 * npx))) return rc;
    if (npx > c->qcap) return fail(HT_ERR_OTHER, "ht_render_cached: image larger than the query capacity");
    if (!npx) return HT_OK;
    ht_ctx *ctx = c->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, c->counters, 4u, 0u);
    HIP_TRY(hipGetLastError());
    void *sp = scene->ptr, *cp = canvas->ptr, *dp = seeds->ptr;
    uint32_t qcap = c->qcap, tcap = c->tcap;
    void *args[] = {&w, &h, &sp, &cp, &dp, &pass_index, &c->counters, &c->q_pixel, &c->q_feat, &c->q_thr, &qcap,
                    &c->t_feat, &c->t_target, &tcap, &train_every, &query_hop};
    uint32_t grid = (uint32_t)((npx + 255) / 256);
    {
        std::lock_guard<std::mutex> lock(prog->launch_mu);
        if ((rc = ht_internal::fill_bounds(prog, sp))) return rc;
        HIP_TRY(hipModuleLaunchKernel(prog->render_cached, grid, 1, 1, 256, 1, 1, 0, ctx->stream, args, nullptr));
    }
    uint32_t blocks = min((uint32_t)((npx + 63) / 64), 2048u);
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, ctx->stream, c->image, c->params, c->counters, 0u,
                       c->qcap, c->q_feat, c->q_pixel, c->q_thr, (float4 *)canvas->ptr, nullptr);
    HIP_TRY(hipGetLastError());
    return HT_OK;
}

extern "C" int ht_cache_train(ht_cache *c, uint32_t steps, uint32_t batch, float lr) {
    if (!c) return fail(HT_ERR_OTHER, "ht_cache_train: null cache");
    if (batch == 0 || batch % 64 || batch > c->max_batch)
        return fail(HT_ERR_OTHER, "ht_cache_train: batch must be a positive multiple of 64 up to max_batch");
    HIP_TRY(hipSetDevice(c->ctx->device));
    for (uint32_t k = 0; k < steps; ++k) {
        int rc = train_batch(c, c->counters + 1, 0, c->tcap, k * batch, batch, c->t_feat, c->t_target, lr, 1);
        if (rc) return rc;
    }
    return HT_OK;
}
