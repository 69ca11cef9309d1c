// The troubled-node indicator network: an MLP 16 -> 64 -> 64 -> 1 with GELU (tanh form), BF16 weights and
// activations, FP32 accumulation, on the CDNA4 matrix cores (v_mfma_f32_32x32x16_bf16, v_mfma_f32_16x16x32_bf16).
// Forward, backward (binary cross entropy on the logit with a weight on the positive class) and Adam run on the
// device; the gradients of a batch are reduced over workgroups in a fixed order, so training is deterministic.
//
// One wave handles a tile of 32 samples. Matrix operand fragments (32x32x16: lane l holds row / column l % 32 and the
// 8 consecutive K values 8 (l / 32) .. +7; result element (r, c) sits in lane c + 32 ((r / 4) % 2), register
// 4 (r / 8) + r % 4) are read from LDS copies laid out so that each fragment is one 16-byte read: activations by
// sample row ([r][k]) and transposed ([k][r]) where the K of a product runs over samples.
//
// Parameters (fp32, this order): W1[16][64] (input feature, hidden unit), b1[64], W2[64][64], b2[64], w3[64], b3.
#include "hcc_mlp.h"
#include "hcc_device.h"
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

using namespace hcc;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int F = kMlpIn, H = kMlpHidden, T = 32;   // features, hidden units, samples per tile
constexpr int P_W1 = 0, P_B1 = F * H, P_W2 = P_B1 + H, P_B2 = P_W2 + H * H, P_W3 = P_B2 + H, P_B3 = P_W3 + H;
static_assert(P_B3 + 1 == kMlpParams, "parameter layout");

// bf16 copies of the weights, laid out as the matrix cores read them (16-byte aligned rows; the size a multiple of
// 16 bytes, copied to LDS 16 bytes at a time)
struct alignas(16) MlpW {
    __bf16 w1t[H][F];   // W1 transposed: [hidden unit][feature]
    __bf16 w2t[H][H];   // W2 transposed: [layer-2 unit][layer-1 unit]
    __bf16 w2[H][H];    // W2: [layer-1 unit][layer-2 unit] (backward: dH1 = dA2 W2^T)
    __bf16 w3[H];
    float b1[H], b2[H], b3;
};

// tanh(u) = 1 - 2 / (exp(2u) + 1): one exponential and one reciprocal (saturates correctly at both ends)
__device__ __forceinline__ float tanh_e(float u) { return 1.0f - 2.0f / (__expf(2.0f * u) + 1.0f); }
__device__ __forceinline__ float gelu(float x) {
    const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
    return 0.5f * x * (1.0f + tanh_e(u));
}
__device__ __forceinline__ float gelu_grad(float x) {
    const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
    const float t = tanh_e(u);
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * x * x);
}
__device__ __forceinline__ float softplus(float z) { return z > 0.0f ? z + log1pf(expf(-z)) : log1pf(expf(z)); }
__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

__device__ __forceinline__ bf16x8 lds8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

// row of result register v of lane l in a 32x32 tile
__device__ __forceinline__ int crow(int v, int l) { return 8 * (v / 4) + 4 * (l / 32) + (v % 4); }

// Stores a 32x32 result tile (rows r, columns c0 + l % 32) as bf16 into a [row][ld] matrix and, if tr, into a
// [col][32] transposed matrix.
__device__ __forceinline__ void store_tile(const float (&v)[16], int l, int c0, __bf16* rm, int ld, __bf16* tr) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int r = crow(q, l), c = c0 + l % 32;
        const __bf16 b = (__bf16)v[q];
        if (rm) rm[r * ld + c] = b;
        if (tr) tr[c * T + r] = b;
    }
}

// Per-wave LDS of a tile (every fragment row 16-byte aligned)
struct alignas(16) TileLds {
    __bf16 x[T][F];     // inputs by sample
    __bf16 xt[F][T];    // inputs transposed
    __bf16 h1[T][H];    // layer-1 activations by sample
    __bf16 h1t[H][T];   // transposed
    __bf16 d2[T][H];    // dL/dA2 by sample
    __bf16 d2t[H][T];   // transposed
    __bf16 d1t[H][T];   // dL/dA1 transposed
    float z[T];
};

__device__ __forceinline__ void load_weights(MlpW& w, const MlpW* gw) {
    const int4* s = reinterpret_cast<const int4*>(gw);
    int4* d = reinterpret_cast<int4*>(&w);
    for (int t = threadIdx.x; t < (int)(sizeof(MlpW) / 16); t += blockDim.x) d[t] = s[t];
}

// Forward of the tile in x ([32][16]): a1 / a2 hold the pre-activations (result layout), h2 the layer-2 activations as
// bf16 values, z the logits; the layer-1 activations go to h1 ([32][64]) and, for training, h1t ([64][32]). Every
// lane of the wave takes part.
template <class Tile>
__device__ __forceinline__ void forward_tile(const MlpW& w, Tile& t, __bf16* h1t, int l, float (&a1)[2][16],
                                             float (&a2)[2][16], float (&h2)[2][16]) {
    const bf16x8 ax = lds8(&t.x[l % 32][8 * (l / 32)]);
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        f32x16 c = {};
        c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ax, lds8(&w.w1t[nt * 32 + l % 32][8 * (l / 32)]), c, 0, 0, 0);
        float h[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) { a1[nt][q] = c[q] + w.b1[nt * 32 + l % 32]; h[q] = gelu(a1[nt][q]); }
        store_tile(h, l, nt * 32, &t.h1[0][0], H, h1t);
    }
    __syncthreads();
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        f32x16 c = {};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
            c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds8(&t.h1[l % 32][kt * 16 + 8 * (l / 32)]),
                                                        lds8(&w.w2t[nt * 32 + l % 32][kt * 16 + 8 * (l / 32)]), c, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            a2[nt][q] = c[q] + w.b2[nt * 32 + l % 32];
            h2[nt][q] = (float)(__bf16)gelu(a2[nt][q]);
        }
    }
    // logit: z_r = sum_c bf16(h2[r][c]) w3[c] + b3, the 64 columns summed over the 32 lanes of a half-wave
    float zp[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) zp[q] = h2[0][q] * (float)w.w3[l % 32] + h2[1][q] * (float)w.w3[32 + l % 32];
#pragma unroll
    for (int off = 16; off > 0; off >>= 1)
#pragma unroll
        for (int q = 0; q < 16; ++q) zp[q] += __shfl_xor(zp[q], off);
    if (l % 32 == 0)
#pragma unroll
        for (int q = 0; q < 16; ++q) t.z[crow(q, l)] = zp[q] + w.b3;
    __syncthreads();
}

// Loads tile `tile` of X (n samples; sample s is row idx[s] when idx is given) into t.x and, if xt, transposed;
// samples past n are zero.
template <class Tile>
__device__ __forceinline__ void load_x(Tile& t, __bf16 (*xt)[T], const __bf16* X, int64_t n, int64_t tile, int l,
                                       const int32_t* idx = nullptr) {
    const int r = l % 32, h = l / 32;
    const int64_t s = tile * T + r;
    bf16x8 v = {};
    if (s < n) v = *reinterpret_cast<const bf16x8*>(X + (idx ? (int64_t)idx[s] : s) * F + 8 * h);
    *reinterpret_cast<bf16x8*>(&t.x[r][8 * h]) = v;
    if (xt)
#pragma unroll
        for (int e = 0; e < 8; ++e) xt[8 * h + e][r] = v[e];
    __syncthreads();
}

// the LDS of an inference tile: inputs and layer-1 activations by sample
struct alignas(16) FwdLds {
    __bf16 x[T][F];
    __bf16 h1[T][H];
    float z[T];
};

// Logits of n samples: four waves per workgroup share the weights, each wave a tile of its own; the workgroup steps
// over groups of four tiles, so every wave reaches each barrier (a tile past the end computes on zeros, unstored).
// With idx, sample s is feature row idx[s] and n is read from *n_dev.
constexpr int kFwdWaves = 4;
// apply (not null): the indicator's decision fused in: a listed particle whose logit exceeds thr takes order 1, and
// the count of such particles is added to apply_count (one atomic per wave with a hit)
__global__ __launch_bounds__(64 * kFwdWaves) void k_mlp_forward(const MlpW* __restrict__ gw, const __bf16* __restrict__ X,
                                                                int64_t n_host, const int32_t* __restrict__ idx,
                                                                const unsigned long long* __restrict__ n_dev,
                                                                float* __restrict__ z, int8_t* __restrict__ apply,
                                                                float thr, unsigned long long* __restrict__ apply_count) {
    __shared__ MlpW w;
    __shared__ FwdLds tiles[kFwdWaves];
    load_weights(w, gw);
    __syncthreads();
    const int l = threadIdx.x % 64, wv = threadIdx.x / 64;
    FwdLds& t = tiles[wv];
    const int64_t n = n_dev ? (int64_t)*n_dev : n_host;
    const int64_t ntile = (n + T - 1) / T;
    for (int64_t base = (int64_t)blockIdx.x * kFwdWaves; base < ntile; base += (int64_t)gridDim.x * kFwdWaves) {
        const int64_t tile = base + wv;
        load_x(t, nullptr, X, n, tile, l, idx);
        float a1[2][16], a2[2][16], h2[2][16];
        forward_tile(w, t, nullptr, l, a1, a2, h2);
        bool hit = false;
        if (l < T && tile * T + l < n) {
            const float zl = t.z[l];
            z[tile * T + l] = zl;
            if (apply && zl > thr) {
                apply[idx[tile * T + l]] = 1;
                hit = true;
            }
        }
        if (apply) {
            const unsigned long long b = __ballot(hit);
            if (b && l == (int)(__ffsll((unsigned long long)b) - 1)) atomicAdd(apply_count, (unsigned long long)__popcll(b));
        }
        __syncthreads();
    }
}

// Gradient partials of one workgroup over its tiles of the batch [0, n): loss sum and the parameter gradients of
// mean_s wt_s (softplus(z_s) - y_s z_s), wt_s = pos_weight for y_s = 1, 1 else; written to part[blockIdx.x][P + 1]
// (loss last).
__global__ __launch_bounds__(64) void k_mlp_grad(const MlpW* __restrict__ gw, const __bf16* __restrict__ X,
                                                 const uint8_t* __restrict__ y, int64_t n, float pos_weight,
                                                 float* __restrict__ part) {
    __shared__ MlpW w;
    __shared__ TileLds t;
    load_weights(w, gw);
    __syncthreads();
    const int l = threadIdx.x;
    const float inv_n = 1.0f / (float)n;
    // accumulators: dW2 as four 32x32 result tiles [i tile][j tile], dW1 as four 16x16 tiles (16x16x32 layout:
    // element (r, c) in lane c + 16 (r / 4), register r % 4), per-lane column sums for the biases and w3
    f32x16 gw2[2][2] = {};
    f32x4 gw1[4] = {};
    float gb1[2] = {}, gb2[2] = {}, gw3[2] = {}, gb3 = 0.0f, loss = 0.0f;
    const int64_t ntile = (n + T - 1) / T;
    for (int64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        load_x(t, t.xt, X, n, tile, l);
        float a1[2][16], a2[2][16], h2[2][16];
        forward_tile(w, t, &t.h1t[0][0], l, a1, a2, h2);
        // dL/dz per row of this lane's result registers; rows past n weigh 0
        float dz[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int r = crow(q, l);
            const int64_t s = tile * T + r;
            const float zz = t.z[r];
            float g = 0.0f;
            if (s < n) {
                const float yy = y[s] ? 1.0f : 0.0f, wt = y[s] ? pos_weight : 1.0f;
                g = wt * (sigmoid(zz) - yy) * inv_n;
                if (l % 32 == 0) loss += wt * (softplus(zz) - yy * zz);
            }
            dz[q] = g;
        }
        float d2[2][16];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const int c = nt * 32 + l % 32;
            const float w3c = (float)w.w3[c];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                gw3[nt] += h2[nt][q] * dz[q];
                d2[nt][q] = dz[q] * w3c * gelu_grad(a2[nt][q]);
                gb2[nt] += d2[nt][q];
            }
            store_tile(d2[nt], l, nt * 32, &t.d2[0][0], H, &t.d2t[0][0]);
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) gb3 += (l % 32 == 0) ? dz[q] : 0.0f;
        __syncthreads();
        // dW2[i][j] += sum_r h1[r][i] d2[r][j]: M = i, N = j, K = the 32 samples
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                for (int kt = 0; kt < 2; ++kt)
                    gw2[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                        lds8(&t.h1t[mt * 32 + l % 32][kt * 16 + 8 * (l / 32)]),
                        lds8(&t.d2t[nt * 32 + l % 32][kt * 16 + 8 * (l / 32)]), gw2[mt][nt], 0, 0, 0);
        // dH1[r][i] = sum_j d2[r][j] W2[i][j]; dA1 = dH1 gelu'(a1)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            f32x16 c = {};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt)
                c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds8(&t.d2[l % 32][kt * 16 + 8 * (l / 32)]),
                                                            lds8(&w.w2[nt * 32 + l % 32][kt * 16 + 8 * (l / 32)]), c, 0, 0, 0);
            float d1[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) { d1[q] = c[q] * gelu_grad(a1[nt][q]); gb1[nt] += d1[q]; }
            store_tile(d1, l, nt * 32, nullptr, 0, &t.d1t[0][0]);
        }
        __syncthreads();
        // dW1[f][i] += sum_r x[r][f] d1[r][i]: 16x16x32 with M = f, N = i (four tiles), K = the 32 samples
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
            gw1[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lds8(&t.xt[l % 16][8 * (l / 16)]),
                                                              lds8(&t.d1t[nt * 16 + l % 16][8 * (l / 16)]), gw1[nt], 0, 0, 0);
        __syncthreads();
    }
    // column sums: lanes l and l + 32 hold other rows of the same column
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        gb1[nt] += __shfl_xor(gb1[nt], 32);
        gb2[nt] += __shfl_xor(gb2[nt], 32);
        gw3[nt] += __shfl_xor(gw3[nt], 32);
    }
    gb3 += __shfl_xor(gb3, 32);
    loss += __shfl_xor(loss, 32);
    float* o = part + (int64_t)blockIdx.x * (kMlpParams + 1);
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int q = 0; q < 16; ++q) o[P_W2 + (mt * 32 + crow(q, l)) * H + nt * 32 + l % 32] = gw2[mt][nt][q];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int q = 0; q < 4; ++q) o[P_W1 + (4 * (l / 16) + q) * H + nt * 16 + l % 16] = gw1[nt][q];
    if (l < 32)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            o[P_B1 + nt * 32 + l] = gb1[nt];
            o[P_B2 + nt * 32 + l] = gb2[nt];
            o[P_W3 + nt * 32 + l] = gw3[nt];
        }
    if (l == 0) { o[P_B3] = gb3; o[kMlpParams] = loss * inv_n; }
}

// Sums the workgroup partials in workgroup order: g[p] = sum_b part[b][p] (p < P + 1).
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ part, int nb, float* __restrict__ g) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p > kMlpParams) return;
    float s = 0.0f;
    for (int b = 0; b < nb; ++b) s += part[(int64_t)b * (kMlpParams + 1) + p];
    g[p] = s;
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) on the fp32 masters, then the bf16 copies.
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ prm, float* __restrict__ m, float* __restrict__ v,
                                                  const float* __restrict__ g, float lr, float c1, float c2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kMlpParams) return;
    const float gp = g[p];
    m[p] = 0.9f * m[p] + 0.1f * gp;
    v[p] = 0.999f * v[p] + 0.001f * gp * gp;
    prm[p] -= lr * (m[p] / c1) / (sqrtf(v[p] / c2) + 1e-8f);
}

// The bf16 operand copies of the fp32 parameters.
__global__ __launch_bounds__(256) void k_mlp_pack(const float* __restrict__ prm, MlpW* __restrict__ w) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < F * H) { const int f = t / H, i = t % H; w->w1t[i][f] = (__bf16)prm[P_W1 + t]; }
    if (t < H * H) {
        const int i = t / H, j = t % H;
        w->w2t[j][i] = (__bf16)prm[P_W2 + t];
        w->w2[i][j] = (__bf16)prm[P_W2 + t];
    }
    if (t < H) {
        w->w3[t] = (__bf16)prm[P_W3 + t];
        w->b1[t] = prm[P_B1 + t];
        w->b2[t] = prm[P_B2 + t];
    }
    if (t == 0) w->b3 = prm[P_B3];
}

}  // namespace

struct hcc_mlp {
    hcc_ctx* ctx = nullptr;
    float* prm = nullptr;   // fp32 masters
    float* m = nullptr;     // Adam moments
    float* v = nullptr;
    int64_t step = 0;
    MlpW* w = nullptr;      // bf16 copies
    float* g = nullptr;     // reduced gradient (+ loss)
    float* part = nullptr;  // per-workgroup partials
};

namespace hcc {
hcc_ctx* mlp_context(const hcc_mlp* m) { return m->ctx; }
// Logits of n feature rows already on the device (the stepper's indicator pass), queued on the context stream.
void mlp_forward_device(const hcc_mlp* m, const uint16_t* X, int64_t n, float* z) {
    if (n <= 0) return;
    const int64_t ngroup = ((n + T - 1) / T + kFwdWaves - 1) / kFwdWaves;
    const unsigned nb = (unsigned)std::min<int64_t>(ngroup, 4096);
    k_mlp_forward<<<nb, 64 * kFwdWaves, 0, m->ctx->stream>>>(m->w, reinterpret_cast<const __bf16*>(X), n, nullptr,
                                                             nullptr, z, nullptr, 0.0f, nullptr);
}
void mlp_forward_list(const hcc_mlp* m, const uint16_t* X, const int32_t* idx, const unsigned long long* n_dev,
                      int64_t n_max, float* z, int8_t* apply, float thr, unsigned long long* apply_count) {
    if (n_max <= 0) return;
    // the list's length is on the device: up to 1024 workgroups looping over its tiles (a quarter of that measured
    // slower: a tile's forward pass is a chain of LDS round trips, so the tiles want waves in flight)
    const int64_t ngroup = ((n_max + T - 1) / T + kFwdWaves - 1) / kFwdWaves;
    const unsigned nb = (unsigned)std::min<int64_t>(ngroup, 1024);
    k_mlp_forward<<<nb, 64 * kFwdWaves, 0, m->ctx->stream>>>(m->w, reinterpret_cast<const __bf16*>(X), 0, idx, n_dev, z,
                                                             apply, thr, apply_count);
}
}  // namespace hcc

static constexpr int kGradBlocks = 512;

// Scratch of hcc_mlp_gradients_noalloc: the per-workgroup gradient partials (kGradBlocks rows of P + 1 floats,
// 10.4 MiB), a file-scope device array that every GPU the library runs on holds until the process ends. Host code
// never passes `scratch` itself to a kernel (on the host that name is a stand-in, not a device address): it passes
// the address hipGetSymbolAddress returned for the current device, resolved once per device by hcc_mlp_create.
constexpr int64_t kScratchRows = kGradBlocks;
static __device__ __align__(256) float scratch[kScratchRows * (kMlpParams + 1)];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static float* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (the current device); resolved on the first call per device.
static hipError_t scratch_address(int device, float** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

extern "C" int hcc_mlp_destroy(hcc_mlp* m) {
    if (!m) return HCC_OK;
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    for (void* p : {(void*)m->prm, (void*)m->m, (void*)m->v, (void*)m->w, (void*)m->g, (void*)m->part}) (void)hipFree(p);
    delete m;
    return HCC_OK;
}

extern "C" int hcc_mlp_create(hcc_ctx* ctx, hcc_mlp** out) {
    if (!ctx || !out) return set_err(ctx, HCC_E_ARG, "hcc_mlp_create: NULL argument");
    *out = nullptr;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    std::unique_ptr<hcc_mlp, int (*)(hcc_mlp*)> m(new hcc_mlp(), hcc_mlp_destroy);
    m->ctx = ctx;
    HCC_TRY_HIP(ctx, dalloc(&m->prm, kMlpParams));
    HCC_TRY_HIP(ctx, dalloc(&m->m, kMlpParams));
    HCC_TRY_HIP(ctx, dalloc(&m->v, kMlpParams));
    HCC_TRY_HIP(ctx, hipMalloc((void**)&m->w, sizeof(MlpW)));
    HCC_TRY_HIP(ctx, dalloc(&m->g, kMlpParams + 1));
    HCC_TRY_HIP(ctx, dalloc(&m->part, (int64_t)kGradBlocks * (kMlpParams + 1)));
    for (float* p : {m->prm, m->m, m->v}) HCC_TRY_HIP(ctx, hipMemset(p, 0, sizeof(float) * kMlpParams));
    HCC_TRY_HIP(ctx, hipMemset(m->w, 0, sizeof(MlpW)));
    float* sp = nullptr;   // the no-allocation gradient's scratch on this device (resolved once per device)
    HCC_TRY_HIP(ctx, scratch_address(ctx->device, &sp));
    *out = m.release();
    return HCC_OK;
}

extern "C" int hcc_mlp_set_params(hcc_mlp* m, const float* params) {
    if (!m || !params) return HCC_E_ARG;
    hcc_ctx* ctx = m->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    HCC_TRY_HIP(ctx, hipMemcpy(m->prm, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    HCC_TRY_HIP(ctx, hipMemset(m->m, 0, sizeof(float) * kMlpParams));
    HCC_TRY_HIP(ctx, hipMemset(m->v, 0, sizeof(float) * kMlpParams));
    m->step = 0;
    k_mlp_pack<<<(H * H + 255) / 256, 256, 0, ctx->stream>>>(m->prm, m->w);
    HCC_TRY_HIP(ctx, hipGetLastError());
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return HCC_OK;
}

extern "C" int hcc_mlp_get_params(hcc_mlp* m, float* params) {
    if (!m || !params) return HCC_E_ARG;
    hcc_ctx* ctx = m->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    HCC_TRY_HIP(ctx, hipMemcpy(params, m->prm, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    return HCC_OK;
}

// the batch's features and labels on the device for the call
struct DevBatch {
    uint16_t* X = nullptr;
    uint8_t* y = nullptr;
    ~DevBatch() { (void)hipFree(X); (void)hipFree(y); }
};
static int upload_batch(hcc_ctx* ctx, const uint16_t* X, const uint8_t* y, int64_t n, DevBatch& b) {
    HCC_TRY_HIP(ctx, dalloc(&b.X, n * F));
    HCC_TRY_HIP(ctx, hipMemcpy(b.X, X, sizeof(uint16_t) * n * F, hipMemcpyHostToDevice));
    if (y) {
        HCC_TRY_HIP(ctx, dalloc(&b.y, n));
        HCC_TRY_HIP(ctx, hipMemcpy(b.y, y, n, hipMemcpyHostToDevice));
    }
    return HCC_OK;
}

extern "C" int hcc_mlp_forward(hcc_mlp* m, const uint16_t* X, int64_t n, float* z) {
    if (!m || n < 0 || (n > 0 && (!X || !z))) return HCC_E_ARG;
    hcc_ctx* ctx = m->ctx;
    if (n == 0) return HCC_OK;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    DevBatch b;
    if (int rc = upload_batch(ctx, X, nullptr, n, b)) return rc;
    float* dz = nullptr;
    HCC_TRY_HIP(ctx, dalloc(&dz, n));
    mlp_forward_device(m, b.X, n, dz);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(z, dz, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dz);
    HCC_TRY_HIP(ctx, e);
    return HCC_OK;
}

// workgroups of the gradient kernel for a batch of n samples (one tile of T samples each, at most kGradBlocks)
static int grad_blocks(int64_t n) { return (int)std::min<int64_t>((n + T - 1) / T, kGradBlocks); }

// gradient of the mean weighted loss over the batch into g (+ loss), on the device, with the workgroup partials in
// part (grad_blocks(n) rows of P + 1 floats)
static void grad_launch(hcc_mlp* m, const uint16_t* X, const uint8_t* y, int64_t n, float pos_weight, float* part,
                        float* g) {
    const int nb = grad_blocks(n);
    hipStream_t st = m->ctx->stream;
    k_mlp_grad<<<nb, 64, 0, st>>>(m->w, reinterpret_cast<const __bf16*>(X), y, n, pos_weight, part);
    k_mlp_reduce<<<(kMlpParams + 1 + 255) / 256, 256, 0, st>>>(part, nb, g);
}
static void grad_device(hcc_mlp* m, const uint16_t* X, const uint8_t* y, int64_t n, float pos_weight) {
    grad_launch(m, X, y, n, pos_weight, m->part, m->g);
}

extern "C" int hcc_mlp_gradients(hcc_mlp* m, const uint16_t* X, const uint8_t* y, int64_t n, float pos_weight,
                                 float* grads, double* loss) {
    if (!m || n <= 0 || !X || !y) return set_err(m ? m->ctx : nullptr, HCC_E_ARG, "hcc_mlp_gradients: bad arguments");
    hcc_ctx* ctx = m->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    DevBatch b;
    if (int rc = upload_batch(ctx, X, y, n, b)) return rc;
    grad_device(m, b.X, b.y, n, pos_weight);
    HCC_TRY_HIP(ctx, hipGetLastError());
    std::vector<float> g(kMlpParams + 1);
    HCC_TRY_HIP(ctx, hipMemcpyAsync(g.data(), m->g, sizeof(float) * g.size(), hipMemcpyDeviceToHost, ctx->stream));
    HCC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (grads) std::memcpy(grads, g.data(), sizeof(float) * kMlpParams);
    if (loss) *loss = g[kMlpParams];
    return HCC_OK;
}

extern "C" int hcc_mlp_gradients_noalloc(hcc_mlp* m, const uint16_t* X, const uint8_t* y, int64_t n, float pos_weight,
                                         float* grads) {
    if (!m || n <= 0 || !X || !y || !grads)
        return set_err(m ? m->ctx : nullptr, HCC_E_ARG, "hcc_mlp_gradients_noalloc: bad arguments");
    hcc_ctx* ctx = m->ctx;
    // the partials take one row per workgroup; every batch size fits (at most kScratchRows workgroups), the check
    // guards the scratch against a change of the launch shape
    if (grad_blocks(n) > kScratchRows)
        return set_err(ctx, HCC_E_CAPACITY, "hcc_mlp_gradients_noalloc: batch needs more partial rows than the "
                                            "static scratch holds");
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    float* sp = nullptr;
    HCC_TRY_HIP(ctx, scratch_address(ctx->device, &sp));   // resolved by hcc_mlp_create: a table read
    grad_launch(m, X, y, n, pos_weight, sp, grads);
    HCC_TRY_HIP(ctx, hipGetLastError());
    return HCC_OK;
}

extern "C" int hcc_mlp_train(hcc_mlp* m, const uint16_t* X, const uint8_t* y, int64_t n, int32_t epochs, int32_t batch,
                             float lr, float pos_weight, double* losses) {
    if (!m || n <= 0 || !X || !y || epochs < 1 || batch < T || !(lr > 0))
        return set_err(m ? m->ctx : nullptr, HCC_E_ARG, "hcc_mlp_train: bad arguments");
    hcc_ctx* ctx = m->ctx;
    HCC_TRY_HIP(ctx, hipSetDevice(ctx->device));
    DevBatch b;
    if (int rc = upload_batch(ctx, X, y, n, b)) return rc;
    hipStream_t st = ctx->stream;
    std::vector<double> el((size_t)epochs, 0.0);
    float* hloss = nullptr;
    HCC_TRY_HIP(ctx, hipHostMalloc((void**)&hloss, sizeof(float), hipHostMallocDefault));
    std::unique_ptr<float, hipError_t (*)(void*)> keep(hloss, hipHostFree);
    for (int ep = 0; ep < epochs; ++ep) {
        double sum = 0.0;
        int64_t nbatch = 0;
        for (int64_t s0 = 0; s0 < n; s0 += batch) {   // the samples in their given order (the caller shuffles)
            const int64_t nb = std::min<int64_t>(batch, n - s0);
            grad_device(m, b.X + s0 * F, b.y + s0, nb, pos_weight);
            ++m->step;
            const float c1 = 1.0f - std::pow(0.9f, (float)m->step), c2 = 1.0f - std::pow(0.999f, (float)m->step);
            k_mlp_adam<<<(kMlpParams + 255) / 256, 256, 0, st>>>(m->prm, m->m, m->v, m->g, lr, c1, c2);
            k_mlp_pack<<<(H * H + 255) / 256, 256, 0, st>>>(m->prm, m->w);
            if (losses) {
                HCC_TRY_HIP(ctx, hipMemcpyAsync(hloss, m->g + kMlpParams, sizeof(float), hipMemcpyDeviceToHost, st));
                HCC_TRY_HIP(ctx, hipStreamSynchronize(st));
                sum += *hloss;
                ++nbatch;
            }
        }
        HCC_TRY_HIP(ctx, hipGetLastError());
        el[(size_t)ep] = nbatch ? sum / (double)nbatch : 0.0;
    }
    HCC_TRY_HIP(ctx, hipStreamSynchronize(st));
    if (losses) std::memcpy(losses, el.data(), sizeof(double) * epochs);
    return HCC_OK;
}
