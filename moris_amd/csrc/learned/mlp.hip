// mlp.hip -- the learned cut-cell quadrature network on MFMA: 8 -> 64 -> 64 -> 27, SiLU.
//
// One wavefront owns a tile of 16 samples and runs every layer as 16x16x16 BF16 MFMA tiles with
// FP32 accumulation (__builtin_amdgcn_mfma_f32_16x16x16bf16_1k); activations move between layers
// through LDS in the layout the next MFMA operand needs.  Training: fused forward + backward per
// tile, FP32 weight gradients summed over the batch with atomics, Adam on FP32 master weights, and
// a conversion kernel that refreshes the BF16 operand copies of the weights.
//
// MFMA 16x16x16 operand layouts (lane l, 64 lanes):
//   A (16 x 16, M x K): row l%16, k = 4*(l/16) + 0..3
//   B (16 x 16, K x N): col l%16, k = 4*(l/16) + 0..3
//   C (16 x 16, M x N): col l%16, rows 4*(l/16) + 0..3
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../setup/internal.hpp"
#include "quad.hpp"

namespace mgpu {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short f2bf(float x) {
    unsigned int u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
    return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short b) { return __uint_as_float(((unsigned int)b) << 16); }

__device__ inline floatx4 mfma16(bf16x4 a, bf16x4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// fragment of a row-major bf16 matrix M[rows][ld]: rows r0 + l%16, cols k0 + 4*(l/16) .. +3
__device__ inline bf16x4 frag(const unsigned short* M, int ld, int r0, int k0, int lane) {
    const unsigned short* p = M + (size_t)(r0 + lane % 16) * ld + k0 + 4 * (lane / 16);
    bf16x4 v;
    v[0] = (short)p[0];
    v[1] = (short)p[1];
    v[2] = (short)p[2];
    v[3] = (short)p[3];
    return v;
}

__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ inline float dsilu(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

constexpr int H = kLqHidden;   // 64
constexpr int KI = 16;         // padded input width
constexpr int NO = 32;         // padded output width

// BF16 operand copies of the weights
struct MlpDev {
    float* params;          // FP32 master [kLqParams]
    float* grads;           // FP32 [kLqParams]
    float* adam_m;
    float* adam_v;
    unsigned short* w1t;    // [64][16]  (out, in) forward B operand
    unsigned short* w2t;    // [64][64]  (out, in)
    unsigned short* w3t;    // [32][64]  (out, in)
    unsigned short* w1;     // [16][64]  (in, out) backward B operand
    unsigned short* w2;     // [64][64]
    unsigned short* w3;     // [64][32]
    float* loss;            // [1]
};

__global__ void k_mlp_convert(MlpDev d) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const float* P = d.params;
    if (t < H * KI) {   // w1t[o][i], w1[i][o]
        const int o = t / KI, i = t % KI;
        const float v = i < kLqInputs ? P[kLqOffW1 + i * H + o] : 0.0f;
        d.w1t[o * KI + i] = f2bf(v);
        d.w1[i * H + o] = f2bf(v);
    }
    if (t < H * H) {
        const int o = t / H, i = t % H;
        const float v = P[kLqOffW2 + i * H + o];
        d.w2t[o * H + i] = f2bf(v);
        d.w2[i * H + o] = f2bf(v);
    }
    if (t < NO * H) {
        const int o = t / H, i = t % H;
        const float v = o < kLqPoints ? P[kLqOffW3 + i * kLqPoints + o] : 0.0f;
        d.w3t[o * H + i] = f2bf(v);
        d.w3[i * NO + o] = f2bf(v);
    }
}

// Per-wave LDS workspace of one 16-sample tile.
struct TileLds {
    unsigned short xs[16 * KI];   // X   sample-major
    unsigned short xf[KI * 16];   // X   feature-major
    float z1[16 * H];             // pre-activations layer 1 (sample-major)
    float z2[16 * H];
    unsigned short h1s[16 * H], h1f[H * 16];
    unsigned short h2s[16 * H], h2f[H * 16];
    unsigned short dys[16 * NO], dyf[NO * 16];
    unsigned short dz2s[16 * H], dz2f[H * 16];
    unsigned short dz1f[H * 16];
};

// Forward of one tile; writes y (FP32, sample-major [16][32]) into the caller's registers via LDS.
__device__ inline void tile_forward(const MlpDev& d, TileLds& L, int lane, float* y_out /* [16][32] in LDS */) {
    const float* P = d.params;
    // layer 1
    for (int nt = 0; nt < H / 16; ++nt) {
        floatx4 c = {0.f, 0.f, 0.f, 0.f};
        c = mfma16(frag(L.xs, KI, 0, 0, lane), frag(d.w1t, KI, nt * 16, 0, lane), c);
        const int col = nt * 16 + lane % 16;
        for (int j = 0; j < 4; ++j) {
            const int row = 4 * (lane / 16) + j;
            const float z = c[j] + P[kLqOffB1 + col];
            L.z1[row * H + col] = z;
            const unsigned short hb = f2bf(silu(z));
            L.h1s[row * H + col] = hb;
            L.h1f[col * 16 + row] = hb;
        }
    }
    __syncthreads();
    // layer 2
    for (int nt = 0; nt < H / 16; ++nt) {
        floatx4 c = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < H / 16; ++ks)
            c = mfma16(frag(L.h1s, H, 0, ks * 16, lane), frag(d.w2t, H, nt * 16, ks * 16, lane), c);
        const int col = nt * 16 + lane % 16;
        for (int j = 0; j < 4; ++j) {
            const int row = 4 * (lane / 16) + j;
            const float z = c[j] + P[kLqOffB2 + col];
            L.z2[row * H + col] = z;
            const unsigned short hb = f2bf(silu(z));
            L.h2s[row * H + col] = hb;
            L.h2f[col * 16 + row] = hb;
        }
    }
    __syncthreads();
    // layer 3 (linear)
    for (int nt = 0; nt < NO / 16; ++nt) {
        floatx4 c = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < H / 16; ++ks)
            c = mfma16(frag(L.h2s, H, 0, ks * 16, lane), frag(d.w3t, H, nt * 16, ks * 16, lane), c);
        const int col = nt * 16 + lane % 16;
        for (int j = 0; j < 4; ++j) {
            const int row = 4 * (lane / 16) + j;
            y_out[row * NO + col] = col < kLqPoints ? c[j] + P[kLqOffB3 + col] : 0.0f;
        }
    }
    __syncthreads();
}

__device__ inline void load_inputs(TileLds& L, const float* x, int64_t s0, int64_t n, int lane) {
    for (int idx = lane; idx < 16 * KI; idx += 64) {
        const int s = idx / KI, i = idx % KI;
        float v = 0.0f;
        if (i < kLqInputs && s0 + s < n) v = x[(s0 + s) * kLqInputs + i];
        const unsigned short b = f2bf(v);
        L.xs[s * KI + i] = b;
        L.xf[i * 16 + s] = b;
    }
    __syncthreads();
}

// Inference workspace of one 16-sample tile (no backward buffers: small enough for high occupancy).
struct InferLds {
    unsigned short xs[16 * KI];
    unsigned short h1s[16 * H];
    unsigned short h2s[16 * H];
    float yo[16 * NO];
};

// Forward of one tile for inference (same arithmetic as tile_forward: BF16 operands, FP32
// accumulation, SiLU on the FP32 pre-activation).
__device__ inline void tile_infer(const MlpDev& d, InferLds& L, int lane) {
    const float* P = d.params;
    for (int nt = 0; nt < H / 16; ++nt) {
        floatx4 c = {0.f, 0.f, 0.f, 0.f};
        c = mfma16(frag(L.xs, KI, 0, 0, lane), frag(d.w1t, KI, nt * 16, 0, lane), c);
        const int col = nt * 16 + lane % 16;
        for (int j = 0; j < 4; ++j) L.h1s[(4 * (lane / 16) + j) * H + col] = f2bf(silu(c[j] + P[kLqOffB1 + col]));
    }
    __syncthreads();
    for (int nt = 0; nt < H / 16; ++nt) {
        floatx4 c = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < H / 16; ++ks)
            c = mfma16(frag(L.h1s, H, 0, ks * 16, lane), frag(d.w2t, H, nt * 16, ks * 16, lane), c);
        const int col = nt * 16 + lane % 16;
        for (int j = 0; j < 4; ++j) L.h2s[(4 * (lane / 16) + j) * H + col] = f2bf(silu(c[j] + P[kLqOffB2 + col]));
    }
    __syncthreads();
    for (int nt = 0; nt < NO / 16; ++nt) {
        floatx4 c = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < H / 16; ++ks)
            c = mfma16(frag(L.h2s, H, 0, ks * 16, lane), frag(d.w3t, H, nt * 16, ks * 16, lane), c);
        const int col = nt * 16 + lane % 16;
        for (int j = 0; j < 4; ++j) L.yo[(4 * (lane / 16) + j) * NO + col] = col < kLqPoints ? c[j] + P[kLqOffB3 + col] : 0.0f;
    }
    __syncthreads();
}

// forward only: y[n][27]; each wavefront loops over 16-sample tiles.  PHI: the inputs are formed
// from the 8 nodal level-set values of each cell, scaled by their largest magnitude (FP64, then
// FP32), instead of being read from x.
template <bool PHI>
__global__ __launch_bounds__(64) void k_mlp_forward(MlpDev d, const float* x, const double* vphi, int64_t n,
                                                    float* y) {
    __shared__ InferLds L;
    const int lane = threadIdx.x;
    const int64_t n_tiles = (n + 15) / 16;
    for (int64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const int64_t s0 = tile * 16;
        for (int idx = lane; idx < 16 * KI; idx += 64) {
            const int sm = idx / KI, i = idx % KI;
            float v = 0.0f;
            if (i < kLqInputs && s0 + sm < n) {
                if constexpr (PHI) {
                    const double* ph = vphi + (s0 + sm) * kLqInputs;
                    double m = 0.0;
                    for (int k = 0; k < kLqInputs; ++k) m = fmax(m, fabs(ph[k]));
                    const double inv = m > 0.0 ? 1.0 / m : 0.0;
                    v = (float)(ph[i] * inv);
                } else {
                    v = x[(s0 + sm) * kLqInputs + i];
                }
            }
            L.xs[sm * KI + i] = f2bf(v);
        }
        __syncthreads();
        tile_infer(d, L, lane);
        for (int idx = lane; idx < 16 * kLqPoints; idx += 64) {
            const int sm = idx / kLqPoints, o = idx % kLqPoints;
            if (s0 + sm < n) y[(s0 + sm) * kLqPoints + o] = L.yo[sm * NO + o];
        }
        __syncthreads();
    }
}

// forward + backward of one tile; loss = sum (y - t)^2 / (n_total * 27); gradients accumulated
__global__ __launch_bounds__(64) void k_mlp_backward(MlpDev d, const float* x, const float* t, int64_t n,
                                                     float inv_count) {
    __shared__ TileLds L;
    __shared__ float yo[16 * NO];
    const int lane = threadIdx.x;
    const int64_t s0 = (int64_t)blockIdx.x * 16;
    load_inputs(L, x, s0, n, lane);
    tile_forward(d, L, lane, yo);
    // dY = 2 (y - t) / count; loss
    float lsum = 0.0f;
    for (int idx = lane; idx < 16 * NO; idx += 64) {
        const int s = idx / NO, o = idx % NO;
        float g = 0.0f;
        if (o < kLqPoints && s0 + s < n) {
            const float e = yo[s * NO + o] - t[(s0 + s) * kLqPoints + o];
            lsum += e * e;
            g = 2.0f * e * inv_count;
        }
        const unsigned short b = f2bf(g);
        L.dys[s * NO + o] = b;
        L.dyf[o * 16 + s] = b;
        yo[s * NO + o] = g;   // keep FP32 dY for the bias gradient
    }
    __syncthreads();
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (lane == 0) atomicAdd(d.loss, lsum * inv_count);
    float* G = d.grads;
    // db3, dW3 = H2^T dY  (M = 64 hidden, N = 32 outputs, K = 16 samples)
    if (lane < NO && lane < kLqPoints) {
        float s = 0.0f;
        for (int r = 0; r < 16; ++r) s += yo[r * NO + lane];
        atomicAdd(&G[kLqOffB3 + lane], s);
    }
    for (int mt = 0; mt < H / 16; ++mt)
        for (int nt = 0; nt < NO / 16; ++nt) {
            floatx4 c = {0.f, 0.f, 0.f, 0.f};
            c = mfma16(frag(L.h2f, 16, mt * 16, 0, lane), frag(L.dyf, 16, nt * 16, 0, lane), c);
            const int col = nt * 16 + lane % 16;
            if (col < kLqPoints)
                for (int j = 0; j < 4; ++j) {
                    const int row = mt * 16 + 4 * (lane / 16) + j;
                    atomicAdd(&G[kLqOffW3 + row * kLqPoints + col], c[j]);
                }
        }
    // dH2 = dY W3^T (M = 16, N = 64, K = 32) ; dZ2 = dH2 * silu'(Z2)
    for (int nt = 0; nt < H / 16; ++nt) {
        floatx4 c = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < NO / 16; ++ks)
            c = mfma16(frag(L.dys, NO, 0, ks * 16, lane), frag(d.w3, NO, nt * 16, ks * 16, lane), c);
        const int col = nt * 16 + lane % 16;
        for (int j = 0; j < 4; ++j) {
            const int row = 4 * (lane / 16) + j;
            const float dz = c[j] * dsilu(L.z2[row * H + col]);
            L.z2[row * H + col] = dz;   // reuse as FP32 dZ2 for the bias gradient
            const unsigned short b = f2bf(dz);
            L.dz2s[row * H + col] = b;
            L.dz2f[col * 16 + row] = b;
        }
    }
    __syncthreads();
    {   // db2
        float s = 0.0f;
        for (int r = 0; r < 16; ++r) s += L.z2[r * H + lane];
        atomicAdd(&G[kLqOffB2 + lane], s);
    }
    // dW2 = H1^T dZ2 (M = 64, N = 64, K = 16)
    for (int mt = 0; mt < H / 16; ++mt)
        for (int nt = 0; nt < H / 16; ++nt) {
            floatx4 c = {0.f, 0.f, 0.f, 0.f};
            c = mfma16(frag(L.h1f, 16, mt * 16, 0, lane), frag(L.dz2f, 16, nt * 16, 0, lane), c);
            const int col = nt * 16 + lane % 16;
            for (int j = 0; j < 4; ++j) {
                const int row = mt * 16 + 4 * (lane / 16) + j;
                atomicAdd(&G[kLqOffW2 + row * H + col], c[j]);
            }
        }
    // dH1 = dZ2 W2^T (M = 16, N = 64, K = 64) ; dZ1 = dH1 * silu'(Z1)
    for (int nt = 0; nt < H / 16; ++nt) {
        floatx4 c = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < H / 16; ++ks)
            c = mfma16(frag(L.dz2s, H, 0, ks * 16, lane), frag(d.w2, H, nt * 16, ks * 16, lane), c);
        const int col = nt * 16 + lane % 16;
        for (int j = 0; j < 4; ++j) {
            const int row = 4 * (lane / 16) + j;
            const float dz = c[j] * dsilu(L.z1[row * H + col]);
            L.z1[row * H + col] = dz;
            L.dz1f[col * 16 + row] = f2bf(dz);
        }
    }
    __syncthreads();
    {   // db1
        float s = 0.0f;
        for (int r = 0; r < 16; ++r) s += L.z1[r * H + lane];
        atomicAdd(&G[kLqOffB1 + lane], s);
    }
    // dW1 = X^T dZ1 (M = 16 inputs, N = 64, K = 16)
    for (int nt = 0; nt < H / 16; ++nt) {
        floatx4 c = {0.f, 0.f, 0.f, 0.f};
        c = mfma16(frag(L.xf, 16, 0, 0, lane), frag(L.dz1f, 16, nt * 16, 0, lane), c);
        const int col = nt * 16 + lane % 16;
        for (int j = 0; j < 4; ++j) {
            const int row = 4 * (lane / 16) + j;
            if (row < kLqInputs) atomicAdd(&G[kLqOffW1 + row * H + col], c[j]);
        }
    }
}

__global__ void k_adam(MlpDev d, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kLqParams) return;
    const float g = d.grads[i];
    const float m = b1 * d.adam_m[i] + (1.0f - b1) * g;
    const float v = b2 * d.adam_v[i] + (1.0f - b2) * g * g;
    d.adam_m[i] = m;
    d.adam_v[i] = v;
    d.params[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
    d.grads[i] = 0.0f;
}

}  // namespace mgpu

// ================================================================================================
using namespace mgpu;

struct mgpu_mlp {
    MlpDev d{};
    std::vector<void*> allocs;
    int64_t steps = 0;
};

namespace {
#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess) return set_error(MGPU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

template <class T>
mgpu_status dalloc(mgpu_mlp* m, size_t n, T** p) {
    void* q = nullptr;
    HIP_TRY(hipMalloc(&q, n * sizeof(T)));
    HIP_TRY(hipMemset(q, 0, n * sizeof(T)));
    m->allocs.push_back(q);
    *p = static_cast<T*>(q);
    return MGPU_OK;
}

mgpu_status refresh(mgpu_mlp* m) {
    hipLaunchKernelGGL(k_mlp_convert, dim3((H * H + 255) / 256), dim3(256), 0, 0, m->d);
    HIP_TRY(hipGetLastError());
    return MGPU_OK;
}

// Glorot-uniform initial weights, zero biases (seeded splitmix64)
void init_params(uint64_t seed, std::vector<float>& p) {
    p.assign(kLqParams, 0.0f);
    uint64_t c = 0;
    auto draw = [&](double lim) {
        uint64_t z = seed + (++c) * 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return (float)(lim * (2.0 * ((double)(z >> 11) / 9007199254740992.0) - 1.0));
    };
    const double l1 = std::sqrt(6.0 / (kLqInputs + H)), l2 = std::sqrt(6.0 / (H + H)), l3 = std::sqrt(6.0 / (H + kLqPoints));
    for (int i = 0; i < kLqInputs * H; ++i) p[kLqOffW1 + i] = draw(l1);
    for (int i = 0; i < H * H; ++i) p[kLqOffW2 + i] = draw(l2);
    for (int i = 0; i < H * kLqPoints; ++i) p[kLqOffW3 + i] = draw(l3);
}
}  // namespace

namespace mgpu {
// Device view of a model's weights, used by the assembly's inference kernel.
const void* mlp_device_view(const mgpu_mlp* m) { return &m->d; }
void mlp_forward_launch(const mgpu_mlp* m, const float* x, int64_t n, float* y, hipStream_t st) {
    if (n > 0) {
        const int64_t n_tiles = (n + 15) / 16;
        // one tile per block: a 6.6 KB block leaves room for every tile of a set resident at once
        const unsigned grid = (unsigned)std::min<int64_t>(n_tiles, 1 << 20);
        hipLaunchKernelGGL(k_mlp_forward<false>, dim3(grid), dim3(64), 0, st, m->d, x, nullptr, n, y);
    }
}
void mlp_forward_phi_launch(const mgpu_mlp* m, const double* vphi, int64_t n, float* y, hipStream_t st) {
    if (n > 0) {
        const int64_t n_tiles = (n + 15) / 16;
        const unsigned grid = (unsigned)std::min<int64_t>(n_tiles, 1 << 20);
        hipLaunchKernelGGL(k_mlp_forward<true>, dim3(grid), dim3(64), 0, st, m->d, nullptr, vphi, n, y);
    }
}
}  // namespace mgpu

extern "C" mgpu_status mgpu_mlp_create(uint64_t seed, mgpu_mlp** out) {
    MGPU_GUARD_BEGIN
    if (!out) return set_error(MGPU_ERR_INVALID_ARG, "null out");
    auto m = std::make_unique<mgpu_mlp>();
    mgpu_status rc;
    if ((rc = dalloc(m.get(), kLqParams, &m->d.params)) || (rc = dalloc(m.get(), kLqParams, &m->d.grads)) ||
        (rc = dalloc(m.get(), kLqParams, &m->d.adam_m)) || (rc = dalloc(m.get(), kLqParams, &m->d.adam_v)) ||
        (rc = dalloc(m.get(), H * KI, &m->d.w1t)) || (rc = dalloc(m.get(), H * H, &m->d.w2t)) ||
        (rc = dalloc(m.get(), NO * H, &m->d.w3t)) || (rc = dalloc(m.get(), KI * H, &m->d.w1)) ||
        (rc = dalloc(m.get(), H * H, &m->d.w2)) || (rc = dalloc(m.get(), H * NO, &m->d.w3)) ||
        (rc = dalloc(m.get(), 1, &m->d.loss)))
        return rc;
    std::vector<float> p;
    init_params(seed, p);
    HIP_TRY(hipMemcpy(m->d.params, p.data(), kLqParams * sizeof(float), hipMemcpyHostToDevice));
    if ((rc = refresh(m.get())) != MGPU_OK) return rc;
    HIP_TRY(hipDeviceSynchronize());
    *out = m.release();
    return MGPU_OK;
    MGPU_GUARD_END
}

extern "C" mgpu_status mgpu_mlp_destroy(mgpu_mlp* m) {
    if (!m) return MGPU_OK;
    for (void* p : m->allocs) (void)hipFree(p);
    delete m;
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_mlp_num_params(int32_t* n) {
    if (!n) return set_error(MGPU_ERR_INVALID_ARG, "null");
    *n = kLqParams;
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_mlp_get_params(const mgpu_mlp* m, float* host) {
    if (!m || !host) return set_error(MGPU_ERR_INVALID_ARG, "null");
    HIP_TRY(hipMemcpy(host, m->d.params, kLqParams * sizeof(float), hipMemcpyDeviceToHost));
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_mlp_set_params(mgpu_mlp* m, const float* host) {
    if (!m || !host) return set_error(MGPU_ERR_INVALID_ARG, "null");
    HIP_TRY(hipMemcpy(m->d.params, host, kLqParams * sizeof(float), hipMemcpyHostToDevice));
    mgpu_status rc = refresh(m);
    if (rc != MGPU_OK) return rc;
    HIP_TRY(hipDeviceSynchronize());
    return MGPU_OK;
}

extern "C" mgpu_status mgpu_mlp_forward(const mgpu_mlp* m, const float* x, int64_t n, float* y) {
    MGPU_GUARD_BEGIN
    if (!m || (n > 0 && (!x || !y)) || n < 0) return set_error(MGPU_ERR_INVALID_ARG, "bad argument");
    if (n == 0) return MGPU_OK;
    float *dx = nullptr, *dy = nullptr;
    HIP_TRY(hipMalloc(&dx, n * kLqInputs * sizeof(float)));
    HIP_TRY(hipMalloc(&dy, n * kLqPoints * sizeof(float)));
    HIP_TRY(hipMemcpy(dx, x, n * kLqInputs * sizeof(float), hipMemcpyHostToDevice));
    mlp_forward_launch(m, dx, n, dy, nullptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(y, dy, n * kLqPoints * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(dx);
    (void)hipFree(dy);
    if (e != hipSuccess) return set_error(MGPU_ERR_HIP, hipGetErrorString(e));
    return MGPU_OK;
    MGPU_GUARD_END
}

// Loss and parameter gradients of the mean squared error over n samples (no update).
extern "C" mgpu_status mgpu_mlp_gradients(mgpu_mlp* m, const float* x, const float* t, int64_t n, float* grad,
                                          float* loss) {
    MGPU_GUARD_BEGIN
    if (!m || !x || !t || n <= 0) return set_error(MGPU_ERR_INVALID_ARG, "bad argument");
    float *dx = nullptr, *dt = nullptr;
    HIP_TRY(hipMalloc(&dx, n * kLqInputs * sizeof(float)));
    HIP_TRY(hipMalloc(&dt, n * kLqPoints * sizeof(float)));
    HIP_TRY(hipMemcpy(dx, x, n * kLqInputs * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dt, t, n * kLqPoints * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(m->d.grads, 0, kLqParams * sizeof(float)));
    HIP_TRY(hipMemset(m->d.loss, 0, sizeof(float)));
    const float inv = 1.0f / (float)(n * kLqPoints);
    hipLaunchKernelGGL(k_mlp_backward, dim3((unsigned)((n + 15) / 16)), dim3(64), 0, 0, m->d, dx, dt, n, inv);
    HIP_TRY(hipGetLastError());
    if (grad) HIP_TRY(hipMemcpy(grad, m->d.grads, kLqParams * sizeof(float), hipMemcpyDeviceToHost));
    if (loss) HIP_TRY(hipMemcpy(loss, m->d.loss, sizeof(float), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemset(m->d.grads, 0, kLqParams * sizeof(float)));
    (void)hipFree(dx);
    (void)hipFree(dt);
    return MGPU_OK;
    MGPU_GUARD_END
}

// Adam training on host arrays (copied to HBM once): `epochs` passes over n samples in batches of
// `batch` (a multiple of 16), samples visited in a seeded order.  history[epochs] receives the
// mean training loss per epoch.
extern "C" mgpu_status mgpu_mlp_train(mgpu_mlp* m, const float* x, const float* t, int64_t n, int32_t epochs,
                                      int32_t batch, float lr, uint64_t seed, float* history) {
    MGPU_GUARD_BEGIN
    if (!m || !x || !t || n < 16 || epochs < 1 || batch < 16 || batch % 16) return set_error(MGPU_ERR_INVALID_ARG, "bad argument");
    float *dx = nullptr, *dt = nullptr;
    HIP_TRY(hipMalloc(&dx, n * kLqInputs * sizeof(float)));
    HIP_TRY(hipMalloc(&dt, n * kLqPoints * sizeof(float)));
    // shuffle once on the host (seeded Fisher-Yates) so a batch is a contiguous slice
    std::vector<int64_t> perm(n);
    for (int64_t i = 0; i < n; ++i) perm[i] = i;
    uint64_t c = 0;
    for (int64_t i = n - 1; i > 0; --i) {
        uint64_t z = seed + (++c) * 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        std::swap(perm[i], perm[(int64_t)(z % (uint64_t)(i + 1))]);
    }
    std::vector<float> xs((size_t)n * kLqInputs), ts((size_t)n * kLqPoints);
    for (int64_t i = 0; i < n; ++i) {
        std::memcpy(&xs[i * kLqInputs], x + perm[i] * kLqInputs, kLqInputs * sizeof(float));
        std::memcpy(&ts[i * kLqPoints], t + perm[i] * kLqPoints, kLqPoints * sizeof(float));
    }
    HIP_TRY(hipMemcpy(dx, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dt, ts.data(), ts.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(m->d.grads, 0, kLqParams * sizeof(float)));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int64_t nb = n / batch;
    for (int ep = 0; ep < epochs; ++ep) {
        HIP_TRY(hipMemset(m->d.loss, 0, sizeof(float)));
        for (int64_t b = 0; b < nb; ++b) {
            const float inv = 1.0f / (float)(batch * kLqPoints);
            hipLaunchKernelGGL(k_mlp_backward, dim3((unsigned)(batch / 16)), dim3(64), 0, 0, m->d,
                               dx + b * batch * kLqInputs, dt + b * batch * kLqPoints, (int64_t)batch, inv);
            ++m->steps;
            const float c1 = 1.0f - std::pow(b1, (float)m->steps), c2 = 1.0f - std::pow(b2, (float)m->steps);
            // cosine decay of the step size over the run
            const float frac = (float)(ep * nb + b) / (float)(epochs * nb);
            const float lr_t = lr * 0.5f * (1.0f + std::cos(3.14159265f * frac));
            hipLaunchKernelGGL(k_adam, dim3((kLqParams + 255) / 256), dim3(256), 0, 0, m->d, lr_t, b1, b2, eps, c1, c2);
            hipLaunchKernelGGL(k_mlp_convert, dim3((H * H + 255) / 256), dim3(256), 0, 0, m->d);
        }
        HIP_TRY(hipGetLastError());
        float l = 0.0f;
        HIP_TRY(hipMemcpy(&l, m->d.loss, sizeof(float), hipMemcpyDeviceToHost));
        if (history) history[ep] = l / (float)nb;
    }
    (void)hipFree(dx);
    (void)hipFree(dt);
    return MGPU_OK;
    MGPU_GUARD_END
}
