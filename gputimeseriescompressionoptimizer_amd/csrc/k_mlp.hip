// Planner network: 32 block features -> 128 -> 128 -> 48 outputs (the
// predicted log2 compressed size of each candidate tree), ReLU on the hidden
// layers.  Every matrix product runs on the matrix cores as
// v_mfma_f32_32x32x8_bf16: BF16 operands, FP32 accumulation.  Master weights,
// gradients and the Adam state stay in FP32.
//
// A wave owns a tile of 32 rows (blocks) and runs a whole pass for it out
// of LDS.  Register layout of the 32x32x8 instruction, lane l of 64:
//   A (32 x 8, M x K): lane holds A[l % 32][4 (l / 32) + j], j = 0..3
//   B (8 x 32, K x N): lane holds B[4 (l / 32) + j][l % 32]
//   D (32 x 32):       register r holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
// so an A operand is 4 consecutive K values of one row and a B operand 4
// consecutive K values of one column: LDS keeps every B matrix transposed
// (one row per output column) and every row padded by 4 elements, which
// spreads the 32 rows of a fetch over the banks.
#include "k_common.h"
#include "kernels.h"

namespace ddj {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kF = kMlpFeatures, kH = kMlpHidden, kO = kMlpOutputs, kOP = kMlpOutputsPadded;
constexpr int LF = kF + 4, LH = kH + 4, LO = kOP + 4, LR = 32 + 4;

// Round to nearest even; NaN stays a (quiet) NaN.
__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ int d_row(int r, int l) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }

// acc += A[32 x K] * B[K x 32]; A row-major (row stride lda), B given as BT
// (row n = column n of B, stride ldb).  K is a multiple of 8.
__device__ __forceinline__ f32x16 mma(const uint16_t* A, int lda, const uint16_t* BT, int ldb, int K, f32x16 acc) {
    const int l = threadIdx.x & 63;
    const int i = l & 31, kq = 4 * (l >> 5);
    for (int k0 = 0; k0 < K; k0 += 8) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(A + i * lda + k0 + kq);
        const bf16x4 b = *reinterpret_cast<const bf16x4*>(BT + i * ldb + k0 + kq);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// Loads W [K][N] (FP32, row-major) as BT [N][K] BF16 into LDS.
__device__ __forceinline__ void load_wt(const float* __restrict__ W, int K, int N, uint16_t* wt, int ldw) {
    for (int e = threadIdx.x; e < K * N; e += 64) wt[(e % N) * ldw + e / N] = to_bf16(W[e]);
}
// Loads W [K][N] (FP32) as W [K][N] BF16 (row k = B column k of W^T).
__device__ __forceinline__ void load_w(const float* __restrict__ W, int K, int N, uint16_t* w, int ldw) {
    for (int e = threadIdx.x; e < K * N; e += 64) w[(e / N) * ldw + e % N] = to_bf16(W[e]);
}

// One tile of a forward layer into LDS: bias, optional ReLU, BF16.
__device__ __forceinline__ void put_act(const f32x16& acc, uint16_t* Hs, int ldh, int col0, const float* bias) {
    const int l = threadIdx.x & 63, c = col0 + (l & 31);
    const float bb = bias[c];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        float v = acc[r] + bb;
        v = v > 0.f ? v : 0.f;
        Hs[d_row(r, l) * ldh + c] = to_bf16(v);
    }
}

}  // namespace

// Forward pass over B rows: Y [B][48] FP32; the hidden activations (BF16)
// are kept for the backward pass when H1s / H2s are given.
__global__ __launch_bounds__(64) void k_mlp_forward(const float* __restrict__ X, int B, MlpParams p,
                                                    float* __restrict__ Y, uint16_t* __restrict__ H1s,
                                                    uint16_t* __restrict__ H2s) {
    __shared__ __attribute__((aligned(16))) uint16_t w1t[kH * LF];
    __shared__ __attribute__((aligned(16))) uint16_t w2t[kH * LH];
    __shared__ __attribute__((aligned(16))) uint16_t w3t[kOP * LH];
    __shared__ __attribute__((aligned(16))) uint16_t xs[32 * LF];
    __shared__ __attribute__((aligned(16))) uint16_t h1[32 * LH];
    __shared__ __attribute__((aligned(16))) uint16_t h2[32 * LH];
    const int l = threadIdx.x;
    load_wt(p.W1, kF, kH, w1t, LF);
    load_wt(p.W2, kH, kH, w2t, LH);
    load_wt(p.W3, kH, kOP, w3t, LH);
    const int r0 = blockIdx.x * 32;
    for (int e = l; e < 32 * kF; e += 64) {
        const int i = e / kF, k = e % kF;
        xs[i * LF + k] = (r0 + i < B) ? to_bf16(X[(size_t)(r0 + i) * kF + k]) : (uint16_t)0;
    }
    __syncthreads();
    for (int nt = 0; nt < kH / 32; nt++) put_act(mma(xs, LF, w1t + nt * 32 * LF, LF, kF, f32x16{}), h1, LH, nt * 32, p.b1);
    __syncthreads();
    for (int nt = 0; nt < kH / 32; nt++) put_act(mma(h1, LH, w2t + nt * 32 * LH, LH, kH, f32x16{}), h2, LH, nt * 32, p.b2);
    __syncthreads();
    for (int nt = 0; nt < kOP / 32; nt++) {
        const f32x16 acc = mma(h2, LH, w3t + nt * 32 * LH, LH, kH, f32x16{});
        const int c = nt * 32 + (l & 31);
        if (c < kO) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int row = r0 + d_row(r, l);
                if (row < B) Y[(size_t)row * kO + c] = acc[r] + p.b3[c];
            }
        }
    }
    if (H1s && H2s) {
        for (int e = l; e < 32 * kH; e += 64) {
            const int i = e / kH, k = e % kH;
            if (r0 + i < B) {
                H1s[(size_t)(r0 + i) * kH + k] = h1[i * LH + k];
                H2s[(size_t)(r0 + i) * kH + k] = h2[i * LH + k];
            }
        }
    }
}

namespace {

// Adds a 32x32 FP32 tile to G [..][ldg] at (m0, n0) with atomics.
__device__ __forceinline__ void add_tile(const f32x16& acc, float* G, int ldg, int m0, int n0, int n_lim) {
    const int l = threadIdx.x & 63, c = n0 + (l & 31);
    if (c >= n_lim) return;
#pragma unroll
    for (int r = 0; r < 16; r++) atomicAdd(&G[(size_t)(m0 + d_row(r, l)) * ldg + c], acc[r]);
}

// Column sums of a masked 32-row tile (bias gradients): lanes l and l + 32
// hold the same column.
__device__ __forceinline__ void add_colsum(float s, float* g, int c) {
    s += __shfl_xor(s, 32, 64);
    if ((threadIdx.x & 63) < 32) atomicAdd(&g[c], s);
}

}  // namespace

// Backward pass of the mean squared error over B rows and 48 outputs:
// dY = 2 (Y - T) / (48 B).  Gradients are added (FP32 atomics) into g.
__global__ __launch_bounds__(64) void k_mlp_backward(const float* __restrict__ X, const uint16_t* __restrict__ H1s,
                                                     const uint16_t* __restrict__ H2s, const float* __restrict__ Y,
                                                     const float* __restrict__ T, int B, MlpParams p, MlpParams g) {
    __shared__ __attribute__((aligned(16))) uint16_t xT[kF * LR];
    __shared__ __attribute__((aligned(16))) uint16_t h1T[kH * LR];
    __shared__ __attribute__((aligned(16))) uint16_t h2T[kH * LR];
    __shared__ __attribute__((aligned(16))) uint16_t dy[32 * LO];
    __shared__ __attribute__((aligned(16))) uint16_t dyT[kOP * LR];
    __shared__ __attribute__((aligned(16))) uint16_t w3[kH * LO];
    __shared__ __attribute__((aligned(16))) uint16_t w2[kH * LH];
    __shared__ __attribute__((aligned(16))) uint16_t dh2[32 * LH];
    __shared__ __attribute__((aligned(16))) uint16_t dh2T[kH * LR];
    __shared__ __attribute__((aligned(16))) uint16_t dh1T[kH * LR];
    const int l = threadIdx.x;
    const int r0 = blockIdx.x * 32;
    const float scale = 2.0f / ((float)kO * (float)B);
    load_w(p.W3, kH, kOP, w3, LO);
    load_w(p.W2, kH, kH, w2, LH);
    for (int e = l; e < 32 * kF; e += 64) {
        const int i = e / kF, k = e % kF;
        xT[k * LR + i] = (r0 + i < B) ? to_bf16(X[(size_t)(r0 + i) * kF + k]) : (uint16_t)0;
    }
    for (int e = l; e < 32 * kH; e += 64) {
        const int i = e / kH, k = e % kH;
        const bool in = r0 + i < B;
        h1T[k * LR + i] = in ? H1s[(size_t)(r0 + i) * kH + k] : (uint16_t)0;
        h2T[k * LR + i] = in ? H2s[(size_t)(r0 + i) * kH + k] : (uint16_t)0;
    }
    // dY (FP32 for the bias, BF16 for the products); rows past B and the
    // padding columns are zero
    for (int c = l; c < kOP; c += 64) {
        float s = 0.f;
        for (int i = 0; i < 32; i++) {
            float d = 0.f;
            if (c < kO && r0 + i < B) d = scale * (Y[(size_t)(r0 + i) * kO + c] - T[(size_t)(r0 + i) * kO + c]);
            s += d;
            const uint16_t hb = to_bf16(d);
            dy[i * LO + c] = hb;
            dyT[c * LR + i] = hb;
        }
        if (c < kO) atomicAdd(&g.b3[c], s);
    }
    __syncthreads();
    // dW3 [128][64] += H2^T dY
    for (int mt = 0; mt < kH / 32; mt++)
        for (int nt = 0; nt < kOP / 32; nt++)
            add_tile(mma(h2T + mt * 32 * LR, LR, dyT + nt * 32 * LR, LR, 32, f32x16{}), g.W3, kOP, mt * 32, nt * 32, kO);
    // dH2 = dY W3^T, masked by H2 > 0
    for (int nt = 0; nt < kH / 32; nt++) {
        const f32x16 acc = mma(dy, LO, w3 + nt * 32 * LO, LO, kOP, f32x16{});
        const int c = nt * 32 + (l & 31);
        float s = 0.f;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int i = d_row(r, l);
            const float v = h2T[c * LR + i] != 0 ? acc[r] : 0.f;  // ReLU output > 0 <=> input > 0
            s += v;
            const uint16_t hb = to_bf16(v);
            dh2[i * LH + c] = hb;
            dh2T[c * LR + i] = hb;
        }
        add_colsum(s, g.b2, c);
    }
    __syncthreads();
    // dW2 [128][128] += H1^T dH2
    for (int mt = 0; mt < kH / 32; mt++)
        for (int nt = 0; nt < kH / 32; nt++)
            add_tile(mma(h1T + mt * 32 * LR, LR, dh2T + nt * 32 * LR, LR, 32, f32x16{}), g.W2, kH, mt * 32, nt * 32, kH);
    // dH1 = dH2 W2^T, masked by H1 > 0
    for (int nt = 0; nt < kH / 32; nt++) {
        const f32x16 acc = mma(dh2, LH, w2 + nt * 32 * LH, LH, kH, f32x16{});
        const int c = nt * 32 + (l & 31);
        float s = 0.f;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int i = d_row(r, l);
            const float v = h1T[c * LR + i] != 0 ? acc[r] : 0.f;
            s += v;
            dh1T[c * LR + i] = to_bf16(v);
        }
        add_colsum(s, g.b1, c);
    }
    __syncthreads();
    // dW1 [32][128] += X^T dH1
    for (int nt = 0; nt < kH / 32; nt++) add_tile(mma(xT, LR, dh1T + nt * 32 * LR, LR, 32, f32x16{}), g.W1, kH, 0, nt * 32, kH);
}

// Adam (Kingma & Ba 2015) on one FP32 tensor; clears the gradient.
__global__ void k_adam(float* __restrict__ w, float* __restrict__ grad, float* __restrict__ m, float* __restrict__ v,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gr = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * gr;
    const float vi = b2 * v[i] + (1.f - b2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    grad[i] = 0.f;
}

void launch_mlp_forward(Ctx* c, const float* X, int B, const MlpParams& p, float* Y, uint16_t* H1s, uint16_t* H2s) {
    if (B <= 0) return;
    { KernelTimer kt_(c, "k_mlp_forward"); hipLaunchKernelGGL(k_mlp_forward, dim3((B + 31) / 32), dim3(64), 0, c->stream, X, B, p, Y, H1s, H2s); }
    HIPC(hipGetLastError());
}

void launch_mlp_backward(Ctx* c, const float* X, const uint16_t* H1s, const uint16_t* H2s, const float* Y,
                         const float* T, int B, const MlpParams& p, const MlpParams& g) {
    if (B <= 0) return;
    { KernelTimer kt_(c, "k_mlp_backward"); hipLaunchKernelGGL(k_mlp_backward, dim3((B + 31) / 32), dim3(64), 0, c->stream, X, H1s, H2s, Y, T, B, p, g); }
    HIPC(hipGetLastError());
}

void launch_adam(Ctx* c, float* w, float* grad, float* m, float* v, int n, float lr, int step) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    { KernelTimer kt_(c, "k_adam"); hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, c->stream, w, grad, m, v, n, lr, b1, b2, eps, c1, c2); }
    HIPC(hipGetLastError());
}

}  // namespace ddj
