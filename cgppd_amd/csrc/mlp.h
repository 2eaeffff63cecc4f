// mlp.h -- the pair-energy surrogate on gfx950 MFMA (device code; included
// by cgppd_hip.hip).  No reference counterpart.
//
// Network: 80 features -> 128 -> 128 -> 1, SiLU on the hidden layers.
// Precision: inputs, weights and hidden activations are BF16 (round to
// nearest even) at every matrix product, products accumulate in FP32 on
// v_mfma_f32_16x16x32_bf16, biases and pre-activations are FP32; FP32
// master weights with Adam.  A workgroup of 4 waves handles a tile of 32
// rows with all activations in LDS.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlp {

constexpr int kFeat = 80;      // features
constexpr int kFeatPad = 96;   // K padded to a multiple of 32
constexpr int kHidden = 128;
constexpr int kRows = 32;      // rows per workgroup tile
constexpr int kThreads = 256;  // 4 waves
constexpr float kTargetClip = 15.0f;

// FP32 master layout: W1[k][n] (80 x 128), b1, W2[k][n] (128 x 128), b2, w3, b3
constexpr int kOW1 = 0;
constexpr int kOb1 = kOW1 + kFeat * kHidden;
constexpr int kOW2 = kOb1 + kHidden;
constexpr int kOb2 = kOW2 + kHidden * kHidden;
constexpr int kOw3 = kOb2 + kHidden;
constexpr int kOb3 = kOw3 + kHidden;
constexpr int kParams = kOb3 + 1;
// BF16 copies for the matrix products, transposed so that K is contiguous:
// W1T[n][k] (128 x 96, zero for k >= 80), W2T[n][k] (128 x 128), w3 (128)
constexpr int kBW1 = 0;
constexpr int kBW2 = kBW1 + kHidden * kFeatPad;
constexpr int kBw3 = kBW2 + kHidden * kHidden;
constexpr int kBf16Params = kBw3 + kHidden;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> BF16 bits, round to nearest even (NaN -> quiet NaN)
__host__ __device__ inline unsigned short f2bf(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf2f(unsigned short h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

__device__ inline float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ inline float silu_grad(float z)
{
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// One 16x16 output tile of C = A * B accumulated over k in [0, K) (K a
// multiple of 32): a(m, k) = A[m * am + k * ak], b(k, n) = B[k * bk + n * bn]
// (BF16 bits).  Lane l supplies row (l % 16) of A and column (l % 16) of B;
// the result lane l holds C[m0 + 4 (l / 16) + i][n0 + l % 16], i < 4.
__device__ inline f32x4 mma_tile(const unsigned short *A, int am, int ak, const unsigned short *B, int bk, int bn,
                                 int m0, int n0, int K, f32x4 acc)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 8;
    if (ak == 1 && bk == 1) {
        // K contiguous in both operands (row strides multiples of 8 elements,
        // 16-byte aligned bases): one 16-byte load per operand
        for (int k0 = 0; k0 < K; k0 += 32) {
            const u16x8 a = *reinterpret_cast<const u16x8 *>(A + (m0 + r) * am + k0 + kq);
            const u16x8 b = *reinterpret_cast<const u16x8 *>(B + (n0 + r) * bn + k0 + kq);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                          __builtin_bit_cast(bf16x8, b), acc, 0, 0, 0);
        }
        return acc;
    }
    for (int k0 = 0; k0 < K; k0 += 32) {
        u16x8 a, b;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            a[i] = A[(m0 + r) * am + (k0 + kq + i) * ak];
            b[i] = B[(k0 + kq + i) * bk + (n0 + r) * bn];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                      acc, 0, 0, 0);
    }
    return acc;
}

struct Tile {
    alignas(16) unsigned short x[kRows * kFeatPad];   // BF16 inputs
    alignas(16) float z1[kRows * kHidden];            // pre-activations (then dZ1)
    alignas(16) unsigned short a1[kRows * kHidden];   // BF16 activations (then BF16 dZ1)
    alignas(16) float z2[kRows * kHidden];            // (then dZ2)
    alignas(16) unsigned short a2[kRows * kHidden];   // (then BF16 dZ2)
    float y[kRows];
};

// Hidden layer: Z = A_in * W + b (FP32 into z), A_out = BF16(SiLU(Z)).
// a_in row-major (rows x K), wt = W transposed (n x K).
__device__ inline void dense_silu(const unsigned short *a_in, int K, const unsigned short *wt, const float *b,
                                  float *z, unsigned short *a_out)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // 2 row tiles x 8 column tiles; 4 per wave
    for (int t = wave; t < 16; t += 4) {
        const int m0 = (t >> 3) * 16, n0 = (t & 7) * 16;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mma_tile(a_in, K, 1, wt, 1, K, m0, n0, K, acc);
        const int col = n0 + (lane & 15);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = m0 + 4 * (lane >> 4) + i;
            const float v = acc[i] + b[col];
            z[row * kHidden + col] = v;
            a_out[row * kHidden + col] = f2bf(silu(v));
        }
    }
}

// Forward pass of the tile (x filled, rows < n valid): z1, a1, z2, a2, y.
__device__ inline void forward(Tile &T, const float *master, const unsigned short *wb, int n)
{
    dense_silu(T.x, kFeatPad, wb + kBW1, master + kOb1, T.z1, T.a1);
    __syncthreads();
    dense_silu(T.a1, kHidden, wb + kBW2, master + kOb2, T.z2, T.a2);
    __syncthreads();
    {
        // output layer: 8 lanes per row, 16 products each, then a butterfly
        const int row = threadIdx.x >> 3, part = threadIdx.x & 7;
        float s = 0.f;
        for (int k = part * 16; k < part * 16 + 16; ++k) s += bf2f(T.a2[row * kHidden + k]) * bf2f(wb[kBw3 + k]);
        s += __shfl_xor(s, 1, 64);
        s += __shfl_xor(s, 2, 64);
        s += __shfl_xor(s, 4, 64);
        if (part == 0) T.y[row] = row < n ? s + master[kOb3] : 0.f;
    }
    __syncthreads();
}

// Backward pass after forward(): dy[k] = (y - t) / batch for the valid rows,
// parameter gradients added (atomically) into grad (FP32, master layout).
__device__ inline void backward(Tile &T, const unsigned short *wb, const float *dy, float *grad)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    // output layer and dZ2 (FP32 in z2)
    if (tid < kHidden) {
        float gw = 0.f;
        for (int k = 0; k < kRows; ++k) gw += bf2f(T.a2[k * kHidden + tid]) * dy[k];
        atomicAdd(grad + kOw3 + tid, gw);
    }
    if (tid == 0) {
        float gb = 0.f;
        for (int k = 0; k < kRows; ++k) gb += dy[k];
        atomicAdd(grad + kOb3, gb);
    }
    __syncthreads();
    for (int e = tid; e < kRows * kHidden; e += kThreads) {
        const int k = e / kHidden, n = e % kHidden;
        const float d = dy[k] * bf2f(wb[kBw3 + n]) * silu_grad(T.z2[e]);
        T.z2[e] = d;
        T.a2[e] = f2bf(d);
    }
    __syncthreads();
    if (tid < kHidden) {
        float gb = 0.f;
        for (int k = 0; k < kRows; ++k) gb += T.z2[k * kHidden + tid];
        atomicAdd(grad + kOb2 + tid, gb);
    }
    // dW2[j][n] = sum_k A1[k][j] dZ2[k][n]: 8 x 8 tiles over K = 32 rows
    for (int t = wave; t < 64; t += 4) {
        const int m0 = (t >> 3) * 16, n0 = (t & 7) * 16;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mma_tile(T.a1, 1, kHidden, T.a2, kHidden, 1, m0, n0, kRows, acc);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            atomicAdd(grad + kOW2 + (m0 + 4 * (lane >> 4) + i) * kHidden + n0 + (lane & 15), acc[i]);
    }
    // dA1 = dZ2 W2^T, then dZ1 = dA1 * SiLU'(Z1) into z1 (FP32)
    for (int t = wave; t < 16; t += 4) {
        const int m0 = (t >> 3) * 16, n0 = (t & 7) * 16;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        // b(k = n_out, n = j) = W2[j][n_out] = W2T[n_out][j]
        acc = mma_tile(T.a2, kHidden, 1, wb + kBW2, kHidden, 1, m0, n0, kHidden, acc);
        const int col = n0 + (lane & 15);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = m0 + 4 * (lane >> 4) + i;
            T.z1[row * kHidden + col] = acc[i] * silu_grad(T.z1[row * kHidden + col]);
        }
    }
    __syncthreads();
    for (int e = tid; e < kRows * kHidden; e += kThreads) T.a1[e] = f2bf(T.z1[e]);
    if (tid < kHidden) {
        float gb = 0.f;
        for (int k = 0; k < kRows; ++k) gb += T.z1[k * kHidden + tid];
        atomicAdd(grad + kOb1 + tid, gb);
    }
    __syncthreads();
    // dW1[f][n] = sum_k X[k][f] dZ1[k][n]: 5 x 8 tiles (features < 80)
    for (int t = wave; t < 40; t += 4) {
        const int m0 = (t >> 3) * 16, n0 = (t & 7) * 16;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mma_tile(T.x, 1, kFeatPad, T.a1, kHidden, 1, m0, n0, kRows, acc);
#pragma unroll
        for (int i = 0; i < 4; ++i)
            atomicAdd(grad + kOW1 + (m0 + 4 * (lane >> 4) + i) * kHidden + n0 + (lane & 15), acc[i]);
    }
}

}  // namespace mlp
