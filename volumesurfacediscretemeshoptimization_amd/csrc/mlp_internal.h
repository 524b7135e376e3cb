// mlp_internal.h -- the learned vertex proposal (internal): a 96-128-128-3
// GELU MLP with BF16 weights and activations on MFMA and FP32 accumulation,
// its training set of one-ring / optimised-position pairs, and the one-ring
// feature encoding shared by training and inference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vsdmo_internal.h"

namespace vsdmo {

constexpr int kNbr = 32;            // neighbours per one-ring feature vector
constexpr int kIn = 3 * kNbr;       // 96 input features
constexpr int kHid = 128;
constexpr int kOut = 3;
// packed parameter layout (FP32 master copy): W1[128x96], b1, W2[128x128], b2, W3[3x128], b3
constexpr int kOW1 = 0;
constexpr int kOB1 = kOW1 + kHid * kIn;
constexpr int kOW2 = kOB1 + kHid;
constexpr int kOB2 = kOW2 + kHid * kHid;
constexpr int kOW3 = kOB2 + kHid;
constexpr int kOB3 = kOW3 + kOut * kHid;
constexpr int kNParams = kOB3 + kOut;
static_assert(kOW1 % 8 == 0 && kOW2 % 8 == 0 && kOW3 % 8 == 0, "weight rows load as 16-byte bf16x8");

// round-to-nearest-even FP32 -> BF16 bits
__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

// ------------------------------------------------------------- inference
// 16 rows (vertices) per block of 4 waves (256 threads), all in LDS.
constexpr int kMlpRows = 16;
constexpr int kLdx = kIn + 8;    // LDS row pitches (bf16), multiples of 8
constexpr int kLdh = kHid + 8;
struct MlpLds {
    __align__(16) uint16_t X[kMlpRows * kLdx];
    __align__(16) uint16_t H1[kMlpRows * kLdh];
    __align__(16) uint16_t H2[kMlpRows * kLdh];
    float rel[kMlpRows][kNbr][3];
    float dist[kMlpRows][kNbr];
    float L[kMlpRows];
    float Y[kMlpRows][kOut];
};

// features of row r = threadIdx.x >> 4 (vertex v; live = it exists): 16
// threads per row gather two table neighbours each; L = mean distance summed
// in neighbour order (the order of k_features); X = bf16((n_k - v) / L),
// zero padded.  All 256 threads call it; ends with a barrier.
__device__ __forceinline__ void mlp_features16(const float* __restrict__ x, const float* __restrict__ y,
                                               const float* __restrict__ z, const int* __restrict__ nbr, int v,
                                               bool live, MlpLds& sm)
{
    const int r = threadIdx.x >> 4, q = threadIdx.x & 15;
    if (live) {
        const int* nb = nbr + (size_t)v * kNbr;
        const float vx = x[v], vy = y[v], vz = z[v];
#pragma unroll
        for (int k = 2 * q; k < 2 * q + 2; ++k) {
            const int u = nb[k];
            if (u < 0) { sm.dist[r][k] = -1.0f; continue; }
            const float dx = x[u] - vx, dy = y[u] - vy, dz = z[u] - vz;
            sm.rel[r][k][0] = dx; sm.rel[r][k][1] = dy; sm.rel[r][k][2] = dz;
            sm.dist[r][k] = sqrtf((dx * dx + dy * dy) + dz * dz);
        }
    }
    // the table's valid entries are a prefix: cnt = their number, counted by
    // a ballot over the row's 16 threads (lanes 16 (r % 4) + q of the wave)
    const int* nbv = nbr + (size_t)v * kNbr + 2 * q;
    const bool v0 = live && nbv[0] >= 0, v1 = live && nbv[1] >= 0;
    const unsigned long long rowMask = 0xffffull << (16 * (r & 3));
    const int cnt = __popcll(__ballot(v0) & rowMask) + __popcll(__ballot(v1) & rowMask);
    __syncthreads();
    if (q == 0) {
        float sum = 0.0f, l = 1.0f;
        if (live) {
#pragma unroll
            for (int k = 0; k < kNbr; ++k)   // neighbour order, as k_features
                if (k < cnt) sum += sm.dist[r][k];
            l = cnt > 0 ? sum / (float)cnt : 1.0f;
            if (!(l > 0.0f)) l = 1.0f;
        }
        sm.L[r] = l;
    }
    __syncthreads();
    const float l = sm.L[r];
    uint16_t* o = sm.X + r * kLdx;
#pragma unroll
    for (int k = 2 * q; k < 2 * q + 2; ++k) {
        if (k < cnt) {
            o[3 * k] = f2bf(sm.rel[r][k][0] / l);
            o[3 * k + 1] = f2bf(sm.rel[r][k][1] / l);
            o[3 * k + 2] = f2bf(sm.rel[r][k][2] / l);
        } else {
            o[3 * k] = 0; o[3 * k + 1] = 0; o[3 * k + 2] = 0;
        }
    }
    __syncthreads();
}

// the three layers of the 16 rows of sm.X into sm.Y.  Layers 1-2: wave w
// computes columns [32w, 32w + 32) -- two 16x16 tiles of
// mfma_f32_16x16x32_bf16 (A: 16 LDS rows, lane l holds row l%16, k = 8(l/16)
// + 0..7; B: 16-byte weight rows from global memory; D: column l%16, rows
// 4(l/16) + 0..3), then bias + GELU to BF16 in LDS.  Layer 3 (128 -> 3):
// one tile on wave 0 with W3 zero padded to 16 columns.  All 256 threads
// call it; ends with a barrier.
__device__ __forceinline__ void mlp_layers16(const uint16_t* __restrict__ wb, const float* __restrict__ w, MlpLds& sm)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    {
        f32x4 acc[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) acc[t] = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int k0 = 0; k0 < kIn; k0 += 32) {
            const bf16x8 av = *reinterpret_cast<const bf16x8*>(sm.X + r16 * kLdx + k0 + kq);
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int col = (2 * wave + t) * 16 + r16;
                const bf16x8 bv = *reinterpret_cast<const bf16x8*>(wb + kOW1 + (size_t)col * kIn + k0 + kq);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[t], 0, 0, 0);
            }
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int n = (2 * wave + t) * 16 + r16;
            const float b = w[kOB1 + n];
#pragma unroll
            for (int r = 0; r < 4; ++r) sm.H1[((lane >> 4) * 4 + r) * kLdh + n] = f2bf(gelu(acc[t][r] + b));
        }
    }
    __syncthreads();
    {
        f32x4 acc[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) acc[t] = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int k0 = 0; k0 < kHid; k0 += 32) {
            const bf16x8 av = *reinterpret_cast<const bf16x8*>(sm.H1 + r16 * kLdh + k0 + kq);
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int col = (2 * wave + t) * 16 + r16;
                const bf16x8 bv = *reinterpret_cast<const bf16x8*>(wb + kOW2 + (size_t)col * kHid + k0 + kq);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[t], 0, 0, 0);
            }
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int n = (2 * wave + t) * 16 + r16;
            const float b = w[kOB2 + n];
#pragma unroll
            for (int r = 0; r < 4; ++r) sm.H2[((lane >> 4) * 4 + r) * kLdh + n] = f2bf(gelu(acc[t][r] + b));
        }
    }
    __syncthreads();
    if (wave == 0) {
        // layer 3 on one 16x16 tile: columns 0..2 are W3's rows, the rest zero
        typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
        f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int k0 = 0; k0 < kHid; k0 += 32) {
            const bf16x8 av = *reinterpret_cast<const bf16x8*>(sm.H2 + r16 * kLdh + k0 + kq);
            bf16x8 bv = __builtin_bit_cast(bf16x8, (u32x4){ 0u, 0u, 0u, 0u });
            if (r16 < kOut) bv = *reinterpret_cast<const bf16x8*>(wb + kOW3 + (size_t)r16 * kHid + k0 + kq);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
        }
        if (r16 < kOut) {
            const float b = w[kOB3 + r16];
#pragma unroll
            for (int r = 0; r < 4; ++r) sm.Y[(lane >> 4) * 4 + r][r16] = acc[r] + b;
        }
    }
    __syncthreads();
}

// canonical neighbour table of the inner vertices: the distinct vertices of
// the vertex's volume one-ring in order of first appearance (ring elements in
// their stored order), at most kNbr, -1 padded; rows of surface vertices are
// all -1
void build_neighbour_table(vsdmo_mesh& m);

// features of vertices [lo, hi): X (bf16, 96 per vertex) = (n_i - v) / L for
// the table's neighbours, zero padded; L = mean |n_i - v|; v0 = start position
void launch_features(const float* x, const float* y, const float* z, const int* nbr, int lo, int hi,
                     uint16_t* X, float* L, float* v0, hipStream_t s);
// targets of the same vertices after their search: T = (p - v0) / L
void launch_targets(const float* x, const float* y, const float* z, int lo, int hi, const float* L,
                    const float* v0, float* T, hipStream_t s);

}  // namespace vsdmo

struct vsdmo_dataset {
    long long cap = 0, n = 0;
    vsdmo::DevBuf<uint16_t> X;   // cap x 96 bf16
    vsdmo::DevBuf<float> T;      // cap x 3
    vsdmo::DevBuf<float> L, v0;  // per-class scratch
    hipStream_t stream = nullptr;
};

struct vsdmo_mlp {
    hipStream_t stream = nullptr;
    vsdmo::DevBuf<float> w, g, m1, m2;   // packed parameters, gradients, Adam moments
    vsdmo::DevBuf<uint16_t> wb;          // bf16 copies of the packed parameters
    int step = 0;
    int cap = 0;                         // batch capacity of the work buffers
    int n_users = 0;                     // optimisers and batches with this network as their proposal
    vsdmo::DevBuf<uint16_t> X, H1, H2, dY16, dZ1, dZ2;
    vsdmo::DevBuf<float> Z1, Z2, Y, T, part;
    vsdmo::DevBuf<double> lossd;
    vsdmo::DevBuf<int> perm;
};

namespace vsdmo {
// forward of n rows of bf16 features X (device) into Y (device, n x 3)
void mlp_forward_dev(vsdmo_mlp* m, const uint16_t* X, int n, float* Y, hipStream_t s);
void mlp_reserve(vsdmo_mlp* m, int n);
// fused inference (features + 3 layers) of vertices [lo, hi) or of rows of
// Xin (n = hi - lo); Y (n x 3) and, when computing features, L (n)
void launch_mlp_infer(vsdmo_mlp* m, const float* x, const float* y, const float* z, const int* nbr, int lo, int hi,
                      const uint16_t* Xin, float* Y, float* L, hipStream_t s);
}  // namespace vsdmo
