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
// ring-local corner positions (vsdmo_mesh::ringLoc): bytes per ring element
// (3 tet / 7 hex corners, padded for aligned 4- / 8-byte loads) and the mark
// of a corner outside the table
constexpr int kRingLocStrideTet = 4, kRingLocStrideHex = 8;
constexpr uint8_t kRingLocNone = 0xff;
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
    float dist[kMlpRows][kNbr];
    float L[kMlpRows];
    float Y[kMlpRows][kOut];
};

// The two table neighbours 2q, 2q+1 (q = threadIdx.x & 15) of row vertex v
// and their positions, gathered into registers; a caller issues this first
// so that its latency overlaps the caller's other loads.
struct MlpGather {
    int u[2];
    float x[2], y[2], z[2];
};
__device__ __forceinline__ MlpGather mlp_gather16(const float* __restrict__ x, const float* __restrict__ y,
                                                  const float* __restrict__ z, const int* __restrict__ nbr, int v,
                                                  bool live)
{
    const int q = threadIdx.x & 15;
    MlpGather g;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        g.u[k] = live ? nbr[(size_t)v * kNbr + 2 * q + k] : -1;
        g.x[k] = 0.0f; g.y[k] = 0.0f; g.z[k] = 0.0f;
        if (g.u[k] >= 0) { g.x[k] = x[g.u[k]]; g.y[k] = y[g.u[k]]; g.z[k] = z[g.u[k]]; }
    }
    return g;
}

// features of row r = threadIdx.x >> 4 (vertex at (vx, vy, vz); live = it
// exists) from the row's gathered neighbours: L = mean distance summed in
// neighbour order (the order of k_features); X = bf16((n_k - v) / L), zero
// padded.  relOut (optional, LDS): n_k - v of every table entry, kNbr x 3
// floats per row (entries past the row's valid prefix hold no position).
// All 256 threads call it; ends with a barrier.
__device__ __forceinline__ void mlp_features16_from(const MlpGather& g, float vx, float vy, float vz, bool live,
                                                    MlpLds& sm, float* relOut = nullptr)
{
    const int r = threadIdx.x >> 4, q = threadIdx.x & 15;
    float rel[2][3];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        rel[k][0] = g.x[k] - vx; rel[k][1] = g.y[k] - vy; rel[k][2] = g.z[k] - vz;
        sm.dist[r][2 * q + k] = sqrtf((rel[k][0] * rel[k][0] + rel[k][1] * rel[k][1]) + rel[k][2] * rel[k][2]);
        if (relOut) {
            float* o = relOut + ((size_t)r * kNbr + 2 * q + k) * 3;
            o[0] = rel[k][0]; o[1] = rel[k][1]; o[2] = rel[k][2];
        }
    }
    // the table's valid entries are a prefix: cnt = their number, counted by
    // a ballot over the row's 16 threads (lanes 16 (r % 4) + q of the wave)
    const bool v0 = live && g.u[0] >= 0, v1 = live && g.u[1] >= 0;
    const unsigned long long rowMask = 0xffffull << (16 * (r & 3));
    const int cnt = __popcll(__ballot(v0) & rowMask) + __popcll(__ballot(v1) & rowMask);
    __syncthreads();
    // every thread of the row sums the row's distances in neighbour order
    // (all 32 read first, then added: no read waits inside the sum)
    float l = 1.0f;
    if (live) {
        float d[kNbr];
#pragma unroll
        for (int k = 0; k < kNbr; ++k) d[k] = sm.dist[r][k];
        float sum = 0.0f;
#pragma unroll
        for (int k = 0; k < kNbr; ++k) sum = (k < cnt) ? sum + d[k] : sum;
        l = cnt > 0 ? sum / (float)cnt : 1.0f;
        if (!(l > 0.0f)) l = 1.0f;
    }
    if (q == 0) sm.L[r] = l;
    uint16_t* o = sm.X + r * kLdx;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int kk = 2 * q + k;
        if (kk < cnt) {
            o[3 * kk] = f2bf(rel[k][0] / l);
            o[3 * kk + 1] = f2bf(rel[k][1] / l);
            o[3 * kk + 2] = f2bf(rel[k][2] / l);
        } else {
            o[3 * kk] = 0; o[3 * kk + 1] = 0; o[3 * kk + 2] = 0;
        }
    }
    __syncthreads();
}

// gather + features of row r = threadIdx.x >> 4 (vertex v)
__device__ __forceinline__ void mlp_features16(const float* __restrict__ x, const float* __restrict__ y,
                                               const float* __restrict__ z, const int* __restrict__ nbr, int v,
                                               bool live, MlpLds& sm)
{
    const MlpGather g = mlp_gather16(x, y, z, nbr, v, live);
    float vx = 0.0f, vy = 0.0f, vz = 0.0f;
    if (live) { vx = x[v]; vy = y[v]; vz = z[v]; }
    mlp_features16_from(g, vx, vy, vz, live, sm);
}

// The layer-1 weight operands that wave w = threadIdx.x / 64 reads (its
// column tiles 2w, 2w+1; lane l: row (2w+t)*16 + l%16, k = 8(l/16) + 32i +
// 0..7) and the wave's layer 1-2 biases, loaded into registers; a caller
// issues this first so that the loads overlap its other work (layer 2's
// operands are loaded while layer 1 runs, layer 3's while layer 2 runs).
struct MlpW {
    bf16x8 b1[kIn / 32][2];
    float bias1[2], bias2[2];
};
__device__ __forceinline__ MlpW mlp_load_w16(const uint16_t* __restrict__ wb, const float* __restrict__ w)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    MlpW m;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int col = (2 * wave + t) * 16 + r16;
#pragma unroll
        for (int i = 0; i < kIn / 32; ++i)
            m.b1[i][t] = *reinterpret_cast<const bf16x8*>(wb + kOW1 + (size_t)col * kIn + 32 * i + kq);
        m.bias1[t] = w[kOB1 + col];
        m.bias2[t] = w[kOB2 + col];
    }
    return m;
}

// the three layers of the 16 rows of sm.X into sm.Y, layer 1 with the
// operands of mlp_load_w16.  Layers 1-2: wave w computes columns [32w, 32w +
// 32) -- two 16x16 tiles of mfma_f32_16x16x32_bf16 (A: 16 LDS rows, lane l
// holds row l%16, k = 8(l/16) + 0..7; B: the wave's weight rows; D: column
// l%16, rows 4(l/16) + 0..3), then bias + GELU to BF16 in LDS.  Layer 3 (128
// -> 3): one tile on wave 0 with W3 zero padded to 16 columns.  Each layer's
// weight operands are loaded while the layer before it runs.  All 256
// threads call it; ends with a barrier.
__device__ __forceinline__ void mlp_layers16_w(const MlpW& m, const uint16_t* __restrict__ wb,
                                               const float* __restrict__ w, MlpLds& sm)
{
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    // layer 2's operands, loaded while layer 1 runs
    bf16x8 b2[kHid / 32][2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int col = (2 * wave + t) * 16 + r16;
#pragma unroll
        for (int i = 0; i < kHid / 32; ++i)
            b2[i][t] = *reinterpret_cast<const bf16x8*>(wb + kOW2 + (size_t)col * kHid + 32 * i + kq);
    }
    {
        f32x4 acc[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) acc[t] = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int i = 0; i < kIn / 32; ++i) {
            const bf16x8 av = *reinterpret_cast<const bf16x8*>(sm.X + r16 * kLdx + 32 * i + kq);
#pragma unroll
            for (int t = 0; t < 2; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, m.b1[i][t], acc[t], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int n = (2 * wave + t) * 16 + r16;
#pragma unroll
            for (int r = 0; r < 4; ++r) sm.H1[((lane >> 4) * 4 + r) * kLdh + n] = f2bf(gelu(acc[t][r] + m.bias1[t]));
        }
    }
    __syncthreads();
    // layer 3's operands (wave 0, columns 0..2), loaded while layer 2 runs
    bf16x8 b3[kHid / 32];
    float bias3 = 0.0f;
#pragma unroll
    for (int i = 0; i < kHid / 32; ++i) b3[i] = __builtin_bit_cast(bf16x8, (u32x4){ 0u, 0u, 0u, 0u });
    if (wave == 0 && r16 < kOut) {
#pragma unroll
        for (int i = 0; i < kHid / 32; ++i)
            b3[i] = *reinterpret_cast<const bf16x8*>(wb + kOW3 + (size_t)r16 * kHid + 32 * i + kq);
        bias3 = w[kOB3 + r16];
    }
    {
        f32x4 acc[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) acc[t] = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int i = 0; i < kHid / 32; ++i) {
            const bf16x8 av = *reinterpret_cast<const bf16x8*>(sm.H1 + r16 * kLdh + 32 * i + kq);
#pragma unroll
            for (int t = 0; t < 2; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b2[i][t], acc[t], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int n = (2 * wave + t) * 16 + r16;
#pragma unroll
            for (int r = 0; r < 4; ++r) sm.H2[((lane >> 4) * 4 + r) * kLdh + n] = f2bf(gelu(acc[t][r] + m.bias2[t]));
        }
    }
    __syncthreads();
    if (wave == 0) {
        f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int i = 0; i < kHid / 32; ++i) {
            const bf16x8 av = *reinterpret_cast<const bf16x8*>(sm.H2 + r16 * kLdh + 32 * i + kq);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b3[i], acc, 0, 0, 0);
        }
        if (r16 < kOut) {
#pragma unroll
            for (int r = 0; r < 4; ++r) sm.Y[(lane >> 4) * 4 + r][r16] = acc[r] + bias3;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void mlp_layers16(const uint16_t* __restrict__ wb, const float* __restrict__ w, MlpLds& sm)
{
    mlp_layers16_w(mlp_load_w16(wb, w), wb, w, sm);
}

// canonical neighbour table of the inner vertices: the distinct vertices of
// the vertex's volume one-ring in order of first appearance (ring elements in
// their stored order), at most kNbr, -1 padded; rows of surface vertices are
// all -1.  Also m.ringLoc: for every corner of a one-ring element, its
// position in the vertex's table row (kRingLocNone when the row is full
// without it, and for surface vertices)
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
