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
