// mlp.hip -- the learned vertex proposal: BF16 MFMA GEMMs with fused
// epilogues, the 96-128-128-3 GELU MLP (forward, backward, Adam), its
// training set and the one-ring feature encoding.
//
// There is no learned component in the reference.  The network predicts the
// displacement DMO gives a vertex (dmo/DMOTetImplCub.h:20-137,
// DMOHexImplCub.h:20-137) from its one-ring; the optimiser then searches a
// small window around the prediction and falls back to the full search.
//
// GEMMs: one wave per 16x16 output tile, __builtin_amdgcn_mfma_f32_16x16x32_bf16
// over K in chunks of 32 (lane l: A row l%16 and B column l%16, k = 8*(l/16)
// + 0..7; D column l%16, rows 4*(l/16) + 0..3), operands read through
// arbitrary (row, column) strides so that the transposed products of the
// backward pass need no copies.
#include <float.h>
#include <math.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "dmo_kernels.h"
#include "mlp_internal.h"

namespace vsdmo {

__device__ __forceinline__ float gelu_d(float z)
{
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum { EPI_F32 = 0, EPI_GELU = 1, EPI_DGELU = 2 };

struct GemmArgs {
    int M, N, K;
    const uint16_t* A;
    long long sam, sak;      // A[m][k] = A[m*sam + k*sak]
    const uint16_t* B;
    long long sbk, sbn;      // B[k][n] = B[k*sbk + n*sbn]
    const float* bias;       // [N] or null
    int mode;
    float* Cf;               // EPI_F32: acc (+bias); EPI_GELU: Z = acc + bias
    long long ldcf;
    uint16_t* Cb;            // EPI_GELU: bf16(gelu(Z)); EPI_DGELU: bf16(acc * gelu'(Zin))
    long long ldcb;
    const float* Zin;
    long long ldz;
};

__global__ __launch_bounds__(256) void k_gemm(GemmArgs g)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tilesN = (g.N + 15) >> 4, tilesM = (g.M + 15) >> 4;
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= tilesM * tilesN) return;   // whole wave; the kernel has no block barrier
    const int mt = tile / tilesN, nt = tile - mt * tilesN;
    const int row = mt * 16 + (lane & 15);
    const int col = nt * 16 + (lane & 15);
    const int kq = (lane >> 4) * 8;
    f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
    for (int k0 = 0; k0 < g.K; k0 += 32) {
        bf16x8 a, b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = k0 + kq + j;
            uint16_t av = (row < g.M && k < g.K) ? g.A[(long long)row * g.sam + (long long)k * g.sak] : (uint16_t)0;
            uint16_t bv = (col < g.N && k < g.K) ? g.B[(long long)k * g.sbk + (long long)col * g.sbn] : (uint16_t)0;
            a[j] = __builtin_bit_cast(__bf16, av);
            b[j] = __builtin_bit_cast(__bf16, bv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    const int n = nt * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int m = mt * 16 + 4 * (lane >> 4) + r;
        if (m >= g.M || n >= g.N) continue;
        float v = acc[r];
        if (g.mode == EPI_F32) {
            if (g.bias) v += g.bias[n];
            g.Cf[(long long)m * g.ldcf + n] = v;
        } else if (g.mode == EPI_GELU) {
            v += g.bias[n];
            g.Cf[(long long)m * g.ldcf + n] = v;
            g.Cb[(long long)m * g.ldcb + n] = f2bf(gelu(v));
        } else {
            g.Cb[(long long)m * g.ldcb + n] = f2bf(v * gelu_d(g.Zin[(long long)m * g.ldz + n]));
        }
    }
}

static void gemm(const GemmArgs& g, hipStream_t s)
{
    const int tiles = ((g.M + 15) / 16) * ((g.N + 15) / 16);
    if (tiles <= 0 || g.K <= 0) return;
    k_gemm<<<(tiles + 3) / 4, 256, 0, s>>>(g);
}

// mean squared error over n rows of 3: loss partial per block (FP64, fixed
// tree), dY = 2 (Y - T) / n in FP32 and BF16
__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ T, int n, float* __restrict__ dYf,
                      uint16_t* __restrict__ dY16, double* __restrict__ part)
{
    __shared__ double s[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double l = 0.0;
    if (i < n) {
        for (int c = 0; c < 3; ++c) {
            float d = Y[3 * i + c] - T[3 * i + c];
            l += (double)d * (double)d;
            float gy = 2.0f * d / (float)n;
            dYf[3 * i + c] = gy;
            dY16[3 * i + c] = f2bf(gy);
        }
    }
    s[threadIdx.x] = l;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// loss = sum(part) / n, added to *acc (one thread, fixed order)
__global__ void k_loss_final(const double* __restrict__ part, int nb, int n, double* __restrict__ acc)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < nb; ++i) s += part[i];
    *acc += s / (double)n;
}

// column sums (bias gradients): one block of 256 per column, strided FP64
// partial sums then a fixed tree, so the order depends only on `rows`
template <class T>
__device__ __forceinline__ double as_f64(T v);
template <>
__device__ __forceinline__ double as_f64<float>(float v) { return (double)v; }
template <>
__device__ __forceinline__ double as_f64<uint16_t>(uint16_t v) { return (double)bf2f(v); }

template <class T>
__global__ __launch_bounds__(256) void k_colsum(const T* __restrict__ a, int rows, int cols, float* __restrict__ out)
{
    __shared__ double s[256];
    const int c = blockIdx.x;
    double acc = 0.0;
    for (int r = threadIdx.x; r < rows; r += 256) acc += as_f64<T>(a[(long long)r * cols + c]);
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = (float)s[0];
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, uint16_t* __restrict__ wb, int n, float lr, float c1, float c2)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float gi = g[i];
    float a = b1 * m1[i] + (1.0f - b1) * gi;
    float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    float mh = a / c1, vh = b / c2;
    float wi = w[i] - lr * mh / (sqrtf(vh) + eps);
    w[i] = wi;
    wb[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float* __restrict__ a, long long n, uint16_t* __restrict__ b)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}

// batch gather through a permutation
__global__ void k_gather_rows(const uint16_t* __restrict__ X, const float* __restrict__ T, const int* __restrict__ perm,
                              int n, uint16_t* __restrict__ Xo, float* __restrict__ To)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * kIn) return;
    int r = i / kIn, c = i - r * kIn;
    long long src = perm[r];
    Xo[i] = X[src * kIn + c];
    if (c < 3) To[3 * r + c] = T[src * 3 + c];
}

// one-ring features of vertex v into o[0..95] (bf16); returns L
__device__ __forceinline__ float vertex_features(const float* __restrict__ x, const float* __restrict__ y,
                                                 const float* __restrict__ z, const int* __restrict__ nbr, int v,
                                                 uint16_t* o)
{
    const int* nb = nbr + (size_t)v * kNbr;
    const float vx = x[v], vy = y[v], vz = z[v];
    float sum = 0.0f;
    int cnt = 0;
    for (int k = 0; k < kNbr; ++k) {
        int u = nb[k];
        if (u < 0) break;
        float dx = x[u] - vx, dy = y[u] - vy, dz = z[u] - vz;
        sum += sqrtf((dx * dx + dy * dy) + dz * dz);
        ++cnt;
    }
    float l = cnt > 0 ? sum / (float)cnt : 1.0f;
    if (!(l > 0.0f)) l = 1.0f;
    for (int k = 0; k < kNbr; ++k) {
        int u = nb[k];
        if (u < 0) {
            o[3 * k] = 0; o[3 * k + 1] = 0; o[3 * k + 2] = 0;
        } else {
            o[3 * k] = f2bf((x[u] - vx) / l);
            o[3 * k + 1] = f2bf((y[u] - vy) / l);
            o[3 * k + 2] = f2bf((z[u] - vz) / l);
        }
    }
    return l;
}

// one-ring features: thread per vertex
__global__ void k_features(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                           const int* __restrict__ nbr, int lo, int hi, uint16_t* __restrict__ X, float* __restrict__ L,
                           float* __restrict__ v0)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int v = lo + i;
    if (v >= hi) return;
    L[i] = vertex_features(x, y, z, nbr, v, X + (size_t)i * kIn);
    v0[3 * i] = x[v]; v0[3 * i + 1] = y[v]; v0[3 * i + 2] = z[v];
}

struct InferArgs {
    const float* x;
    const float* y;
    const float* z;
    const int* nbr;
    int lo, n;
    const uint16_t* Xin;   // optional precomputed features (n x 96)
    const uint16_t* wb;
    const float* w;
    float* Y;
    float* L;
};

// Fused inference, 16 rows per block of 4 waves (mlp_features16 /
// mlp_layers16): features (from the positions, or rows of Xin) -> LDS, the
// three layers on MFMA tiles.  Writes Y and L of rows [0, n).
__global__ __launch_bounds__(256) void k_mlp_infer(InferArgs a)
{
    __shared__ MlpLds sm;
    const int tid = threadIdx.x;
    const int row0 = blockIdx.x * kMlpRows;
    if (a.Xin) {
        for (int i = tid; i < kMlpRows * kIn; i += 256) {
            int r = i / kIn, c = i - r * kIn;
            sm.X[r * kLdx + c] = (row0 + r < a.n) ? a.Xin[(size_t)(row0 + r) * kIn + c] : (uint16_t)0;
        }
    } else {
        const int r = tid >> 4;
        const bool live = row0 + r < a.n;
        mlp_features16(a.x, a.y, a.z, a.nbr, a.lo + row0 + r, live, sm);
        if (live && (tid & 15) == 0) a.L[row0 + r] = sm.L[r];
    }
    __syncthreads();
    mlp_layers16(a.wb, a.w, sm);
    if (tid < kMlpRows * kOut) {
        const int r = tid / kOut, o = tid - r * kOut;
        if (row0 + r < a.n) a.Y[(size_t)(row0 + r) * kOut + o] = sm.Y[r][o];
    }
}

void launch_mlp_infer(vsdmo_mlp* m, const float* x, const float* y, const float* z, const int* nbr, int lo, int hi,
                      const uint16_t* Xin, float* Y, float* L, hipStream_t s)
{
    const int n = hi - lo;
    if (n <= 0) return;
    InferArgs a{ x, y, z, nbr, lo, n, Xin, m->wb.p, m->w.p, Y, L };
    k_mlp_infer<<<(n + kMlpRows - 1) / kMlpRows, 256, 0, s>>>(a);
}

__global__ void k_targets(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z, int lo,
                          int hi, const float* __restrict__ L, const float* __restrict__ v0, float* __restrict__ T)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int v = lo + i;
    if (v >= hi) return;
    const float l = L[i];
    T[3 * i] = (x[v] - v0[3 * i]) / l;
    T[3 * i + 1] = (y[v] - v0[3 * i + 1]) / l;
    T[3 * i + 2] = (z[v] - v0[3 * i + 2]) / l;
}

void launch_features(const float* x, const float* y, const float* z, const int* nbr, int lo, int hi, uint16_t* X,
                     float* L, float* v0, hipStream_t s)
{
    if (hi > lo) k_features<<<(hi - lo + 127) / 128, 128, 0, s>>>(x, y, z, nbr, lo, hi, X, L, v0);
}

void launch_targets(const float* x, const float* y, const float* z, int lo, int hi, const float* L, const float* v0,
                    float* T, hipStream_t s)
{
    if (hi > lo) k_targets<<<(hi - lo + 127) / 128, 128, 0, s>>>(x, y, z, lo, hi, L, v0, T);
}

void build_neighbour_table(vsdmo_mesh& m)
{
    if (m.nbr.p) return;
    const int st = (m.elem == VSDMO_TET) ? 3 : 7;
    const int ls = (m.elem == VSDMO_TET) ? kRingLocStrideTet : kRingLocStrideHex;
    std::vector<int> row = m.ringRow.to_host(m.stream);
    std::vector<int> val = m.ringVal.to_host(m.stream);
    std::vector<int> t((size_t)m.nV * kNbr, -1);
    const size_t nE = (size_t)row[m.nV];
    std::vector<uint8_t> loc(std::max<size_t>(nE * ls, 1), kRingLocNone);
    for (int v = m.nSurf; v < m.nV; ++v) {
        int* o = t.data() + (size_t)v * kNbr;
        int cnt = 0;
        for (int e = row[v]; e < row[v + 1]; ++e)
            for (int k = 0; k < st; ++k) {
                const int u = val[(size_t)e * st + k];
                int q = 0;
                while (q < cnt && o[q] != u) ++q;
                if (q == cnt && cnt < kNbr) o[cnt++] = u;   // first appearance
                if (q < cnt) loc[(size_t)e * ls + k] = (uint8_t)q;
            }
    }
    m.nbr.alloc(t.size());
    m.nbr.upload(t.data(), t.size(), m.stream);
    m.ringLoc.alloc(loc.size());
    m.ringLoc.upload(loc.data(), loc.size(), m.stream);
    VSDMO_HIP(hipStreamSynchronize(m.stream));
}

void mlp_reserve(vsdmo_mlp* m, int n)
{
    if (n <= m->cap) return;
    int c = std::max(n, 2 * m->cap);
    m->X.alloc((size_t)c * kIn);
    m->H1.alloc((size_t)c * kHid);
    m->H2.alloc((size_t)c * kHid);
    m->dZ1.alloc((size_t)c * kHid);
    m->dZ2.alloc((size_t)c * kHid);
    m->dY16.alloc((size_t)c * kOut);
    m->Z1.alloc((size_t)c * kHid);
    m->Z2.alloc((size_t)c * kHid);
    m->Y.alloc((size_t)c * kOut);
    m->T.alloc((size_t)c * kOut);
    m->part.alloc((size_t)c * kOut);   // dY in FP32
    m->lossd.alloc((size_t)(c + 255) / 256 + 1);
    m->cap = c;
}

static GemmArgs gargs(int M, int N, int K, const uint16_t* A, long long sam, long long sak, const uint16_t* B,
                      long long sbk, long long sbn)
{
    GemmArgs g{};
    g.M = M; g.N = N; g.K = K;
    g.A = A; g.sam = sam; g.sak = sak;
    g.B = B; g.sbk = sbk; g.sbn = sbn;
    return g;
}

// forward of n rows (X bf16 device) -> Y (device, FP32); keeps Z1, H1, Z2, H2
void mlp_forward_dev(vsdmo_mlp* m, const uint16_t* X, int n, float* Y, hipStream_t s)
{
    const uint16_t* wb = m->wb.p;
    const float* w = m->w.p;
    // Z1 = X W1^T + b1, H1 = gelu(Z1)
    GemmArgs g = gargs(n, kHid, kIn, X, kIn, 1, wb + kOW1, 1, kIn);
    g.bias = w + kOB1; g.mode = EPI_GELU; g.Cf = m->Z1.p; g.ldcf = kHid; g.Cb = m->H1.p; g.ldcb = kHid;
    gemm(g, s);
    g = gargs(n, kHid, kHid, m->H1.p, kHid, 1, wb + kOW2, 1, kHid);
    g.bias = w + kOB2; g.mode = EPI_GELU; g.Cf = m->Z2.p; g.ldcf = kHid; g.Cb = m->H2.p; g.ldcb = kHid;
    gemm(g, s);
    g = gargs(n, kOut, kHid, m->H2.p, kHid, 1, wb + kOW3, 1, kHid);
    g.bias = w + kOB3; g.mode = EPI_F32; g.Cf = Y; g.ldcf = kOut;
    gemm(g, s);
}

// loss and gradients of n rows (X, T in the work buffers); adds the mean
// loss to *lossAcc
static void mlp_backward_dev(vsdmo_mlp* m, int n, double* lossAcc, hipStream_t s)
{
    const uint16_t* wb = m->wb.p;
    float* gr = m->g.p;
    mlp_forward_dev(m, m->X.p, n, m->Y.p, s);
    const int nb = (n + 255) / 256;
    k_mse<<<nb, 256, 0, s>>>(m->Y.p, m->T.p, n, m->part.p, m->dY16.p, m->lossd.p);
    k_loss_final<<<1, 64, 0, s>>>(m->lossd.p, nb, n, lossAcc);
    // dW3 = dY^T H2 ; db3 = colsum dY
    GemmArgs g = gargs(kOut, kHid, n, m->dY16.p, 1, kOut, m->H2.p, kHid, 1);
    g.mode = EPI_F32; g.Cf = gr + kOW3; g.ldcf = kHid;
    gemm(g, s);
    k_colsum<float><<<kOut, 256, 0, s>>>(m->part.p, n, kOut, gr + kOB3);
    // dZ2 = (dY W3) * gelu'(Z2)
    g = gargs(n, kHid, kOut, m->dY16.p, kOut, 1, wb + kOW3, kHid, 1);
    g.mode = EPI_DGELU; g.Cb = m->dZ2.p; g.ldcb = kHid; g.Zin = m->Z2.p; g.ldz = kHid;
    gemm(g, s);
    // dW2 = dZ2^T H1 ; db2
    g = gargs(kHid, kHid, n, m->dZ2.p, 1, kHid, m->H1.p, kHid, 1);
    g.mode = EPI_F32; g.Cf = gr + kOW2; g.ldcf = kHid;
    gemm(g, s);
    k_colsum<uint16_t><<<kHid, 256, 0, s>>>(m->dZ2.p, n, kHid, gr + kOB2);
    // dZ1 = (dZ2 W2) * gelu'(Z1)
    g = gargs(n, kHid, kHid, m->dZ2.p, kHid, 1, wb + kOW2, kHid, 1);
    g.mode = EPI_DGELU; g.Cb = m->dZ1.p; g.ldcb = kHid; g.Zin = m->Z1.p; g.ldz = kHid;
    gemm(g, s);
    // dW1 = dZ1^T X ; db1
    g = gargs(kHid, kIn, n, m->dZ1.p, 1, kHid, m->X.p, kIn, 1);
    g.mode = EPI_F32; g.Cf = gr + kOW1; g.ldcf = kIn;
    gemm(g, s);
    k_colsum<uint16_t><<<kHid, 256, 0, s>>>(m->dZ1.p, n, kHid, gr + kOB1);
}

static void refresh_bf16(vsdmo_mlp* m, hipStream_t s)
{
    k_to_bf16<<<(kNParams + 255) / 256, 256, 0, s>>>(m->w.p, kNParams, m->wb.p);
}

static vsdmo_status fail_m(vsdmo_status st, const std::string& msg)
{
    set_error(msg);
    return st;
}

template <class F>
static vsdmo_status guarded_m(F&& f)
{
    try {
        return f();
    } catch (const HipError& e) {
        return fail_m(VSDMO_HIP_ERROR, std::string(hipGetErrorString(e.code)) + " in " + e.where);
    } catch (const StatusError& e) {
        return fail_m(e.status, e.msg);
    } catch (const std::bad_alloc&) {
        return fail_m(VSDMO_HIP_ERROR, "host allocation failed");
    }
}

static uint64_t smix(uint64_t& s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace vsdmo

using namespace vsdmo;

extern "C" {

int32_t vsdmo_mlp_num_params(void) { return kNParams; }

vsdmo_status vsdmo_mlp_create(uint64_t seed, vsdmo_mlp** out)
{
    return guarded_m([&]() -> vsdmo_status {
        if (!out) return fail_m(VSDMO_INVALID_ARG, "out is NULL");
        *out = nullptr;
        auto* m = new vsdmo_mlp();
        try {
            VSDMO_HIP(hipStreamCreate(&m->stream));
            // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases
            std::vector<float> w(kNParams);
            uint64_t s = seed;
            auto fill = [&](int off, int cnt, int fan) {
                float b = 1.0f / std::sqrt((float)fan);
                for (int i = 0; i < cnt; ++i) {
                    double u = (double)(smix(s) >> 11) * (1.0 / 9007199254740992.0);
                    w[off + i] = (float)((2.0 * u - 1.0) * b);
                }
            };
            fill(kOW1, kHid * kIn, kIn); fill(kOB1, kHid, kIn);
            fill(kOW2, kHid * kHid, kHid); fill(kOB2, kHid, kHid);
            fill(kOW3, kOut * kHid, kHid); fill(kOB3, kOut, kHid);
            m->w.alloc(kNParams); m->g.alloc(kNParams); m->m1.alloc(kNParams); m->m2.alloc(kNParams);
            m->wb.alloc(kNParams);
            m->w.upload(w.data(), kNParams, m->stream);
            std::vector<float> zero(kNParams, 0.0f);
            m->m1.upload(zero.data(), kNParams, m->stream);
            m->m2.upload(zero.data(), kNParams, m->stream);
            refresh_bf16(m, m->stream);
            VSDMO_HIP(hipStreamSynchronize(m->stream));
        } catch (...) {
            if (m->stream) (void)hipStreamDestroy(m->stream);
            delete m;
            throw;
        }
        *out = m;
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_mlp_destroy(vsdmo_mlp* m)
{
    if (!m) return VSDMO_OK;
    if (m->n_users > 0)
        return fail_m(VSDMO_INVALID_ARG, "network still set as the proposal of an optimiser or batch");
    if (m->stream) {
        (void)hipStreamSynchronize(m->stream);
        (void)hipStreamDestroy(m->stream);
    }
    delete m;
    return VSDMO_OK;
}

vsdmo_status vsdmo_mlp_get_params(vsdmo_mlp* m, float* out)
{
    return guarded_m([&]() -> vsdmo_status {
        if (!m || !out) return fail_m(VSDMO_INVALID_ARG, "NULL argument");
        m->w.download(out, kNParams, m->stream);
        VSDMO_HIP(hipStreamSynchronize(m->stream));
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_mlp_set_params(vsdmo_mlp* m, const float* in)
{
    return guarded_m([&]() -> vsdmo_status {
        if (!m || !in) return fail_m(VSDMO_INVALID_ARG, "NULL argument");
        m->w.upload(in, kNParams, m->stream);
        refresh_bf16(m, m->stream);
        std::vector<float> zero(kNParams, 0.0f);
        m->m1.upload(zero.data(), kNParams, m->stream);
        m->m2.upload(zero.data(), kNParams, m->stream);
        m->step = 0;
        VSDMO_HIP(hipStreamSynchronize(m->stream));
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_mlp_forward(vsdmo_mlp* m, const float* X, int32_t n, float* Y)
{
    return guarded_m([&]() -> vsdmo_status {
        if (!m || !X || !Y || n < 0) return fail_m(VSDMO_INVALID_ARG, "bad argument");
        if (n == 0) return VSDMO_OK;
        mlp_reserve(m, n);
        std::vector<uint16_t> xb((size_t)n * kIn);
        for (size_t i = 0; i < xb.size(); ++i) xb[i] = f2bf(X[i]);
        m->X.upload(xb.data(), xb.size(), m->stream);
        launch_mlp_infer(m, nullptr, nullptr, nullptr, nullptr, 0, n, m->X.p, m->Y.p, nullptr, m->stream);
        m->Y.download(Y, (size_t)n * kOut, m->stream);
        VSDMO_HIP(hipStreamSynchronize(m->stream));
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_mlp_gradients(vsdmo_mlp* m, const float* X, const float* T, int32_t n, double* loss, float* grad)
{
    return guarded_m([&]() -> vsdmo_status {
        if (!m || !X || !T || n <= 0) return fail_m(VSDMO_INVALID_ARG, "bad argument");
        mlp_reserve(m, n);
        std::vector<uint16_t> xb((size_t)n * kIn);
        for (size_t i = 0; i < xb.size(); ++i) xb[i] = f2bf(X[i]);
        m->X.upload(xb.data(), xb.size(), m->stream);
        m->T.upload(T, (size_t)n * kOut, m->stream);
        DevBuf<double> acc(1);
        double z = 0.0;
        acc.upload(&z, 1, m->stream);
        mlp_backward_dev(m, n, acc.p, m->stream);
        if (grad) m->g.download(grad, kNParams, m->stream);
        double l = 0.0;
        acc.download(&l, 1, m->stream);
        VSDMO_HIP(hipStreamSynchronize(m->stream));
        if (loss) *loss = l;
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_mlp_train(vsdmo_mlp* m, vsdmo_dataset* ds, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                             double* epoch_loss)
{
    return guarded_m([&]() -> vsdmo_status {
        if (!m || !ds) return fail_m(VSDMO_INVALID_ARG, "NULL argument");
        if (epochs < 0 || batch <= 0 || !(lr > 0.0f)) return fail_m(VSDMO_INVALID_ARG, "bad training parameters");
        const long long N = ds->n;
        if (N < batch) return fail_m(VSDMO_INVALID_ARG, "dataset smaller than one batch");
        if (N >= (1LL << 31)) return fail_m(VSDMO_INVALID_ARG, "dataset too large");
        hipStream_t s = m->stream;
        VSDMO_HIP(hipStreamSynchronize(ds->stream));
        mlp_reserve(m, batch);
        m->perm.alloc((size_t)N);
        std::vector<int> perm((size_t)N);
        DevBuf<double> acc(1);
        uint64_t rs = seed;
        const long long steps = N / batch;   // the ragged tail of each shuffled epoch is dropped
        for (int ep = 0; ep < epochs; ++ep) {
            for (long long i = 0; i < N; ++i) perm[i] = (int)i;
            for (long long i = N - 1; i > 0; --i) std::swap(perm[i], perm[(long long)(smix(rs) % (uint64_t)(i + 1))]);
            m->perm.upload(perm.data(), (size_t)N, s);
            double z = 0.0;
            acc.upload(&z, 1, s);
            for (long long st = 0; st < steps; ++st) {
                k_gather_rows<<<(batch * kIn + 255) / 256, 256, 0, s>>>(ds->X.p, ds->T.p, m->perm.p + st * batch, batch,
                                                                      m->X.p, m->T.p);
                mlp_backward_dev(m, batch, acc.p, s);
                m->step++;
                const float c1 = 1.0f - std::pow(0.9f, (float)m->step), c2 = 1.0f - std::pow(0.999f, (float)m->step);
                k_adam<<<(kNParams + 255) / 256, 256, 0, s>>>(m->w.p, m->g.p, m->m1.p, m->m2.p, m->wb.p, kNParams, lr,
                                                            c1, c2);
            }
            double l = 0.0;
            acc.download(&l, 1, s);
            VSDMO_HIP(hipStreamSynchronize(s));
            if (epoch_loss) epoch_loss[ep] = l / (double)steps;
        }
        VSDMO_HIP(hipGetLastError());
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_dataset_create(int64_t capacity, vsdmo_dataset** out)
{
    return guarded_m([&]() -> vsdmo_status {
        if (!out || capacity <= 0) return fail_m(VSDMO_INVALID_ARG, "bad argument");
        *out = nullptr;
        auto* d = new vsdmo_dataset();
        try {
            VSDMO_HIP(hipStreamCreate(&d->stream));
            d->cap = capacity;
            d->X.alloc((size_t)capacity * kIn);
            d->T.alloc((size_t)capacity * kOut);
        } catch (...) {
            if (d->stream) (void)hipStreamDestroy(d->stream);
            delete d;
            throw;
        }
        *out = d;
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_dataset_size(vsdmo_dataset* d, int64_t* n)
{
    if (!d || !n) return fail_m(VSDMO_INVALID_ARG, "NULL argument");
    *n = d->n;
    return VSDMO_OK;
}

vsdmo_status vsdmo_dataset_get(vsdmo_dataset* d, int64_t start, int32_t n, float* X, float* T)
{
    return guarded_m([&]() -> vsdmo_status {
        if (!d || start < 0 || n < 0 || start + n > d->n) return fail_m(VSDMO_INVALID_ARG, "range outside the dataset");
        if (X) {
            std::vector<uint16_t> xb((size_t)n * kIn);
            VSDMO_HIP(hipMemcpyAsync(xb.data(), d->X.p + start * kIn, xb.size() * 2, hipMemcpyDeviceToHost, d->stream));
            VSDMO_HIP(hipStreamSynchronize(d->stream));
            for (size_t i = 0; i < xb.size(); ++i) X[i] = bf2f(xb[i]);
        }
        if (T) {
            VSDMO_HIP(hipMemcpyAsync(T, d->T.p + start * kOut, sizeof(float) * n * kOut, hipMemcpyDeviceToHost, d->stream));
            VSDMO_HIP(hipStreamSynchronize(d->stream));
        }
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_dataset_destroy(vsdmo_dataset* d)
{
    if (!d) return VSDMO_OK;
    if (d->stream) {
        (void)hipStreamSynchronize(d->stream);
        (void)hipStreamDestroy(d->stream);
    }
    delete d;
    return VSDMO_OK;
}

}  // extern "C"
