// dmo_kernels.hip -- DMO vertex-search and quality kernels for gfx950.
//
// One wave64 per vertex.  For the 3D searches the nq^3 candidates of a level
// are split so that lane l = j + nq*k owns the nq candidates (i, j, k),
// i = 0..nq-1, i.e. the contiguous candidate indices [nq*l, nq*l + nq): a
// lowest-lane-then-lowest-i argmax is the reference's lowest-index tie rule.
// The one-ring is staged in LDS as per-element constants in a vertex-local
// frame; per element and lane the (y, z) part of det and L is formed once and
// reused for the nq candidates along x.
//
// Reference kernels: k_optimizeHierarchicalInnerTet / Hex
// (VSDMO/src/dmo/DMOTetImplCub.h:20-137, DMOHexImplCub.h:20-137) and
// k_optimizeHierarchical2DFlat (DMOTriFlatImplCub.h:23-152).
#include <float.h>
#include <type_traits>

#include "vsdmo_internal.h"
#include "vsdmo_numerics.h"
#include "dmo_kernels.h"
#include "mlp_internal.h"

namespace vsdmo {

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m));
    return v;
}
__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fminf(v, __shfl_xor(v, m));
    return v;
}
// argmax over the wave: larger score wins, equal scores -> smaller index
__device__ __forceinline__ void wave_argmax(float& best, int& idx)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        float ob = __shfl_xor(best, m);
        int oi = __shfl_xor(idx, m);
        bool take = (ob > best) || (ob == best && oi < idx);
        best = take ? ob : best;
        idx = take ? oi : idx;
    }
}

// Block-level reductions for W waves per vertex (blockDim = 64*W).  With
// W = 1 they are the wave shuffles alone; otherwise the per-wave results are
// combined through LDS.  max/min and the (score, index) argmax are order
// independent, so every W gives the same result.
template <int W>
__device__ __forceinline__ float block_max(float v, float* s)
{
    v = wave_max(v);
    if constexpr (W > 1) {
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) s[w] = v;
        __syncthreads();
        v = s[0];
#pragma unroll
        for (int k = 1; k < W; ++k) v = fmaxf(v, s[k]);
        __syncthreads();
    }
    return v;
}
template <int W>
__device__ __forceinline__ float block_min(float v, float* s)
{
    v = wave_min(v);
    if constexpr (W > 1) {
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) s[w] = v;
        __syncthreads();
        v = s[0];
#pragma unroll
        for (int k = 1; k < W; ++k) v = fminf(v, s[k]);
        __syncthreads();
    }
    return v;
}
template <int W>
__device__ __forceinline__ void block_argmax(float& best, int& idx, float* sf, int* si)
{
    wave_argmax(best, idx);
    if constexpr (W > 1) {
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) { sf[w] = best; si[w] = idx; }
        __syncthreads();
        best = sf[0];
        idx = si[0];
#pragma unroll
        for (int k = 1; k < W; ++k) {
            const float ob = sf[k];
            const int oi = si[k];
            if (ob > best || (ob == best && oi < idx)) { best = ob; idx = oi; }
        }
        __syncthreads();
    }
}

// LDS scratch of the block reductions
template <int W>
struct RedLds { float f[W]; int i[W]; float e[4][W]; };

// The staging reductions in one exchange: per-axis window half-extents (min
// or max of the offsets) and the largest offset.  One barrier: rd.e is read
// here and written again only by the next vertex's staging, after the
// barrier that ends this one.
template <int W>
__device__ __forceinline__ void block_extents(bool emin, float& ex, float& ey, float& ez, float& mabs, RedLds<W>& rd)
{
    if (emin) { ex = wave_min(ex); ey = wave_min(ey); ez = wave_min(ez); }
    else { ex = wave_max(ex); ey = wave_max(ey); ez = wave_max(ez); }
    mabs = wave_max(mabs);
    if constexpr (W > 1) {
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) { rd.e[0][w] = ex; rd.e[1][w] = ey; rd.e[2][w] = ez; rd.e[3][w] = mabs; }
        __syncthreads();
        ex = rd.e[0][0]; ey = rd.e[1][0]; ez = rd.e[2][0]; mabs = rd.e[3][0];
#pragma unroll
        for (int k = 1; k < W; ++k) {
            ex = emin ? fminf(ex, rd.e[0][k]) : fmaxf(ex, rd.e[0][k]);
            ey = emin ? fminf(ey, rd.e[1][k]) : fmaxf(ey, rd.e[1][k]);
            ez = emin ? fminf(ez, rd.e[2][k]) : fmaxf(ez, rd.e[2][k]);
            mabs = fmaxf(mabs, rd.e[3][k]);
        }
    }
}

// ------------------------------------------------------------------ tets
// The staged state of one vertex search: start position, window
// half-extents, frame scale and ring size; the ring's constants are in LDS.
struct TetFrame {
    float x0x, x0y, x0z, ex, ey, ez, sg, is3;
    int n;
};

// Stage vid's one-ring: per-axis window half-extent (max or min offset), the
// frame scale and the constants of every ring element in LDS.  Every thread
// of the block calls it; it ends with a barrier.
template <int W>
__device__ TetFrame tet_stage(const SearchArgs& a, int vid, bool emin, TetC* s_tc, RedLds<W>& rd)
{
    constexpr int T = 64 * W;
    // W = 1 bodies also run one per wave inside wider blocks (k_search_split)
    const int tid = (W == 1) ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
    TetFrame f;
    const int r0 = a.ringRow[vid];
    f.n = a.ringRow[vid + 1] - r0;
    f.x0x = a.x[vid]; f.x0y = a.y[vid]; f.x0z = a.z[vid];
    float ex = emin ? FLT_MAX : -FLT_MAX, ey = ex, ez = ex, mabs = 0.0f;
    auto extent = [&](float dxs, float dys, float dzs) {
        const float dx = fabsf(dxs), dy = fabsf(dys), dz = fabsf(dzs);
        ex = emin ? fminf(ex, dx) : fmaxf(ex, dx);
        ey = emin ? fminf(ey, dy) : fmaxf(ey, dy);
        ez = emin ? fminf(ez, dz) : fmaxf(ez, dz);
        mabs = fmaxf(mabs, fmaxf(dx, fmaxf(dy, dz)));
    };
    // the thread's first element (all of them when the ring has at most T)
    // keeps its corner offsets in registers for the constants, so the
    // positions are read from memory once
    const bool own0 = tid < f.n;
    float d0[3][3];
    if (own0) {
        const int* h = a.ringVal + (size_t)(r0 + tid) * 3;
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const int v = h[t];
            d0[t][0] = a.x[v] - f.x0x; d0[t][1] = a.y[v] - f.x0y; d0[t][2] = a.z[v] - f.x0z;
            extent(d0[t][0], d0[t][1], d0[t][2]);
        }
    }
    for (int e = tid + T; e < f.n; e += T) {
        const int* h = a.ringVal + (size_t)(r0 + e) * 3;
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            const int v = h[t];
            extent(a.x[v] - f.x0x, a.y[v] - f.x0y, a.z[v] - f.x0z);
        }
    }
    block_extents<W>(emin, ex, ey, ez, mabs, rd);
    f.ex = ex; f.ey = ey; f.ez = ez;
    f.sg = pow2_scale(mabs);
    const float isg = 1.0f / f.sg;
    f.is3 = ((1.0f * isg) * isg) * isg;
    if (own0) {
        f3 p[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) p[t] = { d0[t][0] * f.sg, d0[t][1] * f.sg, d0[t][2] * f.sg };
        s_tc[tid] = tet_consts(p[0], p[1], p[2]);
    }
    for (int e = tid + T; e < f.n; e += T) {
        const int* h = a.ringVal + (size_t)(r0 + e) * 3;
        f3 p[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            int v = h[t];
            p[t] = { (a.x[v] - f.x0x) * f.sg, (a.y[v] - f.x0y) * f.sg, (a.z[v] - f.x0z) * f.sg };
        }
        s_tc[e] = tet_consts(p[0], p[1], p[2]);
    }
    __syncthreads();
    return f;
}

// ring score of one candidate p' (frame coordinates)
__device__ __forceinline__ float tet_score(const TetFrame& f, const TetC* s_tc, float px, float py, float pz)
{
    const float Qyz = 3.0f * fmaf(pz, pz, py * py);
    const float t3 = 3.0f * px;
    TetAcc acc = tet_acc0();
    for (int e = 0; e < f.n; ++e) {
        const TetC c = s_tc[e];
        float Tyz = fmaf(c.nny, py, fmaf(c.nnz, pz, c.D));
        float Ryz = fmaf(c.m2sy, py, fmaf(c.m2sz, pz, c.K + Qyz));
        tet_acc_add(acc, fmaf(c.nnx, px, Tyz), fmaf(px, t3 + c.m2sx, Ryz));
    }
    return tet_acc_score(acc, f.is3);
}

// The nested grid search from the start position (depth levels of NQ^3
// candidates) with W waves.  Lane l = j + nq*k of wave w owns the candidates
// (i, j, k) with i in [w*nq/W, (w+1)*nq/W); candidate index i + nq*l, so the
// lowest index wins ties as in the reference.  The start score qcur is
// accumulated in the first level's element loop (at p = 0, det = D and L =
// K exactly: tet_score's terms, in ring order).  Returns whether the vertex
// moved, and its position in (cx, cy, cz).
template <int NQ, int W>
__device__ bool tet_levels(const SearchArgs& a, const TetFrame& f, const TetC* s_tc, RedLds<W>& rd, float& cx,
                           float& cy, float& cz)
{
    constexpr int CPL = NQ / W;
    static_assert(NQ % W == 0, "nq must be a multiple of the waves per vertex");
    const int tid = (W == 1) ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
    const int lane = tid & 63, i0 = (tid >> 6) * CPL;
    const int j = lane % NQ, k = lane / NQ;
    const bool active = lane < NQ * NQ;
    float ds = a.gridScale;
    bool moved = false;
    float qcur = 0.0f;
    for (int d = 0; d < a.depth; ++d) {
        const float gxM = cx + ds * f.ex, gxm = cx - ds * f.ex;
        const float gyM = cy + ds * f.ey, gym = cy - ds * f.ey;
        const float gzM = cz + ds * f.ez, gzm = cz - ds * f.ez;
        const float py = (grid_coord(a.aff, j, NQ, gym, gyM) - f.x0y) * f.sg;
        const float pz = (grid_coord(a.aff, k, NQ, gzm, gzM) - f.x0z) * f.sg;
        float px[CPL], t3[CPL];
#pragma unroll
        for (int i = 0; i < CPL; ++i) {
            px[i] = (grid_coord(a.aff, i0 + i, NQ, gxm, gxM) - f.x0x) * f.sg;
            t3[i] = 3.0f * px[i];
        }
        const float Qyz = 3.0f * fmaf(pz, pz, py * py);
        TetAcc acc[CPL];
#pragma unroll
        for (int i = 0; i < CPL; ++i) acc[i] = tet_acc0();
        // two elements per iteration: their minima fold into v_min3 without
        // canonicalising each accumulator per element (111 -> 97.5 VALU per
        // element at 8 candidates per lane)
        auto ring = [&](auto withQ0) {
            constexpr bool WQ = decltype(withQ0)::value;
            TetAcc acc0 = tet_acc0();
#pragma unroll 2
            for (int e = 0; e < f.n; ++e) {
                const TetC c = s_tc[e];
                const float Tyz = fmaf(c.nny, py, fmaf(c.nnz, pz, c.D));
                const float Ryz = fmaf(c.m2sy, py, fmaf(c.m2sz, pz, c.K + Qyz));
#pragma unroll
                for (int i = 0; i < CPL; ++i)
                    tet_acc_add(acc[i], fmaf(c.nnx, px[i], Tyz), fmaf(px[i], t3[i] + c.m2sx, Ryz));
                if constexpr (WQ) tet_acc_add(acc0, c.D, c.K);
            }
            if constexpr (WQ) qcur = tet_acc_score(acc0, f.is3);
        };
        if (d == 0) ring(std::true_type{});
        else ring(std::false_type{});
        float best = -FLT_MAX;
        int bi = 0;
#pragma unroll
        for (int i = 0; i < CPL; ++i) {
            float q = tet_acc_score(acc[i], f.is3);
            if (q > best) { best = q; bi = i; }
        }
        if (!active) best = -FLT_MAX;
        int cand = lane * NQ + i0 + bi;
        block_argmax<W>(best, cand, rd.f, rd.i);
        if (qcur < best) {
            const int wi = cand % NQ, wj = (cand / NQ) % NQ, wk = cand / (NQ * NQ);
            cx = grid_coord(a.aff, wi, NQ, gxm, gxM);
            cy = grid_coord(a.aff, wj, NQ, gym, gyM);
            cz = grid_coord(a.aff, wk, NQ, gzm, gzM);
            qcur = best;
            moved = true;
        }
        ds = ds * a.shrink;
    }
    return moved;
}

// One vertex per block of W waves (or per wave, W = 1).
template <int NQ, int W>
__device__ void tet_vertex(const SearchArgs& a, int vid, bool emin, TetC* s_tc, RedLds<W>& rd)
{
    const int tid = (W == 1) ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
    const TetFrame f = tet_stage<W>(a, vid, emin, s_tc, rd);
    float cx = f.x0x, cy = f.x0y, cz = f.x0z;
    if (tet_levels<NQ, W>(a, f, s_tc, rd, cx, cy, cz) && tid == 0) {
        a.xw[vid] = cx; a.yw[vid] = cy; a.zw[vid] = cz;
    }
}

// ------------------------------------------------------------------ hexes
struct HexLds { float cd[4], cF[4]; HexP hp[5]; };

// staged state of a hex vertex search (as TetFrame), plus the accumulator of
// the 4 sub-tets per ring element that do not contain the moving vertex
struct HexFrame {
    float x0x, x0y, x0z, ex, ey, ez, sg, epsD, epsF;
    int n;
    bool corner8;   // the principal-axes sub-tet (slot 4 of HexP) left out
    HexAcc acc0;
};

template <int W>
__device__ HexFrame hex_stage(const SearchArgs& a, int vid, bool emin, HexLds* s_hx, RedLds<W>& rd)
{
    constexpr int T = 64 * W;
    const int tid = (W == 1) ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
    HexFrame f;
    const int r0 = a.ringRow[vid];
    f.n = a.ringRow[vid + 1] - r0;
    f.x0x = a.x[vid]; f.x0y = a.y[vid]; f.x0z = a.z[vid];
    f.corner8 = a.hexCorner8 != 0;
    float ex = emin ? FLT_MAX : -FLT_MAX, ey = ex, ez = ex, mabs = 0.0f;
    auto extent = [&](float dxs, float dys, float dzs) {
        const float dx = fabsf(dxs), dy = fabsf(dys), dz = fabsf(dzs);
        ex = emin ? fminf(ex, dx) : fmaxf(ex, dx);
        ey = emin ? fminf(ey, dy) : fmaxf(ey, dy);
        ez = emin ? fminf(ez, dz) : fmaxf(ez, dz);
        mabs = fmaxf(mabs, fmaxf(dx, fmaxf(dy, dz)));
    };
    // the thread's first element keeps its corner offsets in registers for
    // the constants (the positions are read once; tet_stage)
    const bool own0 = tid < f.n;
    float d0[7][3];
    if (own0) {
        const int* h = a.ringVal + (size_t)(r0 + tid) * 7;
#pragma unroll
        for (int t = 0; t < 7; ++t) {
            const int v = h[t];
            d0[t][0] = a.x[v] - f.x0x; d0[t][1] = a.y[v] - f.x0y; d0[t][2] = a.z[v] - f.x0z;
            extent(d0[t][0], d0[t][1], d0[t][2]);
        }
    }
    for (int e = tid + T; e < f.n; e += T) {
        const int* h = a.ringVal + (size_t)(r0 + e) * 7;
        for (int t = 0; t < 7; ++t) {
            const int v = h[t];
            extent(a.x[v] - f.x0x, a.y[v] - f.x0y, a.z[v] - f.x0z);
        }
    }
    block_extents<W>(emin, ex, ey, ez, mabs, rd);
    f.ex = ex; f.ey = ey; f.ez = ez;
    f.sg = pow2_scale(mabs);
    const float s2 = f.sg * f.sg, s3 = s2 * f.sg;
    f.epsD = kHexEps * s3; f.epsF = kHexEps * s2;
    if (own0) {
        f3 r[8];
        r[0] = { 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int t = 0; t < 7; ++t) r[t + 1] = { d0[t][0] * f.sg, d0[t][1] * f.sg, d0[t][2] * f.sg };
        HexLds L;
        hex_consts(r, L.cd, L.cF, L.hp);
        s_hx[tid] = L;
    }
    for (int e = tid + T; e < f.n; e += T) {
        const int* h = a.ringVal + (size_t)(r0 + e) * 7;
        f3 r[8];
        r[0] = { 0.0f, 0.0f, 0.0f };
        for (int t = 0; t < 7; ++t) {
            int v = h[t];
            r[t + 1] = { (a.x[v] - f.x0x) * f.sg, (a.y[v] - f.x0y) * f.sg, (a.z[v] - f.x0z) * f.sg };
        }
        HexLds L;
        hex_consts(r, L.cd, L.cF, L.hp);
        s_hx[e] = L;
    }
    __syncthreads();
    // p-free sub-tets of the whole ring, in ring order
    f.acc0 = hex_acc0();
    for (int e = 0; e < f.n; ++e)
#pragma unroll
        for (int sIdx = 0; sIdx < 4; ++sIdx) hex_acc_add(f.acc0, s_hx[e].cd[sIdx], s_hx[e].cF[sIdx]);
    return f;
}

// Whether the vertex cannot move: when the sub-tets it is not part of
// already score 0 (a determinant below epsD or a Frobenius term below epsF),
// every candidate's accumulator starts from acc0 and takes fminf of those
// minima, so every candidate -- the start included -- scores exactly 0 and
// none is strictly better than the start.  The search is skipped; the result
// is the full search's.
__device__ __forceinline__ bool hex_pinned(const HexFrame& f)
{
    return f.acc0.dmin < f.epsD || f.acc0.fmin < f.epsF;
}

// the p-dependent sub-tets of the ring at one point: NS = 5 (the nine-tet
// metric) or 4 (eight corners).  The slot count is a template argument so
// that the callers branch on the metric once, outside the element loop (a
// branch inside it holds the LDS reads of the last slot back)
template <int NS>
__device__ __forceinline__ void hex_ring_add(HexAcc& acc, const HexLds* s_hx, int n, float px, float py, float pz)
{
    const float Q1 = fmaf(pz, pz, py * py), Q3 = 3.0f * Q1, t3 = 3.0f * px;
    for (int e = 0; e < n; ++e)
#pragma unroll
        for (int sIdx = 0; sIdx < NS; ++sIdx) {
            const HexP h = s_hx[e].hp[sIdx];
            const bool a3 = (sIdx == 1 || sIdx == 4);
            float Tyz = fmaf(h.gy, py, fmaf(h.gz, pz, h.D));
            float Ryz = fmaf(h.sy, py, fmaf(h.sz, pz, h.K + (a3 ? Q3 : Q1)));
            hex_acc_add(acc, fmaf(h.gx, px, Tyz), fmaf(px, (a3 ? t3 : px) + h.sx, Ryz));
        }
}

__device__ __forceinline__ float hex_score(const HexFrame& f, const HexLds* s_hx, float px, float py, float pz)
{
    HexAcc acc = f.acc0;
    if (f.corner8) hex_ring_add<4>(acc, s_hx, f.n, px, py, pz);
    else hex_ring_add<5>(acc, s_hx, f.n, px, py, pz);
    return hex_acc_score(acc, f.epsD, f.epsF);
}

// the p-dependent sub-tets of the ring for the CPL candidates px[i] of a lane
// (shared y, z): NS slots per element as in hex_ring_add
template <int NS, int CPL>
__device__ __forceinline__ void hex_ring_add_cpl(HexAcc* acc, const HexLds* s_hx, int n, const float* px,
                                                 const float* t3, float py, float pz)
{
    const float Q1 = fmaf(pz, pz, py * py), Q3 = 3.0f * Q1;
    // two elements per iteration at 8 candidates per lane: the minima of
    // consecutive elements fold into v_min3 without canonicalising the
    // accumulator each time (533 -> 506 VALU per element); at fewer
    // candidates per lane the register growth costs more than it saves
    constexpr int kUnrollE = CPL >= 8 ? 2 : 1;
#pragma unroll kUnrollE
    for (int e = 0; e < n; ++e) {
#pragma unroll
        for (int sIdx = 0; sIdx < NS; ++sIdx) {
            const HexP h = s_hx[e].hp[sIdx];
            const bool a3 = (sIdx == 1 || sIdx == 4);
            const float Tyz = fmaf(h.gy, py, fmaf(h.gz, pz, h.D));
            const float Ryz = fmaf(h.sy, py, fmaf(h.sz, pz, h.K + (a3 ? Q3 : Q1)));
#pragma unroll
            for (int i = 0; i < CPL; ++i)
                hex_acc_add(acc[i], fmaf(h.gx, px[i], Tyz), fmaf(px[i], (a3 ? t3[i] : px[i]) + h.sx, Ryz));
        }
    }
}

// the nested grid search of a hex vertex (lane layout as tet_levels)
template <int NQ, int W>
__device__ bool hex_levels(const SearchArgs& a, const HexFrame& f, const HexLds* s_hx, RedLds<W>& rd, float qcur,
                           float& cx, float& cy, float& cz)
{
    constexpr int CPL = NQ / W;
    static_assert(NQ % W == 0, "nq must be a multiple of the waves per vertex");
    const int tid = (W == 1) ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
    const int lane = tid & 63, i0 = (tid >> 6) * CPL;
    const int j = lane % NQ, k = lane / NQ;
    const bool active = lane < NQ * NQ;
    float ds = a.gridScale;
    bool moved = false;
    for (int d = 0; d < a.depth; ++d) {
        const float gxM = cx + ds * f.ex, gxm = cx - ds * f.ex;
        const float gyM = cy + ds * f.ey, gym = cy - ds * f.ey;
        const float gzM = cz + ds * f.ez, gzm = cz - ds * f.ez;
        const float py = (grid_coord(a.aff, j, NQ, gym, gyM) - f.x0y) * f.sg;
        const float pz = (grid_coord(a.aff, k, NQ, gzm, gzM) - f.x0z) * f.sg;
        float px[CPL], t3[CPL];
#pragma unroll
        for (int i = 0; i < CPL; ++i) {
            px[i] = (grid_coord(a.aff, i0 + i, NQ, gxm, gxM) - f.x0x) * f.sg;
            t3[i] = 3.0f * px[i];
        }
        HexAcc acc[CPL];
#pragma unroll
        for (int i = 0; i < CPL; ++i) acc[i] = f.acc0;
        if (a.hexCorner8) hex_ring_add_cpl<4, CPL>(acc, s_hx, f.n, px, t3, py, pz);   // uniform: a kernel argument
        else hex_ring_add_cpl<5, CPL>(acc, s_hx, f.n, px, t3, py, pz);
        float best = -FLT_MAX;
        int bi = 0;
#pragma unroll
        for (int i = 0; i < CPL; ++i) {
            float q = hex_acc_score(acc[i], f.epsD, f.epsF);
            if (q > best) { best = q; bi = i; }
        }
        if (!active) best = -FLT_MAX;
        int cand = lane * NQ + i0 + bi;
        block_argmax<W>(best, cand, rd.f, rd.i);
        if (qcur < best) {
            const int wi = cand % NQ, wj = (cand / NQ) % NQ, wk = cand / (NQ * NQ);
            cx = grid_coord(a.aff, wi, NQ, gxm, gxM);
            cy = grid_coord(a.aff, wj, NQ, gym, gyM);
            cz = grid_coord(a.aff, wk, NQ, gzm, gzM);
            qcur = best;
            moved = true;
        }
        ds = ds * a.shrink;
    }
    return moved;
}

template <int NQ, int W>
__device__ void hex_vertex(const SearchArgs& a, int vid, bool emin, HexLds* s_hx, RedLds<W>& rd)
{
    const int tid = (W == 1) ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
    const HexFrame f = hex_stage<W>(a, vid, emin, s_hx, rd);
    if (hex_pinned(f)) return;   // uniform over the block: f is computed by every thread from the LDS
    const float qcur = hex_score(f, s_hx, 0.0f, 0.0f, 0.0f);
    float cx = f.x0x, cy = f.x0y, cz = f.x0z;
    if (hex_levels<NQ, W>(a, f, s_hx, rd, qcur, cx, cy, cz) && tid == 0) {
        a.xw[vid] = cx; a.yw[vid] = cy; a.zw[vid] = cz;
    }
}

// ------------------------------------------------------------- flat tris
// Lane l of wave 0 holds the reference thread t = l/2's candidate B (l even)
// or A (l odd): A = (t/nq, t%nq), B = (t/nq + nq/2, t%nq).  The lowest lane
// among the maxima is the reference's winner (lowest thread; B on an exact
// tie).  Further waves (W > 1) only help stage the ring.
template <int NQ, int W>
__device__ void flat_vertex(const SearchArgs& a, int vid, TriC* s_tr, RedLds<W>& rd)
{
    constexpr int T = 64 * W;
    // W = 1 bodies also run one per wave inside wider blocks (k_search_split)
    const int tid = (W == 1) ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
    const int h0 = a.heRow[vid];
    const int n = a.heRow[vid + 1] - h0;
    const float x0x = a.x[vid], x0y = a.y[vid];

    float ex = -FLT_MAX, ey = -FLT_MAX, mabs = 0.0f;
    for (int e = tid; e < n; e += T) {
        int v = a.heTarget[h0 + e];
        float dx = fabsf(a.x[v] - x0x), dy = fabsf(a.y[v] - x0y);
        ex = fmaxf(ex, dx);
        ey = fmaxf(ey, dy);
        mabs = fmaxf(mabs, fmaxf(dx, dy));
    }
    float ez = 0.0f;
    block_extents<W>(false, ex, ey, ez, mabs, rd);
    const float sg = pow2_scale(mabs);
    const float isg = 1.0f / sg;
    const float is2 = (1.0f * isg) * isg;
    for (int e = tid; e < n; e += T) {
        int va = a.heTarget[h0 + e];
        int vb = a.heTarget[h0 + (e + 1 == n ? 0 : e + 1)];
        s_tr[e] = tri_consts((a.x[va] - x0x) * sg, (a.y[va] - x0y) * sg, (a.x[vb] - x0x) * sg, (a.y[vb] - x0y) * sg);
    }
    __syncthreads();

    auto score = [&](float px, float py) {
        const float Q2 = 2.0f * (py * py);
        const float t2 = 2.0f * px;
        TriAcc acc = tri_acc0();
        for (int e = 0; e < n; ++e) {
            const TriC c = s_tr[e];
            float Ty = fmaf(c.gy, py, c.C);
            float Ry = fmaf(c.sy, py, c.K + Q2);
            tri_acc_add(acc, fmaf(c.gx, px, Ty), fmaf(px, t2 + c.sx, Ry));
        }
        return tri_acc_score(acc, is2);
    };

    float qcur = score(0.0f, 0.0f);
    const float qold = qcur;
    const int t = tid >> 1;
    const bool isA = (tid & 1) != 0;
    const int ci = t / NQ + (isA ? 0 : NQ / 2), cj = t % NQ;
    const bool active = tid < 64 && t < NQ * NQ / 2;
    float cx = x0x, cy = x0y;
    float ds = a.gridScale;
    float qlast = -FLT_MAX;
    for (int d = 0; d < a.depth; ++d) {
        const float gxM = cx + ds * ex, gxm = cx - ds * ex;
        const float gyM = cy + ds * ey, gym = cy - ds * ey;
        float best = active ? score((grid_coord(a.aff, ci, NQ, gxm, gxM) - x0x) * sg,
                                    (grid_coord(a.aff, cj, NQ, gym, gyM) - x0y) * sg)
                            : -FLT_MAX;
        int wl = tid;
        block_argmax<W>(best, wl, rd.f, rd.i);
        if (qcur < best) {
            const int wt = wl >> 1;
            const int wi = wt / NQ + ((wl & 1) ? 0 : NQ / 2), wj = wt % NQ;
            cx = grid_coord(a.aff, wi, NQ, gxm, gxM);
            cy = grid_coord(a.aff, wj, NQ, gym, gyM);
            qcur = best;
        }
        qlast = best;
        ds = ds * a.shrink;
    }
    if (qold < qlast && tid == 0) { a.xw[vid] = cx; a.yw[vid] = cy; }
}

// one vertex per block, vertices [cStart, cEnd) of one colour class.  The
// tet search asks for at least 5 waves per SIMD (launch bounds' second
// argument; 89 VGPRs, no spill): a colour class of the 40^3 cube (~7.4k
// vertices) runs in ~1.45 rounds of the 256 CUs' 5 x 4 wave slots.  6 or 7
// waves per SIMD spill (7 / 18 registers) with the element loop taking two
// elements per iteration.
#ifndef VSDMO_TET_MIN_WAVES
#define VSDMO_TET_MIN_WAVES 5
#endif
template <int NQ, int W>
__global__ __launch_bounds__(64 * W, VSDMO_TET_MIN_WAVES) void k_search_tet(SearchArgs a)
{
    extern __shared__ TetC s_tc[];
    __shared__ RedLds<W> rd;
    tet_vertex<NQ, W>(a, a.cStart + blockIdx.x, a.extentMin != 0, s_tc, rd);
}
template <int NQ, int W>
__global__ __launch_bounds__(64 * W) void k_search_hex(SearchArgs a)
{
    extern __shared__ HexLds s_hx[];
    __shared__ RedLds<W> rd;
    hex_vertex<NQ, W>(a, a.cStart + blockIdx.x, a.extentMin != 0, s_hx, rd);
}
template <int NQ, int W>
__global__ __launch_bounds__(64 * W) void k_search_flat(SearchArgs a)
{
    extern __shared__ TriC s_tr[];
    __shared__ RedLds<W> rd;
    flat_vertex<NQ, W>(a, a.cStart + blockIdx.x, s_tr, rd);
}

KernelEvents*& kernel_events()
{
    static thread_local KernelEvents* slot = nullptr;
    return slot;
}

__global__ void k_fill_int(int* __restrict__ p, int n, int v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------ learned proposal
// Vertex i of the colour class gets the proposal c = v + L_i * Y_i (network
// output in units of its mean edge length L_i; mode 2: c = v).  A 4x4x4 grid
// (candidate index i + 4j + 16k, lowest wins ties) is searched around c on a
// window `frac` times the plain first-level window, for `depthL` levels with
// the window shrunk by 2/3 per level; the centre moves to a level's best
// candidate when it is strictly better than the best quality so far, which
// starts at q(v).  Fallback rule: when the window's best candidate is lower
// than q(v), the vertex goes on the class's fallback list.  Its plain full
// search runs in a second launch before the next class (immediate policy:
// the update is the plain path's, bit for bit) or in the class's next pass,
// beside that pass's windows (deferred policy: the vertex keeps its position
// until then).  A tie at a positive quality (a level's best equals q(v) > 0) leaves
// the vertex where it is: it is the mark of a ring whose minimum is held by
// sub-tets the vertex is not part of (hexes), which no position raises.  At
// quality 0 a tie can also be the vertex's own inverted sub-tet, which the
// full window may still repair: such vertices fall back, unless hex_pinned.
struct LearnArgs {
    const float* Y;
    const float* L;
    int mode;
    int depthL;
    float frac;
    int* fallbacks;   // running total (the list launch adds the class's count)
    int* list;        // this class's fallback vertices (appended)
    int* cnt;         // their number (zero before the class)
    // network run inside the window kernel (mode 1, one 16-vertex tile per
    // block): bf16 / fp32 packed parameters and the neighbour table
    const uint16_t* wb;
    const float* w;
    const int* nbr;
    // deferred fallback (epoch non-null): the vertices appended to list in
    // iteration `it` take the full search in the class's next pass, from
    // listIn / cntIn (the previous iteration's list); epoch[vid] = the
    // iteration that deferred vid, whose window the next pass skips
    const int* listIn;
    const int* cntIn;
    int* epoch;
    int it;
    // class positions -> vertex ids through a work-item list (a batch round:
    // id in the low 30 bits); null: the class is the range a.cStart + position
    const int* items;
    // deferred fallback with the iteration in device memory (a replayed
    // graph): it = *itDev; list / cnt are the parity-0 buffers of the class,
    // parity p at + p * listStride / cntStride; this iteration appends to
    // parity it & 1 and reads the previous iteration's from the other
    const int* itDev;
    int listStride, cntStride;
    // FUSED: the ring corners' positions in the neighbour table (LearnNet.loc)
    const uint8_t* loc;
};

// the concrete iteration and list buffers of a deferred pass
__device__ __forceinline__ LearnArgs learn_resolve(const LearnArgs& la)
{
    if (!la.itDev) return la;
    LearnArgs r = la;
    const int it = *la.itDev;
    const size_t p = (size_t)(it & 1), q = p ^ 1;
    r.it = it;
    r.list = la.list + p * la.listStride;
    r.cnt = la.cnt + p * la.cntStride;
    r.listIn = la.list + q * la.listStride;
    r.cntIn = la.cnt + q * la.cntStride;
    return r;
}

// vertex id of class position ci
__device__ __forceinline__ int learned_vertex(const SearchArgs& a, const LearnArgs& la, int ci)
{
    return la.items ? (la.items[a.cStart + ci] & 0x3fffffff) : a.cStart + ci;
}

// The proposal window runs 4 vertices per wave, 16 lanes each: lane s =
// j + 4k of a group owns the candidates (i, j, k), i = 0..3, which share the
// (y, z) part of every ring element, so one LDS read of an element's
// constants serves 4 candidates (one candidate per lane would spend a third
// of the wave's cycles waiting on LDS).
// vertices per wave in the proposal window (1, 2 and 4 measured on 40^3 tet
// and 48^3 hex cubes: 4 is fastest for hexes and within noise of 2 for tets)
constexpr int kLV = 4;
// mode 3's second window, around the vertex itself, is this fraction of the
// first (the vertices whose proposal window fails are close to a converged
// position: on the seed-7 and seed-8 40^3 tet cubes 1/4 cut the fallbacks
// ~9x against a second window of the first's size, 1/6.7 and 1/5 measured
// the same, 1/2 and 1/1 slower; a power of two keeps the scaling exact)
constexpr float kMode3Window = 0.25f;

template <int G>
__device__ __forceinline__ float group_max(float v)
{
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m));
    return v;
}
template <int G>
__device__ __forceinline__ float group_min(float v)
{
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) v = fminf(v, __shfl_xor(v, m));
    return v;
}
template <int G>
__device__ __forceinline__ void group_argmax(float& best, int& idx)
{
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) {
        float ob = __shfl_xor(best, m);
        int oi = __shfl_xor(idx, m);
        bool take = (ob > best) || (ob == best && oi < idx);
        best = take ? ob : best;
        idx = take ? oi : idx;
    }
}

// The flat search of flat_vertex at nq = 8 for 4 vertices per wave.  The 64
// candidate slots of a vertex are the lanes of the one-wave layout (thread
// t = slot / 2 holds B = (t / 8 + 4, t % 8) at even and A = (t / 8, t % 8) at
// odd slots); here they are spread 4 per lane over a group of 16 lanes.  Each
// lane keeps its first best slot and the group argmax takes the larger
// score, then the lower slot: the wave argmax of the one-wave layout, so
// every decision and every result is the same.  The two slots of a thread
// share the (y) part of each element.  vid < 0: an idle group (no writes);
// s_tr: the group's cap TriC entries of LDS.
constexpr int kFlatGroups = 16;   // vertices per 256-thread block

__device__ void flat_group8(const SearchArgs& a, int vid, TriC* s_tr)
{
    constexpr int NQ = 8, G = 16;
    const int s = (int)(threadIdx.x & (G - 1));
    const bool live = vid >= 0;
    const int h0 = live ? a.heRow[vid] : 0;
    const int n = live ? a.heRow[vid + 1] - h0 : 0;
    const float x0x = live ? a.x[vid] : 0.0f, x0y = live ? a.y[vid] : 0.0f;
    float ex = -FLT_MAX, ey = -FLT_MAX, mabs = 0.0f;
    for (int e = s; e < n; e += G) {
        const int v = a.heTarget[h0 + e];
        const float dx = fabsf(a.x[v] - x0x), dy = fabsf(a.y[v] - x0y);
        ex = fmaxf(ex, dx);
        ey = fmaxf(ey, dy);
        mabs = fmaxf(mabs, fmaxf(dx, dy));
    }
    ex = group_max<G>(ex);
    ey = group_max<G>(ey);
    mabs = group_max<G>(mabs);
    const float sg = pow2_scale(mabs);
    const float isg = 1.0f / sg;
    const float is2 = (1.0f * isg) * isg;
    for (int e = s; e < n; e += G) {
        const int va = a.heTarget[h0 + e];
        const int vb = a.heTarget[h0 + (e + 1 == n ? 0 : e + 1)];
        s_tr[e] = tri_consts((a.x[va] - x0x) * sg, (a.y[va] - x0y) * sg, (a.x[vb] - x0x) * sg, (a.y[vb] - x0y) * sg);
    }
    __syncthreads();

    float qcur;
    {
        TriAcc acc = tri_acc0();
        const float px = 0.0f, py = 0.0f, Q2 = 2.0f * (py * py), t2 = 2.0f * px;
        for (int e = 0; e < n; ++e) {
            const TriC c = s_tr[e];
            const float Ty = fmaf(c.gy, py, c.C);
            const float Ry = fmaf(c.sy, py, c.K + Q2);
            tri_acc_add(acc, fmaf(c.gx, px, Ty), fmaf(px, t2 + c.sx, Ry));
        }
        qcur = tri_acc_score(acc, is2);
    }
    const float qold = qcur;
    // slots 4s + i: threads t0 = 2s (i = 0 B, 1 A) and t1 = 2s + 1 (i = 2 B, 3 A)
    const int t0 = 2 * s, t1 = 2 * s + 1;
    const int cj[2] = { t0 % NQ, t1 % NQ };
    const int ci[4] = { t0 / NQ + NQ / 2, t0 / NQ, t1 / NQ + NQ / 2, t1 / NQ };
    float cx = x0x, cy = x0y;
    float ds = a.gridScale;
    float qlast = -FLT_MAX;
    for (int d = 0; d < a.depth; ++d) {
        const float gxM = cx + ds * ex, gxm = cx - ds * ex;
        const float gyM = cy + ds * ey, gym = cy - ds * ey;
        float py[2], Q2[2], px[4], t2[4];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            py[j] = (grid_coord(a.aff, cj[j], NQ, gym, gyM) - x0y) * sg;
            Q2[j] = 2.0f * (py[j] * py[j]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            px[i] = (grid_coord(a.aff, ci[i], NQ, gxm, gxM) - x0x) * sg;
            t2[i] = 2.0f * px[i];
        }
        TriAcc acc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = tri_acc0();
        for (int e = 0; e < n; ++e) {
            const TriC c = s_tr[e];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const float Ty = fmaf(c.gy, py[j], c.C);
                const float Ry = fmaf(c.sy, py[j], c.K + Q2[j]);
#pragma unroll
                for (int i = 2 * j; i < 2 * j + 2; ++i)
                    tri_acc_add(acc[i], fmaf(c.gx, px[i], Ty), fmaf(px[i], t2[i] + c.sx, Ry));
            }
        }
        float best = -FLT_MAX;
        int bi = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float q = tri_acc_score(acc[i], is2);
            if (q > best) { best = q; bi = i; }
        }
        int wl = 4 * s + bi;
        group_argmax<G>(best, wl);
        if (qcur < best) {
            const int wt = wl >> 1;
            const int wi = wt / NQ + ((wl & 1) ? 0 : NQ / 2), wj = wt % NQ;
            cx = grid_coord(a.aff, wi, NQ, gxm, gxM);
            cy = grid_coord(a.aff, wj, NQ, gym, gyM);
            qcur = best;
        }
        qlast = best;
        ds = ds * a.shrink;
    }
    if (live && qold < qlast && s == 0) { a.xw[vid] = cx; a.yw[vid] = cy; }
}

// 16 vertices of the class per 256-thread block; cap TriC per group
__global__ __launch_bounds__(256) void k_search_flat8x4(SearchArgs a, int cap)
{
    extern __shared__ TriC s_trg[];
    const int g = (int)(threadIdx.x >> 4);
    const int i = (int)blockIdx.x * kFlatGroups + g;
    const int vid = (a.cStart + i < a.cEnd) ? a.cStart + i : -1;
    flat_group8(a, vid, s_trg + (size_t)g * cap);
}

// the flat triangle items [lo, hi) of a batch round, 16 per block
__global__ __launch_bounds__(256) void k_batch_flat8x4(SearchArgs a, const int* __restrict__ items, int lo, int hi,
                                                      int cap)
{
    extern __shared__ TriC s_trb[];
    const int g = (int)(threadIdx.x >> 4);
    const int i = lo + (int)blockIdx.x * kFlatGroups + g;
    const int vid = (i < hi) ? (items[i] & 0x3fffffff) : -1;
    flat_group8(a, vid, s_trb + (size_t)g * cap);
}

// whether the 4-per-wave flat search runs: nq 8, the block's fans fit 64 KB
// of LDS, and at least kFlat4MinVertices vertices (256 blocks: a smaller
// launch spreads over more CUs with one vertex per wave; measured on the 64^2
// grid's ~960-vertex classes, 0.112 against 0.123 ms per iteration)
constexpr int kFlat4MinVertices = 256 * kFlatGroups;
static bool flat8x4_ok(int nq, size_t ldsPerVertex, int nVertices)
{
    return nq == 8 && ldsPerVertex >= sizeof(TriC) && kFlatGroups * ldsPerVertex <= 65536 &&
           nVertices >= kFlat4MinVertices;
}


// proposal displacement of the group's vertex (class position ci, at (vx,
// vy, vz)): c = v + l * y.  FUSED: the block's 16 vertices (class positions
// ci0 + 0..15, group gb = threadIdx.x / 16) are one tile of the network, run
// here in the LDS lds (free again on return) from the inputs the caller
// gathered (ng: the row's table neighbours, nw: the wave's weights);
// otherwise l, y come from the separate inference launch's L, Y (mode 1) or
// are zero (mode 2).
template <bool FUSED>
__device__ __forceinline__ void learned_offset(const SearchArgs& a, const LearnArgs& la, int ci0, int gb, int ci,
                                               float vx, float vy, float vz, const MlpGather& ng, const MlpW& nw,
                                               unsigned char* lds, float* relOut, float& l, float& y0, float& y1,
                                               float& y2)
{
    l = 0.0f; y0 = 0.0f; y1 = 0.0f; y2 = 0.0f;
    if constexpr (FUSED) {
        MlpLds& sm = *reinterpret_cast<MlpLds*>(lds);
        const bool live = ci0 + gb < a.cEnd - a.cStart;
        mlp_features16_from(ng, vx, vy, vz, live, sm, relOut);
        mlp_layers16_w(nw, la.wb, la.w, sm);
        l = sm.L[gb]; y0 = sm.Y[gb][0]; y1 = sm.Y[gb][1]; y2 = sm.Y[gb][2];
        __syncthreads();   // the ring staging reuses the LDS
    } else if (learn_uses_net(la.mode)) {
        l = la.L[ci]; y0 = la.Y[3 * ci]; y1 = la.Y[3 * ci + 1]; y2 = la.Y[3 * ci + 2];
    }
    (void)ci; (void)vx; (void)vy; (void)vz; (void)ng; (void)nw; (void)relOut;
}

// LDS of a FUSED window block: the network's LDS, reused for the 16 rings,
// then (at learned_rel_offset) the 16 rows of table-neighbour positions
// relative to their vertex, which the ring staging reads
__host__ __device__ constexpr size_t learned_rel_offset(size_t per, int ringCap)
{
    return ((sizeof(MlpLds) > (size_t)kMlpRows * ringCap * per ? sizeof(MlpLds) : (size_t)kMlpRows * ringCap * per)
            + 15) / 16 * 16;
}
constexpr size_t kLearnedRelBytes = sizeof(float) * kMlpRows * kNbr * 3;

// append vid to the class's fallback list where need (one lane per group):
// one atomic per wave for the wave's fallbacks (the list order does not
// matter: the class's vertices are independent).  Every lane calls it.  A
// class appends at most its cap = class size vertices (its count starts at
// zero); the bound only guards the list.
__device__ __forceinline__ void fallback_append(const LearnArgs& la, bool need, int vid, int cap)
{
    const unsigned long long bal = __ballot(need);
    if (bal == 0ull) return;
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (lane == 0) base = atomicAdd(la.cnt, __popcll(bal));
    base = __shfl(base, 0);
    const int idx = base + __popcll(bal & ((1ull << lane) - 1ull));
    if (need && idx >= 0 && idx < cap) la.list[idx] = vid;
}

// The window search of the WPB * LV class vertices of block blk: LV
// vertices per wave on WPB waves; with FUSED (WPB = 4, LV = 4) the block
// first runs their 16-vertex network tile.  lds: WPB * LV rings (after the
// network's LDS).
template <int NQ, int LV, int WPB, bool FUSED>
__device__ void tet_window(const SearchArgs& a, const LearnArgs& la, int ringCap, int blk, unsigned char* s_tcl)
{
    constexpr int NL = 4, G = 64 / LV, CPL = NL * NL * NL / G;   // lanes per vertex, candidates per lane
    constexpr int VPB = WPB * LV;
    static_assert(!FUSED || VPB == kMlpRows, "one network tile per block");
    const int s = threadIdx.x % G, gb = threadIdx.x / G;
    const int nc = a.cEnd - a.cStart, ci0 = VPB * blk;
    // a group past the end of the class mirrors the class's last vertex and writes nothing
    const int ci = min(ci0 + gb, nc - 1);
    const int vid = learned_vertex(a, la, ci);
    // a vertex deferred by the previous iteration takes the full search in this launch instead
    const bool own = ci0 + gb < nc && !(la.epoch && la.epoch[vid] == la.it - 1);
    TetFrame f;
    const int r0 = a.ringRow[vid];
    f.n = a.ringRow[vid + 1] - r0;
    f.x0x = a.x[vid]; f.x0y = a.y[vid]; f.x0z = a.z[vid];
    // the network's neighbours, the lane's first PRE ring elements and the
    // network's weights are loaded into registers before the network runs
    // (their latencies overlap; no other vertex of this launch moves these
    // positions: ring and table neighbours belong to other colour classes).
    // FUSED: a ring element is its corners' positions in the neighbour table
    // (one 4-byte word), whose positions relative to v the features leave in
    // LDS; otherwise its corner positions.
    MlpGather ng{};
    if constexpr (FUSED) ng = mlp_gather16(a.x, a.y, a.z, la.nbr, vid, ci0 + gb < nc);
    constexpr int PRE = 2;
    uint32_t lw[PRE];
    float rx[PRE][3], ry[PRE][3], rz[PRE][3];
#pragma unroll
    for (int k = 0; k < PRE; ++k) {
        const int e = s + k * G;
        lw[k] = 0xffffffffu;
#pragma unroll
        for (int t = 0; t < 3; ++t) { rx[k][t] = f.x0x; ry[k][t] = f.x0y; rz[k][t] = f.x0z; }
        if (e < f.n) {
            if constexpr (FUSED) {
                lw[k] = *reinterpret_cast<const uint32_t*>(la.loc + (size_t)(r0 + e) * kRingLocStrideTet);
            } else {
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const int v = a.ringVal[(size_t)(r0 + e) * 3 + t];
                    rx[k][t] = a.x[v]; ry[k][t] = a.y[v]; rz[k][t] = a.z[v];
                }
            }
        }
    }
    MlpW nw{};
    if constexpr (FUSED) nw = mlp_load_w16(la.wb, la.w);
    float* rel = FUSED ? reinterpret_cast<float*>(s_tcl + learned_rel_offset(sizeof(TetC), ringCap)) : nullptr;
    float pl, py0, py1, py2;
    learned_offset<FUSED>(a, la, ci0, gb, ci, f.x0x, f.x0y, f.x0z, ng, nw, s_tcl, rel, pl, py0, py1, py2);
    const float* relRow = FUSED ? rel + (size_t)gb * kNbr * 3 : nullptr;
    TetC* tc = reinterpret_cast<TetC*>(s_tcl) + (size_t)gb * ringCap;
    // corner t of ring element e relative to v: from the table (w: the
    // element's packed table positions) or, for a corner the table does not
    // hold (or without a table), from the positions
    auto corner = [&](int e, int t, uint32_t w, float& dx, float& dy, float& dz) {
        const uint32_t li = (w >> (8 * t)) & 0xffu;
        if (FUSED && li != kRingLocNone) {
            dx = relRow[3 * li]; dy = relRow[3 * li + 1]; dz = relRow[3 * li + 2];
        } else {
            const int v = a.ringVal[(size_t)(r0 + e) * 3 + t];
            dx = a.x[v] - f.x0x; dy = a.y[v] - f.x0y; dz = a.z[v] - f.x0z;
        }
    };
    auto word = [&](int e) -> uint32_t {
        if constexpr (FUSED) return *reinterpret_cast<const uint32_t*>(la.loc + (size_t)(r0 + e) * kRingLocStrideTet);
        else return 0xffffffffu;
    };
    // window half-extents and frame scale (max / min: the order of the
    // elements does not matter), then the ring constants into LDS
    const bool emin = a.extentMin != 0;
    float ex = emin ? FLT_MAX : -FLT_MAX, ey = ex, ez = ex, mabs = 0.0f;
    auto extent = [&](float dxs, float dys, float dzs) {   // signed offsets from v
        const float dx = fabsf(dxs), dy = fabsf(dys), dz = fabsf(dzs);
        ex = emin ? fminf(ex, dx) : fmaxf(ex, dx);
        ey = emin ? fminf(ey, dy) : fmaxf(ey, dy);
        ez = emin ? fminf(ez, dz) : fmaxf(ez, dz);
        mabs = fmaxf(mabs, fmaxf(dx, fmaxf(dy, dz)));
    };
#pragma unroll
    for (int k = 0; k < PRE; ++k) {
        const int e = s + k * G;
        if (e < f.n) {
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                float dx, dy, dz;
                if constexpr (FUSED) corner(e, t, lw[k], dx, dy, dz);
                else { dx = rx[k][t] - f.x0x; dy = ry[k][t] - f.x0y; dz = rz[k][t] - f.x0z; }
                extent(dx, dy, dz);
            }
        }
    }
    for (int e = s + PRE * G; e < f.n; e += G) {
        const uint32_t w = word(e);
        for (int t = 0; t < 3; ++t) {
            float dx, dy, dz;
            corner(e, t, w, dx, dy, dz);
            extent(dx, dy, dz);
        }
    }
    if (emin) { ex = group_min<G>(ex); ey = group_min<G>(ey); ez = group_min<G>(ez); }
    else { ex = group_max<G>(ex); ey = group_max<G>(ey); ez = group_max<G>(ez); }
    mabs = group_max<G>(mabs);
    f.ex = ex; f.ey = ey; f.ez = ez;
    f.sg = pow2_scale(mabs);
    const float isg = 1.0f / f.sg;
    f.is3 = ((1.0f * isg) * isg) * isg;
#pragma unroll
    for (int k = 0; k < PRE; ++k) {
        const int e = s + k * G;
        if (e < f.n) {
            f3 p[3];
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                float dx, dy, dz;
                if constexpr (FUSED) corner(e, t, lw[k], dx, dy, dz);
                else { dx = rx[k][t] - f.x0x; dy = ry[k][t] - f.x0y; dz = rz[k][t] - f.x0z; }
                p[t] = { dx * f.sg, dy * f.sg, dz * f.sg };
            }
            tc[e] = tet_consts(p[0], p[1], p[2]);
        }
    }
    for (int e = s + PRE * G; e < f.n; e += G) {
        const uint32_t w = word(e);
        f3 p[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            float dx, dy, dz;
            corner(e, t, w, dx, dy, dz);
            p[t] = { dx * f.sg, dy * f.sg, dz * f.sg };
        }
        tc[e] = tet_consts(p[0], p[1], p[2]);
    }
    __syncthreads();

    float cx = f.x0x, cy = f.x0y, cz = f.x0z;
    if (learn_uses_net(la.mode)) { cx = f.x0x + pl * py0; cy = f.x0y + pl * py1; cz = f.x0z + pl * py2; }
    // q0, the start position's score, is accumulated beside the candidates
    // of the first level (at p = 0: det = D, L = K exactly, the terms
    // tet_score adds, in the same ring order)
    float q0 = 0.0f, qbest = 0.0f;
    bool haveQ0 = false;
    bool moved = false, tie = false;
    const float affL = 1.0f / (float)(NL - 1);
    const int lj = s % NL, lk = (s / NL) % NL, i0 = (s / (NL * NL)) * CPL;
    // depthL levels of the 4x4x4 window around (cx, cy, cz); the branch
    // conditions are uniform over the vertex's lane group
    auto levels = [&](float scale) {
        float ds = (la.frac * scale) * a.gridScale;
        for (int d = 0; d < la.depthL; ++d) {
            const float gxM = cx + ds * f.ex, gxm = cx - ds * f.ex;
            const float gyM = cy + ds * f.ey, gym = cy - ds * f.ey;
            const float gzM = cz + ds * f.ez, gzm = cz - ds * f.ez;
            const float py = (grid_coord(affL, lj, NL, gym, gyM) - f.x0y) * f.sg;
            const float pz = (grid_coord(affL, lk, NL, gzm, gzM) - f.x0z) * f.sg;
            float px[CPL], t3[CPL];
    #pragma unroll
            for (int i = 0; i < CPL; ++i) {
                px[i] = (grid_coord(affL, i0 + i, NL, gxm, gxM) - f.x0x) * f.sg;
                t3[i] = 3.0f * px[i];
            }
            const float Qyz = 3.0f * fmaf(pz, pz, py * py);
            TetAcc acc[CPL];
    #pragma unroll
            for (int i = 0; i < CPL; ++i) acc[i] = tet_acc0();
            // one pass over the ring; the first level run also accumulates q0
            // (later levels and mode 3's window skip it)
            auto ring = [&](auto withQ0) {
                constexpr bool WQ = decltype(withQ0)::value;
                TetAcc acc0 = tet_acc0();
                // two elements per iteration (the minima fold into v_min3)
    #pragma unroll 2
                for (int e = 0; e < f.n; ++e) {
                    const TetC c = tc[e];
                    const float Tyz = fmaf(c.nny, py, fmaf(c.nnz, pz, c.D));
                    const float Ryz = fmaf(c.m2sy, py, fmaf(c.m2sz, pz, c.K + Qyz));
    #pragma unroll
                    for (int i = 0; i < CPL; ++i)
                        tet_acc_add(acc[i], fmaf(c.nnx, px[i], Tyz), fmaf(px[i], t3[i] + c.m2sx, Ryz));
                    if constexpr (WQ) tet_acc_add(acc0, c.D, c.K);
                }
                if constexpr (WQ) { q0 = tet_acc_score(acc0, f.is3); qbest = q0; haveQ0 = true; }
            };
            if (!haveQ0) ring(std::true_type{});
            else ring(std::false_type{});
            float best = -FLT_MAX;
            int bi = 0;
    #pragma unroll
            for (int i = 0; i < CPL; ++i) {
                float q = tet_acc_score(acc[i], f.is3);
                if (q > best) { best = q; bi = i; }
            }
            int cand = i0 + bi + NL * (s % (NL * NL));
            group_argmax<G>(best, cand);
            tie = tie || (best == q0 && q0 > 0.0f);
            if (qbest < best) {
                cx = grid_coord(affL, cand % NL, NL, gxm, gxM);
                cy = grid_coord(affL, (cand / NL) % NL, NL, gym, gyM);
                cz = grid_coord(affL, cand / (NL * NL), NL, gzm, gzM);
                qbest = best;
                moved = true;
            }
            ds = ds * (2.0f / (float)(NL - 1));
        }
    };
    levels(1.0f);
    // mode 3: where the network's window finds nothing better, a window a
    // quarter of its size around the vertex itself before the full search
    if (la.mode == 3 && !moved && !tie) {
        cx = f.x0x; cy = f.x0y; cz = f.x0z;
        levels(kMode3Window);
    }
    if (moved && own && s == 0) { a.xw[vid] = cx; a.yw[vid] = cy; a.zw[vid] = cz; }
    const bool need = !moved && !tie && own && s == 0;
    fallback_append(la, need, vid, nc);
    if (need && la.epoch) la.epoch[vid] = la.it;
}

// C8: the eight-corner hex metric (a.hexCorner8) as a template argument: a
// runtime branch around the principal-axes slot held its LDS reads back in
// this loop (36.5 -> 38.1 us per 48^3 class)
template <int NQ, int LV, int WPB, bool FUSED, bool C8>
__device__ void hex_window(const SearchArgs& a, const LearnArgs& la, int ringCap, int blk, unsigned char* s_hxl)
{
    constexpr int NL = 4, G = 64 / LV, CPL = NL * NL * NL / G;
    constexpr int VPB = WPB * LV;
    static_assert(!FUSED || VPB == kMlpRows, "one network tile per block");
    const int s = threadIdx.x % G, gb = threadIdx.x / G;
    const int nc = a.cEnd - a.cStart, ci0 = VPB * blk;
    const int ci = min(ci0 + gb, nc - 1);
    const int vid = learned_vertex(a, la, ci);
    const bool own = ci0 + gb < nc && !(la.epoch && la.epoch[vid] == la.it - 1);
    HexFrame f;
    const int r0 = a.ringRow[vid];
    f.n = a.ringRow[vid + 1] - r0;
    f.x0x = a.x[vid]; f.x0y = a.y[vid]; f.x0z = a.z[vid];
    // the network's inputs (table neighbours, weights) and the packed table
    // positions of the lane's first ring element are loaded before it runs
    MlpGather ng{};
    MlpW nw{};
    uint2 lw0 = make_uint2(0xffffffffu, 0xffffffffu);
    if constexpr (FUSED) {
        ng = mlp_gather16(a.x, a.y, a.z, la.nbr, vid, ci0 + gb < nc);
        if (s < f.n) lw0 = *reinterpret_cast<const uint2*>(la.loc + (size_t)(r0 + s) * kRingLocStrideHex);
        nw = mlp_load_w16(la.wb, la.w);
    }
    float* rel = FUSED ? reinterpret_cast<float*>(s_hxl + learned_rel_offset(sizeof(HexLds), ringCap)) : nullptr;
    float pl, py0, py1, py2;
    learned_offset<FUSED>(a, la, ci0, gb, ci, f.x0x, f.x0y, f.x0z, ng, nw, s_hxl, rel, pl, py0, py1, py2);
    const float* relRow = FUSED ? rel + (size_t)gb * kNbr * 3 : nullptr;
    HexLds* hx = reinterpret_cast<HexLds*>(s_hxl) + (size_t)gb * ringCap;
    // corner t of ring element e relative to v: from the table (w: the
    // element's packed table positions) or, for a corner the table does not
    // hold (or without a table), from the positions
    auto word = [&](int e) -> uint2 {
        if constexpr (FUSED)
            return e == s ? lw0 : *reinterpret_cast<const uint2*>(la.loc + (size_t)(r0 + e) * kRingLocStrideHex);
        else return make_uint2(0xffffffffu, 0xffffffffu);
    };
    auto corner = [&](int e, int t, uint2 w, float& dx, float& dy, float& dz) {
        const uint32_t li = ((t < 4 ? w.x : w.y) >> (8 * (t & 3))) & 0xffu;
        if (FUSED && li != kRingLocNone) {
            dx = relRow[3 * li]; dy = relRow[3 * li + 1]; dz = relRow[3 * li + 2];
        } else {
            const int v = a.ringVal[(size_t)(r0 + e) * 7 + t];
            dx = a.x[v] - f.x0x; dy = a.y[v] - f.x0y; dz = a.z[v] - f.x0z;
        }
    };
    // window half-extents (group_extents's rule) and frame scale
    {
        const bool emin = a.extentMin != 0;
        float ex = emin ? FLT_MAX : -FLT_MAX, ey = ex, ez = ex, mabs = 0.0f;
        for (int e = s; e < f.n; e += G) {
            const uint2 w = word(e);
            for (int t = 0; t < 7; ++t) {
                float dx, dy, dz;
                corner(e, t, w, dx, dy, dz);
                dx = fabsf(dx); dy = fabsf(dy); dz = fabsf(dz);
                ex = emin ? fminf(ex, dx) : fmaxf(ex, dx);
                ey = emin ? fminf(ey, dy) : fmaxf(ey, dy);
                ez = emin ? fminf(ez, dz) : fmaxf(ez, dz);
                mabs = fmaxf(mabs, fmaxf(dx, fmaxf(dy, dz)));
            }
        }
        if (emin) { ex = group_min<G>(ex); ey = group_min<G>(ey); ez = group_min<G>(ez); }
        else { ex = group_max<G>(ex); ey = group_max<G>(ey); ez = group_max<G>(ez); }
        f.ex = ex; f.ey = ey; f.ez = ez;
        f.sg = pow2_scale(group_max<G>(mabs));
    }
    const float s2 = f.sg * f.sg, s3 = s2 * f.sg;
    f.epsD = kHexEps * s3; f.epsF = kHexEps * s2;
    for (int e = s; e < f.n; e += G) {
        const uint2 w = word(e);
        f3 r[8];
        r[0] = { 0.0f, 0.0f, 0.0f };
        for (int t = 0; t < 7; ++t) {
            float dx, dy, dz;
            corner(e, t, w, dx, dy, dz);
            r[t + 1] = { dx * f.sg, dy * f.sg, dz * f.sg };
        }
        HexLds L;
        hex_consts(r, L.cd, L.cF, L.hp);
        hx[e] = L;
    }
    __syncthreads();
    f.acc0 = hex_acc0();
    for (int e = 0; e < f.n; ++e)
#pragma unroll
        for (int sIdx = 0; sIdx < 4; ++sIdx) hex_acc_add(f.acc0, hx[e].cd[sIdx], hx[e].cF[sIdx]);

    float cx = f.x0x, cy = f.x0y, cz = f.x0z;
    if (learn_uses_net(la.mode)) { cx = f.x0x + pl * py0; cy = f.x0y + pl * py1; cz = f.x0z + pl * py2; }
    // q0, the start position's score, is accumulated beside the candidates
    // of the first level (at p = 0: det = D, |J|^2 = K exactly, the terms
    // hex_score adds, in the same order)
    float q0 = 0.0f, qbest = 0.0f;
    bool haveQ0 = false;
    bool moved = false, tie = false;
    const float affL = 1.0f / (float)(NL - 1);
    const int lj = s % NL, lk = (s / NL) % NL, i0 = (s / (NL * NL)) * CPL;
    auto levels = [&](float scale) {
        float ds = (la.frac * scale) * a.gridScale;
        for (int d = 0; d < la.depthL; ++d) {
            const float gxM = cx + ds * f.ex, gxm = cx - ds * f.ex;
            const float gyM = cy + ds * f.ey, gym = cy - ds * f.ey;
            const float gzM = cz + ds * f.ez, gzm = cz - ds * f.ez;
            const float py = (grid_coord(affL, lj, NL, gym, gyM) - f.x0y) * f.sg;
            const float pz = (grid_coord(affL, lk, NL, gzm, gzM) - f.x0z) * f.sg;
            float px[CPL], t3[CPL];
    #pragma unroll
            for (int i = 0; i < CPL; ++i) {
                px[i] = (grid_coord(affL, i0 + i, NL, gxm, gxM) - f.x0x) * f.sg;
                t3[i] = 3.0f * px[i];
            }
            const float Q1 = fmaf(pz, pz, py * py), Q3 = 3.0f * Q1;
            HexAcc acc[CPL];
    #pragma unroll
            for (int i = 0; i < CPL; ++i) acc[i] = f.acc0;
            // the first level run also accumulates q0 (later levels and mode
            // 3's window skip it)
            auto ring = [&](auto withQ0) {
                constexpr bool WQ = decltype(withQ0)::value;
                HexAcc accO = f.acc0;
                for (int e = 0; e < f.n; ++e) {
    #pragma unroll
                    for (int sIdx = 0; sIdx < (C8 ? 4 : 5); ++sIdx) {
                        const HexP h = hx[e].hp[sIdx];
                        const bool a3 = (sIdx == 1 || sIdx == 4);
                        const float Tyz = fmaf(h.gy, py, fmaf(h.gz, pz, h.D));
                        const float Ryz = fmaf(h.sy, py, fmaf(h.sz, pz, h.K + (a3 ? Q3 : Q1)));
    #pragma unroll
                        for (int i = 0; i < CPL; ++i)
                            hex_acc_add(acc[i], fmaf(h.gx, px[i], Tyz), fmaf(px[i], (a3 ? t3[i] : px[i]) + h.sx, Ryz));
                        if constexpr (WQ) hex_acc_add(accO, h.D, h.K);
                    }
                }
                if constexpr (WQ) { q0 = hex_acc_score(accO, f.epsD, f.epsF); qbest = q0; haveQ0 = true; }
            };
            if (!haveQ0) ring(std::true_type{});
            else ring(std::false_type{});
            float best = -FLT_MAX;
            int bi = 0;
    #pragma unroll
            for (int i = 0; i < CPL; ++i) {
                float q = hex_acc_score(acc[i], f.epsD, f.epsF);
                if (q > best) { best = q; bi = i; }
            }
            int cand = i0 + bi + NL * (s % (NL * NL));
            group_argmax<G>(best, cand);
            tie = tie || (best == q0 && q0 > 0.0f);
            if (qbest < best) {
                cx = grid_coord(affL, cand % NL, NL, gxm, gxM);
                cy = grid_coord(affL, (cand / NL) % NL, NL, gym, gyM);
                cz = grid_coord(affL, cand / (NL * NL), NL, gzm, gzM);
                qbest = best;
                moved = true;
            }
            ds = ds * (2.0f / (float)(NL - 1));
        }
    };
    levels(1.0f);
    if (la.mode == 3 && !moved && !tie && !hex_pinned(f)) {
        cx = f.x0x; cy = f.x0y; cz = f.x0z;
        levels(kMode3Window);
    }
    if (moved && own && s == 0) { a.xw[vid] = cx; a.yw[vid] = cy; a.zw[vid] = cz; }
    const bool need = !moved && !tie && !hex_pinned(f) && own && s == 0;
    fallback_append(la, need, vid, nc);
    if (need && la.epoch) la.epoch[vid] = la.it;
}

// immediate fallback: one block = LV vertices on one wave, or (FUSED) 4 waves
// of LV = 4 vertices that first run their network tile; the fallbacks go to a
// second launch (k_search_list_*)
template <int NQ, int LV, bool FUSED>
__global__ __launch_bounds__(FUSED ? 256 : 64) void k_search_tet_learned(SearchArgs a, LearnArgs la, int ringCap)
{
    extern __shared__ __align__(16) unsigned char s_tcl[];
    tet_window<NQ, LV, FUSED ? 4 : 1, FUSED>(a, la, ringCap, blockIdx.x, s_tcl);
}
template <int NQ, int LV, bool FUSED, bool C8>
__global__ __launch_bounds__(FUSED ? 256 : 64) void k_search_hex_learned(SearchArgs a, LearnArgs la, int ringCap)
{
    extern __shared__ __align__(16) unsigned char s_hxl[];
    hex_window<NQ, LV, FUSED ? 4 : 1, FUSED, C8>(a, la, ringCap, blockIdx.x, s_hxl);
}

// deferred fallback: the first gridDim.x - nbList blocks take the windows of
// 16 class vertices each (the vertices deferred last iteration excepted),
// the last nbList blocks the full searches (4 waves per vertex) of the
// vertices the previous iteration deferred; both are vertices of one colour
// class, so the two kinds run side by side.  The windows come first so that
// the dispatcher starts them before the list blocks, most of which find no
// vertex (list blocks first: 0.27 against 0.226 ms per learned 40^3 tet
// iteration).
template <int NQ, bool FUSED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void k_search_tet_learned_def(SearchArgs a, LearnArgs la0, int ringCap, int nbList)
{
    extern __shared__ __align__(16) unsigned char s_tcd[];
    __shared__ RedLds<4> rd;
    const LearnArgs la = learn_resolve(la0);
    const int nbWin = (int)gridDim.x - nbList;
    if ((int)blockIdx.x >= nbWin) {
        const int lb = (int)blockIdx.x - nbWin;
        const int n = min(*la.cntIn, a.cEnd - a.cStart);
        if (lb == 0 && threadIdx.x == 0) atomicAdd(la.fallbacks, n);
        for (int i = lb; i < n; i += nbList) {
            tet_vertex<NQ, 4>(a, la.listIn[i], a.extentMin != 0, reinterpret_cast<TetC*>(s_tcd), rd);
            __syncthreads();   // the next vertex restages the LDS ring
        }
        return;
    }
    tet_window<NQ, 4, 4, FUSED>(a, la, ringCap, (int)blockIdx.x, s_tcd);
}
// 4 waves per SIMD (<= 128 registers): a 48^3 class (~13k vertices, 3.3k
// waves) then fits the chip's wave slots in one round
template <int NQ, bool FUSED, bool C8>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void k_search_hex_learned_def(SearchArgs a, LearnArgs la0, int ringCap, int nbList)
{
    extern __shared__ __align__(16) unsigned char s_hxd[];
    __shared__ RedLds<4> rd;
    const LearnArgs la = learn_resolve(la0);
    const int nbWin = (int)gridDim.x - nbList;
    if ((int)blockIdx.x >= nbWin) {
        const int lb = (int)blockIdx.x - nbWin;
        const int n = min(*la.cntIn, a.cEnd - a.cStart);
        if (lb == 0 && threadIdx.x == 0) atomicAdd(la.fallbacks, n);
        for (int i = lb; i < n; i += nbList) {
            hex_vertex<NQ, 4>(a, la.listIn[i], a.extentMin != 0, reinterpret_cast<HexLds*>(s_hxd), rd);
            __syncthreads();
        }
        return;
    }
    hex_window<NQ, 4, 4, FUSED, C8>(a, la, ringCap, (int)blockIdx.x, s_hxd);
}

// plain search of the vertices list[0 .. *cnt) (the fallbacks of a class),
// one vertex per block of W waves at a time, blocks striding over the list
// (the grid is sized for a typical fallback share, the length is only known
// on the device)
template <int NQ, int W>
__global__ __launch_bounds__(64 * W) void k_search_list_tet(SearchArgs a, const int* __restrict__ list,
                                                           const int* __restrict__ cnt, int* fallbacks)
{
    extern __shared__ TetC s_tcf[];
    __shared__ RedLds<W> rd;
    const int n = min(*cnt, a.cEnd - a.cStart);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(fallbacks, n);
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        tet_vertex<NQ, W>(a, list[i], a.extentMin != 0, s_tcf, rd);
        __syncthreads();   // the next vertex restages the LDS ring
    }
}
template <int NQ, int W>
__global__ __launch_bounds__(64 * W) void k_search_list_hex(SearchArgs a, const int* __restrict__ list,
                                                           const int* __restrict__ cnt, int* fallbacks)
{
    extern __shared__ HexLds s_hxf[];
    __shared__ RedLds<W> rd;
    const int n = min(*cnt, a.cEnd - a.cStart);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(fallbacks, n);
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        hex_vertex<NQ, W>(a, list[i], a.extentMin != 0, s_hxf, rd);
        __syncthreads();   // the next vertex restages the LDS ring
    }
}
template <int NQ, int W>
static void list_nqw(int kind, const SearchArgs& a, const int* list, const int* cnt, int* fallbacks, int grid,
                     size_t lds, hipStream_t s)
{
    if (kind == VSDMO_TET) launch_kernel(k_search_list_tet<NQ, W>, grid, 64 * W, lds, s, a, list, cnt, fallbacks);
    else launch_kernel(k_search_list_hex<NQ, W>, grid, 64 * W, lds, s, a, list, cnt, fallbacks);
}

template <int NQ, int LV, bool FUSED>
static void learned_nq_lv(int kind, const SearchArgs& a, const LearnArgs& la, int ringCap, int nb, size_t lds, hipStream_t s)
{
    constexpr int T = FUSED ? 256 : 64;
    if (kind == VSDMO_TET) launch_kernel(k_search_tet_learned<NQ, LV, FUSED>, nb, T, lds, s, a, la, ringCap);
    else if (a.hexCorner8) launch_kernel(k_search_hex_learned<NQ, LV, FUSED, true>, nb, T, lds, s, a, la, ringCap);
    else launch_kernel(k_search_hex_learned<NQ, LV, FUSED, false>, nb, T, lds, s, a, la, ringCap);
}

static size_t learned_per(int kind) { return (kind == VSDMO_TET) ? sizeof(TetC) : sizeof(HexLds); }

// largest ring the window kernel's LDS (kLV rings per block, 64 KB) holds
int learned_ring_cap(int kind) { return (int)((64 * 1024) / (kLV * learned_per(kind))); }

// 16 rings per block in 64 KB of LDS
static bool rings16_fit(int kind, int ringCap)
{
    return (size_t)kMlpRows * std::max(ringCap, 1) * learned_per(kind) <= 64 * 1024;
}

// LDS of a FUSED window block: the network's (reused for the 16 rings) and
// the table-neighbour positions
static size_t fused_lds(int kind, int ringCap)
{
    return learned_rel_offset(learned_per(kind), std::max(ringCap, 1)) + kLearnedRelBytes;
}

// the fused kernel holds 16 rings and the table-neighbour positions per
// block in the same 64 KB
bool learned_fused_ok(int kind, int ringCap)
{
    return fused_lds(kind, ringCap) <= 64 * 1024;
}

// the deferred policy's full searches run 4 waves per vertex, and its blocks
// hold 16 rings in LDS with or without the network
bool learned_deferred_ok(int kind, int nq, int ringCap)
{
    return (nq == 4 || nq == 8) && rings16_fit(kind, ringCap);
}


template <int NQ, bool FUSED>
static void learned_def_nq(int kind, const SearchArgs& a, const LearnArgs& la, int ringCap, int nbList, int nb,
                           size_t lds, hipStream_t s)
{
    if (kind == VSDMO_TET) launch_kernel(k_search_tet_learned_def<NQ, FUSED>, nbList + nb, 256, lds, s, a, la, ringCap, nbList);
    else if (a.hexCorner8)
        launch_kernel(k_search_hex_learned_def<NQ, FUSED, true>, nbList + nb, 256, lds, s, a, la, ringCap, nbList);
    else launch_kernel(k_search_hex_learned_def<NQ, FUSED, false>, nbList + nb, 256, lds, s, a, la, ringCap, nbList);
}

void launch_search_learned(int kind, int nq, const SearchArgs& a, const LearnNet& net, int mode, int depthL,
                           float frac, int* fallbacks, int* list, int* cnt, int ringCap, const LearnDefer* def,
                           hipStream_t s)
{
    const int nc = a.cEnd - a.cStart;
    if (nc <= 0) return;
    ringCap = std::max(ringCap, 1);
    const size_t per = learned_per(kind);
    const bool fused = learn_uses_net(mode) && net.wb && net.loc && learned_fused_ok(kind, ringCap);
    LearnArgs la{ net.Y, net.L, mode, depthL, frac, fallbacks, list, cnt, net.wb, net.w, net.nbr,
                  nullptr, nullptr, nullptr, 0, net.items };
    la.loc = net.loc;
    if (def) {
        // one launch: this iteration's windows, 16 vertices per block, beside
        // the previous iteration's fallbacks (list blocks sized for an eighth
        // of the class, they loop when more were deferred: a sixteenth
        // measured the same, a thirty-second slower on the 40^3 tet cube)
        la.listIn = def->listIn; la.cntIn = def->cntIn; la.epoch = def->epoch; la.it = def->it;
        la.itDev = def->itDev; la.listStride = def->listStride; la.cntStride = def->cntStride;
        const int nbList = std::max(1, std::min(nc, std::max(64, nc / 8)));
        const int nb = (nc + kMlpRows - 1) / kMlpRows;
        const size_t lds = fused ? fused_lds(kind, ringCap) : (size_t)kMlpRows * ringCap * per;
        switch (nq) {   // the full searches use 4 waves: nq 4 and 8 (learned_deferred_ok)
        case 4: fused ? learned_def_nq<4, true>(kind, a, la, ringCap, nbList, nb, lds, s)
                      : learned_def_nq<4, false>(kind, a, la, ringCap, nbList, nb, lds, s); break;
        case 8: fused ? learned_def_nq<8, true>(kind, a, la, ringCap, nbList, nb, lds, s)
                      : learned_def_nq<8, false>(kind, a, la, ringCap, nbList, nb, lds, s); break;
        default: break;
        }
        return;
    }
    const int vpb = fused ? kMlpRows : kLV;
    const int nb = (nc + vpb - 1) / vpb;
    const size_t lds = fused ? fused_lds(kind, ringCap) : (size_t)kLV * ringCap * per;
#define VSDMO_LEARN_CASE(Q)                                                                           \
    case Q:                                                                                          \
        if (fused) learned_nq_lv<Q, kLV, true>(kind, a, la, ringCap, nb, lds, s);                     \
        else learned_nq_lv<Q, kLV, false>(kind, a, la, ringCap, nb, lds, s);                          \
        break;
    switch (nq) {
        VSDMO_LEARN_CASE(2)
        VSDMO_LEARN_CASE(4)
        VSDMO_LEARN_CASE(6)
        VSDMO_LEARN_CASE(8)
    default: break;
    }
#undef VSDMO_LEARN_CASE
    // second launch: the plain search of the fallbacks, W waves per vertex,
    // a quarter of the class's blocks (blocks loop when more fall back)
    const size_t lds1 = (size_t)ringCap * per;
    const int g1 = std::max(1, std::min(nc, std::max(256, nc / 4)));
    switch (nq) {
    case 2: list_nqw<2, 2>(kind, a, list, cnt, fallbacks, g1, lds1, s); break;
    case 4: list_nqw<4, 4>(kind, a, list, cnt, fallbacks, g1, lds1, s); break;
    case 6: list_nqw<6, 2>(kind, a, list, cnt, fallbacks, g1, lds1, s); break;
    case 8:
        // the fallback lists are short (a few per cent of the class): latency
        // bound, 4 waves per vertex (1, 2, 4 and 8 measured on the 40^3 tet
        // and 48^3 hex cubes)
        list_nqw<8, 4>(kind, a, list, cnt, fallbacks, g1, lds1, s);
        break;
    default: break;
    }
}

// ------------------------------------------------------- split launches
// One colour class in one launch of 256-thread blocks: the first nHead
// vertices one per wave (the W = 1 body, its ring staged in the wave's own
// quarter of the LDS), the remaining ones one per block with 4 (or 2) waves
// per vertex.  The end of a launch, where few waves are left on the GPU,
// then drains in finer grains.  Every body gives the same result, so the
// split changes no bits.  All 4 waves of a block pass the same barriers: a
// head wave without a vertex passes the single barrier of a W = 1 search.
template <int NQ, int WT>
__global__ __launch_bounds__(256, 5) void k_search_tet_split(SearchArgs a, int nHead, int ringCap)
{
    extern __shared__ TetC s_tcs[];
    __shared__ RedLds<1> rd1;
    __shared__ RedLds<WT> rdt;
    const bool emin = a.extentMin != 0;
    const int headBlocks = (nHead + 3) >> 2;
    if ((int)blockIdx.x < headBlocks) {
        const int w = threadIdx.x >> 6;
        const int i = 4 * (int)blockIdx.x + w;
        if (i < nHead) tet_vertex<NQ, 1>(a, a.cStart + i, emin, s_tcs + (size_t)w * ringCap, rd1);
        else __syncthreads();   // the one barrier of a W = 1 search
    } else {
        static_assert(WT == 4, "the tail blocks use all 4 waves");
        tet_vertex<NQ, WT>(a, a.cStart + nHead + ((int)blockIdx.x - headBlocks), emin, s_tcs, rdt);
    }
}
template <int NQ, int WT>
__global__ __launch_bounds__(256) void k_search_hex_split(SearchArgs a, int nHead, int ringCap)
{
    extern __shared__ HexLds s_hxs[];
    __shared__ RedLds<1> rd1;
    __shared__ RedLds<WT> rdt;
    const bool emin = a.extentMin != 0;
    const int headBlocks = (nHead + 3) >> 2;
    if ((int)blockIdx.x < headBlocks) {
        const int w = threadIdx.x >> 6;
        const int i = 4 * (int)blockIdx.x + w;
        if (i < nHead) hex_vertex<NQ, 1>(a, a.cStart + i, emin, s_hxs + (size_t)w * ringCap, rd1);
        else __syncthreads();
    } else {
        static_assert(WT == 4, "the tail blocks use all 4 waves");
        hex_vertex<NQ, WT>(a, a.cStart + nHead + ((int)blockIdx.x - headBlocks), emin, s_hxs, rdt);
    }
}

// ------------------------------------------------------- batched sweep
// Persistent kernel of one colour round of a batch of meshes: blocks take
// work items from a queue (one atomic counter per round) until it is empty.
// An item is a global vertex id with its element type in bits 30-31
// (0 tet, 1 hex, 2 flat tri).  Vertices of one round belong to one colour
// class of their mesh, so the order in which they are taken does not change
// any result.
// Items [lo, mid) are tet / hex searches, taken one at a time from the
// round's counter (their cost varies with the ring); items [mid, hi) are flat
// triangle searches, ~100x cheaper and uniform, statically strided over the
// blocks so that they cost no atomics.
template <int NQ, int W>
__device__ __forceinline__ void batch_item(const SearchArgs& a, const int* rowHex, const int* valHex, int code,
                                           float4* s_dyn, RedLds<W>& rd)
{
    const int vid = code & 0x3fffffff, type = (code >> 30) & 3;
    if (type == 0) tet_vertex<NQ, W>(a, vid, false, reinterpret_cast<TetC*>(s_dyn), rd);
    else if (type == 1) {
        SearchArgs ah = a;
        ah.ringRow = rowHex;
        ah.ringVal = valHex;
        hex_vertex<NQ, W>(ah, vid, true, reinterpret_cast<HexLds*>(s_dyn), rd);
    }
    else flat_vertex<NQ, W>(a, vid, reinterpret_cast<TriC*>(s_dyn), rd);
}

template <int NQ, int W>
__global__ __launch_bounds__(64 * W) void k_batch_round(SearchArgs a, const int* __restrict__ rowHex,
                                                       const int* __restrict__ valHex, const int* __restrict__ items,
                                                       int lo, int mid, int hi, int* __restrict__ ctr)
{
    extern __shared__ float4 s_dyn[];
    __shared__ RedLds<W> rd;
    __shared__ int s_item;
    for (;;) {
        __syncthreads();   // previous item's LDS reads are done
        if (threadIdx.x == 0) s_item = atomicAdd(ctr, 1);
        __syncthreads();
        const int it = lo + s_item;
        if (it >= mid) break;
        batch_item<NQ, W>(a, rowHex, valHex, items[it], s_dyn, rd);
    }
    for (int it = mid + (int)blockIdx.x; it < hi; it += (int)gridDim.x) {
        __syncthreads();
        batch_item<NQ, W>(a, rowHex, valHex, items[it], s_dyn, rd);
    }
}

// ------------------------------------------------------------ qualities
// element quality = ring score of the element seen from its vertex 0 at p'=0
__device__ float elem_quality(const float* __restrict__ X, const float* __restrict__ Y, const float* __restrict__ Z,
                              const int* c, int elem, bool corner8)
{
    const float x0x = X[c[0]], x0y = Y[c[0]], x0z = Z[c[0]];
    if (elem == VSDMO_TET) {
        float m = 0.0f;
        for (int i = 1; i < 4; ++i) {
            m = fmaxf(m, fabsf(X[c[i]] - x0x));
            m = fmaxf(m, fabsf(Y[c[i]] - x0y));
            m = fmaxf(m, fabsf(Z[c[i]] - x0z));
        }
        const float sg = pow2_scale(m);
        const float isg = 1.0f / sg;
        f3 p[3];
        for (int i = 0; i < 3; ++i) {
            int v = c[i + 1];
            p[i] = { (X[v] - x0x) * sg, (Y[v] - x0y) * sg, (Z[v] - x0z) * sg };
        }
        TetC tc = tet_consts(p[0], p[1], p[2]);
        const float px = 0.0f, py = 0.0f, pz = 0.0f;
        const float Qyz = 3.0f * fmaf(pz, pz, py * py), t3 = 3.0f * px;
        TetAcc acc = tet_acc0();
        float Tyz = fmaf(tc.nny, py, fmaf(tc.nnz, pz, tc.D));
        float Ryz = fmaf(tc.m2sy, py, fmaf(tc.m2sz, pz, tc.K + Qyz));
        tet_acc_add(acc, fmaf(tc.nnx, px, Tyz), fmaf(px, t3 + tc.m2sx, Ryz));
        return tet_acc_score(acc, ((1.0f * isg) * isg) * isg);
    }
    if (elem == VSDMO_HEX) {
        float m = 0.0f;
        for (int i = 1; i < 8; ++i) {
            m = fmaxf(m, fabsf(X[c[i]] - x0x));
            m = fmaxf(m, fabsf(Y[c[i]] - x0y));
            m = fmaxf(m, fabsf(Z[c[i]] - x0z));
        }
        const float sg = pow2_scale(m);
        const float s2 = sg * sg, s3 = s2 * sg;
        f3 r[8];
        r[0] = { 0.0f, 0.0f, 0.0f };
        for (int i = 1; i < 8; ++i) {
            int v = c[i];
            r[i] = { (X[v] - x0x) * sg, (Y[v] - x0y) * sg, (Z[v] - x0z) * sg };
        }
        float cd[4], cF[4];
        HexP hp[5];
        hex_consts(r, cd, cF, hp);
        HexAcc acc = hex_acc0();
        for (int sIdx = 0; sIdx < 4; ++sIdx) hex_acc_add(acc, cd[sIdx], cF[sIdx]);
        const float px = 0.0f, py = 0.0f, pz = 0.0f;
        const float Q1 = fmaf(pz, pz, py * py), Q3 = 3.0f * Q1, t3 = 3.0f * px;
#pragma unroll
        for (int sIdx = 0; sIdx < 5; ++sIdx) {   // unrolled: hp[] stays in registers
            if (sIdx == 4 && corner8) break;
            const bool a3 = (sIdx == 1 || sIdx == 4);
            float Tyz = fmaf(hp[sIdx].gy, py, fmaf(hp[sIdx].gz, pz, hp[sIdx].D));
            float Ryz = fmaf(hp[sIdx].sy, py, fmaf(hp[sIdx].sz, pz, hp[sIdx].K + (a3 ? Q3 : Q1)));
            hex_acc_add(acc, fmaf(hp[sIdx].gx, px, Tyz), fmaf(px, (a3 ? t3 : px) + hp[sIdx].sx, Ryz));
        }
        return hex_acc_score(acc, kHexEps * s3, kHexEps * s2);
    }
    if (elem == VSDMO_QUAD)
        return mr_quad3(f3{ x0x, x0y, x0z }, f3{ X[c[1]], Y[c[1]], Z[c[1]] }, f3{ X[c[2]], Y[c[2]], Z[c[2]] },
                        f3{ X[c[3]], Y[c[3]], Z[c[3]] }, nullptr);
    if (elem == kElemTri3)
        return mr_tri3(f3{ x0x, x0y, x0z }, f3{ X[c[1]], Y[c[1]], Z[c[1]] }, f3{ X[c[2]], Y[c[2]], Z[c[2]] }, nullptr);
    float m = 0.0f;
    for (int i = 1; i < 3; ++i) {
        m = fmaxf(m, fabsf(X[c[i]] - x0x));
        m = fmaxf(m, fabsf(Y[c[i]] - x0y));
    }
    const float sg = pow2_scale(m);
    const float isg = 1.0f / sg;
    TriC tc = tri_consts((X[c[1]] - x0x) * sg, (Y[c[1]] - x0y) * sg, (X[c[2]] - x0x) * sg, (Y[c[2]] - x0y) * sg);
    const float px = 0.0f, py = 0.0f;
    const float Q2 = 2.0f * (py * py), t2 = 2.0f * px;
    TriAcc acc = tri_acc0();
    float Ty = fmaf(tc.gy, py, tc.C);
    float Ry = fmaf(tc.sy, py, tc.K + Q2);
    tri_acc_add(acc, fmaf(tc.gx, px, Ty), fmaf(px, t2 + tc.sx, Ry));
    return tri_acc_score(acc, (1.0f * isg) * isg);
}

// elements per thread of the quality kernels: up to 4 (strided by 256), as
// long as the launch keeps at least 512 blocks (two per CU)
int quality_per_thread(long long nElements)
{
    int per = 4;
    while (per > 1 && nElements < (long long)per * 256 * 512) per >>= 1;
    return per;
}

// min and FP64 sum over the 256 threads of a block: a butterfly within each
// wave (lane 0's value: the order of its additions is fixed), then the 4 waves
// in order; the result is every thread's.  One call per kernel (static LDS).
__device__ __forceinline__ void block_minsum256(float& m, double& s)
{
    __shared__ float wm[4];
    __shared__ double ws[4];
#pragma unroll
    for (int k = 32; k >= 1; k >>= 1) {
        m = fminf(m, __shfl_xor(m, k));
        s = s + __shfl_xor(s, k);
    }
    const int w = (int)(threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0) { wm[w] = m; ws[w] = s; }
    __syncthreads();
    m = fminf(fminf(wm[0], wm[1]), fminf(wm[2], wm[3]));
    s = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

// the qualities of the block's 256 * per elements (per per thread, strided
// by 256), written to q, and their min and FP64 sum in a fixed order
// (per-block partials for k_reduce_final)
__global__ __launch_bounds__(256) void k_elem_quality(QualityArgs a, float* __restrict__ pmin,
                                                      double* __restrict__ psum, int per)
{
    float v = FLT_MAX;
    double sv = 0.0;
    for (int k = 0; k < per; ++k) {
        const int e = blockIdx.x * 256 * per + k * 256 + (int)threadIdx.x;
        if (e < a.nE) {
            const float q = elem_quality(a.x, a.y, a.z, a.cells + (size_t)e * a.npe, a.elem, a.hexCorner8 != 0);
            a.q[e] = q;
            v = fminf(v, q);
            sv = sv + (double)q;
        }
    }
    block_minsum256(v, sv);
    if (threadIdx.x == 0) { pmin[blockIdx.x] = v; psum[blockIdx.x] = sv; }
}

// batch qualities: one block of 256 per chunk of at most 256 * a.per
// elements of one mesh; per-chunk min and FP64 sum in a fixed order
__global__ __launch_bounds__(256) void k_batch_quality(BatchQArgs a)
{
    const int c = blockIdx.x;
    const int e0 = a.chunkElem0[c], ne = a.chunkElem0[c + 1] - e0;
    const int elem = a.chunkType[c];
    const int npe = elem == VSDMO_TET ? 4 : elem == VSDMO_HEX ? 8 : 3;
    float m = FLT_MAX;
    double s = 0.0;
    for (int k = 0; k < a.per; ++k) {
        const int t = k * 256 + (int)threadIdx.x;
        if (t < ne) {
            const float q = elem_quality(a.x, a.y, a.z, a.cells + a.chunkCell0[c] + (size_t)t * npe, elem,
                                         a.hexCorner8 != 0);
            m = fminf(m, q);
            s = s + (double)q;
        }
    }
    block_minsum256(m, s);
    if (threadIdx.x == 0) { a.cmin[c] = m; a.csum[c] = s; }
}

// per-mesh min / FP64 sum over the mesh's chunks: one block per mesh,
// strided partials then a fixed tree (the order depends only on the chunk
// count); the minimum is also written to history row *hctr
__global__ void k_batch_mesh_reduce(BatchQArgs a)
{
    __shared__ float smin[256];
    __shared__ double ssum[256];
    const int b = blockIdx.x;
    float m = FLT_MAX;
    double s = 0.0;
    for (int c = a.meshChunk0[b] + (int)threadIdx.x; c < a.meshChunk0[b + 1]; c += 256) {
        m = fminf(m, a.cmin[c]);
        s = s + a.csum[c];
    }
    smin[threadIdx.x] = m;
    ssum[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + k]);
            ssum[threadIdx.x] = ssum[threadIdx.x] + ssum[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.mmin[b] = smin[0];
        a.msum[b] = ssum[0];
        const int row = *a.hctr;
        if (row < a.hcap) a.hist[(size_t)row * a.nMeshes + b] = smin[0];
    }
}

// advances the history row after every mesh of the iteration wrote it
__global__ void k_batch_history_next(int* hctr)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *hctr = *hctr + 1;
}

// one block of 256 threads: strided partial sums then a fixed tree (the order
// of additions depends only on n, so the FP64 sum is reproducible)
__global__ void k_reduce_final(const float* __restrict__ pmin, const double* __restrict__ psum, int n,
                               float* __restrict__ omin, double* __restrict__ osum)
{
    __shared__ float smin[256];
    __shared__ double ssum[256];
    float m = FLT_MAX;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) { m = fminf(m, pmin[i]); s = s + psum[i]; }
    smin[threadIdx.x] = m;
    ssum[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + k]);
            ssum[threadIdx.x] = ssum[threadIdx.x] + ssum[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { *omin = smin[0]; *osum = ssum[0]; }
}

__global__ void k_histogram(const float* __restrict__ q, int n, int ncols, int* __restrict__ bins)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = fminf(0.9999f, q[i]);
    long idx = (long)(v * (float)ncols);
    if (idx < 0) idx = 0;
    atomicAdd(&bins[idx], 1);
}

// ------------------------------------------------------------- launchers
template <int NQ, int W>
static void launch_nqw(int kind, const SearchArgs& a, int nBlocks, size_t lds, hipStream_t s)
{
    if (kind == VSDMO_TET) launch_kernel(k_search_tet<NQ, W>, nBlocks, 64 * W, lds, s, a);
    else if (kind == VSDMO_HEX) launch_kernel(k_search_hex<NQ, W>, nBlocks, 64 * W, lds, s, a);
    else launch_kernel(k_search_flat<NQ, W>, nBlocks, 64 * W, lds, s, a);
}
template <int NQ>
static void launch_nq(int kind, int w, const SearchArgs& a, int nBlocks, size_t lds, hipStream_t s)
{
    if (w == 2 && NQ % 2 == 0) launch_nqw<NQ, (NQ % 2 == 0 ? 2 : 1)>(kind, a, nBlocks, lds, s);
    else if (w == 4 && NQ % 4 == 0) launch_nqw<NQ, (NQ % 4 == 0 ? 4 : 1)>(kind, a, nBlocks, lds, s);
    else launch_nqw<NQ, 1>(kind, a, nBlocks, lds, s);
}
// Non-persistent variant of a round: one block per item, scheduled by the
// hardware dispatcher (no queue, no atomics), one launch per element type
// (0 tet, 1 hex, 2 flat tri) with the search body of that type alone, so
// that each keeps the registers and LDS of its own body.  The tet body asks
// for the plain tet search's occupancy (VSDMO_TET_MIN_WAVES waves per SIMD).
template <int NQ, int W, int TYPE>
__global__ __launch_bounds__(64 * W, TYPE == 0 ? VSDMO_TET_MIN_WAVES : 1) void k_batch_typed(SearchArgs a,
                                                                                        const int* __restrict__ rowHex,
                                                       const int* __restrict__ valHex, const int* __restrict__ items,
                                                       int lo)
{
    extern __shared__ float4 s_dyn[];
    __shared__ RedLds<W> rd;
    const int vid = items[lo + blockIdx.x] & 0x3fffffff;
    if constexpr (TYPE == 0) {
        tet_vertex<NQ, W>(a, vid, false, reinterpret_cast<TetC*>(s_dyn), rd);
    } else if constexpr (TYPE == 1) {
        SearchArgs ah = a;
        ah.ringRow = rowHex;
        ah.ringVal = valHex;
        hex_vertex<NQ, W>(ah, vid, true, reinterpret_cast<HexLds*>(s_dyn), rd);
    } else {
        flat_vertex<NQ, W>(a, vid, reinterpret_cast<TriC*>(s_dyn), rd);
    }
}

template <int NQ, int W>
static void batch_typed_nqw(const SearchArgs& a, const int* rowHex, const int* valHex, const int* items, int lo, int hi,
                            int type, size_t lds, hipStream_t s)
{
    const unsigned n = (unsigned)(hi - lo);
    if (type == 0) k_batch_typed<NQ, W, 0><<<n, 64 * W, lds, s>>>(a, rowHex, valHex, items, lo);
    else if (type == 1) k_batch_typed<NQ, W, 1><<<n, 64 * W, lds, s>>>(a, rowHex, valHex, items, lo);
    else k_batch_typed<NQ, W, 2><<<n, 64 * W, lds, s>>>(a, rowHex, valHex, items, lo);
}
template <int NQ>
static void batch_typed_nq(int w, const SearchArgs& a, const int* rowHex, const int* valHex, const int* items, int lo,
                           int hi, int type, size_t lds, hipStream_t s)
{
    if (w == 2 && NQ % 2 == 0) batch_typed_nqw<NQ, (NQ % 2 == 0 ? 2 : 1)>(a, rowHex, valHex, items, lo, hi, type, lds, s);
    else if (w == 4 && NQ % 4 == 0)
        batch_typed_nqw<NQ, (NQ % 4 == 0 ? 4 : 1)>(a, rowHex, valHex, items, lo, hi, type, lds, s);
    else batch_typed_nqw<NQ, 1>(a, rowHex, valHex, items, lo, hi, type, lds, s);
}

struct BatchLaunch {
    SearchArgs a;
    const int* rowHex;
    const int* valHex;
    const int* items;
    int lo, mid, hi;
    int* ctr;
    int grid;
    size_t lds;
    hipStream_t s;
};
template <int NQ, int W>
static void batch_nqw(const BatchLaunch& L)
{
    k_batch_round<NQ, W><<<L.grid, 64 * W, L.lds, L.s>>>(L.a, L.rowHex, L.valHex, L.items, L.lo, L.mid, L.hi, L.ctr);
}
template <int NQ, int W>
static int batch_occ(size_t lds)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch_round<NQ, W>, 64 * W, lds) != hipSuccess) nb = 1;
    return nb;
}
template <int NQ>
static void batch_nq(int w, const BatchLaunch& L)
{
    if (w == 2 && NQ % 2 == 0) batch_nqw<NQ, (NQ % 2 == 0 ? 2 : 1)>(L);
    else if (w == 4 && NQ % 4 == 0) batch_nqw<NQ, (NQ % 4 == 0 ? 4 : 1)>(L);
    else batch_nqw<NQ, 1>(L);
}
template <int NQ>
static int batch_occ_nq(int w, size_t lds)
{
    if (w == 2 && NQ % 2 == 0) return batch_occ<NQ, (NQ % 2 == 0 ? 2 : 1)>(lds);
    if (w == 4 && NQ % 4 == 0) return batch_occ<NQ, (NQ % 4 == 0 ? 4 : 1)>(lds);
    return batch_occ<NQ, 1>(lds);
}

bool waves_supported(int nq, int w)
{
    return w == 0 || w == 1 || (w == 2 && nq % 2 == 0) || (w == 4 && nq % 4 == 0);
}

template <int NQ, int WT>
static void split_nq(int kind, const SearchArgs& a, int nHead, int nTail, int ringCap, hipStream_t s)
{
    const int grid = (nHead + 3) / 4 + nTail;
    if (grid <= 0) return;
    if (kind == VSDMO_TET)
        launch_kernel(k_search_tet_split<NQ, WT>, grid, 256, sizeof(TetC) * 4 * (size_t)ringCap, s, a, nHead, ringCap);
    else
        launch_kernel(k_search_hex_split<NQ, WT>, grid, 256, sizeof(HexLds) * 4 * (size_t)ringCap, s, a, nHead, ringCap);
}

bool split_supported(int kind, int nq, int ringCap)
{
    if (kind != VSDMO_TET && kind != VSDMO_HEX) return false;
    if (nq != 4 && nq != 8) return false;   // the tail needs 4 waves per vertex (nq / 4 candidates per lane)
    const size_t per = (kind == VSDMO_TET) ? sizeof(TetC) : sizeof(HexLds);
    return per * 4 * (size_t)std::max(ringCap, 1) <= 65536;
}

void launch_search_split(int kind, int nq, const SearchArgs& a, int tail, int ringCap, hipStream_t s)
{
    const int n = a.cEnd - a.cStart;
    if (n <= 0) return;
    const int nTail = std::min(n, std::max(tail, 0));
    const int nHead = n - nTail;
    if (nq == 4) split_nq<4, 4>(kind, a, nHead, nTail, ringCap, s);
    else if (nq == 8) split_nq<8, 4>(kind, a, nHead, nTail, ringCap, s);
}

int batch_round_blocks_per_cu(int nq, int w, size_t lds)
{
    switch (nq) {
    case 2: return batch_occ_nq<2>(w, lds);
    case 4: return batch_occ_nq<4>(w, lds);
    case 6: return batch_occ_nq<6>(w, lds);
    default: return batch_occ_nq<8>(w, lds);
    }
}

void launch_batch_round(int nq, int w, const SearchArgs& a, const int* rowHex, const int* valHex, const int* items,
                        int lo, int mid, int hi, int* ctr, int grid, size_t lds, hipStream_t s)
{
    if (hi <= lo || grid <= 0) return;
    BatchLaunch L{ a, rowHex, valHex, items, lo, mid, hi, ctr, grid, lds, s };
    switch (nq) {
    case 2: batch_nq<2>(w, L); break;
    case 4: batch_nq<4>(w, L); break;
    case 6: batch_nq<6>(w, L); break;
    case 8: batch_nq<8>(w, L); break;
    default: break;
    }
}

void launch_batch_typed(int nq, int w, const SearchArgs& a, const int* rowHex, const int* valHex, const int* items,
                        int lo, int hi, int type, size_t lds, hipStream_t s)
{
    if (hi <= lo) return;
    if (type == 2 && w == 1 && flat8x4_ok(nq, lds, hi - lo)) {
        const int groups = (hi - lo + kFlatGroups - 1) / kFlatGroups;
        k_batch_flat8x4<<<groups, 256, kFlatGroups * lds, s>>>(a, items, lo, hi, (int)(lds / sizeof(TriC)));
        return;
    }
    switch (nq) {
    case 2: batch_typed_nq<2>(w, a, rowHex, valHex, items, lo, hi, type, lds, s); break;
    case 4: batch_typed_nq<4>(w, a, rowHex, valHex, items, lo, hi, type, lds, s); break;
    case 6: batch_typed_nq<6>(w, a, rowHex, valHex, items, lo, hi, type, lds, s); break;
    case 8: batch_typed_nq<8>(w, a, rowHex, valHex, items, lo, hi, type, lds, s); break;
    default: break;
    }
}

void launch_fill_int(int* p, int n, int v, hipStream_t s)
{
    if (n > 0) k_fill_int<<<(n + 255) / 256, 256, 0, s>>>(p, n, v);
}

// zero the n counters of parity (*itDev & 1): cnt + parity * n
__global__ void k_parity_zero(int* __restrict__ cnt, int n, const int* __restrict__ itDev)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cnt[(size_t)(*itDev & 1) * n + i] = 0;
}

__global__ void k_add_int(int* __restrict__ p, int v)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *p += v;
}

void launch_parity_zero(int* cnt, int n, const int* itDev, hipStream_t s)
{
    if (n > 0) k_parity_zero<<<(n + 255) / 256, 256, 0, s>>>(cnt, n, itDev);
}

void launch_add_int(int* p, int v, hipStream_t s) { k_add_int<<<1, 64, 0, s>>>(p, v); }

void launch_batch_quality(const BatchQArgs& a, hipStream_t s)
{
    if (a.nChunks > 0) k_batch_quality<<<a.nChunks, 256, 0, s>>>(a);
    if (a.nMeshes > 0) k_batch_mesh_reduce<<<a.nMeshes, 256, 0, s>>>(a);
    k_batch_history_next<<<1, 64, 0, s>>>(a.hctr);
}

size_t batch_lds_bytes(int maxRingTet, int maxRingHex, int maxHETri)
{
    size_t b = std::max(sizeof(TetC) * (size_t)maxRingTet, sizeof(HexLds) * (size_t)maxRingHex);
    b = std::max(b, sizeof(TriC) * (size_t)maxHETri);
    return std::max(b, (size_t)16);
}

size_t search_lds_bytes(int kind, int maxRing)
{
    size_t per = (kind == VSDMO_TET) ? sizeof(TetC) : (kind == VSDMO_HEX) ? sizeof(HexLds) : sizeof(TriC);
    return per * (size_t)std::max(maxRing, 1);
}

void launch_search(int kind, int nq, int w, const SearchArgs& a, size_t lds, hipStream_t s)
{
    int nBlocks = a.cEnd - a.cStart;
    if (nBlocks <= 0) return;
    if (kind == VSDMO_TRI && w == 1 && flat8x4_ok(nq, lds, nBlocks)) {
        const int groups = (nBlocks + kFlatGroups - 1) / kFlatGroups;
        launch_kernel(k_search_flat8x4, groups, 256, kFlatGroups * lds, s, a, (int)(lds / sizeof(TriC)));
        return;
    }
    switch (nq) {
    case 2: launch_nq<2>(kind, w, a, nBlocks, lds, s); break;
    case 4: launch_nq<4>(kind, w, a, nBlocks, lds, s); break;
    case 6: launch_nq<6>(kind, w, a, nBlocks, lds, s); break;
    case 8: launch_nq<8>(kind, w, a, nBlocks, lds, s); break;
    default: break;
    }
}

void launch_quality(const QualityArgs& a, float* pmin, double* psum, float* omin, double* osum, hipStream_t s)
{
    const int per = quality_per_thread(a.nE);
    const int nb = (a.nE + 256 * per - 1) / (256 * per);
    k_elem_quality<<<nb, 256, 0, s>>>(a, pmin, psum, per);
    k_reduce_final<<<1, 256, 0, s>>>(pmin, psum, nb, omin, osum);
}

void launch_histogram(const float* q, int n, int ncols, int* bins, hipStream_t s)
{
    k_histogram<<<(n + 255) / 256, 256, 0, s>>>(q, n, ncols, bins);
}

}  // namespace vsdmo
