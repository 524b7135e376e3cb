// surface.hip -- surface-vertex DMO on the GPU: local quadric fits (own 3x3
// LU), feature-sector fits, 1D feature-curve fits, nearest-neighbour walk,
// and the 2D surface / 1D feature-curve searches with projection onto the
// blended local surfaces.
//
// Reference (VSDMO/src/dmo/...): surfaceP/Surface.h:18-169, Surface.cu:11-121,
// surfaceP/SurfaceFeature.h:79-106, SurfaceFeature.cu:12-223,
// Surface1D.h:13-114, Surface1D.cu:7-103, SurfaceEstimation.h:122-197,
// DMOCommon.cu:11-43, DMOCommon.h:130-165, DMOHexImplCub.h:141-489,
// DMOTetImplCub.h:141-494; surface meshes DMOTriImplCub.h:25-380,
// DMOQuadImplCub.h:23-404, DMOQualityTri.h:257-275, DMOQualityQuad.h:30-50,
// MeshGPUCommon.cu:413-434 (vertex normals), Laplace.cu:8-41.  Expressions follow the reference literally
// (dot = (x*x' + y*y') + z*z', normalize divides by sqrtf(dot)) with
// contraction off, so a host restatement of the same expressions reproduces
// every value.
#include <float.h>

#include "dmo_kernels.h"
#include "vsdmo_internal.h"
#include "vsdmo_numerics.h"

namespace vsdmo {

constexpr int kMaxNeigh = 40;
constexpr int kTableSize = 20;
constexpr float kSEps = 1.19209e-07f;

struct w3 { float x, y, z; };
__device__ __forceinline__ w3 W(const float* p) { return { p[0], p[1], p[2] }; }
__device__ __forceinline__ w3 WP(const float* x, const float* y, const float* z, int v) { return { x[v], y[v], z[v] }; }
__device__ __forceinline__ w3 wsub(w3 a, w3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
__device__ __forceinline__ w3 wadd(w3 a, w3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
__device__ __forceinline__ w3 wmul(float s, w3 a) { return { s * a.x, s * a.y, s * a.z }; }
__device__ __forceinline__ float wdot(w3 a, w3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ w3 wcross(w3 a, w3 b)
{
    return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
__device__ __forceinline__ float wnorm(w3 a) { return sqrtf((a.x * a.x + a.y * a.y) + a.z * a.z); }
__device__ __forceinline__ w3 wnormalize(w3 a)
{
    float l = wnorm(a);
    return { a.x / l, a.y / l, a.z / l };
}

// local surface record: p0 (3), e0 e1 e2 (9), a b c (3), numPoints (1)
struct LSurf { w3 p0, e[3]; float a, b, c, np; };
struct LSurf1 { w3 p0, e[2]; float a; };
static_assert(sizeof(LSurf) == 16 * sizeof(float), "LSurf layout");
static_assert(sizeof(LSurf1) == 10 * sizeof(float), "LSurf1 layout");

__device__ __forceinline__ float poly(const LSurf& s, float u, float v) { return (s.a * u * u + s.b * u * v) + s.c * v * v; }
__device__ __forceinline__ w3 l2w(const LSurf& s, float u, float v, float w)
{
    return wadd(wadd(wadd(s.p0, wmul(u, s.e[0])), wmul(v, s.e[1])), wmul(w, s.e[2]));
}

// 3x3 LU with partial pivoting (first maximal |pivot|) and solve; column-major A
__device__ void lu_solve3(const float* Acm, const float* rhs, float* x)
{
    float A[3][3], b[3];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) A[i][j] = Acm[i + 3 * j];
        b[i] = rhs[i];
    }
    int piv[3] = { 0, 1, 2 };
    for (int j = 0; j < 3; ++j) {
        int p = j;
        float mx = fabsf(A[j][j]);
        for (int i = j + 1; i < 3; ++i)
            if (fabsf(A[i][j]) > mx) { mx = fabsf(A[i][j]); p = i; }
        if (p != j) {
            for (int k = 0; k < 3; ++k) { float t = A[j][k]; A[j][k] = A[p][k]; A[p][k] = t; }
            int t = piv[j]; piv[j] = piv[p]; piv[p] = t;
        }
        for (int i = j + 1; i < 3; ++i) {
            A[i][j] = A[i][j] / A[j][j];
            for (int k = j + 1; k < 3; ++k) A[i][k] = A[i][k] - A[i][j] * A[j][k];
        }
    }
    float y[3];
    for (int i = 0; i < 3; ++i) y[i] = b[piv[i]];
    for (int i = 1; i < 3; ++i)
        for (int k = 0; k < i; ++k) y[i] = y[i] - A[i][k] * y[k];
    for (int i = 2; i >= 0; --i) {
        for (int k = i + 1; k < 3; ++k) y[i] = y[i] - A[i][k] * y[k];
        y[i] = y[i] / A[i][i];
    }
    x[0] = y[0]; x[1] = y[1]; x[2] = y[2];
}

__device__ void frame3(w3 n, w3 r, w3* e0, w3* e1, w3* e2)
{
    if (wnorm(wcross(n, r)) < 10.0f * kSEps) r = { n.x, n.z, -n.y };
    w3 t1 = wcross(n, r);
    w3 t2 = wcross(n, t1);
    *e2 = wnormalize(n);
    *e0 = wnormalize(t1);
    *e1 = wnormalize(t2);
}

__device__ void fit_quadric(LSurf& s, const w3* pts, int np)
{
    float M[9] = { 0, 0, 0, 0, 0, 0, 0, 0, 0 }, R[3] = { 0, 0, 0 };
    for (int i = 0; i < np; ++i) {
        w3 p = wsub(pts[i], s.p0);
        float u = wdot(s.e[0], p), v = wdot(s.e[1], p), w = wdot(s.e[2], p);
        R[0] += u * u * w; R[1] += u * v * w; R[2] += v * v * w;
        M[0] += u * u * u * u; M[1] += u * u * u * v; M[4] += u * u * v * v;
        M[2] += u * u * v * v; M[5] += u * v * v * v; M[8] += v * v * v * v;
    }
    M[3] = M[1]; M[6] = M[2]; M[7] = M[5];
    float abc[3];
    lu_solve3(M, R, abc);
    s.a = abc[0]; s.b = abc[1]; s.c = abc[2];
    s.np = (float)np;
    if (np <= 4) { s.a = 0.0f; s.b = 0.0f; s.c = 0.0f; }
    if (s.a != s.a) s.a = 0.0f;
    if (s.b != s.b) s.b = 0.0f;
    if (s.c != s.c) s.c = 0.0f;
}

// ------------------------------------------------------------- init kernels
__global__ void k_surf_face_normals(SurfArgs a)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nFaces) return;
    const int* F = a.faces + (size_t)f * a.fpe;
    w3 p0 = WP(a.x, a.y, a.z, F[0]), p1 = WP(a.x, a.y, a.z, F[1]), p2 = WP(a.x, a.y, a.z, F[a.fpe == 3 ? 2 : 3]);
    w3 n = wnormalize(wcross(wsub(p1, p0), wsub(p2, p0)));
    a.fnorm[3 * f] = n.x; a.fnorm[3 * f + 1] = n.y; a.fnorm[3 * f + 2] = n.z;
}

__global__ void k_surf_vertex_init(SurfArgs a)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.nS) return;
    // vertex normal: normalized sum of incident face normals (CCW order)
    w3 n = { 0.0f, 0.0f, 0.0f };
    for (int h = a.heRow[v]; h < a.heRow[v + 1]; ++h)
        if (a.heFace[h] != -1) n = wadd(n, W(a.fnorm + 3 * a.heFace[h]));
    n = wnormalize(n);
    a.vnorm[3 * v] = n.x; a.vnorm[3 * v + 1] = n.y; a.vnorm[3 * v + 2] = n.z;
    a.pinit[3 * v] = a.x[v]; a.pinit[3 * v + 1] = a.y[v]; a.pinit[3 * v + 2] = a.z[v];
    a.nn[v] = v;
    w3 pv = WP(a.x, a.y, a.z, v);
    const bool feat = a.isFeature[v] != 0;
    // 1D curve fit of feature vertices (Surface1D.cu:48-91)
    if (feat) {
        int nb[kMaxNeigh + 1], np = 0, vf = -1;
        nb[np++] = v;
        for (int h = a.heRow[v]; h < a.heRow[v + 1]; ++h) {
            int d = a.heTarget[h];
            if (a.isFeature[d] && np < kMaxNeigh) { nb[np++] = d; vf = d; }
        }
        if (vf >= 0) {
            w3 nrm = a.flat ? w3{ 0.0f, 0.0f, 1.0f } : n;
            w3 r = wsub(WP(a.x, a.y, a.z, vf), pv);
            if (wnorm(wcross(nrm, r)) < 10.0f * kSEps) r = { nrm.x, nrm.z, -nrm.y };
            w3 t1 = wcross(nrm, r), t2 = wcross(nrm, t1);
            nrm = wnormalize(nrm);
            t2 = wnormalize(t2);
            LSurf1 q;
            q.p0 = pv; q.e[0] = t2; q.e[1] = nrm;
            float ATA = 0.0f, ATb = 0.0f;
            for (int i = 0; i < np; ++i) {
                w3 p = wsub(WP(a.x, a.y, a.z, nb[i]), q.p0);
                float u = wdot(q.e[0], p), vv = wdot(q.e[1], p);
                ATb += u * u * vv;
                ATA += u * u * u * u;
            }
            q.a = ATb / ATA;
            reinterpret_cast<LSurf1*>(a.ls1)[v] = q;
        }
        // 1D eligibility: two feature neighbours, cos <= cos(160 deg)
        int nf = 0;
        w3 pp[2] = { { 0, 0, 0 }, { 0, 0, 0 } };
        for (int h = a.heRow[v]; h < a.heRow[v + 1]; ++h) {
            int d = a.heTarget[h];
            if (a.isFeature[d]) { if (nf < 2) pp[nf] = WP(a.x, a.y, a.z, d); ++nf; }
        }
        uint8_t ok = 0;
        if (nf == 2) {
            w3 ea = wsub(pp[0], pv), eb = wsub(pp[1], pv);
            float c = wdot(ea, eb) / sqrtf(wdot(ea, ea) * wdot(eb, eb));
            if (c <= a.cos160) ok = 1;
        }
        a.opt1d[v] = ok;
    } else {
        a.opt1d[v] = 0;
    }
    // quadric fit over {v} and its one-ring (Surface.cu:11-105); ring 2 adds
    // the one-ring's CCW targets not yet taken, neighbour by neighbour
    w3 pts[kMaxFitPoints + 1];
    int np = 0;
    pts[np++] = pv;
    for (int h = a.heRow[v]; h < a.heRow[v + 1] && np < kMaxNeigh; ++h) pts[np++] = WP(a.x, a.y, a.z, a.heTarget[h]);
    if (a.ring2) {
        int ids[kMaxFitPoints];
        int ni = 0;
        ids[ni++] = v;
        for (int h = a.heRow[v]; h < a.heRow[v + 1] && ni < np; ++h) ids[ni++] = a.heTarget[h];
        const int n1 = ni;
        bool over = false;
        for (int k = 1; k < n1 && !over; ++k) {
            const int d = ids[k];
            for (int h = a.heRow[d]; h < a.heRow[d + 1]; ++h) {
                const int t = a.heTarget[h];
                bool seen = false;
                for (int q = 0; q < ni && !seen; ++q) seen = ids[q] == t;
                if (seen) continue;
                if (ni == kMaxFitPoints) { over = true; break; }
                ids[ni++] = t;
                pts[np++] = WP(a.x, a.y, a.z, t);
            }
        }
        if (over) atomicOr(a.err, kSurfErrFitPoints);
    }
    if (np == 5) { pts[np] = wmul(0.5f, wadd(pts[2], pts[3])); ++np; }
    LSurf q;
    q.p0 = pts[0];
    int first = a.heTarget[a.heRow[v]];
    frame3(n, wsub(WP(a.x, a.y, a.z, first), pv), &q.e[0], &q.e[1], &q.e[2]);
    fit_quadric(q, pts, np);
    reinterpret_cast<LSurf*>(a.ls)[v] = q;
}

// number of sector surfaces of each feature vertex (k_countLocalSurfacesFeature)
__global__ void k_sector_count(SurfArgs a, int* cnt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nFeature) return;
    int v = a.nFree + i, c = 0;
    for (int h = a.heRow[v] + 1; h < a.heRow[v + 1]; ++h) if (a.isFeature[a.heTarget[h]]) ++c;
    if (!a.isB1D[v]) ++c;
    cnt[i] = c;
}

// Feature sectors of feature vertex v (the local surfaces of its smooth
// pieces between feature edges; reference SurfaceFeature.cu:134-223).  The
// CCW fan t_0 .. t_{n-1} of v is cut at every target t_k, k >= 1, that is a
// feature vertex.  Sector s holds v, its opening target (t_0 for sector 0,
// the previous cut target after that), the targets after it and its closing
// cut target.  On a closed fan (v not on a surface boundary) one more sector
// wraps from the last cut target round to t_0.  The table maps fan position
// k (< kTableSize) to the sector holding t_k after its opening target;
// position 0 maps to sector 0.  Targets after the last cut of an open fan
// belong to no sector: reported as an error.  Lists hold at most kMaxNeigh
// vertices.
__global__ void k_sector_lists(SurfArgs a, int* secN, int* secCnt)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.nFeature) return;
    const int v = a.nFree + f;
    const int* fan = a.heTarget + a.heRow[v];
    const int n = a.heRow[v + 1] - a.heRow[v];
    const int first = a.frow[f], count = a.frow[f + 1] - first;
    int* tab = a.table + (size_t)kTableSize * f;
    for (int k = 0; k < kTableSize; ++k) tab[k] = -1;
    tab[0] = 0;
    // emit sector sec: v, fan[open], fan[open + 1 .. close) and closing target c
    auto emit = [&](int sec, int open, int close, int c) {
        int* list = secN + (size_t)kMaxNeigh * (first + sec);
        int len = 0;
        list[len++] = v;
        list[len++] = fan[open];
        for (int k = open + 1; k < close && len < kMaxNeigh; ++k) list[len++] = fan[k];
        if (len < kMaxNeigh) list[len++] = c;
        secCnt[first + sec] = len;
    };
    int sec = 0, open = 0;
    for (int k = 1; k < n; ++k) {
        if (sec >= count) {                       // past the last cut of an open fan
            atomicOr(a.err, kSurfErrUnsectored);
            break;
        }
        if (k < kTableSize) tab[k] = sec;
        if (a.isFeature[fan[k]]) {
            emit(sec, open, k, fan[k]);
            ++sec;
            open = k;
        }
    }
    if (!a.isB1D[v] && sec < count) emit(sec, open, n, fan[0]);
}

__global__ void k_sector_fit(SurfArgs a, const int* secN, const int* secCnt, int nSec)
{
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nSec) return;
    const int* L = secN + (size_t)kMaxNeigh * q;
    int np = secCnt[q];
    if (np < 2) return;   // not reached for closed surfaces
    int v = L[0];
    w3 pts[kMaxNeigh + 1];
    for (int i = 0; i < np; ++i) pts[i] = WP(a.x, a.y, a.z, L[i]);
    if (np == 5) { pts[np] = wmul(0.5f, wadd(pts[2], pts[3])); ++np; }
    w3 n = { 0.0f, 0.0f, 0.0f };
    for (int h = a.heRow[v]; h < a.heRow[v + 1]; ++h)
        if (a.heTarget[h] == L[1] && a.heFace[h] >= 0) n = W(a.fnorm + 3 * a.heFace[h]);
    LSurf s;
    s.p0 = pts[0];
    frame3(n, wsub(WP(a.x, a.y, a.z, L[1]), WP(a.x, a.y, a.z, v)), &s.e[0], &s.e[1], &s.e[2]);
    fit_quadric(s, pts, np);
    reinterpret_cast<LSurf*>(a.fs)[q] = s;
}

// nearest-neighbour walk over the initial surface positions within the
// vertex's feature class (k_updateNearestNeighbor)
__global__ void k_update_nn(SurfArgs a)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.nS) return;
    w3 p = WP(a.x, a.y, a.z, v);
    int vh = a.nn[v];
    w3 d0 = wsub(p, W(a.pinit + 3 * vh));
    float sq = wdot(d0, d0);
    const uint8_t fv = a.isFeature[v];
    int vnew;
    do {
        vnew = vh;
        int c = vh;
        for (int h = a.heRow[c]; h < a.heRow[c + 1]; ++h) {
            int dst = a.heTarget[h];
            if ((fv != 0) != (a.isFeature[dst] != 0)) continue;
            w3 dd = wsub(p, W(a.pinit + 3 * dst));
            float sd = wdot(dd, dd);
            if (sd < sq) { sq = sd; vh = dst; }
        }
    } while (vnew != vh);
    a.nn[v] = vnew;
}

// ------------------------------------------------------------- projection
// localSurfaceEstimationWithFeature: evaluate the nearest neighbour's quadric
// at (u,v), then blend with its projections onto each one-ring neighbour's
// surface (sector surfaces for feature neighbours), weights 1/(d^2 + eps).
// Two passes over the neighbours (weights, then the weighted sum); the
// arithmetic is the single-pass one.  Everything but (u,v) -- the fits, the
// nearest neighbours' initial positions and the sector choice -- is the same
// for all candidates of a search, so the searches stage it once in LDS
// (Proj2 / Proj1) and blend from there.

// the surface a one-ring neighbour projects onto: the fit of its nearest
// neighbour nn, or for a feature nn the sector fit of the halfedge of nn whose
// target (snapshot position) is closest to vidPos
__device__ LSurf neighbour_fit(const SurfArgs& a, int nn, w3 vidPos)
{
    const LSurf* ls = reinterpret_cast<const LSurf*>(a.ls);
    if (!a.isFeature[nn]) return ls[nn];
    const int off = nn - a.nFree;
    float md = 3.40282347e+38f;
    int arg = 0, idx = 0;
    for (int h = a.heRow[nn]; h < a.heRow[nn + 1]; ++h, ++idx) {
        w3 t = wsub(WP(a.sx, a.sy, a.sz, a.heTarget[h]), vidPos);
        float dd = wdot(t, t);
        if (dd < md) { md = dd; arg = idx; }
    }
    int sid = (arg < kTableSize) ? a.table[kTableSize * off + arg] : -1;
    const int nsec = a.frow[off + 1] - a.frow[off];
    if (sid < 0 || sid >= nsec) {
        // the reference asserts sid != -1 (SurfaceEstimation.h:172); here the
        // optimiser reports VSDMO_TOPOLOGY after the pass
        atomicOr(a.err, kSurfErrNoSector);
        sid = 0;
    }
    return nsec > 0 ? reinterpret_cast<const LSurf*>(a.fs)[a.frow[off] + sid] : ls[nn];
}

// the blend of the 2D projection over a source of L (the NN's fit), q0 (its
// initial position), nh, qn(i) (initial position of neighbour i's NN) and
// LN(i) (the surface neighbour i projects onto)
template <class Src>
__device__ w3 blend2(const Src& s, float u, float v)
{
    const LSurf L = s.L();
    const float w = poly(L, u, v);
    const w3 pw = l2w(L, u, v, w);
    const int nh = s.nh();
    float isum = 0.0f;
    {
        w3 d = wsub(s.q0(), pw);
        isum += 1.0f / (wdot(d, d) + kSEps);
    }
    for (int i = 0; i < nh; ++i) {
        w3 d = wsub(s.qn(i), pw);
        isum += 1.0f / (wdot(d, d) + kSEps);
    }
    const float ii = 1.0f / isum;
    w3 sum = { 0.0f, 0.0f, 0.0f };
    {
        w3 d = wsub(s.q0(), pw);
        float inv = 1.0f / (wdot(d, d) + kSEps);
        sum = wadd(sum, wmul(ii, wmul(inv, pw)));
    }
    for (int i = 0; i < nh; ++i) {
        w3 d = wsub(s.qn(i), pw);
        float inv = 1.0f / (wdot(d, d) + kSEps);
        const LSurf LN = s.LN(i);
        w3 q = wsub(pw, LN.p0);
        float lu = wdot(q, LN.e[0]), lv = wdot(q, LN.e[1]);
        float lw = poly(LN, lu, lv);
        sum = wadd(sum, wmul(ii, wmul(inv, l2w(LN, lu, lv, lw))));
    }
    return sum;
}

// 1D: the curve v = a u^2 of the NN's 1D fit, blended with the projections
// onto the curves of the listed feature neighbours' NNs
template <class Src>
__device__ w3 blend1(const Src& s, float u)
{
    const LSurf1 L = s.L();
    const float v = L.a * u * u;
    const w3 pw = wadd(wadd(L.p0, wmul(u, L.e[0])), wmul(v, L.e[1]));
    const int n = s.nh();
    float isum = 0.0f;
    {
        w3 d = wsub(s.q0(), pw);
        isum += 1.0f / (wdot(d, d) + kSEps);
    }
    for (int i = 0; i < n; ++i) {
        w3 d = wsub(s.qn(i), pw);
        isum += 1.0f / (wdot(d, d) + kSEps);
    }
    const float ii = 1.0f / isum;
    w3 sum = { 0.0f, 0.0f, 0.0f };
    {
        w3 d = wsub(s.q0(), pw);
        float inv = 1.0f / (wdot(d, d) + kSEps);
        sum = wadd(sum, wmul(ii, wmul(inv, pw)));
    }
    for (int i = 0; i < n; ++i) {
        w3 d = wsub(s.qn(i), pw);
        float inv = 1.0f / (wdot(d, d) + kSEps);
        const LSurf1 LN = s.LN(i);
        w3 q = wsub(pw, LN.p0);
        float lu = wdot(q, LN.e[0]);
        float lv = LN.a * lu * lu;
        sum = wadd(sum, wmul(ii, wmul(inv, wadd(wadd(LN.p0, wmul(lu, LN.e[0])), wmul(lv, LN.e[1])))));
    }
    return sum;
}

// staged (LDS) sources
struct Proj2 {
    LSurf L;
    w3 q0;
    int n;
    w3 qn[kMaxNeigh];
    LSurf LN[kMaxNeigh];
};
struct Proj2Src {
    const Proj2* c;
    __device__ LSurf L() const { return c->L; }
    __device__ w3 q0() const { return c->q0; }
    __device__ int nh() const { return c->n; }
    __device__ w3 qn(int i) const { return c->qn[i]; }
    __device__ LSurf LN(int i) const { return c->LN[i]; }
};
struct Proj1 {
    LSurf1 L;
    w3 q0;
    int n;
    int nb[kMaxNeigh];
    w3 qn[kMaxNeigh];
    LSurf1 LN[kMaxNeigh];
};
struct Proj1Src {
    const Proj1* c;
    __device__ LSurf1 L() const { return c->L; }
    __device__ w3 q0() const { return c->q0; }
    __device__ int nh() const { return c->n; }
    __device__ w3 qn(int i) const { return c->qn[i]; }
    __device__ LSurf1 LN(int i) const { return c->LN[i]; }
};

// stage vid's 2D projection data: lane i < nh resolves one-ring neighbour i
// (its NN, that NN's initial position and fit / sector fit); all lanes call
// it, it ends with a barrier
__device__ void stage_proj2(const SurfArgs& a, int vid, w3 vidPos, Proj2* c)
{
    const int lane = threadIdx.x;
    const int h0 = a.heRow[vid];
    int nh = a.heRow[vid + 1] - h0;
    if (nh > kMaxNeigh) nh = kMaxNeigh;
    if (lane == 0) {
        const int n0 = a.nn[vid];
        c->L = reinterpret_cast<const LSurf*>(a.ls)[n0];
        c->q0 = W(a.pinit + 3 * n0);
        c->n = nh;
    }
    if (lane < nh) {
        const int nn = a.nn[a.heTarget[h0 + lane]];
        c->qn[lane] = W(a.pinit + 3 * nn);
        c->LN[lane] = neighbour_fit(a, nn, vidPos);
    }
    __syncthreads();
}

// feature neighbours of vid that have exactly two feature neighbours
// themselves (the curve neighbours of the 1D projection), in halfedge order
__device__ int curve_neighbours(const SurfArgs& a, int vid, int* nb)
{
    int nnb = 0;
    for (int h = a.heRow[vid]; h < a.heRow[vid + 1]; ++h) {
        int d = a.heTarget[h];
        if (!a.isFeature[d]) continue;
        int c = 0;
        for (int h2 = a.heRow[d]; h2 < a.heRow[d + 1]; ++h2) if (a.isFeature[a.heTarget[h2]]) ++c;
        if (c == 2 && nnb < kMaxNeigh) nb[nnb++] = d;
    }
    return nnb;
}

// stage vid's 1D projection data: lane 0 lists the curve neighbours, then
// lane i < n resolves neighbour i; all lanes call it, it ends with a barrier
__device__ void stage_proj1(const SurfArgs& a, int vid, Proj1* c)
{
    const LSurf1* ls1 = reinterpret_cast<const LSurf1*>(a.ls1);
    const int lane = threadIdx.x;
    if (lane == 0) {
        c->n = curve_neighbours(a, vid, c->nb);
        const int n0 = a.nn[vid];
        c->L = ls1[n0];
        c->q0 = W(a.pinit + 3 * n0);
    }
    __syncthreads();
    if (lane < c->n) {
        const int nn = a.nn[c->nb[lane]];
        c->qn[lane] = W(a.pinit + 3 * nn);
        c->LN[lane] = ls1[nn];
    }
    __syncthreads();
}

// unstaged source for single evaluations (surface sampling)
struct Proj2Global {
    const SurfArgs* a;
    int vid, h0, n;
    w3 vidPos;
    __device__ LSurf L() const { return reinterpret_cast<const LSurf*>(a->ls)[a->nn[vid]]; }
    __device__ w3 q0() const { return W(a->pinit + 3 * a->nn[vid]); }
    __device__ int nh() const { return n; }
    __device__ w3 qn(int i) const { return W(a->pinit + 3 * a->nn[a->heTarget[h0 + i]]); }
    __device__ LSurf LN(int i) const { return neighbour_fit(*a, a->nn[a->heTarget[h0 + i]], vidPos); }
};

__device__ w3 project2d(const SurfArgs& a, int vid, w3 vidPos, float u, float v)
{
    const int h0 = a.heRow[vid];
    int nh = a.heRow[vid + 1] - h0;
    if (nh > kMaxNeigh) nh = kMaxNeigh;
    return blend2(Proj2Global{ &a, vid, h0, nh, vidPos }, u, v);
}


// ------------------------------------------------------------- ring scores
struct HexRingLds { float cd[4], cF[4]; HexP hp[5]; };

struct RingFrame {
    float x0x, x0y, x0z, sg, is3, epsD, epsF, minDist, maxDist;
    int n;
    int nsub;   // p-dependent hex sub-tets scored: 5, or 4 without the principal-axes tet
};

// stage the volume one-ring of vid in LDS (same constants as the inner
// search) and return frame data; all lanes call it
__device__ RingFrame stage_ring(const SurfArgs& a, int vid, bool tet, void* lds, HexAcc* hacc)
{
    const int lane = threadIdx.x;
    RingFrame f;
    const int r0 = a.ringRow[vid];
    f.n = a.ringRow[vid + 1] - r0;
    f.x0x = a.x[vid]; f.x0y = a.y[vid]; f.x0z = a.z[vid];
    f.nsub = a.hexCorner8 ? 4 : 5;
    const int st = tet ? 3 : 7;
    float mabs = 0.0f, mn = FLT_MAX, mx = -FLT_MAX;
    for (int e = lane; e < f.n; e += 64) {
        const int* h = a.ringVal + (size_t)(r0 + e) * st;
        for (int t = 0; t < st; ++t) {
            int v = h[t];
            w3 d = { a.x[v] - f.x0x, a.y[v] - f.x0y, a.z[v] - f.x0z };
            mabs = fmaxf(mabs, fmaxf(fabsf(d.x), fmaxf(fabsf(d.y), fabsf(d.z))));
            float dist = wnorm(d);
            mn = fminf(mn, dist);
            mx = fmaxf(mx, dist);
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        mabs = fmaxf(mabs, __shfl_xor(mabs, m));
        mn = fminf(mn, __shfl_xor(mn, m));
        mx = fmaxf(mx, __shfl_xor(mx, m));
    }
    f.minDist = mn; f.maxDist = mx;
    f.sg = pow2_scale(mabs);
    const float isg = 1.0f / f.sg;
    f.is3 = ((1.0f * isg) * isg) * isg;
    const float s2 = f.sg * f.sg, s3 = s2 * f.sg;
    f.epsD = kHexEps * s3; f.epsF = kHexEps * s2;
    for (int e = lane; e < f.n; e += 64) {
        const int* h = a.ringVal + (size_t)(r0 + e) * st;
        if (tet) {
            f3 p[3];
            for (int t = 0; t < 3; ++t) {
                int v = h[t];
                p[t] = { (a.x[v] - f.x0x) * f.sg, (a.y[v] - f.x0y) * f.sg, (a.z[v] - f.x0z) * f.sg };
            }
            reinterpret_cast<TetC*>(lds)[e] = tet_consts(p[0], p[1], p[2]);
        } else {
            f3 r[8];
            r[0] = { 0.0f, 0.0f, 0.0f };
            for (int t = 0; t < 7; ++t) {
                int v = h[t];
                r[t + 1] = { (a.x[v] - f.x0x) * f.sg, (a.y[v] - f.x0y) * f.sg, (a.z[v] - f.x0z) * f.sg };
            }
            HexRingLds L;
            hex_consts(r, L.cd, L.cF, L.hp);
            reinterpret_cast<HexRingLds*>(lds)[e] = L;
        }
    }
    __syncthreads();
    if (!tet) {
        HexAcc acc = hex_acc0();
        const HexRingLds* L = reinterpret_cast<const HexRingLds*>(lds);
        for (int e = 0; e < f.n; ++e)
            for (int s = 0; s < 4; ++s) hex_acc_add(acc, L[e].cd[s], L[e].cF[s]);
        *hacc = acc;
    }
    return f;
}

// the p-dependent hex sub-tets of the ring at one point, NS per element (5:
// the nine-tet metric, 4: eight corners; a compile-time count keeps the
// element loop free of the metric branch)
template <int NS>
__device__ __forceinline__ void ring_hex_add(HexAcc& acc, const HexRingLds* L, int n, float px, float py, float pz)
{
    const float Q1 = fmaf(pz, pz, py * py), Q3 = 3.0f * Q1, t3 = 3.0f * px;
    for (int e = 0; e < n; ++e)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const HexP h = L[e].hp[s];
            const bool a3 = (s == 1 || s == 4);
            float Tyz = fmaf(h.gy, py, fmaf(h.gz, pz, h.D));
            float Ryz = fmaf(h.sy, py, fmaf(h.sz, pz, h.K + (a3 ? Q3 : Q1)));
            hex_acc_add(acc, fmaf(h.gx, px, Tyz), fmaf(px, (a3 ? t3 : px) + h.sx, Ryz));
        }
}

__device__ float ring_score(const RingFrame& f, bool tet, const void* lds, const HexAcc& hacc, w3 p)
{
    const float px = (p.x - f.x0x) * f.sg, py = (p.y - f.x0y) * f.sg, pz = (p.z - f.x0z) * f.sg;
    if (tet) {
        const TetC* tc = reinterpret_cast<const TetC*>(lds);
        const float Qyz = 3.0f * fmaf(pz, pz, py * py);
        const float t3 = 3.0f * px;
        TetAcc acc = tet_acc0();
        for (int e = 0; e < f.n; ++e) {
            const TetC c = tc[e];
            float Tyz = fmaf(c.nny, py, fmaf(c.nnz, pz, c.D));
            float Ryz = fmaf(c.m2sy, py, fmaf(c.m2sz, pz, c.K + Qyz));
            tet_acc_add(acc, fmaf(c.nnx, px, Tyz), fmaf(px, t3 + c.m2sx, Ryz));
        }
        return tet_acc_score(acc, f.is3);
    }
    const HexRingLds* L = reinterpret_cast<const HexRingLds*>(lds);
    HexAcc acc = hacc;
    if (f.nsub == 4) ring_hex_add<4>(acc, L, f.n, px, py, pz);
    else ring_hex_add<5>(acc, L, f.n, px, py, pz);
    return hex_acc_score(acc, f.epsD, f.epsF);
}

__device__ __forceinline__ void wave_argmax_s(float& best, int& idx)
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

// ------------------------------------------------------------- search bodies
// Scorers: score(p) is what the search maximises (one candidate per lane);
// start(p0, qs, qref) gives the start position's score and the reference the
// last level's maximum must beat for the write (all lanes call it).

// volume one-ring of a tet / hex surface vertex
struct VolScore {
    RingFrame f;
    bool tet;
    const void* lds;
    HexAcc hacc;
    __device__ float score(w3 p) const { return ring_score(f, tet, lds, hacc, p); }
    // start score (maximised from) and the reference the last level must beat
    __device__ void start(w3 p, float& qs, float& qref) const { qs = score(p); qref = qs; }
};

// fan of a surface-mesh vertex (DMOQualityTri.h:257-275, DMOQualityQuad.h:
// 30-50): triangles (p, r_k, r_k+1), quads (p, r_k, d_k, r_k+1); score()
// signs each element by the nearest neighbour's vertex normal n, final_ref()
// uses the unsigned metric
template <bool QUAD>
struct FanScore {
    const w3* pts;
    int ne;
    w3 n;
    __device__ float elem(w3 p, int k, bool oriented) const
    {
        const f3 nn = { n.x, n.y, n.z };
        const f3* no = oriented ? &nn : nullptr;
        const f3 q = { p.x, p.y, p.z };
        if (QUAD) {
            const w3 a = pts[2 * k], b = pts[2 * k + 1], c = pts[2 * k + 2];
            return mr_quad3(q, f3{ a.x, a.y, a.z }, f3{ b.x, b.y, b.z }, f3{ c.x, c.y, c.z }, no);
        }
        const w3 a = pts[k], b = pts[k + 1];
        return mr_tri3(q, f3{ a.x, a.y, a.z }, f3{ b.x, b.y, b.z }, no);
    }
    __device__ void elem2(w3 p, int k, float& qs, float& qu) const
    {
        const f3 nn = { n.x, n.y, n.z };
        const f3 q = { p.x, p.y, p.z };
        if (QUAD) {
            const w3 a = pts[2 * k], b = pts[2 * k + 1], c = pts[2 * k + 2];
            mr_quad3_both(q, f3{ a.x, a.y, a.z }, f3{ b.x, b.y, b.z }, f3{ c.x, c.y, c.z }, nn, qs, qu);
        } else {
            const w3 a = pts[k], b = pts[k + 1];
            mr_tri3_both(q, f3{ a.x, a.y, a.z }, f3{ b.x, b.y, b.z }, nn, qs, qu);
        }
    }
    __device__ float score(w3 p) const
    {
        float q = FLT_MAX;
        for (int k = 0; k < ne; ++k) q = fminf(q, elem(p, k, true));
        return q;
    }
    // lane k < ne evaluates fan element k signed and unsigned; two wave
    // min-reductions (fminf is exact, so the order does not matter)
    __device__ void start(w3 p, float& qs, float& qref) const
    {
        const int k = threadIdx.x;
        float s = FLT_MAX, u = FLT_MAX;
        if (k < ne) elem2(p, k, s, u);
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            s = fminf(s, __shfl_xor(s, m));
            u = fminf(u, __shfl_xor(u, m));
        }
        qs = s;
        qref = u;
    }
};

// fan of a flat triangle mesh's boundary vertex in the canonical 2D
// evaluation (frame of the start position x0, scale sg)
struct FlatScore {
    const TriC* tc;
    int ne;
    w3 x0;
    float sg, is2;
    __device__ float score(w3 q) const
    {
        const float px = (q.x - x0.x) * sg, py = (q.y - x0.y) * sg;
        const float Q2 = 2.0f * (py * py);
        const float t2 = 2.0f * px;
        TriAcc acc = tri_acc0();
        for (int e = 0; e < ne; ++e) {
            const TriC c = tc[e];
            float Ty = fmaf(c.gy, py, c.C);
            float Ry = fmaf(c.sy, py, c.K + Q2);
            tri_acc_add(acc, fmaf(c.gx, px, Ty), fmaf(px, t2 + c.sx, Ry));
        }
        return tri_acc_score(acc, is2);
    }
    __device__ void start(w3 p, float& qs, float& qref) const { qs = score(p); qref = qs; }
};

// 2D search on the blended local surfaces.  Lane l holds reference thread
// t = l/2's candidate A (l even) or B (l odd): A = (t/nq, t%nq), B = (t/nq +
// nq/2, t%nq); the surface kernels keep A on an exact tie, so A ranks first.
// Window half-width s_d * dist in the nearest neighbour's (u,v).
template <class S>
__device__ void surf2d_search(const SurfArgs& a, int vid, w3 vidPos, float dist, int nq, const S& sc)
{
    const int lane = threadIdx.x;
    __shared__ Proj2 s_proj;
    stage_proj2(a, vid, vidPos, &s_proj);
    const Proj2Src ps{ &s_proj };
    const LSurf L = s_proj.L;
    w3 cur = vidPos;
    const w3 cvec = wsub(cur, L.p0);
    float cu = wdot(L.e[0], cvec), cv = wdot(L.e[1], cvec);
    float qcur, qold;
    sc.start(cur, qcur, qold);
    float qlast = -FLT_MAX;
    const int t = lane >> 1;
    const bool isB = (lane & 1) != 0;
    const bool active = t < nq * nq / 2;
    const int ci = t / nq + (isB ? nq / 2 : 0), cj = t % nq;
    float ds = a.gridScale;
    for (int d = 0; d < a.depth; ++d) {
        const float uMax = ds * dist, uMin = -(ds * dist);
        float best = -FLT_MAX;
        float u = 0.0f, v = 0.0f;
        w3 pc = { 0.0f, 0.0f, 0.0f };
        if (active) {
            u = cu + a.aff * ((float)ci * uMin + (float)(nq - 1 - ci) * uMax);
            v = cv + a.aff * ((float)cj * uMin + (float)(nq - 1 - cj) * uMax);
            pc = blend2(ps, u, v);
            best = sc.score(pc);
        }
        int wl = lane;
        wave_argmax_s(best, wl);
        if (qcur < best) {
            // the winner's projected point (the projection of (bu, bv))
            cu = __shfl(u, wl); cv = __shfl(v, wl);
            cur = { __shfl(pc.x, wl), __shfl(pc.y, wl), __shfl(pc.z, wl) };
            qcur = best;
        }
        qlast = best;
        ds = ds * a.shrink;
    }
    if (qold < qlast && lane == 0) { a.xw[vid] = cur.x; a.yw[vid] = cur.y; a.zw[vid] = cur.z; }
}

// 1D search along the feature curve: lane i < nq holds candidate i; the
// curve projection blends the fits of the curve neighbours
template <class S>
__device__ void surf1d_search(const SurfArgs& a, int vid, w3 vidPos, float dist, int nq, const S& sc)
{
    const int lane = threadIdx.x;
    __shared__ Proj1 s_proj;
    stage_proj1(a, vid, &s_proj);
    const Proj1Src ps{ &s_proj };
    const LSurf1 L = s_proj.L;
    w3 cur = vidPos;
    const w3 cvec = wsub(cur, L.p0);
    float cu = wdot(L.e[0], cvec);
    float qcur, qold;
    sc.start(cur, qcur, qold);
    float qlast = -FLT_MAX;
    const bool active = lane < nq;
    float ds = a.gridScale;
    for (int d = 0; d < a.depth; ++d) {
        const float uMax = ds * dist, uMin = -(ds * dist);
        float best = -FLT_MAX, u = 0.0f;
        w3 pc = { 0.0f, 0.0f, 0.0f };
        if (active) {
            u = cu + a.aff * ((float)lane * uMin + (float)(nq - 1 - lane) * uMax);
            pc = blend1(ps, u);
            best = sc.score(pc);
        }
        int wl = lane;
        wave_argmax_s(best, wl);
        if (qcur < best) {
            cu = __shfl(u, wl);
            cur = { __shfl(pc.x, wl), __shfl(pc.y, wl), __shfl(pc.z, wl) };
            qcur = best;
        }
        qlast = best;
        ds = ds * a.shrink;
    }
    if (qold < qlast && lane == 0) { a.xw[vid] = cur.x; a.yw[vid] = cur.y; a.zw[vid] = cur.z; }
}

// ------------------------------------------------- tet / hex surface searches
// Window half-width: s_d * (min distance to the ring points for hexes, max
// for tets).  (k_optimizeHierarchical2DTet/Hex, 1DTet/Hex)
__global__ __launch_bounds__(64) void k_search_surf2d(SurfArgs a, int cStart, int nq, int tet)
{
    extern __shared__ float s_lds[];
    const int vid = cStart + blockIdx.x;
    VolScore sc;
    sc.hacc = hex_acc0();
    sc.f = stage_ring(a, vid, tet != 0, s_lds, &sc.hacc);
    if (sc.f.n <= 0) return;
    sc.tet = tet != 0;
    sc.lds = s_lds;
    surf2d_search(a, vid, w3{ sc.f.x0x, sc.f.x0y, sc.f.x0z }, tet ? sc.f.maxDist : sc.f.minDist, nq, sc);
}

__global__ __launch_bounds__(64) void k_search_feat1d(SurfArgs a, int cStart, int nq, int tet)
{
    extern __shared__ float s_lds[];
    const int vid = cStart + blockIdx.x;
    if (!a.opt1d[vid]) return;   // uniform over the block
    VolScore sc;
    sc.hacc = hex_acc0();
    sc.f = stage_ring(a, vid, tet != 0, s_lds, &sc.hacc);
    if (sc.f.n <= 0) return;
    sc.tet = tet != 0;
    sc.lds = s_lds;
    surf1d_search(a, vid, w3{ sc.f.x0x, sc.f.x0y, sc.f.x0z }, tet ? sc.f.maxDist : sc.f.minDist, nq, sc);
}

// ------------------------------------------------ surface-mesh (tri / quad)
// Stage the fan of vid in LDS: lane i < n loads the halfedge target r_i (and
// for quads the diagonal d_i, the target of the halfedge after vid -> r_i in
// its face); closed (r_n = r_0) unless open.  Returns false (uniformly) when
// the vertex has no fan element.  dist: the largest distance from vid to the
// r_i (and d_i).  The d_i come from the pass snapshot: on a surface boundary
// the "diagonal" of the boundary halfedge is the boundary neighbour of the
// next boundary vertex, which can be in the colour class being updated;
// every other fan vertex shares an element with vid, so the snapshot holds
// its current position.
template <bool QUAD>
__device__ bool stage_fan(const SurfArgs& a, int vid, bool open, w3* pts, FanScore<QUAD>& sc, float& dist)
{
    const int lane = threadIdx.x;
    const int h0 = a.heRow[vid], n = a.heRow[vid + 1] - h0;
    if (n <= 0 || n > kMaxNeigh) return false;
    const w3 x0 = WP(a.x, a.y, a.z, vid);
    const int K = QUAD ? 2 : 1;
    float mx = -FLT_MAX;
    if (lane < n) {
        const int h = h0 + lane;
        const w3 r = WP(a.x, a.y, a.z, a.heTarget[h]);
        pts[K * lane] = r;
        mx = wnorm(wsub(r, x0));
        if (QUAD) {
            const int nx = a.heNext[h];
            const w3 dg = WP(a.sx, a.sy, a.sz, a.heTarget[nx >= 0 ? nx : h]);
            pts[K * lane + 1] = dg;
            mx = fmaxf(mx, wnorm(wsub(dg, x0)));
        }
    }
    if (lane == 0 && !open) pts[K * n] = WP(a.x, a.y, a.z, a.heTarget[h0]);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) mx = fmaxf(mx, __shfl_xor(mx, m));
    __syncthreads();
    sc.pts = pts;
    sc.ne = open ? n - 1 : n;
    sc.n = W(a.vnorm + 3 * a.nn[vid]);
    dist = mx;
    return sc.ne > 0;
}

// free vertices: 2D search, fan always closed (k_optimizeHierarchical2D,
// DMOTriImplCub.h:25-200, DMOQuadImplCub.h:23-212)
template <bool QUAD>
__global__ __launch_bounds__(64) void k_search_fan2d(SurfArgs a, int cStart, int nq)
{
    __shared__ w3 s_fan[2 * kMaxNeigh + 1];
    const int vid = cStart + blockIdx.x;
    FanScore<QUAD> sc;
    float dist;
    if (!stage_fan<QUAD>(a, vid, false, s_fan, sc, dist)) return;
    surf2d_search(a, vid, WP(a.x, a.y, a.z, vid), dist, nq, sc);
}

// feature vertices: 1D search, fan open on a surface boundary
// (k_optimizeHierarchical1D, DMOTriImplCub.h:205-378, DMOQuadImplCub.h:216-402)
template <bool QUAD>
__global__ __launch_bounds__(64) void k_search_fan1d(SurfArgs a, int cStart, int nq)
{
    __shared__ w3 s_fan[2 * kMaxNeigh + 1];
    const int vid = cStart + blockIdx.x;
    if (!a.opt1d[vid]) return;   // uniform over the block
    FanScore<QUAD> sc;
    float dist;
    if (!stage_fan<QUAD>(a, vid, a.isB1D[vid] != 0, s_fan, sc, dist)) return;
    surf1d_search(a, vid, WP(a.x, a.y, a.z, vid), dist, nq, sc);
}

// vertex normals from the current face normals (k_calcVertexNormals)
__global__ void k_surf_vertex_normals(SurfArgs a)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.nS) return;
    w3 n = { 0.0f, 0.0f, 0.0f };
    for (int h = a.heRow[v]; h < a.heRow[v + 1]; ++h)
        if (a.heFace[h] != -1) n = wadd(n, W(a.fnorm + 3 * a.heFace[h]));
    n = wnormalize(n);
    a.vnorm[3 * v] = n.x; a.vnorm[3 * v + 1] = n.y; a.vnorm[3 * v + 2] = n.z;
}

// positions snapshot at the start of a colour pass (one launch for x, y, z)
__global__ void k_snapshot(const float* x, const float* y, const float* z, float* sx, float* sy, float* sz, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sx[i] = x[i]; sy[i] = y[i]; sz[i] = z[i];
}

// Laplace smoothing of one colour class of free surface vertices
// (k_laplaceSmoothing, Laplace.cu:8-28): p <- (p + sum of the halfedge
// targets) * (1 / (1 + n))
__global__ void k_laplace(float* x, float* y, float* z, const int* heRow, const int* heTarget, int lo, int hi)
{
    int v = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= hi) return;
    const int h0 = heRow[v], n = heRow[v + 1] - h0;
    w3 q = WP(x, y, z, v);
    for (int h = h0; h < h0 + n; ++h) q = wadd(q, WP(x, y, z, heTarget[h]));
    q = wmul(1.0f / (float)(1 + n), q);
    x[v] = q.x; y[v] = q.y; z[v] = q.z;
}

// ------------------------------------------------------ 1D flat boundary search
// Boundary vertices of a flat triangle mesh follow their boundary curve
// (k_optimizeHierarchical1DFlat): the ring is the halfedge fan (closed only
// off a surface boundary), the window half-width is s_d times the largest
// Euclidean distance to the fan, and candidates are scored by the 2D mean
// ratio of the fan triangles in the frame of the start position.
__global__ __launch_bounds__(64) void k_search_flat1d(SurfArgs a, int cStart, int nq)
{
    __shared__ TriC s_tr[kMaxNeigh];
    const int lane = threadIdx.x;
    const int vid = cStart + blockIdx.x;
    if (!a.opt1d[vid]) return;   // uniform over the block
    const int h0 = a.heRow[vid], n = a.heRow[vid + 1] - h0;
    const int nOne = n + (a.isB1D[vid] ? 0 : 1);
    if (n <= 0 || nOne < 2 || n > kMaxNeigh) return;
    const w3 x0 = WP(a.x, a.y, a.z, vid);
    float mabs = 0.0f, mx = -FLT_MAX;
    if (lane < n) {
        w3 d = wsub(WP(a.x, a.y, a.z, a.heTarget[h0 + lane]), x0);
        mabs = fmaxf(fabsf(d.x), fabsf(d.y));
        mx = wnorm(d);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        mabs = fmaxf(mabs, __shfl_xor(mabs, m));
        mx = fmaxf(mx, __shfl_xor(mx, m));
    }
    const float sg = pow2_scale(mabs);
    const float isg = 1.0f / sg;
    const float is2 = (1.0f * isg) * isg;
    const int ne = nOne - 1;
    if (lane < ne) {
        const int va = a.heTarget[h0 + lane], vb = a.heTarget[h0 + (lane + 1 == n ? 0 : lane + 1)];
        s_tr[lane] = tri_consts((a.x[va] - x0.x) * sg, (a.y[va] - x0.y) * sg, (a.x[vb] - x0.x) * sg, (a.y[vb] - x0.y) * sg);
    }
    __syncthreads();
    surf1d_search(a, vid, x0, mx, nq, FlatScore{ s_tr, ne, x0, sg, is2 });
}

// ------------------------------------------------------- surface sampling
// Points of vid's local surfaces on an nu x nv grid (k_fillEstimateLocalSurfacePoints
// with USE_SURF_OF_NN = false, DMOCommon.h:197-276; k_fillLocalSurfacePoints,
// DMOCommon.cu:73-150): idx = i*nu + j.  mode 0 blends the projections
// (window s * maxDist), mode 1 evaluates the NN's fit or its sector fit sid
// (window 2 s * maxDist).  One thread per point.
__global__ void k_surface_points(SurfArgs a, int vid, int nu, int nv, int mode, int sid, float* out)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nu * nv) return;
    const w3 pv = WP(a.x, a.y, a.z, vid);
    float mx = -FLT_MAX;
    for (int h = a.heRow[vid]; h < a.heRow[vid + 1]; ++h) mx = fmaxf(mx, wnorm(wsub(WP(a.x, a.y, a.z, a.heTarget[h]), pv)));
    const int nn = a.nn[vid];
    LSurf L = reinterpret_cast<const LSurf*>(a.ls)[nn];
    if (mode == 1 && nn >= a.nFree) {
        int off = nn - a.nFree, nsec = a.frow[off + 1] - a.frow[off];
        if (sid >= nsec) sid = nsec - 1;
        if (nsec > 0 && sid >= 0) L = reinterpret_cast<const LSurf*>(a.fs)[a.frow[off] + sid];
    }
    const w3 c = wsub(pv, L.p0);
    const float cu = wdot(L.e[0], c), cv = wdot(L.e[1], c);
    float uMax = a.gridScale * mx, uMin = -(a.gridScale * mx);
    if (mode == 1) { uMax *= 2.0f; uMin *= 2.0f; }
    const float au = 1.0f / (float)(nu - 1), av = 1.0f / (float)(nv - 1);
    const int j = idx % nu, i = idx / nu;
    const float u = cu + au * ((float)i * uMin + (float)(nu - 1 - i) * uMax);
    const float v = cv + av * ((float)j * uMin + (float)(nv - 1 - j) * uMax);
    const w3 q = (mode == 0) ? project2d(a, vid, pv, u, v) : l2w(L, u, v, poly(L, u, v));
    out[3 * idx] = q.x; out[3 * idx + 1] = q.y; out[3 * idx + 2] = q.z;
}

// ------------------------------------------------------------- launchers
static inline int nb256(long long n) { return (int)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

size_t surf_lds_bytes(int elem, int maxRing)
{
    size_t per = (elem == VSDMO_TET) ? sizeof(TetC) : sizeof(HexRingLds);
    return per * (size_t)(maxRing > 0 ? maxRing : 1);
}

void surf_init_launch(const SurfArgs& a, hipStream_t s)
{
    k_surf_face_normals<<<nb256(a.nFaces), 256, 0, s>>>(a);
    k_surf_vertex_init<<<nb256(a.nS), 256, 0, s>>>(a);
}
void surf_sector_count_launch(const SurfArgs& a, int* cnt, hipStream_t s)
{
    k_sector_count<<<nb256(a.nFeature), 256, 0, s>>>(a, cnt);
}
void surf_sector_launch(const SurfArgs& a, int* secN, int* secCnt, int nSec, hipStream_t s)
{
    k_sector_lists<<<nb256(a.nFeature), 256, 0, s>>>(a, secN, secCnt);
    if (nSec > 0) k_sector_fit<<<nb256(nSec), 256, 0, s>>>(a, secN, secCnt, nSec);
}
void surf_nn_launch(const SurfArgs& a, hipStream_t s)
{
    k_update_nn<<<nb256(a.nS), 256, 0, s>>>(a);
}
void surf_search2d_launch(const SurfArgs& a, int cStart, int cEnd, int nq, int tet, size_t lds, hipStream_t s)
{
    if (cEnd > cStart) k_search_surf2d<<<cEnd - cStart, 64, lds, s>>>(a, cStart, nq, tet);
}
void surf_search1d_launch(const SurfArgs& a, int cStart, int cEnd, int nq, int tet, size_t lds, hipStream_t s)
{
    if (cEnd > cStart) k_search_feat1d<<<cEnd - cStart, 64, lds, s>>>(a, cStart, nq, tet);
}
void surf_points_launch(const SurfArgs& a, int vid, int nu, int nv, int mode, int sid, float* out, hipStream_t s)
{
    const int n = nu * nv;
    if (n > 0) k_surface_points<<<(n + 63) / 64, 64, 0, s>>>(a, vid, nu, nv, mode, sid, out);
}
void surf_flat1d_launch(const SurfArgs& a, int cStart, int cEnd, int nq, hipStream_t s)
{
    if (cEnd > cStart) k_search_flat1d<<<cEnd - cStart, 64, 0, s>>>(a, cStart, nq);
}
void surf_fan2d_launch(const SurfArgs& a, int cStart, int cEnd, int nq, hipStream_t s)
{
    if (cEnd <= cStart) return;
    if (a.fpe == 4) launch_kernel(k_search_fan2d<true>, cEnd - cStart, 64, 0, s, a, cStart, nq);
    else launch_kernel(k_search_fan2d<false>, cEnd - cStart, 64, 0, s, a, cStart, nq);
}
void surf_fan1d_launch(const SurfArgs& a, int cStart, int cEnd, int nq, hipStream_t s)
{
    if (cEnd <= cStart) return;
    if (a.fpe == 4) launch_kernel(k_search_fan1d<true>, cEnd - cStart, 64, 0, s, a, cStart, nq);
    else launch_kernel(k_search_fan1d<false>, cEnd - cStart, 64, 0, s, a, cStart, nq);
}
void surf_normals_launch(const SurfArgs& a, hipStream_t s)
{
    k_surf_face_normals<<<nb256(a.nFaces), 256, 0, s>>>(a);
    k_surf_vertex_normals<<<nb256(a.nS), 256, 0, s>>>(a);
}
void snapshot_launch(const SurfArgs& a, hipStream_t s)
{
    if (a.nS > 0) k_snapshot<<<nb256(a.nS), 256, 0, s>>>(a.x, a.y, a.z, const_cast<float*>(a.sx), const_cast<float*>(a.sy),
                                                         const_cast<float*>(a.sz), a.nS);
}
void laplace_launch(float* x, float* y, float* z, const int* heRow, const int* heTarget, int lo, int hi, hipStream_t s)
{
    if (hi > lo) k_laplace<<<nb256(hi - lo), 256, 0, s>>>(x, y, z, heRow, heTarget, lo, hi);
}

}  // namespace vsdmo
