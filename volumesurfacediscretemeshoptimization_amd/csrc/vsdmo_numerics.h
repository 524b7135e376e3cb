// vsdmo_numerics.h -- FP32 element and ring evaluation used by the DMO kernels.
//
// Every operation below rounds on its own (the library is compiled with
// -ffp-contract=off); fused multiply-adds appear only as explicit fmaf().
// The evaluation order is fixed so that a host restatement of the same order
// reproduces every value bit for bit.
//
// Metrics (reference VSDMO/src/dmo):
//   tet mean ratio  12*cbrt(9)*cbrt(V^2)/sum(l^2), signed V if inverted
//                   (DMOQualityTet.h:119-143)
//   hex shape       3*min over 9 sub-tets cbrt(det^2)/|J|_F^2, 0 if any
//                   det or |J|^2 < 1.19209e-7 (DMOQualityHex.h:54-71,
//                   DMOQualityUtil.h:307-349)
//   tri 2D ratio    2*sqrt(3)*cross/sum(l^2), signed cross if inverted
//                   (DMOQualityTri.h:243-257)
// All evaluations run in a vertex-local frame p' = (p - x0) * sigma, sigma a
// power of two (so scaling changes no rounding; it only keeps products such
// as det^2 * L^3 inside the float range).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vsdmo {

// 12/cbrt(4): mean-ratio constant for det = 6V
constexpr float kCTet = 0x1.e3cf48p+2f;
// 2*sqrt(3) formed as 2.f*sqrtf(3.f)
constexpr float kCTri = 0x1.bb67aep+1f;
// hex shape metric threshold
constexpr float kHexEps = 1.19209e-7f;

__host__ __device__ inline uint32_t fbits(float f) { return __builtin_bit_cast(uint32_t, f); }
__host__ __device__ inline float bitsf(uint32_t u) { return __builtin_bit_cast(float, u); }

// Cube root for x >= 0 from the exponent/3 bit split: r0 = bits^-1(0x54a8a9ab -
// bits(x)/3) approximates x^(-1/3) within 7%; three Newton steps on r^-3 = x
// (r += r*(1 - x r^3)/3), then y = x r^2 and one Newton step on y^3 = x with
// r^2 standing in for 1/y^2.  Faithful (< 1 ulp).  Denormals pre-scaled by 2^24.
__device__ __forceinline__ float cbrt_nr(float x)
{
    if (!(x > 0.0f)) return x;
    if (x > 3.40282347e+38f) return x;
    float s = 1.0f;
    if (x < 1.17549435e-38f) { x = x * 16777216.0f; s = 0.00390625f; }
    const float third = 1.0f / 3.0f;
    float r = bitsf(0x54a8a9abu - fbits(x) / 3u);
#pragma unroll
    for (int it = 0; it < 3; ++it) {
        float r3 = (r * r) * r;
        float t = fmaf(-x, r3, 1.0f);
        r = fmaf(t * r, third, r);
    }
    float y = (x * r) * r;
    float c = fmaf(y * y, y, -x);
    y = fmaf(-c, (r * r) * third, y);
    return y * s;
}

// 2^-e for m = f * 2^e, f in [1,2); 1 for 0, denormal, inf, NaN
__host__ __device__ inline float pow2_scale(float m)
{
    uint32_t e = (fbits(m) >> 23) & 0xffu;
    if (e == 0u || e >= 254u) return 1.0f;
    return bitsf((254u - e) << 23);
}

struct f3 { float x, y, z; };
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
__device__ __forceinline__ f3 neg3(f3 a) { return { -a.x, -a.y, -a.z }; }
__device__ __forceinline__ f3 scl3(f3 a, float s) { return { a.x * s, a.y * s, a.z * s }; }
// cross: per component mul, mul, sub (no fusion)
__device__ __forceinline__ f3 cross3(f3 a, f3 b)
{
    return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
// dot: fmaf(x, x', fmaf(y, y', z * z'))
__device__ __forceinline__ float dot3(f3 a, f3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }

// ------------------------------------------------- surface triangles / quads
// Literal evaluation order of the reference's Vec3 (dot = (x*x' + y*y') +
// z*z', norm = sqrtf(dot), normalized divides each component), contraction
// off, for the surface-mesh metrics (DMOQualityTri.h:240-275,
// DMOQualityQuad.h:13-50).
__device__ __forceinline__ float ldot3(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float lnorm3(f3 a) { return sqrtf((a.x * a.x + a.y * a.y) + a.z * a.z); }
__device__ __forceinline__ f3 lunit3(f3 a)
{
    const float l = lnorm3(a);
    return { a.x / l, a.y / l, a.z / l };
}
// 2 sqrt(3), the FP32 product 2 * sqrtf(3)
constexpr float kTwoSqrt3 = 3.46410155296325684f;

// meanRatioMetricTri of (p0, p1, p2): 2 sqrt(3) A / (|e0|^2 + |e1|^2 + |e2|^2)
// with e0 = p1 - p0, e1 = p2 - p0, e2 = p2 - p1 and A = |e0 x e1|, times
// sign(n . unit(e0 x e1)) when n is given
__device__ __forceinline__ float mr_tri3(f3 p0, f3 p1, f3 p2, const f3* n)
{
    const f3 e0 = sub3(p1, p0), e1 = sub3(p2, p0), e2 = sub3(p2, p1);
    const float l = (ldot3(e0, e0) + ldot3(e1, e1)) + ldot3(e2, e2);
    const f3 N = cross3(e0, e1);
    float area = lnorm3(N);
    if (n) {
        const float d = ldot3(lunit3(N), *n);
        area = area * (d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f));
    }
    return (kTwoSqrt3 * area) / l;
}

// signed (by n) and unsigned values of one triangle from one evaluation
__device__ __forceinline__ void mr_tri3_both(f3 p0, f3 p1, f3 p2, f3 n, float& qs, float& qu)
{
    const f3 e0 = sub3(p1, p0), e1 = sub3(p2, p0), e2 = sub3(p2, p1);
    const float l = (ldot3(e0, e0) + ldot3(e1, e1)) + ldot3(e2, e2);
    const f3 N = cross3(e0, e1);
    const float area = lnorm3(N);
    const float d = ldot3(lunit3(N), n);
    qs = (kTwoSqrt3 * (area * (d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f)))) / l;
    qu = (kTwoSqrt3 * area) / l;
}

// meanRatioMetricQuad of (p0, p1, p2, p3): 2 (|e1 x -e0| + |e3 x -e2|) /
// sum |e_i|^2 with e_i = p_(i+1) - p_i; with n: -1 when a corner triangle's
// unit normal u has n . u < 0.7
__device__ __forceinline__ float mr_quad3(f3 p0, f3 p1, f3 p2, f3 p3, const f3* n)
{
    const f3 e0 = sub3(p1, p0), e1 = sub3(p2, p1), e2 = sub3(p3, p2), e3 = sub3(p0, p3);
    const float l = ((ldot3(e0, e0) + ldot3(e1, e1)) + ldot3(e2, e2)) + ldot3(e3, e3);
    const f3 N1 = cross3(e1, neg3(e0)), N2 = cross3(e3, neg3(e2));
    const float a1 = lnorm3(N1), a2 = lnorm3(N2);
    if (n) {
        const float d1 = ldot3(lunit3(N1), *n), d2 = ldot3(lunit3(N2), *n);
        if (d1 < 0.7f || d2 < 0.7f) return -1.0f;
    }
    return (2.0f * (a1 + a2)) / l;
}

__device__ __forceinline__ void mr_quad3_both(f3 p0, f3 p1, f3 p2, f3 p3, f3 n, float& qs, float& qu)
{
    const f3 e0 = sub3(p1, p0), e1 = sub3(p2, p1), e2 = sub3(p3, p2), e3 = sub3(p0, p3);
    const float l = ((ldot3(e0, e0) + ldot3(e1, e1)) + ldot3(e2, e2)) + ldot3(e3, e3);
    const f3 N1 = cross3(e1, neg3(e0)), N2 = cross3(e3, neg3(e2));
    const float a1 = lnorm3(N1), a2 = lnorm3(N2);
    const float d1 = ldot3(lunit3(N1), n), d2 = ldot3(lunit3(N2), n);
    qu = (2.0f * (a1 + a2)) / l;
    qs = (d1 < 0.7f || d2 < 0.7f) ? -1.0f : qu;
}

// ---------------------------------------------------------------- tets
// Ring element (a, b, c) of a vertex in the local frame.  With n = (b-a)x(c-a):
//   det(p) = 6V = D - n.p,            D = n.a
//   L(p)   = sum of squared edges = 3|p|^2 + m2s.p + K,
//            m2s = -2(a+b+c), K = |a|^2+|b|^2+|c|^2+|b-a|^2+|c-a|^2+|c-b|^2
// Candidate evaluation with the (y,z) part split off (reused for the nq
// candidates along x that one lane owns):
//   Qyz = 3*fmaf(pz,pz,py*py)
//   Tyz = fmaf(-ny,py, fmaf(-nz,pz, D))          det = fmaf(-nx,px, Tyz)
//   Ryz = fmaf(m2sy,py, fmaf(m2sz,pz, K + Qyz))  L   = fmaf(px, 3*px + m2sx, Ryz)
struct TetC { float nnx, nny, nnz, D, m2sx, m2sy, m2sz, K; };

__device__ __forceinline__ TetC tet_consts(f3 a, f3 b, f3 c)
{
    TetC t;
    f3 e1 = sub3(b, a), e2 = sub3(c, a), e3 = sub3(c, b);
    f3 n = cross3(e1, e2);
    t.D = dot3(n, a);
    t.nnx = -n.x; t.nny = -n.y; t.nnz = -n.z;
    f3 s = add3(add3(a, b), c);
    t.m2sx = -2.0f * s.x; t.m2sy = -2.0f * s.y; t.m2sz = -2.0f * s.z;
    float la = dot3(a, a), lb = dot3(b, b), lc = dot3(c, c);
    float l1 = dot3(e1, e1), l2 = dot3(e2, e2), l3 = dot3(e3, e3);
    t.K = ((((la + lb) + lc) + l1) + l2) + l3;
    return t;
}

// Running selection of the ring minimum of det^2/L^3 ("d2*bL3 < bd2*L3", ring
// order) plus the minimum det.  Final score: most negative det/6 rescaled to
// world units if any det < 0; 0 if min det == 0; else kCTet*cbrt(bd2/bL3).
struct TetAcc { float bd2, bL3, dmin; };
__device__ __forceinline__ TetAcc tet_acc0() { return { 1.0f, 0.0f, 3.40282347e+38f }; }
__device__ __forceinline__ void tet_acc_add(TetAcc& a, float det, float L)
{
    float d2 = det * det;
    float L3 = (L * L) * L;
    bool take = d2 * a.bL3 < a.bd2 * L3;
    a.bd2 = take ? d2 : a.bd2;
    a.bL3 = take ? L3 : a.bL3;
    a.dmin = fminf(a.dmin, det);
}
__device__ __forceinline__ float tet_acc_score(const TetAcc& a, float invs3)
{
    if (a.dmin < 0.0f) return (a.dmin * (1.0f / 6.0f)) * invs3;
    if (a.dmin == 0.0f) return 0.0f;
    return kCTet * cbrt_nr(a.bd2 / a.bL3);
}

// ---------------------------------------------------------------- hexes
// shapeMetricHex with the candidate p as internal vertex 0 and halfhex points
// r1..r7: the reference evaluates hex_edges / hex_principal_axes on
// (r4,r5,r6,r7,p,r1,r2,r3).  Edge vectors:
//   L0=r5-r4 L1=r6-r5 L2=r7-r6 L3=r7-r4 L5=r1-r5 L6=r2-r6 L7=r3-r7 L9=r2-r1
//   L10=r3-r2 (L4=p-r4, L8=r1-p, L11=r3-p contain p)
// Sub-tets 1,2,3,6 are p-free; 0,4,5,7,8 have det = G.p + D and
// |J|^2 = alpha|p|^2 + S.p + K (alpha = 3 for subs 4 and 8, else 1).
// Selection order: p-free sub-tets (ring order; subs 1,2,3,6) then
// p-dependent sub-tets (ring order; subs 0,4,5,7,8).
struct HexP { float gx, gy, gz, D, sx, sy, sz, K; };   // slot 1 and 4 have alpha 3

__device__ __forceinline__ void subtet(f3 t0, f3 t1, f3 t2, float& det, float& F)
{
    det = dot3(t0, cross3(t1, t2));
    F = (dot3(t0, t0) + dot3(t1, t1)) + dot3(t2, t2);
}

// r[1..7] local ring points; outputs p-free (cd, cF)[4] and p-dependent hp[5]
__device__ __forceinline__ void hex_consts(const f3* r, float* cd, float* cF, HexP* hp)
{
    f3 L0 = sub3(r[5], r[4]), L1 = sub3(r[6], r[5]), L2 = sub3(r[7], r[6]);
    f3 L3 = sub3(r[7], r[4]), L5 = sub3(r[1], r[5]), L6 = sub3(r[2], r[6]);
    f3 L7 = sub3(r[3], r[7]), L9 = sub3(r[2], r[1]), L10 = sub3(r[3], r[2]);
    subtet(L1, neg3(L0), L5, cd[0], cF[0]);
    subtet(L2, neg3(L1), L6, cd[1], cF[1]);
    subtet(neg3(L3), neg3(L2), L7, cd[2], cF[2]);
    subtet(neg3(L9), L10, neg3(L6), cd[3], cF[3]);
    f3 N;
    // sub 0: det = (L0 x L3).(p - r4), |J|^2 = |L0|^2 + |L3|^2 + |p - r4|^2
    N = cross3(L0, L3);
    hp[0] = { N.x, N.y, N.z, -dot3(N, r[4]), -2.0f * r[4].x, -2.0f * r[4].y, -2.0f * r[4].z,
              (dot3(L0, L0) + dot3(L3, L3)) + dot3(r[4], r[4]) };
    // sub 4: (r3-p, r1-p, r4-p): det = N.r3 - N.p with N = (r1-r3)x(r4-r3)
    N = cross3(sub3(r[1], r[3]), sub3(r[4], r[3]));
    {
        f3 s = add3(add3(r[3], r[1]), r[4]);
        hp[1] = { -N.x, -N.y, -N.z, dot3(N, r[3]), -2.0f * s.x, -2.0f * s.y, -2.0f * s.z,
                  (dot3(r[3], r[3]) + dot3(r[1], r[1])) + dot3(r[4], r[4]) };
    }
    // sub 5: (p-r1, L9, r5-r1): det = (L9 x (r5-r1)).(p - r1)
    N = cross3(L9, neg3(L5));
    hp[2] = { N.x, N.y, N.z, -dot3(N, r[1]), -2.0f * r[1].x, -2.0f * r[1].y, -2.0f * r[1].z,
              (dot3(L9, L9) + dot3(L5, L5)) + dot3(r[1], r[1]) };
    // sub 7: (r2-r3, p-r3, r7-r3): det = ((r7-r3) x (r2-r3)).(p - r3)
    N = cross3(neg3(L7), neg3(L10));
    hp[3] = { N.x, N.y, N.z, -dot3(N, r[3]), -2.0f * r[3].x, -2.0f * r[3].y, -2.0f * r[3].z,
              (dot3(L10, L10) + dot3(L7, L7)) + dot3(r[3], r[3]) };
    // sub 8: principal axes A0 - p, A1 - p, A2 + p
    {
        f3 A0 = add3(add3(add3(L0, sub3(r[6], r[7])), sub3(r[2], r[3])), r[1]);
        f3 A1 = add3(add3(add3(L3, L1), L9), r[3]);
        f3 A2 = sub3(add3(add3(L5, L6), L7), r[4]);
        f3 c01 = cross3(A0, A1), c12 = cross3(A1, A2), c20 = cross3(A2, A0);
        f3 s = sub3(sub3(A2, A0), A1);
        hp[4] = { (c01.x - c12.x) - c20.x, (c01.y - c12.y) - c20.y, (c01.z - c12.z) - c20.z,
                  dot3(A0, c12), 2.0f * s.x, 2.0f * s.y, 2.0f * s.z,
                  (dot3(A0, A0) + dot3(A1, A1)) + dot3(A2, A2) };
    }
}

struct HexAcc { float bd2, bF3, dmin, fmin; };
__device__ __forceinline__ HexAcc hex_acc0() { return { 1.0f, 0.0f, 3.40282347e+38f, 3.40282347e+38f }; }
__device__ __forceinline__ void hex_acc_add(HexAcc& a, float det, float F)
{
    float d2 = det * det;
    float F3 = (F * F) * F;
    bool take = d2 * a.bF3 < a.bd2 * F3;
    a.bd2 = take ? d2 : a.bd2;
    a.bF3 = take ? F3 : a.bF3;
    a.dmin = fminf(a.dmin, det);
    a.fmin = fminf(a.fmin, F);
}
__device__ __forceinline__ float hex_acc_score(const HexAcc& a, float epsD, float epsF)
{
    if (a.dmin < epsD || a.fmin < epsF) return 0.0f;
    return 3.0f * cbrt_nr(a.bd2 / a.bF3);
}

// ---------------------------------------------------------------- tri 2D
// Ring element (a, b) of a fan: cross(p) = (a-p)x(b-a) = C + gx px + gy py,
// L(p) = 2|p|^2 + S.p + K with d = b - a, C = ax*dy - ay*dx, g = (-dy, dx),
// S = -2(a+b), K = (|a|^2 + |b|^2) + |d|^2.
//   Ty = fmaf(gy,py,C)  cr = fmaf(gx,px,Ty)
//   Ry = fmaf(sy,py, K + 2*(py*py))  L = fmaf(px, 2*px + sx, Ry)
// Selection of min cr/L; q = (kCTri*cr)/L; negative ring -> most negative cr.
struct TriC { float gx, gy, C, sx, sy, K; };

__device__ __forceinline__ TriC tri_consts(float ax, float ay, float bx, float by)
{
    TriC t;
    float dx = bx - ax, dy = by - ay;
    t.C = ax * dy - ay * dx;
    t.gx = -dy; t.gy = dx;
    t.sx = -2.0f * (ax + bx); t.sy = -2.0f * (ay + by);
    t.K = (fmaf(ax, ax, ay * ay) + fmaf(bx, bx, by * by)) + fmaf(dx, dx, dy * dy);
    return t;
}

struct TriAcc { float bc, bL, cmin; };
__device__ __forceinline__ TriAcc tri_acc0() { return { 1.0f, 0.0f, 3.40282347e+38f }; }
__device__ __forceinline__ void tri_acc_add(TriAcc& a, float cr, float L)
{
    bool take = cr * a.bL < a.bc * L;
    a.bc = take ? cr : a.bc;
    a.bL = take ? L : a.bL;
    a.cmin = fminf(a.cmin, cr);
}
__device__ __forceinline__ float tri_acc_score(const TriAcc& a, float invs2)
{
    if (a.cmin < 0.0f) return a.cmin * invs2;
    if (a.cmin == 0.0f) return 0.0f;
    return (kCTri * a.bc) / a.bL;
}

// grid coordinate of index i on [gmin, gmax] (reference DMOTetImplCub.h:101-103)
__device__ __forceinline__ float grid_coord(float aff, int i, int nq, float gmin, float gmax)
{
    return aff * ((float)i * gmin + (float)(nq - 1 - i) * gmax);
}

}  // namespace vsdmo
