/*
 * or_internal.h -- canonical FP32 element/ring evaluation shared by the
 * oracle's public metric functions and its search (TEST INFRASTRUCTURE).
 * Include only from oracle sources compiled with -ffp-contract=off.
 */
#ifndef OR_INTERNAL_H
#define OR_INTERNAL_H

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

/* OR_MUTATION k > 0 plants deliberate mistake k in the oracle; the mutation
   tests compile such copies and check that the pinning tests reject them.
   The oracle proper is built with 0. */
#ifndef OR_MUTATION
#define OR_MUTATION 0
#endif

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* mean-ratio constant for det = 6V: 12*cbrt(9)/cbrt(36) = 12/cbrt(4) */
#define C_TET 0x1.e3cf48p+2f
/* 2*sqrt(3) as the reference forms it: 2.f*sqrtf(3.f) (DMOQualityTri.h:256) */
#define C_TRI 0x1.bb67aep+1f
/* EPSILON of the hex shape metric (reference SurfaceCommon.h:8) */
#define HEX_EPS 1.19209e-7f

typedef struct { float x, y, z; } v3;
static inline v3 vsub(v3 a, v3 b) { v3 r = { a.x - b.x, a.y - b.y, a.z - b.z }; return r; }
static inline v3 vadd(v3 a, v3 b) { v3 r = { a.x + b.x, a.y + b.y, a.z + b.z }; return r; }
static inline v3 vneg(v3 a) { v3 r = { -a.x, -a.y, -a.z }; return r; }
static inline v3 vscale(v3 a, float s) { v3 r = { a.x * s, a.y * s, a.z * s }; return r; }
/* canonical cross: each component mul, mul, sub */
static inline v3 vcross(v3 a, v3 b)
{
    v3 r = { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
    return r;
}
/* canonical dot: fmaf(x, x', fmaf(y, y', z*z')) */
static inline float vdot(v3 a, v3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }
static inline v3 ld3(const float* p) { v3 r = { p[0], p[1], p[2] }; return r; }
/* local frame: (p - x0) * sigma, sigma a power of two */
static inline v3 loc3(const float* p, v3 x0, float s) { return vscale(vsub(ld3(p), x0), s); }

float or_cbrt(float x);

/* ------------------------------------------------------------------ tets */
/* Per ring element (a,b,c) in the local frame (p' = (p - x0)*sigma):
 *   n = (b-a)x(c-a), D = n.a          ->  det(p') = 6V = D - n.p'
 *   m2s = -2(a+b+c),
 *   K = |a|^2+|b|^2+|c|^2 + |b-a|^2+|c-a|^2+|c-b|^2
 *                                      ->  L(p') = 3|p'|^2 + m2s.p' + K
 * (L = sum of the six squared edge lengths, reference DMOQualityTet.h:132-139).
 * Evaluation of candidate p' = (px,py,pz):
 *   Qyz = 3*fmaf(pz,pz,py*py)
 *   Tyz = fmaf(-ny,py, fmaf(-nz,pz, D));  det = fmaf(-nx,px, Tyz)
 *   Ryz = fmaf(m2sy,py, fmaf(m2sz,pz, K+Qyz));  L = fmaf(px, 3*px + m2sx, Ryz)
 * The ring minimum of q = C_TET*cbrt(det^2/L^3) is found by the running
 * selection  "det^2 * bestL3 < bestDet2 * L^3"  in ring order; only the
 * selected element is cube-rooted.  A ring with a negative det returns the
 * most negative signed volume det/6 (reference DMOQualityTet.h:126-129); a
 * zero det returns 0. */
typedef struct { float nnx, nny, nnz, D, m2sx, m2sy, m2sz, K; } tetc;

static inline tetc tet_consts(v3 a, v3 b, v3 c)
{
    tetc t;
    v3 e1 = vsub(b, a), e2 = vsub(c, a), e3 = vsub(c, b);
    v3 n = vcross(e1, e2);
    t.D = vdot(n, a);
    t.nnx = -n.x; t.nny = -n.y; t.nnz = -n.z;
    v3 s = vadd(vadd(a, b), c);
    t.m2sx = -2.0f * s.x; t.m2sy = -2.0f * s.y; t.m2sz = -2.0f * s.z;
    float la = vdot(a, a), lb = vdot(b, b), lc = vdot(c, c);
    float l1 = vdot(e1, e1), l2 = vdot(e2, e2), l3 = vdot(e3, e3);
    t.K = ((((la + lb) + lc) + l1) + l2) + l3;
    return t;
}

/* invs3 = sigma^-3 (exact power of two) */
static inline float tet_score(int n, const tetc* cs, float px, float py, float pz, float invs3)
{
    float Qyz = 3.0f * fmaf(pz, pz, py * py);
    float t3 = 3.0f * px;
    float bd2 = 1.0f, bL3 = 0.0f, dmin = FLT_MAX;
    for (int e = 0; e < n; ++e) {
        const tetc* c = cs + e;
        float Tyz = fmaf(c->nny, py, fmaf(c->nnz, pz, c->D));
        float Ryz = fmaf(c->m2sy, py, fmaf(c->m2sz, pz, c->K + Qyz));
        float det = fmaf(c->nnx, px, Tyz);
        float L = fmaf(px, t3 + c->m2sx, Ryz);
        float d2 = det * det;
        float L3 = (L * L) * L;
        if (d2 * bL3 < bd2 * L3) { bd2 = d2; bL3 = L3; }
        dmin = fminf(dmin, det);
    }
    if (dmin < 0.0f) return (dmin * (1.0f / 6.0f)) * invs3;
    if (dmin == 0.0f) return 0.0f;
    return C_TET * or_cbrt(bd2 / bL3);
}

/* ------------------------------------------------------------------ hexes */
/* Hex shape metric (reference DMOQualityHex.h:54-71 with
 * DMOQualityUtil.h:307-349): shapeMetricHex(v) evaluates hex_edges and
 * hex_principal_axes on (v4,v5,v6,v7,v0,v1,v2,v3).  With the candidate p as v0
 * and the halfhex points r1..r7 the edge vectors are
 *   L0=r5-r4 L1=r6-r5 L2=r7-r6 L3=r7-r4 L4=p-r4  L5=r1-r5  L6=r2-r6  L7=r3-r7
 *   L8=r1-p  L9=r2-r1 L10=r3-r2 L11=r3-p
 * Sub-tets 1,2,3,6 do not contain p ("const"); sub-tets 0,4,5,7,8 have a det
 * affine in p (det = G.p + D) and a Frobenius norm alpha|p|^2 + S.p + K with
 * alpha 1 (subs 0,5,7) or 3 (subs 4,8).  q_sub = cbrt(det^2)/F; the element
 * is 0 if any det < eps or F < eps; the ring score is 3*min over all sub-tets.
 * Selection order: const sub-tets (ring order, subs 1,2,3,6), then the
 * p-dependent ones (ring order, subs 0,4,5,7,8).  Evaluation of a p-dependent
 * sub-tet at candidate p':
 *   Q1 = fmaf(pz,pz,py*py), Q3 = 3*Q1, t3 = 3*px
 *   Tyz = fmaf(gy,py, fmaf(gz,pz, D));  det = fmaf(gx,px, Tyz)
 *   Ryz = fmaf(sy,py, fmaf(sz,pz, K + Qa));  F = fmaf(px, (a3 ? t3 : px) + sx, Ryz) */
typedef struct { float gx, gy, gz, D, sx, sy, sz, K; } hexp;   /* alpha: 1 for subs 0,2,3 of hp[], 3 for 1,4 */
#define HEX_A3(s) ((s) == 1 || (s) == 4)

static inline void subtet_df(v3 t0, v3 t1, v3 t2, float* det, float* F)
{
    *det = vdot(t0, vcross(t1, t2));
    *F = (vdot(t0, t0) + vdot(t1, t1)) + vdot(t2, t2);
}

/* r[1..7]: ring points in the local frame; outputs const sub-tets (cd, cF)
   in order 1,2,3,6 and p-dependent constants hp[] in order 0,4,5,7,8 */
static inline void hex_consts(const v3 r[8], float* cd, float* cF, hexp* hp)
{
    v3 L0 = vsub(r[5], r[4]), L1 = vsub(r[6], r[5]), L2 = vsub(r[7], r[6]);
    v3 L3 = vsub(r[7], r[4]), L5 = vsub(r[1], r[5]), L6 = vsub(r[2], r[6]);
    v3 L7 = vsub(r[3], r[7]), L9 = vsub(r[2], r[1]), L10 = vsub(r[3], r[2]);
    subtet_df(L1, vneg(L0), L5, &cd[0], &cF[0]);
    subtet_df(L2, vneg(L1), L6, &cd[1], &cF[1]);
    subtet_df(vneg(L3), vneg(L2), L7, &cd[2], &cF[2]);
    subtet_df(vneg(L9), L10, vneg(L6), &cd[3], &cF[3]);
    v3 N;
    /* sub 0: (L0, L3, p - r4): det = (L0 x L3).(p - r4) */
    N = vcross(L0, L3);
    hp[0].gx = N.x; hp[0].gy = N.y; hp[0].gz = N.z; hp[0].D = -vdot(N, r[4]);
    hp[0].sx = -2.0f * r[4].x; hp[0].sy = -2.0f * r[4].y; hp[0].sz = -2.0f * r[4].z;
    hp[0].K = (vdot(L0, L0) + vdot(L3, L3)) + vdot(r[4], r[4]);
    /* sub 4: (r3 - p, r1 - p, r4 - p): det = N.r3 - N.p, N = (r1-r3)x(r4-r3) */
    N = vcross(vsub(r[1], r[3]), vsub(r[4], r[3]));
    hp[1].gx = -N.x; hp[1].gy = -N.y; hp[1].gz = -N.z; hp[1].D = vdot(N, r[3]);
    {
        v3 s = vadd(vadd(r[3], r[1]), r[4]);
        hp[1].sx = -2.0f * s.x; hp[1].sy = -2.0f * s.y; hp[1].sz = -2.0f * s.z;
    }
    hp[1].K = (vdot(r[3], r[3]) + vdot(r[1], r[1])) + vdot(r[4], r[4]);
    /* sub 5: (p - r1, L9, r5 - r1): det = (L9 x (r5-r1)).(p - r1) */
    N = vcross(L9, vneg(L5));
    hp[2].gx = N.x; hp[2].gy = N.y; hp[2].gz = N.z; hp[2].D = -vdot(N, r[1]);
    hp[2].sx = -2.0f * r[1].x; hp[2].sy = -2.0f * r[1].y; hp[2].sz = -2.0f * r[1].z;
    hp[2].K = (vdot(L9, L9) + vdot(L5, L5)) + vdot(r[1], r[1]);
    /* sub 7: (r2 - r3, p - r3, r7 - r3): det = ((r7-r3) x (r2-r3)).(p - r3) */
    N = vcross(vneg(L7), vneg(L10));
    hp[3].gx = N.x; hp[3].gy = N.y; hp[3].gz = N.z; hp[3].D = -vdot(N, r[3]);
    hp[3].sx = -2.0f * r[3].x; hp[3].sy = -2.0f * r[3].y; hp[3].sz = -2.0f * r[3].z;
    hp[3].K = (vdot(L10, L10) + vdot(L7, L7)) + vdot(r[3], r[3]);
    /* sub 8: principal axes X0 = A0 - p, X1 = A1 - p, X2 = A2 + p;
       det = A0.(A1xA2) + p.((A0xA1) - (A1xA2) - (A2xA0)) */
    {
        v3 A0 = vadd(vadd(vadd(L0, vsub(r[6], r[7])), vsub(r[2], r[3])), r[1]);
        v3 A1 = vadd(vadd(vadd(L3, L1), L9), r[3]);
        v3 A2 = vsub(vadd(vadd(L5, L6), L7), r[4]);
        v3 c01 = vcross(A0, A1), c12 = vcross(A1, A2), c20 = vcross(A2, A0);
        hp[4].gx = (c01.x - c12.x) - c20.x;
        hp[4].gy = (c01.y - c12.y) - c20.y;
        hp[4].gz = (c01.z - c12.z) - c20.z;
        hp[4].D = vdot(A0, c12);
        v3 s = vsub(vsub(A2, A0), A1);
        hp[4].sx = 2.0f * s.x; hp[4].sy = 2.0f * s.y; hp[4].sz = 2.0f * s.z;
        hp[4].K = (vdot(A0, A0) + vdot(A1, A1)) + vdot(A2, A2);
    }
}

typedef struct { float bd2, bF3, dmin, fmin; } hexacc;

/* p-dependent sub-tets per element of a hex metric (OR_HEX_SHAPE9 / _CORNER8) */
static inline int hex_nsub(int metric) { return metric == OR_HEX_CORNER8 ? 4 : 5; }

static inline void hex_sel(hexacc* a, float d, float F)
{
    float d2 = d * d;
    float F3 = (F * F) * F;
    if (d2 * a->bF3 < a->bd2 * F3) { a->bd2 = d2; a->bF3 = F3; }
    a->dmin = fminf(a->dmin, d);
    a->fmin = fminf(a->fmin, F);
}

static inline hexacc hex_acc_init(void) { hexacc a = { 1.0f, 0.0f, FLT_MAX, FLT_MAX }; return a; }

/* const part over n elements: cd/cF hold 4 values per element */
static inline hexacc hex_const_acc(int n, const float* cd, const float* cF)
{
    hexacc a = hex_acc_init();
    for (int e = 0; e < n; ++e)
        for (int s = 0; s < 4; ++s) hex_sel(&a, cd[4 * e + s], cF[4 * e + s]);
    return a;
}

/* nsub: the p-dependent sub-tets scored per element, 5 (the reference's 9
   sub-tets) or 4 (the 8 corner tets: the principal-axes tet hp[4] left out) */
static inline float hex_score(int n, hexacc a, const hexp* hp, int nsub, float px, float py, float pz,
                              float epsD, float epsF)
{
    float Q1 = fmaf(pz, pz, py * py);
    float Q3 = 3.0f * Q1;
    float t3 = 3.0f * px;
#if OR_MUTATION == 3
    nsub = 4;                          /* planted: the principal-axes sub-tet dropped */
#endif
    for (int e = 0; e < n; ++e)
        for (int s = 0; s < nsub; ++s) {
            const hexp* h = hp + 5 * e + s;
            int a3 = HEX_A3(s);
            float Tyz = fmaf(h->gy, py, fmaf(h->gz, pz, h->D));
            float Ryz = fmaf(h->sy, py, fmaf(h->sz, pz, h->K + (a3 ? Q3 : Q1)));
            float det = fmaf(h->gx, px, Tyz);
            float F = fmaf(px, (a3 ? t3 : px) + h->sx, Ryz);
            hex_sel(&a, det, F);
        }
    if (a.dmin < epsD || a.fmin < epsF) return 0.0f;
    return 3.0f * or_cbrt(a.bd2 / a.bF3);
}

/* ------------------------------------------------------------------ tri 2D */
/* Reference meanRatioMetricTri2D (DMOQualityTri.h:243-257) on (p, a, b):
 *   cross(p) = (a-p) x (b-a) = C + gx*px + gy*py      (d = b-a)
 *   L(p) = 2|p|^2 + S.p + K,  S = -2(a+b),  K = |a|^2 + |b|^2 + |d|^2
 * Evaluation: Ty = fmaf(gy,py,C); cr = fmaf(gx,px,Ty);
 *             Ry = fmaf(sy,py, K + 2*(py*py)); L = fmaf(px, 2*px + sx, Ry)
 * Running selection of min cr/L ("cr*bestL < bestCr*L"); q = (C_TRI*cr)/L for
 * the selected element; a ring with negative cross returns the most negative
 * cross (the reference's signed "area"), zero returns 0. */
typedef struct { float gx, gy, C, sx, sy, K; } tric;

static inline tric tri_consts(float ax, float ay, float bx, float by)
{
    tric t;
    float dx = bx - ax, dy = by - ay;
    t.C = ax * dy - ay * dx;
    t.gx = -dy; t.gy = dx;
    t.sx = -2.0f * (ax + bx); t.sy = -2.0f * (ay + by);
    t.K = (fmaf(ax, ax, ay * ay) + fmaf(bx, bx, by * by)) + fmaf(dx, dx, dy * dy);
    return t;
}

static inline float tri_score(int n, const tric* cs, float px, float py, float invs2)
{
    float Q2 = 2.0f * (py * py);
    float t2 = 2.0f * px;
    float bc = 1.0f, bL = 0.0f, cmin = FLT_MAX;
    for (int e = 0; e < n; ++e) {
        const tric* c = cs + e;
        float Ty = fmaf(c->gy, py, c->C);
        float Ry = fmaf(c->sy, py, c->K + Q2);
        float cr = fmaf(c->gx, px, Ty);
        float L = fmaf(px, t2 + c->sx, Ry);
        if (cr * bL < bc * L) { bc = cr; bL = L; }
        cmin = fminf(cmin, cr);
    }
    if (cmin < 0.0f) return cmin * invs2;
    if (cmin == 0.0f) return 0.0f;
    return (C_TRI * bc) / bL;
}

#endif
