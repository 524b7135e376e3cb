/*
 * or_canon.c -- canonical FP32 numerics and element metrics of the oracle
 * (TEST INFRASTRUCTURE).
 *
 * Compiled with -ffp-contract=off: every multiply and add rounds on its own;
 * fused operations appear only as explicit fmaf() calls.
 */
#include "vsdmo_oracle.h"
#include "or_internal.h"

/* Cube root for x >= 0.
 * Method: the exponent/3 bit-split initial guess for x^(-1/3)
 *   bits(r0) = 0x54a8a9ab - bits(x)/3
 * (constant fitted by minimising the max relative error of r0 over one period
 * of three binades: 7.0e-2), three Newton-Raphson steps on r^-3 = x
 *   r <- r + r*(1 - x*r^3)/3,
 * then y = x*r*r and one Newton step on y^3 = x using r^2 ~ 1/y^2
 *   y <- y - (y^3 - x) * r^2 / 3.
 * Max error 0.73 ulp over [2^-126, 2^128).  Denormal inputs are pre-scaled by
 * 2^24 (result rescaled by 2^-8).  0 -> 0, inf -> inf. */
float or_cbrt(float x)
{
    if (!(x > 0.0f)) return x;
    if (x > FLT_MAX) return x;
    float s = 1.0f;
    if (x < FLT_MIN) { x = x * 16777216.0f; s = 0.00390625f; }
    const float third = 1.0f / 3.0f;
    uint32_t i = f2u(x);
    float r = u2f(0x54a8a9abu - i / 3u);
    for (int it = 0; it < 3; ++it) {
        float r2 = r * r;
        float r3 = r2 * r;
        float t = fmaf(-x, r3, 1.0f);
        float u = t * r;
        r = fmaf(u, third, r);
    }
    float y = (x * r) * r;
    float c = fmaf(y * y, y, -x);
    float rr = r * r;
    y = fmaf(-c, rr * third, y);
    return y * s;
}

/* 2^-e for m = f*2^e, f in [1,2); 1 for m == 0, denormal, inf or NaN. */
float or_pow2_scale(float m)
{
    uint32_t e = (f2u(m) >> 23) & 0xffu;
    if (e == 0u || e >= 254u) return 1.0f;
    return u2f((254u - e) << 23);
}

static inline float inv_pow(float sigma, int k)
{
    float is = 1.0f / sigma, r = 1.0f;
    for (int i = 0; i < k; ++i) r = r * is;
    return r;
}

float or_ring_score_tet(int n, const float* ring, const float* x0p, float sigma, const float* p)
{
    v3 x0 = ld3(x0p);
    tetc cs[512];
    if (n > 512) return NAN;
    for (int e = 0; e < n; ++e)
        cs[e] = tet_consts(loc3(ring + 9 * e, x0, sigma), loc3(ring + 9 * e + 3, x0, sigma), loc3(ring + 9 * e + 6, x0, sigma));
    v3 pl = loc3(p, x0, sigma);
    return tet_score(n, cs, pl.x, pl.y, pl.z, inv_pow(sigma, 3));
}

float or_ring_score_hex(int n, const float* ring, const float* x0p, float sigma, const float* p)
{
    return or_ring_score_hex_m(n, ring, x0p, sigma, p, OR_HEX_SHAPE9);
}

float or_ring_score_hex_m(int n, const float* ring, const float* x0p, float sigma, const float* p, int metric)
{
    v3 x0 = ld3(x0p);
    float cd[4 * 64], cF[4 * 64];
    hexp hp[5 * 64];
    if (n > 64) return NAN;
    for (int e = 0; e < n; ++e) {
        v3 r[8];
        r[0] = x0;
        for (int k = 1; k < 8; ++k) r[k] = loc3(ring + 21 * e + 3 * (k - 1), x0, sigma);
        hex_consts(r, cd + 4 * e, cF + 4 * e, hp + 5 * e);
    }
    float s2 = sigma * sigma, s3 = s2 * sigma;
    hexacc a = hex_const_acc(n, cd, cF);
    v3 pl = loc3(p, x0, sigma);
    return hex_score(n, a, hp, hex_nsub(metric), pl.x, pl.y, pl.z, HEX_EPS * s3, HEX_EPS * s2);
}

float or_ring_score_tri2d(int n, const float* ring, const float* x0p, float sigma, const float* p)
{
    tric cs[512];
    if (n > 512) return NAN;
    for (int e = 0; e < n; ++e) {
        const float* r = ring + 6 * e;
        cs[e] = tri_consts((r[0] - x0p[0]) * sigma, (r[1] - x0p[1]) * sigma,
                           (r[3] - x0p[0]) * sigma, (r[4] - x0p[1]) * sigma);
    }
    float px = (p[0] - x0p[0]) * sigma, py = (p[1] - x0p[1]) * sigma;
    return tri_score(n, cs, px, py, inv_pow(sigma, 2));
}

/* ------------------------------------------------- element qualities */
/* element quality = ring score of the element seen from its vertex 0 at its
   current position (the frame origin is vertex 0, sigma from the element's
   extent) */
static float ext_scale(const float* pts, int n, const float* x0, int dims)
{
    float m = 0.0f;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < dims; ++k) m = fmaxf(m, fabsf(pts[3 * i + k] - x0[k]));
    return or_pow2_scale(m);
}

float or_q_tet(const float* p) { return or_ring_score_tet(1, p + 3, p, ext_scale(p + 3, 3, p, 3), p); }
float or_q_hex(const float* p) { return or_q_hex_m(p, OR_HEX_SHAPE9); }
float or_q_hex_m(const float* p, int metric)
{
    return or_ring_score_hex_m(1, p + 3, p, ext_scale(p + 3, 7, p, 3), p, metric);
}
float or_q_tri2d(const float* p)
{
    float ring[6] = { p[3], p[4], p[5], p[6], p[7], p[8] };
    return or_ring_score_tri2d(1, ring, p, ext_scale(p + 3, 2, p, 2), p);
}

/* ---------------------------------------------- FP64 literal formulas */
typedef struct { double x, y, z; } d3;
static d3 dld(const float* p) { d3 r = { p[0], p[1], p[2] }; return r; }
static d3 dsub(d3 a, d3 b) { d3 r = { a.x - b.x, a.y - b.y, a.z - b.z }; return r; }
static d3 dadd(d3 a, d3 b) { d3 r = { a.x + b.x, a.y + b.y, a.z + b.z }; return r; }
static d3 dneg(d3 a) { d3 r = { -a.x, -a.y, -a.z }; return r; }
static double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 dcross(d3 a, d3 b) { d3 r = { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; return r; }

/* reference DMOQualityTet.h:119-143 in double */
double or_q_tet64(const float* p)
{
    d3 v[4];
    for (int i = 0; i < 4; ++i) v[i] = dld(p + 3 * i);
    d3 e0 = dsub(v[1], v[0]), e1 = dsub(v[2], v[0]), e2 = dsub(v[3], v[0]);
    double vol = (1.0 / 6.0) * ddot(e0, dcross(e1, e2));
    if (vol < 0.0) return vol;
    double mu = cbrt(vol * vol);
    d3 e3 = dsub(v[2], v[1]), e4 = dsub(v[3], v[1]), e5 = dsub(v[3], v[2]);
    double l = ddot(e0, e0) + ddot(e1, e1) + ddot(e2, e2) + ddot(e3, e3) + ddot(e4, e4) + ddot(e5, e5);
    return 12.0 * cbrt(9.0) * mu / l;
}

/* reference DMOQualityHex.h:54-71 with DMOQualityUtil.h:307-349 in double
   (CORNER8: the first 8 sub-tets, the corner tets, alone) */
double or_q_hex64(const float* p) { return or_q_hex64_m(p, OR_HEX_SHAPE9); }
double or_q_hex64_m(const float* p, int metric)
{
    d3 q[8];
    for (int i = 0; i < 4; ++i) { q[i] = dld(p + 3 * (i + 4)); q[i + 4] = dld(p + 3 * i); }
    d3 L[12], X[3];
    L[0] = dsub(q[1], q[0]); L[4] = dsub(q[4], q[0]); L[8] = dsub(q[5], q[4]);
    L[1] = dsub(q[2], q[1]); L[5] = dsub(q[5], q[1]); L[9] = dsub(q[6], q[5]);
    L[2] = dsub(q[3], q[2]); L[6] = dsub(q[6], q[2]); L[10] = dsub(q[7], q[6]);
    L[3] = dsub(q[3], q[0]); L[7] = dsub(q[7], q[3]); L[11] = dsub(q[7], q[4]);
    X[0] = dadd(dadd(dadd(dsub(q[1], q[0]), dsub(q[2], q[3])), dsub(q[5], q[4])), dsub(q[6], q[7]));
    X[1] = dadd(dadd(dadd(dsub(q[3], q[0]), dsub(q[2], q[1])), dsub(q[7], q[4])), dsub(q[6], q[5]));
    X[2] = dadd(dadd(dadd(dsub(q[4], q[0]), dsub(q[5], q[1])), dsub(q[6], q[2])), dsub(q[7], q[3]));
    d3 T[9][3] = {
        { L[0], L[3], L[4] }, { L[1], dneg(L[0]), L[5] }, { L[2], dneg(L[1]), L[6] },
        { dneg(L[3]), dneg(L[2]), L[7] }, { L[11], L[8], dneg(L[4]) }, { dneg(L[8]), L[9], dneg(L[5]) },
        { dneg(L[9]), L[10], dneg(L[6]) }, { dneg(L[10]), dneg(L[11]), dneg(L[7]) }, { X[0], X[1], X[2] } };
    double ret = 1e30;
    const int nt = metric == OR_HEX_CORNER8 ? 8 : 9;
    for (int i = 0; i < nt; ++i) {
        double sj = ddot(T[i][0], dcross(T[i][1], T[i][2]));
        double Aj = ddot(T[i][0], T[i][0]) + ddot(T[i][1], T[i][1]) + ddot(T[i][2], T[i][2]);
        if (sj < 1.19209e-7 || Aj < 1.19209e-7) return 0.0;
        double v = cbrt(sj * sj) / Aj;
        if (v < ret) ret = v;
    }
    return 3.0 * ret;
}

/* reference DMOQualityTri.h:243-257 in double */
double or_q_tri2d64(const float* p)
{
    double x[3], y[3];
    for (int i = 0; i < 3; ++i) { x[i] = p[3 * i]; y[i] = p[3 * i + 1]; }
    double ex[3], ey[3], l = 0.0;
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3;
        ex[i] = x[j] - x[i]; ey[i] = y[j] - y[i];
        l += ex[i] * ex[i] + ey[i] * ey[i];
    }
    double area = ex[0] * ey[1] - ey[0] * ex[1];
    if (area < 0.0) return area;
    return 2.0 * sqrt(3.0) * area / l;
}
