/*
 * or_surface.c -- oracle restatement of surface-vertex DMO (TEST
 * INFRASTRUCTURE): local quadric fits, feature-sector fits, 1D feature-curve
 * fits, the projection of (u,v) candidates onto the blended local surfaces,
 * the nearest-neighbour walk and the 2D / 1D surface searches.
 *
 * Reference files (VSDMO/src/dmo/...):
 *   surfaceP/Surface.h:18-93, surfaceP/Surface.cu:11-121   quadric fit + frame
 *   surfaceP/SurfaceFeature.h:79-106, .cu:12-223             feature sectors
 *   Surface1D.h:13-114, Surface1D.cu:7-103                    1D curve fit
 *   SurfaceEstimation.h:122-197 (localSurfaceEstimationWithFeature), Surface1D.h:54-93
 *   DMOCommon.cu:11-43 (k_updateNearestNeighbor), DMOCommon.h:130-165 (k_calcOptFeatureVec)
 *   DMOHexImplCub.h:141-318 / DMOTetImplCub.h:141-322 (2D), :320-489 / :326-494 (1D)
 *   DMOHex.cu:131-182, DMOTet.cu:128-170 (iteration order)
 * Arithmetic follows the reference expressions literally with contraction
 * off: dot = (x*x' + y*y') + z*z', norm = sqrtf(dot), normalize divides each
 * component, Vec3 sums left to right.  3x3 systems are solved by an LU with
 * partial pivoting (first maximal |pivot|, division by the pivot, then
 * forward and back substitution), the textbook getrf/getrs recipe.
 */
#include "vsdmo_oracle.h"
#include "or_internal.h"
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAX_NEIGH 40
#define TABLE_SIZE 20
#define SEPS 1.19209e-07f

typedef struct { float x, y, z; } w3;
static inline w3 W(const float* p) { w3 r = { p[0], p[1], p[2] }; return r; }
static inline w3 wsub(w3 a, w3 b) { w3 r = { a.x - b.x, a.y - b.y, a.z - b.z }; return r; }
static inline w3 wadd(w3 a, w3 b) { w3 r = { a.x + b.x, a.y + b.y, a.z + b.z }; return r; }
static inline w3 wmul(float s, w3 a) { w3 r = { s * a.x, s * a.y, s * a.z }; return r; }
static inline float wdot(w3 a, w3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline w3 wcross(w3 a, w3 b) { w3 r = { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; return r; }
static inline float wnorm(w3 a) { return sqrtf((a.x * a.x + a.y * a.y) + a.z * a.z); }
static inline w3 wnormalize(w3 a) { float l = wnorm(a); w3 r = { a.x / l, a.y / l, a.z / l }; return r; }

/* localSurface: p0, e0, e1, e2, a, b, c (15 floats) + numPoints */
typedef struct { w3 p0, e[3]; float a, b, c; int np; } lsurf;
typedef struct { w3 p0, e[2]; float a; } lsurf1;

static inline float poly(const lsurf* s, float u, float v) { return (s->a * u * u + s->b * u * v) + s->c * v * v; }
static inline w3 l2w(const lsurf* s, float u, float v, float w)
{
    return wadd(wadd(wadd(s->p0, wmul(u, s->e[0])), wmul(v, s->e[1])), wmul(w, s->e[2]));
}

/* 3x3 LU with partial pivoting + solve; A column-major as in the reference */
static void lu_solve3(const float* Acm, const float* rhs, float* x)
{
    float A[3][3], b[3];
    for (int i = 0; i < 3; ++i) { for (int j = 0; j < 3; ++j) A[i][j] = Acm[i + 3 * j]; b[i] = rhs[i]; }
    int piv[3] = { 0, 1, 2 };
    for (int j = 0; j < 3; ++j) {
        int p = j;
        float mx = fabsf(A[j][j]);
        for (int i = j + 1; i < 3; ++i) if (fabsf(A[i][j]) > mx) { mx = fabsf(A[i][j]); p = i; }
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
    for (int i = 1; i < 3; ++i) for (int k = 0; k < i; ++k) y[i] = y[i] - A[i][k] * y[k];
    for (int i = 2; i >= 0; --i) {
        for (int k = i + 1; k < 3; ++k) y[i] = y[i] - A[i][k] * y[k];
        y[i] = y[i] / A[i][i];
    }
    x[0] = y[0]; x[1] = y[1]; x[2] = y[2];
}

/* frame from a normal and a neighbour (computeLocalCoordinates, Surface.h:66-93) */
static void frame3(w3 n, w3 r, w3* e0, w3* e1, w3* e2)
{
    if (wnorm(wcross(n, r)) < 10.0f * SEPS) { w3 rr = { n.x, n.z, -n.y }; r = rr; }
    w3 t1 = wcross(n, r);
    w3 t2 = wcross(n, t1);
    *e2 = wnormalize(n);
    *e0 = wnormalize(t1);
    *e1 = wnormalize(t2);
}

/* accumulate the normal equations of w = a u^2 + b uv + c v^2 and solve */
static void fit_quadric(lsurf* s, const w3* pts, int np)
{
    float M[9] = { 0 }, R[3] = { 0 };
    for (int i = 0; i < np; ++i) {
        w3 p = wsub(pts[i], s->p0);
        float u = wdot(s->e[0], p), v = wdot(s->e[1], p), w = wdot(s->e[2], p);
        R[0] += u * u * w; R[1] += u * v * w; R[2] += v * v * w;
        M[0] += u * u * u * u; M[1] += u * u * u * v; M[4] += u * u * v * v;
        M[2] += u * u * v * v; M[5] += u * v * v * v; M[8] += v * v * v * v;
    }
    M[3] = M[1]; M[6] = M[2]; M[7] = M[5];
    float abc[3];
    lu_solve3(M, R, abc);
    s->a = abc[0]; s->b = abc[1]; s->c = abc[2];
    s->np = np;
    if (np <= 4) { s->a = 0.0f; s->b = 0.0f; s->c = 0.0f; }
    if (s->a != s->a) s->a = 0.0f;
    if (s->b != s->b) s->b = 0.0f;
    if (s->c != s->c) s->c = 0.0f;
}

struct or_surf {
    int nS, nFree, nF;
    int err;              /* 1: a projection found no sector, 2: unsectored neighbours, 4: a ring-2 fit
                             would hold more than MAX_FIT points (as the kernels' bits) */
    float* pinit;         /* 3*nS: surface positions at optimiser creation */
    float* vnorm;         /* 3*nS */
    float* fnorm;         /* 3*nFaces */
    lsurf* ls;            /* nS */
    lsurf1* ls1;          /* nS (feature vertices) */
    int* frow;            /* nFeature+1 */
    int* table;           /* nFeature*TABLE_SIZE */
    lsurf* fs;            /* feature sector surfaces */
    int* nn;              /* nS */
    unsigned char* opt1d; /* nS */
};

static const float* P(const or_mesh* m, int v) { return m->xyz + 3 * (size_t)v; }

/* face normals (k_calcFaceNormals) and vertex normals (k_calcVertexNormals):
   normalized sum of incident face normals in CCW order */
static void normals(const or_mesh* m, float* fn, float* vn)
{
    int k = m->fpe;
    for (int f = 0; f < m->nFaces; ++f) {
        const int* F = m->faces + (size_t)f * k;
        w3 p0 = W(P(m, F[0])), p1 = W(P(m, F[1])), p2 = W(P(m, F[k == 3 ? 2 : 3]));
        w3 n = wnormalize(wcross(wsub(p1, p0), wsub(p2, p0)));
        fn[3 * f] = n.x; fn[3 * f + 1] = n.y; fn[3 * f + 2] = n.z;
    }
    for (int v = 0; v < m->nSurf; ++v) {
        w3 n = { 0.0f, 0.0f, 0.0f };
        for (int h = m->heRow[v]; h < m->heRow[v + 1]; ++h)
            if (m->heFace[h] != -1) n = wadd(n, W(fn + 3 * m->heFace[h]));
        n = wnormalize(n);
        vn[3 * v] = n.x; vn[3 * v + 1] = n.y; vn[3 * v + 2] = n.z;
    }
}

#define MAX_FIT 64

/* the points of v's quadric fit: v, its one-ring in CCW order (at most
   MAX_NEIGH points in all, Surface.cu:15-28) and, for ring 2, each one-ring
   neighbour's CCW targets not taken yet, neighbour by neighbour; returns
   the count, or -1 when a ring-2 set exceeds MAX_FIT points */
static int fit_points(const or_mesh* m, int v, int ring, w3* pts)
{
    int ids[MAX_FIT];
    int np = 0;
    ids[np] = v; pts[np++] = W(P(m, v));
    for (int h = m->heRow[v]; h < m->heRow[v + 1] && np < MAX_NEIGH; ++h) {
        ids[np] = m->heTarget[h];
        pts[np++] = W(P(m, m->heTarget[h]));
    }
    if (ring == 2) {
        const int n1 = np;
        for (int k = 1; k < n1; ++k) {
            const int d = ids[k];
            for (int h = m->heRow[d]; h < m->heRow[d + 1]; ++h) {
                const int t = m->heTarget[h];
                int seen = 0;
#if OR_MUTATION == 17
                seen = (t == v);   /* planted: only v is skipped, so shared neighbours enter twice */
#else
                for (int q = 0; q < np && !seen; ++q) seen = ids[q] == t;
#endif
                if (seen) continue;
                if (np == MAX_FIT) return -1;
                ids[np] = t;
                pts[np++] = W(P(m, t));
            }
        }
    }
    return np;
}

or_surf* or_surface_init(const or_mesh* m) { return or_surface_init_ring(m, 1); }

or_surf* or_surface_init_ring(const or_mesh* m, int ring)
{
    or_surf* s = (or_surf*)calloc(1, sizeof(or_surf));
    int nS = m->nSurf;
    s->nS = nS; s->nFree = m->nFree;
    int nF = m->nFeature;
    s->nF = nF;
    s->pinit = (float*)malloc(sizeof(float) * 3 * (size_t)(nS > 0 ? nS : 1));
    memcpy(s->pinit, m->xyz, sizeof(float) * 3 * (size_t)nS);
    s->vnorm = (float*)malloc(sizeof(float) * 3 * (size_t)(nS > 0 ? nS : 1));
    s->fnorm = (float*)malloc(sizeof(float) * 3 * (size_t)(m->nFaces > 0 ? m->nFaces : 1));
    normals(m, s->fnorm, s->vnorm);
    s->ls = (lsurf*)calloc((size_t)(nS > 0 ? nS : 1), sizeof(lsurf));
    s->ls1 = (lsurf1*)calloc((size_t)(nS > 0 ? nS : 1), sizeof(lsurf1));
    s->nn = (int*)malloc(sizeof(int) * (size_t)(nS > 0 ? nS : 1));
    s->opt1d = (unsigned char*)calloc((size_t)(nS > 0 ? nS : 1), 1);
    for (int v = 0; v < nS; ++v) s->nn[v] = v;

    /* 1D curve fits of feature vertices (Surface1D.cu:7-91) */
    for (int v = 0; v < nS; ++v) {
        if (!m->isFeature[v]) continue;
        int nb[MAX_NEIGH + 1], np = 0, vf = -1;
        nb[np++] = v;
        for (int h = m->heRow[v]; h < m->heRow[v + 1]; ++h) {
            int d = m->heTarget[h];
            if (m->isFeature[d] && np < MAX_NEIGH) { nb[np++] = d; vf = d; }
        }
        if (vf < 0) continue;
        w3 n = m->flat ? (w3){ 0.0f, 0.0f, 1.0f } : W(s->vnorm + 3 * v);
        w3 r = wsub(W(P(m, vf)), W(P(m, v)));
        if (wnorm(wcross(n, r)) < 10.0f * SEPS) { w3 rr = { n.x, n.z, -n.y }; r = rr; }
        w3 t1 = wcross(n, r), t2 = wcross(n, t1);
        n = wnormalize(n); t2 = wnormalize(t2);
        lsurf1* q = &s->ls1[v];
        q->p0 = W(P(m, v)); q->e[0] = t2; q->e[1] = n;
        float ATA = 0.0f, ATb = 0.0f;
        for (int i = 0; i < np; ++i) {
            w3 p = wsub(W(P(m, nb[i])), q->p0);
            float u = wdot(q->e[0], p), vv = wdot(q->e[1], p);
            ATb += u * u * vv;
            ATA += u * u * u * u;
        }
        q->a = ATb / ATA;
    }
    /* quadric fits of all surface vertices (Surface.cu:11-105), over the
       one-ring or the 2-ring */
    for (int v = 0; v < nS; ++v) {
        w3 pts[MAX_FIT + 1];
        int np = fit_points(m, v, ring, pts);
        if (np < 0) { s->err |= 4; np = MAX_FIT; }
        if (np == 5) pts[np++] = wmul(0.5f, wadd(pts[2], pts[3]));
        lsurf* q = &s->ls[v];
        q->p0 = pts[0];
        int first = m->heTarget[m->heRow[v]];
        frame3(W(s->vnorm + 3 * v), wsub(W(P(m, first)), W(P(m, v))), &q->e[0], &q->e[1], &q->e[2]);
        fit_quadric(q, pts, np);
    }
    /* feature sectors (SurfaceFeature.cu:134-223) */
    s->frow = (int*)calloc((size_t)nF + 1, sizeof(int));
    s->table = (int*)malloc(sizeof(int) * (size_t)(nF > 0 ? nF : 1) * TABLE_SIZE);
    for (int i = 0; i < nF * TABLE_SIZE; ++i) s->table[i] = -1;
    for (int i = 0; i < nF; ++i) {
        int v = m->nFree + i, cnt = 0;
        for (int h = m->heRow[v] + 1; h < m->heRow[v + 1]; ++h) if (m->isFeature[m->heTarget[h]]) ++cnt;
        if (!m->isBoundary1D[v]) ++cnt;
        s->frow[i + 1] = s->frow[i] + cnt;
    }
    int nSec = s->frow[nF];
    s->fs = (lsurf*)calloc((size_t)(nSec > 0 ? nSec : 1), sizeof(lsurf));
    int* secN = (int*)malloc(sizeof(int) * (size_t)(nSec > 0 ? nSec : 1) * MAX_NEIGH);
    int* secCnt = (int*)calloc((size_t)(nSec > 0 ? nSec : 1), sizeof(int));
    for (int i = 0; i < nF; ++i) {
        int v = m->nFree + i;
        int h0 = m->heRow[v], h1 = m->heRow[v + 1];
        int sid = 0, nneigh = 0, lastWasFeature = 1, last = h0, vindex = 1;
        s->table[TABLE_SIZE * i] = 0;
        const int nsec = s->frow[i + 1] - s->frow[i];
        /* neighbours after the last feature neighbour of an open fan belong
           to no sector (sid == nsec): recorded in err, left out of the lists */
        for (int h = h0 + 1; h < h1; ++h) {
            int d = m->heTarget[h];
            int* L = secN + (size_t)MAX_NEIGH * (s->frow[i] + (sid < nsec ? sid : 0));
            const int own = sid < nsec;
            if (!own) s->err |= 2;
            if (lastWasFeature) { nneigh = 0; if (own) { L[nneigh++] = v; L[nneigh++] = m->heTarget[last]; } }
            if (m->isFeature[d]) {
#if OR_MUTATION == 16                   /* 16 plants: a cut mapped to the sector it opens */
                if (vindex < TABLE_SIZE) s->table[TABLE_SIZE * i + vindex] = sid + 1;
#else
                if (vindex < TABLE_SIZE) s->table[TABLE_SIZE * i + vindex] = sid;
#endif
                if (own && nneigh < MAX_NEIGH) L[nneigh++] = d;
                if (own) secCnt[s->frow[i] + sid] = nneigh;
                ++sid; nneigh = 0; lastWasFeature = 1;
            } else {
                if (vindex < TABLE_SIZE) s->table[TABLE_SIZE * i + vindex] = sid;
                if (own && nneigh < MAX_NEIGH) L[nneigh++] = d;
                lastWasFeature = 0;
            }
            ++last; ++vindex;
        }
        if (!m->isBoundary1D[v] && sid < nsec) {
            int* L = secN + (size_t)MAX_NEIGH * (s->frow[i] + sid);
            if (lastWasFeature) { nneigh = 0; L[nneigh++] = v; L[nneigh++] = m->heTarget[last]; }
            if (nneigh < MAX_NEIGH) L[nneigh++] = m->heTarget[h0];
            secCnt[s->frow[i] + sid] = nneigh;
        }
    }
    for (int q = 0; q < nSec; ++q) {
        const int* L = secN + (size_t)MAX_NEIGH * q;
        int np = secCnt[q];
        if (np < 2) continue;            /* not reached for closed surfaces */
        int v = L[0];
        w3 pts[MAX_NEIGH + 1];
        for (int i = 0; i < np; ++i) pts[i] = W(P(m, L[i]));
        if (np == 5) { pts[np] = wmul(0.5f, wadd(pts[2], pts[3])); ++np; }
        w3 n = { 0.0f, 0.0f, 0.0f };
        for (int h = m->heRow[v]; h < m->heRow[v + 1]; ++h)
            if (m->heTarget[h] == L[1] && m->heFace[h] >= 0) n = W(s->fnorm + 3 * m->heFace[h]);
        lsurf* fq = &s->fs[q];
        fq->p0 = pts[0];
        frame3(n, wsub(W(P(m, L[1])), W(P(m, v))), &fq->e[0], &fq->e[1], &fq->e[2]);
        fit_quadric(fq, pts, np);
    }
    free(secN); free(secCnt);
    /* feature eligibility for the 1D search (DMOCommon.h:130-165): two
       feature neighbours enclosing >= 160 degrees, written as cos <= cos 160 */
    const float c160 = (float)cos(160.0 * 3.14159265358979323846 / 180.0);
    for (int v = 0; v < nS; ++v) {
        if (!m->isFeature[v]) continue;
        int nn = 0; w3 pp[2];
        for (int h = m->heRow[v]; h < m->heRow[v + 1]; ++h) {
            int d = m->heTarget[h];
            if (m->isFeature[d]) { if (nn < 2) pp[nn] = W(P(m, d)); ++nn; }
        }
        if (nn == 2) {
            w3 a = wsub(pp[0], W(P(m, v))), b = wsub(pp[1], W(P(m, v)));
            float c = wdot(a, b) / sqrtf(wdot(a, a) * wdot(b, b));
            if (c <= c160) s->opt1d[v] = 1;
        }
    }
    return s;
}

void or_surface_free(or_surf* s)
{
    if (!s) return;
    free(s->pinit); free(s->vnorm); free(s->fnorm); free(s->ls); free(s->ls1);
    free(s->frow); free(s->table); free(s->fs); free(s->nn); free(s->opt1d);
    free(s);
}

/* projection of (u,v) (localSurfaceEstimationWithFeature, SurfaceEstimation.h:122-197) */
static w3 project2d(const or_mesh* m, const or_surf* s, const float* snap, int vid, const float* vidPos, float u, float v)
{
    const lsurf* L = &s->ls[s->nn[vid]];
    float w = poly(L, u, v);
    w3 pw = l2w(L, u, v, w);
    w3 pl[MAX_NEIGH + 1];
    float inv[MAX_NEIGH + 1];
    float isum = 0.0f;
    pl[0] = pw;
    {
        w3 d = wsub(W(s->pinit + 3 * s->nn[vid]), pw);
        inv[0] = 1.0f / (wdot(d, d) + SEPS);
#if OR_MUTATION == 4
        inv[0] = inv[0] * inv[0];      /* planted: blend weight squared */
#endif
        isum += inv[0];
    }
    int h0 = m->heRow[vid], nh = m->heRow[vid + 1] - h0;
    if (nh > MAX_NEIGH) nh = MAX_NEIGH;
    for (int i = 0; i < nh; ++i) {
        int vn = m->heTarget[h0 + i];
        int nn = s->nn[vn];
        w3 d = wsub(W(s->pinit + 3 * nn), pw);
        inv[i + 1] = 1.0f / (wdot(d, d) + SEPS);
#if OR_MUTATION == 4
        inv[i + 1] = inv[i + 1] * inv[i + 1];
#endif
        isum += inv[i + 1];
        const lsurf* LN;
        if (m->isFeature[nn]) {
            int off = nn - s->nFree;
            float md = 3.40282347e+38f;
            int arg = 0, idx = 0;
            w3 vp = W(vidPos);
            for (int h = m->heRow[nn]; h < m->heRow[nn + 1]; ++h, ++idx) {
                w3 t = wsub(W(snap + 3 * (size_t)m->heTarget[h]), vp);
                float dd = wdot(t, t);
                if (dd < md) { md = dd; arg = idx; }
            }
            int sid = (arg < TABLE_SIZE) ? s->table[TABLE_SIZE * off + arg] : -1;
            const int nsec = s->frow[off + 1] - s->frow[off];
            if (sid < 0 || sid >= nsec) {
                or_surf* sw = (or_surf*)s;
#ifdef _OPENMP
#pragma omp atomic update
#endif
                sw->err |= 1;
                sid = 0;
            }
#if OR_MUTATION == 11
            sid = 0;                   /* planted: the first sector always */
#endif
            LN = nsec > 0 ? &s->fs[s->frow[off] + sid] : &s->ls[nn];
        } else {
            LN = &s->ls[nn];
        }
        w3 q = wsub(pw, LN->p0);
        float lu = wdot(q, LN->e[0]), lv = wdot(q, LN->e[1]);
        float lw = poly(LN, lu, lv);
        pl[i + 1] = l2w(LN, lu, lv, lw);
    }
    w3 sum = { 0.0f, 0.0f, 0.0f };
    float ii = 1.0f / isum;
    for (int i = 0; i < nh + 1; ++i) sum = wadd(sum, wmul(ii, wmul(inv[i], pl[i])));
    return sum;
}

/* projection onto the 1D curves (Surface1D.h:54-93) over the feature
   neighbours nb[0..n) */
static w3 project1d(const or_surf* s, int vid, const int* nb, int n, float u)
{
    const lsurf1* L = &s->ls1[s->nn[vid]];
    float v = L->a * u * u;
    w3 pw = wadd(wadd(L->p0, wmul(u, L->e[0])), wmul(v, L->e[1]));
    w3 pl[MAX_NEIGH + 1];
    float inv[MAX_NEIGH + 1];
    float isum = 0.0f;
    pl[0] = pw;
    {
        w3 d = wsub(W(s->pinit + 3 * s->nn[vid]), pw);
        inv[0] = 1.0f / (wdot(d, d) + SEPS);
        isum += inv[0];
    }
    for (int i = 0; i < n; ++i) {
        int nn = s->nn[nb[i]];
        w3 d = wsub(W(s->pinit + 3 * nn), pw);
        inv[i + 1] = 1.0f / (wdot(d, d) + SEPS);
        isum += inv[i + 1];
        const lsurf1* LN = &s->ls1[nn];
        w3 q = wsub(pw, LN->p0);
        float lu = wdot(q, LN->e[0]);
        float lv = LN->a * lu * lu;
        pl[i + 1] = wadd(wadd(LN->p0, wmul(lu, LN->e[0])), wmul(lv, LN->e[1]));
    }
    w3 sum = { 0.0f, 0.0f, 0.0f };
    float ii = 1.0f / isum;
    for (int i = 0; i < n + 1; ++i) sum = wadd(sum, wmul(ii, wmul(inv[i], pl[i])));
    return sum;
}

/* ring score of a volume vertex at an arbitrary point: canonical evaluation
   (or_internal.h) with the frame of the vertex's start position */
typedef struct {
    int tet, n;
    int nsub;     /* hex: p-dependent sub-tets scored per element (5 or 4) */
    v3 x0;
    float sigma, is3, epsD, epsF;
    tetc tc[512];
    float cd[4 * 64], cF[4 * 64];
    hexp hp[5 * 64];
    hexacc hacc;
    float minDist, maxDist;
} vring;

static void vring_init(const or_mesh* m, int vid, const float* xyz, int hexMetric, vring* r)
{
    r->nsub = hex_nsub(hexMetric);
    r->tet = (m->elem == OR_TET);
    int st = m->ringStride, r0 = m->ringRow[vid];
    r->n = m->ringRow[vid + 1] - r0;
    r->x0 = ld3(xyz + 3 * (size_t)vid);
    float mabs = 0.0f, mn = FLT_MAX, mx = -FLT_MAX;
    for (int e = 0; e < r->n; ++e)
        for (int k = 0; k < st; ++k) {
            w3 q = W(xyz + 3 * (size_t)m->ringVal[(size_t)(r0 + e) * st + k]);
            w3 d = { q.x - r->x0.x, q.y - r->x0.y, q.z - r->x0.z };
            mabs = fmaxf(mabs, fmaxf(fabsf(d.x), fmaxf(fabsf(d.y), fabsf(d.z))));
            float dist = wnorm(d);
            mn = fminf(mn, dist); mx = fmaxf(mx, dist);
        }
    r->minDist = mn; r->maxDist = mx;
    r->sigma = or_pow2_scale(mabs);
    float is = 1.0f / r->sigma;
    r->is3 = ((1.0f * is) * is) * is;
    float s2 = r->sigma * r->sigma, s3 = s2 * r->sigma;
    r->epsD = HEX_EPS * s3; r->epsF = HEX_EPS * s2;
    for (int e = 0; e < r->n; ++e) {
        const int* h = m->ringVal + (size_t)(r0 + e) * st;
        if (r->tet) {
            r->tc[e] = tet_consts(loc3(xyz + 3 * (size_t)h[0], r->x0, r->sigma), loc3(xyz + 3 * (size_t)h[1], r->x0, r->sigma),
                                  loc3(xyz + 3 * (size_t)h[2], r->x0, r->sigma));
        } else {
            v3 q[8];
            q[0] = r->x0;
            for (int k = 1; k < 8; ++k) q[k] = loc3(xyz + 3 * (size_t)h[k - 1], r->x0, r->sigma);
            hex_consts(q, r->cd + 4 * e, r->cF + 4 * e, r->hp + 5 * e);
        }
    }
    if (!r->tet) r->hacc = hex_const_acc(r->n, r->cd, r->cF);
}

static float vring_score(const vring* r, w3 p)
{
    float px = (p.x - r->x0.x) * r->sigma, py = (p.y - r->x0.y) * r->sigma, pz = (p.z - r->x0.z) * r->sigma;
    return r->tet ? tet_score(r->n, r->tc, px, py, pz, r->is3)
                  : hex_score(r->n, r->hacc, r->hp, r->nsub, px, py, pz, r->epsD, r->epsF);
}

/* 2D surface search of a free surface vertex (k_optimizeHierarchical2DHex /
   2DTet): (u,v) window around the current position in the nearest
   neighbour's frame, half-width s_d * (min distance to the ring points for
   hexes, max for tets); thread t holds A = (t/nq, t%nq) and B = (t/nq + nq/2,
   t%nq) and keeps A unless q_B > q_A (A on an exact tie); lowest thread wins;
   write iff q(old) < last level max.  uvOut (optional): the final (u, v). */
static int search2d(const or_mesh* m, const or_surf* s, const float* snap, const or_dmo_params* p, int vid, float* xyz,
                    float* uvOut)
{
    vring* r = (vring*)malloc(sizeof(vring));
    vring_init(m, vid, xyz, p->hex_metric, r);
    if (r->n <= 0) { free(r); return 0; }
#if OR_MUTATION == 7
    const float dist = (m->elem == OR_HEX) ? r->maxDist : r->minDist;   /* planted: window rules swapped */
#else
    const float dist = (m->elem == OR_HEX) ? r->minDist : r->maxDist;
#endif
    const lsurf* L = &s->ls[s->nn[vid]];
    float vidPos[3] = { xyz[3 * (size_t)vid], xyz[3 * (size_t)vid + 1], xyz[3 * (size_t)vid + 2] };
    w3 cur = W(vidPos);
    w3 cv = wsub(cur, L->p0);
    float cu = wdot(L->e[0], cv), cvv = wdot(L->e[1], cv);
    const int nq = p->nq, half = nq / 2, T = nq * nq / 2;
    const float aff = 1.0f / (float)(nq - 1);
    const float shr = (p->shrink > 0.0f) ? p->shrink : 2.0f / (float)(nq - 1);
    float qcur = vring_score(r, cur);
    const float qold = qcur;
    float qlast = -FLT_MAX;
    float ds = p->grid_scale;
    for (int d = 0; d < p->depth; ++d) {
        float uMax = ds * dist, uMin = -(ds * dist);
        float best = -FLT_MAX, bu = 0.0f, bv = 0.0f;
        for (int t = 0; t < T; ++t) {
            int i1 = t / nq, j1 = t % nq, i2 = i1 + half;
            float u1 = cu + aff * ((float)i1 * uMin + (float)(nq - 1 - i1) * uMax);
            float v1 = cvv + aff * ((float)j1 * uMin + (float)(nq - 1 - j1) * uMax);
            float u2 = cu + aff * ((float)i2 * uMin + (float)(nq - 1 - i2) * uMax);
            float v2 = v1;
            float q1 = vring_score(r, project2d(m, s, snap, vid, vidPos, u1, v1));
            float q2 = vring_score(r, project2d(m, s, snap, vid, vidPos, u2, v2));
            float qm = q1, uu = u1, vv = v1;
#if OR_MUTATION == 8
            if (q2 >= q1) { qm = q2; uu = u2; vv = v2; }   /* planted: B wins an exact tie */
#else
            if (q2 > q1) { qm = q2; uu = u2; vv = v2; }
#endif
            if (qm > best) { best = qm; bu = uu; bv = vv; }
        }
        if (qcur < best) {
            cur = project2d(m, s, snap, vid, vidPos, bu, bv);
            cu = bu; cvv = bv;
            qcur = best;
        }
        qlast = best;
        ds = ds * shr;
    }
    free(r);
    if (uvOut) { uvOut[0] = cu; uvOut[1] = cvv; }
    if (qold < qlast) { xyz[3 * (size_t)vid] = cur.x; xyz[3 * (size_t)vid + 1] = cur.y; xyz[3 * (size_t)vid + 2] = cur.z; return 1; }
    return 0;
}

/* 1D feature-curve search (k_optimizeHierarchical1DHex / 1DTet).  uvOut
   (optional): the final u. */
static int search1d(const or_mesh* m, const or_surf* s, const or_dmo_params* p, int vid, float* xyz, float* uvOut)
{
    if (!s->opt1d[vid]) return 0;
    vring* r = (vring*)malloc(sizeof(vring));
    vring_init(m, vid, xyz, p->hex_metric, r);
    if (r->n <= 0) { free(r); return 0; }
    /* feature neighbours that themselves have exactly two feature neighbours */
    int nb[MAX_NEIGH], nnb = 0;
    for (int h = m->heRow[vid]; h < m->heRow[vid + 1]; ++h) {
        int d = m->heTarget[h];
        if (!m->isFeature[d]) continue;
        int c = 0;
        for (int h2 = m->heRow[d]; h2 < m->heRow[d + 1]; ++h2) if (m->isFeature[m->heTarget[h2]]) ++c;
        if (c == 2 && nnb < MAX_NEIGH) nb[nnb++] = d;
    }
#if OR_MUTATION == 10
    const float dist = (m->elem == OR_HEX) ? r->maxDist : r->minDist;   /* planted: window rules swapped */
#else
    const float dist = (m->elem == OR_HEX) ? r->minDist : r->maxDist;
#endif
    const lsurf1* L = &s->ls1[s->nn[vid]];
    w3 cur = W(xyz + 3 * (size_t)vid);
    w3 cv = wsub(cur, L->p0);
    float cu = wdot(L->e[0], cv);
    const int nq = p->nq;
    const float aff = 1.0f / (float)(nq - 1);
    const float shr = (p->shrink > 0.0f) ? p->shrink : 2.0f / (float)(nq - 1);
    float qcur = vring_score(r, cur);
    const float qold = qcur;
    float qlast = -FLT_MAX;
    float ds = p->grid_scale;
    for (int d = 0; d < p->depth; ++d) {
        float uMax = ds * dist, uMin = -(ds * dist);
        float best = -FLT_MAX, bu = 0.0f;
        for (int i = 0; i < nq; ++i) {
            float u = cu + aff * ((float)i * uMin + (float)(nq - 1 - i) * uMax);
            float q = vring_score(r, project1d(s, vid, nb, nnb, u));
            if (q > best) { best = q; bu = u; }
        }
        if (qcur < best) {
            cur = project1d(s, vid, nb, nnb, bu);
            cu = bu;
            qcur = best;
        }
        qlast = best;
        ds = ds * shr;
    }
    free(r);
    if (uvOut) uvOut[0] = cu;
    if (qold < qlast) { xyz[3 * (size_t)vid] = cur.x; xyz[3 * (size_t)vid + 1] = cur.y; xyz[3 * (size_t)vid + 2] = cur.z; return 1; }
    return 0;
}

/* 1D boundary-curve search of a flat triangle mesh
   (k_optimizeHierarchical1DFlat, DMOTriFlatImplCub.h:158-333): the ring is
   the halfedge fan of the vertex, closed only for a vertex that is not on a
   surface boundary; window half-width s_d * (max Euclidean distance to the
   fan); scored by the 2D mean ratio of the fan triangles (p, r_k, r_k+1) in
   the canonical evaluation with the frame of the start position. */
static int search1d_flat(const or_mesh* m, const or_surf* s, const or_dmo_params* p, int vid, float* xyz, float* uvOut)
{
    if (!s->opt1d[vid]) return 0;
    const int h0 = m->heRow[vid], n = m->heRow[vid + 1] - h0;
#if OR_MUTATION == 15
    const int nOne = n + 1;            /* planted: the boundary vertex's open fan closed */
#else
    const int nOne = n + (m->isBoundary1D[vid] ? 0 : 1);
#endif
    if (n <= 0 || nOne < 2 || n > MAX_NEIGH) return 0;
    const w3 x0 = W(xyz + 3 * (size_t)vid);
    float mabs = 0.0f, mx = -FLT_MAX;
    for (int i = 0; i < n; ++i) {
        w3 d = wsub(W(xyz + 3 * (size_t)m->heTarget[h0 + i]), x0);
        mabs = fmaxf(mabs, fmaxf(fabsf(d.x), fabsf(d.y)));
        mx = fmaxf(mx, wnorm(d));
    }
    const float sg = or_pow2_scale(mabs);
    const float isg = 1.0f / sg;
    const float is2 = (1.0f * isg) * isg;
    tric cs[MAX_NEIGH];
    const int ne = nOne - 1;
    for (int e = 0; e < ne; ++e) {
        const float* a = xyz + 3 * (size_t)m->heTarget[h0 + e];
        const float* b = xyz + 3 * (size_t)m->heTarget[h0 + (e + 1 == n ? 0 : e + 1)];
        cs[e] = tri_consts((a[0] - x0.x) * sg, (a[1] - x0.y) * sg, (b[0] - x0.x) * sg, (b[1] - x0.y) * sg);
    }
    int nb[MAX_NEIGH], nnb = 0;
    for (int h = h0; h < h0 + n; ++h) {
        int d = m->heTarget[h];
        if (!m->isFeature[d]) continue;
        int c = 0;
        for (int h2 = m->heRow[d]; h2 < m->heRow[d + 1]; ++h2) if (m->isFeature[m->heTarget[h2]]) ++c;
        if (c == 2 && nnb < MAX_NEIGH) nb[nnb++] = d;
    }
    const lsurf1* L = &s->ls1[s->nn[vid]];
    w3 cur = x0;
    float cu = wdot(L->e[0], wsub(cur, L->p0));
    const int nq = p->nq;
    const float aff = 1.0f / (float)(nq - 1);
    const float shr = (p->shrink > 0.0f) ? p->shrink : 2.0f / (float)(nq - 1);
    float qcur = tri_score(ne, cs, 0.0f, 0.0f, is2);
    const float qold = qcur;
    float qlast = -FLT_MAX;
    float ds = p->grid_scale;
    for (int d = 0; d < p->depth; ++d) {
        float uMax = ds * mx, uMin = -(ds * mx);
        float best = -FLT_MAX, bu = 0.0f;
        for (int i = 0; i < nq; ++i) {
            float u = cu + aff * ((float)i * uMin + (float)(nq - 1 - i) * uMax);
            w3 q = project1d(s, vid, nb, nnb, u);
            float sc = tri_score(ne, cs, (q.x - x0.x) * sg, (q.y - x0.y) * sg, is2);
            if (sc > best) { best = sc; bu = u; }
        }
        if (qcur < best) {
            cur = project1d(s, vid, nb, nnb, bu);
            cu = bu;
            qcur = best;
        }
        qlast = best;
        ds = ds * shr;
    }
    if (uvOut) uvOut[0] = cu;
    if (qold < qlast) { xyz[3 * (size_t)vid] = cur.x; xyz[3 * (size_t)vid + 1] = cur.y; xyz[3 * (size_t)vid + 2] = cur.z; return 1; }
    return 0;
}

/* nearest-neighbour walk over the initial surface positions, within the
   vertex's feature class (k_updateNearestNeighbor, DMOCommon.cu:11-43) */
static void update_nn(const or_mesh* m, or_surf* s)
{
    for (int v = 0; v < m->nSurf; ++v) {
        w3 p = W(P(m, v));
        int vh = s->nn[v];
        w3 d0 = wsub(p, W(s->pinit + 3 * vh));
        float sq = wdot(d0, d0);
        int vnew;
        do {
            vnew = vh;
            int c = vh;
            for (int h = m->heRow[c]; h < m->heRow[c + 1]; ++h) {
                int dst = m->heTarget[h];
#if OR_MUTATION != 9                   /* 9 plants: the feature-class filter dropped */
                if (m->isFeature[v] != m->isFeature[dst]) continue;
#endif
                w3 dd = wsub(p, W(s->pinit + 3 * dst));
                float sd = wdot(dd, dd);
                if (sd < sq) { sq = sd; vh = dst; }
            }
        } while (vnew != vh);
        s->nn[v] = vnew;
    }
}

long long or_dmo_iteration_surface(or_mesh* m, or_surf* s, const or_dmo_params* p, float lastMin, long long* nSurf2d, long long* nFeat1d)
{
    long long cnt = 0;
#ifdef _OPENMP
    int nt = p->nthreads > 0 ? p->nthreads : omp_get_max_threads();
#endif
    *nSurf2d = 0; *nFeat1d = 0;
    cnt += or_dmo_iteration(m, p);
    if (!p->fixed_boundary && lastMin > 0.0f && s && m->elem == OR_TRI) {
        /* flat triangle mesh: the free vertices were searched above; the
           boundary vertices follow their boundary curves */
        for (int c = 0; c < m->nColorsFeature; ++c) {
            int lo = m->colOffFeature[c], hi = m->colOffFeature[c + 1];
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
            for (int v = lo; v < hi; ++v) search1d_flat(m, s, p, v, m->xyz, NULL);
            for (int v = lo; v < hi; ++v) *nFeat1d += s->opt1d[v];
        }
    } else if (!p->fixed_boundary && lastMin > 0.0f && s) {
        /* the feature-sector choice of the projection reads positions of
           vertices up to two hops away, which may belong to the colour class
           being updated: it reads a snapshot taken at the start of the pass */
        float* snap = (float*)malloc(sizeof(float) * 3 * (size_t)(m->nSurf > 0 ? m->nSurf : 1));
        for (int c = 0; c < m->nColorsFree; ++c) {
            int lo = m->colOffFree[c], hi = m->colOffFree[c + 1];
            memcpy(snap, m->xyz, sizeof(float) * 3 * (size_t)m->nSurf);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
            for (int v = lo; v < hi; ++v) search2d(m, s, snap, p, v, m->xyz, NULL);
            *nSurf2d += hi - lo;
        }
        free(snap);
        for (int c = 0; c < m->nColorsFeature; ++c) {
            int lo = m->colOffFeature[c], hi = m->colOffFeature[c + 1];
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
            for (int v = lo; v < hi; ++v) search1d(m, s, p, v, m->xyz, NULL);
            for (int v = lo; v < hi; ++v) *nFeat1d += s->opt1d[v];
        }
    }
    if (s) update_nn(m, s);
    return cnt + *nSurf2d + *nFeat1d;
}

/* ------------------------------------------------------------------------
 * Surface meshes (triangles / quads): DMOTriClass / DMOQuadClass
 * (DMOTri.cu:128-170, DMOQuad.cu:128-170, DMOTriImplCub.h:25-380,
 * DMOQuadImplCub.h:23-404).  The free vertices search in 2D on the blended
 * local surfaces, the feature vertices in 1D on their curves, as for volume
 * meshes; the score is the mean ratio of the fan elements around the vertex,
 * signed by the vertex normal of the nearest neighbour (DMOQualityTri.h:
 * 257-275, DMOQualityQuad.h:30-50), and the final write test uses the
 * unsigned metric.  Vertex normals follow the positions after every
 * iteration (updateNormals).
 * ---------------------------------------------------------------------- */
static float sgnf(float a) { return a > 0.0f ? 1.0f : (a < 0.0f ? -1.0f : 0.0f); }
static w3 wneg(w3 a) { w3 r = { -a.x, -a.y, -a.z }; return r; }
static float two_sqrt3(void) { return 2.0f * sqrtf(3.0f); }

/* meanRatioMetricTri (3D), signed by n when n != NULL: 2 sqrt(3) A / sum l^2,
   e0 = p1 - p0, e1 = p2 - p0, e2 = p2 - p1, A = |e0 x e1| * sign(n . unit(e0 x e1)) */
static float mr_tri3(w3 p0, w3 p1, w3 p2, const w3* n)
{
    w3 e0 = wsub(p1, p0), e1 = wsub(p2, p0), e2 = wsub(p2, p1);
    float l = (wdot(e0, e0) + wdot(e1, e1)) + wdot(e2, e2);
    w3 N = wcross(e0, e1);
    float area = wnorm(N);
    if (n) area = area * sgnf(wdot(wnormalize(N), *n));
    return (two_sqrt3() * area) / l;
}

/* meanRatioMetricQuad: 2 (|e1 x -e0| + |e3 x -e2|) / sum l^2 with e_i =
   p_(i+1) - p_i; with n: -1 when either corner triangle's unit normal has
   n . u < 0.7 */
static float mr_quad3(w3 p0, w3 p1, w3 p2, w3 p3, const w3* n)
{
    w3 e0 = wsub(p1, p0), e1 = wsub(p2, p1), e2 = wsub(p3, p2), e3 = wsub(p0, p3);
    float l = ((wdot(e0, e0) + wdot(e1, e1)) + wdot(e2, e2)) + wdot(e3, e3);
    w3 N1 = wcross(e1, wneg(e0)), N2 = wcross(e3, wneg(e2));
    float a1 = wnorm(N1), a2 = wnorm(N2);
    if (n) {
        float d1 = wdot(wnormalize(N1), *n), d2 = wdot(wnormalize(N2), *n);
        if (d1 < 0.7f || d2 < 0.7f) return -1.0f;
    }
    return (2.0f * (a1 + a2)) / l;
}

float or_q_tri3(const float* p) { return mr_tri3(W(p), W(p + 3), W(p + 6), NULL); }
float or_q_tri3n(const float* p, const float* n) { w3 nn = W(n); return mr_tri3(W(p), W(p + 3), W(p + 6), &nn); }
float or_q_quad(const float* p) { return mr_quad3(W(p), W(p + 3), W(p + 6), W(p + 9), NULL); }
float or_q_quadn(const float* p, const float* n) { w3 nn = W(n); return mr_quad3(W(p), W(p + 3), W(p + 6), W(p + 9), &nn); }

/* the same metrics in FP64 (literal formulas, for the tests' tolerance checks) */
static void d3(const float* p, double* o) { o[0] = p[0]; o[1] = p[1]; o[2] = p[2]; }
static double dnorm3(const double* a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static void dcross3(const double* a, const double* b, double* o)
{
    o[0] = a[1] * b[2] - a[2] * b[1]; o[1] = a[2] * b[0] - a[0] * b[2]; o[2] = a[0] * b[1] - a[1] * b[0];
}
double or_q_tri364(const float* p)
{
    double a[3], b[3], c[3], e0[3], e1[3], e2[3], N[3];
    d3(p, a); d3(p + 3, b); d3(p + 6, c);
    for (int i = 0; i < 3; ++i) { e0[i] = b[i] - a[i]; e1[i] = c[i] - a[i]; e2[i] = c[i] - b[i]; }
    double l = 0.0;
    for (int i = 0; i < 3; ++i) l += e0[i] * e0[i] + e1[i] * e1[i] + e2[i] * e2[i];
    dcross3(e0, e1, N);
    return 2.0 * sqrt(3.0) * dnorm3(N) / l;
}
double or_q_quad64(const float* p)
{
    double v[4][3], e[4][3], m0[3], m2[3], N1[3], N2[3];
    for (int k = 0; k < 4; ++k) d3(p + 3 * k, v[k]);
    double l = 0.0;
    for (int k = 0; k < 4; ++k)
        for (int i = 0; i < 3; ++i) { e[k][i] = v[(k + 1) % 4][i] - v[k][i]; l += e[k][i] * e[k][i]; }
    for (int i = 0; i < 3; ++i) { m0[i] = -e[0][i]; m2[i] = -e[2][i]; }
    dcross3(e[1], m0, N1);
    dcross3(e[3], m2, N2);
    return 2.0 * (dnorm3(N1) + dnorm3(N2)) / l;
}

/* element metric of a surface mesh on the surface path: unsigned 3D mean ratio */
float or_q_surface_elem(const or_mesh* m, int e)
{
    const int* c = m->cells + (size_t)e * m->npe;
    if (m->npe == 4) return mr_quad3(W(P(m, c[0])), W(P(m, c[1])), W(P(m, c[2])), W(P(m, c[3])), NULL);
    return mr_tri3(W(P(m, c[0])), W(P(m, c[1])), W(P(m, c[2])), NULL);
}

void or_surface_mesh_qualities(const or_mesh* m, float* out)
{
    for (int e = 0; e < m->nE; ++e) out[e] = or_q_surface_elem(m, e);
}

float or_surface_mesh_min_quality(const or_mesh* m)
{
    float q = FLT_MAX;
    for (int e = 0; e < m->nE; ++e) q = fminf(q, or_q_surface_elem(m, e));
    return q;
}

/* fan of a surface vertex: triangles keep the halfedge targets r_i and score
   the elements (p, r_k, r_k+1); quads keep (r_i, d_i), d_i the target of the
   halfedge after vid -> r_i in its face, and score (p, r_k, d_k, r_k+1).
   Closed (r_n = r_0, n elements) unless open (n - 1 elements).  maxDist is
   the largest distance from vid to the r_i (and d_i).  The d_i are read from
   the snapshot of the pass: on a surface boundary the "diagonal" of the
   boundary halfedge is the boundary neighbour of the next boundary vertex,
   which can belong to the colour class being updated; every other fan vertex
   shares an element with vid, so the snapshot holds its current position. */
typedef struct {
    int quad, ne;
    w3 pts[2 * MAX_NEIGH + 1];
    w3 n;
    float maxDist;
} fan;

static int fan_init(const or_mesh* m, const or_surf* s, const float* snap, int vid, int open, fan* f)
{
    const int h0 = m->heRow[vid], n = m->heRow[vid + 1] - h0;
    if (n <= 0 || n > MAX_NEIGH) return 0;
    f->quad = (m->fpe == 4);
    const int K = f->quad ? 2 : 1;
    const w3 x0 = W(P(m, vid));
#if OR_MUTATION == 14
    float mx = FLT_MAX;                /* planted: the window from the nearest fan vertex */
#define FAN_EXT fminf
#else
    float mx = -FLT_MAX;
#define FAN_EXT fmaxf
#endif
    for (int i = 0; i < n; ++i) {
        const int h = h0 + i;
        w3 r = W(P(m, m->heTarget[h]));
        f->pts[K * i] = r;
        mx = FAN_EXT(mx, wnorm(wsub(r, x0)));
        if (f->quad) {
            const int nx = m->heNext[h];
            w3 d = W(snap + 3 * (size_t)m->heTarget[nx >= 0 ? nx : h]);
            f->pts[K * i + 1] = d;
            mx = FAN_EXT(mx, wnorm(wsub(d, x0)));
        }
    }
#undef FAN_EXT
    f->ne = open ? n - 1 : n;
    if (!open) f->pts[K * n] = f->pts[0];
    f->maxDist = mx;
    f->n = W(s->vnorm + 3 * (size_t)s->nn[vid]);
    return f->ne > 0;
}

static float fan_score(const fan* f, w3 p, int oriented)
{
    const w3* n = oriented ? &f->n : NULL;
    float q = FLT_MAX;
    for (int k = 0; k < f->ne; ++k) {
        float e = f->quad ? mr_quad3(p, f->pts[2 * k], f->pts[2 * k + 1], f->pts[2 * k + 2], n)
                          : mr_tri3(p, f->pts[k], f->pts[k + 1], n);
        q = fminf(q, e);
    }
    return q;
}

/* 2D search of a free vertex of a surface mesh (k_optimizeHierarchical2D,
   DMOTriImplCub.h:25-200, DMOQuadImplCub.h:23-212): window half-width
   s_d * maxDist in the nearest neighbour's (u, v); candidates, ties and the
   update rule as search2d; write iff q_unsigned(old) < last level max */
static int search2d_fan(const or_mesh* m, const or_surf* s, const float* snap, const or_dmo_params* p, int vid, float* xyz,
                        float* uvOut)
{
    fan f;
    if (!fan_init(m, s, snap, vid, 0, &f)) return 0;
    const lsurf* L = &s->ls[s->nn[vid]];
    float vidPos[3] = { xyz[3 * (size_t)vid], xyz[3 * (size_t)vid + 1], xyz[3 * (size_t)vid + 2] };
    w3 cur = W(vidPos);
    w3 cv = wsub(cur, L->p0);
    float cu = wdot(L->e[0], cv), cvv = wdot(L->e[1], cv);
    const int nq = p->nq, half = nq / 2, T = nq * nq / 2;
    const float aff = 1.0f / (float)(nq - 1);
    const float shr = (p->shrink > 0.0f) ? p->shrink : 2.0f / (float)(nq - 1);
    float qcur = fan_score(&f, cur, 1);
    const float qold = fan_score(&f, cur, 0);
    float qlast = -FLT_MAX;
    float ds = p->grid_scale;
    for (int d = 0; d < p->depth; ++d) {
        float uMax = ds * f.maxDist, uMin = -(ds * f.maxDist);
        float best = -FLT_MAX, bu = 0.0f, bv = 0.0f;
        for (int t = 0; t < T; ++t) {
            int i1 = t / nq, j1 = t % nq, i2 = i1 + half;
            float u1 = cu + aff * ((float)i1 * uMin + (float)(nq - 1 - i1) * uMax);
            float v1 = cvv + aff * ((float)j1 * uMin + (float)(nq - 1 - j1) * uMax);
            float u2 = cu + aff * ((float)i2 * uMin + (float)(nq - 1 - i2) * uMax);
            float q1 = fan_score(&f, project2d(m, s, snap, vid, vidPos, u1, v1), 1);
            float q2 = fan_score(&f, project2d(m, s, snap, vid, vidPos, u2, v1), 1);
            float qm = q1, uu = u1, vv = v1;
            if (q2 > q1) { qm = q2; uu = u2; }
            if (qm > best) { best = qm; bu = uu; bv = vv; }
        }
        if (qcur < best) {
            cur = project2d(m, s, snap, vid, vidPos, bu, bv);
            cu = bu; cvv = bv;
            qcur = best;
        }
        qlast = best;
        ds = ds * shr;
    }
    if (uvOut) { uvOut[0] = cu; uvOut[1] = cvv; }
    if (qold < qlast) { xyz[3 * (size_t)vid] = cur.x; xyz[3 * (size_t)vid + 1] = cur.y; xyz[3 * (size_t)vid + 2] = cur.z; return 1; }
    return 0;
}

/* 1D search of a feature vertex of a surface mesh (k_optimizeHierarchical1D,
   DMOTriImplCub.h:205-378, DMOQuadImplCub.h:216-402): the fan is open on a
   surface boundary */
static int search1d_fan(const or_mesh* m, const or_surf* s, const float* snap, const or_dmo_params* p, int vid, float* xyz,
                        float* uvOut)
{
    if (!s->opt1d[vid]) return 0;
    fan f;
    if (!fan_init(m, s, snap, vid, m->isBoundary1D[vid] != 0, &f)) return 0;
    int nb[MAX_NEIGH], nnb = 0;
    for (int h = m->heRow[vid]; h < m->heRow[vid + 1]; ++h) {
        int d = m->heTarget[h];
        if (!m->isFeature[d]) continue;
        int c = 0;
        for (int h2 = m->heRow[d]; h2 < m->heRow[d + 1]; ++h2) if (m->isFeature[m->heTarget[h2]]) ++c;
        if (c == 2 && nnb < MAX_NEIGH) nb[nnb++] = d;
    }
    const lsurf1* L = &s->ls1[s->nn[vid]];
    w3 cur = W(xyz + 3 * (size_t)vid);
    float cu = wdot(L->e[0], wsub(cur, L->p0));
    const int nq = p->nq;
    const float aff = 1.0f / (float)(nq - 1);
    const float shr = (p->shrink > 0.0f) ? p->shrink : 2.0f / (float)(nq - 1);
    float qcur = fan_score(&f, cur, 1);
    const float qold = fan_score(&f, cur, 0);
    float qlast = -FLT_MAX;
    float ds = p->grid_scale;
    for (int d = 0; d < p->depth; ++d) {
        float uMax = ds * f.maxDist, uMin = -(ds * f.maxDist);
        float best = -FLT_MAX, bu = 0.0f;
        for (int i = 0; i < nq; ++i) {
            float u = cu + aff * ((float)i * uMin + (float)(nq - 1 - i) * uMax);
            float q = fan_score(&f, project1d(s, vid, nb, nnb, u), 1);
            if (q > best) { best = q; bu = u; }
        }
        if (qcur < best) {
            cur = project1d(s, vid, nb, nnb, bu);
            cu = bu;
            qcur = best;
        }
        qlast = best;
        ds = ds * shr;
    }
    if (uvOut) uvOut[0] = cu;
    if (qold < qlast) { xyz[3 * (size_t)vid] = cur.x; xyz[3 * (size_t)vid + 1] = cur.y; xyz[3 * (size_t)vid + 2] = cur.z; return 1; }
    return 0;
}

long long or_dmo_iteration_surface_mesh(or_mesh* m, or_surf* s, const or_dmo_params* p, long long* nSurf2d, long long* nFeat1d)
{
#ifdef _OPENMP
    int nt = p->nthreads > 0 ? p->nthreads : omp_get_max_threads();
#endif
    *nSurf2d = 0; *nFeat1d = 0;
    float* snap = (float*)malloc(sizeof(float) * 3 * (size_t)(m->nSurf > 0 ? m->nSurf : 1));
    for (int c = 0; c < m->nColorsFree; ++c) {
        int lo = m->colOffFree[c], hi = m->colOffFree[c + 1];
        memcpy(snap, m->xyz, sizeof(float) * 3 * (size_t)m->nSurf);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
        for (int v = lo; v < hi; ++v) search2d_fan(m, s, snap, p, v, m->xyz, NULL);
        *nSurf2d += hi - lo;
    }
    if (!p->fixed_boundary) {
        for (int c = 0; c < m->nColorsFeature; ++c) {
            int lo = m->colOffFeature[c], hi = m->colOffFeature[c + 1];
            memcpy(snap, m->xyz, sizeof(float) * 3 * (size_t)m->nSurf);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
            for (int v = lo; v < hi; ++v) search1d_fan(m, s, snap, p, v, m->xyz, NULL);
            for (int v = lo; v < hi; ++v) *nFeat1d += s->opt1d[v];
        }
    }
    free(snap);
    update_nn(m, s);
    normals(m, s->fnorm, s->vnorm);
    return *nSurf2d + *nFeat1d;
}

/* Laplace smoothing of the free vertices of a surface mesh, colour by colour
   (laplaceSmoothing, Laplace.cu:8-41): p <- (p + sum of the halfedge
   targets) * (1 / (1 + n)) */
void or_laplace(or_mesh* m, int n_iter)
{
    for (int it = 0; it < n_iter; ++it)
        for (int c = 0; c < m->nColorsFree; ++c) {
            int lo = m->colOffFree[c], hi = m->colOffFree[c + 1];
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
            for (int v = lo; v < hi; ++v) {
                const int h0 = m->heRow[v], n = m->heRow[v + 1] - h0;
                w3 q = W(P(m, v));
                for (int h = h0; h < h0 + n; ++h) q = wadd(q, W(P(m, m->heTarget[h])));
                q = wmul(1.0f / (float)(1 + n), q);
                float* o = m->xyz + 3 * (size_t)v;
                o[0] = q.x; o[1] = q.y; o[2] = q.z;
            }
        }
}

/* Sample points of the local surfaces of surface vertex vid on an nu x nv
   grid (DMOCommon.h:197-276 k_fillEstimateLocalSurfacePoints with
   USE_SURF_OF_NN = false, and DMOCommon.cu:73-150 k_fillLocalSurfacePoints):
   point idx = i*nu + j, i = idx / nu, j = idx % nu.  mode 0: the blended
   projection of (u, v) in the frame of the NN's fit, window s * maxDist;
   mode 1: the NN's own fit (or its sector fit `sid` when the NN is a
   feature vertex), window 2 s * maxDist.  maxDist = largest distance from
   vid to its surface one-ring. */
void or_surface_points(const or_mesh* m, const or_surf* s, int vid, int nu, int nv, int mode, int sid, float gridScale,
                       float* out)
{
    w3 pv = W(P(m, vid));
    float mx = -FLT_MAX;
    for (int h = m->heRow[vid]; h < m->heRow[vid + 1]; ++h) mx = fmaxf(mx, wnorm(wsub(W(P(m, m->heTarget[h])), pv)));
    const int nn = s->nn[vid];
    const lsurf* L = &s->ls[nn];
    if (mode == 1 && nn >= s->nFree) {
        int off = nn - s->nFree, nsec = s->frow[off + 1] - s->frow[off];
        if (sid >= nsec) sid = nsec - 1;
        if (nsec > 0 && sid >= 0) L = &s->fs[s->frow[off] + sid];
    }
    w3 c = wsub(pv, L->p0);
    float cu = wdot(L->e[0], c), cv = wdot(L->e[1], c);
    float uMax = gridScale * mx, uMin = -(gridScale * mx);
    if (mode == 1) { uMax *= 2.0f; uMin *= 2.0f; }
    const float au = 1.0f / (float)(nu - 1), av = 1.0f / (float)(nv - 1);
    float vidPos[3] = { pv.x, pv.y, pv.z };
    for (int idx = 0; idx < nu * nv; ++idx) {
        int j = idx % nu, i = idx / nu;
        float u = cu + au * ((float)i * uMin + (float)(nu - 1 - i) * uMax);
        float v = cv + av * ((float)j * uMin + (float)(nv - 1 - j) * uMax);
        w3 q = (mode == 0) ? project2d(m, s, m->xyz, vid, vidPos, u, v) : l2w(L, u, v, poly(L, u, v));
        out[3 * idx] = q.x; out[3 * idx + 1] = q.y; out[3 * idx + 2] = q.z;
    }
}

/* test hooks: the local quadric fit of a point set in the frame built from a
   normal and a reference neighbour (computeLocalSurface1/2), and the 3x3 LU
   solve on its own */
void or_fit_local_surface(const float* pts, int np, const float* normal, const float* refp, float* out16)
{
    lsurf q;
    q.p0 = W(pts);
    w3 n = W(normal), r = wsub(W(refp), q.p0);
    frame3(n, r, &q.e[0], &q.e[1], &q.e[2]);
    w3 buf[MAX_NEIGH + 1];
    if (np > MAX_NEIGH + 1) np = MAX_NEIGH + 1;
    for (int i = 0; i < np; ++i) buf[i] = W(pts + 3 * i);
    fit_quadric(&q, buf, np);
    out16[0] = q.p0.x; out16[1] = q.p0.y; out16[2] = q.p0.z;
    for (int k = 0; k < 3; ++k) { out16[3 + 3 * k] = q.e[k].x; out16[4 + 3 * k] = q.e[k].y; out16[5 + 3 * k] = q.e[k].z; }
    out16[12] = q.a; out16[13] = q.b; out16[14] = q.c; out16[15] = (float)q.np;
}

void or_lu_solve3(const float* Acm, const float* rhs, float* x) { lu_solve3(Acm, rhs, x); }

/* accessors for tests */
void or_surface_get(const or_surf* s, int what, void* out)
{
    switch (what) {
    case 0: memcpy(out, s->nn, sizeof(int) * (size_t)s->nS); break;
    case 1: for (int v = 0; v < s->nS; ++v) {
                float* o = (float*)out + 16 * (size_t)v;
                const lsurf* q = &s->ls[v];
                o[0] = q->p0.x; o[1] = q->p0.y; o[2] = q->p0.z;
                for (int k = 0; k < 3; ++k) { o[3 + 3 * k] = q->e[k].x; o[4 + 3 * k] = q->e[k].y; o[5 + 3 * k] = q->e[k].z; }
                o[12] = q->a; o[13] = q->b; o[14] = q->c; o[15] = (float)q->np;
            }
            break;
    case 2: for (int v = 0; v < s->nS; ++v) ((unsigned char*)out)[v] = s->opt1d[v]; break;
    case 3: for (int v = 0; v < s->nS; ++v) {
                float* o = (float*)out + 10 * (size_t)v;
                const lsurf1* q = &s->ls1[v];
                o[0] = q->p0.x; o[1] = q->p0.y; o[2] = q->p0.z;
                o[3] = q->e[0].x; o[4] = q->e[0].y; o[5] = q->e[0].z;
                o[6] = q->e[1].x; o[7] = q->e[1].y; o[8] = q->e[1].z; o[9] = q->a;
            }
            break;
    case 4: memcpy(out, s->vnorm, sizeof(float) * 3 * (size_t)s->nS); break;
    default: break;
    }
}

int or_surface_nsectors(const or_surf* s, int nFeature) { return s->frow[nFeature]; }

int or_surface_error(const or_surf* s) { return s->err; }

void or_surface_get_sectors(const or_surf* s, int* table, int* frow, float* surf16)
{
    if (table) memcpy(table, s->table, sizeof(int) * (size_t)s->nF * TABLE_SIZE);
    if (frow) memcpy(frow, s->frow, sizeof(int) * ((size_t)s->nF + 1));
    if (surf16)
        for (int q = 0; q < s->frow[s->nF]; ++q) {
            float* o = surf16 + 16 * (size_t)q;
            const lsurf* f = &s->fs[q];
            o[0] = f->p0.x; o[1] = f->p0.y; o[2] = f->p0.z;
            for (int k = 0; k < 3; ++k) { o[3 + 3 * k] = f->e[k].x; o[4 + 3 * k] = f->e[k].y; o[5 + 3 * k] = f->e[k].z; }
            o[12] = f->a; o[13] = f->b; o[14] = f->c; o[15] = (float)f->np;
        }
}

int or_surface_search_vertex(const or_mesh* m, const or_surf* s, const or_dmo_params* p, int vid, int dim,
                             float* xyz, float* uvOut)
{
    /* a single vertex: the snapshot of the pass is the current positions */
    const int surfaceMesh = m->elem == OR_QUAD || (m->elem == OR_TRI && !m->flat);
    if (surfaceMesh) return dim == 1 ? search1d_fan(m, s, xyz, p, vid, xyz, uvOut) : search2d_fan(m, s, xyz, p, vid, xyz, uvOut);
    if (m->elem == OR_TRI) return dim == 1 ? search1d_flat(m, s, p, vid, xyz, uvOut) : 0;
    if (dim == 1) return search1d(m, s, p, vid, xyz, uvOut);
    return search2d(m, s, xyz, p, vid, xyz, uvOut);
}

void or_surface_project(const or_mesh* m, const or_surf* s, int vid, float u, float v, float* out3)
{
    w3 q = project2d(m, s, m->xyz, vid, P(m, vid), u, v);
    out3[0] = q.x; out3[1] = q.y; out3[2] = q.z;
}

void or_surface_update_nn(const or_mesh* m, or_surf* s) { update_nn(m, s); }
