/*
 * or_dmo.c -- oracle restatement of the DMO vertex search (TEST
 * INFRASTRUCTURE).
 *
 * Volume inner search: reference k_optimizeHierarchicalInnerTet / Hex
 *   (DMOTetImplCub.h:20-137, DMOHexImplCub.h:20-137):
 *   - window half-extent per axis from the one-ring: max |ring - x| (tet) or
 *     min |ring - x| (hex, DMOHexImplCub.h:87-88), taken once at x0;
 *   - level d: grid p_i = aff*(i*gmin + (nq-1-i)*gmax), gmax/gmin =
 *     curr +- s_d*extent, s_0 = grid_scale, s_{d+1} = s_d*shrink (2/(nq-1));
 *   - candidate index n = i + nq*j + nq^2*k, the lowest index wins ties
 *     (cub::ArgMax keeps the lower key; each thread owns a contiguous range);
 *   - the vertex moves to the level's best candidate iff it is strictly better
 *     than the current position; the position is written iff it moved (the
 *     reference's cumulative per-thread qMax with "qLast < qMax" and the final
 *     "qOld < qMax" reduce to exactly this).
 * Flat 2D search: reference k_optimizeHierarchical2DFlat
 *   (DMOTriFlatImplCub.h:23-152): thread t of nq^2/2 owns candidates
 *   A = (t/nq, t%nq) and B = (t/nq + nq/2, t%nq); per-level max resets; the
 *   thread picks A iff q_A > q_B (B on an exact tie); the lowest thread with the
 *   block maximum wins; the final write compares q(old) with the LAST level's
 *   maximum.
 */
#include "vsdmo_oracle.h"
#include "or_internal.h"
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXR 512

static inline float invpow(float sigma, int k)
{
    float is = 1.0f / sigma, r = 1.0f;
    for (int i = 0; i < k; ++i) r = r * is;
    return r;
}

static inline float shrink_of(const or_dmo_params* p)
{
    return (p->shrink > 0.0f) ? p->shrink : 2.0f / (float)(p->nq - 1);
}

/* grid coordinate (reference DMOTetImplCub.h:101-103) */
static inline float gridc(float aff, int i, int nq, float gmin, float gmax)
{
    return aff * ((float)i * gmin + (float)(nq - 1 - i) * gmax);
}

/* One-ring of a volume vertex in its local frame (p' = (p - x0) * sigma):
   the window half-extents, the scale and the element constants of the
   canonical evaluation (or_internal.h). */
typedef struct {
    int tet, n;
    int nsub;        /* hex: p-dependent sub-tets scored per element (5 or 4) */
    v3 x0;
    float ext[3], sigma, is3, epsD, epsF;
    tetc tcs[MAXR];
    float cd[4 * 64], cF[4 * 64];
    hexp hp[5 * 64];
    hexacc hacc;
} ringctx;

/* ringPts: n elements of 3 (tet halfface) or 7 (halfhex) points, world
   coordinates; emin: per-axis minimum extent instead of the maximum */
static void ring_setup(ringctx* r, int tet, int n, const float* ringPts, const float* x0p, int emin, int hexMetric)
{
    const int st = tet ? 3 : 7;
    r->tet = tet; r->n = n;
    r->nsub = hex_nsub(hexMetric);
#if OR_MUTATION == 18
    r->nsub = 5;                       /* planted: the inner search scores the nine sub-tets whatever the metric */
#endif
    r->x0 = ld3(x0p);
    for (int a = 0; a < 3; ++a) r->ext[a] = emin ? FLT_MAX : -FLT_MAX;
    float mabs = 0.0f;
    for (int e = 0; e < n; ++e)
        for (int k = 0; k < st; ++k) {
            const float* q = ringPts + 3 * (size_t)(e * st + k);
            float d[3] = { fabsf(q[0] - r->x0.x), fabsf(q[1] - r->x0.y), fabsf(q[2] - r->x0.z) };
            for (int a = 0; a < 3; ++a) {
                r->ext[a] = emin ? fminf(r->ext[a], d[a]) : fmaxf(r->ext[a], d[a]);
                mabs = fmaxf(mabs, d[a]);
            }
        }
    r->sigma = or_pow2_scale(mabs);
    float s2 = r->sigma * r->sigma, s3 = s2 * r->sigma;
    r->is3 = invpow(r->sigma, 3);
    r->epsD = HEX_EPS * s3; r->epsF = HEX_EPS * s2;
    r->hacc = hex_acc_init();
    if (tet) {
        for (int e = 0; e < n; ++e) {
            const float* h = ringPts + 9 * (size_t)e;
            r->tcs[e] = tet_consts(loc3(h, r->x0, r->sigma), loc3(h + 3, r->x0, r->sigma), loc3(h + 6, r->x0, r->sigma));
        }
    } else {
        for (int e = 0; e < n; ++e) {
            const float* h = ringPts + 21 * (size_t)e;
            v3 q[8];
            q[0] = r->x0;
            for (int k = 1; k < 8; ++k) q[k] = loc3(h + 3 * (k - 1), r->x0, r->sigma);
            hex_consts(q, r->cd + 4 * e, r->cF + 4 * e, r->hp + 5 * e);
        }
        r->hacc = hex_const_acc(n, r->cd, r->cF);
    }
}

/* ring score at the world point (x, y, z) */
static inline float ring_score_at(const ringctx* r, float x, float y, float z)
{
    float px = (x - r->x0.x) * r->sigma, py = (y - r->x0.y) * r->sigma, pz = (z - r->x0.z) * r->sigma;
    return r->tet ? tet_score(r->n, r->tcs, px, py, pz, r->is3)
                  : hex_score(r->n, r->hacc, r->hp, r->nsub, px, py, pz, r->epsD, r->epsF);
}

/* Search of one vertex on an explicit ring: ringPts holds n elements of st
   points (3 per tet halfface, 7 per halfhex), world coordinates; x0 is the
   vertex's start position.  Writes the final position to out (x0 when the
   vertex does not move) and returns 1 iff it moved. */
static int search_volume_ring(int tet, int n, const float* ringPts, const float* x0p, const or_dmo_params* p,
                              float* out)
{
    out[0] = x0p[0]; out[1] = x0p[1]; out[2] = x0p[2];
    if (n <= 0 || n > (tet ? MAXR : 64)) return 0;
    int emin = (p->extent_mode == OR_EXTENT_MIN);
#if OR_MUTATION == 1
    emin = !emin;                      /* planted: min and max extent swapped */
#endif
    ringctx* r = (ringctx*)malloc(sizeof(ringctx));
    ring_setup(r, tet, n, ringPts, x0p, emin, p->hex_metric);
    const int nq = p->nq;
    const float aff = 1.0f / (float)(nq - 1);
    const float shr = shrink_of(p);
    float cur[3] = { x0p[0], x0p[1], x0p[2] };
    float qcur = ring_score_at(r, cur[0], cur[1], cur[2]);
    float ds = p->grid_scale;
    int moved = 0;
    for (int d = 0; d < p->depth; ++d) {
        float gmax[3], gmin[3];
        for (int a = 0; a < 3; ++a) { gmax[a] = cur[a] + ds * r->ext[a]; gmin[a] = cur[a] - ds * r->ext[a]; }
        float best = -FLT_MAX, bp[3] = { 0, 0, 0 };
        for (int k = 0; k < nq; ++k) {
            float pzg = gridc(aff, k, nq, gmin[2], gmax[2]);
            for (int j = 0; j < nq; ++j) {
                float pyg = gridc(aff, j, nq, gmin[1], gmax[1]);
                for (int i = 0; i < nq; ++i) {
                    float pxg = gridc(aff, i, nq, gmin[0], gmax[0]);
                    float q = ring_score_at(r, pxg, pyg, pzg);
#if OR_MUTATION == 6
                    if (q >= best) {           /* planted: the highest index wins a tie */
#else
                    if (q > best) {
#endif
                        best = q; bp[0] = pxg; bp[1] = pyg; bp[2] = pzg;
                    }
                }
            }
        }
        if (qcur < best) { cur[0] = bp[0]; cur[1] = bp[1]; cur[2] = bp[2]; qcur = best; moved = 1; }
        ds = ds * shr;
    }
    free(r);
    if (moved) { out[0] = cur[0]; out[1] = cur[1]; out[2] = cur[2]; }
    return moved;
}

/* Learned proposal window of one inner vertex (this build's own search, no
   reference counterpart): a 4x4x4 grid (candidate i + 4j + 16k, the lowest
   index wins ties) on frac times the plain first-level window around the
   proposal c, for depthL levels with the window shrunk by 2/3 per level;
   the centre moves to a level's best candidate when it beats the best
   quality so far, which starts at q(x0).  Returns 1 when the vertex moved
   (out = its new position), 0 when it stays because a level's best tied a
   positive q(x0) or a hex vertex is pinned (its sub-tets without it already
   score 0), -1 when it needs the plain full search (fallback). */
static int learned_window_ring(int tet, int n, const float* ringPts, const float* x0p, const float* c,
                               const or_dmo_params* p, int depthL, float frac, float* out)
{
    out[0] = x0p[0]; out[1] = x0p[1]; out[2] = x0p[2];
    if (n <= 0 || n > (tet ? MAXR : 64)) return -1;
    ringctx* r = (ringctx*)malloc(sizeof(ringctx));
    ring_setup(r, tet, n, ringPts, x0p, p->extent_mode == OR_EXTENT_MIN, p->hex_metric);
    const int NL = 4;
    const float aff = 1.0f / (float)(NL - 1);
    const float q0 = ring_score_at(r, x0p[0], x0p[1], x0p[2]);
    float cur[3] = { c[0], c[1], c[2] };
    float qbest = q0;
    int moved = 0, tie = 0;
    float ds = frac * p->grid_scale;
    for (int d = 0; d < depthL; ++d) {
        float gmax[3], gmin[3];
        for (int a = 0; a < 3; ++a) { gmax[a] = cur[a] + ds * r->ext[a]; gmin[a] = cur[a] - ds * r->ext[a]; }
        float best = -FLT_MAX, bp[3] = { 0, 0, 0 };
        for (int k = 0; k < NL; ++k) {
            float pzg = gridc(aff, k, NL, gmin[2], gmax[2]);
            for (int j = 0; j < NL; ++j) {
                float pyg = gridc(aff, j, NL, gmin[1], gmax[1]);
                for (int i = 0; i < NL; ++i) {
                    float pxg = gridc(aff, i, NL, gmin[0], gmax[0]);
                    float q = ring_score_at(r, pxg, pyg, pzg);
                    if (q > best) { best = q; bp[0] = pxg; bp[1] = pyg; bp[2] = pzg; }
                }
            }
        }
#if OR_MUTATION != 12                  /* 12 plants: the positive-tie rule dropped */
        if (best == q0 && q0 > 0.0f) tie = 1;
#endif
        if (qbest < best) { cur[0] = bp[0]; cur[1] = bp[1]; cur[2] = bp[2]; qbest = best; moved = 1; }
#if OR_MUTATION == 13
        ds = ds * 0.5f;                /* planted: the window halved per level */
#else
        ds = ds * (2.0f / (float)(NL - 1));
#endif
    }
    const int pinned = !tet && (r->hacc.dmin < r->epsD || r->hacc.fmin < r->epsF);
    free(r);
    if (moved) { out[0] = cur[0]; out[1] = cur[1]; out[2] = cur[2]; return 1; }
    return (tie || pinned) ? 0 : -1;
}

static int vertex_volume(const or_mesh* m, const or_dmo_params* p, int vid, float* xyz)
{
    int tet = (m->elem == OR_TET);
    int st = m->ringStride;
    int r0 = m->ringRow[vid], n = m->ringRow[vid + 1] - r0;
    if (n <= 0 || n > (tet ? MAXR : 64)) return 0;
    float pts[MAXR * 9];
    for (int e = 0; e < n; ++e)
        for (int k = 0; k < st; ++k) {
            const float* q = xyz + 3 * (size_t)m->ringVal[(size_t)(r0 + e) * st + k];
            float* o = pts + 3 * (size_t)(e * st + k);
            o[0] = q[0]; o[1] = q[1]; o[2] = q[2];
        }
    float x0[3] = { xyz[3 * (size_t)vid], xyz[3 * (size_t)vid + 1], xyz[3 * (size_t)vid + 2] };
    float out[3];
    int moved = search_volume_ring(tet, n, pts, x0, p, out);
    if (moved) { xyz[3 * (size_t)vid] = out[0]; xyz[3 * (size_t)vid + 1] = out[1]; xyz[3 * (size_t)vid + 2] = out[2]; }
    return moved;
}

/* Flat 2D search on an explicit closed fan of n ring points (x, y, z with
   z ignored): elements (ring[k], ring[k+1]), ring[n] = ring[0].  Writes the
   final position to out (x0 when it does not move); returns 1 iff written. */
static int search_flat_fan(int n, const float* fan, const float* x0, const or_dmo_params* p, float* out)
{
    out[0] = x0[0]; out[1] = x0[1]; out[2] = x0[2];
    if (n <= 0 || n >= MAXR) return 0;
    float ext[3] = { -FLT_MAX, -FLT_MAX, -FLT_MAX };
    float mabs = 0.0f;
    for (int i = 0; i < n; ++i) {
        const float* q = fan + 3 * (size_t)i;
        for (int a = 0; a < 3; ++a) ext[a] = fmaxf(ext[a], fabsf(q[a] - x0[a]));
        mabs = fmaxf(mabs, fmaxf(fabsf(q[0] - x0[0]), fabsf(q[1] - x0[1])));
    }
    float sigma = or_pow2_scale(mabs);
    float is2 = invpow(sigma, 2);
    tric cs[MAXR];
    for (int e = 0; e < n; ++e) {
        const float* a = fan + 3 * (size_t)e;
        const float* b = fan + 3 * (size_t)((e + 1) % n);
        cs[e] = tri_consts((a[0] - x0[0]) * sigma, (a[1] - x0[1]) * sigma, (b[0] - x0[0]) * sigma, (b[1] - x0[1]) * sigma);
    }
    const int nq = p->nq, half = nq / 2, T = nq * nq / 2;
    const float aff = 1.0f / (float)(nq - 1);
    const float shr = shrink_of(p);
    float cur[2] = { x0[0], x0[1] };
    float qcur = tri_score(n, cs, 0.0f, 0.0f, is2);
    const float qold = qcur;
    float qmaxLast = -FLT_MAX;
    float ds = p->grid_scale;
    for (int d = 0; d < p->depth; ++d) {
        float gmax[2], gmin[2];
        for (int a = 0; a < 2; ++a) { gmax[a] = cur[a] + ds * ext[a]; gmin[a] = cur[a] - ds * ext[a]; }
        float best = -FLT_MAX, bp[2] = { 0, 0 };
        for (int t = 0; t < T; ++t) {
            int i1 = t / nq, j1 = t % nq, i2 = i1 + half;
            float ax = gridc(aff, i1, nq, gmin[0], gmax[0]), ay = gridc(aff, j1, nq, gmin[1], gmax[1]);
            float bx = gridc(aff, i2, nq, gmin[0], gmax[0]), by = ay;
            float q1 = tri_score(n, cs, (ax - x0[0]) * sigma, (ay - x0[1]) * sigma, is2);
            float q2 = tri_score(n, cs, (bx - x0[0]) * sigma, (by - x0[1]) * sigma, is2);
            float qm = (q2 > q1) ? q2 : q1;
            if (qm > best) {
                best = qm;
#if OR_MUTATION == 2
                if (q1 >= q2) {                /* planted: A wins an exact tie */
#else
                if (q1 > q2) {
#endif
                    bp[0] = ax; bp[1] = ay;
                } else {
                    bp[0] = bx; bp[1] = by;
                }
            }
        }
        if (qcur < best) { cur[0] = bp[0]; cur[1] = bp[1]; qcur = best; }
        qmaxLast = best;
        ds = ds * shr;
    }
#if OR_MUTATION == 5
    if (qold < qcur) {                         /* planted: written whenever a level improved */
#else
    if (qold < qmaxLast) {
#endif
        out[0] = cur[0]; out[1] = cur[1];
        return 1;
    }
    return 0;
}

static int vertex_flat(const or_mesh* m, const or_dmo_params* p, int vid, float* xyz)
{
    int h0 = m->heRow[vid], n = m->heRow[vid + 1] - h0;
    if (n <= 0 || n >= MAXR) return 0;
    float fan[3 * MAXR];
    for (int i = 0; i < n; ++i) {
        const float* q = xyz + 3 * (size_t)m->heTarget[h0 + i];
        fan[3 * i] = q[0]; fan[3 * i + 1] = q[1]; fan[3 * i + 2] = q[2];
    }
    float x0[3] = { xyz[3 * (size_t)vid], xyz[3 * (size_t)vid + 1], xyz[3 * (size_t)vid + 2] };
    float out[3];
    int w = search_flat_fan(n, fan, x0, p, out);
    if (w) { xyz[3 * (size_t)vid] = out[0]; xyz[3 * (size_t)vid + 1] = out[1]; }
    return w;
}

int or_learned_window(const or_mesh* m, const or_dmo_params* p, int vid, const float* xyz, const float* c,
                      int depthL, float frac, float* out)
{
    int tet = (m->elem == OR_TET);
    int st = m->ringStride;
    int r0 = m->ringRow[vid], n = m->ringRow[vid + 1] - r0;
    const float* x0 = xyz + 3 * (size_t)vid;
    if (n <= 0 || n > (tet ? MAXR : 64)) { out[0] = x0[0]; out[1] = x0[1]; out[2] = x0[2]; return -1; }
    float pts[MAXR * 9];
    for (int e = 0; e < n; ++e)
        for (int k = 0; k < st; ++k) {
            const float* q = xyz + 3 * (size_t)m->ringVal[(size_t)(r0 + e) * st + k];
            float* o = pts + 3 * (size_t)(e * st + k);
            o[0] = q[0]; o[1] = q[1]; o[2] = q[2];
        }
    return learned_window_ring(tet, n, pts, x0, c, p, depthL, frac, out);
}

int or_search_ring(int elem, int n, const float* ringPts, const float* x0, const or_dmo_params* p, float* out)
{
    if (elem == OR_TET || elem == OR_HEX) return search_volume_ring(elem == OR_TET, n, ringPts, x0, p, out);
    if (elem == OR_TRI) return search_flat_fan(n, ringPts, x0, p, out);
    out[0] = x0[0]; out[1] = x0[1]; out[2] = x0[2];
    return 0;
}

int or_dmo_vertex(const or_mesh* m, const or_dmo_params* p, int vid, float* xyz)
{
    if (m->elem == OR_TET || m->elem == OR_HEX) return vertex_volume(m, p, vid, xyz);
    if (m->elem == OR_TRI) return vertex_flat(m, p, vid, xyz);
    return 0;
}

long long or_dmo_iteration(or_mesh* m, const or_dmo_params* p)
{
    long long cnt = 0;
#ifdef _OPENMP
    int nt = p->nthreads > 0 ? p->nthreads : omp_get_max_threads();
#endif
    if (m->elem == OR_TET || m->elem == OR_HEX) {
        for (int c = 0; c < m->nColorsInner; ++c) {
            int lo = m->colOffInner[c], hi = m->colOffInner[c + 1];
            if (p->max_per_class > 0 && hi - lo > p->max_per_class) hi = lo + p->max_per_class;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
            for (int v = lo; v < hi; ++v) vertex_volume(m, p, v, m->xyz);
            cnt += hi - lo;
        }
    } else if (m->elem == OR_TRI && m->flat) {
        for (int c = 0; c < m->nColorsFree; ++c) {
            int lo = m->colOffFree[c], hi = m->colOffFree[c + 1];
            if (p->max_per_class > 0 && hi - lo > p->max_per_class) hi = lo + p->max_per_class;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
            for (int v = lo; v < hi; ++v) vertex_flat(m, p, v, m->xyz);
            cnt += hi - lo;
        }
    }
    return cnt;
}

/* ------------------------------------------------------------- quality */
static void elem_pts(const or_mesh* m, int e, float* pts)
{
    for (int i = 0; i < m->npe; ++i) {
        const float* q = m->xyz + 3 * (size_t)m->cells[(size_t)e * m->npe + i];
        pts[3 * i] = q[0]; pts[3 * i + 1] = q[1]; pts[3 * i + 2] = q[2];
    }
}

/* the element metric of each mesh kind: tet / hex shape metrics, the 2D
   mean ratio of flat triangles, the unsigned 3D mean ratio of surface
   triangles and quads */
static float elem_q(const or_mesh* m, const float* pts, int hexMetric)
{
    switch (m->elem) {
    case OR_TET: return or_q_tet(pts);
    case OR_HEX: return or_q_hex_m(pts, hexMetric);
    case OR_QUAD: return or_q_quad(pts);
    default: return m->flat ? or_q_tri2d(pts) : or_q_tri3(pts);
    }
}

static double elem_q64(const or_mesh* m, const float* pts, int hexMetric)
{
    switch (m->elem) {
    case OR_TET: return or_q_tet64(pts);
    case OR_HEX: return or_q_hex64_m(pts, hexMetric);
    case OR_QUAD: return or_q_quad64(pts);
    default: return m->flat ? or_q_tri2d64(pts) : or_q_tri364(pts);
    }
}

void or_element_qualities_m(const or_mesh* m, int hexMetric, float* out)
{
    for (int e = 0; e < m->nE; ++e) {
        float pts[24];
        elem_pts(m, e, pts);
        out[e] = elem_q(m, pts, hexMetric);
    }
}

float or_min_quality_m(const or_mesh* m, int hexMetric)
{
    float q = FLT_MAX;
    for (int e = 0; e < m->nE; ++e) {
        float pts[24];
        elem_pts(m, e, pts);
        float v = elem_q(m, pts, hexMetric);
        q = fminf(q, v);
    }
    return q;
}

void or_quality64_m(const or_mesh* m, int hexMetric, double* qmin, double* qmean)
{
    double mn = 1e300, sum = 0.0;
    for (int e = 0; e < m->nE; ++e) {
        float pts[24];
        elem_pts(m, e, pts);
        double v = elem_q64(m, pts, hexMetric);
        if (v < mn) mn = v;
        sum += v;
    }
    *qmin = mn;
    *qmean = m->nE > 0 ? sum / m->nE : 0.0;
}

void or_element_qualities(const or_mesh* m, float* out) { or_element_qualities_m(m, OR_HEX_SHAPE9, out); }
float or_min_quality(const or_mesh* m) { return or_min_quality_m(m, OR_HEX_SHAPE9); }
void or_quality64(const or_mesh* m, double* qmin, double* qmean) { or_quality64_m(m, OR_HEX_SHAPE9, qmin, qmean); }
