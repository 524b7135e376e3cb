/*
 * or_mesh.c -- oracle restatement of the reference's mesh construction for
 * DMO (TEST INFRASTRUCTURE).  Sequential C; every sort is stable, which is
 * what thrust's radix sort_by_key on integer keys gives the reference.
 *
 * Pipeline for volume meshes (reference MeshTetGPU.cu:35-42, MeshHexGPU.cu:35-43):
 *   1 boundary faces           MeshTetGPU.cu:178-200, MeshGPUCommon.cu:25-94,1376-1430
 *   2 [surface | inner] split  MeshTetGPU.cu:241-262 (stable partition)
 *   3 inner adjacency+colour   MeshTetGPU.cu:298-392, MeshGPUCommon.cu:858-951,
 *                              VertexColoring.cu:76-125
 *   4 surface features         MeshTriGPU.cu:463-510, MeshGPUCommon.cu:363-411,466-495,1227-1239
 *   5 [free | feature] split,  MeshTriGPU.cu:646-671,690-784 (adjacency through
 *     colouring of both          the volume elements, MeshTetGPU.cu:522-604,
 *                                MeshHexGPU.cu:287-430)
 *   6 surface halfedges (CCW)  MeshTriGPU.cu:128-214, MeshQuadGPU.cu:156-243,
 *                              MeshGPUCommon.cu:96-135,211-361,565-621
 *   7 volume one-ring CSR      MeshTetGPU.cu:113-142, MeshHexGPU.cu:116-139,
 *                              MeshGPUCommon.cu:1313-1369
 * Surface (tri) meshes run steps 4-6 on all vertices (MeshTriGPU.cu:37-52).
 */
#include "vsdmo_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_ERR_ARG 1
#define OR_ERR_TOPO 4

static void* xmalloc(size_t n) { void* p = malloc(n ? n : 1); return p; }
static int* ialloc(size_t n) { return (int*)xmalloc(n * sizeof(int)); }

/* ---------------------------------------------------------------- sorting */
/* lexicographic comparison of g_k ints (records end with their emit index,
   which makes the order total and qsort therefore stable) */
static int g_k;
static int cmp_lex(const void* a, const void* b)
{
    const int* x = (const int*)a; const int* y = (const int*)b;
    for (int i = 0; i < g_k; ++i) { if (x[i] != y[i]) return x[i] < y[i] ? -1 : 1; }
    return 0;
}

/* stable counting sort of item indices by integer key in [0,nk): returns
   order[] (items grouped by key, original order within a key) and rowPtr */
static void counting_sort(int n, const int* key, int nk, int* order, int* rowPtr)
{
    memset(rowPtr, 0, (size_t)(nk + 1) * sizeof(int));
    for (int i = 0; i < n; ++i) rowPtr[key[i] + 1]++;
    for (int k = 0; k < nk; ++k) rowPtr[k + 1] += rowPtr[k];
    int* fill = ialloc(nk);
    memcpy(fill, rowPtr, (size_t)nk * sizeof(int));
    for (int i = 0; i < n; ++i) order[fill[key[i]]++] = i;
    free(fill);
}

/* ---------------------------------------------------------- boundary faces */
static const int TET_FACES[4][3] = { { 2, 1, 0 }, { 0, 1, 3 }, { 1, 2, 3 }, { 0, 3, 2 } };
static const int HEX_FACES[6][4] = { { 0, 1, 2, 3 }, { 1, 5, 6, 2 }, { 3, 2, 6, 7 },
                                     { 4, 0, 3, 7 }, { 1, 0, 4, 5 }, { 5, 4, 7, 6 } };

/* rotate so the smallest vertex comes first (reference getTriangleSmallestFirst
   / getQuadSmallestFirst, MeshGPUCommon.cu:1376-1396; ties cannot occur for
   valid elements) */
static void rot_min_first(int* f, int k)
{
    int m = 0;
    for (int i = 1; i < k; ++i) if (f[i] < f[m]) m = i;
    int t[4];
    for (int i = 0; i < k; ++i) t[i] = f[(m + i) % k];
    memcpy(f, t, (size_t)k * sizeof(int));
}

/* emit all element faces, sort lexicographically, cancel each face against
   its opposite-orientation twin (reference k_markDuplicateTris/Quads,
   MeshGPUCommon.cu:56-94: a face (a,b,c) with b<c -- quads: v1<v3 -- removes
   itself and the FIRST (a,c,b) -- quads (a,d,c,b) -- after it in sorted
   order).  Survivors keep sorted order (remove_if is stable). */
int or_boundary_faces(int elem, int nE, const int* cells, int* out)
{
    int k = (elem == OR_TET) ? 3 : 4;
    int nf = (elem == OR_TET) ? 4 : 6;
    int npe = (elem == OR_TET) ? 4 : 8;
    int n = nE * nf;
    int w = k + 1;
    int* rec = ialloc((size_t)n * w);
    for (int e = 0; e < nE; ++e)
        for (int f = 0; f < nf; ++f) {
            int* r = rec + (size_t)(e * nf + f) * w;
            for (int i = 0; i < k; ++i)
                r[i] = cells[(size_t)e * npe + (elem == OR_TET ? TET_FACES[f][i] : HEX_FACES[f][i])];
            rot_min_first(r, k);
            r[k] = e * nf + f;
        }
    /* full lexicographic sort; the trailing emit index makes it a total order */
    g_k = w;
    qsort(rec, (size_t)n, (size_t)w * sizeof(int), cmp_lex);
    unsigned char* dead = (unsigned char*)calloc((size_t)n, 1);
    for (int i = 0; i < n; ++i) {
        const int* a = rec + (size_t)i * w;
        int search = (k == 3) ? (a[1] < a[2]) : (a[1] < a[3]);
        if (!search) continue;
        for (int j = i + 1; j < n; ++j) {
            const int* b = rec + (size_t)j * w;
            if (b[0] > a[0]) break;
            int match = (k == 3) ? (b[0] == a[0] && b[1] == a[2] && b[2] == a[1])
                                 : (b[0] == a[0] && b[1] == a[3] && b[2] == a[2] && b[3] == a[1]);
            if (match) { dead[i] = 1; dead[j] = 1; break; }
        }
    }
    int m = 0;
    for (int i = 0; i < n; ++i) {
        if (dead[i]) continue;
        memcpy(out + (size_t)m * k, rec + (size_t)i * w, (size_t)k * sizeof(int));
        ++m;
    }
    free(dead);
    free(rec);
    return m;
}

/* -------------------------------------------------------------- colouring */
/* Greedy: sequential first-fit in vertex order, i.e. the colouring the
 * Jones-Plassmann rounds with the reference's priority n-i
 * (VertexColoring.cu:107-125) produce when each vertex takes the smallest
 * colour unused by its already-coloured neighbours.
 * JPL: the reference's color_jpl as written: round c colours every uncoloured
 * vertex whose priority n-i exceeds that of all uncoloured in-range
 * neighbours (VertexColoring.cu:76-105), colour = round. */
int or_color(int mode, int n, const int* rowPtr, const int* colInd, int* colors)
{
    if (n <= 0) return 0;
    int nc = 0;
    if (mode == OR_COLOR_GREEDY) {
        int* mark = ialloc((size_t)n + 1);
        for (int i = 0; i <= n; ++i) mark[i] = -1;
        for (int v = 0; v < n; ++v) {
            for (int k = rowPtr[v]; k < rowPtr[v + 1]; ++k) {
                int u = colInd[k];
                if (u < v) { int c = colors[u]; if (c <= n) mark[c] = v; }
            }
            int c = 0;
            while (mark[c] == v) ++c;
            colors[v] = c;
            if (c + 1 > nc) nc = c + 1;
        }
        free(mark);
    } else {
        for (int i = 0; i < n; ++i) colors[i] = -1;
        int left = n;
        unsigned char* take = (unsigned char*)xmalloc((size_t)n);
        for (int c = 0; c < n && left > 0; ++c) {
            for (int i = 0; i < n; ++i) {
                take[i] = 0;
                if (colors[i] != -1) continue;
                int ok = 1;
                for (int k = rowPtr[i]; k < rowPtr[i + 1]; ++k) {
                    int j = colInd[k];
                    if (j == i || colors[j] != -1) continue;
                    /* priority n-i > n-j  <=>  i < j */
                    if (!(i < j)) { ok = 0; break; }
                }
                take[i] = (unsigned char)ok;
            }
            for (int i = 0; i < n; ++i) if (take[i]) { colors[i] = c; --left; }
            nc = c + 1;
        }
        free(take);
    }
    return nc;
}

/* adjacency "shares an element" restricted to vertex range [lo,hi), local ids,
   deduplicated, neighbours ascending (reference fulledge kernels + zip sort +
   unique, MeshTetGPU.cu:460-604, MeshHexGPU.cu:287-430 'Fixed' all-pairs;
   tris: 3 edges, MeshGPUCommon.cu:1046-1071) */
int or_adjacency_range(int elem, int nE, const int* cells, int lo, int hi, int** rowPtrOut, int** colIndOut)
{
    int npe = (elem == OR_TRI) ? 3 : (elem == OR_QUAD || elem == OR_TET) ? 4 : 8;
    int n = hi - lo;
    /* pairs per element */
    int np = 0;
    int pa[28], pb[28];
    if (elem == OR_QUAD) {
        static const int q[6][2] = { { 0, 1 }, { 1, 2 }, { 2, 3 }, { 3, 0 }, { 0, 2 }, { 1, 3 } };
        for (int i = 0; i < 6; ++i) { pa[np] = q[i][0]; pb[np] = q[i][1]; ++np; }
    } else {
        for (int i = 0; i < npe; ++i) for (int j = i + 1; j < npe; ++j) { pa[np] = i; pb[np] = j; ++np; }
    }
    size_t cap = (size_t)nE * np * 2;
    int* key = ialloc(cap); int* val = ialloc(cap);
    size_t m = 0;
    for (int e = 0; e < nE; ++e) {
        const int* c = cells + (size_t)e * npe;
        for (int p = 0; p < np; ++p) {
            int a = c[pa[p]], b = c[pb[p]];
            if (a >= lo && a < hi && b >= lo && b < hi) {
                key[m] = a - lo; val[m] = b - lo; ++m;
                key[m] = b - lo; val[m] = a - lo; ++m;
            }
        }
    }
    int* rowPtr = ialloc((size_t)n + 1);
    int* order = ialloc(m);
    counting_sort((int)m, key, n, order, rowPtr);
    int* tmp = ialloc(m);
    for (size_t i = 0; i < m; ++i) tmp[i] = val[order[i]];
    /* sort each row ascending and unique */
    int* col = ialloc(m);
    int w = 0;
    int* newRow = ialloc((size_t)n + 1);
    newRow[0] = 0;
    for (int v = 0; v < n; ++v) {
        int s = rowPtr[v], t = rowPtr[v + 1];
        /* insertion sort */
        for (int i = s + 1; i < t; ++i) {
            int x = tmp[i], j = i - 1;
            while (j >= s && tmp[j] > x) { tmp[j + 1] = tmp[j]; --j; }
            tmp[j + 1] = x;
        }
        for (int i = s; i < t; ++i) if (i == s || tmp[i] != tmp[i - 1]) col[w++] = tmp[i];
        newRow[v + 1] = w;
    }
    free(key); free(val); free(order); free(tmp); free(rowPtr);
    *rowPtrOut = newRow; *colIndOut = col;
    return w;
}

/* ----------------------------------------------------------- renumbering */
typedef struct {
    int nV;
    float* xyz;
    int* new_of_old;   /* composition of all permutations so far */
    int* old_of_new;
} numbering;

/* apply permutation given as order[newpos] = current id, restricted to range
   [lo, lo+n) (ids outside unchanged); remaps cells/faces; permutes positions
   and the per-vertex byte arrays in 'flags' */
static void apply_perm(numbering* nb, int lo, int n, const int* order,
                       int* cells, size_t nCellIdx, int* faces, size_t nFaceIdx,
                       unsigned char** flags, int nflags, int* colorArr)
{
    int* inv = ialloc((size_t)n);          /* inv[current - lo] = newpos */
    for (int i = 0; i < n; ++i) inv[order[i]] = i;
    for (size_t i = 0; i < nCellIdx; ++i) { int v = cells[i]; if (v >= lo && v < lo + n) cells[i] = lo + inv[v - lo]; }
    for (size_t i = 0; i < nFaceIdx; ++i) { int v = faces[i]; if (v >= lo && v < lo + n) faces[i] = lo + inv[v - lo]; }
    float* nx = (float*)xmalloc((size_t)n * 3 * sizeof(float));
    for (int i = 0; i < n; ++i) memcpy(nx + 3 * i, nb->xyz + 3 * (size_t)(lo + order[i]), 3 * sizeof(float));
    memcpy(nb->xyz + 3 * (size_t)lo, nx, (size_t)n * 3 * sizeof(float));
    free(nx);
    for (int f = 0; f < nflags; ++f) {
        unsigned char* a = flags[f];
        if (!a) continue;
        unsigned char* t = (unsigned char*)xmalloc((size_t)n);
        for (int i = 0; i < n; ++i) t[i] = a[lo + order[i]];
        memcpy(a + lo, t, (size_t)n);
        free(t);
    }
    if (colorArr) {
        int* t = ialloc((size_t)n);
        for (int i = 0; i < n; ++i) t[i] = colorArr[lo + order[i]];
        memcpy(colorArr + lo, t, (size_t)n * sizeof(int));
        free(t);
    }
    /* compose: old id o had current id c = new_of_old[o]; its new id */
    for (int o = 0; o < nb->nV; ++o) {
        int c = nb->new_of_old[o];
        if (c >= lo && c < lo + n) nb->new_of_old[o] = lo + inv[c - lo];
    }
    for (int o = 0; o < nb->nV; ++o) nb->old_of_new[nb->new_of_old[o]] = o;
    free(inv);
}

/* stable partition of [lo,lo+n): vertices with key==first come first */
static int stable_partition(const unsigned char* key, int lo, int n, unsigned char first, int* order)
{
    int m = 0;
    for (int i = 0; i < n; ++i) if (key[lo + i] == first) order[m++] = i;
    int cnt = m;
    for (int i = 0; i < n; ++i) if (key[lo + i] != first) order[m++] = i;
    return cnt;
}

/* colour a vertex range and sort it by colour (stable); returns nColors and
   fills colOff (absolute ids) */
static int colour_and_sort(or_mesh* m, numbering* nb, int lo, int hi, const int* adjCells, int adjElem,
                           int adjNE, int mode, int** colOffOut, int* faces, size_t nFaceIdx,
                           unsigned char** flags, int nflags)
{
    int n = hi - lo;
    if (n <= 0) { int* co = ialloc(1); co[0] = lo; *colOffOut = co; return 0; }
    int *rp, *ci;
    or_adjacency_range(adjElem, adjNE, adjCells, lo, hi, &rp, &ci);
    int* col = ialloc((size_t)n);
    int nc = or_color(mode, n, rp, ci, col);
    free(rp); free(ci);
    /* stable counting sort by colour */
    int* order = ialloc((size_t)n);
    int* off = ialloc((size_t)nc + 1);
    counting_sort(n, col, nc, order, off);
    for (int i = 0; i < n; ++i) m->colorOf[lo + i] = col[i];
    apply_perm(nb, lo, n, order, m->cells, (size_t)m->nE * m->npe, faces, nFaceIdx, flags, nflags, m->colorOf);
    for (int c = 0; c <= nc; ++c) off[c] += lo;
    free(order); free(col);
    *colOffOut = off;
    return nc;
}

/* ------------------------------------------------------- surface halfedges */
typedef struct { int target, opp, face, next; } hedge;

/* face normal (reference k_calcFaceNormals, MeshGPUCommon.cu:363-411; quads use
   v0, v1, v3): canonical cross and IEEE sqrt / division */
static void face_normal(const float* xyz, const int* f, int k, float* n)
{
    const float* p0 = xyz + 3 * (size_t)f[0];
    const float* p1 = xyz + 3 * (size_t)f[1];
    const float* p2 = xyz + 3 * (size_t)f[k == 3 ? 2 : 3];
    float e0x = p1[0] - p0[0], e0y = p1[1] - p0[1], e0z = p1[2] - p0[2];
    float e1x = p2[0] - p0[0], e1y = p2[1] - p0[1], e1z = p2[2] - p0[2];
    float cx = e0y * e1z - e0z * e1y, cy = e0z * e1x - e0x * e1z, cz = e0x * e1y - e0y * e1x;
    float l = sqrtf(fmaf(cx, cx, fmaf(cy, cy, cz * cz)));
    n[0] = cx / l; n[1] = cy / l; n[2] = cz / l;
}

/* halfedges of a face list over vertices [0,nS): CSR by source in face order
   (stable), opposite = LAST match in the target's row
   (k_setOppositeHalfedgeHelper, MeshGPUCommon.cu:107-121) */
static void build_he_raw(int nS, int nF, int k, const int* faces, int** rowOut, hedge** heOut, int* nOut)
{
    int n = nF * k;
    int* key = ialloc((size_t)n);
    hedge* h = (hedge*)xmalloc((size_t)n * sizeof(hedge));
    for (int f = 0; f < nF; ++f)
        for (int i = 0; i < k; ++i) {
            key[f * k + i] = faces[(size_t)f * k + i];
            h[f * k + i].target = faces[(size_t)f * k + (i + 1) % k];
            h[f * k + i].face = f; h[f * k + i].opp = -1; h[f * k + i].next = -1;
        }
    int* order = ialloc((size_t)n);
    int* row = ialloc((size_t)nS + 1);
    counting_sort(n, key, nS, order, row);
    hedge* s = (hedge*)xmalloc((size_t)n * sizeof(hedge));
    for (int i = 0; i < n; ++i) s[i] = h[order[i]];
    for (int v = 0; v < nS; ++v)
        for (int i = row[v]; i < row[v + 1]; ++i) {
            int d = s[i].target, opp = -1;
            for (int j = row[d]; j < row[d + 1]; ++j) if (s[j].target == v) opp = j;
            s[i].opp = opp;
        }
    free(key); free(h); free(order);
    *rowOut = row; *heOut = s; *nOut = n;
}

static int third_vertex(const int* f, int k, int src, int dst)
{
    if (k == 3) {
        if (f[0] != src && f[0] != dst) return f[0];
        if (f[1] != src && f[1] != dst) return f[1];
        return f[2];
    }
    /* quads: the vertex before src (getThirdVertex for Quad, MeshGPUCommon.cu:149-161) */
    if (dst == f[1]) return f[3];
    if (dst == f[2]) return f[0];
    if (dst == f[3]) return f[1];
    return f[2];
}

/* full surface halfedge construction (reference constructTriHalfedgesNew,
   MeshTriGPU.cu:128-214; quads MeshQuadGPU.cu:156-243) */
static int build_surface_halfedges(or_mesh* m)
{
    int nS = m->nSurf, nF = m->nFaces, k = m->fpe;
    int* row; hedge* h; int n;
    build_he_raw(nS, nF, k, m->faces, &row, &h, &n);
    /* dummies: for a halfedge without opposite add dst->src with face -1,
       appended in source-sorted position order (k_addBoundaryHalfedges,
       MeshGPUCommon.cu:565-575), then a second stable sort by source */
    int nd = 0;
    for (int i = 0; i < n; ++i) if (h[i].opp == -1) ++nd;
    int tot = n + nd;
    int* key = ialloc((size_t)tot);
    hedge* all = (hedge*)xmalloc((size_t)tot * sizeof(hedge));
    for (int v = 0; v < nS; ++v) for (int i = row[v]; i < row[v + 1]; ++i) { key[i] = v; all[i] = h[i]; }
    int w = n;
    for (int i = 0; i < n; ++i)
        if (h[i].opp == -1) {
            key[w] = h[i].target;
            all[w].target = key[i]; all[w].face = -1; all[w].opp = -1; all[w].next = -1;
            ++w;
        }
    int* order = ialloc((size_t)tot);
    int* row2 = ialloc((size_t)nS + 1);
    counting_sort(tot, key, nS, order, row2);
    hedge* srt = (hedge*)xmalloc((size_t)tot * sizeof(hedge));
    for (int i = 0; i < tot; ++i) srt[i] = all[order[i]];
    for (int v = 0; v < nS; ++v)
        for (int i = row2[v]; i < row2[v + 1]; ++i) {
            int d = srt[i].target, opp = -1;
            for (int j = row2[d]; j < row2[d + 1]; ++j) if (srt[j].target == v) opp = j;
            srt[i].opp = opp;
        }
    /* CCW order per vertex (k_sortCCWDummyOpposite, MeshGPUCommon.cu:211-361) */
    hedge* out = (hedge*)xmalloc((size_t)tot * sizeof(hedge));
    int err = 0;
    for (int v = 0; v < nS; ++v) {
        int st = row2[v], cnt = row2[v + 1] - st;
        if (cnt == 0) { err = 1; continue; }
        int pos = 0;
        hedge cur = srt[st];
        if (m->isBoundary1D[v]) {
            int j = 0;
            while (cur.opp < 0 || srt[cur.opp].face != -1) {
                ++j;
                if (j >= cnt) { err = 1; break; }
                cur = srt[st + j];
            }
        } else if (m->isFeature[v]) {
            int best = -1, bt = 0x7fffffff;
            for (int j = 0; j < cnt; ++j) {
                int t = srt[st + j].target;
                if (m->isFeature[t] && t < bt) { bt = t; best = j; }
            }
            if (best < 0) { err = 1; continue; }
            cur = srt[st + best];
        } else {
            for (int j = 1; j < cnt; ++j) if (srt[st + j].target < cur.target) cur = srt[st + j];
        }
        out[st + pos++] = cur;
        int found;
        do {
            if (cur.face < 0) { err = 1; break; }
            int nxt = third_vertex(m->faces + (size_t)cur.face * k, k, v, cur.target);
            found = 0;
            for (int j = 0; j < cnt; ++j)
                if (srt[st + j].target == nxt) {
                    found = 1; cur = srt[st + j];
                    if (pos < cnt) out[st + pos] = cur;
                    ++pos;
                }
        } while (found && pos < cnt);
        if (pos != cnt) err = 1;
    }
    /* opposite and next on the final order (MeshGPUCommon.cu:107-135; last match wins) */
    for (int v = 0; v < nS; ++v)
        for (int i = row2[v]; i < row2[v + 1]; ++i) {
            int d = out[i].target, opp = -1, nx = -1;
            for (int j = row2[d]; j < row2[d + 1]; ++j) {
                if (out[j].target == v) opp = j;
                if (out[j].face == out[i].face) nx = j;
            }
            out[i].opp = opp; out[i].next = nx;
        }
    m->nHE = tot;
    m->heRow = row2;
    m->heTarget = ialloc((size_t)tot); m->heOpp = ialloc((size_t)tot);
    m->heFace = ialloc((size_t)tot); m->heNext = ialloc((size_t)tot);
    m->maxHE = 0;
    for (int i = 0; i < tot; ++i) {
        m->heTarget[i] = out[i].target; m->heOpp[i] = out[i].opp;
        m->heFace[i] = out[i].face; m->heNext[i] = out[i].next;
    }
    for (int v = 0; v < nS; ++v) if (row2[v + 1] - row2[v] > m->maxHE) m->maxHE = row2[v + 1] - row2[v];
    free(row); free(h); free(key); free(all); free(order); free(srt); free(out);
    return err ? OR_ERR_TOPO : 0;
}

/* features: boundary-1D vertices (a halfedge without opposite) and, unless
   flat, both ends of every halfedge whose face normals enclose more than the
   feature angle: cos(angle) computed as dot/sqrt(|n1|^2 |n2|^2) and compared
   with cos(threshold) (reference k_initFeaturesRowPtr, MeshGPUCommon.cu:466-495,
   compares acos(...)*180/pi > 20 -- equivalent up to rounding at the threshold) */
static void surface_features(or_mesh* m, float cthr)
{
    int nS = m->nSurf, nF = m->nFaces, k = m->fpe;
    int* row; hedge* h; int n;
    build_he_raw(nS, nF, k, m->faces, &row, &h, &n);
    memset(m->isBoundary1D, 0, (size_t)nS);
    memset(m->isFeature, 0, (size_t)nS);
    for (int v = 0; v < nS; ++v)
        for (int i = row[v]; i < row[v + 1]; ++i) if (h[i].opp == -1) { m->isBoundary1D[v] = 1; break; }
    if (m->flat) {
        memcpy(m->isFeature, m->isBoundary1D, (size_t)nS);
    } else {
        float* fn = (float*)xmalloc((size_t)nF * 3 * sizeof(float));
        for (int f = 0; f < nF; ++f) face_normal(m->xyz, m->faces + (size_t)f * k, k, fn + 3 * f);
        for (int v = 0; v < nS; ++v)
            for (int i = row[v]; i < row[v + 1]; ++i) {
                int d = h[i].target;
                if (h[i].opp == -1) { m->isFeature[v] = 1; m->isFeature[d] = 1; continue; }
                const float* a = fn + 3 * h[i].face;
                const float* b = fn + 3 * h[h[i].opp].face;
                float dot = fmaf(a[0], b[0], fmaf(a[1], b[1], a[2] * b[2]));
                float la = fmaf(a[0], a[0], fmaf(a[1], a[1], a[2] * a[2]));
                float lb = fmaf(b[0], b[0], fmaf(b[1], b[1], b[2] * b[2]));
                float c = dot / sqrtf(la * lb);
                if (c < cthr) { m->isFeature[v] = 1; m->isFeature[d] = 1; }
            }
        free(fn);
    }
    free(row); free(h);
}

/* --------------------------------------------------------- volume one-ring */
static const int HALFHEX[8][7] = {
    { 1, 2, 3, 4, 5, 6, 7 }, { 5, 6, 2, 0, 4, 7, 3 }, { 6, 7, 3, 1, 5, 4, 0 }, { 7, 4, 0, 2, 6, 5, 1 },
    { 0, 3, 7, 5, 1, 2, 6 }, { 4, 7, 6, 1, 0, 3, 2 }, { 5, 4, 7, 2, 1, 0, 3 }, { 6, 5, 4, 3, 2, 1, 0 } };
static const int HALFFACE[4][3] = { { 1, 2, 3 }, { 0, 3, 2 }, { 0, 1, 3 }, { 0, 2, 1 } };

static void build_ring(or_mesh* m)
{
    int tet = (m->elem == OR_TET);
    int per = tet ? 4 : 8, st = tet ? 3 : 7;
    int n = m->nE * per;
    int* key = ialloc((size_t)n);
    int* val = ialloc((size_t)n * st);
    for (int e = 0; e < m->nE; ++e) {
        const int* c = m->cells + (size_t)e * per;
        for (int i = 0; i < per; ++i) {
            key[e * per + i] = c[i];
            for (int j = 0; j < st; ++j)
                val[(size_t)(e * per + i) * st + j] = c[tet ? HALFFACE[i][j] : HALFHEX[i][j]];
        }
    }
    int* order = ialloc((size_t)n);
    m->ringRow = ialloc((size_t)m->nV + 1);
    counting_sort(n, key, m->nV, order, m->ringRow);
    m->ringVal = ialloc((size_t)n * st);
    for (int i = 0; i < n; ++i) memcpy(m->ringVal + (size_t)i * st, val + (size_t)order[i] * st, (size_t)st * sizeof(int));
    m->ringStride = st;
    m->maxRing = 0;
    for (int v = 0; v < m->nV; ++v) if (m->ringRow[v + 1] - m->ringRow[v] > m->maxRing) m->maxRing = m->ringRow[v + 1] - m->ringRow[v];
    free(key); free(val); free(order);
}

/* ------------------------------------------------------------------ build */
void or_mesh_free(or_mesh* m)
{
    if (!m) return;
    free(m->xyz); free(m->cells); free(m->new_of_old); free(m->old_of_new);
    free(m->colOffInner); free(m->colOffFree); free(m->colOffFeature); free(m->colorOf);
    free(m->faces); free(m->ringRow); free(m->ringVal);
    free(m->heRow); free(m->heTarget); free(m->heOpp); free(m->heFace); free(m->heNext);
    free(m->isBoundary1D); free(m->isFeature);
    free(m);
}

int or_mesh_build(int elem, const float* xyz, int nV, const int* cells, int nE,
                  const or_build_params* prm, or_mesh** out)
{
    *out = NULL;
    if (elem != OR_TRI && elem != OR_TET && elem != OR_HEX && elem != OR_QUAD) return OR_ERR_ARG;
    int npe = (elem == OR_TRI) ? 3 : (elem == OR_HEX) ? 8 : 4;
    if (nV <= 0 || nE <= 0) return OR_ERR_ARG;
    for (size_t i = 0; i < (size_t)nE * npe; ++i) if (cells[i] < 0 || cells[i] >= nV) return OR_ERR_ARG;
    int mode = prm ? prm->coloring : OR_COLOR_GREEDY;
    float ang = (prm && prm->feature_angle_deg > 0.0f) ? prm->feature_angle_deg : 20.0f;
    float cthr = (float)cos((double)ang * 3.14159265358979323846 / 180.0);

    or_mesh* m = (or_mesh*)calloc(1, sizeof(or_mesh));
    m->elem = elem; m->npe = npe; m->nV = nV; m->nE = nE;
    m->xyz = (float*)xmalloc((size_t)nV * 3 * sizeof(float));
    memcpy(m->xyz, xyz, (size_t)nV * 3 * sizeof(float));
    m->cells = ialloc((size_t)nE * npe);
    memcpy(m->cells, cells, (size_t)nE * npe * sizeof(int));
    m->new_of_old = ialloc((size_t)nV); m->old_of_new = ialloc((size_t)nV);
    for (int i = 0; i < nV; ++i) { m->new_of_old[i] = i; m->old_of_new[i] = i; }
    m->colorOf = ialloc((size_t)nV);
    for (int i = 0; i < nV; ++i) m->colorOf[i] = 0;
    numbering nb = { nV, m->xyz, m->new_of_old, m->old_of_new };

    int volume = (elem == OR_TET || elem == OR_HEX);
    unsigned char* isB = (unsigned char*)calloc((size_t)nV, 1);
    if (volume) {
        m->fpe = (elem == OR_TET) ? 3 : 4;
        int* fb = ialloc((size_t)nE * (elem == OR_TET ? 12 : 24));
        m->nFaces = or_boundary_faces(elem, nE, cells, fb);
        m->faces = fb;
        for (int i = 0; i < m->nFaces * m->fpe; ++i) isB[fb[i]] = 1;
        /* [surface | inner], stable (sortVerticesByBoundary2D) */
        int* order = ialloc((size_t)nV);
        int nS = stable_partition(isB, 0, nV, 1, order);
        unsigned char* fl[1] = { isB };
        apply_perm(&nb, 0, nV, order, m->cells, (size_t)nE * npe, m->faces, (size_t)m->nFaces * m->fpe, fl, 1, NULL);
        free(order);
        m->nSurf = nS; m->nInner = nV - nS;
        /* inner colouring (colorInnerVerticesAndSort) */
        m->nColorsInner = colour_and_sort(m, &nb, nS, nV, m->cells, elem, nE, mode, &m->colOffInner,
                                          m->faces, (size_t)m->nFaces * m->fpe, NULL, 0);
    } else {
        m->fpe = npe;
        m->nFaces = nE;
        m->faces = ialloc((size_t)nE * npe);
        memcpy(m->faces, m->cells, (size_t)nE * npe * sizeof(int));
        m->nSurf = nV; m->nInner = 0;
        m->nColorsInner = 0;
        m->colOffInner = ialloc(1); m->colOffInner[0] = nV;
        int flat = 1;
        for (int i = 0; i < nV; ++i) if (xyz[3 * i + 2] != 0.0f) { flat = 0; break; }
        m->flat = flat;
    }
    free(isB);

    /* surface processing on [0,nSurf) */
    int nS = m->nSurf;
    m->isBoundary1D = (unsigned char*)calloc((size_t)(nS > 0 ? nS : 1), 1);
    m->isFeature = (unsigned char*)calloc((size_t)(nS > 0 ? nS : 1), 1);
    int rc = 0;
    if (nS > 0 && m->nFaces > 0) {
        surface_features(m, cthr);
        /* [free | feature], stable (sortSurfVerticesByFeature: key false first) */
        int* order = ialloc((size_t)nS);
        int nFree = stable_partition(m->isFeature, 0, nS, 0, order);
        /* the faces array and cells are both remapped (remapElements is virtual) */
        unsigned char* fl[2] = { m->isBoundary1D, m->isFeature };
        if (volume) {
            apply_perm(&nb, 0, nS, order, m->cells, (size_t)nE * npe, m->faces, (size_t)m->nFaces * m->fpe, fl, 2, NULL);
        } else {
            apply_perm(&nb, 0, nS, order, m->cells, (size_t)nE * npe, m->faces, (size_t)m->nFaces * m->fpe, fl, 2, NULL);
        }
        free(order);
        m->nFree = nFree; m->nFeature = nS - nFree;
        /* colour free and feature ranges: adjacency through the volume
           elements for volume meshes, through the surface elements otherwise */
        const int* adjC = m->cells; int adjE = elem; int adjN = nE;
        m->nColorsFree = colour_and_sort(m, &nb, 0, nFree, adjC, adjE, adjN, mode, &m->colOffFree,
                                         m->faces, (size_t)m->nFaces * m->fpe, fl, 2);
        m->nColorsFeature = colour_and_sort(m, &nb, nFree, nS, adjC, adjE, adjN, mode, &m->colOffFeature,
                                            m->faces, (size_t)m->nFaces * m->fpe, fl, 2);
        /* keep face list consistent for surface meshes (cells == faces) */
        if (!volume) memcpy(m->faces, m->cells, (size_t)nE * npe * sizeof(int));
        rc = build_surface_halfedges(m);
    } else {
        m->colOffFree = ialloc(1); m->colOffFree[0] = 0;
        m->colOffFeature = ialloc(1); m->colOffFeature[0] = 0;
    }
    if (volume) build_ring(m);
    *out = m;
    return rc;
}
