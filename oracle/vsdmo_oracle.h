/*
 * vsdmo_oracle.h -- CPU oracle for the DMO hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This is a plain-C restatement of the reference's discrete mesh optimisation
 * (DMO) algorithm, used exclusively by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Nothing in the product links,
 * loads or calls it.
 *
 * Reference files restated (paths relative to the reference repository):
 *   mesh build  : VSDMO/src/mesh/MeshTetGPU.cu, MeshHexGPU.cu, MeshTriGPU.cu,
 *                 MeshGPUCommon.cu, SortUtil.cu, VertexColoring.cu
 *   search      : VSDMO/src/dmo/DMOTetImplCub.h, DMOHexImplCub.h,
 *                 DMOTriFlatImplCub.h
 *   metrics     : VSDMO/src/dmo/DMOQualityTet.h, DMOQualityHex.h,
 *                 DMOQualityTri.h, DMOQualityUtil.h
 *   drivers     : VSDMO/src/dmo/DMOTet.cu, DMOHex.cu, DMOTriFlat.cu, DMOTri.cu,
 *                 DMOQuad.cu, DMOTriImplCub.h, DMOQuadImplCub.h, Laplace.cu
 *   surface metrics: DMOQualityTri.h, DMOQualityQuad.h
 *
 * Canonical FP32 evaluation order.  The reference's own FP32 bits are not
 * reproducible (nvcc contracts mul+add into FMA at its discretion and the
 * reference calls CUDA's cbrtf).  The oracle therefore fixes ONE evaluation
 * order -- the same mathematical quantities, with every rounding step written
 * out (contraction off, explicit fmaf) -- and the HIP kernels reproduce it bit
 * for bit.  The order is documented next to each function below.  Every
 * function is additionally checked against closed forms and FP64 re-evaluation
 * in tests/.
 */
#ifndef VSDMO_ORACLE_H
#define VSDMO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_TRI = 0, OR_QUAD = 1, OR_TET = 2, OR_HEX = 3 };
enum { OR_COLOR_GREEDY = 0, OR_COLOR_JPL = 1 };
enum { OR_EXTENT_MAX = 0, OR_EXTENT_MIN = 1 };
/* hex metric: the reference's 9 sub-tets, or the 8 corner tets alone */
enum { OR_HEX_SHAPE9 = 0, OR_HEX_CORNER8 = 1 };

/* ---- canonical scalar primitives ---- */
float or_cbrt(float x);                 /* own cube root (Newton on x^-1/3) */
float or_pow2_scale(float m);           /* 2^-e with m = f*2^e, f in [1,2)   */

/* ---- element qualities (canonical FP32, vertex-0 frame) ---- */
float or_q_tet(const float* p4x3);      /* mean ratio, signed volume if inverted */
float or_q_hex(const float* p8x3);      /* 9-sub-tet shape metric, internal order */
float or_q_hex_m(const float* p8x3, int metric);   /* OR_HEX_SHAPE9 / OR_HEX_CORNER8 */
float or_q_tri2d(const float* p3x3);    /* 2D mean ratio (x,y), signed if inverted */
/* surface triangles / quads (literal FP32 with contraction off): unsigned
   3D mean ratio, and the variant signed by a unit normal n (quads: -1 when
   a corner triangle's normal has n . u < 0.7) */
float or_q_tri3(const float* p3x3);
float or_q_tri3n(const float* p3x3, const float* n);
float or_q_quad(const float* p4x3);
float or_q_quadn(const float* p4x3, const float* n);
/* ---- FP64 literal re-evaluation (reference formulas in double) ---- */
double or_q_tet64(const float* p4x3);
double or_q_hex64(const float* p8x3);
double or_q_hex64_m(const float* p8x3, int metric);
double or_q_tri2d64(const float* p3x3);
double or_q_tri364(const float* p3x3);
double or_q_quad64(const float* p4x3);

/* ---- ring scores (what the search maximises) ---- */
/* ring: nElem elements of 3 (tet), 7 (hex) or 2 (tri fan pair) points, each
   point xyz; x0: frame origin (vertex position at search start); p: candidate */
float or_ring_score_tet(int nElem, const float* ring, const float* x0, float sigma, const float* p);
float or_ring_score_hex(int nElem, const float* ring, const float* x0, float sigma, const float* p);
float or_ring_score_hex_m(int nElem, const float* ring, const float* x0, float sigma, const float* p, int metric);
float or_ring_score_tri2d(int nElem, const float* ring, const float* x0, float sigma, const float* p);

/* ---- mesh build ---- */
typedef struct or_mesh {
    int elem;            /* OR_TRI / OR_QUAD / OR_TET / OR_HEX */
    int npe;             /* vertices per element */
    int nV, nE;
    int flat;            /* tri meshes: 1 if all z == 0 */
    float* xyz;          /* 3*nV, final (renumbered) order */
    int* cells;          /* npe*nE, final ids */
    int* new_of_old;     /* nV */
    int* old_of_new;     /* nV */
    /* vertex ranges: [0,nFree) free surface | [nFree,nSurf) feature | [nSurf,nV) inner */
    int nSurf, nFeature, nFree, nInner;
    int nColorsInner, nColorsFree, nColorsFeature;
    int* colOffInner;    /* nColorsInner+1, absolute ids */
    int* colOffFree;     /* nColorsFree+1 */
    int* colOffFeature;  /* nColorsFeature+1, absolute ids */
    int* colorOf;        /* nV: colour within its class family */
    /* boundary faces of a volume mesh (tri for tet, quad for hex); for surface
       meshes these are the elements themselves */
    int nFaces, fpe;
    int* faces;          /* fpe*nFaces, final ids */
    /* volume one-ring CSR over all nV vertices: halffaces (3) / halfhexes (7) */
    int ringStride;
    int* ringRow;        /* nV+1 */
    int* ringVal;        /* ringStride*ringRow[nV] */
    int maxRing;
    /* surface halfedges CSR over [0,nSurf) */
    int nHE;
    int* heRow;          /* nSurf+1 */
    int* heTarget; int* heOpp; int* heFace; int* heNext;
    int maxHE;
    unsigned char* isBoundary1D;  /* nSurf */
    unsigned char* isFeature;     /* nSurf */
} or_mesh;

typedef struct or_build_params {
    int coloring;        /* OR_COLOR_GREEDY (default) or OR_COLOR_JPL */
    float feature_angle_deg; /* 20 in the reference */
} or_build_params;

/* returns 0 on success, nonzero error code otherwise */
int or_mesh_build(int elem, const float* xyz, int nV, const int* cells, int nE,
                  const or_build_params* prm, or_mesh** out);
void or_mesh_free(or_mesh* m);

/* individual build steps (exported for step-level tests) */
int or_boundary_faces(int elem, int nE, const int* cells, int* outFaces); /* returns nFaces */
int or_color(int mode, int n, const int* rowPtr, const int* colInd, int* colors); /* returns nColors */
int or_adjacency_range(int elem, int nE, const int* cells, int lo, int hi,
                       int** rowPtr, int** colInd);   /* CSR over [lo,hi), local ids */

/* ---- DMO ---- */
typedef struct or_dmo_params {
    int nq;              /* candidates per axis (8) */
    int depth;           /* refinement levels (3) */
    float grid_scale;    /* initial window scale (0.5) */
    float shrink;        /* per-level window factor; 0 => 2/(nq-1) as the reference */
    int extent_mode;     /* OR_EXTENT_MAX (tet/tri) or OR_EXTENT_MIN (hex) */
    int fixed_boundary;  /* 1: surface/feature passes skipped */
    int nthreads;        /* OpenMP threads, 0 = default */
    int max_per_class;   /* > 0: search only the first max_per_class vertices of
                            each colour class (a timing sample); 0 = all */
    int hex_metric;      /* OR_HEX_SHAPE9 (reference) or OR_HEX_CORNER8 */
} or_dmo_params;

/* one DMO iteration over the inner colour classes (volume meshes) or the free
   colour classes (flat tri meshes); positions updated in place in m->xyz.
   Returns the number of vertex searches executed. */
long long or_dmo_iteration(or_mesh* m, const or_dmo_params* p);
/* search one vertex; exported for brute-force tests.  Returns 1 if moved. */
int or_dmo_vertex(const or_mesh* m, const or_dmo_params* p, int vid, float* xyz_inout);
/* the same search on an explicit ring: elem OR_TET / OR_HEX with nElem
   halffaces (3 points) / halfhexes (7 points), or OR_TRI with a closed fan of
   nElem points; x0 = start position; out = final position (x0 when the vertex
   is not written).  Returns 1 iff the position is written. */
int or_search_ring(int elem, int nElem, const float* ringPts, const float* x0, const or_dmo_params* p, float* out3);
/* the learned proposal's window search of inner vertex vid (tet / hex) around
   the proposal c (world coordinates) on the positions xyz: depthL levels of
   a 4x4x4 grid on frac x the plain first-level window, shrunk by 2/3 per
   level.  Returns 1 when the vertex moves (out = its new position), 0 when
   it stays (a level's best ties a positive q(x0), or a hex vertex is
   pinned), -1 when the vertex takes the plain full search (fallback). */
int or_learned_window(const or_mesh* m, const or_dmo_params* p, int vid, const float* xyz, const float* c,
                      int depthL, float frac, float* out3);

/* ---- surface vertices ---- */
typedef struct or_surf or_surf;
/* local surfaces, feature sectors, 1D curves, NN table, 1D eligibility;
   built from the mesh's current positions */
or_surf* or_surface_init(const or_mesh* m);
/* the same with the quadric fits over {v} + one-ring (ring 1, the
   reference) or {v} + one-ring + 2-ring (ring 2) */
or_surf* or_surface_init_ring(const or_mesh* m, int ring);
void or_surface_free(or_surf* s);
/* inner pass, then (if !fixed_boundary and lastMin > 0) free-surface 2D and
   feature 1D passes, then the nearest-neighbour update; returns searches */
long long or_dmo_iteration_surface(or_mesh* m, or_surf* s, const or_dmo_params* p, float lastMin,
                                   long long* nSurf2d, long long* nFeat1d);
/* surface meshes (tri / quad) on the surface path: free-vertex 2D passes,
   feature 1D passes (unless fixed_boundary), nearest-neighbour update and
   vertex-normal update; no minimum-quality gate */
long long or_dmo_iteration_surface_mesh(or_mesh* m, or_surf* s, const or_dmo_params* p,
                                        long long* nSurf2d, long long* nFeat1d);
/* the surface path's element metric (unsigned 3D mean ratio of triangles /
   quads, also for flat triangle meshes) */
void or_surface_mesh_qualities(const or_mesh* m, float* out);
float or_surface_mesh_min_quality(const or_mesh* m);
/* Laplace smoothing of the free vertices of a surface mesh */
void or_laplace(or_mesh* m, int n_iter);
/* what: 0 NN table (int nS), 1 local surfaces (16 floats each: p0, e0, e1,
   e2, a, b, c, numPoints), 2 1D eligibility (uint8 nS), 3 1D curves (10
   floats: p0, e0, e1, a), 4 vertex normals (3 floats) */
void or_surface_get(const or_surf* s, int what, void* out);
int or_surface_nsectors(const or_surf* s, int nFeature);
/* the surface errors met so far (the optimiser's VSDMO_TOPOLOGY cases): bit 0
   a projection found no sector for a feature neighbour, bit 1 a feature
   vertex has neighbours that belong to no sector */
int or_surface_error(const or_surf* s);
/* feature sectors: the halfedge-to-sector table (TABLE_SIZE = 20 ints per
   feature vertex), the sector row offsets (nFeature + 1) and the sector fits
   (16 floats each, as the local surfaces) */
void or_surface_get_sectors(const or_surf* s, int* table, int* frow, float* surf16);
/* one surface vertex's search on the current positions (dim 2: the free
   vertex 2D search, 1: the feature 1D search); xyz updated in place, uvOut:
   the final local coordinates (u, v) / u.  Returns 1 iff written. */
int or_surface_search_vertex(const or_mesh* m, const or_surf* s, const or_dmo_params* p, int vid, int dim,
                             float* xyz, float* uvOut);
/* the blended projection of (u, v) in vid's nearest neighbour's frame, at
   the current positions */
void or_surface_project(const or_mesh* m, const or_surf* s, int vid, float u, float v, float* out3);
/* the nearest-neighbour walk of all surface vertices from their current table */
void or_surface_update_nn(const or_mesh* m, or_surf* s);

/* element quality statistics in FP32 canonical (min) and FP64 (min, mean) */
float or_min_quality(const or_mesh* m);   /* flat tri: 2D metric; surface tri / quad: unsigned 3D */
void or_quality64(const or_mesh* m, double* qmin, double* qmean);
void or_element_qualities(const or_mesh* m, float* out);
/* the same with the hex metric named (OR_HEX_SHAPE9 / OR_HEX_CORNER8) */
float or_min_quality_m(const or_mesh* m, int hex_metric);
void or_quality64_m(const or_mesh* m, int hex_metric, double* qmin, double* qmean);
void or_element_qualities_m(const or_mesh* m, int hex_metric, float* out);

/* sample points of vid's local surfaces on an nu x nv grid (mode 0: blended
   projection, mode 1: the NN's fit / sector fit sid); out: 3*nu*nv floats */
void or_surface_points(const or_mesh* m, const or_surf* s, int vid, int nu, int nv, int mode, int sid, float gridScale,
                       float* out);
/* test hooks of the surface fit: quadric of pts[0..np) (pts[0] = p0) in the
   frame of (normal, refp - p0); out16 = p0, e0, e1, e2, a, b, c, np */
void or_fit_local_surface(const float* pts, int np, const float* normal, const float* refp, float* out16);
/* 3x3 LU with partial pivoting, A column-major */
void or_lu_solve3(const float* Acm, const float* rhs, float* x);

#ifdef __cplusplus
}
#endif
#endif
