/*
 * vsdmo.h -- C ABI of libvsdmo, the MI355X-native discrete mesh optimisation
 * (DMO) library.  Plain C types only: host pointers, sizes and opaque handles.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository, VSDMO/src/...).  Errors are returned as
 * vsdmo_status; a thread-local message is available from vsdmo_last_error().
 * No call exits the process or throws.
 *
 * Ownership: a vsdmo_mesh owns its device buffers; a vsdmo_opt references a
 * mesh and mutates its vertex positions in place.  A mesh cannot be destroyed
 * while an optimiser references it (VSDMO_INVALID_ARG).  Handles are not
 * thread-safe; use one handle per host thread or lock externally.  All calls
 * are blocking unless stated otherwise.
 */
#ifndef VSDMO_H
#define VSDMO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VSDMO_ABI_VERSION 2

typedef enum vsdmo_status {
    VSDMO_OK = 0,
    VSDMO_INVALID_ARG = 1,
    VSDMO_HIP_ERROR = 2,
    VSDMO_TOPOLOGY = 3,          /* non-manifold input, CCW walk failure */
    VSDMO_QUALITY_DECREASED = 4, /* monotone-quality guard fired */
    VSDMO_LDS_OVERFLOW = 5,      /* one-ring larger than the kernels support */
    VSDMO_NOT_SUPPORTED = 6
} vsdmo_status;

typedef enum vsdmo_elem { VSDMO_TRI = 0, VSDMO_QUAD = 1, VSDMO_TET = 2, VSDMO_HEX = 3 } vsdmo_elem;

/* reference QualityCriterium (dmo/DMOQualityUtil.h:10-18); MEAN_RATIO is the
   supported objective (for hexes the reference's shapeMetricHex) */
typedef enum vsdmo_criterion { VSDMO_MEAN_RATIO = 0 } vsdmo_criterion;

typedef enum vsdmo_coloring {
    VSDMO_COLOR_GREEDY = 0,  /* first-fit in vertex order via Jones-Plassmann rounds, priority n-i */
    VSDMO_COLOR_JPL = 1      /* reference color_jpl (mesh/VertexColoring.cu:76-125): colour = round */
} vsdmo_coloring;

/* which optimiser a triangle / quad mesh gets */
typedef enum vsdmo_surface_path {
    VSDMO_SURFACE_PATH_AUTO = 0,     /* flat path for a flat triangle mesh (z == 0), surface path otherwise */
    VSDMO_SURFACE_PATH_SURFACE = 1,  /* DMOTriClass / DMOQuadClass: vertices move on their local surfaces */
    VSDMO_SURFACE_PATH_FLAT = 2      /* DMOTriFlatClass: 2D search in the plane (flat triangle meshes only) */
} vsdmo_surface_path;

/* the hex element metric.  The reference's shapeMetricHex takes the minimum
   over 9 sub-tets: the 8 corner tets and the principal-axes tet
   (dmo/DMOQualityHex.h:54-71, dmo/DMOQualityUtil.h:307-349).  CORNER8 takes
   the minimum over the 8 corner tets alone (the "min over 8 corner
   Jacobians" objective), with the same scaling and the same zero for a
   corner determinant or Frobenius term below the threshold. */
typedef enum vsdmo_hex_metric {
    VSDMO_HEX_SHAPE9 = 0,   /* reference default */
    VSDMO_HEX_CORNER8 = 1
} vsdmo_hex_metric;

typedef enum vsdmo_extent {
    VSDMO_EXTENT_AUTO = -1,  /* reference choice: max for tet/tri, min for hex */
    VSDMO_EXTENT_MAX = 0,
    VSDMO_EXTENT_MIN = 1
} vsdmo_extent;

typedef struct vsdmo_mesh vsdmo_mesh;
typedef struct vsdmo_opt vsdmo_opt;

typedef struct vsdmo_mesh_params {
    int32_t coloring;          /* vsdmo_coloring */
    float feature_angle_deg;   /* 20 (reference mesh/MeshTriGPU.cu:506) */
    int32_t coloring_polls;    /* greedy colouring: the polls of its lower neighbours' colours a
                                  vertex makes in the one-launch colouring before colouring
                                  rounds take over; 0 = 65536; the colours do not depend on it */
} vsdmo_mesh_params;

typedef struct vsdmo_params {
    int32_t nq;                /* candidates per axis, 2..8 (reference DMO_NQ = 8) */
    int32_t depth;             /* refinement levels (reference DMO_DEPTH = 3) */
    float shrink;              /* per-level window factor; 0 = 2/(nq-1) as the reference */
    int32_t extent_mode;       /* vsdmo_extent */
    int32_t fixed_boundary;    /* 1: surface and feature vertices are not moved;
                                  0: tet/hex surface vertices move on their local
                                  quadric fits and feature vertices along their
                                  feature curves, boundary vertices of flat
                                  triangle meshes along the boundary (needs surface
                                  one-rings of at most 40 vertices) */
    int32_t quality_guard;     /* 1: fail with VSDMO_QUALITY_DECREASED if min quality drops */
    float guard_tolerance;     /* relative slack of the guard (FP32 frame differences) */
    int32_t waves_per_vertex;  /* 1, 2 or 4 wavefronts per vertex search (dividing nq), or 0:
                                  one per vertex except the last 1024 vertices of a tet / hex
                                  colour class, which get 4 (nq 4 or 8); results do not depend
                                  on it */
    int32_t surface_path;      /* vsdmo_surface_path, triangle and quad meshes only; on the
                                  surface path fixed_boundary = 1 keeps the feature vertices */
    int32_t hex_metric;        /* vsdmo_hex_metric: the objective of the hex searches and the
                                  optimiser's hex qualities (min, mean, histogram, guard) */
    int32_t surface_ring;      /* support of the surface vertices' quadric fits: 1 = {v} and its
                                  one-ring (reference dmo/surfaceP/Surface.cu:11-28), 2 = {v},
                                  its one-ring and its 2-ring: the one-ring's CCW targets not yet
                                  taken, neighbour by neighbour in CCW order (at most 64 points,
                                  else VSDMO_NOT_SUPPORTED); feature sectors keep one-ring fits */
} vsdmo_params;

typedef struct vsdmo_counts {
    int32_t n_vertices, n_elements, n_surface, n_free, n_feature, n_inner;
    int32_t n_colors_inner, n_colors_free, n_colors_feature;
    int32_t n_faces, n_halfedges, max_ring, max_halfedges, flat;
} vsdmo_counts;

typedef struct vsdmo_stats {
    int64_t searches_inner, searches_surface, searches_feature;
    int32_t iterations;
    float min_quality;
    double kernel_ms;          /* summed device time of the search kernels (timing on) */
    int64_t kernel_launches;
    int64_t searches_learned;  /* inner searches that started from the learned proposal */
    int64_t fallbacks;         /* of those, the ones that fell back to the full search */
} vsdmo_stats;

/* defaults: greedy colouring, 20 degrees / nq 8, depth 3, reference shrink,
   auto extent, surface and boundary vertices optimised as in the reference
   (fixed_boundary = 0; pass 1 to keep them fixed), guard on, one wave per
   vertex, automatic surface path, the reference hex metric (9 sub-tets) and
   one-ring quadric fits */
void vsdmo_mesh_params_default(vsdmo_mesh_params* p);
void vsdmo_params_default(vsdmo_params* p);

/* Replaces DMOMesh{Tri,Quad,Tet,Hex}Factory::create (mesh/MeshFactory.h:17-41,
 * mesh/MeshFactory.cu:3-111) together with Mesh{Tri,Quad,Tet,Hex}GPU::init
 * (mesh/MeshTriGPU.cu:37-52, mesh/MeshQuadGPU.cu:37-45, mesh/MeshTetGPU.cu:35-42,
 * mesh/MeshHexGPU.cu:35-43).  xyz: 3*nV floats; cells: npe*nCells int32 (tri 3,
 * quad 4, tet 4, hex 8; hexes in the reference's internal order, top face
 * first).  A triangle or quad mesh is flat when every z is 0.  Host arrays are copied in and
 * the mesh structures are built on the GPU.  boundary_or_null must be NULL
 * (boundary detection from the elements, reference findBoundary2D=true). */
vsdmo_status vsdmo_mesh_create(vsdmo_elem type, const float* xyz, int32_t nV, const int32_t* cells,
                               int32_t nCells, const uint8_t* boundary_or_null,
                               const vsdmo_mesh_params* params, vsdmo_mesh** out);
/* Replaces the DMOMesh* destructors (mesh/DMOMeshBase.h:132-142) */
vsdmo_status vsdmo_mesh_destroy(vsdmo_mesh* mesh);

/* Replaces the public counters of DMOMeshBase / DMOMeshTet (mesh/DMOMeshBase.h:80-172) */
vsdmo_status vsdmo_mesh_counts(const vsdmo_mesh* mesh, vsdmo_counts* out);
/* positions: 3*nV floats; original_order=1 returns them indexed by input id
   (the reference writes files in the renumbered order, io/FileWriter.cu:64-102) */
vsdmo_status vsdmo_mesh_get_positions(const vsdmo_mesh* mesh, float* out, int32_t original_order);
vsdmo_status vsdmo_mesh_set_positions(vsdmo_mesh* mesh, const float* in, int32_t original_order);
/* Replaces laplaceSmoothing(DMOMeshTri&, n_iter) (dmo/Laplace.h:10, dmo/Laplace.cu:8-41):
   n_iter sweeps over the free-vertex colour classes of a triangle or quad
   mesh, p <- (p + sum of the one-ring) / (1 + n).  Optimisers created before
   keep their local surfaces. */
vsdmo_status vsdmo_mesh_laplace(vsdmo_mesh* mesh, int32_t n_iter);
/* new_of_old[nV]: renumbering applied by the build (mesh/SortUtil.cu:103-107) */
vsdmo_status vsdmo_mesh_get_permutation(const vsdmo_mesh* mesh, int32_t* new_of_old);
/* elements in the renumbered ids, npe*nE */
vsdmo_status vsdmo_mesh_get_cells(const vsdmo_mesh* mesh, int32_t* out);
/* colour class offsets (absolute ids); which: 0 inner, 1 free surface, 2 feature
   (reference colorOffsets*, mesh/DMOMeshBase.h:80-172); out has n_colors+1 */
vsdmo_status vsdmo_mesh_get_coloring(const vsdmo_mesh* mesh, int32_t which, int32_t* offsets);
/* per-vertex colour within its class family, nV */
vsdmo_status vsdmo_mesh_get_colors(const vsdmo_mesh* mesh, int32_t* colors);
/* Checks a partition of the vertex range [offsets[0], offsets[n_classes]) of
   the current numbering into the classes [offsets[c], offsets[c+1]): no
   element may hold two vertices of one class (the rule of k_checkColoring,
   mesh/MeshGPUCommon.cu:1470-1680, called after each colour sort, e.g.
   mesh/MeshTetGPU.cu:385).  VSDMO_TOPOLOGY names the first offending class.
   vsdmo_mesh_create runs the same check on every colouring it builds. */
vsdmo_status vsdmo_mesh_check_coloring(const vsdmo_mesh* mesh, const int32_t* offsets, int32_t n_classes);
/* boundary faces of a volume mesh (tri/quad, renumbered ids), fpe*n_faces */
vsdmo_status vsdmo_mesh_get_faces(const vsdmo_mesh* mesh, int32_t* out);
/* volume one-ring CSR (halffaces / halfhexes, mesh/MeshTetGPU.cu:113-142,
   mesh/MeshHexGPU.cu:116-139): row_ptr nV+1, values stride*row_ptr[nV] */
vsdmo_status vsdmo_mesh_get_onering(const vsdmo_mesh* mesh, int32_t* row_ptr, int32_t* values);
/* surface halfedge CSR over the surface vertices (mesh/MeshTypes.h:6-11,
   mesh/MeshTriGPU.cu:128-214): row_ptr n_surface+1; target/opposite/face/next n_halfedges */
vsdmo_status vsdmo_mesh_get_halfedges(const vsdmo_mesh* mesh, int32_t* row_ptr, int32_t* target,
                                      int32_t* opposite, int32_t* face, int32_t* next);
/* per surface vertex flags (boundary of an open surface, feature) */
vsdmo_status vsdmo_mesh_get_flags(const vsdmo_mesh* mesh, uint8_t* boundary1d, uint8_t* feature);

/* Replaces DMO::DMO{Tet,Hex,TriFlat,Tri,Quad}Class(mesh, criterion, gridScale, n_iter)
 * (dmo/DMOTet.h:24, dmo/DMOHex.h:18, dmo/DMOTriFlat.h:25, dmo/DMOTri.h:24,
 * dmo/DMOQuad.h:18) and their init() (dmo/DMOTet.cu:92-125, dmo/DMOTri.cu:93-125,
 * dmo/DMOQuad.cu:91-123).  params may be NULL (defaults).  A triangle / quad
 * mesh takes the path params->surface_path selects. */
vsdmo_status vsdmo_opt_create(vsdmo_mesh* mesh, int32_t criterion, float grid_scale, int32_t n_iter,
                              const vsdmo_params* params, vsdmo_opt** out);
/* Replaces DMOBaseClass::doIteration (dmo/DMOTet.cu:128-170, dmo/DMOHex.cu:131-182,
 * dmo/DMOTriFlat.cu:129-164), n times */
vsdmo_status vsdmo_opt_iterate(vsdmo_opt* opt, int32_t n);
/* Replaces DMOBaseClass::optimize (dmo/DMOBase.h:44-52): iterate until n_iter */
vsdmo_status vsdmo_opt_optimize(vsdmo_opt* opt);
/* Replaces DMOBaseClass::isDone */
int32_t vsdmo_opt_is_done(const vsdmo_opt* opt);
/* Replaces DMOBaseClass::getQuality (last minimum element quality, FP32) */
vsdmo_status vsdmo_opt_min_quality(const vsdmo_opt* opt, float* out);
/* mean element quality (FP64 sum of FP32 qualities, fixed reduction order) */
vsdmo_status vsdmo_opt_mean_quality(const vsdmo_opt* opt, double* out);
/* Replaces getQualityHistogram (dmo/DMOTet.cu:202-212): n_cols bins on [0,1) */
vsdmo_status vsdmo_opt_histogram(const vsdmo_opt* opt, int32_t n_cols, int32_t* out);
/* No counterpart in the reference.  vsdmo_opt_histogram without any device
   allocation, for callers that cannot allocate (code under HIP graph
   capture, tight serving loops): the bins live in a file-scope device array
   of VSDMO_HISTOGRAM_NOALLOC_MAX_COLS ints that each GPU holds until the
   process ends.  n_cols above the maximum is refused (VSDMO_INVALID_ARG).
   Calls on one GPU share the array, so calls that overlap on one GPU race.
   The result equals vsdmo_opt_histogram's. */
#define VSDMO_HISTOGRAM_NOALLOC_MAX_COLS 4096
vsdmo_status vsdmo_opt_histogram_noalloc(const vsdmo_opt* opt, int32_t n_cols, int32_t* out);
/* Replaces getElementQualities (dmo/DMOTet.cu:235-240), into host memory */
vsdmo_status vsdmo_opt_element_qualities(const vsdmo_opt* opt, float* host_out);
/* counters and timing; set_timing(1) records device time of every search launch */
vsdmo_status vsdmo_opt_stats(const vsdmo_opt* opt, vsdmo_stats* out);
vsdmo_status vsdmo_opt_set_timing(vsdmo_opt* opt, int32_t on);
/* Replaces the failure dump of updateCurrentQualities (dmo/DMOTet.cu:223-233,
   dmo/DMOHex.cu:235-245, dmo/DMOQuad.cu:218-228: writeTetgen / writeHex /
   writeOFF of "res/fail_debug", then throw): when the quality guard fails,
   the mesh (renumbered order, positions after the failing iteration) is
   written to path in its format (vsdmo_mesh_write) before the call returns
   VSDMO_QUALITY_DECREASED.  path NULL or "" turns it off (the default). */
vsdmo_status vsdmo_opt_set_failure_dump(vsdmo_opt* opt, const char* path);
vsdmo_status vsdmo_opt_destroy(vsdmo_opt* opt);

/* thread-local message of the last failing call on this thread */
const char* vsdmo_last_error(void);

/* Mesh files.  Replace the file overloads of DMOMesh{Tri,Quad,Tet,Hex}Factory::create
 * (mesh/MeshFactory.cu:3-111) with the reference readers io/ParserOFF.cu:15-160
 * (triangle / quad .off), io/ParserTetGen.cu:13-120 (<path>.node + <path>.ele,
 * 1-based numbering detected from the first point index) and
 * io/ParserHex.cu:115-240 (MEDIT .mesh hexahedra, file bottom face = internal
 * top face), and the writers io/FileWriter.cu:9-147 (9 significant digits, so a
 * written file reads back bit-identically).  vsdmo_mesh_file_read with NULL
 * buffers returns the sizes only.  The reference returns nullptr on a parse
 * failure; these calls return VSDMO_INVALID_ARG. */
vsdmo_status vsdmo_mesh_file_read(vsdmo_elem type, const char* path, int32_t* nV, int32_t* nCells, float* xyz,
                                  int32_t* cells);
vsdmo_status vsdmo_mesh_file_write(vsdmo_elem type, const char* path, const float* xyz, int32_t nV,
                                   const int32_t* cells, int32_t nCells);
vsdmo_status vsdmo_mesh_create_from_file(vsdmo_elem type, const char* path, const vsdmo_mesh_params* params,
                                         vsdmo_mesh** out);
/* current positions and elements, in the input numbering (original_order = 1)
 * or the optimiser numbering (0, as the reference's writers) */
vsdmo_status vsdmo_mesh_write(vsdmo_mesh* m, const char* path, int32_t original_order);

/* Replaces DMOBaseClass::getEstimateLocalSurfacePoints (mode 0) and
 * getLocalSurfacePoints (mode 1, feature_sid) (dmo/DMOBase.h:38-39;
 * kernels dmo/DMOCommon.h:197-276, dmo/DMOCommon.cu:73-150): nu x nv points
 * (out: 3*nu*nv floats, host) sampled around surface vertex vid (optimiser
 * numbering) on its blended projection or on its nearest neighbour's local
 * fit.  Needs fixed_boundary = 0. */
vsdmo_status vsdmo_opt_surface_points(vsdmo_opt* o, int32_t vid, int32_t nu, int32_t nv, int32_t mode,
                                      int32_t feature_sid, float* out);

/* ------------------------------------------------------------------------
 * Learned proposal.  No counterpart in the reference.  A 96-128-128-3 GELU
 * MLP (BF16 weights and activations on MFMA, FP32 accumulation) maps a
 * vertex's one-ring -- the distinct one-ring vertices in order of first
 * appearance in its ring elements, relative to the vertex, divided by their
 * mean distance L, zero padded to 32 x 3 -- to the displacement / L that
 * the full DMO search gives it.  The optimiser then searches a 4x4x4 grid on
 * a window `window_frac` times the plain first-level window around the
 * proposal for `depth` levels and falls back to the plain full search when
 * the window's best candidate is lower than the vertex's current quality
 * (a tie at a positive quality -- a ring minimum held by sub-tets the vertex
 * is not part of -- keeps the vertex, as does a hex vertex whose
 * vertex-free sub-tets already score 0).
 * ------------------------------------------------------------------------ */
typedef struct vsdmo_mlp vsdmo_mlp;
typedef struct vsdmo_dataset vsdmo_dataset;

int32_t vsdmo_mlp_num_params(void);   /* 29,315 FP32 parameters: W1[128x96] b1 W2[128x128] b2 W3[3x128] b3 */
vsdmo_status vsdmo_mlp_create(uint64_t seed, vsdmo_mlp** out);
/* refuses (VSDMO_INVALID_ARG) while an optimiser or batch has the network as
 * its proposal: set the proposal to mode 0 or destroy those first */
vsdmo_status vsdmo_mlp_destroy(vsdmo_mlp* m);
vsdmo_status vsdmo_mlp_get_params(vsdmo_mlp* m, float* params);
vsdmo_status vsdmo_mlp_set_params(vsdmo_mlp* m, const float* params);   /* also resets Adam */
/* host rows: X n x 96 (rounded to BF16 on upload), Y n x 3 */
vsdmo_status vsdmo_mlp_forward(vsdmo_mlp* m, const float* X, int32_t n, float* Y);
/* mean squared error over n rows and its gradient w.r.t. the packed parameters */
vsdmo_status vsdmo_mlp_gradients(vsdmo_mlp* m, const float* X, const float* T, int32_t n, double* loss, float* grad);
/* Adam (beta 0.9 / 0.999, eps 1e-8), seeded shuffles per epoch, the ragged
 * tail of each epoch dropped; epoch_loss[epochs] = mean batch loss */
vsdmo_status vsdmo_mlp_train(vsdmo_mlp* m, vsdmo_dataset* ds, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                             double* epoch_loss);

vsdmo_status vsdmo_dataset_create(int64_t capacity, vsdmo_dataset** out);
vsdmo_status vsdmo_dataset_size(vsdmo_dataset* d, int64_t* n);
vsdmo_status vsdmo_dataset_get(vsdmo_dataset* d, int64_t start, int32_t n, float* X, float* T);
vsdmo_status vsdmo_dataset_destroy(vsdmo_dataset* d);

/* run n plain iterations recording one (one-ring features, displacement / L)
 * pair per inner vertex search until the dataset is full */
vsdmo_status vsdmo_opt_record(vsdmo_opt* o, vsdmo_dataset* ds, int32_t n);
/* mode 0: plain search; 1: network proposal; 2: the proposal is the vertex
 * itself (the same windowed search and fallback without the network); 3:
 * the network's proposal and, where its window finds nothing better, a
 * window a quarter of its size around the vertex itself before the full
 * search.
 * One-rings up to 512 tets / 85 hexes (VSDMO_NOT_SUPPORTED above).  With
 * one-rings up to 116 tets / 19 hexes the window kernel runs the network
 * itself; environment variable VSDMO_LEARN_SEPARATE_INFERENCE=1 (read here)
 * keeps it in a launch of its own (same results, for diagnosis). */
vsdmo_status vsdmo_opt_set_proposal(vsdmo_opt* o, vsdmo_mlp* mlp, int32_t mode, int32_t depth, float window_frac);
/* when a vertex's proposal window finds nothing better than its start:
 * IMMEDIATE -- its full search runs in the same class pass (a second launch);
 * DEFERRED -- it keeps its position and its full search runs in the class's
 * next pass, beside that pass's windows (one launch per class; nq 4 or 8 and
 * one-rings of at most 128 tets / 21 hexes, else VSDMO_NOT_SUPPORTED).  Setting
 * it clears pending deferred searches. */
enum { VSDMO_FALLBACK_IMMEDIATE = 0, VSDMO_FALLBACK_DEFERRED = 1 };
vsdmo_status vsdmo_opt_set_fallback(vsdmo_opt* o, int32_t policy);

/* ------------------------------------------------------------------------
 * Batched DMO over many independent meshes.  No counterpart in the
 * reference, which optimises one mesh per DMO*Class object
 * (dmo/DMOBase.h:23-88); each mesh's searches are the reference's and give
 * the same positions as one vsdmo_opt per mesh with fixed_boundary = 1.
 * Colour round r sweeps colour r of every mesh (inner colours of tet / hex
 * meshes, free colours of flat triangle meshes) in one persistent launch; an
 * iteration (rounds + per-mesh quality) is captured once as a hipGraph and
 * replayed.  The meshes must outlive the batch (vsdmo_mesh_destroy refuses
 * while a batch holds them); the batch works on its own copy of the
 * positions until vsdmo_batch_writeback.
 * ------------------------------------------------------------------------ */
typedef struct vsdmo_batch vsdmo_batch;

typedef struct vsdmo_batch_params {
    int32_t nq;                /* candidates per axis: 2, 4, 6, 8 */
    int32_t depth;             /* refinement levels */
    int32_t waves_per_vertex;  /* 1, 2, 4 (divides nq); does not change results */
    float grid_scale;          /* first-level window scale (reference 0.5) */
    int32_t use_graph;         /* 1: replay a captured hipGraph of one iteration */
    uint32_t shuffle_seed;     /* 0: work items in mesh order; else a seeded permutation inside each round */
    int32_t quality_guard;     /* 1: fail with VSDMO_QUALITY_DECREASED when a mesh's min quality drops */
    float guard_tolerance;
    int32_t persistent;        /* 1: persistent blocks take a round's items from a queue; 0: one block per
                                  item; does not change results */
    int32_t hex_metric;        /* vsdmo_hex_metric of the hex meshes' searches and qualities */
} vsdmo_batch_params;

typedef struct vsdmo_batch_stats {
    int64_t searches;              /* vertex searches executed */
    int32_t iterations, rounds, meshes;
    int64_t items_per_iteration;
    double gpu_ms;                 /* device time of the iterations (events around them) */
    int32_t max_grid;              /* blocks of the largest persistent launch */
    int32_t total_vertices;
} vsdmo_batch_stats;

void vsdmo_batch_params_default(vsdmo_batch_params* p);
vsdmo_status vsdmo_batch_create(vsdmo_mesh* const* meshes, int32_t n_meshes, const vsdmo_batch_params* params,
                                vsdmo_batch** out);
vsdmo_status vsdmo_batch_iterate(vsdmo_batch* b, int32_t n);
/* per-mesh minimum quality after the last iteration (n_meshes floats) */
vsdmo_status vsdmo_batch_min_quality(vsdmo_batch* b, float* per_mesh);
vsdmo_status vsdmo_batch_mean_quality(vsdmo_batch* b, double* per_mesh);
/* positions of one mesh in the optimiser numbering (3*nV floats) */
vsdmo_status vsdmo_batch_get_positions(vsdmo_batch* b, int32_t mesh, float* xyz);
/* copy the batch positions into the meshes */
vsdmo_status vsdmo_batch_writeback(vsdmo_batch* b);
/* back to the positions at vsdmo_batch_create */
vsdmo_status vsdmo_batch_restore(vsdmo_batch* b);
vsdmo_status vsdmo_batch_get_params(vsdmo_batch* b, vsdmo_batch_params* p);
vsdmo_status vsdmo_batch_set_params(vsdmo_batch* b, const vsdmo_batch_params* p);
/* Autotuner over (nq, depth, waves_per_vertex, persistent): every trial starts
 * from the current positions and runs `iters` iterations; the schedule
 * (waves_per_vertex, persistent) is chosen by time at (8, 3), then (nq, depth) candidates are accepted when every
 * mesh's final min quality is >= the (8, 3) run's - tol; the fastest
 * accepted configuration is set and returned, the positions are left as
 * they were, and each trial is written to log_path as one JSON line. */
vsdmo_status vsdmo_batch_autotune(vsdmo_batch* b, int32_t iters, float tol, const char* log_path,
                                  vsdmo_batch_params* chosen);
vsdmo_status vsdmo_batch_get_stats(vsdmo_batch* b, vsdmo_batch_stats* out);
/* Learned proposal for the batch's tet or hex vertices, elem = VSDMO_TET or
 * VSDMO_HEX, each type with its own network and window (no counterpart in
 * the reference; the rule of vsdmo_opt_set_proposal with the policy of
 * vsdmo_batch_set_fallback, applied per colour round to the round's items
 * of that type; the other items keep the plain search).  Each such mesh's positions
 * equal a vsdmo_opt of the mesh with fixed_boundary = 1 and the same
 * proposal.  Modes 1 and 3 run the network inside the window kernel and need
 * every ring of the type to fit its LDS tile (VSDMO_NOT_SUPPORTED
 * otherwise); mode 0 switches the type's proposal off.  The network is read
 * at every iteration: vsdmo_mlp_destroy refuses while the batch holds it.
 * The autotuner refuses while a proposal is set. */
vsdmo_status vsdmo_batch_set_proposal(vsdmo_batch* b, int32_t elem, vsdmo_mlp* mlp, int32_t mode, int32_t depth,
                                      float window_frac);
/* Fallback policy of the batch's learned proposals, as vsdmo_opt_set_fallback
 * (VSDMO_FALLBACK_IMMEDIATE, the default, or VSDMO_FALLBACK_DEFERRED: a
 * vertex whose window finds nothing better takes the full search in its
 * round of the next iteration; the iteration number and the double-buffered
 * lists live on the device, so the captured iteration replays unchanged).
 * Clears pending deferrals, as do vsdmo_batch_restore and
 * vsdmo_batch_set_proposal.  Each mesh then ends where a vsdmo_opt with the
 * same policy ends.  Deferred needs nq 4 or 8. */
vsdmo_status vsdmo_batch_set_fallback(vsdmo_batch* b, int32_t policy);
/* learned visits so far and how many of them took the full search */
vsdmo_status vsdmo_batch_learned_stats(vsdmo_batch* b, int64_t* visits, int64_t* fallbacks);
vsdmo_status vsdmo_batch_destroy(vsdmo_batch* b);

#ifdef __cplusplus
}
#endif
#endif
