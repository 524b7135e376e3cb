// api.hip -- C ABI of libvsdmo (declarations and reference citations in
// include/vsdmo.h).
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>
#include <new>
#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

#include "dmo_kernels.h"
#include "mlp_internal.h"
#include "vsdmo_internal.h"

namespace vsdmo {
void build_volume(vsdmo_mesh& m, const vsdmo_mesh_params& prm);
void build_surface_mesh(vsdmo_mesh& m, const vsdmo_mesh_params& prm);
void check_classes(const vsdmo_mesh& m, const std::vector<int>& off);

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

template <class F>
static vsdmo_status guarded(F&& f)
{
    try {
        g_err.clear();
        return f();
    } catch (const HipError& e) {
        set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
        return VSDMO_HIP_ERROR;
    } catch (const StatusError& e) {
        set_error(e.msg);
        return e.status;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return VSDMO_HIP_ERROR;
    } catch (...) {
        set_error("unknown internal error");
        return VSDMO_HIP_ERROR;
    }
}

static vsdmo_status fail(vsdmo_status s, const std::string& msg)
{
    set_error(msg);
    return s;
}

// largest one-ring each search kernel is built for (LDS per 64-thread block)
static int max_ring_supported(int elem)
{
    return elem == VSDMO_HEX ? 256 : 1024;   // <= 48 KB of LDS per block
}

template <class T>
static vsdmo_status copy_dev(const DevBuf<T>& b, size_t n, T* out, hipStream_t s)
{
    return guarded([&]() -> vsdmo_status {
        if (!out) return fail(VSDMO_INVALID_ARG, "NULL output");
        b.download(out, n, s);
        VSDMO_HIP(hipStreamSynchronize(s));
        return VSDMO_OK;
    });
}

static void destroy_mesh(vsdmo_mesh* m)
{
    hipStream_t s = m->stream;
    delete m;
    if (s) (void)hipStreamDestroy(s);
}

}  // namespace vsdmo

using namespace vsdmo;

// 16 bytes of pinned host memory for the per-iteration quality readback
struct PinnedQ {
    double* p = nullptr;
    PinnedQ() = default;
    PinnedQ(const PinnedQ&) = delete;
    PinnedQ& operator=(const PinnedQ&) = delete;
    ~PinnedQ() { if (p) (void)hipHostFree(p); }
    void alloc()
    {
        if (!p) VSDMO_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), 2 * sizeof(double), hipHostMallocDefault));
    }
};

struct vsdmo_opt {
    vsdmo_mesh* m = nullptr;
    int criterion = 0;
    float grid_scale = 0.5f;
    int n_iter = 30;
    vsdmo_params p{};
    float shrink = 0.0f, aff = 0.0f;
    int extent_min = 0;
    int curr_it = 0;
    float last_min = 0.0f;
    double last_sum = 0.0;
    int64_t s_inner = 0, s_surf = 0, s_feat = 0;
    bool timing = false;
    double kernel_ms = 0.0;
    int64_t launches = 0;
    DevBuf<float> q, pmin;
    DevBuf<double> psum;
    DevBuf<double> qout;   // [0]: minimum (float in bytes 0-3, surface error bits in 4-7), [1]: FP64 sum
    PinnedQ hq;            // their host copy
    std::vector<hipEvent_t> ev;   // pool of events, [0, ev_used) in use by the timed launches
    size_t ev_used = 0;
    std::vector<std::pair<size_t, size_t>> ev_pairs;   // (begin, end) event of every timed launch
    // surface-vertex state (fixed_boundary == 0)
    bool surface = false;
    int nSectors = 0;
    DevBuf<float> pinit, fnorm, vnorm, ls, ls1, fs, snapx, snapy, snapz;
    // surface meshes (tri / quad) on the surface path (DMOTriClass / DMOQuadClass)
    bool surf_mesh = false;
    int qelem = 0;                  // element kind of the quality kernels
    DevBuf<int> nn, frow, table;
    DevBuf<uint8_t> opt1d;
    std::vector<uint8_t> opt1d_host;
    // learned proposal (inner tet / hex searches)
    vsdmo_mlp* mlp = nullptr;
    int learn_mode = 0, learn_depth = 4;
    float learn_frac = 0.25f;
    bool learn_separate = false;     // network in its own launch even where the window kernel could run it
    int learn_fallback = VSDMO_FALLBACK_IMMEDIATE;
    // deferred fallback: per-class lists of two iterations (class c at its
    // offset in the inner range, buffer it & 1 written in iteration it), their
    // lengths, and the iteration that deferred each vertex
    DevBuf<int> dList, dCnt, dEpoch;
    bool defer_pending = false;      // the lists hold the last iteration's deferrals
    DevBuf<float> lL, lY;
    DevBuf<int> lFallback;           // running count of fallbacks to the full search
    DevBuf<int> lList, lCnt;         // fallback list of a class; per-class list lengths
    int64_t s_learned = 0;
    vsdmo_dataset* rec = nullptr;   // training pairs are recorded while set
    std::string fail_dump;          // mesh file written when the quality guard fails
};

extern "C" {

void vsdmo_mesh_params_default(vsdmo_mesh_params* p)
{
    if (!p) return;
    p->coloring = VSDMO_COLOR_GREEDY;
    p->feature_angle_deg = 20.0f;
    p->coloring_polls = 0;
}

void vsdmo_params_default(vsdmo_params* p)
{
    if (!p) return;
    p->nq = 8;
    p->depth = 3;
    p->shrink = 0.0f;
    p->extent_mode = VSDMO_EXTENT_AUTO;
    p->fixed_boundary = 0;
    p->quality_guard = 1;
    p->guard_tolerance = 1e-5f;
    p->waves_per_vertex = 1;
    p->surface_path = VSDMO_SURFACE_PATH_AUTO;
    p->hex_metric = VSDMO_HEX_SHAPE9;
    p->surface_ring = 1;
}

const char* vsdmo_last_error(void) { return g_err.c_str(); }

vsdmo_status vsdmo_mesh_create(vsdmo_elem type, const float* xyz, int32_t nV, const int32_t* cells, int32_t nCells,
                               const uint8_t* boundary_or_null, const vsdmo_mesh_params* params, vsdmo_mesh** out)
{
    return guarded([&]() -> vsdmo_status {
        if (!out) return fail(VSDMO_INVALID_ARG, "out is NULL");
        *out = nullptr;
        if (!xyz || !cells || nV <= 0 || nCells <= 0) return fail(VSDMO_INVALID_ARG, "empty or NULL mesh arrays");
        if (boundary_or_null) return fail(VSDMO_NOT_SUPPORTED, "explicit boundary flags are not supported; pass NULL");
        if (type != VSDMO_TET && type != VSDMO_HEX && type != VSDMO_TRI && type != VSDMO_QUAD)
            return fail(VSDMO_INVALID_ARG, "unknown element type");
        int npe = (type == VSDMO_TRI) ? 3 : (type == VSDMO_TET || type == VSDMO_QUAD) ? 4 : 8;
        const bool surfaceMesh = (type == VSDMO_TRI || type == VSDMO_QUAD);
        long long nci = (long long)nCells * npe;
        if (nci > (long long)0x7fffffff / 2) return fail(VSDMO_INVALID_ARG, "mesh too large for int32 indexing");
        for (long long i = 0; i < nci; ++i)
            if (cells[i] < 0 || cells[i] >= nV) return fail(VSDMO_INVALID_ARG, "element vertex index out of range");
        for (long long e = 0; e < nCells; ++e)
            for (int a = 0; a < npe; ++a)
                for (int b = a + 1; b < npe; ++b)
                    if (cells[e * npe + a] == cells[e * npe + b])
                        return fail(VSDMO_INVALID_ARG, "element with repeated vertex " + std::to_string(e));
        vsdmo_mesh_params prm;
        vsdmo_mesh_params_default(&prm);
        if (params) prm = *params;
        if (prm.coloring != VSDMO_COLOR_GREEDY && prm.coloring != VSDMO_COLOR_JPL)
            return fail(VSDMO_INVALID_ARG, "unknown colouring mode");
        if (prm.coloring_polls < 0) return fail(VSDMO_INVALID_ARG, "coloring_polls must be >= 0");

        auto* m = new vsdmo_mesh();
        try {
            m->elem = type; m->npe = npe; m->nV = nV; m->nE = nCells;
            VSDMO_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
            hipStream_t s = m->stream;
            std::vector<float> hx(nV), hy(nV), hz(nV);
            int flat = 1;
            for (int i = 0; i < nV; ++i) {
                hx[i] = xyz[3 * i]; hy[i] = xyz[3 * i + 1]; hz[i] = xyz[3 * i + 2];
                if (hz[i] != 0.0f) flat = 0;
            }
            m->flat = surfaceMesh ? flat : 0;
            m->x.alloc(nV); m->y.alloc(nV); m->z.alloc(nV);
            m->x.upload(hx.data(), nV, s); m->y.upload(hy.data(), nV, s); m->z.upload(hz.data(), nV, s);
            m->cells.alloc(nci);
            m->cells.upload(cells, nci, s);
            m->colorOf.alloc(nV);
            VSDMO_HIP(hipMemsetAsync(m->colorOf.p, 0, sizeof(int) * nV, s));
            m->newOfOld.alloc(nV);
            {
                std::vector<int> id(nV);
                for (int i = 0; i < nV; ++i) id[i] = i;
                m->newOfOld.upload(id.data(), nV, s);
                VSDMO_HIP(hipStreamSynchronize(s));
            }
            if (surfaceMesh) build_surface_mesh(*m, prm);
            else build_volume(*m, prm);
            m->new_of_old = m->newOfOld.to_host(s);
            if (m->maxRing > max_ring_supported(type) || m->maxHE > 1024) {
                int mr = m->maxRing;
                destroy_mesh(m);
                return fail(VSDMO_LDS_OVERFLOW, "one-ring of " + std::to_string(mr) + " elements exceeds the kernel limit");
            }
        } catch (...) {
            destroy_mesh(m);
            throw;
        }
        *out = m;
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_mesh_destroy(vsdmo_mesh* m)
{
    return guarded([&]() -> vsdmo_status {
        if (!m) return VSDMO_OK;
        if (m->n_opts > 0) return fail(VSDMO_INVALID_ARG, "mesh still referenced by an optimiser");
        destroy_mesh(m);
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_mesh_counts(const vsdmo_mesh* m, vsdmo_counts* o)
{
    if (!m || !o) return fail(VSDMO_INVALID_ARG, "NULL argument");
    o->n_vertices = m->nV; o->n_elements = m->nE;
    o->n_surface = m->nSurf; o->n_free = m->nFree; o->n_feature = m->nFeature; o->n_inner = m->nInner;
    o->n_colors_inner = (int)m->colOffInner.size() - 1;
    o->n_colors_free = (int)m->colOffFree.size() - 1;
    o->n_colors_feature = (int)m->colOffFeature.size() - 1;
    o->n_faces = m->nFaces; o->n_halfedges = m->nHE;
    o->max_ring = m->maxRing; o->max_halfedges = m->maxHE; o->flat = m->flat;
    return VSDMO_OK;
}

vsdmo_status vsdmo_mesh_get_positions(const vsdmo_mesh* m, float* out, int32_t original_order)
{
    return guarded([&]() -> vsdmo_status {
        if (!m || !out) return fail(VSDMO_INVALID_ARG, "NULL argument");
        std::vector<float> hx = m->x.to_host(m->stream), hy = m->y.to_host(m->stream), hz = m->z.to_host(m->stream);
        for (int o = 0; o < m->nV; ++o) {
            int v = original_order ? m->new_of_old[o] : o;
            out[3 * o] = hx[v]; out[3 * o + 1] = hy[v]; out[3 * o + 2] = hz[v];
        }
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_mesh_set_positions(vsdmo_mesh* m, const float* in, int32_t original_order)
{
    return guarded([&]() -> vsdmo_status {
        if (!m || !in) return fail(VSDMO_INVALID_ARG, "NULL argument");
        std::vector<float> hx(m->nV), hy(m->nV), hz(m->nV);
        for (int o = 0; o < m->nV; ++o) {
            int v = original_order ? m->new_of_old[o] : o;
            hx[v] = in[3 * o]; hy[v] = in[3 * o + 1]; hz[v] = in[3 * o + 2];
        }
        m->x.upload(hx.data(), m->nV, m->stream);
        m->y.upload(hy.data(), m->nV, m->stream);
        m->z.upload(hz.data(), m->nV, m->stream);
        VSDMO_HIP(hipStreamSynchronize(m->stream));
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_mesh_laplace(vsdmo_mesh* m, int32_t n_iter)
{
    return guarded([&]() -> vsdmo_status {
        if (!m) return fail(VSDMO_INVALID_ARG, "NULL mesh");
        if (n_iter < 0) return fail(VSDMO_INVALID_ARG, "negative iteration count");
        if (m->elem != VSDMO_TRI && m->elem != VSDMO_QUAD)
            return fail(VSDMO_NOT_SUPPORTED, "Laplace smoothing is for triangle and quad surface meshes");
        for (int it = 0; it < n_iter; ++it)
            for (size_t c = 0; c + 1 < m->colOffFree.size(); ++c)
                laplace_launch(m->x.p, m->y.p, m->z.p, m->heRow.p, m->heTarget.p, m->colOffFree[c], m->colOffFree[c + 1],
                               m->stream);
        VSDMO_HIP(hipGetLastError());
        VSDMO_HIP(hipStreamSynchronize(m->stream));
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_mesh_get_permutation(const vsdmo_mesh* m, int32_t* new_of_old)
{
    if (!m || !new_of_old) return fail(VSDMO_INVALID_ARG, "NULL argument");
    std::memcpy(new_of_old, m->new_of_old.data(), sizeof(int) * m->nV);
    return VSDMO_OK;
}

vsdmo_status vsdmo_mesh_get_cells(const vsdmo_mesh* m, int32_t* out)
{
    if (!m) return fail(VSDMO_INVALID_ARG, "NULL mesh");
    return copy_dev(m->cells, (size_t)m->nE * m->npe, out, m->stream);
}

vsdmo_status vsdmo_mesh_get_coloring(const vsdmo_mesh* m, int32_t which, int32_t* offsets)
{
    if (!m || !offsets) return fail(VSDMO_INVALID_ARG, "NULL argument");
    const std::vector<int>* v = which == 0 ? &m->colOffInner : which == 1 ? &m->colOffFree : which == 2 ? &m->colOffFeature : nullptr;
    if (!v) return fail(VSDMO_INVALID_ARG, "which must be 0 (inner), 1 (free) or 2 (feature)");
    std::memcpy(offsets, v->data(), sizeof(int) * v->size());
    return VSDMO_OK;
}

vsdmo_status vsdmo_mesh_check_coloring(const vsdmo_mesh* m, const int32_t* offsets, int32_t n_classes)
{
    return guarded([&]() -> vsdmo_status {
        if (!m || !offsets || n_classes < 1) return fail(VSDMO_INVALID_ARG, "bad argument");
        std::vector<int> off(offsets, offsets + n_classes + 1);
        if (off[0] < 0 || off[n_classes] > m->nV) return fail(VSDMO_INVALID_ARG, "offsets outside [0, nV]");
        for (int c = 0; c < n_classes; ++c)
            if (off[c + 1] < off[c]) return fail(VSDMO_INVALID_ARG, "offsets must not decrease");
        check_classes(*m, off);
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_mesh_get_colors(const vsdmo_mesh* m, int32_t* colors)
{
    if (!m) return fail(VSDMO_INVALID_ARG, "NULL mesh");
    return copy_dev(m->colorOf, (size_t)m->nV, colors, m->stream);
}

vsdmo_status vsdmo_mesh_get_faces(const vsdmo_mesh* m, int32_t* out)
{
    if (!m) return fail(VSDMO_INVALID_ARG, "NULL mesh");
    return copy_dev(m->faces, (size_t)m->nFaces * m->fpe, out, m->stream);
}

vsdmo_status vsdmo_mesh_get_onering(const vsdmo_mesh* m, int32_t* row_ptr, int32_t* values)
{
    return guarded([&]() -> vsdmo_status {
        if (!m || !row_ptr) return fail(VSDMO_INVALID_ARG, "NULL argument");
        m->ringRow.download(row_ptr, (size_t)m->nV + 1, m->stream);
        VSDMO_HIP(hipStreamSynchronize(m->stream));
        int st = m->elem == VSDMO_HEX ? 7 : 3;
        size_t nnz = (size_t)row_ptr[m->nV] * st;
        if (values && nnz) {
            m->ringVal.download(values, nnz, m->stream);
            VSDMO_HIP(hipStreamSynchronize(m->stream));
        }
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_mesh_get_halfedges(const vsdmo_mesh* m, int32_t* row_ptr, int32_t* target, int32_t* opposite,
                                      int32_t* face, int32_t* next)
{
    return guarded([&]() -> vsdmo_status {
        if (!m || !row_ptr) return fail(VSDMO_INVALID_ARG, "NULL argument");
        m->heRow.download(row_ptr, (size_t)m->nSurf + 1, m->stream);
        size_t n = (size_t)m->nHE;
        if (n) {
            if (target) m->heTarget.download(target, n, m->stream);
            if (opposite) m->heOpp.download(opposite, n, m->stream);
            if (face) m->heFace.download(face, n, m->stream);
            if (next) m->heNext.download(next, n, m->stream);
        }
        VSDMO_HIP(hipStreamSynchronize(m->stream));
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_mesh_get_flags(const vsdmo_mesh* m, uint8_t* b1d, uint8_t* feat)
{
    return guarded([&]() -> vsdmo_status {
        if (!m) return fail(VSDMO_INVALID_ARG, "NULL mesh");
        if (m->nSurf > 0) {
            if (b1d) m->isB1D.download(b1d, m->nSurf, m->stream);
            if (feat) m->isFeature.download(feat, m->nSurf, m->stream);
        }
        VSDMO_HIP(hipStreamSynchronize(m->stream));
        return VSDMO_OK;
    });
}

// ---------------------------------------------------------------- optimiser
static void opt_quality(vsdmo_opt* o)
{
    vsdmo_mesh* m = o->m;
    QualityArgs qa{ m->x.p, m->y.p, m->z.p, m->cells.p, m->nE, m->npe, o->qelem, o->q.p,
                    o->p.hex_metric == VSDMO_HEX_CORNER8 };
    launch_quality(qa, o->pmin.p, o->psum.p, reinterpret_cast<float*>(o->qout.p), o->qout.p + 1, m->stream);
    VSDMO_HIP(hipGetLastError());
    // one 16-byte copy into pinned memory (no staging, one synchronisation)
    VSDMO_HIP(hipMemcpyAsync(o->hq.p, o->qout.p, 2 * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    VSDMO_HIP(hipStreamSynchronize(m->stream));
    float mn = 0.0f;
    int err = 0;
    std::memcpy(&mn, o->hq.p, sizeof(float));
    std::memcpy(&err, reinterpret_cast<const char*>(o->hq.p) + sizeof(float), sizeof(int));
    o->last_min = mn;
    o->last_sum = o->hq.p[1];
    if (err & kSurfErrUnsectored)
        throw StatusError{ VSDMO_TOPOLOGY, "a feature vertex has surface neighbours after its last feature "
                                           "neighbour (open fan): they belong to no feature sector" };
    if (err & kSurfErrNoSector)
        throw StatusError{ VSDMO_TOPOLOGY, "projection onto a feature neighbour's sectors found no sector for "
                                           "its halfedge nearest to the vertex" };
    if (err & kSurfErrFitPoints)
        throw StatusError{ VSDMO_NOT_SUPPORTED, "surface_ring 2: a surface vertex, its one-ring and its 2-ring "
                                                "hold more than 64 points" };
}

// empty deferred lists and no deferral epochs (allocated on first use)
static void reset_deferred(vsdmo_opt* o)
{
    vsdmo_mesh* m = o->m;
    if (o->learn_fallback != VSDMO_FALLBACK_DEFERRED || m->colOffInner.size() < 2) return;
    const int nCls = (int)m->colOffInner.size() - 1;
    const int nIn = m->colOffInner.back() - m->colOffInner.front();
    if (!o->dList.p) {
        o->dList.alloc((size_t)2 * std::max(nIn, 1));
        o->dCnt.alloc((size_t)2 * nCls);
        o->dEpoch.alloc(std::max(m->nV, 1));
    }
    launch_fill_int(o->dCnt.p, 2 * nCls, 0, m->stream);
    launch_fill_int(o->dEpoch.p, std::max(m->nV, 1), INT_MIN, m->stream);
}

static SurfArgs surf_args(vsdmo_opt* o)
{
    vsdmo_mesh* m = o->m;
    SurfArgs a{};
    a.x = m->x.p; a.y = m->y.p; a.z = m->z.p;
    a.xw = m->x.p; a.yw = m->y.p; a.zw = m->z.p;
    a.sx = o->snapx.p; a.sy = o->snapy.p; a.sz = o->snapz.p;
    a.ringRow = m->ringRow.p; a.ringVal = m->ringVal.p;
    a.heRow = m->heRow.p; a.heTarget = m->heTarget.p; a.heFace = m->heFace.p; a.heNext = m->heNext.p;
    a.faces = m->faces.p;
    a.isFeature = m->isFeature.p; a.isB1D = m->isB1D.p;
    a.nS = m->nSurf; a.nFree = m->nFree; a.nFeature = m->nFeature; a.nFaces = m->nFaces; a.fpe = m->fpe;
    a.flat = m->flat;
    a.cos160 = (float)std::cos(160.0 * 3.14159265358979323846 / 180.0);
    a.pinit = o->pinit.p; a.fnorm = o->fnorm.p; a.vnorm = o->vnorm.p; a.ls = o->ls.p; a.ls1 = o->ls1.p; a.fs = o->fs.p;
    a.nn = o->nn.p; a.frow = o->frow.p; a.table = o->table.p; a.opt1d = o->opt1d.p;
    a.gridScale = o->grid_scale; a.shrink = o->shrink; a.aff = o->aff; a.depth = o->p.depth;
    a.hexCorner8 = o->p.hex_metric == VSDMO_HEX_CORNER8;
    a.ring2 = o->p.surface_ring == 2;
    a.err = reinterpret_cast<int*>(o->qout.p) + 1;
    return a;
}

// local surfaces, feature sectors, 1D curves, NN table and 1D eligibility
// from the current positions (reference DMOHexClass::init, dmo/DMOHex.cu:95-129)
static void surface_init(vsdmo_opt* o)
{
    vsdmo_mesh* m = o->m;
    hipStream_t s = m->stream;
    const int nS = m->nSurf, nF = m->nFeature;
    o->pinit.alloc((size_t)3 * nS); o->fnorm.alloc((size_t)3 * std::max(m->nFaces, 1)); o->vnorm.alloc((size_t)3 * nS);
    o->ls.alloc((size_t)16 * nS); o->ls1.alloc((size_t)10 * nS);
    o->nn.alloc(nS); o->opt1d.alloc(nS);
    o->frow.alloc((size_t)nF + 1); o->table.alloc((size_t)20 * std::max(nF, 1));
    o->snapx.alloc(nS); o->snapy.alloc(nS); o->snapz.alloc(nS);
    VSDMO_HIP(hipMemsetAsync(o->ls1.p, 0, sizeof(float) * 10 * nS, s));
    SurfArgs a = surf_args(o);
    surf_init_launch(a, s);
    std::vector<int> frow(nF + 1, 0);
    if (nF > 0) {
        DevBuf<int> cnt(nF);
        surf_sector_count_launch(a, cnt.p, s);
        std::vector<int> hc = cnt.to_host(s);
        for (int i = 0; i < nF; ++i) frow[i + 1] = frow[i] + hc[i];
    }
    o->frow.upload(frow.data(), nF + 1, s);
    o->nSectors = frow[nF];
    o->fs.alloc((size_t)16 * std::max(o->nSectors, 1));
    // sectors with fewer than two points keep a zero record, as in the oracle
    VSDMO_HIP(hipMemsetAsync(o->fs.p, 0, sizeof(float) * 16 * std::max(o->nSectors, 1), s));
    a = surf_args(o);
    if (nF > 0) {
        DevBuf<int> secN((size_t)40 * std::max(o->nSectors, 1)), secCnt(std::max(o->nSectors, 1));
        VSDMO_HIP(hipMemsetAsync(secCnt.p, 0, sizeof(int) * std::max(o->nSectors, 1), s));
        surf_sector_launch(a, secN.p, secCnt.p, o->nSectors, s);
        VSDMO_HIP(hipStreamSynchronize(s));
    }
    VSDMO_HIP(hipGetLastError());
    o->opt1d_host = o->opt1d.to_host(s);
    o->surface = true;
}

vsdmo_status vsdmo_opt_create(vsdmo_mesh* m, int32_t criterion, float grid_scale, int32_t n_iter,
                              const vsdmo_params* params, vsdmo_opt** out)
{
    return guarded([&]() -> vsdmo_status {
        if (!out) return fail(VSDMO_INVALID_ARG, "out is NULL");
        *out = nullptr;
        if (!m) return fail(VSDMO_INVALID_ARG, "mesh is NULL");
        if (criterion != VSDMO_MEAN_RATIO) return fail(VSDMO_NOT_SUPPORTED, "only the mean-ratio criterion is supported");
        if (!(grid_scale > 0.0f) || n_iter < 0) return fail(VSDMO_INVALID_ARG, "grid_scale must be > 0 and n_iter >= 0");
        vsdmo_params p;
        vsdmo_params_default(&p);
        if (params) p = *params;
        if (!(p.nq == 2 || p.nq == 4 || p.nq == 6 || p.nq == 8))
            return fail(VSDMO_INVALID_ARG, "nq must be 2, 4, 6 or 8");
        if (p.depth < 1 || p.depth > 16) return fail(VSDMO_INVALID_ARG, "depth must be in [1,16]");
        if (!waves_supported(p.nq, p.waves_per_vertex))
            return fail(VSDMO_INVALID_ARG, "waves_per_vertex must be 0 (auto), 1, 2 or 4 and divide nq");
        if (p.surface_path < VSDMO_SURFACE_PATH_AUTO || p.surface_path > VSDMO_SURFACE_PATH_FLAT)
            return fail(VSDMO_INVALID_ARG, "unknown surface_path");
        if (p.hex_metric != VSDMO_HEX_SHAPE9 && p.hex_metric != VSDMO_HEX_CORNER8)
            return fail(VSDMO_INVALID_ARG, "unknown hex_metric");
        if (p.surface_ring != 1 && p.surface_ring != 2) return fail(VSDMO_INVALID_ARG, "surface_ring must be 1 or 2");
        const bool surfaceMesh = (m->elem == VSDMO_TRI || m->elem == VSDMO_QUAD);
        bool surfMesh = false;   // the surface path of a tri / quad mesh
        if (surfaceMesh) {
            if (p.surface_path == VSDMO_SURFACE_PATH_FLAT && !(m->elem == VSDMO_TRI && m->flat))
                return fail(VSDMO_NOT_SUPPORTED, "the flat path needs a flat triangle mesh (z == 0)");
            surfMesh = p.surface_path == VSDMO_SURFACE_PATH_SURFACE || !(m->elem == VSDMO_TRI && m->flat);
        }
        if ((!p.fixed_boundary || surfMesh) && (m->maxHE > 40 || m->nSurf == 0))
            return fail(VSDMO_NOT_SUPPORTED, "surface optimisation needs surface one-rings of at most 40 vertices");
        auto* o = new vsdmo_opt();
        o->m = m;
        o->criterion = criterion;
        o->grid_scale = grid_scale;
        o->n_iter = n_iter;
        o->p = p;
        o->aff = 1.0f / (float)(p.nq - 1);
        o->shrink = p.shrink > 0.0f ? p.shrink : 2.0f / (float)(p.nq - 1);
        int em = p.extent_mode;
        if (em == VSDMO_EXTENT_AUTO) em = (m->elem == VSDMO_HEX) ? VSDMO_EXTENT_MIN : VSDMO_EXTENT_MAX;
        o->extent_min = (em == VSDMO_EXTENT_MIN);
        o->surf_mesh = surfMesh;
        o->qelem = surfMesh ? (m->elem == VSDMO_QUAD ? VSDMO_QUAD : kElemTri3) : m->elem;
        try {
            int nb = (m->nE + 255) / 256;
            o->q.alloc(m->nE);
            o->pmin.alloc(nb); o->psum.alloc(nb);
            o->qout.alloc(2);
            VSDMO_HIP(hipMemsetAsync(o->qout.p, 0, 2 * sizeof(double), m->stream));
            o->hq.alloc();
            if (!p.fixed_boundary || surfMesh) surface_init(o);
            opt_quality(o);
        } catch (...) {
            delete o;
            throw;
        }
        m->n_opts++;
        *out = o;
        return VSDMO_OK;
    });
}

// waves_per_vertex 0: one launch per class with one vertex per wave and the
// last kSplitTail vertices at 4 waves each (tet / hex, nq 4 or 8)
constexpr int kSplitTail = 1024;
static void dispatch_search(vsdmo_opt* o, int kind, const SearchArgs& a, size_t lds, hipStream_t s)
{
    const int w = o->p.waves_per_vertex;
    if (w == 0 && split_supported(kind, o->p.nq, o->m->maxRing))
        launch_search_split(kind, o->p.nq, a, kSplitTail, o->m->maxRing, s);
    else
        launch_search(kind, o->p.nq, w == 0 ? 1 : w, a, lds, s);
}

// HIP events on every search launch while timing is on: the kernels' own
// begin and end (KernelEvents; every timed launch goes through
// launch_kernel).  The events are pooled and reused: creating them per
// launch would cost host time inside the iteration loop that is being timed.
static void timed(vsdmo_opt* o, const std::function<void()>& launch)
{
    if (o->timing) {
        const size_t k = o->ev_used;
        while (o->ev.size() < k + 2) {
            hipEvent_t e;
            VSDMO_HIP(hipEventCreate(&e));
            o->ev.push_back(e);
        }
        KernelEvents ke;
        ke.start = o->ev[k];
        ke.stop = o->ev[k + 1];
        KernelEvents* outer = kernel_events();
        kernel_events() = &ke;
        try {
            launch();
        } catch (...) {
            kernel_events() = outer;
            throw;
        }
        kernel_events() = outer;
        VSDMO_HIP(hipGetLastError());
        if (!ke.started) throw StatusError{ VSDMO_NOT_SUPPORTED, "timing: a search launch recorded no kernel events" };
        o->ev_pairs.push_back({ k, k + 1 });
        o->ev_used += 2;
    } else {
        launch();
    }
    o->launches++;
}

static void collect_timing(vsdmo_opt* o)
{
    if (o->ev_used == 0) return;
    VSDMO_HIP(hipStreamSynchronize(o->m->stream));
    for (const auto& pr : o->ev_pairs) {
        float ms = 0.0f;
        VSDMO_HIP(hipEventElapsedTime(&ms, o->ev[pr.first], o->ev[pr.second]));
        o->kernel_ms += ms;
    }
    o->ev_pairs.clear();
    o->ev_used = 0;
}

vsdmo_status vsdmo_opt_iterate(vsdmo_opt* o, int32_t n)
{
    return guarded([&]() -> vsdmo_status {
        if (!o) return fail(VSDMO_INVALID_ARG, "NULL optimiser");
        if (n < 0) return fail(VSDMO_INVALID_ARG, "negative iteration count");
        vsdmo_mesh* m = o->m;
        SearchArgs a{};
        a.x = m->x.p; a.y = m->y.p; a.z = m->z.p;
        a.xw = m->x.p; a.yw = m->y.p; a.zw = m->z.p;
        a.ringRow = m->ringRow.p; a.ringVal = m->ringVal.p;
        a.heRow = m->heRow.p; a.heTarget = m->heTarget.p;
        a.gridScale = o->grid_scale; a.shrink = o->shrink; a.aff = o->aff;
        a.depth = o->p.depth; a.extentMin = o->extent_min;
        a.hexCorner8 = o->p.hex_metric == VSDMO_HEX_CORNER8;
        for (int it = 0; it < n; ++it) {
            if (m->elem == VSDMO_TET || m->elem == VSDMO_HEX) {
                size_t lds = search_lds_bytes(m->elem, m->maxRing);
                // the classes' fallback-list lengths start at zero (one fill per iteration)
                const int nCls = (int)m->colOffInner.size() - 1;
                const bool deferred = o->learn_mode != 0 && !o->rec && o->learn_fallback == VSDMO_FALLBACK_DEFERRED;
                const int par = (int)(o->curr_it & 1);
                const int nIn = m->colOffInner.empty() ? 0 : m->colOffInner.back() - m->colOffInner.front();
                if (deferred) launch_fill_int(o->dCnt.p + (size_t)par * std::max(nCls, 1), std::max(nCls, 0), 0, m->stream);
                else if (o->learn_mode != 0 && !o->rec)
                    launch_fill_int(o->lCnt.p, (int)m->colOffInner.size(), 0, m->stream);
                // an iteration that does not keep the deferred lists (recording,
                // learning off, the immediate policy) drops them: a later deferred
                // pass must not take a stale list, whose vertices its windows
                // would not skip
                if (deferred) o->defer_pending = true;
                else if (o->defer_pending) {
                    reset_deferred(o);
                    o->defer_pending = false;
                }
                for (size_t c = 0; c + 1 < m->colOffInner.size(); ++c) {
                    a.cStart = m->colOffInner[c]; a.cEnd = m->colOffInner[c + 1];
                    const int nc = a.cEnd - a.cStart;
                    vsdmo_dataset* ds = o->rec;
                    const int take = ds ? (int)std::min<long long>(nc, ds->cap - ds->n) : 0;
                    if (take > 0)
                        launch_features(m->x.p, m->y.p, m->z.p, m->nbr.p, a.cStart, a.cStart + take,
                                        ds->X.p + ds->n * kIn, ds->L.p, ds->v0.p, m->stream);
                    if (o->learn_mode != 0 && !ds) {
                        LearnNet net;
                        net.Y = o->lY.p; net.L = o->lL.p;
                        if (learn_uses_net(o->learn_mode)) {
                            if (!o->learn_separate && learned_fused_ok(m->elem, m->maxRing)) {
                                net.wb = o->mlp->wb.p; net.w = o->mlp->w.p; net.nbr = m->nbr.p;
                                net.loc = m->ringLoc.p;
                            } else {
                                launch_mlp_infer(o->mlp, m->x.p, m->y.p, m->z.p, m->nbr.p, a.cStart, a.cEnd, nullptr,
                                                 o->lY.p, o->lL.p, m->stream);
                            }
                        }
                        if (deferred) {
                            const size_t off = (size_t)(a.cStart - m->colOffInner.front());
                            const size_t cs = (size_t)std::max(nCls, 1);
                            LearnDefer def{ o->dList.p + (size_t)(1 - par) * nIn + off, o->dCnt.p + (1 - par) * cs + c,
                                            o->dEpoch.p, (int)o->curr_it };
                            timed(o, [&]() {
                                launch_search_learned(m->elem, o->p.nq, a, net, o->learn_mode, o->learn_depth,
                                                      o->learn_frac, o->lFallback.p, o->dList.p + (size_t)par * nIn + off,
                                                      o->dCnt.p + par * cs + c, m->maxRing, &def, m->stream);
                            });
                        } else {
                            timed(o, [&]() {
                                launch_search_learned(m->elem, o->p.nq, a, net, o->learn_mode, o->learn_depth,
                                                      o->learn_frac, o->lFallback.p, o->lList.p, o->lCnt.p + c,
                                                      m->maxRing, nullptr, m->stream);
                            });
                        }
                        o->s_learned += nc;
                    } else {
                        timed(o, [&]() { dispatch_search(o, m->elem, a, lds, m->stream); });
                    }
                    if (take > 0) {
                        launch_targets(m->x.p, m->y.p, m->z.p, a.cStart, a.cStart + take, ds->L.p, ds->v0.p,
                                       ds->T.p + ds->n * kOut, m->stream);
                        ds->n += take;
                    }
                    o->s_inner += nc;
                }
                if (o->surface && o->last_min > 0.0f) {
                    SurfArgs sa = surf_args(o);
                    size_t slds = surf_lds_bytes(m->elem, m->maxRing);
                    const int tet = (m->elem == VSDMO_TET);
                    for (size_t c = 0; c + 1 < m->colOffFree.size(); ++c) {
                        int lo = m->colOffFree[c], hi = m->colOffFree[c + 1];
                        snapshot_launch(sa, m->stream);
                        surf_search2d_launch(sa, lo, hi, o->p.nq, tet, slds, m->stream);
                        o->s_surf += hi - lo;
                    }
                    for (size_t c = 0; c + 1 < m->colOffFeature.size(); ++c) {
                        int lo = m->colOffFeature[c], hi = m->colOffFeature[c + 1];
                        surf_search1d_launch(sa, lo, hi, o->p.nq, tet, slds, m->stream);
                        for (int v = lo; v < hi; ++v) o->s_feat += o->opt1d_host[v];
                    }
                }
                if (o->surface) surf_nn_launch(surf_args(o), m->stream);
            } else if (o->surf_mesh) {
                // surface mesh on the surface path (DMOTri.cu:128-170, DMOQuad.cu:128-170):
                // free 2D passes, feature 1D passes, NN walk, normals; no quality gate
                SurfArgs sa = surf_args(o);
                for (size_t c = 0; c + 1 < m->colOffFree.size(); ++c) {
                    int lo = m->colOffFree[c], hi = m->colOffFree[c + 1];
                    snapshot_launch(sa, m->stream);
                    timed(o, [&]() { surf_fan2d_launch(sa, lo, hi, o->p.nq, m->stream); });
                    o->s_surf += hi - lo;
                }
                if (!o->p.fixed_boundary) {
                    for (size_t c = 0; c + 1 < m->colOffFeature.size(); ++c) {
                        int lo = m->colOffFeature[c], hi = m->colOffFeature[c + 1];
                        if (m->elem == VSDMO_QUAD) {   // quad fan diagonals read the pass snapshot
                            snapshot_launch(sa, m->stream);
                        }
                        timed(o, [&]() { surf_fan1d_launch(sa, lo, hi, o->p.nq, m->stream); });
                        for (int v = lo; v < hi; ++v) o->s_feat += o->opt1d_host[v];
                    }
                }
                surf_nn_launch(sa, m->stream);
                surf_normals_launch(sa, m->stream);
            } else {
                size_t lds = search_lds_bytes(VSDMO_TRI, m->maxHE);
                for (size_t c = 0; c + 1 < m->colOffFree.size(); ++c) {
                    a.cStart = m->colOffFree[c]; a.cEnd = m->colOffFree[c + 1];
                    timed(o, [&]() { dispatch_search(o, VSDMO_TRI, a, lds, m->stream); });
                    o->s_surf += a.cEnd - a.cStart;
                }
                if (o->surface && o->last_min > 0.0f) {
                    SurfArgs sa = surf_args(o);
                    for (size_t c = 0; c + 1 < m->colOffFeature.size(); ++c) {
                        int lo = m->colOffFeature[c], hi = m->colOffFeature[c + 1];
                        surf_flat1d_launch(sa, lo, hi, o->p.nq, m->stream);
                        for (int v = lo; v < hi; ++v) o->s_feat += o->opt1d_host[v];
                    }
                }
                if (o->surface) surf_nn_launch(surf_args(o), m->stream);
            }
            VSDMO_HIP(hipGetLastError());
            float prev = o->last_min;
            opt_quality(o);
            o->curr_it++;
            if (o->p.quality_guard) {
                float tol = o->p.guard_tolerance * std::fabs(prev);
                if (o->last_min < prev - tol) {
                    collect_timing(o);
                    std::string msg = "minimum quality decreased from " + std::to_string(prev) + " to " +
                                      std::to_string(o->last_min);
                    if (!o->fail_dump.empty()) {
                        if (vsdmo_mesh_write(m, o->fail_dump.c_str(), 0) == VSDMO_OK) msg += "; mesh written to " + o->fail_dump;
                        else msg += "; writing " + o->fail_dump + " failed";
                    }
                    return fail(VSDMO_QUALITY_DECREASED, msg);
                }
            }
        }
        collect_timing(o);
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_opt_optimize(vsdmo_opt* o)
{
    if (!o) return fail(VSDMO_INVALID_ARG, "NULL optimiser");
    int left = o->n_iter - o->curr_it;
    return left > 0 ? vsdmo_opt_iterate(o, left) : VSDMO_OK;
}

int32_t vsdmo_opt_is_done(const vsdmo_opt* o) { return o ? (o->curr_it >= o->n_iter) : 1; }

vsdmo_status vsdmo_opt_min_quality(const vsdmo_opt* o, float* out)
{
    if (!o || !out) return fail(VSDMO_INVALID_ARG, "NULL argument");
    *out = o->last_min;
    return VSDMO_OK;
}

vsdmo_status vsdmo_opt_mean_quality(const vsdmo_opt* o, double* out)
{
    if (!o || !out) return fail(VSDMO_INVALID_ARG, "NULL argument");
    *out = o->m->nE > 0 ? o->last_sum / o->m->nE : 0.0;
    return VSDMO_OK;
}

vsdmo_status vsdmo_opt_histogram(const vsdmo_opt* o, int32_t n_cols, int32_t* out)
{
    return guarded([&]() -> vsdmo_status {
        if (!o || !out || n_cols <= 0) return fail(VSDMO_INVALID_ARG, "bad argument");
        DevBuf<int> bins(n_cols);
        VSDMO_HIP(hipMemsetAsync(bins.p, 0, sizeof(int) * n_cols, o->m->stream));
        launch_histogram(o->q.p, o->m->nE, n_cols, bins.p, o->m->stream);
        VSDMO_HIP(hipGetLastError());
        bins.download(out, n_cols, o->m->stream);
        VSDMO_HIP(hipStreamSynchronize(o->m->stream));
        return VSDMO_OK;
    });
}

// ------------------------------------------------ allocation-free histogram
// The bins of vsdmo_opt_histogram_noalloc live in this file-scope device
// array: each GPU the library runs on holds one copy of
// VSDMO_HISTOGRAM_NOALLOC_MAX_COLS ints until the process ends.  Host code
// reaches it only through the per-device address from hipGetSymbolAddress;
// the array's name is not a device address on the host.
static __device__ __align__(256) int scratch[VSDMO_HISTOGRAM_NOALLOC_MAX_COLS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch[kMaxDevices] = {};

// the scratch array's address on device dev, looked up once per device with
// that device current
int* scratch_on(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) throw StatusError{ VSDMO_NOT_SUPPORTED, "device index out of range" };
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        int cur = 0;
        VSDMO_HIP(hipGetDevice(&cur));
        if (cur != dev) VSDMO_HIP(hipSetDevice(dev));
        int* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(reinterpret_cast<void**>(&ptr), HIP_SYMBOL(scratch));
        if (cur != dev) VSDMO_HIP(hipSetDevice(cur));
        if (e != hipSuccess) throw HipError{ e, "hipGetSymbolAddress(scratch)" };
        g_scratch[dev] = ptr;
    }
    return g_scratch[dev];
}
}  // namespace

vsdmo_status vsdmo_opt_histogram_noalloc(const vsdmo_opt* o, int32_t n_cols, int32_t* out)
{
    return guarded([&]() -> vsdmo_status {
        if (!o || !out || n_cols <= 0) return fail(VSDMO_INVALID_ARG, "bad argument");
        if (n_cols > VSDMO_HISTOGRAM_NOALLOC_MAX_COLS)
            return fail(VSDMO_INVALID_ARG, "n_cols above VSDMO_HISTOGRAM_NOALLOC_MAX_COLS (" +
                                               std::to_string(VSDMO_HISTOGRAM_NOALLOC_MAX_COLS) + ")");
        hipStream_t s = o->m->stream;
        int dev = 0;
        VSDMO_HIP(hipStreamGetDevice(s, &dev));
        int* bins = scratch_on(dev);
        // the array keeps the last call's counts: cleared by a kernel first
        launch_fill_int(bins, n_cols, 0, s);
        launch_histogram(o->q.p, o->m->nE, n_cols, bins, s);
        VSDMO_HIP(hipGetLastError());
        VSDMO_HIP(hipMemcpyAsync(out, bins, sizeof(int) * (size_t)n_cols, hipMemcpyDeviceToHost, s));
        VSDMO_HIP(hipStreamSynchronize(s));
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_opt_element_qualities(const vsdmo_opt* o, float* host_out)
{
    if (!o) return fail(VSDMO_INVALID_ARG, "NULL optimiser");
    return copy_dev(o->q, (size_t)o->m->nE, host_out, o->m->stream);
}

vsdmo_status vsdmo_opt_set_fallback(vsdmo_opt* o, int32_t policy)
{
    return guarded([&]() -> vsdmo_status {
        if (!o) return fail(VSDMO_INVALID_ARG, "NULL optimiser");
        if (policy != VSDMO_FALLBACK_IMMEDIATE && policy != VSDMO_FALLBACK_DEFERRED)
            return fail(VSDMO_INVALID_ARG, "policy must be VSDMO_FALLBACK_IMMEDIATE or VSDMO_FALLBACK_DEFERRED");
        if (policy == VSDMO_FALLBACK_DEFERRED && !learned_deferred_ok(o->m->elem, o->p.nq, o->m->maxRing))
            return fail(VSDMO_NOT_SUPPORTED, "the deferred fallback needs nq 4 or 8 and one-rings of at most 128 tets "
                                             "or 21 hexes");
        o->learn_fallback = policy;
        reset_deferred(o);
        VSDMO_HIP(hipStreamSynchronize(o->m->stream));
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_opt_stats(const vsdmo_opt* o, vsdmo_stats* out)
{
    if (!o || !out) return fail(VSDMO_INVALID_ARG, "NULL argument");
    int fb = 0;
    if (o->lFallback.p) {   // the device counter is read here, not per iteration
        vsdmo_status st = copy_dev(o->lFallback, 1, &fb, o->m->stream);
        if (st != VSDMO_OK) return st;
    }
    out->searches_inner = o->s_inner;
    out->searches_surface = o->s_surf;
    out->searches_feature = o->s_feat;
    out->iterations = o->curr_it;
    out->min_quality = o->last_min;
    out->kernel_ms = o->kernel_ms;
    out->kernel_launches = o->launches;
    out->searches_learned = o->s_learned;
    out->fallbacks = fb;
    return VSDMO_OK;
}

vsdmo_status vsdmo_opt_surface_points(vsdmo_opt* o, int32_t vid, int32_t nu, int32_t nv, int32_t mode,
                                      int32_t feature_sid, float* out)
{
    return guarded([&]() -> vsdmo_status {
        if (!o || !out) return fail(VSDMO_INVALID_ARG, "NULL argument");
        vsdmo_mesh* m = o->m;
        if (!o->surface) return fail(VSDMO_NOT_SUPPORTED, "local surfaces exist only with fixed_boundary = 0");
        if (vid < 0 || vid >= m->nSurf) return fail(VSDMO_INVALID_ARG, "vid is not a surface vertex");
        if (nu < 2 || nv < 2 || (long long)nu * nv > (1 << 24)) return fail(VSDMO_INVALID_ARG, "nu, nv must be >= 2");
        if (mode != 0 && mode != 1) return fail(VSDMO_INVALID_ARG, "mode must be 0 or 1");
        SurfArgs a = surf_args(o);
        a.sx = m->x.p; a.sy = m->y.p; a.sz = m->z.p;   // the current positions
        DevBuf<float> d((size_t)3 * nu * nv);
        surf_points_launch(a, vid, nu, nv, mode, feature_sid, d.p, m->stream);
        VSDMO_HIP(hipGetLastError());
        d.download(out, d.n, m->stream);
        VSDMO_HIP(hipStreamSynchronize(m->stream));
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_opt_set_proposal(vsdmo_opt* o, vsdmo_mlp* mlp, int32_t mode, int32_t depth, float window_frac)
{
    return guarded([&]() -> vsdmo_status {
        if (!o) return fail(VSDMO_INVALID_ARG, "NULL optimiser");
        vsdmo_mesh* m = o->m;
        if (mode < 0 || mode > 3)
            return fail(VSDMO_INVALID_ARG, "mode must be 0 (off), 1 (network), 2 (identity) or 3 (network, then identity)");
        if (learn_uses_net(mode) && !mlp) return fail(VSDMO_INVALID_ARG, "modes 1 and 3 need a network");
        if (mode != 0 && m->elem != VSDMO_TET && m->elem != VSDMO_HEX)
            return fail(VSDMO_NOT_SUPPORTED, "the learned proposal is for tet and hex inner vertices");
        if (mode != 0 && (depth < 1 || depth > 8 || !(window_frac > 0.0f)))
            return fail(VSDMO_INVALID_ARG, "depth must be in [1,8] and window_frac > 0");
        if (mode != 0 && m->maxRing > learned_ring_cap(m->elem))
            return fail(VSDMO_NOT_SUPPORTED, "the learned search holds one-rings of at most " +
                                                 std::to_string(learned_ring_cap(m->elem)) + " elements");
        if (mode != 0) {
            // buffers first: a failure here leaves the previous setting in place
            int maxc = 1;
            for (size_t c = 0; c + 1 < m->colOffInner.size(); ++c)
                maxc = std::max(maxc, m->colOffInner[c + 1] - m->colOffInner[c]);
            if (!o->lFallback.p) {
                o->lFallback.alloc(1);
                launch_fill_int(o->lFallback.p, 1, 0, m->stream);
            }
            if ((int)o->lL.n < maxc) {
                o->lL.alloc(maxc); o->lY.alloc((size_t)3 * maxc); o->lList.alloc(maxc);
            }
            if (!o->lCnt.p) o->lCnt.alloc(std::max<size_t>(m->colOffInner.size(), 1));
            reset_deferred(o);
            if (learn_uses_net(mode)) {
                build_neighbour_table(*m);
                VSDMO_HIP(hipStreamSynchronize(mlp->stream));   // the weights are final
            }
            VSDMO_HIP(hipStreamSynchronize(m->stream));
        }
        o->learn_mode = mode;
        if (o->mlp) o->mlp->n_users--;
        o->mlp = learn_uses_net(mode) ? mlp : nullptr;
        if (o->mlp) o->mlp->n_users++;
        o->learn_depth = depth;
        o->learn_frac = window_frac;
        const char* sep = std::getenv("VSDMO_LEARN_SEPARATE_INFERENCE");
        o->learn_separate = sep && sep[0] && sep[0] != '0';
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_opt_record(vsdmo_opt* o, vsdmo_dataset* ds, int32_t n)
{
    return guarded([&]() -> vsdmo_status {
        if (!o || !ds) return fail(VSDMO_INVALID_ARG, "NULL argument");
        vsdmo_mesh* m = o->m;
        if (m->elem != VSDMO_TET && m->elem != VSDMO_HEX)
            return fail(VSDMO_NOT_SUPPORTED, "training pairs are recorded from tet and hex inner searches");
        int maxc = 1;
        for (size_t c = 0; c + 1 < m->colOffInner.size(); ++c)
            maxc = std::max(maxc, m->colOffInner[c + 1] - m->colOffInner[c]);
        build_neighbour_table(*m);
        VSDMO_HIP(hipStreamSynchronize(ds->stream));
        if ((int)ds->L.n < maxc) { ds->L.alloc(maxc); ds->v0.alloc((size_t)3 * maxc); }
        o->rec = ds;
        vsdmo_status st = vsdmo_opt_iterate(o, n);
        o->rec = nullptr;
        VSDMO_HIP(hipStreamSynchronize(m->stream));
        return st;
    });
}

vsdmo_status vsdmo_opt_set_timing(vsdmo_opt* o, int32_t on)
{
    if (!o) return fail(VSDMO_INVALID_ARG, "NULL optimiser");
    o->timing = on != 0;
    o->kernel_ms = 0.0;
    o->launches = 0;
    return VSDMO_OK;
}

vsdmo_status vsdmo_opt_set_failure_dump(vsdmo_opt* o, const char* path)
{
    if (!o) return fail(VSDMO_INVALID_ARG, "NULL optimiser");
    o->fail_dump = path ? path : "";
    return VSDMO_OK;
}

vsdmo_status vsdmo_opt_destroy(vsdmo_opt* o)
{
    return guarded([&]() -> vsdmo_status {
        if (!o) return VSDMO_OK;
        collect_timing(o);
        for (hipEvent_t e : o->ev) (void)hipEventDestroy(e);
        o->m->n_opts--;
        if (o->mlp) o->mlp->n_users--;
        delete o;
        return VSDMO_OK;
    });
}

}  // extern "C"
