// batch.hip -- batched DMO over many independent meshes.
//
// The meshes' positions, one-rings and elements are concatenated into one
// set of device arrays (global vertex ids).  Colour round r holds colour r of
// every mesh (inner colours of tet / hex meshes, free colours of flat
// triangle meshes): a union of independent sets of disjoint meshes is again
// independent, so one persistent launch per round sweeps every mesh.  One
// iteration = a counter reset, the rounds and the per-mesh quality
// reduction; it is captured once as a hipGraph and replayed.  The per-mesh
// minimum quality of every iteration is appended to a device history that
// the host reads after the last replay (quality guard).
//
// There is no batched driver in the reference; each mesh's searches are the
// reference's (dmo/DMOTetImplCub.h:20-137, DMOHexImplCub.h:20-137,
// DMOTriFlatImplCub.h:23-152) and give the same positions as one optimiser
// per mesh.
#include <float.h>

#include <climits>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "dmo_kernels.h"
#include "mlp_internal.h"
#include "vsdmo_internal.h"

namespace vsdmo {

__global__ void k_offset_copy(int* __restrict__ dst, const int* __restrict__ src, long long n, int add)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] + add;
}

static void offset_copy(int* dst, const int* src, long long n, int add, hipStream_t s)
{
    if (n > 0) k_offset_copy<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(dst, src, n, add);
}

static vsdmo_status fail_status(vsdmo_status st, const std::string& msg)
{
    set_error(msg);
    return st;
}

}  // namespace vsdmo

using namespace vsdmo;

struct vsdmo_batch {
    std::vector<vsdmo_mesh*> meshes;
    hipStream_t stream = nullptr;
    vsdmo_batch_params p{};
    float aff = 0.0f, shrink = 0.0f;
    int totalV = 0;
    std::vector<int> vOff;                       // nMeshes+1
    DevBuf<float> x, y, z;                       // working positions
    DevBuf<float> x0, y0, z0;                    // positions at creation (restore)
    // one-ring CSRs per element type over all batch vertices; rows of the
    // other meshes are empty and hold the running total of their own type, so
    // the row end of every mesh's last vertex is its own total
    DevBuf<int> rowTet, ringTet, rowHex, ringHex, heRow, heTarget;
    int maxRingTet = 0, maxRingHex = 0, maxHE = 0;
    // work items per round, in mesh order, and the order used (shuffled)
    // round r: tets [off, midT), hexes [midT, mid), flat triangles [mid, off+1)
    std::vector<int> itemsMeshOrder, roundOff, roundMidT, roundMid;
    DevBuf<int> items, ctr;
    std::vector<int> itemsPerMesh;               // searches per iteration per mesh
    // qualities
    DevBuf<int> cells, chunkElem0, chunkType, meshChunk0;
    DevBuf<long long> chunkCell0;
    DevBuf<float> cmin, mmin, hist;
    DevBuf<double> csum, msum;
    DevBuf<int> hctr;
    int nChunks = 0, hcap = 0, hrow = 0;
    int qper = 1;          // quality elements per thread (chunks of 256 * qper)
    std::vector<float> lastMin;
    std::vector<int> nElem;
    // execution
    size_t lds = 0;
    int cus = 0;
    std::vector<int> grid;                       // per round
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long long searches = 0;
    int iterations = 0;
    double gpu_ms = 0.0;
    // learned proposal of the tet items ([0]) and the hex items ([1]):
    // network and window per type, the global neighbour table; immediate
    // policy: the fallback list (reused by each learned launch in stream
    // order), the list lengths (tets: [0, rounds), hexes: [rounds, 2 rounds))
    struct Learn {
        vsdmo_mlp* mlp = nullptr;
        int mode = 0, depth = 1;
        float frac = 0.025f;
    } learn[2];
    DevBuf<int> nbr, lList, lCnt, lFallback;
    // the meshes' ring-corner table positions in the batch's tet / hex ring
    // index space (built with nbr); first ring element of every mesh in them
    DevBuf<uint8_t> locTet, locHex;
    std::vector<long long> ringBaseTet, ringBaseHex;
    long long learned = 0;
    // deferred fallback: the lists of (type t, round r) at dOff[t * rounds + r]
    // (parity 1 at + dTotal), their counts at t * rounds + r (parity 1 at
    // + 2 rounds), the per-vertex deferral epochs and the iteration counter,
    // all on the device so that one captured iteration serves every replay
    int fallback = VSDMO_FALLBACK_IMMEDIATE;
    std::vector<int> dOff;
    int dTotal = 0;
    DevBuf<int> dList, dCnt, dEpoch, itDev;
};

namespace {

uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

void drop_graph(vsdmo_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

SearchArgs search_args(vsdmo_batch* b)
{
    SearchArgs a{};
    a.x = b->x.p; a.y = b->y.p; a.z = b->z.p;
    a.xw = b->x.p; a.yw = b->y.p; a.zw = b->z.p;
    a.ringRow = b->rowTet.p; a.ringVal = b->ringTet.p;
    a.heRow = b->heRow.p; a.heTarget = b->heTarget.p;
    a.gridScale = b->p.grid_scale; a.shrink = b->shrink; a.aff = b->aff;
    a.depth = b->p.depth;
    a.hexCorner8 = b->p.hex_metric == VSDMO_HEX_CORNER8;
    return a;
}

BatchQArgs quality_args(vsdmo_batch* b)
{
    BatchQArgs q{};
    q.x = b->x.p; q.y = b->y.p; q.z = b->z.p;
    q.cells = b->cells.p;
    q.chunkElem0 = b->chunkElem0.p; q.chunkCell0 = b->chunkCell0.p; q.chunkType = b->chunkType.p;
    q.meshChunk0 = b->meshChunk0.p;
    q.cmin = b->cmin.p; q.csum = b->csum.p; q.mmin = b->mmin.p; q.msum = b->msum.p;
    q.hist = b->hist.p; q.hctr = b->hctr.p;
    q.nChunks = b->nChunks; q.nMeshes = (int)b->meshes.size(); q.hcap = b->hcap;
    q.hexCorner8 = b->p.hex_metric == VSDMO_HEX_CORNER8;
    q.per = b->qper;
    return q;
}

// work order of the rounds: the mesh order, or a seeded permutation of the
// items inside each round (the results must not depend on it)
void upload_items(vsdmo_batch* b)
{
    std::vector<int> it = b->itemsMeshOrder;
    if (b->p.shuffle_seed != 0) {
        uint64_t s = b->p.shuffle_seed;
        auto shuffle = [&](int lo, int hi) {
            for (int i = hi - 1; i > lo; --i) {
                int j = lo + (int)(splitmix(s) % (uint64_t)(i - lo + 1));
                std::swap(it[i], it[j]);
            }
        };
        for (size_t r = 0; r + 1 < b->roundOff.size(); ++r) {
            shuffle(b->roundOff[r], b->roundMidT[r]);
            shuffle(b->roundMidT[r], b->roundMid[r]);
            shuffle(b->roundMid[r], b->roundOff[r + 1]);
        }
    }
    b->items.upload(it.data(), it.size(), b->stream);
    VSDMO_HIP(hipStreamSynchronize(b->stream));
}

void size_grids(vsdmo_batch* b)
{
    b->lds = batch_lds_bytes(b->maxRingTet, b->maxRingHex, b->maxHE);
    int per = std::max(1, batch_round_blocks_per_cu(b->p.nq, b->p.waves_per_vertex, b->lds));
    b->grid.assign(b->roundOff.size() > 0 ? b->roundOff.size() - 1 : 0, 0);
    for (size_t r = 0; r + 1 < b->roundOff.size(); ++r) {
        int n = b->roundOff[r + 1] - b->roundOff[r];
        b->grid[r] = b->p.persistent ? std::min(n, b->cus * per) : 0;
    }
}

bool learning(const vsdmo_batch* b) { return b->learn[0].mode != 0 || b->learn[1].mode != 0; }

bool deferred(const vsdmo_batch* b) { return b->fallback == VSDMO_FALLBACK_DEFERRED; }

void alloc_deferred(vsdmo_batch* b)
{
    if (b->dList.p) return;
    const int nR = (int)b->grid.size();
    b->dOff.assign((size_t)2 * nR, 0);
    int total = 0;
    for (int t = 0; t < 2; ++t)
        for (int r = 0; r < nR; ++r) {
            b->dOff[(size_t)t * nR + r] = total;
            total += t == 0 ? b->roundMidT[r] - b->roundOff[r] : b->roundMid[r] - b->roundMidT[r];
        }
    b->dTotal = std::max(total, 1);
    b->dList.alloc((size_t)2 * b->dTotal);
    b->dCnt.alloc((size_t)std::max(4 * nR, 1));
    b->dEpoch.alloc((size_t)std::max(b->totalV, 1));
    b->itDev.alloc(1);
    launch_fill_int(b->itDev.p, 1, 0, b->stream);
}

// no pending deferrals: both parities' counts zero, no vertex deferred
void reset_deferred(vsdmo_batch* b)
{
    if (!b->dList.p) return;
    launch_fill_int(b->dCnt.p, std::max(4 * (int)b->grid.size(), 1), 0, b->stream);
    launch_fill_int(b->dEpoch.p, std::max(b->totalV, 1), INT_MIN, b->stream);
    VSDMO_HIP(hipStreamSynchronize(b->stream));
}

// the deferred policy's full searches need nq 4 / 8 and every learned type's
// rings within the window kernel's LDS tile
std::string deferred_check(const vsdmo_batch* b, int nq)
{
    for (int t = 0; t < 2; ++t)
        if (b->learn[t].mode != 0 &&
            !learned_deferred_ok(t == 0 ? VSDMO_TET : VSDMO_HEX, nq, t == 0 ? b->maxRingTet : b->maxRingHex))
            return "the deferred fallback needs nq 4 or 8 and one-rings of at most 128 tets or 21 hexes";
    return "";
}

// one type's items [lo, hi) of round r through the learned window kernel;
// the full search of their fallbacks follows (immediate policy) or runs in
// the round's launch of the next iteration (deferred)
void launch_learned_items(vsdmo_batch* b, int t, const SearchArgs& a, int lo, int hi, int r, hipStream_t s)
{
    const vsdmo_batch::Learn& L = b->learn[t];
    if (L.mode == 0 || hi <= lo) return;
    SearchArgs al = a;
    al.cStart = lo; al.cEnd = hi;
    if (t == 1) {   // hexes: their CSR, the minimum-extent window (as the batch's plain hex search)
        al.ringRow = b->rowHex.p; al.ringVal = b->ringHex.p; al.extentMin = 1;
    }
    LearnNet net;
    net.items = b->items.p;
    if (learn_uses_net(L.mode)) {
        net.wb = L.mlp->wb.p; net.w = L.mlp->w.p; net.nbr = b->nbr.p;
        net.loc = t == 0 ? b->locTet.p : b->locHex.p;
    }
    const int nR = (int)b->grid.size();
    const int ring = t == 0 ? b->maxRingTet : b->maxRingHex;
    const int kind = t == 0 ? VSDMO_TET : VSDMO_HEX;
    if (deferred(b)) {
        LearnDefer d{ nullptr, nullptr, b->dEpoch.p, 0 };
        d.itDev = b->itDev.p;
        d.listStride = b->dTotal;
        d.cntStride = 2 * nR;
        launch_search_learned(kind, b->p.nq, al, net, L.mode, L.depth, L.frac, b->lFallback.p,
                              b->dList.p + b->dOff[(size_t)t * nR + r], b->dCnt.p + (size_t)t * nR + r, ring, &d, s);
    } else {
        launch_search_learned(kind, b->p.nq, al, net, L.mode, L.depth, L.frac, b->lFallback.p, b->lList.p,
                              b->lCnt.p + (size_t)t * nR + r, ring, nullptr, s);
    }
}

// the launches of one iteration, in stream order
void record_iteration(vsdmo_batch* b)
{
    hipStream_t s = b->stream;
    const int nR = (int)b->grid.size();
    launch_fill_int(b->ctr.p, nR, 0, s);
    const bool learn = learning(b), def = learn && deferred(b);
    if (def) launch_parity_zero(b->dCnt.p, 2 * nR, b->itDev.p, s);   // this iteration's lists
    else if (learn) launch_fill_int(b->lCnt.p, 2 * nR, 0, s);
    SearchArgs a = search_args(b);
    for (int r = 0; r < nR; ++r) {
        const int off = b->roundOff[r], midT = b->roundMidT[r], mid = b->roundMid[r], hi = b->roundOff[r + 1];
        launch_learned_items(b, 0, a, off, midT, r, s);
        launch_learned_items(b, 1, a, midT, mid, r, s);
        // plain searches of the other items: tets [off, midT) and hexes
        // [midT, mid) from the round's queue, flat triangles [mid, hi) strided
        const bool lt = b->learn[0].mode != 0, lh = b->learn[1].mode != 0;
        if (!b->p.persistent) {
            // one block per item: a launch per element type, each with its
            // own body's registers and LDS
            const int w = b->p.waves_per_vertex;
            if (!lt)
                launch_batch_typed(b->p.nq, w, a, b->rowHex.p, b->ringHex.p, b->items.p, off, midT, 0,
                                   batch_lds_bytes(b->maxRingTet, 0, 0), s);
            if (!lh)
                launch_batch_typed(b->p.nq, w, a, b->rowHex.p, b->ringHex.p, b->items.p, midT, mid, 1,
                                   batch_lds_bytes(0, b->maxRingHex, 0), s);
            launch_batch_typed(b->p.nq, w, a, b->rowHex.p, b->ringHex.p, b->items.p, mid, hi, 2,
                               batch_lds_bytes(0, 0, b->maxHE), s);
            continue;
        }
        auto plain = [&](int lo, int qend, int end) {
            launch_batch_round(b->p.nq, b->p.waves_per_vertex, a, b->rowHex.p, b->ringHex.p, b->items.p, lo, qend, end,
                               b->ctr.p + r, b->grid[r], b->lds, s);
        };
        if (lh && !lt) {   // hexes learned, tets plain: the tets, then the triangles
            plain(off, midT, midT);
            plain(mid, mid, hi);
        } else {
            plain(!lt ? off : (lh ? mid : midT), mid, hi);
        }
    }
    if (def) launch_add_int(b->itDev.p, 1, s);
    launch_batch_quality(quality_args(b), s);
}

void capture(vsdmo_batch* b)
{
    drop_graph(b);
    VSDMO_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    try {
        record_iteration(b);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(b->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    VSDMO_HIP(hipStreamEndCapture(b->stream, &b->graph));
    VSDMO_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
}

void grow_history(vsdmo_batch* b, int need)
{
    if (b->hrow + need <= b->hcap) return;
    int cap = std::max(2 * b->hcap, b->hrow + need);
    const size_t nm = b->meshes.size();
    DevBuf<float> h((size_t)cap * nm);
    if (b->hrow > 0)
        VSDMO_HIP(hipMemcpyAsync(h.p, b->hist.p, sizeof(float) * nm * b->hrow, hipMemcpyDeviceToDevice, b->stream));
    VSDMO_HIP(hipStreamSynchronize(b->stream));
    b->hist = std::move(h);
    b->hcap = cap;
    drop_graph(b);   // the quality node holds the old history pointer
}

// run n iterations; returns the per-mesh minimum qualities after each
// iteration (row-major n x nMeshes) and checks the guard
vsdmo_status run(vsdmo_batch* b, int n, std::vector<float>* rows_out)
{
    if (n <= 0) return VSDMO_OK;
    grow_history(b, n);
    if (b->p.use_graph && !b->exec) capture(b);
    hipEvent_t e0, e1;
    VSDMO_HIP(hipEventCreate(&e0));
    VSDMO_HIP(hipEventCreate(&e1));
    VSDMO_HIP(hipEventRecord(e0, b->stream));
    for (int i = 0; i < n; ++i) {
        if (b->p.use_graph) VSDMO_HIP(hipGraphLaunch(b->exec, b->stream));
        else record_iteration(b);
    }
    VSDMO_HIP(hipEventRecord(e1, b->stream));
    VSDMO_HIP(hipGetLastError());
    VSDMO_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    VSDMO_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->gpu_ms += ms;
    const size_t nm = b->meshes.size();
    std::vector<float> rows((size_t)n * nm);
    VSDMO_HIP(hipMemcpyAsync(rows.data(), b->hist.p + (size_t)b->hrow * nm, sizeof(float) * rows.size(),
                             hipMemcpyDeviceToHost, b->stream));
    VSDMO_HIP(hipStreamSynchronize(b->stream));
    b->hrow += n;
    b->iterations += n;
    b->searches += (long long)n * (long long)b->itemsMeshOrder.size();
    for (size_t r = 0; r + 1 < b->roundOff.size(); ++r) {
        if (b->learn[0].mode != 0) b->learned += (long long)n * (b->roundMidT[r] - b->roundOff[r]);
        if (b->learn[1].mode != 0) b->learned += (long long)n * (b->roundMid[r] - b->roundMidT[r]);
    }
    vsdmo_status st = VSDMO_OK;
    std::string msg;
    for (int i = 0; i < n; ++i)
        for (size_t m = 0; m < nm; ++m) {
            float q = rows[(size_t)i * nm + m];
            float prev = b->lastMin[m];
            if (b->p.quality_guard && st == VSDMO_OK && q < prev - b->p.guard_tolerance * std::fabs(prev)) {
                st = VSDMO_QUALITY_DECREASED;
                msg = "mesh " + std::to_string(m) + ": minimum quality decreased from " + std::to_string(prev) +
                      " to " + std::to_string(q);
            }
            b->lastMin[m] = q;
        }
    if (rows_out) *rows_out = std::move(rows);
    if (st != VSDMO_OK) return fail_status(st, msg);
    return VSDMO_OK;
}

void restore_positions(vsdmo_batch* b)
{
    size_t n = sizeof(float) * (size_t)b->totalV;
    VSDMO_HIP(hipMemcpyAsync(b->x.p, b->x0.p, n, hipMemcpyDeviceToDevice, b->stream));
    VSDMO_HIP(hipMemcpyAsync(b->y.p, b->y0.p, n, hipMemcpyDeviceToDevice, b->stream));
    VSDMO_HIP(hipMemcpyAsync(b->z.p, b->z0.p, n, hipMemcpyDeviceToDevice, b->stream));
    VSDMO_HIP(hipStreamSynchronize(b->stream));
}

// per-mesh minimum of the current positions (outside the history)
std::vector<float> current_min(vsdmo_batch* b)
{
    BatchQArgs q = quality_args(b);
    q.hcap = 0;   // do not append
    // the reduce kernel advances *hctr; keep it consistent by using a scratch counter
    DevBuf<int> tmp(1);
    q.hctr = tmp.p;
    launch_fill_int(tmp.p, 1, 0, b->stream);
    launch_batch_quality(q, b->stream);
    std::vector<float> m(b->meshes.size());
    b->mmin.download(m.data(), m.size(), b->stream);
    VSDMO_HIP(hipStreamSynchronize(b->stream));
    return m;
}

std::string check_params(const vsdmo_batch_params& p)
{
    if (!(p.nq == 2 || p.nq == 4 || p.nq == 6 || p.nq == 8)) return "nq must be 2, 4, 6 or 8";
    if (p.depth < 1 || p.depth > 16) return "depth must be in [1,16]";
    if (p.waves_per_vertex < 1 || !waves_supported(p.nq, p.waves_per_vertex))
        return "waves_per_vertex must be 1, 2 or 4 and divide nq";
    if (!(p.grid_scale > 0.0f)) return "grid_scale must be > 0";
    if (p.hex_metric != VSDMO_HEX_SHAPE9 && p.hex_metric != VSDMO_HEX_CORNER8) return "unknown hex_metric";
    return "";
}

void apply_params(vsdmo_batch* b, const vsdmo_batch_params& p)
{
    bool reorder = p.shuffle_seed != b->p.shuffle_seed;
    b->p = p;
    b->aff = 1.0f / (float)(p.nq - 1);
    b->shrink = 2.0f / (float)(p.nq - 1);
    drop_graph(b);
    size_grids(b);
    if (reorder) upload_items(b);
}

template <class F>
vsdmo_status guarded_b(F&& f)
{
    try {
        return f();
    } catch (const HipError& e) {
        return fail_status(VSDMO_HIP_ERROR, std::string(hipGetErrorString(e.code)) + " in " + e.where);
    } catch (const StatusError& e) {
        return fail_status(e.status, e.msg);
    } catch (const std::bad_alloc&) {
        return fail_status(VSDMO_HIP_ERROR, "host allocation failed");
    } catch (const std::exception& e) {
        return fail_status(VSDMO_HIP_ERROR, e.what());
    }
}

}  // namespace

extern "C" {

void vsdmo_batch_params_default(vsdmo_batch_params* p)
{
    if (!p) return;
    p->nq = 8;
    p->depth = 3;
    p->waves_per_vertex = 1;
    p->grid_scale = 0.5f;
    p->use_graph = 1;
    p->persistent = 1;
    p->shuffle_seed = 0;
    p->quality_guard = 1;
    p->guard_tolerance = 1e-5f;
    p->hex_metric = VSDMO_HEX_SHAPE9;
}

vsdmo_status vsdmo_batch_create(vsdmo_mesh* const* meshes, int32_t n_meshes, const vsdmo_batch_params* params,
                                vsdmo_batch** out)
{
    return guarded_b([&]() -> vsdmo_status {
        if (!out) return fail_status(VSDMO_INVALID_ARG, "out is NULL");
        *out = nullptr;
        if (!meshes || n_meshes <= 0) return fail_status(VSDMO_INVALID_ARG, "no meshes");
        vsdmo_batch_params p;
        vsdmo_batch_params_default(&p);
        if (params) p = *params;
        std::string perr = check_params(p);
        if (!perr.empty()) return fail_status(VSDMO_INVALID_ARG, perr);
        for (int i = 0; i < n_meshes; ++i) {
            if (!meshes[i]) return fail_status(VSDMO_INVALID_ARG, "mesh " + std::to_string(i) + " is NULL");
            for (int j = 0; j < i; ++j)
                if (meshes[j] == meshes[i]) return fail_status(VSDMO_INVALID_ARG, "a mesh appears twice in the batch");
            const vsdmo_mesh* m = meshes[i];
            if (m->elem == VSDMO_QUAD || (m->elem == VSDMO_TRI && !m->flat))
                return fail_status(VSDMO_NOT_SUPPORTED, "batches hold tet, hex and flat triangle meshes (mesh " +
                                                            std::to_string(i) + ")");
        }
        auto* b = new vsdmo_batch();
        try {
            b->meshes.assign(meshes, meshes + n_meshes);
            VSDMO_HIP(hipStreamCreate(&b->stream));
            hipStream_t s = b->stream;
            int dev = 0;
            VSDMO_HIP(hipGetDevice(&dev));
            VSDMO_HIP(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev));
            for (vsdmo_mesh* m : b->meshes) VSDMO_HIP(hipStreamSynchronize(m->stream));
            // vertex offsets, ring / halfedge sizes
            const int nm = n_meshes;
            b->vOff.assign(nm + 1, 0);
            std::vector<long long> ringBase(nm + 1, 0), hexBase(nm + 1, 0), heBase(nm + 1, 0);
            long long cellInts = 0;
            for (int i = 0; i < nm; ++i) {
                vsdmo_mesh* m = b->meshes[i];
                b->vOff[i + 1] = b->vOff[i] + m->nV;
                long long ringN = (m->elem == VSDMO_TRI) ? 0 : (long long)m->ringVal.n / (m->elem == VSDMO_TET ? 3 : 7);
                ringBase[i + 1] = ringBase[i] + (m->elem == VSDMO_TET ? ringN : 0);
                hexBase[i + 1] = hexBase[i] + (m->elem == VSDMO_HEX ? ringN : 0);
                heBase[i + 1] = heBase[i] + (m->elem == VSDMO_TRI ? (long long)m->nHE : 0);
                if (m->elem == VSDMO_TET) b->maxRingTet = std::max(b->maxRingTet, m->maxRing);
                if (m->elem == VSDMO_HEX) b->maxRingHex = std::max(b->maxRingHex, m->maxRing);
                if (m->elem == VSDMO_TRI) b->maxHE = std::max(b->maxHE, m->maxHE);
                cellInts += (long long)m->nE * m->npe;
            }
            if ((long long)b->vOff[nm] >= (1LL << 30) || ringBase[nm] >= (1LL << 31) / 3 ||
                hexBase[nm] >= (1LL << 31) / 7 || heBase[nm] >= (1LL << 31) || cellInts >= (1LL << 31))
                throw StatusError{ VSDMO_INVALID_ARG, "batch too large for 32-bit indices" };
            if (b->maxRingTet > 1024 || b->maxRingHex > 256 || b->maxHE > 1024)
                throw StatusError{ VSDMO_LDS_OVERFLOW, "one-ring too large for the LDS staging" };
            b->totalV = b->vOff[nm];
            b->ringBaseTet = ringBase;
            b->ringBaseHex = hexBase;
            const size_t V = (size_t)b->totalV;
            b->x.alloc(V); b->y.alloc(V); b->z.alloc(V);
            b->x0.alloc(V); b->y0.alloc(V); b->z0.alloc(V);
            b->rowTet.alloc(V + 1); b->rowHex.alloc(V + 1); b->heRow.alloc(V + 1);
            b->ringTet.alloc((size_t)std::max<long long>(1, 3 * ringBase[nm]));
            b->ringHex.alloc((size_t)std::max<long long>(1, 7 * hexBase[nm]));
            b->heTarget.alloc((size_t)std::max<long long>(1, heBase[nm]));
            b->cells.alloc((size_t)std::max<long long>(1, cellInts));
            // concatenate
            long long cellPos = 0;
            for (int i = 0; i < nm; ++i) {
                vsdmo_mesh* m = b->meshes[i];
                const int vo = b->vOff[i];
                const size_t nv = (size_t)m->nV;
                VSDMO_HIP(hipMemcpyAsync(b->x.p + vo, m->x.p, sizeof(float) * nv, hipMemcpyDeviceToDevice, s));
                VSDMO_HIP(hipMemcpyAsync(b->y.p + vo, m->y.p, sizeof(float) * nv, hipMemcpyDeviceToDevice, s));
                VSDMO_HIP(hipMemcpyAsync(b->z.p + vo, m->z.p, sizeof(float) * nv, hipMemcpyDeviceToDevice, s));
                if (m->elem == VSDMO_TET) {
                    offset_copy(b->rowTet.p + vo, m->ringRow.p, m->nV, (int)ringBase[i], s);
                    offset_copy(b->ringTet.p + 3 * ringBase[i], m->ringVal.p, (long long)m->ringVal.n, vo, s);
                } else {
                    launch_fill_int(b->rowTet.p + vo, m->nV, (int)ringBase[i], s);
                }
                if (m->elem == VSDMO_HEX) {
                    offset_copy(b->rowHex.p + vo, m->ringRow.p, m->nV, (int)hexBase[i], s);
                    offset_copy(b->ringHex.p + 7 * hexBase[i], m->ringVal.p, (long long)m->ringVal.n, vo, s);
                } else {
                    launch_fill_int(b->rowHex.p + vo, m->nV, (int)hexBase[i], s);
                }
                if (m->elem == VSDMO_TRI) {
                    offset_copy(b->heRow.p + vo, m->heRow.p, m->nV, (int)heBase[i], s);
                    offset_copy(b->heTarget.p + heBase[i], m->heTarget.p, m->nHE, vo, s);
                } else {
                    launch_fill_int(b->heRow.p + vo, m->nV, (int)heBase[i], s);
                }
                offset_copy(b->cells.p + cellPos, m->cells.p, (long long)m->nE * m->npe, vo, s);
                cellPos += (long long)m->nE * m->npe;
            }
            launch_fill_int(b->rowTet.p + V, 1, (int)ringBase[nm], s);
            launch_fill_int(b->rowHex.p + V, 1, (int)hexBase[nm], s);
            launch_fill_int(b->heRow.p + V, 1, (int)heBase[nm], s);
            size_t vb = sizeof(float) * V;
            VSDMO_HIP(hipMemcpyAsync(b->x0.p, b->x.p, vb, hipMemcpyDeviceToDevice, s));
            VSDMO_HIP(hipMemcpyAsync(b->y0.p, b->y.p, vb, hipMemcpyDeviceToDevice, s));
            VSDMO_HIP(hipMemcpyAsync(b->z0.p, b->z.p, vb, hipMemcpyDeviceToDevice, s));
            // work items: round r = colour r of every mesh; tet/hex items first
            int nRounds = 0;
            for (vsdmo_mesh* m : b->meshes) {
                const auto& co = (m->elem == VSDMO_TRI) ? m->colOffFree : m->colOffInner;
                nRounds = std::max(nRounds, (int)co.size() - 1);
            }
            b->roundOff.assign(nRounds + 1, 0);
            b->roundMidT.assign(nRounds, 0);
            b->roundMid.assign(nRounds, 0);
            b->itemsPerMesh.assign(nm, 0);
            for (int r = 0; r < nRounds; ++r) {
                // tets, then hexes, then flat triangles (mesh order inside each)
                for (int pass = 0; pass < 3; ++pass) {
                    if (pass == 1) b->roundMidT[r] = (int)b->itemsMeshOrder.size();
                    if (pass == 2) b->roundMid[r] = (int)b->itemsMeshOrder.size();
                    for (int i = 0; i < nm; ++i) {
                        vsdmo_mesh* m = b->meshes[i];
                        const bool tri = (m->elem == VSDMO_TRI);
                        const int want = (m->elem == VSDMO_TET) ? 0 : (m->elem == VSDMO_HEX) ? 1 : 2;
                        if (want != pass) continue;
                        const auto& co = tri ? m->colOffFree : m->colOffInner;
                        if (r + 1 >= (int)co.size()) continue;
                        const int type = (m->elem == VSDMO_TET) ? 0 : (m->elem == VSDMO_HEX) ? 1 : 2;
                        for (int v = co[r]; v < co[r + 1]; ++v)
                            b->itemsMeshOrder.push_back((b->vOff[i] + v) | (type << 30));
                        b->itemsPerMesh[i] += co[r + 1] - co[r];
                    }
                }
                b->roundOff[r + 1] = (int)b->itemsMeshOrder.size();
            }
            b->items.alloc(std::max<size_t>(1, b->itemsMeshOrder.size()));
            b->ctr.alloc((size_t)std::max(1, nRounds));
            // quality chunks of <= 256 * qper elements, never across meshes
            long long nAll = 0;
            for (int i = 0; i < nm; ++i) nAll += b->meshes[i]->nE;
            b->qper = quality_per_thread(nAll);
            const int chunk = 256 * b->qper;
            std::vector<int> ce0{ 0 }, ctype, mc0{ 0 };
            std::vector<long long> cc0;
            long long cpos = 0;
            int epos = 0;
            for (int i = 0; i < nm; ++i) {
                vsdmo_mesh* m = b->meshes[i];
                for (int e = 0; e < m->nE; e += chunk) {
                    int ne = std::min(chunk, m->nE - e);
                    cc0.push_back(cpos + (long long)e * m->npe);
                    ctype.push_back(m->elem);
                    epos += ne;
                    ce0.push_back(epos);
                }
                cpos += (long long)m->nE * m->npe;
                mc0.push_back((int)ctype.size());
                b->nElem.push_back(m->nE);
            }
            b->nChunks = (int)ctype.size();
            b->chunkElem0.alloc(ce0.size()); b->chunkElem0.upload(ce0.data(), ce0.size(), s);
            b->chunkCell0.alloc(std::max<size_t>(1, cc0.size())); b->chunkCell0.upload(cc0.data(), cc0.size(), s);
            b->chunkType.alloc(std::max<size_t>(1, ctype.size())); b->chunkType.upload(ctype.data(), ctype.size(), s);
            b->meshChunk0.alloc(mc0.size()); b->meshChunk0.upload(mc0.data(), mc0.size(), s);
            b->cmin.alloc(std::max(1, b->nChunks)); b->csum.alloc(std::max(1, b->nChunks));
            b->mmin.alloc(nm); b->msum.alloc(nm);
            b->hcap = 64;
            b->hist.alloc((size_t)b->hcap * nm);
            b->hctr.alloc(1);
            launch_fill_int(b->hctr.p, 1, 0, s);
            b->p.shuffle_seed = p.shuffle_seed ^ 1u;   // force the first upload
            apply_params(b, p);
            b->lastMin = current_min(b);
            VSDMO_HIP(hipGetLastError());
        } catch (...) {
            drop_graph(b);
            if (b->stream) (void)hipStreamDestroy(b->stream);
            delete b;
            throw;
        }
        for (vsdmo_mesh* m : b->meshes) m->n_opts++;
        *out = b;
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_batch_iterate(vsdmo_batch* b, int32_t n)
{
    return guarded_b([&]() -> vsdmo_status {
        if (!b) return fail_status(VSDMO_INVALID_ARG, "NULL batch");
        if (n < 0) return fail_status(VSDMO_INVALID_ARG, "negative iteration count");
        return run(b, n, nullptr);
    });
}

vsdmo_status vsdmo_batch_min_quality(vsdmo_batch* b, float* per_mesh)
{
    return guarded_b([&]() -> vsdmo_status {
        if (!b || !per_mesh) return fail_status(VSDMO_INVALID_ARG, "NULL argument");
        std::copy(b->lastMin.begin(), b->lastMin.end(), per_mesh);
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_batch_mean_quality(vsdmo_batch* b, double* per_mesh)
{
    return guarded_b([&]() -> vsdmo_status {
        if (!b || !per_mesh) return fail_status(VSDMO_INVALID_ARG, "NULL argument");
        (void)current_min(b);
        std::vector<double> s(b->meshes.size());
        b->msum.download(s.data(), s.size(), b->stream);
        VSDMO_HIP(hipStreamSynchronize(b->stream));
        for (size_t i = 0; i < s.size(); ++i) per_mesh[i] = s[i] / (double)std::max(1, b->nElem[i]);
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_batch_get_positions(vsdmo_batch* b, int32_t mesh, float* xyz)
{
    return guarded_b([&]() -> vsdmo_status {
        if (!b || !xyz) return fail_status(VSDMO_INVALID_ARG, "NULL argument");
        if (mesh < 0 || mesh >= (int)b->meshes.size()) return fail_status(VSDMO_INVALID_ARG, "mesh index out of range");
        const int vo = b->vOff[mesh], nv = b->vOff[mesh + 1] - vo;
        std::vector<float> X(nv), Y(nv), Z(nv);
        VSDMO_HIP(hipMemcpyAsync(X.data(), b->x.p + vo, sizeof(float) * nv, hipMemcpyDeviceToHost, b->stream));
        VSDMO_HIP(hipMemcpyAsync(Y.data(), b->y.p + vo, sizeof(float) * nv, hipMemcpyDeviceToHost, b->stream));
        VSDMO_HIP(hipMemcpyAsync(Z.data(), b->z.p + vo, sizeof(float) * nv, hipMemcpyDeviceToHost, b->stream));
        VSDMO_HIP(hipStreamSynchronize(b->stream));
        for (int v = 0; v < nv; ++v) { xyz[3 * v] = X[v]; xyz[3 * v + 1] = Y[v]; xyz[3 * v + 2] = Z[v]; }
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_batch_writeback(vsdmo_batch* b)
{
    return guarded_b([&]() -> vsdmo_status {
        if (!b) return fail_status(VSDMO_INVALID_ARG, "NULL batch");
        for (size_t i = 0; i < b->meshes.size(); ++i) {
            vsdmo_mesh* m = b->meshes[i];
            const int vo = b->vOff[i];
            size_t n = sizeof(float) * (size_t)m->nV;
            VSDMO_HIP(hipMemcpyAsync(m->x.p, b->x.p + vo, n, hipMemcpyDeviceToDevice, b->stream));
            VSDMO_HIP(hipMemcpyAsync(m->y.p, b->y.p + vo, n, hipMemcpyDeviceToDevice, b->stream));
            VSDMO_HIP(hipMemcpyAsync(m->z.p, b->z.p + vo, n, hipMemcpyDeviceToDevice, b->stream));
        }
        VSDMO_HIP(hipStreamSynchronize(b->stream));
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_batch_restore(vsdmo_batch* b)
{
    return guarded_b([&]() -> vsdmo_status {
        if (!b) return fail_status(VSDMO_INVALID_ARG, "NULL batch");
        restore_positions(b);
        reset_deferred(b);   // deferrals of the abandoned trajectory
        b->lastMin = current_min(b);
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_batch_get_params(vsdmo_batch* b, vsdmo_batch_params* p)
{
    if (!b || !p) return fail_status(VSDMO_INVALID_ARG, "NULL argument");
    *p = b->p;
    return VSDMO_OK;
}

vsdmo_status vsdmo_batch_set_params(vsdmo_batch* b, const vsdmo_batch_params* p)
{
    return guarded_b([&]() -> vsdmo_status {
        if (!b || !p) return fail_status(VSDMO_INVALID_ARG, "NULL argument");
        std::string perr = check_params(*p);
        if (!perr.empty()) return fail_status(VSDMO_INVALID_ARG, perr);
        if (deferred(b) && !deferred_check(b, p->nq).empty())
            return fail_status(VSDMO_NOT_SUPPORTED, deferred_check(b, p->nq));
        apply_params(b, *p);
        return VSDMO_OK;
    });
}

// Learned proposal for the tet or hex items (a round's windows; the full
// search of the vertices whose window does not improve them follows in the
// round, or in its next iteration under the deferred policy), the
// single-mesh rule of vsdmo_opt_set_proposal over a round's items of that
// type; each type has its own network and window.
// The network runs inside the window kernel, so mode 1 needs every ring of
// the type to fit its LDS tile.  The neighbour table is global: each tet and
// hex mesh's table with the mesh's vertex offset added, built once.
vsdmo_status vsdmo_batch_set_proposal(vsdmo_batch* b, int32_t elem, vsdmo_mlp* mlp, int32_t mode, int32_t depth,
                                      float window_frac)
{
    return guarded_b([&]() -> vsdmo_status {
        if (!b) return fail_status(VSDMO_INVALID_ARG, "NULL batch");
        if (elem != VSDMO_TET && elem != VSDMO_HEX)
            return fail_status(VSDMO_NOT_SUPPORTED, "the learned proposal is for the tet and hex meshes of a batch");
        if (mode < 0 || mode > 3)
            return fail_status(VSDMO_INVALID_ARG,
                               "mode must be 0 (off), 1 (network), 2 (identity) or 3 (network, then identity)");
        if (learn_uses_net(mode) && !mlp) return fail_status(VSDMO_INVALID_ARG, "modes 1 and 3 need a network");
        if (mode != 0 && (depth < 1 || depth > 8 || !(window_frac > 0.0f)))
            return fail_status(VSDMO_INVALID_ARG, "depth must be in [1,8] and window_frac > 0");
        const int t = elem == VSDMO_TET ? 0 : 1;
        const int ring = t == 0 ? b->maxRingTet : b->maxRingHex;
        if (learn_uses_net(mode) && !learned_fused_ok(elem, ring))
            return fail_status(VSDMO_NOT_SUPPORTED, "the batched learned search runs the network in the window "
                                                    "kernel, whose LDS tile of 16 rings does not hold one-rings of " +
                                                        std::to_string(ring) + " elements");
        if (mode == 2 && ring > learned_ring_cap(elem))
            return fail_status(VSDMO_NOT_SUPPORTED, "the learned search holds one-rings of at most " +
                                                        std::to_string(learned_ring_cap(elem)) + " elements");
        if (mode != 0 && deferred(b) && !learned_deferred_ok(elem, b->p.nq, ring))
            return fail_status(VSDMO_NOT_SUPPORTED, "the deferred fallback needs nq 4 or 8 and one-rings of at most "
                                                    "128 tets or 21 hexes");
        hipStream_t s = b->stream;
        VSDMO_HIP(hipStreamSynchronize(s));
        drop_graph(b);   // the recorded launches change
        if (mode != 0) {
            // buffers first: a failure here leaves the previous setting in place
            const int nR = (int)b->grid.size();
            int maxItems = 1;
            for (int r = 0; r < nR; ++r)
                maxItems = std::max({ maxItems, b->roundMidT[r] - b->roundOff[r], b->roundMid[r] - b->roundMidT[r] });
            if ((int)b->lList.n < maxItems) b->lList.alloc(maxItems);
            if (!b->lCnt.p) b->lCnt.alloc(std::max(2 * nR, 1));
            if (!b->lFallback.p) {
                b->lFallback.alloc(1);
                launch_fill_int(b->lFallback.p, 1, 0, s);
            }
            if (learn_uses_net(mode) && !b->nbr.p) {
                std::vector<int> tab((size_t)b->totalV * kNbr, -1);
                for (size_t i = 0; i < b->meshes.size(); ++i) {
                    vsdmo_mesh* m = b->meshes[i];
                    if (m->elem != VSDMO_TET && m->elem != VSDMO_HEX) continue;
                    build_neighbour_table(*m);
                    std::vector<int> loc = m->nbr.to_host(m->stream);
                    const int off = b->vOff[i];
                    int* dst = tab.data() + (size_t)off * kNbr;
                    for (size_t k = 0; k < loc.size(); ++k) dst[k] = loc[k] >= 0 ? loc[k] + off : -1;
                }
                b->nbr.alloc(tab.size());
                b->nbr.upload(tab.data(), tab.size(), s);
                const size_t nm = b->meshes.size();
                b->locTet.alloc((size_t)std::max<long long>(1, kRingLocStrideTet * b->ringBaseTet[nm]));
                b->locHex.alloc((size_t)std::max<long long>(1, kRingLocStrideHex * b->ringBaseHex[nm]));
                for (size_t i = 0; i < nm; ++i) {
                    vsdmo_mesh* m = b->meshes[i];
                    const bool tet = m->elem == VSDMO_TET;
                    if (!tet && m->elem != VSDMO_HEX) continue;
                    const std::vector<long long>& base = tet ? b->ringBaseTet : b->ringBaseHex;
                    const size_t stride = tet ? kRingLocStrideTet : kRingLocStrideHex;
                    const size_t bytes = stride * (size_t)(base[i + 1] - base[i]);   // the mesh's ring elements
                    if (bytes > 0)
                        VSDMO_HIP(hipMemcpyAsync((tet ? b->locTet.p : b->locHex.p) + stride * base[i], m->ringLoc.p,
                                                 bytes, hipMemcpyDeviceToDevice, s));
                }
            }
            if (learn_uses_net(mode)) VSDMO_HIP(hipStreamSynchronize(mlp->stream));   // the weights are final
            VSDMO_HIP(hipStreamSynchronize(s));
        }
        vsdmo_batch::Learn& L = b->learn[t];
        L.mode = mode;
        if (L.mlp) L.mlp->n_users--;
        L.mlp = learn_uses_net(mode) ? mlp : nullptr;
        if (L.mlp) L.mlp->n_users++;
        L.depth = depth;
        L.frac = window_frac;
        reset_deferred(b);   // pending deferrals belong to the previous setting
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_batch_set_fallback(vsdmo_batch* b, int32_t policy)
{
    return guarded_b([&]() -> vsdmo_status {
        if (!b) return fail_status(VSDMO_INVALID_ARG, "NULL batch");
        if (policy != VSDMO_FALLBACK_IMMEDIATE && policy != VSDMO_FALLBACK_DEFERRED)
            return fail_status(VSDMO_INVALID_ARG, "policy must be VSDMO_FALLBACK_IMMEDIATE or VSDMO_FALLBACK_DEFERRED");
        if (policy == VSDMO_FALLBACK_DEFERRED) {
            std::string err = deferred_check(b, b->p.nq);
            if (!err.empty()) return fail_status(VSDMO_NOT_SUPPORTED, err);
            alloc_deferred(b);
        }
        VSDMO_HIP(hipStreamSynchronize(b->stream));
        drop_graph(b);   // the recorded launches change
        b->fallback = policy;
        reset_deferred(b);
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_batch_learned_stats(vsdmo_batch* b, int64_t* visits, int64_t* fallbacks)
{
    return guarded_b([&]() -> vsdmo_status {
        if (!b || !visits || !fallbacks) return fail_status(VSDMO_INVALID_ARG, "NULL argument");
        *visits = b->learned;
        int f = 0;
        if (b->lFallback.p) {
            VSDMO_HIP(hipMemcpyAsync(&f, b->lFallback.p, sizeof(int), hipMemcpyDeviceToHost, b->stream));
            VSDMO_HIP(hipStreamSynchronize(b->stream));
        }
        *fallbacks = f;
        return VSDMO_OK;
    });
}

// Autotuner over (nq, depth, waves per vertex).  Every trial starts from the
// positions at the call and runs `iters` iterations.  waves_per_vertex does
// not change results, so it is tuned on time alone at the current (nq,
// depth); the (nq, depth) trials are accepted when every mesh's final
// minimum quality is at least the (8, 3) run's minus `tol`.  The fastest
// accepted configuration is kept; each trial is one JSON line in log_path.
vsdmo_status vsdmo_batch_autotune(vsdmo_batch* b, int32_t iters, float tol, const char* log_path,
                                  vsdmo_batch_params* chosen)
{
    return guarded_b([&]() -> vsdmo_status {
        if (!b) return fail_status(VSDMO_INVALID_ARG, "NULL batch");
        if (iters <= 0 || !(tol >= 0.0f)) return fail_status(VSDMO_INVALID_ARG, "iters must be > 0 and tol >= 0");
        if (learning(b))
            return fail_status(VSDMO_NOT_SUPPORTED, "the autotuner runs plain searches: set the proposals to mode 0 first");
        const size_t nm = b->meshes.size();
        const size_t V = (size_t)b->totalV;
        // snapshot of the start positions, guard state and counters
        DevBuf<float> sx(V), sy(V), sz(V);
        size_t vb = sizeof(float) * V;
        VSDMO_HIP(hipMemcpyAsync(sx.p, b->x.p, vb, hipMemcpyDeviceToDevice, b->stream));
        VSDMO_HIP(hipMemcpyAsync(sy.p, b->y.p, vb, hipMemcpyDeviceToDevice, b->stream));
        VSDMO_HIP(hipMemcpyAsync(sz.p, b->z.p, vb, hipMemcpyDeviceToDevice, b->stream));
        VSDMO_HIP(hipStreamSynchronize(b->stream));
        const std::vector<float> min0 = b->lastMin;
        const long long s0 = b->searches;
        const int it0 = b->iterations;
        const double ms0 = b->gpu_ms;
        const vsdmo_batch_params start = b->p;
        FILE* log = log_path && *log_path ? std::fopen(log_path, "w") : nullptr;
        if (log_path && *log_path && !log) return fail_status(VSDMO_INVALID_ARG, std::string("cannot open ") + log_path);

        struct Trial { vsdmo_batch_params p; double ms; std::vector<float> q; bool ok; };
        auto trial = [&](const vsdmo_batch_params& p) -> Trial {
            VSDMO_HIP(hipMemcpyAsync(b->x.p, sx.p, vb, hipMemcpyDeviceToDevice, b->stream));
            VSDMO_HIP(hipMemcpyAsync(b->y.p, sy.p, vb, hipMemcpyDeviceToDevice, b->stream));
            VSDMO_HIP(hipMemcpyAsync(b->z.p, sz.p, vb, hipMemcpyDeviceToDevice, b->stream));
            VSDMO_HIP(hipStreamSynchronize(b->stream));
            apply_params(b, p);
            b->lastMin = min0;
            vsdmo_batch_params g = p;
            g.quality_guard = 0;
            b->p.quality_guard = 0;
            // one untimed iteration instantiates the graph and warms the caches
            std::vector<float> rows;
            double before = b->gpu_ms;
            vsdmo_status st = run(b, iters, &rows);
            double ms = b->gpu_ms - before;
            b->p.quality_guard = p.quality_guard;
            Trial t{ p, ms / iters, std::vector<float>(rows.end() - nm, rows.end()), st == VSDMO_OK };
            return t;
        };
        auto log_trial = [&](const Trial& t, const char* stage, double dev, bool accepted) {
            if (!log) return;
            std::fprintf(log,
                         "{\"stage\": \"%s\", \"nq\": %d, \"depth\": %d, \"waves_per_vertex\": %d, \"persistent\": %d, "
                         "\"ms_per_iter\": %.6f, \"iters\": %d, \"min_quality_worst\": %.9g, \"max_drop_vs_ref\": %.9g, "
                         "\"accepted\": %s}\n",
                         stage, t.p.nq, t.p.depth, t.p.waves_per_vertex, t.p.persistent, t.ms, iters,
                         (double)*std::min_element(t.q.begin(), t.q.end()), dev, accepted ? "true" : "false");
        };
        // reference configuration (8, 3): waves per vertex by time
        vsdmo_batch_params ref = start;
        ref.nq = 8; ref.depth = 3;
        Trial best{};
        std::vector<float> qref;
        int bestW = 1;
        int bestP = 1;
        for (int pers : { 1, 0 })
            for (int w : { 1, 2, 4 }) {
                vsdmo_batch_params p = ref;
                p.waves_per_vertex = w;
                p.persistent = pers;
                (void)trial(p);            // warm-up of this shape
                Trial t = trial(p);
                if (qref.empty()) qref = t.q;
                double dev = 0.0;
                for (size_t m = 0; m < nm; ++m) dev = std::max(dev, (double)qref[m] - t.q[m]);
                log_trial(t, "schedule", dev, true);
                if (best.q.empty() || t.ms < best.ms) { best = t; bestW = w; bestP = pers; }
            }
        // grid resolution and depth
        const int cand[][2] = { { 8, 2 }, { 6, 3 }, { 6, 2 }, { 4, 4 }, { 4, 3 }, { 6, 4 } };
        for (const auto& c : cand) {
            vsdmo_batch_params p = ref;
            p.nq = c[0];
            p.depth = c[1];
            p.waves_per_vertex = waves_supported(p.nq, bestW) ? bestW : 1;
            p.persistent = bestP;
            (void)trial(p);
            Trial t = trial(p);
            double dev = 0.0;
            bool ok = t.ok;
            for (size_t m = 0; m < nm; ++m) {
                dev = std::max(dev, (double)qref[m] - t.q[m]);
                if (t.q[m] < qref[m] - tol) ok = false;
            }
            log_trial(t, "grid", dev, ok);
            if (ok && t.ms < best.ms) best = t;
        }
        if (log) {
            std::fprintf(log, "{\"stage\": \"chosen\", \"nq\": %d, \"depth\": %d, \"waves_per_vertex\": %d, "
                              "\"persistent\": %d, \"ms_per_iter\": %.6f, \"tol\": %g}\n",
                         best.p.nq, best.p.depth, best.p.waves_per_vertex, best.p.persistent, best.ms, (double)tol);
            std::fclose(log);
        }
        // back to the start state with the chosen parameters
        VSDMO_HIP(hipMemcpyAsync(b->x.p, sx.p, vb, hipMemcpyDeviceToDevice, b->stream));
        VSDMO_HIP(hipMemcpyAsync(b->y.p, sy.p, vb, hipMemcpyDeviceToDevice, b->stream));
        VSDMO_HIP(hipMemcpyAsync(b->z.p, sz.p, vb, hipMemcpyDeviceToDevice, b->stream));
        VSDMO_HIP(hipStreamSynchronize(b->stream));
        vsdmo_batch_params fin = best.p;
        fin.quality_guard = start.quality_guard;
        fin.guard_tolerance = start.guard_tolerance;
        fin.use_graph = start.use_graph;
        fin.shuffle_seed = start.shuffle_seed;
        apply_params(b, fin);
        b->lastMin = min0;
        b->searches = s0;
        b->iterations = it0;
        b->gpu_ms = ms0;
        if (chosen) *chosen = fin;
        return VSDMO_OK;
    });
}

vsdmo_status vsdmo_batch_get_stats(vsdmo_batch* b, vsdmo_batch_stats* out)
{
    if (!b || !out) return fail_status(VSDMO_INVALID_ARG, "NULL argument");
    out->searches = b->searches;
    out->iterations = b->iterations;
    out->rounds = (int32_t)b->grid.size();
    out->meshes = (int32_t)b->meshes.size();
    out->items_per_iteration = (int64_t)b->itemsMeshOrder.size();
    out->gpu_ms = b->gpu_ms;
    out->max_grid = b->grid.empty() ? 0 : *std::max_element(b->grid.begin(), b->grid.end());
    out->total_vertices = b->totalV;
    return VSDMO_OK;
}

vsdmo_status vsdmo_batch_destroy(vsdmo_batch* b)
{
    if (!b) return VSDMO_OK;
    drop_graph(b);
    if (b->stream) {
        (void)hipStreamSynchronize(b->stream);
        (void)hipStreamDestroy(b->stream);
    }
    for (vsdmo_mesh* m : b->meshes) m->n_opts--;
    for (auto& L : b->learn)
        if (L.mlp) L.mlp->n_users--;
    delete b;
    return VSDMO_OK;
}

}  // extern "C"
