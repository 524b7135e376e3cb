// dmo_kernels.h -- launch interface of the DMO kernels (internal).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace vsdmo {

// Kernel-exact timing.  While a KernelEvents is installed in the calling
// thread's slot (the optimiser's timed launches), launch_kernel records
// `start` with the first kernel's dispatch and `stop` with every kernel's
// completion (hipExtLaunchKernel's events: the kernels' own begin and end,
// without the stream's dispatch latency around them); otherwise it is a
// plain launch.
struct KernelEvents {
    hipEvent_t start = nullptr, stop = nullptr;
    bool started = false;
};
KernelEvents*& kernel_events();
template <typename... Params, typename... Args>
inline void launch_kernel(void (*k)(Params...), dim3 grid, dim3 block, size_t lds, hipStream_t s, Args... args)
{
    if (KernelEvents* e = kernel_events()) {
        hipExtLaunchKernelGGL(k, grid, block, (uint32_t)lds, s, e->started ? nullptr : e->start, e->stop, 0u,
                              args...);
        e->started = true;
    } else {
        k<<<grid, block, lds, s>>>(args...);
    }
}

struct SearchArgs {
    const float* __restrict__ x;
    const float* __restrict__ y;
    const float* __restrict__ z;
    float* xw;
    float* yw;
    float* zw;
    const int* __restrict__ ringRow;   // volume one-ring CSR
    const int* __restrict__ ringVal;
    const int* __restrict__ heRow;     // surface halfedge CSR (flat tri)
    const int* __restrict__ heTarget;
    int cStart, cEnd;                  // colour class [cStart, cEnd)
    float gridScale, shrink, aff;
    int depth;
    int extentMin;
    int hexCorner8;                    // 1: hex metric over the 8 corner sub-tets (vsdmo_hex_metric)
};

// element kinds of the quality kernels beyond vsdmo_elem: triangles of a
// surface mesh on the surface path (unsigned 3D mean ratio)
constexpr int kElemTri3 = 4;

struct QualityArgs {
    const float* x;
    const float* y;
    const float* z;
    const int* cells;
    int nE, npe, elem;
    float* q;
    int hexCorner8;
};

// surface-vertex state and mesh views for the surface kernels
struct SurfArgs {
    const float* x;
    const float* y;
    const float* z;
    float* xw;
    float* yw;
    float* zw;
    const float* sx;        // positions snapshot at the start of a surface colour pass
    const float* sy;
    const float* sz;
    const int* ringRow;
    const int* ringVal;
    const int* heRow;
    const int* heTarget;
    const int* heFace;
    const int* heNext;
    const int* faces;
    const uint8_t* isFeature;
    const uint8_t* isB1D;
    int nS, nFree, nFeature, nFaces, fpe, flat;
    float cos160;
    // owned by the optimiser
    float* pinit;     // 3*nS
    float* fnorm;     // 3*nFaces
    float* vnorm;     // 3*nS vertex normals
    float* ls;        // 16*nS
    float* ls1;       // 10*nS
    float* fs;        // 16*nSectors
    int* nn;          // nS
    int* frow;        // nFeature+1
    int* table;       // 20*nFeature
    uint8_t* opt1d;   // nS
    float gridScale, shrink, aff;
    int depth;
    int hexCorner8;   // hex ring scores over the 8 corner sub-tets
    int ring2;        // quadric fits over {v}, one-ring and 2-ring
    // error bits (kSurfErr*), or-ed by the kernels; the host reads them with
    // the quality after every iteration
    int* err;
};
// a feature neighbour's projection found no sector for the halfedge nearest
// to the vertex (table entry missing or past the vertex's sectors)
constexpr int kSurfErrNoSector = 1;
// a feature vertex has neighbours after its last feature neighbour on an
// open fan: they belong to no sector
constexpr int kSurfErrUnsectored = 2;
// a surface vertex's {v} + one-ring + 2-ring holds more than kMaxFitPoints
// points (ring-2 fits only; reported by the init, which then fails)
constexpr int kSurfErrFitPoints = 4;
constexpr int kMaxFitPoints = 64;

size_t surf_lds_bytes(int elem, int maxRing);
void surf_init_launch(const SurfArgs& a, hipStream_t s);
void surf_sector_count_launch(const SurfArgs& a, int* cnt, hipStream_t s);
void surf_sector_launch(const SurfArgs& a, int* secN, int* secCnt, int nSec, hipStream_t s);
void surf_nn_launch(const SurfArgs& a, hipStream_t s);
void surf_search2d_launch(const SurfArgs& a, int cStart, int cEnd, int nq, int tet, size_t lds, hipStream_t s);
void surf_search1d_launch(const SurfArgs& a, int cStart, int cEnd, int nq, int tet, size_t lds, hipStream_t s);
void surf_flat1d_launch(const SurfArgs& a, int cStart, int cEnd, int nq, hipStream_t s);
void surf_points_launch(const SurfArgs& a, int vid, int nu, int nv, int mode, int sid, float* out, hipStream_t s);
// surface meshes (tri / quad, chosen by a.fpe): free-vertex 2D and
// feature-vertex 1D searches of one colour class, normals from the current
// positions, Laplace smoothing of one colour class
void surf_fan2d_launch(const SurfArgs& a, int cStart, int cEnd, int nq, hipStream_t s);
void surf_fan1d_launch(const SurfArgs& a, int cStart, int cEnd, int nq, hipStream_t s);
void surf_normals_launch(const SurfArgs& a, hipStream_t s);
// snapshot of the surface positions into a.sx / a.sy / a.sz
void snapshot_launch(const SurfArgs& a, hipStream_t s);
void laplace_launch(float* x, float* y, float* z, const int* heRow, const int* heTarget, int lo, int hi, hipStream_t s);

// chunked per-mesh qualities of a batch
// elements per thread of the quality kernels for a launch over nElements
// (chunks of 256 * that many elements per block)
int quality_per_thread(long long nElements);

struct BatchQArgs {
    const float* x;
    const float* y;
    const float* z;
    const int* cells;          // global vertex ids, npe per element
    const int* chunkElem0;     // nChunks+1: first element of each chunk
    const long long* chunkCell0;  // nChunks: first cell index (ints) of each chunk
    const int* chunkType;      // nChunks: vsdmo_elem
    const int* meshChunk0;     // nMeshes+1
    float* cmin;
    double* csum;
    float* mmin;
    double* msum;
    float* hist;               // hcap rows of nMeshes minima
    int* hctr;
    int nChunks, nMeshes, hcap;
    int hexCorner8;
    int per;                   // elements per thread: chunks of <= 256 * per elements
};

size_t search_lds_bytes(int kind, int maxRing);
bool waves_supported(int nq, int w);
void launch_search(int kind, int nq, int w, const SearchArgs& a, size_t lds, hipStream_t s);
// tet / hex class in one launch: one vertex per wave for all but the last
// `tail` vertices, which get several waves each (waves_per_vertex = 0)
bool split_supported(int kind, int nq, int ringCap);
void launch_search_split(int kind, int nq, const SearchArgs& a, int tail, int ringCap, hipStream_t s);
// one colour round of a batch, persistent: grid blocks take the tet / hex
// items [lo, mid) from the atomic counter *ctr and stride over the flat
// items [mid, hi) (the one-block-per-item schedule is launch_batch_typed)
void launch_batch_round(int nq, int w, const SearchArgs& a, const int* rowHex, const int* valHex, const int* items,
                        int lo, int mid, int hi, int* ctr, int grid, size_t lds, hipStream_t s);
int batch_round_blocks_per_cu(int nq, int w, size_t lds);
void launch_fill_int(int* p, int n, int v, hipStream_t s);
// one block per item of items[lo, hi), all of element type `type` (0 tet,
// 1 hex, 2 flat tri): that type's search body alone
void launch_batch_typed(int nq, int w, const SearchArgs& a, const int* rowHex, const int* valHex, const int* items,
                        int lo, int hi, int type, size_t lds, hipStream_t s);
// zero the n counters cnt + (*itDev & 1) * n (device-side iteration parity)
void launch_parity_zero(int* cnt, int n, const int* itDev, hipStream_t s);
// *p += v on the device (one thread)
void launch_add_int(int* p, int v, hipStream_t s);
// proposal modes of the learned search: 1 the network's proposal, 2 the
// vertex itself (the same window without the network), 3 the network's
// proposal and, where its window finds nothing better, a window a quarter of
// its size around the vertex itself before the full search
__host__ __device__ constexpr bool learn_uses_net(int mode) { return mode == 1 || mode == 3; }
// where the learned search finds its proposals (modes 1, 3): the network's
// packed parameters (wb bf16, w fp32) and the neighbour table, with which the
// window kernel runs the network itself when learned_fused_ok(); otherwise
// Y, L of a separate inference launch, indexed by the vertex's position in
// the class
struct LearnNet {
    const float* Y = nullptr;
    const float* L = nullptr;
    const uint16_t* wb = nullptr;
    const float* w = nullptr;
    const int* nbr = nullptr;
    // vertex ids of the class positions through a batch round's work items
    // (id in the low 30 bits); null: positions a.cStart + i of one mesh
    const int* items = nullptr;
    // positions of the one-ring corners in the neighbour table (ringLoc of
    // the mesh, in the index space of the SearchArgs ring), with which the
    // window kernel takes corner positions from the network's features
    const uint8_t* loc = nullptr;
};
// deferred fallback of one class pass: the previous iteration's list of the
// class (listIn[0 .. *cntIn)), the per-vertex deferral epochs, the iteration
struct LearnDefer {
    const int* listIn;
    const int* cntIn;
    int* epoch;
    int it;
    // iteration in device memory (graph replay): listIn / cntIn / it are
    // ignored, list / cnt are parity-0 buffers with the parity-1 ones at
    // + listStride / cntStride (see launch_parity_zero, launch_add_int)
    const int* itDev = nullptr;
    int listStride = 0, cntStride = 0;
};
// learned-proposal search of one colour class [a.cStart, a.cEnd) (tet / hex):
// the proposal window (4 vertices per wave) appends the vertices whose
// window finds nothing better than their start to list (*cnt zero on entry);
// def null: a second launch runs their full search now; def set: they keep
// their position and take the full search in the class's next pass, which
// runs def->listIn beside the windows.  *fallbacks counts the full searches
// run; ringCap >= the largest one-ring
void launch_search_learned(int kind, int nq, const SearchArgs& a, const LearnNet& net, int mode, int depthL,
                           float frac, int* fallbacks, int* list, int* cnt, int ringCap, const LearnDefer* def,
                           hipStream_t s);
// largest one-ring (elements) the learned search supports
int learned_ring_cap(int kind);
// whether the deferred fallback policy runs with this candidate grid and ring size
bool learned_deferred_ok(int kind, int nq, int ringCap);
// whether the window kernel can run the network itself for this ring size
bool learned_fused_ok(int kind, int ringCap);
void launch_batch_quality(const BatchQArgs& a, hipStream_t s);
size_t batch_lds_bytes(int maxRingTet, int maxRingHex, int maxHETri);
void launch_quality(const QualityArgs& a, float* pmin, double* psum, float* omin, double* osum, hipStream_t s);
void launch_histogram(const float* q, int n, int ncols, int* bins, hipStream_t s);

}  // namespace vsdmo
