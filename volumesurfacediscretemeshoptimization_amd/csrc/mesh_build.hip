// mesh_build.hip -- GPU construction of the DMO mesh structures.
//
// Steps (reference files under VSDMO/src/mesh):
//   boundary faces    MeshTetGPU.cu:178-200 / MeshHexGPU.cu:141-163,
//                     MeshGPUCommon.cu:25-94 (rotate smallest-first, sort,
//                     cancel opposite-orientation twins), :1376-1430
//   [surface|inner]   MeshTetGPU.cu:241-262 (stable partition)
//   colouring         MeshTetGPU.cu:298-392, VertexColoring.cu:76-125
//   colour sort       MeshTetGPU.cu:264-295 (stable sort of a range by colour)
//   one-ring CSR      MeshTetGPU.cu:113-142, MeshHexGPU.cu:116-139,
//                     MeshGPUCommon.cu:1313-1369
// Every step is deterministic: sorts are counting sorts followed by a per-row
// sort on the original index, so the output never depends on atomic order.
#include <algorithm>
#include <map>
#include <numeric>

#include "vsdmo_internal.h"

namespace vsdmo {

constexpr int TB = 256;
static inline int nblk(long long n, int tb = TB) { return (int)std::max<long long>(1, (n + tb - 1) / tb); }

// ------------------------------------------------------------------ scan
constexpr int SCAN_TB = 256;
constexpr int SCAN_PER = 4;
constexpr int SCAN_CHUNK = SCAN_TB * SCAN_PER;

__global__ void k_scan_chunk(const int* __restrict__ in, int* __restrict__ out, int* __restrict__ sums, int n)
{
    __shared__ int s[SCAN_TB];
    const int base = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_PER;
    int v[SCAN_PER];
    int t = 0;
#pragma unroll
    for (int i = 0; i < SCAN_PER; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        t += v[i];
    }
    s[threadIdx.x] = t;
    __syncthreads();
    for (int off = 1; off < SCAN_TB; off <<= 1) {
        int a = (threadIdx.x >= off) ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += a;
        __syncthreads();
    }
    int run = s[threadIdx.x] - t;  // exclusive prefix of this thread
#pragma unroll
    for (int i = 0; i < SCAN_PER; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == SCAN_TB - 1 && sums) sums[blockIdx.x] = s[threadIdx.x];
}

__global__ void k_scan_add(int* __restrict__ out, const int* __restrict__ sums, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += sums[i / SCAN_CHUNK];
}

// Read-back slots of the build steps, per host thread and device: a few
// device ints, the partial sums of the scans (grown on demand) and pinned
// host ints, so that a scalar comes back with one small DMA and one
// synchronisation.  Allocated on first use and kept for the life of the
// process (freeing them at thread exit could run after the HIP runtime has
// shut down).
struct BuildSlots {
    int* dev = nullptr;      // kSlots ints
    int* host = nullptr;     // kSlots pinned ints
    int* sums = nullptr;     // scan partial sums
    size_t sumsCap = 0;
};
constexpr int kSlots = 16;

static BuildSlots& build_slots(size_t sumsNeed = 0)
{
    thread_local std::map<int, BuildSlots> slots;
    int d = 0;
    VSDMO_HIP(hipGetDevice(&d));
    BuildSlots& b = slots[d];
    if (!b.dev) {
        VSDMO_HIP(hipMalloc(&b.dev, kSlots * sizeof(int)));
        VSDMO_HIP(hipHostMalloc(reinterpret_cast<void**>(&b.host), kSlots * sizeof(int), hipHostMallocDefault));
    }
    if (sumsNeed > b.sumsCap) {
        if (b.sums) VSDMO_HIP(hipFree(b.sums));
        b.sums = nullptr;
        const size_t cap = std::max<size_t>(sumsNeed, 4096);
        VSDMO_HIP(hipMalloc(&b.sums, cap * sizeof(int)));
        b.sumsCap = cap;
    }
    return b;
}

// one int of device memory, read back through the pinned slot
int read_int(const int* d, hipStream_t s)
{
    BuildSlots& b = build_slots();
    VSDMO_HIP(hipMemcpyAsync(b.host, d, sizeof(int), hipMemcpyDeviceToHost, s));
    VSDMO_HIP(hipStreamSynchronize(s));
    return b.host[0];
}

__global__ void k_copy_int(const int* __restrict__ src, int* __restrict__ dst) { *dst = *src; }
__global__ void k_scan_total(const int* __restrict__ out, int n, const int* __restrict__ last, int* __restrict__ total)
{
    *total = out[n - 1] + *last;
}

// partial-sum ints the scan of n values needs over all its levels
static size_t scan_sums_need(int n)
{
    size_t need = 0;
    for (long long m = n; m > SCAN_CHUNK;) {
        m = (m + SCAN_CHUNK - 1) / SCAN_CHUNK;
        need += (size_t)m;
    }
    return need;
}

// exclusive scan on the device, partial sums of each level in sums
static void scan_levels(const int* in, int* out, int n, int* sums, hipStream_t s)
{
    const int nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
    if (nb == 1) {
        k_scan_chunk<<<1, SCAN_TB, 0, s>>>(in, out, nullptr, n);
        return;
    }
    k_scan_chunk<<<nb, SCAN_TB, 0, s>>>(in, out, sums, n);
    scan_levels(sums, sums, nb, sums + nb, s);
    k_scan_add<<<nblk(n), TB, 0, s>>>(out, sums, n);
}

// exclusive scan of n ints (out may alias in); returns the total
int exclusive_scan(const int* in, int* out, int n, hipStream_t s)
{
    if (n <= 0) return 0;
    BuildSlots& b = build_slots(scan_sums_need(n));
    k_copy_int<<<1, 1, 0, s>>>(in + (n - 1), b.dev);   // before an in-place scan overwrites it
    scan_levels(in, out, n, b.sums, s);
    k_scan_total<<<1, 1, 0, s>>>(out, n, b.dev, b.dev + 1);
    VSDMO_HIP(hipGetLastError());
    return read_int(b.dev + 1, s);
}

// ------------------------------------------------------ counting-sort CSR
__global__ void k_count_keys(const int* __restrict__ key, int n, int* __restrict__ cnt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&cnt[key[i]], 1);
}

__global__ void k_scatter_keys(const int* __restrict__ key, int n, int* __restrict__ cursor, int* __restrict__ items)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) items[atomicAdd(&cursor[key[i]], 1)] = i;
}

// sort each row's item indices ascending (restores the stable order)
__global__ void k_sort_rows(const int* __restrict__ row, int nrows, int* __restrict__ items)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    int s = row[r], t = row[r + 1];
    for (int i = s + 1; i < t; ++i) {
        int x = items[i], j = i - 1;
        while (j >= s && items[j] > x) { items[j + 1] = items[j]; --j; }
        items[j + 1] = x;
    }
}

// stable CSR of items 0..n-1 grouped by key in [0,nk): row[nk+1], items[n]
void stable_csr(const int* key, int n, int nk, DevBuf<int>& row, DevBuf<int>& items, hipStream_t s)
{
    row.alloc((size_t)nk + 1);
    items.alloc((size_t)std::max(n, 1));
    VSDMO_HIP(hipMemsetAsync(row.p, 0, sizeof(int) * ((size_t)nk + 1), s));
    if (n > 0) k_count_keys<<<nblk(n), TB, 0, s>>>(key, n, row.p);
    exclusive_scan(row.p, row.p, nk + 1, s);
    DevBuf<int> cursor(nk > 0 ? nk : 1);
    if (nk > 0) VSDMO_HIP(hipMemcpyAsync(cursor.p, row.p, sizeof(int) * nk, hipMemcpyDeviceToDevice, s));
    if (n > 0) k_scatter_keys<<<nblk(n), TB, 0, s>>>(key, n, cursor.p, items.p);
    if (nk > 0) k_sort_rows<<<nblk(nk), TB, 0, s>>>(row.p, nk, items.p);
    VSDMO_HIP(hipGetLastError());
    VSDMO_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------ boundary faces
__constant__ int c_tetFaces[4][3] = { { 2, 1, 0 }, { 0, 1, 3 }, { 1, 2, 3 }, { 0, 3, 2 } };
__constant__ int c_hexFaces[6][4] = { { 0, 1, 2, 3 }, { 1, 5, 6, 2 }, { 3, 2, 6, 7 },
                                      { 4, 0, 3, 7 }, { 1, 0, 4, 5 }, { 5, 4, 7, 6 } };

// all element faces, rotated smallest-first; fk: 4 ints per face (quad slot 3
// unused for tris), v0: first vertex
__global__ void k_emit_faces(const int* __restrict__ cells, int nE, int tet, int* __restrict__ fk, int* __restrict__ v0)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    int nf = tet ? 4 : 6, k = tet ? 3 : 4, npe = tet ? 4 : 8;
    for (int f = 0; f < nf; ++f) {
        int a[4];
        for (int i = 0; i < k; ++i) a[i] = cells[(size_t)e * npe + (tet ? c_tetFaces[f][i] : c_hexFaces[f][i])];
        int m = 0;
        for (int i = 1; i < k; ++i) if (a[i] < a[m]) m = i;
        int id = e * nf + f;
        for (int i = 0; i < 4; ++i) fk[(size_t)id * 4 + i] = (i < k) ? a[(m + i) % k] : 0;
        v0[id] = a[m];
    }
}

__device__ __forceinline__ bool face_less(const int* fk, int a, int b, int k)
{
    for (int i = 1; i < k; ++i) {
        int x = fk[(size_t)a * 4 + i], y = fk[(size_t)b * 4 + i];
        if (x != y) return x < y;
    }
    return a < b;
}

// per bucket (same first vertex): sort by (v1, v2[, v3], emit index), then
// cancel each (a,b,c) with b<c (quads (a,b,c,d) with b<d) against the first
// (a,c,b) (quads (a,d,c,b)) after it (reference k_markDuplicateTris/Quads)
__global__ void k_bucket_faces(const int* __restrict__ row, int nb, int* __restrict__ items,
                               const int* __restrict__ fk, int k, uint8_t* __restrict__ dead)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    int s = row[b], t = row[b + 1];
    for (int i = s + 1; i < t; ++i) {
        int x = items[i], j = i - 1;
        while (j >= s && face_less(fk, x, items[j], k)) { items[j + 1] = items[j]; --j; }
        items[j + 1] = x;
    }
    for (int i = s; i < t; ++i) dead[i] = 0;
    for (int i = s; i < t; ++i) {
        const int* a = fk + (size_t)items[i] * 4;
        bool search = (k == 3) ? (a[1] < a[2]) : (a[1] < a[3]);
        if (!search) continue;
        for (int j = i + 1; j < t; ++j) {
            const int* c = fk + (size_t)items[j] * 4;
            bool match = (k == 3) ? (c[1] == a[2] && c[2] == a[1])
                                  : (c[1] == a[3] && c[2] == a[2] && c[3] == a[1]);
            if (match) { dead[i] = 1; dead[j] = 1; break; }
        }
    }
}

__global__ void k_alive(const uint8_t* __restrict__ dead, int n, int* __restrict__ flag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = dead[i] ? 0 : 1;
}

__global__ void k_compact_faces(const int* __restrict__ items, const uint8_t* __restrict__ dead,
                                const int* __restrict__ pos, int n, const int* __restrict__ fk, int k,
                                int* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || dead[i]) return;
    const int* a = fk + (size_t)items[i] * 4;
    for (int j = 0; j < k; ++j) out[(size_t)pos[i] * k + j] = a[j];
}

__global__ void k_mark_vertices(const int* __restrict__ ids, int n, int* __restrict__ flag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[ids[i]] = 1;
}

// ------------------------------------------------------------ permutations
__global__ void k_partition_ids(const int* __restrict__ flag, const int* __restrict__ scan, int n, int nFirst,
                                int firstVal, int* __restrict__ newid)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    // scan counts flag==1 before v
    int ones = scan[v];
    int isOne = flag[v];
    bool first = (isOne == firstVal);
    int before = firstVal ? ones : (v - ones);
    newid[v] = first ? before : nFirst + (v - before);
}

__global__ void k_permute_f(const int* __restrict__ newid, int n, const float* __restrict__ a, float* __restrict__ b)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) b[newid[v]] = a[v];
}
__global__ void k_permute_i(const int* __restrict__ newid, int n, const int* __restrict__ a, int* __restrict__ b)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) b[newid[v]] = a[v];
}
__global__ void k_permute_u8(const int* __restrict__ newid, int n, const uint8_t* __restrict__ a, uint8_t* __restrict__ b)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) b[newid[v]] = a[v];
}
__global__ void k_remap(int* __restrict__ ids, long long n, const int* __restrict__ newid)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ids[i] = newid[ids[i]];
}
__global__ void k_invert(const int* __restrict__ newOfOld, int n, int* __restrict__ oldOfNew)
{
    int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o < n) oldOfNew[newOfOld[o]] = o;
}

// Applies newid (a permutation of all nV current ids) to positions, elements,
// faces, the composed renumbering and optional per-vertex arrays.
struct Renumber {
    vsdmo_mesh& m;
    DevBuf<int> oldOfCur;   // current id -> input id
    explicit Renumber(vsdmo_mesh& mesh) : m(mesh) {}

    void apply(const int* newid, std::vector<DevBuf<uint8_t>*> u8 = {}, std::vector<DevBuf<int>*> i32 = {})
    {
        hipStream_t s = m.stream;
        int n = m.nV;
        DevBuf<float> t(n);
        k_permute_f<<<nblk(n), TB, 0, s>>>(newid, n, m.x.p, t.p);
        std::swap(m.x, t);
        k_permute_f<<<nblk(n), TB, 0, s>>>(newid, n, m.y.p, t.p);
        std::swap(m.y, t);
        k_permute_f<<<nblk(n), TB, 0, s>>>(newid, n, m.z.p, t.p);
        std::swap(m.z, t);
        long long nc = (long long)m.nE * m.npe;
        k_remap<<<nblk(nc), TB, 0, s>>>(m.cells.p, nc, newid);
        long long nf = (long long)m.nFaces * m.fpe;
        if (nf > 0 && m.faces.p != m.cells.p) k_remap<<<nblk(nf), TB, 0, s>>>(m.faces.p, nf, newid);
        k_remap<<<nblk(n), TB, 0, s>>>(m.newOfOld.p, n, newid);
        for (auto* b : u8) {
            DevBuf<uint8_t> tb(b->n);
            int nn = (int)b->n;
            k_permute_u8<<<nblk(nn), TB, 0, s>>>(newid, nn, b->p, tb.p);
            std::swap(*b, tb);
        }
        for (auto* b : i32) {
            DevBuf<int> tb(b->n);
            int nn = (int)b->n;
            k_permute_i<<<nblk(nn), TB, 0, s>>>(newid, nn, b->p, tb.p);
            std::swap(*b, tb);
        }
        oldOfCur.alloc(n);
        k_invert<<<nblk(n), TB, 0, s>>>(m.newOfOld.p, n, oldOfCur.p);
        VSDMO_HIP(hipGetLastError());
        VSDMO_HIP(hipStreamSynchronize(s));
    }
};

// ------------------------------------------------------------- colouring
// Neighbour walk of current vertex v: incident elements from the incidence
// CSR (input ids), other corners through the current element array.
struct Incidence {
    const int* row;   // nV+1 over input ids
    const int* val;   // element*8 + corner
    const int* oldOfCur;
    const int* cells;
    int npe;
};

// Greedy first-fit in index order by Jones-Plassmann rounds with priority n-i:
// a vertex is coloured once all its lower in-range neighbours are, with the
// smallest colour none of them uses.  The colour of a vertex depends only on
// the final colours of its lower neighbours, so the result equals the
// sequential greedy colouring whatever the round in which a vertex is reached.
__global__ void k_color_greedy_round(Incidence inc, int lo, int hi, int* __restrict__ col, int* __restrict__ left)
{
    const int v = lo + blockIdx.x * blockDim.x + threadIdx.x;
    volatile int* vc = col;
    bool wait = false;
    if (v < hi && vc[v - lo] < 0) {
        const int o = inc.oldOfCur[v];
        uint64_t mask = 0;
        bool big = false;
        for (int k = inc.row[o]; k < inc.row[o + 1] && !wait; ++k) {
            const int e = inc.val[k] >> 3, c = inc.val[k] & 7;
            for (int i = 0; i < inc.npe; ++i) {
                if (i == c) continue;
                const int u = inc.cells[(size_t)e * inc.npe + i];
                if (u < lo || u >= v) continue;
                const int cu = vc[u - lo];
                if (cu < 0) { wait = true; break; }
                if (cu < 64) mask |= (1ull << cu); else big = true;
            }
        }
        if (!wait) {
            int cc = __builtin_ctzll(~mask);
            if (big && cc == 64) {
                // more than 64 colours in use around v: linear mex search
                for (cc = 64;; ++cc) {
                    bool used = false;
                    for (int k = inc.row[o]; k < inc.row[o + 1] && !used; ++k) {
                        const int e = inc.val[k] >> 3, c = inc.val[k] & 7;
                        for (int i = 0; i < inc.npe; ++i) {
                            if (i == c) continue;
                            const int u = inc.cells[(size_t)e * inc.npe + i];
                            if (u >= lo && u < v && vc[u - lo] == cc) { used = true; break; }
                        }
                    }
                    if (!used) break;
                }
            }
            vc[v - lo] = cc;
        }
    }
    // one plain store per wave with a waiting vertex (no atomic per vertex)
    if (__ballot(wait) != 0 && (threadIdx.x & 63) == 0) *left = 1;
}

// Lower neighbours of each vertex v of [lo, hi): the in-range ids u < v of
// the corners of v's elements, as offsets from lo (an id appears once per
// element shared with v), for the single-pass colouring below
__global__ void k_lower_count(Incidence inc, int lo, int hi, int* __restrict__ cnt)
{
    const int v = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= hi) return;
    const int o = inc.oldOfCur[v];
    int c = 0;
    for (int k = inc.row[o]; k < inc.row[o + 1]; ++k) {
        const int e = inc.val[k] >> 3, cr = inc.val[k] & 7;
        for (int i = 0; i < inc.npe; ++i) {
            if (i == cr) continue;
            const int u = inc.cells[(size_t)e * inc.npe + i];
            c += (u >= lo && u < v) ? 1 : 0;
        }
    }
    cnt[v - lo] = c;
}

__global__ void k_lower_fill(Incidence inc, int lo, int hi, const int* __restrict__ row, int* __restrict__ val)
{
    const int v = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= hi) return;
    const int o = inc.oldOfCur[v];
    int w = row[v - lo];
    for (int k = inc.row[o]; k < inc.row[o + 1]; ++k) {
        const int e = inc.val[k] >> 3, cr = inc.val[k] & 7;
        for (int i = 0; i < inc.npe; ++i) {
            if (i == cr) continue;
            const int u = inc.cells[(size_t)e * inc.npe + i];
            if (u >= lo && u < v) val[w++] = u - lo;
        }
    }
}

// The greedy first-fit colouring of k_color_greedy_round in one launch: each
// vertex polls the colours of its lower neighbours until all are set, then
// takes the smallest colour none of them uses.  A colour depends only on the
// final colours of the lower neighbours, so the result is the sequential
// greedy colouring.  A vertex waits only for lower vertices: those of its own
// wave (the wave loops until all its lanes are done, so the lowest waiting
// lane always finds its neighbours set) and those of lower workgroups, which
// are dispatched before it.  The wait is bounded all the same: a vertex that
// has polled spinLimit times stops uncoloured and raises *gaveUp, and the
// round kernel completes the colouring with the same colours.  Colours are
// loaded and stored at device scope, so that a poll sees a store made on
// another XCD.
__global__ __launch_bounds__(256) void k_color_greedy_sweep(const int* __restrict__ row, const int* __restrict__ val,
                                                           int n, int* __restrict__ col, int spinLimit,
                                                           int* __restrict__ gaveUp)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    bool todo = v < n;
    const int r1 = todo ? row[v + 1] : 0;
    int k = todo ? row[v] : 0;        // lower neighbours [row[v], k) are coloured and in mask
    uint64_t mask = 0;
    bool big = false;
    int polls = 0;
    constexpr int B = 8;   // list entries loaded together (their latencies overlap)
    while (__ballot(todo) != 0) {
        if (todo) {
            bool blocked = false;
            while (k < r1 && !blocked) {
                int id[B], cu[B];
                // entries past the list re-read its last one (k < r1), so every
                // load is unconditional and in bounds; their values are ignored
#pragma unroll
                for (int j = 0; j < B; ++j) id[j] = val[min(k + j, r1 - 1)];
#pragma unroll
                for (int j = 0; j < B; ++j) cu[j] = __hip_atomic_load(col + id[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                int done = 0;   // entries taken in order before the first uncoloured one
#pragma unroll
                for (int j = 0; j < B; ++j) {
                    if (blocked || k + j >= r1) continue;
                    if (cu[j] < 0) { blocked = true; continue; }
                    if (cu[j] < 64) mask |= 1ull << cu[j]; else big = true;
                    ++done;
                }
                k += done;
            }
            if (k == r1) {
                int cc = (~mask == 0ull) ? 64 : __builtin_ctzll(~mask);
                if (big && cc == 64) {
                    // more than 64 colours around v: linear mex search
                    for (;; ++cc) {
                        bool used = false;
                        for (int j = row[v]; j < r1 && !used; ++j)
                            used = __hip_atomic_load(col + val[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == cc;
                        if (!used) break;
                    }
                }
                __hip_atomic_store(col + v, cc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                todo = false;
            } else if (++polls >= spinLimit) {
                __hip_atomic_store(gaveUp, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                todo = false;
            }
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Reference color_jpl round c (VertexColoring.cu:76-105): v takes colour c iff
// every in-range neighbour that is uncoloured or coloured c in this round has
// a lower priority n-u, i.e. a larger index.
__global__ void k_color_jpl_round(Incidence inc, int lo, int hi, int c, int* __restrict__ col, int* __restrict__ left)
{
    const int v = lo + blockIdx.x * blockDim.x + threadIdx.x;
    volatile int* vc = col;
    bool wait = false;
    if (v < hi && vc[v - lo] == -1) {
        const int o = inc.oldOfCur[v];
        bool ok = true;
        for (int k = inc.row[o]; k < inc.row[o + 1] && ok; ++k) {
            const int e = inc.val[k] >> 3, cr = inc.val[k] & 7;
            for (int i = 0; i < inc.npe; ++i) {
                if (i == cr) continue;
                const int u = inc.cells[(size_t)e * inc.npe + i];
                if (u < lo || u >= hi || u == v) continue;
                const int cu = vc[u - lo];
                if ((cu != -1) && (cu != c)) continue;
                if (!(v < u)) { ok = false; break; }
            }
        }
        if (ok) vc[v - lo] = c; else wait = true;
    }
    if (__ballot(wait) != 0 && (threadIdx.x & 63) == 0) *left = 1;
}

__global__ void k_max_reduce(const int* __restrict__ a, int n, int* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicMax(out, a[i]);
}

// colours of one sort at most (the per-block counts live in LDS)
constexpr int kMaxSortColours = 16384;

// per block of TB vertices of [0, n): the count of each colour, stored
// colour-major (hist[c * nb + b]) so that one exclusive scan gives every
// (colour, block) its first position in the colour-sorted order
__global__ void k_colour_hist(const int* __restrict__ col, int n, int nc, int nb, int* __restrict__ hist)
{
    extern __shared__ int sh[];
    for (int c = threadIdx.x; c < nc; c += blockDim.x) sh[c] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&sh[col[i]], 1);
    __syncthreads();
    for (int c = threadIdx.x; c < nc; c += blockDim.x) hist[(size_t)c * nb + blockIdx.x] = sh[c];
}

// new id of vertex lo + i: lo + the scanned first position of (its colour,
// its block) + the vertices of its colour before it in the block
__global__ __launch_bounds__(TB) void k_colour_place(const int* __restrict__ col, int n, int nb,
                                                     const int* __restrict__ pos, int lo, int* __restrict__ newid)
{
    __shared__ int sc[TB];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = (i < n) ? col[i] : -1;
    sc[threadIdx.x] = c;
    __syncthreads();
    if (i >= n) return;
    int r = 0;
    for (int j = 0; j < (int)threadIdx.x; ++j) r += (sc[j] == c) ? 1 : 0;
    newid[lo + i] = lo + pos[(size_t)c * nb + blockIdx.x] + r;
}

// first position of each colour: off[c] = pos[c * nb]
__global__ void k_colour_offsets(const int* __restrict__ pos, int nb, int nc, int* __restrict__ off)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) off[c] = pos[(size_t)c * nb];
}

__global__ void k_iota(int* __restrict__ a, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

__global__ void k_fill_range(int* __restrict__ a, int lo, int n, const int* __restrict__ col)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[lo + i] = col[i];
}

// Class partition check (the rule of k_checkColoring, MeshGPUCommon.cu:
// 1470-1680): the vertices [off[0], off[nc]) are split into the classes
// [off[c], off[c+1]); no element may hold two vertices of one class, or
// their searches in one launch would read each other's positions.  One
// thread per element; a corner's class by binary search of the offsets.
__global__ void k_check_classes(const int* __restrict__ cells, int nE, int npe, const int* __restrict__ off, int nc,
                                int* __restrict__ bad)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    const int lo = off[0], hi = off[nc];
    int cls[8];
    for (int i = 0; i < npe; ++i) {
        const int v = cells[(size_t)e * npe + i];
        int c = -1;
        if (v >= lo && v < hi) {
            int a = 0, b = nc - 1;          // largest c with off[c] <= v
            while (a < b) {
                const int mid = (a + b + 1) >> 1;
                if (off[mid] <= v) a = mid; else b = mid - 1;
            }
            c = a;
        }
        cls[i] = c;
        for (int j = 0; j < i; ++j)
            if (c >= 0 && cls[j] == c) atomicMin(bad, c);
    }
}

void check_classes(const vsdmo_mesh& m, const std::vector<int>& off)
{
    const int nc = (int)off.size() - 1;
    if (nc <= 0 || m.nE == 0) return;
    hipStream_t s = m.stream;
    DevBuf<int> doff(off.size()), bad(1);
    doff.upload(off.data(), off.size(), s);
    const int none = INT_MAX;
    bad.upload(&none, 1, s);
    k_check_classes<<<nblk(m.nE), TB, 0, s>>>(m.cells.p, m.nE, m.npe, doff.p, nc, bad.p);
    VSDMO_HIP(hipGetLastError());
    const int h = read_int(bad.p, s);
    if (h != none)
        throw StatusError{ VSDMO_TOPOLOGY, "colour class " + std::to_string(h) + " (vertices " +
                                               std::to_string(off[h]) + ".." + std::to_string(off[h + 1] - 1) +
                                               ") holds two vertices of one element" };
}

// colour [lo,hi) of the current numbering and sort it stably by colour.
// Returns absolute colour offsets, checked to be independent sets.
// default polls of a vertex in the single-pass greedy colouring before it
// leaves its colour to the rounds (a wait of ~0.1 s at most; the chains of
// the meshes here need a few hundred)
constexpr int kSweepPolls = 1 << 16;

std::vector<int> colour_and_sort(vsdmo_mesh& m, Renumber& rn, const int* incRow, const int* incVal, int lo, int hi,
                                 const vsdmo_mesh_params& prm, std::vector<DevBuf<uint8_t>*> u8)
{
    hipStream_t s = m.stream;
    int n = hi - lo;
    if (n <= 0) return { lo };
    const int mode = prm.coloring;
    const int polls = prm.coloring_polls > 0 ? prm.coloring_polls : kSweepPolls;
    Incidence inc{ incRow, incVal, rn.oldOfCur.p, m.cells.p, m.npe };
    // rounds go out in batches of kRoundsPerCheck, each round raising its own
    // flag when a vertex still waits, and the host reads the flag of the
    // batch's last round: a round after the colouring is complete finds every
    // vertex coloured and changes nothing, so the colours are those of one
    // check per round, with 1/kRoundsPerCheck of the syncs
    constexpr int kRoundsPerCheck = 16;
    DevBuf<int> col(n), left(kRoundsPerCheck);
    VSDMO_HIP(hipMemsetAsync(col.p, 0xff, sizeof(int) * n, s));
    int h_left = 1;
    if (mode != VSDMO_COLOR_JPL) {
        // greedy: one sweep over lower-neighbour lists; the rounds below only
        // run if a vertex ran out of polls
        DevBuf<int> lrow((size_t)n + 1), gave(1);
        k_lower_count<<<nblk(n), TB, 0, s>>>(inc, lo, hi, lrow.p);
        VSDMO_HIP(hipMemsetAsync(lrow.p + n, 0, sizeof(int), s));
        const int tot = exclusive_scan(lrow.p, lrow.p, n + 1, s);
        DevBuf<int> lval((size_t)std::max(tot, 1));
        k_lower_fill<<<nblk(n), TB, 0, s>>>(inc, lo, hi, lrow.p, lval.p);
        VSDMO_HIP(hipMemsetAsync(gave.p, 0, sizeof(int), s));
        k_color_greedy_sweep<<<nblk(n), TB, 0, s>>>(lrow.p, lval.p, n, col.p, polls, gave.p);
        VSDMO_HIP(hipGetLastError());
        h_left = read_int(gave.p, s);
    }
    for (int round = 0; h_left != 0;) {
        if (round > n + 1) throw StatusError{ VSDMO_TOPOLOGY, "colouring did not converge" };
        VSDMO_HIP(hipMemsetAsync(left.p, 0, sizeof(int) * kRoundsPerCheck, s));
        for (int k = 0; k < kRoundsPerCheck; ++k, ++round) {
            if (mode == VSDMO_COLOR_JPL)
                k_color_jpl_round<<<nblk(n), TB, 0, s>>>(inc, lo, hi, round, col.p, left.p + k);
            else
                k_color_greedy_round<<<nblk(n), TB, 0, s>>>(inc, lo, hi, col.p, left.p + k);
        }
        VSDMO_HIP(hipGetLastError());
        h_left = read_int(left.p + kRoundsPerCheck - 1, s);
    }
    DevBuf<int> mx(1);
    VSDMO_HIP(hipMemsetAsync(mx.p, 0, sizeof(int), s));
    k_max_reduce<<<nblk(n), TB, 0, s>>>(col.p, n, mx.p);
    const int nc = read_int(mx.p, s) + 1;
    if (nc > kMaxSortColours) throw StatusError{ VSDMO_NOT_SUPPORTED, "more than 16384 colours" };
    // stable counting sort by colour: per block of TB vertices the count of
    // each colour, scanned colour-major, gives each (colour, block) its first
    // position; a vertex adds the same-coloured vertices before it in its block
    const int nb = nblk(n);
    DevBuf<int> newid(m.nV), pos((size_t)nc * nb), doff(nc);
    k_iota<<<nblk(m.nV), TB, 0, s>>>(newid.p, m.nV);
    k_colour_hist<<<nb, TB, sizeof(int) * nc, s>>>(col.p, n, nc, nb, pos.p);
    exclusive_scan(pos.p, pos.p, nc * nb, s);
    k_colour_place<<<nb, TB, 0, s>>>(col.p, n, nb, pos.p, lo, newid.p);
    k_colour_offsets<<<nblk(nc), TB, 0, s>>>(pos.p, nb, nc, doff.p);
    std::vector<int> off(nc + 1, 0);
    doff.download(off.data(), nc, s);
    VSDMO_HIP(hipStreamSynchronize(s));
    off[nc] = n;
    VSDMO_HIP(hipGetLastError());
    k_fill_range<<<nblk(n), TB, 0, s>>>(m.colorOf.p, lo, n, col.p);
    std::vector<DevBuf<int>*> i32 = { &m.colorOf };
    rn.apply(newid.p, u8, i32);
    for (int& o : off) o += lo;
    check_classes(m, off);
    return off;
}

// ------------------------------------------------------------ one-ring CSR
__constant__ int c_halfface[4][3] = { { 1, 2, 3 }, { 0, 3, 2 }, { 0, 1, 3 }, { 0, 2, 1 } };
__constant__ int c_halfhex[8][7] = { { 1, 2, 3, 4, 5, 6, 7 }, { 5, 6, 2, 0, 4, 7, 3 }, { 6, 7, 3, 1, 5, 4, 0 },
                                     { 7, 4, 0, 2, 6, 5, 1 }, { 0, 3, 7, 5, 1, 2, 6 }, { 4, 7, 6, 1, 0, 3, 2 },
                                     { 5, 4, 7, 2, 1, 0, 3 }, { 6, 5, 4, 3, 2, 1, 0 } };

__global__ void k_ring_counts(const int* __restrict__ incRow, const int* __restrict__ oldOfCur, int n, int* __restrict__ cnt)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) { int o = oldOfCur[v]; cnt[v] = incRow[o + 1] - incRow[o]; }
}

__global__ void k_ring_fill(const int* __restrict__ incRow, const int* __restrict__ incVal, const int* __restrict__ oldOfCur,
                            const int* __restrict__ ringRow, const int* __restrict__ cells, int n, int tet,
                            int* __restrict__ ringVal)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    int o = oldOfCur[v];
    int st = tet ? 3 : 7, npe = tet ? 4 : 8;
    int w = ringRow[v];
    for (int k = incRow[o]; k < incRow[o + 1]; ++k, ++w) {
        int e = incVal[k] >> 3, c = incVal[k] & 7;
        for (int j = 0; j < st; ++j)
            ringVal[(size_t)w * st + j] = cells[(size_t)e * npe + (tet ? c_halfface[c][j] : c_halfhex[c][j])];
    }
}

__global__ void k_incidence_keys(const int* __restrict__ cells, long long n, int* __restrict__ key)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) key[i] = cells[i];
}

__global__ void k_incidence_vals(const int* __restrict__ items, long long n, int npe, int* __restrict__ val)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { int it = items[i]; val[i] = (it / npe) * 8 + (it % npe); }
}

struct SurfaceInputs {
    const int* incRow;   // incidence over input ids (colouring adjacency)
    const int* incVal;
};
void build_surface(vsdmo_mesh& m, Renumber& rn, const SurfaceInputs& si, const vsdmo_mesh_params& prm);

// ------------------------------------------------------------------ build
void build_volume(vsdmo_mesh& m, const vsdmo_mesh_params& prm)
{
    hipStream_t s = m.stream;
    const bool tet = (m.elem == VSDMO_TET);
    const int npe = m.npe, nV = m.nV, nE = m.nE;
    const long long nci = (long long)nE * npe;

    // incidence in input ids: vertex -> (element, corner), ordered by element
    DevBuf<int> incRow, incVal;
    {
        DevBuf<int> key(nci), items;
        k_incidence_keys<<<nblk(nci), TB, 0, s>>>(m.cells.p, nci, key.p);
        stable_csr(key.p, (int)nci, nV, incRow, items, s);
        incVal.alloc(nci);
        k_incidence_vals<<<nblk(nci), TB, 0, s>>>(items.p, nci, npe, incVal.p);
    }

    // boundary faces
    {
        int nf = tet ? 4 : 6, k = tet ? 3 : 4;
        int n = nE * nf;
        DevBuf<int> fk((size_t)n * 4), v0(n), row, items, pos(n), alive(n);
        DevBuf<uint8_t> dead(n);
        k_emit_faces<<<nblk(nE), TB, 0, s>>>(m.cells.p, nE, tet, fk.p, v0.p);
        stable_csr(v0.p, n, nV, row, items, s);
        k_bucket_faces<<<nblk(nV), TB, 0, s>>>(row.p, nV, items.p, fk.p, k, dead.p);
        k_alive<<<nblk(n), TB, 0, s>>>(dead.p, n, alive.p);
        int nb = exclusive_scan(alive.p, pos.p, n, s);
        m.fpe = k;
        m.nFaces = nb;
        m.faces.alloc((size_t)std::max(nb, 1) * k);
        k_compact_faces<<<nblk(n), TB, 0, s>>>(items.p, dead.p, pos.p, n, fk.p, k, m.faces.p);
        VSDMO_HIP(hipGetLastError());
    }

    Renumber rn(m);
    // [surface | inner], stable
    {
        DevBuf<int> flag(nV), scan(nV), newid(nV);
        VSDMO_HIP(hipMemsetAsync(flag.p, 0, sizeof(int) * nV, s));
        long long nfi = (long long)m.nFaces * m.fpe;
        if (nfi > 0) k_mark_vertices<<<nblk(nfi), TB, 0, s>>>(m.faces.p, (int)nfi, flag.p);
        int nS = exclusive_scan(flag.p, scan.p, nV, s);
        k_partition_ids<<<nblk(nV), TB, 0, s>>>(flag.p, scan.p, nV, nS, 1, newid.p);
        rn.apply(newid.p);
        m.nSurf = nS;
        m.nInner = nV - nS;
    }
    // inner colouring and colour sort
    m.colOffInner = colour_and_sort(m, rn, incRow.p, incVal.p, m.nSurf, nV, prm, {});

    // surface: features, [free | feature], colouring, halfedges
    SurfaceInputs si{ incRow.p, incVal.p };
    build_surface(m, rn, si, prm);

    // one-ring CSR in the final numbering
    {
        DevBuf<int> cnt(nV);
        k_ring_counts<<<nblk(nV), TB, 0, s>>>(incRow.p, rn.oldOfCur.p, nV, cnt.p);
        m.ringRow.alloc((size_t)nV + 1);
        VSDMO_HIP(hipMemcpyAsync(m.ringRow.p, cnt.p, sizeof(int) * nV, hipMemcpyDeviceToDevice, s));
        VSDMO_HIP(hipMemsetAsync(m.ringRow.p + nV, 0, sizeof(int), s));
        int tot = exclusive_scan(m.ringRow.p, m.ringRow.p, nV + 1, s);
        int st = tet ? 3 : 7;
        m.ringVal.alloc((size_t)std::max(tot, 1) * st);
        k_ring_fill<<<nblk(nV), TB, 0, s>>>(incRow.p, incVal.p, rn.oldOfCur.p, m.ringRow.p, m.cells.p, nV, tet, m.ringVal.p);
        DevBuf<int> mx(1);
        VSDMO_HIP(hipMemsetAsync(mx.p, 0, sizeof(int), s));
        k_max_reduce<<<nblk(nV), TB, 0, s>>>(cnt.p, nV, mx.p);
        VSDMO_HIP(hipGetLastError());
        m.maxRing = read_int(mx.p, s);
    }
}

// surface meshes (tri): all vertices are surface vertices
void build_surface_mesh(vsdmo_mesh& m, const vsdmo_mesh_params& prm)
{
    hipStream_t s = m.stream;
    const int npe = m.npe, nV = m.nV, nE = m.nE;
    const long long nci = (long long)nE * npe;
    DevBuf<int> incRow, incVal;
    {
        DevBuf<int> key(nci), items;
        k_incidence_keys<<<nblk(nci), TB, 0, s>>>(m.cells.p, nci, key.p);
        stable_csr(key.p, (int)nci, nV, incRow, items, s);
        incVal.alloc(nci);
        k_incidence_vals<<<nblk(nci), TB, 0, s>>>(items.p, nci, npe, incVal.p);
    }
    m.nSurf = nV;
    m.nInner = 0;
    m.colOffInner = { nV };
    m.fpe = npe;
    m.nFaces = nE;
    m.faces.alloc((size_t)nE * npe);
    VSDMO_HIP(hipMemcpyAsync(m.faces.p, m.cells.p, sizeof(int) * nci, hipMemcpyDeviceToDevice, s));
    Renumber rn(m);
    {
        DevBuf<int> id(nV);
        k_iota<<<nblk(nV), TB, 0, s>>>(id.p, nV);
        rn.apply(id.p);
    }
    SurfaceInputs si{ incRow.p, incVal.p };
    build_surface(m, rn, si, prm);
    m.ringRow.alloc((size_t)nV + 1);
    VSDMO_HIP(hipMemsetAsync(m.ringRow.p, 0, sizeof(int) * ((size_t)nV + 1), s));
    VSDMO_HIP(hipStreamSynchronize(s));
}

}  // namespace vsdmo

// =====================================================================
// Surface structures (reference VSDMO/src/mesh/MeshTriGPU.cu:37-52,128-214,
// 463-510,646-784; MeshQuadGPU.cu:37-45,156-243,342-390,481-536;
// MeshGPUCommon.cu:96-135,139-161,211-361,363-411,466-495,565-621,1227-1239)
// =====================================================================
namespace vsdmo {

// canonical face normal: quads use v0, v1, v3 (k_calcFaceNormals)
__global__ void k_face_normals(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                               const int* __restrict__ faces, int nF, int k, float* __restrict__ fn)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    int a = faces[(size_t)f * k], b = faces[(size_t)f * k + 1], c = faces[(size_t)f * k + (k == 3 ? 2 : 3)];
    float e0x = x[b] - x[a], e0y = y[b] - y[a], e0z = z[b] - z[a];
    float e1x = x[c] - x[a], e1y = y[c] - y[a], e1z = z[c] - z[a];
    float cx = e0y * e1z - e0z * e1y, cy = e0z * e1x - e0x * e1z, cz = e0x * e1y - e0y * e1x;
    float l = sqrtf(fmaf(cx, cx, fmaf(cy, cy, cz * cz)));
    fn[3 * f] = cx / l;
    fn[3 * f + 1] = cy / l;
    fn[3 * f + 2] = cz / l;
}

__device__ __forceinline__ int he_src(const int* faces, int k, int h) { return faces[h]; }
__device__ __forceinline__ int he_dst(const int* faces, int k, int h)
{
    int f = h / k, i = h - f * k;
    return faces[(size_t)f * k + (i + 1 == k ? 0 : i + 1)];
}

__global__ void k_he_src(const int* __restrict__ faces, int n, int* __restrict__ key)
{
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h < n) key[h] = faces[h];
}

// opposite of every sorted raw halfedge: last match in the target's row
// (k_setOppositeHalfedgeHelper)
__global__ void k_he_opp_raw(const int* __restrict__ row, const int* __restrict__ items, int nS,
                             const int* __restrict__ faces, int k, int* __restrict__ opp)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nS) return;
    for (int p = row[v]; p < row[v + 1]; ++p) {
        int d = he_dst(faces, k, items[p]);
        int o = -1;
        for (int q = row[d]; q < row[d + 1]; ++q)
            if (he_dst(faces, k, items[q]) == v) o = q;
        opp[p] = o;
    }
}

// boundary-1D flags and features (k_initBoundary1DRowPtr, k_initFeaturesRowPtr
// with the angle test written as cos < cos(threshold))
__global__ void k_features(const int* __restrict__ row, const int* __restrict__ items, const int* __restrict__ opp,
                           int nS, const int* __restrict__ faces, int k, const float* __restrict__ fn, int flat,
                           float cthr, uint8_t* __restrict__ b1d, uint8_t* __restrict__ feat)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nS) return;
    bool isb = false;
    for (int p = row[v]; p < row[v + 1]; ++p) if (opp[p] == -1) { isb = true; break; }
    if (isb) b1d[v] = 1;
    if (flat) {
        if (isb) feat[v] = 1;
        return;
    }
    for (int p = row[v]; p < row[v + 1]; ++p) {
        int h = items[p];
        int d = he_dst(faces, k, h);
        if (opp[p] == -1) { feat[v] = 1; feat[d] = 1; continue; }
        const float* a = fn + 3 * (h / k);
        const float* b = fn + 3 * (items[opp[p]] / k);
        float dot = fmaf(a[0], b[0], fmaf(a[1], b[1], a[2] * b[2]));
        float la = fmaf(a[0], a[0], fmaf(a[1], a[1], a[2] * a[2]));
        float lb = fmaf(b[0], b[0], fmaf(b[1], b[1], b[2] * b[2]));
        float c = dot / sqrtf(la * lb);
        if (c < cthr) { feat[v] = 1; feat[d] = 1; }
    }
}

__global__ void k_u8_to_int(const uint8_t* __restrict__ a, int n, int* __restrict__ b)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i] ? 1 : 0;
}

// stable partition of [0,nS): flag==0 first; identity on [nS,nV)
__global__ void k_partition_surface(const int* __restrict__ flag, const int* __restrict__ scan, int nS, int nV,
                                    int nZero, int* __restrict__ newid)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    if (v >= nS) { newid[v] = v; return; }
    int ones = scan[v];
    newid[v] = flag[v] ? nZero + ones : (v - ones);
}

// merged raw + dummy halfedges: raw sorted position p keeps (src, dst, face);
// every raw halfedge without opposite adds (dst -> src, face -1), appended in
// sorted-position order (k_addBoundaryHalfedges + stable compaction)
__global__ void k_he_merge(const int* __restrict__ row, const int* __restrict__ items, const int* __restrict__ opp,
                           const int* __restrict__ dpos, int nS, int nRaw, const int* __restrict__ faces, int k,
                           int* __restrict__ mKey, int* __restrict__ mTgt, int* __restrict__ mFace)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nS) return;
    for (int p = row[v]; p < row[v + 1]; ++p) {
        int h = items[p];
        int d = he_dst(faces, k, h);
        mKey[p] = v; mTgt[p] = d; mFace[p] = h / k;
        if (opp[p] == -1) {
            int q = nRaw + dpos[p];
            mKey[q] = d; mTgt[q] = v; mFace[q] = -1;
        }
    }
}

__global__ void k_noopp(const int* __restrict__ opp, int n, int* __restrict__ flag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = (opp[i] == -1) ? 1 : 0;
}

__global__ void k_gather2(const int* __restrict__ items, int n, const int* __restrict__ a, const int* __restrict__ b,
                          int* __restrict__ oa, int* __restrict__ ob)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { oa[i] = a[items[i]]; ob[i] = b[items[i]]; }
}

__global__ void k_opp_sorted(const int* __restrict__ row, int nS, const int* __restrict__ tgt, int* __restrict__ opp)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nS) return;
    for (int p = row[v]; p < row[v + 1]; ++p) {
        int d = tgt[p], o = -1;
        for (int q = row[d]; q < row[d + 1]; ++q) if (tgt[q] == v) o = q;
        opp[p] = o;
    }
}

__device__ __forceinline__ int third_vertex(const int* f, int k, int src, int dst)
{
    if (k == 3) {
        if (f[0] != src && f[0] != dst) return f[0];
        if (f[1] != src && f[1] != dst) return f[1];
        return f[2];
    }
    if (dst == f[1]) return f[3];
    if (dst == f[2]) return f[0];
    if (dst == f[3]) return f[1];
    return f[2];
}

// counter-clockwise fan order per vertex (k_sortCCWDummyOpposite): start at
// the boundary halfedge whose twin is a dummy, else at the smallest feature
// neighbour (feature vertices), else at the smallest neighbour; then walk to
// the halfedge pointing at the third vertex of the current face.
__global__ void k_ccw(const int* __restrict__ row, int nS, const int* __restrict__ tgt, const int* __restrict__ face,
                      const int* __restrict__ opp, const uint8_t* __restrict__ b1d, const uint8_t* __restrict__ feat,
                      const int* __restrict__ faces, int k, int* __restrict__ order, int* __restrict__ err)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nS) return;
    int st = row[v], cnt = row[v + 1] - st;
    if (cnt == 0) { atomicOr(err, 1); return; }
    int cur = st;
    if (b1d[v]) {
        int j = 0;
        while (opp[cur] < 0 || face[opp[cur]] != -1) {
            ++j;
            if (j >= cnt) { atomicOr(err, 2); return; }
            cur = st + j;
        }
    } else if (feat[v]) {
        int best = -1, bt = 0x7fffffff;
        for (int j = 0; j < cnt; ++j) {
            int t = tgt[st + j];
            if (feat[t] && t < bt) { bt = t; best = j; }
        }
        if (best < 0) { atomicOr(err, 4); return; }
        cur = st + best;
    } else {
        for (int j = 1; j < cnt; ++j) if (tgt[st + j] < tgt[cur]) cur = st + j;
    }
    int pos = 0;
    order[st + pos++] = cur;
    bool found;
    do {
        int f = face[cur];
        if (f < 0) { atomicOr(err, 8); return; }
        int nxt = third_vertex(faces + (size_t)f * k, k, v, tgt[cur]);
        found = false;
        for (int j = 0; j < cnt; ++j)
            if (tgt[st + j] == nxt) {
                found = true;
                cur = st + j;
                if (pos < cnt) order[st + pos] = cur;
                ++pos;
            }
    } while (found && pos < cnt);
    if (pos != cnt) atomicOr(err, 16);
}

// final opposite and next (last match), k_setOppositeHalfedgeHelper /
// k_setNextHalfedgeHelper
__global__ void k_opp_next(const int* __restrict__ row, int nS, const int* __restrict__ tgt, const int* __restrict__ face,
                           int* __restrict__ opp, int* __restrict__ next)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nS) return;
    for (int p = row[v]; p < row[v + 1]; ++p) {
        int d = tgt[p], o = -1, nx = -1, f = face[p];
        for (int q = row[d]; q < row[d + 1]; ++q) {
            if (tgt[q] == v) o = q;
            if (face[q] == f) nx = q;
        }
        opp[p] = o;
        next[p] = nx;
    }
}

__global__ void k_row_sizes(const int* __restrict__ row, int n, int* __restrict__ mx)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) atomicMax(mx, row[v + 1] - row[v]);
}

void build_surface(vsdmo_mesh& m, Renumber& rn, const SurfaceInputs& si, const vsdmo_mesh_params& prm)
{
    hipStream_t s = m.stream;
    const int nS = m.nSurf, nF = m.nFaces, k = m.fpe;
    const int nV = m.nV;
    m.isB1D.alloc(std::max(nS, 1));
    m.isFeature.alloc(std::max(nS, 1));
    VSDMO_HIP(hipMemsetAsync(m.isB1D.p, 0, std::max(nS, 1), s));
    VSDMO_HIP(hipMemsetAsync(m.isFeature.p, 0, std::max(nS, 1), s));
    if (nS == 0 || nF == 0) {
        m.nFree = 0; m.nFeature = 0;
        m.colOffFree = { 0 };
        m.colOffFeature = { 0 };
        m.heRow.alloc(1);
        VSDMO_HIP(hipMemsetAsync(m.heRow.p, 0, sizeof(int), s));
        return;
    }
    const int nRaw = nF * k;
    float ang = prm.feature_angle_deg > 0.0f ? prm.feature_angle_deg : 20.0f;
    const float cthr = (float)std::cos((double)ang * 3.14159265358979323846 / 180.0);
    // features on the current numbering
    {
        DevBuf<float> fn((size_t)nF * 3);
        k_face_normals<<<nblk(nF), TB, 0, s>>>(m.x.p, m.y.p, m.z.p, m.faces.p, nF, k, fn.p);
        DevBuf<int> key(nRaw), row, items, opp(nRaw);
        k_he_src<<<nblk(nRaw), TB, 0, s>>>(m.faces.p, nRaw, key.p);
        stable_csr(key.p, nRaw, nS, row, items, s);
        k_he_opp_raw<<<nblk(nS), TB, 0, s>>>(row.p, items.p, nS, m.faces.p, k, opp.p);
        k_features<<<nblk(nS), TB, 0, s>>>(row.p, items.p, opp.p, nS, m.faces.p, k, fn.p, m.flat, cthr,
                                           m.isB1D.p, m.isFeature.p);
        VSDMO_HIP(hipGetLastError());
    }
    // [free | feature], stable
    {
        DevBuf<int> flag(nS), scan(nS), newid(nV);
        k_u8_to_int<<<nblk(nS), TB, 0, s>>>(m.isFeature.p, nS, flag.p);
        int nFeat = exclusive_scan(flag.p, scan.p, nS, s);
        int nFree = nS - nFeat;
        k_partition_surface<<<nblk(nV), TB, 0, s>>>(flag.p, scan.p, nS, nV, nFree, newid.p);
        rn.apply(newid.p, { &m.isB1D, &m.isFeature });
        m.nFree = nFree;
        m.nFeature = nFeat;
    }
    m.colOffFree = colour_and_sort(m, rn, si.incRow, si.incVal, 0, m.nFree, prm, { &m.isB1D, &m.isFeature });
    m.colOffFeature = colour_and_sort(m, rn, si.incRow, si.incVal, m.nFree, nS, prm, { &m.isB1D, &m.isFeature });
    // halfedges on the final numbering
    {
        DevBuf<int> key(nRaw), row, items, opp(nRaw), noopp(nRaw), dpos(nRaw);
        k_he_src<<<nblk(nRaw), TB, 0, s>>>(m.faces.p, nRaw, key.p);
        stable_csr(key.p, nRaw, nS, row, items, s);
        k_he_opp_raw<<<nblk(nS), TB, 0, s>>>(row.p, items.p, nS, m.faces.p, k, opp.p);
        k_noopp<<<nblk(nRaw), TB, 0, s>>>(opp.p, nRaw, noopp.p);
        int nd = exclusive_scan(noopp.p, dpos.p, nRaw, s);
        int tot = nRaw + nd;
        DevBuf<int> mKey(tot), mTgt(tot), mFace(tot);
        k_he_merge<<<nblk(nS), TB, 0, s>>>(row.p, items.p, opp.p, dpos.p, nS, nRaw, m.faces.p, k, mKey.p, mTgt.p, mFace.p);
        DevBuf<int> items2;
        stable_csr(mKey.p, tot, nS, m.heRow, items2, s);
        DevBuf<int> sTgt(tot), sFace(tot), sOpp(tot), order(tot), err(1);
        k_gather2<<<nblk(tot), TB, 0, s>>>(items2.p, tot, mTgt.p, mFace.p, sTgt.p, sFace.p);
        k_opp_sorted<<<nblk(nS), TB, 0, s>>>(m.heRow.p, nS, sTgt.p, sOpp.p);
        VSDMO_HIP(hipMemsetAsync(err.p, 0, sizeof(int), s));
        VSDMO_HIP(hipMemsetAsync(order.p, 0, sizeof(int) * tot, s));
        k_ccw<<<nblk(nS), TB, 0, s>>>(m.heRow.p, nS, sTgt.p, sFace.p, sOpp.p, m.isB1D.p, m.isFeature.p, m.faces.p, k,
                                      order.p, err.p);
        {
            const int herr = read_int(err.p, s);
            if (herr)
                throw StatusError{ VSDMO_TOPOLOGY, "surface halfedge fan walk failed (non-manifold surface?) code " +
                                                       std::to_string(herr) };
        }
        m.heTarget.alloc(tot); m.heFace.alloc(tot); m.heOpp.alloc(tot); m.heNext.alloc(tot);
        k_gather2<<<nblk(tot), TB, 0, s>>>(order.p, tot, sTgt.p, sFace.p, m.heTarget.p, m.heFace.p);
        k_opp_next<<<nblk(nS), TB, 0, s>>>(m.heRow.p, nS, m.heTarget.p, m.heFace.p, m.heOpp.p, m.heNext.p);
        DevBuf<int> mx(1);
        VSDMO_HIP(hipMemsetAsync(mx.p, 0, sizeof(int), s));
        k_row_sizes<<<nblk(nS), TB, 0, s>>>(m.heRow.p, nS, mx.p);
        VSDMO_HIP(hipGetLastError());
        m.maxHE = read_int(mx.p, s);
        m.nHE = tot;
    }
}

}  // namespace vsdmo
