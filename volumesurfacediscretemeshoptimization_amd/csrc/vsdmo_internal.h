// vsdmo_internal.h -- internal types of libvsdmo (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/vsdmo.h"

namespace vsdmo {

// thread-local last error message (vsdmo_last_error)
void set_error(const std::string& msg);

struct HipError {
    hipError_t code;
    std::string where;
};

#define VSDMO_HIP(expr)                                                                     \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess) throw ::vsdmo::HipError{ _e, std::string(#expr) + " at " +    \
                                                              __FILE__ + ":" + std::to_string(__LINE__) }; \
    } while (0)

struct StatusError {
    vsdmo_status status;
    std::string msg;
};

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count)
    {
        release();
        n = count;
        if (count) VSDMO_HIP(hipMalloc(&p, count * sizeof(T)));
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count, hipStream_t s)
    {
        if (count) VSDMO_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* h, size_t count, hipStream_t s) const
    {
        if (count) VSDMO_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    std::vector<T> to_host(hipStream_t s) const
    {
        std::vector<T> v(n);
        download(v.data(), n, s);
        VSDMO_HIP(hipStreamSynchronize(s));
        return v;
    }
};

}  // namespace vsdmo

// Mesh: device-resident vertex positions (SoA), elements, colouring and the
// one-ring structures, all in the optimiser's vertex numbering:
//   [0, nFree) free surface | [nFree, nSurf) feature surface | [nSurf, nV) inner,
// each range sorted by colour.
struct vsdmo_mesh {
    int elem = 0, npe = 0, nV = 0, nE = 0;
    int flat = 0;
    int nSurf = 0, nFree = 0, nFeature = 0, nInner = 0;
    int nFaces = 0, fpe = 0;
    int maxRing = 0, maxHE = 0, nHE = 0;
    std::vector<int> colOffInner, colOffFree, colOffFeature;  // absolute ids
    std::vector<int> new_of_old;                               // host copy
    vsdmo::DevBuf<float> x, y, z;
    vsdmo::DevBuf<int> cells;        // npe * nE
    vsdmo::DevBuf<int> faces;        // fpe * nFaces
    vsdmo::DevBuf<int> ringRow;      // nV + 1
    vsdmo::DevBuf<int> ringVal;      // stride * nnz (3 tet, 7 hex)
    vsdmo::DevBuf<int> colorOf;      // nV
    vsdmo::DevBuf<int> newOfOld;     // nV
    vsdmo::DevBuf<int> heRow, heTarget, heOpp, heFace, heNext;
    vsdmo::DevBuf<uint8_t> isB1D, isFeature;
    vsdmo::DevBuf<int> nbr;          // canonical neighbour table (learned proposal), built on demand
    // position of every one-ring corner in its vertex's neighbour table
    // (kRingLocNone: not in it), kRingLocStride(elem) bytes per ring element;
    // built with nbr
    vsdmo::DevBuf<uint8_t> ringLoc;
    hipStream_t stream = nullptr;
    int n_opts = 0;   // live optimisers referencing this mesh
};
