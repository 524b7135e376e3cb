// dmo_kernels.h -- launch interface of the DMO kernels (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace vsdmo {

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
};

struct QualityArgs {
    const float* x;
    const float* y;
    const float* z;
    const int* cells;
    int nE, npe, elem;
    float* q;
};

size_t search_lds_bytes(int kind, int maxRing);
void launch_search(int kind, int nq, const SearchArgs& a, size_t lds, hipStream_t s);
void launch_quality(const QualityArgs& a, float* pmin, double* psum, float* omin, double* osum, hipStream_t s);
void launch_histogram(const float* q, int n, int ncols, int* bins, hipStream_t s);

}  // namespace vsdmo
