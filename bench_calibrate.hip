// bench_calibrate.hip -- measurement support for bench.py --calibrate (not
// part of the library): a kernel with a known count of FP32 FMA
// instructions, to measure the VALU issue rate that the DMO search kernels
// are compared with.  Built with the library's flags (no packed FP32, no
// contraction), so every fmaf below is one non-packed v_fma_f32.
#include <hip/hip_runtime.h>

#define CHAINS 8
#define UNROLL 4

// every lane runs iters x UNROLL rounds of CHAINS independent fmaf chains:
// iters * UNROLL * CHAINS v_fma_f32 per lane, 2 FLOP each
__global__ __launch_bounds__(256) void k_fma_probe(float* __restrict__ out, int iters, float a, float b)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float x[CHAINS];
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) x[c] = (float)(gid & 255) * 1e-3f + (float)c;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
#pragma unroll
            for (int c = 0; c < CHAINS; ++c) x[c] = fmaf(x[c], a, b);
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) s = s + x[c];
    out[gid] = s;
}

extern "C" {

// FMAs per lane of one launch
long long calib_fma_per_lane(int iters) { return (long long)iters * UNROLL * CHAINS; }

// Launches the probe `reps` times (after one untimed launch) on `blocks` x
// 256 threads; ms_out[r] = duration of launch r from HIP events.  Returns 0,
// or the HIP error code.
int calib_fma_run(int blocks, int iters, int reps, float* ms_out)
{
    float* out = nullptr;
    hipEvent_t e0, e1;
    const size_t n = (size_t)blocks * 256;
    hipError_t err = hipMalloc(&out, n * sizeof(float));
    if (err != hipSuccess) return (int)err;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    // a = 1 - 2^-20, b = 2^-12: the chains stay finite and normal
    const float a = 0.99999905f, b = 0.000244140625f;
    k_fma_probe<<<blocks, 256>>>(out, iters, a, b);
    for (int r = 0; r < reps && err == hipSuccess; ++r) {
        (void)hipEventRecord(e0, 0);
        k_fma_probe<<<blocks, 256>>>(out, iters, a, b);
        (void)hipEventRecord(e1, 0);
        err = hipEventSynchronize(e1);
        if (err == hipSuccess) err = hipEventElapsedTime(&ms_out[r], e0, e1);
    }
    if (err == hipSuccess) err = hipGetLastError();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    return (int)err;
}

}
