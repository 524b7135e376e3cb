// bench_calibrate.hip -- measurement support for bench.py --calibrate (not
// part of the library): kernels with a known count of VALU instructions of
// one kind each, to measure the issue rate of every instruction class the DMO
// search kernels execute.  Built with the library's flags (no packed FP32, no
// contraction), so every operation below is one non-packed VALU instruction
// (checked in the generated code: the loop bodies hold only the probed
// instruction, plus the loop counter on the scalar unit).
#include <hip/hip_runtime.h>

#define CHAINS 8
#define UNROLL 4

// probed operation per chain and step:
//   0 v_fma_f32   x = fmaf(x, a, b)
//   1 v_mul_f32   x = x * a
//   2 v_add_f32   x = x + b
//   3 v_cmp_lt_f32 + v_cndmask_b32   x = (x < y) ? y : x, y = the register y_{(c+u) mod 8} of chain c's
//     step u (two instructions per operation)
//   4 the tet search's inner step: for 4 candidates, det = fmaf(n, p, T),
//     L = fmaf(p, t + m, R), then the running selection of min det^2 / L^3
//     (mul, mul, mul, mul, mul, cmp, 2 x cndmask) and the running min det;
//     the instruction count per step is read from the generated code by the
//     caller (calib_probe_instr_per_lane)
template <int OP>
__device__ __forceinline__ float probe_op(float x, float a, float b, float y)
{
    if constexpr (OP == 0) return fmaf(x, a, b);
    else if constexpr (OP == 1) return x * a;
    else if constexpr (OP == 2) return x + b;
    else return (x < y) ? y : x;
}

// every lane runs iters x UNROLL rounds of CHAINS independent chains
template <int OP>
__global__ __launch_bounds__(256) void k_valu_probe(float* __restrict__ out, int iters, float a, float b)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    float x[CHAINS], y[CHAINS];
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) {
        x[c] = (float)(gid & 255) * 1e-3f + (float)c;
        y[c] = (float)((gid >> 3) & 7) - (float)c;
    }
    if constexpr (OP == 4) {
        // 4 candidates (x[0..3]) against rotating element constants y[0..7]
        float bd[4], bl[4], dm[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) { bd[c] = 1.0f; bl[c] = 0.0f; dm[c] = 3.0e38f; }
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int u = 0; u < UNROLL * 2; ++u) {
                const float T = y[u & 7], R = y[(u + 3) & 7] + 8.0f, m = y[(u + 5) & 7];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float det = fmaf(a, x[c], T);
                    const float L = fmaf(x[c], x[c + 4] + m, R);
                    const float d2 = det * det, L3 = (L * L) * L;
                    const bool take = d2 * bl[c] < bd[c] * L3;
                    bd[c] = take ? d2 : bd[c];
                    bl[c] = take ? L3 : bl[c];
                    dm[c] = fminf(dm[c], det);
                }
            }
        }
        float s = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c) s = s + (bd[c] / bl[c] + dm[c]);
        out[gid] = s;
        return;
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
#pragma unroll
            for (int c = 0; c < CHAINS; ++c) x[c] = probe_op<OP>(x[c], a, b, y[(c + u) % CHAINS]);
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < CHAINS; ++c) s = s + x[c];
    out[gid] = s;
}

extern "C" {

// probed operations per lane of one launch (op 3: each is two instructions)
long long calib_fma_per_lane(int iters) { return (long long)iters * UNROLL * CHAINS; }

// number of probe kinds
int calib_probe_count() { return 5; }

// Launches probe `op` `reps` times (after one untimed launch) on `blocks` x
// 256 threads; ms_out[r] = duration of launch r from HIP events.  Returns 0,
// or the HIP error code (hipErrorInvalidValue for an unknown op).
int calib_probe_run(int op, int blocks, int iters, int reps, float* ms_out)
{
    if (op < 0 || op > 4) return (int)hipErrorInvalidValue;
    float* out = nullptr;
    hipEvent_t e0, e1;
    const size_t n = (size_t)blocks * 256;
    hipError_t err = hipMalloc(&out, n * sizeof(float));
    if (err != hipSuccess) return (int)err;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    // a = 1 - 2^-20, b = 2^-12: the chains stay finite and normal
    const float a = 0.99999905f, b = 0.000244140625f;
    auto launch = [&]() {
        switch (op) {
        case 0: k_valu_probe<0><<<blocks, 256>>>(out, iters, a, b); break;
        case 1: k_valu_probe<1><<<blocks, 256>>>(out, iters, a, b); break;
        case 2: k_valu_probe<2><<<blocks, 256>>>(out, iters, a, b); break;
        case 3: k_valu_probe<3><<<blocks, 256>>>(out, iters, a, b); break;
        default: k_valu_probe<4><<<blocks, 256>>>(out, iters, a, b); break;
        }
    };
    launch();
    for (int r = 0; r < reps && err == hipSuccess; ++r) {
        (void)hipEventRecord(e0, 0);
        launch();
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

// the FP32 FMA probe alone (kept for its callers)
int calib_fma_run(int blocks, int iters, int reps, float* ms_out)
{
    return calib_probe_run(0, blocks, iters, reps, ms_out);
}

}
