// microbench.hip -- VALU throughput of the operations the grid builder issues
// (v_fma_f32, v_sin_f32, v_cos_f32, v_fma_f64), measured over the whole device.
// Used to state the transcendental weight of the builder's roofline.

#include "device.hpp"

namespace dplus {

constexpr int kChains = 8, kIters = 4096;

template <int kOp>
__global__ __launch_bounds__(256) void valu_loop(float seed, float *sink) {
    float x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; c++) x[c] = seed + 0.001f * (threadIdx.x + c);
    for (int i = 0; i < kIters; i++) {
#pragma unroll
        for (int c = 0; c < kChains; c++) {
            if constexpr (kOp == 0) x[c] = fmaf(x[c], 0.999f, 0.0001f);
            if constexpr (kOp == 1) x[c] = __builtin_amdgcn_sinf(x[c]);
            if constexpr (kOp == 2) x[c] = __builtin_amdgcn_cosf(x[c]);
        }
    }
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < kChains; c++) s += x[c];
    if (s == 12345.678f) sink[threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void valu_loop64(double seed, double *sink) {
    double x[kChains];
#pragma unroll
    for (int c = 0; c < kChains; c++) x[c] = seed + 0.001 * (threadIdx.x + c);
    for (int i = 0; i < kIters; i++) {
#pragma unroll
        for (int c = 0; c < kChains; c++) x[c] = fma(x[c], 0.999, 0.0001);
    }
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < kChains; c++) s += x[c];
    if (s == 12345.678) sink[threadIdx.x] = s;
}

void microbench_valu(hipStream_t s, double *fma32, double *sin32, double *cos32, double *fma64) {
    const int blocks = Device::get().cus * 16;
    const double ops = (double)blocks * 256 * kChains * kIters;
    DevBuf sink;
    sink.ensure(256 * sizeof(double));
    hipEvent_t e0, e1;
    check_hip(hipEventCreate(&e0), "event");
    check_hip(hipEventCreate(&e1), "event");
    double *outs[4] = {fma32, sin32, cos32, fma64};
    for (int op = 0; op < 4; op++) {
        float best = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
            check_hip(hipEventRecord(e0, s), "record");
            if (op == 0) valu_loop<0><<<blocks, 256, 0, s>>>(0.5f, (float *)sink.p);
            if (op == 1) valu_loop<1><<<blocks, 256, 0, s>>>(0.1f, (float *)sink.p);
            if (op == 2) valu_loop<2><<<blocks, 256, 0, s>>>(0.1f, (float *)sink.p);
            if (op == 3) valu_loop64<<<blocks, 256, 0, s>>>(0.5, (double *)sink.p);
            check_hip(hipGetLastError(), "microbench launch");
            check_hip(hipEventRecord(e1, s), "record");
            check_hip(hipEventSynchronize(e1), "sync");
            float ms = 0.f;
            check_hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            best = std::min(best, ms);
        }
        if (outs[op]) *outs[op] = ops / (best * 1e-3);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

}  // namespace dplus

namespace dplus {

__global__ __launch_bounds__(256) void copy_probe_kernel(const float4 *__restrict__ a, float4 *__restrict__ b,
                                                         long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// plain device-to-device copy of `bytes` (multiple of 16): HBM bandwidth probe and
// the known-size reference for hardware byte counters
void copy_probe(hipStream_t s, long long bytes, double *ms_out) {
    DevBuf a, b;
    a.ensure((size_t)bytes);
    b.ensure((size_t)bytes);
    check_hip(hipMemsetAsync(a.p, 1, (size_t)bytes, s), "memset");
    const long long n = bytes / 16;
    hipEvent_t e0, e1;
    check_hip(hipEventCreate(&e0), "event");
    check_hip(hipEventCreate(&e1), "event");
    check_hip(hipEventRecord(e0, s), "record");
    copy_probe_kernel<<<Device::get().cus * 8, 256, 0, s>>>((const float4 *)a.p, (float4 *)b.p, n);
    check_hip(hipGetLastError(), "copy launch");
    check_hip(hipEventRecord(e1, s), "record");
    check_hip(hipEventSynchronize(e1), "sync");
    float ms = 0.f;
    check_hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    if (ms_out) *ms_out = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

}  // namespace dplus
