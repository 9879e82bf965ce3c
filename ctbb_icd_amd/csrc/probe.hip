// probe.hip -- counter calibration kernel.
#include "common.hpp"

// ---------------------------------------------------------------------------
// Counter calibration: a plain device copy of a known number of bytes
// (float4 loads and stores, grid-stride), so that hardware byte counters read
// for the hot-path kernels can be scaled by what they read for a copy.
namespace {
__global__ void ctbb_copy_kernel(const float4 *__restrict__ src,
                                 float4 *__restrict__ dst, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void ctbb_fill_kernel(float4 *dst, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = make_float4(1.0f, 2.0f, 3.0f, 4.0f);
}
}  // namespace

extern "C" int ctbb_copy_probe(int64_t bytes, double *ms)
{
    if (bytes < 16)
        return ctbb::fail(CTBB_E_ARG, "copy probe needs at least 16 bytes");
    const int64_t n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    CTBB_TRY(ctbb::dev_alloc_n(&a, n, "copy source"));
    int rc = ctbb::dev_alloc_n(&b, n, "copy destination");
    if (rc) {
        (void)hipFree(a);
        return rc;
    }
    hipEvent_t e0, e1;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(ctbb_fill_kernel, dim3(8192), dim3(256), 0, 0, a, n);
        hipLaunchKernelGGL(ctbb_fill_kernel, dim3(8192), dim3(256), 0, 0, b, n);
        // best of five copies (the first launches also warm clocks and TLBs)
        float best = 0.0f;
        for (int rep = 0; rep < 5 && e == hipSuccess; rep++) {
            e = hipEventRecord(e0, 0);
            hipLaunchKernelGGL(ctbb_copy_kernel, dim3(8192), dim3(256), 0, 0, a, b, n);
            if (e == hipSuccess) e = hipEventRecord(e1, 0);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float t = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
            if (rep == 0 || t < best)
                best = t;
        }
        if (ms) *ms = best;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    CTBB_HIP(e);
    return CTBB_OK;
}
