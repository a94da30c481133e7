// Plain device-to-device copy used to measure the HBM bandwidth the box
// delivers and to calibrate the FETCH_SIZE / WRITE_SIZE counters against a
// kernel of known traffic (bytes read = bytes written = size).
#include <hip/hip_runtime.h>

#include "../../include/mrlbm.h"

namespace {
// each thread moves 4 float4 (loads issued before stores), one pass
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    const size_t i0 = (size_t)blockIdx.x * 1024 + threadIdx.x;
    float4 v[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (i0 + 256 * k < n) v[k] = src[i0 + 256 * k];
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (i0 + 256 * k < n) dst[i0 + 256 * k] = v[k];
}
}  // namespace

extern "C" int mrlbm_copy_bench(int device, size_t bytes, int iters, double* ms_per_copy) {
    if (!ms_per_copy || iters <= 0 || bytes < 16) return MRLBM_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return MRLBM_EHIP;
    const size_t n = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 16) != hipSuccess) return MRLBM_ENOMEM;
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return MRLBM_ENOMEM;
    }
    int rc = MRLBM_OK;
    hipStream_t s;
    hipEvent_t e0, e1;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) rc = MRLBM_EHIP;
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = MRLBM_EHIP;
    if (!rc) {
        const int grid = (int)((n + 1023) / 1024);
        (void)hipMemsetAsync(a, 0, n * 16, s);
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, s, a, b, n);  // warm-up
        (void)hipEventRecord(e0, s);
        for (int i = 0; i < iters; i++) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, s, a, b, n);
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) rc = MRLBM_EHIP;
        float f = 0;
        if (!rc && hipEventElapsedTime(&f, e0, e1) != hipSuccess) rc = MRLBM_EHIP;
        *ms_per_copy = f / iters;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}
