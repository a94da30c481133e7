// Plain device-to-device copy used to measure the HBM bandwidth the box
// delivers and to calibrate the FETCH_SIZE / WRITE_SIZE counters against a
// kernel of known traffic (bytes read = bytes written = size).
#include <hip/hip_runtime.h>

#include "../../include/mrlbm.h"

namespace {
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
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
        const int grid = 256 * 8 * 4;
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
