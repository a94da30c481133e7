// FP64 lattice reductions: sum of (rho - RHO_0) and of rho |u|^2 / 2 over
// all nodes (solid nodes included, as the reference's sumReductionThread_rho
// / sumReductionThread_KE do, src/reduction.cu:99-155). The reference sums
// in FP32 with repeated block passes; here each thread accumulates in FP64,
// one LDS tree per block, and a single block folds the partials in a fixed
// order, so the result is deterministic for a given launch shape.
#include "kernels.h"

namespace mrlbm {

constexpr int RED_THREADS = 256;

template <int S>
__global__ __launch_bounds__(RED_THREADS) void k_reduce_partial(const void* mom_, size_t ntiles, int T, float rho0,
                                                               float rs0, float rs1, double* partial) {
    using C = Codec<S>;
    const typename C::T* mom = (const typename C::T*)mom_;
    __shared__ double sm[RED_THREADS], sk[RED_THREADS];
    double am = 0.0, ak = 0.0;
    const size_t N = ntiles * (size_t)T;
    for (size_t j = (size_t)blockIdx.x * RED_THREADS + threadIdx.x; j < N; j += (size_t)gridDim.x * RED_THREADS) {
        const size_t tile = j / T, t = j % T;
        const size_t b = tile * 10 * T + t;
        const float dr = C::ld(mom, b, rs0);
        const double ux = C::ld(mom, b + T, rs1), uy = C::ld(mom, b + 2 * T, rs1), uz = C::ld(mom, b + 3 * T, rs1);
        am += (double)dr;
        ak += ((double)rho0 + (double)dr) * (ux * ux + uy * uy + uz * uz) / 18.0;
    }
    sm[threadIdx.x] = am;
    sk[threadIdx.x] = ak;
    __syncthreads();
    for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sm[threadIdx.x] += sm[threadIdx.x + s];
            sk[threadIdx.x] += sk[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sm[0];
        partial[2 * blockIdx.x + 1] = sk[0];
    }
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce_final(const double* partial, int n, double* out2) {
    __shared__ double sm[RED_THREADS], sk[RED_THREADS];
    double am = 0.0, ak = 0.0;
    for (int j = threadIdx.x; j < n; j += RED_THREADS) {
        am += partial[2 * j];
        ak += partial[2 * j + 1];
    }
    sm[threadIdx.x] = am;
    sk[threadIdx.x] = ak;
    __syncthreads();
    for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sm[threadIdx.x] += sm[threadIdx.x + s];
            sk[threadIdx.x] += sk[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out2[0] = sm[0];
        out2[1] = sk[0];
    }
}

hipError_t launch_reduce(int storage, const void* mom, size_t ntiles, int tile_nodes, float rho0, float rs0, float rs1,
                         double* partial, int nblocks, double* out2, hipStream_t s) {
    if (storage == 0)
        hipLaunchKernelGGL(k_reduce_partial<0>, dim3(nblocks), dim3(RED_THREADS), 0, s, mom, ntiles, tile_nodes, rho0,
                           rs0, rs1, partial);
    else if (storage == 1)
        hipLaunchKernelGGL(k_reduce_partial<1>, dim3(nblocks), dim3(RED_THREADS), 0, s, mom, ntiles, tile_nodes, rho0,
                           rs0, rs1, partial);
    else if (storage == 2)
        hipLaunchKernelGGL(k_reduce_partial<2>, dim3(nblocks), dim3(RED_THREADS), 0, s, mom, ntiles, tile_nodes, rho0,
                           rs0, rs1, partial);
    else
        return hipErrorInvalidValue;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_THREADS), 0, s, partial, nblocks, out2);
    return hipGetLastError();
}

}  // namespace mrlbm
