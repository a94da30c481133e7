// FP64 lattice reductions: sum of (rho - RHO_0) and of rho |u|^2 / 2 over
// all nodes (solid nodes included, as the reference's sumReductionThread_rho
// / sumReductionThread_KE do, src/reduction.cu:99-155), with a count of
// non-finite nodes, and the velocity convergence residual. The reference sums
// in FP32 with repeated block passes; here each thread accumulates in FP64,
// one LDS tree per block, and a single block folds the partials in a fixed
// order, so the result is deterministic for a given launch shape.
#include <mutex>
#include <vector>

#include "context.h"

namespace mrlbm {

constexpr int RED_THREADS = 256;

// Per block: sums of (rho - RHO_0) and rho |u|^2 / 2, and the number of
// nodes whose rho - RHO_0 or u is not finite (any such node also makes the
// sums non-finite; the count says how many). partial[3 b + 0..2].
template <int S>
__global__ __launch_bounds__(RED_THREADS) void k_reduce_partial(const void* mom_, size_t ntiles, int T, float rho0,
                                                               float rs0, float rs1, double* partial) {
    using C = Codec<S>;
    const typename C::T* mom = (const typename C::T*)mom_;
    __shared__ double sm[RED_THREADS], sk[RED_THREADS], sb[RED_THREADS];
    double am = 0.0, ak = 0.0, ab = 0.0;
    const size_t N = ntiles * (size_t)T;
    for (size_t j = (size_t)blockIdx.x * RED_THREADS + threadIdx.x; j < N; j += (size_t)gridDim.x * RED_THREADS) {
        const size_t tile = j / T, t = j % T;
        const size_t b = tile * 10 * T + t;
        const float dr = C::ld(mom, b, rs0);
        const double ux = C::ld(mom, b + T, rs1), uy = C::ld(mom, b + 2 * T, rs1), uz = C::ld(mom, b + 3 * T, rs1);
        am += (double)dr;
        ak += ((double)rho0 + (double)dr) * (ux * ux + uy * uy + uz * uz) / 18.0;
        ab += (__builtin_isfinite(dr) && __builtin_isfinite(ux) && __builtin_isfinite(uy) && __builtin_isfinite(uz))
                  ? 0.0
                  : 1.0;
    }
    sm[threadIdx.x] = am;
    sk[threadIdx.x] = ak;
    sb[threadIdx.x] = ab;
    __syncthreads();
    for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sm[threadIdx.x] += sm[threadIdx.x + s];
            sk[threadIdx.x] += sk[threadIdx.x + s];
            sb[threadIdx.x] += sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x] = sm[0];
        partial[3 * blockIdx.x + 1] = sk[0];
        partial[3 * blockIdx.x + 2] = sb[0];
    }
}

// Fold n partial triples in a fixed order into out3; with n_nodes > 0 the
// first two (the sums) are divided by it.
__global__ __launch_bounds__(RED_THREADS) void k_reduce_final(const double* partial, int n, double* out3,
                                                              double n_nodes) {
    __shared__ double s3[3][RED_THREADS];
    double a[3] = {0.0, 0.0, 0.0};
    for (int j = threadIdx.x; j < n; j += RED_THREADS)
        for (int k = 0; k < 3; k++) a[k] += partial[3 * j + k];
    for (int k = 0; k < 3; k++) s3[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 3; k++) s3[k][threadIdx.x] += s3[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 3) out3[threadIdx.x] = (threadIdx.x < 2 && n_nodes > 0.0) ? s3[threadIdx.x][0] / n_nodes
                                                                                 : s3[threadIdx.x][0];
}

hipError_t launch_reduce(int storage, const void* mom, size_t ntiles, int tile_nodes, float rho0, float rs0, float rs1,
                         double* partial, int nblocks, double* out3, hipStream_t s, double n_nodes) {
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
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_THREADS), 0, s, partial, nblocks, out3, n_nodes);
    return hipGetLastError();
}

// Convergence residual: per block, sum over nodes of |u - u_prev|^2 and of
// |u|^2 in FP64 (u from the stored moments, u_prev a saved FP32 copy,
// 3 planes of N values in block layout); u_prev is then overwritten with u.
// partial[3 b + 0..1] (the third slot is zero, so k_reduce_final folds it).
template <int S>
__global__ __launch_bounds__(RED_THREADS) void k_residual_partial(const void* mom_, float* uprev, size_t ntiles, int T,
                                                                 float rs1, double* partial) {
    using C = Codec<S>;
    const typename C::T* mom = (const typename C::T*)mom_;
    __shared__ double sd[RED_THREADS], su[RED_THREADS];
    double ad = 0.0, au = 0.0;
    const size_t N = ntiles * (size_t)T;
    for (size_t j = (size_t)blockIdx.x * RED_THREADS + threadIdx.x; j < N; j += (size_t)gridDim.x * RED_THREADS) {
        const size_t tile = j / T, t = j % T;
        const size_t b = tile * 10 * T + t;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float u = C::ld(mom, b + (size_t)(1 + a) * T, rs1) * (1.0f / 3.0f);
            const double d = (double)u - (double)uprev[(size_t)a * N + j];
            ad += d * d;
            au += (double)u * (double)u;
            uprev[(size_t)a * N + j] = u;
        }
    }
    sd[threadIdx.x] = ad;
    su[threadIdx.x] = au;
    __syncthreads();
    for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sd[threadIdx.x] += sd[threadIdx.x + s];
            su[threadIdx.x] += su[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x] = sd[0];
        partial[3 * blockIdx.x + 1] = su[0];
        partial[3 * blockIdx.x + 2] = 0.0;
    }
}

hipError_t launch_residual(int storage, const void* mom, float* uprev, size_t ntiles, int tile_nodes, float rs1,
                           double* partial, int nblocks, double* out3, hipStream_t s) {
    if (storage == 0)
        hipLaunchKernelGGL(k_residual_partial<0>, dim3(nblocks), dim3(RED_THREADS), 0, s, mom, uprev, ntiles,
                           tile_nodes, rs1, partial);
    else if (storage == 1)
        hipLaunchKernelGGL(k_residual_partial<1>, dim3(nblocks), dim3(RED_THREADS), 0, s, mom, uprev, ntiles,
                           tile_nodes, rs1, partial);
    else if (storage == 2)
        hipLaunchKernelGGL(k_residual_partial<2>, dim3(nblocks), dim3(RED_THREADS), 0, s, mom, uprev, ntiles,
                           tile_nodes, rs1, partial);
    else
        return hipErrorInvalidValue;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_THREADS), 0, s, partial, nblocks, out3, 0.0);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Allocation-free checked reduction (mrlbm_reduce_checked_noalloc). The
// partial sums live in a file-scope device array, one copy per device for
// the life of the process, instead of the context's partial buffer; the
// kernels and the block count are those of mrlbm_reduce_checked, so the
// result is the same bit for bit. In host code `scratch` names a host-side
// stand-in: kernels get the device address that hipGetSymbolAddress returns
// for the current device, looked up once per device and kept.
constexpr int NOALLOC_RED_BLOCKS = 1024;  // 1024 partial triples, 24 KiB
static __device__ __align__(256) double scratch[3 * NOALLOC_RED_BLOCKS];

namespace {
std::mutex g_scratch_mu;
std::vector<double*> g_scratch;  // per device: device address of `scratch`
}  // namespace

int noalloc_scratch_setup(int device, double** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0) return fail(MRLBM_EINVAL, "device %d", device);
    if ((int)g_scratch.size() <= device) g_scratch.resize((size_t)device + 1, nullptr);
    if (!g_scratch[device]) {
        int cur = -1;
        HIPCHK(hipGetDevice(&cur));
        if (cur != device) return fail(MRLBM_EHIP, "scratch set-up for device %d while device %d is current", device, cur);
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MRLBM_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        g_scratch[device] = (double*)p;
    }
    if (out) *out = g_scratch[device];
    return MRLBM_OK;
}

}  // namespace mrlbm

extern "C" int mrlbm_reduce_checked_noalloc(mrlbm_ctx* c, double* d_out3) {
    using namespace mrlbm;
    if (!c || !d_out3) return fail(MRLBM_EINVAL, "null argument");
    if (c->red_blocks > NOALLOC_RED_BLOCKS)
        return fail(MRLBM_EINVAL, "reduction uses %d partial blocks, the scratch holds %d", c->red_blocks,
                    NOALLOC_RED_BLOCKS);
    HIPCHK(hipSetDevice(c->p.device));
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, d_out3) != hipSuccess || at.type != hipMemoryTypeDevice || at.device != c->p.device) {
        (void)hipGetLastError();  // clear the query's error state
        return fail(MRLBM_EINVAL, "d_out3 must be device memory on device %d", c->p.device);
    }
    double* part = nullptr;
    const int rc = noalloc_scratch_setup(c->p.device, &part);
    if (rc) return rc;
    HIPCHK(launch_reduce(c->p.storage, c->d_mom, c->ntiles, c->T, (float)c->p.rho0, c->mrscale[0], c->mrscale[1], part,
                         c->red_blocks, d_out3, c->stream, (double)c->N));
    return MRLBM_OK;
}
