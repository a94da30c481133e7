// Diagnostics on the device: node sampling (probe points and lines),
// per-moment FP64 means, a running mean flow and the turbulent kinetic
// energy against it. They replace the reference's treatData family
// (src/treatData.cu): mean_moment (:36-102), velocityProfile (:383-498) and
// probeExport (:501-584), which copy one node at a time to the host, the
// MEAN_FLOW accumulation of fMom_mean and turbulentKineticEnergy (:236-296).
#include <hip/hip_runtime.h>

#include <vector>

#include "context.h"

using namespace mrlbm;

namespace {

constexpr int DIAG_THREADS = 256;

struct Layout {
    int nx, ny, nz, bx, by, bz, ntx, nty, T;
    __device__ __forceinline__ size_t mom_index(int x, int y, int z, int m) const {
        const int tile = x / bx + ntx * (y / by + nty * (z / bz));
        const int t = x % bx + bx * (y % by + by * (z % bz));
        return (size_t)tile * 10 * T + (size_t)m * T + t;
    }
};

Layout layout_of(const mrlbm_ctx* c) {
    return Layout{c->p.nx, c->p.ny, c->p.nz, c->bx, c->by, c->bz, c->ntx, c->nty, c->T};
}

struct Scales {
    float r[10];
};

Scales scales_of(const mrlbm_ctx* c) {
    Scales s{};
    for (int m = 0; m < 10; m++) s.r[m] = c->mrscale[m];
    return s;
}

// stored-convention moments of n nodes (xyz triples) as FP64
template <int S>
__global__ void k_sample(const void* mom_, Layout L, Scales sc, const int* __restrict__ xyz, int n,
                         double* __restrict__ out) {
    using C = Codec<S>;
    const auto mom = gptr((const typename C::T*)mom_);
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int x = xyz[3 * k], y = xyz[3 * k + 1], z = xyz[3 * k + 2];
#pragma unroll
    for (int m = 0; m < 10; m++) out[10 * (size_t)k + m] = (double)C::ld(mom, L.mom_index(x, y, z, m), sc.r[m]);
}

// per-block FP64 sums of the 10 stored moments
template <int S>
__global__ __launch_bounds__(DIAG_THREADS) void k_moment_sums(const void* mom_, size_t ntiles, int T, Scales sc,
                                                              double* __restrict__ partial) {
    using C = Codec<S>;
    const auto mom = gptr((const typename C::T*)mom_);
    __shared__ double sh[10][DIAG_THREADS];
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const size_t N = ntiles * (size_t)T;
    for (size_t j = (size_t)blockIdx.x * DIAG_THREADS + threadIdx.x; j < N; j += (size_t)gridDim.x * DIAG_THREADS) {
        const size_t b = (j / T) * 10 * T + j % T;
#pragma unroll
        for (int m = 0; m < 10; m++) acc[m] += (double)C::ld(mom, b + (size_t)m * T, sc.r[m]);
    }
#pragma unroll
    for (int m = 0; m < 10; m++) sh[m][threadIdx.x] = acc[m];
    __syncthreads();
    for (int s = DIAG_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int m = 0; m < 10; m++) sh[m][threadIdx.x] += sh[m][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 10) partial[10 * (size_t)blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// acc (4 x N, x-fastest, FP64) += (rho, ux, uy, uz) of every node
template <int S>
__global__ void k_mean_accumulate(const void* mom_, Layout L, Scales sc, float rho0, double* __restrict__ acc) {
    using C = Codec<S>;
    const auto mom = gptr((const typename C::T*)mom_);
    const size_t N = (size_t)L.nx * L.ny * L.nz;
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const int x = (int)(j % L.nx), y = (int)((j / L.nx) % L.ny), z = (int)(j / ((size_t)L.nx * L.ny));
    const double rho = (double)rho0 + (double)C::ld(mom, L.mom_index(x, y, z, 0), sc.r[0]);
    acc[j] += rho;
    for (int a = 0; a < 3; a++) acc[(size_t)(a + 1) * N + j] += (double)C::ld(mom, L.mom_index(x, y, z, 1 + a), sc.r[1 + a]) / 3.0;
}

// per-block sums of rho |u - mean u|^2 / 2, mean u = acc / count
template <int S>
__global__ __launch_bounds__(DIAG_THREADS) void k_tke_partial(const void* mom_, Layout L, Scales sc, float rho0,
                                                              const double* __restrict__ acc, double inv_count,
                                                              double* __restrict__ partial) {
    using C = Codec<S>;
    const auto mom = gptr((const typename C::T*)mom_);
    __shared__ double sh[DIAG_THREADS];
    const size_t N = (size_t)L.nx * L.ny * L.nz;
    double a = 0.0;
    for (size_t j = (size_t)blockIdx.x * DIAG_THREADS + threadIdx.x; j < N; j += (size_t)gridDim.x * DIAG_THREADS) {
        const int x = (int)(j % L.nx), y = (int)((j / L.nx) % L.ny), z = (int)(j / ((size_t)L.nx * L.ny));
        const double rho = (double)rho0 + (double)C::ld(mom, L.mom_index(x, y, z, 0), sc.r[0]);
        double e = 0.0;
        for (int d = 0; d < 3; d++) {
            const double u = (double)C::ld(mom, L.mom_index(x, y, z, 1 + d), sc.r[1 + d]) / 3.0;
            const double du = u - acc[(size_t)(d + 1) * N + j] * inv_count;
            e += du * du;
        }
        a += 0.5 * rho * e;
    }
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int s = DIAG_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// host-side fold of block partials in a fixed order (deterministic)
double fold(const std::vector<double>& p, int stride, int k) {
    double s = 0.0;
    for (size_t b = 0; b * stride + k < p.size(); b++) s += p[b * stride + k];
    return s;
}

unsigned blocks_for(size_t n) { return (unsigned)((n + DIAG_THREADS - 1) / DIAG_THREADS); }

int ensure_mean(mrlbm_ctx* c) {
    if (c->d_mean) return MRLBM_OK;
    HIPCHK(hipMalloc((void**)&c->d_mean, sizeof(double) * 4 * c->N));
    HIPCHK(hipMemsetAsync(c->d_mean, 0, sizeof(double) * 4 * c->N, c->stream));
    c->device_bytes += (double)(sizeof(double) * 4 * c->N);
    c->mean_count = 0;
    return MRLBM_OK;
}

}  // namespace

#define MRLBM_DIAG_LAUNCH(KERNEL, GRID, BLOCK, ...)                                            \
    do {                                                                                       \
        if (c->p.storage == MRLBM_STORE_FP32)                                                  \
            hipLaunchKernelGGL(KERNEL<0>, GRID, BLOCK, 0, c->stream, __VA_ARGS__);             \
        else if (c->p.storage == MRLBM_STORE_FP16)                                             \
            hipLaunchKernelGGL(KERNEL<1>, GRID, BLOCK, 0, c->stream, __VA_ARGS__);             \
        else                                                                                   \
            hipLaunchKernelGGL(KERNEL<2>, GRID, BLOCK, 0, c->stream, __VA_ARGS__);             \
        HIPCHK(hipGetLastError());                                                             \
    } while (0)

extern "C" {

int mrlbm_sample(mrlbm_ctx* c, int n, const int32_t* xyz, double* out) {
    if (!c || n < 0 || (n > 0 && (!xyz || !out))) return fail(MRLBM_EINVAL, "bad argument");
    if (!c->initialised) return fail(MRLBM_ESTATE, "lattice not initialised");
    for (int k = 0; k < n; k++)
        if (xyz[3 * k] < 0 || xyz[3 * k] >= c->p.nx || xyz[3 * k + 1] < 0 || xyz[3 * k + 1] >= c->p.ny ||
            xyz[3 * k + 2] < 0 || xyz[3 * k + 2] >= c->p.nz)
            return fail(MRLBM_EINVAL, "sample point %d (%d,%d,%d) outside the lattice", k, xyz[3 * k], xyz[3 * k + 1],
                        xyz[3 * k + 2]);
    if (n == 0) return MRLBM_OK;
    HIPCHK(hipSetDevice(c->p.device));
    int* d_xyz = nullptr;
    double* d_out = nullptr;
    HIPCHK(hipMallocAsync((void**)&d_xyz, sizeof(int) * 3 * (size_t)n, c->stream));
    HIPCHK(hipMallocAsync((void**)&d_out, sizeof(double) * 10 * (size_t)n, c->stream));
    HIPCHK(hipMemcpyAsync(d_xyz, xyz, sizeof(int) * 3 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    const Layout L = layout_of(c);
    const Scales sc = scales_of(c);
    MRLBM_DIAG_LAUNCH(k_sample, dim3(blocks_for(n)), dim3(DIAG_THREADS), c->d_mom, L, sc, d_xyz, n, d_out);
    HIPCHK(hipMemcpyAsync(out, d_out, sizeof(double) * 10 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipFreeAsync(d_xyz, c->stream));
    HIPCHK(hipFreeAsync(d_out, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MRLBM_OK;
}

int mrlbm_moment_means(mrlbm_ctx* c, double out[10]) {
    if (!c || !out) return fail(MRLBM_EINVAL, "null argument");
    if (!c->initialised) return fail(MRLBM_ESTATE, "lattice not initialised");
    HIPCHK(hipSetDevice(c->p.device));
    const int nb = 2 * c->red_blocks / 10;  // partials fit the reduction scratch
    const Scales sc = scales_of(c);
    MRLBM_DIAG_LAUNCH(k_moment_sums, dim3(nb), dim3(DIAG_THREADS), c->d_mom, (size_t)c->ntiles, c->T, sc,
                      c->d_partial);
    std::vector<double> p(10 * (size_t)nb);
    HIPCHK(hipMemcpyAsync(p.data(), c->d_partial, sizeof(double) * p.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int m = 0; m < 10; m++) out[m] = fold(p, 10, m) / (double)c->N;
    return MRLBM_OK;
}

int mrlbm_mean_flow_reset(mrlbm_ctx* c) {
    if (!c) return fail(MRLBM_EINVAL, "null context");
    HIPCHK(hipSetDevice(c->p.device));
    if (c->d_mean) HIPCHK(hipMemsetAsync(c->d_mean, 0, sizeof(double) * 4 * c->N, c->stream));
    c->mean_count = 0;
    return MRLBM_OK;
}

int mrlbm_mean_flow_accumulate(mrlbm_ctx* c) {
    if (!c) return fail(MRLBM_EINVAL, "null context");
    if (!c->initialised) return fail(MRLBM_ESTATE, "lattice not initialised");
    HIPCHK(hipSetDevice(c->p.device));
    int rc = ensure_mean(c);
    if (rc) return rc;
    const Layout L = layout_of(c);
    const Scales sc = scales_of(c);
    MRLBM_DIAG_LAUNCH(k_mean_accumulate, dim3(blocks_for(c->N)), dim3(DIAG_THREADS), c->d_mom, L, sc,
                      (float)c->p.rho0, c->d_mean);
    c->mean_count++;
    return MRLBM_OK;
}

int mrlbm_mean_flow_read(mrlbm_ctx* c, double* out, int64_t* count) {
    if (!c || !out || !count) return fail(MRLBM_EINVAL, "null argument");
    if (!c->d_mean || c->mean_count == 0) return fail(MRLBM_ESTATE, "no mean-flow samples accumulated");
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(hipMemcpyAsync(out, c->d_mean, sizeof(double) * 4 * c->N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double inv = 1.0 / (double)c->mean_count;
    for (size_t j = 0; j < 4 * c->N; j++) out[j] *= inv;
    *count = c->mean_count;
    return MRLBM_OK;
}

int mrlbm_turbulent_ke(mrlbm_ctx* c, double* tke) {
    if (!c || !tke) return fail(MRLBM_EINVAL, "null argument");
    if (!c->d_mean || c->mean_count == 0) return fail(MRLBM_ESTATE, "no mean-flow samples accumulated");
    HIPCHK(hipSetDevice(c->p.device));
    const int nb = c->red_blocks;
    const Layout L = layout_of(c);
    const Scales sc = scales_of(c);
    MRLBM_DIAG_LAUNCH(k_tke_partial, dim3(nb), dim3(DIAG_THREADS), c->d_mom, L, sc, (float)c->p.rho0, c->d_mean,
                      1.0 / (double)c->mean_count, c->d_partial);
    std::vector<double> p(nb);
    HIPCHK(hipMemcpyAsync(p.data(), c->d_partial, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *tke = fold(p, 1, 0) / (double)c->N;
    return MRLBM_OK;
}

}  // extern "C"
