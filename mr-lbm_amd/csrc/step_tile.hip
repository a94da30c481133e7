// Storage-format dispatch of the tile and batched step kernels (the
// kernels themselves: step_tile_body.h, instantiated per storage format in
// step_tile_fp32.hip, step_tile_fp16.hip and step_tile_bf16.hip), and the
// batch's per-member step/parity advance.
#include "step_tile_launch.h"

namespace mrlbm {

__global__ void k_batch_advance(InstArgs* tab, int ninst, int nsteps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ninst) {
        tab[i].step0 += (uint32_t)nsteps;
        tab[i].cur0 = (tab[i].cur0 + nsteps) & 1;
    }
}

hipError_t launch_batch_advance(InstArgs* tab, int ninst, int nsteps, hipStream_t s) {
    if (ninst <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_advance, dim3((ninst + 255) / 256), dim3(256), 0, s, tab, ninst, nsteps);
    return hipGetLastError();
}

bool tile_shape_supported(int q, int bx, int by, int bz, int storage) {
    if (q != 19 && q != 27) return false;
    if (storage < 0 || storage > 2) return false;
#define X(A, B, C) \
    if (bx == A && by == B && bz == C) return true;
    MRLBM_TILE_SHAPES(X)
#undef X
    return false;
}

hipError_t launch_step_tile(int q, int bx, int by, int bz, bool has_bc, int storage, const StepArgs& a,
                            const uint32_t* list, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    switch (storage) {
        case 0: return launch_step_tile_s<0>(q, bx, by, bz, has_bc, a, list, n, s);
        case 1: return launch_step_tile_s<1>(q, bx, by, bz, has_bc, a, list, n, s);
        case 2: return launch_step_tile_s<2>(q, bx, by, bz, has_bc, a, list, n, s);
        default: return hipErrorInvalidValue;
    }
}

bool mixed_launch_supported(int q, int bx, int by, int bz, int storage) {
    switch (storage) {
        case 0: return mixed_launch_s<0>(q, bx, by, bz);
        case 1: return mixed_launch_s<1>(q, bx, by, bz);
        case 2: return mixed_launch_s<2>(q, bx, by, bz);
        default: return false;
    }
}

hipError_t launch_step_tile_mixed(int q, int bx, int by, int bz, int storage, const StepArgs& a, const uint32_t* list,
                                  int n, int nbulk, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    switch (storage) {
        case 0: return launch_step_mixed_s<0>(q, bx, by, bz, &a, nullptr, list, n, nbulk, 0, s);
        case 1: return launch_step_mixed_s<1>(q, bx, by, bz, &a, nullptr, list, n, nbulk, 0, s);
        case 2: return launch_step_mixed_s<2>(q, bx, by, bz, &a, nullptr, list, n, nbulk, 0, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_step_batch_mixed(int q, int bx, int by, int bz, int storage, const InstArgs* tab,
                                   const uint32_t* work, int nwork, int nbulk, int k, hipStream_t s) {
    if (nwork <= 0) return hipSuccess;
    switch (storage) {
        case 0: return launch_step_mixed_s<0>(q, bx, by, bz, nullptr, tab, work, nwork, nbulk, k, s);
        case 1: return launch_step_mixed_s<1>(q, bx, by, bz, nullptr, tab, work, nwork, nbulk, k, s);
        case 2: return launch_step_mixed_s<2>(q, bx, by, bz, nullptr, tab, work, nwork, nbulk, k, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_init_ghosts(int q, int bx, int by, int bz, int storage, const StepArgs& a, float* const g[6],
                              int ntiles, hipStream_t s) {
    GhostOut go;
    for (int f = 0; f < 6; f++) go.f[f] = g[f];
    switch (storage) {
        case 0: return launch_init_ghosts_s<0>(q, bx, by, bz, a, go, ntiles, s);
        case 1: return launch_init_ghosts_s<1>(q, bx, by, bz, a, go, ntiles, s);
        case 2: return launch_init_ghosts_s<2>(q, bx, by, bz, a, go, ntiles, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_step_batch(int q, int bx, int by, int bz, int storage, bool has_bc, const InstArgs* tab,
                             const uint32_t* work, int nwork, int grid, int k, hipStream_t s) {
    if (nwork <= 0) return hipSuccess;
    switch (storage) {
        case 0: return launch_step_batch_s<0>(q, bx, by, bz, has_bc, tab, work, nwork, grid, k, s);
        case 1: return launch_step_batch_s<1>(q, bx, by, bz, has_bc, tab, work, nwork, grid, k, s);
        case 2: return launch_step_batch_s<2>(q, bx, by, bz, has_bc, tab, work, nwork, grid, k, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace mrlbm
