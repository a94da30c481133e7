// Per-storage launchers of the tile kernels. Each storage format's kernel
// instantiations live in their own translation unit (step_tile_fp32.hip,
// step_tile_fp16.hip, step_tile_bf16.hip) so that they compile in parallel;
// step_tile.hip dispatches on the storage format.
#pragma once

#include "kernels.h"

// Tile shapes the tile and batched kernels are built for.
#define MRLBM_TILE_SHAPES(X) \
    X(8, 8, 8)               \
    X(16, 8, 4)              \
    X(32, 4, 4)              \
    X(64, 4, 2)              \
    X(16, 8, 8)

namespace mrlbm {

template <int S>
hipError_t launch_step_tile_s(int q, int bx, int by, int bz, bool has_bc, const StepArgs& a, const uint32_t* list,
                              int n, hipStream_t s);
template <int S>
hipError_t launch_step_batch_s(int q, int bx, int by, int bz, bool has_bc, const InstArgs* tab, const uint32_t* work,
                               int nwork, int grid, int k, hipStream_t s);
template <int S>
bool mixed_launch_s(int q, int bx, int by, int bz);
template <int S>
hipError_t launch_step_mixed_s(int q, int bx, int by, int bz, const StepArgs* a, const InstArgs* tab,
                               const uint32_t* list, int n, int nbulk, int k, hipStream_t s);
template <int S>
hipError_t launch_init_ghosts_s(int q, int bx, int by, int bz, const StepArgs& a, const GhostOut& go, int ntiles,
                                hipStream_t s);

}  // namespace mrlbm
