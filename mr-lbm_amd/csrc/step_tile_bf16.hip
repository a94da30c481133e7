// Tile and batched step kernels with bfloat16 moment storage (see
// step_tile_body.h); one translation unit per storage format.
#include "step_tile_body.h"

namespace mrlbm {

template hipError_t launch_step_tile_s<2>(int, int, int, int, bool, const StepArgs&, const uint32_t*, int, hipStream_t);
template hipError_t launch_step_batch_s<2>(int, int, int, int, bool, const InstArgs*, const uint32_t*, int, int, int,
                                            hipStream_t);
template bool mixed_launch_s<2>(int, int, int, int);
template hipError_t launch_step_mixed_s<2>(int, int, int, int, const StepArgs*, const InstArgs*, const uint32_t*, int,
                                          int, int, hipStream_t);
template hipError_t launch_init_ghosts_s<2>(int, int, int, int, const StepArgs&, const GhostOut&, int, hipStream_t);

}  // namespace mrlbm
