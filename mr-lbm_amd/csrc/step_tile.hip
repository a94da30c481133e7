// Tile step kernel: one node per thread, one (BX,BY,BZ) tile per workgroup.
//
// Same data flow as the reference's gpuMomCollisionStream
// (src/mlbm.cu:3-714), written for gfx950:
//   1. load the 10 stored moments (block-major SoA: each wave64 reads
//      contiguous 256-byte runs per moment plane)
//   2. reconstruct the post-collision populations of the previous step
//      (as deviations g_i = f_i - w_i RHO_0, see lattice.h)
//   3. write populations 1..Q-1 to LDS planes, barrier, pull from the
//      in-tile upwind neighbour (periodic inside the tile)
//   4. tile-face nodes overwrite the populations that crossed the tile
//      face with the values from the neighbour tile's ghost planes
//      (src/includeFiles/popLoad.inc:2-86; X, then Y, then Z branch, the
//      later branch winning as in the reference)
//   5. boundary class -> moment-based BC, else bulk moments; scale; collide
//   6. store moments in place, reconstruct, write outgoing ghost planes
//      (src/includeFiles/popSave.inc:2-78 with the wall predicates of
//      src/includeFiles/interface.h:6-47)
// Solid nodes run the same code (no early exit before the barrier) and
// simply store nothing.
#include "tile_common.h"

namespace mrlbm {

template <int Q, int BX, int BY, int BZ, bool HAS_BC, int S>
__global__ __launch_bounds__(BX* BY* BZ) void k_step_tile(StepArgs a) {
    constexpr int T = BX * BY * BZ;
    __shared__ float s_pop[(Q - 1) * T];

    const int t = threadIdx.x;
    const int tx = t % BX, ty = (t / BX) % BY, tz = t / (BX * BY);
    const int tile = blockIdx.x;
    const int bx = tile % a.ntx, by = (tile / a.ntx) % a.nty, bz = tile / (a.ntx * a.nty);
    const int x = bx * BX + tx, y = by * BY + ty, z = bz * BZ + tz;

    uint32_t cls = 0;
    if constexpr (HAS_BC) cls = a.cls[(size_t)tile * T + t];
    const bool solid = HAS_BC && (cls == 255u);

    const size_t mb = (size_t)tile * 10 * T + t;
    Node n;
    load_node<Q, BX, BY, BZ, S>(a, mb, n);

    float pop[Q];
    reconstruct<Q>(n, pop);
    static_for<Q - 1>([&](auto I) {
        constexpr int i = decltype(I)::value + 1;
        s_pop[(i - 1) * T + t] = pop[i];
    });
    __syncthreads();
    static_for<Q - 1>([&](auto I) {
        constexpr int i = decltype(I)::value + 1;
        const int sx = (tx - cxv(i) + BX) % BX;
        const int sy = (ty - cyv(i) + BY) % BY;
        const int sz = (tz - czv(i) + BZ) % BZ;
        pop[i] = s_pop[(i - 1) * T + sx + BX * (sy + BY * sz)];
    });

    if (tx == 0)
        get_face<Q, BX, BY, BZ, 1>(a, tx, ty, tz, bx, by, bz, pop);
    else if (tx == BX - 1)
        get_face<Q, BX, BY, BZ, 0>(a, tx, ty, tz, bx, by, bz, pop);
    if (ty == 0)
        get_face<Q, BX, BY, BZ, 3>(a, tx, ty, tz, bx, by, bz, pop);
    else if (ty == BY - 1)
        get_face<Q, BX, BY, BZ, 2>(a, tx, ty, tz, bx, by, bz, pop);
    if (tz == 0)
        get_face<Q, BX, BY, BZ, 5>(a, tx, ty, tz, bx, by, bz, pop);
    else if (tz == BZ - 1)
        get_face<Q, BX, BY, BZ, 4>(a, tx, ty, tz, bx, by, bz, pop);

    Node m;
    if (HAS_BC && cls != 0u && cls != 255u) {
        bc_solve<Q>(a.bc + (size_t)cls * BC_STRIDE, pop, a.rho0, m);
        scale_moments(m);
    } else {
        bulk_moments<Q>(pop, a.rho0, m);
    }
    collide(m, a.R);
    reconstruct<Q>(m, pop);

    if (!solid) {
        store_node<Q, BX, BY, BZ, S>(a, mb, m);
        if (tx == 0 && (!a.wall_x || x != 0)) put_face_c<Q, BX, BY, BZ, 0>(a.gout[0], ty, tz, tile, pop);
        if (tx == BX - 1 && (!a.wall_x || x != a.nx - 1)) put_face_c<Q, BX, BY, BZ, 1>(a.gout[1], ty, tz, tile, pop);
        if (ty == 0 && (!a.wall_y || y != 0)) put_face_c<Q, BX, BY, BZ, 2>(a.gout[2], tx, tz, tile, pop);
        if (ty == BY - 1 && (!a.wall_y || y != a.ny - 1)) put_face_c<Q, BX, BY, BZ, 3>(a.gout[3], tx, tz, tile, pop);
        if (tz == 0 && (!a.wall_z || z != 0)) put_face_c<Q, BX, BY, BZ, 4>(a.gout[4], tx, ty, tile, pop);
        if (tz == BZ - 1 && (!a.wall_z || z != a.nz - 1)) put_face_c<Q, BX, BY, BZ, 5>(a.gout[5], tx, ty, tile, pop);
    }
}

// Initial ghost planes: every tile-face node writes its outgoing
// populations, reconstructed from the stored moments, on every face.
template <int Q, int BX, int BY, int BZ, int S>
__global__ __launch_bounds__(BX* BY* BZ) void k_init_ghosts(StepArgs a, GhostOut g) {
    constexpr int T = BX * BY * BZ;
    const int t = threadIdx.x;
    const int tx = t % BX, ty = (t / BX) % BY, tz = t / (BX * BY);
    const int tile = blockIdx.x;
    Node n;
    load_node<Q, BX, BY, BZ, S>(a, (size_t)tile * 10 * T + t, n);
    float pop[Q];
    reconstruct<Q>(n, pop);
    if (tx == 0) put_face_c<Q, BX, BY, BZ, 0>(g.f[0], ty, tz, tile, pop);
    if (tx == BX - 1) put_face_c<Q, BX, BY, BZ, 1>(g.f[1], ty, tz, tile, pop);
    if (ty == 0) put_face_c<Q, BX, BY, BZ, 2>(g.f[2], tx, tz, tile, pop);
    if (ty == BY - 1) put_face_c<Q, BX, BY, BZ, 3>(g.f[3], tx, tz, tile, pop);
    if (tz == 0) put_face_c<Q, BX, BY, BZ, 4>(g.f[4], tx, ty, tile, pop);
    if (tz == BZ - 1) put_face_c<Q, BX, BY, BZ, 5>(g.f[5], tx, ty, tile, pop);
}

// ---------------------------------------------------------------------------
// dispatch

#define MRLBM_TILE_SHAPES(X) \
    X(8, 8, 8)               \
    X(16, 8, 4)              \
    X(32, 4, 4)              \
    X(64, 4, 2)

bool tile_shape_supported(int q, int bx, int by, int bz, int storage) {
    if (q != 19 && q != 27) return false;
    if (storage != 0) return false;
#define X(A, B, C) \
    if (bx == A && by == B && bz == C) return true;
    MRLBM_TILE_SHAPES(X)
#undef X
    return false;
}

template <int Q, int BX, int BY, int BZ>
static hipError_t launch_q_shape(bool has_bc, int storage, const StepArgs& a, int ntiles, hipStream_t s) {
    dim3 grid(ntiles), block(BX * BY * BZ);
    if (storage == 0) {
        if (has_bc)
            hipLaunchKernelGGL((k_step_tile<Q, BX, BY, BZ, true, 0>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((k_step_tile<Q, BX, BY, BZ, false, 0>), grid, block, 0, s, a);
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

hipError_t launch_step_tile(int q, int bx, int by, int bz, bool has_bc, int storage, const StepArgs& a, int ntiles,
                            hipStream_t s) {
#define X(A, B, C)                                                                    \
    if (bx == A && by == B && bz == C) {                                              \
        if (q == 19) return launch_q_shape<19, A, B, C>(has_bc, storage, a, ntiles, s); \
        if (q == 27) return launch_q_shape<27, A, B, C>(has_bc, storage, a, ntiles, s); \
    }
    MRLBM_TILE_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t launch_init_ghosts(int q, int bx, int by, int bz, int storage, const StepArgs& a, float* const g[6],
                              int ntiles, hipStream_t s) {
    GhostOut go;
    for (int f = 0; f < 6; f++) go.f[f] = g[f];
    dim3 grid(ntiles), block(bx * by * bz);
    if (storage != 0) return hipErrorInvalidValue;
#define X(A, B, C)                                                                                   \
    if (bx == A && by == B && bz == C) {                                                             \
        if (q == 19) hipLaunchKernelGGL((k_init_ghosts<19, A, B, C, 0>), grid, block, 0, s, a, go); \
        else hipLaunchKernelGGL((k_init_ghosts<27, A, B, C, 0>), grid, block, 0, s, a, go);         \
        return hipGetLastError();                                                                    \
    }
    MRLBM_TILE_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace mrlbm
