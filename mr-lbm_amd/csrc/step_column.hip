// Column step kernel: one (x, y) column per thread, the workgroup marching
// through a z-segment of BZ planes of a (BX, BY) in-plane tile.
//
// The data flow per node is the reference's gpuMomCollisionStream
// (src/mlbm.cu:3-714) and the per-node arithmetic is the shared code of
// lattice.h, so results are bit-identical to the tile kernel. What changes
// is how populations meet their destination:
//   - populations moving within the plane go through one LDS plane per
//     z-step (double-buffered: one barrier per plane);
//   - populations moving in z stay in registers: when plane k+1 is
//     reconstructed, each thread files what it pulled from LDS into the
//     incoming arrays of planes k (c_z = -1), k+1 (c_z = 0) and k+2
//     (c_z = +1), so plane k is complete once plane k+1 has been processed;
//   - only the segment's end planes read/write Z ghost planes, so the ghost
//     traffic per update is 4 QF s (1/BX + 1/BY + 1/BZ) with a long BZ.
// Moments are read and written in the block-major SoA layout of the
// reference with tile (BX, BY, BZ): a z-plane of one moment is BX*BY
// contiguous values (one 1 KiB run per 256-thread workgroup). Moments of
// plane k+2 are prefetched while plane k is computed.
#include "tile_common.h"

namespace mrlbm {

// Populations that move within the plane ((c_x, c_y) != (0, 0)), in order.
template <int Q>
__host__ __device__ constexpr int n_inplane() {
    int n = 0;
    for (int i = 0; i < Q; i++)
        if (cxv(i) != 0 || cyv(i) != 0) n++;
    return n;
}
template <int Q>
__host__ __device__ constexpr int inplane_pop(int j) {
    int n = 0;
    for (int i = 0; i < Q; i++)
        if (cxv(i) != 0 || cyv(i) != 0) {
            if (n == j) return i;
            n++;
        }
    return -1;
}

template <int Q, int BX, int BY, int BZ, bool HAS_BC, int S>
__global__ __launch_bounds__(BX* BY) void k_step_column(StepArgs a) {
    constexpr int P = BX * BY;             // nodes per plane
    constexpr size_t T = (size_t)P * BZ;  // nodes per tile
    constexpr int NJ = n_inplane<Q>();
    __shared__ float s_pl[2][NJ][P];

    const int t = threadIdx.x;
    const int tx = t % BX, ty = t / BX;
    const int tile = blockIdx.x;
    const uint32_t tq = a.by_ntx.div((uint32_t)tile), tz = a.by_nty.div(tq);
    const int bx = (int)((uint32_t)tile - tq * (uint32_t)a.ntx), by = (int)(tq - tz * (uint32_t)a.nty), bz = (int)tz;
    const int x = bx * BX + tx, y = by * BY + ty;
    const size_t mbase = (size_t)tile * 10 * T + t;
    const size_t cbase = (size_t)tile * T + t;

    float in0[Q], in1[Q], in2[Q];  // incoming populations of planes k, k+1, k+2
    static_for<Q>([&](auto I) {
        in0[decltype(I)::value] = 0.0f;
        in1[decltype(I)::value] = 0.0f;
        in2[decltype(I)::value] = 0.0f;
    });

    // Scatter the populations of plane p (just reconstructed in g) into the
    // incoming arrays: write in-plane movers to LDS slot `sl`, barrier, pull
    // from the in-plane upwind neighbour. `first` marks the prologue (plane
    // 0: c_z = -1 populations leave the segment and are not kept).
    auto scatter = [&](const float* g, int sl, bool first) {
        static_for<NJ>([&](auto J) {
            constexpr int j = decltype(J)::value;
            s_pl[sl][j][t] = g[inplane_pop<Q>(j)];
        });
        __syncthreads();
        static_for<NJ>([&](auto J) {
            constexpr int j = decltype(J)::value;
            constexpr int i = inplane_pop<Q>(j);
            const int sx = (tx - cxv(i) + BX) % BX;
            const int sy = (ty - cyv(i) + BY) % BY;
            const float v = s_pl[sl][j][sx + BX * sy];
            if constexpr (czv(i) < 0) {
                if (!first) in0[i] = v;
            } else if constexpr (czv(i) == 0) {
                if (first) in0[i] = v; else in1[i] = v;
            } else {
                if (first) in1[i] = v; else in2[i] = v;
            }
        });
        // populations that stay in the column: 0 (rest), 5 (+z), 6 (-z)
        if (first) {
            in0[0] = g[0];
            in1[5] = g[5];
        } else {
            in1[0] = g[0];
            in2[5] = g[5];
            in0[6] = g[6];
        }
    };

    // prologue: plane 0 of the segment
    Node nxt;
    load_node<Q, BX, BY, BZ, S>(a, mbase, nxt);
    {
        float g[Q];
        reconstruct<Q>(nxt, g);
        scatter(g, 0, true);
    }
    if (BZ > 1) load_node<Q, BX, BY, BZ, S>(a, mbase + P, nxt);
    uint32_t cls_next = 0;
    if constexpr (HAS_BC) cls_next = a.cls[cbase];

    // x and y face populations entering plane k, loaded one plane ahead so
    // that their latency overlaps the previous plane's work
    constexpr int QF = VS<Q>::QF;
    float px[QF], py[QF];
    auto fetch_xy = [&](int k) {
        if (tx == 0)
            load_face<Q, BX, BY, BZ, 1>(a, tx, ty, k, bx, by, bz, px);
        else if (tx == BX - 1)
            load_face<Q, BX, BY, BZ, 0>(a, tx, ty, k, bx, by, bz, px);
        if (ty == 0)
            load_face<Q, BX, BY, BZ, 3>(a, tx, ty, k, bx, by, bz, py);
        else if (ty == BY - 1)
            load_face<Q, BX, BY, BZ, 2>(a, tx, ty, k, bx, by, bz, py);
    };
    fetch_xy(0);

    for (int k = 0; k < BZ; k++) {
        const uint32_t cls = cls_next;
        if constexpr (HAS_BC) {
            if (k + 1 < BZ) cls_next = a.cls[cbase + (size_t)(k + 1) * P];
        }
        if (k + 1 < BZ) {
            float g[Q];
            reconstruct<Q>(nxt, g);
            if (k + 2 < BZ) load_node<Q, BX, BY, BZ, S>(a, mbase + (size_t)(k + 2) * P, nxt);
            scatter(g, (k + 1) & 1, false);
        }
        // populations crossing the tile's x/y faces, then the segment's z
        // ends (X, Y, Z order: the later branch wins, as in popLoad.inc)
        if (tx == 0)
            commit_face<Q, 1>(px, in0);
        else if (tx == BX - 1)
            commit_face<Q, 0>(px, in0);
        if (ty == 0)
            commit_face<Q, 3>(py, in0);
        else if (ty == BY - 1)
            commit_face<Q, 2>(py, in0);
        if (k + 1 < BZ) fetch_xy(k + 1);
        if (k == 0) get_face<Q, BX, BY, BZ, 5>(a, tx, ty, k, bx, by, bz, in0);
        if (k == BZ - 1) get_face<Q, BX, BY, BZ, 4>(a, tx, ty, k, bx, by, bz, in0);

        Node m;
        if (HAS_BC && cls != 0u && cls != 255u) {
            bc_solve<Q>(gptr(a.bc) + (size_t)cls * BC_STRIDE, in0, a.rho0, m);
            scale_moments(m);
        } else {
            bulk_moments<Q>(in0, a.rho0, m);
        }
        collide(m, a.R);
        const int z = bz * BZ + k;
        typename Codec<S>::T qv[10];
        quantize<S>(m, a, a.step, (uint32_t)x + (uint32_t)a.nx * ((uint32_t)y + (uint32_t)a.ny * (uint32_t)z), qv);
        float g[Q];
        reconstruct<Q>(m, g);

        const bool solid = HAS_BC && (cls == 255u);
        if (!solid) {
            store_node<Q, BX, BY, BZ, S>(a, mbase + (size_t)k * P, qv);
            if (tx == 0 && (!a.wall_x || x != 0)) put_face_c<Q, BX, BY, BZ, 0>(ghost_face<BX, BY, BZ, VS<Q>::QF>(a.gout, 0, a.ntiles), ty, k, tile, g);
            if (tx == BX - 1 && (!a.wall_x || x != a.nx - 1)) put_face_c<Q, BX, BY, BZ, 1>(ghost_face<BX, BY, BZ, VS<Q>::QF>(a.gout, 1, a.ntiles), ty, k, tile, g);
            if (ty == 0 && (!a.wall_y || y != 0)) put_face_c<Q, BX, BY, BZ, 2>(ghost_face<BX, BY, BZ, VS<Q>::QF>(a.gout, 2, a.ntiles), tx, k, tile, g);
            if (ty == BY - 1 && (!a.wall_y || y != a.ny - 1)) put_face_c<Q, BX, BY, BZ, 3>(ghost_face<BX, BY, BZ, VS<Q>::QF>(a.gout, 3, a.ntiles), tx, k, tile, g);
            if (k == 0 && (!a.wall_z || z != 0)) put_face_c<Q, BX, BY, BZ, 4>(ghost_face<BX, BY, BZ, VS<Q>::QF>(a.gout, 4, a.ntiles), tx, ty, tile, g);
            if (k == BZ - 1 && (!a.wall_z || z != a.nz - 1)) put_face_c<Q, BX, BY, BZ, 5>(ghost_face<BX, BY, BZ, VS<Q>::QF>(a.gout, 5, a.ntiles), tx, ty, tile, g);
        }
        // rotate the incoming arrays
        static_for<Q>([&](auto I) {
            constexpr int i = decltype(I)::value;
            in0[i] = in1[i];
            in1[i] = in2[i];
            in2[i] = 0.0f;
        });
    }
}

// ---------------------------------------------------------------------------
// dispatch: (BX, BY, BZ) with BX*BY = 256 threads

#define MRLBM_COLUMN_SHAPES(X) \
    X(16, 16, 32)              \
    X(16, 16, 64)              \
    X(32, 8, 32)               \
    X(32, 8, 64)               \
    X(64, 4, 64)

bool column_shape_supported(int q, int bx, int by, int bz, int storage) {
    if (q != 19 && q != 27) return false;
    if (storage != 0) return false;
#define X(A, B, C) \
    if (bx == A && by == B && bz == C) return true;
    MRLBM_COLUMN_SHAPES(X)
#undef X
    return false;
}

template <int Q, int BX, int BY, int BZ>
static hipError_t launch_col(bool has_bc, int storage, const StepArgs& a, int ntiles, hipStream_t s) {
    dim3 grid(ntiles), block(BX * BY);
    if (storage == 0) {
        if (has_bc)
            hipLaunchKernelGGL((k_step_column<Q, BX, BY, BZ, true, 0>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((k_step_column<Q, BX, BY, BZ, false, 0>), grid, block, 0, s, a);
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

hipError_t launch_step_column(int q, int bx, int by, int bz, bool has_bc, int storage, const StepArgs& a, int ntiles,
                              hipStream_t s) {
#define X(A, B, C)                                                                 \
    if (bx == A && by == B && bz == C) {                                           \
        if (q == 19) return launch_col<19, A, B, C>(has_bc, storage, a, ntiles, s); \
        if (q == 27) return launch_col<27, A, B, C>(has_bc, storage, a, ntiles, s); \
    }
    MRLBM_COLUMN_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

// Initial ghost planes for a column-tiled lattice: one thread per node of
// the tile (z-segment walked in a loop), every face written.
template <int Q, int BX, int BY, int BZ, int S>
__global__ __launch_bounds__(BX* BY) void k_init_ghosts_column(StepArgs a, GhostOut go) {
    constexpr int P = BX * BY;
    constexpr size_t T = (size_t)P * BZ;
    const int t = threadIdx.x;
    const int tx = t % BX, ty = t / BX;
    const int tile = blockIdx.x;
    for (int k = 0; k < BZ; k++) {
        Node n;
        load_node<Q, BX, BY, BZ, S>(a, (size_t)tile * 10 * T + t + (size_t)k * P, n);
        float g[Q];
        reconstruct<Q>(n, g);
        if (tx == 0) put_face_c<Q, BX, BY, BZ, 0>(go.f[0], ty, k, tile, g);
        if (tx == BX - 1) put_face_c<Q, BX, BY, BZ, 1>(go.f[1], ty, k, tile, g);
        if (ty == 0) put_face_c<Q, BX, BY, BZ, 2>(go.f[2], tx, k, tile, g);
        if (ty == BY - 1) put_face_c<Q, BX, BY, BZ, 3>(go.f[3], tx, k, tile, g);
        if (k == 0) put_face_c<Q, BX, BY, BZ, 4>(go.f[4], tx, ty, tile, g);
        if (k == BZ - 1) put_face_c<Q, BX, BY, BZ, 5>(go.f[5], tx, ty, tile, g);
    }
}

hipError_t launch_init_ghosts_column(int q, int bx, int by, int bz, int storage, const StepArgs& a, float* const g[6],
                                     int ntiles, hipStream_t s) {
    GhostOut go;
    for (int f = 0; f < 6; f++) go.f[f] = g[f];
    dim3 grid(ntiles), block(bx * by);
    if (storage != 0) return hipErrorInvalidValue;
#define X(A, B, C)                                                                                            \
    if (bx == A && by == B && bz == C) {                                                                      \
        if (q == 19) hipLaunchKernelGGL((k_init_ghosts_column<19, A, B, C, 0>), grid, block, 0, s, a, go);   \
        else hipLaunchKernelGGL((k_init_ghosts_column<27, A, B, C, 0>), grid, block, 0, s, a, go);           \
        return hipGetLastError();                                                                             \
    }
    MRLBM_COLUMN_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace mrlbm
