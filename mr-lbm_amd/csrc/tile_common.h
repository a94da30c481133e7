// Helpers shared by the step kernels: moment load/store in the block-major
// SoA layout and ghost-plane reads/writes in the reference layout
// (src/globalFunctions.h:31-124, src/includeFiles/popLoad.inc,
// src/includeFiles/popSave.inc), for a tile of BX x BY x BZ nodes.
#pragma once

#include "kernels.h"
#include "static_for.h"

namespace mrlbm {

// Source of a ghost-plane population: neighbour-tile coordinate along one
// tangential axis. `t` is the local coordinate, `c` the lattice velocity
// component, B the tile extent, b the tile index, nb the tile count.
__device__ __forceinline__ void ghost_src(int t, int c, int B, int b, int nb, int& ts, int& bs) {
    ts = t - c;
    bs = b;
    if (ts < 0) {
        ts += B;
        bs = (b == 0) ? nb - 1 : b - 1;
    } else if (ts >= B) {
        ts -= B;
        bs = (b == nb - 1) ? 0 : b + 1;
    }
}

template <int Q, int BX, int BY, int BZ, int S>
__device__ __forceinline__ void load_node(const StepArgs& a, size_t mb, Node& n) {
    using C = Codec<S>;
    constexpr size_t T = (size_t)BX * BY * BZ;
    const typename C::T* mom = (const typename C::T*)a.mom;
    n.drho = C::ld(mom, mb + 0 * T, a.mrscale[0]);
    n.rho = a.rho0 + n.drho;
    n.ux = C::ld(mom, mb + 1 * T, a.mrscale[1]);
    n.uy = C::ld(mom, mb + 2 * T, a.mrscale[2]);
    n.uz = C::ld(mom, mb + 3 * T, a.mrscale[3]);
    n.mxx = C::ld(mom, mb + 4 * T, a.mrscale[4]);
    n.mxy = C::ld(mom, mb + 5 * T, a.mrscale[5]);
    n.mxz = C::ld(mom, mb + 6 * T, a.mrscale[6]);
    n.myy = C::ld(mom, mb + 7 * T, a.mrscale[7]);
    n.myz = C::ld(mom, mb + 8 * T, a.mrscale[8]);
    n.mzz = C::ld(mom, mb + 9 * T, a.mrscale[9]);
}

template <int Q, int BX, int BY, int BZ, int S>
__device__ __forceinline__ void store_node(const StepArgs& a, size_t mb, const Node& n) {
    using C = Codec<S>;
    constexpr size_t T = (size_t)BX * BY * BZ;
    typename C::T* mom = (typename C::T*)a.mom;
    C::st(mom, mb + 0 * T, n.drho, a.mscale[0]);
    C::st(mom, mb + 1 * T, n.ux, a.mscale[1]);
    C::st(mom, mb + 2 * T, n.uy, a.mscale[2]);
    C::st(mom, mb + 3 * T, n.uz, a.mscale[3]);
    C::st(mom, mb + 4 * T, n.mxx, a.mscale[4]);
    C::st(mom, mb + 5 * T, n.mxy, a.mscale[5]);
    C::st(mom, mb + 6 * T, n.mxz, a.mscale[6]);
    C::st(mom, mb + 7 * T, n.myy, a.mscale[7]);
    C::st(mom, mb + 8 * T, n.myz, a.mscale[8]);
    C::st(mom, mb + 9 * T, n.mzz, a.mscale[9]);
}

// Write the outgoing populations of face FACE of node (u, v) of the face.
template <int Q, int BX, int BY, int BZ, int FACE>
__device__ __forceinline__ void put_face_c(float* __restrict__ g, int u, int v, int tile, const float* pop) {
    constexpr int QF = VS<Q>::QF;
    static_for<QF>([&](auto SI) {
        constexpr int s = decltype(SI)::value;
        constexpr int p = face_pop(FACE, s);
        size_t idx;
        if constexpr (FACE < 2)
            idx = gidx_x(u, v, s, tile, BY, BZ, QF);
        else if constexpr (FACE < 4)
            idx = gidx_y(u, v, s, tile, BX, BZ, QF);
        else
            idx = gidx_z(u, v, s, tile, BX, BY, QF);
        g[idx] = pop[p];
    });
}

// Read the incoming populations through face FACE of this tile. FACE is the
// neighbour's outgoing face: a node at tx==0 reads X_1 of tile bx-1, etc.
template <int Q, int BX, int BY, int BZ, int FACE>
__device__ __forceinline__ void get_face(const StepArgs& a, int tx, int ty, int tz, int bx, int by, int bz,
                                         float* pop) {
    constexpr int QF = VS<Q>::QF;
    const float* __restrict__ g = a.gin[FACE];
    static_for<QF>([&](auto SI) {
        constexpr int s = decltype(SI)::value;
        constexpr int p = face_pop(FACE, s);
        int ux_, bxs, uy_, bys, uz_, bzs;
        ghost_src(tx, cxv(p), BX, bx, a.ntx, ux_, bxs);
        ghost_src(ty, cyv(p), BY, by, a.nty, uy_, bys);
        ghost_src(tz, czv(p), BZ, bz, a.ntz, uz_, bzs);
        const int st = bxs + a.ntx * (bys + a.nty * bzs);
        size_t idx;
        if constexpr (FACE < 2)
            idx = gidx_x(uy_, uz_, s, st, BY, BZ, QF);
        else if constexpr (FACE < 4)
            idx = gidx_y(ux_, uz_, s, st, BX, BZ, QF);
        else
            idx = gidx_z(ux_, uy_, s, st, BX, BY, QF);
        pop[p] = g[idx];
    });
}

}  // namespace mrlbm
