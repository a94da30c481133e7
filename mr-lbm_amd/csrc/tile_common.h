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
    const auto mom = gptr((const typename C::T*)a.mom);
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

// Node moments from the raw stored values (fetched earlier).
template <int S>
__device__ __forceinline__ void decode_node(const StepArgs& a, const typename Codec<S>::T* m, Node& n) {
    using C = Codec<S>;
    n.drho = C::dec(m[0], a.mrscale[0]);
    n.rho = a.rho0 + n.drho;
    n.ux = C::dec(m[1], a.mrscale[1]);
    n.uy = C::dec(m[2], a.mrscale[2]);
    n.uz = C::dec(m[3], a.mrscale[3]);
    n.mxx = C::dec(m[4], a.mrscale[4]);
    n.mxy = C::dec(m[5], a.mrscale[5]);
    n.mxz = C::dec(m[6], a.mrscale[6]);
    n.myy = C::dec(m[7], a.mrscale[7]);
    n.myz = C::dec(m[8], a.mrscale[8]);
    n.mzz = C::dec(m[9], a.mrscale[9]);
}

// Round the post-collision moments to the storage format and replace them
// by the values the storage will hold, so the populations sent through the
// ghost planes are the ones the neighbour would reconstruct from the stored
// moments: results do not depend on the tile shape. FP32 storage is exact.
// Stochastic rounding draws 8 random bits per value from three words:
// one counter-based hash (sr_bits) and two xorshift successors.
__device__ __forceinline__ uint32_t xs32(uint32_t x) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

template <int S>
__device__ __forceinline__ void quantize(Node& n, const StepArgs& a, uint32_t step, uint32_t node_id,
                                         typename Codec<S>::T* q) {
    using C = Codec<S>;
    if constexpr (S == 0) {
        q[0] = n.drho;
        q[1] = n.ux;
        q[2] = n.uy;
        q[3] = n.uz;
        q[4] = n.mxx;
        q[5] = n.mxy;
        q[6] = n.mxz;
        q[7] = n.myy;
        q[8] = n.myz;
        q[9] = n.mzz;
    } else {
        const uint32_t w0 = sr_bits(node_id, step, a.seed);
        const uint32_t w1 = xs32(w0), w2 = xs32(w1);
        constexpr uint32_t SH = S == 1 ? 5u : 8u;            // random byte -> dropped-bit position
        constexpr uint32_t HALF = S == 1 ? 16u : 128u;       // centre of the byte's interval
        auto rb = [&](uint32_t w, int k) { return (((w >> (8 * k)) & 0xFFu) << SH) | HALF; };
#define MRLBM_Q(K, FIELD, W, B)                          \
    q[K] = C::enc(n.FIELD, a.mscale[K], rb(W, B));       \
    n.FIELD = C::dec(q[K], a.mrscale[K]);
        MRLBM_Q(0, drho, w0, 0)
        MRLBM_Q(1, ux, w0, 1)
        MRLBM_Q(2, uy, w0, 2)
        MRLBM_Q(3, uz, w0, 3)
        MRLBM_Q(4, mxx, w1, 0)
        MRLBM_Q(5, mxy, w1, 1)
        MRLBM_Q(6, mxz, w1, 2)
        MRLBM_Q(7, myy, w1, 3)
        MRLBM_Q(8, myz, w2, 0)
        MRLBM_Q(9, mzz, w2, 1)
#undef MRLBM_Q
        n.rho = a.rho0 + n.drho;
    }
}

template <int Q, int BX, int BY, int BZ, int S>
__device__ __forceinline__ void store_node(const StepArgs& a, size_t mb, const typename Codec<S>::T* q) {
    constexpr size_t T = (size_t)BX * BY * BZ;
    const auto mom = gptr((typename Codec<S>::T*)a.mom);
#pragma unroll
    for (int k = 0; k < 10; k++) mom[mb + k * T] = q[k];
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
    const auto g = gptr(a.gin[FACE]);
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
