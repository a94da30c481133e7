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

// Node moments from the raw stored values (fetched earlier; 16-bit formats
// arrive zero-extended to 32 bits).
template <int S>
__device__ __forceinline__ void decode_node(const StepArgs& a, const typename Fetched<S>::T* m, Node& n) {
    using C = Codec<S>;
    float* v[10] = {&n.drho, &n.ux, &n.uy, &n.uz, &n.mxx, &n.mxy, &n.mxz, &n.myy, &n.myz, &n.mzz};
#pragma unroll
    for (int k = 0; k < 10; k++) *v[k] = C::dec(m[k], store_scale(S, k) == 1.0f ? 1.0f : 1.0f / store_scale(S, k));
    n.rho = a.rho0 + n.drho;
}

// Round the post-collision moments to the storage format and replace them
// by the values the storage holds, so the populations a node streams are
// the ones its neighbours would reconstruct from the stored moments: results
// do not depend on the tile shape. FP32 storage is exact.
//
// Stochastic rounding by gfx950's conversion instructions
// (v_cvt_sr_f16_f32, v_cvt_sr_bf16_f32). They round up when the dropped
// fraction plus the random operand, read as a fraction of 2^32, reaches one
// (measured on MI355X: for a value halfway between two neighbours only
// operands with bit 31 set round up), so the operand's top bits decide and
// each value needs its own top bits. Ten 32-bit operands come from three
// words: a counter-based hash of (node, step, seed) (sr_bits) and two
// 24-bit multiplicative successors of it (of bytes 0-2 and of bytes 1-3;
// full-rate multiplies), each operand a funnel shift whose top byte is one
// of w0.b3..b0, w1.b3..b1, w2.b3..b1. A successor's low byte depends on a
// single byte of the hash, so it never leads an operand: any two operands'
// top bits are close to independent (chi-square of their top four bits
// ~1.4x its degrees of freedom over 2^20 nodes, where leading with the low
// bytes of full multiplicative successors put three operands' top bytes on
// one byte of the hash). The same node, step and seed round alike in any
// launch shape or batch. binary16 stores v times a power-of-two scale
// (exact) and the moments are decoded back from the packed halves;
// bfloat16 has FP32's exponent range and is stored unscaled.
template <int K>
__device__ __forceinline__ uint32_t sr_operand(uint32_t w0, uint32_t w1, uint32_t w2) {
    // top bytes: w0.b3, w0.b2, w0.b1, w0.b0, w1.b3, w1.b2, w1.b1, w2.b3, w2.b2, w2.b1
    if constexpr (K == 0) return w0;
    else if constexpr (K == 1) return __builtin_amdgcn_alignbit(w0, w1, 24);
    else if constexpr (K == 2) return __builtin_amdgcn_alignbit(w0, w1, 16);
    else if constexpr (K == 3) return __builtin_amdgcn_alignbit(w0, w1, 8);
    else if constexpr (K == 4) return w1;
    else if constexpr (K == 5) return __builtin_amdgcn_alignbit(w1, w2, 24);
    else if constexpr (K == 6) return __builtin_amdgcn_alignbit(w1, w2, 16);
    else if constexpr (K == 7) return w2;
    else if constexpr (K == 8) return __builtin_amdgcn_alignbit(w2, w0, 24);
    else return __builtin_amdgcn_alignbit(w2, w0, 16);
}

// Reconstruction terms straight from binary16 moment words (storage format
// 1): word k holds moment k in its low half (fetched words, PACKED false),
// or moments 2j and 2j+1 in the low and high halves of word j (quantized
// pairs, PACKED true). Each product of a term with a stored moment is one
// v_fma_mix_f32 on the binary16 value, which it converts exactly, with the
// moment's power-of-two storage scale folded into the other factor; sums
// of stored moments are formed the same way. Scaling by a power of two
// commutes with FP32 rounding, so the terms, re0 and drho are bit for bit
// those recon_terms / recon_re0 give for the decoded moments: the step
// kernels and the ghost initialisation form the same populations whichever
// of the two they use.
template <bool HI>
__device__ __forceinline__ float mix_mul(float a, uint32_t b) {  // a * b, one rounding (b: binary16 half)
    // addend -0 (the inline constant 0 with the negate modifier): the result
    // is the product itself, signed zero included
    float r;
    if constexpr (HI)
        asm("v_fma_mix_f32 %0, %1, %2, neg(0) op_sel:[0,1,0] op_sel_hi:[0,1,0]" : "=v"(r) : "v"(a), "v"(b));
    else
        asm("v_fma_mix_f32 %0, %1, %2, neg(0) op_sel_hi:[0,1,0]" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
template <bool HI>
__device__ __forceinline__ float mix_mul_k(float k, uint32_t b) {  // k * b with a uniform factor k (in an SGPR)
    float r;
    if constexpr (HI)
        asm("v_fma_mix_f32 %0, %1, %2, neg(0) op_sel:[0,1,0] op_sel_hi:[0,1,0]" : "=v"(r) : "s"(k), "v"(b));
    else
        asm("v_fma_mix_f32 %0, %1, %2, neg(0) op_sel_hi:[0,1,0]" : "=v"(r) : "s"(k), "v"(b));
    return r;
}
template <bool HA, bool HB>
__device__ __forceinline__ float mix_add(uint32_t a, uint32_t b) {  // a + b of two binary16 halves
    float r;
    if constexpr (!HA && !HB)
        asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(a), "v"(b));
    else if constexpr (HA && !HB)
        asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,1]" : "=v"(r) : "v"(a), "v"(b));
    else if constexpr (!HA && HB)
        asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel:[0,0,1] op_sel_hi:[1,0,1]" : "=v"(r) : "v"(a), "v"(b));
    else
        asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel:[1,0,1] op_sel_hi:[1,0,1]" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
template <bool HB>
__device__ __forceinline__ float mix_addf(float x, uint32_t b) {  // x + b, b a binary16 half
    float r;
    if constexpr (HB)
        asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(b), "v"(x));
    else
        asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(b), "v"(x));
    return r;
}

template <int Q, bool PACKED>
__device__ __forceinline__ void recon_terms_f16(const uint32_t* w, float rho0, float& drho, float& rho, float& re0,
                                                Recon& t) {
    constexpr float sd = store_scale(1, 0), su = store_scale(1, 1), sm = store_scale(1, 4);
    static_assert(store_scale(1, 2) == su && store_scale(1, 3) == su && store_scale(1, 9) == sm, "one scale per group");
    constexpr float w1 = (float)VS<Q>::wd(1), w2 = (float)VS<Q>::wd(7);
    auto W = [&](auto K) { return PACKED ? w[decltype(K)::value >> 1] : w[decltype(K)::value]; };
    auto mul = [&](float a, auto K) {
        return mix_mul<PACKED && (decltype(K)::value & 1)>(a, W(K));
    };
    auto mulk = [&](float k, auto K) {  // k: a compile-time constant
        return mix_mul_k<PACKED && (decltype(K)::value & 1)>(k, W(K));
    };
    using std::integral_constant;
    const integral_constant<int, 0> k0;
    const integral_constant<int, 1> k1;
    const integral_constant<int, 2> k2;
    const integral_constant<int, 3> k3;
    const integral_constant<int, 4> k4;
    const integral_constant<int, 5> k5;
    const integral_constant<int, 6> k6;
    const integral_constant<int, 7> k7;
    const integral_constant<int, 8> k8;
    const integral_constant<int, 9> k9;
    drho = mulk(1.0f / sd, k0);
    rho = rho0 + drho;
    const float ru = rho * (1.0f / su), rm = rho * (1.0f / sm);
    // (mxx + myy) + mzz in the storage scale, then as recon_re0
    const float s2 = mix_add<PACKED && (4 & 1), PACKED && (7 & 1)>(W(k4), W(k7));
    const float s3 = mix_addf<PACKED && (9 & 1)>(s2, W(k9));
    re0 = rm * (-(1.0f / 3.0f) * s3);
    t.rux = mul(ru, k1);
    t.ruy = mul(ru, k2);
    t.ruz = mul(ru, k3);
    t.ex = re0 + mul(rm, k4);
    t.ey = re0 + (t.ryy = mul(rm, k7));
    t.ez = re0 + (t.rzz = mul(rm, k9));
    t.rxy = mul(rm, k5);
    t.rxz = mul(rm, k6);
    t.ryz = mul(rm, k8);
    t.d1 = mulk(w1 / sd, k0);
    t.d2 = mulk(w2 / sd, k0);
    if constexpr (Q == 27) t.d3 = mulk((float)VS<Q>::wd(26) / sd, k0);
    else t.d3 = 0.0f;
}

// FP16 storage: the moments U and M arrive already times their storage
// scales (bulk_moments / scale_moments / collide carry them), so only the
// density deviation is multiplied before rounding; pk[j] receives the
// rounded pair (2j, 2j+1) for recon_terms_f16, and n.drho, n.rho the
// stored density. FP32 and BF16 (scale 1): n is replaced by the stored
// values.
template <int S>
__device__ __forceinline__ void quantize(Node& n, const StepArgs& a, uint32_t step, uint32_t node_id,
                                         typename Codec<S>::T* q, uint32_t* pk = nullptr) {
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
        typedef _Float16 __attribute__((ext_vector_type(2))) half2_t;
        typedef __bf16 __attribute__((ext_vector_type(2))) bf162_t;
        const uint32_t w0 = sr_bits(node_id, step, a.seed);
        const uint32_t w1 = __umul24(w0, 0x9E3779u), w2 = __umul24(w0 >> 8, 0xEBCA77u);
        float* v[10] = {&n.drho, &n.ux, &n.uy, &n.uz, &n.mxx, &n.mxy, &n.mxz, &n.myy, &n.myz, &n.mzz};
        static_for<5>([&](auto PP) {
            constexpr int k = 2 * decltype(PP)::value;
            const uint32_t r0 = sr_operand<k>(w0, w1, w2), r1 = sr_operand<k + 1>(w0, w1, w2);
            uint32_t b;
            if constexpr (S == 1) {
                // both halves are written: start from any register (no zeroing move)
                half2_t h = __builtin_bit_cast(half2_t, r0);
                const float x0 = k == 0 ? *v[k] * store_scale(1, 0) : *v[k];  // U, M: already scaled
                h = __builtin_amdgcn_cvt_sr_f16_f32(h, x0, r0, false);
                h = __builtin_amdgcn_cvt_sr_f16_f32(h, *v[k + 1], r1, true);
                b = __builtin_bit_cast(uint32_t, h);
                pk[k / 2] = b;
            } else {
                bf162_t h = __builtin_bit_cast(bf162_t, r0);
                h = __builtin_amdgcn_cvt_sr_bf16_f32(h, *v[k], r0, false);
                h = __builtin_amdgcn_cvt_sr_bf16_f32(h, *v[k + 1], r1, true);
                b = __builtin_bit_cast(uint32_t, h);
                *v[k] = __uint_as_float(b << 16);
                *v[k + 1] = __uint_as_float(b & 0xFFFF0000u);
            }
            q[k] = (uint16_t)(b & 0xFFFFu);
            q[k + 1] = (uint16_t)(b >> 16);
        });
        if constexpr (S == 2) n.rho = a.rho0 + n.drho;
    }
}

// Stored moments of node t of tile `tile`: a uniform base per plane and the
// lane's 32-bit offset (no vector address arithmetic).
template <int BX, int BY, int BZ, int S>
__device__ __forceinline__ void store_tile_node(const StepArgs& a, int tile, int t, const typename Codec<S>::T* q) {
    constexpr size_t T = (size_t)BX * BY * BZ;
    const auto mom = gptr((typename Codec<S>::T*)a.mom) + (size_t)tile * 10 * T;
#pragma unroll
    for (int k = 0; k < 10; k++) __builtin_nontemporal_store(q[k], &sbase(mom + k * T)[(uint32_t)t]);  // streamed once per step
}

template <int Q, int BX, int BY, int BZ, int S>
__device__ __forceinline__ void store_node(const StepArgs& a, size_t mb, const typename Codec<S>::T* q) {
    constexpr size_t T = (size_t)BX * BY * BZ;
    const auto mom = gptr((typename Codec<S>::T*)a.mom);
#pragma unroll
    for (int k = 0; k < 10; k++) __builtin_nontemporal_store(q[k], &mom[mb + k * T]);  // streamed once per step
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

// Read the incoming populations through face FACE of this tile, in slot
// order (vals[s] is population face_pop(FACE, s)). FACE is the neighbour's
// outgoing face: a node at tx==0 reads X_1 of tile bx-1, etc.
template <int Q, int BX, int BY, int BZ, int FACE>
__device__ __forceinline__ void load_face(const StepArgs& a, int tx, int ty, int tz, int bx, int by, int bz,
                                          float* vals) {
    constexpr int QF = VS<Q>::QF;
    const auto g = gptr(ghost_face<BX, BY, BZ, QF>(a.gin, FACE, a.ntiles));
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
        vals[s] = g[idx];
    });
}

// File face values (load_face order) into the population array.
template <int Q, int FACE>
__device__ __forceinline__ void commit_face(const float* vals, float* pop) {
    static_for<VS<Q>::QF>([&](auto SI) {
        constexpr int s = decltype(SI)::value;
        pop[face_pop(FACE, s)] = vals[s];
    });
}

template <int Q, int BX, int BY, int BZ, int FACE>
__device__ __forceinline__ void get_face(const StepArgs& a, int tx, int ty, int tz, int bx, int by, int bz,
                                         float* pop) {
    float vals[VS<Q>::QF];
    load_face<Q, BX, BY, BZ, FACE>(a, tx, ty, tz, bx, by, bz, vals);
    commit_face<Q, FACE>(vals, pop);
}

}  // namespace mrlbm
