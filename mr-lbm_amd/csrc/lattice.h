// Velocity sets and per-node arithmetic of the moment-representation step.
//
// Everything here is shared by every step-kernel variant (tile kernel,
// column kernel, batched kernel) so that all variants perform the same FP32
// operations in the same order: results are bit-identical across tile
// shapes and between batched and single-instance runs. The library is built
// with -ffp-contract=off, so no variant can fuse differently.
//
// Moment convention (reference src/arrayIndex.h:8-17,186 and
// src/includeFiles/velocitySets/D3Q19.inc:21-24): stored moments are
//   rho - RHO_0, U = 3u, Mxx = 4.5 m_xx, Mxy = 9 m_xy, Mxz = 9 m_xz,
//   Myy = 4.5 m_yy, Myz = 9 m_yz, Mzz = 4.5 m_zz,
// with m_ab = sum_i f_i c_ia c_ib / rho - cs2 delta_ab.
//
// Populations travel (LDS, ghost planes) as deviations from the rest state,
// g_i = f_i - w_i RHO_0, and density as drho = rho - RHO_0. Every sum the
// reference forms from O(1) numbers (rho = sum f, m_aa = sum f c_a^2/rho -
// 1/3) is then formed from small numbers, which removes its FP32
// cancellation; the arithmetic is otherwise the reference's.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <utility>

#include "static_for.h"

namespace mrlbm {

// Global-address-space view of a device pointer. Pointers that kernels read
// from device tables are generic to the compiler, which then emits flat_*
// accesses; those also count against the LDS wait counter, so every LDS
// wait would stall on outstanding HBM traffic. Accesses through this view
// are global_* instructions.
#define MRLBM_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ MRLBM_GLOBAL T* gptr(T* p) {
    return (MRLBM_GLOBAL T*)p;
}

// A uniform global pointer held in scalar registers as an opaque value: an
// access p[lane offset] is then one instruction with the scalar base and the
// lane's 32-bit offset, however far p lies from the kernel's other bases
// (without it, bases that differ by more than the 12-bit immediate offset
// are formed with 64-bit vector adds in every lane).
template <typename T>
__device__ __forceinline__ MRLBM_GLOBAL T* sbase(MRLBM_GLOBAL T* p) {
    asm volatile("" : "+s"(p));
    return p;
}

// Population ordering of src/includeFiles/velocitySets/D3Q27.inc:44-46;
// the first 19 entries are the D3Q19 set (D3Q19.inc:17-19).
__host__ __device__ constexpr int cxv(int i) {
    constexpr int t[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
    return t[i];
}
__host__ __device__ constexpr int cyv(int i) {
    constexpr int t[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
    return t[i];
}
__host__ __device__ constexpr int czv(int i) {
    constexpr int t[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
    return t[i];
}
// weight class: 0 rest, 1 face, 2 edge, 3 corner
__host__ __device__ constexpr int wclass(int i) { return i == 0 ? 0 : (i < 7 ? 1 : (i < 19 ? 2 : 3)); }

template <int Q> struct VS;
template <> struct VS<19> {
    static constexpr int Q = 19, QF = 5;
    __host__ __device__ static constexpr double wd(int i) {
        return wclass(i) == 0 ? 1.0 / 3 : (wclass(i) == 1 ? 1.0 / 18 : 1.0 / 36);
    }
};
template <> struct VS<27> {
    static constexpr int Q = 27, QF = 9;
    __host__ __device__ static constexpr double wd(int i) {
        return wclass(i) == 0 ? 8.0 / 27 : (wclass(i) == 1 ? 2.0 / 27 : (wclass(i) == 2 ? 1.0 / 54 : 1.0 / 216));
    }
};

// Ghost-face slot tables: face f (0 X_0, 1 X_1, 2 Y_0, 3 Y_1, 4 Z_0, 5 Z_1),
// slot s -> population. The face "a_0" carries the populations leaving the
// tile through its low side (c_a = -1), "a_1" those with c_a = +1.
// Same slot order as src/includeFiles/popSave.inc:2-78 so that ghost
// buffers (and checkpoints holding them) keep the reference layout.
__host__ __device__ constexpr int face_pop(int face, int slot) {
    constexpr int t[6][9] = {
        {2, 8, 10, 14, 16, 20, 22, 24, 25},   // X_0
        {1, 7, 9, 13, 15, 19, 21, 23, 26},    // X_1
        {4, 8, 12, 13, 18, 20, 22, 23, 26},   // Y_0
        {3, 7, 11, 14, 17, 19, 21, 24, 25},   // Y_1
        {6, 10, 12, 15, 17, 20, 21, 24, 26},  // Z_0
        {5, 9, 11, 16, 18, 19, 22, 23, 25},   // Z_1
    };
    return t[face][slot];
}

// ---------------------------------------------------------------------------
// FP32 node state in the scaled variables of the reference kernel.
struct Node {
    float drho;             // rho - RHO_0
    float rho;              // full density
    float inv;              // 1 / rho (set by the moment evaluation)
    float ux, uy, uz;       // 3u
    float mxx, mxy, mxz, myy, myz, mzz;  // 4.5 m_aa / 9 m_ab
};

// Second-order Hermite reconstruction (src/colrec/MR_LBM/reconstruction.inc:1-47):
// f_i = rho w_i [1 - cs2 (Mxx+Myy+Mzz) + c_i.U + sum_a c_ia^2 M_aa + sum_{a<b} c_ia c_ib M_ab],
// returned as g_i = f_i - w_i RHO_0 = w_i drho + w_i rho e_i (e_i = bracket - 1).
// Opposite populations share the even part of the bracket, so each pair
// costs two adds and two FMAs once the rho-scaled moments are formed.
//
// The work is split in two: recon_terms forms the rho-scaled moments once
// per node, recon_pop<Q, I> evaluates one population from them. A kernel
// that needs only a few populations of a node (the ones leaving a tile
// face) evaluates just those from terms staged in LDS, with the same
// operation tree, so the values are bit-identical to reconstruct()'s.
struct Recon {
    float ex, ey, ez;       // R (-(Mxx+Myy+Mzz)/3 + M_aa)
    float ryy, rzz;         // R Myy, R Mzz
    float rxy, rxz, ryz;    // R M_ab
    float rux, ruy, ruz;    // R U_a
    float d1, d2, d3;       // w_k drho for |c|^2 = 1, 2, 3
    static constexpr int N = 14;
    __device__ __forceinline__ float& at(int k) {
        float* f[N] = {&ex, &ey, &ez, &ryy, &rzz, &rxy, &rxz, &ryz, &rux, &ruy, &ruz, &d1, &d2, &d3};
        return *f[k];
    }
};

// Terms of recon_pop<Q, I> that population I reads (bit k = Recon::at(k)).
template <int Q>
__host__ __device__ constexpr uint32_t recon_uses(int i) {
    const int cx = cxv(i), cy = cyv(i), cz = czv(i);
    const int n = (cx != 0) + (cy != 0) + (cz != 0);
    uint32_t m = 0;
    if (n == 1) {
        const int a = cx ? 0 : (cy ? 1 : 2);
        m = (1u << a) | (1u << (8 + a)) | (1u << 11);
    } else if (n == 2) {
        if (cz == 0) m = (1u << 0) | (1u << 3) | (1u << 5) | (1u << 8) | (1u << 9);
        else if (cy == 0) m = (1u << 0) | (1u << 4) | (1u << 6) | (1u << 8) | (1u << 10);
        else m = (1u << 1) | (1u << 4) | (1u << 7) | (1u << 9) | (1u << 10);
        m |= 1u << 12;
    } else if (n == 3) {
        m = (1u << 0) | (1u << 3) | (1u << 4) | (1u << 5) | (1u << 6) | (1u << 7) | (1u << 8) | (1u << 9) |
            (1u << 10) | (1u << 13);
    }
    return m;
}

template <int Q>
__device__ __forceinline__ float recon_re0(const Node& n) {
    return n.rho * (-(1.0f / 3.0f) * ((n.mxx + n.myy) + n.mzz));
}

template <int Q>
__device__ __forceinline__ void recon_terms(const Node& n, float re0, Recon& t) {
    constexpr float w1 = (float)VS<Q>::wd(1), w2 = (float)VS<Q>::wd(7);
    const float R = n.rho;
    t.rux = R * n.ux;
    t.ruy = R * n.uy;
    t.ruz = R * n.uz;
    t.ex = re0 + R * n.mxx;
    t.ey = re0 + (t.ryy = R * n.myy);
    t.ez = re0 + (t.rzz = R * n.mzz);
    t.rxy = R * n.mxy;
    t.rxz = R * n.mxz;
    t.ryz = R * n.myz;
    t.d1 = w1 * n.drho;
    t.d2 = w2 * n.drho;
    if constexpr (Q == 27) t.d3 = (float)VS<Q>::wd(26) * n.drho;
    else t.d3 = 0.0f;
}

// Population I (1 <= I < Q) from the terms.
template <int Q, int I>
__device__ __forceinline__ float recon_pop(const Recon& t) {
    constexpr int cx = cxv(I), cy = cyv(I), cz = czv(I);
    constexpr int nz = (cx != 0) + (cy != 0) + (cz != 0);
    constexpr float w1 = (float)VS<Q>::wd(1), w2 = (float)VS<Q>::wd(7);
    if constexpr (nz == 1) {
        const float e = cx ? t.ex : (cy ? t.ey : t.ez);
        const float u = cx ? t.rux : (cy ? t.ruy : t.ruz);
        return __builtin_fmaf(w1, (cx + cy + cz) > 0 ? e + u : e - u, t.d1);
    } else if constexpr (nz == 2) {
        // plane ab (a < b): e_ab = (re0 + R M_aa) + R M_bb; populations
        // (+a,+b), (-a,-b) use e_ab + R M_ab and R U_a + R U_b; (+a,-b),
        // (-a,+b) use e_ab - R M_ab and R U_a - R U_b
        constexpr int ca = cx ? cx : cy;
        constexpr int cb = cz ? cz : cy;
        float e, rab, ua, ub;
        if constexpr (cz == 0) e = t.ex + t.ryy, rab = t.rxy, ua = t.rux, ub = t.ruy;
        else if constexpr (cy == 0) e = t.ex + t.rzz, rab = t.rxz, ua = t.rux, ub = t.ruz;
        else e = t.ey + t.rzz, rab = t.ryz, ua = t.ruy, ub = t.ruz;
        if constexpr (ca == cb) {
            const float p = e + rab, s = ua + ub;
            return __builtin_fmaf(w2, ca > 0 ? p + s : p - s, t.d2);
        } else {
            const float m = e - rab, d = ua - ub;
            return __builtin_fmaf(w2, ca > 0 ? m + d : m - d, t.d2);
        }
    } else {
        static_assert(Q == 27 && nz == 3, "corner populations exist in D3Q27 only");
        constexpr float w3 = (float)VS<Q>::wd(26);
        // x sign fixes the +/- of the final combination; the xy pair sign
        // picks a/b, the (xz, yz) pair signs pick e1..e4 and o1..o4
        const float exyz = (t.ex + t.ryy) + t.rzz;
        const bool sxy = cx == cy, sxz = cx == cz;
        const float ab = sxy ? exyz + t.rxy : exyz - t.rxy;
        const float st = sxy ? t.rux + t.ruy : t.rux - t.ruy;
        float e, o;
        if (sxy) {
            e = sxz ? ab + (t.rxz + t.ryz) : ab - (t.rxz + t.ryz);
        } else {
            e = sxz ? ab + (t.rxz - t.ryz) : ab - (t.rxz - t.ryz);
        }
        o = sxz ? st + t.ruz : st - t.ruz;
        return __builtin_fmaf(w3, cx > 0 ? e + o : e - o, t.d3);
    }
}

// All Q populations from the terms (re0 and drho as recon_terms used them).
template <int Q>
__device__ __forceinline__ void reconstruct_t(const Recon& t, float re0, float drho, float* g) {
    constexpr float w0 = (float)VS<Q>::wd(0);
    g[0] = __builtin_fmaf(w0, re0, w0 * drho);
    static_for<Q - 1>([&](auto I) { g[decltype(I)::value + 1] = recon_pop<Q, decltype(I)::value + 1>(t); });
}

template <int Q>
__device__ __forceinline__ void reconstruct(const Node& n, float* g) {
    const float re0 = recon_re0<Q>(n);
    Recon t;
    recon_terms<Q>(n, re0, t);
    reconstruct_t<Q>(t, re0, n.drho, g);
}

// Sums of streamed deviations: r = sum g, j = sum g c, s_ab = sum g c_a c_b,
// formed from the pair sums p = g_i + g_i' and differences d = g_i - g_i'.
struct Sums {
    float r, jx, jy, jz, sxx, syy, szz, sxy, sxz, syz;
};

template <int Q>
__device__ __forceinline__ void moment_sums(const float* g, Sums& S) {
    const float p12 = g[1] + g[2], d12 = g[1] - g[2];
    const float p34 = g[3] + g[4], d34 = g[3] - g[4];
    const float p56 = g[5] + g[6], d56 = g[5] - g[6];
    const float p78 = g[7] + g[8], d78 = g[7] - g[8];
    const float p910 = g[9] + g[10], d910 = g[9] - g[10];
    const float p1112 = g[11] + g[12], d1112 = g[11] - g[12];
    const float p1314 = g[13] + g[14], d1314 = g[13] - g[14];
    const float p1516 = g[15] + g[16], d1516 = g[15] - g[16];
    const float p1718 = g[17] + g[18], d1718 = g[17] - g[18];
    float sxx = ((p12 + p78) + (p910 + p1314)) + p1516;
    float syy = ((p34 + p78) + (p1112 + p1314)) + p1718;
    float szz = ((p56 + p910) + (p1112 + p1516)) + p1718;
    float sxy = p78 - p1314, sxz = p910 - p1516, syz = p1112 - p1718;
    float jx = ((d12 + d78) + (d910 + d1314)) + d1516;
    float jy = ((d34 + d78) + (d1112 - d1314)) + d1718;
    float jz = ((d56 + d910) + (d1112 - d1516)) - d1718;
    float r = (((g[0] + p12) + (p34 + p56)) + ((p78 + p910) + (p1112 + p1314))) + (p1516 + p1718);
    if constexpr (Q == 27) {
        const float pa = g[19] + g[20], da = g[19] - g[20];  // (+,+,+)
        const float pb = g[21] + g[22], db = g[21] - g[22];  // (+,+,-)
        const float pc = g[23] + g[24], dc = g[23] - g[24];  // (+,-,+)
        const float pd = g[26] + g[25], dd = g[26] - g[25];  // (+,-,-)
        const float pall = (pa + pb) + (pc + pd);
        r = r + pall;
        sxx = sxx + pall;
        syy = syy + pall;
        szz = szz + pall;
        sxy = sxy + ((pa + pb) - (pc + pd));
        sxz = sxz + ((pa - pb) + (pc - pd));
        syz = syz + ((pa - pb) - (pc - pd));
        jx = jx + ((da + db) + (dc + dd));
        jy = jy + ((da + db) - (dc + dd));
        jz = jz + ((da - db) + (dc - dd));
    }
    S.r = r;
    S.jx = jx, S.jy = jy, S.jz = jz;
    S.sxx = sxx, S.syy = syy, S.szz = szz;
    S.sxy = sxy, S.sxz = sxz, S.syz = syz;
}

// Moments of the streamed populations at a bulk node
// (src/mlbm.cu:464-509): rho, u = sum f c / rho, m_ab = sum f c_a c_b / rho - cs2 d_ab,
// then scaled by 3, 4.5, 9. From deviations: drho = sum g, j = sum g c,
// m_aa = (sum g c_a^2 - drho/3) / rho, m_ab = sum g c_a c_b / rho.
// su, sm: power-of-two scales of the outputs U and M (the storage scales
// of a 16-bit format, whose rounding then needs no further multiply); the
// density deviation is not scaled.
__device__ __forceinline__ void bulk_from_sums(const Sums& S, float rho0, Node& n, float su = 1.0f, float sm = 1.0f) {
    const float r = S.r;
    n.drho = r;
    n.rho = rho0 + r;
    const float inv = __builtin_amdgcn_rcpf(n.rho);
    n.inv = inv;
    const float u3 = (3.0f * su) * inv, m45 = (4.5f * sm) * inv, m9 = (9.0f * sm) * inv;
    const float r3 = r * (1.0f / 3.0f);
    n.ux = S.jx * u3;
    n.uy = S.jy * u3;
    n.uz = S.jz * u3;
    n.mxx = (S.sxx - r3) * m45;
    n.myy = (S.syy - r3) * m45;
    n.mzz = (S.szz - r3) * m45;
    n.mxy = S.sxy * m9;
    n.mxz = S.sxz * m9;
    n.myz = S.syz * m9;
}

template <int Q>
__device__ __forceinline__ void bulk_moments(const float* g, float rho0, Node& n, float su = 1.0f, float sm = 1.0f) {
    Sums S;
    moment_sums<Q>(g, S);
    bulk_from_sums(S, rho0, n, su, sm);
}

// Boundary-condition class record (device table, BC_STRIDE floats per class).
//   [0]       bit mask of known populations (bit i set = population i known)
//   [1..3]    wall velocity (unscaled)
//   [4..52]   A^-1, inverse of the 7x7 moment-matching system, row-major
//   [53]      W_K = sum_{i in K} w_i
//   [54..59]  c_k = sum_{i in K} w_i H_i,k / W_K   (k = xx, xy, xz, yy, yz, zz)
//   [60]      d0  = W_K y_rest,0 - 1
//   [61..67]  y_rest = A^-1 (1, c)  (the solution for the rest state)
// The system (one per class, built on the host in FP64) is the one whose
// closed-form solutions are the reference's generated bc_definition.inc
// formulas: with P = rho m,
//   sum_{i in K} f^R_i(rho, u_w, (1-omega) m + omega u_w u_w) = rho_I
//   sum_{i in K} f^R_i(rho, u_w, m) H_i,ab                     = rho_I m_I,ab
// where K is the known set, rho_I = sum_K f_i and m_I = sum_K f_i H_i / rho_I;
// x = (rho, P) = rho_I A^-1 (1, m_I).
constexpr int BC_STRIDE = 72;

// Moment-based boundary condition at a node of class `c`, from the sums of
// its known streamed deviations (f_i = w_i RHO_0 + g_i; moment_sums over g
// with the unknown populations zeroed, see known_only). The solve is split
// around the rest state: (1, m_I) = (1, c) + (0, db) with
// db = (T - c r) / rho_I, where r = sum_K g and T = sum_K g H, so every term
// the FP32 arithmetic forms is a small deviation. Outputs unscaled rho, u, m
// exactly as a bc_definition.inc fragment does (src/mlbm.cu:456-461); the
// caller scales.
template <typename P>
__device__ __forceinline__ void bc_from_sums(P c, const Sums& S, float rho0, Node& n) {
    const float r = S.r, sxx = S.sxx, syy = S.syy, szz = S.szz, sxy = S.sxy, sxz = S.sxz, syz = S.syz;
    const float rI = rho0 * c[53] + r;
    const float inv = __builtin_amdgcn_rcpf(rI);  // v_rcp_f32 (1 ulp): the BC moments' rounding level
    const float r3 = r * (1.0f / 3.0f);
    float db[6];
    db[0] = ((sxx - r3) - c[54] * r) * inv;
    db[1] = (sxy - c[55] * r) * inv;
    db[2] = (sxz - c[56] * r) * inv;
    db[3] = ((syy - r3) - c[57] * r) * inv;
    db[4] = (syz - c[58] * r) * inv;
    db[5] = ((szz - r3) - c[59] * r) * inv;
    float dy[7];
#pragma unroll
    for (int k = 0; k < 7; k++) {
        float acc = 0.0f;
#pragma unroll
        for (int j = 0; j < 6; j++) acc = acc + c[4 + 7 * k + 1 + j] * db[j];
        dy[k] = acc;
    }
    const float y0 = c[61] + dy[0];
    const float iy0 = __builtin_amdgcn_rcpf(y0);
    // rho - RHO_0 = rho_I y0 - RHO_0 = RHO_0 d0 + r y_rest,0 + rho_I dy0
    n.drho = (rho0 * c[60] + r * c[61]) + rI * dy[0];
    n.rho = rho0 + n.drho;
    n.inv = __builtin_amdgcn_rcpf(n.rho);
    n.ux = c[1];
    n.uy = c[2];
    n.uz = c[3];
    n.mxx = (c[62] + dy[1]) * iy0;
    n.mxy = (c[63] + dy[2]) * iy0;
    n.mxz = (c[64] + dy[3]) * iy0;
    n.myy = (c[65] + dy[4]) * iy0;
    n.myz = (c[66] + dy[5]) * iy0;
    n.mzz = (c[67] + dy[6]) * iy0;
}

// Zero the populations a boundary node does not know (bit i of `known`
// clear): one signed bit-field extract (0 or all ones) and one AND each.
template <int Q>
__device__ __forceinline__ void known_only(uint32_t known, float* g) {
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const int keep = __builtin_amdgcn_sbfe((int)known, i, 1);
        g[i] = __int_as_float(__float_as_int(g[i]) & keep);
    }
}

template <int Q, typename P>
__device__ __forceinline__ void bc_solve(P c, float* g, float rho0, Node& n) {
    known_only<Q>(__float_as_uint(c[0]), g);
    Sums S;
    moment_sums<Q>(g, S);
    bc_from_sums(c, S, rho0, n);
}

// Scale unscaled u, m to the stored convention (src/mlbm.cu:500-509),
// times the power-of-two scales su (U) and sm (M) of bulk_moments.
__device__ __forceinline__ void scale_moments(Node& n, float su = 1.0f, float sm = 1.0f) {
    n.ux = (3.0f * su) * n.ux;
    n.uy = (3.0f * su) * n.uy;
    n.uz = (3.0f * su) * n.uz;
    n.mxx = (4.5f * sm) * n.mxx;
    n.mxy = (9.0f * sm) * n.mxy;
    n.mxz = (9.0f * sm) * n.mxz;
    n.myy = (4.5f * sm) * n.myy;
    n.myz = (9.0f * sm) * n.myz;
    n.mzz = (4.5f * sm) * n.mzz;
}

// Relaxation constants of collide(). With U and M carried times
// power-of-two scales su and sm (the 16-bit storages' scales, 1 for FP32
// and BF16), the host folds the scales in: the force fx, fy, fz carries su
// (it enters U), and omega, omega_d2, tt_omega, tt_omega_t3 carry sm / su^2
// (each multiplies a product of two U-scaled values into a second moment).
// Every product is then the unscaled one times a power of two, exactly.
struct Relax {
    float omega, t_omega, tt_omega, omega_d2, tt_omega_t3;
    float fx, fy, fz;
    int has_force;
};

// Second-order moment collision with force terms
// (src/colrec/MR_LBM/collision.inc:2-14), in scaled variables:
//   U += 3 F / rho
//   M_aa = (1-w) M_aa + (w/2) U_a^2 + (3/(2 rho)) (1-w/2) 2 F_a U_a
//   M_ab = (1-w) M_ab +  w U_a U_b  + (3/rho) (1-w/2) (F_a U_b + F_b U_a)
// The force terms are skipped (uniform branch) when F = 0.
__device__ __forceinline__ void collide(Node& n, const Relax& R) {
    if (R.has_force) {
        const float i3 = 3.0f * n.inv;
        n.ux = __builtin_fmaf(i3, R.fx, n.ux);
        n.uy = __builtin_fmaf(i3, R.fy, n.uy);
        n.uz = __builtin_fmaf(i3, R.fz, n.uz);
    }
    const float hx = R.omega_d2 * n.ux, hy = R.omega_d2 * n.uy, hz = R.omega_d2 * n.uz;
    const float ox = R.omega * n.ux, oy = R.omega * n.uy;
    float mxx = __builtin_fmaf(R.t_omega, n.mxx, hx * n.ux);
    float myy = __builtin_fmaf(R.t_omega, n.myy, hy * n.uy);
    float mzz = __builtin_fmaf(R.t_omega, n.mzz, hz * n.uz);
    float mxy = __builtin_fmaf(R.t_omega, n.mxy, ox * n.uy);
    float mxz = __builtin_fmaf(R.t_omega, n.mxz, ox * n.uz);
    float myz = __builtin_fmaf(R.t_omega, n.myz, oy * n.uz);
    if (R.has_force) {
        const float i15 = (1.5f * n.inv) * R.tt_omega;
        const float i3 = R.tt_omega_t3 * n.inv;
        mxx = __builtin_fmaf(i15, 2.0f * (R.fx * n.ux), mxx);
        myy = __builtin_fmaf(i15, 2.0f * (R.fy * n.uy), myy);
        mzz = __builtin_fmaf(i15, 2.0f * (R.fz * n.uz), mzz);
        mxy = __builtin_fmaf(i3, R.fx * n.uy + R.fy * n.ux, mxy);
        mxz = __builtin_fmaf(i3, R.fx * n.uz + R.fz * n.ux, mxz);
        myz = __builtin_fmaf(i3, R.fy * n.uz + R.fz * n.uy, myz);
    }
    n.mxx = mxx;
    n.myy = myy;
    n.mzz = mzz;
    n.mxy = mxy;
    n.mxz = mxz;
    n.myz = myz;
}

}  // namespace mrlbm
