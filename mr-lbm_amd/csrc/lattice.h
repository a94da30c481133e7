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

namespace mrlbm {

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
    float ux, uy, uz;       // 3u
    float mxx, mxy, mxz, myy, myz, mzz;  // 4.5 m_aa / 9 m_ab
};

template <int Q, int I>
__device__ __forceinline__ float recon_one(float e0, const Node& n, float w) {
    constexpr int cx = cxv(I), cy = cyv(I), cz = czv(I);
    float a = e0;
    if constexpr (cx > 0) a = a + n.ux;
    if constexpr (cx < 0) a = a - n.ux;
    if constexpr (cy > 0) a = a + n.uy;
    if constexpr (cy < 0) a = a - n.uy;
    if constexpr (cz > 0) a = a + n.uz;
    if constexpr (cz < 0) a = a - n.uz;
    if constexpr (cx != 0) a = a + n.mxx;
    if constexpr (cy != 0) a = a + n.myy;
    if constexpr (cz != 0) a = a + n.mzz;
    if constexpr (cx * cy > 0) a = a + n.mxy;
    if constexpr (cx * cy < 0) a = a - n.mxy;
    if constexpr (cx * cz > 0) a = a + n.mxz;
    if constexpr (cx * cz < 0) a = a - n.mxz;
    if constexpr (cy * cz > 0) a = a + n.myz;
    if constexpr (cy * cz < 0) a = a - n.myz;
    // f_i - w_i RHO_0 = w_i (drho + rho e_i), e_i = bracket - 1
    return w * (n.drho + n.rho * a);
}

// Second-order Hermite reconstruction (src/colrec/MR_LBM/reconstruction.inc:1-47):
// f_i = rho w_i [1 - cs2 (Mxx+Myy+Mzz) + c_i.U + sum_a c_ia^2 M_aa + sum_{a<b} c_ia c_ib M_ab],
// returned as g_i = f_i - w_i RHO_0.
template <int Q, size_t... Is>
__device__ __forceinline__ void reconstruct_impl(const Node& n, float* g, std::index_sequence<Is...>) {
    const float e0 = -(1.0f / 3.0f) * ((n.mxx + n.myy) + n.mzz);
    const float w[4] = {(float)VS<Q>::wd(0), (float)VS<Q>::wd(1), (float)VS<Q>::wd(7),
                        Q == 27 ? (float)VS<Q>::wd(Q - 1) : 0.0f};
    ((g[Is] = recon_one<Q, (int)Is>(e0, n, w[wclass((int)Is)])), ...);
}
template <int Q>
__device__ __forceinline__ void reconstruct(const Node& n, float* g) {
    reconstruct_impl<Q>(n, g, std::make_index_sequence<Q>{});
}

// Moments of the streamed populations at a bulk node
// (src/mlbm.cu:464-509): rho, u = sum f c / rho, m_ab = sum f c_a c_b / rho - cs2 d_ab,
// then scaled by 3, 4.5, 9. From deviations: drho = sum g, j = sum g c,
// m_aa = (sum g c_a^2 - drho/3) / rho, m_ab = sum g c_a c_b / rho.
template <int Q>
__device__ __forceinline__ void bulk_moments(const float* g, float rho0, Node& n) {
    float r = 0.f, jx = 0.f, jy = 0.f, jz = 0.f, sxx = 0.f, syy = 0.f, szz = 0.f, sxy = 0.f, sxz = 0.f, syz = 0.f;
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const float f = g[i];
        r = r + f;
        if (cxv(i) > 0) jx = jx + f;
        if (cxv(i) < 0) jx = jx - f;
        if (cyv(i) > 0) jy = jy + f;
        if (cyv(i) < 0) jy = jy - f;
        if (czv(i) > 0) jz = jz + f;
        if (czv(i) < 0) jz = jz - f;
        if (cxv(i) != 0) sxx = sxx + f;
        if (cyv(i) != 0) syy = syy + f;
        if (czv(i) != 0) szz = szz + f;
        if (cxv(i) * cyv(i) > 0) sxy = sxy + f;
        if (cxv(i) * cyv(i) < 0) sxy = sxy - f;
        if (cxv(i) * czv(i) > 0) sxz = sxz + f;
        if (cxv(i) * czv(i) < 0) sxz = sxz - f;
        if (cyv(i) * czv(i) > 0) syz = syz + f;
        if (cyv(i) * czv(i) < 0) syz = syz - f;
    }
    n.drho = r;
    n.rho = rho0 + r;
    const float inv = 1.0f / n.rho;
    const float r3 = r * (1.0f / 3.0f);
    n.ux = 3.0f * (jx * inv);
    n.uy = 3.0f * (jy * inv);
    n.uz = 3.0f * (jz * inv);
    n.mxx = 4.5f * ((sxx - r3) * inv);
    n.myy = 4.5f * ((syy - r3) * inv);
    n.mzz = 4.5f * ((szz - r3) * inv);
    n.mxy = 9.0f * (sxy * inv);
    n.mxz = 9.0f * (sxz * inv);
    n.myz = 9.0f * (syz * inv);
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

// Moment-based boundary condition at a node of class `c` from the streamed
// deviations g (f_i = w_i RHO_0 + g_i). The solve is split around the rest
// state: (1, m_I) = (1, c) + (0, db) with db = (T - c r) / rho_I, where
// r = sum_K g and T = sum_K g H, so every term the FP32 arithmetic forms is
// a small deviation. Outputs unscaled rho, u, m exactly as a
// bc_definition.inc fragment does (src/mlbm.cu:456-461); the caller scales.
template <int Q>
__device__ __forceinline__ void bc_solve(const float* __restrict__ c, const float* g, float rho0, Node& n) {
    const uint32_t known = __float_as_uint(c[0]);
    float r = 0.f, sxx = 0.f, syy = 0.f, szz = 0.f, sxy = 0.f, sxz = 0.f, syz = 0.f;
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const float f = ((known >> i) & 1u) ? g[i] : 0.0f;
        r = r + f;
        if (cxv(i) != 0) sxx = sxx + f;
        if (cyv(i) != 0) syy = syy + f;
        if (czv(i) != 0) szz = szz + f;
        if (cxv(i) * cyv(i) > 0) sxy = sxy + f;
        if (cxv(i) * cyv(i) < 0) sxy = sxy - f;
        if (cxv(i) * czv(i) > 0) sxz = sxz + f;
        if (cxv(i) * czv(i) < 0) sxz = sxz - f;
        if (cyv(i) * czv(i) > 0) syz = syz + f;
        if (cyv(i) * czv(i) < 0) syz = syz - f;
    }
    const float rI = rho0 * c[53] + r;
    const float inv = 1.0f / rI;
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
    const float iy0 = 1.0f / y0;
    // rho - RHO_0 = rho_I y0 - RHO_0 = RHO_0 d0 + r y_rest,0 + rho_I dy0
    n.drho = (rho0 * c[60] + r * c[61]) + rI * dy[0];
    n.rho = rho0 + n.drho;
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

// Scale unscaled u, m to the stored convention (src/mlbm.cu:500-509).
__device__ __forceinline__ void scale_moments(Node& n) {
    n.ux = 3.0f * n.ux;
    n.uy = 3.0f * n.uy;
    n.uz = 3.0f * n.uz;
    n.mxx = 4.5f * n.mxx;
    n.mxy = 9.0f * n.mxy;
    n.mxz = 9.0f * n.mxz;
    n.myy = 4.5f * n.myy;
    n.myz = 9.0f * n.myz;
    n.mzz = 4.5f * n.mzz;
}

struct Relax {
    float omega, t_omega, tt_omega, omega_d2, tt_omega_t3;
    float fx, fy, fz;
};

// Second-order moment collision with force terms
// (src/colrec/MR_LBM/collision.inc:2-14), in scaled variables.
__device__ __forceinline__ void collide(Node& n, const Relax& R) {
    const float invRho = 1.0f / n.rho;
    n.ux = n.ux + (3.0f * invRho) * R.fx;
    n.uy = n.uy + (3.0f * invRho) * R.fy;
    n.uz = n.uz + (3.0f * invRho) * R.fz;
    const float i15 = (3.0f * invRho) / 2.0f;
    const float i3 = R.tt_omega_t3 * invRho;
    n.mxx = (R.t_omega * n.mxx + R.omega_d2 * n.ux * n.ux) + i15 * R.tt_omega * (R.fx * n.ux + R.fx * n.ux);
    n.myy = (R.t_omega * n.myy + R.omega_d2 * n.uy * n.uy) + i15 * R.tt_omega * (R.fy * n.uy + R.fy * n.uy);
    n.mzz = (R.t_omega * n.mzz + R.omega_d2 * n.uz * n.uz) + i15 * R.tt_omega * (R.fz * n.uz + R.fz * n.uz);
    n.mxy = (R.t_omega * n.mxy + R.omega * n.ux * n.uy) + i3 * (R.fx * n.uy + R.fy * n.ux);
    n.mxz = (R.t_omega * n.mxz + R.omega * n.ux * n.uz) + i3 * (R.fx * n.uz + R.fz * n.ux);
    n.myz = (R.t_omega * n.myz + R.omega * n.uy * n.uz) + i3 * (R.fy * n.uz + R.fz * n.uy);
}

}  // namespace mrlbm
