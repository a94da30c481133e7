// Tile step kernel: one node per thread, one (BX,BY,BZ) tile per workgroup.
//
// Same data flow as the reference's gpuMomCollisionStream
// (src/mlbm.cu:3-714), written for gfx950:
//   1. load the 10 stored moments (block-major SoA: each wave64 reads
//      contiguous runs per moment plane; non-temporal)
//   2. reconstruct the post-collision populations of the previous step
//      (as deviations g_i = f_i - w_i RHO_0, see lattice.h)
//   3. write populations 1..Q-1 to per-population LDS arrays with the
//      entering plane on the upwind side (x-only movers: DPP lane shifts);
//      whole waves copy the neighbours' ghost planes into those planes
//      (halo.h; the X, Y, Z winner rule of src/includeFiles/popLoad.inc:2-86);
//      barrier; every node pulls its populations at compile-time offsets
//      from its base address
//   4. boundary class -> moment-based BC, else bulk moments; scale; collide;
//      round to the storage format
//   5. store moments in place; stage each node's reconstruction terms in
//      LDS and let whole waves evaluate and write the outgoing ghost
//      populations in contiguous runs (src/includeFiles/popSave.inc:2-78
//      with the wall predicates of src/includeFiles/interface.h:6-47)
// Solid nodes run the same code (no early exit before a barrier) and store
// nothing.
#pragma once

#include "halo.h"
#include "step_tile_launch.h"
#include "tile_common.h"

namespace mrlbm {

// Neighbour-tile indices of a workgroup (uniform): wrapped coordinates
// along each axis for offsets -1, 0, +1.
struct Nbr {
    int x[3], y[3], z[3];
    int ntx, nty, ntiles;
    __device__ __forceinline__ int tile(int dx, int dy, int dz) const {
        return x[dx + 1] + ntx * (y[dy + 1] + nty * z[dz + 1]);
    }
};

// Ghost value of incoming chunk K for this lane. Every lane loads (lanes
// past the window re-read it), so a wave's loads of all its chunks issue
// back to back with no exec-mask branch between them and one wait covers
// them all. The address is the owner's plane base (uniform) + lane + a
// constant: no per-lane index arithmetic.
template <int Q, int BX, int BY, int BZ, int K>
__device__ __forceinline__ float halo_fetch(const float* gin, int lane, const Nbr& nb) {
    constexpr Chunk c = ChunkTables<Q, BX, BY, BZ>::in[K];
    constexpr int QF = VS<Q>::QF;
    constexpr int Bu = face_bu<BX, BY, BZ>(c.face);
    constexpr int Bv = face_bv<BX, BY, BZ>(c.face);
    constexpr int NC = Bu * c.rows;
    static_assert(64 % Bu == 0 && (NC & (NC - 1)) == 0, "face rows must tile a wave");
    const int otile = nb.tile(c.dx, c.dy, c.dz);
    const auto src = gptr(ghost_face<BX, BY, BZ, QF>(gin, c.face, nb.ntiles)) +
                     (size_t)(Bu * Bv) * (c.slot + QF * (size_t)otile) + Bu * c.w0;
    return src[NC >= 64 ? (uint32_t)lane : ((uint32_t)lane & (NC - 1))];
}

// Face cell (u, row r of the window) of this lane in chunk K of the
// incoming (IN) or outgoing table; false if the lane moves no cell.
template <int Q, int BX, int BY, int BZ, bool IN, int K>
__device__ __forceinline__ bool chunk_cell(int lane, int& u, int& r) {
    using CT = ChunkTables<Q, BX, BY, BZ>;
    constexpr Chunk c = IN ? CT::in[K] : CT::out[K];
    constexpr int Bu = face_bu<BX, BY, BZ>(c.face);
    u = lane % Bu;
    r = lane / Bu;
    const int v = c.w0 + r;
    return (Bu * c.rows >= 64 || lane < Bu * c.rows) && u >= c.ulo && u <= c.uhi && v >= c.vlo && v <= c.vhi;
}

template <int Q, int BX, int BY, int BZ, int K>
__device__ __forceinline__ void halo_put(int lane, float v, float* s) {
    constexpr Chunk c = ChunkTables<Q, BX, BY, BZ>::in[K];
    int u, r;
    if (chunk_cell<Q, BX, BY, BZ, true, K>(lane, u, r)) s[c.lds0 + u * c.su + r * c.sv] = v;
}

// Outgoing chunk K: the lane's face node evaluates the chunk's population
// from the reconstruction terms the node staged in LDS (term k of the node
// at base nb is s[k * E + nb]) and writes it to the ghost plane.
template <int Q, int BX, int BY, int BZ, bool HAS_BC, int K>
__device__ __forceinline__ void halo_store(const StepArgs& a, float* gout, int lane, int tile, int bx, int by,
                                           int bz, const float* s, const uint8_t* s_cls) {
    constexpr Chunk c = ChunkTables<Q, BX, BY, BZ>::out[K];
    constexpr int QF = VS<Q>::QF;
    constexpr int Bu = face_bu<BX, BY, BZ>(c.face);
    constexpr int Bv = face_bv<BX, BY, BZ>(c.face);
    using G = HaloGeom<BX, BY, BZ>;
    // wall predicate of src/includeFiles/interface.h:6-47: a tile face on a
    // wall-axis domain boundary is not written
    bool skip = false;
    if constexpr (c.face == 0) skip = a.wall_x && bx == 0;
    if constexpr (c.face == 1) skip = a.wall_x && bx == a.ntx - 1;
    if constexpr (c.face == 2) skip = a.wall_y && by == 0;
    if constexpr (c.face == 3) skip = a.wall_y && by == a.nty - 1;
    if constexpr (c.face == 4) skip = a.wall_z && bz == 0;
    if constexpr (c.face == 5) skip = a.wall_z && bz == a.ntz - 1;
    if (skip) return;
    const auto dst = gptr(ghost_face<BX, BY, BZ, QF>(gout, c.face, a.ntiles)) +
                     (size_t)(Bu * Bv) * (c.slot + QF * (size_t)tile) + Bu * c.w0;
    int u, r;
    if (chunk_cell<Q, BX, BY, BZ, false, K>(lane, u, r)) {
        if constexpr (HAS_BC) {
            // node of the face cell; solid nodes write no ghost values
            constexpr int ax = c.face / 2;
            constexpr int fix = (c.face & 1) ? (ax == 0 ? BX : (ax == 1 ? BY : BZ)) - 1 : 0;
            const int v = c.w0 + r;
            const int nx = ax == 0 ? fix : u;
            const int ny = ax == 1 ? fix : (ax == 0 ? u : v);
            const int nz = ax == 2 ? fix : v;
            if (s_cls[nx + BX * (ny + BY * nz)] == 255u) return;
        }
        const int nb = c.lds0 + u * c.su + r * c.sv;  // staged-term base of the face node
        constexpr uint32_t uses = recon_uses<Q>(c.pop);
        Recon t;
        static_for<Recon::N>([&](auto KK) {
            constexpr int k = decltype(KK)::value;
            t.at(k) = ((uses >> k) & 1u) ? s[k * G::E + nb] : 0.0f;
        });
        dst[(uint32_t)lane] = recon_pop<Q, c.pop>(t);
    }
}

// Terms staged per node for the outgoing ghost values (d3 is D3Q27-only).
template <int Q>
constexpr int staged_terms() { return Q == 27 ? Recon::N : Recon::N - 1; }

// LDS of one tile update: the population arrays (PopLds), reused for the
// staged terms (term k of node base b at k * E + b), and the node classes
// the outgoing-ghost writes read (boundary body only). The classes are
// written with the staged terms, after the populations are consumed, so
// where the population arrays are larger than the staged terms by a byte
// per node they live in that slack (D3Q19: 8x8x8 keeps 40,836 bytes, four
// workgroups per CU, in both bodies); otherwise in an array of their own.
template <int Q, int BX, int BY, int BZ, bool HAS_BC>
struct TileLds {
    static constexpr int T = BX * BY * BZ;
    static constexpr int NP = PopLds<Q, BX, BY, BZ>::total();
    static constexpr int NS = staged_terms<Q>() * HaloGeom<BX, BY, BZ>::E;
    static constexpr bool CLS_SLACK = HAS_BC && (NP - NS) * 4 >= T;
    float s[NP > NS ? NP : NS];
    uint8_t cls_own[(HAS_BC && !CLS_SLACK) ? T : 4];
    __device__ __forceinline__ uint8_t* cls() { return CLS_SLACK ? (uint8_t*)(s + NS) : cls_own; }
};

// HBM reads of one tile update (class bytes, stored moments, this wave's
// ghost chunks), issued ahead of the tile's computation: a workgroup that
// walks several tiles loads tile k+1 while it computes tile k.
template <int Q, int BX, int BY, int BZ, int S>
struct TileIn {
    static constexpr int NW = BX * BY * BZ / 64;
    static constexpr int NIN = (int)ChunkTables<Q, BX, BY, BZ>::in.size();
    static constexpr int RMAX = (NIN + NW - 1) / NW;
    typename Fetched<S>::T m[10];
    float h[RMAX];
    uint32_t cls;
};

// threadIdx.x as a value the compiler cannot see through: per-lane index
// arithmetic derived from it is recomputed in each tile of a tile loop
// instead of being hoisted out of the loop and held in registers (which
// would cost more occupancy than the arithmetic costs issue slots).
template <int T>
__device__ __forceinline__ int thread_index() {
    uint32_t t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return (int)(t & (uint32_t)(T - 1));  // T is a power of two: 0 <= t < T stays known
}

// Tile coordinates of a tile index (x fastest): scalar multiply-high
// divisions (FastDiv), the tile index being uniform.
__device__ __forceinline__ void tile_xyz(const StepArgs& a, int tile, int& bx, int& by, int& bz) {
    const uint32_t t = (uint32_t)tile;
    const uint32_t q = a.by_ntx.div(t);
    const uint32_t z = a.by_nty.div(q);
    bx = (int)(t - q * (uint32_t)a.ntx);
    by = (int)(q - z * (uint32_t)a.nty);
    bz = (int)z;
}

// Tile coordinates of a tile index and its periodic neighbours, as the
// reference's bxm1/bxp1 (src/mlbm.cu:115-122).
__device__ __forceinline__ void tile_coords(const StepArgs& a, int tile, int& bx, int& by, int& bz, Nbr& nb) {
    tile_xyz(a, tile, bx, by, bz);
    nb.ntx = a.ntx;
    nb.nty = a.nty;
    nb.ntiles = a.ntiles;
    nb.x[0] = bx == 0 ? a.ntx - 1 : bx - 1, nb.x[1] = bx, nb.x[2] = bx == a.ntx - 1 ? 0 : bx + 1;
    nb.y[0] = by == 0 ? a.nty - 1 : by - 1, nb.y[1] = by, nb.y[2] = by == a.nty - 1 ? 0 : by + 1;
    nb.z[0] = bz == 0 ? a.ntz - 1 : bz - 1, nb.z[1] = bz, nb.z[2] = bz == a.ntz - 1 ? 0 : bz + 1;
}

template <int Q, int BX, int BY, int BZ, bool HAS_BC, int S>
__device__ __forceinline__ void tile_fetch(const StepArgs& a, const float* gin, int tile,
                                           TileIn<Q, BX, BY, BZ, S>& in) {
    constexpr int T = BX * BY * BZ;
    constexpr int NW = T / 64;
    constexpr int NIN = TileIn<Q, BX, BY, BZ, S>::NIN;
    const int t = thread_index<T>();
    const int lane = t % 64;
    const int w = __builtin_amdgcn_readfirstlane(t / 64);
    int bx, by, bz;
    Nbr nb;
    tile_coords(a, tile, bx, by, bz, nb);
    in.cls = 0;
    if constexpr (HAS_BC) {
        // uniform test: batched lattices may have no class map
        if (a.cls) in.cls = (gptr(a.cls) + (size_t)tile * T)[(uint32_t)t];
    }
    // this tile's moment block: a uniform base (scalar registers), the
    // lane's 32-bit byte offset and an immediate. Planes lie PB bytes apart;
    // those past the 12-bit immediate take a window offset (lane offset +
    // w * 4 KiB, one 32-bit add, opaque so that the compiler does not fold
    // it back into 64-bit vector address adds per plane)
    using MT = typename Codec<S>::T;
    const auto tb = (const MRLBM_GLOBAL char*)(gptr((const MT*)a.mom) + (size_t)tile * 10 * T);
    constexpr uint32_t PB = (uint32_t)(T * sizeof(MT));
    constexpr int PW = PB >= 4096 ? 1 : 4096 / (int)PB;  // planes per window
    constexpr int NWIN = (10 + PW - 1) / PW;
    uint32_t woff[NWIN];
    woff[0] = (uint32_t)t * (uint32_t)sizeof(MT);
    static_for<NWIN - 1>([&](auto WW) {
        constexpr int w = decltype(WW)::value + 1;
        woff[w] = woff[0] + (uint32_t)(w * PW) * PB;
        asm("" : "+v"(woff[w]));
    });
#pragma unroll
    // moments are read and written exactly once per step: non-temporal
    // accesses keep them from displacing the ghost-plane lines that
    // neighbouring tiles read (measured +8 % on 256^3 FP32)
    for (int k = 0; k < 10; k++) {
        const auto plane = (const MRLBM_GLOBAL MT*)(tb + woff[k / PW] + (k % PW) * PB);
        if constexpr (S == 1) {
            // binary16: the reconstruction reads only the low half of the
            // word, so the high half is left undefined (no zero extension)
            typedef _Float16 __attribute__((ext_vector_type(1))) half1_t;
            const half1_t x = {__builtin_nontemporal_load((const MRLBM_GLOBAL _Float16*)plane)};
            in.m[k] = __builtin_bit_cast(uint32_t, __builtin_shufflevector(x, x, 0, -1));
        } else {
            in.m[k] = (typename Fetched<S>::T)__builtin_nontemporal_load(plane);
        }
    }
    // wave W owns incoming chunks W, W + NW, ...; every wave fills all
    // RMAX slots (zeros past its last chunk) so that no slot index depends
    // on the wave (a wave-dependent index would move the array to scratch)
    constexpr int RMAX = TileIn<Q, BX, BY, BZ, S>::RMAX;
    dispatch<NW>(w, [&](auto WW) {
        constexpr int W = decltype(WW)::value;
        static_for<RMAX>([&](auto RR) {
            constexpr int r = decltype(RR)::value;
            if constexpr (W + NW * r < NIN)
                in.h[r] = halo_fetch<Q, BX, BY, BZ, W + NW * r>(gin, lane, nb);
            else
                in.h[r] = 0.0f;
        });
    });
}

// One time step of one tile from its fetched inputs (the body of
// k_step_tile and of the batched kernel). The whole workgroup calls it
// with the same tile. `ahead` is called once the inputs are consumed
// (before the first barrier): it issues the next tile's fetch.
template <int Q, int BX, int BY, int BZ, bool HAS_BC, int S, typename Lds, typename Ahead>
__device__ __forceinline__ void tile_update(const StepArgs& a, float* gout, uint32_t step, int tile, Lds& L,
                                            const TileIn<Q, BX, BY, BZ, S>& in, Ahead&& ahead) {
    constexpr int T = BX * BY * BZ;
    constexpr int NW = T / 64;
    using G = HaloGeom<BX, BY, BZ>;
    constexpr int E = G::E;
    constexpr int NIN = TileIn<Q, BX, BY, BZ, S>::NIN;
    constexpr int NOUT = (int)ChunkTables<Q, BX, BY, BZ>::out.size();
    float* s = L.s;
    uint8_t* s_cls = L.cls();

    const int t = thread_index<T>();
    const int tx = t % BX, ty = (t / BX) % BY, tz = t / (BX * BY);
    const int lane = t % 64;
    const int w = __builtin_amdgcn_readfirstlane(t / 64);
    int bx, by, bz;
    tile_xyz(a, tile, bx, by, bz);
    const int base = tx + G::DX * (ty + G::DY * tz);

    const uint32_t cls = in.cls;
    const bool solid = HAS_BC && (cls == 255u);

    // FP16: the reconstruction reads the fetched binary16 words directly
    // (recon_terms_f16); FP32, BF16: decoded moments
    constexpr bool F16 = S == 1;
    constexpr float SU = store_scale(S, 1), SM = store_scale(S, 4);
    Node n;
    if constexpr (!F16) decode_node<S>(a, in.m, n);
    using P = PopLds<Q, BX, BY, BZ>;
    constexpr int PH = P::PHASES;
    // populations entering the tile in phase PP: ghost chunks -> LDS halo
    auto halo_phase = [&](auto PP) {
        dispatch<NW>(w, [&](auto WW) {
            constexpr int W = decltype(WW)::value;
            constexpr int R = (NIN - W + NW - 1) / NW;
            static_for<R>([&](auto RR) {
                constexpr int r = decltype(RR)::value;
                constexpr int K = W + NW * r;
                if constexpr (P::phase(ChunkTables<Q, BX, BY, BZ>::in[K].pop) == decltype(PP)::value)
                    halo_put<Q, BX, BY, BZ, K>(lane, in.h[r], s);
            });
        });
    };
    halo_phase(std::integral_constant<int, 0>());

    // population arrays: node base per (x extent, y extent) pair of the
    // arrays, so every access is a base register plus an immediate
    auto pbase = [&](auto I) {
        constexpr int i = decltype(I)::value + 1;
        return tx + P::dx(i) * (ty + P::dy(i) * tz);
    };
    float g[Q];
    if constexpr (F16) {
        Recon t0;
        float re0;
        recon_terms_f16<Q, false>(in.m, a.rho0, n.drho, n.rho, re0, t0);
        reconstruct_t<Q>(t0, re0, n.drho, g);
    } else {
        reconstruct<Q>(n, g);
    }
    // populations moving along x only: from the x neighbour's lane (row_shr:1
    // for +x, row_shl:1 for -x: x rows lie inside 16-lane DPP rows); the
    // first node of a row takes the entering value from LDS after the barrier
    float gx[Q];
    static_for<Q - 1>([&](auto I) {
        constexpr int i = decltype(I)::value + 1;
        if constexpr (P::x_only(i)) {
            constexpr int ctrl = cxv(i) > 0 ? 0x111 : 0x101;
            // bound_ctrl: a lane whose source lies outside its DPP row reads
            // 0 (no old value to initialise); it is a row's first node,
            // which takes the entering value instead
            gx[i] = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(g[i]), ctrl, 0xF, 0xF, true));
        }
    });
    // each phase: write its populations, barrier, pull them (D3Q27 runs two
    // phases through the same LDS, with a barrier between the first phase's
    // reads and the second phase's writes)
    static_for<PH>([&](auto PP) {
        constexpr int ph = decltype(PP)::value;
        if constexpr (ph > 0) {
            __syncthreads();
            halo_phase(PP);
        }
        static_for<Q - 1>([&](auto I) {
            constexpr int i = decltype(I)::value + 1;
            if constexpr (!P::x_only(i) && P::phase(i) == ph) {
                constexpr int wo = P::cell(i, cxv(i) > 0, cyv(i) > 0, czv(i) > 0);
                s[pbase(I) + wo] = g[i];
            }
        });
        if constexpr (ph == 0) ahead();
        __syncthreads();
        static_for<Q - 1>([&](auto I) {
            constexpr int i = decltype(I)::value + 1;
            if constexpr (P::phase(i) == ph) {
                if constexpr (P::x_only(i)) {
                    const bool edge = cxv(i) > 0 ? tx == 0 : tx == BX - 1;
                    const float h = s[P::off(i) + ty + BY * tz];
                    g[i] = edge ? h : gx[i];
                } else {
                    constexpr int ro = P::cell(i, cxv(i) < 0, cyv(i) < 0, czv(i) < 0);
                    g[i] = s[pbase(I) + ro];
                }
            }
        });
    });

    Node m;
    if constexpr (HAS_BC) {
        // boundary nodes: one pass of moment sums shared with the bulk nodes
        // of the wave, over their known populations only (the others zeroed)
        const bool bcn = cls != 0u && cls != 255u;
        const uint64_t bcw = __builtin_amdgcn_ballot_w64(bcn);
        if (bcw != 0) {  // uniform: waves without boundary nodes skip
            uint32_t known = ~0u;
            if (bcn) known = __float_as_uint(gptr(a.bc)[(size_t)cls * BC_STRIDE]);
            known_only<Q>(known, g);
        }
        Sums sums;
        moment_sums<Q>(g, sums);
        if (bcn) {
            bc_from_sums(gptr(a.bc) + (size_t)cls * BC_STRIDE, sums, a.rho0, m);
            scale_moments(m, SU, SM);
        } else {
            bulk_from_sums(sums, a.rho0, m, SU, SM);
        }
    } else {
        bulk_moments<Q>(g, a.rho0, m, SU, SM);
    }
    collide(m, a.R);
    typename Codec<S>::T q[10];
    // global node index x + nx (y + ny z), mod 2^32: tz is uniform in a wave
    // (waves cover whole x rows, BX * BY is a multiple of 64), so only
    // tx + nx ty is per-lane arithmetic (24-bit multiply: nx < 2^24)
    static_assert((BX * BY) % 64 == 0, "a wave spans whole xy-plane rows");
    const uint32_t nxu = (uint32_t)a.nx, nxy = nxu * (uint32_t)a.ny;
    const uint32_t zw = (uint32_t)(bz * BZ) + (uint32_t)__builtin_amdgcn_readfirstlane(tz);
    const uint32_t node_id = ((uint32_t)(bx * BX) + nxu * (uint32_t)(by * BY) + nxy * zw) + ((uint32_t)tx + __umul24(nxu, (uint32_t)ty));
    Recon rt;
    if constexpr (F16) {
        uint32_t pk[5];
        quantize<S>(m, a, step, node_id, q, pk);
        float re0;
        recon_terms_f16<Q, true>(pk, a.rho0, m.drho, m.rho, re0, rt);
    } else {
        quantize<S>(m, a, step, node_id, q);
        recon_terms<Q>(m, recon_re0<Q>(m), rt);
    }
    if (!solid) store_tile_node<BX, BY, BZ, S>(a, tile, t, q);

    // populations leaving the tile (popSave.inc): every node stages its
    // reconstruction terms in LDS, then whole waves evaluate and write the
    // ghost planes in contiguous runs, each lane only the one population
    // of its chunk (the other populations of a face node stay unevaluated)
    constexpr int NT = staged_terms<Q>();
    __syncthreads();
    static_for<NT>([&](auto KK) {
        constexpr int k = decltype(KK)::value;
        s[k * E + base] = rt.at(k);
    });
    if constexpr (HAS_BC) s_cls[t] = (uint8_t)cls;
    __syncthreads();
    dispatch<NW>(w, [&](auto WW) {
        constexpr int W = decltype(WW)::value;
        static_for<(NOUT + NW - 1) / NW>([&](auto RR) {
            constexpr int K = W + NW * decltype(RR)::value;
            if constexpr (K < NOUT)
                halo_store<Q, BX, BY, BZ, HAS_BC, K>(a, gout, lane, tile, bx, by, bz, s, s_cls);
        });
    });
}

// Waves per SIMD the LDS of a workgroup allows (160 KB per CU, 4 SIMDs),
// capped at 8 (gfx950 holds at most 8 waves per SIMD). The register allocator is asked for that many waves, so that
// a shape whose LDS admits four 512-thread workgroups per CU is not held
// back by scalar registers (the step arguments alone take ~70 of them; a
// few are kept in VGPR lanes instead).
template <int Q, int BX, int BY, int BZ, bool HAS_BC>
constexpr int lds_waves_per_simd() {
    constexpr int bytes = (int)sizeof(TileLds<Q, BX, BY, BZ, HAS_BC>);
    constexpr int granule = 512;
    constexpr int groups = (160 * 1024) / (((bytes + granule - 1) / granule) * granule);
    constexpr int w = groups * (BX * BY * BZ / 64) / 4;
    return w > 8 ? 8 : (w < 1 ? 1 : w);
}

// Register budget of a step body: the bulk body, the FP32 D3Q19 boundary
// bodies with x rows of at most 16 nodes and the 16-bit D3Q19 boundary
// bodies with 8-node rows (~64 VGPRs at 8 waves per SIMD) take the waves the
// LDS allows; the D3Q27 boundary bodies with rows of at most 16 nodes at most
// 6 waves (80 VGPRs; at 8 waves they spill); other boundary bodies keep the
// compiler's choice (a tighter budget spills VGPRs to memory). The ranged
// batched kernel also holds the next item's fetch: its boundary bodies and
// its D3Q27 bodies keep the compiler's choice.
template <int Q, int BX, int BY, int BZ, bool HAS_BC, int S, bool RANGED = false>
constexpr int step_body_waves() {
    constexpr int w = lds_waves_per_simd<Q, BX, BY, BZ, HAS_BC>();
    if constexpr (HAS_BC) {
        if constexpr (RANGED) return 1;
        if constexpr (Q == 19 && BX <= (S == 0 ? 16 : 8)) return w;
        if constexpr (Q == 27 && BX <= 16) return w > 6 ? 6 : w;
        return 1;
    }
    return (RANGED && Q == 27) ? 1 : w;
}

// Tile step: block b updates tile list[b] (tile b when `list` is null), one
// tile per block: workgroup turnover hides the load latency better than a
// resident grid walking tiles with a prefetch of the next one (measured on
// 256^3), and the single-tile body holds fewer scalar registers. For
// lattices with boundary nodes the library launches the all-bulk tiles and
// the rest separately.
template <int Q, int BX, int BY, int BZ, bool HAS_BC, int S>
__global__ __launch_bounds__(BX* BY* BZ) __attribute__((amdgpu_waves_per_eu(step_body_waves<Q, BX, BY, BZ, HAS_BC, S>())))
void k_step_tile(StepArgs a, const uint32_t* __restrict__ list) {
    __shared__ TileLds<Q, BX, BY, BZ, HAS_BC> L;
    const int tile = list ? (int)__builtin_amdgcn_readfirstlane(list[blockIdx.x]) : (int)blockIdx.x;
    TileIn<Q, BX, BY, BZ, S> in;
    tile_fetch<Q, BX, BY, BZ, HAS_BC, S>(a, a.gin, tile, in);
    tile_update<Q, BX, BY, BZ, HAS_BC, S>(a, a.gout, a.step, tile, L, in, []() {});
}

// Batched step: each block takes a contiguous range of work items (the
// item's work word and instance arguments then mostly hit the scalar cache).
// HAS_BC = false runs tiles whose nodes are all bulk (the lighter body, more
// workgroups per CU); HAS_BC = true runs the tiles that hold boundary or
// solid nodes, and tiles of lattices without a class map read class 0.
template <int Q, int BX, int BY, int BZ, int S, bool HAS_BC>
__global__ __launch_bounds__(BX* BY* BZ) __attribute__((amdgpu_waves_per_eu(step_body_waves<Q, BX, BY, BZ, HAS_BC, S, true>())))
void k_step_batch(const InstArgs* __restrict__ tab,
                                                           const uint32_t* __restrict__ work, int nwork, int k) {
    __shared__ TileLds<Q, BX, BY, BZ, HAS_BC> L;
    using In = TileIn<Q, BX, BY, BZ, S>;
    const int per = (nwork + (int)gridDim.x - 1) / (int)gridDim.x;
    const int begin = (int)blockIdx.x * per;
    const int end = begin + per < nwork ? begin + per : nwork;
    if (begin >= end) return;
    auto item_of = [&](int item, int& inst, int& tile) {
        const uint32_t wv = __builtin_amdgcn_readfirstlane(work[item]);
        inst = (int)(wv >> 20);
        tile = (int)(wv & 0xFFFFFu);
    };
    int inst, tile;
    item_of(begin, inst, tile);
    In in;
    {
        const InstArgs& ia = tab[inst];  // read in place (scalar loads)
        tile_fetch<Q, BX, BY, BZ, HAS_BC, S>(ia.a, ia.g[(ia.cur0 + k) & 1], tile, in);
    }
    for (int item = begin; item < end; item++) {
        const InstArgs& ia = tab[inst];
        const int p = (ia.cur0 + k) & 1;
        int ninst = inst, ntile = tile;
        if (item + 1 < end) item_of(item + 1, ninst, ntile);
        In nx;
        tile_update<Q, BX, BY, BZ, HAS_BC, S>(ia.a, ia.g[1 - p], ia.step0 + (uint32_t)k, tile, L, in, [&]() {
            if (item + 1 < end) {
                const InstArgs& nia = tab[ninst];
                tile_fetch<Q, BX, BY, BZ, HAS_BC, S>(nia.a, nia.g[(nia.cur0 + k) & 1], ntile, nx);
            }
        });
        __syncthreads();  // LDS is reused by the next item
        in = nx;
        inst = ninst;
        tile = ntile;
    }
}

// Batched step with one work item per block (the default grid): like
// k_step_tile, the single-item body holds far fewer registers than the
// ranged kernel's loop with its prefetch (D3Q19 8x8x8 boundary body: ~76
// instead of ~110 VGPRs, three workgroups per CU instead of two).
template <int Q, int BX, int BY, int BZ, int S, bool HAS_BC>
__global__ __launch_bounds__(BX* BY* BZ) __attribute__((amdgpu_waves_per_eu(step_body_waves<Q, BX, BY, BZ, HAS_BC, S>())))
void k_step_batch1(const InstArgs* __restrict__ tab, const uint32_t* __restrict__ work, int k) {
    __shared__ TileLds<Q, BX, BY, BZ, HAS_BC> L;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(work[blockIdx.x]);
    const int inst = (int)(wv >> 20), tile = (int)(wv & 0xFFFFFu);
    const InstArgs& ia = tab[inst];  // read in place (scalar loads)
    const int p = (ia.cur0 + k) & 1;
    TileIn<Q, BX, BY, BZ, S> in;
    tile_fetch<Q, BX, BY, BZ, HAS_BC, S>(ia.a, ia.g[p], tile, in);
    tile_update<Q, BX, BY, BZ, HAS_BC, S>(ia.a, ia.g[1 - p], ia.step0 + (uint32_t)k, tile, L, in, []() {});
}

// Both bodies in one launch, where they run at the same occupancy (the same
// LDS size and register budget: D3Q19 8x8x8 in every storage format, FP32
// 16x8x8): blocks [0, nbulk) take all-bulk tiles with the bulk body, the
// others tiles with boundary or solid nodes. One launch per step instead of
// two: the second launch's ramp-up and the first one's tail overlap.
template <int Q, int BX, int BY, int BZ, int S>
constexpr bool mixed_launch() {
    return step_body_waves<Q, BX, BY, BZ, false, S>() == step_body_waves<Q, BX, BY, BZ, true, S>() &&
           sizeof(TileLds<Q, BX, BY, BZ, false>) == sizeof(TileLds<Q, BX, BY, BZ, true>);
}

template <int Q, int BX, int BY, int BZ, int S>
__global__ __launch_bounds__(BX* BY* BZ) __attribute__((amdgpu_waves_per_eu(step_body_waves<Q, BX, BY, BZ, true, S>())))
void k_step_tile_mixed(StepArgs a, const uint32_t* __restrict__ list, int nbulk) {
    __shared__ TileLds<Q, BX, BY, BZ, true> L;
    const int tile = (int)__builtin_amdgcn_readfirstlane(list[blockIdx.x]);
    if ((int)blockIdx.x < nbulk) {
        TileIn<Q, BX, BY, BZ, S> in;
        tile_fetch<Q, BX, BY, BZ, false, S>(a, a.gin, tile, in);
        tile_update<Q, BX, BY, BZ, false, S>(a, a.gout, a.step, tile, L, in, []() {});
    } else {
        TileIn<Q, BX, BY, BZ, S> in;
        tile_fetch<Q, BX, BY, BZ, true, S>(a, a.gin, tile, in);
        tile_update<Q, BX, BY, BZ, true, S>(a, a.gout, a.step, tile, L, in, []() {});
    }
}

template <int Q, int BX, int BY, int BZ, int S>
__global__ __launch_bounds__(BX* BY* BZ) __attribute__((amdgpu_waves_per_eu(step_body_waves<Q, BX, BY, BZ, true, S>())))
void k_step_batch1_mixed(const InstArgs* __restrict__ tab, const uint32_t* __restrict__ work, int nbulk, int k) {
    __shared__ TileLds<Q, BX, BY, BZ, true> L;
    const uint32_t wv = __builtin_amdgcn_readfirstlane(work[blockIdx.x]);
    const int inst = (int)(wv >> 20), tile = (int)(wv & 0xFFFFFu);
    const InstArgs& ia = tab[inst];  // read in place (scalar loads)
    const int p = (ia.cur0 + k) & 1;
    if ((int)blockIdx.x < nbulk) {
        TileIn<Q, BX, BY, BZ, S> in;
        tile_fetch<Q, BX, BY, BZ, false, S>(ia.a, ia.g[p], tile, in);
        tile_update<Q, BX, BY, BZ, false, S>(ia.a, ia.g[1 - p], ia.step0 + (uint32_t)k, tile, L, in, []() {});
    } else {
        TileIn<Q, BX, BY, BZ, S> in;
        tile_fetch<Q, BX, BY, BZ, true, S>(ia.a, ia.g[p], tile, in);
        tile_update<Q, BX, BY, BZ, true, S>(ia.a, ia.g[1 - p], ia.step0 + (uint32_t)k, tile, L, in, []() {});
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
    float pop[Q];
    if constexpr (S == 1) {  // as the step kernels reconstruct from fetched binary16 words
        const auto mom = gptr((const uint16_t*)a.mom);
        uint32_t w[10];
        for (int k = 0; k < 10; k++) w[k] = mom[(size_t)tile * 10 * T + (size_t)k * T + t];
        Recon rt;
        float drho, rho, re0;
        recon_terms_f16<Q, false>(w, a.rho0, drho, rho, re0, rt);
        reconstruct_t<Q>(rt, re0, drho, pop);
    } else {
        Node n;
        load_node<Q, BX, BY, BZ, S>(a, (size_t)tile * 10 * T + t, n);
        reconstruct<Q>(n, pop);
    }
    if (tx == 0) put_face_c<Q, BX, BY, BZ, 0>(g.f[0], ty, tz, tile, pop);
    if (tx == BX - 1) put_face_c<Q, BX, BY, BZ, 1>(g.f[1], ty, tz, tile, pop);
    if (ty == 0) put_face_c<Q, BX, BY, BZ, 2>(g.f[2], tx, tz, tile, pop);
    if (ty == BY - 1) put_face_c<Q, BX, BY, BZ, 3>(g.f[3], tx, tz, tile, pop);
    if (tz == 0) put_face_c<Q, BX, BY, BZ, 4>(g.f[4], tx, ty, tile, pop);
    if (tz == BZ - 1) put_face_c<Q, BX, BY, BZ, 5>(g.f[5], tx, ty, tile, pop);
}


template <int Q, int BX, int BY, int BZ, bool HAS_BC, int S>
static hipError_t launch_qsb(const StepArgs& a, const uint32_t* list, int n, hipStream_t s) {
    constexpr int T = BX * BY * BZ;
    hipLaunchKernelGGL((k_step_tile<Q, BX, BY, BZ, HAS_BC, S>), dim3(n), dim3(T), 0, s, a, list);
    return hipGetLastError();
}

template <int S>
hipError_t launch_step_tile_s(int q, int bx, int by, int bz, bool has_bc, const StepArgs& a, const uint32_t* list,
                              int n, hipStream_t s) {
#define X(A, B, C)                                                                                        \
    if (bx == A && by == B && bz == C) {                                                                  \
        if (q == 19)                                                                                      \
            return has_bc ? launch_qsb<19, A, B, C, true, S>(a, list, n, s)                               \
                          : launch_qsb<19, A, B, C, false, S>(a, list, n, s);                             \
        if (q == 27)                                                                                      \
            return has_bc ? launch_qsb<27, A, B, C, true, S>(a, list, n, s)                               \
                          : launch_qsb<27, A, B, C, false, S>(a, list, n, s);                             \
    }
    MRLBM_TILE_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

template <int Q, int BX, int BY, int BZ, int S, bool HAS_BC>
static hipError_t launch_batch_qb(const InstArgs* tab, const uint32_t* work, int nwork, int grid, int k,
                                  hipStream_t s) {
    dim3 b(BX * BY * BZ);
    if (grid <= 0 || grid >= nwork) {  // one item per block
        hipLaunchKernelGGL((k_step_batch1<Q, BX, BY, BZ, S, HAS_BC>), dim3(nwork), b, 0, s, tab, work, k);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_step_batch<Q, BX, BY, BZ, S, HAS_BC>), dim3(grid), b, 0, s, tab, work, nwork, k);
    return hipGetLastError();
}

template <int S>
hipError_t launch_step_batch_s(int q, int bx, int by, int bz, bool has_bc, const InstArgs* tab, const uint32_t* work,
                               int nwork, int grid, int k, hipStream_t s) {
#define X(A, B, C)                                                                                        \
    if (bx == A && by == B && bz == C) {                                                                  \
        if (q == 19)                                                                                      \
            return has_bc ? launch_batch_qb<19, A, B, C, S, true>(tab, work, nwork, grid, k, s)           \
                          : launch_batch_qb<19, A, B, C, S, false>(tab, work, nwork, grid, k, s);         \
        if (q == 27)                                                                                      \
            return has_bc ? launch_batch_qb<27, A, B, C, S, true>(tab, work, nwork, grid, k, s)           \
                          : launch_batch_qb<27, A, B, C, S, false>(tab, work, nwork, grid, k, s);         \
    }
    MRLBM_TILE_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

template <int S>
bool mixed_launch_s(int q, int bx, int by, int bz) {
#define X(A, B, C)                                                                                        \
    if (bx == A && by == B && bz == C) return q == 19 ? mixed_launch<19, A, B, C, S>() : mixed_launch<27, A, B, C, S>();
    MRLBM_TILE_SHAPES(X)
#undef X
    return false;
}

template <int Q, int BX, int BY, int BZ, int S>
static hipError_t launch_mixed_q(const StepArgs* a, const InstArgs* tab, const uint32_t* list, int n, int nbulk, int k,
                                 hipStream_t s) {
    if constexpr (mixed_launch<Q, BX, BY, BZ, S>()) {
        if (a)
            hipLaunchKernelGGL((k_step_tile_mixed<Q, BX, BY, BZ, S>), dim3(n), dim3(BX * BY * BZ), 0, s, *a, list, nbulk);
        else
            hipLaunchKernelGGL((k_step_batch1_mixed<Q, BX, BY, BZ, S>), dim3(n), dim3(BX * BY * BZ), 0, s, tab, list,
                               nbulk, k);
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

// One launch over n tiles (a != null: list of one lattice's tiles) or work
// items (a == null: the batch's work table), the first nbulk all-bulk.
template <int S>
hipError_t launch_step_mixed_s(int q, int bx, int by, int bz, const StepArgs* a, const InstArgs* tab,
                               const uint32_t* list, int n, int nbulk, int k, hipStream_t s) {
#define X(A, B, C)                                                                                        \
    if (bx == A && by == B && bz == C)                                                                    \
        return q == 19 ? launch_mixed_q<19, A, B, C, S>(a, tab, list, n, nbulk, k, s)                     \
                       : launch_mixed_q<27, A, B, C, S>(a, tab, list, n, nbulk, k, s);
    MRLBM_TILE_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

template <int S>
hipError_t launch_init_ghosts_s(int q, int bx, int by, int bz, const StepArgs& a, const GhostOut& go, int ntiles,
                                hipStream_t s) {
#define X(A, B, C)                                                                                        \
    if (bx == A && by == B && bz == C) {                                                                  \
        if (q == 19)                                                                                      \
            hipLaunchKernelGGL((k_init_ghosts<19, A, B, C, S>), dim3(ntiles), dim3(A * B * C), 0, s, a, go); \
        else                                                                                              \
            hipLaunchKernelGGL((k_init_ghosts<27, A, B, C, S>), dim3(ntiles), dim3(A * B * C), 0, s, a, go); \
        return hipGetLastError();                                                                         \
    }
    MRLBM_TILE_SHAPES(X)
#undef X
    return hipErrorInvalidValue;
}

}  // namespace mrlbm
