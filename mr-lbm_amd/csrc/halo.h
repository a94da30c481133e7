// Halo layout of the tile kernel's LDS and the compile-time tables that
// move ghost-plane data between HBM and LDS.
//
// Population i of the tile lives in an LDS array of extents B_a + |c_ia|
// per axis (PopLds). A node n writes its outgoing population at
// n + [c_i > 0] and reads its incoming one at n + [c_i < 0] (per axis), so
// both are a base address (one per pair of x, y extents) plus a
// compile-time offset: no wrap arithmetic, no branch. The extra plane on the
// upwind side of an array holds the populations that enter the tile; they
// are copied from the neighbour tiles' ghost planes by whole waves
// ("chunks"), each chunk a window of whole rows of a ghost plane (at most 64
// cells, contiguous in HBM). When a source node lies outside the tile along
// several axes, the value comes from the face of the last of those axes in
// X, Y, Z order, which is the branch that wins in the reference's
// popLoad.inc (src/includeFiles/popLoad.inc:2-86).
// Populations that move along x only stay in their wave when x rows lie in
// 16-lane DPP rows (BX = 8 or 16): they are shifted between lanes, and their
// LDS array is only the entering plane.
#pragma once

#include <array>

#include "lattice.h"

namespace mrlbm {

// A chunk is a window of whole rows of one face plane of one slot: `rows`
// rows of Bu cells (rows * Bu <= 64), lane l <-> face cell (u, v) =
// (l % Bu, w0 + l / Bu), so its ghost-plane index is l + Bu * w0 (a
// constant offset from the lane; lanes past the window re-read it). Lanes
// outside [ulo, uhi] x [vlo, vhi] only load. Bu (the face's first
// coordinate extent) must divide 64.
struct Chunk {
    int face;        // ghost face array (0 X_0, 1 X_1, 2 Y_0, 3 Y_1, 4 Z_0, 5 Z_1)
    int dx, dy, dz;  // offset of the ghost-plane owner tile (incoming chunks)
    int slot;        // slot within the face
    int pop;         // population
    int w0;          // first row of the window
    int rows;        // rows of the window (rows * Bu <= 64, a power of two)
    int ulo, uhi;    // cells of a row that are moved
    int vlo, vhi;    // rows that are moved (within the window)
    int lds0;        // LDS index (floats) of cell (0, w0) under the map below
    int su, sv;      // LDS index = lds0 + u * su + (v - w0) * sv
};
// Incoming chunks index the population arrays (PopLds); outgoing chunks
// index the face node in the staged-term layout, node base
// x + (BX+1) (y + BY z) of term array k at k * HaloGeom::E. The one
// padding column makes the row stride odd: a wave's staged-term reads of
// an x face (rows along y, 8 x 8 cells) and of a y face (rows along x, z
// stride (BX+1) BY = 72 at 8x8x8) fall into 64 distinct LDS banks.

template <int BX, int BY, int BZ>
struct HaloGeom {
    static constexpr int DX = BX + 1, DY = BY, DZ = BZ;
    static constexpr int E = DX * DY * DZ;  // cells per staged-term array
    static constexpr int B[3] = {BX, BY, BZ};
    static constexpr int STRIDE[3] = {1, DX, DX * DY};
};

// LDS arrays of the populations: extents B_a + |c_a| (the entering plane on
// the upwind side), or the entering x plane only for populations that move
// along x alone when BX = 8 or 16 (DPP lane shifts carry the rest).
template <int BX>
__host__ __device__ constexpr bool pop_x_only(int i) {
    return (BX == 8 || BX == 16) && cyv(i) == 0 && czv(i) == 0 && cxv(i) != 0;
}
template <int BX, int BY, int BZ>
__host__ __device__ constexpr int pop_cells(int i) {
    return (pop_x_only<BX>(i) ? 1 : BX + (cxv(i) != 0)) * (BY + (cyv(i) != 0)) * (BZ + (czv(i) != 0));
}

// Where each population's array lives. D3Q27 streams its populations
// through LDS in two phases that reuse the same space (two barriers more),
// so that its arrays take about half the LDS: 8x8x8 fits four workgroups per
// CU instead of two. Opposite pairs (2k-1, 2k) go, largest first, to the
// phase holding fewer cells so far. D3Q19 streams in one phase.
struct PopLayout {
    int phase[27];
    int off[27];    // array offset within its phase's space
    int total[2];   // cells of each phase
};
template <int Q, int BX, int BY, int BZ>
constexpr PopLayout make_pop_layout() {
    PopLayout L{};
    constexpr int PHASES = Q == 27 ? 2 : 1;
    if (PHASES == 2) {
        bool done[14] = {};
        for (int n = 0; n < (Q - 1) / 2; n++) {
            int best = 0, bs = -1;
            for (int k = 1; k <= (Q - 1) / 2; k++) {
                const int sz = pop_cells<BX, BY, BZ>(2 * k - 1) + pop_cells<BX, BY, BZ>(2 * k);
                if (!done[k] && sz > bs) bs = sz, best = k;
            }
            done[best] = true;
            const int p = L.total[0] <= L.total[1] ? 0 : 1;
            L.total[p] += bs;
            L.phase[2 * best - 1] = L.phase[2 * best] = p;
        }
    }
    int used[2] = {0, 0};
    for (int i = 1; i < Q; i++) {
        L.off[i] = used[L.phase[i]];
        used[L.phase[i]] += pop_cells<BX, BY, BZ>(i);
    }
    L.total[0] = used[0];
    L.total[1] = used[1];
    return L;
}
template <int Q, int BX, int BY, int BZ>
inline constexpr PopLayout pop_layout_v = make_pop_layout<Q, BX, BY, BZ>();

template <int Q, int BX, int BY, int BZ>
struct PopLds {
    static constexpr bool DPP_X = BX == 8 || BX == 16;
    static constexpr int PHASES = Q == 27 ? 2 : 1;
    __host__ __device__ static constexpr bool x_only(int i) { return pop_x_only<BX>(i); }
    __host__ __device__ static constexpr int dx(int i) { return x_only(i) ? 1 : BX + (cxv(i) != 0); }
    __host__ __device__ static constexpr int dy(int i) { return BY + (cyv(i) != 0); }
    __host__ __device__ static constexpr int dz(int i) { return BZ + (czv(i) != 0); }
    __host__ __device__ static constexpr int size(int i) { return dx(i) * dy(i) * dz(i); }
    __host__ __device__ static constexpr int phase(int i) { return pop_layout_v<Q, BX, BY, BZ>.phase[i]; }
    // offset of population i's array within its phase's space
    __host__ __device__ static constexpr int off(int i) { return pop_layout_v<Q, BX, BY, BZ>.off[i]; }
    // cells of the largest phase (the LDS the arrays need)
    __host__ __device__ static constexpr int total() {
        constexpr PopLayout L = pop_layout_v<Q, BX, BY, BZ>;
        return L.total[0] > L.total[1] ? L.total[0] : L.total[1];
    }
    // stride of array axis a
    __host__ __device__ static constexpr int stride(int i, int a) {
        return a == 0 ? (x_only(i) ? 0 : 1) : (a == 1 ? dx(i) : dx(i) * dy(i));
    }
    // cell at array coordinates e (x coordinate ignored for x-only arrays)
    __host__ __device__ static constexpr int cell(int i, int ex, int ey, int ez) {
        return off(i) + (x_only(i) ? 0 : ex) + dx(i) * (ey + dy(i) * ez);
    }
};

__host__ __device__ constexpr int face_slot(int face, int pop, int qf) {
    for (int s = 0; s < qf; s++)
        if (face_pop(face, s) == pop) return s;
    return -1;
}

__host__ __device__ constexpr int cvec(int i, int a) { return a == 0 ? cxv(i) : (a == 1 ? cyv(i) : czv(i)); }

// face coordinate axes of a face on axis a: u is x unless the face is an
// x face (then y); v is z unless the face is a z face (then y)
__host__ __device__ constexpr int face_u_axis(int a) { return a == 0 ? 1 : 0; }
__host__ __device__ constexpr int face_v_axis(int a) { return a == 2 ? 1 : 2; }

template <int BX, int BY, int BZ>
__host__ __device__ constexpr int face_bu(int face) {
    return HaloGeom<BX, BY, BZ>::B[face_u_axis(face / 2)];
}
template <int BX, int BY, int BZ>
__host__ __device__ constexpr int face_bv(int face) {
    return HaloGeom<BX, BY, BZ>::B[face_v_axis(face / 2)];
}
template <int BX, int BY, int BZ>
__host__ __device__ constexpr int window_rows(int face) {
    const int bu = face_bu<BX, BY, BZ>(face), bv = face_bv<BX, BY, BZ>(face);
    const int r = 64 / bu > 0 ? 64 / bu : 1;
    return r < bv ? r : bv;
}

// Split rows [vlo, vhi] of a face into windows; calls emit(w0, vlo', vhi').
template <typename F>
constexpr void for_windows(int bv, int rows, int vlo, int vhi, F&& emit) {
    int w = vlo;
    while (w <= vhi) {
        const int w0 = w + rows <= bv ? w : bv - rows;
        const int hi = (w0 + rows - 1) < vhi ? (w0 + rows - 1) : vhi;
        emit(w0, w, hi);
        w = hi + 1;
    }
}

// Incoming chunks of population i (0 < i < Q). Called twice: with out ==
// nullptr to count, then to fill.
template <int Q, int BX, int BY, int BZ>
constexpr int incoming_chunks_of(int i, Chunk* out) {
    using G = HaloGeom<BX, BY, BZ>;
    constexpr int QF = VS<Q>::QF;
    const int c[3] = {cvec(i, 0), cvec(i, 1), cvec(i, 2)};
    int n = 0;
    for (int A = 1; A < 8; A++) {  // axes where the source lies outside the tile
        bool ok = true;
        for (int a = 0; a < 3; a++)
            if (((A >> a) & 1) && c[a] == 0) ok = false;
        if (!ok) continue;
        const int win = (A & 4) ? 2 : ((A & 2) ? 1 : 0);
        const int face = 2 * win + (c[win] > 0 ? 1 : 0);
        const int ua = face_u_axis(win), va = face_v_axis(win);
        // per axis: source range in the owner tile, LDS cell of its first
        // source, and whether the LDS cell moves with the source
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, e0[3] = {0, 0, 0}, mv[3] = {0, 0, 0}, d[3] = {0, 0, 0};
        for (int a = 0; a < 3; a++) {
            const int Ba = G::B[a];
            if ((A >> a) & 1) {
                // outside: source coordinate -1 (c>0) or B (c<0); in the owner
                // tile it is B-1 or 0; halo index 0 or B
                lo[a] = hi[a] = c[a] > 0 ? Ba - 1 : 0;
                e0[a] = c[a] > 0 ? 0 : Ba;
                d[a] = -c[a];
            } else {
                // inside: sources whose destination stays in the tile
                lo[a] = c[a] < 0 ? 1 : 0;
                hi[a] = c[a] > 0 ? Ba - 2 : Ba - 1;
                e0[a] = c[a] > 0 ? 1 : 0;  // LDS cell = source + [c > 0]
                mv[a] = 1;
            }
        }
        if (hi[ua] < lo[ua] || hi[va] < lo[va]) continue;
        const int bv = G::B[va];
        // a single fixed row is read alone (a taller window would fetch
        // lines no cell needs)
        const int rows = mv[va] ? window_rows<BX, BY, BZ>(face) : 1;
        for_windows(bv, rows, lo[va], hi[va], [&](int w0, int vlo, int vhi) {
            if (out) {
                Chunk ch{};
                ch.face = face;
                ch.dx = d[0];
                ch.dy = d[1];
                ch.dz = d[2];
                ch.slot = face_slot(face, i, QF);
                ch.pop = i;
                ch.w0 = w0;
                ch.rows = rows;
                ch.ulo = lo[ua];
                ch.uhi = hi[ua];
                ch.vlo = vlo;
                ch.vhi = vhi;
                using P = PopLds<Q, BX, BY, BZ>;
                ch.su = mv[ua] ? P::stride(i, ua) : 0;
                ch.sv = mv[va] ? P::stride(i, va) : 0;
                // LDS cell of face cell (0, w0): moving axes at source 0 / w0
                int e[3] = {e0[0], e0[1], e0[2]};
                if (mv[va]) e[va] += w0;
                ch.lds0 = P::cell(i, e[0], e[1], e[2]);
                out[n] = ch;
            }
            n++;
        });
    }
    return n;
}

template <int Q, int BX, int BY, int BZ>
constexpr int n_incoming_chunks() {
    int n = 0;
    for (int i = 1; i < Q; i++) n += incoming_chunks_of<Q, BX, BY, BZ>(i, nullptr);
    return n;
}

template <int Q, int BX, int BY, int BZ>
constexpr std::array<Chunk, n_incoming_chunks<Q, BX, BY, BZ>()> incoming_chunks() {
    std::array<Chunk, n_incoming_chunks<Q, BX, BY, BZ>()> arr{};
    int n = 0;
    for (int i = 1; i < Q; i++) n += incoming_chunks_of<Q, BX, BY, BZ>(i, &arr[n]);
    return arr;
}

// Outgoing chunks: face f of this tile, slot s, all face nodes; the LDS
// index is the face node's base in the staged-term layout.
template <int Q, int BX, int BY, int BZ>
constexpr int outgoing_chunks_impl(Chunk* out) {
    using G = HaloGeom<BX, BY, BZ>;
    constexpr int QF = VS<Q>::QF;
    int n = 0;
    for (int face = 0; face < 6; face++) {
        const int a = face / 2;
        const int ua = face_u_axis(a), va = face_v_axis(a);
        const int bv = G::B[va];
        const int rows = window_rows<BX, BY, BZ>(face);
        for (int s = 0; s < QF; s++) {
            const int i = face_pop(face, s);
            for_windows(bv, rows, 0, bv - 1, [&](int w0, int vlo, int vhi) {
                if (out) {
                    Chunk ch{};
                    ch.face = face;
                    ch.slot = s;
                    ch.pop = i;
                    ch.w0 = w0;
                    ch.rows = rows;
                    ch.ulo = 0;
                    ch.uhi = G::B[ua] - 1;
                    ch.vlo = vlo;
                    ch.vhi = vhi;
                    // staged-term base of the face node (u, v) = (0, w0)
                    int node[3] = {0, 0, 0};
                    node[a] = (face & 1) ? G::B[a] - 1 : 0;
                    node[va] = w0;
                    ch.lds0 = node[0] + G::DX * (node[1] + G::DY * node[2]);
                    ch.su = G::STRIDE[ua];
                    ch.sv = G::STRIDE[va];
                    out[n] = ch;
                }
                n++;
            });
        }
    }
    return n;
}

template <int Q, int BX, int BY, int BZ>
constexpr int n_outgoing_chunks() {
    return outgoing_chunks_impl<Q, BX, BY, BZ>(nullptr);
}

template <int Q, int BX, int BY, int BZ>
constexpr std::array<Chunk, n_outgoing_chunks<Q, BX, BY, BZ>()> outgoing_chunks() {
    std::array<Chunk, n_outgoing_chunks<Q, BX, BY, BZ>()> arr{};
    outgoing_chunks_impl<Q, BX, BY, BZ>(&arr[0]);
    return arr;
}

// Both tables of a tile shape, evaluated once per shape.
template <int Q, int BX, int BY, int BZ>
struct ChunkTables {
    static constexpr auto in = incoming_chunks<Q, BX, BY, BZ>();
    static constexpr auto out = outgoing_chunks<Q, BX, BY, BZ>();
};

}  // namespace mrlbm
