// Halo layout of the tile kernel's LDS and the compile-time tables that
// move ghost-plane data between HBM and LDS.
//
// Population i of the tile lives in an LDS array of (BX+1)(BY+1)(BZ+1)
// cells. A node n writes its outgoing population at n + [c_i > 0] and reads
// its incoming one at n + [c_i < 0] (per axis), so both are the thread's
// base address plus a compile-time offset: no wrap arithmetic, no branch.
// The extra plane on the upwind side of each array holds the populations
// that enter the tile; they are copied from the neighbour tiles' ghost
// planes by whole waves ("chunks"), each chunk a box of at most 64 cells
// that is contiguous in the ghost plane. When a source node lies outside
// the tile along several axes, the value comes from the face of the last of
// those axes in X, Y, Z order, which is the branch that wins in the
// reference's popLoad.inc (src/includeFiles/popLoad.inc:2-86).
#pragma once

#include <array>

#include "lattice.h"

namespace mrlbm {

struct Chunk {
    int face;        // ghost face array (0 X_0, 1 X_1, 2 Y_0, 3 Y_1, 4 Z_0, 5 Z_1)
    int dx, dy, dz;  // offset of the ghost-plane owner tile
    int slot;        // slot within the face
    int pop;         // population
    int u0, v0;      // first cell in face coordinates of the owner tile
    int nu, nv;      // extent (nu * nv <= 64)
    int lds0;        // LDS index (floats) of cell (u0, v0)
    int su, sv;      // LDS strides along u and v
};

template <int BX, int BY, int BZ>
struct HaloGeom {
    static constexpr int DX = BX + 1, DY = BY + 1, DZ = BZ + 1;
    static constexpr int E = DX * DY * DZ;  // cells per population array
    static constexpr int B[3] = {BX, BY, BZ};
    static constexpr int STRIDE[3] = {1, DX, DX * DY};
};

__host__ __device__ constexpr int face_slot(int face, int pop, int qf) {
    for (int s = 0; s < qf; s++)
        if (face_pop(face, s) == pop) return s;
    return -1;
}

__host__ __device__ constexpr int cvec(int i, int a) { return a == 0 ? cxv(i) : (a == 1 ? cyv(i) : czv(i)); }

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
        const int ua = win == 0 ? 1 : 0, va = win == 2 ? 1 : 2;  // face coordinate axes
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, e0[3] = {0, 0, 0}, d[3] = {0, 0, 0};
        for (int a = 0; a < 3; a++) {
            const int Ba = G::B[a];
            if ((A >> a) & 1) {
                // outside: source coordinate -1 (c>0) or B (c<0); in the owner
                // tile it is B-1 or 0; halo index 0 or B
                lo[a] = hi[a] = c[a] > 0 ? Ba - 1 : 0;
                e0[a] = c[a] > 0 ? 0 : Ba;
                d[a] = -c[a];
            } else {
                // inside: source range whose destination stays in the tile
                lo[a] = c[a] < 0 ? 1 : 0;
                hi[a] = c[a] > 0 ? Ba - 2 : Ba - 1;
                e0[a] = lo[a] + (c[a] > 0 ? 1 : 0);
            }
        }
        const int nu = hi[ua] - lo[ua] + 1, nvt = hi[va] - lo[va] + 1;
        if (nu <= 0 || nvt <= 0) continue;
        const int vstep = 64 / nu > 0 ? 64 / nu : 1;
        for (int v = 0; v < nvt; v += vstep) {
            if (out) {
                Chunk ch{};
                ch.face = face;
                ch.dx = d[0];
                ch.dy = d[1];
                ch.dz = d[2];
                ch.slot = face_slot(face, i, QF);
                ch.pop = i;
                ch.u0 = lo[ua];
                ch.v0 = lo[va] + v;
                ch.nu = nu;
                ch.nv = (nvt - v) < vstep ? (nvt - v) : vstep;
                int e[3] = {e0[0], e0[1], e0[2]};
                e[va] += v;
                ch.lds0 = (i - 1) * G::E + e[0] + G::DX * (e[1] + G::DY * e[2]);
                ch.su = G::STRIDE[ua];
                ch.sv = G::STRIDE[va];
                out[n] = ch;
            }
            n++;
        }
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
// index is the node's write position of that population.
template <int Q, int BX, int BY, int BZ>
constexpr int outgoing_chunks_impl(Chunk* out) {
    using G = HaloGeom<BX, BY, BZ>;
    constexpr int QF = VS<Q>::QF;
    int n = 0;
    for (int face = 0; face < 6; face++) {
        const int a = face / 2;
        const int ua = a == 0 ? 1 : 0, va = a == 2 ? 1 : 2;
        for (int s = 0; s < QF; s++) {
            const int i = face_pop(face, s);
            const int c[3] = {cvec(i, 0), cvec(i, 1), cvec(i, 2)};
            const int nu = G::B[ua], nvt = G::B[va];
            const int vstep = 64 / nu > 0 ? 64 / nu : 1;
            for (int v = 0; v < nvt; v += vstep) {
                if (out) {
                    Chunk ch{};
                    ch.face = face;
                    ch.slot = s;
                    ch.pop = i;
                    ch.u0 = 0;
                    ch.v0 = v;
                    ch.nu = nu;
                    ch.nv = (nvt - v) < vstep ? (nvt - v) : vstep;
                    int node[3] = {0, 0, 0};
                    node[a] = (face & 1) ? G::B[a] - 1 : 0;
                    node[va] = v;
                    int e[3] = {0, 0, 0};
                    for (int b = 0; b < 3; b++) e[b] = node[b] + (c[b] > 0 ? 1 : 0);
                    ch.lds0 = (i - 1) * G::E + e[0] + G::DX * (e[1] + G::DY * e[2]);
                    ch.su = G::STRIDE[ua];
                    ch.sv = G::STRIDE[va];
                    out[n] = ch;
                }
                n++;
            }
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

}  // namespace mrlbm
