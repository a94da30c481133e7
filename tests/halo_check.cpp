// Host check of the tile kernel's ghost-chunk tables (mr-lbm_amd/csrc/halo.h):
// simulates every lane of every chunk and compares with a brute-force map.
//
// LDS geometry, restated here independently of halo.h: population i has an
// array of extents (B_a + |c_ia|) per axis, stored one after another in
// population order, x fastest, within the LDS space of its streaming phase
// (the phase assignment itself is taken from halo.h: D3Q27 streams in two
// phases that reuse one space, D3Q19 in one); when BX is 8 or 16, a
// population moving along x only keeps just its entering x plane (extent 1
// along x). A cell is identified by (phase, index).
// Incoming: for each node n and population i whose source n - c_i lies
// outside the tile along a set A of axes, the LDS read cell n + [c_i < 0]
// must receive, exactly once, the ghost value of face 2*win + [c_win > 0]
// (win = last axis of A in X, Y, Z order, the popLoad.inc branch order) of
// the owner tile at offset -c along A, slot of i, at the source's face
// coordinates; no other cell may be written. Every load index must lie in
// the owner's plane of that slot.
// Outgoing: every face cell of every slot is written exactly once, by a lane
// whose LDS index is the face node's base x + (BX+1)(y + BY z) in the
// staged-term layout; at 8x8x8 the 64 lanes of an x- or y-face chunk read
// 64 distinct LDS banks (address mod 64).
// Prints "ok" and exits 0 on success.
#include <cstdio>
#include <map>
#include <set>
#include <tuple>

#include "../mr-lbm_amd/csrc/halo.h"

using namespace mrlbm;

static int failures = 0;
#define CHECK(cond, ...)                     \
    do {                                     \
        if (!(cond)) {                       \
            if (failures < 20) {             \
                std::printf(__VA_ARGS__);    \
                std::printf("\n");           \
            }                                \
            failures++;                      \
        }                                    \
    } while (0)

constexpr int PHASE_KEY = 1 << 20;  // cell key = phase * PHASE_KEY + index within the phase

// LDS cell of population i at array coordinates e (the geometry above)
template <int Q, int BX, int BY, int BZ>
int pop_cell(int i, const int e[3]) {
    const int B[3] = {BX, BY, BZ};
    auto ext = [&](int j, int a) {
        const int c[3] = {cxv(j), cyv(j), czv(j)};
        const bool xonly = (BX == 8 || BX == 16) && c[0] != 0 && c[1] == 0 && c[2] == 0;
        if (a == 0 && xonly) return 1;
        return B[a] + (c[a] != 0 ? 1 : 0);
    };
    using P = PopLds<Q, BX, BY, BZ>;
    int off = 0;
    for (int j = 1; j < i; j++)
        if (P::phase(j) == P::phase(i)) off += ext(j, 0) * ext(j, 1) * ext(j, 2);
    const int ex = ext(i, 0) == 1 ? 0 : e[0];
    return PHASE_KEY * P::phase(i) + off + ex + ext(i, 0) * (e[1] + ext(i, 1) * e[2]);
}

template <int Q, int BX, int BY, int BZ>
void check_shape() {
    using G = HaloGeom<BX, BY, BZ>;
    constexpr int QF = VS<Q>::QF;
    const int B[3] = {BX, BY, BZ};
    // expected incoming: LDS cell -> (face, dx, dy, dz, slot, u, v)
    using Src = std::tuple<int, int, int, int, int, int, int>;
    std::map<int, Src> expect;
    for (int i = 1; i < Q; i++) {
        const int c[3] = {cxv(i), cyv(i), czv(i)};
        for (int z = 0; z < BZ; z++)
            for (int y = 0; y < BY; y++)
                for (int x = 0; x < BX; x++) {
                    const int n[3] = {x, y, z};
                    int src[3], A = 0;
                    for (int a = 0; a < 3; a++) {
                        src[a] = n[a] - c[a];
                        if (src[a] < 0 || src[a] >= B[a]) A |= 1 << a;
                    }
                    if (!A) continue;
                    const int win = (A & 4) ? 2 : ((A & 2) ? 1 : 0);
                    const int face = 2 * win + (c[win] > 0 ? 1 : 0);
                    const int ua = face_u_axis(win), va = face_v_axis(win);
                    int own[3], d[3] = {0, 0, 0};
                    for (int a = 0; a < 3; a++) {
                        own[a] = src[a];
                        if ((A >> a) & 1) {
                            d[a] = -c[a];
                            own[a] = (src[a] + B[a]) % B[a];
                        }
                    }
                    int e[3];
                    for (int a = 0; a < 3; a++) e[a] = n[a] + (c[a] < 0 ? 1 : 0);
                    const int cell = pop_cell<Q, BX, BY, BZ>(i, e);
                    CHECK(expect.find(cell) == expect.end(), "Q%d %dx%dx%d two sources for LDS cell %d", Q, BX, BY, BZ,
                          cell);
                    expect[cell] = Src(face, d[0], d[1], d[2], face_slot(face, i, QF), own[ua], own[va]);
                }
    }
    std::map<int, int> written;
    for (const Chunk& ch : ChunkTables<Q, BX, BY, BZ>::in) {
        const int Bu = face_bu<BX, BY, BZ>(ch.face), Bv = face_bv<BX, BY, BZ>(ch.face);
        const int nc = Bu * ch.rows;
        CHECK(64 % Bu == 0 && (nc & (nc - 1)) == 0 && nc <= 64, "bad window Bu=%d rows=%d", Bu, ch.rows);
        for (int lane = 0; lane < 64; lane++) {
            const int idx = nc >= 64 ? lane : (lane & (nc - 1));
            const int plane = idx + Bu * ch.w0;
            CHECK(plane >= 0 && plane < Bu * Bv, "load outside plane Q%d %dx%dx%d pop %d", Q, BX, BY, BZ, ch.pop);
            const int u = lane % Bu, r = lane / Bu, v = ch.w0 + r;
            const bool valid = (nc >= 64 || lane < nc) && u >= ch.ulo && u <= ch.uhi && v >= ch.vlo && v <= ch.vhi;
            if (!valid) continue;
            CHECK(plane == u + Bu * v, "lane/cell mismatch");
            const int cell = PHASE_KEY * PopLds<Q, BX, BY, BZ>::phase(ch.pop) + ch.lds0 + u * ch.su + r * ch.sv;
            written[cell]++;
            auto it = expect.find(cell);
            CHECK(it != expect.end(), "Q%d %dx%dx%d pop %d face %d writes unexpected cell %d", Q, BX, BY, BZ, ch.pop,
                  ch.face, cell);
            if (it == expect.end()) continue;
            const Src got(ch.face, ch.dx, ch.dy, ch.dz, ch.slot, u, v);
            CHECK(got == it->second, "Q%d %dx%dx%d pop %d cell %d: wrong source", Q, BX, BY, BZ, ch.pop, cell);
        }
    }
    for (auto& kv : expect) CHECK(written[kv.first] == 1, "cell %d written %d times", kv.first, written[kv.first]);

    // outgoing
    std::map<std::tuple<int, int, int>, int> out_written;  // (face, slot, plane index) -> count
    for (const Chunk& ch : ChunkTables<Q, BX, BY, BZ>::out) {
        const int Bu = face_bu<BX, BY, BZ>(ch.face), Bv = face_bv<BX, BY, BZ>(ch.face);
        const int nc = Bu * ch.rows;
        const int a = ch.face / 2, ua = face_u_axis(a), va = face_v_axis(a);
        const int i = ch.pop;
        const int c[3] = {cxv(i), cyv(i), czv(i)};
        CHECK(face_pop(ch.face, ch.slot) == i, "slot/pop mismatch");
        for (int lane = 0; lane < 64; lane++) {
            const int u = lane % Bu, r = lane / Bu, v = ch.w0 + r;
            const bool valid = (nc >= 64 || lane < nc) && u >= ch.ulo && u <= ch.uhi && v >= ch.vlo && v <= ch.vhi;
            if (!valid) continue;
            const int plane = lane + Bu * ch.w0;
            CHECK(plane >= 0 && plane < Bu * Bv && plane == u + Bu * v, "store outside plane");
            int node[3];
            node[a] = (ch.face & 1) ? B[a] - 1 : 0;
            node[ua] = u;
            node[va] = v;
            // staged-term layout: one padding column in x, rows of BY, planes of BZ
            const int nodebase = node[0] + (BX + 1) * (node[1] + BY * node[2]);
            CHECK(nodebase < (HaloGeom<BX, BY, BZ>::E), "staged base outside its term array");
            CHECK(ch.lds0 + u * ch.su + r * ch.sv == nodebase, "outgoing LDS cell mismatch Q%d pop %d", Q, i);
            (void)c;
            out_written[std::make_tuple(ch.face, ch.slot, plane)]++;
        }
    }
    // 8x8x8: a wave's staged-term reads of an x or y face hit 64 distinct
    // LDS banks (the odd row stride of the staged layout)
    if (BX == 8 && BY == 8 && BZ == 8) {
        for (const Chunk& ch : ChunkTables<Q, BX, BY, BZ>::out) {
            if (ch.face >= 4) continue;
            std::set<int> banks;
            const int Bu = face_bu<BX, BY, BZ>(ch.face);
            for (int lane = 0; lane < 64 && lane < Bu * ch.rows; lane++)
                banks.insert((ch.lds0 + (lane % Bu) * ch.su + (lane / Bu) * ch.sv) % 64);
            CHECK((int)banks.size() == (Bu * ch.rows < 64 ? Bu * ch.rows : 64),
                  "Q%d face %d chunk: %d distinct banks", Q, ch.face, (int)banks.size());
        }
    }
    for (int f = 0; f < 6; f++) {
        const int n = face_bu<BX, BY, BZ>(f) * face_bv<BX, BY, BZ>(f);
        for (int s = 0; s < QF; s++)
            for (int p = 0; p < n; p++)
                CHECK(out_written[std::make_tuple(f, s, p)] == 1, "face %d slot %d cell %d written %d times", f, s, p,
                      out_written[std::make_tuple(f, s, p)]);
    }
}

int main() {
#define SHAPE(A, B, C)          \
    check_shape<19, A, B, C>(); \
    check_shape<27, A, B, C>();
    SHAPE(8, 8, 8)
    SHAPE(16, 8, 4)
    SHAPE(32, 4, 4)
    SHAPE(64, 4, 2)
    SHAPE(16, 8, 8)
    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
