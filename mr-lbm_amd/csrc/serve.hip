// Serving: a request loop over independent duct solves with continuous
// batching (no reference counterpart; the reference runs one compile-time
// case per process).
//
// One host thread runs the loop. Requests become visible at their arrival
// time (wall clock, or immediately in as-fast-as-possible mode); a visible
// request is admitted when the HBM budget and the batch-size limit allow:
// its lattice is allocated, classified (duct walls on y and z, voxelised
// spheres) and initialised at rest, and it joins the batch. All active
// lattices advance together, one launch per step, in rounds of at most
// `quantum` steps (hipGraph-captured when a full quantum fits); a round
// never passes a member's step budget, so members leave exactly at their
// budget. A finished request is reduced (FP64 mass), optionally read back,
// and its memory is released before the next admission check. Latency is
// arrival to results-on-host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "context.h"

using namespace mrlbm;

extern "C" int mrlbm_classify(int nx, int ny, int nz, int wall_x, int wall_y, int wall_z, const uint8_t* solid,
                              uint32_t* types);
extern "C" int mrlbm_voxelize_spheres(int nx, int ny, int nz, int nsph, const double* sph, uint8_t* solid);

namespace {

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
    int uniform_int(int a, int b) { return a + (int)(next() % (uint64_t)(b - a + 1)); }
};

double state_bytes(const mrlbm_request& r, int q, int storage, int bx, int by, int bz) {
    const double n = (double)r.nx * r.ny * r.nz;
    const double qf = q == 19 ? 5 : 9;
    const double ghosts = 2.0 * 2.0 * qf * 4.0 * n * (1.0 / bx + 1.0 / by + 1.0 / bz);
    return n * (10.0 * storage_bytes(storage) + 1.0) + ghosts;
}

struct Active {
    int idx;
    mrlbm_ctx* ctx;
    int remaining;
    double bytes;
};

}  // namespace

extern "C" {

// The seeded duct request stream: Poisson arrivals at `rate` per second;
// Nx in [64,256], Ny = Nz in [32,128] (multiples of 16); 1..8 spheres of
// radius U[3, Ny/6] at uniform centres; gx = 10^U[-7,-5]; nu = U[0.01, 0.1];
// steps = U{500..5000}.
int mrlbm_generate_duct_requests(uint64_t seed, int n, double rate, mrlbm_request* out) {
    if (!out || n < 0 || !(rate > 0)) return fail(MRLBM_EINVAL, "bad argument");
    SplitMix64 g{seed};
    double t = 0;
    for (int i = 0; i < n; i++) {
        mrlbm_request r{};
        t += -std::log(1.0 - g.uniform()) / rate;
        r.arrival_s = t;
        r.nx = 16 * g.uniform_int(4, 16);
        r.ny = r.nz = 16 * g.uniform_int(2, 8);
        r.nspheres = g.uniform_int(1, 8);
        for (int s = 0; s < r.nspheres; s++) {
            const double rad = 3.0 + g.uniform() * (r.ny / 6.0 - 3.0);
            r.sphere[s][0] = g.uniform() * r.nx;
            r.sphere[s][1] = g.uniform() * r.ny;
            r.sphere[s][2] = g.uniform() * r.nz;
            r.sphere[s][3] = rad;
        }
        r.gx = std::pow(10.0, -7.0 + 2.0 * g.uniform());
        r.nu = 0.01 + 0.09 * g.uniform();
        r.steps = g.uniform_int(500, 5000);
        r.seed = (uint32_t)g.next();
        out[i] = r;
    }
    return MRLBM_OK;
}

int mrlbm_serve_run(const mrlbm_serve_params* sp, mrlbm_request* reqs, int nreq, mrlbm_serve_stats* st) {
    if (!sp || (!reqs && nreq) || !st || nreq < 0) return fail(MRLBM_EINVAL, "null argument");
    const int bx = sp->bx ? sp->bx : 8, by = sp->by ? sp->by : 8, bz = sp->bz ? sp->bz : 8;
    const int quantum = sp->quantum_steps > 0 ? sp->quantum_steps : 100;
    const int graph = sp->graph_steps > 0 && quantum % sp->graph_steps == 0 ? sp->graph_steps : 0;
    const int max_batch = sp->max_batch > 0 ? std::min(sp->max_batch, 4095) : 4095;
    std::memset(st, 0, sizeof *st);
    for (int i = 0; i < nreq; i++) {
        if (reqs[i].nx % bx || reqs[i].ny % by || reqs[i].nz % bz)
            return fail(MRLBM_EINVAL, "request %d extents are not multiples of the tile", i);
        if (state_bytes(reqs[i], sp->q, sp->storage, bx, by, bz) > sp->hbm_budget_bytes)
            return fail(MRLBM_ENOMEM, "request %d alone exceeds the HBM budget", i);
        if (i && reqs[i].arrival_s < reqs[i - 1].arrival_s)
            return fail(MRLBM_EINVAL, "requests must be ordered by arrival time");
        reqs[i].status = 1;  // pending
    }
    mrlbm_batch* batch = nullptr;
    int rc = mrlbm_batch_create(sp->device, sp->q, bx, by, bz, sp->storage, &batch);
    if (rc) return rc;

    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    std::vector<Active> active;
    double used = 0, peak = 0, busy = 0, lu = 0;
    int next = 0, done = 0;
    std::vector<uint8_t> solid;
    std::vector<uint32_t> types;

    auto fail_all = [&](int code) {
        for (auto& a : active) mrlbm_destroy(a.ctx);
        mrlbm_batch_destroy(batch);
        return code;
    };

    while (done < nreq) {
        // admissions
        const double tnow = now();
        while (next < nreq && (!sp->realtime || reqs[next].arrival_s <= tnow) && (int)active.size() < max_batch) {
            mrlbm_request& r = reqs[next];
            const double bytes = state_bytes(r, sp->q, sp->storage, bx, by, bz);
            if (used + bytes > sp->hbm_budget_bytes) break;
            mrlbm_params p{};
            p.nx = r.nx;
            p.ny = r.ny;
            p.nz = r.nz;
            p.q = sp->q;
            p.bx = bx;
            p.by = by;
            p.bz = bz;
            p.storage = sp->storage;
            p.wall_y = p.wall_z = 1;
            p.kernel = MRLBM_KERNEL_TILE;
            p.tau = 0.5 + 3.0 * r.nu;
            p.rho0 = 1.0;
            p.fx = r.gx;
            p.device = sp->device;
            p.seed = r.seed;
            mrlbm_ctx* c = nullptr;
            if ((rc = mrlbm_create(&p, &c))) return fail_all(rc);
            const size_t n = (size_t)r.nx * r.ny * r.nz;
            solid.assign(n, 0);
            types.assign(n, 0);
            mrlbm_voxelize_spheres(r.nx, r.ny, r.nz, r.nspheres, &r.sphere[0][0], solid.data());
            mrlbm_classify(r.nx, r.ny, r.nz, 0, 1, 1, solid.data(), types.data());
            if ((rc = mrlbm_set_node_types(c, types.data())) ||
                (rc = mrlbm_init_uniform(c, 1.0, 0.0, 0.0, 0.0, MRLBM_INIT_EQUILIBRIUM)) ||
                (rc = mrlbm_reduce(c, &r.mass0, &r.ke0)) || (rc = mrlbm_batch_add(batch, c))) {
                mrlbm_destroy(c);
                return fail_all(rc);
            }
            r.start_s = now();
            r.status = 2;  // running
            active.push_back({next, c, r.steps, bytes});
            used += bytes;
            peak = std::max(peak, used);
            next++;
        }
        if (active.empty()) {
            if (next < nreq && sp->realtime) {
                const double wait = reqs[next].arrival_s - now();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
            }
            continue;
        }
        // one round for every active lattice
        int q = quantum;
        for (auto& a : active) q = std::min(q, a.remaining);
        const auto r0 = clk::now();
        rc = mrlbm_batch_step(batch, q, (graph && q == quantum) ? graph : 0);
        if (!rc) rc = mrlbm_batch_synchronize(batch);
        if (rc) return fail_all(rc);
        busy += std::chrono::duration<double>(clk::now() - r0).count();
        st->rounds++;
        // completions
        for (size_t i = 0; i < active.size();) {
            Active& a = active[i];
            a.remaining -= q;
            lu += (double)q * a.ctx->N;
            if (a.remaining > 0) {
                i++;
                continue;
            }
            mrlbm_request& r = reqs[a.idx];
            if ((rc = mrlbm_reduce(a.ctx, &r.mass1, &r.ke1))) return fail_all(rc);
            if (r.out_moments && (rc = mrlbm_read_moments(a.ctx, r.out_moments))) return fail_all(rc);
            r.finish_s = now();
            r.latency_s = r.finish_s - (sp->realtime ? r.arrival_s : 0.0);
            r.status = 0;
            mrlbm_batch_remove(batch, a.ctx);
            mrlbm_destroy(a.ctx);
            used -= a.bytes;
            done++;
            active.erase(active.begin() + i);
        }
    }
    mrlbm_batch_destroy(batch);
    const double wall = now();
    std::vector<double> lat;
    for (int i = 0; i < nreq; i++) lat.push_back(reqs[i].latency_s);
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
        if (lat.empty()) return 0.0;
        const double pos = p * (lat.size() - 1);
        const size_t i = (size_t)pos;
        const double f = pos - i;
        return i + 1 < lat.size() ? lat[i] * (1 - f) + lat[i + 1] * f : lat[i];
    };
    st->wall_s = wall;
    st->busy_s = busy;
    st->lattice_updates = lu;
    st->mlups = lu / wall / 1e6;
    st->mlups_busy = busy > 0 ? lu / busy / 1e6 : 0;
    st->p50_s = pct(0.50);
    st->p99_s = pct(0.99);
    st->max_s = lat.empty() ? 0 : lat.back();
    st->peak_bytes = peak;
    return MRLBM_OK;
}

}  // extern "C"
