// Serving: a request loop over independent duct solves with continuous
// batching (no reference counterpart; the reference runs one compile-time
// case per process).
//
// One host thread runs the loop. Requests become visible at their arrival
// time (wall clock, or immediately in as-fast-as-possible mode); a visible
// request is admitted when the HBM budget and the batch-size limit allow:
// its lattice is allocated, classified on the device (duct walls on y and
// z, voxelised spheres) and initialised at rest on its own stream while the
// current round runs, and it joins the batch at the next round boundary.
// All active lattices advance together, one launch per step, in rounds of
// at most `quantum` steps (hipGraph-captured when a full quantum fits); a
// round never passes a member's step budget, so members leave exactly at
// their budget. A finished request leaves the batch at the round boundary;
// its FP64 mass (and optionally its moments) are read during the next round
// and its memory is released at the boundary after that. Latency is arrival
// to results-on-host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "context.h"

using namespace mrlbm;

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
        if (reqs[i].nspheres < 0 || reqs[i].nspheres > 8)
            return fail(MRLBM_EINVAL, "request %d has %d spheres (0..8)", i, reqs[i].nspheres);
        if (state_bytes(reqs[i], sp->q, sp->storage, bx, by, bz) > sp->hbm_budget_bytes)
            return fail(MRLBM_ENOMEM, "request %d alone exceeds the HBM budget", i);
        if (i && reqs[i].arrival_s < reqs[i - 1].arrival_s)
            return fail(MRLBM_EINVAL, "requests must be ordered by arrival time");
        reqs[i].status = 1;  // pending
    }
    mrlbm_batch* batch = nullptr;
    int rc = mrlbm_batch_create(sp->device, sp->q, bx, by, bz, sp->storage, &batch);
    if (rc) return rc;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (hipEventCreate(&ev0) != hipSuccess || hipEventCreate(&ev1) != hipSuccess) {
        mrlbm_batch_destroy(batch);
        return fail(MRLBM_EHIP, "cannot create events");
    }

    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    // ready: admitted and initialised, joins at the next round boundary;
    // active: batch members; leaving: out of the batch, results being read
    std::vector<Active> ready, active, leaving;
    double used = 0, peak = 0, busy = 0, lu = 0;
    int next = 0, done = 0, round_q = 0;
    bool in_flight = false;

    auto release = [&]() {
        for (auto* v : {&ready, &active, &leaving})
            for (auto& a : *v) mrlbm_destroy(a.ctx);
        mrlbm_batch_destroy(batch);
        (void)hipEventDestroy(ev0);
        (void)hipEventDestroy(ev1);
    };
    auto fail_all = [&](int code) {
        release();
        return code;
    };
    // Admit visible requests while the budget and the batch limit allow;
    // runs while a round is in flight (its work goes to each lattice's own
    // stream), so set-up overlaps stepping.
    auto admit = [&]() -> int {
        const double tnow = now();
        while (next < nreq && (!sp->realtime || reqs[next].arrival_s <= tnow) &&
               (int)(active.size() + ready.size()) < max_batch) {
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
            int e = mrlbm_create(&p, &c);
            if (!e) e = mrlbm_set_geometry(c, nullptr, r.nspheres, &r.sphere[0][0]);
            if (!e) e = mrlbm_init_uniform(c, 1.0, 0.0, 0.0, 0.0, MRLBM_INIT_EQUILIBRIUM);
            if (!e) e = mrlbm_reduce(c, &r.mass0, &r.ke0);
            if (e) {
                mrlbm_destroy(c);
                return e;
            }
            r.start_s = now();
            r.status = 2;  // running
            ready.push_back({next, c, r.steps, bytes});
            used += bytes;
            peak = std::max(peak, used);
            next++;
        }
        return MRLBM_OK;
    };
    // Results of lattices that left the batch: FP64 mass, optional moments.
    auto finish = [&]() -> int {
        for (auto& a : leaving) {
            mrlbm_request& r = reqs[a.idx];
            int e = mrlbm_reduce(a.ctx, &r.mass1, &r.ke1);
            if (!e && r.out_moments) e = mrlbm_read_moments(a.ctx, r.out_moments);
            if (e) return e;
            r.finish_s = now();
            r.latency_s = r.finish_s - (sp->realtime ? r.arrival_s : 0.0);
            r.status = 0;
            done++;
        }
        return MRLBM_OK;
    };

    while (done < nreq) {
        if ((rc = admit())) return fail_all(rc);
        if ((rc = finish())) return fail_all(rc);
        if (in_flight) {
            // end of the round in flight
            if (hipEventSynchronize(ev1) != hipSuccess) return fail_all(fail(MRLBM_EHIP, "round failed"));
            float ms = 0;
            (void)hipEventElapsedTime(&ms, ev0, ev1);
            busy += ms * 1e-3;
            st->rounds++;
            in_flight = false;
            for (auto& a : active) {
                a.remaining -= round_q;
                lu += (double)round_q * a.ctx->N;
            }
        }
        // memory of finished requests is released while the GPU is idle
        for (auto& a : leaving) {
            mrlbm_destroy(a.ctx);
            used -= a.bytes;
        }
        leaving.clear();
        for (size_t i = 0; i < active.size();) {
            if (active[i].remaining > 0) {
                i++;
                continue;
            }
            if ((rc = mrlbm_batch_remove(batch, active[i].ctx))) return fail_all(rc);
            leaving.push_back(active[i]);
            active.erase(active.begin() + i);
        }
        for (auto& a : ready) {
            if ((rc = mrlbm_batch_add(batch, a.ctx))) return fail_all(rc);
            active.push_back(a);
        }
        ready.clear();
        if (active.empty()) {
            if (!leaving.empty()) continue;  // read their results now
            if ((rc = admit())) return fail_all(rc);
            if (ready.empty() && next < nreq && sp->realtime) {
                const double wait = reqs[next].arrival_s - now();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
            }
            continue;
        }
        // next round: at most `quantum` steps, never past a member's budget
        int q = quantum;
        for (auto& a : active) q = std::min(q, a.remaining);
        if (hipEventRecord(ev0, batch_stream(batch)) != hipSuccess) return fail_all(fail(MRLBM_EHIP, "event record"));
        if ((rc = mrlbm_batch_step(batch, q, (graph && q == quantum) ? graph : 0))) return fail_all(rc);
        if (hipEventRecord(ev1, batch_stream(batch)) != hipSuccess) return fail_all(fail(MRLBM_EHIP, "event record"));
        round_q = q;
        in_flight = true;
    }
    release();
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
