// Serving: a request loop over independent duct solves with continuous
// batching (no reference counterpart; the reference runs one compile-time
// case per process).
//
// One host thread runs the loop. Requests become visible at their arrival
// time (wall clock, or immediately in as-fast-as-possible mode); a visible
// request is admitted when the HBM budget and the batch-size limit allow:
// its lattice is allocated, classified on the device (duct walls on y and
// z, voxelised spheres) and initialised at rest on its own stream while the
// current round runs (or, if it arrived during the round, right after it),
// and it joins the batch at the next round boundary.
// All active lattices advance together, one launch per step, in rounds of
// at most `quantum` steps (hipGraph-captured when a full quantum fits); a
// round never passes a member's step budget, so members leave exactly at
// their budget. A finished request leaves the batch at the round boundary;
// its FP64 mass (and optionally its moments) are read and its memory is
// released during the next round. Latency is arrival to results-on-host.
// Failures are per request: a request that cannot be served (extents not a
// multiple of the tile, more than 8 spheres, no steps, larger than the whole
// budget, an admission that fails) or whose lattice turns non-finite (an
// FP64 reduction of every member after every round counts non-finite
// nodes) is marked failed with its status code and its memory released,
// while the others continue. Only device-level errors (a launch, a round,
// batch membership) end the loop for all.
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
    mrlbm_ctx* ctx;  // null on the simulated device
    int remaining;
    double bytes;
    double nodes;
};

// Request that cannot be served at all (checked before the loop runs).
int request_error(const mrlbm_request& r, const mrlbm_serve_params* sp, int bx, int by, int bz) {
    if (r.nx <= 0 || r.ny <= 0 || r.nz <= 0 || r.nx % bx || r.ny % by || r.nz % bz) return MRLBM_EINVAL;
    if (r.nspheres < 0 || r.nspheres > 8) return MRLBM_EINVAL;
    if (r.steps <= 0 || !(r.nu > 0.0) || !std::isfinite(r.nu)) return MRLBM_EINVAL;
    if (state_bytes(r, sp->q, sp->storage, bx, by, bz) > sp->hbm_budget_bytes) return MRLBM_ENOMEM;
    return MRLBM_OK;
}

// The serving loop, shared by the GPU and a simulated device. Dev supplies:
//   double now();                           seconds since the loop started
//   int admit(mrlbm_request&, Active&);     allocate, classify, initialise
//   int add(Active&), remove(Active&);      batch membership
//   int launch(int q, bool graph, const std::vector<Active>&)
//                                           start a round of q steps, then a
//                                           health check of every member
//   int wait_round(double* busy_s, std::vector<char>& bad)
//                                           end of the round in flight; bad[i]
//                                           = member i turned non-finite
//   int finish(Active&, mrlbm_request&);    read results (mass, moments)
//   void release(Active&);                  free its memory
//   void idle(double seconds);              nothing to run until then
template <typename Dev>
int serve_loop(Dev& dev, const mrlbm_serve_params* sp, mrlbm_request* reqs, int nreq, mrlbm_serve_stats* st) {
    const int bx = sp->bx ? sp->bx : 8, by = sp->by ? sp->by : 8, bz = sp->bz ? sp->bz : 8;
    const int quantum = sp->quantum_steps > 0 ? sp->quantum_steps : 100;
    const int graph = sp->graph_steps > 0 && quantum % sp->graph_steps == 0 ? sp->graph_steps : 0;
    const int max_batch = sp->max_batch > 0 ? std::min(sp->max_batch, 4095) : 4095;
    std::memset(st, 0, sizeof *st);
    for (int i = 0; i < nreq; i++)
        if (i && reqs[i].arrival_s < reqs[i - 1].arrival_s)
            return fail(MRLBM_EINVAL, "requests must be ordered by arrival time");
    for (int i = 0; i < nreq; i++) {
        mrlbm_request& r = reqs[i];
        r.start_s = r.finish_s = r.latency_s = 0.0;
        r.status = 1;  // pending
    }
    // ready: admitted and initialised, joins at the next round boundary;
    // active: batch members; leaving: out of the batch, results being read
    std::vector<Active> ready, active, leaving;
    double used = 0, peak = 0, busy = 0, lu = 0;
    int next = 0, done = 0, round_q = 0, rc = 0;
    int64_t failed = 0;
    bool in_flight = false;
    std::vector<char> bad;
    // a request leaves the loop failed with status `code` (negative)
    auto fail_request = [&](mrlbm_request& r, int code) {
        r.status = code;
        r.finish_s = dev.now();
        r.latency_s = 0.0;
        done++;
        failed++;
    };

    auto release_all = [&]() {
        for (auto* v : {&ready, &active, &leaving})
            for (auto& a : *v) dev.release(a);
    };
    auto fail_all = [&](int code) {
        release_all();
        return code;
    };
    // Admit visible requests (arrival order) while the budget and the batch
    // limit allow; runs while a round is in flight.
    auto admit = [&]() -> int {
        const double tnow = dev.now();
        while (next < nreq && (!sp->realtime || reqs[next].arrival_s <= tnow) &&
               (int)(active.size() + ready.size()) < max_batch) {
            mrlbm_request& r = reqs[next];
            const int bad_req = request_error(r, sp, bx, by, bz);
            if (bad_req) {
                next++;
                fail_request(r, bad_req);
                continue;
            }
            const double bytes = state_bytes(r, sp->q, sp->storage, bx, by, bz);
            if (used + bytes > sp->hbm_budget_bytes) break;
            Active a{next, nullptr, r.steps, bytes, (double)r.nx * r.ny * r.nz};
            const int e = dev.admit(r, a);
            if (e) {
                next++;
                fail_request(r, e);
                continue;
            }
            r.start_s = dev.now();
            r.status = 2;  // running
            ready.push_back(a);
            used += bytes;
            peak = std::max(peak, used);
            next++;
        }
        return MRLBM_OK;
    };
    auto finish = [&]() -> int {
        for (auto& a : leaving) {
            mrlbm_request& r = reqs[a.idx];
            const int e = dev.finish(a, r);
            if (e) {
                fail_request(r, e);
                continue;
            }
            r.finish_s = dev.now();
            r.latency_s = r.finish_s - (sp->realtime ? r.arrival_s : 0.0);
            r.status = 0;
            done++;
        }
        return MRLBM_OK;
    };

    while (done < nreq) {
        if ((rc = admit())) return fail_all(rc);
        if ((rc = finish())) return fail_all(rc);
        // memory of finished requests is released while the round runs (the
        // frees cost ~1 ms of host time per lattice, which at the round
        // boundary would leave the device idle)
        for (auto& a : leaving) {
            dev.release(a);
            used -= a.bytes;
        }
        leaving.clear();
        if (in_flight) {
            double b = 0;
            if ((rc = dev.wait_round(&b, bad))) return fail_all(rc);
            busy += b;
            st->rounds++;
            in_flight = false;
            for (auto& a : active) {
                a.remaining -= round_q;
                lu += (double)round_q * a.nodes;
            }
            // members that turned non-finite leave failed; the others go on
            for (size_t i = active.size(); i-- > 0;) {
                if (!bad[i]) continue;
                if ((rc = dev.remove(active[i]))) return fail_all(rc);
                dev.release(active[i]);
                used -= active[i].bytes;
                fail_request(reqs[active[i].idx], MRLBM_ENUMERIC);
                active.erase(active.begin() + i);
            }
            // requests that arrived during the round join the next one
            // (the first pass ran as the round started)
            if ((rc = admit())) return fail_all(rc);
        }
        for (size_t i = 0; i < active.size();) {
            if (active[i].remaining > 0) {
                i++;
                continue;
            }
            if ((rc = dev.remove(active[i]))) return fail_all(rc);
            leaving.push_back(active[i]);
            active.erase(active.begin() + i);
        }
        for (auto& a : ready) {
            if ((rc = dev.add(a))) return fail_all(rc);
            active.push_back(a);
        }
        ready.clear();
        if (active.empty()) {
            if (!leaving.empty()) continue;  // read their results now
            if ((rc = admit())) return fail_all(rc);
            if (ready.empty() && next < nreq && sp->realtime) {
                const double wait = reqs[next].arrival_s - dev.now();
                if (wait > 0) dev.idle(wait);
            }
            continue;
        }
        // next round: at most `quantum` steps, never past a member's budget
        int q = quantum;
        for (auto& a : active) q = std::min(q, a.remaining);
        if ((rc = dev.launch(q, graph && q == quantum, active))) return fail_all(rc);
        round_q = q;
        in_flight = true;
    }
    release_all();
    const double wall = dev.now();
    std::vector<double> lat;
    for (int i = 0; i < nreq; i++)
        if (reqs[i].status == 0) lat.push_back(reqs[i].latency_s);
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
    st->mlups = wall > 0 ? lu / wall / 1e6 : 0;
    st->mlups_busy = busy > 0 ? lu / busy / 1e6 : 0;
    st->p50_s = pct(0.50);
    st->p99_s = pct(0.99);
    st->max_s = lat.empty() ? 0 : lat.back();
    st->peak_bytes = peak;
    st->failed = failed;
    return MRLBM_OK;
}

// The GPU: one context per request in one batch, wall clock.
struct GpuDevice {
    const mrlbm_serve_params* sp;
    int bx, by, bz;
    mrlbm_batch* batch = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    static constexpr int kMaxMembers = 4096;
    double* d_health = nullptr;  // 3 doubles per member: FP64 mass, KE, non-finite nodes
    double* h_health = nullptr;  // pinned copy
    int n_checked = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
    int admit(mrlbm_request& r, Active& a) {
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
        a.ctx = c;
        return MRLBM_OK;
    }
    int add(Active& a) { return mrlbm_batch_add(batch, a.ctx); }
    int remove(Active& a) { return mrlbm_batch_remove(batch, a.ctx); }
    int launch(int q, bool use_graph, const std::vector<Active>& active) {
        const hipStream_t s = batch_stream(batch);
        if (hipEventRecord(ev0, s) != hipSuccess) return fail(MRLBM_EHIP, "event record");
        const int e = mrlbm_batch_step(batch, q, use_graph ? sp->graph_steps : 0);
        if (e) return e;
        // health of every member after the round, on the batch stream
        n_checked = (int)active.size();
        for (int i = 0; i < n_checked; i++) {
            const mrlbm_ctx* c = active[i].ctx;
            const hipError_t he = launch_reduce(c->p.storage, c->d_mom, c->ntiles, c->T, (float)c->p.rho0,
                                                c->mrscale[0], c->mrscale[1], c->d_partial, c->red_blocks,
                                                d_health + 3 * i, s);
            if (he != hipSuccess) return fail(MRLBM_EHIP, "health reduction: %s", hipGetErrorString(he));
        }
        if (n_checked &&
            hipMemcpyAsync(h_health, d_health, sizeof(double) * 3 * n_checked, hipMemcpyDeviceToHost, s) != hipSuccess)
            return fail(MRLBM_EHIP, "health read-back");
        if (hipEventRecord(ev1, s) != hipSuccess) return fail(MRLBM_EHIP, "event record");
        return MRLBM_OK;
    }
    int wait_round(double* busy_s, std::vector<char>& bad) {
        if (hipEventSynchronize(ev1) != hipSuccess) return fail(MRLBM_EHIP, "round failed");
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ev0, ev1);
        *busy_s = ms * 1e-3;
        bad.assign(n_checked, 0);
        for (int i = 0; i < n_checked; i++)
            bad[i] = h_health[3 * i + 2] != 0.0 || !std::isfinite(h_health[3 * i]) || !std::isfinite(h_health[3 * i + 1]);
        return MRLBM_OK;
    }
    int finish(Active& a, mrlbm_request& r) {
        int64_t nonfinite = 0;
        int e = mrlbm_reduce_checked(a.ctx, &r.mass1, &r.ke1, &nonfinite);
        if (!e && nonfinite) e = fail(MRLBM_ENUMERIC, "request %d: %lld non-finite nodes", a.idx, (long long)nonfinite);
        if (!e && r.out_moments) e = mrlbm_read_moments(a.ctx, r.out_moments);
        return e;
    }
    void release(Active& a) {
        mrlbm_destroy(a.ctx);
        a.ctx = nullptr;
    }
    void idle(double seconds) { std::this_thread::sleep_for(std::chrono::duration<double>(std::min(seconds, 0.01))); }
};

// A simulated device for testing the loop without a GPU: a virtual clock;
// a round of q steps over nodes N takes N q / rate seconds, admission
// admit_s seconds of host time; results are not computed.
struct SimDevice {
    double rate_lups, admit_s;
    const mrlbm_request* reqs;
    double clock = 0, round_end = 0, round_len = 0;
    double batch_nodes = 0;
    std::vector<char> round_bad;
    double now() const { return clock; }
    int admit(mrlbm_request& r, Active&) {
        clock += admit_s;
        r.mass0 = r.mass1 = r.ke0 = r.ke1 = 0;
        return MRLBM_OK;
    }
    int add(Active& a) {
        batch_nodes += a.nodes;
        return MRLBM_OK;
    }
    int remove(Active& a) {
        batch_nodes -= a.nodes;
        return MRLBM_OK;
    }
    // a lattice driven by a non-finite force is non-finite after its first
    // round, as on the GPU
    int launch(int q, bool, const std::vector<Active>& active) {
        round_len = batch_nodes * q / rate_lups;
        round_end = clock + round_len;
        round_bad.assign(active.size(), 0);
        for (size_t i = 0; i < active.size(); i++) round_bad[i] = !std::isfinite(reqs[active[i].idx].gx);
        return MRLBM_OK;
    }
    int wait_round(double* busy_s, std::vector<char>& bad) {
        clock = std::max(clock, round_end);
        *busy_s = round_len;
        bad = round_bad;
        return MRLBM_OK;
    }
    int finish(Active&, mrlbm_request&) { return MRLBM_OK; }
    void release(Active&) {}
    void idle(double seconds) { clock += seconds; }
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
    GpuDevice dev;
    dev.sp = sp;
    dev.bx = sp->bx ? sp->bx : 8;
    dev.by = sp->by ? sp->by : 8;
    dev.bz = sp->bz ? sp->bz : 8;
    int rc = mrlbm_batch_create(sp->device, sp->q, dev.bx, dev.by, dev.bz, sp->storage, &dev.batch);
    if (rc) return rc;
    auto cleanup = [&]() {
        mrlbm_batch_destroy(dev.batch);
        if (dev.ev0) (void)hipEventDestroy(dev.ev0);
        if (dev.ev1) (void)hipEventDestroy(dev.ev1);
        if (dev.d_health) (void)hipFree(dev.d_health);
        if (dev.h_health) (void)hipHostFree(dev.h_health);
    };
    if (hipEventCreate(&dev.ev0) != hipSuccess || hipEventCreate(&dev.ev1) != hipSuccess ||
        hipMalloc((void**)&dev.d_health, sizeof(double) * 3 * GpuDevice::kMaxMembers) != hipSuccess ||
        hipHostMalloc((void**)&dev.h_health, sizeof(double) * 3 * GpuDevice::kMaxMembers) != hipSuccess) {
        cleanup();
        return fail(MRLBM_EHIP, "cannot create the serving loop's events and health buffers");
    }
    dev.t0 = std::chrono::steady_clock::now();
    rc = serve_loop(dev, sp, reqs, nreq, st);
    cleanup();
    return rc;
}

int mrlbm_serve_simulate(const mrlbm_serve_params* sp, mrlbm_request* reqs, int nreq, double rate_mlups,
                         double admit_s, mrlbm_serve_stats* st) {
    if (!sp || (!reqs && nreq) || !st || nreq < 0 || !(rate_mlups > 0) || !(admit_s >= 0))
        return fail(MRLBM_EINVAL, "bad argument");
    SimDevice dev{rate_mlups * 1e6, admit_s, reqs};
    return serve_loop(dev, sp, reqs, nreq, st);
}

}  // extern "C"
