// Batches: many lattices advanced together, one launch per time step.
//
// A batch is a set of ordinary contexts of the same lattice type (q, tile
// shape, storage, device). Each launch of k_step_batch covers every tile of
// every member through a block -> (instance, tile) work table; per-instance
// arguments (pointers, extents, relaxation, ghost parity, step) live in a
// device table. Members adopt the batch's stream, so their own calls
// (read-back, reductions, re-initialisation) stay ordered with batch work.
// The reference steps one lattice per launch with a device sync per step
// (src/main.cu:223-230, src/main.cuh:353-363); small lattices are then
// launch-bound, which is what batching removes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "context.h"

using namespace mrlbm;

struct mrlbm_batch {
    int device, q, bx, by, bz, storage;
    hipStream_t stream = nullptr;
    std::vector<mrlbm_ctx*> members;
    std::vector<uint64_t> seen;  // member state versions the device table reflects
    InstArgs* d_tab = nullptr;
    int tab_cap = 0;
    uint32_t* d_work = nullptr;  // bulk-tile items, then boundary-tile items
    int work_cap = 0;
    int nwork = 0;
    int nbulk = 0;
    bool dirty = true;
    int grid_limit = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_steps = 0;
    bool graph_valid = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

void drop_graph(mrlbm_batch* b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
}

bool stale(mrlbm_batch* b) {
    if (b->dirty) return true;
    for (size_t i = 0; i < b->members.size(); i++)
        if (b->members[i]->version != b->seen[i]) return true;
    return false;
}

// Upload the instance and work tables from the members' host state.
int refresh(mrlbm_batch* b) {
    if (!stale(b)) return MRLBM_OK;
    const int n = (int)b->members.size();
    if (n > 4095) return fail(MRLBM_EINVAL, "a batch holds at most 4095 lattices");
    std::vector<InstArgs> tab(n);
    std::vector<uint32_t> work, bc_work;
    for (int i = 0; i < n; i++) {
        mrlbm_ctx* c = b->members[i];
        if (!c->initialised) return fail(MRLBM_ESTATE, "batch member %d is not initialised", i);
        if (c->ntiles > (1 << 20)) return fail(MRLBM_EINVAL, "lattice of %d tiles too large for a batch", c->ntiles);
        InstArgs ia{};
        ia.a = c->args();
        for (int s = 0; s < 2; s++)
            ia.g[s] = c->d_gbase[s];
        ia.cur0 = c->cur;
        ia.step0 = (uint32_t)c->step;
        ia.has_bc = c->has_bc ? 1 : 0;
        tab[i] = ia;
        const std::vector<uint8_t>* flags = nullptr;
        int rc = tile_flags(c, &flags);
        if (rc) return rc;
        for (int t = 0; t < c->ntiles; t++)
            ((*flags)[t] ? bc_work : work).push_back(((uint32_t)i << 20) | (uint32_t)t);
    }
    const int nbulk = (int)work.size();
    work.insert(work.end(), bc_work.begin(), bc_work.end());
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    if (n > b->tab_cap) {
        (void)hipFree(b->d_tab);
        b->d_tab = nullptr;
        b->tab_cap = std::max(n, 2 * b->tab_cap);
        HIPCHK(hipMalloc((void**)&b->d_tab, sizeof(InstArgs) * b->tab_cap));
    }
    if ((int)work.size() > b->work_cap) {
        (void)hipFree(b->d_work);
        b->d_work = nullptr;
        b->work_cap = std::max((int)work.size(), 2 * b->work_cap);
        HIPCHK(hipMalloc((void**)&b->d_work, sizeof(uint32_t) * b->work_cap));
    }
    if (n) HIPCHK(hipMemcpy(b->d_tab, tab.data(), sizeof(InstArgs) * n, hipMemcpyHostToDevice));
    if (!work.empty()) HIPCHK(hipMemcpy(b->d_work, work.data(), sizeof(uint32_t) * work.size(), hipMemcpyHostToDevice));
    b->nwork = (int)work.size();
    b->nbulk = nbulk;
    for (int i = 0; i < n; i++) b->seen[i] = b->members[i]->version;
    b->dirty = false;
    drop_graph(b);
    return MRLBM_OK;
}

int enqueue_steps(mrlbm_batch* b, int nsteps) {
    // one launch over the all-bulk tiles, one over the boundary tiles (both
    // in one launch where the two bodies run at the same occupancy and there
    // is no grid limit); they read the same ghost set and write disjoint data
    const int nb = b->nbulk, nc = b->nwork - b->nbulk;
    const int gb = b->grid_limit > 0 ? std::min(b->grid_limit, nb) : nb;
    const int gc = b->grid_limit > 0 ? std::min(b->grid_limit, nc) : nc;
    const bool mixed = b->grid_limit <= 0 && nb > 0 && nc > 0 &&
                       mixed_launch_supported(b->q, b->bx, b->by, b->bz, b->storage);
    for (int k = 0; k < nsteps; k++) {
        hipError_t e;
        if (mixed) {
            e = launch_step_batch_mixed(b->q, b->bx, b->by, b->bz, b->storage, b->d_tab, b->d_work, b->nwork, nb, k,
                                        b->stream);
        } else {
            e = launch_step_batch(b->q, b->bx, b->by, b->bz, b->storage, false, b->d_tab, b->d_work, nb, gb, k,
                                  b->stream);
            if (e == hipSuccess)
                e = launch_step_batch(b->q, b->bx, b->by, b->bz, b->storage, true, b->d_tab, b->d_work + nb, nc, gc, k,
                                      b->stream);
        }
        if (e != hipSuccess) return fail(MRLBM_EHIP, "batched step launch failed: %s", hipGetErrorString(e));
    }
    hipError_t e = launch_batch_advance(b->d_tab, (int)b->members.size(), nsteps, b->stream);
    if (e != hipSuccess) return fail(MRLBM_EHIP, "batch advance launch failed: %s", hipGetErrorString(e));
    return MRLBM_OK;
}

void advance_members(mrlbm_batch* b, int nsteps) {
    for (size_t i = 0; i < b->members.size(); i++) {
        mrlbm_ctx* c = b->members[i];
        c->step += nsteps;
        c->cur = (c->cur + nsteps) & 1;
    }
}

// Graph of `per` steps for the current membership (captured, not run: the
// members do not advance). Tables must be current (refresh).
int ensure_graph(mrlbm_batch* b, int per) {
    if (b->graph_valid && b->graph_steps == per) return MRLBM_OK;
    drop_graph(b);
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_steps(b, per);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : fail(MRLBM_EHIP, "graph capture failed: %s", hipGetErrorString(e));
    }
    b->graph = g;
    const hipError_t ie = hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        b->gexec = nullptr;
        drop_graph(b);
        return fail(MRLBM_EHIP, "graph instantiation failed: %s", hipGetErrorString(ie));
    }
    b->graph_steps = per;
    b->graph_valid = true;
    return MRLBM_OK;
}

int do_batch_steps(mrlbm_batch* b, int nsteps, int use_graph) {
    if (use_graph && nsteps % use_graph)
        return fail(MRLBM_EINVAL, "%d steps is not a multiple of the graph length %d", nsteps, use_graph);
    int rc = refresh(b);
    if (rc) return rc;
    if (b->members.empty() || nsteps == 0) return MRLBM_OK;
    if (!use_graph) {
        rc = enqueue_steps(b, nsteps);
        if (rc) return rc;
        advance_members(b, nsteps);
        return MRLBM_OK;
    }
    // graph of use_graph steps (captured once per membership), replayed
    const int per = use_graph;
    if ((rc = ensure_graph(b, per))) return rc;
    for (int r = 0; r < nsteps / per; r++) HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    advance_members(b, nsteps);
    return MRLBM_OK;
}

}  // namespace

hipStream_t mrlbm::batch_stream(const mrlbm_batch* b) { return b->stream; }

extern "C" {

int mrlbm_batch_create(int device, int q, int bx, int by, int bz, int storage, mrlbm_batch** out) {
    if (!out) return fail(MRLBM_EINVAL, "null argument");
    *out = nullptr;
    if (!tile_shape_supported(q, bx, by, bz, storage))
        return fail(MRLBM_EINVAL, "tile shape %dx%dx%d not built for q=%d storage=%d", bx, by, bz, q, storage);
    HIPCHK(hipSetDevice(device));
    mrlbm_batch* b = new mrlbm_batch();
    b->device = device;
    b->q = q;
    b->bx = bx;
    b->by = by;
    b->bz = bz;
    b->storage = storage;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        delete b;
        return fail(MRLBM_EHIP, "cannot create the batch stream/events");
    }
    *out = b;
    return MRLBM_OK;
}

int mrlbm_batch_add(mrlbm_batch* b, mrlbm_ctx* c) {
    if (!b || !c) return fail(MRLBM_EINVAL, "null argument");
    if (c->batch) return fail(MRLBM_ESTATE, "lattice already belongs to a batch");
    if (c->p.q != b->q || c->bx != b->bx || c->by != b->by || c->bz != b->bz || c->p.storage != b->storage ||
        c->p.device != b->device || c->p.kernel != MRLBM_KERNEL_TILE)
        return fail(MRLBM_EINVAL, "lattice type (q, tile, storage, device, tile kernel) differs from the batch's");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->own_stream = c->stream;
    c->stream = b->stream;
    c->batch = b;
    b->members.push_back(c);
    b->seen.push_back(~0ull);
    b->dirty = true;
    return MRLBM_OK;
}

int mrlbm_batch_remove(mrlbm_batch* b, mrlbm_ctx* c) {
    if (!b || !c) return fail(MRLBM_EINVAL, "null argument");
    auto it = std::find(b->members.begin(), b->members.end(), c);
    if (it == b->members.end()) return fail(MRLBM_EINVAL, "lattice is not in this batch");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    b->seen.erase(b->seen.begin() + (it - b->members.begin()));
    b->members.erase(it);
    c->stream = c->own_stream;
    c->own_stream = nullptr;
    c->batch = nullptr;
    b->dirty = true;
    return MRLBM_OK;
}

int mrlbm_batch_set_grid(mrlbm_batch* b, int max_blocks) {
    if (!b || max_blocks < 0) return fail(MRLBM_EINVAL, "bad argument");
    b->grid_limit = max_blocks;
    drop_graph(b);
    return MRLBM_OK;
}

int mrlbm_batch_step(mrlbm_batch* b, int nsteps, int graph_steps) {
    if (!b || nsteps < 0 || graph_steps < 0) return fail(MRLBM_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(b->device));
    return do_batch_steps(b, nsteps, graph_steps);
}

int mrlbm_batch_time_steps(mrlbm_batch* b, int nsteps, int graph_steps, double* ms) {
    if (!b || !ms || nsteps < 0) return fail(MRLBM_EINVAL, "bad argument");
    HIPCHK(hipSetDevice(b->device));
    if (graph_steps > 0 && nsteps % graph_steps)
        return fail(MRLBM_EINVAL, "%d steps is not a multiple of the graph length %d", nsteps, graph_steps);
    int rc = refresh(b);
    if (rc) return rc;
    // the graph is captured (not run) outside the timed region
    if (graph_steps > 0 && !b->members.empty() && nsteps > 0 && (rc = ensure_graph(b, graph_steps))) return rc;
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    rc = do_batch_steps(b, nsteps, graph_steps);
    if (rc) return rc;
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    HIPCHK(hipEventSynchronize(b->ev1));
    float f = 0;
    HIPCHK(hipEventElapsedTime(&f, b->ev0, b->ev1));
    *ms = f;
    return MRLBM_OK;
}

int mrlbm_batch_synchronize(mrlbm_batch* b) {
    if (!b) return fail(MRLBM_EINVAL, "null batch");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    return MRLBM_OK;
}

int mrlbm_batch_size(mrlbm_batch* b, int* members, int64_t* nodes) {
    if (!b || !members || !nodes) return fail(MRLBM_EINVAL, "null argument");
    *members = (int)b->members.size();
    int64_t n = 0;
    for (auto* c : b->members) n += (int64_t)c->N;
    *nodes = n;
    return MRLBM_OK;
}

int mrlbm_batch_destroy(mrlbm_batch* b) {
    if (!b) return MRLBM_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    for (auto* c : b->members) {
        c->stream = c->own_stream;
        c->own_stream = nullptr;
        c->batch = nullptr;
    }
    drop_graph(b);
    (void)hipFree(b->d_tab);
    (void)hipFree(b->d_work);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return MRLBM_OK;
}

// Autotuner: time `steps` graph-captured batched steps of lattices with the
// given extents (all-bulk, uniform state) for each candidate tile shape that
// divides every extent; report the fastest.
int mrlbm_batch_autotune(int device, int q, int storage, int nshapes, const int* shapes, int ninst, const int* dims,
                         int steps, int* best, double* ms_out) {
    return mrlbm_batch_autotune_walls(device, q, storage, nshapes, shapes, ninst, dims, nullptr, steps, best, ms_out);
}

int mrlbm_batch_autotune_walls(int device, int q, int storage, int nshapes, const int* shapes, int ninst,
                               const int* dims, const int* walls, int steps, int* best, double* ms_out) {
    if (!shapes || !dims || !best || nshapes <= 0 || ninst <= 0 || steps <= 0)
        return fail(MRLBM_EINVAL, "bad argument");
    *best = -1;
    double best_ms = 1e300;
    for (int s = 0; s < nshapes; s++) {
        const int bx = shapes[3 * s], by = shapes[3 * s + 1], bz = shapes[3 * s + 2];
        if (ms_out) ms_out[s] = -1.0;
        bool fits = tile_shape_supported(q, bx, by, bz, storage);
        for (int i = 0; i < ninst && fits; i++)
            if (dims[3 * i] % bx || dims[3 * i + 1] % by || dims[3 * i + 2] % bz) fits = false;
        if (!fits) continue;
        mrlbm_batch* b = nullptr;
        int rc = mrlbm_batch_create(device, q, bx, by, bz, storage, &b);
        if (rc) return rc;
        std::vector<mrlbm_ctx*> ctxs;
        for (int i = 0; i < ninst && !rc; i++) {
            mrlbm_params p{};
            p.nx = dims[3 * i];
            p.ny = dims[3 * i + 1];
            p.nz = dims[3 * i + 2];
            p.q = q;
            p.bx = bx;
            p.by = by;
            p.bz = bz;
            p.storage = storage;
            p.kernel = MRLBM_KERNEL_TILE;
            p.tau = 0.6;
            p.rho0 = 1.0;
            p.device = device;
            if (walls) {
                p.wall_x = walls[3 * i] != 0;
                p.wall_y = walls[3 * i + 1] != 0;
                p.wall_z = walls[3 * i + 2] != 0;
            }
            mrlbm_ctx* c = nullptr;
            rc = mrlbm_create(&p, &c);
            // wall axes: domain walls classified on the device, so the boundary
            // tiles run the boundary body as in the real lattices
            if (!rc && walls && (p.wall_x || p.wall_y || p.wall_z)) rc = mrlbm_set_geometry(c, nullptr, 0, nullptr);
            if (!rc) rc = mrlbm_init_uniform(c, 1.0, 0.0, 0.0, 0.0, MRLBM_INIT_EQUILIBRIUM);
            if (!rc) rc = mrlbm_batch_add(b, c);
            if (c) ctxs.push_back(c);
        }
        double ms = 0, again = 0;
        const int per = steps % 10 == 0 ? 10 : 1;
        if (!rc) rc = mrlbm_batch_time_steps(b, steps, per, &ms);  // warm-up + capture
        if (!rc) rc = mrlbm_batch_time_steps(b, steps, per, &ms);
        if (!rc) rc = mrlbm_batch_time_steps(b, steps, per, &again);  // best of two timed runs
        if (!rc && again < ms) ms = again;
        mrlbm_batch_destroy(b);
        for (auto* c : ctxs) mrlbm_destroy(c);
        if (rc) return rc;
        if (ms_out) ms_out[s] = ms;
        if (ms < best_ms) {
            best_ms = ms;
            *best = s;
        }
    }
    if (*best < 0) return fail(MRLBM_EINVAL, "no candidate tile shape divides every lattice");
    return MRLBM_OK;
}

}  // extern "C"
