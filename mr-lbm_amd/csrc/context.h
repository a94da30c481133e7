// Host-side state of one lattice (a context) and the helpers shared by the
// C-ABI implementation files.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mrlbm.h"
#include "kernels.h"

struct mrlbm_batch;

namespace mrlbm {

int fail(int code, const char* fmt, ...);
size_t storage_bytes(int storage);

}  // namespace mrlbm

#define HIPCHK(expr)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) return mrlbm::fail(MRLBM_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

struct mrlbm_ctx {
    mrlbm_params p;
    int bx, by, bz, T, ntx, nty, ntz, ntiles, qf;
    size_t N;
    hipStream_t stream = nullptr;
    void* d_mom = nullptr;
    uint8_t* d_cls = nullptr;
    float* d_bc = nullptr;
    size_t bc_bytes = 0;  // capacity of d_bc
    float* d_gbase[2] = {};       // ghost sets, six faces one after another (ghost_face)
    float* d_ghost[2][6] = {};    // the faces inside them
    size_t face_elems[6] = {};
    int cur = 0;  // ghost set read by the next step
    int64_t step = 0;
    bool has_bc = false, has_solid = false, initialised = false;
    int n_classes = 0;
    double bc_vel[4][3] = {};
    double* d_partial = nullptr;
    double* d_red = nullptr;
    int red_blocks = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double device_bytes = 0;
    mrlbm_batch* batch = nullptr;  // batch this lattice is stepped in (its stream is then the batch's)
    uint64_t version = 0;          // bumped when buffers or state are replaced outside a step
    hipStream_t own_stream = nullptr;
    float mscale[10], mrscale[10];  // storage scales (powers of two)
    uint64_t cls_version = 0;        // bumped when the class map is replaced
    std::vector<uint8_t> tile_flags;  // per tile: 1 = holds a boundary or solid node
    uint64_t tile_flags_version = ~0ull;
    uint32_t* d_tiles = nullptr;      // all-bulk tiles, then the others (tile kernel with boundaries)
    int n_bulk_tiles = 0;
    double* d_mean = nullptr;         // running sums of rho, ux, uy, uz (4 x N, x-fastest), allocated on first use
    float* d_uprev = nullptr;         // velocity of the last residual call (3 x N, block layout), on first use
    int64_t mean_count = 0;

    size_t mom_index(int x, int y, int z, int m) const {
        const int tile = (x / bx) + ntx * ((y / by) + nty * (z / bz));
        const int t = (x % bx) + bx * ((y % by) + by * (z % bz));
        return (size_t)tile * 10 * T + (size_t)m * T + t;
    }
    size_t node_index(int x, int y, int z) const {
        const int tile = (x / bx) + ntx * ((y / by) + nty * (z / bz));
        const int t = (x % bx) + bx * ((y % by) + by * (z % bz));
        return (size_t)tile * T + t;
    }
    mrlbm::StepArgs args() const {
        mrlbm::StepArgs a{};
        a.mom = d_mom;
        a.cls = d_cls;
        a.bc = d_bc;
        a.gin = d_gbase[cur];
        a.gout = d_gbase[1 - cur];
        a.ntiles = ntiles;
        a.ntx = ntx;
        a.nty = nty;
        a.by_ntx = mrlbm::make_fastdiv((uint32_t)ntx);
        a.by_nty = mrlbm::make_fastdiv((uint32_t)nty);
        a.ntz = ntz;
        a.nx = p.nx;
        a.ny = p.ny;
        a.nz = p.nz;
        a.wall_x = p.wall_x;
        a.wall_y = p.wall_y;
        a.wall_z = p.wall_z;
        a.rho0 = (float)p.rho0;
        const double omega = 1.0 / p.tau;
        // relaxation constants as src/mlbm.cu:56-60 derives them
        const float om = (float)omega;
        // storage scales of U and M folded into the constants (powers of two,
        // 1 for FP32 and BF16: the products are exact rescalings)
        const float su = mscale[1], sm = mscale[4];
        const float uu = sm / (su * su);
        a.R.omega = om * uu;
        a.R.t_omega = 1.0f - om;
        a.R.tt_omega = (1.0f - om / 2.0f) * uu;
        a.R.omega_d2 = (om / 2.0f) * uu;
        a.R.tt_omega_t3 = ((1.0f - om / 2.0f) * 3.0f) * uu;
        a.R.fx = (float)p.fx * su;
        a.R.fy = (float)p.fy * su;
        a.R.fz = (float)p.fz * su;
        a.R.has_force = (p.fx != 0.0 || p.fy != 0.0 || p.fz != 0.0) ? 1 : 0;
        for (int m = 0; m < 10; m++) {
            a.mscale[m] = mscale[m];
            a.mrscale[m] = mrscale[m];
        }
        a.step = (uint32_t)step;
        a.seed = p.seed;
        return a;
    }
    double bytes_per_lu() const {
        const double sm = (double)mrlbm::storage_bytes(p.storage);
        const double sg = 4.0;
        double b = 2.0 * 10.0 * sm + 4.0 * qf * sg * (1.0 / bx + 1.0 / by + 1.0 / bz);
        if (has_bc) b += 1.0;
        return b;
    }
};

namespace mrlbm {

// Every node's ten stored moments set on the device to `enc` (storage words).
int fill_uniform_moments(mrlbm_ctx* c, const uint32_t enc[10]);

// Per-tile boundary flags of the current class map (computed on the device
// once per class map).
int tile_flags(mrlbm_ctx* c, const std::vector<uint8_t>** out);

// Device list of the lattice's tiles, all-bulk ones first (n_bulk_tiles),
// for the split bulk/boundary launches of a lattice with boundary nodes.
int build_tile_list(mrlbm_ctx* c);

// The stream a batch launches on (members' work is ordered on it).
hipStream_t batch_stream(const mrlbm_batch* b);

// Device address of the allocation-free reduction's scratch array on
// `device` (which must be current), looked up on the first call for that
// device and kept (reduce.hip).
int noalloc_scratch_setup(int device, double** out);

}  // namespace mrlbm
