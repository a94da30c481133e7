// C ABI of the library (include/mrlbm.h): context ownership, node map,
// initial state, step loop, read-back, reductions, checkpoint and VTK.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mrlbm.h"
#include "bc_classes.h"
#include "context.h"
#include "kernels.h"

using namespace mrlbm;

namespace mrlbm {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

size_t storage_bytes(int storage) { return storage == MRLBM_STORE_FP32 ? 4 : 2; }

// host-side storage conversion (round to nearest even; the kernels round
// stochastically)
uint16_t to_half(double v) { return __builtin_bit_cast(uint16_t, (_Float16)(float)v); }
double from_half(uint16_t b) { return (double)(float)__builtin_bit_cast(_Float16, b); }
uint16_t to_bf16(double v) {
    const float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
double from_bf16(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

}  // namespace mrlbm

extern "C" {

const char* mrlbm_last_error(void) { return g_err.c_str(); }

int mrlbm_destroy(mrlbm_ctx* c) {
    if (!c) return MRLBM_OK;
    if (c->batch) mrlbm_batch_remove(c->batch, c);
    hipSetDevice(c->p.device);
    if (c->stream) hipStreamSynchronize(c->stream);
    hipFree(c->d_mom);
    hipFree(c->d_cls);
    hipFree(c->d_bc);
    hipFree(c->d_tiles);
    hipFree(c->d_mean);
    for (int s = 0; s < 2; s++)
        hipFree(c->d_gbase[s]);
    hipFree(c->d_partial);
    hipFree(c->d_red);
    hipFree(c->d_uprev);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return MRLBM_OK;
}

int mrlbm_create(const mrlbm_params* pp, mrlbm_ctx** out) {
    if (!pp || !out) return fail(MRLBM_EINVAL, "null argument");
    *out = nullptr;
    const mrlbm_params& p = *pp;
    if (p.q != 19 && p.q != 27) return fail(MRLBM_EINVAL, "q must be 19 or 27 (got %d)", p.q);
    if (p.nx <= 0 || p.ny <= 0 || p.nz <= 0) return fail(MRLBM_EINVAL, "lattice extents must be positive");
    if (p.nx >= (1 << 24)) return fail(MRLBM_EINVAL, "nx must be below 2^24");  // 24-bit node-key multiply
    if (!(p.tau > 0.5)) return fail(MRLBM_EINVAL, "tau must exceed 0.5 (got %g)", p.tau);
    if (p.storage != MRLBM_STORE_FP32 && p.storage != MRLBM_STORE_FP16 && p.storage != MRLBM_STORE_BF16)
        return fail(MRLBM_EINVAL, "unknown storage precision %d", p.storage);
    if (p.kernel != MRLBM_KERNEL_AUTO && p.kernel != MRLBM_KERNEL_TILE && p.kernel != MRLBM_KERNEL_COLUMN)
        return fail(MRLBM_EINVAL, "unknown kernel family %d", p.kernel);
    int bx = p.bx, by = p.by, bz = p.bz;
    int kernel = p.kernel;
    if (bx == 0 && by == 0 && bz == 0) {
        // default: the fastest measured tile that divides the lattice, or for
        // the column kernel the longest segment that does. FP32 D3Q19 is
        // HBM-bound, and 16x8x8 tiles move 105 instead of 110 B per update
        // (256^3 TGV 54.9 -> 57.7 GLUPS, cavity 50.9 -> 51.3); the 16-bit
        // storages and D3Q27 are bound by instruction issue and latency,
        // which the 1024-thread tiles' two workgroups per CU hide worse
        // (FP16 TGV 74.9 -> 68.0 GLUPS), so they keep 8x8x8.
        const bool big = p.q == 19 && p.storage == MRLBM_STORE_FP32;
        const int cand[4][3] = {{16, 8, 8}, {8, 8, 8}, {16, 16, 32}, {16, 16, 64}};
        for (int c = big ? 0 : 1; c < 4; c++) {
            const bool col = cand[c][0] * cand[c][1] == 256;
            if (kernel != MRLBM_KERNEL_COLUMN && col) continue;
            if (kernel == MRLBM_KERNEL_COLUMN && !col) continue;
            if (p.nx % cand[c][0] == 0 && p.ny % cand[c][1] == 0 && p.nz % cand[c][2] == 0) {
                bx = cand[c][0];
                by = cand[c][1];
                bz = cand[c][2];
                break;
            }
        }
        if (bx == 0) return fail(MRLBM_EINVAL, "no built tile shape divides %dx%dx%d", p.nx, p.ny, p.nz);
    }
    if (kernel == MRLBM_KERNEL_AUTO) kernel = (bx * by == 256 && bz > 8) ? MRLBM_KERNEL_COLUMN : MRLBM_KERNEL_TILE;
    if (kernel == MRLBM_KERNEL_TILE && !tile_shape_supported(p.q, bx, by, bz, p.storage))
        return fail(MRLBM_EINVAL, "tile shape %dx%dx%d not built for q=%d storage=%d", bx, by, bz, p.q, p.storage);
    if (kernel == MRLBM_KERNEL_COLUMN && !column_shape_supported(p.q, bx, by, bz, p.storage))
        return fail(MRLBM_EINVAL, "column shape %dx%dx%d not built for q=%d storage=%d", bx, by, bz, p.q, p.storage);
    if (p.nx % bx || p.ny % by || p.nz % bz)
        return fail(MRLBM_EINVAL, "lattice %dx%dx%d is not a multiple of the tile %dx%dx%d", p.nx, p.ny, p.nz, bx, by,
                    bz);
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (p.device < 0 || p.device >= ndev) return fail(MRLBM_EINVAL, "device %d not present", p.device);
    HIPCHK(hipSetDevice(p.device));
    // set-up of the device's scratch array for mrlbm_reduce_checked_noalloc
    // (once per device: the symbol lookup is not an allocation)
    if (int rc0 = noalloc_scratch_setup(p.device, nullptr)) return rc0;

    mrlbm_ctx* c = new mrlbm_ctx();
    c->p = p;
    if (c->p.rho0 == 0.0) c->p.rho0 = 1.0;
    c->p.bx = c->bx = bx;
    c->p.by = c->by = by;
    c->p.bz = c->bz = bz;
    c->p.kernel = kernel;
    c->T = bx * by * bz;
    c->ntx = p.nx / bx;
    c->nty = p.ny / by;
    c->ntz = p.nz / bz;
    c->ntiles = c->ntx * c->nty * c->ntz;
    c->qf = p.q == 19 ? 5 : 9;
    c->N = (size_t)p.nx * p.ny * p.nz;
    // storage scales (powers of two), the same constants the step kernels use
    static_assert(MRLBM_STORE_FP16 == 1 && MRLBM_STORE_BF16 == 2, "storage codes");
    for (int m = 0; m < 10; m++) {
        const float s = store_scale(p.storage, m);
        c->mscale[m] = s;
        c->mrscale[m] = 1.0f / s;
    }

    auto alloc = [&](void** ptr, size_t bytes) -> int {
        hipError_t e = hipMalloc(ptr, bytes);
        if (e != hipSuccess) {
            *ptr = nullptr;
            return fail(MRLBM_ENOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        }
        c->device_bytes += (double)bytes;
        return MRLBM_OK;
    };
    int rc;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        mrlbm_destroy(c);
        return fail(MRLBM_EHIP, "hipStreamCreate failed");
    }
    if ((rc = alloc(&c->d_mom, c->N * 10 * storage_bytes(p.storage)))) {
        mrlbm_destroy(c);
        return rc;
    }
    c->face_elems[0] = c->face_elems[1] = (size_t)by * bz * c->qf * c->ntiles;
    c->face_elems[2] = c->face_elems[3] = (size_t)bx * bz * c->qf * c->ntiles;
    c->face_elems[4] = c->face_elems[5] = (size_t)bx * by * c->qf * c->ntiles;
    for (int s = 0; s < 2; s++) {
        size_t set = 0;
        for (int f = 0; f < 6; f++) set += c->face_elems[f];
        if ((rc = alloc((void**)&c->d_gbase[s], set * sizeof(float)))) {
            mrlbm_destroy(c);
            return rc;
        }
        size_t o = 0;
        for (int f = 0; f < 6; f++) {
            c->d_ghost[s][f] = c->d_gbase[s] + o;
            o += c->face_elems[f];
        }
    }
    c->red_blocks = 1024;
    if ((rc = alloc((void**)&c->d_partial, sizeof(double) * 3 * c->red_blocks)) ||
        (rc = alloc((void**)&c->d_red, sizeof(double) * 4))) {
        mrlbm_destroy(c);
        return rc;
    }
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        mrlbm_destroy(c);
        return fail(MRLBM_EHIP, "hipEventCreate failed");
    }
    *out = c;
    return MRLBM_OK;
}

int mrlbm_set_bc_velocity(mrlbm_ctx* c, int k, double ux, double uy, double uz) {
    if (!c) return fail(MRLBM_EINVAL, "null context");
    if (k < 0 || k > 3) return fail(MRLBM_EINVAL, "wall-velocity index %d outside 0..3", k);
    c->bc_vel[k][0] = ux;
    c->bc_vel[k][1] = uy;
    c->bc_vel[k][2] = uz;
    return MRLBM_OK;
}

int mrlbm_set_node_types(mrlbm_ctx* c, const uint32_t* types) {
    if (!c) return fail(MRLBM_EINVAL, "null context");
    HIPCHK(hipSetDevice(c->p.device));
    BcClassTable tab;
    std::string err = build_class_table(c->p.q, 1.0 / c->p.tau, c->bc_vel, types, c->N, tab);
    if (!err.empty()) return fail(MRLBM_EINVAL, "%s", err.c_str());
    HIPCHK(hipStreamSynchronize(c->stream));
    hipFree(c->d_cls);
    hipFree(c->d_bc);
    c->d_cls = nullptr;
    c->d_bc = nullptr;
    c->bc_bytes = 0;
    c->has_bc = tab.has_bc;
    c->has_solid = tab.has_solid;
    c->n_classes = (int)tab.keys.size() - 1;
    c->version++;
    c->cls_version++;
    if (!c->has_bc) return build_tile_list(c);
    std::vector<uint8_t> blk(c->N);
    const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) blk[c->node_index(x, y, z)] = tab.class_of[x + (size_t)nx * (y + (size_t)ny * z)];
    HIPCHK(hipMalloc((void**)&c->d_cls, c->N));
    HIPCHK(hipMalloc((void**)&c->d_bc, tab.table.size() * sizeof(float)));
    c->bc_bytes = tab.table.size() * sizeof(float);
    HIPCHK(hipMemcpy(c->d_cls, blk.data(), c->N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_bc, tab.table.data(), tab.table.size() * sizeof(float), hipMemcpyHostToDevice));
    return build_tile_list(c);
}

}  // extern "C"

namespace {

// Stored (scaled) moments of one node from unscaled macroscopic fields.
void node_moments(int q, int mode, double rho0, double rho, double ux, double uy, double uz, double m[10]) {
    double pxx, pxy, pxz, pyy, pyz, pzz;
    if (mode == MRLBM_INIT_EQUILIBRIUM) {
        pxx = ux * ux;
        pxy = ux * uy;
        pxz = ux * uz;
        pyy = uy * uy;
        pyz = uy * uz;
        pzz = uz * uz;
    } else {
        // f_eq(w_i RHO_0, u), summed over the D3Q19 index subsets and divided
        // by rho, as src/lbmInitialization.cu:102-130 does for either lattice.
        static const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0};
        static const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1};
        static const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1};
        double f[19];
        const double uu = ux * ux + uy * uy + uz * uz;
        for (int i = 0; i < 19; i++) {
            const int cls = i == 0 ? 0 : (i < 7 ? 1 : 2);
            const double w = q == 19 ? (cls == 0 ? 1.0 / 3 : (cls == 1 ? 1.0 / 18 : 1.0 / 36))
                                     : (cls == 0 ? 8.0 / 27 : (cls == 1 ? 2.0 / 27 : 1.0 / 54));
            const double uc3 = 3.0 * (ux * CX[i] + uy * CY[i] + uz * CZ[i]);
            f[i] = (w * rho0) * ((1.0 - 1.5 * uu) + uc3 * (1.0 + uc3 * 0.5));
        }
        const double ir = 1.0 / rho, cs2 = 1.0 / 3.0;
        pxx = (f[1] + f[2] + f[7] + f[8] + f[9] + f[10] + f[13] + f[14] + f[15] + f[16]) * ir - cs2;
        pxy = ((f[7] + f[8]) - (f[13] + f[14])) * ir;
        pxz = ((f[9] + f[10]) - (f[15] + f[16])) * ir;
        pyy = (f[3] + f[4] + f[7] + f[8] + f[11] + f[12] + f[13] + f[14] + f[17] + f[18]) * ir - cs2;
        pyz = ((f[11] + f[12]) - (f[17] + f[18])) * ir;
        pzz = (f[5] + f[6] + f[9] + f[10] + f[11] + f[12] + f[15] + f[16] + f[17] + f[18]) * ir - cs2;
    }
    m[0] = rho - rho0;
    m[1] = 3.0 * ux;
    m[2] = 3.0 * uy;
    m[3] = 3.0 * uz;
    m[4] = 4.5 * pxx;
    m[5] = 9.0 * pxy;
    m[6] = 9.0 * pxz;
    m[7] = 4.5 * pyy;
    m[8] = 9.0 * pyz;
    m[9] = 4.5 * pzz;
}

// One stored moment (index m, stored convention) -> its storage word
uint32_t encode_value(const mrlbm_ctx* c, int m, double x) {
    const double v = x * c->mscale[m];
    if (c->p.storage == MRLBM_STORE_FP32) {
        const float f = (float)v;
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return u;
    }
    return c->p.storage == MRLBM_STORE_FP16 ? to_half(v) : to_bf16(v);
}

// Moments (block layout, doubles, stored convention) -> storage bytes
std::vector<char> encode_block(const mrlbm_ctx* c, const std::vector<double>& blk) {
    const size_t sb = storage_bytes(c->p.storage);
    std::vector<char> raw(blk.size() * sb);
    const size_t T = c->T;
    for (size_t j = 0; j < blk.size(); j++) {
        const uint32_t u = encode_value(c, (int)((j / T) % 10), blk[j]);
        if (sb == 4) {
            std::memcpy(&raw[j * 4], &u, 4);
        } else {
            const uint16_t h = (uint16_t)u;
            std::memcpy(&raw[j * 2], &h, 2);
        }
    }
    return raw;
}

std::vector<double> decode_block(const mrlbm_ctx* c, const std::vector<char>& raw) {
    const size_t sb = storage_bytes(c->p.storage);
    std::vector<double> blk(raw.size() / sb);
    const size_t T = c->T;
    for (size_t j = 0; j < blk.size(); j++) {
        const int m = (int)((j / T) % 10);
        double v;
        if (c->p.storage == MRLBM_STORE_FP32) {
            float f;
            std::memcpy(&f, &raw[j * 4], 4);
            v = f;
        } else {
            uint16_t h;
            std::memcpy(&h, &raw[j * 2], 2);
            v = c->p.storage == MRLBM_STORE_FP16 ? from_half(h) : from_bf16(h);
        }
        blk[j] = v * (double)c->mrscale[m];
    }
    return blk;
}

// Both ghost sets from the stored moments; the lattice restarts at step 0.
int ghosts_from_moments(mrlbm_ctx* c) {
    StepArgs a = c->args();
    for (int s = 0; s < 2; s++) {
        if (c->p.kernel == MRLBM_KERNEL_COLUMN)
            HIPCHK(launch_init_ghosts_column(c->p.q, c->bx, c->by, c->bz, c->p.storage, a, c->d_ghost[s], c->ntiles,
                                             c->stream));
        else
            HIPCHK(launch_init_ghosts(c->p.q, c->bx, c->by, c->bz, c->p.storage, a, c->d_ghost[s], c->ntiles,
                                      c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->cur = 0;
    c->step = 0;
    c->initialised = true;
    c->version++;
    return MRLBM_OK;
}

int upload_and_ghosts(mrlbm_ctx* c, const std::vector<double>& blk) {
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(hipStreamSynchronize(c->stream));
    const std::vector<char> raw = encode_block(c, blk);
    HIPCHK(hipMemcpy(c->d_mom, raw.data(), raw.size(), hipMemcpyHostToDevice));
    return ghosts_from_moments(c);
}

template <typename F>
int init_with(mrlbm_ctx* c, int mode, F&& field) {
    if (mode != MRLBM_INIT_EQUILIBRIUM && mode != MRLBM_INIT_REFERENCE)
        return fail(MRLBM_EINVAL, "unknown second-moment mode %d", mode);
    std::vector<double> blk(c->N * 10);
    const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                double r, ux, uy, uz, m[10];
                field(x, y, z, r, ux, uy, uz);
                if (!(r > 0)) return fail(MRLBM_EINVAL, "non-positive density at (%d,%d,%d)", x, y, z);
                node_moments(c->p.q, mode, c->p.rho0, r, ux, uy, uz, m);
                for (int k = 0; k < 10; k++) blk[c->mom_index(x, y, z, k)] = m[k];
            }
    return upload_and_ghosts(c, blk);
}

}  // namespace

extern "C" {

int mrlbm_init_fields(mrlbm_ctx* c, const double* rho, const double* ux, const double* uy, const double* uz,
                      int mode) {
    if (!c || !rho || !ux || !uy || !uz) return fail(MRLBM_EINVAL, "null argument");
    const size_t nx = c->p.nx, ny = c->p.ny;
    return init_with(c, mode, [&](int x, int y, int z, double& r, double& a, double& b, double& d) {
        const size_t j = x + nx * (y + ny * z);
        r = rho[j];
        a = ux[j];
        b = uy[j];
        d = uz[j];
    });
}

int mrlbm_init_uniform(mrlbm_ctx* c, double rho, double ux, double uy, double uz, int mode) {
    if (!c) return fail(MRLBM_EINVAL, "null context");
    if (mode != MRLBM_INIT_EQUILIBRIUM && mode != MRLBM_INIT_REFERENCE)
        return fail(MRLBM_EINVAL, "unknown second-moment mode %d", mode);
    if (!(rho > 0)) return fail(MRLBM_EINVAL, "non-positive density");
    // one node's stored moments, encoded once and filled on the device
    double m[10];
    node_moments(c->p.q, mode, c->p.rho0, rho, ux, uy, uz, m);
    uint32_t enc[10];
    for (int k = 0; k < 10; k++) enc[k] = encode_value(c, k, m[k]);
    HIPCHK(hipSetDevice(c->p.device));
    int rc = fill_uniform_moments(c, enc);
    if (rc) return rc;
    return ghosts_from_moments(c);
}

int mrlbm_init_taylor_green(mrlbm_ctx* c, double u0, int mode) {
    if (!c) return fail(MRLBM_EINVAL, "null context");
    // cases/taylorGreen/flow_initialization.inc:1-4 with L = N/(2 pi)
    const double kx = 2.0 * M_PI / c->p.nx, ky = 2.0 * M_PI / c->p.ny, kz = 2.0 * M_PI / c->p.nz;
    const double rho0 = c->p.rho0;
    return init_with(c, mode, [&](int x, int y, int z, double& r, double& a, double& b, double& d) {
        r = rho0 + (3.0 / 16.0) * rho0 * u0 * u0 * (std::cos(2 * kx * x) + std::cos(2 * ky * y)) *
                       (std::cos(2 * kz * z) + 2.0);
        a = u0 * std::sin(kx * x) * std::cos(ky * y) * std::cos(kz * z);
        b = -u0 * std::cos(kx * x) * std::sin(ky * y) * std::cos(kz * z);
        d = 0.0;
    });
}

static int do_steps(mrlbm_ctx* c, int nsteps) {
    const bool split = c->p.kernel == MRLBM_KERNEL_TILE && c->has_bc && c->d_tiles;
    const bool mixed = split && mixed_launch_supported(c->p.q, c->bx, c->by, c->bz, c->p.storage);
    for (int s = 0; s < nsteps; s++) {
        StepArgs a = c->args();
        hipError_t e;
        if (c->p.kernel == MRLBM_KERNEL_COLUMN) {
            e = launch_step_column(c->p.q, c->bx, c->by, c->bz, c->has_bc, c->p.storage, a, c->ntiles, c->stream);
        } else if (split && mixed) {
            // one launch: the all-bulk tiles with the light body, the rest
            // with the boundary body
            e = launch_step_tile_mixed(c->p.q, c->bx, c->by, c->bz, c->p.storage, a, c->d_tiles, c->ntiles,
                                       c->n_bulk_tiles, c->stream);
        } else if (split) {
            // all-bulk tiles with the light body, the rest with boundaries;
            // both read the same ghost set and write disjoint data
            e = launch_step_tile(c->p.q, c->bx, c->by, c->bz, false, c->p.storage, a, c->d_tiles, c->n_bulk_tiles,
                                 c->stream);
            if (e == hipSuccess)
                e = launch_step_tile(c->p.q, c->bx, c->by, c->bz, true, c->p.storage, a, c->d_tiles + c->n_bulk_tiles,
                                     c->ntiles - c->n_bulk_tiles, c->stream);
        } else {
            e = launch_step_tile(c->p.q, c->bx, c->by, c->bz, c->has_bc, c->p.storage, a, nullptr, c->ntiles,
                                 c->stream);
        }
        if (e != hipSuccess) return fail(MRLBM_EHIP, "step launch failed: %s", hipGetErrorString(e));
        c->cur = 1 - c->cur;
        c->step++;
    }
    return MRLBM_OK;
}

int mrlbm_step(mrlbm_ctx* c, int nsteps) {
    if (!c) return fail(MRLBM_EINVAL, "null context");
    if (c->batch) return fail(MRLBM_ESTATE, "lattice is stepped by its batch (mrlbm_batch_step)");
    if (!c->initialised) return fail(MRLBM_ESTATE, "lattice not initialised");
    if (nsteps < 0) return fail(MRLBM_EINVAL, "negative step count");
    HIPCHK(hipSetDevice(c->p.device));
    return do_steps(c, nsteps);
}

int mrlbm_time_steps(mrlbm_ctx* c, int nsteps, double* ms) {
    if (!c || !ms) return fail(MRLBM_EINVAL, "null argument");
    if (c->batch) return fail(MRLBM_ESTATE, "lattice is stepped by its batch (mrlbm_batch_time_steps)");
    if (!c->initialised) return fail(MRLBM_ESTATE, "lattice not initialised");
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    int rc = do_steps(c, nsteps);
    if (rc) return rc;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipEventSynchronize(c->ev1));
    float f = 0;
    HIPCHK(hipEventElapsedTime(&f, c->ev0, c->ev1));
    *ms = f;
    return MRLBM_OK;
}

int mrlbm_synchronize(mrlbm_ctx* c) {
    if (!c) return fail(MRLBM_EINVAL, "null context");
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MRLBM_OK;
}

int mrlbm_read_moments(mrlbm_ctx* c, double* out) {
    if (!c || !out) return fail(MRLBM_EINVAL, "null argument");
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<char> raw(c->N * 10 * storage_bytes(c->p.storage));
    HIPCHK(hipMemcpy(raw.data(), c->d_mom, raw.size(), hipMemcpyDeviceToHost));
    const std::vector<double> blk = decode_block(c, raw);
    const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    for (int k = 0; k < 10; k++)
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++)
                    out[(size_t)k * c->N + x + (size_t)nx * (y + (size_t)ny * z)] = blk[c->mom_index(x, y, z, k)];
    return MRLBM_OK;
}

int mrlbm_write_moments(mrlbm_ctx* c, const double* in) {
    if (!c || !in) return fail(MRLBM_EINVAL, "null argument");
    std::vector<double> blk(c->N * 10);
    const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    for (int k = 0; k < 10; k++)
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++)
                    blk[c->mom_index(x, y, z, k)] = in[(size_t)k * c->N + x + (size_t)nx * (y + (size_t)ny * z)];
    return upload_and_ghosts(c, blk);
}

int mrlbm_reduce_checked(mrlbm_ctx* c, double* mass, double* ke, int64_t* nonfinite) {
    if (!c || !mass || !ke || !nonfinite) return fail(MRLBM_EINVAL, "null argument");
    HIPCHK(hipSetDevice(c->p.device));
    HIPCHK(launch_reduce(c->p.storage, c->d_mom, c->ntiles, c->T, (float)c->p.rho0, c->mrscale[0], c->mrscale[1],
                         c->d_partial, c->red_blocks, c->d_red, c->stream, (double)c->N));
    double h[3];
    HIPCHK(hipMemcpyAsync(h, c->d_red, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *mass = h[0];
    *ke = h[1];
    *nonfinite = (int64_t)h[2];
    return MRLBM_OK;
}

int mrlbm_reduce(mrlbm_ctx* c, double* mass, double* ke) {
    int64_t bad;
    return mrlbm_reduce_checked(c, mass, ke, &bad);
}

int mrlbm_velocity_residual(mrlbm_ctx* c, double* residual) {
    if (!c || !residual) return fail(MRLBM_EINVAL, "null argument");
    if (!c->initialised) return fail(MRLBM_ESTATE, "lattice not initialised");
    HIPCHK(hipSetDevice(c->p.device));
    const bool first = c->d_uprev == nullptr;
    if (first) {
        const size_t bytes = sizeof(float) * 3 * c->N;
        if (hipMalloc((void**)&c->d_uprev, bytes) != hipSuccess) {
            c->d_uprev = nullptr;
            return fail(MRLBM_ENOMEM, "hipMalloc(%zu bytes) for the residual reference failed", bytes);
        }
        c->device_bytes += (double)bytes;
        HIPCHK(hipMemsetAsync(c->d_uprev, 0, bytes, c->stream));
    }
    HIPCHK(launch_residual(c->p.storage, c->d_mom, c->d_uprev, c->ntiles, c->T, c->mrscale[1], c->d_partial,
                           c->red_blocks, c->d_red, c->stream));
    double h[3];
    HIPCHK(hipMemcpyAsync(h, c->d_red, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *residual = first ? -1.0 : (h[1] > 0.0 ? std::sqrt(h[0] / h[1]) : (h[0] > 0.0 ? INFINITY : 0.0));
    return MRLBM_OK;
}

int mrlbm_get_info(mrlbm_ctx* c, mrlbm_info* o) {
    if (!c || !o) return fail(MRLBM_EINVAL, "null argument");
    o->nx = c->p.nx;
    o->ny = c->p.ny;
    o->nz = c->p.nz;
    o->q = c->p.q;
    o->bx = c->bx;
    o->by = c->by;
    o->bz = c->bz;
    o->storage = c->p.storage;
    o->kernel = c->p.kernel;
    o->step = c->step;
    o->n_bc_classes = c->n_classes;
    o->has_solid = c->has_solid;
    o->bytes_per_lu = c->bytes_per_lu();
    o->device_bytes = c->device_bytes;
    return MRLBM_OK;
}

int mrlbm_bc_class_info(int q, uint32_t node_type, double omega, const double u[3], uint32_t* known,
                        double* ainv49) {
    if (!u || !known || !ainv49) return fail(MRLBM_EINVAL, "null argument");
    if (q != 19 && q != 27) return fail(MRLBM_EINVAL, "q must be 19 or 27");
    *known = known_mask(q, node_type & 0xFFu);
    if (!build_bc_inverse(q, *known, omega, u, ainv49, nullptr, closure_lattice(q, node_type & 0xFFu)))
        return fail(MRLBM_ENUMERIC, "singular boundary system");
    return MRLBM_OK;
}

}  // extern "C"
