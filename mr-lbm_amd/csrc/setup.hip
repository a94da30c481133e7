// Device-side set-up of a lattice: geometry classification and uniform
// initialisation without host passes over the nodes.
//
// Classification follows the same rules as the host path (geometry.cpp),
// which restates the reference's hostInitialization_nodeType, read_xyz_file
// and define_voxel_bc (src/lbmInitialization.cu:958-1227): a node next to a
// wall plane or a solid voxel gets the octant byte of the directions that
// point into the wall or the solid (a solid neighbour across a wall axis does
// not wrap), a solid voxel is SOLID_NODE. The class map is written in the
// tile layout with class id == octant byte, so one table of every no-slip
// octant pattern (built on the host from the relaxation rate alone) serves
// any geometry; the host only reads back which patterns occur.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "bc_classes.h"
#include "context.h"

using namespace mrlbm;

namespace {

constexpr int MAX_SPHERES_PER_LAUNCH = 16;

struct SphereBox {
    double cx, cy, cz, r;
    int x0, x1, y0, y1, z0, z1;  // clipped bounding box (inclusive)
};

struct SphereSet {
    int n;
    SphereBox s[MAX_SPHERES_PER_LAUNCH];
};

struct Extents {
    int nx, ny, nz;
    int wall[3];
    int bx, by, bz, ntx, nty;
};

// solid[j] = (init ? 0 : solid[j]) | inside any sphere of the set; the
// sphere test is the host's expression, evaluated in FP64 (the library is
// built without floating-point contraction), inside the same clipped box
__global__ void k_voxelize(Extents e, SphereSet set, int init, uint8_t* __restrict__ solid) {
    const size_t n = (size_t)e.nx * e.ny * e.nz;
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int x = (int)(j % e.nx), y = (int)((j / e.nx) % e.ny), z = (int)(j / ((size_t)e.nx * e.ny));
    uint8_t v = init ? 0 : solid[j];
    for (int s = 0; s < set.n; s++) {
        const SphereBox& b = set.s[s];
        if (x < b.x0 || x > b.x1 || y < b.y0 || y > b.y1 || z < b.z0 || z > b.z1) continue;
        const double dx = x - b.cx, dy = y - b.cy, dz = z - b.cz;
        if (dx * dx + dy * dy + dz * dz <= b.r * b.r) v = 1;
    }
    solid[j] = v;
}

// octants (bit k = corner (k&1, k>>1&1, k>>2&1), 1 = + side) containing direction d
__device__ __forceinline__ uint32_t octants_of(int dx, int dy, int dz) {
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int cx = (k & 1) ? 1 : -1, cy = (k & 2) ? 1 : -1, cz = (k & 4) ? 1 : -1;
        if ((dx == 0 || dx == cx) && (dy == 0 || dy == cy) && (dz == 0 || dz == cz)) m |= 1u << k;
    }
    return m;
}

// Class map (tile layout) and the 256-bit set of classes that occur.
__global__ void k_classify(Extents e, const uint8_t* __restrict__ solid, uint8_t* __restrict__ cls,
                           uint32_t* __restrict__ occ) {
    __shared__ uint32_t seen[8];
    if (threadIdx.x < 8) seen[threadIdx.x] = 0;
    __syncthreads();
    const size_t n = (size_t)e.nx * e.ny * e.nz;
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) {
        const int p[3] = {(int)(j % e.nx), (int)((j / e.nx) % e.ny), (int)(j / ((size_t)e.nx * e.ny))};
        const int ext[3] = {e.nx, e.ny, e.nz};
        uint32_t bits = 0;
        if (solid && solid[j]) {
            bits = 0xFFu;
        } else {
            for (int a = 0; a < 3; a++) {
                if (!e.wall[a]) continue;
                if (p[a] == 0) bits |= octants_of(a == 0 ? -1 : 0, a == 1 ? -1 : 0, a == 2 ? -1 : 0);
                if (p[a] == ext[a] - 1) bits |= octants_of(a == 0 ? 1 : 0, a == 1 ? 1 : 0, a == 2 ? 1 : 0);
            }
            if (solid) {
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            if (!dx && !dy && !dz) continue;
                            int q[3] = {p[0] + dx, p[1] + dy, p[2] + dz};
                            bool ok = true;
                            for (int a = 0; a < 3; a++)
                                if (q[a] < 0 || q[a] >= ext[a]) {
                                    if (e.wall[a]) ok = false;
                                    q[a] = (q[a] + ext[a]) % ext[a];
                                }
                            if (ok && solid[q[0] + (size_t)e.nx * (q[1] + (size_t)e.ny * q[2])])
                                bits |= octants_of(dx, dy, dz);
                        }
            }
        }
        const int tile = p[0] / e.bx + e.ntx * (p[1] / e.by + e.nty * (p[2] / e.bz));
        const int t = p[0] % e.bx + e.bx * (p[1] % e.by + e.by * (p[2] % e.bz));
        cls[(size_t)tile * (e.bx * e.by * e.bz) + t] = (uint8_t)bits;
        if (bits) atomicOr(&seen[bits >> 5], 1u << (bits & 31));
    }
    __syncthreads();
    if (threadIdx.x < 8 && seen[threadIdx.x]) atomicOr(&occ[threadIdx.x], seen[threadIdx.x]);
}

__global__ void k_zero_words(uint32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// every stored moment of every node set to the encoded value of its moment
// index: element j of the block-major array holds moment (j / T) % 10
template <typename W>
__global__ void k_fill_moments(W* __restrict__ mom, size_t n, int T, W v0, W v1, W v2, W v3, W v4, W v5, W v6, W v7,
                               W v8, W v9) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int m = (int)((j / (size_t)T) % 10);
    const W v[10] = {v0, v1, v2, v3, v4, v5, v6, v7, v8, v9};
    W out = v[0];
#pragma unroll
    for (int k = 1; k < 10; k++)
        if (m == k) out = v[k];
    mom[j] = out;
}

// flags[tile] = 1 if any node of the tile is not bulk (one block per tile)
__global__ void k_tile_flags(const uint8_t* __restrict__ cls, int T, uint8_t* __restrict__ flags) {
    const uint8_t* c = cls + (size_t)blockIdx.x * T;
    int any = 0;
    for (int t = threadIdx.x; t < T; t += blockDim.x) any |= c[t];
    any = __syncthreads_or(any);
    if (threadIdx.x == 0) flags[blockIdx.x] = any ? 1 : 0;
}

Extents extents_of(const mrlbm_ctx* c) {
    Extents e{};
    e.nx = c->p.nx;
    e.ny = c->p.ny;
    e.nz = c->p.nz;
    e.wall[0] = c->p.wall_x;
    e.wall[1] = c->p.wall_y;
    e.wall[2] = c->p.wall_z;
    e.bx = c->bx;
    e.by = c->by;
    e.bz = c->bz;
    e.ntx = c->ntx;
    e.nty = c->nty;
    return e;
}

unsigned blocks_for(size_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace

namespace mrlbm {

// Uniform state on the device: `enc` holds the ten stored moments already
// encoded in the lattice's storage format (4- or 2-byte words).
int fill_uniform_moments(mrlbm_ctx* c, const uint32_t enc[10]) {
    const size_t n = c->N * 10;
    const unsigned grid = blocks_for(n, 256);
    if (storage_bytes(c->p.storage) == 4) {
        k_fill_moments<uint32_t><<<grid, 256, 0, c->stream>>>((uint32_t*)c->d_mom, n, c->T, enc[0], enc[1], enc[2],
                                                               enc[3], enc[4], enc[5], enc[6], enc[7], enc[8], enc[9]);
    } else {
        uint16_t h[10];
        for (int k = 0; k < 10; k++) h[k] = (uint16_t)enc[k];
        k_fill_moments<uint16_t><<<grid, 256, 0, c->stream>>>((uint16_t*)c->d_mom, n, c->T, h[0], h[1], h[2], h[3],
                                                               h[4], h[5], h[6], h[7], h[8], h[9]);
    }
    HIPCHK(hipGetLastError());
    return MRLBM_OK;
}

int tile_flags(mrlbm_ctx* c, const std::vector<uint8_t>** out) {
    if (c->tile_flags_version != c->cls_version || c->tile_flags.size() != (size_t)c->ntiles) {
        c->tile_flags.assign(c->ntiles, 0);
        if (c->has_bc && c->d_cls) {
            HIPCHK(hipSetDevice(c->p.device));
            uint8_t* d = nullptr;
            HIPCHK(hipMallocAsync((void**)&d, c->ntiles, c->stream));
            k_tile_flags<<<c->ntiles, 256, 0, c->stream>>>(c->d_cls, c->T, d);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(c->tile_flags.data(), d, c->ntiles, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipFreeAsync(d, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        c->tile_flags_version = c->cls_version;
    }
    *out = &c->tile_flags;
    return MRLBM_OK;
}

int build_tile_list(mrlbm_ctx* c) {
    c->n_bulk_tiles = c->ntiles;
    if (!c->has_bc || !c->d_cls) return MRLBM_OK;
    const std::vector<uint8_t>* flags = nullptr;
    int rc = tile_flags(c, &flags);
    if (rc) return rc;
    std::vector<uint32_t> list;
    list.reserve(c->ntiles);
    for (int t = 0; t < c->ntiles; t++)
        if (!(*flags)[t]) list.push_back((uint32_t)t);
    c->n_bulk_tiles = (int)list.size();
    for (int t = 0; t < c->ntiles; t++)
        if ((*flags)[t]) list.push_back((uint32_t)t);
    if (!c->d_tiles) HIPCHK(hipMalloc((void**)&c->d_tiles, sizeof(uint32_t) * c->ntiles));
    HIPCHK(hipMemcpyAsync(c->d_tiles, list.data(), sizeof(uint32_t) * c->ntiles, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MRLBM_OK;
}

}  // namespace mrlbm

extern "C" {

int mrlbm_set_geometry(mrlbm_ctx* c, const uint8_t* solid, int nspheres, const double* spheres) {
    if (!c) return fail(MRLBM_EINVAL, "null context");
    if (nspheres < 0 || (nspheres > 0 && !spheres)) return fail(MRLBM_EINVAL, "bad sphere list");
    for (int s = 0; s < nspheres; s++)
        if (!(spheres[4 * s + 3] >= 0) || !std::isfinite(spheres[4 * s]) || !std::isfinite(spheres[4 * s + 1]) ||
            !std::isfinite(spheres[4 * s + 2]) || !std::isfinite(spheres[4 * s + 3]))
            return fail(MRLBM_EINVAL, "sphere %d is not finite with a non-negative radius", s);
    HIPCHK(hipSetDevice(c->p.device));
    // the class records of every octant pattern depend only on q and omega
    std::vector<float> table;
    uint32_t singular[8];
    build_octant_table(c->p.q, 1.0 / c->p.tau, table, singular);

    HIPCHK(hipStreamSynchronize(c->stream));
    if (!c->d_cls) HIPCHK(hipMalloc((void**)&c->d_cls, c->N));
    if (c->bc_bytes < table.size() * sizeof(float)) {  // a table from mrlbm_set_node_types may be smaller
        (void)hipFree(c->d_bc);
        c->d_bc = nullptr;
        c->bc_bytes = 0;
        HIPCHK(hipMalloc((void**)&c->d_bc, table.size() * sizeof(float)));
        c->bc_bytes = table.size() * sizeof(float);
    }
    HIPCHK(hipMemcpyAsync(c->d_bc, table.data(), table.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));

    const Extents e = extents_of(c);
    const bool any_solid = solid || nspheres > 0;
    uint8_t* d_solid = nullptr;
    if (any_solid) {
        HIPCHK(hipMallocAsync((void**)&d_solid, c->N, c->stream));
        if (solid) HIPCHK(hipMemcpyAsync(d_solid, solid, c->N, hipMemcpyHostToDevice, c->stream));
        const unsigned grid = blocks_for(c->N, 256);
        int first = solid ? 0 : 1;
        for (int s0 = 0; s0 < nspheres || first; s0 += MAX_SPHERES_PER_LAUNCH) {
            SphereSet set{};
            set.n = std::max(0, std::min(MAX_SPHERES_PER_LAUNCH, nspheres - s0));
            for (int k = 0; k < set.n; k++) {
                const double* sp = spheres + 4 * (s0 + k);
                SphereBox& b = set.s[k];
                b.cx = sp[0];
                b.cy = sp[1];
                b.cz = sp[2];
                b.r = sp[3];
                b.x0 = std::max(0, (int)std::floor(b.cx - b.r));
                b.x1 = std::min(e.nx - 1, (int)std::ceil(b.cx + b.r));
                b.y0 = std::max(0, (int)std::floor(b.cy - b.r));
                b.y1 = std::min(e.ny - 1, (int)std::ceil(b.cy + b.r));
                b.z0 = std::max(0, (int)std::floor(b.cz - b.r));
                b.z1 = std::min(e.nz - 1, (int)std::ceil(b.cz + b.r));
            }
            k_voxelize<<<grid, 256, 0, c->stream>>>(e, set, first, d_solid);
            HIPCHK(hipGetLastError());
            first = 0;
        }
    }
    uint32_t* d_occ = (uint32_t*)c->d_partial;  // reduction scratch, 2 * red_blocks doubles
    k_zero_words<<<1, 64, 0, c->stream>>>(d_occ, 8);
    k_classify<<<blocks_for(c->N, 256), 256, 0, c->stream>>>(e, d_solid, c->d_cls, d_occ);
    HIPCHK(hipGetLastError());
    uint32_t occ[8];
    HIPCHK(hipMemcpyAsync(occ, d_occ, sizeof occ, hipMemcpyDeviceToHost, c->stream));
    if (d_solid) HIPCHK(hipFreeAsync(d_solid, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));

    for (int w = 0; w < 8; w++)
        if (occ[w] & singular[w])
            for (int b = 0; b < 32; b++)
                if ((occ[w] & singular[w]) >> b & 1u)
                    return fail(MRLBM_EINVAL, "singular boundary system for node type %d", 32 * w + b);
    c->has_solid = (occ[7] >> 31) & 1u;
    int n_classes = 0;
    bool any = false;
    for (int w = 0; w < 8; w++) {
        const uint32_t bc = occ[w] & (w == 7 ? 0x7FFFFFFFu : 0xFFFFFFFFu);
        n_classes += __builtin_popcount(bc);
        any |= occ[w] != 0;
    }
    c->has_bc = any;
    c->n_classes = n_classes;
    c->version++;
    c->cls_version++;
    return build_tile_list(c);
}

}  // extern "C"
