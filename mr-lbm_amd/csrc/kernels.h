// Device-side argument blocks and launcher declarations shared by the
// step kernels and the host library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fastdiv.h"
#include "lattice.h"

namespace mrlbm {

// Storage codecs. Arithmetic is always FP32; a codec loads a stored value
// and decodes it to float. Loads are non-temporal: moments are streamed
// once per step. Reduced-precision formats hold v * scale (a power of two,
// so the scaling is exact), rounded stochastically when the step kernels
// store them (quantize in tile_common.h: the hardware stochastic
// conversions with counter-based random operands, see sr_bits below), so
// per-step changes smaller than half an ulp are not lost on average.
// Storage scale of moment k (0 drho, 1-3 U, 4-9 M), a power of two:
// binary16 keeps the expected magnitudes (|drho|, |M| up to ~0.1, |U| up to
// ~0.3 for Mach < 0.2) near 2^8..2^9, far from both its overflow (65504) and
// its subnormal range (< 2^-14); bfloat16 has FP32's exponent range and is
// stored unscaled. The step kernels use these constants; the host uses the
// same function for encoding and read-back.
__host__ __device__ constexpr float store_scale(int storage, int k) {
    return storage == 1 ? ((k >= 1 && k <= 3) ? 1024.0f : 4096.0f) : 1.0f;
}

// binary16 value (low or high half of `b`) times a power-of-two scale in one
// mixed-precision FMA: v_fma_mix_f32 converts the half exactly and the
// addend -0 (the inline constant 0 with the negate modifier) makes the
// result x * s bit for bit, signed zeros included (the same value as a
// conversion followed by a multiply, one instruction less). The scale is
// uniform (an SGPR operand).
template <bool HI>
__device__ __forceinline__ float f16_scaled(uint32_t b, float s) {
    float r;
    if constexpr (HI)
        asm("v_fma_mix_f32 %0, %1, %2, neg(0) op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(b), "s"(s));
    else
        asm("v_fma_mix_f32 %0, %1, %2, neg(0) op_sel_hi:[1,0,0]" : "=v"(r) : "v"(b), "s"(s));
    return r;
}

template <int S> struct Codec;
template <> struct Codec<0> {
    using T = float;
    template <typename P>
    __device__ __forceinline__ static float ld(P p, size_t i, float) { return __builtin_nontemporal_load(&p[i]); }
    __device__ __forceinline__ static float dec(T b, float) { return b; }
};
template <> struct Codec<1> {  // IEEE binary16, 13 dropped bits rounded stochastically
    using T = uint16_t;
    __device__ __forceinline__ static float dec(uint32_t b, float rscale) { return f16_scaled<false>(b, rscale); }
    template <typename P>
    __device__ __forceinline__ static float ld(P p, size_t i, float rscale) { return dec(__builtin_nontemporal_load(&p[i]), rscale); }
};
// bfloat16, 16 dropped bits rounded stochastically. FP32's exponent range
// needs no storage scale: the library's bfloat16 scales are 1, and the
// codec ignores them (no multiplies by 1 in the step kernels).
template <> struct Codec<2> {
    using T = uint16_t;
    __device__ __forceinline__ static float dec(uint32_t b, float) { return __uint_as_float(b << 16); }
    template <typename P>
    __device__ __forceinline__ static float ld(P p, size_t i, float rscale) { return dec(__builtin_nontemporal_load(&p[i]), rscale); }
};

// Register type of a fetched stored moment: the 16-bit formats are held in
// 32-bit registers (bfloat16 zero-extended by the load; binary16 with an
// undefined high half, which the reconstruction does not read).
template <int S> struct Fetched { using T = uint32_t; };
template <> struct Fetched<0> { using T = float; };

// Counter-based random bits for stochastic rounding: a 32-bit mix of the
// node's global index, the time step and the lattice seed. The same node,
// step and seed give the same bits in any launch shape or batch, so results
// do not depend on how lattices are batched.
__device__ __forceinline__ uint32_t sr_bits(uint32_t node, uint32_t step, uint32_t seed) {
    uint32_t x = node ^ (step * 0x9E3779B9u) ^ seed;
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

// Arguments of one lattice step. Ghost faces: index 0 X_0, 1 X_1, 2 Y_0,
// 3 Y_1, 4 Z_0, 5 Z_1. `gin` is read (previous step's outgoing
// populations), `gout` is written; the host swaps them after each step as
// swapGhostInterfaces does (src/main.cuh:353-363).
struct StepArgs {
    void* mom;               // 10 moment planes per tile, block-major SoA
    const uint8_t* cls;      // node class per node (block layout) or null
    const float* bc;         // boundary class table (BC_STRIDE floats each)
    const float* gin;        // ghost set read: the six faces one after another (ghost_face)
    float* gout;             // ghost set written
    int ntiles;
    int ntx, nty, ntz;       // tiles per axis
    FastDiv by_ntx, by_nty;  // division by ntx, by nty (tile coordinates)
    int nx, ny, nz;
    int wall_x, wall_y, wall_z;
    float rho0;
    Relax R;
    float mscale[10];        // storage scale per moment (reduced precision)
    float mrscale[10];       // inverse scale
    uint32_t step;           // time step being computed (stochastic rounding key)
    uint32_t seed;           // lattice seed (stochastic rounding key)
};

// Ghost-plane indices of the reference layout with runtime tile shape
// (src/globalFunctions.h:45-124): X faces (ty,tz), Y faces (tx,tz),
// Z faces (tx,ty); slot-major within a tile, tiles x-fastest.
__host__ __device__ __forceinline__ size_t gidx_x(int ty, int tz, int slot, int tile, int BY, int BZ, int QF) {
    return (size_t)ty + BY * ((size_t)tz + BZ * ((size_t)slot + (size_t)QF * tile));
}
__host__ __device__ __forceinline__ size_t gidx_y(int tx, int tz, int slot, int tile, int BX, int BZ, int QF) {
    return (size_t)tx + BX * ((size_t)tz + BZ * ((size_t)slot + (size_t)QF * tile));
}
__host__ __device__ __forceinline__ size_t gidx_z(int tx, int ty, int slot, int tile, int BX, int BY, int QF) {
    return (size_t)tx + BX * ((size_t)ty + BY * ((size_t)slot + (size_t)QF * tile));
}

// A ghost set holds the six faces one after another: face f keeps
// face_tile_cells(f) values per tile (slot-major per tile, tiles x-fastest,
// the reference's idxPopX/Y/Z order within the face). One base pointer per
// set keeps the step kernels' scalar registers down (12 fewer than six
// face pointers per set).
__host__ __device__ constexpr size_t face_tile_cells(int f, int bx, int by, int bz, int qf) {
    return (size_t)(f < 2 ? by * bz : (f < 4 ? bx * bz : bx * by)) * (size_t)qf;
}
__host__ __device__ constexpr size_t face_tile_offset(int f, int bx, int by, int bz, int qf) {
    size_t o = 0;
    for (int j = 0; j < f; j++) o += face_tile_cells(j, bx, by, bz, qf);
    return o;
}
// Face f of the ghost set at `base` (compile-time tile shape).
template <int BX, int BY, int BZ, int QF, typename T>
__host__ __device__ __forceinline__ T* ghost_face(T* base, int f, int ntiles) {
    return base + face_tile_offset(f, BX, BY, BZ, QF) * (size_t)ntiles;
}

struct GhostOut {
    float* f[6];
};

// One lattice of a batch (device table entry). The kernel of launch offset
// k reads ghost set (cur0 + k) & 1 and computes time step step0 + k; after
// a sequence of n launches an advance kernel adds n to both.
struct InstArgs {
    StepArgs a;      // gin/gout/step are filled per launch
    float* g[2];     // both ghost sets (base pointers)
    int cur0;
    uint32_t step0;
    int has_bc;
    int pad;
};

// Batched step (step_tile_body.h): one launch advances every lattice of the
// work table by one step. work[i] = instance << 20 | tile. grid <= 0 or
// grid >= nwork: one item per block; otherwise grid blocks, block b taking
// the contiguous items [b*per, (b+1)*per), per = ceil(nwork/grid), with the
// next item's reads issued while the current one computes. has_bc = false
// only for tiles whose nodes are all bulk.
hipError_t launch_step_batch(int q, int bx, int by, int bz, int storage, bool has_bc, const InstArgs* tab,
                             const uint32_t* work, int nwork, int grid, int k, hipStream_t s);
hipError_t launch_batch_advance(InstArgs* tab, int ninst, int nsteps, hipStream_t s);

// Launchers (step_tile.hip, dispatching to the per-storage translation units
// of step_tile_body.h). Return hipError_t of the launch.
// Step of the n tiles list[0..n) (tiles 0..n-1 when list is null), one
// workgroup per tile. has_bc = false only for tiles whose nodes are all
// bulk.
hipError_t launch_step_tile(int q, int bx, int by, int bz, bool has_bc, int storage, const StepArgs& a,
                            const uint32_t* list, int n, hipStream_t s);
bool tile_shape_supported(int q, int bx, int by, int bz, int storage);
// Both bodies in one launch: the first nbulk of the n tiles (work items) are
// all-bulk, the rest hold boundary or solid nodes. Only where the two bodies
// run at the same occupancy (mixed_launch_supported).
bool mixed_launch_supported(int q, int bx, int by, int bz, int storage);
hipError_t launch_step_tile_mixed(int q, int bx, int by, int bz, int storage, const StepArgs& a, const uint32_t* list,
                                  int n, int nbulk, hipStream_t s);
hipError_t launch_step_batch_mixed(int q, int bx, int by, int bz, int storage, const InstArgs* tab,
                                   const uint32_t* work, int nwork, int nbulk, int k, hipStream_t s);

// Column kernel (step_column.hip): one (x,y) column per thread over a
// z-segment of bz planes; bx*by = 256.
hipError_t launch_step_column(int q, int bx, int by, int bz, bool has_bc, int storage, const StepArgs& a,
                              int ntiles, hipStream_t s);
bool column_shape_supported(int q, int bx, int by, int bz, int storage);
hipError_t launch_init_ghosts_column(int q, int bx, int by, int bz, int storage, const StepArgs& a,
                                     float* const g[6], int ntiles, hipStream_t s);

// Rebuild all six ghost faces of buffer `g` from the stored moments,
// unconditionally on every tile face (src/lbmInitialization.cu:220-335).
hipError_t launch_init_ghosts(int q, int bx, int by, int bz, int storage, const StepArgs& a, float* const g[6],
                              int ntiles, hipStream_t s);

// FP64 partial sums of (rho - RHO_0), rho |u|^2 / 2 and the count of
// non-finite nodes per block (partial: 3 doubles per block); then one block
// folds the partials into out3, dividing the two sums by n_nodes when it is
// positive. Deterministic (no atomics).
hipError_t launch_reduce(int storage, const void* mom, size_t ntiles, int tile_nodes, float rho0, float rs0, float rs1,
                         double* partial, int nblocks, double* out3, hipStream_t s, double n_nodes = 0.0);
// FP64 sums of |u - u_prev|^2 and |u|^2 into out3[0..1] (u_prev: 3 FP32
// planes of N values, block layout, overwritten with u).
hipError_t launch_residual(int storage, const void* mom, float* uprev, size_t ntiles, int tile_nodes, float rs1,
                           double* partial, int nblocks, double* out3, hipStream_t s);

}  // namespace mrlbm
