/*
 * mrlbm_case.h - a reference case directory on the C ABI.
 *
 * The reference configures a simulation by textual inclusion of a case
 * directory, src/cases/<BC_PROBLEM>/ (src/var.h:27,39-51): constants.inc and
 * model.inc (compile-time constants), flow_initialization.inc (initial rho,
 * u per node, src/lbmInitialization.cu:94), bc_initialization.inc (node type
 * per node, src/lbmInitialization.cu:958-975) and bc_definition.inc (the
 * boundary formulas, src/mlbm.cu:456-461). This header takes the same files,
 * unchanged, into a host C++ translation unit and turns them into calls of
 * include/mrlbm.h:
 *
 *     #define CASE_DIRECTORY cases            // as src/var.h (default "cases")
 *     #define BC_PROBLEM lidDrivenCavity_3D   // as src/var.h:27
 *     #include "mrlbm_case.h"                 // compile with -I<dir holding cases/>
 *     ...
 *     mrlbm_ctx* ctx;
 *     int rc = mrlbm_case::create(0, MRLBM_STORE_FP32, &ctx);  // lattice ready to step
 *
 * What each file becomes:
 *   constants.inc, model.inc -> mrlbm_params (NX, NY, NZ_TOTAL, D3Q19/D3Q27,
 *       TAU, RHO_0, FX/FY/FZ, BC_{X,Y,Z}_WALL) and the wall-velocity table
 *       (UX_BC/UY_BC/UZ_BC[4]).
 *   bc_initialization.inc -> the node-type map, evaluated on the host for
 *       every node as hostInitialization_nodeType does; with VOXEL_FILENAME
 *       defined, the voxel CSV is read and classified first and non-BULK
 *       case types overwrite it (src/main.cuh:996-1008).
 *   bc_definition.inc -> the wall velocity of each boundary node: the
 *       fragment is evaluated once per boundary node and its velocity
 *       outputs (ux_t30, uy_t30, uz_t30) become BC_VELOCITY_WALL entries of
 *       the wall-velocity table (at most four distinct velocities). The
 *       moments themselves come from the library's moment-matching solve,
 *       which reproduces these fragments' formulas (the golden vectors of
 *       the reference's cases pin it).
 *   flow_initialization.inc -> rho, ux, uy, uz per node, uploaded with
 *       mrlbm_init_fields in the reference's second-moment mode
 *       (gpuInitialization_mom, src/lbmInitialization.cu:28-218).
 *
 * Case constants are evaluated in FP64 (dfloat = double). Fragments that
 * need a variable this header does not provide fail to compile; the
 * variables provided are those of the reference's contexts: x, y, z,
 * nodeType; rho, ux, uy, uz, cs2; pop[Q], omegaVar, rhoVar, ux_t30, uy_t30,
 * uz_t30, m_xx_t45, m_xy_t90, m_xz_t90, m_yy_t45, m_yz_t90, m_zz_t45.
 */
#ifndef MRLBM_CASE_H
#define MRLBM_CASE_H

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "mrlbm.h"
#include "mrlbm_nodetype.h"

#ifndef BC_PROBLEM
#error "define BC_PROBLEM (the case directory name, as src/var.h:27) before including mrlbm_case.h"
#endif
#ifndef CASE_DIRECTORY
#define CASE_DIRECTORY cases
#endif
#define MRLBM_CASE_STR_IMPL(A) #A
#define MRLBM_CASE_STR(A) MRLBM_CASE_STR_IMPL(A)
#define MRLBM_CASE_FILE(F) MRLBM_CASE_STR(CASE_DIRECTORY/BC_PROBLEM/F)

#ifndef __device__
#define __device__
#define MRLBM_CASE_UNDEF_DEVICE
#endif

namespace mrlbm_case {

typedef double dfloat;

#include MRLBM_CASE_FILE(constants.inc)
#include MRLBM_CASE_FILE(model.inc)

#if defined(D3Q27)
constexpr int Q = 27;
#else
constexpr int Q = 19;
#endif

#ifdef BC_X_WALL
constexpr int WALL_X = 1;
#else
constexpr int WALL_X = 0;
#endif
#ifdef BC_Y_WALL
constexpr int WALL_Y = 1;
#else
constexpr int WALL_Y = 0;
#endif
#ifdef BC_Z_WALL
constexpr int WALL_Z = 1;
#else
constexpr int WALL_Z = 0;
#endif

/* Lattice parameters of the case (device 0). The tile is the reference's
 * block, BLOCK_NX x BLOCK_NY x BLOCK_NZ = 8x8x8 (src/definitions.h:90-92):
 * a case whose node types read populations from outside the domain runs
 * as the reference's kernel runs it only at that block shape. */
inline mrlbm_params params(int storage = MRLBM_STORE_FP32, int device = 0) {
    mrlbm_params p{};
    p.nx = NX;
    p.ny = NY;
    p.nz = NZ_TOTAL;
    p.bx = 8;
    p.by = 8;
    p.bz = 8;
    p.q = Q;
    p.storage = storage;
    p.wall_x = WALL_X;
    p.wall_y = WALL_Y;
    p.wall_z = WALL_Z;
    p.kernel = MRLBM_KERNEL_AUTO;
    p.tau = TAU;
    p.rho0 = RHO_0;
    p.fx = FX;
    p.fy = FY;
    p.fz = FZ;
    p.device = device;
    p.seed = 0;
    return p;
}

/* bc_initialization.inc at node (x, y, z). */
inline uint32_t node_type_at(int x, int y, int z) {
    unsigned int nodeType = BULK;
    {
#include MRLBM_CASE_FILE(bc_initialization.inc)
    }
    return (uint32_t)nodeType;
}

/* flow_initialization.inc at node (x, y, z): unscaled rho, u. */
inline void flow_at(int x, int y, int z, double& rho_out, double& ux_out, double& uy_out, double& uz_out) {
    const dfloat cs2 = 1.0 / 3.0;
    (void)cs2;
    dfloat rho = RHO_0, ux = 0.0, uy = 0.0, uz = 0.0;
    {
#include MRLBM_CASE_FILE(flow_initialization.inc)
    }
    rho_out = rho;
    ux_out = ux;
    uy_out = uy;
    uz_out = uz;
}

/* bc_definition.inc for one node: the velocity the fragment imposes
 * (unscaled), evaluated on the rest-state populations. */
inline void wall_velocity_at(int x, int y, int z, uint32_t node_type, double u[3]) {
    static const double W19[4] = {1.0 / 3, 1.0 / 18, 1.0 / 36, 0.0};
    static const double W27[4] = {8.0 / 27, 2.0 / 27, 1.0 / 54, 1.0 / 216};
    const double* W = Q == 19 ? W19 : W27;
    dfloat pop[Q];
    for (int i = 0; i < Q; i++) pop[i] = RHO_0 * W[i == 0 ? 0 : (i < 7 ? 1 : (i < 19 ? 2 : 3))];
    const unsigned int nodeType = node_type;
    const dfloat omegaVar = 1.0 / TAU;
    const dfloat nan = NAN;
    dfloat rhoVar = nan, ux_t30 = 0.0, uy_t30 = 0.0, uz_t30 = 0.0;
    dfloat m_xx_t45 = nan, m_xy_t90 = nan, m_xz_t90 = nan, m_yy_t45 = nan, m_yz_t90 = nan, m_zz_t45 = nan;
    {
#include MRLBM_CASE_FILE(bc_definition.inc)
    }
    (void)x, (void)y, (void)z, (void)omegaVar, (void)rhoVar;
    (void)m_xx_t45, (void)m_xy_t90, (void)m_xz_t90, (void)m_yy_t45, (void)m_yz_t90, (void)m_zz_t45;
    u[0] = ux_t30;
    u[1] = uy_t30;
    u[2] = uz_t30;
}

/* Node-type map of the case's NX x NY x NZ_TOTAL lattice (x-fastest), in
 * the reference encoding, and the wall-velocity table vel[4][3]. Nodes the case itself marks BC_VELOCITY_WALL | index
 * keep their index into UX_BC/UY_BC/UZ_BC; other boundary nodes whose
 * bc_definition.inc imposes a nonzero velocity are marked BC_VELOCITY_WALL
 * with an index the case leaves unused. Returns 0 or a negative MRLBM_E*
 * status (MRLBM_EINVAL: more distinct wall velocities than four). */
inline int node_types(uint32_t* types, double vel[4][3]) {
    const int nx = NX, ny = NY, nz = NZ_TOTAL;
    const size_t n = (size_t)nx * ny * nz;
    for (size_t j = 0; j < n; j++) types[j] = BULK;
#ifdef VOXEL_FILENAME
    {
        std::vector<uint8_t> solid(n, 0);
        const int rc = mrlbm_read_voxel_csv(VOXEL_FILENAME, nx, ny, nz, solid.data());
        if (rc < 0) return rc;
        const int rc2 = mrlbm_classify(nx, ny, nz, WALL_X, WALL_Y, WALL_Z, solid.data(), types);
        if (rc2) return rc2;
    }
#endif
    bool taken[4] = {false, false, false, false};
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const uint32_t t = node_type_at(x, y, z);
                if (t == BULK) continue;
                types[x + (size_t)nx * (y + (size_t)ny * z)] = t;
                if ((t & MRLBM_NT_KIND_MASK) == BC_VELOCITY_WALL) taken[(t & MRLBM_NT_VEL_MASK) >> MRLBM_NT_VEL_SHIFT] = true;
            }
    for (int k = 0; k < 4; k++) {
        vel[k][0] = UX_BC[k];
        vel[k][1] = UY_BC[k];
        vel[k][2] = UZ_BC[k];
    }
    for (size_t j = 0; j < n; j++) {
        uint32_t& t = types[j];
        const uint32_t oct = t & MRLBM_NT_OCTANT_MASK;
        if (oct == 0 || oct == SOLID_NODE || (t & MRLBM_NT_KIND_MASK) != BC_ZERO_VELOCITY_WALL) continue;
        const int x = (int)(j % nx), y = (int)((j / nx) % ny), z = (int)(j / ((size_t)nx * ny));
        double u[3];
        wall_velocity_at(x, y, z, t, u);
        if (u[0] == 0.0 && u[1] == 0.0 && u[2] == 0.0) continue;
        int k = 0;
        while (k < 4 && !(taken[k] && vel[k][0] == u[0] && vel[k][1] == u[1] && vel[k][2] == u[2])) k++;
        if (k == 4) {
            k = 0;
            while (k < 4 && taken[k]) k++;
            if (k == 4) return MRLBM_EINVAL;
            vel[k][0] = u[0];
            vel[k][1] = u[1];
            vel[k][2] = u[2];
            taken[k] = true;
        }
        t = (t & ~(MRLBM_NT_KIND_MASK | MRLBM_NT_VEL_MASK)) | BC_VELOCITY_WALL | ((uint32_t)k << MRLBM_NT_VEL_SHIFT);
    }
    return MRLBM_OK;
}

/* Initial fields of the case's lattice (x-fastest, unscaled). */
inline void flow_fields(double* rho, double* ux, double* uy, double* uz) {
    const int nx = NX, ny = NY, nz = NZ_TOTAL;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const size_t j = x + (size_t)nx * (y + (size_t)ny * z);
                flow_at(x, y, z, rho[j], ux[j], uy[j], uz[j]);
            }
}

/* The case's lattice, set up as initializeDomain does (src/main.cuh:904-1070):
 * parameters, wall velocities, node types, initial fields. */
inline int create(int device, int storage, mrlbm_ctx** out) {
    *out = nullptr;
    const mrlbm_params p = params(storage, device);
    const size_t n = (size_t)p.nx * p.ny * p.nz;
    std::vector<uint32_t> types(n);
    double vel[4][3];
    int rc = node_types(types.data(), vel);
    if (rc) return rc;
    mrlbm_ctx* c = nullptr;
    if ((rc = mrlbm_create(&p, &c))) return rc;
    for (int k = 0; k < 4 && !rc; k++) rc = mrlbm_set_bc_velocity(c, k, vel[k][0], vel[k][1], vel[k][2]);
    if (!rc) rc = mrlbm_set_node_types(c, types.data());
    if (!rc) {
        std::vector<double> rho(n), ux(n), uy(n), uz(n);
        flow_fields(rho.data(), ux.data(), uy.data(), uz.data());
        rc = mrlbm_init_fields(c, rho.data(), ux.data(), uy.data(), uz.data(), MRLBM_INIT_REFERENCE);
    }
    if (rc) {
        mrlbm_destroy(c);
        return rc;
    }
    *out = c;
    return MRLBM_OK;
}

}  // namespace mrlbm_case

#ifdef MRLBM_CASE_UNDEF_DEVICE
#undef __device__
#undef MRLBM_CASE_UNDEF_DEVICE
#endif

#endif /* MRLBM_CASE_H */
