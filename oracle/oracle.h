/*
 * FP64 CPU oracle of the moment-representation LBM step.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker / the reported CPU baseline. The product (mr-lbm_amd/) never
 * calls it.
 *
 * It restates the reference algorithm of gpuMomCollisionStream
 * (src/mlbm.cu:3-714) without tiles or ghost planes: each step is
 *   pass 1  f*_i(x) = R(moments(x))            (reconstruction.inc:1-47)
 *   pass 2  f_i(x)  = f*_i(x - c_i)  (global pull; periodic axes wrap,
 *                     sources outside a wall axis or inside a solid node
 *                     are NaN, proving the boundary rule never reads them)
 *           boundary node -> moment-based BC, else bulk moments
 *           (src/mlbm.cu:456-509), scale, collide (collision.inc:2-14),
 *           store moments.
 * Boundary nodes solve the reference's moment-matching system per node by
 * Gaussian elimination in FP64 (the reference's case files hold its
 * closed-form solutions, e.g. cases/lidDrivenCavity_3D/bc_definition.inc).
 *
 * Moments use the reference's stored convention: rho - RHO_0, 3u,
 * 4.5 m_aa, 9 m_ab (src/arrayIndex.h:8-17,186), planes of N values,
 * x-fastest.
 */
#ifndef MRLBM_ORACLE_H
#define MRLBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

orc* orc_create(int nx, int ny, int nz, int q, double tau, double rho0, double fx, double fy, double fz, int wall_x,
                int wall_y, int wall_z);
void orc_destroy(orc* o);
int orc_set_bc_velocity(orc* o, int k, double ux, double uy, double uz);
/* reference-encoded node types, x-fastest; NULL = all bulk. Returns 0 or -1 */
int orc_set_node_types(orc* o, const uint32_t* types);
/* mode 0: m_ab = u_a u_b; mode 1: src/lbmInitialization.cu:102-130 */
int orc_init_fields(orc* o, const double* rho, const double* ux, const double* uy, const double* uz, int mode);
int orc_init_taylor_green(orc* o, double u0, int mode);
int orc_init_uniform(orc* o, double rho, double ux, double uy, double uz, int mode);
int orc_step(orc* o, int nsteps);
/* seconds of steady_clock time spent in nsteps (the step loop only) */
double orc_step_timed(orc* o, int nsteps);
void orc_get_moments(const orc* o, double* out10n);
/* also the moments the ghost planes are rebuilt from (ghost semantics) */
void orc_set_moments(orc* o, const double* in10n);
/* Reference ghost semantics for tile shape bx x by x bz (0,0,0: off, the
 * default). Off, a population streamed from outside a wall axis is NaN, so
 * a boundary rule that reads one poisons its node. On, it is the value the
 * reference's kernel reads (src/includeFiles/popLoad.inc:2-86,
 * interface.h:6-47): the ghost plane of the face winning in popLoad.inc's
 * X, Y, Z order, of the tile holding the periodically wrapped source;
 * never rewritten after initialisation when that face is on a wall-axis
 * domain boundary (initial population), else the wrapped source's current
 * one. Call before the initialisation. Returns 0, or -1 if the tile does
 * not divide the lattice. */
int orc_set_ghost_tiles(orc* o, int bx, int by, int bz);
/* mean(rho - RHO_0) and mean(rho |u|^2 / 2), all nodes, long double sums */
void orc_reduce(const orc* o, double* mass, double* ke);
/* number of NaN values in the stored moments of non-solid nodes */
int64_t orc_count_nan(const orc* o);
void orc_set_threads(int n);
int orc_get_threads(void);

/* Single-node kernels for golden-vector checks (unscaled in/out).
 * orc_bc_node: pop[q] streamed populations (unknown entries ignored),
 *   node type word, omega, wall velocity -> out = rho, ux, uy, uz, mxx,
 *   mxy, mxz, myy, myz, mzz (unscaled, as a bc_definition.inc fragment
 *   leaves them). Returns 0, or -1 if the system is singular.
 * orc_known_mask: bit i set if population i is known for node type. */
int orc_bc_node(int q, uint32_t node_type, double omega, const double u[3], const double* pop, double out[10]);
uint32_t orc_known_mask(int q, uint32_t node_type);
/* reconstruction of scaled moments m[10] (rho full) -> pop[q] */
void orc_reconstruct(int q, const double m[10], double* pop);
/* collision of scaled moments in place (m[0] = rho full), force unscaled */
void orc_collide(double m[10], double omega, double fx, double fy, double fz);

#ifdef __cplusplus
}
#endif

#endif
