// Evaluates one of the reference's case bc_definition.inc fragments on a
// single node, in FP64 (dfloat = double), through the fragment's own
// variable contract (src/mlbm.cu:456-461): in pop[Q], nodeType, omegaVar,
// U_MAX; out rhoVar, ux_t30..uz_t30, m_xx_t45..m_zz_t45 (unscaled).
// The node-type macros are the reference's own: the #define lines of
// src/nodeTypeMap.h, extracted by oracle/ref/Makefile (NODETYPE_HEADER).
//
// TEST INFRASTRUCTURE ONLY: built by oracle/ref/Makefile into oracle/_ref/
// from the fragment files where they lie under the reference tree, and used
// only by tests/golden/gen_bc_vectors.py to write golden vectors.
#include <cmath>
#include <limits>

#ifndef NODETYPE_HEADER
#error "NODETYPE_HEADER must name the node-type constants extracted from src/nodeTypeMap.h"
#endif
#include NODETYPE_HEADER

#ifndef BC_Q
#define BC_Q 19
#endif
#ifndef BC_FRAGMENT
#error "BC_FRAGMENT must name the bc_definition.inc to evaluate"
#endif

typedef double dfloat;

extern "C" int bcref_q(void) { return BC_Q; }

extern "C" int bcref_eval(unsigned int nodeType, double omega_in, double umax_in, const double* pop_in,
                          double* out) {
    const dfloat nan = std::numeric_limits<double>::quiet_NaN();
    dfloat pop[BC_Q];
    for (int i = 0; i < BC_Q; i++) pop[i] = pop_in[i];
    const dfloat omegaVar = omega_in;
    const dfloat U_MAX = umax_in;
    const dfloat RHO_0 = 1.0;
    const dfloat cs2 = 1.0 / 3.0;
    (void)RHO_0;
    (void)cs2;
    (void)U_MAX;
    dfloat rhoVar = nan, ux_t30 = nan, uy_t30 = nan, uz_t30 = nan;
    dfloat m_xx_t45 = nan, m_xy_t90 = nan, m_xz_t90 = nan, m_yy_t45 = nan, m_yz_t90 = nan, m_zz_t45 = nan;
#define VOXEL_FILENAME "harness"
#include BC_FRAGMENT
    out[0] = rhoVar;
    out[1] = ux_t30;
    out[2] = uy_t30;
    out[3] = uz_t30;
    out[4] = m_xx_t45;
    out[5] = m_xy_t90;
    out[6] = m_xz_t90;
    out[7] = m_yy_t45;
    out[8] = m_yz_t90;
    out[9] = m_zz_t45;
    return 0;
}
