// Evaluates a reference case directory through include/mrlbm_case.h - the
// same adapter a user compiles against src/cases/<BC_PROBLEM>/ - and
// exports its parameters, node-type map, wall-velocity table and initial
// fields for tests/golden/gen_case_vectors.py to record as fixtures.
//
// TEST INFRASTRUCTURE ONLY: built by oracle/ref/Makefile into oracle/_ref/
// against the case files where they lie under the reference tree (a copy
// of constants.inc with the size constant N changed, for the cases recorded
// at a smaller size), never shipped.
#include "mrlbm_case.h"

extern "C" int case_params(mrlbm_params* p) {
    *p = mrlbm_case::params();
    return 0;
}

extern "C" int case_node_types(uint32_t* types, double* vel12) {
    double vel[4][3];
    const int rc = mrlbm_case::node_types(types, vel);
    for (int k = 0; k < 4; k++)
        for (int a = 0; a < 3; a++) vel12[3 * k + a] = vel[k][a];
    return rc;
}

extern "C" int case_flow_fields(double* rho, double* ux, double* uy, double* uz) {
    mrlbm_case::flow_fields(rho, ux, uy, uz);
    return 0;
}
