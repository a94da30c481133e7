// Node-type constants of the reference (the #define lines of
// src/nodeTypeMap.h, extracted by oracle/ref/Makefile into NODETYPE_HEADER),
// by name, for tests/golden/gen_bc_vectors.py to record as a fixture.
// Every name this project's include/mrlbm_nodetype.h defines with a
// counterpart in the reference is listed; a name missing from the
// reference header fails the build.
//
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/, never shipped.
#include <cstdint>

#ifndef NODETYPE_HEADER
#error "NODETYPE_HEADER must name the extracted node-type constants"
#endif
#include NODETYPE_HEADER

#define NODE_TYPE_NAMES(X)                                                                                  \
    X(BULK) X(NORTH) X(SOUTH) X(WEST) X(EAST) X(FRONT) X(BACK) X(NORTH_WEST) X(NORTH_EAST) X(NORTH_FRONT)   \
    X(NORTH_BACK) X(SOUTH_WEST) X(SOUTH_EAST) X(SOUTH_FRONT) X(SOUTH_BACK) X(WEST_FRONT) X(WEST_BACK)       \
    X(EAST_FRONT) X(EAST_BACK) X(NORTH_WEST_FRONT) X(NORTH_WEST_BACK) X(NORTH_EAST_FRONT) X(NORTH_EAST_BACK) \
    X(SOUTH_WEST_FRONT) X(SOUTH_WEST_BACK) X(SOUTH_EAST_FRONT) X(SOUTH_EAST_BACK) X(SOLID_NODE)             \
    X(BC_ZERO_VELOCITY_WALL) X(BC_VELOCITY_WALL) X(BC_OUTFLOW) X(BC_SYMMETRY) X(BC_FREESLIP)               \
    X(BC_VELOCITY_INDEX_0) X(BC_VELOCITY_INDEX_1) X(BC_VELOCITY_INDEX_2) X(BC_VELOCITY_INDEX_3)            \
    X(MISSING_DEFINITION)

#define NAME_OF(n) #n,
#define VALUE_OF(n) (uint32_t)(n),
static const char* const kNames[] = {NODE_TYPE_NAMES(NAME_OF)};
static const uint32_t kValues[] = {NODE_TYPE_NAMES(VALUE_OF)};

extern "C" int nodetype_ref_count(void) { return (int)(sizeof kValues / sizeof kValues[0]); }
extern "C" const char* nodetype_ref_name(int i) { return kNames[i]; }
extern "C" uint32_t nodetype_ref_value(int i) { return kValues[i]; }
