/*
 * Node-type word of the MR-LBM plugin surface.
 *
 * Case fragments (bc_initialization.inc / bc_definition.inc) name node types
 * with these macros, so the bit values must equal the reference's
 * (src/nodeTypeMap.h:10-80) for existing case files to drop in unchanged.
 *
 * Layout of the 32-bit word (reference src/nodeTypeMap.h:56-77):
 *   bits 0-7   solid octants around the node. Bit k is the octant whose
 *              corner is (k&1, (k>>1)&1, (k>>2)&1), 1 meaning the + side.
 *              0x00 = bulk fluid node, 0xFF = solid node.
 *   bits 8-10  boundary-condition kind (000 no-slip wall, 001 fixed velocity).
 *   bits 11-12 index into the wall-velocity table UX_BC/UY_BC/UZ_BC[4].
 *
 * A population i at a boundary node is "known" (streamed from a fluid
 * neighbour) unless every octant touched by its source direction -c_i is
 * solid. This is the rule the reference's generated formulas follow
 * (e.g. cases/voxel/bc_definition.inc, cases/lidDrivenCavity_3D/bc_definition.inc).
 */
#ifndef MRLBM_NODETYPE_H
#define MRLBM_NODETYPE_H

#define BULK                  (0x00u)
/* faces */
#define NORTH                 (0xCCu) /* y = NY-1 */
#define SOUTH                 (0x33u) /* y = 0    */
#define WEST                  (0x55u) /* x = 0    */
#define EAST                  (0xAAu) /* x = NX-1 */
#define FRONT                 (0xF0u) /* z = NZ-1 */
#define BACK                  (0x0Fu) /* z = 0    */
/* edges */
#define NORTH_WEST            (0xDDu)
#define NORTH_EAST            (0xEEu)
#define NORTH_FRONT           (0xFCu)
#define NORTH_BACK            (0xCFu)
#define SOUTH_WEST            (0x77u)
#define SOUTH_EAST            (0xBBu)
#define SOUTH_FRONT           (0xF3u)
#define SOUTH_BACK            (0x3Fu)
#define WEST_FRONT            (0xF5u)
#define WEST_BACK             (0x5Fu)
#define EAST_FRONT            (0xFAu)
#define EAST_BACK             (0xAFu)
/* corners */
#define NORTH_WEST_FRONT      (0xFDu)
#define NORTH_WEST_BACK       (0xDFu)
#define NORTH_EAST_FRONT      (0xFEu)
#define NORTH_EAST_BACK       (0xEFu)
#define SOUTH_WEST_FRONT      (0xF7u)
#define SOUTH_WEST_BACK       (0x7Fu)
#define SOUTH_EAST_FRONT      (0xFBu)
#define SOUTH_EAST_BACK       (0xBFu)

#define SOLID_NODE            (0xFFu)

/* boundary-condition kind, bits 8-10 */
#define BC_ZERO_VELOCITY_WALL (0x000u)
#define BC_VELOCITY_WALL      (0x100u)
#define BC_OUTFLOW            (0x200u)
#define BC_SYMMETRY           (0x300u)
#define BC_FREESLIP           (0x400u)
/* wall-velocity table index, bits 11-12 */
#define BC_VELOCITY_INDEX_0   (0x0000u)
#define BC_VELOCITY_INDEX_1   (0x0800u)
#define BC_VELOCITY_INDEX_2   (0x1000u)
#define BC_VELOCITY_INDEX_3   (0x1800u)

#define MISSING_DEFINITION    (0xFFFFFFFFu)

#define MRLBM_NT_OCTANT_MASK  (0xFFu)
#define MRLBM_NT_KIND_MASK    (0x700u)
#define MRLBM_NT_KIND_SHIFT   8
#define MRLBM_NT_VEL_MASK     (0x1800u)
#define MRLBM_NT_VEL_SHIFT    11

#endif /* MRLBM_NODETYPE_H */
