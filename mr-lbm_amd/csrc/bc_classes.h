// Boundary-condition classes: host-side construction of the per-class
// moment-matching system used by bc_solve (lattice.h).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace mrlbm {

// Known-population mask of a node from the octant byte of its node type
// (include/mrlbm_nodetype.h): population i is unknown iff every octant its
// source direction -c_i touches is solid.
uint32_t known_mask(int q, uint32_t octants);

// Build the 7x7 system A for known set `known`, relaxation omega and wall
// velocity u (unscaled), in FP64, and invert it. Returns false if singular.
// Unknown vector x = (rho, P_xx, P_xy, P_xz, P_yy, P_yz, P_zz), P = rho m;
// right-hand side rho_I (1, m_I). The sums of A run over the known
// populations of lattice closure_q (see closure_lattice).
bool build_bc_inverse(int q, uint32_t known, double omega, const double u[3], double ainv[49], double* cond,
                      int closure_q);

// Lattice whose weights and velocities form a boundary node's closure: the
// lattice itself, except D3Q27 nodes on a single plane wall (octant byte
// NORTH, SOUTH, WEST, EAST, FRONT or BACK), whose closure is the D3Q19 one
// as in the reference's only D3Q27 wall formulas
// (cases/parallelPlates_D3Q27/bc_definition.inc:18-72). Edges, corners and
// voxel patterns, for which the reference has no D3Q27 formula, use the
// consistent D3Q27 closure: with the D3Q19 one a D3Q27 duct does not
// converge (steady duct error 1.3 % -> 1.9 % as the cross-section doubles,
// against 0.84 % -> 0.20 % with the D3Q27 closure).
int closure_lattice(int q, uint32_t octants);

struct BcClassTable {
    std::vector<uint32_t> keys;      // node-type key (octants | kind | velocity index) per class, index 1..
    std::vector<float> table;        // 72 floats per class (lattice.h BC_STRIDE; class 0 unused)
    std::vector<uint8_t> class_of;   // per node in the caller's x-fastest order
    bool has_bc = false, has_solid = false;
};

// Class record (BC_STRIDE floats, layout of lattice.h) of a boundary node
// with the given octant byte and wall velocity u. False if singular.
bool class_record(int q, uint32_t octants, double omega, const double u[3], float* rec);

// Records of every no-slip octant pattern, indexed by the octant byte
// itself (256 records; 0 bulk and 255 solid unused), for class maps built
// on the device where class id == octant byte. Patterns whose system is
// singular are flagged in `singular` (bit oct of the 256-bit set).
void build_octant_table(int q, double omega, std::vector<float>& table, uint32_t singular[8]);

// Map a reference-encoded node-type map to class ids (0 bulk, 255 solid,
// 1..254 boundary classes) and build the device table. bc_vel is the
// UX_BC/UY_BC/UZ_BC table (4 x 3). Returns an empty string on success.
std::string build_class_table(int q, double omega, const double bc_vel[4][3], const uint32_t* types, size_t n,
                              BcClassTable& out);

}  // namespace mrlbm
