// Node-type maps from geometry (host): domain walls, solid voxels and their
// boundary classification, sphere voxelisation and the reference's voxel
// CSV import. Same rules as the reference's hostInitialization_nodeType,
// read_xyz_file and define_voxel_bc/bc_id (src/lbmInitialization.cu:958-1227),
// with one difference: a solid neighbour across a wall axis does not wrap.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "../../include/mrlbm.h"
#include "../../include/mrlbm_nodetype.h"

namespace {

// octants (bit k = corner (k&1, k>>1&1, k>>2&1), 1 = + side) containing direction d
uint32_t octants_of(int dx, int dy, int dz) {
    uint32_t m = 0;
    for (int k = 0; k < 8; k++) {
        const int cx = (k & 1) ? 1 : -1, cy = (k & 2) ? 1 : -1, cz = (k & 4) ? 1 : -1;
        if ((dx == 0 || dx == cx) && (dy == 0 || dy == cy) && (dz == 0 || dz == cz)) m |= 1u << k;
    }
    return m;
}

}  // namespace

extern "C" {

int mrlbm_classify(int nx, int ny, int nz, int wall_x, int wall_y, int wall_z, const uint8_t* solid,
                   uint32_t* types) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !types) return MRLBM_EINVAL;
    const int n[3] = {nx, ny, nz};
    const int wall[3] = {wall_x, wall_y, wall_z};
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const size_t j = x + (size_t)nx * (y + (size_t)ny * z);
                if (solid && solid[j]) {
                    types[j] = SOLID_NODE;
                    continue;
                }
                const int p[3] = {x, y, z};
                uint32_t bits = 0;
                for (int a = 0; a < 3; a++) {
                    if (!wall[a]) continue;
                    int d[3] = {0, 0, 0};
                    if (p[a] == 0) {
                        d[a] = -1;
                        bits |= octants_of(d[0], d[1], d[2]);
                    }
                    if (p[a] == n[a] - 1) {
                        d[a] = 1;
                        bits |= octants_of(d[0], d[1], d[2]);
                    }
                }
                if (solid) {
                    for (int dz = -1; dz <= 1; dz++)
                        for (int dy = -1; dy <= 1; dy++)
                            for (int dx = -1; dx <= 1; dx++) {
                                if (!dx && !dy && !dz) continue;
                                int q[3] = {x + dx, y + dy, z + dz};
                                bool ok = true;
                                for (int a = 0; a < 3; a++)
                                    if (q[a] < 0 || q[a] >= n[a]) {
                                        if (wall[a]) ok = false;
                                        q[a] = (q[a] + n[a]) % n[a];
                                    }
                                if (ok && solid[q[0] + (size_t)nx * (q[1] + (size_t)ny * q[2])])
                                    bits |= octants_of(dx, dy, dz);
                            }
                }
                types[j] = bits;
            }
    return MRLBM_OK;
}

int mrlbm_set_moving_wall(int nx, int ny, int nz, int axis, int side, int vel_index, uint32_t* types) {
    if (!types || axis < 0 || axis > 2 || vel_index < 0 || vel_index > 3) return MRLBM_EINVAL;
    const int n[3] = {nx, ny, nz};
    const int fixed = side < 0 ? 0 : n[axis] - 1;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const int p[3] = {x, y, z};
                if (p[axis] != fixed) continue;
                uint32_t& t = types[x + (size_t)nx * (y + (size_t)ny * z)];
                if ((t & MRLBM_NT_OCTANT_MASK) == SOLID_NODE) continue;
                t |= BC_VELOCITY_WALL | ((uint32_t)vel_index << MRLBM_NT_VEL_SHIFT);
            }
    return MRLBM_OK;
}

int mrlbm_voxelize_spheres(int nx, int ny, int nz, int nsph, const double* sph, uint8_t* solid) {
    if (!solid || (nsph > 0 && !sph)) return MRLBM_EINVAL;
    std::memset(solid, 0, (size_t)nx * ny * nz);
    for (int s = 0; s < nsph; s++) {
        const double cx = sph[4 * s], cy = sph[4 * s + 1], cz = sph[4 * s + 2], r = sph[4 * s + 3];
        const int x0 = std::max(0, (int)std::floor(cx - r)), x1 = std::min(nx - 1, (int)std::ceil(cx + r));
        const int y0 = std::max(0, (int)std::floor(cy - r)), y1 = std::min(ny - 1, (int)std::ceil(cy + r));
        const int z0 = std::max(0, (int)std::floor(cz - r)), z1 = std::min(nz - 1, (int)std::ceil(cz + r));
        for (int z = z0; z <= z1; z++)
            for (int y = y0; y <= y1; y++)
                for (int x = x0; x <= x1; x++)
                    if ((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz) <= r * r)
                        solid[x + (size_t)nx * (y + (size_t)ny * z)] = 1;
    }
    return MRLBM_OK;
}

// Voxel CSV of the reference (read_xyz_file, src/lbmInitialization.cu:978-1039;
// written by aSupportFiles/stl_voxel.m): one "x,y,z" per line; points
// outside the lattice are skipped. Returns the number of solid voxels set,
// or a negative status.
int mrlbm_read_voxel_csv(const char* path, int nx, int ny, int nz, uint8_t* solid) {
    if (!path || !solid) return MRLBM_EINVAL;
    std::ifstream f(path);
    if (!f) return MRLBM_EIO;
    std::string line;
    int count = 0;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        std::stringstream ss(line);
        std::string a, b, c;
        if (!std::getline(ss, a, ',') || !std::getline(ss, b, ',') || !std::getline(ss, c, ',')) continue;
        int x, y, z;
        try {
            x = std::stoi(a);
            y = std::stoi(b);
            z = std::stoi(c);
        } catch (...) {
            continue;
        }
        if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) continue;
        solid[x + (size_t)nx * (y + (size_t)ny * z)] = 1;
        count++;
    }
    return count;
}

}  // extern "C"
