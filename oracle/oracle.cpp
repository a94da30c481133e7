// FP64 CPU oracle - see oracle.h. TEST INFRASTRUCTURE ONLY.
#include "oracle.h"

#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

namespace {

// Velocity sets: src/includeFiles/velocitySets/D3Q19.inc:1-24 (first 19
// entries) and D3Q27.inc:1-28.
const int CX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
const int CY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
const int CZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};

double W(int q, int i) {
    int k = i == 0 ? 0 : (i < 7 ? 1 : (i < 19 ? 2 : 3));
    if (q == 19) {
        const double w[3] = {1.0 / 3.0, 1.0 / 18.0, 1.0 / 36.0};
        return w[k];
    }
    const double w[4] = {8.0 / 27.0, 2.0 / 27.0, 1.0 / 54.0, 1.0 / 216.0};
    return w[k];
}

const double NaN = std::numeric_limits<double>::quiet_NaN();

// H_i components in the order xx, xy, xz, yy, yz, zz
void herm(int i, double h[6]) {
    h[0] = CX[i] * CX[i] - 1.0 / 3.0;
    h[1] = CX[i] * CY[i];
    h[2] = CX[i] * CZ[i];
    h[3] = CY[i] * CY[i] - 1.0 / 3.0;
    h[4] = CY[i] * CZ[i];
    h[5] = CZ[i] * CZ[i] - 1.0 / 3.0;
}

// Octant rule of the node-type byte (src/nodeTypeMap.h:56-77, bc_id
// src/lbmInitialization.cu:1080-1227): the source direction s = -c_i is
// blocked when every octant containing s is solid.
uint32_t known_mask_impl(int q, uint32_t octants) {
    uint32_t m = 0;
    for (int i = 0; i < q; i++) {
        bool blocked = true;
        for (int k = 0; k < 8; k++) {
            const int sx = -CX[i], sy = -CY[i], sz = -CZ[i];
            const int ox = (k & 1) ? 1 : -1, oy = (k & 2) ? 1 : -1, oz = (k & 4) ? 1 : -1;
            const bool inside = (sx == 0 || sx == ox) && (sy == 0 || sy == oy) && (sz == 0 || sz == oz);
            if (inside && !((octants >> k) & 1u)) {
                blocked = false;
                break;
            }
        }
        if (!blocked) m |= 1u << i;
    }
    return m;
}

// Solve A x = b (7x7) by Gaussian elimination with partial pivoting.
bool solve7(double A[7][7], double b[7], double x[7]) {
    for (int c = 0; c < 7; c++) {
        int p = c;
        for (int r = c + 1; r < 7; r++)
            if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
        if (std::fabs(A[p][c]) < 1e-14) return false;
        if (p != c) {
            for (int k = 0; k < 7; k++) std::swap(A[c][k], A[p][k]);
            std::swap(b[c], b[p]);
        }
        for (int r = c + 1; r < 7; r++) {
            const double f = A[r][c] / A[c][c];
            for (int k = c; k < 7; k++) A[r][k] -= f * A[c][k];
            b[r] -= f * b[c];
        }
    }
    for (int r = 6; r >= 0; r--) {
        double s = b[r];
        for (int k = r + 1; k < 7; k++) s -= A[r][k] * x[k];
        x[r] = s / A[r][r];
    }
    return true;
}

// Moment-based boundary condition at one node, unscaled in and out.
// Unknowns: rho and P_ab = rho m_ab (m pre-collision). Equations (the
// system the reference's generated bc_definition.inc formulas solve):
//   sum_{K} w_i [rho (1 + 3 c.u) + 4.5 H_i:((1-omega) P + omega rho u u)] = rho_I
//   sum_{K} w_i H_i,ab [rho (1 + 3 c.u) + 4.5 H_i:P]                     = rho_I m_I,ab
// with rho_I = sum_K f_i, rho_I m_I = sum_K f_i H_i.
// Lattice of the closure: D3Q19 coefficients for D3Q27 nodes on a single
// plane wall, as the reference's only D3Q27 wall formulas have them
// (cases/parallelPlates_D3Q27/bc_definition.inc:18-72); the lattice's own
// otherwise (edges, corners and voxel patterns have no D3Q27 formula in the
// reference, and the D3Q19 closure there keeps a D3Q27 duct from
// converging). Octant bytes of the six plane walls: src/nodeTypeMap.h.
int closure_q(int q, uint32_t octants) {
    const bool face = octants == 0xCC || octants == 0x33 || octants == 0x55 || octants == 0xAA ||
                      octants == 0xF0 || octants == 0x0F;
    return (q == 27 && face) ? 19 : q;
}

bool bc_solve(int q, int qc, uint32_t known, double omega, const double u[3], const double* pop, double out[10]) {
    double rI = 0, S[6] = {0, 0, 0, 0, 0, 0};
    double A[7][7] = {};
    const double coef[6] = {1, 2, 2, 1, 2, 1};
    const double uu[6] = {u[0] * u[0], u[0] * u[1], u[0] * u[2], u[1] * u[1], u[1] * u[2], u[2] * u[2]};
    // Incoming moments run over every known population of the lattice; the
    // closure matrix over the known populations of lattice qc (closure_q).
    for (int i = 0; i < q; i++) {
        if (!((known >> i) & 1u)) continue;
        double h[6];
        herm(i, h);
        rI += pop[i];
        for (int k = 0; k < 6; k++) S[k] += pop[i] * h[k];
    }
    for (int i = 0; i < qc; i++) {
        if (!((known >> i) & 1u)) continue;
        double h[6];
        herm(i, h);
        const double w = W(qc, i);
        const double cu = CX[i] * u[0] + CY[i] * u[1] + CZ[i] * u[2];
        double huu = 0;
        for (int k = 0; k < 6; k++) huu += coef[k] * h[k] * uu[k];
        A[0][0] += w * (1.0 + 3.0 * cu + 4.5 * omega * huu);
        for (int k = 0; k < 6; k++) A[0][1 + k] += w * 4.5 * (1.0 - omega) * coef[k] * h[k];
        for (int r = 0; r < 6; r++) {
            A[1 + r][0] += w * h[r] * (1.0 + 3.0 * cu);
            for (int k = 0; k < 6; k++) A[1 + r][1 + k] += w * h[r] * 4.5 * coef[k] * h[k];
        }
    }
    double b[7] = {rI, S[0], S[1], S[2], S[3], S[4], S[5]};
    double x[7];
    if (!solve7(A, b, x)) return false;
    out[0] = x[0];
    out[1] = u[0];
    out[2] = u[1];
    out[3] = u[2];
    for (int k = 0; k < 6; k++) out[4 + k] = x[1 + k] / x[0];
    return true;
}

// Reconstruction (src/colrec/MR_LBM/reconstruction.inc): scaled moments.
void reconstruct(int q, const double m[10], double* f) {
    const double rho = m[0], ux = m[1], uy = m[2], uz = m[3];
    const double mxx = m[4], mxy = m[5], mxz = m[6], myy = m[7], myz = m[8], mzz = m[9];
    const double pics2 = 1.0 - (mxx + myy + mzz) / 3.0;
    for (int i = 0; i < q; i++) {
        const double cx = CX[i], cy = CY[i], cz = CZ[i];
        f[i] = rho * W(q, i) *
               (pics2 + cx * ux + cy * uy + cz * uz + cx * cx * mxx + cy * cy * myy + cz * cz * mzz + cx * cy * mxy +
                cx * cz * mxz + cy * cz * myz);
    }
}

// Collision in scaled variables (src/colrec/MR_LBM/collision.inc:2-14).
void collide(double m[10], double omega, double fx, double fy, double fz) {
    const double rho = m[0], ir = 1.0 / rho;
    const double t = 1.0 - omega, tt = 1.0 - omega / 2.0, od2 = omega / 2.0, tt3 = tt * 3.0;
    double& ux = m[1];
    double& uy = m[2];
    double& uz = m[3];
    ux += 3.0 * ir * fx;
    uy += 3.0 * ir * fy;
    uz += 3.0 * ir * fz;
    const double i15 = 3.0 * ir / 2.0;
    m[4] = t * m[4] + od2 * ux * ux + i15 * tt * (2.0 * fx * ux);
    m[7] = t * m[7] + od2 * uy * uy + i15 * tt * (2.0 * fy * uy);
    m[9] = t * m[9] + od2 * uz * uz + i15 * tt * (2.0 * fz * uz);
    m[5] = t * m[5] + omega * ux * uy + tt3 * ir * (fx * uy + fy * ux);
    m[6] = t * m[6] + omega * ux * uz + tt3 * ir * (fx * uz + fz * ux);
    m[8] = t * m[8] + omega * uy * uz + tt3 * ir * (fy * uz + fz * uy);
}

// Stored moments from unscaled fields (src/lbmInitialization.cu:28-130).
void init_node(int q, int mode, double rho0, double rho, double ux, double uy, double uz, double m[10]) {
    double p[6];
    if (mode == 0) {
        p[0] = ux * ux;
        p[1] = ux * uy;
        p[2] = ux * uz;
        p[3] = uy * uy;
        p[4] = uy * uz;
        p[5] = uz * uz;
    } else {
        // f_eq with weight w_i RHO_0 summed over the D3Q19 subsets for
        // either lattice, divided by rho (the reference's init, including
        // its D3Q27 subset quirk).
        double S[6] = {0, 0, 0, 0, 0, 0};
        const double uu = ux * ux + uy * uy + uz * uz;
        for (int i = 0; i < 19; i++) {
            const double cu = CX[i] * ux + CY[i] * uy + CZ[i] * uz;
            const double f = W(q, i) * rho0 * (1.0 - 1.5 * uu + 3.0 * cu + 4.5 * cu * cu);
            S[0] += f * CX[i] * CX[i];
            S[1] += f * CX[i] * CY[i];
            S[2] += f * CX[i] * CZ[i];
            S[3] += f * CY[i] * CY[i];
            S[4] += f * CY[i] * CZ[i];
            S[5] += f * CZ[i] * CZ[i];
        }
        p[0] = S[0] / rho - 1.0 / 3.0;
        p[1] = S[1] / rho;
        p[2] = S[2] / rho;
        p[3] = S[3] / rho - 1.0 / 3.0;
        p[4] = S[4] / rho;
        p[5] = S[5] / rho - 1.0 / 3.0;
    }
    m[0] = rho - rho0;
    m[1] = 3 * ux;
    m[2] = 3 * uy;
    m[3] = 3 * uz;
    m[4] = 4.5 * p[0];
    m[5] = 9 * p[1];
    m[6] = 9 * p[2];
    m[7] = 4.5 * p[3];
    m[8] = 9 * p[4];
    m[9] = 4.5 * p[5];
}

}  // namespace

struct orc {
    int nx, ny, nz, q;
    size_t N;
    double tau, omega, rho0, F[3];
    int wall[3];
    double vel[4][3];
    std::vector<double> mom;    // 10 * N scaled stored moments
    std::vector<double> fpost;  // q * N post-collision populations
    std::vector<uint32_t> type;  // node types (0 = bulk)
    bool any_bc = false;
    int gtile[3] = {0, 0, 0};     // reference ghost semantics for this tile shape (0: off)
    std::vector<double> mom0;     // moments the ghost planes were built from (ghost semantics only)
};

extern "C" {

orc* orc_create(int nx, int ny, int nz, int q, double tau, double rho0, double fx, double fy, double fz, int wall_x,
                int wall_y, int wall_z) {
    if ((q != 19 && q != 27) || nx <= 0 || ny <= 0 || nz <= 0 || !(tau > 0.5)) return nullptr;
    orc* o = new orc();
    o->nx = nx;
    o->ny = ny;
    o->nz = nz;
    o->q = q;
    o->N = (size_t)nx * ny * nz;
    o->tau = tau;
    o->omega = 1.0 / tau;
    o->rho0 = rho0;
    o->F[0] = fx;
    o->F[1] = fy;
    o->F[2] = fz;
    o->wall[0] = wall_x;
    o->wall[1] = wall_y;
    o->wall[2] = wall_z;
    std::memset(o->vel, 0, sizeof o->vel);
    o->mom.assign(10 * o->N, 0.0);
    o->fpost.assign((size_t)q * o->N, 0.0);
    o->type.assign(o->N, 0u);
    return o;
}

void orc_destroy(orc* o) { delete o; }

int orc_set_bc_velocity(orc* o, int k, double ux, double uy, double uz) {
    if (k < 0 || k > 3) return -1;
    o->vel[k][0] = ux;
    o->vel[k][1] = uy;
    o->vel[k][2] = uz;
    return 0;
}

int orc_set_node_types(orc* o, const uint32_t* t) {
    o->any_bc = false;
    if (!t) {
        std::fill(o->type.begin(), o->type.end(), 0u);
        return 0;
    }
    for (size_t j = 0; j < o->N; j++) {
        o->type[j] = t[j];
        if (t[j] == 0xFFFFFFFFu) return -1;
        const uint32_t kind = (t[j] >> 8) & 7u;
        if ((t[j] & 0xFFu) != 0xFFu && kind > 1) return -1;
        if (t[j]) o->any_bc = true;
    }
    return 0;
}

static int init_any(orc* o, int mode, const double* rho, const double* ux, const double* uy, const double* uz,
                    bool tgv, double u0, const double* uni) {
    if (mode != 0 && mode != 1) return -1;
    const double kx = 2 * M_PI / o->nx, ky = 2 * M_PI / o->ny, kz = 2 * M_PI / o->nz;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < o->nz; z++)
        for (int y = 0; y < o->ny; y++)
            for (int x = 0; x < o->nx; x++) {
                const size_t j = x + (size_t)o->nx * (y + (size_t)o->ny * z);
                double r, a, b, c, m[10];
                if (tgv) {
                    // cases/taylorGreen/flow_initialization.inc:1-4
                    r = o->rho0 + (1.0 / (16.0 / 3.0)) * o->rho0 * u0 * u0 *
                                      (std::cos(2 * kx * x) + std::cos(2 * ky * y)) * (std::cos(2 * kz * z) + 2.0);
                    a = u0 * std::sin(kx * x) * std::cos(ky * y) * std::cos(kz * z);
                    b = -u0 * std::cos(kx * x) * std::sin(ky * y) * std::cos(kz * z);
                    c = 0.0;
                } else if (uni) {
                    r = uni[0];
                    a = uni[1];
                    b = uni[2];
                    c = uni[3];
                } else {
                    r = rho[j];
                    a = ux[j];
                    b = uy[j];
                    c = uz[j];
                }
                init_node(o->q, mode, o->rho0, r, a, b, c, m);
                for (int k = 0; k < 10; k++) o->mom[(size_t)k * o->N + j] = m[k];
            }
    if (o->gtile[0]) o->mom0 = o->mom;
    return 0;
}

int orc_init_fields(orc* o, const double* rho, const double* ux, const double* uy, const double* uz, int mode) {
    return init_any(o, mode, rho, ux, uy, uz, false, 0.0, nullptr);
}
int orc_init_taylor_green(orc* o, double u0, int mode) {
    return init_any(o, mode, nullptr, nullptr, nullptr, nullptr, true, u0, nullptr);
}
int orc_init_uniform(orc* o, double rho, double ux, double uy, double uz, int mode) {
    const double u[4] = {rho, ux, uy, uz};
    return init_any(o, mode, nullptr, nullptr, nullptr, nullptr, false, 0.0, u);
}

// A population whose source lies outside the domain along a wall axis,
// read as the reference's kernel reads it (src/includeFiles/popLoad.inc:2-86
// with the wall predicates of src/includeFiles/interface.h:6-47), for the
// tile shape o->gtile: from the ghost plane of the face that wins in
// popLoad.inc - the last axis, in X, Y, Z order, along which the source
// leaves the puller's tile - of the tile holding the periodically wrapped
// source (sx, sy, sz). A face on a wall-axis domain boundary is written
// only at initialisation (gpuInitialization_pop, then gGhost <- fGhost,
// src/main.cuh:1016) and holds the source's initial population; any other
// face holds its current post-collision population.
static double ghost_value(const orc* o, int i, int x, int y, int z, int sx, int sy, int sz) {
    const int c[3] = {CX[i], CY[i], CZ[i]};
    const int p[3] = {x, y, z}, s[3] = {sx, sy, sz}, n[3] = {o->nx, o->ny, o->nz};
    int axis = -1;
    for (int a = 0; a < 3; a++) {
        const int l = p[a] % o->gtile[a] - c[a];
        if (l < 0 || l >= o->gtile[a]) axis = a;
    }
    const size_t N = o->N;
    const size_t src = sx + (size_t)o->nx * (sy + (size_t)o->ny * sz);
    bool stale = false;
    if (axis >= 0 && o->wall[axis]) {
        const int tile = s[axis] / o->gtile[axis], ntile = n[axis] / o->gtile[axis];
        stale = c[axis] > 0 ? tile == ntile - 1 : tile == 0;
    }
    if (!stale) return o->fpost[(size_t)i * N + src];
    if ((o->type[src] & 0xFFu) == 0xFFu) return NaN;
    double m[10], f[27];
    for (int k = 0; k < 10; k++) m[k] = o->mom0[(size_t)k * N + src];
    m[0] += o->rho0;
    reconstruct(o->q, m, f);
    return f[i];
}

static void one_step(orc* o) {
    const int nx = o->nx, ny = o->ny, nz = o->nz, q = o->q;
    const size_t N = o->N;
    double* fp = o->fpost.data();
    double* mom = o->mom.data();
    const uint32_t* type = o->type.data();
    // pass 1: post-collision populations of the previous step
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const size_t j = x + (size_t)nx * (y + (size_t)ny * z);
                double f[27];
                if ((type[j] & 0xFFu) == 0xFFu) {
                    for (int i = 0; i < q; i++) fp[(size_t)i * N + j] = NaN;
                    continue;
                }
                double m[10];
                for (int k = 0; k < 10; k++) m[k] = mom[(size_t)k * N + j];
                m[0] += o->rho0;
                reconstruct(q, m, f);
                for (int i = 0; i < q; i++) fp[(size_t)i * N + j] = f[i];
            }
    // pass 2: pull, moments (bulk or boundary), collide, store
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const size_t j = x + (size_t)nx * (y + (size_t)ny * z);
                const uint32_t t = type[j];
                if ((t & 0xFFu) == 0xFFu) continue;
                double f[27];
                for (int i = 0; i < q; i++) {
                    int sx = x - CX[i], sy = y - CY[i], sz = z - CZ[i];
                    bool out = false;
                    if (sx < 0 || sx >= nx) {
                        if (o->wall[0]) out = true;
                        sx = (sx + nx) % nx;
                    }
                    if (sy < 0 || sy >= ny) {
                        if (o->wall[1]) out = true;
                        sy = (sy + ny) % ny;
                    }
                    if (sz < 0 || sz >= nz) {
                        if (o->wall[2]) out = true;
                        sz = (sz + nz) % nz;
                    }
                    const size_t src = sx + (size_t)nx * (sy + (size_t)ny * sz);
                    if (!out) {
                        f[i] = fp[(size_t)i * N + src];
                    } else if (!o->gtile[0]) {
                        f[i] = NaN;
                    } else {
                        f[i] = ghost_value(o, i, x, y, z, sx, sy, sz);
                    }
                }
                double m[10];
                if (t == 0u) {
                    double r = 0, j3[3] = {0, 0, 0}, S[6] = {0, 0, 0, 0, 0, 0};
                    for (int i = 0; i < q; i++) {
                        double h[6];
                        herm(i, h);
                        r += f[i];
                        j3[0] += f[i] * CX[i];
                        j3[1] += f[i] * CY[i];
                        j3[2] += f[i] * CZ[i];
                        for (int k = 0; k < 6; k++) S[k] += f[i] * h[k];
                    }
                    m[0] = r;
                    m[1] = j3[0] / r;
                    m[2] = j3[1] / r;
                    m[3] = j3[2] / r;
                    for (int k = 0; k < 6; k++) m[4 + k] = S[k] / r;
                } else {
                    const uint32_t kind = (t >> 8) & 7u;
                    const int vi = (int)((t >> 11) & 3u);
                    const double zero[3] = {0, 0, 0};
                    const double* u = kind == 1 ? o->vel[vi] : zero;
                    const uint32_t known = known_mask_impl(q, t & 0xFFu);
                    if (!bc_solve(q, closure_q(q, t & 0xFFu), known, o->omega, u, f, m)) {
                        for (int k = 0; k < 10; k++) m[k] = NaN;
                    }
                }
                // scale (src/mlbm.cu:500-509): 3u, 4.5 m_aa, 9 m_ab
                double s[10] = {m[0], 3 * m[1], 3 * m[2], 3 * m[3], 4.5 * m[4], 9 * m[5], 9 * m[6],
                                4.5 * m[7], 9 * m[8], 4.5 * m[9]};
                collide(s, o->omega, o->F[0], o->F[1], o->F[2]);
                s[0] -= o->rho0;
                for (int k = 0; k < 10; k++) mom[(size_t)k * N + j] = s[k];
            }
}

int orc_step(orc* o, int nsteps) {
    for (int s = 0; s < nsteps; s++) one_step(o);
    return 0;
}

double orc_step_timed(orc* o, int nsteps) {
    auto t0 = std::chrono::steady_clock::now();
    orc_step(o, nsteps);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

void orc_get_moments(const orc* o, double* out) { std::memcpy(out, o->mom.data(), o->mom.size() * sizeof(double)); }
void orc_set_moments(orc* o, const double* in) {
    std::memcpy(o->mom.data(), in, o->mom.size() * sizeof(double));
    if (o->gtile[0]) o->mom0 = o->mom;
}

int orc_set_ghost_tiles(orc* o, int bx, int by, int bz) {
    if (bx == 0 && by == 0 && bz == 0) {
        o->gtile[0] = o->gtile[1] = o->gtile[2] = 0;
        o->mom0.clear();
        return 0;
    }
    if (bx <= 0 || by <= 0 || bz <= 0 || o->nx % bx || o->ny % by || o->nz % bz) return -1;
    o->gtile[0] = bx;
    o->gtile[1] = by;
    o->gtile[2] = bz;
    o->mom0 = o->mom;
    return 0;
}

void orc_reduce(const orc* o, double* mass, double* ke) {
    long double sm = 0, sk = 0;
    for (size_t j = 0; j < o->N; j++) {
        const double dr = o->mom[j];
        const double ux = o->mom[o->N + j], uy = o->mom[2 * o->N + j], uz = o->mom[3 * o->N + j];
        sm += dr;
        // src/reduction.cu:147: rho (Ux^2+Uy^2+Uz^2) / (2 F_M_I_SCALE^2)
        sk += (long double)(o->rho0 + dr) * (ux * ux + uy * uy + uz * uz) / 18.0L;
    }
    *mass = (double)(sm / o->N);
    *ke = (double)(sk / o->N);
}

int64_t orc_count_nan(const orc* o) {
    int64_t n = 0;
    for (size_t j = 0; j < o->N; j++) {
        if ((o->type[j] & 0xFFu) == 0xFFu) continue;
        for (int k = 0; k < 10; k++)
            if (std::isnan(o->mom[(size_t)k * o->N + j])) n++;
    }
    return n;
}

void orc_set_threads(int n) { omp_set_num_threads(n); }
int orc_get_threads(void) { return omp_get_max_threads(); }

int orc_bc_node(int q, uint32_t t, double omega, const double u[3], const double* pop, double out[10]) {
    const uint32_t known = known_mask_impl(q, t & 0xFFu);
    return bc_solve(q, closure_q(q, t & 0xFFu), known, omega, u, pop, out) ? 0 : -1;
}
uint32_t orc_known_mask(int q, uint32_t t) { return known_mask_impl(q, t & 0xFFu); }
void orc_reconstruct(int q, const double m[10], double* pop) { reconstruct(q, m, pop); }
void orc_collide(double m[10], double omega, double fx, double fy, double fz) { collide(m, omega, fx, fy, fz); }

}  // extern "C"
