#include "bc_classes.h"

#include <cmath>
#include <cstring>
#include <map>

#include "../../include/mrlbm_nodetype.h"

namespace mrlbm {

namespace {
const int CX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
const int CY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
const int CZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};

double weight(int q, int i) {
    const int cls = i == 0 ? 0 : (i < 7 ? 1 : (i < 19 ? 2 : 3));
    if (q == 19) return cls == 0 ? 1.0 / 3 : (cls == 1 ? 1.0 / 18 : 1.0 / 36);
    return cls == 0 ? 8.0 / 27 : (cls == 1 ? 2.0 / 27 : (cls == 2 ? 1.0 / 54 : 1.0 / 216));
}

// second-order Hermite components (xx, xy, xz, yy, yz, zz) of population i
void hermite(int i, double h[6]) {
    const double t = 1.0 / 3.0;
    h[0] = CX[i] * CX[i] - t;
    h[1] = CX[i] * CY[i];
    h[2] = CX[i] * CZ[i];
    h[3] = CY[i] * CY[i] - t;
    h[4] = CY[i] * CZ[i];
    h[5] = CZ[i] * CZ[i] - t;
}
const double COEF[6] = {1, 2, 2, 1, 2, 1};
constexpr int BC_REC = 72;  // floats per class record (lattice.h BC_STRIDE)
}  // namespace

uint32_t known_mask(int q, uint32_t octants) {
    uint32_t mask = 0;
    for (int i = 0; i < q; i++) {
        const int s[3] = {-CX[i], -CY[i], -CZ[i]};
        bool all_solid = true;
        for (int k = 0; k < 8 && all_solid; k++) {
            const int corner[3] = {(k & 1) ? 1 : -1, (k & 2) ? 1 : -1, (k & 4) ? 1 : -1};
            bool touches = true;
            for (int a = 0; a < 3; a++)
                if (s[a] != 0 && s[a] != corner[a]) touches = false;
            if (touches && !((octants >> k) & 1u)) all_solid = false;
        }
        if (!all_solid) mask |= (1u << i);
    }
    return mask;
}

bool build_bc_inverse(int q, uint32_t known, double omega, const double u[3], double ainv[49], double* cond,
                      int closure_q) {
    double A[7][7] = {};
    const double uu[6] = {u[0] * u[0], u[0] * u[1], u[0] * u[2], u[1] * u[1], u[1] * u[2], u[2] * u[2]};
    // The closure runs over the known populations of lattice closure_q; the
    // incoming moments summed in the kernel run over all known populations
    // of the lattice (see closure_lattice).
    (void)q;
    for (int i = 0; i < closure_q; i++) {
        if (!((known >> i) & 1u)) continue;
        const double w = weight(closure_q, i);
        double h[6];
        hermite(i, h);
        const double cu = CX[i] * u[0] + CY[i] * u[1] + CZ[i] * u[2];
        double huu = 0;
        for (int j = 0; j < 6; j++) huu += COEF[j] * h[j] * uu[j];
        A[0][0] += w * (1.0 + 3.0 * cu + 4.5 * omega * huu);
        for (int j = 0; j < 6; j++) A[0][1 + j] += w * 4.5 * (1.0 - omega) * COEF[j] * h[j];
        for (int k = 0; k < 6; k++) {
            A[1 + k][0] += w * h[k] * (1.0 + 3.0 * cu);
            for (int j = 0; j < 6; j++) A[1 + k][1 + j] += w * h[k] * 4.5 * COEF[j] * h[j];
        }
    }
    // Gauss-Jordan with partial pivoting on [A | I]
    double M[7][14];
    double anorm = 0;
    for (int r = 0; r < 7; r++) {
        double rs = 0;
        for (int c = 0; c < 7; c++) {
            M[r][c] = A[r][c];
            rs += std::fabs(A[r][c]);
        }
        anorm = std::fmax(anorm, rs);
        for (int c = 0; c < 7; c++) M[r][7 + c] = (r == c) ? 1.0 : 0.0;
    }
    for (int c = 0; c < 7; c++) {
        int piv = c;
        for (int r = c + 1; r < 7; r++)
            if (std::fabs(M[r][c]) > std::fabs(M[piv][c])) piv = r;
        if (std::fabs(M[piv][c]) < 1e-13) return false;
        if (piv != c)
            for (int k = 0; k < 14; k++) std::swap(M[c][k], M[piv][k]);
        const double d = M[c][c];
        for (int k = 0; k < 14; k++) M[c][k] /= d;
        for (int r = 0; r < 7; r++) {
            if (r == c) continue;
            const double f = M[r][c];
            if (f == 0.0) continue;
            for (int k = 0; k < 14; k++) M[r][k] -= f * M[c][k];
        }
    }
    double inorm = 0;
    for (int r = 0; r < 7; r++) {
        double rs = 0;
        for (int c = 0; c < 7; c++) {
            ainv[7 * r + c] = M[r][7 + c];
            rs += std::fabs(M[r][7 + c]);
        }
        inorm = std::fmax(inorm, rs);
    }
    if (cond) *cond = anorm * inorm;
    return true;
}

int closure_lattice(int q, uint32_t octants) {
    const bool face = octants == NORTH || octants == SOUTH || octants == WEST || octants == EAST ||
                      octants == FRONT || octants == BACK;
    return (q == 27 && face) ? 19 : q;
}

bool class_record(int q, uint32_t octants, double omega, const double u[3], float* rec) {
    const uint32_t known = known_mask(q, octants);
    double ainv[49];
    if (!build_bc_inverse(q, known, omega, u, ainv, nullptr, closure_lattice(q, octants))) return false;
    std::memset(rec, 0, sizeof(float) * BC_REC);
    std::memcpy(&rec[0], &known, 4);
    rec[1] = (float)u[0];
    rec[2] = (float)u[1];
    rec[3] = (float)u[2];
    for (int k = 0; k < 49; k++) rec[4 + k] = (float)ainv[k];
    // rest-state split of the right-hand side (populations travel
    // as f_i - w_i RHO_0): W_K, c = sum_K w_i H_i / W_K, and the rest
    // solution y_rest = A^-1 (1, c), d0 = W_K y_rest,0 - 1
    double wk = 0, wh[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < q; i++) {
        if (!((known >> i) & 1u)) continue;
        double h[6];
        hermite(i, h);
        wk += weight(q, i);
        for (int j = 0; j < 6; j++) wh[j] += weight(q, i) * h[j];
    }
    double brest[7] = {1.0, wh[0] / wk, wh[1] / wk, wh[2] / wk, wh[3] / wk, wh[4] / wk, wh[5] / wk};
    double yrest[7];
    for (int k = 0; k < 7; k++) {
        yrest[k] = 0;
        for (int j = 0; j < 7; j++) yrest[k] += ainv[7 * k + j] * brest[j];
    }
    rec[53] = (float)wk;
    for (int j = 0; j < 6; j++) rec[54 + j] = (float)brest[1 + j];
    rec[60] = (float)(wk * yrest[0] - 1.0);
    for (int k = 0; k < 7; k++) rec[61 + k] = (float)yrest[k];
    return true;
}

std::string build_class_table(int q, double omega, const double bc_vel[4][3], const uint32_t* types, size_t n,
                              BcClassTable& out) {
    out = BcClassTable();
    out.class_of.assign(n, 0);
    out.keys.push_back(0);
    out.table.assign(BC_REC, 0.0f);  // class 0 unused (bulk)
    if (!types) return "";
    std::map<uint32_t, int> ids;
    for (size_t j = 0; j < n; j++) {
        const uint32_t w = types[j];
        if (w == BULK) continue;
        if (w == MISSING_DEFINITION) return "node type MISSING_DEFINITION left unresolved (classify voxels first)";
        const uint32_t oct = w & MRLBM_NT_OCTANT_MASK;
        if (oct == SOLID_NODE) {
            out.class_of[j] = 255;
            out.has_solid = true;
            continue;
        }
        const uint32_t key = w & (MRLBM_NT_OCTANT_MASK | MRLBM_NT_KIND_MASK | MRLBM_NT_VEL_MASK);
        auto it = ids.find(key);
        int id;
        if (it == ids.end()) {
            const uint32_t kind = (key & MRLBM_NT_KIND_MASK) >> MRLBM_NT_KIND_SHIFT;
            const int vi = (int)((key & MRLBM_NT_VEL_MASK) >> MRLBM_NT_VEL_SHIFT);
            double u[3] = {0, 0, 0};
            if (kind == 1) {
                u[0] = bc_vel[vi][0];
                u[1] = bc_vel[vi][1];
                u[2] = bc_vel[vi][2];
            } else if (kind != 0) {
                return "unsupported boundary-condition kind " + std::to_string(kind) +
                       " (only no-slip and fixed-velocity walls)";
            }
            if ((int)ids.size() >= 254) return "more than 254 distinct boundary classes";
            float rec[BC_REC];
            if (!class_record(q, oct, omega, u, rec))
                return "singular boundary system for node type " + std::to_string(key);
            id = (int)ids.size() + 1;
            ids[key] = id;
            out.keys.push_back(key);
            out.table.insert(out.table.end(), rec, rec + BC_REC);
        } else {
            id = it->second;
        }
        out.class_of[j] = (uint8_t)id;
        out.has_bc = true;
    }
    if (out.has_solid) out.has_bc = true;
    return "";
}

void build_octant_table(int q, double omega, std::vector<float>& table, uint32_t singular[8]) {
    table.assign(256 * BC_REC, 0.0f);
    for (int w = 0; w < 8; w++) singular[w] = 0;
    const double u[3] = {0, 0, 0};
    for (uint32_t oct = 1; oct < 255; oct++)
        if (!class_record(q, oct, omega, u, &table[oct * BC_REC])) singular[oct >> 5] |= 1u << (oct & 31);
}

}  // namespace mrlbm
