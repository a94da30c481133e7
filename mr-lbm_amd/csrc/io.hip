// Files of a lattice: the versioned checkpoint (replaces the raw dumps of
// src/checkpoint.cu:129-382), the legacy-VTK field file
// (src/saveData.cu:303-406), the raw per-field .bin files
// (src/saveData.cu:183-282) and the <ID>_info.txt simulation summary
// (src/saveData.cu:469-613) that the reference's post/ tools read.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/mrlbm.h"
#include "context.h"

using namespace mrlbm;

namespace {

// Checkpoint layout (little-endian): this header, then the stored moments
// (block layout, storage words, mom_bytes), then the six faces of the ghost
// set the next step reads (FP32, face_bytes[f] each).
struct CkHeader {
    char magic[8];      // "MRLBMCK2"
    uint32_t version;   // 2
    int32_t nx, ny, nz, q, bx, by, bz, storage, wall_x, wall_y, wall_z;
    uint32_t seed;      // stochastic-rounding key of FP16/BF16 storage
    uint32_t reserved;  // 0
    int64_t step;
    double tau, rho0, fx, fy, fz;
    uint64_t mom_bytes, face_bytes[6];
};
static_assert(sizeof(CkHeader) == MRLBM_CHECKPOINT_HEADER_BYTES, "checkpoint header layout");
constexpr char kMagic[8] = {'M', 'R', 'L', 'B', 'M', 'C', 'K', '2'};

// Payload sizes a lattice of these parameters has.
void payload_sizes(int nx, int ny, int nz, int q, int bx, int by, int bz, int storage, uint64_t& mom,
                   uint64_t face[6]) {
    const uint64_t ntiles = (uint64_t)(nx / bx) * (ny / by) * (nz / bz);
    const uint64_t qf = q == 19 ? 5 : 9;
    mom = (uint64_t)nx * ny * nz * 10 * storage_bytes(storage);
    face[0] = face[1] = (uint64_t)by * bz * qf * ntiles * 4;
    face[2] = face[3] = (uint64_t)bx * bz * qf * ntiles * 4;
    face[4] = face[5] = (uint64_t)bx * by * qf * ntiles * 4;
}

// Open `path` and check its header against the file's own structure:
// magic and version, a lattice the library can hold, payload sizes that
// lattice implies, and a file length of exactly header + payloads.
int open_checkpoint(const char* path, CkHeader& h, FILE** out) {
    *out = nullptr;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fail(MRLBM_EIO, "cannot open '%s'", path);
    auto bad = [&](const char* why) {
        std::fclose(fp);
        return fail(MRLBM_EIO, "'%s' is not a valid checkpoint: %s", path, why);
    };
    if (std::fread(&h, sizeof h, 1, fp) != 1) return bad("shorter than its header");
    if (std::memcmp(h.magic, kMagic, 8) != 0 || h.version != 2) return bad("not a version-2 checkpoint");
    if (h.q != 19 && h.q != 27) return bad("velocity set is neither D3Q19 nor D3Q27");
    if (h.storage < MRLBM_STORE_FP32 || h.storage > MRLBM_STORE_BF16) return bad("unknown storage format");
    if (h.nx <= 0 || h.ny <= 0 || h.nz <= 0 || h.bx <= 0 || h.by <= 0 || h.bz <= 0 || h.nx % h.bx ||
        h.ny % h.by || h.nz % h.bz)
        return bad("lattice extents are not positive multiples of the tile");
    if ((double)h.nx * h.ny * h.nz > 4.0e10) return bad("lattice larger than any device holds");
    if (h.step < 0) return bad("negative step");
    uint64_t mom, face[6];
    payload_sizes(h.nx, h.ny, h.nz, h.q, h.bx, h.by, h.bz, h.storage, mom, face);
    if (h.mom_bytes != mom) return bad("moment payload size does not match the lattice");
    uint64_t total = sizeof h + mom;
    for (int f = 0; f < 6; f++) {
        if (h.face_bytes[f] != face[f]) return bad("ghost-plane payload size does not match the lattice");
        total += face[f];
    }
    if (std::fseek(fp, 0, SEEK_END) != 0) return bad("cannot seek");
    const long len = std::ftell(fp);
    if (len < 0 || (uint64_t)len != total) return bad("file length differs from header + payloads (truncated?)");
    if (std::fseek(fp, (long)sizeof h, SEEK_SET) != 0) return bad("cannot seek");
    *out = fp;
    return MRLBM_OK;
}

void be_float(FILE* fp, float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u = (u >> 24) | ((u >> 8) & 0xFF00u) | ((u << 8) & 0xFF0000u) | (u << 24);
    std::fwrite(&u, 4, 1, fp);
}

// "PATH/ID_<name><step>.<ext>" with the step zero-padded as
// getVarFilename pads it (src/saveData.cu:439-467): to six digits, and
// seven zeros for step 0.
std::string var_filename(const char* dir, const char* id, const char* name, int64_t step, const char* ext) {
    int zeros = 0;
    if (step != 0) {
        for (int64_t p = 10; step * p < 1000000; p *= 10) zeros++;
    } else {
        zeros = 6;
    }
    std::string s = std::string(dir) + "/" + id + "_" + name;
    s.append((size_t)zeros, '0');
    s += std::to_string(step);
    s += ext;
    return s;
}

}  // namespace

extern "C" {

int mrlbm_checkpoint_save(mrlbm_ctx* c, const char* path) {
    if (!c || !path) return fail(MRLBM_EINVAL, "null argument");
    if (!c->initialised) return fail(MRLBM_ESTATE, "lattice not initialised");
    try {
        HIPCHK(hipSetDevice(c->p.device));
        HIPCHK(hipStreamSynchronize(c->stream));
        CkHeader h{};
        std::memcpy(h.magic, kMagic, 8);
        h.version = 2;
        h.nx = c->p.nx;
        h.ny = c->p.ny;
        h.nz = c->p.nz;
        h.q = c->p.q;
        h.bx = c->bx;
        h.by = c->by;
        h.bz = c->bz;
        h.storage = c->p.storage;
        h.wall_x = c->p.wall_x;
        h.wall_y = c->p.wall_y;
        h.wall_z = c->p.wall_z;
        h.seed = c->p.seed;
        h.step = c->step;
        h.tau = c->p.tau;
        h.rho0 = c->p.rho0;
        h.fx = c->p.fx;
        h.fy = c->p.fy;
        h.fz = c->p.fz;
        payload_sizes(h.nx, h.ny, h.nz, h.q, h.bx, h.by, h.bz, h.storage, h.mom_bytes, h.face_bytes);
        // read everything back before touching the file
        std::vector<char> mom(h.mom_bytes);
        HIPCHK(hipMemcpy(mom.data(), c->d_mom, h.mom_bytes, hipMemcpyDeviceToHost));
        std::vector<std::vector<char>> faces(6);
        for (int f = 0; f < 6; f++) {
            faces[f].resize(h.face_bytes[f]);
            HIPCHK(hipMemcpy(faces[f].data(), c->d_ghost[c->cur][f], h.face_bytes[f], hipMemcpyDeviceToHost));
        }
        FILE* fp = std::fopen(path, "wb");
        if (!fp) return fail(MRLBM_EIO, "cannot open '%s' for writing", path);
        bool ok = std::fwrite(&h, sizeof h, 1, fp) == 1 && std::fwrite(mom.data(), 1, mom.size(), fp) == mom.size();
        for (int f = 0; f < 6 && ok; f++) ok = std::fwrite(faces[f].data(), 1, faces[f].size(), fp) == faces[f].size();
        if (std::fclose(fp) != 0) ok = false;
        if (!ok) return fail(MRLBM_EIO, "writing checkpoint '%s' failed", path);
    } catch (const std::bad_alloc&) {
        return fail(MRLBM_ENOMEM, "host memory for the checkpoint of '%s' exhausted", path);
    }
    return MRLBM_OK;
}

int mrlbm_checkpoint_info(const char* path, mrlbm_params* out, int64_t* step) {
    if (!path || !out || !step) return fail(MRLBM_EINVAL, "null argument");
    CkHeader h;
    FILE* fp;
    int rc = open_checkpoint(path, h, &fp);
    if (rc) return rc;
    std::fclose(fp);
    mrlbm_params p{};
    p.nx = h.nx;
    p.ny = h.ny;
    p.nz = h.nz;
    p.q = h.q;
    p.bx = h.bx;
    p.by = h.by;
    p.bz = h.bz;
    p.storage = h.storage;
    p.wall_x = h.wall_x;
    p.wall_y = h.wall_y;
    p.wall_z = h.wall_z;
    p.kernel = MRLBM_KERNEL_AUTO;
    p.tau = h.tau;
    p.rho0 = h.rho0;
    p.fx = h.fx;
    p.fy = h.fy;
    p.fz = h.fz;
    p.device = 0;
    p.seed = h.seed;
    *out = p;
    *step = h.step;
    return MRLBM_OK;
}

int mrlbm_checkpoint_load(mrlbm_ctx* c, const char* path) {
    if (!c || !path) return fail(MRLBM_EINVAL, "null argument");
    CkHeader h;
    FILE* fp;
    int rc = open_checkpoint(path, h, &fp);
    if (rc) return rc;
    auto reject = [&](const char* what) {
        std::fclose(fp);
        return fail(MRLBM_EINVAL, "checkpoint '%s' does not match the context: %s", path, what);
    };
    if (h.nx != c->p.nx || h.ny != c->p.ny || h.nz != c->p.nz) return reject("lattice extents");
    if (h.q != c->p.q) return reject("velocity set");
    if (h.bx != c->bx || h.by != c->by || h.bz != c->bz) return reject("tile shape (moment and ghost layout)");
    if (h.storage != c->p.storage) return reject("storage format");
    if (h.wall_x != c->p.wall_x || h.wall_y != c->p.wall_y || h.wall_z != c->p.wall_z) return reject("wall axes");
    if (h.rho0 != c->p.rho0) return reject("RHO_0 (stored moments are rho - RHO_0, ghosts f - w RHO_0)");
    if (h.tau != c->p.tau) return reject("tau");
    if (h.fx != c->p.fx || h.fy != c->p.fy || h.fz != c->p.fz) return reject("body force");
    if (h.seed != c->p.seed) return reject("stochastic-rounding seed");
    try {
        // whole payload into host memory first: a failed read leaves the
        // device state untouched
        std::vector<char> mom(h.mom_bytes);
        bool ok = std::fread(mom.data(), 1, mom.size(), fp) == mom.size();
        std::vector<std::vector<char>> faces(6);
        for (int f = 0; f < 6 && ok; f++) {
            faces[f].resize(h.face_bytes[f]);
            ok = std::fread(faces[f].data(), 1, faces[f].size(), fp) == faces[f].size();
        }
        std::fclose(fp);
        if (!ok) return fail(MRLBM_EIO, "reading checkpoint '%s' failed", path);
        HIPCHK(hipSetDevice(c->p.device));
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(c->d_mom, mom.data(), mom.size(), hipMemcpyHostToDevice));
        // both sets: the one read next and the one written next (boundary
        // slots that are never rewritten keep their saved contents)
        for (int f = 0; f < 6; f++)
            for (int s = 0; s < 2; s++)
                HIPCHK(hipMemcpy(c->d_ghost[s][f], faces[f].data(), faces[f].size(), hipMemcpyHostToDevice));
    } catch (const std::bad_alloc&) {
        return fail(MRLBM_ENOMEM, "host memory for reading '%s' exhausted", path);
    }
    c->cur = 0;
    c->step = h.step;
    c->initialised = true;
    c->version++;
    return MRLBM_OK;
}

int mrlbm_save_vtk(mrlbm_ctx* c, const char* path) {
    if (!c || !path) return fail(MRLBM_EINVAL, "null argument");
    try {
        std::vector<double> m(c->N * 10);
        int rc = mrlbm_read_moments(c, m.data());
        if (rc) return rc;
        FILE* fp = std::fopen(path, "wb");
        if (!fp) return fail(MRLBM_EIO, "cannot open '%s' for writing", path);
        std::fprintf(fp,
                     "# vtk DataFile Version 3.0\nLBM output (binary)\nBINARY\nDATASET STRUCTURED_POINTS\n"
                     "DIMENSIONS %d %d %d\nORIGIN 0 0 0\nSPACING 1 1 1\nPOINT_DATA %zu\n",
                     c->p.nx, c->p.ny, c->p.nz, c->N);
        std::fprintf(fp, "SCALARS rho float 1\nLOOKUP_TABLE default\n");
        for (size_t j = 0; j < c->N; j++) be_float(fp, (float)(c->p.rho0 + m[j]));
        std::fprintf(fp, "VECTORS velocity float\n");
        for (size_t j = 0; j < c->N; j++) {
            be_float(fp, (float)(m[c->N + j] / 3.0));
            be_float(fp, (float)(m[2 * c->N + j] / 3.0));
            be_float(fp, (float)(m[3 * c->N + j] / 3.0));
        }
        if (std::ferror(fp)) {
            std::fclose(fp);
            return fail(MRLBM_EIO, "writing '%s' failed", path);
        }
        if (std::fclose(fp) != 0) return fail(MRLBM_EIO, "writing '%s' failed", path);
    } catch (const std::bad_alloc&) {
        return fail(MRLBM_ENOMEM, "host memory for '%s' exhausted", path);
    }
    return MRLBM_OK;
}

int mrlbm_save_bin(mrlbm_ctx* c, const char* dir, const char* id) {
    if (!c || !dir || !id) return fail(MRLBM_EINVAL, "null argument");
    try {
        std::vector<double> m(c->N * 10);
        int rc = mrlbm_read_moments(c, m.data());
        if (rc) return rc;
        const char* names[4] = {"rho", "ux", "uy", "uz"};
        std::vector<float> v(c->N);
        for (int k = 0; k < 4; k++) {
            // saveMacr's arrays: rho = RHO_0 + stored, u = the stored 3u
            for (size_t j = 0; j < c->N; j++) v[j] = (float)(k == 0 ? c->p.rho0 + m[j] : m[(size_t)k * c->N + j]);
            const std::string path = var_filename(dir, id, names[k], c->step, ".bin");
            FILE* fp = std::fopen(path.c_str(), "wb");
            if (!fp) return fail(MRLBM_EIO, "cannot open '%s' for writing", path.c_str());
            const bool ok = std::fwrite(v.data(), sizeof(float), v.size(), fp) == v.size();
            if (std::fclose(fp) != 0 || !ok) return fail(MRLBM_EIO, "writing '%s' failed", path.c_str());
        }
    } catch (const std::bad_alloc&) {
        return fail(MRLBM_ENOMEM, "host memory for the field files exhausted");
    }
    return MRLBM_OK;
}

int mrlbm_save_info(mrlbm_ctx* c, const char* dir, const char* id, const char* case_name, double u_max,
                    int64_t save_steps, double mlups) {
    if (!c || !dir || !id) return fail(MRLBM_EINVAL, "null argument");
    const std::string path = std::string(dir) + "/" + id + "_info.txt";
    FILE* fp = std::fopen(path.c_str(), "w");
    if (!fp) return fail(MRLBM_EIO, "cannot open '%s' for writing", path.c_str());
    // the lines post/fileTreat.py reads take the value after the last space
    const char* rule = "--------------------------------------------------------------------------------\n";
    std::fprintf(fp, "---------------------------- SIMULATION INFORMATION ----------------------------\n");
    std::fprintf(fp, "      Simulation ID: %s\n", id);
    std::fprintf(fp, "       Velocity set: D3Q%d\n", c->p.q);
    std::fprintf(fp, "          Precision: float\n");
    std::fprintf(fp, "                 NX: %d\n", c->p.nx);
    std::fprintf(fp, "                 NY: %d\n", c->p.ny);
    std::fprintf(fp, "                 NZ: %d\n", c->p.nz);
    std::fprintf(fp, "           NZ_TOTAL: %d\n", c->p.nz);
    std::fprintf(fp, "                Tau: %.6e\n", c->p.tau);
    std::fprintf(fp, "               Umax: %.6e\n", u_max);
    std::fprintf(fp, "                 FX: %.6e\n", c->p.fx);
    std::fprintf(fp, "                 FY: %.6e\n", c->p.fy);
    std::fprintf(fp, "                 FZ: %.6e\n", c->p.fz);
    std::fprintf(fp, "         Save steps: %lld\n", (long long)save_steps);
    std::fprintf(fp, "             Nsteps: %lld\n", (long long)c->step);
    std::fprintf(fp, "              MLUPS: %.6e\n", mlups);
    std::fprintf(fp, "                 BX: %d\n", c->bx);
    std::fprintf(fp, "                 BY: %d\n", c->by);
    std::fprintf(fp, "                 BZ: %d\n", c->bz);
    std::fprintf(fp, "%s", rule);
    std::fprintf(fp, "\n------------------------------ BOUNDARY CONDITIONS -----------------------------\n");
    std::fprintf(fp, "            BC mode: Moment Based \n");
    std::fprintf(fp, "            BC type: %s\n", case_name ? case_name : "");
    std::fprintf(fp, "%s", rule);
    std::fprintf(fp, "\n--------------------------------- STORAGE --------------------------------------\n");
    std::fprintf(fp, "     Moment storage: %s\n",
                 c->p.storage == MRLBM_STORE_FP32 ? "fp32" : (c->p.storage == MRLBM_STORE_FP16 ? "fp16" : "bf16"));
    std::fprintf(fp, "%s", rule);
    const bool err = std::ferror(fp) != 0;
    if (std::fclose(fp) != 0 || err) return fail(MRLBM_EIO, "writing '%s' failed", path.c_str());
    return MRLBM_OK;
}

}  // extern "C"
