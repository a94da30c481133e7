/*
 * mrlbm.h - C ABI of the MI355X moment-representation LBM step.
 *
 * One context = one lattice (an "instance") resident in HBM: the 10 stored
 * moments per node, two sets of six ghost-interface planes, the node-type
 * map and the boundary-condition class table. Every call returns 0 on
 * success and a negative MRLBM_E* code on failure; mrlbm_last_error() gives
 * the text of the last failure on the calling thread. Nothing calls exit().
 * A context is thread-compatible (not thread-safe); all device work of a
 * context is ordered on its own HIP stream.
 *
 * Host arrays crossing this boundary are plain x-fastest arrays
 * (index = x + NX*(y + NY*z)), as written by the reference's saveMacr
 * (src/saveData.cu:7-260) and read by its bc_initialization loop
 * (src/lbmInitialization.cu:958-975).
 */
#ifndef MRLBM_H
#define MRLBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define MRLBM_OK          0
#define MRLBM_EINVAL     -1  /* bad argument or unsupported parameter set */
#define MRLBM_EHIP       -2  /* HIP runtime error */
#define MRLBM_ENOMEM     -3  /* device allocation failed / over budget */
#define MRLBM_EIO        -4  /* file error */
#define MRLBM_ESTATE     -5  /* call not valid in the current state */
#define MRLBM_ENUMERIC   -6  /* singular boundary system, NaN detected */

/* moment storage precision (arithmetic is always FP32). FP16 stores
 * value * 2^k (k = 12 for rho-RHO_0 and the second moments, 10 for U),
 * BF16 stores the value; both round stochastically (counter-based bits
 * keyed by node, step and seed). Ghost planes stay FP32. */
#define MRLBM_STORE_FP32  0
#define MRLBM_STORE_FP16  1
#define MRLBM_STORE_BF16  2

/* second-moment initialisation */
#define MRLBM_INIT_EQUILIBRIUM 0 /* m_ab = u_a u_b                                  */
#define MRLBM_INIT_REFERENCE   1 /* m_ab from f_eq(RHO_0,u)/rho - cs2 d_ab, as in
                                    src/lbmInitialization.cu:102-130              */

/* step kernel family */
#define MRLBM_KERNEL_AUTO   0
#define MRLBM_KERNEL_TILE   1 /* one node per thread, 3-D tile, LDS streaming   */
#define MRLBM_KERNEL_COLUMN 2 /* one (x,y) column per thread marching in z      */

typedef struct mrlbm_ctx mrlbm_ctx;

/* Lattice parameters. Replaces the compile-time constants of a case's
 * constants.inc/model.inc (reference src/var.h:184-186, src/definitions.h). */
typedef struct {
    int nx, ny, nz;        /* NX, NY, NZ                                        */
    int q;                 /* 19 (D3Q19) or 27 (D3Q27)                          */
    int bx, by, bz;        /* tile shape (8x8x8, 16x8x4, 32x4x4, 64x4x2, 16x8x8;
                              16x16x32, 16x16x64 with the column kernel); 0,0,0 =
                              library default: 16x8x8 for FP32 D3Q19 where it
                              divides the lattice, else 8x8x8 */
    int storage;           /* MRLBM_STORE_*                                     */
    int wall_x, wall_y, wall_z; /* 0 periodic (BC_*_PERIODIC), 1 wall (BC_*_WALL) */
    int kernel;            /* MRLBM_KERNEL_*                                    */
    double tau;            /* TAU; omega = 1/TAU (src/definitions.h:8)          */
    double rho0;           /* RHO_0                                             */
    double fx, fy, fz;     /* body force L_Fx/L_Fy/L_Fz (src/mlbm.cu:76-78)     */
    int device;            /* HIP device ordinal                                */
    unsigned int seed;     /* key of the stochastic rounding of FP16/BF16 storage */
} mrlbm_params;

typedef struct {
    int nx, ny, nz, q, bx, by, bz, storage, kernel;
    int64_t step;             /* steps taken since initialisation/resume     */
    int n_bc_classes;         /* distinct boundary classes in the node map   */
    int has_solid;            /* any SOLID_NODE present                      */
    double bytes_per_lu;      /* algorithmic HBM bytes per lattice update    */
    double device_bytes;      /* HBM held by the context                     */
} mrlbm_info;

/* Last error text of the calling thread ("" if none). */
const char* mrlbm_last_error(void);

/* Allocate a lattice on the device.
 * Replaces allocateDeviceMemory/interfaceMalloc (src/main.cuh:874, :455). */
int mrlbm_create(const mrlbm_params* p, mrlbm_ctx** out);

/* Release everything (src/main.cuh:74 interfaceFree). NULL is accepted. */
int mrlbm_destroy(mrlbm_ctx* ctx);

/* Wall-velocity table entry k (0..3), unscaled lattice units.
 * Replaces UX_BC/UY_BC/UZ_BC[4] of a case's constants.inc. Call before
 * mrlbm_set_node_types. */
int mrlbm_set_bc_velocity(mrlbm_ctx* ctx, int k, double ux, double uy, double uz);

/* Node-type map, x-fastest, reference encoding (include/mrlbm_nodetype.h).
 * Replaces hostInitialization_nodeType + its H2D copy
 * (src/lbmInitialization.cu:958-975, src/main.cuh:996-1008).
 * NULL = every node BULK (fully periodic lattice). */
int mrlbm_set_node_types(mrlbm_ctx* ctx, const uint32_t* node_types);

/* No-slip geometry classified on the device: the lattice's wall axes
 * (params wall_x/y/z) plus solid voxels, from a host mask (x-fastest, N
 * bytes, nonzero = solid; NULL for none) and/or spheres (nspheres records
 * cx, cy, cz, r; a node is solid when its squared distance to a centre is
 * at most r^2). Node types and boundary rules are those that
 * mrlbm_voxelize_spheres + mrlbm_classify + mrlbm_set_node_types give, so
 * the stepped fields are identical, without host passes over the lattice.
 * Replaces the host triple loops of hostInitialization_nodeType and
 * define_voxel_bc (src/lbmInitialization.cu:958-975, :1107-1227) for
 * wall and voxel cases without moving walls. */
int mrlbm_set_geometry(mrlbm_ctx* ctx, const uint8_t* solid, int nspheres, const double* spheres);

/* Initial macroscopic fields (x-fastest, unscaled): moments and both ghost
 * sets. Replaces gpuInitialization_mom + gpuInitialization_pop
 * (src/lbmInitialization.cu:28-218, :220-335). Resets the step counter. */
int mrlbm_init_fields(mrlbm_ctx* ctx, const double* rho, const double* ux,
                      const double* uy, const double* uz, int second_moment_mode);

/* Built-in initial states evaluated in FP64 on the host (no N-sized caller
 * arrays): uniform state, and the Taylor-Green vortex of
 * cases/taylorGreen/flow_initialization.inc:1-4 with k = 2*pi/NX. */
int mrlbm_init_uniform(mrlbm_ctx* ctx, double rho, double ux, double uy, double uz,
                       int second_moment_mode);
int mrlbm_init_taylor_green(mrlbm_ctx* ctx, double u0, int second_moment_mode);

/* Advance nsteps time steps (stream-ordered, asynchronous).
 * Replaces the step loop body gpuMomCollisionStream + swapGhostInterfaces
 * (src/main.cu:223-230, src/mlbm.cu:3-714); no host sync per step. */
int mrlbm_step(mrlbm_ctx* ctx, int nsteps);

/* Advance nsteps steps bracketed by HIP events on the context stream;
 * *elapsed_ms = event time (src/main.cuh:56-67 recordElapsedTime). */
int mrlbm_time_steps(mrlbm_ctx* ctx, int nsteps, double* elapsed_ms);

/* Wait for all work of the context. */
int mrlbm_synchronize(mrlbm_ctx* ctx);

/* Stored moments as doubles, x-fastest, 10 planes of N values in the
 * reference order and scaling: rho-RHO_0, 3ux, 3uy, 3uz, 4.5mxx, 9mxy,
 * 9mxz, 4.5myy, 9myz, 4.5mzz (src/arrayIndex.h:8-17,186). */
int mrlbm_read_moments(mrlbm_ctx* ctx, double* out10n);

/* Overwrite the stored moments (same format) and rebuild both ghost sets
 * from them, as gpuInitialization_pop does. */
int mrlbm_write_moments(mrlbm_ctx* ctx, const double* in10n);

/* FP64 reductions over all nodes (solid included), divided by N:
 * mean(rho - RHO_0) and mean(rho |u|^2 / 2).
 * Replaces mean_rho / totalKineticEnergy (src/auxFunctions.cu:8-67,
 * src/treatData.cu:105-165, src/reduction.cu:99-155). */
int mrlbm_reduce(mrlbm_ctx* ctx, double* mean_mass, double* mean_ke);
/* Same sums, and the number of nodes whose stored rho - RHO_0 or u is not
 * finite (a NaN or Inf anywhere also makes the sums non-finite). */
int mrlbm_reduce_checked(mrlbm_ctx* ctx, double* mean_mass, double* mean_ke, int64_t* nonfinite_nodes);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid
 * hipMalloc). mrlbm_reduce_checked without host synchronisation and without
 * the context's partial-sum buffer: enqueues on the context's stream the
 * same FP64 sums (same kernels, same block count, so the same bits) and
 * writes d_out3[0] = mean(rho - RHO_0), d_out3[1] = mean(rho |u|^2 / 2),
 * d_out3[2] = the number of non-finite nodes (as a double) to the caller's
 * device buffer, which must be device memory on the context's device. The
 * partial sums use a per-device array of 1024 FP64 triples (24 KiB) held
 * until the process ends; the sums stride over the nodes, so every lattice
 * fits. Calls on one device share that array: calls ordered on one stream
 * are safe, calls that overlap on different streams of one device race.
 * The results are ready once the context's stream reaches them
 * (mrlbm_synchronize). */
int mrlbm_reduce_checked_noalloc(mrlbm_ctx* ctx, double* d_out3);

/* FP64 convergence residual of the velocity field: ||u - u_prev|| / ||u||
 * (L2 over all nodes), u_prev being the velocity at the previous call; the
 * first call stores the reference (12 B/node, allocated then) and returns
 * -1. A steady state drives it to zero. No counterpart in the reference,
 * whose steady-state cases run a fixed N_STEPS. */
int mrlbm_velocity_residual(mrlbm_ctx* ctx, double* residual);

/* Diagnostics (device-side; replace the reference's treatData family,
 * src/treatData.cu, which copies one node at a time to the host). */
/* Stored-convention moments (10 per point, FP64) at n nodes given as x,y,z
 * triples: probe points and lines. Replaces probeExport and
 * velocityProfile (src/treatData.cu:501-584, :383-498). */
int mrlbm_sample(mrlbm_ctx* ctx, int n, const int32_t* xyz, double* out10n);
/* Mean over all nodes of each stored moment, FP64.
 * Replaces mean_moment (src/treatData.cu:36-102). */
int mrlbm_moment_means(mrlbm_ctx* ctx, double out[10]);
/* Running mean flow: accumulate the current rho, ux, uy, uz of every node
 * (FP64 sums, 32 B/node allocated on first use), reset, and read the mean
 * (4 x N, x-fastest) with the sample count. Replaces the MEAN_FLOW
 * accumulation of fMom_mean (src/main.cu, src/treatData.cu). */
int mrlbm_mean_flow_accumulate(mrlbm_ctx* ctx);
int mrlbm_mean_flow_reset(mrlbm_ctx* ctx);
int mrlbm_mean_flow_read(mrlbm_ctx* ctx, double* out4n, int64_t* count);
/* Turbulent kinetic energy of the current state against the running mean:
 * mean over nodes of rho |u - mean u|^2 / 2.
 * Replaces turbulentKineticEnergy (src/treatData.cu:236-296). */
int mrlbm_turbulent_ke(mrlbm_ctx* ctx, double* tke);

/* Checkpoint = parameters + step + stored moments + the ghost set read
 * next, one little-endian file. Replaces saveSimCheckpoint /
 * loadSimCheckpoint / operateSimCheckpoint (src/checkpoint.cu:372-382,
 * :353-369, :129-316), whose state is the same (fMom, fGhost, step).
 * Layout: a MRLBM_CHECKPOINT_HEADER_BYTES header
 *   char magic[8] = "MRLBMCK2"; uint32 version = 2;
 *   int32 nx, ny, nz, q, bx, by, bz, storage, wall_x, wall_y, wall_z;
 *   uint32 seed, reserved; int64 step; double tau, rho0, fx, fy, fz;
 *   uint64 mom_bytes, face_bytes[6];
 * then mom_bytes of moments (block layout, storage words: FP16/BF16 state
 * is saved bit-exactly) and face_bytes[f] of ghost face f (FP32, f_i -
 * w_i RHO_0). Load rejects (MRLBM_EIO) a file whose sizes do not follow
 * from its lattice or whose length is not header + payloads, and
 * (MRLBM_EINVAL) one whose lattice, tile, storage, walls, RHO_0, tau,
 * force or seed differ from the context's; it reads the whole file before
 * it changes the device state. Node types are not part of the state: set
 * them on the context as before the save. A lattice stepped in a batch
 * may be saved and loaded between batch steps. */
#define MRLBM_CHECKPOINT_HEADER_BYTES 168
int mrlbm_checkpoint_save(mrlbm_ctx* ctx, const char* path);
int mrlbm_checkpoint_load(mrlbm_ctx* ctx, const char* path);
/* Host only: validate a checkpoint file's structure and return the
 * parameters it was saved with (kernel = AUTO, device = 0) and its step. */
int mrlbm_checkpoint_info(const char* path, mrlbm_params* out, int64_t* step);

/* Legacy-VTK STRUCTURED_POINTS file (big-endian float rho, velocity),
 * as saveVarVTK (src/saveData.cu:303-406). */
int mrlbm_save_vtk(mrlbm_ctx* ctx, const char* path);

/* Raw field files for the reference's post-processing tools
 * (post/fileTreat.py): dir/ID_rho<step>.bin, _ux, _uy, _uz, float32,
 * x-fastest, step zero-padded as getVarFilename does; rho = RHO_0 +
 * stored, u = the stored 3u, as saveMacr + saveVarBin write them
 * (src/saveData.cu:65-70, :183-228, :262-282). The directory must exist. */
int mrlbm_save_bin(mrlbm_ctx* ctx, const char* dir, const char* id);
/* dir/ID_info.txt in the reference's layout (getSimInfoString/saveSimInfo,
 * src/saveData.cu:469-613): lattice, tau, u_max (the case's U_MAX), force,
 * save interval, current step, MLUPS, tile, case name. */
int mrlbm_save_info(mrlbm_ctx* ctx, const char* dir, const char* id, const char* case_name, double u_max,
                    int64_t save_steps, double mlups);

int mrlbm_get_info(mrlbm_ctx* ctx, mrlbm_info* out);

/* Host-only diagnostic (no device work): the known-population mask and the
 * 7x7 inverse moment-matching matrix the kernels use for a boundary node of
 * type `node_type` (relaxation omega, wall velocity u). The reference has no
 * counterpart: its case files hold the closed-form solutions of this system
 * (e.g. cases/lidDrivenCavity_3D/bc_definition.inc). */
int mrlbm_bc_class_info(int q, uint32_t node_type, double omega, const double u[3], uint32_t* known,
                        double* ainv49);

/* ---------------------------------------------------------------------
 * Batches (no reference counterpart: the reference steps one lattice per
 * launch with a device sync per step, src/main.cu:223-230). A batch
 * advances many lattices of the same type (q, tile shape, storage, device,
 * tile kernel) with one launch per step through a block -> (lattice, tile)
 * work table. Members are ordinary contexts: they adopt the batch's stream,
 * keep their own parameters, node types and state, and are read back,
 * reduced, checkpointed or re-initialised with the mrlbm_* calls above;
 * mrlbm_step on a member is refused. Results are bit-identical to stepping
 * each lattice alone.
 */
typedef struct mrlbm_batch mrlbm_batch;
int mrlbm_batch_create(int device, int q, int bx, int by, int bz, int storage, mrlbm_batch** out);
int mrlbm_batch_add(mrlbm_batch* b, mrlbm_ctx* ctx);
int mrlbm_batch_remove(mrlbm_batch* b, mrlbm_ctx* ctx);
/* 0 (default): one workgroup per work item; n > 0: at most n workgroups
 * walking the work table (persistent launch) */
int mrlbm_batch_set_grid(mrlbm_batch* b, int max_blocks);
/* Advance every member by nsteps. graph_steps = 0: plain launches;
 * graph_steps = g > 0: a hipGraph of g steps is captured (once per
 * membership) and replayed nsteps/g times (nsteps must be a multiple). */
int mrlbm_batch_step(mrlbm_batch* b, int nsteps, int graph_steps);
/* Same, bracketed by HIP events on the batch stream; *elapsed_ms. A graph
 * that does not exist yet is captured (not run) before the timed region;
 * nsteps not a multiple of graph_steps is refused before anything runs. */
int mrlbm_batch_time_steps(mrlbm_batch* b, int nsteps, int graph_steps, double* elapsed_ms);
int mrlbm_batch_synchronize(mrlbm_batch* b);
int mrlbm_batch_size(mrlbm_batch* b, int* members, int64_t* nodes);
int mrlbm_batch_destroy(mrlbm_batch* b);
/* Autotuner: for each candidate tile shape (shapes = 3*nshapes ints) that
 * divides every lattice extent (dims = 3*ninst ints), time `steps` batched
 * graph steps of all-bulk lattices; *best = index of the fastest,
 * ms_out[s] = its time (-1 if the shape does not fit). */
int mrlbm_batch_autotune(int device, int q, int storage, int nshapes, const int* shapes, int ninst, const int* dims,
                         int steps, int* best, double* ms_out);
/* Same, with the lattices' wall axes (walls = 3*ninst ints, nonzero = wall;
 * null = all periodic): wall lattices get device-classified domain walls,
 * so the boundary tiles are timed with the boundary body. */
int mrlbm_batch_autotune_walls(int device, int q, int storage, int nshapes, const int* shapes, int ninst,
                               const int* dims, const int* walls, int steps, int* best, double* ms_out);

/* ---------------------------------------------------------------------
 * Geometry (host): node-type maps in the reference encoding.
 * mrlbm_classify: domain walls on the wall axes, SOLID_NODE for solid
 *   voxels, and for fluid nodes next to solid voxels the octants of every
 *   solid neighbour direction (bc_id, src/lbmInitialization.cu:1080-1227;
 *   walls and voxels combine by OR; no wrap across a wall axis).
 *   solid may be NULL. Arrays are x-fastest.
 * mrlbm_set_moving_wall: non-solid nodes of the domain face `side` (-1/+1)
 *   of `axis` get BC_VELOCITY_WALL with velocity-table index vel_index.
 * mrlbm_voxelize_spheres: solid[] = 1 inside any sphere (cx,cy,cz,r) x nsph.
 * mrlbm_read_voxel_csv: the reference's voxel file (read_xyz_file,
 *   src/lbmInitialization.cu:978-1039): "x,y,z" per line; returns the
 *   number of voxels set or a negative status. */
int mrlbm_classify(int nx, int ny, int nz, int wall_x, int wall_y, int wall_z, const uint8_t* solid,
                   uint32_t* types);
int mrlbm_set_moving_wall(int nx, int ny, int nz, int axis, int side, int vel_index, uint32_t* types);
int mrlbm_voxelize_spheres(int nx, int ny, int nz, int nsph, const double* spheres, uint8_t* solid);
int mrlbm_read_voxel_csv(const char* path, int nx, int ny, int nz, uint8_t* solid);

/* ---------------------------------------------------------------------
 * Serving (no reference counterpart): a request loop over independent duct
 * solves (x periodic, walls on y and z, voxelised spheres, body force gx,
 * viscosity nu, step budget) with continuous batching. Requests become
 * visible at arrival_s (realtime = 1) or immediately (realtime = 0), are
 * admitted when the HBM budget and max_batch allow, keep their fields
 * resident, advance together one launch per step in rounds of at most
 * quantum_steps (graph-captured when graph_steps divides a full round) and
 * leave exactly at their step budget. Failures are per request (status <
 * 0): after every round an FP64 reduction of each member counts non-finite
 * nodes, and a member with any leaves the batch failed while the others
 * continue. Percentiles are over the requests that completed. */
typedef struct {
    double arrival_s;           /* in: arrival time, seconds from start         */
    int nx, ny, nz;             /* in: lattice (multiples of the tile)          */
    int nspheres;               /* in: 0..8                                     */
    double sphere[8][4];        /* in: centre x, y, z and radius                */
    double gx;                  /* in: body force along x                       */
    double nu;                  /* in: viscosity, tau = 0.5 + 3 nu              */
    int steps;                  /* in: step budget                              */
    unsigned int seed;          /* in: stochastic-rounding key                  */
    double* out_moments;        /* in: optional 10*N buffer for the final state */
    double start_s, finish_s;   /* out: admission / completion times            */
    double latency_s;           /* out: completion - arrival                    */
    double mass0, mass1;        /* out: FP64 mean(rho - 1) before / after       */
    double ke0, ke1;            /* out: FP64 mean kinetic energy before / after */
    int status;                 /* out: 0 done; < 0 failed, an MRLBM_E* code:
                                   EINVAL (extents not a multiple of the tile,
                                   0 or > 8 spheres, no steps, nu <= 0), ENOMEM
                                   (larger than the whole budget, or allocation
                                   failed), ENUMERIC (the lattice turned
                                   non-finite: the body force is not checked,
                                   so a non-finite one ends here)              */
} mrlbm_request;

typedef struct {
    int device, q, storage;
    int bx, by, bz;             /* 0 = 8x8x8                                    */
    double hbm_budget_bytes;
    int quantum_steps;          /* 0 = 100                                      */
    int graph_steps;            /* 0 = plain launches                           */
    int max_batch;              /* 0 = no limit (4095)                          */
    int realtime;
} mrlbm_serve_params;

typedef struct {
    double wall_s, busy_s;      /* loop wall time, time inside batch rounds     */
    double lattice_updates;
    double mlups, mlups_busy;   /* aggregate over wall / busy time              */
    double p50_s, p99_s, max_s; /* latency percentiles                          */
    double peak_bytes;          /* peak admitted state                          */
    int64_t rounds;
    int64_t failed;             /* requests that ended with status < 0          */
} mrlbm_serve_stats;

/* The seeded duct request stream (Poisson arrivals at `rate`/s; Nx in
 * [64,256], Ny = Nz in [32,128], multiples of 16; 1..8 spheres of radius
 * U[3, Ny/6]; gx = 10^U[-7,-5]; nu = U[0.01,0.1]; steps = U{500..5000}). */
int mrlbm_generate_duct_requests(uint64_t seed, int n, double rate, mrlbm_request* out);
int mrlbm_serve_run(const mrlbm_serve_params* params, mrlbm_request* reqs, int nreq, mrlbm_serve_stats* stats);
/* The same serving loop on a simulated device (no GPU needed): a virtual
 * clock, rounds of q steps over N batched nodes taking N q / rate_mlups us,
 * each admission admit_s seconds. Results (mass, moments) are not computed;
 * start/finish/latency and the statistics are. For testing and sizing the
 * scheduler. */
int mrlbm_serve_simulate(const mrlbm_serve_params* params, mrlbm_request* reqs, int nreq, double rate_mlups,
                         double admit_s, mrlbm_serve_stats* stats);

/* Measurement utility (no reference counterpart): time `iters` plain
 * device-to-device copies of `bytes` on `device`; *ms_per_copy = mean event
 * time of one copy (it reads and writes `bytes` each). */
int mrlbm_copy_bench(int device, size_t bytes, int iters, double* ms_per_copy);

#ifdef __cplusplus
}
#endif

#endif /* MRLBM_H */
