/*
 * libmpmx - C ABI of the MI355X (gfx950) hot path of MPMWarpYDW.
 *
 * Every entry point replaces one Warp kernel launch (or one Warp runtime
 * facility) of the reference; the comment above each declaration cites the
 * reference file:line it replaces (paths relative to the MPMWarpYDW tree).
 *
 * Conventions
 *  - All pointers are device pointers (hipMalloc'd, or torch CUDA tensors'
 *    data_ptr()), unless named host_*.
 *  - Structure-of-arrays layout: a per-particle vec3 array of n items is three
 *    float planes [3][n]; a mat33 array is nine planes [9][n] in row-major
 *    component order; grid arrays hold G = gx*gy*gz nodes with node index
 *    (ix*gy + iy)*gz + iz, vec3 grid arrays are [3][G].
 *  - Labels are int32; "active" means activeLabel == 1.
 *  - stream is a hipStream_t (0 = the null stream).  Calls are asynchronous on
 *    that stream unless stated.  No entry point allocates device memory: the
 *    ones that need scratch take a caller-provided workspace whose size is
 *    returned by the matching *_workspace_bytes() query.
 *  - Return value: 0 on success, a negative MPMX_E* code for an argument error,
 *    or a positive hipError_t.  mpmx_last_error() returns a thread-local text.
 *  - Environment MPMX_DEBUG=1 synchronises and checks after every launch (the
 *    analogue of the reference's wp.config.verify_cuda, runMPMYDW.py:7).
 */
#ifndef MPMX_H
#define MPMX_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPMX_EINVAL (-1)   /* bad argument (null pointer, negative size, ...) */
#define MPMX_ERANGE (-2)   /* size beyond a documented maximum */

/* boundary-condition selector of mpmx_grid_update (simulationRoutines.py:124-228):
 * boundaryCondition "friction", "friction_gradual", "restitution", "absorbing",
 * "plane_strain_y", "plane_strain_y_global" */
#define MPMX_BC_NONE 0
#define MPMX_BC_FRICTION 1
#define MPMX_BC_FRICTION_GRADUAL 2
#define MPMX_BC_RESTITUTION 3
#define MPMX_BC_ABSORBING 4
#define MPMX_BC_PLANE_STRAIN_Y 5
#define MPMX_BC_PLANE_STRAIN_Y_GLOBAL 6

/* Device pointers of the per-particle state arrays (the MPMState/SimState/
 * XPBDState fields of simStates.py, SoA layout) for the fused substep. */
typedef struct mpmx_particles {
    int n;
    int32_t* activeLabel;
    int32_t* materialLabel;
    float* x;            /* vec3 */
    float* v;            /* vec3 */
    float* x_init_xpbd;  /* vec3, particle_x_initial_xpbd */
    float* v_init_xpbd;  /* vec3, particle_v_initial_xpbd */
    float* F;            /* mat33 */
    float* F_trial;      /* mat33 */
    float* C;            /* mat33, APIC affine matrix */
    float* stress;       /* mat33, Kirchhoff */
    float* cov;          /* float[6n] */
    float* mu;
    float* lam;
    float* ys;
    float* alpha;
    float* hardening;
    float* softening;
    float* density;
    float* strainCriteria;
    float* eta_shear;
    float* eta_bulk;
    float* acc_strain;
    float* damage;
    float* vol;
    float* mass;
    /* optional (may be NULL): a particle numbering, e.g. the particle's index
     * before the arrays were permuted into spatial order (mpmx_reorder_
     * permutation).  Inside a grid cell particles are summed in ascending
     * order[id] (ascending id when NULL), which makes the results independent
     * of where each particle sits in the arrays.  No reference counterpart. */
    int32_t* order;
    /* optional (may be NULL = scene 0): the scene of every particle when
     * several scenes advance together (mpmx_ensemble_*).  No reference
     * counterpart. */
    int32_t* scene;
} mpmx_particles;

/* One scene of an ensemble: its grid (gx*gy*gz nodes from the origin min*,
 * spacing dx of the step parameters).  No reference counterpart. */
typedef struct mpmx_scene {
    int gx, gy, gz;
    float minx, miny, minz;
} mpmx_scene;

/* Grid arrays of MPMState (simStates.py:117-121). */
typedef struct mpmx_grid {
    int gx, gy, gz;
    float* m;     /* [G] */
    float* v_in;  /* [3][G] */
    float* v_out; /* [3][G] */
} mpmx_grid;

/* Scalars of one substep (SimState/MPMState fields used by mpmSimulationStep). */
typedef struct mpmx_step_params {
    float dt, dx, inv_dx;
    float minx, miny, minz;
    float gravity_x, gravity_y, gravity_z;
    float rpic_damping;
    float eff;
    float damping_scale;      /* grid_v_damping_scale (applied when < 1) */
    float friction;           /* boundFriction */
    float gravity_magnitude;  /* |gravity_z| */
    int bc;                   /* MPMX_BC_* */
    int boundary_padding;
    int update_cov;
    int constitutive_model;   /* passed through, never read (mpmRoutines.py:253) */
    float restitution;        /* boundRestitution (MPMX_BC_RESTITUTION) */
    int vlc;                  /* volumetric_locking_correction after G2P; the fused
                                 substep keeps the cell sums in its workspace */
} mpmx_step_params;

const char* mpmx_last_error(void);

/* Launch-shape knobs, process-wide (no reference counterpart), set by the
 * autotuner.  MPMX_TUNE_XCD_TILES = 1 deals the tiles of the tile kernels to
 * the 8 XCDs in contiguous ranges (neighbouring tiles share an L2), 2 in runs
 * of 16 tiles in turn (each XCD gets tiles from the whole range), 0 in block
 * order.  MPMX_TUNE_SPLIT_STRESS = 1 runs the constitutive update of the
 * fused substep as its own particle-parallel kernel before a P2G-only tile
 * kernel, 0 inside the tile kernel.  MPMX_TUNE_G2P_TILES = 1 runs G2P per tile
 * with the tile's grid velocities staged in LDS, 2 likewise with two blocks
 * sharing each tile's particles, 0 one thread per particle with the grid read
 * through the caches.  MPMX_TUNE_SPLIT_STRESS takes 0 or 1.  Results do not depend on
 * any knob. */
#define MPMX_TUNE_XCD_TILES 0
#define MPMX_TUNE_SPLIT_STRESS 1
#define MPMX_TUNE_G2P_TILES 2
/* MPMX_TUNE_P2G_CHUNK: particles the P2G tile kernel stages per LDS round,
 * 0 = 512, 1 = 384, 2 = 256 (the smaller, the more workgroups per CU) */
#define MPMX_TUNE_P2G_CHUNK 3
#define MPMX_TUNE_COUNT 4
int mpmx_set_tuning(int knob, int value);
int mpmx_get_tuning(int knob);

/* ---- contexts, device buffers, captured substep sequences ----------------
 * For callers without a device-memory runtime of their own (the Python layer
 * uses torch tensors and skips these).  A context is one device and one HIP
 * stream (default priority); buffers are hipMalloc'd and zeroed at allocation
 * (the analogue of wp.zeros in SimState construction, simStates.py); upload
 * and download are synchronising (download mirrors wp.array.numpy()).
 * mpmx_graph_begin/_end capture everything issued on the context's stream in
 * between (e.g. ten mpmx_mpm_substep calls and one XPBD step) into a graph
 * that mpmx_graph_launch replays; no entry point allocates while capturing,
 * and every buffer clear inside the library is a kernel, not a memset. */
typedef struct mpmx_ctx mpmx_ctx;
typedef struct mpmx_graph mpmx_graph;
int mpmx_ctx_create(int device, mpmx_ctx** out);
int mpmx_ctx_destroy(mpmx_ctx* ctx);
void* mpmx_ctx_stream(mpmx_ctx* ctx);
int mpmx_buf_alloc(mpmx_ctx* ctx, size_t bytes, void** out);
int mpmx_buf_free(mpmx_ctx* ctx, void* p);
int mpmx_buf_upload(mpmx_ctx* ctx, void* dst, const void* host, size_t bytes);
int mpmx_buf_download(mpmx_ctx* ctx, void* host, const void* src, size_t bytes);
int mpmx_graph_begin(mpmx_ctx* ctx);
int mpmx_graph_end(mpmx_ctx* ctx, mpmx_graph** out);
int mpmx_graph_launch(mpmx_graph* graph, void* stream);
int mpmx_graph_destroy(mpmx_graph* graph);
int mpmx_version(void);
/* Layout of the structs of this header as this build sees them, for bindings
 * to check theirs against: out[0..] = sizeof(mpmx_particles), offsetof(
 * mpmx_particles, order), sizeof(mpmx_grid), sizeof(mpmx_step_params),
 * offsetof(mpmx_step_params, bc), offsetof(mpmx_particles, scene),
 * sizeof(mpmx_scene), sizeof(mpmx_xpbd_bounds), offsetof(mpmx_xpbd_bounds,
 * delta), offsetof(mpmx_xpbd_bounds, xpbd_count); returns the number of
 * values written (<= cap). */
int mpmx_struct_layout(long* out, int cap);
/* Launch profiler (no counterpart in the reference; replaces its wall-clock
 * prints, runMPMYDW.py:762): between begin and end, every kernel the library
 * issues carries a start/stop event pair that its own dispatch timestamps
 * (hipExtLaunchKernelGGL), so the times are the kernels' own (not while the
 * stream is being captured into a graph: those launches are not timed).
 * report() writes "name launches total_ms" lines. */
int mpmx_profile_begin(int capacity);
int mpmx_profile_mark(void* stream);
/* Queue a one-thread kernel that keeps the stream busy for the given wall
 * time (at most 1 s), so that the launches issued behind it reach the GPU
 * back to back and their event brackets time the kernels alone. */
int mpmx_profile_hold(void* stream, int microseconds);
int mpmx_profile_end(void);
int mpmx_profile_report(char* buf, size_t buflen);
/* device synchronisation used by the host wrappers (mirrors .numpy() syncs) */
int mpmx_stream_sync(void* stream);
/* fill n int32 / float words with a value (wp.array.zero_()/fill_()) */
int mpmx_fill_i32(int32_t* dst, long n, int32_t value, void* stream);
int mpmx_fill_f32(float* dst, long n, float value, void* stream);
/* wp.copy / wp.array.assign on float planes */
int mpmx_copy_f32(float* dst, const float* src, long n, void* stream);

/* ---- MPM kernels (utils/mpmRoutines.py) ---------------------------------- */

/* initialize_geostatic_F (mpmRoutines.py:52-124) or, with plane_strain = 1,
 * initialize_geostatic_F_planestrain (:128-203), as launched at runMPMYDW.py:415-438:
 * F = diag(exp eps) with eps the linear-elastic strain of the geostatic stress
 * sigma_v = -rho g (z_top - z), sigma_h = K0 sigma_v.  Overwrites F [9][n]. */
int mpmx_initialize_geostatic_F(int n, const float* x, float* F, const float* mu, const float* lam,
                                const float* density, float gravity, float z_top, float K0,
                                int plane_strain, void* stream);
/* deactivate_and_teleport_xpbd_below_datum, mpmRoutines.py:1446-1475 (launched at
 * the end of every big step, runMPMYDW.py:864-877) */
int mpmx_deactivate_and_teleport_xpbd_below_datum(int n, float* x, float* v, int32_t* activeLabel,
                                                  const int32_t* materialLabel, float z_datum,
                                                  float teleport_distance, void* stream);
/* compute_mu_lam_bulk_from_E_nu, mpmRoutines.py:206-226 */
int mpmx_compute_mu_lam_bulk(int n, const float* E, const float* nu, float* mu, float* lam,
                             float* bulk, void* stream);

/* compute_stress_from_F_trial (+ drucker_prager_return_mapping, kirchoff_stress,
 * wp.svd3), mpmRoutines.py:229-471; launched at simulationRoutines.py:15-45.
 * x, v, x_init, v_init: vec3[n]; F, F_trial, C, stress: mat33[n]. */
int mpmx_compute_stress_from_F_trial(int n, const int32_t* activeLabel, int32_t* materialLabel,
                                     const float* x, float* v, float* x_init, float* v_init,
                                     float* F, const float* F_trial, float* mu, float* lam,
                                     float* ys, const float* hardening, const float* softening,
                                     const float* density, const float* strainCriteria,
                                     float eff, const float* eta_shear, const float* eta_bulk,
                                     const float* C, float* stress, float* acc_strain,
                                     float* damage, int constitutive_model, float* alpha,
                                     void* stream);

/* p2g_apic_with_stress, mpmRoutines.py:485-569 (launch simulationRoutines.py:48-69),
 * as a deterministic binned reduction: particles are counting-sorted by base cell
 * into 4x4x4-cell tiles, each tile is reduced in LDS in a fixed order and the
 * 6x6x6-node tile partials are summed per node in a fixed order.
 * Writes grid_m[G] and grid_v_in[3][G] completely (no prior zeroing needed).
 * The workspace must hold mpmx_p2g_workspace_bytes(n, gx, gy, gz) bytes.
 * *host_out_of_grid (optional, synchronising when non-null) receives the number of
 * active MPM particles whose stencil left the grid; they are skipped (the
 * reference has no bounds check there). */
size_t mpmx_p2g_workspace_bytes(int n, int gx, int gy, int gz);
int mpmx_p2g_apic_with_stress(int n, const int32_t* activeLabel, const int32_t* materialLabel,
                              const float* stress, const float* x, const float* v,
                              const float* C, const float* vol, const float* mass, float dx,
                              float inv_dx, float minx, float miny, float minz,
                              float rpic_damping, int gx, int gy, int gz, float* grid_m,
                              float* grid_v_in, float dt, void* workspace, size_t ws_bytes,
                              int* host_out_of_grid, void* stream);

/* p2g_apic_with_stress followed by steps 4-6 of mpmSimulationStep
 * (simulationRoutines.py:48-140) in the same pass over the grid: also writes
 * grid_v_out (normalisation + gravity, damping when damping_scale < 1, then the
 * boundary condition bc), so no separate zeroing or grid launch is needed.
 * Same workspace as mpmx_p2g_apic_with_stress. */
int mpmx_p2g_grid_update(int n, const int32_t* activeLabel, const int32_t* materialLabel,
                         const float* stress, const float* x, const float* v, const float* C,
                         const float* vol, const float* mass, float dx, float inv_dx, float minx,
                         float miny, float minz, float rpic_damping, int gx, int gy, int gz,
                         float* grid_m, float* grid_v_in, float* grid_v_out, float gxv, float gyv,
                         float gzv, float dt, float damping_scale, int bc, float friction,
                         float gravity_magnitude, int boundary_padding, float restitution, void* workspace,
                         size_t ws_bytes, void* stream);

/* mpmSimulationStep (simulationRoutines.py:6-250: stress, P2G, grid update,
 * G2P) as three passes over the state: a tile pass running the constitutive
 * update of every particle and its deterministic P2G reduction, a node pass
 * for the grid, and a tile pass for G2P that also bins each particle for the
 * next substep.  Results equal compute_stress_from_F_trial + p2g + grid update
 * + g2p.  Particles whose stencil leaves the grid get the constitutive update
 * and no transfer.  Call mpmx_substep_bin() once before the first substep and
 * again after any change of x, labels or activity other than this function's
 * own (the XPBD step between substeps restores MPM positions and needs none).
 * Workspace: mpmx_substep_workspace_bytes(n, gx, gy, gz), kept between calls. */
size_t mpmx_substep_workspace_bytes(int n, int gx, int gy, int gz);
int mpmx_substep_bin(const mpmx_particles* particles, const mpmx_grid* grid,
                     const mpmx_step_params* params, void* workspace, size_t ws_bytes,
                     void* stream);
int mpmx_mpm_substep(const mpmx_particles* particles, const mpmx_grid* grid,
                     const mpmx_step_params* params, void* workspace, size_t ws_bytes,
                     void* stream);
/* Number of particles found outside their bin's cell since the last
 * mpmx_substep_bin() (0 when bins were kept valid; synchronising). */
int mpmx_substep_misbinned(void* workspace, size_t ws_bytes, int gx, int gy, int gz, int n,
                           int* host_count, void* stream);

/* Ensemble: several independent scenes (each with its own grid and walls, the
 * same dx, dt and gravity) advanced by ONE launch of every kernel of the
 * fused substep above.  Tiles and grid nodes are numbered scene after scene:
 * scene s owns the nodes node0(s) .. node0(s) + gx*gy*gz of the grid arrays,
 * node0(s) being the node count of the scenes before it; particles carry their
 * scene in particles->scene.  The step parameters' grid origin is not used (each
 * scene has its own).  Each scene's results equal running it alone through
 * mpmx_mpm_substep.  Workspace: mpmx_ensemble_workspace_bytes(); call
 * mpmx_ensemble_bin() first, as mpmx_substep_bin() for one scene. */
size_t mpmx_ensemble_workspace_bytes(int n, const mpmx_scene* scenes, int nscenes);
int mpmx_ensemble_bin(const mpmx_particles* particles, const mpmx_scene* scenes, int nscenes,
                      const mpmx_step_params* params, void* workspace, size_t ws_bytes,
                      void* stream);
int mpmx_ensemble_substep(const mpmx_particles* particles, const mpmx_scene* scenes, int nscenes,
                          float* grid_m, float* grid_v_in, float* grid_v_out,
                          const mpmx_step_params* params, void* workspace, size_t ws_bytes,
                          void* stream);
int mpmx_ensemble_reorder_permutation(const mpmx_particles* particles, const mpmx_scene* scenes,
                                      int nscenes, const mpmx_step_params* params, int32_t* perm,
                                      void* workspace, size_t ws_bytes, void* stream);

/* Spatial order of the particle arrays.  perm[n] receives a permutation that
 * lists the active MPM particles tile by tile and cell by cell of the
 * substep's binning (ascending particles->order inside a cell, or ascending id
 * when order is NULL), then the out-of-grid ones, then every other particle in
 * ascending order (or id).  Gathering every per-particle array with it
 * (mpmx_gather_words) makes each tile's particles contiguous in HBM, so the
 * substep's particle loads coalesce; call mpmx_substep_bin() afterwards.  Same
 * workspace as the substep.  No reference counterpart (the reference keeps
 * the input order). */
int mpmx_reorder_permutation(const mpmx_particles* particles, const mpmx_grid* grid,
                             const mpmx_step_params* params, int32_t* perm, void* workspace,
                             size_t ws_bytes, void* stream);
/* dst[(p*n + i)*item_words + k] = src[(p*n + perm[i])*item_words + k] for
 * planes p < planes, items i < n, 32-bit words k < item_words: one SoA array
 * (planes = components, item_words = 1) or an AoS one (planes = 1, item_words
 * = words per item) permuted by perm; dst and src must not overlap. */
int mpmx_gather_words(void* dst, const void* src, const int32_t* perm, long n, int planes,
                      int item_words, void* stream);
/* inv[perm[i]] = i for a permutation perm of 0..n-1 (e.g. the array index of
 * each input index from the particle numbering after a reorder) */
int mpmx_invert_permutation(const int32_t* perm, int32_t* inv, long n, void* stream);

/* grid_normalization_and_gravity, mpmRoutines.py:691-704: writes v_out only
 * where m > 1e-15 (the caller zeroes v_out first, simulationRoutines.py:13) */
int mpmx_grid_normalization_and_gravity(int gx, int gy, int gz, const float* grid_m,
                                        const float* grid_v_in, float* grid_v_out, float gxv,
                                        float gyv, float gzv, float dt, void* stream);
/* add_damping_via_grid, mpmRoutines.py:706-711 */
int mpmx_add_damping_via_grid(int gx, int gy, int gz, float* grid_v_out, float scale,
                              void* stream);
/* collideBounds (+ apply_coulomb_friction), mpmRoutines.py:915-993 */
int mpmx_collide_bounds(float* grid_v_out, int gx, int gy, int gz, float friction, float dt,
                        float gravity_magnitude, int boundary_padding, void* stream);
/* steps 4-6 of mpmSimulationStep (simulationRoutines.py:94-140) fused into one
 * pass: normalisation + gravity, damping when damping_scale < 1, then the
 * boundary condition bc (MPMX_BC_*). */
int mpmx_grid_update(int gx, int gy, int gz, const float* grid_m, const float* grid_v_in,
                     float* grid_v_out, float gxv, float gyv, float gzv, float dt,
                     float damping_scale, int bc, float friction, float gravity_magnitude,
                     int boundary_padding, float restitution, void* stream);
/* the other boundary conditions of simulationRoutines.py:141-228, one kernel
 * each in the reference: collideBoundsGradualFriction (mpmRoutines.py:1249-1340),
 * collideBoundsRestitution (:1179-1247), collideBoundsAbsorbing (:1126-1176),
 * collideBoundsPlaneStrainY (:996-1062), collideBoundsPlaneStrainYGlobal
 * (:1065-1123); bc selects one (MPMX_BC_*, MPMX_BC_FRICTION = collideBounds) */
int mpmx_collide_bounds_variant(float* grid_v_out, int gx, int gy, int gz, int bc, float friction,
                                float dt, float gravity_magnitude, int boundary_padding,
                                float restitution, void* stream);
/* Volumetric locking correction (mpmRoutines.py:805-911, after G2P,
 * simulationRoutines.py:255-293).  accumulate_cell_divergence: per grid cell
 * int((x - min) inv_dx) the sums of tr(C) vol and vol over active MPM particles,
 * into cell_div / cell_vol [gx*gy*gz] (overwritten; they are summed in 64-bit
 * fixed point, so the result does not depend on the order of the atomics).
 * apply: C += (cell_div/cell_vol - tr C) / 3 I.  Particles whose cell is outside
 * the grid are skipped (the reference would index out of bounds).  Workspace:
 * mpmx_vlc_workspace_bytes(gx, gy, gz). */
size_t mpmx_vlc_workspace_bytes(int gx, int gy, int gz);
int mpmx_accumulate_cell_divergence(int n, const int32_t* activeLabel, const int32_t* materialLabel,
                                    const float* x, const float* C, const float* vol, float inv_dx,
                                    float minx, float miny, float minz, int gx, int gy, int gz,
                                    float* cell_div, float* cell_vol, void* workspace,
                                    size_t ws_bytes, void* stream);
int mpmx_apply_volumetric_locking_correction(int n, const int32_t* activeLabel,
                                             const int32_t* materialLabel, const float* x, float* C,
                                             float inv_dx, float minx, float miny, float minz,
                                             int gx, int gy, int gz, const float* cell_div,
                                             const float* cell_vol, void* stream);

/* g2p (+ update_cov), mpmRoutines.py:715-802; launched simulationRoutines.py:231-250.
 * cov is float[6*n] packed per particle [xx,xy,xz,yy,yz,zz] (the reference layout). */
int mpmx_g2p(float dt, int n, const int32_t* activeLabel, const int32_t* materialLabel,
             float* x, float* v, float* C, const float* F, float* F_trial, float* cov,
             float inv_dx, const float* grid_v_out, int gx, int gy, int gz, int update_cov,
             float minx, float miny, float minz, void* stream);

/* ---- reductions (utils/simulationRoutines.py:538-675), FP64 on the device,
 * two-level and in a fixed order, so the results are run-to-run identical.
 * All results are written to device memory; read them with a synchronising
 * copy.  The workspace must hold mpmx_reduce_workspace_bytes(n) bytes. */
size_t mpmx_reduce_workspace_bytes(int n);
/* velocityConvergence :538-549 and countActiveParticles :551-558:
 * out_sum = sum over active of |v|/radius, out_count = number of active */
int mpmx_velocity_convergence(int n, const int32_t* activeLabel, const float* v,
                              const float* radius, double* out_sum, int32_t* out_count,
                              void* workspace, size_t ws_bytes, void* stream);
/* No counterpart in the reference: mpmx_velocity_convergence without a
 * workspace, for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same kernels, same
 * results bit for bit.  Its block partials live in a 12 KB device array of
 * the library (one per GPU, held until the process ends); it allocates
 * nothing.  Any n >= 0 fits (n beyond the scratch would return MPMX_ERANGE).
 * Calls that overlap on one GPU race on the array. */
int mpmx_velocity_convergence_noalloc(int n, const int32_t* activeLabel, const float* v,
                                      const float* radius, double* out_sum, int32_t* out_count,
                                      void* stream);
/* sum_float_array_active :598-610 (compute_mean_gpu filtered=True, :633-675) */
int mpmx_sum_float_array_active(int n, const float* arr, const int32_t* activeLabel,
                                double* out_sum, int32_t* out_count, void* workspace,
                                size_t ws_bytes, void* stream);
/* countParticlesByType :560-585: out4 = {activeMPM, activeXPBD, inactiveMPM, inactiveXPBD} */
int mpmx_count_particles_by_type(int n, const int32_t* activeLabel,
                                 const int32_t* materialLabel, int32_t* out4, void* workspace,
                                 size_t ws_bytes, void* stream);

/* ---- XPBD (utils/xpbdRoutines.py) ---------------------------------------- */

/* wp.HashGrid(dx,dy,dz).build(points, cell_width) (simStates.py:174,
 * simulationRoutines.py:311): point_ids[n] lists the points sorted by hash cell
 * (stable: ascending id inside a cell); cell_start[dims+1] with cell c holding
 * point_ids[cell_start[c] .. cell_start[c+1]).  Workspace:
 * mpmx_hashgrid_workspace_bytes(n, dims). */
size_t mpmx_hashgrid_workspace_bytes(int n, int dimx, int dimy, int dimz);
int mpmx_hashgrid_build(int n, const float* x, float cell_width, int dimx, int dimy, int dimz,
                        int32_t* point_ids, int32_t* cell_start, void* workspace,
                        size_t ws_bytes, void* stream);

/* The same table over the points whose materialLabel equals `label` only (the
 * others are left out; cell_start[dims] = number of binned points).  With
 * label 2 this is the XPBD-only table used by the split contact solve below. */
int mpmx_hashgrid_build_label(int n, const float* x, float cell_width, int dimx, int dimy,
                              int dimz, const int32_t* materialLabel, int label,
                              int32_t* point_ids, int32_t* cell_start, void* workspace,
                              size_t ws_bytes, void* stream);

/* Either table above with the ids inside a cell ascending in order[id]
 * instead of id (materialLabel NULL = all points).  With order = the
 * particles' numbering before a spatial permutation of the arrays, the
 * candidate order of the contact solve, and so its results, are those of the
 * unpermuted arrays.  No reference counterpart (the reference never permutes). */
int mpmx_hashgrid_build_ordered(int n, const float* x, float cell_width, int dimx, int dimy,
                                int dimz, const int32_t* materialLabel, int label,
                                const int32_t* order, int32_t* point_ids, int32_t* cell_start,
                                void* workspace, size_t ws_bytes, void* stream);

/* The table of all points (simulationRoutines.py:311, the grid of the XPBD
 * step, dims dimx x dimy x dimz) and the table of the points whose
 * materialLabel equals `label` (the split contact solve's, dims label_dim*), in
 * one set of launches: the same contents as mpmx_hashgrid_build_ordered called
 * once per table with its dims, with materialLabel NULL for the first (order may
 * be NULL).  point_pos / label_point_pos (both or neither may be NULL; n each):
 * each point's position in point_ids / label_point_ids, -1 where it is not in
 * the second table; written only when the second table holds a point (else
 * left as they were).  lazy != 0: while no point has the label, the first
 * table (point_ids, cell_start) is left as it was; the fused XPBD step's
 * contact solve then reads neither table.  After a lazy build, read the first
 * table only when the second holds a point (label_cell_start[label_dimx *
 * label_dimy * label_dimz] > 0): otherwise it may list the ids of an earlier
 * build, stale after a reorder of the particle arrays.  Workspace:
 * mpmx_hashgrid_pair_workspace_bytes. */
size_t mpmx_hashgrid_pair_workspace_bytes(int n, int dimx, int dimy, int dimz, int label_dimx,
                                          int label_dimy, int label_dimz);
int mpmx_hashgrid_build_pair(int n, const float* x, float cell_width, int dimx, int dimy,
                             int dimz, const int32_t* materialLabel, int label, int label_dimx,
                             int label_dimy, int label_dimz, const int32_t* order,
                             int32_t* point_ids, int32_t* cell_start, int32_t* label_point_ids,
                             int32_t* label_cell_start, int32_t* point_pos, int32_t* label_point_pos,
                             int lazy, void* workspace, size_t ws_bytes, void* stream);

/* reset_mpm_contact_labels, xpbdRoutines.py:3-21 */
int mpmx_reset_mpm_contact_labels(int n, const int32_t* activeLabel, int32_t* materialLabel,
                                  void* stream);
/* integrateParticlesXPBD, xpbdRoutines.py:23-48 (x_new/v_new may alias x/v) */
int mpmx_integrate_particles_xpbd(int n, const int32_t* activeLabel,
                                  const int32_t* materialLabel, const float* x, const float* v,
                                  float gxv, float gyv, float gzv, float dt, float* x_new,
                                  float* v_new, float v_max, void* stream);
/* my_solve_particle_bound_contacts, xpbdRoutines.py:49-136 (adds into delta) */
int mpmx_solve_particle_bound_contacts(int n, const int32_t* activeLabel,
                                       const int32_t* materialLabel, const float* x,
                                       const float* v, const float* mass, const float* radius,
                                       float mu, float static_threshold, float static_friction,
                                       float minx, float miny, float minz, float maxx, float maxy,
                                       float maxz, float dt, float relaxation, float* delta,
                                       void* stream);
/* my_solve_particle_particle_contacts, xpbdRoutines.py:138-219 (adds into deltas);
 * candidates are visited in wp.hash_grid_query order (cells x fastest, then y,
 * then z; ids ascending inside a cell). */
int mpmx_solve_particle_particle_contacts(int n, const int32_t* activeLabel,
                                          int32_t* materialLabel, const int32_t* point_ids,
                                          const int32_t* cell_start, float cell_width, int dimx,
                                          int dimy, int dimz, const float* x, const float* v,
                                          const float* mass, const float* radius, float k_mu,
                                          float static_threshold, float static_friction,
                                          float k_cohesion, float max_radius, float dt,
                                          float relaxation, int contact_lock, float* deltas,
                                          void* stream);
/* my_solve_particle_particle_contacts with a second table holding only the
 * label-2 points (mpmx_hashgrid_build_label(..., 2, ...) on the same positions,
 * dims xpbd_dim*, which may be smaller than the first table's): MPM particles
 * (labels 0/1), which skip every MPM candidate in the reference
 * (xpbdRoutines.py:185-186), walk the XPBD-only table, whose candidates in
 * contact reach come in the same relative order.  A particle whose query box
 * comes within its contact reach of wrapping in the smaller table walks the
 * full one.  Results equal mpmx_solve_particle_particle_contacts.
 * slot_of_id (may be NULL): after a spatial permutation of the particle arrays,
 * slot_of_id[k] is the array index of the particle whose input index is k; the
 * reference's tests of activeLabel/materialLabel/mass at the thread index
 * (xpbdRoutines.py:151-175) then read that particle, as without the permutation. */
int mpmx_solve_particle_particle_contacts_split(
    int n, const int32_t* activeLabel, int32_t* materialLabel, const int32_t* point_ids,
    const int32_t* cell_start, const int32_t* xpbd_point_ids, const int32_t* xpbd_cell_start,
    float cell_width, int dimx, int dimy, int dimz, int xpbd_dimx, int xpbd_dimy, int xpbd_dimz,
    const float* x, const float* v, const float* mass, const float* radius, float k_mu,
    float static_threshold, float static_friction, float k_cohesion, float max_radius, float dt,
    float relaxation, int contact_lock, float* deltas, const int32_t* slot_of_id, void* stream);

/* Candidate records for mpmx_solve_particle_particle_contacts_packed in
 * hash-table order: record s < n describes particle point_ids[s] of the full
 * table, record n + s particle xpbd_point_ids[s] of the XPBD-only table (its
 * point count is xpbd_cell_start[cells]); xr[8n] = (x, y, z, radius),
 * vw[8n] = (vx, vy, vz, 1/mass), flags[2n] = 1 (activeLabel == 1 and mass > 0)
 * | 2 (materialLabel <= 1) | 4 (materialLabel == 2).  x and v are the SoA
 * arrays the contact solve reads.  With particle_xr (may be NULL; then also
 * particle_vw, particle_flags) the records are gathered from records in
 * particle order (mpmx_xpbd_bounds) instead of from activeLabel ... v, which
 * may then be NULL.  Nothing is written while the XPBD-only table is empty
 * (the contact solve then reads no record).  Records need 16-byte alignment.  No counterpart in the reference: a layout of the data
 * that my_solve_particle_particle_contacts gathers per candidate
 * (xpbdRoutines.py:178-196). */
int mpmx_xpbd_contact_pack(int n, const int32_t* activeLabel, const int32_t* materialLabel,
                           const float* mass, const float* radius, const float* x, const float* v,
                           const float* particle_xr, const float* particle_vw,
                           const int32_t* particle_flags, const int32_t* point_ids,
                           const int32_t* xpbd_point_ids,
                           const int32_t* xpbd_cell_start, int xpbd_dimx, int xpbd_dimy, int xpbd_dimz,
                           float* xr, float* vw, int32_t* flags, void* stream);

/* mpmx_solve_particle_particle_contacts_split reading each candidate from the
 * records of mpmx_xpbd_contact_pack, packed from the same tables, x, v, labels,
 * mass and radius after their last change (my_solve_particle_particle_contacts,
 * xpbdRoutines.py:138-219).  The results equal those of the split form: lock
 * writes turn label 1 into 0 only, which changes no flag. */
int mpmx_solve_particle_particle_contacts_packed(
    int n, const int32_t* activeLabel, int32_t* materialLabel, const int32_t* point_ids,
    const int32_t* cell_start, const int32_t* xpbd_point_ids, const int32_t* xpbd_cell_start,
    float cell_width, int dimx, int dimy, int dimz, int xpbd_dimx, int xpbd_dimy, int xpbd_dimz,
    const float* x, const float* v, const float* mass, const float* radius, const float* xr,
    const float* vw, const int32_t* flags, float k_mu, float static_threshold, float static_friction,
    float k_cohesion, float max_radius, float dt, float relaxation, int contact_lock, float* deltas,
    const int32_t* slot_of_id, const uint64_t* xpbd_near, const int32_t* contact_list,
    const int32_t* contact_count, void* stream);

/* The table positions whose contact walk can meet a contact in this XPBD step,
 * for mpmx_solve_particle_particle_contacts_packed (contact_list, contact_count;
 * both NULL solves every position).  No counterpart in the reference.  It drops
 * the MPM particles that walk the XPBD-only table from a cell whose
 * neighbourhood bitmap bit (mpmx_hashgrid_near_cells) is clear: such a walk
 * meets no XPBD point in any iteration, so the particle gets no delta and its
 * query box stays put.  point_pos is the full table's position of every
 * particle (mpmx_hashgrid_build_pair); list holds up to n positions, count gets
 * their number (0 while the XPBD-only table is empty).  With more than n/16
 * XPBD points no list is written and count gets n + 1; the solve walks every
 * table position whenever count exceeds n/8 (a long list in particle-array
 * order walks slower than the table order; the results are the same bits).
 * Build it after the tables of the step and before its first contact solve. */
int mpmx_xpbd_contact_list(int n, const int32_t* materialLabel, const int32_t* point_pos,
                           const int32_t* xpbd_cell_start, float cell_width, int xpbd_dimx,
                           int xpbd_dimy, int xpbd_dimz, const float* x, const float* radius,
                           float k_cohesion, float max_radius, const uint64_t* xpbd_near,
                           int32_t* list, int32_t* count, void* stream);

/* Neighbourhood bitmap of a hash table for the contact solve (no counterpart
 * in the reference): bit c of near (cells / 64 words, mpmx_hashgrid_near_words)
 * is set when table cell c or one of its 26 neighbours in the table's wrapped
 * cell coordinates holds a point; occ is scratch of the same size.  Given to
 * mpmx_solve_particle_particle_contacts_packed as xpbd_near (built on the
 * XPBD-only table, may be NULL), it lets an MPM particle whose query box lies
 * within one cell of its own cell skip a walk that the bitmap shows to meet no
 * XPBD point (the walk would solve no contact and write no lock label).
 * Nothing is written while the table is empty. */
size_t mpmx_hashgrid_near_words(int dimx, int dimy, int dimz);
int mpmx_hashgrid_near_cells(const int32_t* cell_start, int dimx, int dimy, int dimz, uint64_t* occ,
                             uint64_t* near, void* stream);

/* my_apply_particle_deltas, xpbdRoutines.py:221-257 (x_out/v_out may alias x_pred) */
int mpmx_apply_particle_deltas(int n, const int32_t* activeLabel, const int32_t* materialLabel,
                               const float* x_orig, const float* x_pred, const float* delta,
                               float dt, float v_max, float* x_out, float* v_out, void* stream);
/* sleepParticles, xpbdRoutines.py:313-334; counters[2] = {total, sleeping} are
 * incremented (zero them first, simulationRoutines.py:424-425) */
int mpmx_sleep_particles(int n, const int32_t* activeLabel, const int32_t* materialLabel,
                         float threshold, const float* radius, const float* x_prev, float* x_after,
                         float* dist_total, float dt, int32_t* counters, void* stream);
/* apply_mpm_contact_impulses, xpbdRoutines.py:259-310 */
int mpmx_apply_mpm_contact_impulses(int n, const int32_t* activeLabel,
                                    const int32_t* materialLabel, const float* x_orig,
                                    const float* v_orig, float* x_final, const float* v_final,
                                    float dt, float v_max, float coupling, int xpbd_only,
                                    float* v_out, void* stream);
/* ---- fused XPBD step: xpbdSimulationStep (simulationRoutines.py:295-526) with
 * its copies folded in.  Sequence: mpmx_xpbd_begin; per iteration
 * mpmx_xpbd_bound_contacts_write, mpmx_solve_particle_particle_contacts[_split],
 * mpmx_xpbd_apply_deltas_inplace; then mpmx_xpbd_finish (its two counters are
 * zeroed by mpmx_xpbd_begin when it is given them, else by the caller).
 * x, v, labels, the cumulative distance and the counters end as after the launch
 * sequence; the scratch arrays particle_x_deltaInt/particle_v_deltaInt/
 * particle_v_before_contact are not written. */
/* The parameters of my_solve_particle_bound_contacts (xpbdRoutines.py:49-136)
 * and the delta array it writes, for the kernels below that fold it in. */
typedef struct {
    const float* mass;
    const float* radius;
    float mu, static_threshold, static_friction;
    float minx, miny, minz, maxx, maxy, maxz; /* the XPBD box */
    float dt, relaxation;
    float* delta; /* [3][n]: every particle's delta is written */
    /* The candidate records of the next contact solve, in table order (2n
     * records, the layout of mpmx_xpbd_contact_pack), written from the new
     * x_int / v_int and labels by the kernel that takes these bounds at each
     * particle's positions pos[p] and n + xpbd_pos[p] (point_pos and
     * label_point_pos of mpmx_hashgrid_build_pair; -1: not in that table).
     * Not written when xr is NULL, nor while xpbd_count (may be NULL: the
     * XPBD-only table's point count, cell_start[cells]) reads 0, when the
     * contact solve reads no record.  While xpbd_count reads 0,
     * mpmx_xpbd_apply_deltas_inplace given these bounds applies nothing: with
     * no label-2 particle every delta is 0, and mpmx_xpbd_finish gives the same
     * x and v whether the zero deltas were applied or not. */
    float* xr;
    float* vw;
    int32_t* flags;
    const int32_t* pos;
    const int32_t* xpbd_pos;
    const int32_t* xpbd_count;
} mpmx_xpbd_bounds;

/* x_int, v_int <- x, v then integrateParticlesXPBD (xpbdRoutines.py:23-48); with
 * reset_labels also reset_mpm_contact_labels (:3-21); counters[0..1] (the sleep
 * counts of mpmx_xpbd_finish, simulationRoutines.py:424-425) are zeroed when
 * counters is not NULL.  With bounds (may be NULL), also the first iteration's
 * particle_delta.zero_() + my_solve_particle_bound_contacts from the new
 * x_int / v_int, as mpmx_xpbd_bound_contacts_write would give. */
int mpmx_xpbd_begin(int n, const int32_t* activeLabel, int32_t* materialLabel, const float* x,
                    const float* v, float gxv, float gyv, float gzv, float dt, float v_max,
                    int reset_labels, float* x_int, float* v_int, int32_t* counters,
                    const mpmx_xpbd_bounds* bounds, void* stream);
/* particle_delta.zero_() + my_solve_particle_bound_contacts (xpbdRoutines.py:49-136):
 * writes every particle's delta */
int mpmx_xpbd_bound_contacts_write(int n, const int32_t* activeLabel,
                                   const int32_t* materialLabel, const float* x, const float* v,
                                   const float* mass, const float* radius, float mu,
                                   float static_threshold, float static_friction, float minx,
                                   float miny, float minz, float maxx, float maxy, float maxz,
                                   float dt, float relaxation, float* delta, void* stream);
/* the four copies and my_apply_particle_deltas (xpbdRoutines.py:221-257) as one
 * in-place update of x_int / v_int.  With next_bounds (may be NULL; it then
 * needs materialLabel), also the next iteration's bound contacts from the new
 * x_int / v_int into next_bounds->delta (which may be delta itself). */
int mpmx_xpbd_apply_deltas_inplace(int n, const int32_t* activeLabel, const float* x_orig,
                                   const float* delta, float dt, float v_max, float* x_int,
                                   float* v_int, const int32_t* materialLabel,
                                   const mpmx_xpbd_bounds* next_bounds, void* stream);
/* sleepParticles (xpbdRoutines.py:313-334), apply_mpm_contact_impulses (:259-310)
 * and the commit (simulationRoutines.py:469-470); counters = {total, sleeping}.
 * With delta (may be NULL), the last iteration's my_apply_particle_deltas
 * (x_orig = x, the same dt and v_max) comes first. */
int mpmx_xpbd_finish(int n, const int32_t* activeLabel, const int32_t* materialLabel,
                     float sleep_threshold, const float* radius, float* dist_total, float dt,
                     float v_max, float coupling, int xpbd_only, float* x, float* v, float* x_int,
                     float* v_int, int32_t* counters, const float* delta, void* stream);

/* swellParticlesType2, xpbdRoutines.py:354-376 */
int mpmx_swell_particles_type2(int n, const int32_t* activeLabel, const int32_t* materialLabel,
                               const float* x_init, const float* x, float* radius, float factor,
                               float factor_max, const float* swell_radius,
                               const float* base_radius, void* stream);

/* ---- learned return map (no reference counterpart) ----------------------
 * A 10-64-64-3 GELU network stands in for the projection step of the
 * Drucker-Prager return map (drucker_prager_return_mapping, mpmRoutines.py:
 * 360-413), with bfloat16 weights on the matrix cores and FP32 accumulation.
 * Parameters: one FP32 vector of mpmx_mlp_nparams() floats, W1[64][10], b1[64],
 * W2[64][64], b2[64], W3[3][64], b3[3] (out x in).  Samples: rows of
 * X[B][10]; predictions and targets: rows of Y[B][3], T[B][3].  Features and
 * targets are defined with mpmx_lrm_harvest below. */
int mpmx_mlp_nparams(void);
int mpmx_mlp_forward(int B, const float* params, const float* X, float* Y, void* stream);
/* mean squared error over the B*3 outputs and its gradient w.r.t. every
 * parameter (FP32 grad[nparams], FP64 *loss, may be NULL); reproducible bit for
 * bit for a given B.  Workspace: mpmx_mlp_workspace_bytes(B). */
size_t mpmx_mlp_workspace_bytes(int B);
int mpmx_mlp_loss_grad(int B, const float* params, const float* X, const float* T, float* grad,
                       double* loss, void* workspace, size_t ws_bytes, void* stream);
/* one Adam step (Kingma & Ba 2015) on n parameters, step number t >= 1 */
int mpmx_mlp_adam_step(int n, float* params, const float* grad, float* m, float* v, float lr,
                       float beta1, float beta2, float eps, int t, void* stream);
/* Training pairs from the simulator's own state after G2P: for every particle,
 * flag = 1 (active MPM, plastic trial state), 0 (active MPM, elastic) or -1.
 * With c = max(ys, 1), eref = c / (2 mu), d the deviator of the principal
 * trial log strains e, Y the yield value, the overshoot phi = (2 mu |d| - Y)/c,
 * y0 = Y / c and kap = 2 mu 1e-6 / c, X[p] = the 10 features 0.1 ln(phi + kap),
 * 0.1 ln(max(y0, 1e-12)), 0.1 ln kap, d/|d| (3), tan friction angle, 4 nu - 1,
 * log10 E - 8, damage; T[p] = (s, 20 (alpha - 0.3), 0) for plastic states (zeros for
 * elastic ones), with s = Y / (2 mu (|d| + 1e-6)), so that the exact projection
 * is e' = e - (1 - Y/(2 mu |d|)) d - eref kap s d/|d| + 0.3 alpha eref
 * (phi + kap); e_trial[p] = e.  The state is not modified. */
int mpmx_lrm_harvest(const mpmx_particles* particles, float* X, float* T, float* e_trial,
                     int32_t* flag, void* stream);
/* compute_stress_from_F_trial (mpmRoutines.py:229-312) with the projection of
 * plastic states predicted by the network (params; NULL = exact projection for
 * all) and judged by the exact map: a prediction whose yield-surface/flow
 * residual 2 mu |dev e' - Y/(2 mu edn) d| or dilatancy residual
 * K |tr e' - tr e - 0.9 alpha dgam| exceeds tol * max(cohesion, 1) is replaced
 * by the exact projection.  counters[3] (64-bit, may be NULL) += {plastic states,
 * accepted predictions, particles updated}.  Workspace (with params):
 * mpmx_lrm_workspace_bytes(n). */
size_t mpmx_lrm_workspace_bytes(int n);
int mpmx_compute_stress_learned(const mpmx_particles* particles, float eff, const float* params,
                                float tol, int64_t* counters, void* workspace, size_t ws_bytes,
                                void* stream);
/* mpmx_mpm_substep and mpmx_ensemble_substep with the constitutive update of
 * mpmx_compute_stress_learned (network mlp_params, tolerance tol, counters as
 * there) run as its own kernel before the P2G tile kernel. */
int mpmx_mpm_substep_learned(const mpmx_particles* particles, const mpmx_grid* grid,
                             const mpmx_step_params* params, const float* mlp_params, float tol,
                             int64_t* counters, void* workspace, size_t ws_bytes, void* stream);
int mpmx_ensemble_substep_learned(const mpmx_particles* particles, const mpmx_scene* scenes,
                                  int nscenes, float* grid_m, float* grid_v_in, float* grid_v_out,
                                  const mpmx_step_params* params, const float* mlp_params, float tol,
                                  int64_t* counters, void* workspace, size_t ws_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MPMX_H */
