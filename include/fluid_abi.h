/*
 * fluid_abi.h — C ABI of the MI355X (gfx950) SPH step library `libfluidhip.so`.
 *
 * The library is a drop-in for the compute side of fluid-toy's
 * `SphSimulation3DGrid` (src/simulation/sph_3d_grid.rs:33-1570): it owns every
 * device buffer of the step and a HIP stream, runs the same 13 compute entry
 * points per step (grid clear/build, prefix scan, reorder, density, XSPH,
 * force, PCISPH predict/solve xN/finalize, integrate) as hand-written HIP
 * kernels, and exposes the struct's public methods as plain C functions.
 *
 * Conventions
 *  - Every function returns an int status: FLUID_OK (0) or a negative
 *    FLUID_ERR_* code; fluid_last_error() returns a thread-local message.
 *    The library never aborts the process (the reference panics instead,
 *    src/gpu/mod.rs:33,45).
 *  - The caller owns host arrays; they are copied in and out during the call
 *    and no pointer is retained.  Structs are passed by const pointer.
 *  - One handle is externally serialised.  GPU work is stream-ordered and
 *    asynchronous; readbacks synchronise the handle's stream (the reference's
 *    only sync point is device.poll in read_rigid_body_accum,
 *    sph_3d_grid.rs:1486).
 *  - Plain pointers and sizes only; no C++ or torch types cross the boundary.
 *
 * Every POD struct below is byte-identical to its Rust/WGSL counterpart
 * (static_asserts in fluid-toy_amd/csrc/fluid_internal.h).
 */
#ifndef FLUID_ABI_H
#define FLUID_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLUID_ABI_VERSION 1

/* ---- status codes ---------------------------------------------------- */
#define FLUID_OK 0
#define FLUID_ERR_INVALID_ARG (-1) /* null pointer, bad size, bad value   */
#define FLUID_ERR_CAPACITY (-2)    /* request exceeds max_particles/cap    */
#define FLUID_ERR_HIP (-3)         /* a HIP runtime call failed            */
#define FLUID_ERR_NO_DEVICE (-4)   /* no usable gfx950 device             */
#define FLUID_ERR_STATE (-5)       /* call not valid in the current state  */

/* ---- POD structs (byte layouts of the reference) --------------------- */

/* src/simulation/particle.rs:6-17 (64 B, WGSL vec3 padding) */
typedef struct SphParticle3D {
  float position[3];
  float _pad0;
  float velocity[3];
  float _pad1;
  float force[3];
  float density;
  float near_density;
  float normal[3];
} SphParticle3D;

/* src/state/simulation.rs:342-361 (64 B) */
typedef struct GpuSphParams3D {
  float kernel_radius;
  float kernel_radius_sq;
  float kernel_radius_pow5;
  float kernel_radius_pow6;
  float kernel_radius_pow9;
  float mass;
  float rest_density;
  float stiffness; /* copied into the uniform, read by no kernel */
  float near_stiffness;
  float viscosity;
  float dt;
  uint32_t num_particles;
  float surface_tension;
  float pcisph_delta;
  float xsph_epsilon;
  float _padding_st;
} GpuSphParams3D;

/* src/simulation/sph_3d_grid.rs:16-30 (48 B) */
typedef struct GpuGridParams {
  uint32_t grid_size_x;
  uint32_t grid_size_y;
  uint32_t grid_size_z;
  uint32_t total_cells;
  float cell_size;
  float inv_cell_size;
  float grid_origin_x;
  float grid_origin_y;
  float grid_origin_z;
  uint32_t num_particles;
  uint32_t _padding[2];
} GpuGridParams;

/* src/state/simulation.rs:365-386 (128 B) */
typedef struct GpuContainerGeometry {
  float half_width;
  float half_height;
  float half_depth;
  float center_y;
  float forward_row0[4];
  float forward_row1[4];
  float forward_row2[4];
  float inverse_row0[4];
  float inverse_row1[4];
  float inverse_row2[4];
  float wall_stiffness;
  float damping;
  uint32_t clip_enabled;
  float clip_margin;
} GpuContainerGeometry;

/* src/state/simulation.rs:389-394 (16 B) */
typedef struct GpuGravity {
  float direction[3];
  float _padding;
} GpuGravity;

/* src/state/interaction.rs:102-111 (48 B) */
typedef struct GpuMouseForce {
  float position[3];
  float radius;
  float strength;
  uint32_t is_active;
  uint32_t mode; /* 0 push, 1 pull, 2 vortex, 3 explode, 4 drain */
  float _pad;
  float direction[3];
  float _pad2;
} GpuMouseForce;

/* src/state/rigid_body.rs:212-224 (96 B) */
typedef struct GpuRigidBody {
  float position[3];
  float half_extent;
  float velocity[3];
  uint32_t is_active;
  float stiffness;
  uint32_t shape; /* 0 cube, 1 sphere, 2 cylinder, 3 torus, 4 custom SDF */
  float _pad1;
  float _pad2;
  float rot_row0[4];
  float rot_row1[4];
  float rot_row2[4];
} GpuRigidBody;

/* src/state/rigid_body.rs:247-256 (32 B; fixed point x1000) */
typedef struct GpuRigidBodyAccum {
  int32_t force_x;
  int32_t force_y;
  int32_t force_z;
  uint32_t contact_count;
  int32_t torque_x;
  int32_t torque_y;
  int32_t torque_z;
  uint32_t _pad;
} GpuRigidBodyAccum;

/* src/shaders/pcisph_predict.wgsl:33-42 (32 B) */
typedef struct PredictedState {
  float pred_pos[3];
  float pressure;
  float pred_vel[3];
  float pred_density;
} PredictedState;

/* Parameters of SphSimulation3DGrid::new (sph_3d_grid.rs:104-112).
 * `bounds_extent` sizes the dense cubic grid exactly as the reference does
 * (sph_3d_grid.rs:121-143): cell = h, G = ceil(2*extent/h)+2 per axis,
 * origin = -(extent+h).  0 selects the reference's fixed 2.0.  Larger scenes
 * (boxes wider than 4.0) pass a larger extent.  `sdf_voxels` (res^3 floats,
 * may be null) is the custom-shape rigid-body SDF of sph_3d_grid.rs:346-384. */
typedef struct fluid_create_info {
  const SphParticle3D *particles;
  uint32_t num_particles;
  uint32_t max_particles; /* raised to num_particles if smaller (:115) */
  GpuSphParams3D sph_params;
  GpuContainerGeometry container;
  float bounds_extent;
  int device; /* HIP device ordinal */
  const float *sdf_voxels;
  uint32_t sdf_resolution;
} fluid_create_info;

typedef struct fluid_sim fluid_sim;

/* ---- stage ids of one step (sph_3d_grid.rs:1269-1432, in order) ------ */
#define FLUID_STAGE_CLEAR 0     /* grid_clear.wgsl + clear_buffer(:1281)     */
#define FLUID_STAGE_BUILD 1     /* grid_build.wgsl                           */
#define FLUID_STAGE_SCAN 2      /* copy(:1295) + prefix_scan.wgsl x3         */
#define FLUID_STAGE_REORDER 3   /* grid_reorder.wgsl (canonical order)       */
#define FLUID_STAGE_DENSITY 4   /* sph_density_3d_grid.wgsl                  */
#define FLUID_STAGE_XSPH 5      /* xsph.wgsl                                 */
#define FLUID_STAGE_FORCE 6     /* sph_force_3d_grid.wgsl                    */
#define FLUID_STAGE_PREDICT 7   /* pcisph_predict.wgsl                       */
#define FLUID_STAGE_SOLVE 8     /* pcisph_solve.wgsl x pcisph_iterations     */
#define FLUID_STAGE_FINALIZE 9  /* pcisph_finalize.wgsl                      */
#define FLUID_STAGE_INTEGRATE 10 /* sph_integrate_3d.wgsl                    */
#define FLUID_STAGE_COUNT 11

/* ---- library --------------------------------------------------------- */

/* Thread-local message of the last failing call ("" if none). */
const char *fluid_last_error(void);

/* FLUID_ABI_VERSION of the loaded library. */
int fluid_abi_version(void);

/* Parameter derivation of SphConfig::to_gpu_params_3d
 * (src/state/simulation.rs:315-336): powers of h by repeated f32 multiply
 * (Rust powi), rest density from compute_rest_density (:402-425) unless
 * rest_density_override > 0, PCISPH delta from compute_pcisph_delta
 * (:429-480) evaluated with that rest density. */
int fluid_derive_sph_params(float kernel_radius, float stiffness,
                            float near_stiffness, float viscosity, float mass,
                            float surface_tension, float xsph_epsilon,
                            float dt, uint32_t num_particles,
                            float rest_density_override, GpuSphParams3D *out);

/* ContainerConfig::to_gpu_geometry (src/state/simulation.rs:139-184). */
int fluid_derive_container_geometry(float width, float height, float depth,
                                    float floor_y, float tilt_x, float tilt_z,
                                    float wall_stiffness, float damping,
                                    uint32_t clip_enabled, float clip_margin,
                                    GpuContainerGeometry *out);

/* ---- SphSimulation3DGrid surface -------------------------------------- */

/* SphSimulation3DGrid::new (sph_3d_grid.rs:104-1258).  Gravity starts at
 * (0,-9.8,0) (:277-280), mouse force and rigid body at their inactive
 * defaults (:270-292), pcisph_iterations at 5 (:1253). */
int fluid_create(const fluid_create_info *info, fluid_sim **out);

/* Drop the handle and all device memory. Null is accepted. */
int fluid_destroy(fluid_sim *sim);

/* SphSimulation3DGrid::step (sph_3d_grid.rs:1261-1444): one full step,
 * launched as a captured HIP graph on the handle's stream. */
int fluid_step(fluid_sim *sim);

/* `n` consecutive steps (the substep loop of app.rs:1211-1213). */
int fluid_step_n(fluid_sim *sim, uint32_t n);

/* Run stages [first, last] of one step with plain launches (no graph), for
 * per-stage inspection; a full step is first=0, last=FLUID_STAGE_COUNT-1. */
int fluid_step_stages(fluid_sim *sim, uint32_t first, uint32_t last);

/* update_sph_params (sph_3d_grid.rs:1446-1453).  num_particles in the
 * struct is ignored: the handle's own count drives every stage. */
int fluid_update_sph_params(fluid_sim *sim, const GpuSphParams3D *params);
/* update_container_geometry (:1455) */
int fluid_update_container_geometry(fluid_sim *sim,
                                    const GpuContainerGeometry *geom);
/* update_mouse_force (:1459) */
int fluid_update_mouse_force(fluid_sim *sim, const GpuMouseForce *mouse);
/* update_gravity (:1463) */
int fluid_update_gravity(fluid_sim *sim, const GpuGravity *gravity);
/* update_rigid_body (:1467) */
int fluid_update_rigid_body(fluid_sim *sim, const GpuRigidBody *body);
/* set_pcisph_iterations (:1471-1473), clamped to >= 1 */
int fluid_set_pcisph_iterations(fluid_sim *sim, uint32_t iterations);
uint32_t fluid_pcisph_iterations(const fluid_sim *sim);
/* clear_rigid_body_accum (:1475-1481) */
int fluid_clear_rigid_body_accum(fluid_sim *sim);
/* read_rigid_body_accum + rigid_body_accum (:1483-1498), blocking. */
int fluid_read_rigid_body_accum(fluid_sim *sim, GpuRigidBodyAccum *out);

/* spawn_particles (:1534-1569): golden-angle spiral, clamps to capacity;
 * *spawned receives the count actually added (may be 0). */
int fluid_spawn_particles(fluid_sim *sim, const float position[3],
                          uint32_t count, float spread, uint32_t *spawned);

/* num_particles (:1528) */
uint32_t fluid_num_particles(const fluid_sim *sim);
uint32_t fluid_max_particles(const fluid_sim *sim);

/* ---- buffer access (the getters of :1500-1526) ------------------------ */
/* All readbacks synchronise the stream.  `cap` is the element capacity of
 * `out`; FLUID_ERR_CAPACITY if it is smaller than the buffer. */

/* particle_buffer: canonical (original index) order, 64-B records. */
int fluid_read_particles(fluid_sim *sim, SphParticle3D *out, uint32_t cap);
/* queue.write_buffer(particle_buffer, 0, ..) of `n` records (n <= max). */
int fluid_write_particles(fluid_sim *sim, const SphParticle3D *in,
                          uint32_t n);
/* sorted_particle_buffer: grid-sorted order as left by the last reorder. */
int fluid_read_sorted_particles(fluid_sim *sim, SphParticle3D *out,
                                uint32_t cap);
/* cell_counts_buffer / cell_starts_buffer (inclusive scan: cell range end). */
int fluid_read_cell_counts(fluid_sim *sim, uint32_t *out, uint32_t cap);
int fluid_read_cell_starts(fluid_sim *sim, uint32_t *out, uint32_t cap);
/* sorted slot -> original index, and its inverse (the reorder's two maps,
 * grid_reorder.wgsl:56-62). */
int fluid_read_sorted_to_orig(fluid_sim *sim, uint32_t *out, uint32_t cap);
int fluid_read_orig_to_sorted(fluid_sim *sim, uint32_t *out, uint32_t cap);
/* sorted_predicted_{a,b} (which = 0 or 1), sorted order. */
int fluid_read_predicted(fluid_sim *sim, uint32_t which, PredictedState *out,
                         uint32_t cap);
/* grid_params_buffer / sph_params_buffer / container_geom_buffer */
int fluid_read_grid_params(const fluid_sim *sim, GpuGridParams *out);
int fluid_read_sph_params(const fluid_sim *sim, GpuSphParams3D *out);
int fluid_read_container_geometry(const fluid_sim *sim,
                                  GpuContainerGeometry *out);

/* ---- device-side access (interop) -------------------------------------
 * The reference hands its wgpu::Buffers to the spray system and the
 * renderers (the getters of sph_3d_grid.rs:1500-1526, read by spray.rs:
 * 105-113 and app.rs:1365,1430,1523-1528).  These are the same buffers on
 * the device, for a consumer that binds them without a host round trip
 * (e.g. memory it imported from a graphics API with hipImportExternalMemory
 * or allocated with hipMalloc).  Work is enqueued on the handle's stream
 * (fluid_stream) and the calls return without waiting: order consumers
 * after it with a HIP event or fluid_sync. */

/* particle_buffer: the canonical (original index order) SphParticle3D
 * records written into `dst`, device memory of at least num_particles
 * records, 16-byte aligned. */
int fluid_export_particles_device(fluid_sim *sim, SphParticle3D *dst, uint32_t cap);
/* sorted_particle_buffer: the records in the grid-sorted order of the last
 * reorder, with the density, near density and normal of the last density
 * pass (what the reference's density pass writes there). */
int fluid_export_sorted_particles_device(fluid_sim *sim, SphParticle3D *dst, uint32_t cap);

/* Device pointers the handle owns (valid until fluid_destroy; the step
 * rewrites their contents).  For a batched handle, `scene` selects one
 * scene's grid and uniform blocks; its cell_starts and sorted_to_orig hold
 * indices into the whole concatenated particle array (subtract
 * first_particle for the scene's own).  A plain handle has scene 0 only. */
typedef struct fluid_device_view {
  const uint32_t *cell_counts;        /* cell_counts_buffer, num_cells u32     */
  const uint32_t *cell_starts;        /* cell_starts_buffer (inclusive scan)   */
  uint32_t num_cells;
  const uint32_t *sorted_to_orig;     /* the reorder's slot -> original index  */
  uint32_t num_particles;
  uint32_t first_particle;
  const GpuSphParams3D *sph_params;   /* sph_params_buffer (64 B)              */
  const GpuGridParams *grid_params;   /* grid_params_buffer (48 B)             */
  const GpuContainerGeometry *container_geometry; /* container_geom_buffer     */
  const GpuGravity *gravity;          /* the gravity uniform (16 B)            */
  void *stream;                       /* the hipStream_t that writes them      */
} fluid_device_view;
int fluid_device_buffers(const fluid_sim *sim, uint32_t scene, fluid_device_view *out);

/* Wait for all queued work of the handle. */
int fluid_sync(fluid_sim *sim);

/* The handle's hipStream_t, as an opaque pointer (for interop). */
void *fluid_stream(fluid_sim *sim);

/* ---- batched form ------------------------------------------------------ */

/* One scene of a batch: its particles, uniforms and grid extent (as in
 * fluid_create_info; the scene's grid is make_grid(h, bounds_extent)). */
typedef struct fluid_scene_desc {
  const SphParticle3D *particles;
  uint32_t num_particles;
  GpuSphParams3D sph_params;
  GpuContainerGeometry container;
  GpuGravity gravity;
  float bounds_extent;
} fluid_scene_desc;

/* Many independent scenes advanced together: every stage is one launch over
 * all scenes' particles (concatenated scene-major, each scene keeping a
 * fixed slot range) and one scan over all scenes' grids (laid end to end).
 * The handle works with fluid_step/_n, fluid_time_*, fluid_autotune,
 * fluid_pair_stats, fluid_sync and the readbacks, which return all scenes
 * concatenated (original index = scene's first particle + index in scene).
 * Per-scene uniforms change through fluid_batch_update_scene; the single-
 * scene fluid_update_sph_params/_container_geometry/_gravity and
 * fluid_spawn_particles return FLUID_ERR_STATE on a batch.  No counterpart
 * in the reference (which simulates one scene per SphSimulation3DGrid). */
int fluid_batch_create(const fluid_scene_desc *scenes, uint32_t num_scenes,
                       uint32_t pcisph_iterations, int device, fluid_sim **out);
uint32_t fluid_batch_num_scenes(const fluid_sim *sim);
int fluid_batch_scene_info(const fluid_sim *sim, uint32_t scene,
                           uint32_t *first_particle, uint32_t *num_particles,
                           uint32_t *first_cell, GpuGridParams *grid);
int fluid_batch_update_scene(fluid_sim *sim, uint32_t scene,
                             const GpuSphParams3D *params,
                             const GpuContainerGeometry *container,
                             const GpuGravity *gravity);

/* ---- measurement ------------------------------------------------------ */

/* Time `steps` graph-launched steps with HIP events on the handle's stream
 * (after `warmup` untimed steps); *ms_total receives the elapsed time. */
int fluid_time_steps(fluid_sim *sim, uint32_t warmup, uint32_t steps,
                     float *ms_total);

/* Run `steps` steps with plain launches and a HIP event pair around every
 * stage kernel on the handle's stream; ms_per_stage[FLUID_STAGE_COUNT]
 * receives the mean duration of each stage per step (SOLVE summed over its
 * iterations). */
int fluid_time_stages(fluid_sim *sim, uint32_t steps, float *ms_per_stage);

/* Re-run stages [first, last] (within FLUID_STAGE_DENSITY ..
 * FLUID_STAGE_PREDICT) of the last step `reps` times back to back between
 * one HIP event pair on the handle's stream, launched as a full step would
 * launch them (XSPH + force as one kernel under FLUID_FUSE_XSPH_FORCE);
 * *ms_per_rep receives the mean time of one repetition.  These stages read
 * the step's grid and sorted state and rewrite what they compute, so every
 * repetition does the same work; the velocity, force and predicted state
 * they overwrite are put back afterwards.  FLUID_ERR_STATE before the
 * first step.  No counterpart in the reference: it times a kernel's
 * back-to-back launches (as a graph replay runs them), for the roofline of
 * the step's dominant kernel. */
int fluid_time_range(fluid_sim *sim, uint32_t first, uint32_t last, uint32_t reps,
                     float *ms_per_rep);

/* Launch configuration of the neighbour sweeps, cfg[3]:
 *   cfg[0] lanes cooperating on one particle in the density sweep: 1, 2,
 *          4, 8 or 16 (it writes the in-range list for 8 walker lanes
 *          either way; XSPH and force walk it with 8, 16 or 32 lanes:
 *          FLUID_WALK_16 / _32);
 *          cfg[1] the same for the PCISPH solve: 1, 2, 4, 8 or 16 (0 keeps
 *          the current value; the LDS-tiled sweeps run at most 8);
 *   cfg[2] workgroup-to-XCD mapping: 1 gives each of the 8 XCDs one
 *          contiguous run of tiles (shared L2 between neighbours), 0 keeps
 *          the hardware's round-robin order (spreads dense regions). */
int fluid_set_sweep_config(fluid_sim *sim, const uint32_t cfg[3]);
int fluid_get_sweep_config(const fluid_sim *sim, uint32_t out[3]);

/* LDS-staged sweeps, cfg[3]:
 *   cfg[0] 1: density and the PCISPH solve run as persistent workgroups
 *          over the non-empty cell tiles, each staging its tile's 27-cell
 *          neighbourhood (nine x-rows) in LDS; 0: global-memory gathers
 *          (the same sums either way, bit for bit);
 *   cfg[1] cells per tile along x, 1..16 (0 keeps the current value);
 *   cfg[2] persistent workgroups (0 keeps the current value). */
int fluid_set_tile_config(fluid_sim *sim, const uint32_t cfg[3]);

/* Work-sharing switches (flags; default the first two set).  These two give
 * the same bits either way:
 *   FLUID_FUSE_XSPH_FORCE: in full-step ranges, XSPH and force as one kernel
 *     that walks the neighbour list twice with the corrected velocity kept
 *     in registers;
 *   FLUID_FUSE_SOLVE_LIST: the first PCISPH iteration records, per lane in
 *     its walk order, the candidates within h + 2 dt^2 max_accel of the
 *     particle (no later iteration's in-range pair lies outside), and the
 *     later iterations walk that list instead of the 27 cells.
 * These regroup the XSPH and force sums (results within rounding):
 *   FLUID_WALK_16, FLUID_WALK_32 (at most one): XSPH and force walk the
 *     in-range lists with 16 or 32 lanes per particle (two or four per list
 *     lane) instead of 8;
 *   FLUID_SOLVE_WALK_X2, FLUID_SOLVE_WALK_X4 (at most one; with skin lists,
 *     global sweeps and 2 to 8 solve lanes): the later PCISPH iterations walk
 *     the skin lists with two or four walkers per list lane;
 *   FLUID_SOLVE_FIRST_X2, FLUID_SOLVE_FIRST_X4 (at most one; with skin lists,
 *     global sweeps and at most 8 solve lanes): the first PCISPH iteration
 *     sweeps with two or four lanes per skin-list lane and ranks its entries
 *     into the lists by ballot;
 *   FLUID_LIST_SKIP: a particle with more than 2048 candidates (twice the
 *     in-range list's capacity of 1024) records no list; XSPH and force
 *     sweep its candidates, as for an overflowed list (dense clusters). */
#define FLUID_FUSE_XSPH_FORCE 1u
#define FLUID_FUSE_SOLVE_LIST 2u
#define FLUID_WALK_16 4u
#define FLUID_WALK_32 8u
#define FLUID_SOLVE_WALK_X2 16u
#define FLUID_SOLVE_WALK_X4 32u
#define FLUID_SOLVE_FIRST_X2 64u
#define FLUID_SOLVE_FIRST_X4 128u
#define FLUID_LIST_SKIP 256u
int fluid_set_fusion(fluid_sim *sim, uint32_t flags);
uint32_t fluid_fusion(const fluid_sim *sim);
int fluid_get_tile_config(const fluid_sim *sim, uint32_t out[3]);

/* Autotuner: times every launch configuration (sweep mode: global gathers
 * with either XCD mapping or LDS tiles of 2, 4, 8 cells; lanes per particle)
 * on the current state (each trial runs one step from a saved copy of the
 * state, which is restored afterwards, so the simulation does not advance)
 * and keeps the fastest.  out[3] receives the chosen sweep configuration
 * (may be null); fluid_get_tile_config reports the chosen tiling. */
int fluid_autotune(fluid_sim *sim, uint32_t trials, uint32_t out[3]);

/* Candidate and in-range neighbour pair totals of the current sorted state
 * (the 27-cell sweep of sph_density_3d_grid.wgsl:129-167 counted, not
 * summed); used to turn stage times into flop rates. */
int fluid_count_pairs(fluid_sim *sim, uint64_t *candidates,
                      uint64_t *in_range);

/* As fluid_count_pairs, plus the largest per-particle counts:
 * out[0] candidates, out[1] in-range pairs (self included),
 * out[2] max candidates of one particle, out[3] max in-range of one. */
int fluid_pair_stats(fluid_sim *sim, uint64_t out[4]);

/* ---- Differentiable rollout (reverse mode) -----------------------------
 * No counterpart in the reference, whose step has no gradients; these entry
 * points serve optimisation and inverse-design callers of the step of
 * sph_3d_grid.rs:1342-1408.  The reverse pass differentiates every stage's
 * arithmetic (integrate with its wall forces and backstop, finalize, each
 * PCISPH iteration including the CFL clamp and max(0, rho* - rho0),
 * predict, force, XSPH, density with its wall factor); the grid, neighbour
 * sets and sort order are piecewise constant in the state and are not
 * differentiated.  Mouse force and rigid body must be inactive. */

/* Gradient slots per scene in dl_dparams: viscosity, surface tension,
 * gravity x, y, z. */
#define FLUID_GRAD_PARAMS 5

/* Advance `steps` steps exactly as fluid_step_n does, keeping a state
 * checkpoint every `checkpoint_every` steps (0: ceil(sqrt(steps))) for the
 * reverse pass.  Device memory for the checkpoints and for one segment's
 * tape (about 300 bytes per particle per step at 4 PCISPH iterations) is
 * allocated here and reused. */
int fluid_forward_rollout(fluid_sim *sim, uint32_t steps,
                          uint32_t checkpoint_every);

/* Reverse the last fluid_forward_rollout.  Inputs: dL/d(final positions)
 * and dL/d(final velocities), 3 floats per particle in original index
 * order (host or device memory; null = zero).  Outputs (each may be null):
 * dL/d(initial positions), dL/d(initial velocities) in the same layout, and
 * dl_dparams[FLUID_GRAD_PARAMS * num_scenes] (double).  Each checkpoint
 * segment is replayed to record its tape, then reversed.  The handle's
 * state is the rollout's final state again on return.  Fails with
 * FLUID_ERR_STATE when the state, parameters or sweep configuration changed
 * after the forward rollout. */
int fluid_backward_rollout(fluid_sim *sim, const float *dl_dpos,
                           const float *dl_dvel, float *dl_dpos0,
                           float *dl_dvel0, double *dl_dparams);

/* As fluid_backward_rollout, with the state gradients returned as
 * mantissa and exponent: dL/d(initial positions) = dl_dpos0 x 2^log2_scale
 * [scene] (likewise dl_dvel0), log2_scale being one int32 per scene (not
 * null).  The reverse of a violent flow grows the gradient by 10^1-10^2 per
 * step, past f32's range within ~16 steps; the carried adjoint is kept as
 * mantissa x 2^E throughout (fluid_backward_rollout does the same and only
 * applies 2^E at the end, so its values overflow where the true gradient
 * exceeds f32's range).  dl_dparams receives FLUID_GRAD_PARAMS_ALL values
 * (double) per scene: the five of FLUID_GRAD_PARAMS, then stiffness, which
 * no stage of the step reads (its uniform is copied and never used), so its
 * gradient is exactly zero. */
#define FLUID_GRAD_PARAMS_ALL 6
int fluid_backward_rollout_scaled(fluid_sim *sim, const float *dl_dpos,
                                  const float *dl_dvel, float *dl_dpos0,
                                  float *dl_dvel0, double *dl_dparams,
                                  int32_t *log2_scale);

/* Timing of the last fluid_backward_rollout on the handle's stream (HIP
 * events): ms[0] replays, ms[1] reverse kernels, ms[2] whole call. */
int fluid_rollout_timing(const fluid_sim *sim, float ms[3]);

/* ---- Whitewater: spray, foam and bubble particles ---------------------
 * The reference's SpraySystem (src/simulation/spray.rs:82-491), the step's
 * first consumer: an emit pass over the grid-sorted particles of the last
 * step (spray_emit.wgsl) and a simulate pass over a ring of diffuse
 * particles (spray_simulate.wgsl), both reading the simulation's sorted
 * particles and grid.  A spray handle is bound to one plain (not batched)
 * simulation handle, shares its stream, and must be destroyed before it.
 *
 * Differences from the reference, both removing races:
 *  - spawn slots are handed out in sorted-particle order (particle i's
 *    spawns follow those of particles < i), the order of one sequential
 *    run of the reference's racing atomicAdd (spray_emit.wgsl:335); when
 *    one frame spawns more than the ring holds, the later spawns win;
 *  - the potential ceilings' auto-calibration (spray.rs:40-51, 431-455) runs
 *    on the device after every emit pass, so frame f's emit sees the maxima
 *    of every frame before f (the reference harvests them asynchronously,
 *    one or more frames late, and skips frames while a readback is in
 *    flight).  The ta_limit / wc_limit fields of GpuSprayParams are
 *    therefore not read: the app fills them from auto_limits() anyway
 *    (app.rs:1051-1057). */

/* src/state/interaction.rs:131-153 (80 B) */
typedef struct GpuSprayParams {
  float min_speed;
  float emission_rate;
  float lifetime;
  float lifetime_variation;
  float drag;
  float speed_multiplier;
  float velocity_jitter;
  float dt;
  uint32_t max_particles;
  uint32_t num_sph_particles;
  uint32_t frame_count;
  float gravity_y;
  float k_trapped_air;
  float k_wave_crest;
  float ta_limit;
  float bubble_buoyancy;
  float bubble_drag;
  float wc_limit;
  float _pad[2];
} GpuSprayParams;

/* src/state/interaction.rs:160-174 (48 B); kind 0 spray, 1 foam, 2 bubble */
typedef struct GpuSprayParticle {
  float pos_x, pos_y, pos_z;
  float lifetime;
  float vel_x, vel_y, vel_z;
  float max_lifetime;
  uint32_t kind;
  float age;
  float _pad[2];
} GpuSprayParticle;

/* Device-resident state of a spray handle. */
typedef struct fluid_spray_state {
  uint32_t write_head;  /* spawns so far (mod 2^32); slot = head % max    */
  float ta_limit;       /* auto-calibrated trapped-air ceiling            */
  float wc_limit;       /* auto-calibrated wave-crest ceiling             */
  uint32_t live[4];     /* last simulate pass: total, spray, foam, bubble */
  uint32_t emitted;     /* spawns of the last emit pass                   */
} fluid_spray_state;

typedef struct fluid_spray fluid_spray;

/* SpraySystem::new (spray.rs:105-343): a zeroed (all dead) ring of
 * max_spray particles, write head 0, ceilings at their initial guesses
 * (AUTO_TA_INIT, AUTO_WC_INIT, spray.rs:30-31).  params->max_particles must
 * equal max_spray. */
int fluid_spray_create(fluid_sim *sim, uint32_t max_spray,
                       const GpuSprayParams *params, fluid_spray **out);
int fluid_spray_destroy(fluid_spray *spray);
/* update_params (spray.rs:345-347) */
int fluid_spray_update_params(fluid_spray *spray, const GpuSprayParams *params);
/* step (spray.rs:349-428): emit over the first
 * min(num_sph_particles, params.num_sph_particles, particles of the
 * simulation) sorted particles, then simulate every ring slot. */
int fluid_spray_step(fluid_spray *spray, uint32_t num_sph_particles);
/* The two passes of a step on their own (spray.rs:368-399): emit (with the
 * slot assignment and spawn writes) and simulate (with the ceilings'
 * update from the emit pass's maxima). */
int fluid_spray_emit(fluid_spray *spray, uint32_t num_sph_particles);
int fluid_spray_simulate(fluid_spray *spray);
/* auto_limits (spray.rs:458-460): out = {trapped air, wave crest} */
int fluid_spray_auto_limits(fluid_spray *spray, float out[2]);
/* read_stats (spray.rs:464-475): live counts of the last simulate pass */
int fluid_spray_read_stats(fluid_spray *spray, uint32_t out[4]);
int fluid_spray_read_state(fluid_spray *spray, fluid_spray_state *out);
/* spray_buffer (spray.rs:477-479) as a readback; the write restores a ring
 * (count must be the ring size). */
int fluid_spray_read_particles(fluid_spray *spray, GpuSprayParticle *out,
                               uint32_t count);
int fluid_spray_write_particles(fluid_spray *spray, const GpuSprayParticle *in,
                                uint32_t count);
/* The last emit pass per sorted particle: potentials[4 * i ..] = trapped
 * air, wave crest (after its velocity gate), divergence, in-range neighbour
 * count; spawns[i] = particles spawned.  Zero for particles below
 * min_speed. */
int fluid_spray_read_emit(fluid_spray *spray, float *potentials,
                          uint32_t *spawns, uint32_t count);
/* reset (spray.rs:482-490) */
int fluid_spray_reset(fluid_spray *spray);
/* Wall time of `frames` spray steps after `warmup` untimed ones (HIP
 * events on the stream), over the simulation's current sorted state. */
int fluid_spray_time_steps(fluid_spray *spray, uint32_t num_sph_particles,
                           uint32_t warmup, uint32_t frames, float *ms);

/* ---- Marching-cubes density field: anisotropy and splat ------------------
 * The compute passes of the reference's MarchingCubesRenderer::generate
 * (src/render/marching_cubes.rs:2349-2470) that read the step's sorted
 * particles and grid: pass 0, the per-particle anisotropic kernel fit
 * (mc_anisotropy.wgsl, Yu & Turk), and pass 1, the density field on the MC
 * voxel grid (mc_density.wgsl).  The blur, the surface extraction and the
 * shading are rendering and stay with the caller.  A field handle is bound
 * to one plain simulation handle, shares its stream and must be destroyed
 * before it.  The density field is grid_size^3 floats, x fastest
 * (texture texel (x, y, z) at x + y*G + z*G*G); voxels outside a clipping
 * container hold -1. */

/* src/render/marching_cubes.rs:55-66 (48 B; named GpuGridParams there) */
typedef struct GpuMcGridParams {
  float grid_min[3];
  uint32_t grid_size;
  float grid_max[3];
  float cell_size;
  float kernel_radius;
  float iso_value;
  uint32_t num_particles;
  uint32_t max_vertices;
} GpuMcGridParams;

/* src/render/marching_cubes.rs:41-51 (32 B) */
typedef struct GpuAnisoParams {
  uint32_t enabled;
  float strength;
  float support_radius;
  float h_mc;
  float kr;
  float lambda;
  float max_stretch;
  float max_shift;
} GpuAnisoParams;

/* ParticleAniso, mc_anisotropy.wgsl:46-50 (48 B): (Gxx, Gxy, Gxz, cx),
 * (Gyy, Gyz, Gzz, cy), (cz, reach, amplitude, 0) */
typedef struct GpuParticleAniso {
  float q0[4];
  float q1[4];
  float q2[4];
} GpuParticleAniso;

typedef struct fluid_mc fluid_mc;

/* The field's device resources of MarchingCubesRenderer::new
 * (marching_cubes.rs:388-520): a grid_size^3 density field (1 to 512 per
 * axis) and per-particle ellipsoid records for the simulation's
 * max_particles. */
int fluid_mc_create(fluid_sim *sim, uint32_t grid_size, fluid_mc **out);
int fluid_mc_destroy(fluid_mc *mc);
/* update_params + set_bounds (marching_cubes.rs:2099-2131, 2170-2174);
 * params->grid_size must equal the handle's. */
int fluid_mc_update_params(fluid_mc *mc, const GpuMcGridParams *params);
/* update_aniso_params (marching_cubes.rs:2133-2146) */
int fluid_mc_update_aniso_params(fluid_mc *mc, const GpuAnisoParams *params);
/* update_container_geometry: the MC pass's own copy, with its clip margin
 * (app.rs:1487-1501) */
int fluid_mc_update_container_geometry(fluid_mc *mc, const GpuContainerGeometry *geom);
/* generate, passes 0 and 1 (marching_cubes.rs:2402-2470): the anisotropy
 * fit of the first num_particles sorted particles when enabled, then the
 * density field. */
int fluid_mc_generate(fluid_mc *mc, uint32_t num_particles);
int fluid_mc_read_density(fluid_mc *mc, float *out, uint64_t count);
int fluid_mc_read_aniso(fluid_mc *mc, GpuParticleAniso *out, uint32_t count);
/* Replace the ellipsoid records (count = num_particles of the next
 * generate) and run only the density pass on them. */
int fluid_mc_density_from_aniso(fluid_mc *mc, const GpuParticleAniso *in, uint32_t count);
/* Wall time of `iters` generate calls after `warmup` untimed ones (HIP
 * events on the stream). */
int fluid_mc_time(fluid_mc *mc, uint32_t num_particles, uint32_t warmup, uint32_t iters,
                  float *ms);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): fluid_mc_generate's two passes over all of the
 * simulation's particles without a field handle, allocating nothing.  The
 * ellipsoid records between the passes live in a fixed array of the
 * library (48 bytes x FLUID_MC_NOALLOC_MAX_PARTICLES = 3 MiB per device,
 * held until the process ends), so with anisotropy enabled the simulation
 * may hold at most FLUID_MC_NOALLOC_MAX_PARTICLES particles
 * (FLUID_ERR_CAPACITY beyond; the isotropic splat has no limit).  `field`
 * is a device pointer to grid->grid_size^3 floats (x fastest), written on
 * the simulation's stream; the call returns without waiting.  Every call on
 * one device shares that device's array: calls that overlap on one GPU
 * (other streams, other threads) race. */
#define FLUID_MC_NOALLOC_MAX_PARTICLES 65536u
int fluid_mc_generate_noalloc(fluid_sim *sim, const GpuMcGridParams *grid,
                              const GpuAnisoParams *aniso,
                              const GpuContainerGeometry *container, float *field);

#ifdef __cplusplus
}
#endif

#endif /* FLUID_ABI_H */
