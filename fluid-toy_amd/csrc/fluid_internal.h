// fluid_internal.h — device-side data layout of the SPH step and the launch
// entry points shared by the kernels and the host orchestration.
//
// Layout in HBM (all float4 arrays 16-B aligned, one element per slot):
//   Particle state lives permanently in the cell-sorted order produced by the
//   last reorder ("slot order"); the original particle index of each slot is
//   `sid[slot]`.  The reference's canonical array (original index order,
//   src/simulation/sph_3d_grid.rs:196-201) is materialised only at the ABI
//   edge, by scattering through `sid`.
//     pos, vel, frc   canonical position / velocity / force (.w unused)
//     spos, svel, sfrc  the reference's sorted copy (sorted_particle_buffer)
//     srec            per sorted slot, three float4 written by the density
//                     pass and gathered by XSPH and force:
//                       (pos, density), (vel, near_density),
//                       (normal, 1 / max(near_density, 1))
//     nbr, ncnt       in-range neighbour lists of the density pass, written
//                     for kListLanes = 8 walker lanes: entry e of slot s is
//                     list lane e % 8's entry at depth e / 8, at
//                     nbr[lane_block<8>(s, kListCapacity) + (e / 8) * 64 +
//                     e % 8] (blocks of the 8 slots a wave of 8-lane walkers
//                     covers, so a wave's depth-d reads are one contiguous
//                     run); ncnt[s] is the entry count, > kListCapacity
//                     marking an overflow
//     slist, scnt     the solve's per-lane skin lists (same block layout
//                     for the solve's lane count) and per-lane counts
//     pred[2]         sorted_predicted_{a,b} split into
//                       xq = (x*, q) with q = p / max(rho*,1)^2,
//                       pr = (p, rho*), v = (v*, 0)
//   Grid: counts / starts (inclusive scan) over G^3 cells; cell_of / slot_of
//   per old slot; tmp = (old slot, original index, x key, cell) per new slot.
//   Within a cell, slots are ordered by position x (its order key, ties by
//   original index), so every x-row of cells is one x-sorted slot range
//   and neighbouring slots are close in space.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fluid_abi.h"

static_assert(sizeof(SphParticle3D) == 64, "SphParticle3D is 64 B");
static_assert(offsetof(SphParticle3D, velocity) == 16, "velocity @16");
static_assert(offsetof(SphParticle3D, force) == 32, "force @32");
static_assert(offsetof(SphParticle3D, density) == 44, "density @44");
static_assert(offsetof(SphParticle3D, near_density) == 48, "near @48");
static_assert(offsetof(SphParticle3D, normal) == 52, "normal @52");
static_assert(sizeof(GpuSphParams3D) == 64, "GpuSphParams3D is 64 B");
static_assert(offsetof(GpuSphParams3D, num_particles) == 44, "n @44");
static_assert(sizeof(GpuGridParams) == 48, "GpuGridParams is 48 B");
static_assert(sizeof(GpuContainerGeometry) == 128, "container is 128 B");
static_assert(offsetof(GpuContainerGeometry, wall_stiffness) == 112, "@112");
static_assert(sizeof(GpuGravity) == 16, "GpuGravity is 16 B");
static_assert(sizeof(GpuMouseForce) == 48, "GpuMouseForce is 48 B");
static_assert(offsetof(GpuMouseForce, direction) == 32, "direction @32");
static_assert(sizeof(GpuRigidBody) == 96, "GpuRigidBody is 96 B");
static_assert(offsetof(GpuRigidBody, rot_row0) == 48, "rot @48");
static_assert(sizeof(GpuRigidBodyAccum) == 32, "accum is 32 B");
static_assert(sizeof(PredictedState) == 32, "PredictedState is 32 B");
static_assert(sizeof(GpuSprayParams) == 80, "GpuSprayParams is 80 B");
static_assert(offsetof(GpuSprayParams, wc_limit) == 68, "wc_limit @68");
static_assert(sizeof(GpuSprayParticle) == 48, "GpuSprayParticle is 48 B");
static_assert(offsetof(GpuSprayParticle, kind) == 32, "kind @32");
static_assert(sizeof(GpuMcGridParams) == 48, "GpuMcGridParams is 48 B");
static_assert(sizeof(GpuAnisoParams) == 32, "GpuAnisoParams is 32 B");
static_assert(sizeof(GpuParticleAniso) == 48, "GpuParticleAniso is 48 B");

namespace fluid {

constexpr float kPI = 3.14159265359f;  // const PI of every WGSL file

// Uniform block read by every kernel through one pointer (scalar loads).
// Kernel-constant factors are evaluated on the host in f32 with the same
// operation order the WGSL uses per invocation, so they are bit-identical.
struct DevUniforms {
  // GpuSphParams3D
  float h, h2, h6, h9, mass, rho0, near_k, visc, dt, st, delta, eps;
  // hoisted constants
  float poly6_scale;   // 315 / (64*PI*h9)          sph_density_3d_grid.wgsl:57
  float near_scale;    // 15 / (PI*h6)                                   :63
  float grad_scale;    // 45 / (PI*h6)              sph_force_3d_grid.wgsl:64
  float akinci_scale;  // 32 / (PI*h9)                                   :79
  float h6_64;         // h6 / 64                                        :86
  float half_h;        // h * 0.5                                        :82
  float normal_scale;  // 945 / (32*PI*h8), h8 = h6*h*h  sph_density:177-178
  float m_over_rho0;   // mass / rest_density                            :161
  float min_dist;      // h * 0.1                   sph_force_3d_grid.wgsl:162
  float visc_eps;      // 0.01 * h2                                      :198
  float max_accel;     // 0.25*h / (dt*dt)          pcisph_solve.wgsl:259
  float max_accel_sq;  // max_accel^2                                    :261
  float boundary_layer;  // h * 0.4                 sph_integrate_3d.wgsl:113
  float interact_range;  // h * 0.7                                      :303
  float close_r2;        // smallest float r2 with sqrtf(r2) >= 0.1 h: the
                         // correctly rounded test sqrtf(r2) < 0.1 h (:162)
                         // is exactly r2 < close_r2, sqrtf being monotone
  float solve_skin2;     // (h + 2 dt^2 max_accel)^2 (1 + 1e-3): pairs farther
                         // apart at the first PCISPH iteration stay out of
                         // range in every later one
  // GpuGridParams (cell_offset: first cell of this grid in a batch's
  // concatenated cell array, 0 for a single scene)
  uint32_t gx, gy, gz, cells, gxy, cell_offset;
  float inv_cell, ox, oy, oz;
  // GpuContainerGeometry
  float hw, hh, hd, cy;
  float fw[3][4];
  float iv[3][4];
  float wall_k, damping;
  // GpuGravity
  float grav[3];
  uint32_t sdf_res;
  GpuMouseForce mouse;
  GpuRigidBody rb;
};

// Uniform source of a launch: one block, or (batched handles) one block per
// scene indexed by the slot's scene.  Scenes keep fixed slot ranges, since
// each owns a contiguous run of cells in the concatenated grid.
struct Uni {
  const DevUniforms *u;
  const uint16_t *scene_of;  // null for a single scene
};

// Cell tiles of the LDS-staged sweeps (see sph_device.h).
constexpr uint32_t kTileWMax = 16;
struct TileMap {
  const uint32_t *scene_first;  // first tile id of each scene, nscenes + 1 entries
  uint32_t nscenes;
  uint32_t total;               // tiles of all scenes
  uint32_t w;                   // cells per tile along x
  uint32_t cap;                 // staged entries per workgroup
  uint32_t chunk;               // particles per work item (256 / lanes)
  uint4 *list;                  // work items: (scene, scene-local tile, chunk, 0)
  uint32_t *count;              // entries in list
};

struct PredBuf {
  float4 *xq;
  float2 *pr;
  float4 *v;
};

struct Bufs {
  float4 *pos, *vel, *frc;
  float4 *spos, *svel, *sfrc;
  float4 *srec;
  uint32_t *nbr, *ncnt;
  uint32_t *slist, *scnt;  // per-lane candidate lists of the later PCISPH iterations
  uint32_t *sid;
  uint32_t *gmap;  // new slot -> previous slot of the last reorder, or null
  uint32_t *counts, *starts;  // the step's cell counts (as the scan publishes them), inclusive starts
  uint32_t *cwork;            // the build's per-cell counters, zeroed again by the scan
  uint32_t *cell_of, *slot_of;
  uint4 *tmp;  // reorder entries: (old slot, original index, x order key, cell)
  PredBuf pred[2];
  unsigned long long *tile_status;
  uint32_t *tile_counter;
  uint32_t *err;
  int32_t *accum;
  const float *sdf;
  TileMap tiles;
};

// Scan tiling: 256 threads x 16 cells (long look-back chains are the scan's
// critical path on multi-million-cell batched grids).
constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanItems = 16;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;
// grids of at least this many cells scan with 4x larger tiles (the tile
// status array is sized for the smaller tiles, so either fits)
constexpr uint32_t kScanBigGrid = 1u << 20;

// In-range neighbour entries stored per particle (split over its lanes).
constexpr uint32_t kListCapacity = 1024;
// Candidate count above which a particle records no in-range list under
// FLUID_LIST_SKIP (the density pass marks it overflowed; XSPH and force
// sweep it), and the value that turns the skip off.
constexpr uint32_t kListSkip = 2 * kListCapacity;
constexpr uint32_t kListSkipOff = 0xFFFFFFFFu;
// Solve candidates kept per particle for the later iterations (split over
// its lanes; a lane that overflows sweeps its cells instead).
constexpr uint32_t kSolveListCapacity = 1024;
// Most lanes per particle any sweep runs with (per-lane counts are sized by it).
constexpr uint32_t kMaxLanes = 16;
// Walker lanes the density pass writes the in-range lists for (XSPH and force
// walk them with 8, 16 or 32 lanes per particle).
constexpr int kListLanes = 8;
constexpr uint32_t kSolveListOverflow = 0xFFFFFFFFu;
constexpr uint32_t kListOverflow = 0xFFFFFFFFu;

// Error bits written by kernels into Bufs::err.
constexpr uint32_t kErrScanSpin = 1u;

// Launch configuration of the neighbour sweeps.
struct SweepCfg {
  int lanes;        // lanes per particle in the density sweep (1, 2, 4, 8 or
                    // 16; its list is written for 8 walkers either way)
  int solve_lanes;  // lanes per particle in the PCISPH solve (1 .. 16)
  int xcd_tiles;    // 1: each XCD takes one contiguous run of tiles;
                    // 0: hardware round-robin order
  int tiled;        // 1: density and solve stage cell tiles in LDS, both with
                    // `lanes` lanes per particle, 256 / lanes particles per
                    // work item (Bufs::tiles.chunk)
  int tile_wgs;     // persistent workgroups of the tiled sweeps
  int fuse_xf;      // 1: XSPH and force run as one kernel in full-step ranges
  int solve_list;   // 1: later PCISPH iterations walk the first one's skin list
  int wide_walk;    // 0 / 1 / 2: XSPH and force walk the 8-lane in-range
                    // lists with 8 / 16 / 32 lanes per particle (1, 2 or 4
                    // per list lane, interleaved depths)
  int solve_walk;   // 0 / 1 / 2 (global sweeps, skin lists on, 2 <= solve
                    // lanes <= 8): the later PCISPH iterations walk the skin
                    // lists with 1 / 2 / 4 walkers per list lane
  int solve_first;  // 0 / 1 / 2: the first PCISPH iteration sweeps with 1 / 2 / 4
                    // lanes per skin-list lane (skin lists on, global sweeps,
                    // solve_lanes <= 8)
  uint32_t list_skip = kListSkipOff;  // density records no list for particles
                                      // with more candidates than this
};

// Host-side launchers (sph_kernels.hip).  All enqueue on `st`.
hipError_t launch_clear(const Bufs &b, uint32_t cells, uint32_t ntiles,
                        hipStream_t st);
hipError_t launch_build(const Uni &u, const Bufs &b, uint32_t n,
                        hipStream_t st);
// With c.tiled, also lists the non-empty cell tiles (b.tiles).
hipError_t launch_scan(const Uni &u, const Bufs &b, uint32_t cells, SweepCfg c,
                       hipStream_t st);
hipError_t launch_reorder(const Uni &u, const Bufs &b, uint32_t n,
                          hipStream_t st);
hipError_t launch_density(const Uni &u, const Bufs &b, uint32_t n,
                          SweepCfg c, hipStream_t st);
hipError_t launch_xsph(const Uni &u, const Bufs &b, uint32_t n,
                       SweepCfg c, hipStream_t st);
// fuse_predict: also write pcisph_predict's output for each particle.
hipError_t launch_force(const Uni &u, const Bufs &b, uint32_t n,
                        SweepCfg c, bool fuse_predict, hipStream_t st);
// XSPH and force as one kernel (the same bits as the two launches).
hipError_t launch_xsph_force(const Uni &u, const Bufs &b, uint32_t n,
                             SweepCfg c, bool fuse_predict, hipStream_t st);
hipError_t launch_predict(const Uni &u, const Bufs &b, uint32_t n,
                          hipStream_t st);
// first_iter: the input holds predict's zero pressures (density sum only).
// ap_out (may be null): the iteration's pressure acceleration before the
// CFL clamp, per slot.
hipError_t launch_solve(const Uni &u, const Bufs &b, uint32_t n, const PredBuf &in,
                        const PredBuf &out, float4 *ap_out, SweepCfg c, bool first_iter,
                        hipStream_t st);
hipError_t launch_finalize(const Uni &u, const Bufs &b, uint32_t n,
                           const PredBuf &fin, hipStream_t st);
// fin_v non-null: apply pcisph_finalize from that predicted velocity first.
hipError_t launch_integrate(const Uni &u, const Bufs &b, uint32_t n,
                            const float4 *fin_v, uint32_t ntiles, hipStream_t st);
hipError_t launch_count_pairs(const Uni &u, const Bufs &b, uint32_t n,
                              unsigned long long *out4, hipStream_t st);

// Device state of a spray handle.  The reference keeps the write head in a
// storage buffer and the potential ceilings on the host (spray.rs:85-100);
// here both live on the device.  `emitted` is the last emit pass's spawn
// count, added to `head` by the next emit pass (its spawn kernel reads the
// head while the pass runs).
struct SprayDev {
  uint32_t head;
  float ta_limit, wc_limit;
  uint32_t live[4];        // total, spray, foam, bubble (k_spray_live_sum)
  uint32_t emitted;
  uint32_t frame_max[2];   // f32 bits of the potential maxima since the last simulate
  uint32_t pad[6];
};

struct SprayBufs {
  SprayDev *d;
  float4 *ring;            // GpuSprayParticle as three float4 per slot
  uint32_t cap;            // ring slots
  uint32_t *emit_n;        // spawns per sorted particle
  uint32_t *emit_incl;     // their inclusive scan
  float4 *pot;             // trapped air, crest, divergence, neighbours
  unsigned long long *tile_status;  // scan of emit_n
  uint32_t *tile_counter;
  uint32_t *live_part;     // live counts per simulate block (4 each)
};

// Whitewater passes (spray.hip): emit = potentials, slot scan and spawns
// over the first n sorted particles; simulate = every ring slot.
hipError_t launch_spray_emit(const Uni &u, const Bufs &b, const SprayBufs &sb, uint32_t n,
                             const GpuSprayParams &p, hipStream_t st);
hipError_t launch_spray_simulate(const Uni &u, const Bufs &b, const SprayBufs &sb,
                                 const GpuSprayParams &p, hipStream_t st);
// Sums the last simulate pass's per-block live counts into SprayDev::live.
hipError_t launch_spray_live_sum(const SprayBufs &sb, hipStream_t st);
uint32_t spray_simulate_blocks(uint32_t cap);

// Marching-cubes field passes (mc.hip).  Per-launch constants are derived on
// the host in f32, in the WGSL expressions' order.
struct McLaunch {
  GpuMcGridParams g;
  GpuAnisoParams a;
  GpuContainerGeometry c;
  float peak;        // 315 / (64 pi h_mc^3)          mc_anisotropy.wgsl:154
  float coeff;       // 315 / (64 pi pow(h, 9))       mc_density.wgsl:81
  int aniso_cells;   // ceil(support_radius * inv_cell)   mc_anisotropy.wgsl:173
  int density_cells; // ceil(search_reach * inv_cell)     mc_density.wgsl:151-155
};
hipError_t launch_mc_aniso(const Uni &u, const Bufs &b, const McLaunch &m, float4 *aniso,
                           uint32_t n, hipStream_t st);
hipError_t launch_mc_density(const Uni &u, const Bufs &b, const McLaunch &m,
                             const float4 *aniso, float *field, hipStream_t st);

// The MC passes' per-launch constants (params.cpp, f32 without
// contraction, in the WGSL order), and the app's default parameters.
McLaunch make_mc_launch(const GpuMcGridParams &g, const GpuAnisoParams &a,
                        const GpuContainerGeometry &c, float sph_inv_cell);
GpuAnisoParams default_aniso_params(float h);
GpuMcGridParams default_mc_grid(const GpuContainerGeometry &c, float h, uint32_t grid_size);

// Host-side parameter helpers (params.cpp).
void fill_uniforms(DevUniforms &u, const GpuSphParams3D &p,
                   const GpuGridParams &g, const GpuContainerGeometry &c,
                   const GpuGravity &grav, const GpuMouseForce &mouse,
                   const GpuRigidBody &rb, uint32_t sdf_res,
                   uint32_t cell_offset = 0);
GpuGridParams make_grid(float h, float bounds_extent, uint32_t n);

}  // namespace fluid
