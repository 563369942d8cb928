// fluid_sim.hip — the C ABI of include/fluid_abi.h: host orchestration of the
// SPH step (the role of SphSimulation3DGrid, src/simulation/sph_3d_grid.rs).
//
// Owns every device buffer and one HIP stream per handle.  A full step is
// captured once into a HIP graph (re-captured when the particle count or the
// PCISPH iteration count changes) and replayed per step; uniform updates are
// stream-ordered copies into one device block that the graph reads, so they
// never invalidate it.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "adjoint.h"
#include "fluid_internal.h"

using namespace fluid;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                         \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess)                                                     \
      return fail(FLUID_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

constexpr int kUniformRing = 16;

}  // namespace

// The uniform buffers of one scene in the reference's layouts (the blocks
// behind sph_params_buffer, grid_params_buffer and container_geom_buffer,
// sph_3d_grid.rs:1512-1526, plus gravity), kept on the device for callers
// that bind them.
struct RefBlock {
  GpuSphParams3D sph;
  GpuGridParams grid;
  GpuContainerGeometry container;
  GpuGravity gravity;
};

// A forward rollout kept for its reverse: state checkpoints every `every`
// steps, the final state, one segment's tape and the adjoint buffers.
struct RolloutState {
  bool valid = false;  // cleared by anything that changes the state or uniforms
  uint32_t steps = 0, every = 0, n = 0, iters = 0;
  SweepCfg cfg{};
  // allocation sizes
  uint32_t cap_n = 0, cap_iters = 0, cap_cells = 0, cap_every = 0, cap_ckpt = 0;
  void *ckpt = nullptr;  // cap_ckpt x (pos, vel, sid)
  void *fin = nullptr;   // pos, vel, frc, sid after the rollout
  void *tape = nullptr;  // cap_every tape steps
  void *adj = nullptr;   // adjoint buffers and staging
  size_t ckpt_bytes = 0, tape_bytes = 0;
  float ms[3] = {0.f, 0.f, 0.f};  // last backward: replay, reverse, total
};

// One scene of a handle: its uniforms, its dense grid's place in the
// concatenated cell array, and its fixed range of particle slots.
struct SceneState {
  GpuSphParams3D params{};
  GpuGridParams grid{};
  GpuContainerGeometry cont{};
  GpuGravity grav{};
  uint32_t cell_offset = 0, first = 0, count = 0;
};

struct fluid_sim {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t n = 0, max_n = 0, iters = 5;
  std::vector<SceneState> sc;  // one entry for a plain handle
  uint32_t total_cells = 0;
  uint16_t *d_scene_of = nullptr;  // slot -> scene (batched handles only)
  GpuMouseForce mouse{};
  GpuRigidBody rb{};
  uint32_t sdf_res = 0;
  DevUniforms *d_u = nullptr;     // one block per scene
  DevUniforms *h_ring = nullptr;  // pinned staging ring: kUniformRing x scenes
  RefBlock *d_ref = nullptr;      // one block per scene, reference layouts
  RefBlock *h_ref = nullptr;      // pinned staging ring of d_ref, same slots
  hipEvent_t ring_ev[kUniformRing] = {};
  bool ring_used[kUniformRing] = {};
  int ring_pos = 0;
  Bufs b{};
  std::vector<void *> allocs;
  unsigned long long *d_pairs = nullptr;
  // canonical density / near density / normal per original index; the step
  // never writes these canonical fields (sph_density_3d_grid.wgsl:184-193)
  std::vector<float> extra;
  float *d_extra = nullptr;  // the same five floats per original index, on the device
  hipGraphExec_t graph = nullptr;
  uint32_t graph_n = 0, graph_iters = 0;
  int graph_lanes = 0, graph_solve_lanes = 0, graph_xcd = 0, graph_tiled = 0, graph_wgs = 0;
  uint32_t graph_tile_w = 0;
  int graph_fuse_xf = 0, graph_solve_list = 0, graph_wide = 0, graph_solve_walk = 0;
  int graph_solve_first = 0;
  uint32_t graph_list_skip = kListSkipOff;
  uint32_t num_cu = 256;
  uint32_t *d_scene_first = nullptr;  // first tile id of each scene (+ total)
  int cursor = -1;  // last completed stage of a partial step, -1 = none
  bool sorted_ready = false;  // a grid build and reorder have been enqueued
  // What the grid clear would clear, as far as enqueued work goes: the
  // build's counters (zeroed by every scan) and the scan's tile states
  // (reset by every integrate pass).  A full-step range launches no clear
  // when both are clear; the captured step records whether it has one.
  bool cwork_zero = true, scan_reset = true;
  bool graph_has_clear = false;
  bool fold_clear = true;  // separate work counters (grids below 2^20 cells)
  GpuRigidBodyAccum last_accum{};
  SweepCfg cfg{8, 8, 0};
  bool tuned = false;  // cfg chosen by fluid_set_sweep_lanes / fluid_autotune
  uint32_t ntiles = 0;
  RolloutState roll;
  uint32_t bound = 0;  // spray and MC field handles bound to this simulation

  Uni uni() const { return Uni{d_u, d_scene_of}; }
  bool batched() const { return sc.size() > 1; }
};

namespace {

void *dev_alloc(fluid_sim *s, size_t bytes, hipError_t *err) {
  void *p = nullptr;
  *err = hipMalloc(&p, bytes ? bytes : 16);
  if (*err == hipSuccess) {
    s->allocs.push_back(p);
    (void)hipMemset(p, 0, bytes ? bytes : 16);
  }
  return p;
}

SweepCfg pick_cfg(uint32_t) {
  // Default before tuning: 8 lanes per particle, the fastest setting the
  // autotuner found at every size measured (8k to 1M particles, compressed
  // and relaxed states); global-memory sweeps.
  return SweepCfg{8, 8, 0, 0, 1024, 1, 1, 0, 0, 0};
}

// Per-scene tile counts for tiles of w cells along x; uploads the scene
// offsets and sets b.tiles (the list buffer holds one entry per cell, the
// w = 1 bound).
int setup_tiles(fluid_sim *s, uint32_t w) {
  std::vector<uint32_t> first(s->sc.size() + 1, 0);
  uint64_t total = 0;
  for (size_t k = 0; k < s->sc.size(); ++k) {
    const GpuGridParams &g = s->sc[k].grid;
    first[k] = (uint32_t)total;
    total += (uint64_t)g.grid_size_y * g.grid_size_z * ((g.grid_size_x + w - 1) / w);
  }
  first[s->sc.size()] = (uint32_t)total;
  HIP_TRY(hipMemcpy(s->d_scene_first, first.data(), first.size() * 4, hipMemcpyHostToDevice));
  s->b.tiles.scene_first = s->d_scene_first;
  s->b.tiles.nscenes = (uint32_t)s->sc.size();
  s->b.tiles.total = (uint32_t)total;
  s->b.tiles.w = w;
  return FLUID_OK;
}

void update_cfg(fluid_sim *s) {
  if (s->tuned) return;
  const int wgs = s->cfg.tile_wgs, fuse = s->cfg.fuse_xf, sl = s->cfg.solve_list;
  const int wide = s->cfg.wide_walk, sw = s->cfg.solve_walk, sf = s->cfg.solve_first;
  const uint32_t skip = s->cfg.list_skip;
  s->cfg = pick_cfg(s->n);
  s->cfg.tile_wgs = wgs;
  s->cfg.fuse_xf = fuse;
  s->cfg.solve_list = sl;
  s->cfg.wide_walk = wide;
  s->cfg.solve_walk = sw;
  s->cfg.solve_first = sf;
  s->cfg.list_skip = skip;
}

int push_uniforms(fluid_sim *s) {
  s->roll.valid = false;
  const int slot = s->ring_pos;
  s->ring_pos = (s->ring_pos + 1) % kUniformRing;
  if (s->ring_used[slot]) HIP_TRY(hipEventSynchronize(s->ring_ev[slot]));
  const size_t S = s->sc.size();
  DevUniforms *h = s->h_ring + (size_t)slot * S;
  RefBlock *r = s->h_ref + (size_t)slot * S;
  for (size_t k = 0; k < S; ++k) {
    const SceneState &c = s->sc[k];
    fill_uniforms(h[k], c.params, c.grid, c.cont, c.grav, s->mouse, s->rb, s->sdf_res,
                  c.cell_offset);
    r[k] = RefBlock{c.params, c.grid, c.cont, c.grav};
  }
  HIP_TRY(hipMemcpyAsync(s->d_u, h, S * sizeof(DevUniforms), hipMemcpyHostToDevice,
                         s->stream));
  HIP_TRY(hipMemcpyAsync(s->d_ref, r, S * sizeof(RefBlock), hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipEventRecord(s->ring_ev[slot], s->stream));
  s->ring_used[slot] = true;
  return FLUID_OK;
}

// Enqueue stages [first, last].  Where the range holds both stages of a
// pair, the pair runs as one kernel: force + predict (the force epilogue
// writes the predicted state) and finalize + integrate.
int enqueue_range(fluid_sim *s, uint32_t first, uint32_t last) {
  s->roll.valid = false;
  s->b.tiles.chunk = 256u / (uint32_t)std::min(s->cfg.lanes, 8);
  const uint32_t n = s->n;
  hipStream_t st = s->stream;
  const Bufs &b = s->b;
  const Uni u = s->uni();
  const bool fuse_fp = first <= FLUID_STAGE_FORCE && last >= FLUID_STAGE_PREDICT;
  const bool fuse_xf = s->cfg.fuse_xf && first <= FLUID_STAGE_XSPH && last >= FLUID_STAGE_FORCE;
  const bool fuse_fi = first <= FLUID_STAGE_FINALIZE && last >= FLUID_STAGE_INTEGRATE;
  const int fin_buf = (s->iters % 2 == 0) ? 0 : 1;
  for (uint32_t stage = first; stage <= last; ++stage) {
    switch (stage) {
      case FLUID_STAGE_CLEAR:
        // skipped only where the same range rebuilds and republishes the
        // counts; on its own the stage clears them, as the reference does
        if (!(last >= FLUID_STAGE_SCAN && s->cwork_zero && s->scan_reset)) {
          HIP_TRY(launch_clear(b, s->total_cells, s->ntiles, st));
          s->cwork_zero = s->scan_reset = true;
        }
        break;
      case FLUID_STAGE_BUILD:
        HIP_TRY(launch_build(u, b, n, st));
        s->cwork_zero = false;
        break;
      case FLUID_STAGE_SCAN:
        HIP_TRY(launch_scan(u, b, s->total_cells, s->cfg, st));
        s->cwork_zero = s->fold_clear;  // counted in place: left set
        s->scan_reset = false;
        break;
      case FLUID_STAGE_REORDER:
        HIP_TRY(launch_reorder(u, b, n, st));
        s->sorted_ready = true;
        break;
      case FLUID_STAGE_DENSITY:
        HIP_TRY(launch_density(u, b, n, s->cfg, st));
        break;
      case FLUID_STAGE_XSPH:
        if (fuse_xf)
          HIP_TRY(launch_xsph_force(u, b, n, s->cfg, fuse_fp, st));
        else
          HIP_TRY(launch_xsph(u, b, n, s->cfg, st));
        break;
      case FLUID_STAGE_FORCE:
        if (!fuse_xf) HIP_TRY(launch_force(u, b, n, s->cfg, fuse_fp, st));
        break;
      case FLUID_STAGE_PREDICT:
        if (!fuse_fp) HIP_TRY(launch_predict(u, b, n, st));
        break;
      case FLUID_STAGE_SOLVE:
        for (uint32_t it = 0; it < s->iters; ++it)
          HIP_TRY(launch_solve(u, b, n, b.pred[it % 2], b.pred[1 - it % 2], nullptr, s->cfg,
                               it == 0, st));
        break;
      case FLUID_STAGE_FINALIZE:
        if (!fuse_fi) HIP_TRY(launch_finalize(u, b, n, b.pred[fin_buf], st));
        break;
      case FLUID_STAGE_INTEGRATE:
        HIP_TRY(launch_integrate(u, b, n, fuse_fi ? b.pred[fin_buf].v : nullptr, s->ntiles, st));
        s->scan_reset = true;
        break;
      default:
        return fail(FLUID_ERR_INVALID_ARG, "unknown stage");
    }
  }
  return FLUID_OK;
}

int enqueue_stage(fluid_sim *s, uint32_t stage) { return enqueue_range(s, stage, stage); }

int enqueue_step(fluid_sim *s) { return enqueue_range(s, 0, FLUID_STAGE_COUNT - 1); }

void drop_graph(fluid_sim *s) {
  if (s->graph) (void)hipGraphExecDestroy(s->graph);
  s->graph = nullptr;
}

int ensure_graph(fluid_sim *s) {
  if (s->graph && s->graph_n == s->n && s->graph_iters == s->iters &&
      s->graph_lanes == s->cfg.lanes && s->graph_solve_lanes == s->cfg.solve_lanes &&
      s->graph_xcd == s->cfg.xcd_tiles && s->graph_tiled == s->cfg.tiled &&
      s->graph_wgs == s->cfg.tile_wgs && s->graph_tile_w == s->b.tiles.w &&
      s->graph_fuse_xf == s->cfg.fuse_xf && s->graph_solve_list == s->cfg.solve_list &&
      s->graph_wide == s->cfg.wide_walk && s->graph_solve_walk == s->cfg.solve_walk &&
      s->graph_solve_first == s->cfg.solve_first && s->graph_list_skip == s->cfg.list_skip)
    return FLUID_OK;
  drop_graph(s);
  hipGraph_t g = nullptr;
  // capturing runs nothing: the grid's state stays what it was
  const bool cz = s->cwork_zero, sr = s->scan_reset;
  s->graph_has_clear = !(cz && sr);
  HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_step(s);
  const hipError_t e = hipStreamEndCapture(s->stream, &g);
  s->cwork_zero = cz;
  s->scan_reset = sr;
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(FLUID_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
  const hipError_t ie = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) {
    s->graph = nullptr;
    return fail(FLUID_ERR_HIP, std::string("instantiate: ") + hipGetErrorString(ie));
  }
  s->graph_n = s->n;
  s->graph_iters = s->iters;
  s->graph_lanes = s->cfg.lanes;
  s->graph_solve_lanes = s->cfg.solve_lanes;
  s->graph_xcd = s->cfg.xcd_tiles;
  s->graph_tiled = s->cfg.tiled;
  s->graph_wgs = s->cfg.tile_wgs;
  s->graph_tile_w = s->b.tiles.w;
  s->graph_fuse_xf = s->cfg.fuse_xf;
  s->graph_solve_list = s->cfg.solve_list;
  s->graph_wide = s->cfg.wide_walk;
  s->graph_solve_walk = s->cfg.solve_walk;
  s->graph_solve_first = s->cfg.solve_first;
  s->graph_list_skip = s->cfg.list_skip;
  return FLUID_OK;
}

// One graph-launched step.  A step captured without the grid clear (the
// grid was clear then) gets one first if other work has since left the
// build's counters or the scan's tile states set.
int replay(fluid_sim *s) {
  if (!s->graph_has_clear && !(s->cwork_zero && s->scan_reset))
    HIP_TRY(launch_clear(s->b, s->total_cells, s->ntiles, s->stream));
  HIP_TRY(hipGraphLaunch(s->graph, s->stream));
  s->cwork_zero = s->fold_clear;
  s->scan_reset = true;
  return FLUID_OK;
}

int check_errors(fluid_sim *s) {
  uint32_t err = 0;
  HIP_TRY(hipMemcpyAsync(&err, s->b.err, 4, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (err & kErrScanSpin) return fail(FLUID_ERR_HIP, "prefix scan look-back exceeded its spin cap");
  return FLUID_OK;
}

// Upload particles[0..n) in original-index order: slot i holds index i.
int upload_canonical(fluid_sim *s, const SphParticle3D *p, uint32_t n) {
  std::vector<float4> pos(n), vel(n), frc(n);
  std::vector<uint32_t> sid(n);
  s->extra.assign((size_t)n * 5, 0.0f);
  for (uint32_t i = 0; i < n; ++i) {
    pos[i] = make_float4(p[i].position[0], p[i].position[1], p[i].position[2], 0.f);
    vel[i] = make_float4(p[i].velocity[0], p[i].velocity[1], p[i].velocity[2], 0.f);
    frc[i] = make_float4(p[i].force[0], p[i].force[1], p[i].force[2], 0.f);
    sid[i] = i;
    float *e = &s->extra[(size_t)i * 5];
    e[0] = p[i].density;
    e[1] = p[i].near_density;
    e[2] = p[i].normal[0];
    e[3] = p[i].normal[1];
    e[4] = p[i].normal[2];
  }
  if (n) {
    HIP_TRY(hipMemcpyAsync(s->b.pos, pos.data(), n * 16, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->b.vel, vel.data(), n * 16, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->b.frc, frc.data(), n * 16, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->b.sid, sid.data(), n * 4, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_extra, s->extra.data(), (size_t)n * 5 * 4,
                           hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
  }
  return FLUID_OK;
}

// ------------------------------------------------------------------------
// Rollout storage.  Every piece of a block is 256-byte aligned; a Carve
// with a null base only measures.
struct Carve {
  char *base = nullptr;
  size_t off = 0;
  template <class T>
  T *take(size_t count) {
    const size_t at = (off + 255) & ~(size_t)255;
    off = at + count * sizeof(T);
    return base ? reinterpret_cast<T *>(base + at) : nullptr;
  }
};

struct CkptView {
  float4 *pos, *vel, *frc;
  uint32_t *sid;
};

CkptView carve_state(Carve &c, size_t n, bool with_frc) {
  CkptView v{};
  v.pos = c.take<float4>(n);
  v.vel = c.take<float4>(n);
  v.frc = with_frc ? c.take<float4>(n) : nullptr;
  v.sid = c.take<uint32_t>(n);
  return v;
}

// One step's tape: x, u, w, the density record, the final predicted
// velocity, the grid, the slot map and every PCISPH iterate.
struct TapeSlot {
  float4 *x, *u, *w, *rec, *vK;
  uint32_t *counts, *starts, *gmap;
  std::vector<float4 *> xq;  // iters + 1
  std::vector<float2 *> pr;  // iters + 1
  std::vector<float4 *> ap;  // iters; ap[k] = iteration k's unclamped ap, ap[0] unused
};

TapeSlot carve_tape(Carve &c, size_t n, size_t cells, uint32_t iters) {
  TapeSlot t;
  t.x = c.take<float4>(n);
  t.u = c.take<float4>(n);
  t.w = c.take<float4>(n);
  t.rec = c.take<float4>(3 * n);
  t.vK = c.take<float4>(n);
  t.counts = c.take<uint32_t>(cells);
  t.starts = c.take<uint32_t>(cells);
  t.gmap = c.take<uint32_t>(n);
  for (uint32_t k = 0; k <= iters; ++k) {
    t.xq.push_back(c.take<float4>(n));
    t.pr.push_back(c.take<float2>(n));
  }
  t.ap.push_back(nullptr);
  for (uint32_t k = 1; k < iters; ++k) t.ap.push_back(c.take<float4>(n));
  return t;
}

struct AdjViews {
  float4 *lam_x[2], *lam_v[2];
  float4 *lx, *lw, *la, *lu, *ln;
  float2 *lrho;
  float4 *lap[2];
  float *lP;
  double *acc;
  float4 *scratch_v;                  // non-final solve velocities of a replay
  float *in_x, *in_v, *out_x, *out_v;  // canonical (3 floats per particle) staging
  uint32_t *slot_first;               // scenes + 1 slot offsets
  double *grads;                      // scenes x kGradParams
  uint32_t *pcnt;                     // in-range pair count per slot
  uint32_t *maxbits;                  // 2 x scenes: largest |lambda| per scene (bits)
  int32_t *e_total, *e_step;          // scenes: lambda scale exponents
};

AdjViews carve_adj(Carve &c, size_t n, size_t scenes) {
  AdjViews a{};
  for (int k = 0; k < 2; ++k) {
    a.lam_x[k] = c.take<float4>(n);
    a.lam_v[k] = c.take<float4>(n);
  }
  a.lx = c.take<float4>(n);
  a.lw = c.take<float4>(n);
  a.la = c.take<float4>(n);
  a.lu = c.take<float4>(n);
  a.ln = c.take<float4>(n);
  a.lrho = c.take<float2>(n);
  a.lap[0] = c.take<float4>(n);
  a.lap[1] = c.take<float4>(n);
  a.lP = c.take<float>(n);
  a.acc = c.take<double>((size_t)kAccStride * n);
  a.scratch_v = c.take<float4>(n);
  a.in_x = c.take<float>(3 * n);
  a.in_v = c.take<float>(3 * n);
  a.out_x = c.take<float>(3 * n);
  a.out_v = c.take<float>(3 * n);
  a.slot_first = c.take<uint32_t>(scenes + 1);
  a.grads = c.take<double>(scenes * kGradParams);
  a.pcnt = c.take<uint32_t>(n);
  a.maxbits = c.take<uint32_t>(2 * scenes);
  a.e_total = c.take<int32_t>(scenes);
  a.e_step = c.take<int32_t>(scenes);
  return a;
}

void free_rollout(fluid_sim *s) {
  RolloutState &R = s->roll;
  for (void *p : {R.ckpt, R.fin, R.tape, R.adj})
    if (p) (void)hipFree(p);
  R.ckpt = R.fin = R.tape = R.adj = nullptr;
  R.cap_n = R.cap_iters = R.cap_cells = R.cap_every = R.cap_ckpt = 0;
  R.valid = false;
}

CkptView ckpt_view(fluid_sim *s, uint32_t k) {
  Carve c{(char *)s->roll.ckpt + (size_t)k * s->roll.ckpt_bytes};
  return carve_state(c, s->roll.cap_n, false);
}

CkptView fin_view(fluid_sim *s) {
  Carve c{(char *)s->roll.fin};
  return carve_state(c, s->roll.cap_n, true);
}

TapeSlot tape_slot(fluid_sim *s, uint32_t t) {
  Carve c{(char *)s->roll.tape + (size_t)t * s->roll.tape_bytes};
  return carve_tape(c, s->roll.cap_n, s->roll.cap_cells, s->roll.cap_iters);
}

AdjViews adj_views(fluid_sim *s) {
  Carve c{(char *)s->roll.adj};
  return carve_adj(c, s->roll.cap_n, s->sc.size());
}

// Size the rollout storage for the handle's current n / iterations / grid.
int ensure_rollout(fluid_sim *s, uint32_t every, uint32_t nckpt) {
  RolloutState &R = s->roll;
  const uint32_t n = s->n, iters = s->iters, cells = s->total_cells;
  if (R.adj && R.cap_n == n && R.cap_iters == iters && R.cap_cells == cells &&
      R.cap_every >= every && R.cap_ckpt >= nckpt)
    return FLUID_OK;
  free_rollout(s);
  Carve m;
  carve_state(m, n, false);
  const size_t ck = (m.off + 255) & ~(size_t)255;
  m = Carve{};
  carve_state(m, n, true);
  const size_t fin = m.off;
  m = Carve{};
  carve_tape(m, n, cells, iters);
  const size_t tp = (m.off + 255) & ~(size_t)255;
  m = Carve{};
  carve_adj(m, n, s->sc.size());
  const size_t adj = m.off;
  void *p[4] = {nullptr, nullptr, nullptr, nullptr};
  const size_t bytes[4] = {ck * std::max(nckpt, 1u), fin, tp * std::max(every, 1u), adj};
  for (int k = 0; k < 4; ++k) {
    const hipError_t e = hipMalloc(&p[k], std::max(bytes[k], (size_t)256));
    if (e != hipSuccess) {
      for (int j = 0; j < k; ++j) (void)hipFree(p[j]);
      return fail(FLUID_ERR_HIP, std::string("hipMalloc (rollout storage, ") +
                                     std::to_string(bytes[k] >> 20) + " MiB): " +
                                     hipGetErrorString(e));
    }
  }
  R.ckpt = p[0];
  R.fin = p[1];
  R.tape = p[2];
  R.adj = p[3];
  R.ckpt_bytes = ck;
  R.tape_bytes = tp;
  R.cap_n = n;
  R.cap_iters = iters;
  R.cap_cells = cells;
  R.cap_every = std::max(every, 1u);
  R.cap_ckpt = std::max(nckpt, 1u);
  return FLUID_OK;
}

hipError_t copy_state(fluid_sim *s, const CkptView &dst, const CkptView &src, size_t n) {
  const hipMemcpyKind k = hipMemcpyDeviceToDevice;
  hipError_t e = hipMemcpyAsync(dst.pos, src.pos, n * 16, k, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dst.vel, src.vel, n * 16, k, s->stream);
  if (e == hipSuccess && dst.frc && src.frc)
    e = hipMemcpyAsync(dst.frc, src.frc, n * 16, k, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dst.sid, src.sid, n * 4, k, s->stream);
  return e;
}

CkptView live_state(fluid_sim *s) { return CkptView{s->b.pos, s->b.vel, s->b.frc, s->b.sid}; }

// Replay one full step (the same kernels, fusions and sweep configuration
// as the captured step) with its sorted arrays, grid, slot map and PCISPH
// iterates redirected into tape slot T.
int enqueue_tape_step(fluid_sim *s, const TapeSlot &T, float4 *scratch_v) {
  const RolloutState &R = s->roll;
  const uint32_t n = R.n, iters = R.iters;
  const SweepCfg c = R.cfg;
  const Uni u = s->uni();
  hipStream_t st = s->stream;
  Bufs rb = s->b;
  rb.tiles.chunk = 256u / (uint32_t)std::min(c.lanes, 8);
  rb.spos = T.x;
  rb.svel = T.u;
  rb.srec = T.rec;
  rb.counts = T.counts;
  if (!s->fold_clear) rb.cwork = T.counts;  // counted in place, as the live grid is
  rb.starts = T.starts;
  rb.gmap = T.gmap;
  HIP_TRY(launch_clear(rb, s->total_cells, s->ntiles, st));
  HIP_TRY(launch_build(u, rb, n, st));
  HIP_TRY(launch_scan(u, rb, s->total_cells, c, st));
  HIP_TRY(launch_reorder(u, rb, n, st));
  HIP_TRY(launch_density(u, rb, n, c, st));
  Bufs mb = rb;  // XSPH's output w stays in the tape
  mb.vel = T.w;
  mb.pred[0] = PredBuf{T.xq[0], T.pr[0], scratch_v};
  if (c.fuse_xf) {
    HIP_TRY(launch_xsph_force(u, mb, n, c, true, st));
  } else {
    HIP_TRY(launch_xsph(u, mb, n, c, st));
    HIP_TRY(launch_force(u, mb, n, c, true, st));
  }
  for (uint32_t k = 0; k < iters; ++k) {
    const PredBuf in{T.xq[k], T.pr[k], scratch_v};
    const PredBuf out{T.xq[k + 1], T.pr[k + 1], k + 1 == iters ? T.vK : scratch_v};
    HIP_TRY(launch_solve(u, mb, n, in, out, k >= 1 ? T.ap[k] : nullptr, c, k == 0, st));
  }
  HIP_TRY(launch_integrate(u, rb, n, T.vK, s->ntiles, st));
  return FLUID_OK;
}

int enqueue_adjoint_step(fluid_sim *s, const TapeSlot &T, const AdjViews &A, int cur) {
  const RolloutState &R = s->roll;
  const uint32_t K = R.iters;
  AdjBufs a{};
  a.lin_x = A.lam_x[cur];
  a.lin_v = A.lam_v[cur];
  a.lout_x = A.lam_x[1 - cur];
  a.lout_v = A.lam_v[1 - cur];
  a.lx = A.lx;
  a.lw = A.lw;
  a.la = A.la;
  a.lu = A.lu;
  a.ln = A.ln;
  a.lrho = A.lrho;
  a.lap[0] = A.lap[0];
  a.lap[1] = A.lap[1];
  a.lP = A.lP;
  a.acc = A.acc;
  a.slist = s->b.slist;  // forward scratch, free while a step is reversed
  a.scnt = s->b.scnt;
  a.plist = s->b.nbr;
  a.pcnt = A.pcnt;
  const TapeStep t{T.x, T.u, T.w, T.rec, T.vK, T.counts, T.starts, T.gmap};
  std::vector<TapeIter> it(K);
  for (uint32_t k = 0; k < K; ++k)
    it[k] = TapeIter{T.xq[k], T.pr[k], k >= 2 ? T.ap[k - 1] : nullptr};
  const TapeIter last{T.xq[K], T.pr[K], K >= 2 ? T.ap[K - 1] : nullptr};
  HIP_TRY(launch_adjoint_step(s->uni(), a, t, it.data(), last, K, R.n, R.cfg.lanes, s->stream));
  return FLUID_OK;
}

template <class T>
int d2h(fluid_sim *s, std::vector<T> &out, const void *src, size_t count) {
  out.resize(count);
  if (count)
    HIP_TRY(hipMemcpyAsync(out.data(), src, count * sizeof(T), hipMemcpyDeviceToHost, s->stream));
  return FLUID_OK;
}

// SphParticle3D records (64 B, four float4 stores each) written from slot-
// order state into a caller's device buffer.  Canonical: slot k's particle
// goes to record sid[k], with the canonical density / near density / normal
// kept per original index (the step never writes those fields,
// sph_density_3d_grid.wgsl:184-193).  Sorted: record k is slot k, with the
// fields the density pass left in the neighbour records.
__global__ __launch_bounds__(256) void k_export_canonical(
    const float4 *__restrict__ pos, const float4 *__restrict__ vel,
    const float4 *__restrict__ frc, const uint32_t *__restrict__ sid,
    const float *__restrict__ extra, float4 *__restrict__ out, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = sid[k];
  if (i >= n) return;
  const float4 p = pos[k], v = vel[k], f = frc[k];
  const float *e = extra + (size_t)i * 5;
  float4 *o = out + (size_t)i * 4;
  o[0] = make_float4(p.x, p.y, p.z, 0.0f);
  o[1] = make_float4(v.x, v.y, v.z, 0.0f);
  o[2] = make_float4(f.x, f.y, f.z, e[0]);
  o[3] = make_float4(e[1], e[2], e[3], e[4]);
}

__global__ __launch_bounds__(256) void k_export_sorted(
    const float4 *__restrict__ spos, const float4 *__restrict__ svel,
    const float4 *__restrict__ sfrc, const float4 *__restrict__ srec,
    float4 *__restrict__ out, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float4 p = spos[k], v = svel[k], f = sfrc[k];
  const float rho = srec[3 * (size_t)k].w, near = srec[3 * (size_t)k + 1].w;
  const float4 nr = srec[3 * (size_t)k + 2];
  float4 *o = out + (size_t)k * 4;
  o[0] = make_float4(p.x, p.y, p.z, 0.0f);
  o[1] = make_float4(v.x, v.y, v.z, 0.0f);
  o[2] = make_float4(f.x, f.y, f.z, rho);
  o[3] = make_float4(near, nr.x, nr.y, nr.z);
}

int export_check(fluid_sim *s, const void *dst, uint32_t cap) {
  if (!s || (!dst && s->n)) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (cap < s->n) return fail(FLUID_ERR_CAPACITY, "output smaller than num_particles");
  if (((uintptr_t)dst & 15u) != 0)
    return fail(FLUID_ERR_INVALID_ARG, "output must be 16-byte aligned device memory");
  HIP_TRY(hipSetDevice(s->device));
  return FLUID_OK;
}

}  // namespace

extern "C" {

const char *fluid_last_error(void) { return g_last_error.c_str(); }

int fluid_abi_version(void) { return FLUID_ABI_VERSION; }

static int check_device(int device) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(FLUID_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= ndev) return fail(FLUID_ERR_NO_DEVICE, "device ordinal out of range");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess)
    return fail(FLUID_ERR_NO_DEVICE, "hipGetDeviceProperties failed");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(FLUID_ERR_NO_DEVICE,
                std::string("library is built for gfx950, device is ") + prop.gcnArchName);
  return FLUID_OK;
}

// Shared by fluid_create (one scene) and fluid_batch_create (many): the
// scenes' grids are laid end to end in one cell array and their particles
// (given concatenated, scene-major) end to end in one slot array.
static int create_handle(int device, std::vector<SceneState> scenes,
                         const SphParticle3D *particles, uint32_t n, uint32_t max_n,
                         const float *sdf, uint32_t sdf_res, uint32_t iters, fluid_sim **out) {
  int rc = check_device(device);
  if (rc) return rc;
  uint64_t cells = 0;
  for (SceneState &c : scenes) {
    c.cell_offset = (uint32_t)cells;
    cells += (uint64_t)c.grid.grid_size_x * c.grid.grid_size_y * c.grid.grid_size_z;
  }
  if (cells > (1ull << 28)) return fail(FLUID_ERR_CAPACITY, "grid exceeds 2^28 cells");

  fluid_sim *s = new fluid_sim();
  s->device = device;
  s->n = n;
  s->max_n = std::max(max_n, n);
  s->sc = std::move(scenes);
  s->total_cells = (uint32_t)cells;
  s->mouse = GpuMouseForce{{0.f, 0.f, 0.f}, 0.5f, 30.0f, 0u, 0u, 0.f, {0.f, 0.f, -1.0f}, 0.f};
  s->rb = GpuRigidBody{{0.f, 0.f, 0.f}, 0.15f, {0.f, 0.f, 0.f}, 0u, 200.0f, 0u, 0.f, 0.f,
                       {1.f, 0.f, 0.f, 0.f}, {0.f, 1.f, 0.f, 0.f}, {0.f, 0.f, 1.f, 0.f}};
  s->iters = std::max(iters, 1u);
  s->cfg = pick_cfg(s->n);
  s->ntiles = (s->total_cells + kScanTile - 1) / kScanTile;
  const size_t S = s->sc.size();

  auto bail = [&](int code) {
    fluid_destroy(s);
    return code;
  };
  if (hipSetDevice(s->device) != hipSuccess) return bail(fail(FLUID_ERR_HIP, "hipSetDevice"));
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(FLUID_ERR_HIP, "hipStreamCreate"));
  hipError_t e = hipSuccess;
  const size_t N = s->max_n, C = s->total_cells;
#define ALLOC(field, type, count)                                                 \
  s->b.field = (type)dev_alloc(s, (size_t)(count) * sizeof(*s->b.field), &e);     \
  if (e != hipSuccess) return bail(fail(FLUID_ERR_HIP, std::string("hipMalloc " #field ": ") + hipGetErrorString(e)));
  ALLOC(pos, float4 *, N);
  ALLOC(vel, float4 *, N);
  ALLOC(frc, float4 *, N);
  ALLOC(spos, float4 *, N);
  ALLOC(svel, float4 *, N);
  ALLOC(sfrc, float4 *, N);
  ALLOC(srec, float4 *, 3 * N);
  // per-lane list blocks cover whole waves: round the slots up by 64
  ALLOC(nbr, uint32_t *, (size_t)kListCapacity * (N + 64));
  ALLOC(ncnt, uint32_t *, N);
  ALLOC(slist, uint32_t *, (size_t)kSolveListCapacity * (N + 64));
  ALLOC(scnt, uint32_t *, (size_t)kMaxLanes * N);  // one count per solve lane
  ALLOC(sid, uint32_t *, N);
  ALLOC(counts, uint32_t *, C);
  if (C < kScanBigGrid) {
    ALLOC(cwork, uint32_t *, C);
  } else {
    s->b.cwork = s->b.counts;  // counted in place, cleared every step
    s->fold_clear = false;
  }
  ALLOC(starts, uint32_t *, C);
  ALLOC(cell_of, uint32_t *, N);
  ALLOC(slot_of, uint32_t *, N);
  ALLOC(tmp, uint4 *, N);
  for (int k = 0; k < 2; ++k) {
    ALLOC(pred[k].xq, float4 *, N);
    ALLOC(pred[k].pr, float2 *, N);
    ALLOC(pred[k].v, float4 *, N);
  }
  ALLOC(tile_status, unsigned long long *, s->ntiles + 1);
  ALLOC(tile_counter, uint32_t *, 1);
  ALLOC(err, uint32_t *, 1);
  ALLOC(accum, int32_t *, 8);
  // work items: at most one per chunk of particles plus one per tile
  ALLOC(tiles.list, uint4 *, N + std::max<size_t>(C, 1));
  ALLOC(tiles.count, uint32_t *, 1);
#undef ALLOC
  s->d_scene_first = (uint32_t *)dev_alloc(s, (S + 1) * 4, &e);
  if (e != hipSuccess) return bail(fail(FLUID_ERR_HIP, "hipMalloc tile offsets"));
  s->b.tiles.cap = 2048;
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device) ==
            hipSuccess && cus > 0)
      s->num_cu = (uint32_t)cus;
    s->cfg.tile_wgs = (int)(4 * s->num_cu);
  }
  rc = setup_tiles(s, 4);
  if (rc) return bail(rc);
  s->d_pairs = (unsigned long long *)dev_alloc(s, 32, &e);
  if (e != hipSuccess) return bail(fail(FLUID_ERR_HIP, "hipMalloc pairs"));
  s->d_u = (DevUniforms *)dev_alloc(s, S * sizeof(DevUniforms), &e);
  if (e != hipSuccess) return bail(fail(FLUID_ERR_HIP, "hipMalloc uniforms"));
  s->d_ref = (RefBlock *)dev_alloc(s, S * sizeof(RefBlock), &e);
  if (e != hipSuccess) return bail(fail(FLUID_ERR_HIP, "hipMalloc uniform blocks"));
  s->d_extra = (float *)dev_alloc(s, N * 5 * sizeof(float), &e);
  if (e != hipSuccess) return bail(fail(FLUID_ERR_HIP, "hipMalloc canonical fields"));
  if (S > 1) {
    std::vector<uint16_t> scene_of(N, 0);
    for (size_t k = 0; k < S; ++k)
      for (uint32_t i = 0; i < s->sc[k].count; ++i) scene_of[s->sc[k].first + i] = (uint16_t)k;
    s->d_scene_of = (uint16_t *)dev_alloc(s, N * 2, &e);
    if (e != hipSuccess) return bail(fail(FLUID_ERR_HIP, "hipMalloc scene map"));
    if (hipMemcpy(s->d_scene_of, scene_of.data(), N * 2, hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(FLUID_ERR_HIP, "scene map upload"));
  }
  if (sdf) {
    const size_t nv = (size_t)sdf_res * sdf_res * sdf_res;
    float *d = (float *)dev_alloc(s, nv * 4, &e);
    if (e != hipSuccess) return bail(fail(FLUID_ERR_HIP, "hipMalloc sdf"));
    if (hipMemcpy(d, sdf, nv * 4, hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(FLUID_ERR_HIP, "sdf upload"));
    s->b.sdf = d;
    s->sdf_res = sdf_res;
  }
  if (hipHostMalloc((void **)&s->h_ring, sizeof(DevUniforms) * kUniformRing * S) != hipSuccess ||
      hipHostMalloc((void **)&s->h_ref, sizeof(RefBlock) * kUniformRing * S) != hipSuccess)
    return bail(fail(FLUID_ERR_HIP, "hipHostMalloc"));
  for (int k = 0; k < kUniformRing; ++k)
    if (hipEventCreateWithFlags(&s->ring_ev[k], hipEventDisableTiming) != hipSuccess)
      return bail(fail(FLUID_ERR_HIP, "hipEventCreate"));
  if (hipDeviceSynchronize() != hipSuccess) return bail(fail(FLUID_ERR_HIP, "memset"));
  rc = upload_canonical(s, particles, s->n);
  if (rc) return bail(rc);
  rc = push_uniforms(s);
  if (rc) return bail(rc);
  if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(fail(FLUID_ERR_HIP, "sync"));
  *out = s;
  return FLUID_OK;
}

int fluid_create(const fluid_create_info *info, fluid_sim **out) {
  if (!info || !out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (info->num_particles && !info->particles)
    return fail(FLUID_ERR_INVALID_ARG, "particles is null");
  const GpuSphParams3D &p = info->sph_params;
  if (!(p.kernel_radius > 0.0f) || !(p.dt > 0.0f))
    return fail(FLUID_ERR_INVALID_ARG, "kernel_radius and dt must be > 0");
  if (info->sdf_voxels && info->sdf_resolution == 0)
    return fail(FLUID_ERR_INVALID_ARG, "sdf_resolution is 0");
  const float extent = info->bounds_extent > 0.0f ? info->bounds_extent : 2.0f;
  SceneState c;
  c.params = p;
  c.params.num_particles = info->num_particles;
  c.grid = make_grid(p.kernel_radius, extent, info->num_particles);
  c.cont = info->container;
  c.grav = GpuGravity{{0.0f, -9.8f, 0.0f}, 0.0f};  // sph_3d_grid.rs:277-280
  c.first = 0;
  c.count = info->num_particles;
  return create_handle(info->device, {c}, info->particles, info->num_particles,
                       info->max_particles, info->sdf_voxels, info->sdf_resolution,
                       5 /* sph_3d_grid.rs:1253 */, out);
}

int fluid_batch_create(const fluid_scene_desc *scenes, uint32_t num_scenes,
                       uint32_t pcisph_iterations, int device, fluid_sim **out) {
  if (!scenes || !out || num_scenes == 0) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (num_scenes > 65535) return fail(FLUID_ERR_CAPACITY, "at most 65535 scenes");
  std::vector<SceneState> st(num_scenes);
  uint64_t total = 0;
  for (uint32_t k = 0; k < num_scenes; ++k) {
    const fluid_scene_desc &d = scenes[k];
    if (d.num_particles && !d.particles)
      return fail(FLUID_ERR_INVALID_ARG, "scene particles is null");
    if (!(d.sph_params.kernel_radius > 0.0f) || !(d.sph_params.dt > 0.0f))
      return fail(FLUID_ERR_INVALID_ARG, "kernel_radius and dt must be > 0");
    const float extent = d.bounds_extent > 0.0f ? d.bounds_extent : 2.0f;
    st[k].params = d.sph_params;
    st[k].params.num_particles = d.num_particles;
    st[k].grid = make_grid(d.sph_params.kernel_radius, extent, d.num_particles);
    st[k].cont = d.container;
    st[k].grav = d.gravity;
    st[k].first = (uint32_t)total;
    st[k].count = d.num_particles;
    total += d.num_particles;
  }
  if (total > 0xffffffffull) return fail(FLUID_ERR_CAPACITY, "too many particles");
  std::vector<SphParticle3D> all((size_t)total);
  for (uint32_t k = 0; k < num_scenes; ++k)
    if (scenes[k].num_particles)
      std::memcpy(all.data() + st[k].first, scenes[k].particles,
                  sizeof(SphParticle3D) * scenes[k].num_particles);
  return create_handle(device, std::move(st), all.data(), (uint32_t)total, (uint32_t)total,
                       nullptr, 0, pcisph_iterations, out);
}

uint32_t fluid_batch_num_scenes(const fluid_sim *s) { return s ? (uint32_t)s->sc.size() : 0; }

int fluid_batch_scene_info(const fluid_sim *s, uint32_t scene, uint32_t *first_particle,
                           uint32_t *num_particles, uint32_t *first_cell,
                           GpuGridParams *grid) {
  if (!s || scene >= s->sc.size()) return fail(FLUID_ERR_INVALID_ARG, "bad scene");
  const SceneState &c = s->sc[scene];
  if (first_particle) *first_particle = c.first;
  if (num_particles) *num_particles = c.count;
  if (first_cell) *first_cell = c.cell_offset;
  if (grid) *grid = c.grid;
  return FLUID_OK;
}

int fluid_batch_update_scene(fluid_sim *s, uint32_t scene, const GpuSphParams3D *params,
                             const GpuContainerGeometry *container, const GpuGravity *gravity) {
  if (!s || scene >= s->sc.size()) return fail(FLUID_ERR_INVALID_ARG, "bad scene");
  SceneState &c = s->sc[scene];
  if (params) {
    if (!(params->kernel_radius > 0.0f) || !(params->dt > 0.0f))
      return fail(FLUID_ERR_INVALID_ARG, "kernel_radius and dt must be > 0");
    if (params->kernel_radius != c.grid.cell_size)
      return fail(FLUID_ERR_INVALID_ARG,
                  "kernel_radius differs from the grid cell size fixed at creation");
    c.params = *params;
    c.params.num_particles = c.count;
  }
  if (container) c.cont = *container;
  if (gravity) c.grav = *gravity;
  HIP_TRY(hipSetDevice(s->device));
  return push_uniforms(s);
}

int fluid_destroy(fluid_sim *s) {
  if (!s) return FLUID_OK;
  if (s->bound)
    return fail(FLUID_ERR_STATE, "destroy the simulation's spray and MC field handles first");
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  drop_graph(s);
  free_rollout(s);
  for (void *p : s->allocs) (void)hipFree(p);
  for (int k = 0; k < kUniformRing; ++k)
    if (s->ring_ev[k]) (void)hipEventDestroy(s->ring_ev[k]);
  if (s->h_ring) (void)hipHostFree(s->h_ring);
  if (s->h_ref) (void)hipHostFree(s->h_ref);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return FLUID_OK;
}

int fluid_step(fluid_sim *s) { return fluid_step_n(s, 1); }

int fluid_step_n(fluid_sim *s, uint32_t steps) {
  if (!s) return fail(FLUID_ERR_INVALID_ARG, "null handle");
  if (s->cursor != -1) return fail(FLUID_ERR_STATE, "a partial step is in progress");
  HIP_TRY(hipSetDevice(s->device));
  int rc = ensure_graph(s);
  if (rc) return rc;
  s->roll.valid = false;
  for (uint32_t k = 0; k < steps; ++k) {
    rc = replay(s);
    if (rc) return rc;
  }
  return FLUID_OK;
}

int fluid_step_stages(fluid_sim *s, uint32_t first, uint32_t last) {
  if (!s) return fail(FLUID_ERR_INVALID_ARG, "null handle");
  if (last >= FLUID_STAGE_COUNT || first > last)
    return fail(FLUID_ERR_INVALID_ARG, "bad stage range");
  if ((int)first != s->cursor + 1)
    return fail(FLUID_ERR_STATE, "stages must continue the current step in order");
  HIP_TRY(hipSetDevice(s->device));
  const int rc = enqueue_range(s, first, last);
  if (rc) return rc;
  s->cursor = (last == FLUID_STAGE_COUNT - 1) ? -1 : (int)last;
  return FLUID_OK;
}

int fluid_update_sph_params(fluid_sim *s, const GpuSphParams3D *p) {
  if (!s || !p) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (!(p->kernel_radius > 0.0f) || !(p->dt > 0.0f))
    return fail(FLUID_ERR_INVALID_ARG, "kernel_radius and dt must be > 0");
  if (s->batched()) return fail(FLUID_ERR_STATE, "batched handle: use fluid_batch_update_scene");
  if (p->kernel_radius != s->sc[0].grid.cell_size)
    return fail(FLUID_ERR_INVALID_ARG,
                "kernel_radius differs from the grid cell size fixed at creation");
  HIP_TRY(hipSetDevice(s->device));
  s->sc[0].params = *p;
  s->sc[0].params.num_particles = s->n;
  return push_uniforms(s);
}

int fluid_update_container_geometry(fluid_sim *s, const GpuContainerGeometry *g) {
  if (!s || !g) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  HIP_TRY(hipSetDevice(s->device));
  if (s->batched()) return fail(FLUID_ERR_STATE, "batched handle: use fluid_batch_update_scene");
  s->sc[0].cont = *g;
  return push_uniforms(s);
}

int fluid_update_mouse_force(fluid_sim *s, const GpuMouseForce *m) {
  if (!s || !m) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  HIP_TRY(hipSetDevice(s->device));
  s->mouse = *m;
  return push_uniforms(s);
}

int fluid_update_gravity(fluid_sim *s, const GpuGravity *g) {
  if (!s || !g) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  HIP_TRY(hipSetDevice(s->device));
  if (s->batched()) return fail(FLUID_ERR_STATE, "batched handle: use fluid_batch_update_scene");
  s->sc[0].grav = *g;
  return push_uniforms(s);
}

int fluid_update_rigid_body(fluid_sim *s, const GpuRigidBody *r) {
  if (!s || !r) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  HIP_TRY(hipSetDevice(s->device));
  s->rb = *r;
  return push_uniforms(s);
}

int fluid_set_pcisph_iterations(fluid_sim *s, uint32_t it) {
  if (!s) return fail(FLUID_ERR_INVALID_ARG, "null handle");
  s->iters = std::max(it, 1u);
  s->roll.valid = false;
  return FLUID_OK;
}

uint32_t fluid_pcisph_iterations(const fluid_sim *s) { return s ? s->iters : 0; }

int fluid_clear_rigid_body_accum(fluid_sim *s) {
  if (!s) return fail(FLUID_ERR_INVALID_ARG, "null handle");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipMemsetAsync(s->b.accum, 0, 32, s->stream));
  return FLUID_OK;
}

int fluid_read_rigid_body_accum(fluid_sim *s, GpuRigidBodyAccum *out) {
  if (!s || !out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipMemcpyAsync(&s->last_accum, s->b.accum, 32, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  *out = s->last_accum;
  return check_errors(s);
}

int fluid_spawn_particles(fluid_sim *s, const float position[3], uint32_t count, float spread,
                          uint32_t *spawned) {
  if (!s || !position) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (spawned) *spawned = 0;
  if (s->cursor != -1) return fail(FLUID_ERR_STATE, "a partial step is in progress");
  if (s->batched()) return fail(FLUID_ERR_STATE, "batched handle: scenes have fixed counts");
  const uint32_t avail = s->max_n - s->n;
  const uint32_t c = std::min(count, avail);
  if (c == 0) return FLUID_OK;
  HIP_TRY(hipSetDevice(s->device));
  // golden-angle spiral of sph_3d_grid.rs:1543-1556, in f32
  std::vector<float4> pos(c), zero(c, make_float4(0.f, 0.f, 0.f, 0.f));
  std::vector<uint32_t> sid(c);
  for (uint32_t i = 0; i < c; ++i) {
    const float angle = (float)i * 2.39996f;
    const float r = spread * std::sqrt((float)i / (float)c);
    const uint32_t layer = i / 20;
    pos[i] = make_float4(position[0] + r * std::cos(angle),
                         position[1] + (float)layer * spread * 0.3f,
                         position[2] + r * std::sin(angle), 0.f);
    sid[i] = s->n + i;
  }
  HIP_TRY(hipMemcpyAsync(s->b.pos + s->n, pos.data(), c * 16, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(s->b.vel + s->n, zero.data(), c * 16, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(s->b.frc + s->n, zero.data(), c * 16, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipMemcpyAsync(s->b.sid + s->n, sid.data(), c * 4, hipMemcpyHostToDevice, s->stream));
  s->extra.resize((size_t)(s->n + c) * 5, 0.0f);
  HIP_TRY(hipMemcpyAsync(s->d_extra + (size_t)s->n * 5, s->extra.data() + (size_t)s->n * 5,
                         (size_t)c * 5 * 4, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->n += c;
  s->sc[0].count = s->n;
  s->sc[0].grid.num_particles = s->n;
  s->sc[0].params.num_particles = s->n;
  update_cfg(s);
  if (spawned) *spawned = c;
  return push_uniforms(s);
}

uint32_t fluid_num_particles(const fluid_sim *s) { return s ? s->n : 0; }
uint32_t fluid_max_particles(const fluid_sim *s) { return s ? s->max_n : 0; }

int fluid_read_particles(fluid_sim *s, SphParticle3D *out, uint32_t cap) {
  if (!s || (!out && s->n)) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (cap < s->n) return fail(FLUID_ERR_CAPACITY, "output smaller than num_particles");
  HIP_TRY(hipSetDevice(s->device));
  const int c = s->cursor;
  // Which buffer holds each canonical field at the current stage boundary.
  const bool reordered = c >= FLUID_STAGE_REORDER;
  const float4 *pos = reordered ? s->b.spos : s->b.pos;
  const float4 *vel = !reordered ? s->b.vel : (c >= FLUID_STAGE_XSPH ? s->b.vel : s->b.svel);
  const float4 *frc = !reordered ? s->b.frc : (c >= FLUID_STAGE_FORCE ? s->b.frc : s->b.sfrc);
  std::vector<float4> hp, hv, hf;
  std::vector<uint32_t> sid;
  int rc;
  if ((rc = d2h(s, hp, pos, s->n)) || (rc = d2h(s, hv, vel, s->n)) ||
      (rc = d2h(s, hf, frc, s->n)) || (rc = d2h(s, sid, s->b.sid, s->n)))
    return rc;
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (uint32_t k = 0; k < s->n; ++k) {
    const uint32_t i = sid[k];
    if (i >= s->n) return fail(FLUID_ERR_STATE, "corrupt slot map");
    SphParticle3D &o = out[i];
    std::memset(&o, 0, sizeof(o));
    o.position[0] = hp[k].x; o.position[1] = hp[k].y; o.position[2] = hp[k].z;
    o.velocity[0] = hv[k].x; o.velocity[1] = hv[k].y; o.velocity[2] = hv[k].z;
    o.force[0] = hf[k].x; o.force[1] = hf[k].y; o.force[2] = hf[k].z;
    const float *e = &s->extra[(size_t)i * 5];
    o.density = e[0];
    o.near_density = e[1];
    o.normal[0] = e[2]; o.normal[1] = e[3]; o.normal[2] = e[4];
  }
  return check_errors(s);
}

int fluid_write_particles(fluid_sim *s, const SphParticle3D *in, uint32_t n) {
  if (!s || (!in && n)) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (n > s->max_n) return fail(FLUID_ERR_CAPACITY, "more particles than max_particles");
  if (s->cursor != -1) return fail(FLUID_ERR_STATE, "a partial step is in progress");
  if (s->batched() && n != s->n)
    return fail(FLUID_ERR_STATE, "batched handle: write all scenes' particles at once");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipStreamSynchronize(s->stream));
  const int rc = upload_canonical(s, in, n);
  if (rc) return rc;
  s->n = n;
  s->sc[0].count = n;
  s->sc[0].grid.num_particles = n;
  s->sc[0].params.num_particles = n;
  update_cfg(s);
  return push_uniforms(s);
}

int fluid_read_sorted_particles(fluid_sim *s, SphParticle3D *out, uint32_t cap) {
  if (!s || (!out && s->n)) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (cap < s->n) return fail(FLUID_ERR_CAPACITY, "output smaller than num_particles");
  HIP_TRY(hipSetDevice(s->device));
  std::vector<float4> hp, hv, hf, hr;
  int rc;
  if ((rc = d2h(s, hp, s->b.spos, s->n)) || (rc = d2h(s, hv, s->b.svel, s->n)) ||
      (rc = d2h(s, hf, s->b.sfrc, s->n)) || (rc = d2h(s, hr, s->b.srec, 3 * (size_t)s->n)))
    return rc;
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (uint32_t k = 0; k < s->n; ++k) {
    SphParticle3D &o = out[k];
    std::memset(&o, 0, sizeof(o));
    o.position[0] = hp[k].x; o.position[1] = hp[k].y; o.position[2] = hp[k].z;
    o.velocity[0] = hv[k].x; o.velocity[1] = hv[k].y; o.velocity[2] = hv[k].z;
    o.force[0] = hf[k].x; o.force[1] = hf[k].y; o.force[2] = hf[k].z;
    o.density = hr[3 * (size_t)k].w;
    o.near_density = hr[3 * (size_t)k + 1].w;
    o.normal[0] = hr[3 * (size_t)k + 2].x;
    o.normal[1] = hr[3 * (size_t)k + 2].y;
    o.normal[2] = hr[3 * (size_t)k + 2].z;
  }
  return check_errors(s);
}

int fluid_export_particles_device(fluid_sim *s, SphParticle3D *dst, uint32_t cap) {
  const int rc = export_check(s, dst, cap);
  if (rc || s->n == 0) return rc;
  // the buffers holding each canonical field at the current stage boundary,
  // as fluid_read_particles picks them
  const int c = s->cursor;
  const bool reordered = c >= FLUID_STAGE_REORDER;
  const float4 *pos = reordered ? s->b.spos : s->b.pos;
  const float4 *vel = !reordered ? s->b.vel : (c >= FLUID_STAGE_XSPH ? s->b.vel : s->b.svel);
  const float4 *frc = !reordered ? s->b.frc : (c >= FLUID_STAGE_FORCE ? s->b.frc : s->b.sfrc);
  hipLaunchKernelGGL(k_export_canonical, dim3((s->n + 255) / 256), dim3(256), 0, s->stream, pos,
                     vel, frc, (const uint32_t *)s->b.sid, (const float *)s->d_extra,
                     (float4 *)dst, s->n);
  HIP_TRY(hipGetLastError());
  return FLUID_OK;
}

int fluid_export_sorted_particles_device(fluid_sim *s, SphParticle3D *dst, uint32_t cap) {
  const int rc = export_check(s, dst, cap);
  if (rc || s->n == 0) return rc;
  hipLaunchKernelGGL(k_export_sorted, dim3((s->n + 255) / 256), dim3(256), 0, s->stream,
                     (const float4 *)s->b.spos, (const float4 *)s->b.svel,
                     (const float4 *)s->b.sfrc, (const float4 *)s->b.srec, (float4 *)dst, s->n);
  HIP_TRY(hipGetLastError());
  return FLUID_OK;
}

int fluid_device_buffers(const fluid_sim *s, uint32_t scene, fluid_device_view *out) {
  if (!s || !out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (scene >= s->sc.size()) return fail(FLUID_ERR_INVALID_ARG, "scene out of range");
  const SceneState &c = s->sc[scene];
  out->cell_counts = s->b.counts + c.cell_offset;
  out->cell_starts = s->b.starts + c.cell_offset;
  out->num_cells = c.grid.total_cells;
  out->sorted_to_orig = s->b.sid + c.first;
  out->num_particles = c.count;
  out->first_particle = c.first;
  out->sph_params = &s->d_ref[scene].sph;
  out->grid_params = &s->d_ref[scene].grid;
  out->container_geometry = &s->d_ref[scene].container;
  out->gravity = &s->d_ref[scene].gravity;
  out->stream = (void *)s->stream;
  return FLUID_OK;
}

static int read_u32(fluid_sim *s, const uint32_t *src, size_t count, uint32_t *out,
                    uint32_t cap) {
  if (!s || (!out && count)) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (cap < count) return fail(FLUID_ERR_CAPACITY, "output too small");
  HIP_TRY(hipSetDevice(s->device));
  if (count)
    HIP_TRY(hipMemcpyAsync(out, src, count * 4, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return check_errors(s);
}

int fluid_read_cell_counts(fluid_sim *s, uint32_t *out, uint32_t cap) {
  if (!s) return fail(FLUID_ERR_INVALID_ARG, "null handle");
  // between a build and its scan the counts are still the build's counters
  return read_u32(s, s->cwork_zero || !s->fold_clear ? s->b.counts : s->b.cwork, s->total_cells,
                  out, cap);
}

int fluid_read_cell_starts(fluid_sim *s, uint32_t *out, uint32_t cap) {
  if (!s) return fail(FLUID_ERR_INVALID_ARG, "null handle");
  return read_u32(s, s->b.starts, s->total_cells, out, cap);
}

int fluid_read_sorted_to_orig(fluid_sim *s, uint32_t *out, uint32_t cap) {
  if (!s) return fail(FLUID_ERR_INVALID_ARG, "null handle");
  return read_u32(s, s->b.sid, s->n, out, cap);
}

int fluid_read_orig_to_sorted(fluid_sim *s, uint32_t *out, uint32_t cap) {
  if (!s || (!out && s->n)) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (cap < s->n) return fail(FLUID_ERR_CAPACITY, "output too small");
  std::vector<uint32_t> sid(s->n);
  const int rc = read_u32(s, s->b.sid, s->n, sid.data(), s->n);
  if (rc) return rc;
  for (uint32_t k = 0; k < s->n; ++k) {
    if (sid[k] >= s->n) return fail(FLUID_ERR_STATE, "corrupt slot map");
    out[sid[k]] = k;
  }
  return FLUID_OK;
}

int fluid_read_predicted(fluid_sim *s, uint32_t which, PredictedState *out, uint32_t cap) {
  if (!s || (!out && s->n) || which > 1) return fail(FLUID_ERR_INVALID_ARG, "bad argument");
  if (cap < s->n) return fail(FLUID_ERR_CAPACITY, "output too small");
  HIP_TRY(hipSetDevice(s->device));
  std::vector<float4> xq, v;
  std::vector<float2> pr;
  int rc;
  if ((rc = d2h(s, xq, s->b.pred[which].xq, s->n)) || (rc = d2h(s, pr, s->b.pred[which].pr, s->n)) ||
      (rc = d2h(s, v, s->b.pred[which].v, s->n)))
    return rc;
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (uint32_t k = 0; k < s->n; ++k) {
    PredictedState &o = out[k];
    o.pred_pos[0] = xq[k].x; o.pred_pos[1] = xq[k].y; o.pred_pos[2] = xq[k].z;
    o.pressure = pr[k].x;
    o.pred_vel[0] = v[k].x; o.pred_vel[1] = v[k].y; o.pred_vel[2] = v[k].z;
    o.pred_density = pr[k].y;
  }
  return check_errors(s);
}

int fluid_read_grid_params(const fluid_sim *s, GpuGridParams *out) {
  if (!s || !out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  *out = s->sc[0].grid;
  return FLUID_OK;
}

int fluid_read_sph_params(const fluid_sim *s, GpuSphParams3D *out) {
  if (!s || !out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  *out = s->sc[0].params;
  return FLUID_OK;
}

int fluid_read_container_geometry(const fluid_sim *s, GpuContainerGeometry *out) {
  if (!s || !out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  *out = s->sc[0].cont;
  return FLUID_OK;
}

int fluid_sync(fluid_sim *s) {
  if (!s) return fail(FLUID_ERR_INVALID_ARG, "null handle");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return check_errors(s);
}

void *fluid_stream(fluid_sim *s) { return s ? (void *)s->stream : nullptr; }

int fluid_time_steps(fluid_sim *s, uint32_t warmup, uint32_t steps, float *ms_total) {
  if (!s || !ms_total) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  int rc = fluid_step_n(s, warmup);
  if (rc) return rc;
  rc = ensure_graph(s);
  if (rc) return rc;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (!s->graph_has_clear && !(s->cwork_zero && s->scan_reset)) {
    HIP_TRY(launch_clear(s->b, s->total_cells, s->ntiles, s->stream));
    s->cwork_zero = s->scan_reset = true;
  }
  HIP_TRY(hipEventRecord(e0, s->stream));
  for (uint32_t k = 0; k < steps; ++k) {
    rc = replay(s);
    if (rc) return rc;
  }
  HIP_TRY(hipEventRecord(e1, s->stream));
  HIP_TRY(hipEventSynchronize(e1));
  HIP_TRY(hipEventElapsedTime(ms_total, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return check_errors(s);
}

int fluid_time_stages(fluid_sim *s, uint32_t steps, float *ms_per_stage) {
  if (!s || !ms_per_stage) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (s->cursor != -1) return fail(FLUID_ERR_STATE, "a partial step is in progress");
  HIP_TRY(hipSetDevice(s->device));
  std::vector<hipEvent_t> ev(2 * FLUID_STAGE_COUNT * std::max(steps, 1u));
  for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
  for (uint32_t k = 0; k < FLUID_STAGE_COUNT; ++k) ms_per_stage[k] = 0.f;
  size_t q = 0;
  for (uint32_t step = 0; step < steps; ++step) {
    for (uint32_t st = 0; st < FLUID_STAGE_COUNT; ++st) {
      HIP_TRY(hipEventRecord(ev[q++], s->stream));
      const int rc = enqueue_stage(s, st);
      if (rc) return rc;
      HIP_TRY(hipEventRecord(ev[q++], s->stream));
    }
  }
  HIP_TRY(hipStreamSynchronize(s->stream));
  q = 0;
  for (uint32_t step = 0; step < steps; ++step)
    for (uint32_t st = 0; st < FLUID_STAGE_COUNT; ++st) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, ev[q], ev[q + 1]));
      q += 2;
      ms_per_stage[st] += ms / (float)steps;
    }
  for (auto &e : ev) (void)hipEventDestroy(e);
  return check_errors(s);
}

int fluid_time_range(fluid_sim *s, uint32_t first, uint32_t last, uint32_t reps,
                     float *ms_per_rep) {
  if (!s || !ms_per_rep) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (first < FLUID_STAGE_DENSITY || last > FLUID_STAGE_PREDICT || first > last || reps == 0)
    return fail(FLUID_ERR_INVALID_ARG, "time range: stages within DENSITY..PREDICT, reps > 0");
  if (s->cursor != -1) return fail(FLUID_ERR_STATE, "a partial step is in progress");
  if (!s->sorted_ready) return fail(FLUID_ERR_STATE, "no step to re-run stages of");
  HIP_TRY(hipSetDevice(s->device));
  const size_t n = s->n;
  *ms_per_rep = 0.f;
  if (n == 0) return FLUID_OK;
  // velocity, force and the first predicted buffer are rewritten by the
  // range: kept and put back
  float4 *keep = nullptr;
  HIP_TRY(hipMalloc(&keep, 4 * n * sizeof(float4) + n * sizeof(float2)));
  float2 *keep_pr = reinterpret_cast<float2 *>(keep + 4 * n);
  const Bufs &b = s->b;
  float4 *const live[4] = {b.vel, b.frc, b.pred[0].xq, b.pred[0].v};
  hipError_t e = hipSuccess;
  for (int k = 0; k < 4 && e == hipSuccess; ++k)
    e = hipMemcpyAsync(keep + k * n, live[k], n * sizeof(float4), hipMemcpyDeviceToDevice,
                       s->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(keep_pr, b.pred[0].pr, n * sizeof(float2), hipMemcpyDeviceToDevice,
                       s->stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  int rc = FLUID_OK;
  if (e == hipSuccess) e = hipEventRecord(e0, s->stream);
  for (uint32_t r = 0; r < reps && e == hipSuccess && rc == FLUID_OK; ++r)
    rc = enqueue_range(s, first, last);
  if (e == hipSuccess && rc == FLUID_OK) e = hipEventRecord(e1, s->stream);
  if (e == hipSuccess && rc == FLUID_OK) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess && rc == FLUID_OK) e = hipEventElapsedTime(&ms, e0, e1);
  for (int k = 0; k < 4; ++k)
    (void)hipMemcpyAsync(live[k], keep + k * n, n * sizeof(float4), hipMemcpyDeviceToDevice,
                         s->stream);
  (void)hipMemcpyAsync(b.pred[0].pr, keep_pr, n * sizeof(float2), hipMemcpyDeviceToDevice,
                       s->stream);
  const hipError_t es = hipStreamSynchronize(s->stream);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(keep);
  if (rc) return rc;
  if (e != hipSuccess || es != hipSuccess)
    return fail(FLUID_ERR_HIP, std::string("time range: ") +
                                   hipGetErrorString(e != hipSuccess ? e : es));
  *ms_per_rep = ms / (float)reps;
  return check_errors(s);
}

static bool valid_lanes(uint32_t l) { return l == 1 || l == 2 || l == 4 || l == 8 || l == 16; }

int fluid_set_sweep_config(fluid_sim *s, const uint32_t cfg[3]) {
  if (!s || !cfg) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if ((cfg[0] && !valid_lanes(cfg[0])) || (cfg[1] && !valid_lanes(cfg[1])))
    return fail(FLUID_ERR_INVALID_ARG, "lanes must be 1, 2, 4, 8 or 16");
  if (cfg[2] > 1) return fail(FLUID_ERR_INVALID_ARG, "xcd mapping must be 0 or 1");
  if (s->cursor != -1) return fail(FLUID_ERR_STATE, "a partial step is in progress");
  if (cfg[0]) s->cfg.lanes = (int)cfg[0];
  if (cfg[1]) s->cfg.solve_lanes = (int)cfg[1];
  s->cfg.xcd_tiles = (int)cfg[2];
  s->tuned = true;
  s->roll.valid = false;
  return FLUID_OK;
}

int fluid_set_tile_config(fluid_sim *s, const uint32_t cfg[3]) {
  if (!s || !cfg) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (cfg[0] > 1) return fail(FLUID_ERR_INVALID_ARG, "tiled must be 0 or 1");
  if (cfg[1] > kTileWMax) return fail(FLUID_ERR_INVALID_ARG, "tile width above 16 cells");
  if (s->cursor != -1) return fail(FLUID_ERR_STATE, "a partial step is in progress");
  HIP_TRY(hipSetDevice(s->device));
  s->cfg.tiled = (int)cfg[0];
  if (cfg[1] && cfg[1] != s->b.tiles.w) {
    HIP_TRY(hipStreamSynchronize(s->stream));
    const int rc = setup_tiles(s, cfg[1]);
    if (rc) return rc;
  }
  if (cfg[2]) s->cfg.tile_wgs = (int)cfg[2];
  s->tuned = true;
  s->roll.valid = false;
  return FLUID_OK;
}

int fluid_set_fusion(fluid_sim *s, uint32_t flags) {
  if (!s) return fail(FLUID_ERR_INVALID_ARG, "null handle");
  if (flags & ~(FLUID_FUSE_XSPH_FORCE | FLUID_FUSE_SOLVE_LIST | FLUID_WALK_16 | FLUID_WALK_32 |
                FLUID_SOLVE_WALK_X2 | FLUID_SOLVE_WALK_X4 | FLUID_SOLVE_FIRST_X2 |
                FLUID_SOLVE_FIRST_X4 | FLUID_LIST_SKIP))
    return fail(FLUID_ERR_INVALID_ARG, "unknown fusion flag");
  if ((flags & FLUID_WALK_16) && (flags & FLUID_WALK_32))
    return fail(FLUID_ERR_INVALID_ARG, "FLUID_WALK_16 and FLUID_WALK_32 exclude each other");
  if ((flags & FLUID_SOLVE_WALK_X2) && (flags & FLUID_SOLVE_WALK_X4))
    return fail(FLUID_ERR_INVALID_ARG,
                "FLUID_SOLVE_WALK_X2 and FLUID_SOLVE_WALK_X4 exclude each other");
  if ((flags & FLUID_SOLVE_FIRST_X2) && (flags & FLUID_SOLVE_FIRST_X4))
    return fail(FLUID_ERR_INVALID_ARG,
                "FLUID_SOLVE_FIRST_X2 and FLUID_SOLVE_FIRST_X4 exclude each other");
  if (s->cursor != -1) return fail(FLUID_ERR_STATE, "a partial step is in progress");
  s->cfg.fuse_xf = (flags & FLUID_FUSE_XSPH_FORCE) ? 1 : 0;
  s->cfg.solve_list = (flags & FLUID_FUSE_SOLVE_LIST) ? 1 : 0;
  s->cfg.wide_walk = (flags & FLUID_WALK_16) ? 1 : (flags & FLUID_WALK_32) ? 2 : 0;
  s->cfg.solve_walk = (flags & FLUID_SOLVE_WALK_X2) ? 1 : (flags & FLUID_SOLVE_WALK_X4) ? 2 : 0;
  s->cfg.solve_first = (flags & FLUID_SOLVE_FIRST_X2) ? 1 : (flags & FLUID_SOLVE_FIRST_X4) ? 2 : 0;
  s->cfg.list_skip = (flags & FLUID_LIST_SKIP) ? kListSkip : kListSkipOff;
  s->roll.valid = false;
  return FLUID_OK;
}

uint32_t fluid_fusion(const fluid_sim *s) {
  if (!s) return 0u;
  return (s->cfg.fuse_xf ? FLUID_FUSE_XSPH_FORCE : 0u) |
         (s->cfg.solve_list ? FLUID_FUSE_SOLVE_LIST : 0u) |
         (s->cfg.wide_walk == 1 ? FLUID_WALK_16 : 0u) |
         (s->cfg.wide_walk == 2 ? FLUID_WALK_32 : 0u) |
         (s->cfg.solve_walk == 1 ? FLUID_SOLVE_WALK_X2 : 0u) |
         (s->cfg.solve_walk == 2 ? FLUID_SOLVE_WALK_X4 : 0u) |
         (s->cfg.solve_first == 1 ? FLUID_SOLVE_FIRST_X2 : 0u) |
         (s->cfg.solve_first == 2 ? FLUID_SOLVE_FIRST_X4 : 0u) |
         (s->cfg.list_skip != kListSkipOff ? FLUID_LIST_SKIP : 0u);
}

int fluid_get_tile_config(const fluid_sim *s, uint32_t out[3]) {
  if (!s || !out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  out[0] = (uint32_t)s->cfg.tiled;
  out[1] = s->b.tiles.w;
  out[2] = (uint32_t)s->cfg.tile_wgs;
  return FLUID_OK;
}

int fluid_get_sweep_config(const fluid_sim *s, uint32_t out[3]) {
  if (!s || !out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  out[0] = (uint32_t)s->cfg.lanes;
  out[1] = (uint32_t)s->cfg.solve_lanes;
  out[2] = (uint32_t)s->cfg.xcd_tiles;
  return FLUID_OK;
}

int fluid_autotune(fluid_sim *s, uint32_t trials, uint32_t out[3]) {
  if (!s) return fail(FLUID_ERR_INVALID_ARG, "null handle");
  if (s->cursor != -1) return fail(FLUID_ERR_STATE, "a partial step is in progress");
  HIP_TRY(hipSetDevice(s->device));
  trials = std::max(trials, 1u);
  const size_t n = s->n;
  if (n == 0) return FLUID_OK;
  // Saved state: canonical position, velocity, force and the slot map.
  float4 *save = nullptr;
  uint32_t *save_sid = nullptr;
  HIP_TRY(hipMalloc(&save, 3 * n * sizeof(float4)));
  if (hipMalloc(&save_sid, n * 4) != hipSuccess) {
    (void)hipFree(save);
    return fail(FLUID_ERR_HIP, "hipMalloc (autotune)");
  }
  auto copy = [&](bool to_save) -> hipError_t {
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    float4 *a[3] = {s->b.pos, s->b.vel, s->b.frc};
    for (int i = 0; i < 3; ++i) {
      hipError_t e = to_save ? hipMemcpyAsync(save + i * n, a[i], n * 16, k, s->stream)
                             : hipMemcpyAsync(a[i], save + i * n, n * 16, k, s->stream);
      if (e != hipSuccess) return e;
    }
    return to_save ? hipMemcpyAsync(save_sid, s->b.sid, n * 4, k, s->stream)
                   : hipMemcpyAsync(s->b.sid, save_sid, n * 4, k, s->stream);
  };
  int32_t accum_keep[8];
  HIP_TRY(hipMemcpyAsync(accum_keep, s->b.accum, 32, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  const SweepCfg keep = s->cfg;
  const uint32_t keep_w = s->b.tiles.w;
  const int lanes[5] = {1, 2, 4, 8, 16};
  // Sweep modes: global gathers with either workgroup-to-XCD mapping, and
  // LDS-staged tiles of 2, 4 or 8 cells.  Within a mode, lanes per particle
  // for density/XSPH/force and for the solve are timed independently.
  struct Mode { int tiled, xcd; uint32_t w; };
  const Mode modes[5] = {{0, 0, 0}, {0, 1, 0}, {1, 0, 2}, {1, 0, 4}, {1, 0, 8}};
  float best_total = 1e30f;
  SweepCfg best = keep;
  uint32_t best_w = keep_w;
  int rc = FLUID_OK;
  hipError_t e = copy(true);
  for (int mi = 0; mi < 5 && e == hipSuccess && rc == FLUID_OK; ++mi) {
    const Mode &m = modes[mi];
    if (m.tiled && m.w != s->b.tiles.w) {
      e = hipStreamSynchronize(s->stream);
      if (e != hipSuccess || (rc = setup_tiles(s, m.w)) != FLUID_OK) break;
    }
    float ml = 1e30f, ms_solve = 1e30f;
    SweepCfg mc = keep;
    mc.tiled = m.tiled;
    mc.xcd_tiles = m.xcd;
    mc.wide_walk = 0;  // decided with the fusion below
    mc.solve_walk = 0;
    mc.solve_first = 0;
    mc.list_skip = kListSkipOff;  // decided with the walks below
    for (int li = 0; li < 5 && e == hipSuccess && rc == FLUID_OK; ++li) {
      if (m.tiled && lanes[li] > 8) continue;  // tiled sweeps run at most 8 lanes
      s->cfg = mc;
      s->cfg.lanes = lanes[li];
      s->cfg.solve_lanes = lanes[li];
      float ms[FLUID_STAGE_COUNT];
      float tl = 0.f, ts = 0.f;
      for (uint32_t t = 0; t < trials && rc == FLUID_OK; ++t) {
        if ((e = copy(false)) != hipSuccess) break;
        rc = fluid_time_stages(s, 1, ms);
        tl += ms[FLUID_STAGE_DENSITY] + ms[FLUID_STAGE_XSPH] + ms[FLUID_STAGE_FORCE];
        ts += ms[FLUID_STAGE_SOLVE];
      }
      if (m.tiled) {
        // one lane count for density, XSPH, force and solve
        if (tl + ts < ml + ms_solve) {
          ml = tl;
          ms_solve = ts;
          mc.lanes = mc.solve_lanes = lanes[li];
        }
      } else {
        if (tl < ml) { ml = tl; mc.lanes = lanes[li]; }
        if (ts < ms_solve) { ms_solve = ts; mc.solve_lanes = lanes[li]; }
      }
    }
    if (ml + ms_solve < best_total) {
      best_total = ml + ms_solve;
      best = mc;
      best_w = m.tiled ? m.w : keep_w;
    }
  }
  if (e == hipSuccess && rc == FLUID_OK && best_w != s->b.tiles.w) {
    e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) rc = setup_tiles(s, best_w);
  }
  // XSPH + force as one kernel or two, and their 8-, 16- or 32-lane list
  // walk: timed as whole graph-launched steps (the fusion only exists in
  // full-step ranges).  Every whole-step trial below first launches its
  // graph once untimed: a graph's first launch after its capture costs more
  // than the rest and would tilt near-equal choices.
  {
    float best_f = 1e30f;
    int fuse = best.fuse_xf, wide = 0;
    const int nwide = 3;
    for (int f = 0; f < 2 && e == hipSuccess && rc == FLUID_OK; ++f) {
      for (int wv = 0; wv < nwide && e == hipSuccess && rc == FLUID_OK; ++wv) {
        s->cfg = best;
        s->cfg.fuse_xf = f;
        s->cfg.wide_walk = wv;
        float tot = 0.f, ms = 0.f;
        for (uint32_t t = 0; t < trials && rc == FLUID_OK; ++t) {
          if ((e = copy(false)) != hipSuccess) break;
          rc = fluid_time_steps(s, 1, 2, &ms);
          tot += ms;
        }
        if (tot < best_f) {
          best_f = tot;
          fuse = f;
          wide = wv;
        }
      }
    }
    best.fuse_xf = fuse;
    best.wide_walk = wide;
  }
  // lists for candidate-heavy particles or not, as whole steps: skipping
  // them saves the density pass's writes and moves those particles' XSPH
  // and force work to sweeps
  {
    float tm[2] = {0.f, 0.f}, ms = 0.f;
    for (int v = 0; v < 2 && e == hipSuccess && rc == FLUID_OK; ++v) {
      s->cfg = best;
      s->cfg.list_skip = v ? kListSkip : kListSkipOff;
      for (uint32_t t = 0; t < trials && rc == FLUID_OK; ++t) {
        if ((e = copy(false)) != hipSuccess) break;
        rc = fluid_time_steps(s, 1, 2, &ms);
        tm[v] += ms;
      }
    }
    best.list_skip = tm[1] < tm[0] ? kListSkip : kListSkipOff;
  }
  // 16 solve lanes against 8, as whole steps: the per-stage timing above
  // often puts them within a few per cent, and only 8 lanes go on to the
  // wider walks and first-iteration sweeps tuned below
  if (best.solve_list && !best.tiled && best.solve_lanes == 16) {
    float t16 = 0.f, t8 = 0.f, ms = 0.f;
    for (int v = 0; v < 2 && e == hipSuccess && rc == FLUID_OK; ++v) {
      s->cfg = best;
      s->cfg.solve_lanes = v ? 8 : 16;
      s->cfg.solve_first = v ? 1 : 0;
      for (uint32_t t = 0; t < trials && rc == FLUID_OK; ++t) {
        if ((e = copy(false)) != hipSuccess) break;
        rc = fluid_time_steps(s, 1, 2, &ms);
        (v ? t8 : t16) += ms;
      }
    }
    if (t8 < t16) best.solve_lanes = 8;
  }
  // the later PCISPH iterations' walkers per skin-list lane, as whole steps
  if (best.solve_list && !best.tiled && best.solve_lanes >= 2 && best.solve_lanes <= 8) {
    float best_f = 1e30f;
    int sw = 0;
    for (int v = 0; v < 3 && e == hipSuccess && rc == FLUID_OK; ++v) {
      s->cfg = best;
      s->cfg.solve_walk = v;
      float tot = 0.f, ms = 0.f;
      for (uint32_t t = 0; t < trials && rc == FLUID_OK; ++t) {
        if ((e = copy(false)) != hipSuccess) break;
        rc = fluid_time_steps(s, 1, 2, &ms);
        tot += ms;
      }
      // one walker per list lane sums in the sweep's order; the wider walks
      // regroup the sums, so they must win by 2% to be taken
      if (v == 0 || tot < 0.98f * best_f) {
        best_f = tot;
        sw = v;
      }
    }
    best.solve_walk = sw;
  }
  // the first PCISPH iteration's sweep lanes per skin-list lane, as whole
  // steps; the wider sweeps regroup the sums, so they must win by 1% (their
  // measured gains are 0.7-1.5%), four lanes by 3%: their gain at the tuned
  // state has not held over the 50k dam break's later flow (1-2% slower over
  // 1000 steps)
  if (best.solve_list && !best.tiled && best.solve_lanes <= 8) {
    float best_f = 1e30f;
    int sf = 0;
    for (int v = 0; v < 3 && e == hipSuccess && rc == FLUID_OK; ++v) {
      s->cfg = best;
      s->cfg.solve_first = v;
      float tot = 0.f, ms = 0.f;
      for (uint32_t t = 0; t < trials && rc == FLUID_OK; ++t) {
        if ((e = copy(false)) != hipSuccess) break;
        rc = fluid_time_steps(s, 1, 2, &ms);
        tot += ms;
      }
      if (v == 0 || tot < (v == 2 ? 0.97f : 0.99f) * best_f) {
        best_f = tot;
        sf = v;
      }
    }
    best.solve_first = sf;
  }
  // the workgroup-to-XCD mapping last, as whole steps with everything else
  // chosen: the per-stage sums that picked it differ by about as much as
  // the mapping changes the step
  if (!best.tiled && e == hipSuccess && rc == FLUID_OK) {
    float tm[2] = {0.f, 0.f}, ms = 0.f;
    for (int v = 0; v < 2 && e == hipSuccess && rc == FLUID_OK; ++v) {
      s->cfg = best;
      s->cfg.xcd_tiles = v;
      for (uint32_t t = 0; t < trials && rc == FLUID_OK; ++t) {
        if ((e = copy(false)) != hipSuccess) break;
        rc = fluid_time_steps(s, 1, 2, &ms);
        tm[v] += ms;
      }
    }
    best.xcd_tiles = tm[1] < tm[0] ? 1 : 0;
  }
  if (e == hipSuccess) e = copy(false);
  if (e == hipSuccess)
    e = hipMemcpyAsync(s->b.accum, accum_keep, 32, hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(save);
  (void)hipFree(save_sid);
  if (rc || e != hipSuccess) {
    s->cfg = keep;
    if (s->b.tiles.w != keep_w) (void)setup_tiles(s, keep_w);
    if (rc) return rc;
    return fail(FLUID_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
  }
  s->cfg = best;
  s->tuned = true;
  if (out) {
    out[0] = (uint32_t)best.lanes;
    out[1] = (uint32_t)best.solve_lanes;
    out[2] = (uint32_t)best.xcd_tiles;
  }
  return FLUID_OK;
}

int fluid_pair_stats(fluid_sim *s, uint64_t out[4]) {
  if (!s || !out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  HIP_TRY(hipSetDevice(s->device));
  HIP_TRY(hipMemsetAsync(s->d_pairs, 0, 32, s->stream));
  HIP_TRY(launch_count_pairs(s->uni(), s->b, s->n, s->d_pairs, s->stream));
  unsigned long long h[4] = {0, 0, 0, 0};
  HIP_TRY(hipMemcpyAsync(h, s->d_pairs, 32, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (int k = 0; k < 4; ++k) out[k] = h[k];
  return FLUID_OK;
}

int fluid_count_pairs(fluid_sim *s, uint64_t *candidates, uint64_t *in_range) {
  if (!candidates || !in_range) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  uint64_t st[4];
  const int rc = fluid_pair_stats(s, st);
  if (rc) return rc;
  *candidates = st[0];
  *in_range = st[1];
  return FLUID_OK;
}

// ------------------------------------------------------------------------
// Differentiable rollout.
int fluid_forward_rollout(fluid_sim *s, uint32_t steps, uint32_t checkpoint_every) {
  if (!s) return fail(FLUID_ERR_INVALID_ARG, "null handle");
  if (s->cursor != -1) return fail(FLUID_ERR_STATE, "a partial step is in progress");
  if (s->mouse.is_active == 1u || s->rb.is_active != 0u)
    return fail(FLUID_ERR_STATE, "mouse force and rigid body are not differentiated: "
                                 "deactivate them before a rollout");
  HIP_TRY(hipSetDevice(s->device));
  uint32_t every = checkpoint_every;
  if (every == 0) every = std::max(1u, (uint32_t)std::ceil(std::sqrt((double)steps)));
  every = std::min(every, std::max(steps, 1u));
  const uint32_t nck = (steps + every - 1) / every;
  int rc = ensure_rollout(s, every, nck);
  if (rc) return rc;
  rc = ensure_graph(s);
  if (rc) return rc;
  RolloutState &R = s->roll;
  R.valid = false;
  const CkptView live = live_state(s);
  for (uint32_t t = 0; t < steps; ++t) {
    if (t % every == 0) HIP_TRY(copy_state(s, ckpt_view(s, t / every), live, s->n));
    rc = replay(s);
    if (rc) return rc;
  }
  HIP_TRY(copy_state(s, fin_view(s), live, s->n));
  rc = check_errors(s);
  if (rc) return rc;
  R.steps = steps;
  R.every = every;
  R.n = s->n;
  R.iters = s->iters;
  R.cfg = s->cfg;
  R.valid = true;
  return FLUID_OK;
}

}  // extern "C"

namespace {

// The reverse of the last forward rollout.  The carried adjoint state is
// kept as mantissa x 2^E per scene (rescaled after every reversed step);
// with log2_scale, the state gradients are returned as mantissas and E,
// otherwise as values (2^E applied in f32).  Parameter gradients are
// values (double).
int backward_rollout(fluid_sim *s, const float *dl_dpos, const float *dl_dvel,
                     float *dl_dpos0, float *dl_dvel0, double *dl_dparams,
                     int32_t *log2_scale) {
  if (!s) return fail(FLUID_ERR_INVALID_ARG, "null handle");
  RolloutState &R = s->roll;
  if (!R.valid)
    return fail(FLUID_ERR_STATE, "no forward rollout to reverse (or the state, parameters or "
                                 "configuration changed since)");
  HIP_TRY(hipSetDevice(s->device));
  const uint32_t n = R.n;
  hipStream_t st = s->stream;
  const AdjViews A = adj_views(s);
  const uint32_t nseg = (R.steps + R.every - 1) / R.every;
  std::vector<hipEvent_t> ev(2 + 3 * (size_t)nseg);
  for (auto &e : ev) HIP_TRY(hipEventCreate(&e));
  auto drop_events = [&]() {
    for (auto &e : ev) (void)hipEventDestroy(e);
  };
  const size_t S = s->sc.size();
  int par = 0;  // parity of the rescale maxima
  auto rescale = [&](int c) -> int {
    HIP_TRY(launch_adj_rescale_step(s->d_scene_of, A.lam_x[c], A.lam_v[c], A.acc, n,
                                    (uint32_t)S, A.maxbits + par * S, A.maxbits + (1 - par) * S,
                                    A.e_total, A.e_step, st));
    par = 1 - par;
    return FLUID_OK;
  };
  HIP_TRY(hipEventRecord(ev[0], st));
  HIP_TRY(hipMemsetAsync(A.maxbits, 0, 2 * S * sizeof(uint32_t), st));
  HIP_TRY(hipMemsetAsync(A.e_total, 0, S * sizeof(int32_t), st));
  if (n) {
    if (dl_dpos) HIP_TRY(hipMemcpyAsync(A.in_x, dl_dpos, (size_t)n * 12, hipMemcpyDefault, st));
    if (dl_dvel) HIP_TRY(hipMemcpyAsync(A.in_v, dl_dvel, (size_t)n * 12, hipMemcpyDefault, st));
    HIP_TRY(launch_adj_load(s->b.sid, dl_dpos ? A.in_x : nullptr, dl_dvel ? A.in_v : nullptr,
                            A.lam_x[0], A.lam_v[0], n, st));
    HIP_TRY(hipMemsetAsync(A.acc, 0, (size_t)n * kAccStride * sizeof(double), st));
    const int rc = rescale(0);
    if (rc) return drop_events(), rc;
  }
  int cur = 0;
  const CkptView live = live_state(s);
  for (int seg = (int)nseg - 1; seg >= 0; --seg) {
    const uint32_t t0 = (uint32_t)seg * R.every;
    const uint32_t len = std::min(R.every, R.steps - t0);
    HIP_TRY(copy_state(s, live, ckpt_view(s, (uint32_t)seg), n));
    HIP_TRY(hipEventRecord(ev[1 + 3 * seg], st));
    for (uint32_t t = 0; t < len; ++t) {
      const int rc = enqueue_tape_step(s, tape_slot(s, t), A.scratch_v);
      if (rc) return drop_events(), rc;
    }
    HIP_TRY(hipEventRecord(ev[2 + 3 * seg], st));
    for (int t = (int)len - 1; t >= 0; --t) {
      int rc = enqueue_adjoint_step(s, tape_slot(s, (uint32_t)t), A, cur);
      if (rc) return drop_events(), rc;
      cur = 1 - cur;
      if (n && (rc = rescale(cur))) return drop_events(), rc;
    }
    HIP_TRY(hipEventRecord(ev[3 + 3 * seg], st));
  }
  std::vector<uint32_t> slot_first(S + 1, 0);
  for (size_t k = 0; k < S; ++k) slot_first[k] = s->sc[k].first;
  slot_first[S] = n;
  if (S == 1) slot_first[0] = 0;
  std::vector<double> grads(S * kGradParams, 0.0);
  std::vector<int32_t> e_total(S, 0);
  if (n) {
    HIP_TRY(hipMemcpyAsync(e_total.data(), A.e_total, S * sizeof(int32_t),
                           hipMemcpyDeviceToHost, st));
    if (!log2_scale)  // values: mantissa x 2^E, rounded to f32 once
      HIP_TRY(launch_adj_apply_scale(s->d_scene_of, A.lam_x[cur], A.lam_v[cur], nullptr, n,
                                     A.e_total, 1, st));
    const uint32_t *sid0 = nseg ? ckpt_view(s, 0).sid : s->b.sid;
    HIP_TRY(launch_adj_store(sid0, A.lam_x[cur], A.lam_v[cur], A.out_x, A.out_v, n, st));
    if (dl_dpos0)
      HIP_TRY(hipMemcpyAsync(dl_dpos0, A.out_x, (size_t)n * 12, hipMemcpyDefault, st));
    if (dl_dvel0)
      HIP_TRY(hipMemcpyAsync(dl_dvel0, A.out_v, (size_t)n * 12, hipMemcpyDefault, st));
    if (dl_dparams) {
      HIP_TRY(hipMemcpyAsync(A.slot_first, slot_first.data(), (S + 1) * 4,
                             hipMemcpyHostToDevice, st));
      HIP_TRY(launch_adj_reduce(A.acc, A.slot_first, (uint32_t)S, A.grads, st));
      HIP_TRY(hipMemcpyAsync(grads.data(), A.grads, grads.size() * sizeof(double),
                             hipMemcpyDeviceToHost, st));
    }
  }
  if (nseg) HIP_TRY(copy_state(s, live, fin_view(s), n));
  HIP_TRY(hipEventRecord(ev[1 + 3 * nseg], st));
  HIP_TRY(hipStreamSynchronize(st));
  float replay = 0.f, reverse = 0.f, total = 0.f, ms = 0.f;
  for (uint32_t k = 0; k < nseg; ++k) {
    HIP_TRY(hipEventElapsedTime(&ms, ev[1 + 3 * k], ev[2 + 3 * k]));
    replay += ms;
    HIP_TRY(hipEventElapsedTime(&ms, ev[2 + 3 * k], ev[3 + 3 * k]));
    reverse += ms;
  }
  HIP_TRY(hipEventElapsedTime(&total, ev[0], ev[1 + 3 * nseg]));
  drop_events();
  R.ms[0] = replay;
  R.ms[1] = reverse;
  R.ms[2] = total;
  // per scene: viscosity, surface tension, gravity x, y, z; with the scaled
  // form also stiffness, which no stage of the step reads (its uniform is
  // copied and never used), so its gradient is exactly zero
  const int stride = log2_scale ? FLUID_GRAD_PARAMS_ALL : FLUID_GRAD_PARAMS;
  if (dl_dparams)
    for (size_t k = 0; k < S; ++k) {
      for (int q = 0; q < kGradParams; ++q)
        dl_dparams[k * stride + q] = std::ldexp(grads[k * kGradParams + q], e_total[k]);
      if (log2_scale) dl_dparams[k * stride + kGradParams] = 0.0;
    }
  if (log2_scale) std::memcpy(log2_scale, e_total.data(), S * sizeof(int32_t));
  return check_errors(s);
}

}  // namespace

extern "C" {

int fluid_backward_rollout(fluid_sim *s, const float *dl_dpos, const float *dl_dvel,
                           float *dl_dpos0, float *dl_dvel0, double *dl_dparams) {
  return backward_rollout(s, dl_dpos, dl_dvel, dl_dpos0, dl_dvel0, dl_dparams, nullptr);
}

int fluid_backward_rollout_scaled(fluid_sim *s, const float *dl_dpos, const float *dl_dvel,
                                  float *dl_dpos0, float *dl_dvel0, double *dl_dparams,
                                  int32_t *log2_scale) {
  if (!log2_scale) return fail(FLUID_ERR_INVALID_ARG, "null log2_scale");
  return backward_rollout(s, dl_dpos, dl_dvel, dl_dpos0, dl_dvel0, dl_dparams, log2_scale);
}

int fluid_rollout_timing(const fluid_sim *s, float ms[3]) {
  if (!s || !ms) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  for (int k = 0; k < 3; ++k) ms[k] = s->roll.ms[k];
  return FLUID_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------
// Whitewater (SpraySystem, src/simulation/spray.rs).  The handle borrows its
// simulation's sorted particles, grid, uniforms and stream; its own device
// memory is the ring, the per-particle emit outputs with their scan, and the
// SprayDev state block.
struct fluid_spray {
  fluid_sim *sim = nullptr;
  GpuSprayParams params{};
  SprayBufs sb{};
  uint32_t last_n = 0;  // particles of the last emit pass
  std::vector<void *> allocs;
};

namespace {

constexpr float kAutoTaInit = 3.0f;  // spray.rs:30
constexpr float kAutoWcInit = 1.0f;  // spray.rs:31

int spray_check(fluid_spray *sp) {
  if (!sp || !sp->sim) return fail(FLUID_ERR_INVALID_ARG, "null spray handle");
  if (sp->sim->cursor != -1)
    return fail(FLUID_ERR_STATE, "a partial step of the simulation is in progress");
  HIP_TRY(hipSetDevice(sp->sim->device));
  return FLUID_OK;
}

int spray_params_ok(const fluid_spray *sp, const GpuSprayParams *p) {
  if (!p) return fail(FLUID_ERR_INVALID_ARG, "null spray params");
  if (p->max_particles != sp->sb.cap)
    return fail(FLUID_ERR_INVALID_ARG, "params.max_particles differs from the ring size");
  return FLUID_OK;
}

// Initial state: head 0, no pending spawns, the ceilings' initial guesses.
hipError_t spray_init_state(fluid_spray *sp) {
  SprayDev d{};
  d.ta_limit = kAutoTaInit;
  d.wc_limit = kAutoWcInit;
  hipError_t e = hipMemcpyAsync(sp->sb.d, &d, sizeof d, hipMemcpyHostToDevice,
                                sp->sim->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(sp->sim->stream);
  return e;
}

uint32_t spray_emitters(const fluid_spray *sp, uint32_t num_sph) {
  return std::min(std::min(num_sph, sp->params.num_sph_particles), sp->sim->n);
}

int spray_emit(fluid_spray *sp, uint32_t num_sph) {
  fluid_sim *s = sp->sim;
  const uint32_t n = spray_emitters(sp, num_sph);
  HIP_TRY(launch_spray_emit(s->uni(), s->b, sp->sb, n, sp->params, s->stream));
  sp->last_n = n;
  return FLUID_OK;
}

int spray_simulate(fluid_spray *sp) {
  fluid_sim *s = sp->sim;
  HIP_TRY(launch_spray_simulate(s->uni(), s->b, sp->sb, sp->params, s->stream));
  return FLUID_OK;
}

}  // namespace

extern "C" {

int fluid_spray_create(fluid_sim *s, uint32_t max_spray, const GpuSprayParams *params,
                       fluid_spray **out) {
  if (!s || !out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (s->batched()) return fail(FLUID_ERR_INVALID_ARG, "spray needs a plain (not batched) simulation");
  if (max_spray == 0) return fail(FLUID_ERR_INVALID_ARG, "max_spray must be positive");
  auto sp = std::make_unique<fluid_spray>();
  sp->sim = s;
  sp->sb.cap = max_spray;
  int rc = spray_params_ok(sp.get(), params);
  if (rc) return rc;
  sp->params = *params;
  HIP_TRY(hipSetDevice(s->device));
  auto alloc = [&](size_t bytes, hipError_t *e) {
    void *p = nullptr;
    *e = hipMalloc(&p, bytes ? bytes : 16);
    if (*e == hipSuccess) {
      sp->allocs.push_back(p);
      *e = hipMemset(p, 0, bytes ? bytes : 16);
    }
    return p;
  };
  const size_t nmax = std::max<uint32_t>(s->max_n, 1);
  hipError_t e = hipSuccess;
  sp->sb.d = (SprayDev *)alloc(sizeof(SprayDev), &e);
  if (e == hipSuccess) sp->sb.ring = (float4 *)alloc((size_t)max_spray * sizeof(GpuSprayParticle), &e);
  if (e == hipSuccess) sp->sb.emit_n = (uint32_t *)alloc(nmax * 4, &e);
  if (e == hipSuccess) sp->sb.emit_incl = (uint32_t *)alloc(nmax * 4, &e);
  if (e == hipSuccess) sp->sb.pot = (float4 *)alloc(nmax * sizeof(float4), &e);
  if (e == hipSuccess)
    sp->sb.tile_status =
        (unsigned long long *)alloc((nmax / kScanTile + 2) * sizeof(unsigned long long), &e);
  if (e == hipSuccess) sp->sb.tile_counter = (uint32_t *)alloc(4, &e);
  if (e == hipSuccess)
    sp->sb.live_part = (uint32_t *)alloc((size_t)spray_simulate_blocks(max_spray) * 16, &e);
  if (e == hipSuccess) e = spray_init_state(sp.get());
  if (e != hipSuccess) {
    for (void *p : sp->allocs) (void)hipFree(p);
    return fail(FLUID_ERR_HIP, std::string("spray allocation: ") + hipGetErrorString(e));
  }
  ++s->bound;
  *out = sp.release();
  return FLUID_OK;
}

int fluid_spray_destroy(fluid_spray *sp) {
  if (!sp) return FLUID_OK;
  if (sp->sim) {
    (void)hipSetDevice(sp->sim->device);
    (void)hipStreamSynchronize(sp->sim->stream);
    --sp->sim->bound;
  }
  for (void *p : sp->allocs) (void)hipFree(p);
  delete sp;
  return FLUID_OK;
}

int fluid_spray_update_params(fluid_spray *sp, const GpuSprayParams *params) {
  if (!sp) return fail(FLUID_ERR_INVALID_ARG, "null spray handle");
  int rc = spray_params_ok(sp, params);
  if (rc) return rc;
  sp->params = *params;  // kernel arguments: every later launch sees them
  return FLUID_OK;
}

int fluid_spray_emit(fluid_spray *sp, uint32_t num_sph_particles) {
  int rc = spray_check(sp);
  if (rc) return rc;
  return spray_emit(sp, num_sph_particles);
}

int fluid_spray_simulate(fluid_spray *sp) {
  int rc = spray_check(sp);
  if (rc) return rc;
  return spray_simulate(sp);
}

int fluid_spray_step(fluid_spray *sp, uint32_t num_sph_particles) {
  int rc = spray_check(sp);
  if (rc) return rc;
  rc = spray_emit(sp, num_sph_particles);
  if (rc) return rc;
  return spray_simulate(sp);
}

int fluid_spray_read_state(fluid_spray *sp, fluid_spray_state *out) {
  int rc = spray_check(sp);
  if (rc) return rc;
  if (!out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  SprayDev d{};
  HIP_TRY(launch_spray_live_sum(sp->sb, sp->sim->stream));
  HIP_TRY(hipMemcpyAsync(&d, sp->sb.d, sizeof d, hipMemcpyDeviceToHost, sp->sim->stream));
  HIP_TRY(hipStreamSynchronize(sp->sim->stream));
  out->write_head = d.head + d.emitted;
  out->ta_limit = d.ta_limit;
  out->wc_limit = d.wc_limit;
  for (int k = 0; k < 4; ++k) out->live[k] = d.live[k];
  out->emitted = d.emitted;
  return check_errors(sp->sim);
}

int fluid_spray_auto_limits(fluid_spray *sp, float out[2]) {
  if (!out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  fluid_spray_state st{};
  int rc = fluid_spray_read_state(sp, &st);
  if (rc) return rc;
  out[0] = st.ta_limit;
  out[1] = st.wc_limit;
  return FLUID_OK;
}

int fluid_spray_read_stats(fluid_spray *sp, uint32_t out[4]) {
  if (!out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  fluid_spray_state st{};
  int rc = fluid_spray_read_state(sp, &st);
  if (rc) return rc;
  for (int k = 0; k < 4; ++k) out[k] = st.live[k];
  return FLUID_OK;
}

int fluid_spray_read_particles(fluid_spray *sp, GpuSprayParticle *out, uint32_t count) {
  int rc = spray_check(sp);
  if (rc) return rc;
  if (!out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (count < sp->sb.cap) return fail(FLUID_ERR_CAPACITY, "output smaller than the ring");
  HIP_TRY(hipMemcpyAsync(out, sp->sb.ring, (size_t)sp->sb.cap * sizeof(GpuSprayParticle),
                         hipMemcpyDeviceToHost, sp->sim->stream));
  HIP_TRY(hipStreamSynchronize(sp->sim->stream));
  return check_errors(sp->sim);
}

int fluid_spray_write_particles(fluid_spray *sp, const GpuSprayParticle *in, uint32_t count) {
  int rc = spray_check(sp);
  if (rc) return rc;
  if (!in) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (count != sp->sb.cap) return fail(FLUID_ERR_INVALID_ARG, "count must equal the ring size");
  HIP_TRY(hipMemcpyAsync(sp->sb.ring, in, (size_t)sp->sb.cap * sizeof(GpuSprayParticle),
                         hipMemcpyHostToDevice, sp->sim->stream));
  HIP_TRY(hipStreamSynchronize(sp->sim->stream));
  return FLUID_OK;
}

int fluid_spray_read_emit(fluid_spray *sp, float *potentials, uint32_t *spawns,
                          uint32_t count) {
  int rc = spray_check(sp);
  if (rc) return rc;
  const uint32_t n = sp->last_n;
  if ((!potentials && !spawns) || count < n)
    return fail(count < n ? FLUID_ERR_CAPACITY : FLUID_ERR_INVALID_ARG,
                "outputs must hold the last emit pass's particles");
  hipStream_t st = sp->sim->stream;
  if (potentials && n)
    HIP_TRY(hipMemcpyAsync(potentials, sp->sb.pot, (size_t)n * sizeof(float4),
                           hipMemcpyDeviceToHost, st));
  if (spawns && n)
    HIP_TRY(hipMemcpyAsync(spawns, sp->sb.emit_n, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return check_errors(sp->sim);
}

int fluid_spray_reset(fluid_spray *sp) {
  int rc = spray_check(sp);
  if (rc) return rc;
  HIP_TRY(hipMemsetAsync(sp->sb.ring, 0, (size_t)sp->sb.cap * sizeof(GpuSprayParticle),
                         sp->sim->stream));
  HIP_TRY(hipMemsetAsync(sp->sb.live_part, 0, (size_t)spray_simulate_blocks(sp->sb.cap) * 16,
                         sp->sim->stream));
  HIP_TRY(spray_init_state(sp));
  return FLUID_OK;
}

int fluid_spray_time_steps(fluid_spray *sp, uint32_t num_sph_particles, uint32_t warmup,
                           uint32_t frames, float *ms) {
  int rc = spray_check(sp);
  if (rc) return rc;
  if (!ms) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  for (uint32_t k = 0; k < warmup; ++k) {
    if ((rc = spray_emit(sp, num_sph_particles)) || (rc = spray_simulate(sp))) return rc;
  }
  hipStream_t st = sp->sim->stream;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipStreamSynchronize(st));
  HIP_TRY(hipEventRecord(e0, st));
  for (uint32_t k = 0; k < frames; ++k) {
    if ((rc = spray_emit(sp, num_sph_particles)) || (rc = spray_simulate(sp))) break;
  }
  HIP_TRY(hipEventRecord(e1, st));
  HIP_TRY(hipEventSynchronize(e1));
  HIP_TRY(hipEventElapsedTime(ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  return check_errors(sp->sim);
}

}  // extern "C"

// ------------------------------------------------------------------------
// Marching-cubes density field (MarchingCubesRenderer's compute passes 0 and
// 1, src/render/marching_cubes.rs:2349-2470), bound to a simulation like a
// spray handle.
struct fluid_mc {
  fluid_sim *sim = nullptr;
  uint32_t G = 0;
  GpuMcGridParams g{};
  GpuAnisoParams a{};
  GpuContainerGeometry c{};
  float *field = nullptr;
  float4 *aniso = nullptr;
  std::vector<void *> allocs;
};

namespace {

int mc_check(fluid_mc *mc) {
  if (!mc || !mc->sim) return fail(FLUID_ERR_INVALID_ARG, "null MC field handle");
  if (mc->sim->cursor != -1)
    return fail(FLUID_ERR_STATE, "a partial step of the simulation is in progress");
  HIP_TRY(hipSetDevice(mc->sim->device));
  return FLUID_OK;
}

int mc_launch(fluid_mc *mc, uint32_t n, bool aniso_pass) {
  fluid_sim *s = mc->sim;
  const McLaunch m = make_mc_launch(mc->g, mc->a, mc->c, s->sc[0].grid.inv_cell_size);
  if (aniso_pass && mc->a.enabled && n)
    HIP_TRY(launch_mc_aniso(s->uni(), s->b, m, mc->aniso, n, s->stream));
  HIP_TRY(launch_mc_density(s->uni(), s->b, m, mc->aniso, mc->field, s->stream));
  return FLUID_OK;
}

}  // namespace

extern "C" {

int fluid_mc_create(fluid_sim *s, uint32_t grid_size, fluid_mc **out) {
  if (!s || !out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  if (s->batched()) return fail(FLUID_ERR_INVALID_ARG, "the MC field needs a plain simulation");
  if (grid_size == 0 || grid_size > 512)
    return fail(FLUID_ERR_INVALID_ARG, "grid_size must be 1 to 512");
  auto mc = std::make_unique<fluid_mc>();
  mc->sim = s;
  mc->G = grid_size;
  // the app's defaults (app.rs:1472-1519, marching_cubes.rs:2133-2146):
  // isotropic h_mc = h, anisotropy on, bounds = the container +- margin
  const float h = s->sc[0].params.kernel_radius;
  const GpuContainerGeometry &cg = s->sc[0].cont;
  mc->a = default_aniso_params(h);
  mc->c = cg;
  mc->c.clip_enabled = 0u;
  mc->g = default_mc_grid(cg, h, grid_size);
  HIP_TRY(hipSetDevice(s->device));
  hipError_t e = hipSuccess;
  auto alloc = [&](size_t bytes) {
    void *p = nullptr;
    if (e == hipSuccess) e = hipMalloc(&p, bytes ? bytes : 16);
    if (e == hipSuccess) {
      mc->allocs.push_back(p);
      e = hipMemset(p, 0, bytes ? bytes : 16);
    }
    return p;
  };
  const size_t g3 = (size_t)grid_size * grid_size * grid_size;
  mc->field = (float *)alloc(g3 * sizeof(float));
  mc->aniso = (float4 *)alloc((size_t)std::max<uint32_t>(s->max_n, 1) * sizeof(GpuParticleAniso));
  if (e != hipSuccess) {
    for (void *p : mc->allocs) (void)hipFree(p);
    return fail(FLUID_ERR_HIP, std::string("MC field allocation: ") + hipGetErrorString(e));
  }
  ++s->bound;
  *out = mc.release();
  return FLUID_OK;
}

int fluid_mc_destroy(fluid_mc *mc) {
  if (!mc) return FLUID_OK;
  if (mc->sim) {
    (void)hipSetDevice(mc->sim->device);
    (void)hipStreamSynchronize(mc->sim->stream);
    --mc->sim->bound;
  }
  for (void *p : mc->allocs) (void)hipFree(p);
  delete mc;
  return FLUID_OK;
}

int fluid_mc_update_params(fluid_mc *mc, const GpuMcGridParams *p) {
  if (!mc || !p) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (p->grid_size != mc->G)
    return fail(FLUID_ERR_INVALID_ARG, "params.grid_size differs from the field's");
  if (!(p->kernel_radius > 0.0f) || !(p->cell_size > 0.0f))
    return fail(FLUID_ERR_INVALID_ARG, "kernel_radius and cell_size must be positive");
  mc->g = *p;
  return FLUID_OK;
}

int fluid_mc_update_aniso_params(fluid_mc *mc, const GpuAnisoParams *p) {
  if (!mc || !p) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (p->enabled && !(p->support_radius > 0.0f && p->h_mc > 0.0f && p->kr > 0.0f))
    return fail(FLUID_ERR_INVALID_ARG, "support_radius, h_mc and kr must be positive");
  mc->a = *p;
  return FLUID_OK;
}

int fluid_mc_update_container_geometry(fluid_mc *mc, const GpuContainerGeometry *geom) {
  if (!mc || !geom) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  mc->c = *geom;
  return FLUID_OK;
}

int fluid_mc_generate(fluid_mc *mc, uint32_t num_particles) {
  int rc = mc_check(mc);
  if (rc) return rc;
  return mc_launch(mc, std::min(num_particles, mc->sim->n), true);
}

int fluid_mc_density_from_aniso(fluid_mc *mc, const GpuParticleAniso *in, uint32_t count) {
  int rc = mc_check(mc);
  if (rc) return rc;
  if (!in || count != mc->sim->n)
    return fail(FLUID_ERR_INVALID_ARG, "one record per particle of the simulation");
  HIP_TRY(hipMemcpyAsync(mc->aniso, in, (size_t)count * sizeof(GpuParticleAniso),
                         hipMemcpyHostToDevice, mc->sim->stream));
  rc = mc_launch(mc, count, false);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(mc->sim->stream));
  return FLUID_OK;
}

int fluid_mc_read_density(fluid_mc *mc, float *out, uint64_t count) {
  int rc = mc_check(mc);
  if (rc) return rc;
  const uint64_t g3 = (uint64_t)mc->G * mc->G * mc->G;
  if (!out) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (count < g3) return fail(FLUID_ERR_CAPACITY, "output smaller than grid_size^3");
  HIP_TRY(hipMemcpyAsync(out, mc->field, g3 * sizeof(float), hipMemcpyDeviceToHost,
                         mc->sim->stream));
  HIP_TRY(hipStreamSynchronize(mc->sim->stream));
  return FLUID_OK;
}

int fluid_mc_read_aniso(fluid_mc *mc, GpuParticleAniso *out, uint32_t count) {
  int rc = mc_check(mc);
  if (rc) return rc;
  const uint32_t n = mc->sim->n;
  if (!out && n) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (count < n) return fail(FLUID_ERR_CAPACITY, "output smaller than num_particles");
  if (n)
    HIP_TRY(hipMemcpyAsync(out, mc->aniso, (size_t)n * sizeof(GpuParticleAniso),
                           hipMemcpyDeviceToHost, mc->sim->stream));
  HIP_TRY(hipStreamSynchronize(mc->sim->stream));
  return FLUID_OK;
}

int fluid_mc_time(fluid_mc *mc, uint32_t num_particles, uint32_t warmup, uint32_t iters,
                  float *ms) {
  int rc = mc_check(mc);
  if (rc) return rc;
  if (!ms) return fail(FLUID_ERR_INVALID_ARG, "null argument");
  const uint32_t n = std::min(num_particles, mc->sim->n);
  for (uint32_t k = 0; k < warmup; ++k)
    if ((rc = mc_launch(mc, n, true))) return rc;
  hipStream_t st = mc->sim->stream;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  HIP_TRY(hipStreamSynchronize(st));
  HIP_TRY(hipEventRecord(e0, st));
  for (uint32_t k = 0; k < iters && rc == FLUID_OK; ++k) rc = mc_launch(mc, n, true);
  HIP_TRY(hipEventRecord(e1, st));
  HIP_TRY(hipEventSynchronize(e1));
  HIP_TRY(hipEventElapsedTime(ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

}  // extern "C"

// ------------------------------------------------------------------------
// The marching-cubes field without a field handle and without allocating:
// the ellipsoid records between the two passes live in `scratch`, a fixed
// array of this library (one copy per device, held until the process
// ends).  Host code never passes the array's name to a kernel: the name is
// a host-side stand-in.  Each device's address comes from
// hipGetSymbolAddress, taken once per device with that device current.
static __device__ __align__(256) float4 scratch[3 * (size_t)FLUID_MC_NOALLOC_MAX_PARTICLES];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
float4 *g_scratch_ptr[kMaxScratchDevices] = {};

// The calling thread's current device must be `device`.
int scratch_on(int device, float4 **out) {
  if (device < 0 || device >= kMaxScratchDevices)
    return fail(FLUID_ERR_INVALID_ARG, "device ordinal out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FLUID_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(e));
    g_scratch_ptr[device] = (float4 *)p;
  }
  *out = g_scratch_ptr[device];
  return FLUID_OK;
}

}  // namespace

extern "C" int fluid_mc_generate_noalloc(fluid_sim *s, const GpuMcGridParams *grid,
                                         const GpuAnisoParams *aniso,
                                         const GpuContainerGeometry *container,
                                         float *field) {
  if (!s || !grid || !aniso || !container || !field)
    return fail(FLUID_ERR_INVALID_ARG, "null argument");
  if (s->batched()) return fail(FLUID_ERR_INVALID_ARG, "the MC field needs a plain simulation");
  if (s->cursor != -1)
    return fail(FLUID_ERR_STATE, "a partial step of the simulation is in progress");
  if (grid->grid_size == 0 || grid->grid_size > 512)
    return fail(FLUID_ERR_INVALID_ARG, "grid_size must be 1 to 512");
  if (!(grid->kernel_radius > 0.0f) || !(grid->cell_size > 0.0f))
    return fail(FLUID_ERR_INVALID_ARG, "kernel_radius and cell_size must be positive");
  if (aniso->enabled && !(aniso->support_radius > 0.0f && aniso->h_mc > 0.0f && aniso->kr > 0.0f))
    return fail(FLUID_ERR_INVALID_ARG, "support_radius, h_mc and kr must be positive");
  const uint32_t n = s->n;
  if (aniso->enabled && n > FLUID_MC_NOALLOC_MAX_PARTICLES)
    return fail(FLUID_ERR_CAPACITY, "anisotropy without allocation takes at most " +
                                        std::to_string(FLUID_MC_NOALLOC_MAX_PARTICLES) +
                                        " particles (the simulation holds " +
                                        std::to_string(n) + ")");
  HIP_TRY(hipSetDevice(s->device));
  float4 *recs = nullptr;
  const int rc = scratch_on(s->device, &recs);
  if (rc) return rc;
  const McLaunch m = make_mc_launch(*grid, *aniso, *container, s->sc[0].grid.inv_cell_size);
  // the fit writes every record the splat reads (particles 0 .. n-1, the
  // slots of the simulation's grid), so nothing in the array needs clearing
  if (aniso->enabled && n) HIP_TRY(launch_mc_aniso(s->uni(), s->b, m, recs, n, s->stream));
  HIP_TRY(launch_mc_density(s->uni(), s->b, m, recs, field, s->stream));
  return FLUID_OK;
}
