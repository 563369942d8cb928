// sph_kernels.hip — the per-step SPH kernels for gfx950 (wave64).
//
// One kernel (or kernel pair) per WGSL entry point dispatched by
// SphSimulation3DGrid::step (src/simulation/sph_3d_grid.rs:1261-1444).  The
// arithmetic of every stage follows the WGSL expression structure; three
// kinds of value are made to agree bit-for-bit with the f32 CPU oracle
// because a branch or an integer depends on them:
//   - cell coordinates: floor((p - origin) * inv_cell), no fusion possible;
//   - squared distances: dot(a, b) evaluated as fma(z, z, fma(y, y, x*x));
//   - the force pass's close-range test r < 0.1 h, on a correctly rounded
//     sqrt of that squared distance.
// Elsewhere the compiler may contract multiply-adds and the neighbour sums
// may be split across lanes, so floating-point fields agree to a tolerance.
//
// Integer/ordering work (build, scan, reorder) is exact and deterministic:
// the within-cell order is ascending original particle index, which the
// reference leaves to a racing atomicAdd (grid_reorder.wgsl:51).
#include "sph_device.h"

namespace fluid {
namespace {

constexpr uint64_t kFlagAgg = 1ull << 32;
constexpr uint64_t kFlagPrefix = 2ull << 32;
constexpr uint32_t kSpinCap = 1u << 24;

// ------------------------------------------------------------------------
// Stage 0 — grid_clear.wgsl:23-29 and clear_buffer(cell_offsets)
// (sph_3d_grid.rs:1281).  The offsets buffer does not exist here (build
// hands out the in-cell slots); the scan's tile states are cleared instead.
// The build counts into `work`, which the scan zeroes again as it publishes
// the counts, and the integrate pass resets the scan's tile states: a full
// step that follows a full step finds all of it clear and launches no
// clear (the step's range decides, fluid_sim.hip).  Grids of 2^20 cells or
// more (batches) count in place (work == counts) and clear every step:
// there the clear's streaming writes cost less than the scan's publishing.
__global__ __launch_bounds__(256) void k_clear(uint32_t *__restrict__ counts,
                                               uint32_t *__restrict__ work,
                                               uint32_t cells,
                                               unsigned long long *status,
                                               uint32_t ntiles,
                                               uint32_t *tile_counter,
                                               uint32_t *cell_tile_count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t q = cells / 4;
  if (i < q) {
    reinterpret_cast<uint4 *>(counts)[i] = make_uint4(0, 0, 0, 0);
    if (work != counts) reinterpret_cast<uint4 *>(work)[i] = make_uint4(0, 0, 0, 0);
  }
  if (i < cells - 4 * q) {
    counts[4 * q + i] = 0u;
    if (work != counts) work[4 * q + i] = 0u;
  }
  if (i < ntiles) status[i] = 0ull;
  if (i == 0) {
    *tile_counter = 0u;
    if (cell_tile_count) *cell_tile_count = 0u;
  }
}

// ------------------------------------------------------------------------
// Stage 1 — grid_build.wgsl:46-62.  cell = clamp(floor((p-o)/h)) linearised
// x + y G + z G^2; the per-cell count is an atomic add.  The particles of a
// workgroup are first grouped by cell in an LDS hash table (open addressing,
// twice as many slots as threads, so every probe sequence ends), which
// gives each particle its rank within its cell's group; then one global
// atomic per distinct cell of the workgroup adds the group's size, and the
// returned base + rank is the particle's provisional in-cell slot, used
// only to scatter before ranking.  Where the flow piles hundreds of
// particles into one cell, the atomics on that cell's counter serialise in
// the L2; grouping cuts them to one per workgroup that holds the cell.
constexpr uint32_t kBuildHash = 512;
constexpr uint32_t kBuildEmpty = 0xFFFFFFFFu;

template <bool B>
__global__ __launch_bounds__(256) void k_build(Uni un, Bufs b, uint32_t n) {
  __shared__ uint32_t s_key[kBuildHash], s_cnt[kBuildHash], s_base[kBuildHash];
  static_assert(kBuildHash >= 2 * 256, "the table must hold every thread's key");
  for (uint32_t i = threadIdx.x; i < kBuildHash; i += blockDim.x) {
    s_key[i] = kBuildEmpty;
    s_cnt[i] = 0u;
  }
  __syncthreads();
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = t < n;
  uint32_t ci = 0, h = 0, rank = 0;
  if (valid) {
    const DevUniforms *__restrict__ U = uni_at<B>(un, t);
    const float4 p = b.pos[t];
    const int cx = min(max(cell_coord(p.x, U->ox, U->inv_cell), 0), (int)U->gx - 1);
    const int cy = min(max(cell_coord(p.y, U->oy, U->inv_cell), 0), (int)U->gy - 1);
    const int cz = min(max(cell_coord(p.z, U->oz, U->inv_cell), 0), (int)U->gz - 1);
    ci = U->cell_offset + (uint32_t)cx + (uint32_t)cy * U->gx + (uint32_t)cz * U->gxy;
    b.cell_of[t] = ci;
    // at most 256 distinct keys in 512 slots: a free or matching slot is
    // always found within the table's length
    h = (ci * 2654435761u) >> 23;
    for (uint32_t probe = 0; probe < kBuildHash; ++probe) {
      const uint32_t k = atomicCAS(&s_key[h], kBuildEmpty, ci);
      if (k == kBuildEmpty || k == ci) break;
      h = (h + 1u) & (kBuildHash - 1u);
    }
    rank = atomicAdd(&s_cnt[h], 1u);
  }
  __syncthreads();
  if (valid && rank == 0) s_base[h] = atomicAdd(&b.cwork[ci], s_cnt[h]);
  __syncthreads();
  if (valid) b.slot_of[t] = s_base[h] + rank;
}

// ------------------------------------------------------------------------
// Stage 2 — copy(counts -> starts) + prefix_scan.wgsl's three dispatches,
// as one single-pass inclusive scan with decoupled look-back.  Tiles take
// ids in launch order from an atomic counter, so a tile only waits on tiles
// that are already running; every spin is capped (error bit on overflow).
// ITEMS cells per thread: the look-back costs a roughly fixed time per tile,
// so large grids use 4x larger tiles (16 -> 64 items: 5M cells 0.16 ->
// 0.05 ms), small ones the smaller tiles' shorter per-thread scan.
template <uint32_t ITEMS>
__global__ __launch_bounds__(256) void k_scan(Bufs b, uint32_t cells) {
  constexpr uint32_t TILE = kScanThreads * ITEMS;
  __shared__ uint32_t s_tile, s_prefix;
  __shared__ uint32_t s_wave[4];
  if (threadIdx.x == 0) s_tile = atomicAdd(b.tile_counter, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t base = tile * TILE + threadIdx.x * ITEMS;
  // counts from b.cwork; with a separate b.counts they are published there
  // and b.cwork is zeroed for the next build (after the look-back, with the
  // starts: stores before it would delay the tile's aggregate)
  const bool publish = b.counts && b.counts != b.cwork;
  uint32_t v[ITEMS];
  if (base + ITEMS <= cells) {
#pragma unroll
    for (uint32_t q = 0; q < ITEMS / 4; ++q) {
      const uint4 x = *reinterpret_cast<const uint4 *>(b.cwork + base + 4 * q);
      v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
    }
  } else {
#pragma unroll
    for (uint32_t q = 0; q < ITEMS; ++q) v[q] = base + q < cells ? b.cwork[base + q] : 0u;
  }
#pragma unroll
  for (uint32_t q = 1; q < ITEMS; ++q) v[q] += v[q - 1];
  const uint32_t tsum = v[ITEMS - 1];
  const uint32_t incl = wave_incl_scan(tsum);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 63) s_wave[wid] = incl;
  __syncthreads();
  uint32_t wave_off = 0;
  for (int w = 0; w < wid; ++w) wave_off += s_wave[w];
  const uint32_t agg = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
  if (wid == 0) {
    uint32_t excl = 0;
    if (tile == 0) {
      if (lane == 0)
        __hip_atomic_store(&b.tile_status[0], kFlagPrefix | agg,
                           __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (lane == 0)
        __hip_atomic_store(&b.tile_status[tile], kFlagAgg | agg,
                           __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      long long look = (long long)tile - 1;
      bool failed = false;
      while (true) {
        const long long idx = look - lane;
        unsigned long long st;
        uint32_t spins = 0;
        do {
          st = idx >= 0 ? __hip_atomic_load(&b.tile_status[idx], __ATOMIC_ACQUIRE,
                                            __HIP_MEMORY_SCOPE_AGENT)
                        : kFlagPrefix;
          if (++spins > kSpinCap) { failed = true; break; }
        } while (__ballot((st >> 32) == 0ull) != 0ull);
        if (__ballot(failed) != 0ull) {
          if (lane == 0) atomicOr(b.err, kErrScanSpin);
          break;
        }
        const unsigned long long pm = __ballot((st >> 32) == 2ull);
        if (pm) {
          const int first = __ffsll((long long)pm) - 1;
          excl += wave_sum(lane <= first ? (uint32_t)st : 0u);
          break;
        }
        excl += wave_sum((uint32_t)st);
        look -= 64;
      }
      if (lane == 0)
        __hip_atomic_store(&b.tile_status[tile], kFlagPrefix | (uint32_t)(excl + agg),
                           __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0) s_prefix = excl;
  }
  __syncthreads();
  const uint32_t off = s_prefix + wave_off + (incl - tsum);
  if (base + ITEMS <= cells) {
#pragma unroll
    for (uint32_t q = 0; q < ITEMS / 4; ++q)
      *reinterpret_cast<uint4 *>(b.starts + base + 4 * q) =
          make_uint4(off + v[4 * q], off + v[4 * q + 1], off + v[4 * q + 2], off + v[4 * q + 3]);
    if (publish) {
#pragma unroll
      for (uint32_t q = 0; q < ITEMS / 4; ++q) {
        const uint32_t p = q ? v[4 * q - 1] : 0u;  // the counts from the running sums
        *reinterpret_cast<uint4 *>(b.counts + base + 4 * q) =
            make_uint4(v[4 * q] - p, v[4 * q + 1] - v[4 * q], v[4 * q + 2] - v[4 * q + 1],
                       v[4 * q + 3] - v[4 * q + 2]);
        *reinterpret_cast<uint4 *>(b.cwork + base + 4 * q) = make_uint4(0, 0, 0, 0);
      }
    }
  } else {
#pragma unroll
    for (uint32_t q = 0; q < ITEMS; ++q)
      if (base + q < cells) {
        b.starts[base + q] = off + v[q];
        if (publish) {
          b.counts[base + q] = v[q] - (q ? v[q - 1] : 0u);
          b.cwork[base + q] = 0u;
        }
      }
  }
}

// ------------------------------------------------------------------------
// After the scan: the work items of the LDS-staged sweeps.  Every
// non-empty cell tile of every scene contributes one item per `chunk` of
// its own particles (dense tiles spread over several workgroups, each
// staging the same rows), appended with one atomic per wave.  The list order
// is arbitrary; items are processed independently of each other.
template <bool B>
__global__ __launch_bounds__(256) void k_tiles(Uni un, Bufs b) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const TileMap &tm = b.tiles;
  uint32_t items = 0, scene = 0, lt = 0;
  if (g < tm.total) {
    if (B) {
      uint32_t lo = 0, hi = tm.nscenes;  // last scene with scene_first <= g
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (tm.scene_first[mid] <= g) lo = mid; else hi = mid;
      }
      scene = lo;
    }
    lt = g - tm.scene_first[scene];
    const DevUniforms *U = un.u + scene;
    const uint32_t gx = U->gx, gy = U->gy;
    const uint32_t nxc = (gx + tm.w - 1) / tm.w;
    const uint32_t xc = lt % nxc, row = lt / nxc;
    const uint32_t y = row % gy, z = row / gy;
    const uint32_t x0 = xc * tm.w, x1 = min(x0 + tm.w, gx) - 1;
    const uint32_t rc = U->cell_offset + y * gx + z * U->gxy;
    const uint32_t own = b.starts[rc + x1] - (b.starts[rc + x0] - b.counts[rc + x0]);
    items = (own + tm.chunk - 1) / tm.chunk;
  }
  const uint32_t incl = wave_incl_scan(items);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t total = (uint32_t)__shfl((int)incl, 63);
  if (total == 0) return;
  uint32_t base = 0;
  if (lane == 63) base = atomicAdd(tm.count, total);
  base = (uint32_t)__shfl((int)base, 63) + incl - items;
  for (uint32_t k = 0; k < items; ++k) tm.list[base + k] = make_uint4(scene, lt, k, 0u);
}

// Stage 3 — grid_reorder.wgsl:37-63 as a deterministic counting sort.  The
// reference's within-cell order comes from a racing atomicAdd
// (grid_reorder.wgsl:51); here it is fixed to ascending position x (the
// x_key of the position the build binned, ties by original index): each
// x-row of cells is one slot range sorted by x, so neighbouring slots are
// close in space and the list walks' record gathers share cache lines.
// (a) scatter (old slot, original index, x key, cell) to cell start +
// provisional slot;
__global__ __launch_bounds__(256) void k_scatter(Bufs b, uint32_t n) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t ci = b.cell_of[t];
  b.tmp[cell_begin(b.starts, ci) + b.slot_of[t]] = make_uint4(t, b.sid[t], x_key(b.pos[t].x), ci);
}

// (b) rank every entry by (x key, original index) within its cell segment
// and copy the particle (position, velocity, force) to its final sorted
// slot.  A workgroup's 256 consecutive entries cover consecutive cell
// segments; their 64-bit order keys are staged in LDS, kRankLds at a time,
// and the O(count) rank loop reads broadcast LDS words.  A range longer than
// one stage (a cluster cell of thousands of particles) is staged in turns,
// each thread counting over the part of its cell in the stage.
constexpr uint32_t kRankLds = 4096;

__global__ __launch_bounds__(256) void k_rank_gather(Bufs b, uint32_t n) {
  __shared__ unsigned long long s_ids[kRankLds];
  __shared__ uint32_t s_lo, s_hi;
  const uint32_t u0 = blockIdx.x * blockDim.x;
  const uint32_t u = u0 + threadIdx.x;
  const uint32_t last = min(n, u0 + blockDim.x) - 1;
  uint4 e = make_uint4(0, 0, 0, 0);
  uint32_t start = 0, end = 0;
  if (u < n) {
    e = b.tmp[u];
    const uint32_t ci = e.w;  // the entry's cell, carried by the scatter
    end = b.starts[ci];
    start = end - b.counts[ci];
    if (u == u0) s_lo = start;
    if (u == last) s_hi = end;
  }
  __syncthreads();
  const uint32_t lo = s_lo, hi = s_hi;
  // rank = entries of the cell with a smaller order key (x key, then
  // original index); eight independent reads per trip so they overlap
  uint32_t r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const unsigned long long key = (unsigned long long)e.z << 32 | e.y;
  for (uint32_t c0 = lo; c0 < hi; c0 += kRankLds) {
    const uint32_t len = min(hi - c0, kRankLds);
    if (c0 != lo) __syncthreads();  // the previous stage has been read
    for (uint32_t v = threadIdx.x; v < len; v += blockDim.x) {
      const uint4 t = b.tmp[c0 + v];
      s_ids[v] = (unsigned long long)t.z << 32 | t.y;
    }
    __syncthreads();
    // this thread's cell within the stage: [vb, ve), empty when the cell
    // lies wholly before or after it
    const uint32_t vb = max(start, c0), vend = min(end, c0 + len);
    if (u < n && vb < vend) {
      uint32_t v = vb - c0;
      const uint32_t ve = vend - c0;
      for (; v + 8 <= ve; v += 8)
#pragma unroll
        for (int q = 0; q < 8; ++q) r[q] += (s_ids[v + q] < key) ? 1u : 0u;
      for (; v < ve; ++v) r[0] += (s_ids[v] < key) ? 1u : 0u;
    }
  }
  if (u >= n) return;
  const uint32_t rank = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  const uint32_t dest = start + rank;
  b.sid[dest] = e.y;
  if (b.gmap) b.gmap[dest] = e.x;
  b.spos[dest] = b.pos[e.x];
  b.svel[dest] = b.vel[e.x];
  b.sfrc[dest] = b.frc[e.x];
}

// ------------------------------------------------------------------------
// Stage 4 — sph_density_3d_grid.wgsl:111-194 (+ boundary_gamma :86-109).
// Besides density, near density and the Akinci normal, the particle's L
// lanes record the in-range neighbours they meet (0 < r^2 < h^2, the pairs
// XSPH and force sum over) in one ballot-balanced list for 8 walker lanes,
// and the particle's neighbour record is written for the later gathers.
// Per-particle body of the density stage; sweep(f) walks the particle's
// candidates group-synchronously (sweep27 or tile_sweep with SYNC).
template <int L, class Sweep>
__device__ __forceinline__ void density_particle(const DevUniforms *__restrict__ U,
                                                 const Bufs &b, uint32_t n, uint32_t s,
                                                 uint32_t sub, float4 pi, uint32_t total,
                                                 uint32_t list_skip, Sweep &&sweep) {
  const float h = U->h, h2 = U->h2;
  const float mps = U->mass * U->poly6_scale, mns = U->mass * U->near_scale;
  const float mr = U->m_over_rho0;
  // the list is written for 8 walker lanes, whatever L
  uint32_t *__restrict__ nbr = b.nbr + lane_block<kListLanes>(s, kListCapacity);
  uint32_t run = 0;  // pairs recorded so far by the whole group
  // A particle with more than list_skip candidates (the group's stencil
  // total) records no list: it is marked overflowed, and XSPH and force
  // sweep its candidates, as they would for a list that overflowed.  Such
  // particles sit in dense clusters, where most candidates are pairs:
  // their lists would mostly overflow, and writing the first kListCapacity
  // entries of a list that is never read can cost more than the sweep
  // saves for the few that would have fit.
  const bool record = total <= list_skip;
  float dens = 0.f, near = 0.f, nx = 0.f, ny = 0.f, nz = 0.f;
  const PairTest in_pair(h2);
  // Branch-free body: a pair outside the support contributes exact zeros
  // through selects, so the lanes of a wave stay converged.
  sweep([&](uint32_t j, float4 pj, bool valid) {
    const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
    const float r2 = dot3(rx, ry, rz, rx, ry, rz);
    const bool inr = valid && r2 < h2;
    const bool pair = valid && in_pair(r2);
    const float d = inr ? h2 - r2 : 0.f;
    const float d2 = d * d;
    dens = __builtin_fmaf(mps, d2 * d, dens);
    const float dn = inr ? h - fast_sqrt(r2) : 0.f;
    near = __builtin_fmaf(mns, dn * dn * dn, near);
    const float fn = pair ? mr * d2 : 0.f;
    nx += fn * rx;
    ny += fn * ry;
    nz += fn * rz;
    if (record) group_list_append<L, kListLanes>(nbr, sub, run, pair, j);
  });
  if (sub == 0) b.ncnt[s] = record ? run : kListOverflow;
  dens = group_sum<L>(dens);
  near = group_sum<L>(near);
  nx = group_sum<L>(nx);
  ny = group_sum<L>(ny);
  nz = group_sum<L>(nz);
  if (sub != 0) return;
  // boundary_gamma in the container frame (container_common.wgsl:27-35)
  const float px = pi.x, py = pi.y - U->cy, pz = pi.z;
  const float lx = dot3(U->iv[0][0], U->iv[0][1], U->iv[0][2], px, py, pz);
  const float ly = dot3(U->iv[1][0], U->iv[1][1], U->iv[1][2], px, py, pz);
  const float lz = dot3(U->iv[2][0], U->iv[2][1], U->iv[2][2], px, py, pz);
  const float dist[6] = {lx - (-U->hw), U->hw - lx, ly - (-U->hh),
                         U->hh - ly,    lz - (-U->hd), U->hd - lz};
  float gamma = 1.0f;
#pragma unroll
  for (int k = 0; k < 6; ++k)
    if (dist[k] < h) gamma *= 0.5f + 0.5f * fminf(fmaxf(dist[k] / h, 0.f), 1.f);
  gamma = fmaxf(gamma, 0.1f);
  const float sc = U->normal_scale;
  const float4 vi = b.svel[s];
  float4 *rec = b.srec + 3 * (size_t)s;
  rec[0] = make_float4(pi.x, pi.y, pi.z, dens / gamma);
  rec[1] = make_float4(vi.x, vi.y, vi.z, near);
  // .w: 1 / max(near density, 1), the force's per-neighbour reciprocal
  rec[2] = make_float4(nx * sc, ny * sc, nz * sc, __builtin_amdgcn_rcpf(max1(near)));
}

template <int L, bool B>
__global__ __launch_bounds__(256) void k_density(Uni un, Bufs b, uint32_t n, uint32_t xcd,
                                                 uint32_t list_skip) {
  __shared__ RowEnt rows[256 / L][9];
  const uint32_t g = xcd_block(xcd) * blockDim.x + threadIdx.x;
  const uint32_t s = g / L, sub = g % L;
  if (s >= n) return;
  const DevUniforms *__restrict__ U = uni_at<B>(un, s);
  const float4 pi = b.spos[s];
  const int cx = cell_coord(pi.x, U->ox, U->inv_cell);
  const int cy = cell_coord(pi.y, U->oy, U->inv_cell);
  const int cz = cell_coord(pi.z, U->oz, U->inv_cell);
  RowEnt *tab = rows[threadIdx.x / L];
  const uint32_t total = rows27<L>(U, b.counts, b.starts, cx, cy, cz, sub, tab);
  density_particle<L>(U, b, n, s, sub, pi, total, list_skip, [&](auto &&f) {
    walk_rows<L, true>([&](uint32_t r) { return tab[r]; }, total, b.spos, sub, s, f);
  });
}

// Tiled form: persistent workgroups take non-empty cell tiles in turn, stage
// the tile's nine candidate rows in LDS and run the same per-particle body
// over them (bitwise the same sums as k_density).  Particles whose stencil
// is not the tile's (outside the grid) and tiles too dense to stage sweep
// from global memory.
template <int L, bool B>
__global__ __launch_bounds__(256) void k_density_t(Uni un, Bufs b, uint32_t n,
                                                   uint32_t list_skip) {
  extern __shared__ float4 stage[];
  __shared__ TileShared S;
  __shared__ RowEnt rows[256 / L][9];
  const uint32_t count = *b.tiles.count;
  for (uint32_t t = blockIdx.x; t < count; t += gridDim.x) {
    const uint4 e = b.tiles.list[t];
    const DevUniforms *__restrict__ U = B ? un.u + e.x : un.u;
    tile_setup(U, b.counts, b.starts, e.y, b.tiles.w, S);
    const bool staged = tile_stage(S, b.spos, stage, b.tiles.cap);
    const uint32_t first = S.own_lo + e.z * b.tiles.chunk;
    const uint32_t own = min(S.own_hi - first, b.tiles.chunk);
    for (uint32_t p0 = 0; p0 < own; p0 += 256 / L) {
      const uint32_t p = p0 + threadIdx.x / L, sub = threadIdx.x % L;
      if (p >= own) continue;
      const uint32_t s = first + p;
      const float4 pi = b.spos[s];
      const int cx = cell_coord(pi.x, U->ox, U->inv_cell);
      const int cy = cell_coord(pi.y, U->oy, U->inv_cell);
      const int cz = cell_coord(pi.z, U->oz, U->inv_cell);
      if (staged && tile_owns(S, U, cx, cy, cz)) {
        const TileRows tr(S, cx, true);
        const uint32_t total = tr.total();
        density_particle<L>(U, b, n, s, sub, pi, total, list_skip, [&](auto &&f) {
          walk_rows<L, true>(tr, total, stage, sub, tr.own(s), f);
        });
      } else {
        RowEnt *tab = rows[threadIdx.x / L];
        const uint32_t total = rows27<L>(U, b.counts, b.starts, cx, cy, cz, sub, tab);
        density_particle<L>(U, b, n, s, sub, pi, total, list_skip, [&](auto &&f) {
          walk_rows<L, true>([&](uint32_t r) { return tab[r]; }, total, b.spos, sub, s, f);
        });
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------
// Pairs per trip of the XSPH and force list walks (their records' loads in
// flight together).
#ifndef FLUID_XSPH_LOADS
#define FLUID_XSPH_LOADS 4
#endif
#ifndef FLUID_FORCE_LOADS
#define FLUID_FORCE_LOADS 2
#endif

// Stage 5 — xsph.wgsl:81-138: the particle's XSPH-corrected velocity w_i
// (in every lane of its group).  With epsilon <= 0 the reference leaves the
// velocity untouched, i.e. equal to the sorted copy.
template <int L>
__device__ __forceinline__ float4 xsph_particle(const DevUniforms *__restrict__ U,
                                                const Bufs &b, uint32_t n, uint32_t s,
                                                uint32_t sub, float4 vi) {
  const float eps = U->eps;
  if (!(eps > 0.0f)) return vi;
  const float4 pi = b.spos[s];
  const float h2 = U->h2, mass = U->mass, ps = U->poly6_scale;
  float cxv = 0.f, cyv = 0.f, czv = 0.f;
  for_each_neighbour<L, FLUID_XSPH_LOADS, 2>(U, b, n, s, sub, pi, [&](uint32_t, const NbrRec<2> &q) {
    const float4 pj = q.r[0], vj = q.r[1];
    const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
    const float r2 = dot3(rx, ry, rz, rx, ry, rz);
    const float d = h2 - r2;
    const float w = ps * d * d * d;
    const float f = mass * __builtin_amdgcn_rcpf(max1(pj.w)) * w;
    cxv += f * (vj.x - vi.x);
    cyv += f * (vj.y - vi.y);
    czv += f * (vj.z - vi.z);
  });
  cxv = group_sum<L>(cxv);
  cyv = group_sum<L>(cyv);
  czv = group_sum<L>(czv);
  return make_float4(vi.x + eps * cxv, vi.y + eps * cyv, vi.z + eps * czv, 0.f);
}

template <int L, bool B>
__global__ __launch_bounds__(256) void k_xsph(Uni un, Bufs b, uint32_t n, uint32_t xcd) {
  const uint32_t g = xcd_block(xcd) * blockDim.x + threadIdx.x;
  const uint32_t s = g / L, sub = g % L;
  if (s >= n) return;
  const DevUniforms *__restrict__ U = uni_at<B>(un, s);
  const float4 w = xsph_particle<L>(U, b, n, s, sub, b.svel[s]);
  if (sub == 0) b.vel[s] = w;
}

// ------------------------------------------------------------------------
// Stage 6 — sph_force_3d_grid.wgsl:105-227: near pressure + close-range
// separation, Monaghan viscosity, Akinci cohesion and curvature, gravity.
// The per-pair quotients are regrouped around one reciprocal each
// (sp = -m/(2 rho_i) (P_i/rho^n_j + k_near), the viscosity and K_ij share
// 1/(rho_i + rho_j)); r < 0.1 h, which switches the direction, is decided on
// the exact bits the oracle uses.  vi is the particle's post-XSPH velocity;
// neighbours contribute their pre-XSPH velocity (sph_force_3d_grid.wgsl
// :107-119 reads the sorted copy).
template <int L>
__device__ __forceinline__ void force_particle(const DevUniforms *__restrict__ U, const Bufs &b,
                                               uint32_t n, uint32_t s, uint32_t sub,
                                               float4 vi, uint32_t fuse_predict) {
  const float4 *rec_i = b.srec + 3 * (size_t)s;
  const float4 pi = rec_i[0];
  const float4 ni = rec_i[2];
  const float di = max1(pi.w);
  const float ndi = max1(rec_i[1].w);
  const float h = U->h, mass = U->mass, near_k = U->near_k;
  const float gs = U->grad_scale, md = U->min_dist;
  const float veps = U->visc_eps, aks = U->akinci_scale;
  const float hh = U->half_h, h6_64 = U->h6_64;
  const float close_r2 = U->close_r2;
  const float npi = near_k * ndi;
  const float a_sp = -mass / (2.0f * di);
  const float c_visc = 2.0f * mass * U->visc * h;  // -m * (-visc * h / (avg))
  const float st = U->st;
  const float c_k = 2.0f * U->rho0 * mass * st;    // surface tension folded in
  const float inv_md = 1.0f / md;
  // one accumulator: a = sum (sp + st mk coh) dir - st mk (n_i - n_j)
  float fx = 0.f, fy = 0.f, fz = 0.f;
  for_each_neighbour<L, FLUID_FORCE_LOADS, 3>(U, b, n, s, sub, pi, [&](uint32_t j, const NbrRec<3> &q) {
    const float4 pj = q.r[0], vj = q.r[1], nj = q.r[2];
    const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
    const float r2 = dot3(rx, ry, rz, rx, ry, rz);
    const bool close = r2 < close_r2;
    const float il = __builtin_amdgcn_rsqf(r2);  // 1 / r
    const float r = r2 * il;
    const float dj = max1(pj.w);
    const float dr = h - r;
    const float grad = gs * dr * dr;
    // near pressure (:183) and separation (:186-189)
    float sp = a_sp * (npi * nj.w + near_k) * grad;  // nj.w = 1 / max(rho_near_j, 1)
    if (close) sp -= 200.0f * (1.0f - r * inv_md);
    // Monaghan viscosity (:193-202): -m * pi_visc * grad, only if v.x < 0
    const float vdx = dot3(vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, rx, ry, rz);
    const float inv_dd = __builtin_amdgcn_rcpf(di + dj);
    sp += c_visc * fminf(vdx, 0.0f) * inv_dd * __builtin_amdgcn_rcpf(r2 + veps) * grad;
    // Akinci cohesion + curvature (:206-214), times surface tension
    const float mk = c_k * inv_dd;
    const float r3 = r * r * r;
    const float dr3 = dr * dr * dr;
    // both pieces, then a select (2 dr3 r3 = 2 (dr3 r3) exactly)
    const float t3 = dr3 * r3;
    const float coh = aks * ((r >= hh) ? t3 : __builtin_fmaf(2.0f, t3, -h6_64));
    const float coef = __builtin_fmaf(mk, coh, sp);
    // dir = -r_vec / r, or the slot-difference direction below 0.1 h, which
    // is rare: the common update runs with a zero scale for close pairs
    const float c = close ? 0.0f : -coef * il;
    fx = __builtin_fmaf(c, rx, fx);
    fy = __builtin_fmaf(c, ry, fy);
    fz = __builtin_fmaf(c, rz, fz);
    if (close) {
      const float3 d3 = close_dir(s, j);
      fx = __builtin_fmaf(coef, d3.x, fx);
      fy = __builtin_fmaf(coef, d3.y, fy);
      fz = __builtin_fmaf(coef, d3.z, fz);
    }
    fx = __builtin_fmaf(-mk, ni.x - nj.x, fx);
    fy = __builtin_fmaf(-mk, ni.y - nj.y, fy);
    fz = __builtin_fmaf(-mk, ni.z - nj.z, fz);
  });
  fx = group_sum<L>(fx); fy = group_sum<L>(fy); fz = group_sum<L>(fz);
  if (sub != 0) return;
  fx += U->grav[0];
  fy += U->grav[1];
  fz += U->grav[2];
  b.frc[s] = make_float4(fx, fy, fz, 0.f);
  if (fuse_predict) {
    // pcisph_predict.wgsl:57-72 for this particle, on the values in hand
    const float dt = U->dt;
    const float vx = vi.x + dt * fx, vy = vi.y + dt * fy, vz = vi.z + dt * fz;
    b.pred[0].xq[s] = make_float4(pi.x + dt * vx, pi.y + dt * vy, pi.z + dt * vz, 0.f);
    b.pred[0].pr[s] = make_float2(0.f, 0.f);
    b.pred[0].v[s] = make_float4(vx, vy, vz, 0.f);
  }
}

template <int L, bool B>
__global__ __launch_bounds__(256) void k_force(Uni un, Bufs b, uint32_t n, uint32_t xcd,
                                               uint32_t fuse_predict) {
  const uint32_t g = xcd_block(xcd) * blockDim.x + threadIdx.x;
  const uint32_t s = g / L, sub = g % L;
  if (s >= n) return;
  const DevUniforms *__restrict__ U = uni_at<B>(un, s);
  force_particle<L>(U, b, n, s, sub, b.vel[s], fuse_predict);
}

// XSPH and force in one pass per particle: the list is walked for XSPH,
// the corrected velocity stays in registers (and is stored for the later
// stages) and the same list, now in cache, is walked for the force.  Force
// reads only the particle's own corrected velocity, so this gives the bits
// of the two separate kernels.
template <int L, bool B>
__global__ __launch_bounds__(256) void k_xsph_force(Uni un, Bufs b, uint32_t n, uint32_t xcd,
                                                    uint32_t fuse_predict) {
  const uint32_t g = xcd_block(xcd) * blockDim.x + threadIdx.x;
  const uint32_t s = g / L, sub = g % L;
  if (s >= n) return;
  const DevUniforms *__restrict__ U = uni_at<B>(un, s);
  const float4 w = xsph_particle<L>(U, b, n, s, sub, b.svel[s]);
  if (sub == 0) b.vel[s] = w;
  force_particle<L>(U, b, n, s, sub, w, fuse_predict);
}

// ------------------------------------------------------------------------
// Stage 7 — pcisph_predict.wgsl:50-73, run per sorted slot (the canonical
// state already lives in slot order).
template <bool B>
__global__ __launch_bounds__(256) void k_predict(Uni un, Bufs b, uint32_t n) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const DevUniforms *__restrict__ U = uni_at<B>(un, s);
  const float dt = U->dt;
  const float4 p = b.spos[s], v = b.vel[s], a = b.frc[s];
  const float vx = v.x + dt * a.x, vy = v.y + dt * a.y, vz = v.z + dt * a.z;
  b.pred[0].xq[s] = make_float4(p.x + dt * vx, p.y + dt * vy, p.z + dt * vz, 0.f);
  b.pred[0].pr[s] = make_float2(0.f, 0.f);
  b.pred[0].v[s] = make_float4(vx, vy, vz, 0.f);
}

// ------------------------------------------------------------------------
// Stage 8 — pcisph_solve.wgsl:172-208, one iteration, in -> out.  The
// neighbour's p / max(rho*,1)^2 is precomputed by the producing iteration
// (same operations as the reference evaluates per pair).  sweep(f) walks
// the candidates of the particle's cell (the grid of the current positions)
// over the iteration's predicted positions, lane-independently.
//
// The first iteration also records, per lane and in its walk order, the
// candidates within the solve skin of the particle (1.5 h plus a margin at
// the first iteration's predicted positions); every later iteration's
// in-range pairs are among them, so the later iterations walk that list and
// sum the same pairs in the same order per lane (the same bits as a sweep).
// With LS < L (later iterations only), L / LS walkers share each of the LS
// list lanes, taking interleaved depths (the same pairs, grouped otherwise;
// an overflowed list lane's walkers sweep its candidates between them).
template <int L, class Sweep, int LS = L>
__device__ __forceinline__ void solve_particle(const DevUniforms *__restrict__ U, const Bufs &b,
                                               uint32_t n, const PredBuf &in, const PredBuf &out,
                                               float4 *ap_out, uint32_t first_iter,
                                               uint32_t use_list, uint32_t s, uint32_t sub,
                                               float4 xi, float4 p0, Sweep &&sweep) {
  static_assert(L % LS == 0, "walkers are a multiple of the list lanes");
  constexpr uint32_t M = L / LS;
  const float h = U->h, h2 = U->h2, mass = U->mass, ps = U->poly6_scale;
  const float gs = U->grad_scale;
  const uint32_t kl = sub % LS, d0 = sub / LS;
  uint32_t *__restrict__ lst = b.slist + lane_block<LS>(s, kSolveListCapacity) + kl;
  constexpr uint32_t kCap = kSolveListCapacity / LS;
  float pd = 0.f, ax = 0.f, ay = 0.f, az = 0.f;
  if (M == 1 && first_iter) {
    // First iteration: every pressure of the input is 0 (pcisph_predict.wgsl
    // :69), so each pressure term is a signed zero and a_total = a_np
    // exactly; only the predicted density needs the sum.
    const float skin2 = U->solve_skin2, mps = mass * ps;
    uint32_t cnt = 0;
    sweep([&](uint32_t j, float4 xj, bool) {
      const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
      const float r2 = dot3(rx, ry, rz, rx, ry, rz);
      // h2 - r2 > 0 exactly when r2 < h2, so the max is the support cut
      const float d = fmaxf(h2 - r2, 0.f);
      pd = __builtin_fmaf(mps, d * d * d, pd);
      const bool keep = use_list && r2 < skin2;
      if (keep && cnt < kCap) lst[cnt * 64u] = j;
      cnt += keep ? 1u : 0u;
    });
    if (use_list) b.scnt[(size_t)s * LS + kl] = cnt <= kCap ? cnt : kSolveListOverflow;
  } else if (M > 1 && first_iter) {
    // First iteration with M sweep lanes per list lane (a group-synchronous
    // sweep): the group's skin entries are ranked by ballot into the LS list
    // lanes, entry e at lane e % LS, depth e / LS, the layout the later
    // iterations walk.
    const float skin2 = U->solve_skin2, mps = mass * ps;
    uint32_t *__restrict__ blk = b.slist + lane_block<LS>(s, kSolveListCapacity);
    uint32_t run = 0;
    sweep([&](uint32_t j, float4 xj, bool valid) {
      const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
      const float r2 = dot3(rx, ry, rz, rx, ry, rz);
      const float d = valid ? fmaxf(h2 - r2, 0.f) : 0.f;
      pd = __builtin_fmaf(mps, d * d * d, pd);
      group_list_append<L, LS, kSolveListCapacity>(blk, sub, run, valid && r2 < skin2, j);
    });
    if (sub < LS)
      b.scnt[(size_t)s * LS + sub] = run > kSolveListCapacity ? kSolveListOverflow
                                     : run > sub ? (run - 1 - sub) / LS + 1 : 0u;
  } else {
    auto body = [&](uint32_t, float4 xj, bool) {
      const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
      const float r2 = dot3(rx, ry, rz, rx, ry, rz);
      if (r2 < h2) {
        const float d = h2 - r2;
        pd += mass * (ps * d * d * d);
        if (r2 > 1e-12f) {
          // -m (q_i + q_j) grad * dir with dir = -r_vec / r
          const float il = __builtin_amdgcn_rsqf(r2);
          const float dr = h - r2 * il;
          const float sc = mass * (xi.w + xj.w) * (gs * dr * dr) * il;
          ax += sc * rx;
          ay += sc * ry;
          az += sc * rz;
        }
      }
    };
    const uint32_t cntk = use_list ? b.scnt[(size_t)s * LS + kl] : kSolveListOverflow;
    if (cntk != kSolveListOverflow) {
      // this walker's depths d0, d0 + M, ... of list lane kl
      const uint32_t cnt = cntk > d0 ? (cntk - 1 - d0) / M + 1 : 0;
      const uint32_t *__restrict__ p = lst + d0 * 64u;
      constexpr uint32_t kStep = 64u * M;
      const float4 *__restrict__ xq = in.xq;
      uint32_t k = 0;
      for (; k + 3 < cnt; k += 4, p += 4 * kStep) {
        const uint32_t j0 = p[0], j1 = p[kStep], j2 = p[2 * kStep], j3 = p[3 * kStep];
        const float4 x0 = xq[j0], x1 = xq[j1], x2 = xq[j2], x3 = xq[j3];
        body(j0, x0, true);
        body(j1, x1, true);
        body(j2, x2, true);
        body(j3, x3, true);
      }
      if (k < cnt) {  // the last (fewer than four) entries as one predicated trip
        uint32_t jt[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) jt[q] = k + q < cnt ? p[q * kStep] : s;
        float4 xt[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) xt[q] = xq[jt[q]];
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (k + q < cnt) body(jt[q], xt[q], true);
      }
    } else {
      sweep(body);
    }
  }
  pd = group_sum<L>(pd);
  ax = group_sum<L>(ax);
  ay = group_sum<L>(ay);
  az = group_sum<L>(az);
  if (sub != 0) return;
  const float old_p = in.pr[s].x;
  const float err = fmaxf(0.0f, pd - U->rho0);
  const float new_p = old_p + U->delta * err;
  if (ap_out) ap_out[s] = make_float4(ax, ay, az, 0.f);
  const float amag = dot3(ax, ay, az, ax, ay, az);
  if (amag > U->max_accel_sq) {
    const float f = U->max_accel / sqrtf(amag);
    ax *= f; ay *= f; az *= f;
  }
  const float dt = U->dt;
  const float4 v0 = b.vel[s], a0 = b.frc[s];
  const float vx = v0.x + dt * (a0.x + ax);
  const float vy = v0.y + dt * (a0.y + ay);
  const float vz = v0.z + dt * (a0.z + az);
  const float nd = fmaxf(pd, 1.0f);
  out.xq[s] = make_float4(p0.x + dt * vx, p0.y + dt * vy, p0.z + dt * vz,
                          new_p / (nd * nd));
  out.pr[s] = make_float2(new_p, pd);
  out.v[s] = make_float4(vx, vy, vz, 0.f);
}

// The row table of a solve sweep, built (by the whole group) only when some
// lane of the group sweeps: in the first iteration, and in a later one for
// a lane whose skin list (list lane sub % LS) overflowed.
template <int L, int LS>
__device__ __forceinline__ uint32_t solve_rows(const DevUniforms *__restrict__ U, const Bufs &b,
                                               uint32_t s, uint32_t sub, int cx, int cy, int cz,
                                               uint32_t first_iter, uint32_t use_list,
                                               RowEnt *tab) {
  const bool need = first_iter || !use_list ||
                    b.scnt[(size_t)s * LS + sub % LS] == kSolveListOverflow;
  if (!group_any<L>(need)) return 0u;
  return rows27<L>(U, b.counts, b.starts, cx, cy, cz, sub, tab);
}

template <int L, bool B>
__global__ __launch_bounds__(256) void k_solve(Uni un, Bufs b, uint32_t n, PredBuf in,
                                               PredBuf out, float4 *ap_out, uint32_t xcd,
                                               uint32_t first_iter, uint32_t use_list) {
  __shared__ RowEnt rows[256 / L][9];
  const uint32_t g = xcd_block(xcd) * blockDim.x + threadIdx.x;
  const uint32_t s = g / L, sub = g % L;
  if (s >= n) return;
  const DevUniforms *__restrict__ U = uni_at<B>(un, s);
  const float4 xi = in.xq[s];
  const float4 p0 = b.spos[s];  // current position (grid lookup)
  const int cx = cell_coord(p0.x, U->ox, U->inv_cell);
  const int cy = cell_coord(p0.y, U->oy, U->inv_cell);
  const int cz = cell_coord(p0.z, U->oz, U->inv_cell);
  RowEnt *tab = rows[threadIdx.x / L];
  const uint32_t total =
      solve_rows<L, L>(U, b, s, sub, cx, cy, cz, first_iter, use_list, tab);
  solve_particle<L>(U, b, n, in, out, ap_out, first_iter, use_list, s, sub, xi, p0, [&](auto &&f) {
    walk_rows<L, false>([&](uint32_t r) { return tab[r]; }, total, in.xq, sub, s, f);
  });
}

// Later PCISPH iterations with L = LS * M walkers over the LS-lane skin lists
// the first iteration recorded (no recording here).
template <int L, int LS, bool B>
__global__ __launch_bounds__(256) void k_solve_w(Uni un, Bufs b, uint32_t n, PredBuf in,
                                                 PredBuf out, float4 *ap_out, uint32_t xcd) {
  __shared__ RowEnt rows[256 / L][9];
  const uint32_t g = xcd_block(xcd) * blockDim.x + threadIdx.x;
  const uint32_t s = g / L, sub = g % L;
  if (s >= n) return;
  const DevUniforms *__restrict__ U = uni_at<B>(un, s);
  const float4 xi = in.xq[s];
  const float4 p0 = b.spos[s];
  const int cx = cell_coord(p0.x, U->ox, U->inv_cell);
  const int cy = cell_coord(p0.y, U->oy, U->inv_cell);
  const int cz = cell_coord(p0.z, U->oz, U->inv_cell);
  RowEnt *tab = rows[threadIdx.x / L];
  const uint32_t total = solve_rows<L, LS>(U, b, s, sub, cx, cy, cz, 0u, 1u, tab);
  auto sweep = [&](auto &&f) {
    walk_rows<L, false>([&](uint32_t r) { return tab[r]; }, total, in.xq, sub, s, f);
  };
  solve_particle<L, decltype(sweep) &, LS>(U, b, n, in, out, ap_out, 0u, 1u, s, sub, xi, p0,
                                           sweep);
}

// The first PCISPH iteration with L = LS * M lanes sweeping (group-
// synchronously) and recording LS-lane skin lists for the later iterations.
template <int L, int LS, bool B>
__global__ __launch_bounds__(256) void k_solve_fw(Uni un, Bufs b, uint32_t n, PredBuf in,
                                                  PredBuf out, float4 *ap_out, uint32_t xcd) {
  __shared__ RowEnt rows[256 / L][9];
  const uint32_t g = xcd_block(xcd) * blockDim.x + threadIdx.x;
  const uint32_t s = g / L, sub = g % L;
  if (s >= n) return;
  const DevUniforms *__restrict__ U = uni_at<B>(un, s);
  const float4 xi = in.xq[s];
  const float4 p0 = b.spos[s];
  const int cx = cell_coord(p0.x, U->ox, U->inv_cell);
  const int cy = cell_coord(p0.y, U->oy, U->inv_cell);
  const int cz = cell_coord(p0.z, U->oz, U->inv_cell);
  RowEnt *tab = rows[threadIdx.x / L];
  const uint32_t total = rows27<L>(U, b.counts, b.starts, cx, cy, cz, sub, tab);
  auto sweep = [&](auto &&f) {
    walk_rows<L, true>([&](uint32_t r) { return tab[r]; }, total, in.xq, sub, s, f);
  };
  solve_particle<L, decltype(sweep) &, LS>(U, b, n, in, out, ap_out, 1u, 1u, s, sub, xi, p0,
                                           sweep);
}

// Tiled form of k_solve (see k_density_t): the tile's candidate rows of the
// predicted state are staged in LDS.
template <int L, bool B>
__global__ __launch_bounds__(256) void k_solve_t(Uni un, Bufs b, uint32_t n, PredBuf in,
                                                 PredBuf out, float4 *ap_out,
                                                 uint32_t first_iter, uint32_t use_list) {
  extern __shared__ float4 stage[];
  __shared__ TileShared S;
  __shared__ RowEnt rows[256 / L][9];
  const uint32_t count = *b.tiles.count;
  for (uint32_t t = blockIdx.x; t < count; t += gridDim.x) {
    const uint4 e = b.tiles.list[t];
    const DevUniforms *__restrict__ U = B ? un.u + e.x : un.u;
    tile_setup(U, b.counts, b.starts, e.y, b.tiles.w, S);
    const bool staged = tile_stage(S, in.xq, stage, b.tiles.cap);
    const uint32_t first = S.own_lo + e.z * b.tiles.chunk;
    const uint32_t own = min(S.own_hi - first, b.tiles.chunk);
    for (uint32_t q0 = 0; q0 < own; q0 += 256 / L) {
      const uint32_t q = q0 + threadIdx.x / L, sub = threadIdx.x % L;
      if (q >= own) continue;
      const uint32_t s = first + q;
      const float4 xi = in.xq[s];
      const float4 p0 = b.spos[s];
      const int cx = cell_coord(p0.x, U->ox, U->inv_cell);
      const int cy = cell_coord(p0.y, U->oy, U->inv_cell);
      const int cz = cell_coord(p0.z, U->oz, U->inv_cell);
      if (staged && tile_owns(S, U, cx, cy, cz)) {
        const TileRows tr(S, cx, true);
        const uint32_t total = tr.total();
        solve_particle<L>(U, b, n, in, out, ap_out, first_iter, use_list, s, sub, xi, p0,
                          [&](auto &&f) {
          walk_rows<L, false>(tr, total, stage, sub, tr.own(s), f);
        });
      } else {
        RowEnt *tab = rows[threadIdx.x / L];
        const uint32_t total =
            solve_rows<L, L>(U, b, s, sub, cx, cy, cz, first_iter, use_list, tab);
        solve_particle<L>(U, b, n, in, out, ap_out, first_iter, use_list, s, sub, xi, p0,
                          [&](auto &&f) {
          walk_rows<L, false>([&](uint32_t r) { return tab[r]; }, total, in.xq, sub, s, f);
        });
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------
// Stage 9 — pcisph_finalize.wgsl:50-74: force <- v* - v, velocity <- v*.
__global__ __launch_bounds__(256) void k_finalize(Bufs b, uint32_t n, PredBuf fin) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const float4 cv = fin.v[s], ov = b.vel[s];
  b.frc[s] = make_float4(cv.x - ov.x, cv.y - ov.y, cv.z - ov.z, 0.f);
  b.vel[s] = make_float4(cv.x, cv.y, cv.z, 0.f);
}

// ------------------------------------------------------------------------
// Stage 10 — sph_integrate_3d.wgsl:91-388.
__device__ __forceinline__ float wsign(float v) {
  return v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f);
}
__device__ __forceinline__ int32_t to_i32(float v) {
  // WGSL f32 -> i32: truncate, saturate, NaN -> 0 (v_cvt_i32_f32)
  return (int32_t)v;
}
__device__ float sdf_fetch(const float *sdf, int r, int x, int y, int z) {
  x = min(max(x, 0), r - 1);
  y = min(max(y, 0), r - 1);
  z = min(max(z, 0), r - 1);
  return sdf[(size_t)x + (size_t)y * r + (size_t)z * r * r];
}
// Linear, clamp-to-edge texture sample (sph_3d_grid.rs:336-344).
__device__ float sdf_sample(const float *sdf, uint32_t res, float ux, float uy,
                            float uz) {
  if (res == 0) return 1.0f;  // the reference's 1x1x1 dummy texture of 1.0
  const float u = ux * (float)res - 0.5f, v = uy * (float)res - 0.5f,
              w = uz * (float)res - 0.5f;
  const float fu = floorf(u), fv = floorf(v), fw = floorf(w);
  const int x0 = (int)fu, y0 = (int)fv, z0 = (int)fw, r = (int)res;
  const float a = u - fu, bb = v - fv, c = w - fw;
  const float c000 = sdf_fetch(sdf, r, x0, y0, z0), c100 = sdf_fetch(sdf, r, x0 + 1, y0, z0);
  const float c010 = sdf_fetch(sdf, r, x0, y0 + 1, z0), c110 = sdf_fetch(sdf, r, x0 + 1, y0 + 1, z0);
  const float c001 = sdf_fetch(sdf, r, x0, y0, z0 + 1), c101 = sdf_fetch(sdf, r, x0 + 1, y0, z0 + 1);
  const float c011 = sdf_fetch(sdf, r, x0, y0 + 1, z0 + 1), c111 = sdf_fetch(sdf, r, x0 + 1, y0 + 1, z0 + 1);
  const float x00 = c000 + (c100 - c000) * a, x10 = c010 + (c110 - c010) * a;
  const float x01 = c001 + (c101 - c001) * a, x11 = c011 + (c111 - c011) * a;
  const float y0v = x00 + (x10 - x00) * bb, y1v = x01 + (x11 - x01) * bb;
  return y0v + (y1v - y0v) * c;
}

template <bool B>
__global__ __launch_bounds__(256) void k_integrate(Uni un, Bufs b, uint32_t n,
                                                   const float4 *__restrict__ fin_v,
                                                   uint32_t ntiles) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  // the scan's tile states and the tile-item count, cleared for the next
  // step's scan (every reader of them ran before this pass)
  if (s < ntiles) b.tile_status[s] = 0ull;
  if (s == ntiles) {
    *b.tile_counter = 0u;
    if (b.tiles.count) *b.tiles.count = 0u;
  }
  if (s >= n) return;
  const DevUniforms *__restrict__ U = uni_at<B>(un, s);
  const float4 p4 = b.spos[s];
  float4 v4 = b.vel[s];
  if (fin_v) {
    // pcisph_finalize.wgsl:50-74 for this particle first: force <- v* - v,
    // velocity <- v*
    const float4 cv = fin_v[s];
    b.frc[s] = make_float4(cv.x - v4.x, cv.y - v4.y, cv.z - v4.z, 0.f);
    v4 = make_float4(cv.x, cv.y, cv.z, 0.f);
  }
  float px = p4.x, py = p4.y, pz = p4.z, vx = v4.x, vy = v4.y, vz = v4.z;
  const float(*iv)[4] = U->iv;
  const float(*fw)[4] = U->fw;
  // world_to_local / world_dir_to_local (container_common.wgsl:27-56)
  const float qy = py - U->cy;
  const float lpx = dot3(iv[0][0], iv[0][1], iv[0][2], px, qy, pz);
  const float lpy = dot3(iv[1][0], iv[1][1], iv[1][2], px, qy, pz);
  const float lpz = dot3(iv[2][0], iv[2][1], iv[2][2], px, qy, pz);
  const float lvx = dot3(iv[0][0], iv[0][1], iv[0][2], vx, vy, vz);
  const float lvy = dot3(iv[1][0], iv[1][1], iv[1][2], vx, vy, vz);
  const float lvz = dot3(iv[2][0], iv[2][1], iv[2][2], vx, vy, vz);
  const float bl = U->boundary_layer, k = U->wall_k, wd = 8.0f;
  const float hw = U->hw, hh = U->hh, hd = U->hd;
  float wx = 0.f, wy = 0.f, wz = 0.f;
  float d = lpx - (-hw);
  if (d < bl) { const float t = 1.0f - d / bl; wx += k * t * t; if (lvx < 0.f) wx -= lvx * wd * t; }
  d = hw - lpx;
  if (d < bl) { const float t = 1.0f - d / bl; wx -= k * t * t; if (lvx > 0.f) wx -= lvx * wd * t; }
  d = lpy - (-hh);
  if (d < bl) { const float t = 1.0f - d / bl; wy += k * t * t; if (lvy < 0.f) wy -= lvy * wd * t; }
  d = hh - lpy;
  if (d < bl) { const float t = 1.0f - d / bl; wy -= k * t * t; if (lvy > 0.f) wy -= lvy * wd * t; }
  d = lpz - (-hd);
  if (d < bl) { const float t = 1.0f - d / bl; wz += k * t * t; if (lvz < 0.f) wz -= lvz * wd * t; }
  d = hd - lpz;
  if (d < bl) { const float t = 1.0f - d / bl; wz -= k * t * t; if (lvz > 0.f) wz -= lvz * wd * t; }
  float ax = 0.f + dot3(fw[0][0], fw[0][1], fw[0][2], wx, wy, wz);
  float ay = 0.f + dot3(fw[1][0], fw[1][1], fw[1][2], wx, wy, wz);
  float az = 0.f + dot3(fw[2][0], fw[2][1], fw[2][2], wx, wy, wz);

  const GpuMouseForce &m = U->mouse;
  if (m.is_active == 1u) {  // :163-200
    const float tx = m.position[0] - px, ty = m.position[1] - py, tz = m.position[2] - pz;
    const float dist = sqrtf(dot3(tx, ty, tz, tx, ty, tz));
    if (dist < m.radius && dist > 0.001f) {
      const float fall = 1.0f - (dist / m.radius);
      const float fx = tx / dist, fy = ty / dist, fz = tz / dist;
      const float sm = m.strength * fall;
      if (m.mode == 1u) {
        ax += fx * sm; ay += fy * sm; az += fz * sm;
      } else if (m.mode == 2u) {
        const float dl = sqrtf(dot3(m.direction[0], m.direction[1], m.direction[2],
                                    m.direction[0], m.direction[1], m.direction[2]));
        const float kx = m.direction[0] / dl, ky = m.direction[1] / dl, kz = m.direction[2] / dl;
        const float gx = fy * kz - fz * ky, gy = fz * kx - fx * kz, gz = fx * ky - fy * kx;
        const float tl = sqrtf(dot3(gx, gy, gz, gx, gy, gz));
        if (tl > 0.001f) { ax += (gx / tl) * sm; ay += (gy / tl) * sm; az += (gz / tl) * sm; }
      } else if (m.mode == 3u) {
        ax -= fx * sm * 3.0f; ay -= fy * sm * 3.0f; az -= fz * sm * 3.0f;
      } else if (m.mode == 4u) {
        const float qx = fx, qy2 = fy + -1.0f, qz = fz;
        const float ql = sqrtf(dot3(qx, qy2, qz, qx, qy2, qz));
        ax += (qx / ql) * sm; ay += (qy2 / ql) * sm; az += (qz / ql) * sm;
      } else {
        ax -= fx * sm; ay -= fy * sm; az -= fz * sm;
      }
    }
  }

  const GpuRigidBody &rb = U->rb;
  if (rb.is_active != 0u) {  // :203-341
    const float wrx = px - rb.position[0], wry = py - rb.position[1], wrz = pz - rb.position[2];
    const float rlx = dot3(rb.rot_row0[0], rb.rot_row0[1], rb.rot_row0[2], wrx, wry, wrz);
    const float rly = dot3(rb.rot_row1[0], rb.rot_row1[1], rb.rot_row1[2], wrx, wry, wrz);
    const float rlz = dot3(rb.rot_row2[0], rb.rot_row2[1], rb.rot_row2[2], wrx, wry, wrz);
    const float he = rb.half_extent;
    float sdf, lnx = 0.f, lny = 0.f, lnz = 0.f;
    if (rb.shape == 1u) {
      const float dist = sqrtf(dot3(rlx, rly, rlz, rlx, rly, rlz));
      sdf = dist - he;
      if (dist > 0.001f) { lnx = rlx / dist; lny = rly / dist; lnz = rlz / dist; }
      else { lny = 1.f; }
    } else if (rb.shape == 2u) {
      const float rd = sqrtf(__builtin_fmaf(rlz, rlz, rlx * rlx));
      const float d_rad = rd - he, d_cap = fabsf(rly) - he;
      sdf = fmaxf(d_rad, d_cap);
      if (d_rad > d_cap) {
        if (rd > 0.001f) { lnx = rlx / rd; lnz = rlz / rd; } else { lnx = 1.f; }
      } else {
        lny = wsign(rly);
      }
    } else if (rb.shape == 3u) {
      const float minor = he * 0.3f;
      const float xz = sqrtf(__builtin_fmaf(rlz, rlz, rlx * rlx));
      const float qx = xz - he, qy3 = rly;
      const float ql = sqrtf(__builtin_fmaf(qy3, qy3, qx * qx));
      sdf = ql - minor;
      if (ql > 0.001f && xz > 0.001f) {
        const float rdx = qx / ql, rdy = qy3 / ql;
        lnx = rdx * rlx / xz; lny = rdy; lnz = rdx * rlz / xz;
      } else {
        lny = 1.f;
      }
    } else if (rb.shape == 4u) {
      const float ux = (rlx / he) * 0.5f + 0.5f, uy = (rly / he) * 0.5f + 0.5f,
                  uz = (rlz / he) * 0.5f + 0.5f;
      const uint32_t res = U->sdf_res;
      sdf = sdf_sample(b.sdf, res, ux, uy, uz) * he;
      const float e = 1.0f / 32.0f;
      const float gxv = sdf_sample(b.sdf, res, ux + e, uy, uz) - sdf_sample(b.sdf, res, ux - e, uy, uz);
      const float gyv = sdf_sample(b.sdf, res, ux, uy + e, uz) - sdf_sample(b.sdf, res, ux, uy - e, uz);
      const float gzv = sdf_sample(b.sdf, res, ux, uy, uz + e) - sdf_sample(b.sdf, res, ux, uy, uz - e);
      const float gl = sqrtf(dot3(gxv, gyv, gzv, gxv, gyv, gzv));
      if (gl > 1e-6f) { lnx = gxv / gl; lny = gyv / gl; lnz = gzv / gl; }
      else {
        const float qy4 = rly + 1e-6f;
        const float l = sqrtf(dot3(rlx, qy4, rlz, rlx, qy4, rlz));
        lnx = rlx / l; lny = qy4 / l; lnz = rlz / l;
      }
    } else {
      const float dx = fabsf(rlx) - he, dy = fabsf(rly) - he, dz = fabsf(rlz) - he;
      sdf = fmaxf(dx, fmaxf(dy, dz));
      if (dx >= dy && dx >= dz) lnx = wsign(rlx);
      else if (dy >= dx && dy >= dz) lny = wsign(rly);
      else lnz = wsign(rlz);
    }
    const float ir = U->interact_range;
    if (sdf < ir) {
      const float nwx = rb.rot_row0[0] * lnx + rb.rot_row1[0] * lny + rb.rot_row2[0] * lnz;
      const float nwy = rb.rot_row0[1] * lnx + rb.rot_row1[1] * lny + rb.rot_row2[1] * lnz;
      const float nwz = rb.rot_row0[2] * lnx + rb.rot_row1[2] * lny + rb.rot_row2[2] * lnz;
      const float t = (ir - sdf) / ir;
      const float pnx = nwx * rb.stiffness * t * t, pny = nwy * rb.stiffness * t * t,
                  pnz = nwz * rb.stiffness * t * t;
      ax += pnx; ay += pny; az += pnz;
      float dmx = 0.f, dmy = 0.f, dmz = 0.f;
      const float vn = dot3(vx - rb.velocity[0], vy - rb.velocity[1], vz - rb.velocity[2],
                            nwx, nwy, nwz);
      if (vn < 0.f) {
        dmx = -nwx * vn * 5.0f; dmy = -nwy * vn * 5.0f; dmz = -nwz * vn * 5.0f;
        ax += dmx; ay += dmy; az += dmz;
      }
      const float mass = U->mass;
      const float rex = -(pnx + dmx) * mass, rey = -(pny + dmy) * mass, rez = -(pnz + dmz) * mass;
      atomicAdd(&b.accum[0], to_i32(rex * 1000.0f));
      atomicAdd(&b.accum[1], to_i32(rey * 1000.0f));
      atomicAdd(&b.accum[2], to_i32(rez * 1000.0f));
      atomicAdd(reinterpret_cast<uint32_t *>(&b.accum[3]), 1u);
      const float tqx = wry * rez - wrz * rey, tqy = wrz * rex - wrx * rez,
                  tqz = wrx * rey - wry * rex;
      atomicAdd(&b.accum[4], to_i32(tqx * 1000.0f));
      atomicAdd(&b.accum[5], to_i32(tqy * 1000.0f));
      atomicAdd(&b.accum[6], to_i32(tqz * 1000.0f));
    }
  }

  const float dt = U->dt;
  vx = vx + dt * ax; vy = vy + dt * ay; vz = vz + dt * az;
  px = px + dt * vx; py = py + dt * vy; pz = pz + dt * vz;
  // hard boundary backstop (:350-384)
  const float qy5 = py - U->cy;
  float nlx = dot3(iv[0][0], iv[0][1], iv[0][2], px, qy5, pz);
  float nly = dot3(iv[1][0], iv[1][1], iv[1][2], px, qy5, pz);
  float nlz = dot3(iv[2][0], iv[2][1], iv[2][2], px, qy5, pz);
  float mvx = dot3(iv[0][0], iv[0][1], iv[0][2], vx, vy, vz);
  float mvy = dot3(iv[1][0], iv[1][1], iv[1][2], vx, vy, vz);
  float mvz = dot3(iv[2][0], iv[2][1], iv[2][2], vx, vy, vz);
  const float rest = U->damping;
  if (nlx < -hw) { nlx = -hw; mvx = fabsf(mvx) * rest; }
  else if (nlx > hw) { nlx = hw; mvx = -fabsf(mvx) * rest; }
  if (nly < -hh) { nly = -hh; mvy = fabsf(mvy) * rest; }
  else if (nly > hh) { nly = hh; mvy = -fabsf(mvy) * rest; }
  if (nlz < -hd) { nlz = -hd; mvz = fabsf(mvz) * rest; }
  else if (nlz > hd) { nlz = hd; mvz = -fabsf(mvz) * rest; }
  const float ox = dot3(fw[0][0], fw[0][1], fw[0][2], nlx, nly, nlz);
  const float oy = dot3(fw[1][0], fw[1][1], fw[1][2], nlx, nly, nlz) + U->cy;
  const float oz = dot3(fw[2][0], fw[2][1], fw[2][2], nlx, nly, nlz);
  b.pos[s] = make_float4(ox, oy, oz, 0.f);
  b.vel[s] = make_float4(dot3(fw[0][0], fw[0][1], fw[0][2], mvx, mvy, mvz),
                         dot3(fw[1][0], fw[1][1], fw[1][2], mvx, mvy, mvz),
                         dot3(fw[2][0], fw[2][1], fw[2][2], mvx, mvy, mvz), 0.f);
}

// Candidate / in-range pair totals (and per-particle maxima) of the density
// stencil: out4 = {sum candidates, sum in range, max candidates, max in range}.
template <bool B>
__global__ __launch_bounds__(256) void k_count_pairs(Uni un, Bufs b, uint32_t n,
                                                     unsigned long long *out4) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long cand = 0, inr = 0;
  if (s < n) {
    const DevUniforms *__restrict__ U = uni_at<B>(un, s);
    const float4 pi = b.spos[s];
    const int cx = cell_coord(pi.x, U->ox, U->inv_cell);
    const int cy = cell_coord(pi.y, U->oy, U->inv_cell);
    const int cz = cell_coord(pi.z, U->oz, U->inv_cell);
    const float h2 = U->h2;
    sweep27<1, false>(U, b.counts, b.starts, b.spos, cx, cy, cz, 0,
                      [&](uint32_t, float4 pj, bool valid) {
      const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
      cand += valid ? 1 : 0;
      if (valid && dot3(rx, ry, rz, rx, ry, rz) < h2) ++inr;
    });
  }
  unsigned long long mc = cand, mi = inr;
  for (int o = 32; o >= 1; o >>= 1) {
    cand += __shfl_xor(cand, o);
    inr += __shfl_xor(inr, o);
    mc = max(mc, (unsigned long long)__shfl_xor(mc, o));
    mi = max(mi, (unsigned long long)__shfl_xor(mi, o));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&out4[0], cand);
    atomicAdd(&out4[1], inr);
    atomicMax(&out4[2], mc);
    atomicMax(&out4[3], mi);
  }
}

inline uint32_t blocks_for(uint64_t threads, uint32_t per) {
  return (uint32_t)((threads + per - 1) / per);
}

}  // namespace

hipError_t launch_clear(const Bufs &b, uint32_t cells, uint32_t ntiles, hipStream_t st) {
  const uint32_t work = std::max(cells / 4 + 4, ntiles + 1);
  hipLaunchKernelGGL(k_clear, dim3(blocks_for(work, 256)), dim3(256), 0, st, b.counts, b.cwork,
                     cells, b.tile_status, ntiles, b.tile_counter, b.tiles.count);
  return hipGetLastError();
}

#define FLUID_LAUNCH_B(KERNEL, GRID, UN, ...)                                             \
  do {                                                                                    \
    if ((UN).scene_of)                                                                    \
      hipLaunchKernelGGL(KERNEL<true>, dim3(GRID), dim3(256), 0, st, UN, __VA_ARGS__);    \
    else                                                                                  \
      hipLaunchKernelGGL(KERNEL<false>, dim3(GRID), dim3(256), 0, st, UN, __VA_ARGS__);   \
  } while (0)

#define FLUID_CASE_L(KERNEL, LL, BB, GRID, UN, ...) \
  case LL: hipLaunchKernelGGL((KERNEL<LL, BB>), dim3(GRID), dim3(256), 0, st, UN, __VA_ARGS__); break;

#define FLUID_DISPATCH_L(KERNEL, LANES, N, UN, ...)                                        \
  do {                                                                                     \
    const uint32_t _grid = blocks_for((uint64_t)(N) * (LANES), 256);                       \
    if ((UN).scene_of) {                                                                   \
      switch (LANES) {                                                                     \
        FLUID_CASE_L(KERNEL, 1, true, _grid, UN, __VA_ARGS__)                              \
        FLUID_CASE_L(KERNEL, 2, true, _grid, UN, __VA_ARGS__)                              \
        FLUID_CASE_L(KERNEL, 4, true, _grid, UN, __VA_ARGS__)                              \
        default: hipLaunchKernelGGL((KERNEL<8, true>), dim3(_grid), dim3(256), 0, st, UN,  \
                                    __VA_ARGS__); break;                                   \
      }                                                                                    \
    } else {                                                                               \
      switch (LANES) {                                                                     \
        FLUID_CASE_L(KERNEL, 1, false, _grid, UN, __VA_ARGS__)                             \
        FLUID_CASE_L(KERNEL, 2, false, _grid, UN, __VA_ARGS__)                             \
        FLUID_CASE_L(KERNEL, 4, false, _grid, UN, __VA_ARGS__)                             \
        default: hipLaunchKernelGGL((KERNEL<8, false>), dim3(_grid), dim3(256), 0, st, UN, \
                                    __VA_ARGS__); break;                                   \
      }                                                                                    \
    }                                                                                      \
  } while (0)

hipError_t launch_build(const Uni &u, const Bufs &b, uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  FLUID_LAUNCH_B(k_build, blocks_for(n, 256), u, b, n);
  return hipGetLastError();
}

hipError_t launch_scan(const Uni &u, const Bufs &b, uint32_t cells, SweepCfg c,
                       hipStream_t st) {
  if (cells >= kScanBigGrid)
    hipLaunchKernelGGL(k_scan<4 * kScanItems>, dim3(blocks_for(cells, 4 * kScanTile)),
                       dim3(kScanThreads), 0, st, b, cells);
  else
    hipLaunchKernelGGL(k_scan<kScanItems>, dim3(blocks_for(cells, kScanTile)),
                       dim3(kScanThreads), 0, st, b, cells);
  if (c.tiled && b.tiles.total)
    FLUID_LAUNCH_B(k_tiles, blocks_for(b.tiles.total, 256), u, b);
  return hipGetLastError();
}

hipError_t launch_reorder(const Uni &, const Bufs &b, uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, st, b, n);
  hipLaunchKernelGGL(k_rank_gather, dim3(blocks_for(n, 256)), dim3(256), 0, st, b, n);
  return hipGetLastError();
}

// Persistent tiled launches: c.tile_wgs workgroups, LDS for the staged rows.
#define FLUID_DISPATCH_TILED(KERNEL, LANES, UN, ...)                                      \
  do {                                                                                    \
    const size_t _lds = (size_t)b.tiles.cap * sizeof(float4);                             \
    const dim3 _g((uint32_t)c.tile_wgs);                                                  \
    if ((UN).scene_of) {                                                                  \
      switch (LANES) {                                                                    \
        case 1: hipLaunchKernelGGL((KERNEL<1, true>), _g, dim3(256), _lds, st, UN, __VA_ARGS__); break; \
        case 2: hipLaunchKernelGGL((KERNEL<2, true>), _g, dim3(256), _lds, st, UN, __VA_ARGS__); break; \
        case 4: hipLaunchKernelGGL((KERNEL<4, true>), _g, dim3(256), _lds, st, UN, __VA_ARGS__); break; \
        default: hipLaunchKernelGGL((KERNEL<8, true>), _g, dim3(256), _lds, st, UN, __VA_ARGS__); break; \
      }                                                                                   \
    } else {                                                                              \
      switch (LANES) {                                                                    \
        case 1: hipLaunchKernelGGL((KERNEL<1, false>), _g, dim3(256), _lds, st, UN, __VA_ARGS__); break; \
        case 2: hipLaunchKernelGGL((KERNEL<2, false>), _g, dim3(256), _lds, st, UN, __VA_ARGS__); break; \
        case 4: hipLaunchKernelGGL((KERNEL<4, false>), _g, dim3(256), _lds, st, UN, __VA_ARGS__); break; \
        default: hipLaunchKernelGGL((KERNEL<8, false>), _g, dim3(256), _lds, st, UN, __VA_ARGS__); break; \
      }                                                                                   \
    }                                                                                     \
  } while (0)

// The list walkers (XSPH, force, both): 8, 16 or 32 lanes per particle over
// the 8-lane list.
#define FLUID_WALK_CASE(KERNEL, LL, N, UN, ...)                                               \
  do {                                                                                        \
    const uint32_t _grid = blocks_for((uint64_t)(N) * (LL), 256);                             \
    if ((UN).scene_of)                                                                        \
      hipLaunchKernelGGL((KERNEL<LL, true>), dim3(_grid), dim3(256), 0, st, UN, __VA_ARGS__);  \
    else                                                                                      \
      hipLaunchKernelGGL((KERNEL<LL, false>), dim3(_grid), dim3(256), 0, st, UN, __VA_ARGS__); \
  } while (0)
#define FLUID_DISPATCH_WALK(KERNEL, C, N, UN, ...)                                            \
  do {                                                                                        \
    if ((C).wide_walk == 1)                                                                   \
      FLUID_WALK_CASE(KERNEL, 16, N, UN, __VA_ARGS__);                                        \
    else if ((C).wide_walk == 2)                                                              \
      FLUID_WALK_CASE(KERNEL, 32, N, UN, __VA_ARGS__);                                        \
    else                                                                                      \
      FLUID_WALK_CASE(KERNEL, 8, N, UN, __VA_ARGS__);                                         \
  } while (0)

hipError_t launch_density(const Uni &u, const Bufs &b, uint32_t n, SweepCfg c, hipStream_t st) {
  if (n == 0) return hipSuccess;
  if (c.tiled)
    FLUID_DISPATCH_TILED(k_density_t, c.lanes, u, b, n, c.list_skip);
  else if (c.lanes == 16)
    FLUID_WALK_CASE(k_density, 16, n, u, b, n, (uint32_t)c.xcd_tiles, c.list_skip);
  else
    FLUID_DISPATCH_L(k_density, c.lanes, n, u, b, n, (uint32_t)c.xcd_tiles, c.list_skip);
  return hipGetLastError();
}

hipError_t launch_xsph(const Uni &u, const Bufs &b, uint32_t n, SweepCfg c, hipStream_t st) {
  if (n == 0) return hipSuccess;
  FLUID_DISPATCH_WALK(k_xsph, c, n, u, b, n, (uint32_t)c.xcd_tiles);
  return hipGetLastError();
}

hipError_t launch_force(const Uni &u, const Bufs &b, uint32_t n, SweepCfg c, bool fuse_predict,
                        hipStream_t st) {
  if (n == 0) return hipSuccess;
  FLUID_DISPATCH_WALK(k_force, c, n, u, b, n, (uint32_t)c.xcd_tiles,
                      (uint32_t)(fuse_predict ? 1 : 0));
  return hipGetLastError();
}

hipError_t launch_xsph_force(const Uni &u, const Bufs &b, uint32_t n, SweepCfg c,
                             bool fuse_predict, hipStream_t st) {
  if (n == 0) return hipSuccess;
  FLUID_DISPATCH_WALK(k_xsph_force, c, n, u, b, n, (uint32_t)c.xcd_tiles,
                      (uint32_t)(fuse_predict ? 1 : 0));
  return hipGetLastError();
}

hipError_t launch_predict(const Uni &u, const Bufs &b, uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  FLUID_LAUNCH_B(k_predict, blocks_for(n, 256), u, b, n);
  return hipGetLastError();
}

hipError_t launch_solve(const Uni &u, const Bufs &b, uint32_t n, const PredBuf &in,
                        const PredBuf &out, float4 *ap_out, SweepCfg c, bool first_iter,
                        hipStream_t st) {
  if (n == 0) return hipSuccess;
  // per-lane skin-list counts are sized for kMaxLanes lanes per particle
  if (c.solve_lanes > (int)kMaxLanes || c.lanes > (int)kMaxLanes) return hipErrorInvalidValue;
  if (c.tiled)
    FLUID_DISPATCH_TILED(k_solve_t, c.lanes, u, b, n, in, out, ap_out,
                         (uint32_t)(first_iter ? 1 : 0), (uint32_t)c.solve_list);
  else if (first_iter && c.solve_list && c.solve_first && c.solve_lanes <= 8) {
    // first iteration: 2 or 4 sweep lanes per skin-list lane
    const int M = 1 << c.solve_first, LS = c.solve_lanes;
    const uint32_t grid = blocks_for((uint64_t)n * (uint64_t)(LS * M), 256);
#define FLUID_SOLVE_FW(LL, LSS)                                                                 \
  if (u.scene_of)                                                                               \
    hipLaunchKernelGGL((k_solve_fw<LL, LSS, true>), dim3(grid), dim3(256), 0, st, u, b, n, in,  \
                       out, ap_out, (uint32_t)c.xcd_tiles);                                      \
  else                                                                                          \
    hipLaunchKernelGGL((k_solve_fw<LL, LSS, false>), dim3(grid), dim3(256), 0, st, u, b, n,     \
                       in, out, ap_out, (uint32_t)c.xcd_tiles);
    if (LS == 1 && M == 2) { FLUID_SOLVE_FW(2, 1) }
    else if (LS == 1) { FLUID_SOLVE_FW(4, 1) }
    else if (LS == 2 && M == 2) { FLUID_SOLVE_FW(4, 2) }
    else if (LS == 2) { FLUID_SOLVE_FW(8, 2) }
    else if (LS == 4 && M == 2) { FLUID_SOLVE_FW(8, 4) }
    else if (LS == 4) { FLUID_SOLVE_FW(16, 4) }
    else if (M == 2) { FLUID_SOLVE_FW(16, 8) }
    else { FLUID_SOLVE_FW(32, 8) }
#undef FLUID_SOLVE_FW
  } else if (!first_iter && c.solve_list && c.solve_walk && c.solve_lanes >= 2 &&
             c.solve_lanes <= 8) {
    // later iterations: 2 or 4 walkers per skin-list lane
    const int M = 1 << c.solve_walk, LS = c.solve_lanes;
    const uint32_t grid = blocks_for((uint64_t)n * (uint64_t)(LS * M), 256);
#define FLUID_SOLVE_W(LL, LSS)                                                                  \
  if (u.scene_of)                                                                               \
    hipLaunchKernelGGL((k_solve_w<LL, LSS, true>), dim3(grid), dim3(256), 0, st, u, b, n, in,   \
                       out, ap_out, (uint32_t)c.xcd_tiles);                                      \
  else                                                                                          \
    hipLaunchKernelGGL((k_solve_w<LL, LSS, false>), dim3(grid), dim3(256), 0, st, u, b, n, in,  \
                       out, ap_out, (uint32_t)c.xcd_tiles);
    if (LS == 2 && M == 2) { FLUID_SOLVE_W(4, 2) }
    else if (LS == 2) { FLUID_SOLVE_W(8, 2) }
    else if (LS == 4 && M == 2) { FLUID_SOLVE_W(8, 4) }
    else if (LS == 4) { FLUID_SOLVE_W(16, 4) }
    else if (M == 2) { FLUID_SOLVE_W(16, 8) }
    else { FLUID_SOLVE_W(32, 8) }
#undef FLUID_SOLVE_W
  } else if (c.solve_lanes == 16)
    FLUID_WALK_CASE(k_solve, 16, n, u, b, n, in, out, ap_out, (uint32_t)c.xcd_tiles,
                    (uint32_t)(first_iter ? 1 : 0), (uint32_t)c.solve_list);
  else
    FLUID_DISPATCH_L(k_solve, c.solve_lanes, n, u, b, n, in, out, ap_out,
                     (uint32_t)c.xcd_tiles, (uint32_t)(first_iter ? 1 : 0),
                     (uint32_t)c.solve_list);
  return hipGetLastError();
}

hipError_t launch_finalize(const Uni &, const Bufs &b, uint32_t n, const PredBuf &fin,
                           hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize, dim3(blocks_for(n, 256)), dim3(256), 0, st, b, n, fin);
  return hipGetLastError();
}

hipError_t launch_integrate(const Uni &u, const Bufs &b, uint32_t n, const float4 *fin_v,
                            uint32_t ntiles, hipStream_t st) {
  FLUID_LAUNCH_B(k_integrate, blocks_for(std::max(n, ntiles + 1), 256), u, b, n, fin_v, ntiles);
  return hipGetLastError();
}

hipError_t launch_count_pairs(const Uni &u, const Bufs &b, uint32_t n,
                              unsigned long long *out4, hipStream_t st) {
  if (n == 0) return hipSuccess;
  FLUID_LAUNCH_B(k_count_pairs, blocks_for(n, 256), u, b, n, out4);
  return hipGetLastError();
}

}  // namespace fluid
