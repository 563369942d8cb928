// spray.hip — whitewater (spray, foam, bubble) particles over the sorted SPH
// state of the last step: the reference's SpraySystem
// (src/simulation/spray.rs:349-428) with its two compute passes,
// spray_emit.wgsl and spray_simulate.wgsl.
//
// Emit is three kernels: the per-particle potentials and spawn counts (a walk
// of the density pass's in-range list, which holds exactly the pairs the
// WGSL loop keeps, spray_emit.wgsl:226-283), a decoupled look-back scan of
// the counts, and the spawn writes at head + exclusive prefix.  The
// reference hands out slots with one atomicAdd per spawn on a single
// counter (:335), so its slot order is a race; the scan gives the order of
// one sequential run (particle i's spawns after those of every j < i) and
// removes the contention.  Simulate is one kernel over the ring (:147-281).
//
// This file is compiled without FMA contraction and with correctly rounded
// division and square root, so every per-particle value that a gate, a
// count or a spawn depends on (speed, normals, remaps, rates, the wall-cap
// extrapolation, the integration) is the f32 oracle's to the bit, given the
// same neighbour sums.  The sums themselves are split over lanes, so their
// terms may round differently at no cost to that: the pair terms use the
// hardware reciprocal square root and reciprocal (a few ulp) and
// algebraically equal forms, e.g. |u| (1 - u_hat . x_hat) = |u| - u . x_hat.
// Neighbour counts and the in-range tests stay exact.
#include "sph_device.h"

namespace fluid {
namespace {

// spray_emit.wgsl:114-154
constexpr float kTwoPi = 6.28318530718f;
constexpr float kAutoLimitMinFrac = 0.1f;
constexpr uint32_t kEmitMinNeighbours = 10u;
constexpr float kKineticMaxSpeed = 2.5f;
constexpr float kCrestDensityFrac = 0.9f;
constexpr float kCrestVelNormalMin = 0.6f;
constexpr float kCrestDivergenceMin = 0.0f;
constexpr float kCrestDivergenceMax = 0.5f;
constexpr uint32_t kMaxEmitPerFrame = 16u;
// spray_simulate.wgsl:118-131
constexpr uint32_t kKindSpray = 0u, kKindFoam = 1u, kKindBubble = 2u;
constexpr uint32_t kSprayMaxNeighbours = 6u;
constexpr uint32_t kBubbleMinNeighbours = 20u;
constexpr float kNewbornSprayTime = 0.15f;
// spray.rs:26-33
constexpr float kAutoLimitRise = 0.05f;
constexpr float kAutoLimitDecay = 0.01f;
constexpr float kAutoTaFloor = 0.2f;
constexpr float kAutoWcFloor = 0.05f;

// Walkers per emitter over its 8-lane in-range list (two per list lane),
// and lanes per diffuse particle over its 27 cells: a diffuse particle deep
// in the fluid meets hundreds of candidates, and its lanes' dependent
// gathers are the pass's critical path.
constexpr int kEmitLanes = 2 * kListLanes;
constexpr int kSimLanes = 32;  // 8 / 16 / 32 / 64 lanes: 97 / 72 / 67 / 66 us per 50k frame

inline uint32_t blocks_for(uint64_t threads, uint32_t per) {
  return (uint32_t)((threads + per - 1) / per);
}

// spray_emit.wgsl:156-164
__device__ __forceinline__ uint32_t spray_hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x45d9f3bu;
  x ^= x >> 16;
  x *= 0x45d9f3bu;
  x ^= x >> 16;
  return x;
}

// :166-168
__device__ __forceinline__ float hash_float(uint32_t seed) {
  return (float)(spray_hash(seed) & 0xFFFFu) / 65535.0f;
}

// clamp(e, lo, hi) = min(max(e, lo), hi)
__device__ __forceinline__ float clampf(float x, float lo, float hi) {
  return fminf(fmaxf(x, lo), hi);
}

// :185-187
__device__ __forceinline__ float remap_clamp(float x, float lo, float hi) {
  return clampf((x - lo) / (hi - lo), 0.0f, 1.0f);
}

// WGSL u32(f32): saturating conversion
__device__ __forceinline__ uint32_t to_u32(float x) {
  if (!(x > 0.0f)) return 0u;
  if (x >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

// spray.rs:40-51
__device__ __forceinline__ float update_auto_limit(float current, float frame_max,
                                                   float floor_v) {
  if (frame_max <= 1e-6f) return current;
  const float alpha = frame_max > current ? kAutoLimitRise : kAutoLimitDecay;
  return fmaxf(current + alpha * (frame_max - current), floor_v);
}

// world_to_local (container_common.wgsl:28-35)
__device__ __forceinline__ float3 world_to_local(const DevUniforms *__restrict__ U, float x,
                                                 float y, float z) {
  const float py = y - U->cy;
  return make_float3(dot3(U->iv[0][0], U->iv[0][1], U->iv[0][2], x, py, z),
                     dot3(U->iv[1][0], U->iv[1][1], U->iv[1][2], x, py, z),
                     dot3(U->iv[2][0], U->iv[2][1], U->iv[2][2], x, py, z));
}

// spray_simulate.wgsl:142-145
__device__ __forceinline__ float cap_volume_fraction(float t) {
  const float tc = clampf(t, 0.0f, 1.0f);
  return tc * tc * (3.0f - tc) * 0.25f;
}

__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

// The 27-cell stencil as nine x-rows in sweep27's order (z, y, then slot),
// with all nine rows' slot ranges read up front: a diffuse particle deep in
// the fluid walks hundreds of candidates, and with few live slots in a
// launch the row-by-row bound reads would sit on its critical path.  Lane
// `sub` of L takes every L-th candidate of each row, four in flight.
template <int L, class F>
__device__ __forceinline__ void sweep27_rows_first(const DevUniforms *__restrict__ U,
                                                   const uint32_t *__restrict__ counts,
                                                   const uint32_t *__restrict__ starts,
                                                   const float4 *__restrict__ pts, int cx,
                                                   int cy, int cz, uint32_t sub, F &&f) {
  const int gx = (int)U->gx, gy = (int)U->gy, gz = (int)U->gz;
  cx = min(max(cx, -2), gx + 1);
  cy = min(max(cy, -2), gy + 1);
  cz = min(max(cz, -2), gz + 1);
  const int x0 = max(cx - 1, 0), x1 = min(cx + 1, gx - 1);
  if (x0 > x1) return;
  uint32_t beg[9], end[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int z = cz + k / 3 - 1, y = cy + k % 3 - 1;
    beg[k] = end[k] = 0u;
    if (z >= 0 && z < gz && y >= 0 && y < gy) {
      const uint32_t row = U->cell_offset + (uint32_t)y * U->gx + (uint32_t)z * U->gxy;
      const uint32_t c0 = row + (uint32_t)x0, c1 = row + (uint32_t)x1;
      beg[k] = starts[c0] - counts[c0];
      end[k] = starts[c1];
    }
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    uint32_t j = beg[k] + sub;
    const uint32_t e = end[k];
    for (; j + 3 * L < e; j += 4 * L) {
      const float4 p0 = pts[j], p1 = pts[j + L], p2 = pts[j + 2 * L], p3 = pts[j + 3 * L];
      f(j, p0);
      f(j + L, p1);
      f(j + 2 * L, p2);
      f(j + 3 * L, p3);
    }
    for (; j < e; j += L) f(j, pts[j]);
  }
}

template <int L>
__device__ __forceinline__ uint32_t group_sum_u(uint32_t v) {
#pragma unroll
  for (int o = L / 2; o >= 1; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, L);
  return v;
}

// ------------------------------------------------------------------------
// Emit, part 1 (spray_emit.wgsl:190-320): per sorted particle, the trapped-
// air, wave-crest and divergence sums over its in-range neighbours, the
// gates, and the number of particles to spawn.  kEmitLanes lanes per
// particle; no lane returns early, so the block-wide maxima and the group
// sums see every lane.  Also clears the slot scan's tile flags and, in
// thread 0, commits the previous pass's spawns to the head.
__global__ __launch_bounds__(256) void k_spray_potentials(Uni un, Bufs b, SprayBufs sb,
                                                          uint32_t n, GpuSprayParams p,
                                                          uint32_t nclear) {
  const uint32_t g = blockIdx.x * 256u + threadIdx.x;
  SprayDev *__restrict__ d = sb.d;
  if (g < nclear) sb.tile_status[g] = 0ull;
  if (g == 0) {
    *sb.tile_counter = 0u;
    d->head += d->emitted;
    d->emitted = 0u;
  }
  const DevUniforms *__restrict__ U = un.u;
  const uint32_t s = g / kEmitLanes, sub = g % kEmitLanes;
  const bool in = s < n;
  float4 r0 = make_float4(0.f, 0.f, 0.f, 0.f), r1 = r0, r2 = r0;
  if (in) {
    const float4 *rec = b.srec + 3 * (size_t)s;
    r0 = rec[0];  // position, density
    r1 = rec[1];  // velocity, near density
    r2 = rec[2];  // normal
  }
  const float vx = r1.x, vy = r1.y, vz = r1.z;
  const float speed = sqrtf(dot3(vx, vy, vz, vx, vy, vz));
  // kinetic gate (:199-201); a NaN speed passes it, as in WGSL
  const bool active = in && !(speed < p.min_speed);
  const float h = U->h, h2 = U->h2, mass = U->mass, norm = U->poly6_scale;
  const float nlen_i = sqrtf(dot3(r2.x, r2.y, r2.z, r2.x, r2.y, r2.z));
  const bool crest_ok = nlen_i > 1e-4f && r0.w < U->rho0 * kCrestDensityFrac;  // :211
  float nhx = 0.f, nhy = 1.f, nhz = 0.f;
  if (crest_ok) {
    nhx = r2.x / nlen_i;
    nhy = r2.y / nlen_i;
    nhz = r2.z / nlen_i;
  }
  float ta = 0.f, crest = 0.f, dv = 0.f;
  uint32_t nn = 0u;
  if (active) {
    const float inv_h = 1.0f / h, mnorm = mass * norm;
    // the list holds exactly the pairs the WGSL loop keeps, 1e-12 < r^2 < h^2
    for_each_neighbour<kEmitLanes, 2, 2>(U, b, n, s, sub, r0,
                                         [&](uint32_t j, const NbrRec<2> &q) {
      const float4 pj = q.r[0], vj = q.r[1];
      const float rx = r0.x - pj.x, ry = r0.y - pj.y, rz = r0.z - pj.z;
      const float rsq = dot3(rx, ry, rz, rx, ry, rz);
      const float il = __builtin_amdgcn_rsqf(rsq);
      const float w = 1.0f - rsq * il * inv_h;  // 1 - r / h
      const float xx = rx * il, xy = ry * il, xz = rz * il;
      nn += 1u;
      const float d2 = h2 - rsq;
      const float wvol = mnorm * __builtin_amdgcn_rcpf(fmaxf(pj.w, 1e-6f)) * (d2 * d2 * d2);
      const float ux = vx - vj.x, uy = vy - vj.y, uz = vz - vj.z;
      const float uu = dot3(ux, uy, uz, ux, uy, uz), ux_hat = dot3(ux, uy, uz, xx, xy, xz);
      // |u| > 1e-6: |u| (1 - u_hat . x_hat) = |u| - u . x_hat
      if (uu > 1e-12f) ta += (uu * __builtin_amdgcn_rsqf(uu) - ux_hat) * wvol;
      dv += ux_hat * w;
      // -x_hat . n_hat < 0 exactly when r_vec . n_hat > 0
      if (crest_ok && dot3(rx, ry, rz, nhx, nhy, nhz) > 0.0f) {
        const float4 nj = b.srec[3 * (size_t)j + 2];
        const float nsq = dot3(nj.x, nj.y, nj.z, nj.x, nj.y, nj.z);
        if (nsq > 1e-8f)  // |n_j| > 1e-4
          crest += (1.0f - dot3(nhx, nhy, nhz, nj.x, nj.y, nj.z) * __builtin_amdgcn_rsqf(nsq)) *
                   wvol;
      }
    });
  }
  ta = group_sum<kEmitLanes>(ta);
  crest = group_sum<kEmitLanes>(crest);
  dv = group_sum<kEmitLanes>(dv);
  nn = group_sum_u<kEmitLanes>(nn);
  uint32_t ne = 0u;
  float ta_max = 0.f, wc_max = 0.f;
  if (active && nn >= kEmitMinNeighbours) {  // embeddedness gate (:285-287)
    if (dot3(vx / speed, vy / speed, vz / speed, nhx, nhy, nhz) < kCrestVelNormalMin)
      crest = 0.0f;  // :290-292
    if (ta > 0.0f) ta_max = ta;  // :296-301
    if (crest > 0.0f) wc_max = crest;
    const float phi_div = remap_clamp(dv, kCrestDivergenceMin, kCrestDivergenceMax);
    const float ta_hi = fmaxf(d->ta_limit, 1e-4f);
    const float wc_hi = fmaxf(d->wc_limit, 1e-4f);
    const float phi_ta = remap_clamp(ta, kAutoLimitMinFrac * ta_hi, ta_hi);
    const float phi_wc = remap_clamp(crest, kAutoLimitMinFrac * wc_hi, wc_hi) * phi_div;
    const float phi_k = remap_clamp(speed, p.min_speed, kKineticMaxSpeed);
    const float rate =
        phi_k * (p.k_trapped_air * phi_ta + p.k_wave_crest * phi_wc) * p.emission_rate;
    const float expected = rate * p.dt;
    const uint32_t base_seed = s + p.frame_count * 7919u;
    const float fl = floorf(expected);
    ne = to_u32(fl);
    if (hash_float(base_seed * 13u + 5u) < expected - fl) ne += 1u;
    ne = min(ne, kMaxEmitPerFrame);
  }
  // Block maxima, then one atomicMax per block and value, skipped when the
  // frame's maximum is already as large: atomics on one address serialise,
  // and the maxima settle within a few blocks.
  __shared__ float s_max[2][4];
  ta_max = wave_max(ta_max);
  wc_max = wave_max(wc_max);
  if ((threadIdx.x & 63u) == 0u) {
    s_max[0][threadIdx.x >> 6] = ta_max;
    s_max[1][threadIdx.x >> 6] = wc_max;
  }
  __syncthreads();
  if (threadIdx.x < 2u) {
    const float *m4 = s_max[threadIdx.x];
    const float m = fmaxf(fmaxf(m4[0], m4[1]), fmaxf(m4[2], m4[3]));
    uint32_t *slot = &d->frame_max[threadIdx.x];
    // non-negative floats order as their bit patterns
    if (m > 0.f && __float_as_uint(m) > __hip_atomic_load(slot, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT))
      atomicMax(slot, __float_as_uint(m));
  }
  if (in && sub == 0) {
    sb.emit_n[s] = ne;
    sb.pot[s] = make_float4(ta, crest, dv, (float)nn);
  }
}

// Emit, part 3 (spray_emit.wgsl:322-372): particle s writes its spawns to
// slots head + (spawns of particles < s) + e.  When the pass spawns more
// than the ring holds, only the last `cap` spawns are written (a later
// spawn would overwrite the same slot).
__global__ __launch_bounds__(256) void k_spray_spawn(Uni un, Bufs b, SprayBufs sb, uint32_t n,
                                                     GpuSprayParams p) {
  const uint32_t s = blockIdx.x * 256u + threadIdx.x;
  if (s >= n) return;
  const uint32_t total = sb.emit_incl[n - 1];
  if (s == n - 1) sb.d->emitted = total;
  const uint32_t ne = sb.emit_n[s];
  if (ne == 0u) return;
  const DevUniforms *__restrict__ U = un.u;
  const float4 *rec = b.srec + 3 * (size_t)s;
  const float4 r0 = rec[0], r1 = rec[1];
  const float vx = r1.x, vy = r1.y, vz = r1.z;
  const float speed = sqrtf(dot3(vx, vy, vz, vx, vy, vz));
  const float phi_k = remap_clamp(speed, p.min_speed, kKineticMaxSpeed);
  // cylinder frame along the velocity (:326-333)
  const float dx = vx / speed, dy = vy / speed, dz = vz / speed;
  float ax = 0.f, ay = 1.f, az = 0.f;
  if (fabsf(dy) > 0.99f) {
    ax = 1.f;
    ay = 0.f;
  }
  const float cx = dy * az - dz * ay, cy = dz * ax - dx * az, cz = dx * ay - dy * ax;
  const float cl = sqrtf(dot3(cx, cy, cz, cx, cy, cz));
  const float e1x = cx / cl, e1y = cy / cl, e1z = cz / cl;
  const float e2x = dy * e1z - dz * e1y, e2y = dz * e1x - dx * e1z, e2z = dx * e1y - dy * e1x;
  const float h = U->h;
  const float lt_min = p.lifetime * (1.0f - p.lifetime_variation);
  const float lt_span = 2.0f * p.lifetime * p.lifetime_variation;
  const uint32_t head = sb.d->head, cap = sb.cap;
  const uint32_t first = sb.emit_incl[s] - ne;
  const uint32_t base_seed = s + p.frame_count * 7919u;
  for (uint32_t e = 0; e < ne; ++e) {
    const uint32_t k = first + e;
    if (total > cap && k < total - cap) continue;
    const uint32_t slot = (head + k) % cap;
    const uint32_t seed = base_seed + e * 3571u;
    const float q0 = hash_float(seed), q1 = hash_float(seed + 1u);
    const float q2 = hash_float(seed + 2u), q3 = hash_float(seed + 3u);
    const float q4 = hash_float(seed + 4u) * 2.0f - 1.0f;
    const float q5 = hash_float(seed + 5u) * 2.0f - 1.0f;
    const float q6 = hash_float(seed + 6u) * 2.0f - 1.0f;
    const float disc_r = h * 0.5f * sqrtf(q0);
    const float theta = kTwoPi * q1;
    const float along = q2 * p.dt * speed;
    const float c = disc_r * cosf(theta), sn = disc_r * sinf(theta);
    const float px = r0.x + e1x * c + e2x * sn + dx * along;
    const float py = r0.y + e1y * c + e2y * sn + dy * along;
    const float pz = r0.z + e1z * c + e2z * sn + dz * along;
    const float wx = vx * p.speed_multiplier + q4 * p.velocity_jitter;
    const float wy = vy * p.speed_multiplier + q5 * p.velocity_jitter;
    const float wz = vz * p.speed_multiplier + q6 * p.velocity_jitter;
    const float lt = lt_min + phi_k * q3 * lt_span;
    float4 *out = sb.ring + 3 * (size_t)slot;
    out[0] = make_float4(px, py, pz, lt);
    out[1] = make_float4(wx, wy, wz, lt);
    out[2] = make_float4(__uint_as_float(kKindSpray), 0.f, 0.f, 0.f);
  }
}

// Simulate (spray_simulate.wgsl:147-281): every live ring slot counts its
// fluid neighbours over the 27 cells (kSimLanes lanes per slot) with the
// radially weighted mean velocity, is reclassified and advanced.  Thread 0
// folds the maxima of the emit passes since the last simulate into the
// potential ceilings (spray.rs:40-51) and clears them.  Each block leaves
// its live counts in live_part.
__global__ __launch_bounds__(256) void k_spray_simulate(Uni un, Bufs b, SprayBufs sb,
                                                        GpuSprayParams p) {
  const uint32_t g = blockIdx.x * 256u + threadIdx.x;
  SprayDev *__restrict__ d = sb.d;
  if (g == 0) {
    d->ta_limit = update_auto_limit(d->ta_limit, __uint_as_float(d->frame_max[0]), kAutoTaFloor);
    d->wc_limit = update_auto_limit(d->wc_limit, __uint_as_float(d->frame_max[1]), kAutoWcFloor);
    d->frame_max[0] = 0u;
    d->frame_max[1] = 0u;
  }
  const DevUniforms *__restrict__ U = un.u;
  const uint32_t idx = g / kSimLanes, sub = g % kSimLanes;
  const bool in = idx < sb.cap;
  float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0, q2 = q0;
  if (in) {
    const float4 *rec = sb.ring + 3 * (size_t)idx;
    q0 = rec[0];  // position, lifetime
    q1 = rec[1];  // velocity, max lifetime
    q2 = rec[2];  // kind, age, pad
  }
  const bool alive = in && !(q0.w <= 0.0f);  // :154-156
  // a block without a live slot only reports zero counts
  __shared__ uint32_t s_live[4][4];
  if (!__syncthreads_or(alive)) {
    if (threadIdx.x < 4u) sb.live_part[4 * (size_t)blockIdx.x + threadIdx.x] = 0u;
    return;
  }
  const float h = U->h, h2 = U->h2;
  uint32_t cnt = 0u;
  float ws = 0.f, ax = 0.f, ay = 0.f, az = 0.f;
  if (alive) {
    const int cx = cell_coord(q0.x, U->ox, U->inv_cell);
    const int cy = cell_coord(q0.y, U->oy, U->inv_cell);
    const int cz = cell_coord(q0.z, U->oz, U->inv_cell);
    const float inv_h = 1.0f / h;
    sweep27_rows_first<kSimLanes>(U, b.counts, b.starts, b.spos, cx, cy, cz, sub,
                                  [&](uint32_t j, float4 pj) {
      // loaded whether or not the pair is in range, so it need not wait for
      // the test (the gathers of a trip go out together)
      const float4 vj = b.svel[j];
      const float rx = q0.x - pj.x, ry = q0.y - pj.y, rz = q0.z - pj.z;
      const float rsq = dot3(rx, ry, rz, rx, ry, rz);
      if (rsq < h2) {
        const float w = 1.0f - fast_sqrt(rsq) * inv_h;
        cnt += 1u;
        ws += w;
        ax += vj.x * w;
        ay += vj.y * w;
        az += vj.z * w;
      }
    });
  }
  cnt = group_sum_u<kSimLanes>(cnt);
  ws = group_sum<kSimLanes>(ws);
  ax = group_sum<kSimLanes>(ax);
  ay = group_sum<kSimLanes>(ay);
  az = group_sum<kSimLanes>(az);
  uint32_t kind = kKindSpray;
  float life = 0.0f;
  // the slot's update (:197-272), for live slots only
  if (alive && sub == 0) {
    if (ws > 1e-6f) {  // :197-199
      ax = ax / ws;
      ay = ay / ws;
      az = az / ws;
    }
    kind = kKindFoam;  // :202-207
    if (cnt <= kSprayMaxNeighbours) kind = kKindSpray;
    else if (cnt >= kBubbleMinNeighbours) kind = kKindBubble;
    if (kind == kKindFoam) {
      // wall-truncated neighbourhoods (:213-233)
      const float3 l = world_to_local(U, q0.x, q0.y, q0.z);
      const float lo[3] = {U->hw + l.x, U->hh + l.y, U->hd + l.z};
      const float hi[3] = {U->hw - l.x, U->hh - l.y, U->hd - l.z};
      float missing = 0.0f;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (lo[a] < h) missing += cap_volume_fraction(1.0f - lo[a] / h);
        if (hi[a] < h) missing += cap_volume_fraction(1.0f - hi[a] / h);
      }
      const float visible = fmaxf(1.0f - missing, 0.25f);
      if (to_u32(rintf((float)cnt / visible)) >= kBubbleMinNeighbours) kind = kKindBubble;
    }
    const float age = q2.y;
    if (kind == kKindFoam && age < kNewbornSprayTime) kind = kKindSpray;  // :235-237
    const float dt = p.dt, gy = p.gravity_y;
    float vx = q1.x, vy = q1.y, vz = q1.z, decay = 0.0f;
    if (kind == kKindSpray) {  // :244-256
      vx = vx + dt * (0.0f - p.drag * vx);
      vy = vy + dt * (gy - p.drag * vy);
      vz = vz + dt * (0.0f - p.drag * vz);
    } else if (kind == kKindFoam) {
      vx = ax;
      vy = ay;
      vz = az;
      decay = 1.0f;
    } else {
      const float nb = -p.bubble_buoyancy;
      vx = vx + dt * (nb * 0.0f);
      vy = vy + dt * (nb * gy);
      vz = vz + dt * (nb * 0.0f);
      const float t = clampf(p.bubble_drag, 0.0f, 1.0f);
      vx = vx * (1.0f - t) + ax * t;
      vy = vy * (1.0f - t) + ay * t;
      vz = vz * (1.0f - t) + az * t;
    }
    const float nx = q0.x + dt * vx, ny = q0.y + dt * vy, nz = q0.z + dt * vz;  // :257
    life = q0.w - dt * decay;
    const float3 ln = world_to_local(U, nx, ny, nz);  // :268-272
    if (!(ln.x >= -U->hw && ln.x <= U->hw && ln.y >= -U->hh && ln.y <= U->hh &&
          ln.z >= -U->hd && ln.z <= U->hd))
      life = 0.0f;
    float4 *out = sb.ring + 3 * (size_t)idx;
    out[0] = make_float4(nx, ny, nz, life);
    out[1] = make_float4(vx, vy, vz, q1.w);
    out[2] = make_float4(__uint_as_float(kind), age + dt, q2.z, q2.w);
  }
  // live counts by kind (:275-278): per block, summed on demand by
  // k_spray_live_sum (one partial per block, no atomics on shared addresses)
  const bool live = alive && sub == 0 && life > 0.0f;
  const uint32_t c_all = (uint32_t)__popcll(__ballot(live));
  const uint32_t c_spray = (uint32_t)__popcll(__ballot(live && kind == kKindSpray));
  const uint32_t c_foam = (uint32_t)__popcll(__ballot(live && kind == kKindFoam));
  const uint32_t c_bubble = (uint32_t)__popcll(__ballot(live && kind == kKindBubble));
  if ((threadIdx.x & 63u) == 0u) {
    const uint32_t w = threadIdx.x >> 6;
    s_live[0][w] = c_all;
    s_live[1][w] = c_spray;
    s_live[2][w] = c_foam;
    s_live[3][w] = c_bubble;
  }
  __syncthreads();
  if (threadIdx.x < 4u) {
    const uint32_t *c4 = s_live[threadIdx.x];
    sb.live_part[4 * (size_t)blockIdx.x + threadIdx.x] = c4[0] + c4[1] + c4[2] + c4[3];
  }
}

// Sum of the simulate pass's per-block live counts into the state block
// (one workgroup; launched before a state readback).
__global__ __launch_bounds__(256) void k_spray_live_sum(SprayBufs sb, uint32_t nblocks) {
  __shared__ uint32_t s_acc[4][256];
  uint32_t acc[4] = {0u, 0u, 0u, 0u};
  for (uint32_t i = threadIdx.x; i < nblocks; i += 256u)
    for (int k = 0; k < 4; ++k) acc[k] += sb.live_part[4 * (size_t)i + k];
  for (int k = 0; k < 4; ++k) s_acc[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (uint32_t o = 128u; o >= 1u; o >>= 1) {
    if (threadIdx.x < o)
      for (int k = 0; k < 4; ++k) s_acc[k][threadIdx.x] += s_acc[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 4u) sb.d->live[threadIdx.x] = s_acc[threadIdx.x][0];
}

}  // namespace

hipError_t launch_spray_emit(const Uni &u, const Bufs &b, const SprayBufs &sb, uint32_t n,
                             const GpuSprayParams &p, hipStream_t st) {
  // one block at least: it commits the previous spawns
  const uint32_t nclear = blocks_for(n, kScanTile);
  const uint32_t grid = n ? blocks_for((uint64_t)n * kEmitLanes, 256) : 1u;
  hipLaunchKernelGGL(k_spray_potentials, dim3(grid), dim3(256), 0, st, u, b, sb, n, p, nclear);
  if (n == 0) return hipGetLastError();
  Bufs sc = b;
  sc.cwork = sb.emit_n;  // scanned in place: no published copy, not zeroed
  sc.counts = nullptr;
  sc.starts = sb.emit_incl;
  sc.tile_status = sb.tile_status;
  sc.tile_counter = sb.tile_counter;
  SweepCfg c{};
  c.tiled = 0;
  hipError_t e = launch_scan(u, sc, n, c, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_spray_spawn, dim3(blocks_for(n, 256)), dim3(256), 0, st, u, b, sb, n, p);
  return hipGetLastError();
}

uint32_t spray_simulate_blocks(uint32_t cap) {
  return blocks_for((uint64_t)cap * kSimLanes, 256);
}

hipError_t launch_spray_simulate(const Uni &u, const Bufs &b, const SprayBufs &sb,
                                 const GpuSprayParams &p, hipStream_t st) {
  hipLaunchKernelGGL(k_spray_simulate, dim3(spray_simulate_blocks(sb.cap)), dim3(256), 0, st, u,
                     b, sb, p);
  return hipGetLastError();
}

hipError_t launch_spray_live_sum(const SprayBufs &sb, hipStream_t st) {
  hipLaunchKernelGGL(k_spray_live_sum, dim3(1), dim3(256), 0, st, sb,
                     spray_simulate_blocks(sb.cap));
  return hipGetLastError();
}

}  // namespace fluid
