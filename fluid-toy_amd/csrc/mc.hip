// mc.hip — the marching-cubes density field over the sorted SPH state: the
// anisotropic kernel fit (mc_anisotropy.wgsl, Yu & Turk) and the density
// splat onto the MC voxel grid (mc_density.wgsl), the compute passes 0 and 1
// of MarchingCubesRenderer::generate (src/render/marching_cubes.rs:2402-2470).
//
// The density pass runs one 64-voxel 4x4x4 block per wave (the reference's
// workgroup shape) and walks the union of its voxels' SPH-cell stencils
// with a wave-uniform candidate index: every lane reads the same candidate,
// so each record is one broadcast load per wave instead of one per voxel.
// Each lane keeps, per stencil row, the slot range of its own voxel's
// stencil and tests only those candidates, so every voxel sums exactly the
// reference's candidates in the reference's order (z, y, x, then slot).
//
// Compiled without FMA contraction and with correctly rounded division and
// square root, so that the density field equals the f32 oracle's to the bit
// for the same particles and ellipsoid records; the anisotropy fit's
// neighbour sums are split over lanes and agree to rounding.
#include "sph_device.h"

namespace fluid {
namespace {

// mc_anisotropy.wgsl:55-60
constexpr float kNLo = 4.0f;
constexpr float kNHi = 12.0f;
constexpr float kMinStretch = 0.25f;
// Lanes per particle of the anisotropy fit (its 2h neighbourhood holds
// several hundred candidates).
constexpr int kAnisoLanes = 16;

inline uint32_t blocks_for(uint64_t threads, uint32_t per) {
  return (uint32_t)((threads + per - 1) / per);
}

__device__ __forceinline__ float clampf(float x, float lo, float hi) {
  return fminf(fmaxf(x, lo), hi);
}

// WGSL smoothstep(e0, e1, x)
__device__ __forceinline__ float smoothstep(float e0, float e1, float x) {
  const float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

// mat3x3 as columns: m[c][r]
struct M3 {
  float m[3][3];
};

// jacobi_rotate (mc_anisotropy.wgsl:79-107): one rotation in the (p, q)
// plane zeroing A[p][q]; k is the remaining index.
__device__ __forceinline__ void jacobi_rotate(M3 &A, M3 &V, int p, int q, int k) {
  const float apq = A.m[p][q];
  if (fabsf(apq) < 1e-12f) return;
  const float app = A.m[p][p], aqq = A.m[q][q];
  const float tau = (aqq - app) / (2.0f * apq);
  const float sign_tau = tau < 0.0f ? -1.0f : 1.0f;
  const float t = sign_tau / (fabsf(tau) + sqrtf(1.0f + tau * tau));
  const float c = 1.0f / sqrtf(1.0f + t * t);
  const float s = t * c;
  A.m[p][p] = app - t * apq;
  A.m[q][q] = aqq + t * apq;
  A.m[p][q] = 0.0f;
  A.m[q][p] = 0.0f;
  const float akp = A.m[k][p], akq = A.m[k][q];
  A.m[k][p] = c * akp - s * akq;
  A.m[p][k] = A.m[k][p];
  A.m[k][q] = s * akp + c * akq;
  A.m[q][k] = A.m[k][q];
  for (int r = 0; r < 3; ++r) {
    const float vp = V.m[p][r], vq = V.m[q][r];
    V.m[p][r] = c * vp - s * vq;
    V.m[q][r] = s * vp + c * vq;
  }
}

// Pass 0 (mc_anisotropy.wgsl:110-276): per sorted particle, the weighted
// covariance of its neighbours within support_radius (kAnisoLanes lanes over
// the (2R+1)^3 stencil, R = aniso_cells), its eigen-decomposition by five
// cyclic Jacobi sweeps, the smoothed centre and the world-to-kernel matrix.
__global__ __launch_bounds__(256) void k_mc_aniso(Uni un, Bufs b, McLaunch m, float4 *aniso,
                                                  uint32_t n) {
  const uint32_t g = blockIdx.x * 256u + threadIdx.x;
  const uint32_t i = g / kAnisoLanes, sub = g % kAnisoLanes;
  if (i >= n) return;  // whole lane groups leave together
  const DevUniforms *__restrict__ U = un.u;
  const float4 xi = b.spos[i];
  const float rs = m.a.support_radius, inv_rs = 1.0f / rs;
  float w_sum = 0.f, n_count = 0.f, m1x = 0.f, m1y = 0.f, m1z = 0.f;
  float cxx = 0.f, cyy = 0.f, czz = 0.f, cxy = 0.f, cxz = 0.f, cyz = 0.f;
  const int cx = cell_coord(xi.x, U->ox, U->inv_cell);
  const int cy = cell_coord(xi.y, U->oy, U->inv_cell);
  const int cz = cell_coord(xi.z, U->oz, U->inv_cell);
  const int R = m.aniso_cells, gx = (int)U->gx, gy = (int)U->gy, gz = (int)U->gz;
  const int x0 = max(cx - R, 0), x1 = min(cx + R, gx - 1);
  const int y0 = max(cy - R, 0), y1 = min(cy + R, gy - 1);
  const int z0 = max(cz - R, 0), z1 = min(cz + R, gz - 1);
  if (x0 <= x1) {
    for (int z = z0; z <= z1; ++z) {
      for (int y = y0; y <= y1; ++y) {
        const uint32_t row = (uint32_t)y * U->gx + (uint32_t)z * U->gxy;
        const uint32_t c0 = row + (uint32_t)x0, c1 = row + (uint32_t)x1;
        const uint32_t end = b.starts[c1];
        for (uint32_t j = b.starts[c0] - b.counts[c0] + sub; j < end; j += kAnisoLanes) {
          const float4 pj = b.spos[j];
          const float dx = pj.x - xi.x, dy = pj.y - xi.y, dz = pj.z - xi.z;
          const float r = sqrtf(dot3(dx, dy, dz, dx, dy, dz));
          if (r < rs) {
            const float u = r * inv_rs;
            const float w = 1.0f - u * u * u;
            w_sum += w;
            n_count += 1.0f;
            m1x += w * dx;
            m1y += w * dy;
            m1z += w * dz;
            cxx += w * dx * dx;
            cyy += w * dy * dy;
            czz += w * dz * dz;
            cxy += w * dx * dy;
            cxz += w * dx * dz;
            cyz += w * dy * dz;
          }
        }
      }
    }
  }
  w_sum = group_sum<kAnisoLanes>(w_sum);
  n_count = group_sum<kAnisoLanes>(n_count);
  m1x = group_sum<kAnisoLanes>(m1x);
  m1y = group_sum<kAnisoLanes>(m1y);
  m1z = group_sum<kAnisoLanes>(m1z);
  cxx = group_sum<kAnisoLanes>(cxx);
  cyy = group_sum<kAnisoLanes>(cyy);
  czz = group_sum<kAnisoLanes>(czz);
  cxy = group_sum<kAnisoLanes>(cxy);
  cxz = group_sum<kAnisoLanes>(cxz);
  cyz = group_sum<kAnisoLanes>(cyz);
  if (sub != 0) return;
  // fade anisotropy and centre smoothing out for sparse neighbourhoods
  const float t_n = smoothstep(kNLo, kNHi, n_count);
  const float s_eff = m.a.strength * t_n;
  const float inv_w = 1.0f / fmaxf(w_sum, 1e-9f);
  const float mux = m1x * inv_w, muy = m1y * inv_w, muz = m1z * inv_w;
  M3 A, V;
  A.m[0][0] = cxx * inv_w - mux * mux;
  A.m[0][1] = cxy * inv_w - mux * muy;
  A.m[0][2] = cxz * inv_w - mux * muz;
  A.m[1][0] = cxy * inv_w - mux * muy;
  A.m[1][1] = cyy * inv_w - muy * muy;
  A.m[1][2] = cyz * inv_w - muy * muz;
  A.m[2][0] = cxz * inv_w - mux * muz;
  A.m[2][1] = cyz * inv_w - muy * muz;
  A.m[2][2] = czz * inv_w - muz * muz;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) V.m[c][r] = c == r ? 1.0f : 0.0f;
  for (int sweep = 0; sweep < 5; ++sweep) {
    jacobi_rotate(A, V, 0, 1, 2);
    jacobi_rotate(A, V, 0, 2, 1);
    jacobi_rotate(A, V, 1, 2, 0);
  }
  // eigenpairs in descending order (:204-221)
  float e[3] = {A.m[0][0], A.m[1][1], A.m[2][2]};
  float v[3][3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) v[c][r] = V.m[c][r];
  auto swap_pair = [&](int a, int c) {
    const float te = e[a];
    e[a] = e[c];
    e[c] = te;
    for (int r = 0; r < 3; ++r) {
      const float tv = v[a][r];
      v[a][r] = v[c][r];
      v[c][r] = tv;
    }
  };
  if (e[0] < e[1]) swap_pair(0, 1);
  if (e[0] < e[2]) swap_pair(0, 2);
  if (e[1] < e[2]) swap_pair(1, 2);
  // shape-gated centre smoothing (:223-242)
  const float s1_raw = sqrtf(fmaxf(e[0], 0.0f));
  const float s3_raw = sqrtf(fmaxf(e[2], 0.0f));
  float flatness = 0.0f;
  if (s1_raw > 1e-7f) {
    const float roundness = s3_raw / s1_raw;
    flatness = 1.0f - roundness * roundness;
  }
  const float k = m.a.lambda * s_eff * flatness;
  float shx = k * mux, shy = k * muy, shz = k * muz;
  const float shift_len = sqrtf(dot3(shx, shy, shz, shx, shy, shz));
  if (shift_len > m.a.max_shift) {
    const float f = m.a.max_shift / shift_len;
    shx *= f;
    shy *= f;
    shz *= f;
  }
  const float cenx = xi.x + shx, ceny = xi.y + shy, cenz = xi.z + shz;
  // axis scales (:244-256)
  float a0 = 1.0f, a1 = 1.0f, a2 = 1.0f;
  const float s1 = s1_raw;
  if (s1 > 1e-7f && s_eff > 0.0f) {
    const float s_min = s1 / m.a.kr;
    const float s2 = fmaxf(sqrtf(fmaxf(e[1], 0.0f)), s_min);
    const float s3 = fmaxf(s3_raw, s_min);
    const float norm = powf(s1 * s2 * s3, 1.0f / 3.0f);
    a0 = s1 / norm;
    a1 = s2 / norm;
    a2 = s3 / norm;
    a0 = powf(a0, s_eff);
    a1 = powf(a1, s_eff);
    a2 = powf(a2, s_eff);
    a0 = clampf(a0, kMinStretch, m.a.max_stretch);
    a1 = clampf(a1, kMinStretch, m.a.max_stretch);
    a2 = clampf(a2, kMinStretch, m.a.max_stretch);
  }
  // G = sum_k v_k v_k^T / (a_k h) (:258-266)
  const float h = m.a.h_mc;
  const float ir0 = 1.0f / (a0 * h), ir1 = 1.0f / (a1 * h), ir2 = 1.0f / (a2 * h);
  auto gent = [&](int r, int c) {
    return ir0 * v[0][r] * v[0][c] + ir1 * v[1][r] * v[1][c] + ir2 * v[2][r] * v[2][c];
  };
  const float gxx = gent(0, 0), gyy = gent(1, 1), gzz = gent(2, 2);
  const float gxy = gent(0, 1), gxz = gent(0, 2), gyz = gent(1, 2);
  const float amplitude = m.peak / fmaxf(a0 * a1 * a2, 1.0f);
  const float reach = h * fmaxf(a0, fmaxf(a1, a2));
  float4 *out = aniso + 3 * (size_t)i;
  out[0] = make_float4(gxx, gxy, gxz, cenx);
  out[1] = make_float4(gyy, gyz, gzz, ceny);
  out[2] = make_float4(cenz, reach, amplitude, 0.0f);
}

// world_to_local (container_common.wgsl:28-35) of the MC pass's container
__device__ __forceinline__ float3 mc_to_local(const GpuContainerGeometry &c, float x, float y,
                                              float z) {
  const float py = y - c.center_y;
  return make_float3(dot3(c.inverse_row0[0], c.inverse_row0[1], c.inverse_row0[2], x, py, z),
                     dot3(c.inverse_row1[0], c.inverse_row1[1], c.inverse_row1[2], x, py, z),
                     dot3(c.inverse_row2[0], c.inverse_row2[1], c.inverse_row2[2], x, py, z));
}

__device__ __forceinline__ int wave_min_i(int v) {
  for (int o = 32; o >= 1; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
  for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// Pass 1 (mc_density.wgsl:133-230): one 4x4x4 voxel block per wave, four
// blocks per workgroup along z.  No lane leaves before the candidate walk
// (the union bounds and the walk are wave-wide).
template <bool ANISO>
__global__ __launch_bounds__(256) void k_mc_density(Uni un, Bufs b, McLaunch m,
                                                    const float4 *__restrict__ aniso,
                                                    float *__restrict__ field) {
  const DevUniforms *__restrict__ U = un.u;
  const uint32_t G = m.g.grid_size;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if ((blockIdx.z * 4u + wave) * 4u >= G) return;  // a whole wave past the grid
  const uint32_t vx = blockIdx.x * 4u + (lane & 3u);
  const uint32_t vy = blockIdx.y * 4u + ((lane >> 2) & 3u);
  const uint32_t vz = (blockIdx.z * 4u + wave) * 4u + (lane >> 4);
  const bool in = vx < G && vy < G && vz < G;
  // grid_to_world (:87-90)
  const float cs = m.g.cell_size;
  const float wx = m.g.grid_min[0] + ((float)vx + 0.5f) * cs;
  const float wy = m.g.grid_min[1] + ((float)vy + 0.5f) * cs;
  const float wz = m.g.grid_min[2] + ((float)vz + 0.5f) * cs;
  const int R = m.density_cells;
  const int gx = (int)U->gx, gy = (int)U->gy, gz = (int)U->gz;
  const int cx = cell_coord(wx, U->ox, U->inv_cell);
  const int cy = cell_coord(wy, U->oy, U->inv_cell);
  const int cz = cell_coord(wz, U->oz, U->inv_cell);
  // union of the block's stencils, clipped to the grid (wave-uniform: read
  // from lane 0 so the walk below runs on scalar registers and loads)
  const int x0 = __builtin_amdgcn_readfirstlane(max(wave_min_i(cx) - R, 0));
  const int x1 = __builtin_amdgcn_readfirstlane(min(wave_max_i(cx) + R, gx - 1));
  const int y0 = __builtin_amdgcn_readfirstlane(max(wave_min_i(cy) - R, 0));
  const int y1 = __builtin_amdgcn_readfirstlane(min(wave_max_i(cy) + R, gy - 1));
  const int z0 = __builtin_amdgcn_readfirstlane(max(wave_min_i(cz) - R, 0));
  const int z1 = __builtin_amdgcn_readfirstlane(min(wave_max_i(cz) + R, gz - 1));
  const float h = m.g.kernel_radius, h_sq = h * h;
  float density = 0.0f;
  // this voxel's own stencil (:157-166), clipped to the grid
  const int ox0 = max(cx - R, 0), ox1 = min(cx + R, gx - 1);
  // Cells whose box is farther than the largest splat reach (plus the
  // centre shift) from every voxel of the block hold no contributing
  // particle and are skipped; boundary cells extend to infinity (the build
  // clamps particles outside the grid into them).  The bound carries a
  // 1e-4 relative margin over f32 rounding.
  const float cell = 1.0f / U->inv_cell;
  const float L = (ANISO ? h * m.a.max_stretch + m.a.max_shift : h) * 1.0001f + 1e-6f;
  const uint32_t bx = blockIdx.x * 4u, by = blockIdx.y * 4u, bz = (blockIdx.z * 4u + wave) * 4u;
  const float bx_lo = m.g.grid_min[0] + ((float)bx + 0.5f) * cs, bx_hi = bx_lo + 3.0f * cs;
  const float by_lo = m.g.grid_min[1] + ((float)by + 0.5f) * cs, by_hi = by_lo + 3.0f * cs;
  const float bz_lo = m.g.grid_min[2] + ((float)bz + 0.5f) * cs, bz_hi = bz_lo + 3.0f * cs;
  auto gap = [&](int k, int kmax, float o, float lo, float hi) {
    const float c_lo = o + (float)k * cell, c_hi = c_lo + cell;
    if (k > 0 && hi < c_lo) return c_lo - hi;
    if (k < kmax && lo > c_hi) return lo - c_hi;
    return 0.0f;
  };
  if (x0 <= x1) {
    for (int z = z0; z <= z1; ++z) {
      const float gz2 = gap(z, gz - 1, U->oz, bz_lo, bz_hi);
      for (int y = y0; y <= y1; ++y) {
        const float gy2 = gap(y, gy - 1, U->oy, by_lo, by_hi);
        // (the builtin moves 32-bit integers: pass the float's bits)
        const float rem = __int_as_float(
            __builtin_amdgcn_readfirstlane(__float_as_int(L * L - gz2 * gz2 - gy2 * gy2)));
        if (!(rem > 0.0f)) continue;
        const float allow = sqrtf(rem);
        // the cells within reach along x; the border cells 0 and gx - 1
        // extend to infinity, as in gap()
        const int xa = __builtin_amdgcn_readfirstlane(
            max(x0, min((int)floorf((bx_lo - allow - U->ox) * U->inv_cell), gx - 1)));
        const int xb = __builtin_amdgcn_readfirstlane(
            min(x1, max((int)floorf((bx_hi + allow - U->ox) * U->inv_cell), 0)));
        if (xa > xb) continue;
        const uint32_t row = (uint32_t)y * U->gx + (uint32_t)z * U->gxy;
        const uint32_t c0 = row + (uint32_t)xa, c1 = row + (uint32_t)xb;
        const uint32_t begin = __builtin_amdgcn_readfirstlane(b.starts[c0] - b.counts[c0]);
        const uint32_t end = __builtin_amdgcn_readfirstlane(b.starts[c1]);
        // the slots of this row that lie in this voxel's stencil
        uint32_t lo = 0u, hi = 0u;
        if (ox0 <= ox1 && y >= cy - R && y <= cy + R && z >= cz - R && z <= cz + R) {
          const uint32_t a0 = row + (uint32_t)ox0, a1 = row + (uint32_t)ox1;
          lo = b.starts[a0] - b.counts[a0];
          hi = b.starts[a1];
        }
        // four candidates per trip, their records loaded together (scalar
        // loads: the index is wave-uniform); terms are added in slot order
        uint32_t j = begin;
        if (ANISO) {
          auto term = [&](uint32_t jj, float4 q0, float4 q1, float4 q2) {
            const float dx = wx - q0.w, dy = wy - q1.w, dz = wz - q2.x;
            const float r_sq = dot3(dx, dy, dz, dx, dy, dz);
            const float reach = q2.y;
            if (jj >= lo && jj < hi && r_sq < reach * reach) {
              const float gdx = q0.x * dx + q0.y * dy + q0.z * dz;
              const float gdy = q0.y * dx + q1.x * dy + q1.y * dz;
              const float gdz = q0.z * dx + q1.y * dy + q1.z * dz;
              const float q_sq = dot3(gdx, gdy, gdz, gdx, gdy, gdz);
              if (q_sq < 1.0f) {
                const float f = 1.0f - q_sq;
                density += q2.z * f * f * f;
              }
            }
          };
          for (; j + 4 <= end; j += 4) {
            float4 q[12];
#pragma unroll
            for (int t = 0; t < 12; ++t) q[t] = aniso[3 * (size_t)j + t];
#pragma unroll
            for (int t = 0; t < 4; ++t) term(j + t, q[3 * t], q[3 * t + 1], q[3 * t + 2]);
          }
          for (; j < end; ++j)
            term(j, aniso[3 * (size_t)j], aniso[3 * (size_t)j + 1], aniso[3 * (size_t)j + 2]);
        } else {
          auto term = [&](uint32_t jj, float4 p) {
            const float dx = wx - p.x, dy = wy - p.y, dz = wz - p.z;
            const float r_sq = dot3(dx, dy, dz, dx, dy, dz);
            if (jj >= lo && jj < hi && r_sq < h_sq) {
              const float diff = h_sq - r_sq;
              density += m.coeff * diff * diff * diff;
            }
          };
          for (; j + 4 <= end; j += 4) {
            float4 q[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) q[t] = b.spos[j + t];
#pragma unroll
            for (int t = 0; t < 4; ++t) term(j + t, q[t]);
          }
          for (; j < end; ++j) term(j, b.spos[j]);
        }
      }
    }
  }
  if (!in) return;
  const float3 l = mc_to_local(m.c, wx, wy, wz);
  float out;
  if (m.c.clip_enabled != 0u &&
      !(l.x >= -m.c.half_width && l.x <= m.c.half_width && l.y >= -m.c.half_height &&
        l.y <= m.c.half_height && l.z >= -m.c.half_depth && l.z <= m.c.half_depth)) {
    out = -1.0f;  // sentinel outside the container (:209-214)
  } else {
    // boundary_gamma, five walls (no ceiling) (:109-129)
    const float d[5] = {l.x + m.c.half_width, m.c.half_width - l.x, l.y + m.c.half_height,
                        l.z + m.c.half_depth, m.c.half_depth - l.z};
    float gamma = 1.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k)
      if (d[k] < h) gamma *= 0.5f + 0.5f * clampf(d[k] / h, 0.0f, 1.0f);
    gamma = fmaxf(gamma, 0.15f);
    out = density / gamma;
  }
  field[(size_t)vx + (size_t)vy * G + (size_t)vz * G * G] = out;
}

}  // namespace

hipError_t launch_mc_aniso(const Uni &u, const Bufs &b, const McLaunch &m, float4 *aniso,
                           uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mc_aniso, dim3(blocks_for((uint64_t)n * kAnisoLanes, 256)), dim3(256),
                     0, st, u, b, m, aniso, n);
  return hipGetLastError();
}

hipError_t launch_mc_density(const Uni &u, const Bufs &b, const McLaunch &m,
                             const float4 *aniso, float *field, hipStream_t st) {
  const uint32_t nb = (m.g.grid_size + 3u) / 4u;  // 4-voxel blocks per axis
  const dim3 grid(nb, nb, (nb + 3u) / 4u);
  if (m.a.enabled)
    hipLaunchKernelGGL(k_mc_density<true>, grid, dim3(256), 0, st, u, b, m, aniso, field);
  else
    hipLaunchKernelGGL(k_mc_density<false>, grid, dim3(256), 0, st, u, b, m, aniso, field);
  return hipGetLastError();
}

}  // namespace fluid
