// sph_device.h — device helpers shared by the step kernels and their
// adjoints: fused dot product, cell coordinates, wave reductions, the
// uniform lookup of batched handles and the 27-cell stencil walk.
#pragma once

#include "fluid_internal.h"

namespace fluid {


__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx,
                                      float by, float bz) {
  return __builtin_fmaf(az, bz, __builtin_fmaf(ay, by, ax * bx));
}

// position_to_cell (grid_build.wgsl:33-36): i32(floor((p - o) * inv)).
// v_cvt_i32_f32 saturates and maps NaN to 0, the same as the oracle.
__device__ __forceinline__ int cell_coord(float p, float o, float inv) {
  return (int)floorf((p - o) * inv);
}

__device__ __forceinline__ float fast_sqrt(float x) {
  return __builtin_amdgcn_sqrtf(x);
}

// Inclusive wave64 scan on DPP: row_shr 1,2,4,8 within each 16-lane row,
// then row_bcast:15 into rows 1 and 3 and row_bcast:31 into rows 2 and 3.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);
  return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

template <int L>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int o = L / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o, L);
  return v;
}

// The uniform block of slot s: the handle's only block, or its scene's.
template <bool B>
__device__ __forceinline__ const DevUniforms *uni_at(const Uni &un, uint32_t s) {
  if constexpr (B) return un.u + un.scene_of[s];
  else return un.u;
}

// Workgroup index remapped so that each XCD (workgroups are dealt to the 8
// XCDs round-robin) receives one contiguous run of tiles: neighbouring
// particles then share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t remap) {
  const uint32_t nb = gridDim.x, i = blockIdx.x;
  if (!remap) return i;
  const uint32_t q = nb >> 3, r = nb & 7u, x = i & 7u, k = i >> 3;
  return x * q + min(x, r) + k;
}

// The 27-cell stencil of sph_density_3d_grid.wgsl:129-167, walked as nine
// x-rows: cells (x-1..x+1, y, z) are consecutive cell indices, so their
// particles form one contiguous sorted range.  The visit order (dz, dy, dx,
// then slot) is the reference's.  Lane `sub` of an L-lane group takes every
// L-th candidate of each row, four gathers in flight per trip.
//   SYNC = false: each lane stops at its own last candidate and f(j, p_j,
//                 true) only sees real candidates.
//   SYNC = true:  all L lanes of a particle make the same trips, the tail
//                 predicated by `valid`, so group-wide ballots inside
//                 f(j, p_j, valid) see every lane of the group.
template <int L, bool SYNC, class F>
__device__ __forceinline__ void sweep27(const DevUniforms *__restrict__ U,
                                        const uint32_t *__restrict__ counts,
                                        const uint32_t *__restrict__ starts,
                                        const float4 *__restrict__ pts,
                                        int cx, int cy, int cz, uint32_t sub,
                                        F &&f) {
  const int gx = (int)U->gx, gy = (int)U->gy, gz = (int)U->gz;
  cx = min(max(cx, -2), gx + 1);
  cy = min(max(cy, -2), gy + 1);
  cz = min(max(cz, -2), gz + 1);
  const int x0 = max(cx - 1, 0), x1 = min(cx + 1, gx - 1);
  if (x0 > x1) return;
  const int z0 = max(cz - 1, 0), z1 = min(cz + 1, gz - 1);
  const int y0 = max(cy - 1, 0), y1 = min(cy + 1, gy - 1);
  for (int z = z0; z <= z1; ++z) {
    for (int y = y0; y <= y1; ++y) {
      const uint32_t row = U->cell_offset + (uint32_t)y * U->gx + (uint32_t)z * U->gxy;
      const uint32_t c0 = row + (uint32_t)x0, c1 = row + (uint32_t)x1;
      const uint32_t begin = starts[c0] - counts[c0];
      const uint32_t end = starts[c1];
      if (SYNC) {
        if (begin >= end) continue;
        const uint32_t last = end - 1;
        for (uint32_t base = begin; base < end; base += 4 * L) {
          const uint32_t j0 = base + sub, j1 = j0 + L, j2 = j1 + L, j3 = j2 + L;
          const float4 p0 = pts[min(j0, last)], p1 = pts[min(j1, last)];
          const float4 p2 = pts[min(j2, last)], p3 = pts[min(j3, last)];
          f(j0, p0, j0 < end);
          f(j1, p1, j1 < end);
          f(j2, p2, j2 < end);
          f(j3, p3, j3 < end);
        }
      } else {
        uint32_t j = begin + sub;
        for (; j + 3 * L < end; j += 4 * L) {
          const float4 p0 = pts[j], p1 = pts[j + L], p2 = pts[j + 2 * L], p3 = pts[j + 3 * L];
          f(j, p0, true);
          f(j + L, p1, true);
          f(j + 2 * L, p2, true);
          f(j + 3 * L, p3, true);
        }
        for (; j < end; j += L) f(j, pts[j], true);
      }
    }
  }
}

}  // namespace fluid
