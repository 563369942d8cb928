// sph_device.h — device helpers shared by the step kernels and their
// adjoints: fused dot product, cell coordinates, wave reductions, the
// uniform lookup of batched handles, the 27-cell stencil walk and the walk
// of the density pass's in-range lists.
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

// Direction of a pair closer than 0.1 h: normalize(sin(d), cos(2d), sin(3d))
// of the sorted-slot difference d = f32(s) - f32(j), an integer
// (sph_force_3d_grid.wgsl:163-172).  d (|d| < 2^24) is reduced to revolutions
// f = d / (2 pi) - rint(d / (2 pi)) in double (error ~5e-10), and the
// hardware sin / cos of 2 pi f (the instructions take revolutions; within
// 1.3e-7 of sin, cos on [-1/2, 1/2], measured on 4M points) give sin d and
// cos d; then cos 2d = (cos d - sin d)(cos d + sin d) and
// sin 3d = sin d (3 - 4 sin^2 d).  Each component is within ~1e-6 of the
// library functions of 2d and 3d, at a few dozen instructions instead of
// three large-argument reductions.
__device__ __forceinline__ float3 close_dir(uint32_t s, uint32_t j) {
  const float idx = (float)s - (float)j;
  const double q = (double)idx * 0.15915494309189535;  // 1 / (2 pi)
  const float f = (float)(q - rint(q));
  const float sn = __builtin_amdgcn_sinf(f), cs = __builtin_amdgcn_cosf(f);
  const float a = sn;
  const float c = (cs - sn) * (cs + sn);
  const float e = sn * __builtin_fmaf(-4.0f * sn, sn, 3.0f);
  const float il = __builtin_amdgcn_rsqf(dot3(a, c, e, a, c, e));
  return make_float3(a * il, c * il, e * il);
}

// Per-lane lists (the in-range lists, the solve's skin lists) live in
// blocks of the P = 64 / L slots one wave covers: entry k of lane `sub` of
// slot s is at list[lane_block<L>(s, cap) + k * 64 + sub], cap entries per
// slot.  A wave's k-th entries are one contiguous 64-word run, and offsets
// within a block are 32-bit (the block base is computed once per slot).
template <int L>
__device__ __forceinline__ size_t lane_block(uint32_t s, uint32_t cap) {
  constexpr uint32_t P = 64 / L;
  return (size_t)(s / P) * (P * cap) + (s % P) * L;
}

// Appends j to a lane group's in-range list when `pair`.  Every lane of the
// group (L lanes) calls it at the same point of a group-synchronous walk; the
// group's k-th entry (ranked across its lanes by a ballot, lanes in order)
// goes to list lane k % LW at depth k / LW, i.e. blk[(k / LW) * 64 + k % LW]
// with blk = list + lane_block<LW>(slot, CAP), so each lane of a later walk
// takes an equal share.  run counts the group's entries (above CAP entries
// per slot the list is incomplete and the count marks it).  For L <= 8
// the group's ballot bits are byte lane / 8 of the wave's ballot (one byte
// permute), shifted down for L < 8.
template <int L, int LW, uint32_t CAP = kListCapacity>
__device__ __forceinline__ void group_list_append(uint32_t *__restrict__ blk, uint32_t sub,
                                                  uint32_t &run, bool pair, uint32_t j) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t bal = __ballot(pair);
  uint32_t gm;
  if constexpr (L > 8) {
    gm = (uint32_t)(bal >> (lane & ~(uint32_t)(L - 1))) & (uint32_t)((1ull << L) - 1ull);
  } else {
    const uint32_t byte = __builtin_amdgcn_perm((uint32_t)(bal >> 32), (uint32_t)bal,
                                                0x0c0c0c00u | (lane >> 3));
    gm = L == 8 ? byte : (byte >> (lane & 7u & ~(uint32_t)(L - 1))) & ((1u << L) - 1u);
  }
  if (pair) {
    const uint32_t e = run + (uint32_t)__popc(gm & ((1u << sub) - 1u));
    // (e / LW) * 64 + e % LW, as one multiply-add
    if (e < CAP) blk[e + (e & ~(uint32_t)(LW - 1)) * (64u / LW - 1u)] = j;
  }
  run += (uint32_t)__popc(gm);
}

// The pair test 1e-12 < r^2 < h^2 as one unsigned compare of bit patterns:
// r^2, a sum of squares, is +0 or more and orders as its bits, and NaN falls
// outside.  A single compare is a mask a ballot takes as it is (two compares
// and an AND are turned back into a lane value and compared again).
struct PairTest {
  uint32_t lo, span;
  __device__ __forceinline__ explicit PairTest(float h2) {
    lo = __float_as_uint(1e-12f) + 1u;
    const uint32_t hb = __float_as_uint(h2);
    span = (h2 > 0.f && hb > lo) ? hb - lo : 0u;
  }
  __device__ __forceinline__ bool operator()(float r2) const {
    return __float_as_uint(r2) - lo < span;
  }
};

// First slot of cell c: the inclusive scan's end of the cell before it (the
// scan runs over the whole concatenated cell array, so across scenes too).
// One load, where starts[c] - counts[c] reads two arrays.
__device__ __forceinline__ uint32_t cell_begin(const uint32_t *__restrict__ starts, uint32_t c) {
  return c ? starts[c - 1u] : 0u;
}

// Order key of a float: unsigned order of the keys is the float order of
// the values (-0 before +0), a total order over every bit pattern (NaNs at
// the two ends).
__device__ __forceinline__ uint32_t x_key(float x) {
  const uint32_t u = __float_as_uint(x);
  return u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
}

// fmaxf(x, 1.0f) for every non-NaN x in one integer max: non-negative floats
// order as their bit patterns, and negative ones (sign bit set) are negative
// as int32, below 1.0f's pattern.  fmaxf itself costs two instructions here
// (its operand is first canonicalized).
__device__ __forceinline__ float max1(float x) {
  return __int_as_float(max(__float_as_int(x), 0x3f800000));
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
// L-th candidate of each row, four gathers in flight per trip (UN with SYNC
// = false).
//   SYNC = false: each lane stops at its own last candidate and f(j, p_j,
//                 true) only sees real candidates.
//   SYNC = true:  all L lanes of a particle make the same trips, the tail
//                 predicated by `valid`, so group-wide ballots inside
//                 f(j, p_j, valid) see every lane of the group.
template <int L, bool SYNC, int UN = 4, class T, class F>
__device__ __forceinline__ void sweep27(const DevUniforms *__restrict__ U,
                                        const uint32_t *__restrict__ counts,
                                        const uint32_t *__restrict__ starts,
                                        const T *__restrict__ pts,
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
      const uint32_t begin = cell_begin(starts, c0);
      const uint32_t end = starts[c1];
      // Full trips address through one running pointer (the four loads use
      // immediate offsets); only a SYNC tail trip clamps its indices.
      if (SYNC) {
        if (begin >= end) continue;
        const uint32_t last = end - 1;
        uint32_t base = begin;
        const T *p = pts + begin + sub;
        for (; base + 4 * L <= end; base += 4 * L, p += 4 * L) {
          const uint32_t j0 = base + sub;
          const T p0 = p[0], p1 = p[L], p2 = p[2 * L], p3 = p[3 * L];
          f(j0, p0, true);
          f(j0 + L, p1, true);
          f(j0 + 2 * L, p2, true);
          f(j0 + 3 * L, p3, true);
        }
        if (base < end) {
          const uint32_t j0 = base + sub, j1 = j0 + L, j2 = j1 + L, j3 = j2 + L;
          const T p0 = pts[min(j0, last)], p1 = pts[min(j1, last)];
          const T p2 = pts[min(j2, last)], p3 = pts[min(j3, last)];
          f(j0, p0, j0 < end);
          f(j1, p1, j1 < end);
          f(j2, p2, j2 < end);
          f(j3, p3, j3 < end);
        }
      } else {
        uint32_t j = begin + sub;
        const T *p = pts + j;
        for (; j + (UN - 1) * L < end; j += UN * L, p += UN * L) {
          T c[UN];
#pragma unroll
          for (int q = 0; q < UN; ++q) c[q] = p[q * L];
#pragma unroll
          for (int q = 0; q < UN; ++q) f(j + q * L, c[q], true);
        }
        for (; j < end; j += L, p += L) f(j, p[0], true);
      }
    }
  }
}

// ------------------------------------------------------------------------
// The 27-cell stencil as one concatenated run of candidates: the nine
// x-rows of sweep27, in its order, end to end.  Lane `sub` of an L-lane
// group takes candidates sub, sub + L, sub + 2L, ... of the concatenation,
// so its share runs on across row boundaries: rows shorter than the group
// leave no lane idle, every trip's four loads are independent of the rows,
// and a lane's trips number ceil(candidates / 4L) instead of the sum of
// every row's.  Each row is (first slot, length, index of the first slot in
// the array the candidates are loaded from); a lane reads the next row only
// when its share crosses into it.
struct RowEnt {
  uint32_t gb, len, lb;
};

// Candidates loaded together per trip of walk_rows.
#ifndef FLUID_WALK_LOADS
#define FLUID_WALK_LOADS 4
#endif
constexpr int kWalkLoads = FLUID_WALK_LOADS;

template <int L>
__device__ __forceinline__ uint32_t group_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = L / 2; o >= 1; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o, L);
  return v;
}

// Whether any lane of this lane's L-lane group has `p` (all of them call it).
template <int L>
__device__ __forceinline__ bool group_any(bool p) {
  const uint64_t bal = __ballot(p);
  if constexpr (L >= 64) return bal != 0;
  const uint32_t lane = threadIdx.x & 63u;
  return ((bal >> (lane & ~(uint32_t)(L - 1))) & ((1ull << L) - 1ull)) != 0;
}

// The row table of the 27-cell stencil of cell (cx, cy, cz), written into
// the group's nine LDS entries `tab` by the group's lanes; every lane of the
// group must call it.  Returns the group's candidate count.
template <int L>
__device__ __forceinline__ uint32_t rows27(const DevUniforms *__restrict__ U,
                                           const uint32_t *__restrict__ counts,
                                           const uint32_t *__restrict__ starts, int cx, int cy,
                                           int cz, uint32_t sub, RowEnt *tab) {
  const int gx = (int)U->gx, gy = (int)U->gy, gz = (int)U->gz;
  cx = min(max(cx, -2), gx + 1);
  cy = min(max(cy, -2), gy + 1);
  cz = min(max(cz, -2), gz + 1);
  const int x0 = max(cx - 1, 0), x1 = min(cx + 1, gx - 1);
  uint32_t mine = 0;
  for (uint32_t q = sub; q < 9u; q += L) {
    // row q: dz = q / 3 - 1 (outer), dy = q % 3 - 1, the reference's order
    const int z = cz + (int)(q / 3u) - 1, y = cy + (int)(q % 3u) - 1;
    uint32_t gb = 0, len = 0;
    if (x0 <= x1 && z >= 0 && z < gz && y >= 0 && y < gy) {
      const uint32_t row = U->cell_offset + (uint32_t)y * U->gx + (uint32_t)z * U->gxy;
      const uint32_t c0 = row + (uint32_t)x0, c1 = row + (uint32_t)x1;
      gb = cell_begin(starts, c0);
      len = starts[c1] - gb;
    }
    tab[q] = RowEnt{gb, len, gb};
    mine += len;
  }
  // the group's lanes read each other's entries (one wave: in order)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return group_sum_u32<L>(mine);
}

// Walk a group's concatenated rows (row(r) gives row r, total their summed
// length).  pts is indexed by RowEnt::lb + offset; `safe` is any valid index
// of pts, loaded in place of the candidates past the end.
//   SYNC = false: each lane walks its own share and f(j, p_j, true) sees
//                 real candidates only.
//   SYNC = true:  the group's lanes make the same ceil(total / 4L) trips,
//                 candidates past the end passed with valid = false, so
//                 group-wide ballots inside f(j, p_j, valid) see every lane.
template <int L, bool SYNC, class T, class Row, class F, int U = kWalkLoads>
__device__ __forceinline__ void walk_rows(Row &&row, uint32_t total, const T *__restrict__ pts,
                                          uint32_t sub, uint32_t safe, F &&f) {
  uint32_t r = 0, o = sub;
  RowEnt e = row(0u);
  // move (r, o) on to the row holding offset o (past the last row: r = 9,
  // e.len = 0); a loop, not unrolled: crossings are rare
  auto settle = [&]() {
#pragma clang loop unroll(disable)
    while (r < 9u && o >= e.len) {
      o -= e.len;
      ++r;
      e = r < 9u ? row(r) : RowEnt{0u, 0u, 0u};
    }
  };
  settle();
  auto trip = [&](uint32_t (&j)[U], T (&p)[U], bool (&v)[U]) {
    if (o + (uint32_t)(U - 1) * L < e.len) {
      // the whole trip lies in the current row: one address, the U loads
      // at fixed offsets from it
      const T *p0 = pts + (e.lb + o);
#pragma unroll
      for (int q = 0; q < U; ++q) {
        v[q] = true;
        j[q] = e.gb + o + (uint32_t)q * L;
        p[q] = p0[q * L];
      }
      o += (uint32_t)U * L;
      settle();
    } else {
      uint32_t li[U];
#pragma unroll
      for (int q = 0; q < U; ++q) {
        v[q] = r < 9u;
        j[q] = e.gb + o;
        li[q] = v[q] ? e.lb + o : safe;
        o += L;
        settle();
      }
#pragma unroll
      for (int q = 0; q < U; ++q) p[q] = pts[li[q]];
    }
  };
  if (SYNC) {
    for (uint32_t base = 0; base < total; base += (uint32_t)U * L) {
      uint32_t j[U];
      bool v[U];
      T p[U];
      trip(j, p, v);
#pragma unroll
      for (int q = 0; q < U; ++q) f(j[q], p[q], v[q]);
    }
  } else {
    while (r < 9u) {
      uint32_t j[U];
      bool v[U];
      T p[U];
      trip(j, p, v);
#pragma unroll
      for (int q = 0; q < U; ++q)
        if (v[q]) f(j[q], p[q], true);
    }
  }
}

// ------------------------------------------------------------------------
// Cell tiles.  A tile is a run of up to w (<= kTileWMax) cells along x of
// one grid row (y, z) of one scene.  Its particles' 27-cell stencils cover
// nine x-rows (y-1..y+1, z-1..z+1) of cells xa = x0-1 .. xb = x1+1, each a
// contiguous range of sorted slots; a workgroup stages those ranges in LDS
// once and every particle of the tile sweeps its candidates from there.
struct TileShared {
  uint32_t goff[9][kTileWMax + 3];  // slot where each row cell starts; [nx] = row end
  uint32_t rowl[10];                // staged index of each row's first slot; [9] = total
  int x0, x1, y, z, xa, xb;
  uint32_t own_lo, own_hi;          // the tile's own slots
};

// Decode scene-local tile `lt` and fill S (all threads of the workgroup
// call this; it synchronises).  Rows outside the grid have length 0.
__device__ __forceinline__ void tile_setup(const DevUniforms *__restrict__ U,
                                           const uint32_t *__restrict__ counts,
                                           const uint32_t *__restrict__ starts, uint32_t lt,
                                           uint32_t w, TileShared &S) {
  const int gx = (int)U->gx, gy = (int)U->gy, gz = (int)U->gz;
  const uint32_t nxc = ((uint32_t)gx + w - 1) / w;
  const int xc = (int)(lt % nxc), row = (int)(lt / nxc);
  const int y = row % gy, z = row / gy;
  const int x0 = xc * (int)w, x1 = min(x0 + (int)w, gx) - 1;
  const int xa = max(x0 - 1, 0), xb = min(x1 + 1, gx - 1);
  const int nx = xb - xa + 1;
  for (int t = (int)threadIdx.x; t < 9 * (nx + 1); t += (int)blockDim.x) {
    const int r = t / (nx + 1), k = t % (nx + 1);
    const int yy = y + (r % 3) - 1, zz = z + (r / 3) - 1;
    uint32_t v = 0;
    if (yy >= 0 && yy < gy && zz >= 0 && zz < gz) {
      const uint32_t rc = U->cell_offset + (uint32_t)yy * U->gx + (uint32_t)zz * U->gxy;
      v = k < nx ? starts[rc + xa + k] - counts[rc + xa + k] : starts[rc + xb];
    }
    S.goff[r][k] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int r = 0; r < 9; ++r) {
      S.rowl[r] = acc;
      acc += S.goff[r][nx] - S.goff[r][0];
    }
    S.rowl[9] = acc;
    S.x0 = x0;
    S.x1 = x1;
    S.y = y;
    S.z = z;
    S.xa = xa;
    S.xb = xb;
    S.own_lo = S.goff[4][x0 - xa];
    S.own_hi = S.goff[4][x1 - xa + 1];
  }
  __syncthreads();
}

// Copy the nine rows of src into dst (LDS) when they fit in `cap`
// entries; returns whether they were staged.
template <class T>
__device__ __forceinline__ bool tile_stage(const TileShared &S, const T *__restrict__ src,
                                           T *dst, uint32_t cap) {
  const uint32_t total = S.rowl[9];
  if (total > cap) return false;
  for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
    int r = 0;
    while (i >= S.rowl[r + 1]) ++r;
    dst[i] = src[S.goff[r][0] + (i - S.rowl[r])];
  }
  __syncthreads();
  return true;
}

// Whether a particle with cell coordinates (cx, cy, cz) sweeps this tile's
// rows: it lies in one of the tile's own cells (a particle outside the grid
// lives in a border cell but sweeps around its own coordinates).
__device__ __forceinline__ bool tile_owns(const TileShared &S, const DevUniforms *U, int cx,
                                          int cy, int cz) {
  cx = min(max(cx, -2), (int)U->gx + 1);
  cy = min(max(cy, -2), (int)U->gy + 1);
  cz = min(max(cz, -2), (int)U->gz + 1);
  return cy == S.y && cz == S.z && cx >= S.x0 && cx <= S.x1;
}

// The same walk as sweep27 (same rows, same trips, same candidate order per
// lane) over a tile's rows; pts is the staged copy (staged = true, indexed
// from 0) or the global array (indexed by slot).  f(j, p_j, valid) receives
// the global slot j.
template <int L, bool SYNC, class T, class F>
__device__ __forceinline__ void tile_sweep(const TileShared &S, const T *pts, bool staged,
                                           int cx, uint32_t sub, F &&f) {
  const int lo = max(cx - 1, S.xa) - S.xa, hi = min(cx + 1, S.xb) - S.xa + 1;
  for (int r = 0; r < 9; ++r) {
    const uint32_t begin = S.goff[r][lo], end = S.goff[r][hi];
    if (begin >= end) continue;
    // staged index of slot j: j + off (mod 2^32) = rowl[r] + (j - goff[r][0])
    const uint32_t off = staged ? S.rowl[r] - S.goff[r][0] : 0u;
    if (SYNC) {
      const uint32_t last = end - 1;
      for (uint32_t base = begin; base < end; base += 4 * L) {
        const uint32_t j0 = base + sub, j1 = j0 + L, j2 = j1 + L, j3 = j2 + L;
        const T p0 = pts[min(j0, last) + off], p1 = pts[min(j1, last) + off];
        const T p2 = pts[min(j2, last) + off], p3 = pts[min(j3, last) + off];
        f(j0, p0, j0 < end);
        f(j1, p1, j1 < end);
        f(j2, p2, j2 < end);
        f(j3, p3, j3 < end);
      }
    } else {
      uint32_t j = begin + sub;
      for (; j + 3 * L < end; j += 4 * L) {
        const T p0 = pts[j + off], p1 = pts[j + L + off];
        const T p2 = pts[j + 2 * L + off], p3 = pts[j + 3 * L + off];
        f(j, p0, true);
        f(j + L, p1, true);
        f(j + 2 * L, p2, true);
        f(j + 3 * L, p3, true);
      }
      for (; j < end; j += L) f(j, pts[j + off], true);
    }
  }
}

// Row r of the 27-cell stencil of a particle in the tile with cell x
// coordinate cx, for walk_rows: slots from the tile's row table, loaded from
// the staged copy (staged) or the global array.
struct TileRows {
  const TileShared &S;
  int lo, hi;
  bool staged;
  __device__ __forceinline__ TileRows(const TileShared &s_, int cx, bool st)
      : S(s_), lo(max(cx - 1, s_.xa) - s_.xa), hi(min(cx + 1, s_.xb) - s_.xa + 1), staged(st) {}
  __device__ __forceinline__ RowEnt operator()(uint32_t r) const {
    const uint32_t b0 = S.goff[r][lo], b1 = S.goff[r][hi];
    return RowEnt{b0, b1 - b0, staged ? b0 + S.rowl[r] - S.goff[r][0] : b0};
  }
  __device__ __forceinline__ uint32_t total() const {
    uint32_t t = 0;
#pragma unroll
    for (int r = 0; r < 9; ++r) t += S.goff[r][hi] - S.goff[r][lo];
    return t;
  }
  // the load index of slot s, one of the tile's own particles (row 4)
  __device__ __forceinline__ uint32_t own(uint32_t s) const {
    return staged ? s + S.rowl[4] - S.goff[4][0] : s;
  }
};

// The neighbour record the density pass writes for slot j: srec[3 j + k]
// is position and rho / gamma (k = 0), velocity and near density (1), the
// scaled normal and 1 / max(near density, 1) (2).
struct SlotRec {
  float4 r[3];
};

// The first NR parts of a neighbour record, as the list walks hand it on.
template <int NR>
struct NbrRec {
  float4 r[NR];
};

// Walk this lane's share of particle s's in-range list, or, if the list
// overflowed, the lane's candidate subset again with the same test;
// f(j, rec) sees every pair once, with the first NR parts of j's record.
// The list is written for LW = 8 lanes; with L = 16 or 32 walkers, M = L / 8
// of them share each list lane, taking interleaved depths.  Full chunks of
// U pairs per trip, the next chunk's list entries loaded a trip ahead so
// only the record gathers wait on memory.  Every record of a trip is loaded
// before any pair is summed: a load inside the per-pair test of the sweep
// would wait on memory once per pair.
template <int L, int U, int NR, class F>
__device__ __forceinline__ void for_each_neighbour(const DevUniforms *__restrict__ Un,
                                                   const Bufs &b, uint32_t n, uint32_t s,
                                                   uint32_t sub, float4 pi, F &&f) {
  static_assert(L % kListLanes == 0, "walkers are a multiple of the list lanes");
  static_assert(NR >= 1 && NR <= 3, "a record has three parts");
  constexpr uint32_t LW = kListLanes, M = L / LW;
  const uint32_t total = b.ncnt[s];
  const float4 *__restrict__ srec = b.srec;
  auto load = [&](uint32_t j) {
    NbrRec<NR> q;
#pragma unroll
    for (int k = 0; k < NR; ++k) q.r[k] = srec[3 * (size_t)j + k];
    return q;
  };
  if (total <= kListCapacity) {
    const uint32_t kl = sub % LW, d0 = sub / LW;
    const uint32_t nd = total > kl ? (total - 1 - kl) / LW + 1 : 0;  // depths of list lane kl
    const uint32_t cnt = nd > d0 ? (nd - 1 - d0) / M + 1 : 0;
    const uint32_t *__restrict__ nbr =
        b.nbr + lane_block<LW>(s, kListCapacity) + d0 * 64 + kl;
    constexpr uint32_t kStep = 64 * M;
    uint32_t k = 0;
    if (cnt >= U) {
      uint32_t jn[U];
#pragma unroll
      for (int q = 0; q < U; ++q) jn[q] = nbr[q * kStep];
      nbr += U * kStep;
      k = U;
      while (true) {
        uint32_t jc[U];
#pragma unroll
        for (int q = 0; q < U; ++q) jc[q] = jn[q];
        const bool more = k + U <= cnt;
        if (more) {
#pragma unroll
          for (int q = 0; q < U; ++q) jn[q] = nbr[q * kStep];
          nbr += U * kStep;
          k += U;
        }
        NbrRec<NR> p[U];
#pragma unroll
        for (int q = 0; q < U; ++q) p[q] = load(jc[q]);
#pragma unroll
        for (int q = 0; q < U; ++q) f(jc[q], p[q]);
        if (!more) break;
      }
    }
    // the remaining (fewer than U) pairs as one predicated chunk: their list
    // entries, then their records, each in one round trip
    if (k < cnt) {
      uint32_t jt[U];
#pragma unroll
      for (int q = 0; q < U; ++q) jt[q] = k + q < cnt ? nbr[q * kStep] : s;
      NbrRec<NR> p[U];
#pragma unroll
      for (int q = 0; q < U; ++q) p[q] = load(jt[q]);
#pragma unroll
      for (int q = 0; q < U; ++q)
        if (k + q < cnt) f(jt[q], p[q]);
    }
    return;
  }
  // overflow: sweep the records themselves (part 0 is the position the
  // density pass tested), the parts a pair needs loaded with it
  const float h2 = Un->h2;
  const int cx = cell_coord(pi.x, Un->ox, Un->inv_cell);
  const int cy = cell_coord(pi.y, Un->oy, Un->inv_cell);
  const int cz = cell_coord(pi.z, Un->oz, Un->inv_cell);
  const SlotRec *__restrict__ recs = reinterpret_cast<const SlotRec *>(srec);
  // candidates per trip: as many parts in flight as four positions and
  // velocities (two whole records with all three parts)
  constexpr int UN = NR == 3 ? 2 : 4;
  sweep27<L, false, UN>(Un, b.counts, b.starts, recs, cx, cy, cz, sub,
                        [&](uint32_t j, const SlotRec &c, bool) {
    NbrRec<NR> q;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
      // in registers before the test: the compiler would otherwise move the
      // load of the parts only the pair uses behind it
      q.r[k] = c.r[k];
      asm volatile("" : "+v"(q.r[k].x), "+v"(q.r[k].y), "+v"(q.r[k].z), "+v"(q.r[k].w));
    }
    const float4 pj = q.r[0];
    const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
    const float r2 = dot3(rx, ry, rz, rx, ry, rz);
    if (r2 < h2 && r2 > 1e-12f) f(j, q);
  });
}

}  // namespace fluid
