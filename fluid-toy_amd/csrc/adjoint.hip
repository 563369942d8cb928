// adjoint.hip — reverse mode (vector-Jacobian products) of one SPH step.
//
// Given the adjoint of the state a step produced (dL/dx, dL/dv per slot of
// that step's sorted order) and the step's tape (its sorted inputs and the
// intermediates the forward kept), these kernels produce the adjoint of the
// state the step consumed and accumulate the gradients of viscosity, surface
// tension and gravity.  The grid, the neighbour sets and the sort order are
// the tape's: they are piecewise constant in the state and are not
// differentiated.  Every kernel is a gather: particle i sums the terms of the
// pairs (i, j) of its own outputs and of the pairs (j, i) in which it is the
// neighbour, so there are no floating-point atomics and the result is
// bitwise reproducible.
//
// Forward maps reversed (the WGSL the step kernels restate):
//   integrate  sph_integrate_3d.wgsl:91-388 (walls, Euler, backstop)
//   finalize   pcisph_finalize.wgsl:50-74
//   solve      pcisph_solve.wgsl:172-208 (one iteration, CFL clamp)
//   predict    pcisph_predict.wgsl:50-73
//   force      sph_force_3d_grid.wgsl:105-227
//   xsph       xsph.wgsl:81-138
//   density    sph_density_3d_grid.wgsl:111-194 (+ boundary_gamma :86-109)
// Mouse forces and rigid bodies are not differentiated; the host refuses a
// rollout while either is active.
#include "adjoint.h"
#include "sph_device.h"

namespace fluid {
namespace {

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 v3(float4 a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ float4 f4(V3 a) { return make_float4(a.x, a.y, a.z, 0.f); }
__device__ __forceinline__ float4 f4(V3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
__device__ __forceinline__ V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 scl(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
// a + s b
__device__ __forceinline__ V3 axpy(V3 a, float s, V3 b) {
  return v3(__builtin_fmaf(s, b.x, a.x), __builtin_fmaf(s, b.y, a.y),
            __builtin_fmaf(s, b.z, a.z));
}
__device__ __forceinline__ float dotv(V3 a, V3 b) { return dot3(a.x, a.y, a.z, b.x, b.y, b.z); }
__device__ __forceinline__ V3 zero3() { return V3{0.f, 0.f, 0.f}; }
// m^T a for a row-major 3x3 (rows padded to 4)
__device__ __forceinline__ V3 mat_tv(const float (*m)[4], V3 a) {
  return v3(dot3(m[0][0], m[1][0], m[2][0], a.x, a.y, a.z),
            dot3(m[0][1], m[1][1], m[2][1], a.x, a.y, a.z),
            dot3(m[0][2], m[1][2], m[2][2], a.x, a.y, a.z));
}
template <int L>
__device__ __forceinline__ V3 group_sum3(V3 a) {
  return v3(group_sum<L>(a.x), group_sum<L>(a.y), group_sum<L>(a.z));
}

__device__ __forceinline__ void cell_of(const DevUniforms *U, float4 p, int &cx, int &cy,
                                        int &cz) {
  cx = cell_coord(p.x, U->ox, U->inv_cell);
  cy = cell_coord(p.y, U->oy, U->inv_cell);
  cz = cell_coord(p.z, U->oz, U->inv_cell);
}

// Reverse of the CFL clamp a -> a (max / |a|) applied when |a|^2 > max^2.
__device__ __forceinline__ V3 clamp_vjp(const DevUniforms *U, V3 ap, V3 lout) {
  const float am2 = dotv(ap, ap);
  if (!(am2 > U->max_accel_sq)) return lout;
  const float am = sqrtf(am2);
  const V3 nh = scl(1.0f / am, ap);
  return scl(U->max_accel / am, axpy(lout, -dotv(nh, lout), nh));
}

// Reverse of one PCISPH iteration's per-particle epilogue,
//   V = w + dt (a + clamp(ap)),  X = x + dt V,  P' = P + delta max(0, rho* - rho0),
// given lam_X (adjoint of X), lam_V (adjoint reaching V from finalize, last
// iteration only), lam_rho (adjoint of rho* through the next q) and lam_P
// (adjoint of P').  Accumulates into x, w, a and returns (adjoint of the
// unclamped ap, total adjoint of rho*).
__device__ __forceinline__ float4 solve_epilogue_vjp(const DevUniforms *U, V3 lam_X, V3 lam_V,
                                                     float lam_rho, float lam_P,
                                                     float rho_pred, V3 ap, V3 &lx, V3 &lw,
                                                     V3 &la) {
  const float dt = U->dt;
  const V3 lv = axpy(lam_V, dt, lam_X);
  lx = vadd(lx, lam_X);
  lw = vadd(lw, lv);
  la = axpy(la, dt, lv);
  const V3 lap = clamp_vjp(U, ap, scl(dt, lv));
  const float lr = lam_rho + ((rho_pred - U->rho0) > 0.0f ? U->delta * lam_P : 0.0f);
  return f4(lap, lr);
}

// ------------------------------------------------------------------------
// Canonical (original index, 3 floats) <-> slot order.
__global__ __launch_bounds__(256) void k_adj_load(const uint32_t *__restrict__ sid,
                                                  const float *__restrict__ dx,
                                                  const float *__restrict__ dv,
                                                  float4 *__restrict__ lx,
                                                  float4 *__restrict__ lv, uint32_t n) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const size_t i = 3 * (size_t)sid[s];
  lx[s] = dx ? make_float4(dx[i], dx[i + 1], dx[i + 2], 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
  lv[s] = dv ? make_float4(dv[i], dv[i + 1], dv[i + 2], 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(256) void k_adj_store(const uint32_t *__restrict__ sid,
                                                   const float4 *__restrict__ lx,
                                                   const float4 *__restrict__ lv,
                                                   float *__restrict__ ox,
                                                   float *__restrict__ ov, uint32_t n) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const size_t i = 3 * (size_t)sid[s];
  if (ox) {
    const float4 a = lx[s];
    ox[i] = a.x;
    ox[i + 1] = a.y;
    ox[i + 2] = a.z;
  }
  if (ov) {
    const float4 a = lv[s];
    ov[i] = a.x;
    ov[i + 1] = a.y;
    ov[i + 2] = a.z;
  }
}

// ------------------------------------------------------------------------
// Integrate (and finalize's v <- V_K) reversed, then the epilogue of the
// last PCISPH iteration.  The forward expressions are re-evaluated as
// k_integrate writes them, so every branch (wall layer, approach direction,
// backstop) takes the forward's side.
template <bool B>
__global__ __launch_bounds__(256) void k_adj_integrate(Uni un, AdjBufs a, TapeStep t,
                                                       TapeIter last, uint32_t n) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const DevUniforms *__restrict__ U = uni_at<B>(un, s);
  const float(*iv)[4] = U->iv;
  const float(*fw)[4] = U->fw;
  const float4 p4 = t.x[s], v4 = t.vK[s];
  const float px = p4.x, py = p4.y, pz = p4.z;
  const float vx0 = v4.x, vy0 = v4.y, vz0 = v4.z;
  const float qy = py - U->cy;
  const float lp[3] = {dot3(iv[0][0], iv[0][1], iv[0][2], px, qy, pz),
                       dot3(iv[1][0], iv[1][1], iv[1][2], px, qy, pz),
                       dot3(iv[2][0], iv[2][1], iv[2][2], px, qy, pz)};
  const float lvl[3] = {dot3(iv[0][0], iv[0][1], iv[0][2], vx0, vy0, vz0),
                        dot3(iv[1][0], iv[1][1], iv[1][2], vx0, vy0, vz0),
                        dot3(iv[2][0], iv[2][1], iv[2][2], vx0, vy0, vz0)};
  const float bl = U->boundary_layer, k = U->wall_k, wd = 8.0f;
  const float half[3] = {U->hw, U->hh, U->hd};
  float wa[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float d = lp[c] - (-half[c]);
    if (d < bl) {
      const float tt = 1.0f - d / bl;
      wa[c] += k * tt * tt;
      if (lvl[c] < 0.f) wa[c] -= lvl[c] * wd * tt;
    }
    d = half[c] - lp[c];
    if (d < bl) {
      const float tt = 1.0f - d / bl;
      wa[c] -= k * tt * tt;
      if (lvl[c] > 0.f) wa[c] -= lvl[c] * wd * tt;
    }
  }
  const float ax = 0.f + dot3(fw[0][0], fw[0][1], fw[0][2], wa[0], wa[1], wa[2]);
  const float ay = 0.f + dot3(fw[1][0], fw[1][1], fw[1][2], wa[0], wa[1], wa[2]);
  const float az = 0.f + dot3(fw[2][0], fw[2][1], fw[2][2], wa[0], wa[1], wa[2]);
  const float dt = U->dt;
  const float vx = vx0 + dt * ax, vy = vy0 + dt * ay, vz = vz0 + dt * az;
  const float nx = px + dt * vx, ny = py + dt * vy, nz = pz + dt * vz;
  const float qy5 = ny - U->cy;
  const float nl[3] = {dot3(iv[0][0], iv[0][1], iv[0][2], nx, qy5, nz),
                       dot3(iv[1][0], iv[1][1], iv[1][2], nx, qy5, nz),
                       dot3(iv[2][0], iv[2][1], iv[2][2], nx, qy5, nz)};
  const float mv[3] = {dot3(iv[0][0], iv[0][1], iv[0][2], vx, vy, vz),
                       dot3(iv[1][0], iv[1][1], iv[1][2], vx, vy, vz),
                       dot3(iv[2][0], iv[2][1], iv[2][2], vx, vy, vz)};
  // reverse; the outputs are fw (backstopped local) (+ cy)
  const V3 gp = mat_tv(fw, v3(a.lin_x[s])), gv = mat_tv(fw, v3(a.lin_v[s]));
  float g_nl[3] = {gp.x, gp.y, gp.z}, g_mv[3] = {gv.x, gv.y, gv.z};
  const float rest = U->damping;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float sg = mv[c] > 0.f ? 1.f : (mv[c] < 0.f ? -1.f : 0.f);
    if (nl[c] < -half[c]) {
      g_nl[c] = 0.f;  // position pinned to the wall, velocity |v| rest
      g_mv[c] = g_mv[c] * rest * sg;
    } else if (nl[c] > half[c]) {
      g_nl[c] = 0.f;  // -|v| rest
      g_mv[c] = -g_mv[c] * rest * sg;
    }
  }
  const V3 g_n = mat_tv(iv, v3(g_nl[0], g_nl[1], g_nl[2]));                // new position
  const V3 g_v = axpy(mat_tv(iv, v3(g_mv[0], g_mv[1], g_mv[2])), dt, g_n);  // new velocity
  V3 lx = g_n;
  V3 lvk = g_v;
  const V3 g_wa = mat_tv(fw, scl(dt, g_v));
  const float gwa[3] = {g_wa.x, g_wa.y, g_wa.z};
  float g_lp[3] = {0.f, 0.f, 0.f}, g_lv[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float d = lp[c] - (-half[c]);
    if (d < bl) {
      const float tt = 1.0f - d / bl;
      float gt = gwa[c] * 2.0f * k * tt;
      if (lvl[c] < 0.f) {
        gt -= gwa[c] * lvl[c] * wd;
        g_lv[c] -= gwa[c] * wd * tt;
      }
      g_lp[c] -= gt / bl;
    }
    d = half[c] - lp[c];
    if (d < bl) {
      const float tt = 1.0f - d / bl;
      float gt = -gwa[c] * 2.0f * k * tt;
      if (lvl[c] > 0.f) {
        gt -= gwa[c] * lvl[c] * wd;
        g_lv[c] -= gwa[c] * wd * tt;
      }
      g_lp[c] += gt / bl;
    }
  }
  lx = vadd(lx, mat_tv(iv, v3(g_lp[0], g_lp[1], g_lp[2])));
  lvk = vadd(lvk, mat_tv(iv, v3(g_lv[0], g_lv[1], g_lv[2])));
  // the last PCISPH iteration: its X, q and P outputs are not read again
  V3 lw = zero3(), la = zero3();
  const V3 ap = last.ap ? v3(last.ap[s]) : zero3();
  const float4 lap =
      solve_epilogue_vjp(U, zero3(), lvk, 0.f, 0.f, last.pr[s].y, ap, lx, lw, la);
  a.lx[s] = f4(lx);
  a.lw[s] = f4(lw);
  a.la[s] = f4(la);
  a.lP[s] = 0.f;
  a.lap[0][s] = lap;
}

// ------------------------------------------------------------------------
// PCISPH iteration k reversed (pcisph_solve.wgsl:172-208).  Forward, with X
// and q this iteration's inputs:
//   rho*_i = sum_j m W(|X_i - X_j|^2)                         (r^2 < h^2)
//   ap_i   = sum_j m (q_i + q_j) gs (h - r)^2 / r (X_i - X_j)   (0 < r^2 < h^2)
// lap_in holds every particle's (adjoint of ap, of rho*).  The pairs give
// the adjoint of X_k and q_k; the epilogue takes q_k = P_k / max(rho*_{k-1},
// 1)^2 back to P_k and rho*_{k-1} and reverses iteration k-1's epilogue
// (k > 0) or predict (k = 0).
//
// mode 1 (the first reversed iteration of K >= 2) sweeps the 27 cells and
// records, per lane, the candidates within the solve skin of the particle at
// the first forward iteration's predicted positions xq0 (every iteration's
// in-range pairs lie among them); mode 2 (the other reversed iterations)
// walks that list; mode 0 (K = 1) only sweeps.
template <int L, bool B>
__global__ __launch_bounds__(256) void k_adj_solve(Uni un, AdjBufs a, TapeStep t, uint32_t n,
                                                   TapeIter in, const float4 *__restrict__ lap_in,
                                                   float4 *__restrict__ lap_out, uint32_t k,
                                                   const float4 *__restrict__ xq0, uint32_t mode) {
  __shared__ RowEnt rows[256 / L][9];
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t s = g / L, sub = g % L;
  if (s >= n) return;
  const DevUniforms *__restrict__ U = uni_at<B>(un, s);
  const float4 xi = in.xq[s];
  const float4 li = lap_in[s];
  int cx, cy, cz;
  cell_of(U, t.x[s], cx, cy, cz);
  const float h = U->h, h2 = U->h2, mass = U->mass, gs = U->grad_scale;
  const float c_rho = mass * (-6.0f * U->poly6_scale);  // 2 d(m W)/d(r^2) = c_rho d^2
  float lXx = 0.f, lXy = 0.f, lXz = 0.f, lq = 0.f;
  uint32_t *__restrict__ lst = a.slist + lane_block<L>(s, kSolveListCapacity) + sub;
  constexpr uint32_t kCap = kSolveListCapacity / L;
  auto pair = [&](uint32_t j, float4 xj, bool) {
    const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
    const float r2 = dot3(rx, ry, rz, rx, ry, rz);
    if (!(r2 < h2)) return;
    const float4 lj = lap_in[j];
    const float d = h2 - r2;
    float cr = (li.w + lj.w) * c_rho * d * d;  // coefficient of r_vec
    if (r2 > 1e-12f) {
      const float ir = __builtin_amdgcn_rsqf(r2);
      const float r = r2 * ir, dr = h - r;
      const float dlx = li.x - lj.x, dly = li.y - lj.y, dlz = li.z - lj.z;
      const float rdl = dot3(rx, ry, rz, dlx, dly, dlz);
      const float c1 = mass * gs * dr * dr * ir;  // the pair vector is c1 r_vec
      lq += c1 * rdl;
      const float qq = xi.w + xj.w;
      const float c2 = mass * gs * dr * (h + r) * ir * ir * ir;
      const float qc1 = qq * c1;
      cr -= qq * c2 * rdl;
      lXx += qc1 * dlx;
      lXy += qc1 * dly;
      lXz += qc1 * dlz;
    }
    lXx += cr * rx;
    lXy += cr * ry;
    lXz += cr * rz;
  };
  // The sweeps walk the 27-cell stencil as one concatenated run of its nine
  // x-rows; the row table is built (by the whole group) only where a lane
  // of the group sweeps.
  const uint32_t cntk = mode == 2 ? a.scnt[(size_t)s * L + sub] : kSolveListOverflow;
  RowEnt *tab = rows[threadIdx.x / L];
  const uint32_t total =
      group_any<L>(cntk == kSolveListOverflow) ? rows27<L>(U, t.counts, t.starts, cx, cy, cz, sub, tab)
                                               : 0u;
  auto sweep = [&](auto &&f) {
    walk_rows<L, false>([&](uint32_t r) { return tab[r]; }, total, in.xq, sub, s, f);
  };
  if (mode == 1) {
    const float4 x0i = xq0[s];
    const float skin2 = U->solve_skin2;
    uint32_t cnt = 0;
    sweep([&](uint32_t j, float4 xj, bool v) {
      const float4 x0j = xq0[j];
      const float ex = x0i.x - x0j.x, ey = x0i.y - x0j.y, ez = x0i.z - x0j.z;
      if (dot3(ex, ey, ez, ex, ey, ez) < skin2) {
        if (cnt < kCap) lst[cnt * 64u] = j;
        ++cnt;
      }
      pair(j, xj, v);
    });
    a.scnt[(size_t)s * L + sub] = cnt <= kCap ? cnt : kSolveListOverflow;
  } else {
    const uint32_t cnt = cntk;
    if (cnt != kSolveListOverflow) {
      uint32_t e = 0;
      for (; e + 3 < cnt; e += 4) {
        const uint32_t j0 = lst[e * 64u], j1 = lst[e * 64u + 64u];
        const uint32_t j2 = lst[e * 64u + 128u], j3 = lst[e * 64u + 192u];
        const float4 x0 = in.xq[j0], x1 = in.xq[j1], x2 = in.xq[j2], x3 = in.xq[j3];
        pair(j0, x0, true);
        pair(j1, x1, true);
        pair(j2, x2, true);
        pair(j3, x3, true);
      }
      for (; e < cnt; ++e) {
        const uint32_t j = lst[e * 64u];
        pair(j, in.xq[j], true);
      }
    } else {
      sweep(pair);
    }
  }
  V3 lX = group_sum3<L>(v3(lXx, lXy, lXz));
  lq = group_sum<L>(lq);
  if (sub != 0) return;
  V3 lx = v3(a.lx[s]), lw = v3(a.lw[s]), la = v3(a.la[s]);
  if (k > 0) {
    const float2 pr = in.pr[s];  // (P_k, rho*_{k-1})
    const float rb = fmaxf(pr.y, 1.0f);
    const float irb2 = 1.0f / (rb * rb);
    const float lP = a.lP[s] + lq * irb2;
    const float lrho = pr.y > 1.0f ? -2.0f * pr.x * lq * irb2 / rb : 0.f;
    const V3 ap = in.ap ? v3(in.ap[s]) : zero3();
    lap_out[s] = solve_epilogue_vjp(U, lX, zero3(), lrho, lP, pr.y, ap, lx, lw, la);
    a.lP[s] = lP;
  } else {
    // predict (pcisph_predict.wgsl:57-72): X_0 = x + dt (w + dt a)
    const float dt = U->dt;
    const V3 lv = scl(dt, lX);
    lx = vadd(lx, lX);
    lw = vadd(lw, lv);
    la = axpy(la, dt, lv);
  }
  a.lx[s] = f4(lx);
  a.lw[s] = f4(lw);
  a.la[s] = f4(la);
}

// ------------------------------------------------------------------------
// Force (sph_force_3d_grid.wgsl:105-227).  The term owner A's acceleration
// receives from neighbour B is
//   T = (S1 g + sep + sv + st m K C) dir - st m K (n_A - n_B)
// with r_vec = x_A - x_B, g = gs (h-r)^2, S1 = -m k_near (nd_A + nd_B) /
// (2 d_A nd_B), sep = -200 (1 - r/md) when r < md, sv = visc mu g m / avg
// when v.r_vec < 0 (v = w_A - u_B), mu = h v.r_vec / (r^2 + eps), avg =
// (d_A + d_B)/2, K = 2 rho0 / (d_A + d_B), C the Akinci cohesion spline,
// dir = -r_vec/r (or the close direction), d = max(rho, 1), nd =
// max(rho_near, 1).
struct PairSide {
  float4 x;     // position
  float4 v;     // velocity (w for the owner, u for the neighbour)
  float4 nrm;   // normal
  float d, nd;  // density, near density (before the max with 1)
  uint32_t slot;
};
struct PairGrad {
  V3 x_a;  // the neighbour's is -x_a
  V3 v_a, n_a, v_b, n_b;
  float d_a, nd_a, d_b, nd_b;
  float visc, st;
};

// What the two terms of a pair (A's from B and B's from A) share: every
// function of the distance alone.  r_vec flips sign between them, r^2 (a
// sum of squares) keeps its bits.
struct PairGeom {
  float r2, ir, r, dr, gk, iden, dr2, r3, C, dCdr;
  bool close, outer;
};

__device__ __forceinline__ PairGeom pair_geom(const DevUniforms *__restrict__ U, float rx,
                                              float ry, float rz) {
  PairGeom q;
  q.r2 = dot3(rx, ry, rz, rx, ry, rz);
  q.close = q.r2 < U->close_r2;
  q.ir = __builtin_amdgcn_rsqf(q.r2);
  q.r = q.r2 * q.ir;
  q.dr = U->h - q.r;
  q.gk = U->grad_scale * q.dr * q.dr;
  q.iden = 1.0f / (q.r2 + U->visc_eps);
  q.dr2 = q.dr * q.dr;
  q.r3 = q.r2 * q.r;
  q.outer = q.r >= U->half_h;
  const float t3 = q.dr2 * q.dr * q.r3;  // 2 dr3 r3 = 2 t3 exactly
  const float aks = U->akinci_scale;
  q.C = aks * (q.outer ? t3 : __builtin_fmaf(2.0f, t3, -U->h6_64));
  q.dCdr = aks * 3.0f * q.dr2 * q.r2 * (q.dr - q.r) * (q.outer ? 1.0f : 2.0f);
  return q;
}

// The reverse of A's term from B; (rx, ry, rz) = x_A - x_B and q its
// geometry.
template <bool NEED_A, bool NEED_B>
__device__ __forceinline__ void force_pair_vjp(const DevUniforms *__restrict__ U,
                                               const PairSide &A, const PairSide &Bn, V3 lam,
                                               PairGrad &G, const PairGeom &q, float rx, float ry,
                                               float rz) {
  const float h = U->h, mass = U->mass, md = U->min_dist;
  const bool close = q.close;
  const float ir = q.ir;
  const float r = q.r;
  V3 dir = v3(-rx * ir, -ry * ir, -rz * ir);
  if (close) {  // rare: the slot-difference direction
    const float3 c = close_dir(A.slot, Bn.slot);
    dir = v3(c.x, c.y, c.z);
  }
  const float da = max1(A.d), nda = max1(A.nd);
  const float db = max1(Bn.d), ndb = max1(Bn.nd);
  const float kn = U->near_k;
  const float dr = q.dr, gs = U->grad_scale;
  const float gk = q.gk;
  const float i2dn = 1.0f / (2.0f * da * ndb);
  const float S1 = -mass * kn * (nda + ndb) * i2dn;
  const float vx = A.v.x - Bn.v.x, vy = A.v.y - Bn.v.y, vz = A.v.z - Bn.v.z;
  const float vdx = dot3(vx, vy, vz, rx, ry, rz);
  const bool von = vdx < 0.f;
  const float iden = q.iden;
  const float mu = h * vdx * iden;
  const float idd = 1.0f / (da + db);
  const float visc = U->visc, st = U->st;
  const float cv = 2.0f * mass * idd;  // m / avg
  const float sv = von ? visc * mu * gk * cv : 0.f;
  const float K = 2.0f * U->rho0 * idd;
  const float C = q.C;
  const float smk = st * mass * K;
  const float coef = S1 * gk + (close ? -200.0f * (1.0f - r / md) : 0.f) + sv + smk * C;
  const float dnx = A.nrm.x - Bn.nrm.x, dny = A.nrm.y - Bn.nrm.y, dnz = A.nrm.z - Bn.nrm.z;
  // reverse
  const float lA = dotv(lam, dir);
  const float ldn = dot3(lam.x, lam.y, lam.z, dnx, dny, dnz);
  if (NEED_A) G.n_a = scl(-smk, lam);
  if (NEED_B) G.n_b = scl(smk, lam);
  const float lK = st * mass * (lA * C - ldn);
  if (NEED_A) {
    G.st = mass * K * (lA * C - ldn);
    G.visc = von ? lA * mu * gk * cv : 0.f;
  }
  float lg = lA * S1;
  float lr = close ? lA * 200.0f / md : 0.f;
  // the viscosity terms through selects (sv = 0 when !von)
  const float lmu = von ? lA * visc * gk * cv : 0.f;
  lg += von ? lA * visc * mu * cv : 0.f;
  const float ldd = -lK * K * idd - lA * sv * idd;  // d_A, d_B enter K and avg
  const float lvdx = lmu * h * iden;
  float lr2 = -(lmu * mu * iden);
  const float lS1 = lA * gk;
  const float lsum = -lS1 * mass * kn * i2dn;  // through the numerator nd_A + nd_B
  if (NEED_A) {
    G.d_a = A.d > 1.0f ? ldd - lS1 * S1 / da : 0.f;
    G.nd_a = A.nd > 1.0f ? lsum : 0.f;
    G.v_a = v3(lvdx * rx, lvdx * ry, lvdx * rz);
  }
  if (NEED_B) {
    G.d_b = Bn.d > 1.0f ? ldd : 0.f;
    G.nd_b = Bn.nd > 1.0f ? lsum - lS1 * S1 / ndb : 0.f;
    G.v_b = v3(-lvdx * rx, -lvdx * ry, -lvdx * rz);
  }
  // dependence on r: g, C, sep; on r_vec: dir and v.r_vec
  lr += lA * smk * q.dCdr - lg * 2.0f * gs * dr;
  float gx = lvdx * vx, gy = lvdx * vy, gz = lvdx * vz;
  {
    // dir = -r_vec / r:  d dir / d r_vec = -(I - rh rh^T) / r; the close
    // direction does not depend on r_vec (zero scale)
    const float cd = close ? 0.f : coef;
    const float ldx = cd * lam.x, ldy = cd * lam.y, ldz = cd * lam.z;
    const float pr = dot3(rx, ry, rz, ldx, ldy, ldz) * ir * ir;
    gx -= (ldx - pr * rx) * ir;
    gy -= (ldy - pr * ry) * ir;
    gz -= (ldz - pr * rz) * ir;
  }
  lr2 += lr * 0.5f * ir;
  gx += 2.0f * lr2 * rx;
  gy += 2.0f * lr2 * ry;
  gz += 2.0f * lr2 * rz;
  G.x_a = v3(gx, gy, gz);
}

// The in-range pairs of this lane: its share (entries sub, sub + L, ...) of
// the balanced list the force adjoint recorded (total entries: pcnt), or
// (after an overflow) the lane's candidates of the 27-cell sweep with the
// same test.  f(j, x_j).
template <int L, class F>
__device__ __forceinline__ void for_each_pair(const DevUniforms *__restrict__ U, const AdjBufs &a,
                                              const TapeStep &t, uint32_t n, uint32_t s,
                                              uint32_t sub, float4 xi, uint32_t total, F &&f) {
  if (total <= kListCapacity) {
    const uint32_t cnt = total > sub ? (total - 1 - sub) / L + 1 : 0;
    const uint32_t *__restrict__ lst = a.plist + lane_block<L>(s, kListCapacity) + sub;
    uint32_t e = 0;
    for (; e + 1 < cnt; e += 2) {
      const uint32_t j0 = lst[e * 64u], j1 = lst[e * 64u + 64u];
      const float4 x0 = t.x[j0], x1 = t.x[j1];
      f(j0, x0);
      f(j1, x1);
    }
    if (e < cnt) {
      const uint32_t j0 = lst[e * 64u];
      f(j0, t.x[j0]);
    }
    return;
  }
  int cx, cy, cz;
  cell_of(U, xi, cx, cy, cz);
  const float h2 = U->h2;
  sweep27<L, false>(U, t.counts, t.starts, t.x, cx, cy, cz, sub,
                    [&](uint32_t j, float4 xj, bool) {
    const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
    const float r2 = dot3(rx, ry, rz, rx, ry, rz);
    if (r2 < h2 && r2 > 1e-12f) f(j, xj);
  });
}

template <int L, bool B>
__global__ __launch_bounds__(256) void k_adj_force(Uni un, AdjBufs a, TapeStep t, uint32_t n) {
  __shared__ RowEnt rows[256 / L][9];
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t s = g / L, sub = g % L;
  if (s >= n) return;
  const DevUniforms *__restrict__ U = uni_at<B>(un, s);
  const float4 *ri = t.rec + 3 * (size_t)s;
  const float4 xi = t.x[s];
  const float4 rec0 = ri[0], rec1 = ri[1], rec2 = ri[2];
  const PairSide own{xi, t.w[s], rec2, rec0.w, rec1.w, s};
  const PairSide asn{xi, rec1, rec2, rec0.w, rec1.w, s};  // i as a neighbour: u_i
  const V3 lai = v3(a.la[s]);
  int cx, cy, cz;
  cell_of(U, xi, cx, cy, cz);
  const float h2 = U->h2;
  V3 gx = zero3(), gw = zero3(), gu = zero3(), gn = zero3();
  float gd = 0.f, gnd = 0.f, gv = 0.f, gst = 0.f;
  // Pass 1: the in-range pairs into a ballot-balanced list (a cheap
  // group-synchronous sweep); pass 2 walks it, so the heavy per-pair body
  // runs with the group's lanes evenly loaded instead of on the sweep's
  // in-range lanes only.
  // Every lane of the group ends pass 1 with the group's count in `run`, and
  // pass 2 uses that, not pcnt (lane 0's store).  Pass 2 reads entries other
  // lanes of the wave stored in pass 1: the workgroup fence between the
  // passes waits for those stores and keeps the loads after them.
  // The sweep walks the 27-cell stencil as one concatenated run of its nine
  // x-rows (the forward density's walk).
  uint32_t run = 0;
  const PairTest in_pair(h2);
  RowEnt *tab = rows[threadIdx.x / L];
  const uint32_t total = rows27<L>(U, t.counts, t.starts, cx, cy, cz, sub, tab);
  walk_rows<L, true>([&](uint32_t r) { return tab[r]; }, total, t.x, sub, s,
                     [&](uint32_t j, float4 xj, bool valid) {
    const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
    const float r2 = dot3(rx, ry, rz, rx, ry, rz);
    group_list_append<L, L>(a.plist + lane_block<L>(s, kListCapacity), sub, run,
                         valid && in_pair(r2), j);
  });
  if (sub == 0) a.pcnt[s] = run;
  __threadfence_block();
  for_each_pair<L>(U, a, t, n, s, sub, xi, run, [&](uint32_t j, float4 xj) {
    const float4 *rj = t.rec + 3 * (size_t)j;
    const float4 rj0 = rj[0], rj1 = rj[1], rj2 = rj[2];
    PairGrad G;
    const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
    const PairGeom q = pair_geom(U, rx, ry, rz);
    // (i, j): the term of a_i
    const PairSide nb{xj, rj1, rj2, rj0.w, rj1.w, j};
    force_pair_vjp<true, false>(U, own, nb, lai, G, q, rx, ry, rz);
    gx = vadd(gx, G.x_a);
    gw = vadd(gw, G.v_a);
    gn = vadd(gn, G.n_a);
    gd += G.d_a;
    gnd += G.nd_a;
    gv += G.visc;
    gst += G.st;
    // (j, i): the term of a_j that reads particle i
    const PairSide oj{xj, t.w[j], rj2, rj0.w, rj1.w, j};
    force_pair_vjp<false, true>(U, oj, asn, v3(a.la[j]), G, q, -rx, -ry, -rz);
    gx = vsub(gx, G.x_a);
    gu = vadd(gu, G.v_b);
    gn = vadd(gn, G.n_b);
    gd += G.d_b;
    gnd += G.nd_b;
  });
  gx = group_sum3<L>(gx);
  gw = group_sum3<L>(gw);
  gu = group_sum3<L>(gu);
  gn = group_sum3<L>(gn);
  gd = group_sum<L>(gd);
  gnd = group_sum<L>(gnd);
  gv = group_sum<L>(gv);
  gst = group_sum<L>(gst);
  if (sub != 0) return;
  a.lx[s] = f4(vadd(v3(a.lx[s]), gx));
  a.lw[s] = f4(vadd(v3(a.lw[s]), gw));
  a.lu[s] = f4(gu);
  a.ln[s] = f4(gn);
  a.lrho[s] = make_float2(gd, gnd);
  double *acc = a.acc + (size_t)kAccStride * s;
  acc[0] += (double)gv;
  acc[1] += (double)gst;
  acc[2] += (double)lai.x;  // the force adds gravity to every particle
  acc[3] += (double)lai.y;
  acc[4] += (double)lai.z;
}

// ------------------------------------------------------------------------
// XSPH (xsph.wgsl:81-138): w_i = u_i + eps sum_j m W(r^2) / max(rho_j, 1)
// (u_j - u_i).  The epilogue reverses the density stage's rho = S / gamma(x)
// for this particle (its density adjoint is complete here: force and XSPH
// are its only readers).
template <int L, bool B>
__global__ __launch_bounds__(256) void k_adj_xsph(Uni un, AdjBufs a, TapeStep t, uint32_t n) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t s = g / L, sub = g % L;
  if (s >= n) return;
  const DevUniforms *__restrict__ U = uni_at<B>(un, s);
  const float eps = U->eps;
  const float4 xi = t.x[s];
  const float4 ri0 = t.rec[3 * (size_t)s];
  const float4 ui = t.rec[3 * (size_t)s + 1];
  const V3 lwi = v3(a.lw[s]);
  float gxx = 0.f, gxy = 0.f, gxz = 0.f, gux = 0.f, guy = 0.f, guz = 0.f, gd = 0.f;
  if (eps > 0.0f) {
    const float h2 = U->h2, mass = U->mass, ps = U->poly6_scale;
    const float idib = 1.0f / max1(ri0.w);
    for_each_pair<L>(U, a, t, n, s, sub, xi, a.pcnt[s], [&](uint32_t j, float4 xj) {
      const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
      const float r2 = dot3(rx, ry, rz, rx, ry, rz);
      const float4 rj0 = t.rec[3 * (size_t)j], uj = t.rec[3 * (size_t)j + 1];
      const float4 lwj = a.lw[j];
      const float d = h2 - r2;
      const float mW = mass * ps * d * d * d;
      const float mdW = mass * (-3.0f * ps) * d * d;  // d(m W)/d(r^2)
      const float idjb = 1.0f / max1(rj0.w);
      const float dux = uj.x - ui.x, duy = uj.y - ui.y, duz = uj.z - ui.z;
      // (i, j): eps m W / d_j (u_j - u_i) in w_i
      const float e1 = eps * dot3(lwi.x, lwi.y, lwi.z, dux, duy, duz);
      const float f_ij = eps * mW * idjb;
      // (j, i): eps m W / d_i (u_i - u_j) in w_j
      const float e2 = -eps * dot3(lwj.x, lwj.y, lwj.z, dux, duy, duz);
      const float f_ji = eps * mW * idib;
      gux += f_ji * lwj.x - f_ij * lwi.x;
      guy += f_ji * lwj.y - f_ij * lwi.y;
      guz += f_ji * lwj.z - f_ij * lwi.z;
      gd -= e2 * mW * idib * idib;
      const float c = 2.0f * mdW * (e1 * idjb + e2 * idib);
      gxx += c * rx;
      gxy += c * ry;
      gxz += c * rz;
    });
    gxx = group_sum<L>(gxx);
    gxy = group_sum<L>(gxy);
    gxz = group_sum<L>(gxz);
    gux = group_sum<L>(gux);
    guy = group_sum<L>(guy);
    guz = group_sum<L>(guz);
    gd = group_sum<L>(gd);
    if (!(ri0.w > 1.0f)) gd = 0.f;
  }
  if (sub != 0) return;
  V3 lx = vadd(v3(a.lx[s]), v3(gxx, gxy, gxz));
  a.lu[s] = f4(vadd(vadd(v3(a.lu[s]), lwi), v3(gux, guy, guz)));
  // density epilogue: rho = S / gamma, gamma = max(prod_w f_w(x), 0.1)
  const float2 lr = a.lrho[s];
  const float lrho = lr.x + gd;
  const float h = U->h;
  const float px = xi.x, py = xi.y - U->cy, pz = xi.z;
  const float(*iv)[4] = U->iv;
  const float lpx = dot3(iv[0][0], iv[0][1], iv[0][2], px, py, pz);
  const float lpy = dot3(iv[1][0], iv[1][1], iv[1][2], px, py, pz);
  const float lpz = dot3(iv[2][0], iv[2][1], iv[2][2], px, py, pz);
  const float dist[6] = {lpx - (-U->hw), U->hw - lpx, lpy - (-U->hh),
                         U->hh - lpy,    lpz - (-U->hd), U->hd - lpz};
  float f[6], prod = 1.0f;
#pragma unroll
  for (int w = 0; w < 6; ++w) {
    f[w] = 1.0f;
    if (dist[w] < h) {
      f[w] = 0.5f + 0.5f * fminf(fmaxf(dist[w] / h, 0.f), 1.f);
      prod *= f[w];
    }
  }
  const float gamma = fmaxf(prod, 0.1f);
  if (prod > 0.1f) {
    const float lg = -lrho * ri0.w / gamma;  // d(S / gamma)/d gamma = -rho / gamma
    float gl[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int w = 0; w < 6; ++w) {
      const float q = dist[w] / h;
      if (dist[w] < h && q > 0.f && q < 1.0f) {
        const float gq = lg * (prod / f[w]) * (0.5f / h);
        gl[w / 2] += (w % 2 == 0) ? gq : -gq;
      }
    }
    lx = vadd(lx, mat_tv(iv, v3(gl[0], gl[1], gl[2])));
  }
  a.lx[s] = f4(lx);
  a.lrho[s] = make_float2(lrho / gamma, lr.y);  // adjoints of (S, near)
}

// ------------------------------------------------------------------------
// Density (sph_density_3d_grid.wgsl:129-193): S_i = sum_j m W(r^2),
// rho_near_i = sum_j m ns (h - r)^3, n_i = cn sum_{r>0} (h^2 - r^2)^2 r_vec.
// The epilogue writes the state adjoint into the previous step's slot order.
template <int L, bool B>
__global__ __launch_bounds__(256) void k_adj_density(Uni un, AdjBufs a, TapeStep t,
                                                     uint32_t n) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t s = g / L, sub = g % L;
  if (s >= n) return;
  const DevUniforms *__restrict__ U = uni_at<B>(un, s);
  const float4 xi = t.x[s];
  const float2 lsi = a.lrho[s];
  const float4 lni = a.ln[s];
  const float h = U->h, h2 = U->h2, mass = U->mass;
  const float c_s = mass * (-6.0f * U->poly6_scale);  // 2 d(m W)/d(r^2) / d^2
  const float c_n = mass * (-3.0f * U->near_scale);   // d(m ns (h-r)^3)/dr / (h-r)^2
  const float cn = U->normal_scale * U->m_over_rho0;
  float gx = 0.f, gy = 0.f, gz = 0.f;
  for_each_pair<L>(U, a, t, n, s, sub, xi, a.pcnt[s], [&](uint32_t j, float4 xj) {
    const float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
    const float r2 = dot3(rx, ry, rz, rx, ry, rz);
    const float2 lsj = a.lrho[j];
    const float4 lnj = a.ln[j];
    const float d = h2 - r2;
    const float ir = __builtin_amdgcn_rsqf(r2);
    const float dr = h - r2 * ir;
    const float dlx = lni.x - lnj.x, dly = lni.y - lnj.y, dlz = lni.z - lnj.z;
    // n_i and n_j hold +-cn d^2 r_vec: d(d^2 r_vec)/d r_vec = d^2 I - 4 d r_vec r_vec^T
    const float c = (lsi.x + lsj.x) * c_s * d * d + (lsi.y + lsj.y) * c_n * dr * dr * ir -
                    4.0f * cn * d * dot3(rx, ry, rz, dlx, dly, dlz);
    const float cd = cn * d * d;
    gx += c * rx + cd * dlx;
    gy += c * ry + cd * dly;
    gz += c * rz + cd * dlz;
  });
  gx = group_sum<L>(gx);
  gy = group_sum<L>(gy);
  gz = group_sum<L>(gz);
  if (sub != 0) return;
  const float4 lx = a.lx[s];
  const uint32_t old = t.gmap[s];
  a.lout_x[old] = make_float4(lx.x + gx, lx.y + gy, lx.z + gz, 0.f);
  a.lout_v[old] = a.lu[s];
}

// Per-scene sums of the per-slot parameter gradients: one workgroup per
// scene, a strided pass then a fixed tree, so the order (and the result) is
// the same on every run.
__global__ __launch_bounds__(256) void k_adj_reduce(const double *__restrict__ acc,
                                                    const uint32_t *__restrict__ slot_first,
                                                    double *__restrict__ out) {
  __shared__ double sh[kGradParams][256];
  const uint32_t k = blockIdx.x, t = threadIdx.x;
  const uint32_t lo = slot_first[k], hi = slot_first[k + 1];
  double v[kGradParams] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (uint32_t i = lo + t; i < hi; i += 256)
#pragma unroll
    for (int q = 0; q < kGradParams; ++q) v[q] += acc[(size_t)i * kAccStride + q];
#pragma unroll
  for (int q = 0; q < kGradParams; ++q) sh[q][t] = v[q];
  __syncthreads();
  for (uint32_t o = 128; o > 0; o >>= 1) {
    if (t < o)
#pragma unroll
      for (int q = 0; q < kGradParams; ++q) sh[q][t] += sh[q][t + o];
    __syncthreads();
  }
  if (t < kGradParams) out[(size_t)k * kGradParams + t] = sh[t][0];
}

inline uint32_t nblocks(uint64_t threads) { return (uint32_t)((threads + 255) / 256); }

// ------------------------------------------------------------------------
// Power-of-two rescaling of the carried adjoint state, per scene.  The
// reverse of a violent rollout grows by 10^1-10^2 per step (the forward
// map's stretching), past f32's range within ~16 steps.  The state carried
// from step to step -- lambda of position and velocity, and the parameter
// accumulators -- is therefore kept as mantissa x 2^E(scene): after every
// reversed step its largest magnitude m is brought into [1, 2) by 2^-k,
// k = floor(log2 m), and k is added to E on the device.  Scaling by a power
// of two is exact (barring underflow of entries 2^-126 below the largest).

// maxbits[scene] = bits of the largest |component| of the scene's lambdas
// (non-negative floats order as their bit patterns).
template <bool B>
__global__ __launch_bounds__(256) void k_adj_absmax(const uint16_t *__restrict__ scene_of,
                                                    const float4 *__restrict__ lx,
                                                    const float4 *__restrict__ lv, uint32_t n,
                                                    uint32_t *__restrict__ maxbits) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  float m = 0.0f;
  uint32_t scene = 0;
  if (s < n) {
    const float4 a = lx[s], b = lv[s];
    m = fmaxf(fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(b.x))),
              fmaxf(fabsf(b.y), fabsf(b.z)));
    if (B) scene = scene_of[s];
  }
  uint32_t bits = __float_as_uint(m);
  const uint32_t first = (uint32_t)__builtin_amdgcn_readfirstlane((int)scene);
  if (!B || __ballot(s < n && scene != first) == 0) {
    // one scene in this wave: reduce, then one atomic
    for (int o = 32; o >= 1; o >>= 1) bits = max(bits, (uint32_t)__shfl_xor((int)bits, o));
    if ((threadIdx.x & 63u) == 0 && bits) atomicMax(&maxbits[first], bits);
  } else if (s < n && bits) {
    atomicMax(&maxbits[scene], bits);
  }
}

// Per scene: k = floor(log2 m) (0 when the scene's lambdas are all zero or
// not finite), E += k; clears the other parity's maxima for the next step.
__global__ void k_adj_scale_update(const uint32_t *__restrict__ cur, uint32_t *__restrict__ next,
                                   int32_t *__restrict__ e_total, int32_t *__restrict__ e_step,
                                   uint32_t scenes) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= scenes) return;
  const uint32_t bits = cur[s];
  const uint32_t ex = (bits >> 23) & 0xFFu;
  int32_t k = 0;
  if (bits != 0u && ex != 0xFFu) k = ex == 0u ? -126 : (int32_t)ex - 127;
  e_step[s] = k;
  e_total[s] += k;
  next[s] = 0u;
}

// lambda *= 2^(mul * e[scene]); the accumulators too when acc is given.
template <bool B>
__global__ __launch_bounds__(256) void k_adj_rescale(const uint16_t *__restrict__ scene_of,
                                                     float4 *__restrict__ lx,
                                                     float4 *__restrict__ lv,
                                                     double *__restrict__ acc, uint32_t n,
                                                     const int32_t *__restrict__ e, int mul) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int k = mul * e[B ? scene_of[s] : 0];
  if (k == 0) return;
  float4 a = lx[s], b = lv[s];
  a = make_float4(ldexpf(a.x, k), ldexpf(a.y, k), ldexpf(a.z, k), a.w);
  b = make_float4(ldexpf(b.x, k), ldexpf(b.y, k), ldexpf(b.z, k), b.w);
  lx[s] = a;
  lv[s] = b;
  if (acc) {
    double *q = acc + (size_t)s * kAccStride;
#pragma unroll
    for (int j = 0; j < kGradParams; ++j) q[j] = ldexp(q[j], k);
  }
}

}  // namespace

hipError_t launch_adj_rescale_step(const uint16_t *scene_of, float4 *lx, float4 *lv,
                                   double *acc, uint32_t n, uint32_t scenes, uint32_t *max_cur,
                                   uint32_t *max_next, int32_t *e_total, int32_t *e_step,
                                   hipStream_t st) {
  if (n == 0) return hipSuccess;
  if (scene_of) {
    hipLaunchKernelGGL(k_adj_absmax<true>, dim3(nblocks(n)), dim3(256), 0, st, scene_of, lx, lv,
                       n, max_cur);
  } else {
    hipLaunchKernelGGL(k_adj_absmax<false>, dim3(nblocks(n)), dim3(256), 0, st, scene_of, lx,
                       lv, n, max_cur);
  }
  hipLaunchKernelGGL(k_adj_scale_update, dim3(nblocks(scenes)), dim3(256), 0, st, max_cur,
                     max_next, e_total, e_step, scenes);
  return launch_adj_apply_scale(scene_of, lx, lv, acc, n, e_step, -1, st);
}

hipError_t launch_adj_apply_scale(const uint16_t *scene_of, float4 *lx, float4 *lv, double *acc,
                                  uint32_t n, const int32_t *e, int mul, hipStream_t st) {
  if (n == 0) return hipSuccess;
  if (scene_of)
    hipLaunchKernelGGL(k_adj_rescale<true>, dim3(nblocks(n)), dim3(256), 0, st, scene_of, lx, lv,
                       acc, n, e, mul);
  else
    hipLaunchKernelGGL(k_adj_rescale<false>, dim3(nblocks(n)), dim3(256), 0, st, scene_of, lx,
                       lv, acc, n, e, mul);
  return hipGetLastError();
}

hipError_t launch_adj_reduce(const double *acc, const uint32_t *slot_first, uint32_t scenes,
                             double *out, hipStream_t st) {
  if (scenes == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_reduce, dim3(scenes), dim3(256), 0, st, acc, slot_first, out);
  return hipGetLastError();
}

hipError_t launch_adj_load(const uint32_t *sid, const float *dl_x3, const float *dl_v3,
                           float4 *lx, float4 *lv, uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_load, dim3(nblocks(n)), dim3(256), 0, st, sid, dl_x3, dl_v3, lx, lv,
                     n);
  return hipGetLastError();
}

hipError_t launch_adj_store(const uint32_t *sid, const float4 *lx, const float4 *lv,
                            float *out_x3, float *out_v3, uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_store, dim3(nblocks(n)), dim3(256), 0, st, sid, lx, lv, out_x3,
                     out_v3, n);
  return hipGetLastError();
}

#define ADJ_LAUNCH_B(KERNEL, GRID, UN, ...)                                             \
  do {                                                                                  \
    if ((UN).scene_of)                                                                  \
      hipLaunchKernelGGL(KERNEL<true>, dim3(GRID), dim3(256), 0, st, UN, __VA_ARGS__);  \
    else                                                                                \
      hipLaunchKernelGGL(KERNEL<false>, dim3(GRID), dim3(256), 0, st, UN, __VA_ARGS__); \
  } while (0)

#define ADJ_LAUNCH_LB(KERNEL, L, N, UN, ...)                                                    \
  do {                                                                                          \
    const uint32_t _g = nblocks((uint64_t)(N) * (L));                                           \
    const bool _b = (UN).scene_of != nullptr;                                                   \
    if ((L) == 8) {                                                                             \
      if (_b)                                                                                   \
        hipLaunchKernelGGL((KERNEL<8, true>), dim3(_g), dim3(256), 0, st, UN, __VA_ARGS__);     \
      else                                                                                      \
        hipLaunchKernelGGL((KERNEL<8, false>), dim3(_g), dim3(256), 0, st, UN, __VA_ARGS__);    \
    } else {                                                                                    \
      if (_b)                                                                                   \
        hipLaunchKernelGGL((KERNEL<4, true>), dim3(_g), dim3(256), 0, st, UN, __VA_ARGS__);     \
      else                                                                                      \
        hipLaunchKernelGGL((KERNEL<4, false>), dim3(_g), dim3(256), 0, st, UN, __VA_ARGS__);    \
    }                                                                                           \
  } while (0)

hipError_t launch_adjoint_step(const Uni &u, AdjBufs a, const TapeStep &t, const TapeIter *it,
                               const TapeIter &last, uint32_t iters, uint32_t n, int lanes,
                               hipStream_t st) {
  if (n == 0) return hipSuccess;
  const int L = lanes >= 8 ? 8 : 4;
  ADJ_LAUNCH_B(k_adj_integrate, nblocks(n), u, a, t, last, n);
  int cur = 0;
  for (int k = (int)iters - 1; k >= 0; --k) {
    const uint32_t mode = iters == 1 ? 0u : (k == (int)iters - 1 ? 1u : 2u);
    ADJ_LAUNCH_LB(k_adj_solve, L, n, u, a, t, n, it[k], a.lap[cur], a.lap[1 - cur],
                  (uint32_t)k, it[0].xq, mode);
    cur = 1 - cur;
  }
  ADJ_LAUNCH_LB(k_adj_force, L, n, u, a, t, n);
  ADJ_LAUNCH_LB(k_adj_xsph, L, n, u, a, t, n);
  ADJ_LAUNCH_LB(k_adj_density, L, n, u, a, t, n);
  return hipGetLastError();
}

}  // namespace fluid
