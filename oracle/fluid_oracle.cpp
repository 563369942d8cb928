// fluid_oracle.cpp — CPU restatement of fluid-toy's SPH step, used ONLY as a
// test oracle and as the reported CPU baseline.  It is test infrastructure:
// nothing in the product library links or calls it.
//
// It restates, stage by stage and in the same floating-point operation order,
// the WGSL compute entry points that SphSimulation3DGrid::step dispatches
// (src/simulation/sph_3d_grid.rs:1261-1444):
//   grid_clear.wgsl, grid_build.wgsl, prefix_scan.wgsl, grid_reorder.wgsl,
//   sph_density_3d_grid.wgsl (+ container_common.wgsl), xsph.wgsl,
//   sph_force_3d_grid.wgsl, pcisph_predict.wgsl, pcisph_solve.wgsl,
//   pcisph_finalize.wgsl, sph_integrate_3d.wgsl.
// One deliberate difference: the reorder's within-cell order, which the
// reference takes from a racing atomicAdd (grid_reorder.wgsl:51), is fixed
// to ascending position x (compared as the float32 value's order key, ties
// by particle index): any order is the reference's, and this one is the
// HIP build's (neighbouring slots close in space for its gathers).
//
// It is built twice from one template: T=float (compile with
// -ffp-contract=off so every WGSL operation rounds once, in source order) and
// T=double (the FP64 transcription used for rollout divergence and
// gradients).  The parameter structs are the reference's f32 uniforms; the
// FP64 variant promotes their values.
//
// Parity status of each function is recorded next to it ("pinned by ...").
// The file also restates, in f32 only, the passes that consume the step's
// sorted state: the whitewater passes (spray_emit.wgsl, spray_simulate.wgsl,
// src/simulation/spray.rs) and the marching-cubes density field
// (mc_anisotropy.wgsl, mc_density.wgsl).
//
// What only running the real wgpu pipeline could pin — naga/driver rounding
// of the sums, sin/cos of the close-range fallback direction, the racing
// reorder order — is not pinned here (no Rust/Vulkan toolchain exists here).

#include "../include/fluid_abi.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// Branch-decision record (for the gradient checks): every non-smooth
// decision of the force, solve and integrate stages that comes out "taken"
// adds a hash of (kind, particle, partner or iteration, detail) to its
// particle's slot of Oracle::sig.  Two runs whose sums agree took the same
// branches (up to hash collisions), so a central difference between them
// differentiates one smooth piece.
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t decision(uint32_t kind, uint32_t a, uint32_t b, int64_t c) {
  return mix64(mix64(((uint64_t)kind << 32) ^ a) ^ ((uint64_t)b << 20) ^ (uint64_t)c);
}

template <class T> struct V3 {
  T x, y, z;
};
template <class T> inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> add(V3<T> a, V3<T> b) {
  return mk(a.x + b.x, a.y + b.y, a.z + b.z);
}
template <class T> inline V3<T> sub(V3<T> a, V3<T> b) {
  return mk(a.x - b.x, a.y - b.y, a.z - b.z);
}
// scalar * vector (WGSL `s * v`)
template <class T> inline V3<T> smul(T s, V3<T> v) {
  return mk(s * v.x, s * v.y, s * v.z);
}
// vector * scalar (WGSL `v * s`)
template <class T> inline V3<T> muls(V3<T> v, T s) {
  return mk(v.x * s, v.y * s, v.z * s);
}
template <class T> inline V3<T> divs(V3<T> v, T s) {
  return mk(v.x / s, v.y / s, v.z / s);
}
// WGSL dot() is evaluated fused, as fma(z, z', fma(y, y', x*x')): the
// lowering GPU compilers give it, and the one the HIP kernels use, so that
// every in-range test (r^2 < h^2) is decided on identical bits.
template <class T> inline T dot(V3<T> a, V3<T> b) {
  return std::fma(a.z, b.z, std::fma(a.y, b.y, a.x * b.x));
}
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a) {
  return divs(a, length(a));
}
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
            a.x * b.y - a.y * b.x);
}
template <class T> inline T clampv(T v, T lo, T hi) {
  return std::min(std::max(v, lo), hi);
}
template <class T> inline T wsign(T v) {
  return v > T(0) ? T(1) : (v < T(0) ? T(-1) : T(0));
}

// WGSL f32 -> i32 conversion: truncate toward zero, saturate, NaN -> 0.
template <class T> inline int32_t to_i32(T v) {
  if (!(v == v)) return 0;
  if (v >= T(2147483520.0)) return 2147483647;
  if (v <= T(-2147483648.0)) return INT32_MIN;
  return (int32_t)v;
}

// Particle record in the precision T, same field order as SphParticle3D
// (src/simulation/particle.rs:6-17).
template <class T> struct Part {
  T pos[3];
  T pad0;
  T vel[3];
  T pad1;
  T force[3];
  T density;
  T near_density;
  T normal[3];
};
static_assert(sizeof(Part<float>) == 64, "particle record");

template <class T> struct Pred {  // pcisph_predict.wgsl:33-42
  T x[3];
  T pressure;
  T v[3];
  T density;
};

template <class T> inline V3<T> P3(const T *p) { return mk(p[0], p[1], p[2]); }
template <class T> inline void S3(T *p, V3<T> v) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}

template <class T> struct Cont {  // container_common.wgsl:5-24
  T hw, hh, hd, cy;
  V3<T> f0, f1, f2, i0, i1, i2;
  T wall_stiffness, damping;
};

template <class T> struct Oracle {
  uint32_t n = 0;
  uint32_t iters = 4;
  // uniforms, promoted to T
  T h, h2, h5, h6, h9, mass, rho0, near_k, visc, dt, st, delta, eps;
  uint32_t gx, gy, gz, cells;
  T inv_cell, ox, oy, oz;
  Cont<T> c;
  V3<T> gravity;
  GpuMouseForce mouse;
  GpuRigidBody rb;
  std::vector<T> sdf;
  uint32_t sdf_res = 0;
  // buffers (sph_3d_grid.rs:196-305)
  std::vector<Part<T>> particles, sorted;
  std::vector<uint32_t> counts, starts, offsets, pci, s2o;
  std::vector<Pred<T>> pa, pb;
  int32_t accum[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<uint64_t> sig;  // branch-decision record per original index
  uint32_t solve_it = 0;      // PCISPH iteration being run (for the record)

  // ---- container_common.wgsl:27-54 -------------------------------------
  V3<T> world_to_local(V3<T> pos) const {
    V3<T> p = mk(pos.x, pos.y - c.cy, pos.z);
    return mk(dot(c.i0, p), dot(c.i1, p), dot(c.i2, p));
  }
  V3<T> local_to_world(V3<T> l) const {
    V3<T> r = mk(dot(c.f0, l), dot(c.f1, l), dot(c.f2, l));
    return mk(r.x, r.y + c.cy, r.z);
  }
  V3<T> world_dir_to_local(V3<T> d) const {
    return mk(dot(c.i0, d), dot(c.i1, d), dot(c.i2, d));
  }
  V3<T> local_dir_to_world(V3<T> d) const {
    return mk(dot(c.f0, d), dot(c.f1, d), dot(c.f2, d));
  }

  // ---- grid helpers (grid_build.wgsl:33-44; sph_density_3d_grid.wgsl:68-81)
  // Cell coordinates are a discrete decision taken on the f32 value of the
  // position in f32 arithmetic at both precisions, so the float64 oracle
  // sorts exactly like the float32 one and differs only in the sums.
  void position_to_cell(V3<T> pos, int32_t out[3]) const {
    const float fi = (float)inv_cell;
    out[0] = to_i32(std::floor(((float)pos.x - (float)ox) * fi));
    out[1] = to_i32(std::floor(((float)pos.y - (float)oy) * fi));
    out[2] = to_i32(std::floor(((float)pos.z - (float)oz) * fi));
  }
  bool is_valid_cell(int32_t x, int32_t y, int32_t z) const {
    return x >= 0 && x < (int32_t)gx && y >= 0 && y < (int32_t)gy && z >= 0 &&
           z < (int32_t)gz;
  }
  uint32_t cell_index(int32_t x, int32_t y, int32_t z) const {
    return (uint32_t)x + (uint32_t)y * gx + (uint32_t)z * gx * gy;
  }

  // kernels (sph_density_3d_grid.wgsl:56-66, sph_force_3d_grid.wgsl:63-88)
  T PI() const { return T(3.14159265359f); }
  T poly6(T r_sq) const {
    T scale = T(315.0) / (T(64.0) * PI() * h9);
    T diff = h2 - r_sq;
    return scale * diff * diff * diff;
  }
  T near_kernel(T r) const {
    T scale = T(15.0) / (PI() * h6);
    T diff = h - r;
    return scale * diff * diff * diff;
  }
  T spiky_grad(T r) const {
    T scale = T(45.0) / (PI() * h6);
    T diff = h - r;
    return scale * diff * diff;
  }
  T akinci(T r) const {
    T scale = T(32.0) / (PI() * h9);
    T diff = h - r;
    T r3 = r * r * r;
    if (r >= h * T(0.5)) return scale * diff * diff * diff * r3;
    return scale * (T(2.0) * diff * diff * diff * r3 - h6 / T(64.0));
  }

  // Iterate the 27-cell stencil of `cell_i` in the reference order
  // (dz, dy, dx, then slot) and call f(j) for every candidate slot j.
  // Coordinates beyond one cell outside the grid see no valid cell either
  // way; clamping them first keeps cell_i + d from overflowing.
  template <class F> void sweep(const int32_t cell_in[3], F &&f) const {
    const int32_t cell_i[3] = {clampv<int32_t>(cell_in[0], -2, (int32_t)gx + 1),
                               clampv<int32_t>(cell_in[1], -2, (int32_t)gy + 1),
                               clampv<int32_t>(cell_in[2], -2, (int32_t)gz + 1)};
    for (int32_t dz = -1; dz <= 1; ++dz)
      for (int32_t dy = -1; dy <= 1; ++dy)
        for (int32_t dx = -1; dx <= 1; ++dx) {
          int32_t x = cell_i[0] + dx, y = cell_i[1] + dy, z = cell_i[2] + dz;
          if (!is_valid_cell(x, y, z)) continue;
          uint32_t ci = cell_index(x, y, z);
          uint32_t count = counts[ci];
          uint32_t start = starts[ci] - count;
          for (uint32_t k = 0; k < count; ++k) f(start + k);
        }
  }

  // ---- stage 0: grid_clear.wgsl:23-29 + clear_buffer(sph_3d_grid.rs:1281)
  void clear() {
    std::fill(counts.begin(), counts.end(), 0u);
    std::fill(offsets.begin(), offsets.end(), 0u);
  }

  // ---- stage 1: grid_build.wgsl:46-62 ----------------------------------
  void build() {
    for (uint32_t i = 0; i < n; ++i) {
      int32_t cl[3];
      position_to_cell(P3(particles[i].pos), cl);
      int32_t cx = clampv<int32_t>(cl[0], 0, (int32_t)gx - 1);
      int32_t cy = clampv<int32_t>(cl[1], 0, (int32_t)gy - 1);
      int32_t cz = clampv<int32_t>(cl[2], 0, (int32_t)gz - 1);
      uint32_t ci = cell_index(cx, cy, cz);
      pci[i] = ci;
      counts[ci] += 1u;
    }
  }

  // ---- stage 2: copy (sph_3d_grid.rs:1295) + prefix_scan.wgsl:50-122 ----
  // The three dispatches compose to an inclusive scan (u32, wrapping).
  void scan() {
    uint32_t acc = 0;
    for (uint32_t ci = 0; ci < cells; ++ci) {
      acc += counts[ci];
      starts[ci] = acc;
    }
  }

  // ---- stage 3: grid_reorder.wgsl:37-63 ---------------------------------
  // Within a cell by (order key of x as float32, index): the key maps the
  // float order to unsigned order over every bit pattern.
  static uint32_t x_key(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    return u ^ ((u >> 31) ? 0xFFFFFFFFu : 0x80000000u);
  }
  void reorder() {
    std::vector<uint64_t> key(n);
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) {
      key[i] = (uint64_t)x_key((float)particles[i].pos[0]) << 32 | i;
      order[i] = i;
    }
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
      return pci[a] != pci[b] ? pci[a] < pci[b] : key[a] < key[b];
    });
    for (uint32_t k = 0; k < n; ++k) {
      const uint32_t i = order[k];
      const uint32_t ci = pci[i];
      offsets[ci] += 1u;
      sorted[k] = particles[i];
      s2o[k] = i;
    }
    for (uint32_t k = 0; k < n; ++k) pci[s2o[k]] = k;
  }

  // ---- stage 4: sph_density_3d_grid.wgsl:86-194 ------------------------
  T boundary_gamma(V3<T> pos) const {
    V3<T> l = world_to_local(pos);
    T g = T(1.0);
    T d[6] = {l.x - (-c.hw), c.hw - l.x, l.y - (-c.hh),
              c.hh - l.y,    l.z - (-c.hd), c.hd - l.z};
    for (int k = 0; k < 6; ++k)
      if (d[k] < h) g *= T(0.5) + T(0.5) * clampv(d[k] / h, T(0), T(1));
    return std::max(g, T(0.1));
  }
  void density() {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t si = 0; si < (int64_t)n; ++si) {
      uint32_t s = (uint32_t)si;
      V3<T> pos_i = P3(sorted[s].pos);
      int32_t cl[3];
      position_to_cell(pos_i, cl);
      T dens = T(0), near = T(0);
      V3<T> nrm = mk(T(0), T(0), T(0));
      sweep(cl, [&](uint32_t j) {
        V3<T> r_vec = sub(pos_i, P3(sorted[j].pos));
        T r_sq = dot(r_vec, r_vec);
        if (r_sq < h2) {
          T r = std::sqrt(r_sq);
          dens += mass * poly6(r_sq);
          near += mass * near_kernel(r);
          if (r_sq > T(1e-12f)) {
            T diff = h2 - r_sq;
            nrm = add(nrm, muls(muls(smul(mass / rho0, r_vec), diff), diff));
          }
        }
      });
      T gamma = boundary_gamma(pos_i);
      dens = dens / gamma;
      T h8 = h6 * h * h;
      T nscale = T(945.0) / (T(32.0) * PI() * h8);
      nrm = muls(nrm, nscale);
      sorted[s].density = dens;
      sorted[s].near_density = near;
      S3(sorted[s].normal, nrm);
    }
  }

  // ---- stage 5: xsph.wgsl:81-138 ---------------------------------------
  void xsph() {
    if (eps <= T(0)) return;  // xsph.wgsl:94-97 (every invocation returns)
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t si = 0; si < (int64_t)n; ++si) {
      uint32_t s = (uint32_t)si;
      V3<T> pos_i = P3(sorted[s].pos);
      V3<T> vel_i = P3(sorted[s].vel);
      int32_t cl[3];
      position_to_cell(pos_i, cl);
      V3<T> corr = mk(T(0), T(0), T(0));
      sweep(cl, [&](uint32_t j) {
        V3<T> r_vec = sub(pos_i, P3(sorted[j].pos));
        T r_sq = dot(r_vec, r_vec);
        if (r_sq < h2 && r_sq > T(1e-12f)) {
          T dj = std::max(sorted[j].density, T(1.0));
          V3<T> vj = P3(sorted[j].vel);
          T w = poly6(r_sq);
          corr = add(corr, muls(smul(mass / dj, sub(vj, vel_i)), w));
        }
      });
      S3(particles[s2o[s]].vel, add(vel_i, smul(eps, corr)));
    }
  }

  // ---- stage 6: sph_force_3d_grid.wgsl:105-227 --------------------------
  void force() {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t si = 0; si < (int64_t)n; ++si) {
      uint32_t s = (uint32_t)si;
      uint32_t orig = s2o[s];
      V3<T> pos_i = P3(sorted[s].pos);
      V3<T> vel_i = P3(particles[orig].vel);
      int32_t cl[3];
      position_to_cell(pos_i, cl);
      T di = std::max(sorted[s].density, T(1.0));
      T ndi = std::max(sorted[s].near_density, T(1.0));
      V3<T> n_i = P3(sorted[s].normal);
      T npi = near_k * ndi;
      V3<T> ap = mk(T(0), T(0), T(0)), av = ap, ac = ap;
      uint64_t dec = 0;
      sweep(cl, [&](uint32_t j) {
        V3<T> pos_j = P3(sorted[j].pos);
        V3<T> r_vec = sub(pos_i, pos_j);
        T r_sq = dot(r_vec, r_vec);
        if (!(r_sq < h2 && r_sq > T(1e-12f))) return;
        T r = std::sqrt(r_sq);
        V3<T> dir;
        T min_dist = h * T(0.1);
        if (r < min_dist) dec += decision(1, orig, s2o[j], (int64_t)s - (int64_t)j);
        if (r < h * T(0.5)) dec += decision(3, orig, s2o[j], 0);
        if (r < min_dist) {
          T idx = T((float)s) - T((float)j);
          dir = normalize(mk(std::sin(idx * T(1.0)), std::cos(idx * T(2.0)),
                             std::sin(idx * T(3.0))));
        } else {
          dir = normalize(sub(pos_j, pos_i));
        }
        T dj = std::max(sorted[j].density, T(1.0));
        T ndj = std::max(sorted[j].near_density, T(1.0));
        T npj = near_k * ndj;
        T sp = -mass * (npi + npj) / (T(2.0) * di * ndj);
        ap = add(ap, muls(smul(sp, dir), spiky_grad(r)));
        if (r < min_dist) {
          T sep = T(200.0) * (T(1.0) - r / min_dist);
          ap = add(ap, smul(-sep, dir));
        }
        V3<T> vel_j = P3(sorted[j].vel);
        V3<T> v_ij = sub(vel_i, vel_j);
        V3<T> x_ij = sub(pos_i, pos_j);
        T vdx = dot(v_ij, x_ij);
        if (vdx < T(0)) dec += decision(2, orig, s2o[j], 0);
        if (vdx < T(0)) {
          T mu = h * vdx / (r_sq + T(0.01) * h2);
          T avg = (di + dj) * T(0.5);
          T pi_visc = -visc * mu / avg;
          av = add(av, smul(-mass * pi_visc * spiky_grad(r), dir));
        }
        T K = T(2.0) * rho0 / (di + dj);
        T coh = akinci(r);
        ac = add(ac, smul(mass * K * coh, dir));
        V3<T> n_j = P3(sorted[j].normal);
        ac = sub(ac, smul(mass * K, sub(n_i, n_j)));
      });
      ac = muls(ac, st);
      S3(particles[orig].force, add(add(add(ap, av), ac), gravity));
      sig[orig] += dec;
    }
  }

  // ---- stage 7: pcisph_predict.wgsl:50-73 ------------------------------
  void predict() {
    for (uint32_t i = 0; i < n; ++i) {
      V3<T> pos = P3(particles[i].pos), vel = P3(particles[i].vel);
      V3<T> a = P3(particles[i].force);
      V3<T> vs = add(vel, smul(dt, a));
      V3<T> xs = add(pos, smul(dt, vs));
      Pred<T> &o = pa[pci[i]];
      S3(o.x, xs);
      o.pressure = T(0);
      S3(o.v, vs);
      o.density = T(0);
    }
  }

  // ---- stage 8: pcisph_solve.wgsl:99-208, ping-pong (sph_3d_grid.rs:1388)
  void solve_iter(const std::vector<Pred<T>> &in, std::vector<Pred<T>> &out) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t si = 0; si < (int64_t)n; ++si) {
      uint32_t s = (uint32_t)si;
      uint32_t orig = s2o[s];
      const Pred<T> &self = in[s];
      V3<T> pp = P3(self.x);
      T old_p = self.pressure;
      int32_t cl[3];
      position_to_cell(P3(particles[orig].pos), cl);
      T pd = T(0);
      V3<T> ap = mk(T(0), T(0), T(0));
      sweep(cl, [&](uint32_t j) {
        const Pred<T> &nb = in[j];
        V3<T> r_vec = sub(pp, P3(nb.x));
        T r_sq = dot(r_vec, r_vec);
        if (r_sq < h2) {
          pd += mass * poly6(r_sq);
          if (r_sq > T(1e-12f)) {
            T r = std::sqrt(r_sq);
            T nd = std::max(nb.density, T(1.0));
            T np = nb.pressure;
            T sd = std::max(self.density, T(1.0));
            V3<T> dir = divs(mk(-r_vec.x, -r_vec.y, -r_vec.z), r);
            T sc = -mass * (old_p / (sd * sd) + np / (nd * nd)) * spiky_grad(r);
            ap = add(ap, smul(sc, dir));
          }
        }
      });
      if (pd - rho0 > T(0)) sig[orig] += decision(4, orig, solve_it, 0);
      T err = std::max(T(0), pd - rho0);
      T new_p = old_p + delta * err;
      T max_a = T(0.25) * h / (dt * dt);
      T amag = dot(ap, ap);
      T max_a2 = max_a * max_a;
      if (amag > max_a2) sig[orig] += decision(5, orig, solve_it, 0);
      if (amag > max_a2) ap = muls(ap, max_a / std::sqrt(amag));
      V3<T> ov = P3(particles[orig].vel), op = P3(particles[orig].pos);
      V3<T> anp = P3(particles[orig].force);
      V3<T> at = add(anp, ap);
      V3<T> vs = add(ov, smul(dt, at));
      V3<T> xs = add(op, smul(dt, vs));
      Pred<T> &o = out[s];
      S3(o.x, xs);
      o.pressure = new_p;
      S3(o.v, vs);
      o.density = pd;
    }
  }
  void solve() {
    for (uint32_t it = 0; it < iters; ++it) {
      solve_it = it;
      if (it % 2 == 0)
        solve_iter(pa, pb);
      else
        solve_iter(pb, pa);
    }
  }

  // ---- stage 9: pcisph_finalize.wgsl:50-74 (buffer: sph_3d_grid.rs:1409)
  void finalize() {
    const std::vector<Pred<T>> &fin = (iters % 2 == 0) ? pa : pb;
    for (uint32_t i = 0; i < n; ++i) {
      const Pred<T> &f = fin[pci[i]];
      V3<T> cv = P3(f.v);
      V3<T> ov = P3(particles[i].vel);
      S3(particles[i].force, sub(cv, ov));
      S3(particles[i].vel, cv);
    }
  }

  // ---- stage 10: sph_integrate_3d.wgsl:91-388 --------------------------
  T sdf_fetch(int x, int y, int z) const {
    int r = (int)sdf_res;
    x = std::min(std::max(x, 0), r - 1);
    y = std::min(std::max(y, 0), r - 1);
    z = std::min(std::max(z, 0), r - 1);
    return sdf[(size_t)x + (size_t)y * r + (size_t)z * r * r];
  }
  // textureSampleLevel with a linear, clamp-to-edge sampler
  // (sph_3d_grid.rs:336-344), exact (unquantised) trilinear weights.
  T sdf_sample(V3<T> uvw) const {
    if (sdf_res == 0) return T(1.0);  // dummy 1x1x1 texture of 1.0 (:347)
    T u = uvw.x * T(sdf_res) - T(0.5), v = uvw.y * T(sdf_res) - T(0.5),
      w = uvw.z * T(sdf_res) - T(0.5);
    T fu = std::floor(u), fv = std::floor(v), fw = std::floor(w);
    int x0 = (int)fu, y0 = (int)fv, z0 = (int)fw;
    T a = u - fu, b = v - fv, cz = w - fw;
    T c000 = sdf_fetch(x0, y0, z0), c100 = sdf_fetch(x0 + 1, y0, z0);
    T c010 = sdf_fetch(x0, y0 + 1, z0), c110 = sdf_fetch(x0 + 1, y0 + 1, z0);
    T c001 = sdf_fetch(x0, y0, z0 + 1), c101 = sdf_fetch(x0 + 1, y0, z0 + 1);
    T c011 = sdf_fetch(x0, y0 + 1, z0 + 1);
    T c111 = sdf_fetch(x0 + 1, y0 + 1, z0 + 1);
    T x00 = c000 + (c100 - c000) * a, x10 = c010 + (c110 - c010) * a;
    T x01 = c001 + (c101 - c001) * a, x11 = c011 + (c111 - c011) * a;
    T y0v = x00 + (x10 - x00) * b, y1v = x01 + (x11 - x01) * b;
    return y0v + (y1v - y0v) * cz;
  }

  void integrate() {
    for (uint32_t i = 0; i < n; ++i) {
      V3<T> pos = P3(particles[i].pos), vel = P3(particles[i].vel);
      V3<T> accel = mk(T(0), T(0), T(0));
      V3<T> lp = world_to_local(pos), lv = world_dir_to_local(vel);
      T bl = h * T(0.4);
      T wd = T(8.0);
      V3<T> wa = mk(T(0), T(0), T(0));
      T k = c.wall_stiffness;
      {  // branch record: wall layers entered, their damping switches
        const T dw[6] = {lp.x + c.hw, c.hw - lp.x, lp.y + c.hh, c.hh - lp.y, lp.z + c.hd,
                         c.hd - lp.z};
        const T into[6] = {-lv.x, lv.x, -lv.y, lv.y, -lv.z, lv.z};
        for (uint32_t f = 0; f < 6; ++f) {
          if (dw[f] < bl) sig[i] += decision(6, i, f, 0);
          if (dw[f] < bl && into[f] > T(0)) sig[i] += decision(7, i, f, 0);
        }
      }
      {
        T d = lp.x - (-c.hw);
        if (d < bl) {
          T t = T(1.0) - d / bl;
          wa.x += k * t * t;
          if (lv.x < T(0)) wa.x -= lv.x * wd * t;
        }
        d = c.hw - lp.x;
        if (d < bl) {
          T t = T(1.0) - d / bl;
          wa.x -= k * t * t;
          if (lv.x > T(0)) wa.x -= lv.x * wd * t;
        }
        d = lp.y - (-c.hh);
        if (d < bl) {
          T t = T(1.0) - d / bl;
          wa.y += k * t * t;
          if (lv.y < T(0)) wa.y -= lv.y * wd * t;
        }
        d = c.hh - lp.y;
        if (d < bl) {
          T t = T(1.0) - d / bl;
          wa.y -= k * t * t;
          if (lv.y > T(0)) wa.y -= lv.y * wd * t;
        }
        d = lp.z - (-c.hd);
        if (d < bl) {
          T t = T(1.0) - d / bl;
          wa.z += k * t * t;
          if (lv.z < T(0)) wa.z -= lv.z * wd * t;
        }
        d = c.hd - lp.z;
        if (d < bl) {
          T t = T(1.0) - d / bl;
          wa.z -= k * t * t;
          if (lv.z > T(0)) wa.z -= lv.z * wd * t;
        }
      }
      accel = add(accel, local_dir_to_world(wa));

      if (mouse.is_active == 1u) {  // :163-200
        V3<T> mp = mk(T(mouse.position[0]), T(mouse.position[1]),
                      T(mouse.position[2]));
        V3<T> tm = sub(mp, pos);
        T dist = length(tm);
        T rad = T(mouse.radius);
        if (dist < rad && dist > T(0.001f)) {
          T fall = T(1.0) - (dist / rad);
          V3<T> fd = normalize(tm);
          T s = T(mouse.strength) * fall;
          switch (mouse.mode) {
            case 1:
              accel = add(accel, muls(fd, s));
              break;
            case 2: {
              V3<T> axis = normalize(mk(T(mouse.direction[0]),
                                        T(mouse.direction[1]),
                                        T(mouse.direction[2])));
              V3<T> tg = cross(fd, axis);
              T tl = length(tg);
              if (tl > T(0.001f)) accel = add(accel, muls(divs(tg, tl), s));
              break;
            }
            case 3:
              accel = sub(accel, muls(muls(fd, s), T(3.0)));
              break;
            case 4: {
              V3<T> dd = normalize(add(fd, mk(T(0), T(-1.0), T(0))));
              accel = add(accel, muls(dd, s));
              break;
            }
            default:
              accel = sub(accel, muls(fd, s));
              break;
          }
        }
      }

      if (rb.is_active != 0u) {  // :203-341
        V3<T> rpos = mk(T(rb.position[0]), T(rb.position[1]),
                        T(rb.position[2]));
        V3<T> r0 = mk(T(rb.rot_row0[0]), T(rb.rot_row0[1]), T(rb.rot_row0[2]));
        V3<T> r1 = mk(T(rb.rot_row1[0]), T(rb.rot_row1[1]), T(rb.rot_row1[2]));
        V3<T> r2 = mk(T(rb.rot_row2[0]), T(rb.rot_row2[1]), T(rb.rot_row2[2]));
        V3<T> wr = sub(pos, rpos);
        V3<T> rl = mk(dot(r0, wr), dot(r1, wr), dot(r2, wr));
        T he = T(rb.half_extent);
        T sdfv;
        V3<T> ln = mk(T(0), T(0), T(0));
        switch (rb.shape) {
          case 1: {
            T dist = length(rl);
            sdfv = dist - he;
            ln = dist > T(0.001f) ? divs(rl, dist) : mk(T(0), T(1), T(0));
            break;
          }
          case 2: {
            T rd = std::sqrt(std::fma(rl.z, rl.z, rl.x * rl.x));
            T d_rad = rd - he;
            T d_cap = std::fabs(rl.y) - he;
            sdfv = std::max(d_rad, d_cap);
            if (d_rad > d_cap)
              ln = rd > T(0.001f) ? mk(rl.x / rd, T(0), rl.z / rd)
                                  : mk(T(1), T(0), T(0));
            else
              ln = mk(T(0), wsign(rl.y), T(0));
            break;
          }
          case 3: {
            T minor = he * T(0.3);
            T xz = std::sqrt(std::fma(rl.z, rl.z, rl.x * rl.x));
            T qx = xz - he, qy = rl.y;
            T ql = std::sqrt(std::fma(qy, qy, qx * qx));
            sdfv = ql - minor;
            if (ql > T(0.001f) && xz > T(0.001f)) {
              T rdx = qx / ql, rdy = qy / ql;
              ln = mk(rdx * rl.x / xz, rdy, rdx * rl.z / xz);
            } else {
              ln = mk(T(0), T(1), T(0));
            }
            break;
          }
          case 4: {
            V3<T> nz = divs(rl, he);
            V3<T> uvw = add(muls(nz, T(0.5)), mk(T(0.5), T(0.5), T(0.5)));
            sdfv = sdf_sample(uvw) * he;
            T e = T(1.0) / T(32.0);
            T gxv = sdf_sample(add(uvw, mk(e, T(0), T(0)))) -
                    sdf_sample(sub(uvw, mk(e, T(0), T(0))));
            T gyv = sdf_sample(add(uvw, mk(T(0), e, T(0)))) -
                    sdf_sample(sub(uvw, mk(T(0), e, T(0))));
            T gzv = sdf_sample(add(uvw, mk(T(0), T(0), e))) -
                    sdf_sample(sub(uvw, mk(T(0), T(0), e)));
            V3<T> g = mk(gxv, gyv, gzv);
            T gl = length(g);
            ln = gl > T(1e-6f) ? divs(g, gl)
                               : normalize(add(rl, mk(T(0), T(1e-6f), T(0))));
            break;
          }
          default: {
            V3<T> d = mk(std::fabs(rl.x) - he, std::fabs(rl.y) - he,
                         std::fabs(rl.z) - he);
            sdfv = std::max(d.x, std::max(d.y, d.z));
            if (d.x >= d.y && d.x >= d.z)
              ln = mk(wsign(rl.x), T(0), T(0));
            else if (d.y >= d.x && d.y >= d.z)
              ln = mk(T(0), wsign(rl.y), T(0));
            else
              ln = mk(T(0), T(0), wsign(rl.z));
            break;
          }
        }
        T ir = h * T(0.7);
        if (sdfv < ir) {
          V3<T> nw = mk(r0.x * ln.x + r1.x * ln.y + r2.x * ln.z,
                        r0.y * ln.x + r1.y * ln.y + r2.y * ln.z,
                        r0.z * ln.x + r1.z * ln.y + r2.z * ln.z);
          T pen = ir - sdfv;
          T t = pen / ir;
          V3<T> pen_a = muls(muls(muls(nw, T(rb.stiffness)), t), t);
          accel = add(accel, pen_a);
          V3<T> damp = mk(T(0), T(0), T(0));
          V3<T> rv = sub(vel, mk(T(rb.velocity[0]), T(rb.velocity[1]),
                                 T(rb.velocity[2])));
          T vn = dot(rv, nw);
          if (vn < T(0)) {
            damp = muls(muls(mk(-nw.x, -nw.y, -nw.z), vn), T(5.0));
            accel = add(accel, damp);
          }
          V3<T> sum = add(pen_a, damp);
          V3<T> re = muls(mk(-sum.x, -sum.y, -sum.z), mass);
          auto acc_add = [&](int k2, int32_t v) {
            accum[k2] = (int32_t)((uint32_t)accum[k2] + (uint32_t)v);
          };
          acc_add(0, to_i32(re.x * T(1000.0)));
          acc_add(1, to_i32(re.y * T(1000.0)));
          acc_add(2, to_i32(re.z * T(1000.0)));
          accum[3] = (int32_t)((uint32_t)accum[3] + 1u);
          V3<T> tq = cross(wr, re);
          acc_add(4, to_i32(tq.x * T(1000.0)));
          acc_add(5, to_i32(tq.y * T(1000.0)));
          acc_add(6, to_i32(tq.z * T(1000.0)));
        }
      }

      vel = add(vel, smul(dt, accel));
      pos = add(pos, smul(dt, vel));

      V3<T> lpn = world_to_local(pos), lvn = world_dir_to_local(vel);
      {  // branch record: backstop contacts
        const T lq[3] = {lpn.x, lpn.y, lpn.z}, hq[3] = {c.hw, c.hh, c.hd};
        for (uint32_t a = 0; a < 3; ++a) {
          if (lq[a] < -hq[a]) sig[i] += decision(8, i, 2 * a, 0);
          if (lq[a] > hq[a]) sig[i] += decision(8, i, 2 * a + 1, 0);
        }
      }
      T rest = c.damping;
      if (lpn.x < -c.hw) {
        lpn.x = -c.hw;
        lvn.x = std::fabs(lvn.x) * rest;
      } else if (lpn.x > c.hw) {
        lpn.x = c.hw;
        lvn.x = -std::fabs(lvn.x) * rest;
      }
      if (lpn.y < -c.hh) {
        lpn.y = -c.hh;
        lvn.y = std::fabs(lvn.y) * rest;
      } else if (lpn.y > c.hh) {
        lpn.y = c.hh;
        lvn.y = -std::fabs(lvn.y) * rest;
      }
      if (lpn.z < -c.hd) {
        lpn.z = -c.hd;
        lvn.z = std::fabs(lvn.z) * rest;
      } else if (lpn.z > c.hd) {
        lpn.z = c.hd;
        lvn.z = -std::fabs(lvn.z) * rest;
      }
      S3(particles[i].pos, local_to_world(lpn));
      S3(particles[i].vel, local_dir_to_world(lvn));
    }
  }

  void stage(uint32_t s) {
    switch (s) {
      case FLUID_STAGE_CLEAR: clear(); break;
      case FLUID_STAGE_BUILD: build(); break;
      case FLUID_STAGE_SCAN: scan(); break;
      case FLUID_STAGE_REORDER: reorder(); break;
      case FLUID_STAGE_DENSITY: density(); break;
      case FLUID_STAGE_XSPH: xsph(); break;
      case FLUID_STAGE_FORCE: force(); break;
      case FLUID_STAGE_PREDICT: predict(); break;
      case FLUID_STAGE_SOLVE: solve(); break;
      case FLUID_STAGE_FINALIZE: finalize(); break;
      case FLUID_STAGE_INTEGRATE: integrate(); break;
      default: break;
    }
  }

  // Candidate / in-range counts of the density sweep of the current state.
  void count_pairs(uint64_t *cand, uint64_t *inr) const {
    uint64_t c0 = 0, c1 = 0;
#pragma omp parallel for reduction(+ : c0, c1) schedule(dynamic, 256)
    for (int64_t si = 0; si < (int64_t)n; ++si) {
      V3<T> pos_i = P3(sorted[si].pos);
      int32_t cl[3];
      position_to_cell(pos_i, cl);
      sweep(cl, [&](uint32_t j) {
        V3<T> r_vec = sub(pos_i, P3(sorted[j].pos));
        c0++;
        if (dot(r_vec, r_vec) < h2) c1++;
      });
    }
    *cand = c0;
    *inr = c1;
  }
};

template <class T>
void set_uniforms(Oracle<T> &o, const GpuSphParams3D *p,
                  const GpuGridParams *g, const GpuContainerGeometry *c,
                  const GpuGravity *grav) {
  o.h = T(p->kernel_radius);
  o.h2 = T(p->kernel_radius_sq);
  o.h5 = T(p->kernel_radius_pow5);
  o.h6 = T(p->kernel_radius_pow6);
  o.h9 = T(p->kernel_radius_pow9);
  o.mass = T(p->mass);
  o.rho0 = T(p->rest_density);
  o.near_k = T(p->near_stiffness);
  o.visc = T(p->viscosity);
  o.dt = T(p->dt);
  o.st = T(p->surface_tension);
  o.delta = T(p->pcisph_delta);
  o.eps = T(p->xsph_epsilon);
  o.gx = g->grid_size_x;
  o.gy = g->grid_size_y;
  o.gz = g->grid_size_z;
  o.cells = g->total_cells;
  o.inv_cell = T(g->inv_cell_size);
  o.ox = T(g->grid_origin_x);
  o.oy = T(g->grid_origin_y);
  o.oz = T(g->grid_origin_z);
  auto v = [](const float *r) { return mk(T(r[0]), T(r[1]), T(r[2])); };
  o.c.hw = T(c->half_width);
  o.c.hh = T(c->half_height);
  o.c.hd = T(c->half_depth);
  o.c.cy = T(c->center_y);
  o.c.f0 = v(c->forward_row0);
  o.c.f1 = v(c->forward_row1);
  o.c.f2 = v(c->forward_row2);
  o.c.i0 = v(c->inverse_row0);
  o.c.i1 = v(c->inverse_row1);
  o.c.i2 = v(c->inverse_row2);
  o.c.wall_stiffness = T(c->wall_stiffness);
  o.c.damping = T(c->damping);
  o.gravity = v(grav->direction);
}

struct Handle {
  int fp64;
  Oracle<float> f;
  Oracle<double> d;
};

// ---- restatement of the reference's host-side parameter derivation ------
// Rust powi(x, n) with a constant n: square-and-multiply, low bit first.
float powi_f32(float x, int n) {
  float res = 1.0f, cur = x;
  bool first = true;
  while (n) {
    if (n & 1) {
      res = first ? cur : res * cur;
      first = false;
    }
    n >>= 1;
    if (n) cur = cur * cur;
  }
  return res;
}

// ---------------------------------------------------------------------------
// Whitewater: spray_emit.wgsl and spray_simulate.wgsl over a sorted particle
// array and its grid (f32 only), plus the ceilings' update of spray.rs:40-51.
// Slots are handed out in sorted order, one spawn at a time: one sequential
// run of the reference's atomicAdd (spray_emit.wgsl:335).
namespace spray {

constexpr float kTwoPi = 6.28318530718f;  // spray_emit.wgsl:114-154
constexpr float kAutoLimitMinFrac = 0.1f;
constexpr uint32_t kEmitMinNeighbors = 10u;
constexpr float kKineticMaxSpeed = 2.5f;
constexpr float kCrestDensityFrac = 0.9f;
constexpr float kCrestVelNormalMin = 0.6f;
constexpr float kCrestDivergenceMin = 0.0f;
constexpr float kCrestDivergenceMax = 0.5f;
constexpr uint32_t kMaxEmitPerFrame = 16u;
constexpr uint32_t kSprayMaxNeighbors = 6u;  // spray_simulate.wgsl:123-135
constexpr uint32_t kBubbleMinNeighbors = 20u;
constexpr float kNewbornSprayTime = 0.15f;
constexpr float kPi = 3.14159265359f;

// spray_emit.wgsl:156-164
uint32_t hash(uint32_t seed) {
  uint32_t x = seed;
  x = x ^ (x >> 16u);
  x = x * 0x45d9f3bu;
  x = x ^ (x >> 16u);
  x = x * 0x45d9f3bu;
  x = x ^ (x >> 16u);
  return x;
}

// :166-168
float hash_float(uint32_t seed) { return (float)(hash(seed) & 0xFFFFu) / 65535.0f; }

// :185-187
float remap_clamp(float x, float lo, float hi) {
  return clampv((x - lo) / (hi - lo), 0.0f, 1.0f);
}

// WGSL u32(f32): saturating
uint32_t to_u32(float v) {
  if (!(v > 0.0f)) return 0u;
  if (v >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)v;
}

struct Grid {
  const GpuGridParams &g;
  // position_to_cell (spray_emit.wgsl:171-174)
  void cell(V3<float> pos, int32_t c[3]) const {
    c[0] = to_i32(std::floor((pos.x - g.grid_origin_x) * g.inv_cell_size));
    c[1] = to_i32(std::floor((pos.y - g.grid_origin_y) * g.inv_cell_size));
    c[2] = to_i32(std::floor((pos.z - g.grid_origin_z) * g.inv_cell_size));
  }
  // is_valid_cell (:180-183)
  bool valid(const int32_t c[3]) const {
    return c[0] >= 0 && c[0] < (int32_t)g.grid_size_x && c[1] >= 0 &&
           c[1] < (int32_t)g.grid_size_y && c[2] >= 0 && c[2] < (int32_t)g.grid_size_z;
  }
  // cell_to_index (:176-178), u32 arithmetic
  uint32_t index(const int32_t c[3]) const {
    return (uint32_t)c[0] + (uint32_t)c[1] * g.grid_size_x +
           (uint32_t)c[2] * g.grid_size_x * g.grid_size_y;
  }
  // the 27-cell loop (dz, dy, dx, then the cell's particles in order)
  template <class F>
  void visit(V3<float> pos, const uint32_t *counts, const uint32_t *starts, F &&f) const {
    int32_t ci[3];
    cell(pos, ci);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int32_t nc[3] = {ci[0] + dx, ci[1] + dy, ci[2] + dz};
          if (!valid(nc)) continue;
          const uint32_t idx = index(nc);
          const uint32_t count = counts[idx];
          const uint32_t start = starts[idx] - count;
          for (uint32_t k = 0; k < count; ++k) f(start + k);
        }
  }
};

// world_to_local (container_common.wgsl:28-35)
V3<float> world_to_local(const GpuContainerGeometry &c, V3<float> pos) {
  const V3<float> p = mk(pos.x, pos.y - c.center_y, pos.z);
  return mk(dot(P3(c.inverse_row0), p), dot(P3(c.inverse_row1), p), dot(P3(c.inverse_row2), p));
}

// spray_simulate.wgsl:142-145
float cap_volume_fraction(float t) {
  const float tc = clampv(t, 0.0f, 1.0f);
  return tc * tc * (3.0f - tc) * 0.25f;
}

}  // namespace spray

// ---------------------------------------------------------------------------
// Marching-cubes density field: mc_anisotropy.wgsl and mc_density.wgsl
// (f32, per voxel / per particle in the reference's loop order).
namespace mc {

constexpr float kPi = 3.14159265359f;

struct M3 {
  float m[3][3];  // m[column][row], as WGSL mat3x3
};

// jacobi_rotate (mc_anisotropy.wgsl:79-107)
void jacobi_rotate(M3 &A, M3 &V, int p, int q, int k) {
  const float apq = A.m[p][q];
  if (std::fabs(apq) < 1e-12f) return;
  const float app = A.m[p][p], aqq = A.m[q][q];
  const float tau = (aqq - app) / (2.0f * apq);
  const float sign_tau = tau < 0.0f ? -1.0f : 1.0f;
  const float t = sign_tau / (std::fabs(tau) + std::sqrt(1.0f + tau * tau));
  const float c = 1.0f / std::sqrt(1.0f + t * t);
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

float smoothstep(float e0, float e1, float x) {
  const float t = clampv((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

}  // namespace mc

}  // namespace

extern "C" {

// spray_emit.wgsl:190-320 for sorted particles 0..n-1: potentials[4i..] =
// trapped air, crest (after its velocity gate), divergence, neighbour count
// (zeros below min_speed); n_emit[i] = spawns; frame_max = the maxima the
// pass reports (f32, 0 when none).  ta_limit / wc_limit are the ceilings.
void oracle_spray_potentials(const void *sorted_particles, uint32_t n, const uint32_t *counts,
                             const uint32_t *starts, const GpuGridParams *grid,
                             const GpuSphParams3D *sph, const GpuSprayParams *prm,
                             float ta_limit, float wc_limit, float *potentials,
                             uint32_t *n_emit, float *frame_max) {
  using namespace spray;
  const Part<float> *sp = (const Part<float> *)sorted_particles;
  const Grid G{*grid};
  frame_max[0] = frame_max[1] = 0.0f;
  const float h = sph->kernel_radius, h_sq = sph->kernel_radius_sq;
  const float poly6_norm = 315.0f / (64.0f * kPi * sph->kernel_radius_pow9);
  for (uint32_t i = 0; i < n; ++i) {
    float *pot = potentials + 4 * (size_t)i;
    pot[0] = pot[1] = pot[2] = pot[3] = 0.0f;
    n_emit[i] = 0u;
    const Part<float> &p = sp[i];
    const V3<float> pos = P3(p.pos), vel = P3(p.vel);
    const float speed = length(vel);
    if (speed < prm->min_speed) continue;
    const V3<float> normal_i = P3(p.normal);
    const float n_len_i = length(normal_i);
    const bool crest_eligible =
        n_len_i > 1e-4f && p.density < sph->rest_density * kCrestDensityFrac;
    V3<float> n_hat_i = mk(0.0f, 1.0f, 0.0f);
    if (crest_eligible) n_hat_i = divs(normal_i, n_len_i);
    float trapped_air = 0.0f, crest = 0.0f, divergence = 0.0f;
    uint32_t n_neighbors = 0u;
    G.visit(pos, counts, starts, [&](uint32_t j) {
      const Part<float> &pj = sp[j];
      const V3<float> r_vec = sub(pos, P3(pj.pos));
      const float r_sq = dot(r_vec, r_vec);
      if (r_sq >= h_sq || r_sq < 1e-12f) return;
      const float r = std::sqrt(r_sq);
      const float w = 1.0f - r / h;
      const V3<float> x_hat = divs(r_vec, r);
      n_neighbors += 1u;
      const float d2 = h_sq - r_sq;
      const float w_vol = (sph->mass / std::max(pj.density, 1e-6f)) * poly6_norm * d2 * d2 * d2;
      const V3<float> v_ij = sub(vel, P3(pj.vel));
      const float v_ij_len = length(v_ij);
      if (v_ij_len > 1e-6f)
        trapped_air += v_ij_len * (1.0f - dot(divs(v_ij, v_ij_len), x_hat)) * w_vol;
      divergence += dot(v_ij, x_hat) * w;
      if (crest_eligible && dot(mk(-x_hat.x, -x_hat.y, -x_hat.z), n_hat_i) < 0.0f) {
        const V3<float> normal_j = P3(pj.normal);
        const float n_len_j = length(normal_j);
        if (n_len_j > 1e-4f)
          crest += (1.0f - dot(n_hat_i, divs(normal_j, n_len_j))) * w_vol;
      }
    });
    pot[0] = trapped_air;
    pot[1] = crest;
    pot[2] = divergence;
    pot[3] = (float)n_neighbors;
    if (n_neighbors < kEmitMinNeighbors) continue;
    if (dot(divs(vel, speed), n_hat_i) < kCrestVelNormalMin) crest = 0.0f;
    pot[1] = crest;
    if (trapped_air > 0.0f) frame_max[0] = std::max(frame_max[0], trapped_air);
    if (crest > 0.0f) frame_max[1] = std::max(frame_max[1], crest);
    const float phi_div = remap_clamp(divergence, kCrestDivergenceMin, kCrestDivergenceMax);
    const float ta_hi = std::max(ta_limit, 1e-4f);
    const float wc_hi = std::max(wc_limit, 1e-4f);
    const float phi_ta = remap_clamp(trapped_air, kAutoLimitMinFrac * ta_hi, ta_hi);
    const float phi_wc = remap_clamp(crest, kAutoLimitMinFrac * wc_hi, wc_hi) * phi_div;
    const float phi_k = remap_clamp(speed, prm->min_speed, kKineticMaxSpeed);
    const float rate =
        phi_k * (prm->k_trapped_air * phi_ta + prm->k_wave_crest * phi_wc) * prm->emission_rate;
    const float expected = rate * prm->dt;
    const uint32_t base_seed = i + prm->frame_count * 7919u;
    uint32_t ne = to_u32(std::floor(expected));
    if (hash_float(base_seed * 13u + 5u) < expected - std::floor(expected)) ne = ne + 1u;
    n_emit[i] = std::min(ne, kMaxEmitPerFrame);
  }
}

// spray_emit.wgsl:322-372: the spawns of sorted particles 0..n-1 (n_emit[i]
// each), written to ring slots (head + k) % max_particles in sorted order;
// returns the new head.
uint32_t oracle_spray_spawn(const void *sorted_particles, uint32_t n,
                            const GpuSphParams3D *sph, const GpuSprayParams *prm,
                            const uint32_t *n_emit, uint32_t head, GpuSprayParticle *ring) {
  using namespace spray;
  const Part<float> *sp = (const Part<float> *)sorted_particles;
  const float h = sph->kernel_radius;
  for (uint32_t i = 0; i < n; ++i) {
    if (n_emit[i] == 0u) continue;
    const Part<float> &p = sp[i];
    const V3<float> pos0 = P3(p.pos), vel = P3(p.vel);
    const float speed = length(vel);
    const float phi_k = remap_clamp(speed, prm->min_speed, kKineticMaxSpeed);
    const uint32_t base_seed = i + prm->frame_count * 7919u;
    const V3<float> v_dir = divs(vel, speed);
    V3<float> ref_axis = mk(0.0f, 1.0f, 0.0f);
    if (std::fabs(v_dir.y) > 0.99f) ref_axis = mk(1.0f, 0.0f, 0.0f);
    const V3<float> e1 = normalize(cross(v_dir, ref_axis));
    const V3<float> e2 = cross(v_dir, e1);
    for (uint32_t e = 0; e < n_emit[i]; ++e) {
      const uint32_t slot = head % prm->max_particles;
      head = head + 1u;
      const uint32_t seed = base_seed + e * 3571u;
      const float r0 = hash_float(seed), r1 = hash_float(seed + 1u);
      const float r2 = hash_float(seed + 2u), r3 = hash_float(seed + 3u);
      const float r4 = hash_float(seed + 4u) * 2.0f - 1.0f;
      const float r5 = hash_float(seed + 5u) * 2.0f - 1.0f;
      const float r6 = hash_float(seed + 6u) * 2.0f - 1.0f;
      const float disc_r = h * 0.5f * std::sqrt(r0);
      const float theta = kTwoPi * r1;
      const float along = r2 * prm->dt * speed;
      const V3<float> pos = add(add(add(pos0, muls(e1, disc_r * std::cos(theta))),
                                    muls(e2, disc_r * std::sin(theta))),
                                muls(v_dir, along));
      const V3<float> jitter = muls(mk(r4, r5, r6), prm->velocity_jitter);
      const V3<float> spawn_vel = add(muls(vel, prm->speed_multiplier), jitter);
      const float lt_min = prm->lifetime * (1.0f - prm->lifetime_variation);
      const float lt_span = 2.0f * prm->lifetime * prm->lifetime_variation;
      const float lt = lt_min + phi_k * r3 * lt_span;
      GpuSprayParticle &o = ring[slot];
      o.pos_x = pos.x;
      o.pos_y = pos.y;
      o.pos_z = pos.z;
      o.lifetime = lt;
      o.vel_x = spawn_vel.x;
      o.vel_y = spawn_vel.y;
      o.vel_z = spawn_vel.z;
      o.max_lifetime = lt;
      o.kind = 0u;
      o.age = 0.0f;
      o._pad[0] = o._pad[1] = 0.0f;
    }
  }
  return head;
}

// spray_simulate.wgsl:147-281 over ring[0..max_particles); live = total,
// spray, foam, bubble counts after the pass.
void oracle_spray_simulate(GpuSprayParticle *ring, const void *sorted_particles,
                           const uint32_t *counts, const uint32_t *starts,
                           const GpuGridParams *grid, const GpuSphParams3D *sph,
                           const GpuContainerGeometry *cont, const GpuSprayParams *prm,
                           uint32_t live[4]) {
  using namespace spray;
  const Part<float> *sp = (const Part<float> *)sorted_particles;
  const Grid G{*grid};
  live[0] = live[1] = live[2] = live[3] = 0u;
  const float dt = prm->dt, h = sph->kernel_radius, h_sq = sph->kernel_radius_sq;
  for (uint32_t idx = 0; idx < prm->max_particles; ++idx) {
    GpuSprayParticle &p = ring[idx];
    if (p.lifetime <= 0.0f) continue;
    const V3<float> pos = mk(p.pos_x, p.pos_y, p.pos_z);
    V3<float> vel = mk(p.vel_x, p.vel_y, p.vel_z);
    uint32_t neighbor_count = 0u;
    float w_sum = 0.0f;
    V3<float> v_avg = mk(0.0f, 0.0f, 0.0f);
    G.visit(pos, counts, starts, [&](uint32_t j) {
      const V3<float> r_vec = sub(pos, P3(sp[j].pos));
      const float r_sq = dot(r_vec, r_vec);
      if (r_sq < h_sq) {
        const float w = 1.0f - std::sqrt(r_sq) / h;
        neighbor_count = neighbor_count + 1u;
        w_sum = w_sum + w;
        v_avg = add(v_avg, muls(P3(sp[j].vel), w));
      }
    });
    if (w_sum > 1e-6f) v_avg = divs(v_avg, w_sum);
    uint32_t kind = 1u;
    if (neighbor_count <= kSprayMaxNeighbors)
      kind = 0u;
    else if (neighbor_count >= kBubbleMinNeighbors)
      kind = 2u;
    if (kind == 1u) {
      const V3<float> lp = world_to_local(*cont, pos);
      const float half[3] = {cont->half_width, cont->half_height, cont->half_depth};
      const float l[3] = {lp.x, lp.y, lp.z};
      float missing = 0.0f;
      for (int a = 0; a < 3; ++a) {
        const float dist_lo = half[a] + l[a], dist_hi = half[a] - l[a];
        if (dist_lo < h) missing += cap_volume_fraction(1.0f - dist_lo / h);
        if (dist_hi < h) missing += cap_volume_fraction(1.0f - dist_hi / h);
      }
      const float visible = std::max(1.0f - missing, 0.25f);
      if (to_u32(std::nearbyint((float)neighbor_count / visible)) >= kBubbleMinNeighbors)
        kind = 2u;
    }
    if (kind == 1u && p.age < kNewbornSprayTime) kind = 0u;
    const V3<float> gravity = mk(0.0f, prm->gravity_y, 0.0f);
    float lifetime_decay = 0.0f;
    if (kind == 0u) {
      vel = add(vel, smul(dt, sub(gravity, smul(prm->drag, vel))));
    } else if (kind == 1u) {
      vel = v_avg;
      lifetime_decay = 1.0f;
    } else {
      vel = add(vel, smul(dt, smul(-prm->bubble_buoyancy, gravity)));
      const float t = clampv(prm->bubble_drag, 0.0f, 1.0f);
      vel = add(muls(vel, 1.0f - t), muls(v_avg, t));
    }
    const V3<float> new_pos = add(pos, smul(dt, vel));
    p.pos_x = new_pos.x;
    p.pos_y = new_pos.y;
    p.pos_z = new_pos.z;
    p.vel_x = vel.x;
    p.vel_y = vel.y;
    p.vel_z = vel.z;
    p.kind = kind;
    p.age = p.age + dt;
    p.lifetime = p.lifetime - dt * lifetime_decay;
    const V3<float> lp = world_to_local(*cont, new_pos);
    const bool inside = lp.x >= -cont->half_width && lp.x <= cont->half_width &&
                        lp.y >= -cont->half_height && lp.y <= cont->half_height &&
                        lp.z >= -cont->half_depth && lp.z <= cont->half_depth;
    if (!inside) p.lifetime = 0.0f;
    if (p.lifetime > 0.0f) {
      live[0] += 1u;
      live[1u + kind] += 1u;
    }
  }
}

// mc_anisotropy.wgsl:110-276 for sorted particles 0..n-1: out = 12 floats
// per particle (ParticleAniso).
void oracle_mc_aniso(const void *sorted_particles, uint32_t n, const uint32_t *counts,
                     const uint32_t *starts, const GpuGridParams *grid,
                     const GpuAnisoParams *prm, float *out) {
  using namespace mc;
  const Part<float> *sp = (const Part<float> *)sorted_particles;
  const spray::Grid G{*grid};
  const float r_support = prm->support_radius, h = prm->h_mc;
  const float peak = 315.0f / (64.0f * kPi * h * h * h);
  const int cell_radius = (int)std::ceil(r_support * grid->inv_cell_size);
  for (uint32_t i = 0; i < n; ++i) {
    const V3<float> xi = P3(sp[i].pos);
    float w_sum = 0.0f, n_count = 0.0f;
    V3<float> m1 = mk(0.0f, 0.0f, 0.0f);
    float cxx = 0.0f, cyy = 0.0f, czz = 0.0f, cxy = 0.0f, cxz = 0.0f, cyz = 0.0f;
    int32_t cc[3];
    G.cell(xi, cc);
    for (int dz = -cell_radius; dz <= cell_radius; ++dz)
      for (int dy = -cell_radius; dy <= cell_radius; ++dy)
        for (int dx = -cell_radius; dx <= cell_radius; ++dx) {
          const int32_t c[3] = {cc[0] + dx, cc[1] + dy, cc[2] + dz};
          if (!G.valid(c)) continue;
          const uint32_t idx = G.index(c), count = counts[idx];
          if (count == 0u) continue;
          const uint32_t start = starts[idx] - count;
          for (uint32_t k = 0; k < count; ++k) {
            const V3<float> d = sub(P3(sp[start + k].pos), xi);
            const float r = length(d);
            if (r < r_support) {
              const float u = r / r_support;
              const float w = 1.0f - u * u * u;
              w_sum += w;
              n_count += 1.0f;
              m1 = add(m1, smul(w, d));
              cxx += w * d.x * d.x;
              cyy += w * d.y * d.y;
              czz += w * d.z * d.z;
              cxy += w * d.x * d.y;
              cxz += w * d.x * d.z;
              cyz += w * d.y * d.z;
            }
          }
        }
    const float t_n = smoothstep(4.0f, 12.0f, n_count);
    const float s_eff = prm->strength * t_n;
    const float inv_w = 1.0f / std::max(w_sum, 1e-9f);
    const V3<float> mu = muls(m1, inv_w);
    M3 A, V;
    A.m[0][0] = cxx * inv_w - mu.x * mu.x;
    A.m[0][1] = cxy * inv_w - mu.x * mu.y;
    A.m[0][2] = cxz * inv_w - mu.x * mu.z;
    A.m[1][0] = cxy * inv_w - mu.x * mu.y;
    A.m[1][1] = cyy * inv_w - mu.y * mu.y;
    A.m[1][2] = cyz * inv_w - mu.y * mu.z;
    A.m[2][0] = cxz * inv_w - mu.x * mu.z;
    A.m[2][1] = cyz * inv_w - mu.y * mu.z;
    A.m[2][2] = czz * inv_w - mu.z * mu.z;
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) V.m[c][r] = c == r ? 1.0f : 0.0f;
    for (int sweep = 0; sweep < 5; ++sweep) {
      jacobi_rotate(A, V, 0, 1, 2);
      jacobi_rotate(A, V, 0, 2, 1);
      jacobi_rotate(A, V, 1, 2, 0);
    }
    float e[3] = {A.m[0][0], A.m[1][1], A.m[2][2]};
    V3<float> v[3] = {P3(V.m[0]), P3(V.m[1]), P3(V.m[2])};
    if (e[0] < e[1]) { std::swap(e[0], e[1]); std::swap(v[0], v[1]); }
    if (e[0] < e[2]) { std::swap(e[0], e[2]); std::swap(v[0], v[2]); }
    if (e[1] < e[2]) { std::swap(e[1], e[2]); std::swap(v[1], v[2]); }
    const float s1_raw = std::sqrt(std::max(e[0], 0.0f));
    const float s3_raw = std::sqrt(std::max(e[2], 0.0f));
    float flatness = 0.0f;
    if (s1_raw > 1e-7f) {
      const float roundness = s3_raw / s1_raw;
      flatness = 1.0f - roundness * roundness;
    }
    V3<float> shift = smul(prm->lambda * s_eff * flatness, mu);
    const float shift_len = length(shift);
    if (shift_len > prm->max_shift) shift = muls(shift, prm->max_shift / shift_len);
    const V3<float> center = add(xi, shift);
    float a[3] = {1.0f, 1.0f, 1.0f};
    const float s1 = s1_raw;
    if (s1 > 1e-7f && s_eff > 0.0f) {
      const float s_min = s1 / prm->kr;
      const float s2 = std::max(std::sqrt(std::max(e[1], 0.0f)), s_min);
      const float s3 = std::max(s3_raw, s_min);
      const float norm = std::pow(s1 * s2 * s3, 1.0f / 3.0f);
      const float sv[3] = {s1, s2, s3};
      for (int k = 0; k < 3; ++k) {
        a[k] = sv[k] / norm;
        a[k] = std::pow(a[k], s_eff);
        a[k] = clampv(a[k], 0.25f, prm->max_stretch);
      }
    }
    const float ir[3] = {1.0f / (a[0] * h), 1.0f / (a[1] * h), 1.0f / (a[2] * h)};
    auto gent = [&](int r, int c) {
      const float *vr0 = &v[0].x, *vr1 = &v[1].x, *vr2 = &v[2].x;
      return ir[0] * vr0[r] * vr0[c] + ir[1] * vr1[r] * vr1[c] + ir[2] * vr2[r] * vr2[c];
    };
    const float amplitude = peak / std::max(a[0] * a[1] * a[2], 1.0f);
    const float reach = h * std::max(a[0], std::max(a[1], a[2]));
    float *o = out + 12 * (size_t)i;
    const float rec[12] = {gent(0, 0), gent(0, 1), gent(0, 2), center.x,
                           gent(1, 1), gent(1, 2), gent(2, 2), center.y,
                           center.z,   reach,      amplitude,  0.0f};
    std::memcpy(o, rec, sizeof rec);
  }
}

// mc_density.wgsl:133-230 over the G^3 voxels (x fastest); aniso = 12
// floats per sorted particle (read when aniso_params->enabled).
void oracle_mc_density(const void *sorted_particles, const uint32_t *counts,
                       const uint32_t *starts, const GpuGridParams *grid,
                       const GpuMcGridParams *prm, const GpuAnisoParams *aprm,
                       const GpuContainerGeometry *cont, const float *aniso, float *field) {
  const Part<float> *sp = (const Part<float> *)sorted_particles;
  const spray::Grid G{*grid};
  const uint32_t N = prm->grid_size;
  const float h = prm->kernel_radius, h_sq = h * h;
  const bool aniso_on = aprm->enabled != 0u;
  const float coeff = 315.0f / (64.0f * 3.14159265359f * std::pow(h, 9.0f));
  float search_reach = h;
  if (aniso_on) search_reach = h * aprm->max_stretch + aprm->max_shift;
  const int cell_radius = (int)std::ceil(search_reach * grid->inv_cell_size);
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t vz = 0; vz < (int64_t)N; ++vz)
    for (uint32_t vy = 0; vy < N; ++vy)
      for (uint32_t vx = 0; vx < N; ++vx) {
        const float cs = prm->cell_size;
        const V3<float> wp = mk(prm->grid_min[0] + ((float)vx + 0.5f) * cs,
                                prm->grid_min[1] + ((float)vy + 0.5f) * cs,
                                prm->grid_min[2] + ((float)vz + 0.5f) * cs);
        int32_t cc[3];
        G.cell(wp, cc);
        float density = 0.0f;
        for (int dz = -cell_radius; dz <= cell_radius; ++dz)
          for (int dy = -cell_radius; dy <= cell_radius; ++dy)
            for (int dx = -cell_radius; dx <= cell_radius; ++dx) {
              const int32_t c[3] = {cc[0] + dx, cc[1] + dy, cc[2] + dz};
              if (!G.valid(c)) continue;
              const uint32_t idx = G.index(c), count = counts[idx];
              if (count == 0u) continue;
              const uint32_t start = starts[idx] - count;
              for (uint32_t k = 0; k < count; ++k) {
                const uint32_t j = start + k;
                if (aniso_on) {
                  const float *an = aniso + 12 * (size_t)j;
                  const V3<float> diff = sub(wp, mk(an[3], an[7], an[8]));
                  const float r_sq = dot(diff, diff);
                  const float reach = an[9];
                  if (r_sq < reach * reach) {
                    const V3<float> gd = mk(an[0] * diff.x + an[1] * diff.y + an[2] * diff.z,
                                            an[1] * diff.x + an[4] * diff.y + an[5] * diff.z,
                                            an[2] * diff.x + an[5] * diff.y + an[6] * diff.z);
                    const float q_sq = dot(gd, gd);
                    if (q_sq < 1.0f) {
                      const float f = 1.0f - q_sq;
                      density += an[10] * f * f * f;
                    }
                  }
                } else {
                  const V3<float> diff = sub(wp, P3(sp[j].pos));
                  const float r_sq = dot(diff, diff);
                  if (r_sq < h_sq) {
                    const float d = h_sq - r_sq;
                    density += coeff * d * d * d;
                  }
                }
              }
            }
        const V3<float> local = spray::world_to_local(*cont, wp);
        float out;
        const bool inside = local.x >= -cont->half_width && local.x <= cont->half_width &&
                            local.y >= -cont->half_height && local.y <= cont->half_height &&
                            local.z >= -cont->half_depth && local.z <= cont->half_depth;
        if (cont->clip_enabled != 0u && !inside) {
          out = -1.0f;
        } else {
          const float dist[5] = {local.x + cont->half_width, cont->half_width - local.x,
                                 local.y + cont->half_height, local.z + cont->half_depth,
                                 cont->half_depth - local.z};
          float gamma = 1.0f;
          for (int k = 0; k < 5; ++k)
            if (dist[k] < h) gamma *= 0.5f + 0.5f * clampv(dist[k] / h, 0.0f, 1.0f);
          out = density / std::max(gamma, 0.15f);
        }
        field[(size_t)vx + (size_t)vy * N + (size_t)vz * N * N] = out;
      }
}

// update_auto_limit (src/simulation/spray.rs:40-51)
float oracle_spray_update_limit(float current, float frame_max, float floor_v) {
  if (frame_max <= 1e-6f) return current;
  const float alpha = frame_max > current ? 0.05f : 0.01f;
  return std::max(current + alpha * (frame_max - current), floor_v);
}

// hash_float (spray_emit.wgsl:166-168)
float oracle_spray_hash_float(uint32_t seed) { return spray::hash_float(seed); }
uint32_t oracle_spray_hash(uint32_t seed) { return spray::hash(seed); }

// compute_rest_density (src/state/simulation.rs:402-425), f32.
float oracle_rest_density_f32(float h, float mass) {
  float spacing = h * 0.6f;
  int cells = (int)std::ceil(h / spacing) + 1;
  float pi = 3.14159265359f;
  float h2 = h * h;
  float h9 = powi_f32(h, 9);
  float scale = 315.0f / (64.0f * pi * h9);
  float density = 0.0f;
  for (int dz = -cells; dz <= cells; ++dz)
    for (int dy = -cells; dy <= cells; ++dy)
      for (int dx = -cells; dx <= cells; ++dx) {
        float a = (float)dx * spacing, b = (float)dy * spacing,
              c = (float)dz * spacing;
        float r_sq = a * a + b * b + c * c;
        if (r_sq < h2) {
          float diff = h2 - r_sq;
          density += mass * scale * diff * diff * diff;
        }
      }
  return density;
}

// compute_pcisph_delta (src/state/simulation.rs:429-480), f32.
float oracle_pcisph_delta_f32(float h, float mass, float rest_density,
                              float dt) {
  float spacing = h * 0.6f;
  int cells = (int)std::ceil(h / spacing) + 1;
  float sg[3] = {0.0f, 0.0f, 0.0f};
  float sgsq = 0.0f;
  float pi = 3.14159265359f;
  float h6 = powi_f32(h, 6);
  for (int dz = -cells; dz <= cells; ++dz)
    for (int dy = -cells; dy <= cells; ++dy)
      for (int dx = -cells; dx <= cells; ++dx) {
        if (dx == 0 && dy == 0 && dz == 0) continue;
        float rv[3] = {(float)dx * spacing, (float)dy * spacing,
                       (float)dz * spacing};
        float r_sq = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
        if (r_sq < h * h && r_sq > 1e-12f) {
          float r = std::sqrt(r_sq);
          float gm = 45.0f / (pi * h6) * (h - r) * (h - r);
          float g[3] = {gm * rv[0] / r, gm * rv[1] / r, gm * rv[2] / r};
          sg[0] += g[0];
          sg[1] += g[1];
          sg[2] += g[2];
          sgsq += g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
        }
      }
  float sgd = sg[0] * sg[0] + sg[1] * sg[1] + sg[2] * sg[2];
  float beta = dt * dt * mass * mass / (rest_density * rest_density);
  float denom = beta * (sgd + sgsq);
  float omega = 0.2f;
  return std::fabs(denom) > 1e-10f ? omega / denom : 0.0f;
}

float oracle_powi_f32(float x, int n) { return powi_f32(x, n); }

// particles: n records of 16 values in the handle's precision.
void *oracle_create(int fp64, const void *particles, uint32_t n,
                    const GpuSphParams3D *p, const GpuGridParams *g,
                    const GpuContainerGeometry *c, const GpuGravity *grav,
                    const GpuMouseForce *mouse, const GpuRigidBody *rb,
                    uint32_t iters, const float *sdf, uint32_t sdf_res) {
  Handle *hd = new Handle();
  hd->fp64 = fp64;
  auto init = [&](auto &o, auto tag) {
    using T = decltype(tag);
    set_uniforms(o, p, g, c, grav);
    o.n = n;
    o.iters = iters < 1 ? 1 : iters;
    o.mouse = *mouse;
    o.rb = *rb;
    o.sdf_res = sdf ? sdf_res : 0;
    if (sdf) o.sdf.assign(sdf, sdf + (size_t)sdf_res * sdf_res * sdf_res);
    o.particles.resize(n);
    std::memcpy(o.particles.data(), particles, sizeof(Part<T>) * n);
    o.sorted.assign(n, Part<T>{});
    o.counts.assign(o.cells, 0u);
    o.starts.assign(o.cells, 0u);
    o.offsets.assign(o.cells, 0u);
    o.pci.assign(n, 0u);
    o.s2o.assign(n, 0u);
    o.sig.assign(n, 0ull);
    o.pa.assign(n, Pred<T>{});
    o.pb.assign(n, Pred<T>{});
  };
  if (fp64)
    init(hd->d, double());
  else
    init(hd->f, float());
  return hd;
}

void oracle_destroy(void *h) { delete (Handle *)h; }

// The branch-decision record per original index (n values).
void oracle_decision_record(void *h, uint64_t *out) {
  Handle *hd = (Handle *)h;
  const std::vector<uint64_t> &v = hd->fp64 ? hd->d.sig : hd->f.sig;
  std::copy(v.begin(), v.end(), out);
}

// Sum of the branch-decision record since creation (see decision()).
uint64_t oracle_decision_signature(void *h) {
  Handle *hd = (Handle *)h;
  const std::vector<uint64_t> &v = hd->fp64 ? hd->d.sig : hd->f.sig;
  uint64_t sum = 0;
  for (uint64_t x : v) sum += x;
  return sum;
}

void oracle_stage(void *h, uint32_t stage) {
  Handle *hd = (Handle *)h;
  if (hd->fp64)
    hd->d.stage(stage);
  else
    hd->f.stage(stage);
}

void oracle_step(void *h, uint32_t steps) {
  for (uint32_t k = 0; k < steps; ++k)
    for (uint32_t s = 0; s < FLUID_STAGE_COUNT; ++s) oracle_stage(h, s);
}

void oracle_set_params(void *h, const GpuSphParams3D *p,
                       const GpuContainerGeometry *c, const GpuGravity *grav,
                       uint32_t iters) {
  Handle *hd = (Handle *)h;
  auto upd = [&](auto &o) {
    GpuGridParams g{};
    g.grid_size_x = o.gx;
    g.grid_size_y = o.gy;
    g.grid_size_z = o.gz;
    g.total_cells = o.cells;
    g.inv_cell_size = (float)o.inv_cell;
    g.grid_origin_x = (float)o.ox;
    g.grid_origin_y = (float)o.oy;
    g.grid_origin_z = (float)o.oz;
    set_uniforms(o, p, &g, c, grav);
    o.iters = iters < 1 ? 1 : iters;
  };
  if (hd->fp64)
    upd(hd->d);
  else
    upd(hd->f);
}

// which: 0 canonical particles, 1 sorted particles (n x 16 values)
void oracle_get_particles(void *h, int which, void *out) {
  Handle *hd = (Handle *)h;
  if (hd->fp64) {
    auto &v = which ? hd->d.sorted : hd->d.particles;
    std::memcpy(out, v.data(), v.size() * sizeof(v[0]));
  } else {
    auto &v = which ? hd->f.sorted : hd->f.particles;
    std::memcpy(out, v.data(), v.size() * sizeof(v[0]));
  }
}

void oracle_set_particles(void *h, const void *in) {
  Handle *hd = (Handle *)h;
  if (hd->fp64)
    std::memcpy(hd->d.particles.data(), in,
                hd->d.particles.size() * sizeof(hd->d.particles[0]));
  else
    std::memcpy(hd->f.particles.data(), in,
                hd->f.particles.size() * sizeof(hd->f.particles[0]));
}

// which: 0 counts, 1 starts, 2 particle_cell_indices, 3 sorted_to_orig,
//        4 offsets
void oracle_get_u32(void *h, int which, uint32_t *out) {
  Handle *hd = (Handle *)h;
  auto get = [&](auto &o) {
    const std::vector<uint32_t> *v[5] = {&o.counts, &o.starts, &o.pci, &o.s2o,
                                         &o.offsets};
    std::memcpy(out, v[which]->data(), v[which]->size() * 4);
  };
  if (hd->fp64)
    get(hd->d);
  else
    get(hd->f);
}

// which: 0 = sorted_predicted_a, 1 = b (n x 8 values)
void oracle_get_pred(void *h, int which, void *out) {
  Handle *hd = (Handle *)h;
  if (hd->fp64) {
    auto &v = which ? hd->d.pb : hd->d.pa;
    std::memcpy(out, v.data(), v.size() * sizeof(v[0]));
  } else {
    auto &v = which ? hd->f.pb : hd->f.pa;
    std::memcpy(out, v.data(), v.size() * sizeof(v[0]));
  }
}

void oracle_get_accum(void *h, int32_t *out8) {
  Handle *hd = (Handle *)h;
  std::memcpy(out8, hd->fp64 ? hd->d.accum : hd->f.accum, 32);
}

void oracle_count_pairs(void *h, uint64_t *cand, uint64_t *inr) {
  Handle *hd = (Handle *)h;
  if (hd->fp64)
    hd->d.count_pairs(cand, inr);
  else
    hd->f.count_pairs(cand, inr);
}

}  // extern "C"
