// params.cpp — host-side parameter derivation of the reference
// (src/state/simulation.rs) and the device uniform block.  Everything here is
// f32 arithmetic in the reference's operation order.  Built without FMA
// contraction (see the Makefile) so each expression rounds as written.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "fluid_internal.h"

namespace fluid {

namespace {

// Rust's f32::powi with a constant exponent: square-and-multiply from the
// low bit (the LLVM expansion of llvm.powi), e.g. x^9 = x * ((x^2)^2)^2.
float powi(float x, int n) {
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

// compute_rest_density (src/state/simulation.rs:402-425)
float rest_density(float h, float mass) {
  const float spacing = h * 0.6f;
  const int cells = (int)std::ceil(h / spacing) + 1;
  const float h2 = h * h;
  const float scale = 315.0f / (64.0f * kPI * powi(h, 9));
  float density = 0.0f;
  for (int dz = -cells; dz <= cells; ++dz)
    for (int dy = -cells; dy <= cells; ++dy)
      for (int dx = -cells; dx <= cells; ++dx) {
        const float a = (float)dx * spacing, b = (float)dy * spacing,
                    c = (float)dz * spacing;
        const float r_sq = a * a + b * b + c * c;
        if (r_sq < h2) {
          const float diff = h2 - r_sq;
          density += mass * scale * diff * diff * diff;
        }
      }
  return density;
}

// compute_pcisph_delta (src/state/simulation.rs:429-480)
float pcisph_delta(float h, float mass, float rho0, float dt) {
  const float spacing = h * 0.6f;
  const int cells = (int)std::ceil(h / spacing) + 1;
  const float h6 = powi(h, 6);
  float sg[3] = {0.0f, 0.0f, 0.0f}, sgsq = 0.0f;
  for (int dz = -cells; dz <= cells; ++dz)
    for (int dy = -cells; dy <= cells; ++dy)
      for (int dx = -cells; dx <= cells; ++dx) {
        if (dx == 0 && dy == 0 && dz == 0) continue;
        const float rv[3] = {(float)dx * spacing, (float)dy * spacing,
                             (float)dz * spacing};
        const float r_sq = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
        if (r_sq < h * h && r_sq > 1e-12f) {
          const float r = std::sqrt(r_sq);
          const float gm = 45.0f / (kPI * h6) * (h - r) * (h - r);
          const float g[3] = {gm * rv[0] / r, gm * rv[1] / r, gm * rv[2] / r};
          sg[0] += g[0];
          sg[1] += g[1];
          sg[2] += g[2];
          sgsq += g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
        }
      }
  const float sgd = sg[0] * sg[0] + sg[1] * sg[1] + sg[2] * sg[2];
  const float beta = dt * dt * mass * mass / (rho0 * rho0);
  const float denom = beta * (sgd + sgsq);
  return std::fabs(denom) > 1e-10f ? 0.2f / denom : 0.0f;
}

}  // namespace

// Grid sizing of SphSimulation3DGrid::new (sph_3d_grid.rs:121-143).
GpuGridParams make_grid(float h, float bounds_extent, uint32_t n) {
  GpuGridParams g;
  std::memset(&g, 0, sizeof(g));
  const float cell = h;
  const uint32_t G = (uint32_t)std::ceil((2.0f * bounds_extent) / cell) + 2u;
  g.grid_size_x = g.grid_size_y = g.grid_size_z = G;
  g.total_cells = G * G * G;
  g.cell_size = cell;
  g.inv_cell_size = 1.0f / cell;
  g.grid_origin_x = g.grid_origin_y = g.grid_origin_z = -bounds_extent - cell;
  g.num_particles = n;
  return g;
}

void fill_uniforms(DevUniforms &u, const GpuSphParams3D &p,
                   const GpuGridParams &g, const GpuContainerGeometry &c,
                   const GpuGravity &grav, const GpuMouseForce &mouse,
                   const GpuRigidBody &rb, uint32_t sdf_res, uint32_t cell_offset) {
  std::memset(&u, 0, sizeof(u));
  u.cell_offset = cell_offset;
  u.h = p.kernel_radius;
  u.h2 = p.kernel_radius_sq;
  u.h6 = p.kernel_radius_pow6;
  u.h9 = p.kernel_radius_pow9;
  u.mass = p.mass;
  u.rho0 = p.rest_density;
  u.near_k = p.near_stiffness;
  u.visc = p.viscosity;
  u.dt = p.dt;
  u.st = p.surface_tension;
  u.delta = p.pcisph_delta;
  u.eps = p.xsph_epsilon;
  u.poly6_scale = 315.0f / (64.0f * kPI * p.kernel_radius_pow9);
  u.near_scale = 15.0f / (kPI * p.kernel_radius_pow6);
  u.grad_scale = 45.0f / (kPI * p.kernel_radius_pow6);
  u.akinci_scale = 32.0f / (kPI * p.kernel_radius_pow9);
  u.h6_64 = p.kernel_radius_pow6 / 64.0f;
  u.half_h = p.kernel_radius * 0.5f;
  const float h8 =
      p.kernel_radius_pow6 * p.kernel_radius * p.kernel_radius;
  u.normal_scale = 945.0f / (32.0f * kPI * h8);
  u.m_over_rho0 = p.mass / p.rest_density;
  u.min_dist = p.kernel_radius * 0.1f;
  u.visc_eps = 0.01f * p.kernel_radius_sq;
  u.max_accel = 0.25f * p.kernel_radius / (p.dt * p.dt);
  u.max_accel_sq = u.max_accel * u.max_accel;
  u.boundary_layer = p.kernel_radius * 0.4f;
  u.interact_range = p.kernel_radius * 0.7f;
  {
    // step from min_dist^2 to the first float whose rounded root reaches
    // min_dist (a few ulps either way)
    float t = u.min_dist * u.min_dist;
    while (t > 0.0f && std::sqrt(std::nextafter(t, 0.0f)) >= u.min_dist) t = std::nextafter(t, 0.0f);
    while (std::sqrt(t) < u.min_dist) t = std::nextafter(t, INFINITY);
    u.close_r2 = t;
  }
  // Predicted positions of iteration k >= 1 differ from iteration 0's by
  // dt^2 clamp(a_p), |clamp(a_p)| <= max_accel, for both particles of a pair.
  const double skin = (double)p.kernel_radius +
                      2.0 * (double)p.dt * (double)p.dt * (double)u.max_accel;
  u.solve_skin2 = (float)(skin * skin * (1.0 + 1e-3));
  u.gx = g.grid_size_x;
  u.gy = g.grid_size_y;
  u.gz = g.grid_size_z;
  u.cells = g.total_cells;
  u.gxy = g.grid_size_x * g.grid_size_y;
  u.inv_cell = g.inv_cell_size;
  u.ox = g.grid_origin_x;
  u.oy = g.grid_origin_y;
  u.oz = g.grid_origin_z;
  u.hw = c.half_width;
  u.hh = c.half_height;
  u.hd = c.half_depth;
  u.cy = c.center_y;
  std::memcpy(u.fw[0], c.forward_row0, 16);
  std::memcpy(u.fw[1], c.forward_row1, 16);
  std::memcpy(u.fw[2], c.forward_row2, 16);
  std::memcpy(u.iv[0], c.inverse_row0, 16);
  std::memcpy(u.iv[1], c.inverse_row1, 16);
  std::memcpy(u.iv[2], c.inverse_row2, 16);
  u.wall_k = c.wall_stiffness;
  u.damping = c.damping;
  u.grav[0] = grav.direction[0];
  u.grav[1] = grav.direction[1];
  u.grav[2] = grav.direction[2];
  u.sdf_res = sdf_res;
  u.mouse = mouse;
  u.rb = rb;
}

}  // namespace fluid

extern "C" {

int fluid_derive_sph_params(float kernel_radius, float stiffness,
                            float near_stiffness, float viscosity, float mass,
                            float surface_tension, float xsph_epsilon,
                            float dt, uint32_t num_particles,
                            float rest_density_override, GpuSphParams3D *out) {
  if (!out || !(kernel_radius > 0.0f) || !(dt > 0.0f)) return FLUID_ERR_INVALID_ARG;
  using fluid::powi;
  const float h = kernel_radius;
  const float rho0 = rest_density_override > 0.0f
                         ? rest_density_override
                         : fluid::rest_density(h, mass);
  GpuSphParams3D p;
  std::memset(&p, 0, sizeof(p));
  p.kernel_radius = h;
  p.kernel_radius_sq = h * h;
  p.kernel_radius_pow5 = powi(h, 5);
  p.kernel_radius_pow6 = powi(h, 6);
  p.kernel_radius_pow9 = powi(h, 9);
  p.mass = mass;
  p.rest_density = rho0;
  p.stiffness = stiffness;
  p.near_stiffness = near_stiffness;
  p.viscosity = viscosity;
  p.dt = dt;
  p.num_particles = num_particles;
  p.surface_tension = surface_tension;
  p.pcisph_delta = fluid::pcisph_delta(h, mass, rho0, dt);
  p.xsph_epsilon = xsph_epsilon;
  *out = p;
  return FLUID_OK;
}

int fluid_derive_container_geometry(float width, float height, float depth,
                                    float floor_y, float tilt_x, float tilt_z,
                                    float wall_stiffness, float damping,
                                    uint32_t clip_enabled, float clip_margin,
                                    GpuContainerGeometry *out) {
  if (!out) return FLUID_ERR_INVALID_ARG;
  // rotation_matrices (src/state/simulation.rs:139-159): R = Rz * Rx
  const float sx = std::sin(tilt_x), cx = std::cos(tilt_x);
  const float sz = std::sin(tilt_z), cz = std::cos(tilt_z);
  GpuContainerGeometry g;
  std::memset(&g, 0, sizeof(g));
  const float fw[3][4] = {{cz, -sz * cx, sz * sx, 0.0f},
                          {sz, cz * cx, -cz * sx, 0.0f},
                          {0.0f, sx, cx, 0.0f}};
  const float iv[3][4] = {{cz, sz, 0.0f, 0.0f},
                          {-sz * cx, cz * cx, sx, 0.0f},
                          {sz * sx, -cz * sx, cx, 0.0f}};
  // to_gpu_geometry (:163-184)
  g.half_width = width / 2.0f;
  g.half_height = height / 2.0f;
  g.half_depth = depth / 2.0f;
  g.center_y = floor_y + height / 2.0f;
  std::memcpy(g.forward_row0, fw[0], 16);
  std::memcpy(g.forward_row1, fw[1], 16);
  std::memcpy(g.forward_row2, fw[2], 16);
  std::memcpy(g.inverse_row0, iv[0], 16);
  std::memcpy(g.inverse_row1, iv[1], 16);
  std::memcpy(g.inverse_row2, iv[2], 16);
  g.wall_stiffness = wall_stiffness;
  g.damping = damping;
  g.clip_enabled = clip_enabled ? 1u : 0u;
  g.clip_margin = clip_margin;
  *out = g;
  return FLUID_OK;
}

}  // extern "C"

namespace fluid {

// Per-launch constants of the MC passes, each in its WGSL expression order.
McLaunch make_mc_launch(const GpuMcGridParams &g, const GpuAnisoParams &a,
                        const GpuContainerGeometry &c, float sph_inv_cell) {
  McLaunch m;
  std::memset(&m, 0, sizeof(m));
  m.g = g;
  m.a = a;
  m.c = c;
  const float hm = a.h_mc;
  m.peak = 315.0f / (64.0f * kPI * hm * hm * hm);                            // mc_anisotropy.wgsl:154
  m.coeff = 315.0f / (64.0f * 3.14159265359f * std::pow(g.kernel_radius, 9.0f));  // mc_density.wgsl:81
  m.aniso_cells = (int)std::ceil(a.support_radius * sph_inv_cell);          // mc_anisotropy.wgsl:173
  float reach = g.kernel_radius;                                            // mc_density.wgsl:151-155
  if (a.enabled) reach = g.kernel_radius * a.max_stretch + a.max_shift;
  m.density_cells = (int)std::ceil(reach * sph_inv_cell);
  return m;
}

// update_aniso_params with the app's settings (marching_cubes.rs:26-34,
// 2133-2146; app.rs:1514-1519; mc_anisotropy on, strength 1, h_mc = h).
GpuAnisoParams default_aniso_params(float h) {
  GpuAnisoParams a;
  a.enabled = 1u;
  a.strength = 1.0f;
  a.support_radius = 2.0f * h;
  a.h_mc = h;
  a.kr = 4.0f;
  a.lambda = 0.9f;
  a.max_stretch = 1.6f;
  a.max_shift = 0.4f * h;
  return a;
}

// update_params after set_bounds as the app calls them (app.rs:1472-1513,
// marching_cubes.rs:2099-2131) for an untilted container: its box plus a
// margin of 1.6 h + 0.05, cubic cells of the largest extent / grid_size,
// iso value 0.95 of the poly6 peak (rendering.rs:313-319).
GpuMcGridParams default_mc_grid(const GpuContainerGeometry &c, float h, uint32_t grid_size) {
  GpuMcGridParams g;
  std::memset(&g, 0, sizeof(g));
  const float margin = h * 1.0f * 1.6f + 0.05f;
  const float lo[3] = {-c.half_width, c.center_y - c.half_height, -c.half_depth};
  const float hi[3] = {c.half_width, c.center_y + c.half_height, c.half_depth};
  for (int k = 0; k < 3; ++k) {
    g.grid_min[k] = lo[k] - margin;
    g.grid_max[k] = hi[k] + margin;
  }
  g.grid_size = grid_size;
  const float ex = g.grid_max[0] - g.grid_min[0], ey = g.grid_max[1] - g.grid_min[1];
  const float ez = g.grid_max[2] - g.grid_min[2];
  g.cell_size = std::max(ex, std::max(ey, ez)) / (float)grid_size;
  g.kernel_radius = h;
  const float pi = 3.14159274101257324f;  // std::f32::consts::PI
  g.iso_value = 0.95f * (315.0f / (64.0f * pi * h * h * h));
  g.max_vertices = 4000000u;
  return g;
}

}  // namespace fluid
