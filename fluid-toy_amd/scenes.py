"""Seeded scene generators for the benchmark and parity workloads.

The reference seeds its initial block from the wall clock
(src/simulation/particle.rs:41-76); these scenes replace that generator with
a counter-based splitmix64 stream so every run (GPU, oracle, CPU baseline)
starts from the same bits.  Lattices are emitted y-major, then z, then x,
the loop order of create_particle_block (particle.rs:48-58).

Workloads (lattice spacing 0.025, h = 0.08, jitter U(-0.002, 0.002)/axis):
  dam_break_8k    16x32x16 block in a 1.6x1.2x0.8 box (seed 0)
  dam_break_50k   50,000 particles filling the left 30% of a 2.0x1.2x1.0 box
                  layer by layer (seed 1)
  block_1m        128x64x128 block, 40% of a 3.2x4.0x3.2 box (seed 2)
  block_4m        256x64x256 block, 40% of a 6.4x4.0x6.4 box (seed 3)
  dam_break_32k   32^3 block at the left of the 2.0x1.2x1.0 box, initial
                  velocity N(0, 0.1^2) per axis (seeds 4/5)
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """count successive outputs of splitmix64 started at `seed`."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int) -> np.ndarray:
    """float64 in [0, 1) with 53 random bits."""
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, count: int) -> np.ndarray:
    """Box-Muller normals from the same stream (2 uniforms per pair)."""
    m = (count + 1) // 2
    u = uniform01(seed, 2 * m)
    u1 = np.maximum(u[0::2], 1e-300)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    out = np.empty(2 * m)
    out[0::2] = z[:m]
    out[1::2] = z[m:]
    return out[:count]


def lattice(nx: int, ny: int, nz: int, spacing: float, count: int | None = None) -> np.ndarray:
    """Cell-centred lattice points ((i+1/2) s), y-major, then z, then x."""
    y, z, x = np.meshgrid(np.arange(ny), np.arange(nz), np.arange(nx), indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.float64)
    pts = (pts + 0.5) * spacing
    return pts if count is None else pts[:count]


def particles_from_positions(pos: np.ndarray, vel: np.ndarray | None = None) -> np.ndarray:
    """(N, 16) float32 SphParticle3D records (SphParticle3D::new + velocity)."""
    n = len(pos)
    out = np.zeros((n, 16), dtype=np.float32)
    out[:, 0:3] = pos.astype(np.float32)
    if vel is not None:
        out[:, 4:7] = vel.astype(np.float32)
    return out


@dataclass
class Scene:
    """Everything needed to build a simulation handle for one workload."""
    name: str
    particles: np.ndarray            # (N, 16) float32
    box: tuple                       # width, height, depth
    floor_y: float
    h: float = 0.08
    rest_density: float = 8000.0
    stiffness: float = 35.0
    near_stiffness: float = 0.45
    viscosity: float = 15.0
    surface_tension: float = 6.5
    wall_stiffness: float = 250.0
    damping: float = 0.55
    dt: float = 0.008
    mass: float = 1.0
    xsph_epsilon: float = 0.305
    gravity: tuple = (0.0, -9.81, 0.0)
    tilt: tuple = (0.0, 0.0)         # container tilt about x, then z (radians)
    pcisph_iterations: int = 4
    bounds_extent: float = 2.0
    steps: int = 200
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return len(self.particles)


def _jittered_block(pts: np.ndarray, seed: int, shift) -> np.ndarray:
    jit = uniform01(seed, pts.size).reshape(pts.shape) * 0.004 - 0.002
    return pts + jit + np.asarray(shift, dtype=np.float64)


def dam_break_8k() -> Scene:
    pts = lattice(16, 32, 16, 0.025)
    pos = _jittered_block(pts, 0, (-0.8, 0.0, -0.4))
    return Scene("dam_break_8k", particles_from_positions(pos), (1.6, 1.2, 0.8), 0.0,
                 steps=200)


def dam_break_50k() -> Scene:
    pts = lattice(24, 53, 40, 0.025, count=50_000)
    pos = _jittered_block(pts, 1, (-1.0, 0.0, -0.5))
    return Scene("dam_break_50k", particles_from_positions(pos), (2.0, 1.2, 1.0), 0.0,
                 steps=1000)


def block(n_side: int, seed: int) -> Scene:
    """Uniform block of n_side x 64 x n_side at spacing 0.025 resting on the
    floor of a box whose volume it fills to 40%."""
    w = n_side * 0.025
    pts = lattice(n_side, 64, n_side, 0.025)
    pos = _jittered_block(pts, seed, (-w / 2, -2.0, -w / 2))
    extent = max(2.0, w / 2 + 0.1)
    return Scene(f"block_{n_side * 64 * n_side}", particles_from_positions(pos),
                 (w, 4.0, w), -2.0, bounds_extent=extent, steps=100)


def block_1m() -> Scene:
    return block(128, 2)


def block_4m() -> Scene:
    return block(256, 3)


def dam_break_block(n_side: int, seed: int) -> Scene:
    """n_side^3 block in the corner of the 2.0x1.2x1.0 box (left wall and
    floor inside its boundary layer), initial velocity N(0, 0.1^2) per axis."""
    pts = lattice(n_side, n_side, n_side, 0.025)
    pos = _jittered_block(pts, seed, (-1.0, 0.0, -0.5))
    vel = normal(seed + 1000, pts.size).reshape(pts.shape) * 0.1
    return Scene(f"dam_break_block{n_side}", particles_from_positions(pos, vel),
                 (2.0, 1.2, 1.0), 0.0, steps=64)


def relaxed_block(n_side: int, seed: int) -> Scene:
    """n_side^3 block at the reference's spawn spacing 0.6 h (app.rs:179),
    where the density is near rest, in the corner of the 2.0x1.2x1.0 box;
    initial velocity N(0, 0.1^2) per axis and rest density from the
    reference's lattice rule (compute_rest_density, 8678.83 at h = 0.08)."""
    s = 0.6 * 0.08
    pts = lattice(n_side, n_side, n_side, s)
    pos = _jittered_block(pts, seed, (-1.0 + 0.25 * s, 0.25 * s, -0.5 + 0.25 * s))
    vel = normal(seed + 1000, pts.size).reshape(pts.shape) * 0.1
    return Scene(f"relaxed_block{n_side}", particles_from_positions(pos, vel),
                 (2.0, 1.2, 1.0), 0.0, rest_density=0.0, steps=64)


def dam_break_32k(seed: int = 4) -> Scene:
    sc = dam_break_block(32, seed)
    sc.name = "dam_break_32k"
    return sc


def tilted(scene: Scene, tilt_x: float, tilt_z: float) -> Scene:
    """The scene in its container tilted by tilt_x about x, then tilt_z about
    z (ContainerConfig tilt, src/state/simulation.rs:139-159): particle
    positions and velocities are carried into the tilted frame about the
    container centre, so the block keeps its place against the local walls."""
    cx, sx, cz, sz = np.cos(tilt_x), np.sin(tilt_x), np.cos(tilt_z), np.sin(tilt_z)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    r = rz @ rx
    centre = np.array([0.0, scene.floor_y + scene.box[1] / 2, 0.0])
    parts = scene.particles.copy()
    pos = parts[:, 0:3].astype(np.float64)
    parts[:, 0:3] = ((pos - centre) @ r.T + centre).astype(np.float32)
    parts[:, 4:7] = (parts[:, 4:7].astype(np.float64) @ r.T).astype(np.float32)
    out = Scene(f"{scene.name}_tilted", parts, scene.box, scene.floor_y,
                **{k: getattr(scene, k) for k in (
                    "h", "rest_density", "stiffness", "near_stiffness", "viscosity",
                    "surface_tension", "wall_stiffness", "damping", "dt", "mass",
                    "xsph_epsilon", "gravity", "pcisph_iterations", "bounds_extent", "steps")})
    out.tilt = (tilt_x, tilt_z)
    return out


def small_block(nx: int, ny: int, nz: int, seed: int = 7, box=(1.6, 1.2, 0.8)) -> Scene:
    """A small dam-break block for fast parity cases."""
    pts = lattice(nx, ny, nz, 0.025)
    pos = _jittered_block(pts, seed, (-box[0] / 2, 0.0, -box[2] / 2))
    return Scene(f"block_{nx}x{ny}x{nz}", particles_from_positions(pos), box, 0.0)


def batch_scenes(count: int = 256, seed: int = 6, n_min: int = 2048,
                 n_max: int = 65536) -> list:
    """Independent dam-break scenes of uneven size: particle counts
    log-uniform in [n_min, n_max], box edges uniform in [0.8, 2.0] per axis,
    and h, rest density, near stiffness, viscosity, surface tension, wall
    stiffness and XSPH epsilon each uniform within +-30% of the defaults.
    The block fills the box floor layer by layer at spacing 0.3125 h (the
    default 0.025 / 0.08 ratio), closer where the box would be more than 60%
    full.  The box is centred on the origin so the scene's cubic grid extent
    is its largest half edge plus one cell."""
    u = uniform01(seed, count * 12).reshape(count, 12)
    out = []
    for k in range(count):
        r = u[k]
        n = int(round(n_min * (n_max / n_min) ** r[0]))
        box = tuple(float(0.8 + 1.2 * r[i]) for i in (1, 2, 3))
        f = [0.7 + 0.6 * r[i] for i in range(4, 11)]
        h = 0.08 * f[0]
        s = min(0.3125 * h, (0.6 * box[0] * box[1] * box[2] / n) ** (1.0 / 3.0))
        nx, nz = max(1, int(box[0] / s) - 1), max(1, int(box[2] / s) - 1)
        ny = -(-n // (nx * nz))
        pts = lattice(nx, ny, nz, s, count=n)
        pos = pts + (uniform01(seed * 1000 + k, pts.size).reshape(pts.shape) * 2 - 1) * 0.08 * s
        pos = pos + np.array([-box[0] / 2, -box[1] / 2, -box[2] / 2])
        sc = Scene(f"batch{k}", particles_from_positions(pos), box, -box[1] / 2, h=h,
                   rest_density=8000.0 * f[1], near_stiffness=0.45 * f[2],
                   viscosity=15.0 * f[3], surface_tension=6.5 * f[4],
                   wall_stiffness=250.0 * f[5], xsph_epsilon=0.305 * f[6],
                   bounds_extent=max(box) / 2 + h, steps=200)
        out.append(sc)
    return out


SCENES = {
    "dam_break_8k": dam_break_8k,
    "dam_break_50k": dam_break_50k,
    "block_1m": block_1m,
    "block_4m": block_4m,
    "dam_break_32k": dam_break_32k,
}


def make_params(scene: Scene, abi_mod, derive_sph_params, derive_container_geometry):
    """GpuSphParams3D, GpuContainerGeometry and GpuGravity of a scene."""
    params = derive_sph_params(scene.h, scene.stiffness, scene.near_stiffness,
                               scene.viscosity, scene.mass, scene.surface_tension,
                               scene.xsph_epsilon, scene.dt, scene.n,
                               rest_density=scene.rest_density)
    w, hgt, d = scene.box
    cont = derive_container_geometry(w, hgt, d, scene.floor_y, scene.tilt[0], scene.tilt[1],
                                     scene.wall_stiffness, scene.damping)
    g = abi_mod.GpuGravity()
    g.direction[0], g.direction[1], g.direction[2] = scene.gravity
    return params, cont, g
