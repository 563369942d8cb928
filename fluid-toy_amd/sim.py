"""Python mirror of SphSimulation3DGrid (src/simulation/sph_3d_grid.rs:33-1570)
over the C ABI of libfluidhip.so.

Method names follow the reference struct; buffers that the reference exposes
as wgpu::Buffer getters (:1500-1526) are exposed as readbacks returning
numpy arrays.  Particle records are (N, 16) float32 arrays in the 64-byte
SphParticle3D layout (columns: abi.POS, abi.VEL, abi.FORCE, abi.DENSITY,
abi.NEAR, abi.NORMAL).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


def derive_sph_params(kernel_radius: float, stiffness: float, near_stiffness: float,
                      viscosity: float, mass: float, surface_tension: float,
                      xsph_epsilon: float, dt: float, num_particles: int,
                      rest_density: float | None = None) -> abi.GpuSphParams3D:
    """SphConfig::to_gpu_params_3d (src/state/simulation.rs:315-336); a given
    rest_density replaces compute_rest_density and feeds compute_pcisph_delta."""
    lib = abi.load_library()
    out = abi.GpuSphParams3D()
    abi.check(lib, lib.fluid_derive_sph_params(
        kernel_radius, stiffness, near_stiffness, viscosity, mass, surface_tension,
        xsph_epsilon, dt, num_particles, float(rest_density or 0.0), C.byref(out)))
    return out


def derive_container_geometry(width: float, height: float, depth: float, floor_y: float,
                              tilt_x: float = 0.0, tilt_z: float = 0.0,
                              wall_stiffness: float = 200.0, damping: float = 0.55,
                              clip_enabled: bool = False, clip_margin: float = 0.0
                              ) -> abi.GpuContainerGeometry:
    """ContainerConfig::to_gpu_geometry (src/state/simulation.rs:139-184)."""
    lib = abi.load_library()
    out = abi.GpuContainerGeometry()
    abi.check(lib, lib.fluid_derive_container_geometry(
        width, height, depth, floor_y, tilt_x, tilt_z, wall_stiffness, damping,
        1 if clip_enabled else 0, clip_margin, C.byref(out)))
    return out


def gravity(x: float, y: float, z: float) -> abi.GpuGravity:
    g = abi.GpuGravity()
    g.direction[0], g.direction[1], g.direction[2] = x, y, z
    return g


class SphSimulation3DGrid:
    """One simulation handle on one HIP device (SphSimulation3DGrid::new,
    sph_3d_grid.rs:104-1258)."""

    def __init__(self, particles: np.ndarray, sph_params: abi.GpuSphParams3D,
                 container_geom: abi.GpuContainerGeometry, max_particles: int = 0,
                 sdf: np.ndarray | None = None, bounds_extent: float = 0.0,
                 device: int = 0):
        self._lib = abi.load_library()
        parts = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, 16)
        self._keep = parts
        info = abi.FluidCreateInfo()
        info.particles = parts.ctypes.data if len(parts) else None
        info.num_particles = len(parts)
        info.max_particles = max(int(max_particles), len(parts))
        info.sph_params = sph_params
        info.container = container_geom
        info.bounds_extent = float(bounds_extent)
        info.device = int(device)
        if sdf is not None:
            vox = np.ascontiguousarray(sdf, dtype=np.float32)
            if vox.ndim != 3 or len(set(vox.shape)) != 1:
                raise ValueError("sdf must be a cubic (res, res, res) array indexed [z, y, x]")
            self._sdf = vox
            info.sdf_voxels = vox.ctypes.data
            info.sdf_resolution = vox.shape[0]
        handle = C.c_void_p()
        abi.check(self._lib, self._lib.fluid_create(C.byref(info), C.byref(handle)))
        self._h = handle

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        for sp in list(getattr(self, "_bound", ())):
            sp.close()
        if getattr(self, "_h", None):
            self._lib.fluid_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int) -> None:
        abi.check(self._lib, rc)

    # -- the reference's surface ------------------------------------------
    def step(self, n: int = 1) -> None:
        """step() (sph_3d_grid.rs:1261); n > 1 runs the substep loop."""
        self._chk(self._lib.fluid_step_n(self._h, n))

    def step_stages(self, first: int | str, last: int | str) -> None:
        f = abi.STAGE[first] if isinstance(first, str) else int(first)
        l = abi.STAGE[last] if isinstance(last, str) else int(last)
        self._chk(self._lib.fluid_step_stages(self._h, f, l))

    def update_sph_params(self, params: abi.GpuSphParams3D) -> None:
        self._chk(self._lib.fluid_update_sph_params(self._h, C.byref(params)))

    def update_container_geometry(self, geom: abi.GpuContainerGeometry) -> None:
        self._chk(self._lib.fluid_update_container_geometry(self._h, C.byref(geom)))

    def update_mouse_force(self, mouse: abi.GpuMouseForce) -> None:
        self._chk(self._lib.fluid_update_mouse_force(self._h, C.byref(mouse)))

    def update_gravity(self, g: abi.GpuGravity) -> None:
        self._chk(self._lib.fluid_update_gravity(self._h, C.byref(g)))

    def update_rigid_body(self, body: abi.GpuRigidBody) -> None:
        self._chk(self._lib.fluid_update_rigid_body(self._h, C.byref(body)))

    def set_pcisph_iterations(self, iterations: int) -> None:
        self._chk(self._lib.fluid_set_pcisph_iterations(self._h, iterations))

    @property
    def pcisph_iterations(self) -> int:
        return self._lib.fluid_pcisph_iterations(self._h)

    def clear_rigid_body_accum(self) -> None:
        self._chk(self._lib.fluid_clear_rigid_body_accum(self._h))

    def read_rigid_body_accum(self) -> abi.GpuRigidBodyAccum:
        out = abi.GpuRigidBodyAccum()
        self._chk(self._lib.fluid_read_rigid_body_accum(self._h, C.byref(out)))
        self._last_accum = out
        return out

    def rigid_body_accum(self) -> abi.GpuRigidBodyAccum:
        return getattr(self, "_last_accum", abi.GpuRigidBodyAccum())

    def spawn_particles(self, position, count: int, spread: float) -> int:
        pos = (C.c_float * 3)(*[float(v) for v in position])
        got = C.c_uint32(0)
        self._chk(self._lib.fluid_spawn_particles(self._h, pos, count, spread, C.byref(got)))
        return got.value

    def num_particles(self) -> int:
        return self._lib.fluid_num_particles(self._h)

    def max_particles(self) -> int:
        return self._lib.fluid_max_particles(self._h)

    # -- buffer access -----------------------------------------------------
    def _read_records(self, fn) -> np.ndarray:
        n = self.num_particles()
        out = np.zeros((n, 16), dtype=np.float32)
        self._chk(fn(self._h, out.ctypes.data if n else None, n))
        return out

    def read_particles(self) -> np.ndarray:
        return self._read_records(self._lib.fluid_read_particles)

    def write_particles(self, particles: np.ndarray) -> None:
        parts = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, 16)
        self._chk(self._lib.fluid_write_particles(
            self._h, parts.ctypes.data if len(parts) else None, len(parts)))

    def read_sorted_particles(self) -> np.ndarray:
        return self._read_records(self._lib.fluid_read_sorted_particles)

    def grid_params(self) -> abi.GpuGridParams:
        g = abi.GpuGridParams()
        self._chk(self._lib.fluid_read_grid_params(self._h, C.byref(g)))
        return g

    def sph_params(self) -> abi.GpuSphParams3D:
        p = abi.GpuSphParams3D()
        self._chk(self._lib.fluid_read_sph_params(self._h, C.byref(p)))
        return p

    def container_geometry(self) -> abi.GpuContainerGeometry:
        g = abi.GpuContainerGeometry()
        self._chk(self._lib.fluid_read_container_geometry(self._h, C.byref(g)))
        return g

    def _read_u32(self, fn, count: int) -> np.ndarray:
        out = np.zeros(count, dtype=np.uint32)
        self._chk(fn(self._h, out.ctypes.data if count else None, count))
        return out

    def read_cell_counts(self) -> np.ndarray:
        return self._read_u32(self._lib.fluid_read_cell_counts, self.grid_params().total_cells)

    def read_cell_starts(self) -> np.ndarray:
        return self._read_u32(self._lib.fluid_read_cell_starts, self.grid_params().total_cells)

    def read_sorted_to_orig(self) -> np.ndarray:
        return self._read_u32(self._lib.fluid_read_sorted_to_orig, self.num_particles())

    def read_orig_to_sorted(self) -> np.ndarray:
        return self._read_u32(self._lib.fluid_read_orig_to_sorted, self.num_particles())

    def read_predicted(self, which: int) -> np.ndarray:
        """sorted_predicted_a (0) / _b (1) as (N, 8) float32:
        pred_pos xyz, pressure, pred_vel xyz, pred_density."""
        n = self.num_particles()
        out = np.zeros((n, 8), dtype=np.float32)
        self._chk(self._lib.fluid_read_predicted(self._h, which, out.ctypes.data if n else None, n))
        return out

    def sync(self) -> None:
        self._chk(self._lib.fluid_sync(self._h))

    # -- device-side access (the buffer getters, sph_3d_grid.rs:1500-1526) --
    def export_particles_device(self, dst_ptr: int, cap: int) -> None:
        """particle_buffer's records written to device memory at dst_ptr
        (asynchronous on the handle's stream)."""
        self._chk(self._lib.fluid_export_particles_device(self._h, C.c_void_p(dst_ptr), cap))

    def export_sorted_particles_device(self, dst_ptr: int, cap: int) -> None:
        """sorted_particle_buffer's records written to device memory."""
        self._chk(self._lib.fluid_export_sorted_particles_device(self._h, C.c_void_p(dst_ptr),
                                                                 cap))

    def device_buffers(self, scene: int = 0) -> abi.FluidDeviceView:
        """Device pointers of the cell counts / starts, the slot map and the
        uniform blocks (reference layouts) of one scene."""
        v = abi.FluidDeviceView()
        self._chk(self._lib.fluid_device_buffers(self._h, scene, C.byref(v)))
        return v

    # -- measurement -------------------------------------------------------
    def time_steps(self, warmup: int, steps: int) -> float:
        ms = C.c_float(0.0)
        self._chk(self._lib.fluid_time_steps(self._h, warmup, steps, C.byref(ms)))
        return ms.value

    def time_range(self, first: str, last: str, reps: int) -> float:
        """Mean ms of one re-run of stages first..last (names of abi.STAGES,
        within density..predict) of the last step (fluid_time_range)."""
        ms = C.c_float()
        self._chk(self._lib.fluid_time_range(self._h, abi.STAGE[first], abi.STAGE[last], reps,
                                             C.byref(ms)))
        return ms.value

    def time_stages(self, steps: int) -> dict:
        arr = (C.c_float * len(abi.STAGES))()
        self._chk(self._lib.fluid_time_stages(self._h, steps, arr))
        return {name: arr[i] for i, name in enumerate(abi.STAGES)}

    def set_sweep_config(self, list_lanes: int = 0, solve_lanes: int = 0,
                         xcd_tiles: int = 0) -> None:
        arr = (C.c_uint32 * 3)(list_lanes, solve_lanes, xcd_tiles)
        self._chk(self._lib.fluid_set_sweep_config(self._h, arr))

    def sweep_config(self) -> tuple[int, int, int]:
        arr = (C.c_uint32 * 3)()
        self._chk(self._lib.fluid_get_sweep_config(self._h, arr))
        return tuple(arr)

    def set_tile_config(self, tiled: int, width: int = 0, workgroups: int = 0) -> None:
        """LDS-staged tiled sweeps on/off, tile width (cells), persistent
        workgroups (fluid_set_tile_config; 0 keeps a value)."""
        arr = (C.c_uint32 * 3)(tiled, width, workgroups)
        self._chk(self._lib.fluid_set_tile_config(self._h, arr))

    def set_fusion(self, flags: int) -> None:
        """Work-sharing flags (fluid_set_fusion): abi.FLUID_FUSE_XSPH_FORCE,
        abi.FLUID_FUSE_SOLVE_LIST, abi.FLUID_WALK_16 or abi.FLUID_WALK_32,
        abi.FLUID_SOLVE_WALK_X2 or abi.FLUID_SOLVE_WALK_X4,
        abi.FLUID_SOLVE_FIRST_X2 or abi.FLUID_SOLVE_FIRST_X4, abi.FLUID_LIST_SKIP."""
        self._chk(self._lib.fluid_set_fusion(self._h, int(flags)))

    def fusion(self) -> int:
        return int(self._lib.fluid_fusion(self._h))

    def tile_config(self) -> tuple[int, int, int]:
        arr = (C.c_uint32 * 3)()
        self._chk(self._lib.fluid_get_tile_config(self._h, arr))
        return tuple(arr)

    def autotune(self, trials: int = 2) -> tuple[int, int, int]:
        arr = (C.c_uint32 * 3)()
        self._chk(self._lib.fluid_autotune(self._h, trials, arr))
        return tuple(arr)

    def pair_stats(self) -> dict:
        arr = (C.c_uint64 * 4)()
        self._chk(self._lib.fluid_pair_stats(self._h, arr))
        n = max(self.num_particles(), 1)
        return {"candidates": arr[0] / n, "in_range": arr[1] / n,
                "max_candidates": arr[2], "max_in_range": arr[3]}

    def count_pairs(self) -> tuple[int, int]:
        a, b = C.c_uint64(0), C.c_uint64(0)
        self._chk(self._lib.fluid_count_pairs(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # -- differentiable rollout ------------------------------------------
    def forward_rollout(self, steps: int, checkpoint_every: int = 0) -> None:
        """Advance `steps` steps (as step(steps)) keeping checkpoints for
        backward_rollout (fluid_forward_rollout)."""
        self._chk(self._lib.fluid_forward_rollout(self._h, int(steps), int(checkpoint_every)))

    def backward_rollout(self, dl_dpos=None, dl_dvel=None, want_state: bool = True):
        """Reverse the last forward_rollout (fluid_backward_rollout).

        dl_dpos / dl_dvel: (n, 3) float32 arrays in original index order, or
        None for zero.  Returns (dL/dpos0 (n,3), dL/dvel0 (n,3), dL/dparams
        (num_scenes, 5): viscosity, surface tension, gravity x, y, z)."""
        n = self.num_particles()
        keep = []

        def src(a):
            if a is None:
                return None
            arr = np.ascontiguousarray(a, dtype=np.float32).reshape(n, 3)
            keep.append(arr)
            return arr.ctypes.data
        gx = np.zeros((n, 3), np.float32) if want_state else None
        gv = np.zeros((n, 3), np.float32) if want_state else None
        scenes = self._lib.fluid_batch_num_scenes(self._h)
        gp = np.zeros((max(scenes, 1), abi.FLUID_GRAD_PARAMS), np.float64)
        self._chk(self._lib.fluid_backward_rollout(
            self._h, src(dl_dpos), src(dl_dvel),
            gx.ctypes.data if gx is not None else None,
            gv.ctypes.data if gv is not None else None, gp.ctypes.data))
        return gx, gv, gp

    def backward_rollout_scaled(self, dl_dpos=None, dl_dvel=None):
        """Reverse the last forward_rollout keeping the state gradients as
        mantissa and exponent (fluid_backward_rollout_scaled).  Returns
        (mantissa of dL/dpos0 (n,3) float32, mantissa of dL/dvel0,
        dL/dparams (num_scenes, 6): viscosity, surface tension, gravity x,
        y, z, stiffness, and log2_scale (num_scenes,) int32): dL/dpos0 =
        mantissa * 2**log2_scale[scene]."""
        n = self.num_particles()
        keep = []

        def src(a):
            if a is None:
                return None
            arr = np.ascontiguousarray(a, dtype=np.float32).reshape(n, 3)
            keep.append(arr)
            return arr.ctypes.data
        gx = np.zeros((n, 3), np.float32)
        gv = np.zeros((n, 3), np.float32)
        scenes = max(self._lib.fluid_batch_num_scenes(self._h), 1)
        gp = np.zeros((scenes, abi.FLUID_GRAD_PARAMS_ALL), np.float64)
        e = np.zeros(scenes, np.int32)
        self._chk(self._lib.fluid_backward_rollout_scaled(
            self._h, src(dl_dpos), src(dl_dvel), gx.ctypes.data, gv.ctypes.data,
            gp.ctypes.data, e.ctypes.data))
        return gx, gv, gp, e

    def backward_rollout_device(self, dl_dpos_ptr: int, dl_dvel_ptr: int,
                                out_pos_ptr: int, out_vel_ptr: int) -> np.ndarray:
        """backward_rollout on device buffers (addresses of (n, 3) float32
        arrays on this handle's device; 0 = null).  Returns dL/dparams."""
        scenes = self._lib.fluid_batch_num_scenes(self._h)
        gp = np.zeros((max(scenes, 1), abi.FLUID_GRAD_PARAMS), np.float64)
        self._chk(self._lib.fluid_backward_rollout(
            self._h, dl_dpos_ptr or None, dl_dvel_ptr or None, out_pos_ptr or None,
            out_vel_ptr or None, gp.ctypes.data))
        return gp

    def rollout_timing(self) -> dict:
        """HIP-event timing of the last backward_rollout (ms)."""
        arr = (C.c_float * 3)()
        self._chk(self._lib.fluid_rollout_timing(self._h, arr))
        return {"replay": arr[0], "reverse": arr[1], "total": arr[2]}


class SphBatch(SphSimulation3DGrid):
    """Many independent scenes advanced together (fluid_batch_create): one
    launch per stage over all scenes.  Readbacks return all scenes
    concatenated; `scene_slice(k)` selects scene k's rows."""

    def __init__(self, scenes, pcisph_iterations: int = 4, device: int = 0):
        """scenes: iterable of (particles (n,16) f32, GpuSphParams3D,
        GpuContainerGeometry, GpuGravity, bounds_extent)."""
        self._lib = abi.load_library()
        scenes = list(scenes)
        descs = (abi.FluidSceneDesc * len(scenes))()
        self._keep = []
        for d, (parts, params, cont, grav, extent) in zip(descs, scenes):
            arr = np.ascontiguousarray(parts, dtype=np.float32).reshape(-1, 16)
            self._keep.append(arr)
            d.particles = arr.ctypes.data if len(arr) else None
            d.num_particles = len(arr)
            d.sph_params = params
            d.container = cont
            d.gravity = grav
            d.bounds_extent = float(extent)
        handle = C.c_void_p()
        abi.check(self._lib, self._lib.fluid_batch_create(descs, len(scenes), pcisph_iterations,
                                                          device, C.byref(handle)))
        self._h = handle

    def num_scenes(self) -> int:
        return self._lib.fluid_batch_num_scenes(self._h)

    def scene_info(self, k: int) -> dict:
        first, count, cell0 = C.c_uint32(), C.c_uint32(), C.c_uint32()
        grid = abi.GpuGridParams()
        self._chk(self._lib.fluid_batch_scene_info(self._h, k, C.byref(first), C.byref(count),
                                                   C.byref(cell0), C.byref(grid)))
        return {"first": first.value, "count": count.value, "first_cell": cell0.value,
                "cells": grid.total_cells, "grid": grid}

    def scene_slice(self, k: int) -> slice:
        i = self.scene_info(k)
        return slice(i["first"], i["first"] + i["count"])

    def update_scene(self, k: int, params=None, container=None, gravity=None) -> None:
        self._chk(self._lib.fluid_batch_update_scene(
            self._h, k, C.byref(params) if params is not None else None,
            C.byref(container) if container is not None else None,
            C.byref(gravity) if gravity is not None else None))

    def read_cell_counts(self) -> np.ndarray:
        total = sum(self.scene_info(k)["cells"] for k in range(self.num_scenes()))
        return self._read_u32(self._lib.fluid_read_cell_counts, total)

    def read_cell_starts(self) -> np.ndarray:
        total = sum(self.scene_info(k)["cells"] for k in range(self.num_scenes()))
        return self._read_u32(self._lib.fluid_read_cell_starts, total)


def _bind(sim, obj) -> None:
    if not hasattr(sim, "_bound"):
        sim._bound = []
    sim._bound.append(obj)


def _unbind(sim, obj) -> None:
    if obj in getattr(sim, "_bound", []):
        sim._bound.remove(obj)


class SpraySystem:
    """Mirror of the reference's SpraySystem (src/simulation/spray.rs:82-491)
    over fluid_spray_*: whitewater particles emitted from and advected by the
    simulation's sorted particles of its last step.  Ring readbacks are
    (max, 12) uint32 arrays of the GpuSprayParticle words (abi.SPRAY_*); use
    `.view(np.float32)` for the float fields."""

    def __init__(self, sim: SphSimulation3DGrid, max_spray: int,
                 params: abi.GpuSprayParams | None = None):
        self._lib = sim._lib
        self._sim = sim
        self.max_particles = int(max_spray)
        if params is None:
            params = abi.default_spray_params(dt=1.0 / 60.0, max_particles=max_spray,
                                              num_sph_particles=sim.num_particles())
        self.params = params
        handle = C.c_void_p()
        abi.check(self._lib, self._lib.fluid_spray_create(sim._h, max_spray, C.byref(params),
                                                          C.byref(handle)))
        self._h = handle
        _bind(sim, self)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.fluid_spray_destroy(self._h)
            self._h = None
            _unbind(self._sim, self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int) -> None:
        abi.check(self._lib, rc)

    def update_params(self, params: abi.GpuSprayParams) -> None:
        """update_params (spray.rs:345-347)."""
        self._chk(self._lib.fluid_spray_update_params(self._h, C.byref(params)))
        self.params = params

    def step(self, num_sph_particles: int | None = None) -> None:
        """step (spray.rs:349-428): emit, then simulate."""
        n = self._sim.num_particles() if num_sph_particles is None else num_sph_particles
        self._chk(self._lib.fluid_spray_step(self._h, n))

    def emit(self, num_sph_particles: int | None = None) -> None:
        n = self._sim.num_particles() if num_sph_particles is None else num_sph_particles
        self._chk(self._lib.fluid_spray_emit(self._h, n))

    def simulate(self) -> None:
        self._chk(self._lib.fluid_spray_simulate(self._h))

    def auto_limits(self) -> tuple[float, float]:
        """auto_limits (spray.rs:458-460): (trapped air, wave crest)."""
        out = (C.c_float * 2)()
        self._chk(self._lib.fluid_spray_auto_limits(self._h, out))
        return float(out[0]), float(out[1])

    def read_stats(self) -> np.ndarray:
        """read_stats (spray.rs:464-475): total, spray, foam, bubble."""
        out = (C.c_uint32 * 4)()
        self._chk(self._lib.fluid_spray_read_stats(self._h, out))
        return np.array(out, dtype=np.uint32)

    def state(self) -> dict:
        st = abi.FluidSprayState()
        self._chk(self._lib.fluid_spray_read_state(self._h, C.byref(st)))
        return {"write_head": st.write_head, "ta_limit": st.ta_limit, "wc_limit": st.wc_limit,
                "live": list(st.live), "emitted": st.emitted}

    def read_particles(self) -> np.ndarray:
        """spray_buffer (spray.rs:477-479) as a (max, 12) uint32 array."""
        out = np.zeros((self.max_particles, 12), np.uint32)
        self._chk(self._lib.fluid_spray_read_particles(self._h, out.ctypes.data,
                                                       self.max_particles))
        return out

    def write_particles(self, ring: np.ndarray) -> None:
        ring = np.ascontiguousarray(ring, dtype=np.uint32).reshape(self.max_particles, 12)
        self._chk(self._lib.fluid_spray_write_particles(self._h, ring.ctypes.data,
                                                        self.max_particles))

    def read_emit(self, n: int) -> tuple[np.ndarray, np.ndarray]:
        """The last emit pass: potentials (n, 4) float32 (trapped air, crest,
        divergence, neighbours) and spawns (n,) uint32."""
        pot = np.zeros((n, 4), np.float32)
        ne = np.zeros(n, np.uint32)
        self._chk(self._lib.fluid_spray_read_emit(self._h, pot.ctypes.data, ne.ctypes.data, n))
        return pot, ne

    def reset(self) -> None:
        """reset (spray.rs:482-490)."""
        self._chk(self._lib.fluid_spray_reset(self._h))

    def time_steps(self, warmup: int, frames: int, num_sph_particles: int | None = None) -> float:
        n = self._sim.num_particles() if num_sph_particles is None else num_sph_particles
        ms = C.c_float(0.0)
        self._chk(self._lib.fluid_spray_time_steps(self._h, n, warmup, frames, C.byref(ms)))
        return ms.value


MC_NOALLOC_MAX_PARTICLES = 65536


def mc_generate_noalloc(sim: SphSimulation3DGrid, grid: abi.GpuMcGridParams,
                        aniso: abi.GpuAnisoParams, container: abi.GpuContainerGeometry,
                        field_ptr: int) -> None:
    """The marching-cubes field of all of `sim`'s particles written to the
    caller's device buffer at `field_ptr` (grid.grid_size^3 float32, x
    fastest) without a field handle and without allocating device memory
    (fluid_mc_generate_noalloc).  Asynchronous on the simulation's stream."""
    sim._chk(sim._lib.fluid_mc_generate_noalloc(sim._h, C.byref(grid), C.byref(aniso),
                                                C.byref(container), C.c_void_p(field_ptr)))


class MarchingCubesField:
    """The compute passes 0 and 1 of the reference's
    MarchingCubesRenderer::generate (src/render/marching_cubes.rs:2349-2470)
    over fluid_mc_*: the anisotropic kernel fit of the sorted particles and
    the density field on a grid_size^3 voxel grid (x fastest, as the
    reference's 3D texture)."""

    def __init__(self, sim: SphSimulation3DGrid, grid_size: int = 200):
        self._lib = sim._lib
        self._sim = sim
        self.grid_size = int(grid_size)
        handle = C.c_void_p()
        abi.check(self._lib, self._lib.fluid_mc_create(sim._h, grid_size, C.byref(handle)))
        self._h = handle
        _bind(sim, self)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.fluid_mc_destroy(self._h)
            self._h = None
            _unbind(self._sim, self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int) -> None:
        abi.check(self._lib, rc)

    def update_params(self, params: abi.GpuMcGridParams) -> None:
        """update_params + set_bounds (marching_cubes.rs:2099-2131, 2170-2174)."""
        self._chk(self._lib.fluid_mc_update_params(self._h, C.byref(params)))

    def update_aniso_params(self, params: abi.GpuAnisoParams) -> None:
        """update_aniso_params (marching_cubes.rs:2133-2146)."""
        self._chk(self._lib.fluid_mc_update_aniso_params(self._h, C.byref(params)))

    def update_container_geometry(self, geom: abi.GpuContainerGeometry) -> None:
        self._chk(self._lib.fluid_mc_update_container_geometry(self._h, C.byref(geom)))

    def generate(self, num_particles: int | None = None) -> None:
        """generate, passes 0 (anisotropy, when enabled) and 1 (density)."""
        n = self._sim.num_particles() if num_particles is None else num_particles
        self._chk(self._lib.fluid_mc_generate(self._h, n))

    def read_density(self) -> np.ndarray:
        """The density field as a (G, G, G) float32 array indexed [z, y, x]."""
        G = self.grid_size
        out = np.zeros((G, G, G), np.float32)
        self._chk(self._lib.fluid_mc_read_density(self._h, out.ctypes.data, out.size))
        return out

    def read_aniso(self) -> np.ndarray:
        """Ellipsoid records of the sorted particles, (N, 12) float32."""
        n = self._sim.num_particles()
        out = np.zeros((n, 12), np.float32)
        self._chk(self._lib.fluid_mc_read_aniso(self._h, out.ctypes.data if n else None, n))
        return out

    def density_from_aniso(self, aniso: np.ndarray) -> None:
        a = np.ascontiguousarray(aniso, dtype=np.float32).reshape(-1, 12)
        self._chk(self._lib.fluid_mc_density_from_aniso(self._h, a.ctypes.data, len(a)))

    def time(self, warmup: int, iters: int, num_particles: int | None = None) -> float:
        n = self._sim.num_particles() if num_particles is None else num_particles
        ms = C.c_float(0.0)
        self._chk(self._lib.fluid_mc_time(self._h, n, warmup, iters, C.byref(ms)))
        return ms.value

