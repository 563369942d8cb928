"""ctypes mirror of include/fluid_abi.h and the loader of libfluidhip.so.

The structs are byte-identical to the reference's Rust/WGSL uniforms
(src/simulation/particle.rs:6-17, src/state/simulation.rs:342-394,
src/simulation/sph_3d_grid.rs:16-30, src/state/interaction.rs:102-111,
src/state/rigid_body.rs:212-256, src/shaders/pcisph_predict.wgsl:33-42).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libfluidhip.so"

FLUID_OK = 0
FLUID_ERR_INVALID_ARG = -1
FLUID_ERR_CAPACITY = -2
FLUID_ERR_HIP = -3
FLUID_ERR_NO_DEVICE = -4
FLUID_ERR_STATE = -5

STAGES = ("clear", "build", "scan", "reorder", "density", "xsph", "force",
          "predict", "solve", "finalize", "integrate")
STAGE = {name: i for i, name in enumerate(STAGES)}

f3 = C.c_float * 3
f4 = C.c_float * 4


class SphParticle3D(C.Structure):
    _fields_ = [("position", f3), ("_pad0", C.c_float), ("velocity", f3),
                ("_pad1", C.c_float), ("force", f3), ("density", C.c_float),
                ("near_density", C.c_float), ("normal", f3)]


class GpuSphParams3D(C.Structure):
    _fields_ = [("kernel_radius", C.c_float), ("kernel_radius_sq", C.c_float),
                ("kernel_radius_pow5", C.c_float), ("kernel_radius_pow6", C.c_float),
                ("kernel_radius_pow9", C.c_float), ("mass", C.c_float),
                ("rest_density", C.c_float), ("stiffness", C.c_float),
                ("near_stiffness", C.c_float), ("viscosity", C.c_float),
                ("dt", C.c_float), ("num_particles", C.c_uint32),
                ("surface_tension", C.c_float), ("pcisph_delta", C.c_float),
                ("xsph_epsilon", C.c_float), ("_padding_st", C.c_float)]


class GpuGridParams(C.Structure):
    _fields_ = [("grid_size_x", C.c_uint32), ("grid_size_y", C.c_uint32),
                ("grid_size_z", C.c_uint32), ("total_cells", C.c_uint32),
                ("cell_size", C.c_float), ("inv_cell_size", C.c_float),
                ("grid_origin_x", C.c_float), ("grid_origin_y", C.c_float),
                ("grid_origin_z", C.c_float), ("num_particles", C.c_uint32),
                ("_padding", C.c_uint32 * 2)]


class GpuContainerGeometry(C.Structure):
    _fields_ = [("half_width", C.c_float), ("half_height", C.c_float),
                ("half_depth", C.c_float), ("center_y", C.c_float),
                ("forward_row0", f4), ("forward_row1", f4), ("forward_row2", f4),
                ("inverse_row0", f4), ("inverse_row1", f4), ("inverse_row2", f4),
                ("wall_stiffness", C.c_float), ("damping", C.c_float),
                ("clip_enabled", C.c_uint32), ("clip_margin", C.c_float)]


class GpuGravity(C.Structure):
    _fields_ = [("direction", f3), ("_padding", C.c_float)]


class GpuMouseForce(C.Structure):
    _fields_ = [("position", f3), ("radius", C.c_float), ("strength", C.c_float),
                ("is_active", C.c_uint32), ("mode", C.c_uint32), ("_pad", C.c_float),
                ("direction", f3), ("_pad2", C.c_float)]


class GpuRigidBody(C.Structure):
    _fields_ = [("position", f3), ("half_extent", C.c_float), ("velocity", f3),
                ("is_active", C.c_uint32), ("stiffness", C.c_float), ("shape", C.c_uint32),
                ("_pad1", C.c_float), ("_pad2", C.c_float), ("rot_row0", f4),
                ("rot_row1", f4), ("rot_row2", f4)]


class GpuRigidBodyAccum(C.Structure):
    _fields_ = [("force_x", C.c_int32), ("force_y", C.c_int32), ("force_z", C.c_int32),
                ("contact_count", C.c_uint32), ("torque_x", C.c_int32),
                ("torque_y", C.c_int32), ("torque_z", C.c_int32), ("_pad", C.c_uint32)]


class PredictedState(C.Structure):
    _fields_ = [("pred_pos", f3), ("pressure", C.c_float), ("pred_vel", f3),
                ("pred_density", C.c_float)]


class FluidCreateInfo(C.Structure):
    _fields_ = [("particles", C.c_void_p), ("num_particles", C.c_uint32),
                ("max_particles", C.c_uint32), ("sph_params", GpuSphParams3D),
                ("container", GpuContainerGeometry), ("bounds_extent", C.c_float),
                ("device", C.c_int), ("sdf_voxels", C.c_void_p),
                ("sdf_resolution", C.c_uint32)]


class FluidSceneDesc(C.Structure):
    _fields_ = [("particles", C.c_void_p), ("num_particles", C.c_uint32),
                ("sph_params", GpuSphParams3D), ("container", GpuContainerGeometry),
                ("gravity", GpuGravity), ("bounds_extent", C.c_float)]


class GpuSprayParams(C.Structure):
    """src/state/interaction.rs:131-153 (80 B)."""
    _fields_ = [("min_speed", C.c_float), ("emission_rate", C.c_float),
                ("lifetime", C.c_float), ("lifetime_variation", C.c_float),
                ("drag", C.c_float), ("speed_multiplier", C.c_float),
                ("velocity_jitter", C.c_float), ("dt", C.c_float),
                ("max_particles", C.c_uint32), ("num_sph_particles", C.c_uint32),
                ("frame_count", C.c_uint32), ("gravity_y", C.c_float),
                ("k_trapped_air", C.c_float), ("k_wave_crest", C.c_float),
                ("ta_limit", C.c_float), ("bubble_buoyancy", C.c_float),
                ("bubble_drag", C.c_float), ("wc_limit", C.c_float),
                ("_pad", C.c_float * 2)]


class GpuSprayParticle(C.Structure):
    """src/state/interaction.rs:160-174 (48 B)."""
    _fields_ = [("pos_x", C.c_float), ("pos_y", C.c_float), ("pos_z", C.c_float),
                ("lifetime", C.c_float), ("vel_x", C.c_float), ("vel_y", C.c_float),
                ("vel_z", C.c_float), ("max_lifetime", C.c_float), ("kind", C.c_uint32),
                ("age", C.c_float), ("_pad", C.c_float * 2)]


class GpuMcGridParams(C.Structure):
    """src/render/marching_cubes.rs:55-66 (48 B; GpuGridParams there)."""
    _fields_ = [("grid_min", f3), ("grid_size", C.c_uint32), ("grid_max", f3),
                ("cell_size", C.c_float), ("kernel_radius", C.c_float), ("iso_value", C.c_float),
                ("num_particles", C.c_uint32), ("max_vertices", C.c_uint32)]


class GpuAnisoParams(C.Structure):
    """src/render/marching_cubes.rs:41-51 (32 B)."""
    _fields_ = [("enabled", C.c_uint32), ("strength", C.c_float),
                ("support_radius", C.c_float), ("h_mc", C.c_float), ("kr", C.c_float),
                ("lambda_", C.c_float), ("max_stretch", C.c_float), ("max_shift", C.c_float)]


class GpuParticleAniso(C.Structure):
    """ParticleAniso, mc_anisotropy.wgsl:46-50 (48 B)."""
    _fields_ = [("q0", C.c_float * 4), ("q1", C.c_float * 4), ("q2", C.c_float * 4)]


class FluidDeviceView(C.Structure):
    """fluid_device_view: device pointers of a handle's buffers (interop)."""
    _fields_ = [("cell_counts", C.c_void_p), ("cell_starts", C.c_void_p),
                ("num_cells", C.c_uint32), ("sorted_to_orig", C.c_void_p),
                ("num_particles", C.c_uint32), ("first_particle", C.c_uint32),
                ("sph_params", C.c_void_p), ("grid_params", C.c_void_p),
                ("container_geometry", C.c_void_p), ("gravity", C.c_void_p),
                ("stream", C.c_void_p)]


def default_aniso_params(h: float, enabled: bool = True, strength: float = 1.0,
                         h_mc: float | None = None) -> GpuAnisoParams:
    """update_aniso_params (marching_cubes.rs:2133-2146) with its constants
    (:26-34): support 2h, k_r 4, lambda 0.9, stretch cap 1.6, shift cap 0.4h."""
    a = GpuAnisoParams()
    a.enabled, a.strength = int(enabled), min(max(strength, 0.0), 1.0)
    a.support_radius, a.h_mc = 2.0 * h, h if h_mc is None else h_mc
    a.kr, a.lambda_, a.max_stretch, a.max_shift = 4.0, 0.9, 1.6, 0.4 * h
    return a


class FluidSprayState(C.Structure):
    _fields_ = [("write_head", C.c_uint32), ("ta_limit", C.c_float), ("wc_limit", C.c_float),
                ("live", C.c_uint32 * 4), ("emitted", C.c_uint32)]


# GpuSprayParticle (src/state/interaction.rs:160-174): 12 words per slot;
# word 8 (kind) is a u32, the rest f32.
SPRAY_POS = slice(0, 3)
SPRAY_LIFETIME = 3
SPRAY_VEL = slice(4, 7)
SPRAY_MAX_LIFETIME = 7
SPRAY_KIND = 8
SPRAY_AGE = 9
SPRAY_KIND_NAMES = ("spray", "foam", "bubble")


def default_spray_params(dt: float, gravity: float = 9.81, max_particles: int = 100_000,
                         num_sph_particles: int = 0, frame_count: int = 0) -> GpuSprayParams:
    """SprayConfig::default (src/state/interaction.rs:67-92) as the app packs
    it (app.rs:1048-1079): dt is the frame time, gravity_y = -gravity, and the
    ceilings at their initial guesses (spray.rs:30-31)."""
    p = GpuSprayParams()
    p.min_speed, p.emission_rate, p.lifetime, p.lifetime_variation = 0.8, 30.0, 2.8, 0.5
    p.drag, p.speed_multiplier, p.velocity_jitter, p.dt = 1.5, 0.85, 0.9, dt
    p.max_particles, p.num_sph_particles, p.frame_count = max_particles, num_sph_particles, frame_count
    p.gravity_y, p.k_trapped_air, p.k_wave_crest = -gravity, 1.25, 1.4
    p.ta_limit, p.bubble_buoyancy, p.bubble_drag, p.wc_limit = 3.0, 2.5, 0.6, 1.0
    return p


def default_mouse_force() -> GpuMouseForce:
    """GpuMouseForce::default (src/state/interaction.rs:113-126)."""
    m = GpuMouseForce()
    m.radius, m.strength, m.direction[2] = 0.5, 30.0, -1.0
    return m


def default_rigid_body() -> GpuRigidBody:
    """GpuRigidBody::default (src/state/rigid_body.rs:226-242)."""
    r = GpuRigidBody()
    r.half_extent, r.stiffness = 0.15, 200.0
    r.rot_row0[0] = r.rot_row1[1] = r.rot_row2[2] = 1.0
    return r


# name -> (restype, argtypes) of every exported function of fluid_abi.h
P = C.c_void_p
u32 = C.c_uint32
PROTOTYPES = {
    "fluid_last_error": (C.c_char_p, []),
    "fluid_abi_version": (C.c_int, []),
    "fluid_derive_sph_params": (C.c_int, [C.c_float] * 8 + [u32, C.c_float, C.POINTER(GpuSphParams3D)]),
    "fluid_derive_container_geometry": (C.c_int, [C.c_float] * 8 + [u32, C.c_float, C.POINTER(GpuContainerGeometry)]),
    "fluid_create": (C.c_int, [C.POINTER(FluidCreateInfo), C.POINTER(P)]),
    "fluid_destroy": (C.c_int, [P]),
    "fluid_step": (C.c_int, [P]),
    "fluid_step_n": (C.c_int, [P, u32]),
    "fluid_step_stages": (C.c_int, [P, u32, u32]),
    "fluid_update_sph_params": (C.c_int, [P, C.POINTER(GpuSphParams3D)]),
    "fluid_update_container_geometry": (C.c_int, [P, C.POINTER(GpuContainerGeometry)]),
    "fluid_update_mouse_force": (C.c_int, [P, C.POINTER(GpuMouseForce)]),
    "fluid_update_gravity": (C.c_int, [P, C.POINTER(GpuGravity)]),
    "fluid_update_rigid_body": (C.c_int, [P, C.POINTER(GpuRigidBody)]),
    "fluid_set_pcisph_iterations": (C.c_int, [P, u32]),
    "fluid_pcisph_iterations": (u32, [P]),
    "fluid_clear_rigid_body_accum": (C.c_int, [P]),
    "fluid_read_rigid_body_accum": (C.c_int, [P, C.POINTER(GpuRigidBodyAccum)]),
    "fluid_spawn_particles": (C.c_int, [P, C.POINTER(C.c_float), u32, C.c_float, C.POINTER(u32)]),
    "fluid_num_particles": (u32, [P]),
    "fluid_max_particles": (u32, [P]),
    "fluid_read_particles": (C.c_int, [P, P, u32]),
    "fluid_write_particles": (C.c_int, [P, P, u32]),
    "fluid_read_sorted_particles": (C.c_int, [P, P, u32]),
    "fluid_read_cell_counts": (C.c_int, [P, P, u32]),
    "fluid_read_cell_starts": (C.c_int, [P, P, u32]),
    "fluid_read_sorted_to_orig": (C.c_int, [P, P, u32]),
    "fluid_read_orig_to_sorted": (C.c_int, [P, P, u32]),
    "fluid_read_predicted": (C.c_int, [P, u32, P, u32]),
    "fluid_read_grid_params": (C.c_int, [P, C.POINTER(GpuGridParams)]),
    "fluid_read_sph_params": (C.c_int, [P, C.POINTER(GpuSphParams3D)]),
    "fluid_read_container_geometry": (C.c_int, [P, C.POINTER(GpuContainerGeometry)]),
    "fluid_sync": (C.c_int, [P]),
    "fluid_export_particles_device": (C.c_int, [P, P, u32]),
    "fluid_export_sorted_particles_device": (C.c_int, [P, P, u32]),
    "fluid_device_buffers": (C.c_int, [P, u32, C.POINTER(FluidDeviceView)]),
    "fluid_stream": (P, [P]),
    "fluid_time_steps": (C.c_int, [P, u32, u32, C.POINTER(C.c_float)]),
    "fluid_time_stages": (C.c_int, [P, u32, C.POINTER(C.c_float)]),
    "fluid_time_range": (C.c_int, [P, u32, u32, u32, C.POINTER(C.c_float)]),
    "fluid_count_pairs": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "fluid_pair_stats": (C.c_int, [P, C.POINTER(C.c_uint64)]),
    "fluid_batch_create": (C.c_int, [C.POINTER(FluidSceneDesc), u32, u32, C.c_int, C.POINTER(P)]),
    "fluid_batch_num_scenes": (u32, [P]),
    "fluid_batch_scene_info": (C.c_int, [P, u32, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32),
                                         C.POINTER(GpuGridParams)]),
    "fluid_batch_update_scene": (C.c_int, [P, u32, C.POINTER(GpuSphParams3D),
                                           C.POINTER(GpuContainerGeometry), C.POINTER(GpuGravity)]),
    "fluid_set_sweep_config": (C.c_int, [P, C.POINTER(u32)]),
    "fluid_get_sweep_config": (C.c_int, [P, C.POINTER(u32)]),
    "fluid_autotune": (C.c_int, [P, u32, C.POINTER(u32)]),
    "fluid_set_tile_config": (C.c_int, [P, C.POINTER(u32)]),
    "fluid_get_tile_config": (C.c_int, [P, C.POINTER(u32)]),
    "fluid_set_fusion": (C.c_int, [P, u32]),
    "fluid_fusion": (u32, [P]),
    "fluid_forward_rollout": (C.c_int, [P, u32, u32]),
    "fluid_backward_rollout": (C.c_int, [P, P, P, P, P, P]),
    "fluid_backward_rollout_scaled": (C.c_int, [P, P, P, P, P, P, P]),
    "fluid_rollout_timing": (C.c_int, [P, C.POINTER(C.c_float)]),
    "fluid_spray_create": (C.c_int, [P, u32, C.POINTER(GpuSprayParams), C.POINTER(P)]),
    "fluid_spray_destroy": (C.c_int, [P]),
    "fluid_spray_update_params": (C.c_int, [P, C.POINTER(GpuSprayParams)]),
    "fluid_spray_step": (C.c_int, [P, u32]),
    "fluid_spray_emit": (C.c_int, [P, u32]),
    "fluid_spray_simulate": (C.c_int, [P]),
    "fluid_spray_auto_limits": (C.c_int, [P, C.POINTER(C.c_float)]),
    "fluid_spray_read_stats": (C.c_int, [P, C.POINTER(u32)]),
    "fluid_spray_read_state": (C.c_int, [P, C.POINTER(FluidSprayState)]),
    "fluid_spray_read_particles": (C.c_int, [P, P, u32]),
    "fluid_spray_write_particles": (C.c_int, [P, P, u32]),
    "fluid_spray_read_emit": (C.c_int, [P, P, P, u32]),
    "fluid_spray_reset": (C.c_int, [P]),
    "fluid_spray_time_steps": (C.c_int, [P, u32, u32, u32, C.POINTER(C.c_float)]),
    "fluid_mc_create": (C.c_int, [P, u32, C.POINTER(P)]),
    "fluid_mc_destroy": (C.c_int, [P]),
    "fluid_mc_update_params": (C.c_int, [P, C.POINTER(GpuMcGridParams)]),
    "fluid_mc_update_aniso_params": (C.c_int, [P, C.POINTER(GpuAnisoParams)]),
    "fluid_mc_update_container_geometry": (C.c_int, [P, C.POINTER(GpuContainerGeometry)]),
    "fluid_mc_generate": (C.c_int, [P, u32]),
    "fluid_mc_read_density": (C.c_int, [P, P, C.c_uint64]),
    "fluid_mc_read_aniso": (C.c_int, [P, P, u32]),
    "fluid_mc_density_from_aniso": (C.c_int, [P, P, u32]),
    "fluid_mc_time": (C.c_int, [P, u32, u32, u32, C.POINTER(C.c_float)]),
    "fluid_mc_generate_noalloc": (C.c_int, [P, C.POINTER(GpuMcGridParams),
                                            C.POINTER(GpuAnisoParams),
                                            C.POINTER(GpuContainerGeometry), P]),
}

FLUID_FUSE_XSPH_FORCE = 1
FLUID_FUSE_SOLVE_LIST = 2
FLUID_WALK_16 = 4
FLUID_WALK_32 = 8
FLUID_SOLVE_WALK_X2 = 16
FLUID_SOLVE_WALK_X4 = 32
FLUID_SOLVE_FIRST_X2 = 64
FLUID_SOLVE_FIRST_X4 = 128
FLUID_LIST_SKIP = 256

# Gradient slots per scene written by fluid_backward_rollout.
FLUID_GRAD_PARAMS = 5
GRAD_PARAM_NAMES = ("viscosity", "surface_tension", "gravity_x", "gravity_y", "gravity_z")
# ... and by fluid_backward_rollout_scaled (stiffness appended: always zero,
# no stage of the step reads it).
FLUID_GRAD_PARAMS_ALL = 6

_LIB = None


class FluidError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"fluid error {code}: {msg}")
        self.code = code


def load_library(path: os.PathLike | str | None = None) -> C.CDLL:
    """Load libfluidhip.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(
            f"{p} is missing: build it first (make -C fluid-toy_amd or __graft_entry__.build())")
    lib = C.CDLL(str(p))
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def check(lib: C.CDLL, rc: int) -> None:
    if rc != FLUID_OK:
        raise FluidError(rc, lib.fluid_last_error().decode())


def particles_to_array(recs) -> np.ndarray:
    """(N,16) float32 view of a ctypes SphParticle3D array (copy)."""
    return np.frombuffer(recs, dtype=np.float32).reshape(-1, 16).copy()


# Column indices of the 16-float SphParticle3D record.
POS = slice(0, 3)
VEL = slice(4, 7)
FORCE = slice(8, 11)
DENSITY = 11
NEAR = 12
NORMAL = slice(13, 16)
