"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the parity tests, __graft_entry__.smoke() and the
bench's cpu_baseline leg use it as the checker.  The product library never
loads it.  See fluid_oracle.cpp for what it restates and how it is pinned.
"""
from __future__ import annotations

import contextlib
import ctypes as C
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "liboracle.so"

_LIB = None


def load():
    global _LIB
    if _LIB is None:
        _LIB = _open(LIB_PATH)
    return _LIB


@contextlib.contextmanager
def library(path):
    """Use another build of fluid_oracle.cpp (a mutant built by the pins'
    sensitivity test) for every oracle call inside the block."""
    global _LIB
    saved = _LIB
    _LIB = _open(Path(path))
    try:
        yield _LIB
    finally:
        _LIB = saved


def _open(path: Path):
    if not path.exists():
        raise FileNotFoundError(f"{path} missing: run make -C oracle")
    lib = C.CDLL(str(path))
    P = C.c_void_p
    lib.oracle_create.restype = P
    lib.oracle_create.argtypes = [C.c_int, P, C.c_uint32, P, P, P, P, P, P, C.c_uint32,
                                  P, C.c_uint32]
    lib.oracle_destroy.argtypes = [P]
    lib.oracle_stage.argtypes = [P, C.c_uint32]
    lib.oracle_step.argtypes = [P, C.c_uint32]
    lib.oracle_set_params.argtypes = [P, P, P, P, C.c_uint32]
    lib.oracle_get_particles.argtypes = [P, C.c_int, P]
    lib.oracle_set_particles.argtypes = [P, P]
    lib.oracle_get_u32.argtypes = [P, C.c_int, P]
    lib.oracle_get_pred.argtypes = [P, C.c_int, P]
    lib.oracle_get_accum.argtypes = [P, P]
    lib.oracle_count_pairs.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    lib.oracle_decision_signature.restype = C.c_uint64
    lib.oracle_decision_signature.argtypes = [P]
    lib.oracle_decision_record.argtypes = [P, P]
    lib.oracle_rest_density_f32.restype = C.c_float
    lib.oracle_rest_density_f32.argtypes = [C.c_float, C.c_float]
    lib.oracle_pcisph_delta_f32.restype = C.c_float
    lib.oracle_pcisph_delta_f32.argtypes = [C.c_float] * 4
    lib.oracle_powi_f32.restype = C.c_float
    lib.oracle_powi_f32.argtypes = [C.c_float, C.c_int]
    lib.oracle_spray_potentials.argtypes = [P, C.c_uint32, P, P, P, P, P, C.c_float,
                                            C.c_float, P, P, P]
    lib.oracle_spray_spawn.restype = C.c_uint32
    lib.oracle_spray_spawn.argtypes = [P, C.c_uint32, P, P, P, C.c_uint32, P]
    lib.oracle_spray_simulate.argtypes = [P, P, P, P, P, P, P, P, P]
    lib.oracle_mc_aniso.argtypes = [P, C.c_uint32, P, P, P, P, P]
    lib.oracle_mc_density.argtypes = [P, P, P, P, P, P, P, P, P]
    lib.oracle_spray_update_limit.restype = C.c_float
    lib.oracle_spray_update_limit.argtypes = [C.c_float] * 3
    lib.oracle_spray_hash_float.restype = C.c_float
    lib.oracle_spray_hash_float.argtypes = [C.c_uint32]
    lib.oracle_spray_hash.restype = C.c_uint32
    lib.oracle_spray_hash.argtypes = [C.c_uint32]
    return lib


def grid_for(h: float, bounds_extent: float, n: int, abi_mod):
    """GpuGridParams of SphSimulation3DGrid::new (sph_3d_grid.rs:121-143),
    restated in float32 arithmetic."""
    f = np.float32
    cell = f(h)
    G = int(np.ceil(f(f(2.0) * f(bounds_extent)) / cell)) + 2
    g = abi_mod.GpuGridParams()
    g.grid_size_x = g.grid_size_y = g.grid_size_z = G
    g.total_cells = G * G * G
    g.cell_size = float(cell)
    g.inv_cell_size = float(f(1.0) / cell)
    o = float(f(-f(bounds_extent)) - cell)
    g.grid_origin_x = g.grid_origin_y = g.grid_origin_z = o
    g.num_particles = n
    return g


class Oracle:
    """One oracle state: canonical + sorted particles, grid, predicted
    buffers, in float32 (WGSL order) or float64."""

    def __init__(self, particles: np.ndarray, params, grid, container, gravity, mouse,
                 rigid_body, iterations: int, fp64: bool = False, sdf: np.ndarray | None = None):
        self._lib = load()
        self.fp64 = bool(fp64)
        self.dtype = np.float64 if fp64 else np.float32
        parts = np.ascontiguousarray(particles, dtype=self.dtype).reshape(-1, 16)
        self.n = len(parts)
        self.cells = grid.total_cells
        sdf_arr = None if sdf is None else np.ascontiguousarray(sdf, dtype=np.float32)
        self._sdf = sdf_arr
        self._h = self._lib.oracle_create(
            int(self.fp64), parts.ctypes.data, self.n, C.addressof(params), C.addressof(grid),
            C.addressof(container), C.addressof(gravity), C.addressof(mouse),
            C.addressof(rigid_body), iterations,
            None if sdf_arr is None else sdf_arr.ctypes.data,
            0 if sdf_arr is None else sdf_arr.shape[0])

    def close(self):
        if self._h:
            self._lib.oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stage(self, s: int) -> None:
        self._lib.oracle_stage(self._h, s)

    def stages(self, first: int, last: int) -> None:
        for s in range(first, last + 1):
            self.stage(s)

    def step(self, steps: int = 1) -> None:
        self._lib.oracle_step(self._h, steps)

    def set_params(self, params, container, gravity, iterations: int) -> None:
        self._lib.oracle_set_params(self._h, C.addressof(params), C.addressof(container),
                                    C.addressof(gravity), iterations)

    def particles(self) -> np.ndarray:
        out = np.zeros((self.n, 16), dtype=self.dtype)
        self._lib.oracle_get_particles(self._h, 0, out.ctypes.data)
        return out

    def set_particles(self, parts: np.ndarray) -> None:
        p = np.ascontiguousarray(parts, dtype=self.dtype).reshape(self.n, 16)
        self._lib.oracle_set_particles(self._h, p.ctypes.data)

    def sorted_particles(self) -> np.ndarray:
        out = np.zeros((self.n, 16), dtype=self.dtype)
        self._lib.oracle_get_particles(self._h, 1, out.ctypes.data)
        return out

    def _u32(self, which: int, count: int) -> np.ndarray:
        out = np.zeros(count, dtype=np.uint32)
        self._lib.oracle_get_u32(self._h, which, out.ctypes.data)
        return out

    def cell_counts(self) -> np.ndarray:
        return self._u32(0, self.cells)

    def cell_starts(self) -> np.ndarray:
        return self._u32(1, self.cells)

    def particle_cell_indices(self) -> np.ndarray:
        return self._u32(2, self.n)

    def sorted_to_orig(self) -> np.ndarray:
        return self._u32(3, self.n)

    def predicted(self, which: int) -> np.ndarray:
        out = np.zeros((self.n, 8), dtype=self.dtype)
        self._lib.oracle_get_pred(self._h, which, out.ctypes.data)
        return out

    def accum(self) -> np.ndarray:
        out = np.zeros(8, dtype=np.int32)
        self._lib.oracle_get_accum(self._h, out.ctypes.data)
        return out

    def decision_signature(self) -> int:
        """Hash-sum of the non-smooth branch decisions taken since creation
        (close-range switch and its slot difference, viscosity on, Akinci
        inner piece, positive density error, CFL clamp, wall layer and its
        damping, backstop): equal for two runs that took the same branches."""
        return int(self._lib.oracle_decision_signature(self._h))

    def decision_record(self) -> np.ndarray:
        """The branch-decision record per original index (uint64)."""
        out = np.zeros(self.n, dtype=np.uint64)
        self._lib.oracle_decision_record(self._h, out.ctypes.data)
        return out

    def count_pairs(self) -> tuple[int, int]:
        a, b = C.c_uint64(0), C.c_uint64(0)
        self._lib.oracle_count_pairs(self._h, C.byref(a), C.byref(b))
        return a.value, b.value


def rest_density_f32(h: float, mass: float) -> float:
    return load().oracle_rest_density_f32(h, mass)


def pcisph_delta_f32(h: float, mass: float, rho0: float, dt: float) -> float:
    return load().oracle_pcisph_delta_f32(h, mass, rho0, dt)


def powi_f32(x: float, n: int) -> float:
    return load().oracle_powi_f32(x, n)


def from_scene(scene, abi_mod, params, container, gravity, fp64: bool = False,
               particles: np.ndarray | None = None, mouse=None, rigid_body=None, sdf=None):
    """An Oracle for a scenes.Scene with the given uniforms."""
    from importlib import import_module  # noqa: F401
    g = grid_for(scene.h, scene.bounds_extent, scene.n, abi_mod)
    if mouse is None:
        mouse = abi_mod.GpuMouseForce()
        mouse.radius, mouse.strength, mouse.direction[2] = 0.5, 30.0, -1.0
    if rigid_body is None:
        rigid_body = abi_mod.GpuRigidBody()
        rigid_body.half_extent, rigid_body.stiffness = 0.15, 200.0
        rigid_body.rot_row0[0] = rigid_body.rot_row1[1] = rigid_body.rot_row2[2] = 1.0
    parts = scene.particles if particles is None else particles
    return Oracle(parts, params, g, container, gravity, mouse, rigid_body,
                  scene.pcisph_iterations, fp64=fp64, sdf=sdf)


# ---------------------------------------------------------------------------
# Whitewater passes (spray_emit.wgsl, spray_simulate.wgsl) over a sorted
# particle array (N, 16) float32 and its grid.  Rings are (max, 12) uint32
# arrays holding the GpuSprayParticle words (kind is word 8).

def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def spray_potentials(sorted_particles: np.ndarray, counts: np.ndarray, starts: np.ndarray,
                     grid, sph, spray_params, ta_limit: float, wc_limit: float):
    """(potentials (N, 4) float32, n_emit (N,) uint32, frame_max (2,) float32)."""
    lib = load()
    sp = np.ascontiguousarray(sorted_particles, dtype=np.float32).reshape(-1, 16)
    cnt = np.ascontiguousarray(counts, dtype=np.uint32)
    st = np.ascontiguousarray(starts, dtype=np.uint32)
    n = len(sp)
    pot = np.zeros((n, 4), np.float32)
    ne = np.zeros(n, np.uint32)
    fm = np.zeros(2, np.float32)
    lib.oracle_spray_potentials(_ptr(sp), n, _ptr(cnt), _ptr(st), C.byref(grid), C.byref(sph),
                                C.byref(spray_params), ta_limit, wc_limit, _ptr(pot), _ptr(ne),
                                _ptr(fm))
    return pot, ne, fm


def spray_spawn(sorted_particles: np.ndarray, sph, spray_params, n_emit: np.ndarray, head: int,
                ring: np.ndarray) -> int:
    """Writes the spawns into `ring` (modified in place); returns the new head."""
    lib = load()
    sp = np.ascontiguousarray(sorted_particles, dtype=np.float32).reshape(-1, 16)
    ne = np.ascontiguousarray(n_emit, dtype=np.uint32)
    assert ring.dtype == np.uint32 and ring.flags.c_contiguous and ring.shape[1] == 12
    return int(lib.oracle_spray_spawn(_ptr(sp), len(sp), C.byref(sph), C.byref(spray_params),
                                      _ptr(ne), head & 0xFFFFFFFF, _ptr(ring)))


def spray_simulate(ring: np.ndarray, sorted_particles: np.ndarray, counts: np.ndarray,
                   starts: np.ndarray, grid, sph, container, spray_params) -> np.ndarray:
    """Advances `ring` in place; returns the live counts (total, spray, foam, bubble)."""
    lib = load()
    sp = np.ascontiguousarray(sorted_particles, dtype=np.float32).reshape(-1, 16)
    cnt = np.ascontiguousarray(counts, dtype=np.uint32)
    st = np.ascontiguousarray(starts, dtype=np.uint32)
    assert ring.dtype == np.uint32 and ring.flags.c_contiguous and ring.shape[1] == 12
    live = np.zeros(4, np.uint32)
    lib.oracle_spray_simulate(_ptr(ring), _ptr(sp), _ptr(cnt), _ptr(st), C.byref(grid),
                              C.byref(sph), C.byref(container), C.byref(spray_params),
                              _ptr(live))
    return live


def spray_update_limit(current: float, frame_max: float, floor: float) -> float:
    return float(load().oracle_spray_update_limit(current, frame_max, floor))


# ---------------------------------------------------------------------------
# Marching-cubes density field (mc_anisotropy.wgsl, mc_density.wgsl).

def mc_aniso(sorted_particles: np.ndarray, counts: np.ndarray, starts: np.ndarray, grid,
             aniso_params) -> np.ndarray:
    """(N, 12) float32 ellipsoid records of the sorted particles."""
    lib = load()
    sp = np.ascontiguousarray(sorted_particles, dtype=np.float32).reshape(-1, 16)
    cnt = np.ascontiguousarray(counts, dtype=np.uint32)
    st = np.ascontiguousarray(starts, dtype=np.uint32)
    out = np.zeros((len(sp), 12), np.float32)
    lib.oracle_mc_aniso(_ptr(sp), len(sp), _ptr(cnt), _ptr(st), C.byref(grid),
                        C.byref(aniso_params), _ptr(out))
    return out


def mc_density(sorted_particles: np.ndarray, counts: np.ndarray, starts: np.ndarray, grid,
               mc_params, aniso_params, container, aniso: np.ndarray | None = None) -> np.ndarray:
    """The (G, G, G) float32 field indexed [z, y, x]."""
    lib = load()
    sp = np.ascontiguousarray(sorted_particles, dtype=np.float32).reshape(-1, 16)
    cnt = np.ascontiguousarray(counts, dtype=np.uint32)
    st = np.ascontiguousarray(starts, dtype=np.uint32)
    an = np.zeros((max(len(sp), 1), 12), np.float32) if aniso is None else \
        np.ascontiguousarray(aniso, dtype=np.float32)
    G = mc_params.grid_size
    out = np.zeros((G, G, G), np.float32)
    lib.oracle_mc_density(_ptr(sp), _ptr(cnt), _ptr(st), C.byref(grid), C.byref(mc_params),
                          C.byref(aniso_params), C.byref(container), _ptr(an), _ptr(out))
    return out

