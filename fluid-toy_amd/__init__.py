"""fluid-toy_amd — MI355X (gfx950) HIP implementation of fluid-toy's SPH step.

The compute library is libfluidhip.so (C ABI, include/fluid_abi.h); this
package is its Python host mirror: `SphSimulation3DGrid` follows the
reference's struct of the same name (src/simulation/sph_3d_grid.rs), and
`SpraySystem` follows src/simulation/spray.rs, `MarchingCubesField` the
compute passes of src/render/marching_cubes.rs, and `scenes` holds the
seeded workloads.  Every call goes to the HIP library;
there is no CPU fallback, and a missing library raises FileNotFoundError.
"""
from . import abi, scenes
from .sim import (MarchingCubesField, SphBatch, SphSimulation3DGrid, SpraySystem,
                  derive_container_geometry, derive_sph_params, gravity, mc_generate_noalloc)


def make_sim(scene, device: int = 0, max_particles: int = 0) -> SphSimulation3DGrid:
    """Create a handle for a scenes.Scene with its parameters pushed the way
    the app does each frame (src/app.rs:963-1019)."""
    params, cont, g = scenes.make_params(scene, abi, derive_sph_params,
                                         derive_container_geometry)
    sim = SphSimulation3DGrid(scene.particles, params, cont, max_particles=max_particles,
                              bounds_extent=scene.bounds_extent, device=device)
    sim.set_pcisph_iterations(scene.pcisph_iterations)
    sim.update_gravity(g)
    return sim


def scene_tuple(scene):
    """(particles, params, container, gravity, bounds_extent) of a Scene, the
    per-scene description SphBatch takes."""
    params, cont, g = scenes.make_params(scene, abi, derive_sph_params,
                                         derive_container_geometry)
    return scene.particles, params, cont, g, scene.bounds_extent


def make_batch(scene_list, device: int = 0) -> SphBatch:
    """One SphBatch over a list of scenes.Scene (all with the same PCISPH
    iteration count)."""
    its = {sc.pcisph_iterations for sc in scene_list}
    if len(its) != 1:
        raise ValueError("scenes of a batch share one PCISPH iteration count")
    return SphBatch([scene_tuple(sc) for sc in scene_list], its.pop(), device=device)


__all__ = ["abi", "scenes", "SphSimulation3DGrid", "SphBatch", "SpraySystem",
           "MarchingCubesField", "derive_sph_params",
           "derive_container_geometry", "gravity", "mc_generate_noalloc", "make_sim", "make_batch", "scene_tuple"]
