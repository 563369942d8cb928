"""Device-side access to the step's buffers (the reference's getters,
src/simulation/sph_3d_grid.rs:1500-1526): the records exported into a
caller's device memory, and the device pointers of the grid, the slot map
and the uniform blocks, all equal bit for bit to the host readbacks of the
same state."""
import ctypes as C
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

_HIP = None


def d2h(ptr, nbytes):
    """Bytes at a device address, copied with the HIP runtime torch loaded."""
    global _HIP
    if _HIP is None:
        _HIP = C.CDLL(os.path.join(os.path.dirname(torch.__file__), "lib", "libamdhip64.so"))
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP.hipMemcpy.restype = C.c_int
    out = np.zeros(nbytes, np.uint8)
    assert _HIP.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), nbytes, 2) == 0  # device to host
    return out


def exported(sim, fn):
    n = sim.num_particles()
    buf = torch.full((n, 16), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    fn(buf.data_ptr(), n)
    sim.sync()
    return buf.cpu().numpy()


def same_bits(a, b, what):
    assert a.shape == b.shape, what
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), what


def test_export_records_equal_readbacks(pkg):
    sim = pkg.make_sim(pkg.scenes.dam_break_8k())
    sim.step(3)
    print("case: export after 3 steps", flush=True)
    same_bits(exported(sim, sim.export_particles_device), sim.read_particles(), "canonical")
    same_bits(exported(sim, sim.export_sorted_particles_device), sim.read_sorted_particles(),
              "sorted")
    # at stage boundaries inside a step the canonical fields come from
    # different buffers (sorted copy, XSPH output, force output)
    S = pkg.abi.STAGE
    done = -1
    for stage in ("reorder", "density", "xsph", "force", "integrate"):
        sim.step_stages(done + 1, S[stage])
        done = S[stage]
        print(f"case: export at stage {stage}", flush=True)
        same_bits(exported(sim, sim.export_particles_device), sim.read_particles(), stage)
    sim.close()


def test_export_keeps_canonical_fields_and_spawns(pkg):
    """Density, near density and normal of the canonical records are the
    values written in (the step never writes them), also for spawned
    particles (zero)."""
    sc = pkg.scenes.small_block(8, 8, 8)
    parts = sc.particles.copy()
    rng = np.random.default_rng(3)
    parts[:, 11:16] = rng.normal(size=(sc.n, 5)).astype(np.float32)
    sc.particles = parts
    sim = pkg.make_sim(sc, max_particles=sc.n + 40)
    sim.step(2)
    assert sim.spawn_particles([0.1, 0.5, 0.0], 40, 0.05) == 40
    sim.step(1)
    got = exported(sim, sim.export_particles_device)
    same_bits(got, sim.read_particles(), "canonical after spawn")
    np.testing.assert_array_equal(got[:sc.n, 11:16], parts[:, 11:16])
    assert np.all(got[sc.n:, 11:16] == 0)
    sim.close()


def test_export_guards(pkg):
    sim = pkg.make_sim(pkg.scenes.small_block(4, 4, 4))
    buf = torch.zeros((64, 16), dtype=torch.float32, device="cuda")
    with pytest.raises(pkg.abi.FluidError):
        sim.export_particles_device(buf.data_ptr(), 63)  # too small
    with pytest.raises(pkg.abi.FluidError):
        sim.export_particles_device(buf.data_ptr() + 4, 64)  # misaligned
    with pytest.raises(pkg.abi.FluidError):
        sim.device_buffers(1)  # a plain handle has scene 0 only
    sim.close()


def check_view(pkg, sim, v, counts, starts, sid, params, grid, cont, grav):
    same_bits(d2h(v.cell_counts, 4 * v.num_cells).view(np.uint32), counts, "counts")
    same_bits(d2h(v.cell_starts, 4 * v.num_cells).view(np.uint32), starts, "starts")
    same_bits(d2h(v.sorted_to_orig, 4 * v.num_particles).view(np.uint32), sid, "sid")
    for ptr, st in ((v.sph_params, params), (v.grid_params, grid),
                    (v.container_geometry, cont), (v.gravity, grav)):
        want = np.frombuffer(bytes(st), np.uint8)
        assert np.array_equal(d2h(ptr, C.sizeof(st)), want), type(st).__name__


def test_device_buffers_plain(pkg):
    sc = pkg.scenes.dam_break_8k()
    sim = pkg.make_sim(sc)
    sim.step(2)
    g = pkg.gravity(0.5, -9.0, 0.25)
    sim.update_gravity(g)
    v = sim.device_buffers()
    assert v.num_particles == sc.n and v.first_particle == 0
    assert v.stream == pkg.abi.load_library().fluid_stream(sim._h)
    sim.sync()
    check_view(pkg, sim, v, sim.read_cell_counts(), sim.read_cell_starts(),
               sim.read_sorted_to_orig(), sim.sph_params(), sim.grid_params(),
               sim.container_geometry(), g)
    sim.close()


def test_device_buffers_batch(pkg):
    scenes = pkg.scenes.batch_scenes(3, seed=11, n_min=1000, n_max=3000)
    batch = pkg.make_batch(scenes)
    batch.step(2)
    counts, starts = batch.read_cell_counts(), batch.read_cell_starts()
    sid = batch.read_sorted_to_orig()
    for k, sc in enumerate(scenes):
        info = batch.scene_info(k)
        v = batch.device_buffers(k)
        assert v.first_particle == info["first"] and v.num_particles == sc.n
        cl = slice(info["first_cell"], info["first_cell"] + info["cells"])
        sl = slice(info["first"], info["first"] + sc.n)
        params, cont, g = pkg.scenes.make_params(sc, pkg.abi, pkg.derive_sph_params,
                                                 pkg.derive_container_geometry)
        params.num_particles = sc.n
        check_view(pkg, batch, v, counts[cl], starts[cl], sid[sl], params, info["grid"], cont, g)
    batch.close()
