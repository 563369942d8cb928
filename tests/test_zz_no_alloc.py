"""fluid_mc_generate_noalloc: the marching-cubes field of a simulation written
into the caller's device buffer without allocating, against the existing
entry point (a fluid_mc_* field handle's generate + read_density) on the same
state and parameters.  The field is exact (the same kernels visit the same
candidates in the same order), so the bar is bit equality.  Two calls in a
row, the second on a smaller simulation and a smaller grid, then an
isotropic call, the capacity error and the argument checks."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _params(pkg, sim, G, aniso):
    c = sim.container_geometry()
    h = sim.sph_params().kernel_radius
    f = np.float32
    m = f(h) * f(1.6) + f(0.05)
    lo = [-f(c.half_width), f(c.center_y) - f(c.half_height), -f(c.half_depth)]
    hi = [f(c.half_width), f(c.center_y) + f(c.half_height), f(c.half_depth)]
    g = pkg.abi.GpuMcGridParams()
    for k in range(3):
        g.grid_min[k] = float(f(lo[k] - m))
        g.grid_max[k] = float(f(hi[k] + m))
    ext = max(f(g.grid_max[k]) - f(g.grid_min[k]) for k in range(3))
    g.grid_size, g.cell_size, g.kernel_radius = G, float(f(ext) / f(G)), h
    a = pkg.abi.default_aniso_params(h, enabled=aniso)
    return g, a, c


def _handle_field(pkg, sim, g, a, c):
    mc = pkg.MarchingCubesField(sim, g.grid_size)
    mc.update_params(g)
    mc.update_aniso_params(a)
    mc.update_container_geometry(c)
    mc.generate()
    want = mc.read_density()
    mc.close()
    return want


def _noalloc_field(pkg, sim, g, a, c):
    G = g.grid_size
    buf = torch.full((G, G, G), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    pkg.mc_generate_noalloc(sim, g, a, c, buf.data_ptr())
    sim.sync()
    return buf.cpu().numpy()


def _same(got, want, what):
    assert (want > 0).sum() > 1000, what
    diff = np.argwhere(got.view(np.uint32) != want.view(np.uint32))
    print(f"  {what}: {len(diff)} voxels differ", flush=True)
    assert len(diff) == 0, (what, len(diff), diff[:3])


def test_noalloc_field_equals_handle(pkg):
    # first call: the 50k dam break, 100^3 voxels, anisotropy on
    big = pkg.make_sim(pkg.scenes.dam_break_50k())
    big.step(20)
    print(f"case: noalloc MC, {big.num_particles()} particles, G=100, aniso", flush=True)
    g, a, c = _params(pkg, big, 100, True)
    _same(_noalloc_field(pkg, big, g, a, c), _handle_field(pkg, big, g, a, c), "50k G=100")
    # second call: a smaller simulation and grid (the array still holds the
    # first call's records beyond this one's particles)
    small = pkg.make_sim(pkg.scenes.dam_break_8k())
    small.step(30)
    print(f"case: noalloc MC, {small.num_particles()} particles, G=64, aniso", flush=True)
    g2, a2, c2 = _params(pkg, small, 64, True)
    _same(_noalloc_field(pkg, small, g2, a2, c2), _handle_field(pkg, small, g2, a2, c2),
          "8k G=64")
    # isotropic splat
    print("case: noalloc MC, 8k, G=48, isotropic", flush=True)
    g3, a3, c3 = _params(pkg, small, 48, False)
    _same(_noalloc_field(pkg, small, g3, a3, c3), _handle_field(pkg, small, g3, a3, c3),
          "8k G=48 isotropic")
    small.close()
    big.close()


def test_noalloc_limits(pkg):
    # 42 x 42 x 40 = 70,560 particles: more than the array's 65,536 records
    sc = pkg.scenes.small_block(42, 40, 42, box=(2.0, 1.2, 2.0))
    assert sc.n > 65536
    sim = pkg.make_sim(sc)
    sim.step(1)
    print(f"case: noalloc MC limits, {sc.n} particles", flush=True)
    g, a, c = _params(pkg, sim, 32, True)
    buf = torch.zeros((32, 32, 32), dtype=torch.float32, device="cuda")
    with pytest.raises(pkg.abi.FluidError) as e:
        pkg.mc_generate_noalloc(sim, g, a, c, buf.data_ptr())
    assert e.value.code == pkg.abi.FLUID_ERR_CAPACITY
    # the isotropic splat reads no records: no limit
    g, a, c = _params(pkg, sim, 32, False)
    _same(_noalloc_field(pkg, sim, g, a, c), _handle_field(pkg, sim, g, a, c), "70k isotropic")
    with pytest.raises(pkg.abi.FluidError):
        pkg.mc_generate_noalloc(sim, g, a, c, 0)  # null field
    g.grid_size = 513
    with pytest.raises(pkg.abi.FluidError):
        pkg.mc_generate_noalloc(sim, g, a, c, buf.data_ptr())
    sim.close()
