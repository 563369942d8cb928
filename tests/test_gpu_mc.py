"""Marching-cubes density field on the GPU (fluid_mc_*) against the oracle,
on the simulation's own sorted particles and grid.

Bars: the density field is the oracle's to the bit, isotropic or from the
GPU's own ellipsoid records (each voxel sums the same candidates in the
same order with the same f32 operations); the ellipsoid records (their
covariance sums are split over 16 lanes) agree to 1e-4 of the kernel scale
1/h in G, 1e-5 h in reach, 1e-4 relative in amplitude, and 5e-5 h in the
centre (1e-4 of the largest centre shift 0.4 h: the shift scales with the
flatness 1 - (s3/s1)^2, an eigenvalue ratio that carries the sums' rounding
into it).  Observed at 8k: 3e-6, 4e-6, 6e-7, 2e-6; at 50k: 2e-5 (G), 7e-6
(centre).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(sim):
    return (sim.read_sorted_particles(), sim.read_cell_counts(), sim.read_cell_starts(),
            sim.grid_params())


def _default_params(pkg, sim, G):
    """The handle's defaults restated: the container box +- (1.6 h + 0.05),
    cubic cells of the largest extent / G."""
    c = sim.container_geometry()
    h = sim.sph_params().kernel_radius
    f = np.float32
    m = f(h) * f(1.0) * f(1.6) + f(0.05)
    lo = [-f(c.half_width), f(c.center_y) - f(c.half_height), -f(c.half_depth)]
    hi = [f(c.half_width), f(c.center_y) + f(c.half_height), f(c.half_depth)]
    g = pkg.abi.GpuMcGridParams()
    for k in range(3):
        g.grid_min[k] = float(f(lo[k] - m))
        g.grid_max[k] = float(f(hi[k] + m))
    ext = max(f(g.grid_max[k]) - f(g.grid_min[k]) for k in range(3))
    g.grid_size, g.cell_size, g.kernel_radius = G, float(f(ext) / f(G)), h
    return g, c


@pytest.mark.parametrize("G", [37, 64])
def test_isotropic_field_bitwise(pkg, orc, G):
    sim = pkg.make_sim(pkg.scenes.dam_break_8k())
    sim.step(30)
    mc = pkg.MarchingCubesField(sim, G)
    h = sim.sph_params().kernel_radius
    a = pkg.abi.default_aniso_params(h, enabled=False)
    mc.update_aniso_params(a)
    g, cont = _default_params(pkg, sim, G)
    mc.update_params(g)
    print(f"case: isotropic MC field G={G}", flush=True)
    mc.generate()
    got = mc.read_density()
    recs, cnt, st, grid = _state(sim)
    want = orc.mc_density(recs, cnt, st, grid, g, a, cont)
    assert (want > 0).sum() > 1000
    diff = np.argwhere(got != want)
    assert len(diff) == 0, (len(diff), diff[:3])
    mc.close()
    sim.close()


def test_aniso_records_and_field(pkg, orc):
    sim = pkg.make_sim(pkg.scenes.dam_break_8k())
    sim.step(30)
    G = 48
    mc = pkg.MarchingCubesField(sim, G)
    h = sim.sph_params().kernel_radius
    a = pkg.abi.default_aniso_params(h)
    g, cont = _default_params(pkg, sim, G)
    mc.update_params(g)
    print("case: anisotropic MC field", flush=True)
    mc.generate()
    rec = mc.read_aniso()
    recs, cnt, st, grid = _state(sim)
    want = orc.mc_aniso(recs, cnt, st, grid, a)
    gcols = [0, 1, 2, 4, 5, 6]
    eg = np.abs(rec[:, gcols] - want[:, gcols]).max() * h
    ec = np.abs(rec[:, [3, 7, 8]] - want[:, [3, 7, 8]]).max() / h
    er = np.abs(rec[:, 9] - want[:, 9]).max() / h
    ea = (np.abs(rec[:, 10] - want[:, 10]) / want[:, 10]).max()
    print(f"  G {eg:.2e} (x 1/h)  centre {ec:.2e} h  reach {er:.2e} h  amplitude {ea:.2e}",
          flush=True)
    assert eg < 1e-4 and ec < 5e-5 and er < 1e-5 and ea < 1e-4
    # the field from the GPU's own records, to the bit
    got = mc.read_density()
    field = orc.mc_density(recs, cnt, st, grid, g, a, cont, rec)
    assert np.array_equal(got, field)
    # and from records handed in
    rec2 = rec.copy()
    rec2[:, 10] *= 0.5
    mc.density_from_aniso(rec2)
    field2 = orc.mc_density(recs, cnt, st, grid, g, a, cont, rec2)
    assert np.array_equal(mc.read_density(), field2)
    mc.close()
    sim.close()


def test_clipped_tilted_container(pkg, orc):
    """The MC pass's own container (tilted 0.2 rad about x, clipping on):
    voxels outside hold -1, the rest are the oracle's to the bit."""
    sim = pkg.make_sim(pkg.scenes.dam_break_8k())
    sim.step(10)
    G = 40
    mc = pkg.MarchingCubesField(sim, G)
    h = sim.sph_params().kernel_radius
    a = pkg.abi.default_aniso_params(h, enabled=False)
    mc.update_aniso_params(a)
    g, _ = _default_params(pkg, sim, G)
    mc.update_params(g)
    cont = pkg.derive_container_geometry(1.6, 1.2, 0.8, 0.0, 0.2, 0.0, clip_enabled=True,
                                         clip_margin=0.02)
    mc.update_container_geometry(cont)
    mc.generate()
    got = mc.read_density()
    recs, cnt, st, grid = _state(sim)
    want = orc.mc_density(recs, cnt, st, grid, g, a, cont)
    assert (want == -1).sum() > 100 and np.array_equal(got, want)
    mc.close()
    sim.close()


@pytest.mark.parametrize("aniso", [False, True])
def test_field_beyond_particle_grid(pkg, orc, aniso):
    """A simulation grid smaller than the container (extent 0.5: cells span
    [-0.58, 0.62] per axis): the build clamps the particles outside it into
    the border cells, which the field's stencils and its per-block cull
    treat as reaching to infinity.  The field covers the whole container, so
    voxel blocks lie beyond the grid on every side; it must still be the
    oracle's to the bit."""
    scene = pkg.scenes.dam_break_8k()
    scene.bounds_extent = 0.5
    sim = pkg.make_sim(scene)
    sim.step(10)
    grid = sim.grid_params()
    lo = grid.grid_origin_x
    hi = lo + grid.grid_size_x * grid.cell_size
    pos = sim.read_particles()[:, 0:3]
    outside = (pos < lo).any(1) | (pos >= hi).any(1)
    print(f"case: MC beyond the particle grid, aniso={aniso}, "
          f"{int(outside.sum())} particles outside [{lo:.2f}, {hi:.2f})", flush=True)
    assert outside.sum() > 100
    G = 40
    mc = pkg.MarchingCubesField(sim, G)
    h = sim.sph_params().kernel_radius
    a = pkg.abi.default_aniso_params(h, enabled=aniso)
    mc.update_aniso_params(a)
    g, cont = _default_params(pkg, sim, G)
    mc.update_params(g)
    mc.generate()
    got = mc.read_density()
    recs, cnt, st, grid = _state(sim)
    want = orc.mc_density(recs, cnt, st, grid, g, a, cont, mc.read_aniso() if aniso else None)
    assert (want > 0).sum() > 1000
    diff = np.argwhere(got != want)
    assert len(diff) == 0, (len(diff), diff[:3])
    mc.close()
    sim.close()


def test_mc_guards(pkg):
    import ctypes as C
    sim = pkg.make_sim(pkg.scenes.dam_break_8k())
    lib = pkg.abi.load_library()
    h = C.c_void_p()
    assert lib.fluid_mc_create(sim._h, 0, C.byref(h)) == pkg.abi.FLUID_ERR_INVALID_ARG
    assert lib.fluid_mc_create(sim._h, 513, C.byref(h)) == pkg.abi.FLUID_ERR_INVALID_ARG
    mc = pkg.MarchingCubesField(sim, 16)
    g = pkg.abi.GpuMcGridParams()
    g.grid_size, g.cell_size, g.kernel_radius = 17, 0.1, 0.08
    with pytest.raises(pkg.abi.FluidError):
        mc.update_params(g)
    assert lib.fluid_destroy(sim._h) == pkg.abi.FLUID_ERR_STATE
    mc.close()
    sim.close()


def test_mc_full_size_50k(pkg, orc):
    """The 50k dam break after 60 steps (the bench workload) on a 100^3
    grid: isotropic field and the field from the GPU's records to the bit,
    records within the bars above."""
    sim = pkg.make_sim(pkg.scenes.dam_break_50k())
    sim.step(60)
    G = 100
    mc = pkg.MarchingCubesField(sim, G)
    h = sim.sph_params().kernel_radius
    g, cont = _default_params(pkg, sim, G)
    mc.update_params(g)
    recs, cnt, st, grid = _state(sim)
    print("case: MC full size, isotropic", flush=True)
    iso = pkg.abi.default_aniso_params(h, enabled=False)
    mc.update_aniso_params(iso)
    mc.generate()
    assert np.array_equal(mc.read_density(), orc.mc_density(recs, cnt, st, grid, g, iso, cont))
    print("case: MC full size, anisotropic", flush=True)
    a = pkg.abi.default_aniso_params(h)
    mc.update_aniso_params(a)
    mc.generate()
    rec = mc.read_aniso()
    want = orc.mc_aniso(recs, cnt, st, grid, a)
    eg = np.abs(rec[:, [0, 1, 2, 4, 5, 6]] - want[:, [0, 1, 2, 4, 5, 6]]).max() * h
    ec = np.abs(rec[:, [3, 7, 8]] - want[:, [3, 7, 8]]).max() / h
    print(f"  G {eg:.2e} (x 1/h)  centre {ec:.2e} h", flush=True)
    assert eg < 1e-4 and ec < 5e-5
    field = orc.mc_density(recs, cnt, st, grid, g, a, cont, rec)
    assert np.array_equal(mc.read_density(), field)
    mc.close()
    sim.close()


def test_mc_and_spray_on_an_empty_then_spawned_simulation(pkg, orc):
    """No particles: the field is zero (with the wall factor, still zero)
    and nothing spawns; after a spawn and a step both passes see the new
    particles, the field again the oracle's to the bit; a 1-voxel grid."""
    scene = pkg.scenes.small_block(1, 1, 1)
    scene.particles = scene.particles[:0]
    sim = pkg.make_sim(scene, max_particles=64)
    sim.step(1)
    mc = pkg.MarchingCubesField(sim, 1)
    mc.generate()
    assert mc.read_density().shape == (1, 1, 1) and mc.read_density()[0, 0, 0] == 0.0
    mc.close()
    cap = 32
    prm = pkg.abi.default_spray_params(dt=0.016, max_particles=cap, num_sph_particles=0)
    spray = pkg.SpraySystem(sim, cap, prm)
    spray.step(0)
    assert spray.read_stats()[0] == 0 and spray.state()["write_head"] == 0
    assert sim.spawn_particles([0.0, 0.5, 0.0], 40, 0.05) == 40
    sim.step(2)
    prm.num_sph_particles = sim.num_particles()
    spray.update_params(prm)
    spray.step(sim.num_particles())
    pot, ne = spray.read_emit(sim.num_particles())
    assert pot.shape == (40, 4)
    spray.close()
    mc = pkg.MarchingCubesField(sim, 24)
    h = sim.sph_params().kernel_radius
    a = pkg.abi.default_aniso_params(h, enabled=False)
    mc.update_aniso_params(a)
    g, cont = _default_params(pkg, sim, 24)
    mc.update_params(g)
    mc.generate()
    recs, cnt, st, grid = _state(sim)
    want = orc.mc_density(recs, cnt, st, grid, g, a, cont)
    assert (want > 0).any() and np.array_equal(mc.read_density(), want)
    mc.close()
    sim.close()
