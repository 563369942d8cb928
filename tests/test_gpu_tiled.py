"""LDS-staged tiled sweeps (fluid_set_tile_config) against the global-memory
sweeps: the tiled density and solve visit every particle's candidates in the
same order with the same arithmetic, so every field must agree bit for bit
- for every tile width (16 cells overflows the LDS stage on this lattice and
exercises the per-particle global fallback), every lane count, particles
outside the grid (which sweep around their own coordinates, not the
tile's), a batched handle, and the graph-launched step."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(pkg, scene, w, lanes):
    a = pkg.make_sim(scene)
    b = pkg.make_sim(scene)
    for s in (a, b):
        s.set_sweep_config(lanes, lanes, 0)
    a.set_tile_config(0)
    b.set_tile_config(1, w)
    return a, b


def _same(a, b, what):
    pa, pb = a.read_particles(), b.read_particles()
    assert np.array_equal(pa, pb), what
    sa, sb = a.read_sorted_particles(), b.read_sorted_particles()
    assert np.array_equal(sa, sb), what
    for k in (0, 1):
        assert np.array_equal(a.read_predicted(k), b.read_predicted(k)), what


@pytest.mark.parametrize("w", [1, 2, 4, 8, 16])
def test_tiled_equals_global_widths(pkg, w):
    scene = pkg.scenes.dam_break_8k()
    a, b = _pair(pkg, scene, w, 8)
    print(f"case: dam_break_8k w={w}", flush=True)
    for step in range(6):
        a.step(1)
        b.step(1)
        _same(a, b, f"w={w} step={step}")
    a.close()
    b.close()


@pytest.mark.parametrize("lanes", [1, 2, 4])
def test_tiled_equals_global_lanes(pkg, lanes):
    scene = pkg.scenes.dam_break_8k()
    a, b = _pair(pkg, scene, 4, lanes)
    print(f"case: dam_break_8k lanes={lanes}", flush=True)
    for step in range(3):
        a.step(1)
        b.step(1)
        _same(a, b, f"lanes={lanes} step={step}")
    a.close()
    b.close()


def test_tiled_out_of_grid(pkg):
    scene = pkg.scenes.dam_break_8k()
    parts = scene.particles.copy()
    rng = np.random.default_rng(1)
    parts[:64, 0:3] = rng.uniform([-2.5, 2.1, -2.5], [2.5, 2.6, 2.5], size=(64, 3))
    scene.particles = parts
    a, b = _pair(pkg, scene, 4, 8)
    print("case: out-of-grid particles, tiled", flush=True)
    for step in range(3):
        a.step(1)
        b.step(1)
        _same(a, b, f"oog step={step}")
    a.close()
    b.close()


def test_tiled_batch_equals_global(pkg):
    scs = pkg.scenes.batch_scenes(count=6, seed=11, n_min=1000, n_max=6000)
    a = pkg.make_batch(scs)
    b = pkg.make_batch(scs)
    a.set_tile_config(0)
    b.set_tile_config(1, 4)
    print("case: 6-scene batch, tiled", flush=True)
    for step in range(4):
        a.step(1)
        b.step(1)
        assert np.array_equal(a.read_particles(), b.read_particles()), step
    a.close()
    b.close()


def test_tiled_stagewise_matches_graph(pkg):
    """Stage-by-stage launches and the captured step agree in tiled mode."""
    scene = pkg.scenes.dam_break_8k()
    a = pkg.make_sim(scene)
    b = pkg.make_sim(scene)
    a.set_tile_config(1, 4)
    b.set_tile_config(1, 4)
    for _ in range(3):
        a.step(1)
        b.step_stages(0, 10)
    assert np.array_equal(a.read_particles(), b.read_particles())
    a.close()
    b.close()


@pytest.mark.parametrize("wide", [0, 4, 8])
def test_fused_xsph_force_equals_separate(pkg, wide):
    """The fused XSPH + force kernel of the captured step gives the bits of
    the two separate kernels, with global and with tiled sweeps, and with
    the 16- and 32-lane walks (FLUID_WALK_16 = 4, FLUID_WALK_32 = 8)."""
    scene = pkg.scenes.dam_break_8k()
    for tiled in (0, 1):
        a = pkg.make_sim(scene)
        b = pkg.make_sim(scene)
        for s in (a, b):
            s.set_sweep_config(8, 8, 0)
        a.set_fusion(wide)
        b.set_fusion(pkg.abi.FLUID_FUSE_XSPH_FORCE | wide)
        for s in (a, b):
            s.set_tile_config(tiled, 4)
        print(f"case: fused xsph+force tiled={tiled} wide={wide}", flush=True)
        for step in range(4):
            a.step(1)
            b.step(1)
            _same(a, b, f"fused tiled={tiled} step={step}")
        a.close()
        b.close()


@pytest.mark.parametrize("lanes", [1, 4, 8, 16])
def test_solve_skin_list_equals_sweep(pkg, lanes):
    """Later PCISPH iterations over the first iteration's skin list give the
    bits of the 27-cell sweep (plain and tiled), on the compressed lattice and
    after it has exploded into fast clusters."""
    scene = pkg.scenes.dam_break_8k()
    full = pkg.abi.FLUID_FUSE_XSPH_FORCE | pkg.abi.FLUID_FUSE_SOLVE_LIST
    for tiled in (0, 1):
        a = pkg.make_sim(scene)
        b = pkg.make_sim(scene)
        for s in (a, b):
            s.set_sweep_config(lanes, lanes, 0)
            s.set_tile_config(tiled, 4)
        a.set_fusion(pkg.abi.FLUID_FUSE_XSPH_FORCE)
        b.set_fusion(full)
        print(f"case: solve skin list lanes={lanes} tiled={tiled}", flush=True)
        for step in range(12):
            a.step(1)
            b.step(1)
            _same(a, b, f"skin lanes={lanes} tiled={tiled} step={step}")
        a.close()
        b.close()


def test_wide_walk_flags(pkg):
    """The two wide walks exclude each other; the flags read back as set."""
    sim = pkg.make_sim(pkg.scenes.dam_break_8k())
    with pytest.raises(pkg.abi.FluidError):
        sim.set_fusion(pkg.abi.FLUID_WALK_16 | pkg.abi.FLUID_WALK_32)
    for flags in (pkg.abi.FLUID_WALK_16, pkg.abi.FLUID_WALK_32 | pkg.abi.FLUID_FUSE_XSPH_FORCE):
        sim.set_fusion(flags)
        assert sim.fusion() == flags
    sim.close()


@pytest.mark.parametrize("wide", [0, 4, 8])
def test_density_16_lanes(pkg, wide):
    """Density with 16 lanes per particle writes the same 8-lane in-range
    lists (the same pairs in the same order) as with 8: from the same state,
    one step differs only through the density sums' grouping.  Density and
    near density agree to float rounding (1e-5); velocity and position to
    the PCISPH solve's own rounding noise (2e-3, per particle with a floor
    of 1% of the field RMS, as the oracle parity tests)."""
    from conftest import rowwise_rel_err
    scene = pkg.scenes.dam_break_8k()
    a = pkg.make_sim(scene)
    b = pkg.make_sim(scene)
    a.set_sweep_config(8, 8, 0)
    b.set_sweep_config(16, 8, 0)
    for s in (a, b):
        s.set_fusion(s.fusion() | wide)
    print(f"case: density 16 lanes, walk flag={wide}", flush=True)
    for step in range(3):
        a.write_particles(b.read_particles())
        a.step(1)
        b.step(1)
        pa, pb = a.read_sorted_particles(), b.read_sorted_particles()
        assert np.array_equal(a.read_sorted_to_orig(), b.read_sorted_to_orig()), step
        for cols, name, tol in ((slice(11, 12), "density", 1e-5), (slice(12, 13), "near", 1e-5),
                                (slice(4, 7), "velocity", 2e-3), (slice(0, 3), "position", 2e-3)):
            err = rowwise_rel_err(pb[:, cols], pa[:, cols])
            print(f"  step {step} {name}: max {err.max():.2e}", flush=True)
            assert err.max() < tol, (step, name, err.max())
    a.close()
    b.close()


# walkers per list lane need at least 2 solve lanes; the wide first iteration
# works from 1
_SOLVE_CASES = [(ls, w) for ls in (1, 2, 4, 8)
                for w in ("x2", "x4", "first_x2", "first_x4", "both_x2")
                if ls > 1 or w.startswith("first")]


@pytest.mark.parametrize("solve_lanes,walk", _SOLVE_CASES)
def test_solve_wide_walk(pkg, solve_lanes, walk):
    """The later PCISPH iterations with two or four walkers per skin-list
    lane, and the first iteration sweeping with two or four lanes per list
    lane (its entries ranked into the lists by ballot), sum the same pairs
    as one lane per list lane, grouped otherwise: from the same state, one
    step agrees to the solve's rounding noise (predicted density 1e-5;
    velocity and position 2e-3, per particle with a floor of 1% of the
    field RMS, as the oracle parity tests; pressure 2e-3 likewise)."""
    from conftest import rowwise_rel_err
    A = pkg.abi
    flag = {"x2": A.FLUID_SOLVE_WALK_X2, "x4": A.FLUID_SOLVE_WALK_X4,
            "first_x2": A.FLUID_SOLVE_FIRST_X2, "first_x4": A.FLUID_SOLVE_FIRST_X4,
            "both_x2": A.FLUID_SOLVE_WALK_X2 | A.FLUID_SOLVE_FIRST_X2}[walk]
    base = pkg.abi.FLUID_FUSE_XSPH_FORCE | pkg.abi.FLUID_FUSE_SOLVE_LIST
    scene = pkg.scenes.dam_break_8k()
    a = pkg.make_sim(scene)
    b = pkg.make_sim(scene)
    for s in (a, b):
        s.set_sweep_config(8, solve_lanes, 0)
        s.set_tile_config(0)
    a.set_fusion(base)
    b.set_fusion(base | flag)
    assert b.fusion() == base | flag
    print(f"case: solve walk {walk} solve_lanes={solve_lanes}", flush=True)
    for step in range(4):
        a.write_particles(b.read_particles())
        a.step(1)
        b.step(1)
        pa, pb = a.read_sorted_particles(), b.read_sorted_particles()
        assert np.array_equal(a.read_sorted_to_orig(), b.read_sorted_to_orig()), step
        rows = [(pb[:, sl], pa[:, sl], name, tol)
                for sl, name, tol in ((slice(4, 7), "velocity", 2e-3),
                                      (slice(0, 3), "position", 2e-3))]
        for k in (0, 1):
            qa, qb = a.read_predicted(k), b.read_predicted(k)
            rows += [(qb[:, 7:8], qa[:, 7:8], f"predicted density {k}", 1e-5),
                     (qb[:, 3:4], qa[:, 3:4], f"pressure {k}", 2e-3)]
        for got, want, name, tol in rows:
            err = rowwise_rel_err(got, want)
            print(f"  step {step} {name}: max {err.max():.2e}", flush=True)
            assert err.max() < tol, (step, name, err.max())
    a.close()
    b.close()


def test_solve_walk_flags(pkg):
    """The two solve walks exclude each other, and so do the two wide first
    iterations; the flags read back as set."""
    sim = pkg.make_sim(pkg.scenes.dam_break_8k())
    with pytest.raises(pkg.abi.FluidError):
        sim.set_fusion(pkg.abi.FLUID_SOLVE_WALK_X2 | pkg.abi.FLUID_SOLVE_WALK_X4)
    with pytest.raises(pkg.abi.FluidError):
        sim.set_fusion(pkg.abi.FLUID_SOLVE_FIRST_X2 | pkg.abi.FLUID_SOLVE_FIRST_X4)
    flags = pkg.abi.FLUID_SOLVE_WALK_X4 | pkg.abi.FLUID_FUSE_SOLVE_LIST
    sim.set_fusion(flags)
    assert sim.fusion() == flags
    sim.close()
