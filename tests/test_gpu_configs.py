"""GPU parity of the configurations the benchmark times, against the oracle.

  * the 50,000-particle dam break after the bench's warm-up, with the launch
    configuration the autotuner picks, and with each of the sweep variants it
    chooses between set explicitly (16-lane density, the 16- and 32-lane list
    walks, the wide first PCISPH iteration, the multi-walker later
    iterations, the XCD-contiguous tile mapping, the list skip for
    candidate-heavy particles);
  * a tilted container (tilt 0.3 about x, -0.2 about z) with the block
    against its walls, stage by stage and as graph steps, so the device-side
    frame changes of density's wall factor, integrate's wall layer and the
    backstop are compared with the oracle;
  * the batched form at its full size: 256 scenes with particle counts
    log-uniform in [2,048, 65,536], autotuned, per-scene grid invariants on
    every scene and the oracle bar on a sample of scenes.

Every comparison uses the plain step's bar (check_fp in test_gpu_parity.py):
grid data and permutations exact, fields within max(1e-4, 2x the float32
oracle's own distance from float64) at the 99.9th percentile and at the
maximum.
"""
import numpy as np
import pytest

from test_gpu_parity import Lockstep, check_fp, compare_grid, make_pair, one_step_compare

pytestmark = pytest.mark.gpu

FUSE = 1 | 2  # FLUID_FUSE_XSPH_FORCE | FLUID_FUSE_SOLVE_LIST
WALK_16, WALK_32 = 4, 8
SOLVE_WALK_X2, SOLVE_WALK_X4 = 16, 32
SOLVE_FIRST_X2, SOLVE_FIRST_X4 = 64, 128
LIST_SKIP = 256


def warmed_50k(pkg, orc, warm=60):
    """The 50k dam break after `warm` graph steps, with an oracle pair
    started from the GPU's state."""
    scene = pkg.scenes.dam_break_50k()
    sim = pkg.make_sim(scene)
    sim.step(warm)
    state = sim.read_particles()
    params, cont, g = pkg.scenes.make_params(scene, pkg.abi, pkg.derive_sph_params,
                                             pkg.derive_container_geometry)
    o = Lockstep.__new__(Lockstep)
    o.o32 = orc.from_scene(scene, pkg.abi, params, cont, g, particles=state)
    o.o64 = orc.from_scene(scene, pkg.abi, params, cont, g, fp64=True, particles=state)
    return sim, o


def teacher_forced(sim, o, steps, tag):
    for k in range(steps):
        print(f"case: {tag} teacher-forced step {k}", flush=True)
        sim.write_particles(o.particles())
        o.sync64()
        one_step_compare(sim, o, f"{tag} step{k}")
        gs, os_, o6 = sim.read_sorted_particles(), o.sorted_particles(), \
            o.sorted64()
        check_fp(f"{tag} step{k} density", gs[:, 11], os_[:, 11], o6[:, 11])
        check_fp(f"{tag} step{k} near density", gs[:, 12], os_[:, 12], o6[:, 12])


def test_dam_break_50k_autotuned(pkg, orc):
    """The headline workload in the configuration the bench times: warm-up,
    autotune (which restores the state), then graph steps."""
    sim, o = warmed_50k(pkg, orc)
    cfg = sim.autotune(2)
    print(f"  autotuned sweep config {cfg}, tiles {sim.tile_config()}, "
          f"fusion {sim.fusion():#x}", flush=True)
    np.testing.assert_array_equal(sim.read_particles(), o.particles())  # not advanced
    teacher_forced(sim, o, 3, "50k autotuned")
    sim.close()


CONFIGS = {
    # (density lanes, solve lanes, XCD-contiguous tiles, fusion flags)
    "d16_walk32_first2_xcd": (16, 8, 1, FUSE | WALK_32 | SOLVE_FIRST_X2),
    "d16_walk32_solve16": (16, 16, 1, FUSE | WALK_32),
    "d8_walk16_s4": (8, 4, 0, FUSE | WALK_16),
    "d8_s4_walkx2": (8, 4, 0, FUSE | SOLVE_WALK_X2),
    "d4_s2_walkx4_first4": (4, 2, 0, FUSE | SOLVE_WALK_X4 | SOLVE_FIRST_X4),
    "d16_unfused": (16, 8, 0, 0),
    "d4_s4_skip": (4, 4, 0, FUSE | LIST_SKIP),
}


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_dam_break_50k_sweep_variants(pkg, orc, name):
    dl, sl, xcd, flags = CONFIGS[name]
    sim, o = warmed_50k(pkg, orc)
    sim.set_tile_config(0)
    sim.set_sweep_config(dl, sl, xcd)
    sim.set_fusion(flags)
    assert sim.sweep_config() == (dl, sl, xcd) and sim.fusion() == flags
    teacher_forced(sim, o, 2, f"50k {name}")
    sim.close()


# ---------------------------------------------------------------- tilted
def tilted_scene(pkg):
    return pkg.scenes.tilted(pkg.scenes.dam_break_8k(), 0.3, -0.2)


def test_tilted_container_stagewise(pkg, orc):
    scene = tilted_scene(pkg)
    sim, o = make_pair(pkg, orc, scene)
    g = sim.container_geometry()
    assert abs(g.forward_row2[1] - np.sin(0.3)) < 1e-6  # the handle's container is tilted
    S = pkg.abi.STAGE
    print("case: tilted container, stage by stage", flush=True)
    sim.step_stages(0, S["reorder"])
    o.stages(0, S["reorder"])
    compare_grid(sim, o)
    np.testing.assert_array_equal(sim.read_sorted_to_orig(), o.sorted_to_orig())
    sim.step_stages(S["density"], S["density"])
    o.stage(S["density"])
    gs, os_, o6 = sim.read_sorted_particles(), o.sorted_particles(), o.sorted64()
    # the block sits in the wall layers of three faces: the wall factor is active
    assert np.sum(os_[:, 11] > 1.2 * np.median(os_[:, 11])) > 0
    check_fp("tilted density", gs[:, 11], os_[:, 11], o6[:, 11])
    check_fp("tilted normal", gs[:, 13:16], os_[:, 13:16], o6[:, 13:16])
    sim.step_stages(S["xsph"], S["finalize"])
    o.stages(S["xsph"], S["finalize"])
    gp, op, o6 = sim.read_particles(), o.particles(), o.o64.particles()
    check_fp("tilted finalize velocity", gp[:, 4:7], op[:, 4:7], o6[:, 4:7])
    sim.step_stages(S["integrate"], S["integrate"])
    o.stage(S["integrate"])
    gp, op, o6 = sim.read_particles(), o.particles(), o.o64.particles()
    check_fp("tilted integrate position", gp[:, 0:3], op[:, 0:3], o6[:, 0:3])
    check_fp("tilted integrate velocity", gp[:, 4:7], op[:, 4:7], o6[:, 4:7])
    sim.close()


def test_tilted_container_graph_steps(pkg, orc):
    scene = tilted_scene(pkg)
    sim, o = make_pair(pkg, orc, scene)
    for k in range(4):
        print(f"case: tilted container graph step {k}", flush=True)
        sim.write_particles(o.particles())
        o.sync64()
        one_step_compare(sim, o, f"tilted step{k}")
    # after a few steps particles have reached the backstop on the tilted
    # walls: every one lies inside the container in its local frame
    g = sim.container_geometry()
    inv = np.array([g.inverse_row0[:3], g.inverse_row1[:3], g.inverse_row2[:3]], np.float64)
    loc = (sim.read_particles()[:, 0:3] - [0, g.center_y, 0]) @ inv.T
    half = np.array([g.half_width, g.half_height, g.half_depth])
    assert np.all(np.abs(loc) <= half * (1 + 1e-6))
    sim.close()


# ---------------------------------------------------------------- batch 256
def test_batch_256_full_size(pkg, orc):
    scenes = pkg.scenes.batch_scenes(256)
    n = [sc.n for sc in scenes]
    print(f"case: 256-scene batch, {sum(n)} particles, n in [{min(n)}, {max(n)}]", flush=True)
    assert min(n) >= 2048 and max(n) <= 65536
    batch = pkg.make_batch(scenes)
    cfg = batch.autotune(1)
    print(f"  autotuned {cfg}", flush=True)
    batch.step()
    counts = batch.read_cell_counts().astype(np.int64)
    starts = batch.read_cell_starts().astype(np.int64)
    s2o = batch.read_sorted_to_orig().astype(np.int64)
    parts = batch.read_particles()
    np.testing.assert_array_equal(starts, np.cumsum(counts))
    for k, sc in enumerate(scenes):
        info = batch.scene_info(k)
        sl = slice(info["first"], info["first"] + info["count"])
        cl = slice(info["first_cell"], info["first_cell"] + info["cells"])
        assert info["count"] == sc.n
        assert counts[cl].sum() == sc.n, k
        assert np.array_equal(np.sort(s2o[sl]), np.arange(info["first"], sl.stop)), k
    assert np.isfinite(parts).all()
    order = np.argsort(n)
    rng = np.random.default_rng(6)
    sample = sorted({int(order[0]), int(order[-1])} |
                    set(int(x) for x in rng.choice(256, 4, replace=False)))
    for k in sample:
        sc = scenes[k]
        info = batch.scene_info(k)
        sl = slice(info["first"], info["first"] + info["count"])
        cl = slice(info["first_cell"], info["first_cell"] + info["cells"])
        print(f"  scene {k}: n={sc.n}", flush=True)
        params, cont, g = pkg.scenes.make_params(sc, pkg.abi, pkg.derive_sph_params,
                                                 pkg.derive_container_geometry)
        o = Lockstep(orc, pkg, sc, params, cont, g)
        o.step()
        np.testing.assert_array_equal(counts[cl], o.cell_counts())
        np.testing.assert_array_equal(s2o[sl] - info["first"], o.sorted_to_orig())
        op, o6 = o.particles(), o.o64.particles()
        check_fp(f"scene {k} position", parts[sl, 0:3], op[:, 0:3], o6[:, 0:3])
        check_fp(f"scene {k} velocity", parts[sl, 4:7], op[:, 4:7], o6[:, 4:7])
        check_fp(f"scene {k} force", parts[sl, 8:11], op[:, 8:11], o6[:, 8:11])
    batch.step(19)
    parts = batch.read_particles()
    assert np.isfinite(parts).all()
    for k, sc in enumerate(scenes):
        info = batch.scene_info(k)
        p = parts[info["first"]:info["first"] + info["count"], 0:3]
        half = np.array(sc.box) / 2 * (1 + 1e-6)
        assert np.all(np.abs(p[:, [0, 2]]) <= half[[0, 2]]), k
        assert np.all(np.abs(p[:, 1]) <= half[1]), k
    batch.close()
