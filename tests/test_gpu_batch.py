"""Batched form: many independent scenes in one handle (fluid_batch_create)
against the same scenes stepped one by one by plain handles and by the CPU
oracle.  Grid data and permutations are exact per scene, the fields are the
plain handle's bit for bit, and against the float32 and float64 oracles
they meet the plain step's parity bar (check_fp)."""
import numpy as np
import pytest

from test_gpu_parity import Lockstep, check_fp

pytestmark = pytest.mark.gpu


def small_batch(pkg, count=6, seed=11):
    return pkg.scenes.batch_scenes(count, seed=seed, n_min=1000, n_max=6000)


@pytest.mark.parametrize("flags", [0, 256])
def test_batch_matches_plain_and_oracle(pkg, orc, flags):
    """flags = FLUID_LIST_SKIP (256) on both sides."""
    scenes = small_batch(pkg)
    print(f"case: batch of {len(scenes)} scenes, n={[s.n for s in scenes]}, flags={flags}",
          flush=True)
    batch = pkg.make_batch(scenes)
    assert batch.num_scenes() == len(scenes)
    if flags:
        batch.set_fusion(batch.fusion() | flags)
    batch.step()
    counts, starts = batch.read_cell_counts(), batch.read_cell_starts()
    s2o = batch.read_sorted_to_orig()
    parts = batch.read_particles()
    for k, sc in enumerate(scenes):
        info = batch.scene_info(k)
        sl = slice(info["first"], info["first"] + info["count"])
        cl = slice(info["first_cell"], info["first_cell"] + info["cells"])
        plain = pkg.make_sim(sc)
        if flags:
            plain.set_fusion(plain.fusion() | flags)
        plain.step()
        np.testing.assert_array_equal(counts[cl], plain.read_cell_counts())
        np.testing.assert_array_equal(starts[cl] - info["first"], plain.read_cell_starts())
        np.testing.assert_array_equal(s2o[sl] - info["first"], plain.read_sorted_to_orig())
        pp = plain.read_particles()
        # the same kernels on the same data: the batched step is the plain
        # step bit for bit
        np.testing.assert_array_equal(parts[sl], pp)
        params, cont, g = pkg.scenes.make_params(sc, pkg.abi, pkg.derive_sph_params,
                                                 pkg.derive_container_geometry)
        o = Lockstep(orc, pkg, sc, params, cont, g)
        o.step()
        np.testing.assert_array_equal(counts[cl], o.cell_counts())
        op, o6 = o.particles(), o.o64.particles()
        for name, cols in (("position", slice(0, 3)), ("velocity", slice(4, 7)),
                           ("force", slice(8, 11))):
            check_fp(f"scene {k} {name}", parts[sl, cols], op[:, cols], o6[:, cols])
        plain.close()
    batch.close()


def test_batch_rollout_invariants(pkg):
    scenes = pkg.scenes.batch_scenes(32, seed=12, n_min=2048, n_max=16384)
    print(f"case: 32-scene rollout, total n={sum(s.n for s in scenes)}", flush=True)
    batch = pkg.make_batch(scenes)
    cfg = batch.autotune(1)
    print(f"  autotuned {cfg}", flush=True)
    batch.step(20)
    counts = batch.read_cell_counts().astype(np.int64)
    parts = batch.read_particles()
    assert np.isfinite(parts).all()
    for k, sc in enumerate(scenes):
        info = batch.scene_info(k)
        assert counts[info["first_cell"]:info["first_cell"] + info["cells"]].sum() == sc.n
        p = parts[info["first"]:info["first"] + info["count"], 0:3]
        half = np.array(sc.box) / 2 + 1e-4
        # integrate's backstop keeps every particle inside its own box
        assert np.all(np.abs(p[:, [0, 2]]) <= half[[0, 2]])
        assert np.all(np.abs(p[:, 1]) <= half[1])
    with pytest.raises(pkg.abi.FluidError):
        batch.spawn_particles([0, 0, 0], 4, 0.1)
    batch.close()


def test_batch_large_grid_scan(pkg):
    """A batch whose concatenated grid has over 2^20 cells (the scan's large
    tiles): inclusive starts equal the cumulative counts over the whole
    concatenation, and scenes stepped by plain handles (small grids, the
    small tiles) give the same cell counts and particles."""
    scenes = pkg.scenes.batch_scenes(64, seed=21, n_min=2048, n_max=4096)
    batch = pkg.make_batch(scenes)
    cells = sum(batch.scene_info(k)["cells"] for k in range(len(scenes)))
    print(f"case: 64-scene batch, {cells} cells", flush=True)
    assert cells >= 1 << 20
    batch.step(2)
    counts = batch.read_cell_counts().astype(np.int64)
    starts = batch.read_cell_starts().astype(np.int64)
    np.testing.assert_array_equal(starts, np.cumsum(counts))
    for k in (0, 31, 63):
        info = batch.scene_info(k)
        plain = pkg.make_sim(scenes[k])
        plain.step(2)
        cl = slice(info["first_cell"], info["first_cell"] + info["cells"])
        np.testing.assert_array_equal(counts[cl], plain.read_cell_counts())
        got = batch.read_particles()[info["first"]:info["first"] + info["count"]]
        np.testing.assert_array_equal(got, plain.read_particles())
        plain.close()
    batch.close()
