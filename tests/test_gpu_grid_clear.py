"""The grid clear of a full step is left out when the previous step's scan
and integrate passes have already cleared what it clears (the build's
counters and the scan's tile states).  Graph steps, stage-by-stage steps,
per-stage timing, kernel re-runs and the autotuner interleaved on one handle
must give the bits of a handle that only ever ran graph steps; and the cell
counts read between a build and its scan are the build's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def grid(sim):
    return sim.read_cell_counts(), sim.read_cell_starts()


def test_interleaved_steps_match_graph_steps(pkg):
    sc = pkg.scenes.dam_break_8k()
    a = pkg.make_sim(sc)
    b = pkg.make_sim(sc)
    print("case: interleaved graph / stage / timing work vs graph steps, dam_break_8k", flush=True)
    a.step(2)
    # one step stage by stage, with the counts read between build and scan
    a.step_stages("clear", "build")
    built = a.read_cell_counts()
    a.step_stages("scan", "integrate")
    counts, starts = grid(a)
    np.testing.assert_array_equal(built, counts)
    assert counts.sum() == sc.n and starts[-1] == sc.n
    a.step(1)
    a.time_stages(1)  # one more step, every stage launched alone
    a.time_range("density", "force", 3)  # re-runs only
    a.step(2)
    b.step(7)
    np.testing.assert_array_equal(a.read_particles(), b.read_particles())
    for x, y in zip(grid(a), grid(b)):
        np.testing.assert_array_equal(x, y)
    # a clear on its own zeroes the published counts, as grid_clear does
    a.step_stages("clear", "clear")
    assert not a.read_cell_counts().any()
    a.step_stages("build", "integrate")
    b.step(1)
    np.testing.assert_array_equal(a.read_particles(), b.read_particles())
    a.close()
    b.close()


def test_autotune_then_steps(pkg):
    sc = pkg.scenes.dam_break_8k()
    a = pkg.make_sim(sc)
    b = pkg.make_sim(sc)
    print("case: autotune between graph steps, dam_break_8k", flush=True)
    a.step(2)
    b.step(2)
    cfg = a.autotune(1)
    b.set_sweep_config(*cfg)
    b.set_tile_config(*a.tile_config())
    b.set_fusion(a.fusion())
    a.step(3)
    b.step(3)
    np.testing.assert_array_equal(a.read_particles(), b.read_particles())
    for x, y in zip(grid(a), grid(b)):
        np.testing.assert_array_equal(x, y)
    a.close()
    b.close()
