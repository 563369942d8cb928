"""fluid_time_range: re-running stages density .. predict of the last step
(the bench's per-kernel timing) leaves the handle as the step left it, and
the next step gives the bits of a handle that never re-ran anything."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_time_range_leaves_state(pkg):
    sc = pkg.scenes.dam_break_8k()
    sim = pkg.make_sim(sc)
    ref = pkg.make_sim(sc)
    print("case: time_range on dam_break_8k", flush=True)
    with pytest.raises(Exception):
        sim.time_range("density", "density", 1)  # no step yet
    sim.step(3)
    ref.step(3)
    before = sim.read_particles()
    pred = sim.read_predicted(0)
    srt = sim.read_sorted_particles()
    for first, last in (("density", "density"), ("xsph", "force"), ("density", "predict")):
        ms = sim.time_range(first, last, 5)
        print(f"  {first}..{last}: {ms:.4f} ms per repetition", flush=True)
        assert ms > 0.0
    for first, last in (("solve", "solve"), ("build", "density"), ("force", "xsph")):
        with pytest.raises(Exception):
            sim.time_range(first, last, 1)
    np.testing.assert_array_equal(sim.read_particles(), before)
    np.testing.assert_array_equal(sim.read_predicted(0), pred)
    np.testing.assert_array_equal(sim.read_sorted_particles(), srt)
    sim.step(2)
    ref.step(2)
    np.testing.assert_array_equal(sim.read_particles(), ref.read_particles())
    sim.close()
    ref.close()
