"""Whitewater passes on the GPU (fluid_spray_*) against the oracle, teacher-
forced on the simulation's own sorted state, grid and spray ring.

Bars:
  - neighbour counts and spawn slots/integer fields: exact;
  - trapped-air, crest and divergence sums (summed over 8 lanes on the GPU,
    in stencil order in the oracle): 1e-4 relative to the particle's value
    with a floor of 1e-3 of the field's RMS, for all but at most 0.1% of
    the emitters (a pair whose convexity test dot(-x_hat, n) < 0 sits at 0
    can flip, adding or dropping one neighbour's crest term);
  - spawn counts: equal for all but at most 0.1% of the emitters (a
    fractional draw within rounding of its threshold);
  - spawned and simulated records: lifetimes and ages 1e-6, spawn
    positions 2e-6 (the spawn disc's cos/sin); foam and bubble velocities,
    weighted means of up to 64 neighbour velocities summed in another order,
    within 2 * 64 * 2^-24 of the fastest fluid speed, and positions within
    dt times that more; kinds equal for all but 0.1% of the live particles.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spray_params(pkg, sim, cap, frame=1, rate=300.0):
    p = pkg.abi.default_spray_params(dt=1.0 / 60.0, max_particles=cap,
                                     num_sph_particles=sim.num_particles(), frame_count=frame)
    p.emission_rate = rate
    return p


def _splashing_sim(pkg, steps=70):
    sim = pkg.make_sim(pkg.scenes.dam_break_8k())
    sim.step(steps)
    return sim


def _sim_state(sim):
    return (sim.read_sorted_particles(), sim.read_cell_counts(), sim.read_cell_starts(),
            sim.grid_params(), sim.sph_params(), sim.container_geometry())


def _close_frac(a, b, rtol, floor_frac):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    floor = floor_frac * np.sqrt(np.mean(b ** 2)) if b.size else 0.0
    return np.abs(a - b) <= rtol * np.maximum(np.abs(b), max(floor, 1e-30))


def _check_emit(pot, ne, pot_o, ne_o, tag):
    active = pot_o[:, 3] > 0
    assert np.array_equal(pot[:, 3], pot_o[:, 3]), tag  # neighbour counts
    n_act = max(int(active.sum()), 1)
    for c, name in ((0, "trapped air"), (1, "crest"), (2, "divergence")):
        ok = _close_frac(pot[active, c], pot_o[active, c], 1e-4, 1e-3)
        bad = int((~ok).sum())
        print(f"  {tag} {name}: {bad} of {n_act} outside tolerance", flush=True)
        assert bad <= max(1, n_act // 1000), (tag, name, bad)
    diff = int((ne != ne_o).sum())
    print(f"  {tag} spawns: gpu {int(ne.sum())} oracle {int(ne_o.sum())}, "
          f"{diff} particles differ", flush=True)
    assert diff <= max(1, n_act // 1000), (tag, diff)


def _check_ring(ring, ring_o, tag, recs=None, dt=0.0):
    """Integer words and kinds exactly (up to 0.1% of kinds); float fields
    within the rounding bounds stated in the module docstring, in absolute
    terms.  The velocity bound for foam and bubbles is 2 * 64 * 2^-24 times
    the fastest fluid particle (a weighted mean of up to 64 neighbour
    velocities summed in another order); positions add dt times that."""
    f, fo = ring.view(np.float32), ring_o.view(np.float32)
    live = fo[:, 3] > 0
    kinds_differ = ring[:, 8] != ring_o[:, 8]
    nk = int(kinds_differ.sum())
    print(f"  {tag}: {int(live.sum())} live, {nk} kinds differ", flush=True)
    assert nk <= int(live.sum()) // 1000, tag
    same = ~kinds_differ
    vmax = float(np.linalg.norm(recs[:, 4:7], axis=1).max()) if recs is not None else 0.0
    vtol = 2 * 64 * 2.0 ** -24 * max(vmax, 1.0)
    for cols, name, tol in ((slice(0, 3), "position", 2e-6 + dt * vtol),
                            (slice(3, 4), "lifetime", 1e-6), (slice(4, 7), "velocity", vtol),
                            (slice(7, 8), "max lifetime", 1e-6), (slice(9, 10), "age", 1e-6)):
        a, b = f[same, cols].astype(np.float64), fo[same, cols].astype(np.float64)
        err = np.abs(a - b).max(axis=1) if a.size else np.zeros(1)
        k = int(np.argmax(err))
        kind = int(ring[same][k, 8]) if a.size else -1
        print(f"    {name}: max abs {err[k]:.2e} (bound {tol:.2e}, kind {kind}, "
              f"|value| {np.abs(b[k]).max() if a.size else 0:.3f}, vmax {vmax:.3f})", flush=True)
        assert err[k] <= tol, (tag, name, float(err[k]), tol)


def test_spray_emit_and_spawn_match_oracle(pkg, orc):
    sim = _splashing_sim(pkg)
    n = sim.num_particles()
    cap = 8192
    prm = _spray_params(pkg, sim, cap)
    spray = pkg.SpraySystem(sim, cap, prm)
    recs, cnt, st, grid, sph, cont = _sim_state(sim)
    s0 = spray.state()
    print(f"case: spray emit n={n} cap={cap} head={s0['write_head']}", flush=True)
    spray.emit(n)
    pot, ne = spray.read_emit(n)
    pot_o, ne_o, fm_o = orc.spray_potentials(recs, cnt, st, grid, sph, prm,
                                             s0["ta_limit"], s0["wc_limit"])
    assert (pot_o[:, 3] >= 10).sum() > 100 and ne_o.sum() > 50  # the flow splashes
    _check_emit(pot, ne, pot_o, ne_o, "emit")
    # spawns teacher-forced on the GPU's counts: same slots, same records
    ring_o = np.zeros((cap, 12), np.uint32)
    head_o = orc.spray_spawn(recs, sph, prm, ne, s0["write_head"], ring_o)
    ring = spray.read_particles()
    assert spray.state()["write_head"] == head_o == int(ne.sum())
    assert np.array_equal(ring[:, 8], ring_o[:, 8])
    _check_ring(ring, ring_o, "spawn", recs)
    # the simulate pass folds this frame's maxima into the ceilings
    spray.simulate()
    s1 = spray.state()
    want_ta = orc.spray_update_limit(s0["ta_limit"], float(fm_o[0]), 0.2)
    want_wc = orc.spray_update_limit(s0["wc_limit"], float(fm_o[1]), 0.05)
    assert abs(s1["ta_limit"] - want_ta) <= 1e-5 * want_ta
    assert abs(s1["wc_limit"] - want_wc) <= 1e-5 * want_wc
    spray.close()
    sim.close()


def test_spray_frames_match_oracle(pkg, orc):
    """The app's loop (app.rs:1209-1238): two SPH substeps, then one spray
    step, for several frames; each frame is checked from the GPU's own state
    before it (sorted particles, grid, ring, head, ceilings)."""
    sim = _splashing_sim(pkg, steps=60)
    n = sim.num_particles()
    cap = 16384
    prm = _spray_params(pkg, sim, cap)
    spray = pkg.SpraySystem(sim, cap, prm)
    for frame in range(1, 9):
        sim.step(2)
        prm = _spray_params(pkg, sim, cap, frame=frame)
        spray.update_params(prm)
        recs, cnt, st, grid, sph, cont = _sim_state(sim)
        s0 = spray.state()
        ring0 = spray.read_particles()
        print(f"case: spray frame {frame} head={s0['write_head']} "
              f"ceilings=({s0['ta_limit']:.4f}, {s0['wc_limit']:.4f})", flush=True)
        spray.step(n)
        pot, ne = spray.read_emit(n)
        pot_o, ne_o, fm_o = orc.spray_potentials(recs, cnt, st, grid, sph, prm,
                                                 s0["ta_limit"], s0["wc_limit"])
        _check_emit(pot, ne, pot_o, ne_o, f"frame {frame}")
        ring_o = ring0.copy()
        head_o = orc.spray_spawn(recs, sph, prm, ne, s0["write_head"], ring_o)
        live_o = orc.spray_simulate(ring_o, recs, cnt, st, grid, sph, cont, prm)
        ring = spray.read_particles()
        s1 = spray.state()
        assert s1["write_head"] == head_o
        _check_ring(ring, ring_o, f"frame {frame}", recs, prm.dt)
        stats = spray.read_stats()
        assert abs(int(stats[0]) - int(live_o[0])) <= max(1, int(live_o[0]) // 1000)
        f = ring.view(np.float32)
        alive = f[:, 3] > 0
        assert stats[0] == alive.sum()
        for k in range(3):
            assert stats[1 + k] == (alive & (ring[:, 8] == k)).sum()
        want_ta = orc.spray_update_limit(s0["ta_limit"], float(fm_o[0]), 0.2)
        assert abs(s1["ta_limit"] - want_ta) <= 1e-5 * want_ta
    assert stats[0] > 0 and (stats[2] + stats[3]) > 0  # foam or bubbles formed
    spray.close()
    sim.close()


def test_spray_ring_wrap_keeps_last_spawns(pkg, orc):
    """A frame that spawns more than the ring holds keeps the last `cap`
    spawns, in the slots a sequential run of the reference's atomic counter
    would leave them in."""
    sim = _splashing_sim(pkg)
    n = sim.num_particles()
    cap = 64
    prm = _spray_params(pkg, sim, cap, rate=3000.0)
    spray = pkg.SpraySystem(sim, cap, prm)
    recs, cnt, st, grid, sph, cont = _sim_state(sim)
    spray.emit(n)
    _, ne = spray.read_emit(n)
    total = int(ne.sum())
    assert total > 3 * cap
    ring_o = np.zeros((cap, 12), np.uint32)
    head_o = orc.spray_spawn(recs, sph, prm, ne, 0, ring_o)
    ring = spray.read_particles()
    assert spray.state()["write_head"] == head_o == total
    _check_ring(ring, ring_o, "wrap", recs)
    spray.close()
    sim.close()


def test_spray_reset_guards_and_no_step(pkg):
    """reset zeroes the ring, head and ceilings; a simulation that has not
    stepped has an empty grid (nothing emits, everything is ballistic); the
    ring size and params must agree; a batched handle takes no spray; the
    simulation refuses destruction while a spray handle is bound."""
    import ctypes as C
    sim = pkg.make_sim(pkg.scenes.dam_break_8k())
    n = sim.num_particles()
    cap = 256
    prm = _spray_params(pkg, sim, cap)
    spray = pkg.SpraySystem(sim, cap, prm)
    spray.step(n)
    _, ne = spray.read_emit(n)
    assert ne.sum() == 0 and spray.read_stats()[0] == 0
    ring = np.zeros((cap, 12), np.uint32)
    f = ring.view(np.float32)
    f[:8, 1] = 0.6
    f[:8, 3] = 1.0
    f[:8, 4] = 0.5
    spray.write_particles(ring)
    spray.simulate()
    out = spray.read_particles()
    assert np.all(out[:8, 8] == 0) and spray.read_stats()[1] == 8
    spray.reset()
    st = spray.state()
    assert not spray.read_particles().any()
    assert st["write_head"] == 0 and st["ta_limit"] == 3.0 and st["wc_limit"] == 1.0
    bad = _spray_params(pkg, sim, cap + 1)
    with pytest.raises(pkg.abi.FluidError):
        spray.update_params(bad)
    lib = pkg.abi.load_library()
    assert lib.fluid_destroy(sim._h) == pkg.abi.FLUID_ERR_STATE
    spray.close()
    batch = pkg.make_batch(pkg.scenes.batch_scenes(2, n_min=256, n_max=512))
    h = C.c_void_p()
    assert lib.fluid_spray_create(batch._h, cap, C.byref(prm), C.byref(h)) == \
        pkg.abi.FLUID_ERR_INVALID_ARG
    batch.close()
    sim.close()


def test_spray_full_size_50k(pkg, orc):
    """The 50k dam break after 60 steps (the bench workload), default
    SprayConfig with a 100k ring: one frame against the oracle, teacher-
    forced as above, at full size."""
    sim = pkg.make_sim(pkg.scenes.dam_break_50k())
    sim.step(60)
    n = sim.num_particles()
    cap = 100_000
    prm = pkg.abi.default_spray_params(dt=2 * 0.008, max_particles=cap, num_sph_particles=n,
                                       frame_count=3)
    spray = pkg.SpraySystem(sim, cap, prm)
    recs, cnt, st, grid, sph, cont = _sim_state(sim)
    s0 = spray.state()
    ring0 = spray.read_particles()
    print(f"case: spray full size n={n} cap={cap}", flush=True)
    spray.step(n)
    pot, ne = spray.read_emit(n)
    pot_o, ne_o, fm_o = orc.spray_potentials(recs, cnt, st, grid, sph, prm,
                                             s0["ta_limit"], s0["wc_limit"])
    assert ne_o.sum() > 100
    _check_emit(pot, ne, pot_o, ne_o, "50k")
    ring_o = ring0.copy()
    head_o = orc.spray_spawn(recs, sph, prm, ne, s0["write_head"], ring_o)
    orc.spray_simulate(ring_o, recs, cnt, st, grid, sph, cont, prm)
    assert spray.state()["write_head"] == head_o
    _check_ring(spray.read_particles(), ring_o, "50k", recs, prm.dt)
    spray.close()
    sim.close()
