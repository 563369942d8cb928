"""GPU parity: the HIP step against the CPU oracle on the same seeded inputs.

Integer work (cell counts, inclusive cell starts, the sorted->original
permutation, the sorted copy of every particle) must be bit-exact.

Floating-point fields are compared per particle, normwise, with a floor of 1%
of the field RMS, against the float32 oracle and against the float64 oracle
run in lockstep on the same state.  The bound is 1e-4 relative, raised to
twice the float32 oracle's own distance from float64 on that field where
summation-order noise alone exceeds 1e-4 (heavily cancelling sums, e.g. the
PCISPH pressure acceleration): the HIP result must be as close to the exact
(float64) value as the reference's own float32 arithmetic is.
"""
import numpy as np
import pytest

from conftest import rowwise_rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-4


class Lockstep:
    """float32 and float64 oracles advanced together on the same state."""

    def __init__(self, orc, pkg, scene, params, cont, g):
        self.o32 = orc.from_scene(scene, pkg.abi, params, cont, g)
        self.o64 = orc.from_scene(scene, pkg.abi, params, cont, g, fp64=True)

    def stage(self, s):
        self.o32.stage(s)
        self.o64.stage(s)

    def stages(self, a, b):
        for s in range(a, b + 1):
            self.stage(s)

    def step(self, k=1):
        self.o32.step(k)
        self.o64.step(k)

    def sync64(self):
        """Restart the float64 oracle from the float32 oracle's state."""
        self.o64.set_particles(self.o32.particles())

    def _to32(self, a64):
        """Rows of a float64-oracle array in sorted-slot order, put in the
        float32 oracle's slot order: the within-cell order follows position
        x, which the two oracles can round apart."""
        o2s = np.empty(self.o64.n, dtype=np.int64)
        o2s[self.o64.sorted_to_orig()] = np.arange(self.o64.n)
        return a64[o2s[self.o32.sorted_to_orig()]]

    def sorted64(self):
        return self._to32(self.o64.sorted_particles())

    def predicted64(self, which):
        return self._to32(self.o64.predicted(which))

    def __getattr__(self, name):
        return getattr(self.o32, name)


def make_pair(pkg, orc, scene, iterations=None, params_edit=None):
    if iterations is not None:
        scene.pcisph_iterations = iterations
    params, cont, g = pkg.scenes.make_params(scene, pkg.abi, pkg.derive_sph_params,
                                             pkg.derive_container_geometry)
    if params_edit:
        params_edit(params)
    sim = pkg.make_sim(scene)
    sim.update_sph_params(params)
    return sim, Lockstep(orc, pkg, scene, params, cont, g)


def report(name, err):
    print(f"  {name}: max {err.max():.3e} p99.9 {np.quantile(err, 0.999):.3e} "
          f"median {np.median(err):.3e}", flush=True)


def check_fp(name, got, want, want64=None, tol=TOL):
    """The HIP error against the float32 oracle, at the 99.9th percentile
    and at the maximum, within max(tol, 2 x the float32 oracle's own error
    against float64 at that same statistic); the same maximum bound against
    float64."""
    err = rowwise_rel_err(got, want)
    report(name, err)
    bound_q = bound_max = tol
    if want64 is not None:
        noise = rowwise_rel_err(want, want64)
        bound_q = max(tol, 2.0 * np.quantile(noise, 0.999))
        bound_max = max(tol, 2.0 * noise.max())
        e64 = rowwise_rel_err(got, want64).max()
        print(f"    f32-oracle noise p99.9 {np.quantile(noise, 0.999):.3e} max {noise.max():.3e}"
              f"  gpu-vs-f64 max {e64:.3e}", flush=True)
        assert e64 <= bound_max, f"{name}: gpu vs f64 {e64:.3e} > {bound_max:.3e}"
    q = np.quantile(err, 0.999)
    assert q <= bound_q, f"{name}: p99.9 rel err {q:.3e} > {bound_q:.3e}"
    assert err.max() <= bound_max, f"{name}: max rel err {err.max():.3e} > {bound_max:.3e}"


def compare_grid(sim, o):
    np.testing.assert_array_equal(sim.read_cell_counts(), o.cell_counts())
    np.testing.assert_array_equal(sim.read_cell_starts(), o.cell_starts())


def wide_walk(pkg, sim, flag):
    """XSPH and force walk the 8-lane lists with 16 or 32 lanes per particle
    (FLUID_WALK_16 = 4, FLUID_WALK_32 = 8), and/or candidate-heavy particles
    record no list (FLUID_LIST_SKIP = 256)."""
    sim.set_sweep_config(8, 8, 0)
    sim.set_fusion(sim.fusion() | flag)


@pytest.mark.parametrize("wide", [0, 4, 8])
@pytest.mark.parametrize("scene_name", ["dam_break_8k"])
def test_stagewise_parity(pkg, orc, scene_name, wide):
    scene = pkg.scenes.SCENES[scene_name]()
    print(f"case: stagewise {scene_name} n={scene.n} walk flag={wide}", flush=True)
    sim, o = make_pair(pkg, orc, scene)
    if wide:
        wide_walk(pkg, sim, wide)
    S = pkg.abi.STAGE
    sim.step_stages(0, S["scan"]); o.stages(0, S["scan"])
    compare_grid(sim, o)
    sim.step_stages(S["reorder"], S["reorder"]); o.stage(S["reorder"])
    np.testing.assert_array_equal(sim.read_sorted_to_orig(), o.sorted_to_orig())
    np.testing.assert_array_equal(sim.read_orig_to_sorted(), o.particle_cell_indices())
    gs, os_ = sim.read_sorted_particles(), o.sorted_particles()
    np.testing.assert_array_equal(gs[:, 0:11], os_[:, 0:11])
    sim.step_stages(S["density"], S["density"]); o.stage(S["density"])
    gs, os_, o64s = sim.read_sorted_particles(), o.sorted_particles(), o.sorted64()
    check_fp("density", gs[:, 11], os_[:, 11], o64s[:, 11])
    check_fp("near_density", gs[:, 12], os_[:, 12], o64s[:, 12])
    check_fp("normal", gs[:, 13:16], os_[:, 13:16], o64s[:, 13:16])
    sim.step_stages(S["xsph"], S["xsph"]); o.stage(S["xsph"])
    check_fp("xsph velocity", sim.read_particles()[:, 4:7], o.particles()[:, 4:7],
             o.o64.particles()[:, 4:7])
    sim.step_stages(S["force"], S["force"]); o.stage(S["force"])
    check_fp("force", sim.read_particles()[:, 8:11], o.particles()[:, 8:11],
             o.o64.particles()[:, 8:11])
    sim.step_stages(S["predict"], S["predict"]); o.stage(S["predict"])
    check_fp("predicted x*", sim.read_predicted(0)[:, 0:3], o.predicted(0)[:, 0:3],
             o.predicted64(0)[:, 0:3])
    check_fp("predicted v*", sim.read_predicted(0)[:, 4:7], o.predicted(0)[:, 4:7],
             o.predicted64(0)[:, 4:7])
    sim.step_stages(S["solve"], S["solve"]); o.stage(S["solve"])
    for w in (0, 1):
        gp, op, o6 = sim.read_predicted(w), o.predicted(w), o.predicted64(w)
        check_fp(f"solve[{w}] x*", gp[:, 0:3], op[:, 0:3], o6[:, 0:3])
        check_fp(f"solve[{w}] pressure", gp[:, 3], op[:, 3], o6[:, 3])
        check_fp(f"solve[{w}] v*", gp[:, 4:7], op[:, 4:7], o6[:, 4:7])
        check_fp(f"solve[{w}] rho*", gp[:, 7], op[:, 7], o6[:, 7])
    sim.step_stages(S["finalize"], S["finalize"]); o.stage(S["finalize"])
    gp, op, o6 = sim.read_particles(), o.particles(), o.o64.particles()
    check_fp("finalize velocity", gp[:, 4:7], op[:, 4:7], o6[:, 4:7])
    check_fp("finalize force (dv)", gp[:, 8:11], op[:, 8:11], o6[:, 8:11])
    sim.step_stages(S["integrate"], S["integrate"]); o.stage(S["integrate"])
    gp, op, o6 = sim.read_particles(), o.particles(), o.o64.particles()
    check_fp("integrate position", gp[:, 0:3], op[:, 0:3], o6[:, 0:3])
    check_fp("integrate velocity", gp[:, 4:7], op[:, 4:7], o6[:, 4:7])
    sim.close()


@pytest.mark.parametrize("wide", [0, 4, 8])
def test_graph_step_teacher_forced(pkg, orc, wide):
    """Five graph-launched steps, each started from the oracle's state (the
    16- and 32-lane list walks: FLUID_WALK_16 = 4, FLUID_WALK_32 = 8)."""
    scene = pkg.scenes.dam_break_8k()
    sim, o = make_pair(pkg, orc, scene)
    if wide:
        wide_walk(pkg, sim, wide)
    for k in range(5):
        print(f"case: teacher-forced step {k} walk flag={wide}", flush=True)
        sim.write_particles(o.particles())
        o.sync64()
        sim.step()
        o.step()
        compare_grid(sim, o)
        np.testing.assert_array_equal(sim.read_sorted_to_orig(), o.sorted_to_orig())
        gp, op, o6 = sim.read_particles(), o.particles(), o.o64.particles()
        check_fp(f"step{k} position", gp[:, 0:3], op[:, 0:3], o6[:, 0:3])
        check_fp(f"step{k} velocity", gp[:, 4:7], op[:, 4:7], o6[:, 4:7])
        check_fp(f"step{k} force", gp[:, 8:11], op[:, 8:11], o6[:, 8:11])
        gs, os_, o6s = sim.read_sorted_particles(), o.sorted_particles(), o.sorted64()
        check_fp(f"step{k} sorted density", gs[:, 11], os_[:, 11], o6s[:, 11])
    sim.close()


def test_free_running_rollout_graph(pkg, orc):
    """20 free-running graph steps: the HIP trajectory stays as close to the
    float64 oracle as the float32 oracle does (chaotic flow: reported, and
    bounded by 5x the float32 oracle's own divergence)."""
    scene = pkg.scenes.dam_break_8k()
    sim, o = make_pair(pkg, orc, scene)
    sim.step(20)
    o.step(20)
    gp, op, o6 = sim.read_particles(), o.particles(), o.o64.particles()
    d_gpu = np.linalg.norm(gp[:, 0:3] - o6[:, 0:3], axis=1)
    d_32 = np.linalg.norm(op[:, 0:3] - o6[:, 0:3], axis=1)
    print(f"  rollout20 |x_gpu - x_f64| max {d_gpu.max():.3e} median {np.median(d_gpu):.3e}; "
          f"|x_f32 - x_f64| max {d_32.max():.3e} median {np.median(d_32):.3e}", flush=True)
    assert np.median(d_gpu) <= 5 * np.median(d_32) + 1e-6
    assert d_gpu.max() <= 5 * d_32.max() + 1e-5
    sim.close()


def test_rollout_200_divergence(pkg, orc):
    """The 8,192-particle dam break over its full 200 steps, free running:
    the divergence from the float64 oracle is reported next to the float32
    oracle's own, at steps 1 to 200.  The over-compressed start explodes
    (speeds of hundreds of m/s within a few steps), so the flow is chaotic:
    both float32 trajectories leave the float64 one at the same exponential
    rate and are decorrelated from it (median distance ~1 m, the box's
    scale) by step 10.  Per-step parity is the teacher-forced tests' job;
    here the bound is the float32 oracle's divergence times 5, at the
    median and the 99th percentile."""
    scene = pkg.scenes.dam_break_8k()
    sim, o = make_pair(pkg, orc, scene)
    out = []
    done = 0
    for k in (1, 2, 3, 5, 10, 20, 50, 200):
        sim.step(k - done)
        o.step(k - done)
        done = k
        gp, op, o6 = sim.read_particles(), o.particles(), o.o64.particles()
        d_gpu = np.linalg.norm(gp[:, 0:3] - o6[:, 0:3], axis=1)
        d_32 = np.linalg.norm(op[:, 0:3] - o6[:, 0:3], axis=1)
        out.append((k, np.median(d_gpu), np.quantile(d_gpu, 0.99), np.median(d_32),
                    np.quantile(d_32, 0.99)))
        print(f"  step {k}: |x_gpu - x_f64| median {out[-1][1]:.3e} p99 {out[-1][2]:.3e}; "
              f"|x_f32 - x_f64| median {out[-1][3]:.3e} p99 {out[-1][4]:.3e}", flush=True)
        assert np.all(np.isfinite(gp))
    for k, mg, qg, m3, q3 in out:
        assert mg <= 5 * m3 + 1e-6 and qg <= 5 * q3 + 1e-5, k
    sim.close()


def one_step_compare(sim, o, tag, accum=False):
    sim.step()
    o.step()
    compare_grid(sim, o)
    np.testing.assert_array_equal(sim.read_sorted_to_orig(), o.sorted_to_orig())
    gp, op, o6 = sim.read_particles(), o.particles(), o.o64.particles()
    check_fp(f"{tag} position", gp[:, 0:3], op[:, 0:3], o6[:, 0:3])
    check_fp(f"{tag} velocity", gp[:, 4:7], op[:, 4:7], o6[:, 4:7])
    check_fp(f"{tag} force", gp[:, 8:11], op[:, 8:11], o6[:, 8:11])


def test_odd_iterations_and_initial_velocity(pkg, orc):
    """5 PCISPH iterations (finalize reads buffer B) on a block with random
    initial velocities (XSPH and viscosity active from the first step)."""
    scene = pkg.scenes.dam_break_32k()
    print(f"case: 32k, 5 iterations, n={scene.n}", flush=True)
    sim, o = make_pair(pkg, orc, scene, iterations=5)
    for k in range(2):
        sim.write_particles(o.particles())
        o.sync64()
        one_step_compare(sim, o, f"iter5 step{k}")
    sim.close()


def test_out_of_grid_particles_and_no_xsph(pkg, orc):
    scene = pkg.scenes.dam_break_8k()
    parts = scene.particles.copy()
    rng = np.random.default_rng(1)
    # 64 particles outside the fixed grid span (+-2.08): build clamps them into
    # boundary cells, the sweeps see them through the unclamped cell test
    parts[:64, 0:3] = rng.uniform([-2.5, 2.1, -2.5], [2.5, 2.6, 2.5], size=(64, 3))
    scene.particles = parts

    def no_xsph(p):
        p.xsph_epsilon = 0.0
    sim, o = make_pair(pkg, orc, scene, params_edit=no_xsph)
    print("case: out-of-grid + xsph off", flush=True)
    one_step_compare(sim, o, "oog")
    sim.close()


def test_pair_test_lower_bound(pkg, orc):
    """The pair test 1e-12 < r^2 < h^2 (sph_density_3d_grid.wgsl:150-159,
    xsph.wgsl:126, sph_force_3d_grid.wgsl:157) at its lower edge: a coincident pair
    (r^2 = 0), a pair at r = 2^-20 (r^2 = 2^-40 < 1e-12: density only, no pair
    terms) and one at r = 2^-19 (r^2 = 2^-38 > 1e-12: a pair, inside the
    separation range).  The offsets are exact in float32 for |x| < 1, so r^2
    is exact on both sides; a misplaced bound adds or drops a separation
    force and fails the force comparison."""
    scene = pkg.scenes.dam_break_8k()
    parts = scene.particles.copy()
    pairs = [(4000, 4001, 0.0), (5000, 5001, 2.0 ** -20), (6000, 6001, 2.0 ** -19)]
    for a, b, dx in pairs:
        assert np.all(np.abs(parts[a, 0:3]) < 1.0)
        parts[b, 0:3] = parts[a, 0:3]
        parts[b, 0] = np.float32(parts[a, 0] + np.float32(dx))
        assert float(parts[b, 0]) - float(parts[a, 0]) == dx
    scene.particles = parts
    sim, o = make_pair(pkg, orc, scene)
    print("case: pair-test lower bound (r^2 = 0, 2^-40, 2^-38)", flush=True)
    one_step_compare(sim, o, "r2 edge")
    sim.close()


@pytest.mark.parametrize("wide", [0, 4, 8, 256, 256 | 8])
def test_list_overflow_cluster(pkg, orc, wide):
    """1,500 particles packed into a ball of radius 0.03 (< h / 2) inside the
    8k block, the collapsed clusters of the 1M and 4M blocks in small: each
    has more in-range neighbours than a list holds (1024), so XSPH and force
    sweep its candidates again (with the 8-lane walk and with 16 or 32
    walkers); with FLUID_LIST_SKIP (256) the cluster's particles, each with
    more than 2048 candidates, record no list at all.  The step must still
    match the oracle."""
    scene = pkg.scenes.dam_break_8k()
    parts = scene.particles.copy()
    rng = np.random.default_rng(7)
    k = 1500
    centre = parts[:, 0:3].mean(axis=0)
    d = rng.normal(size=(k, 3))
    d *= (0.03 * rng.uniform(size=(k, 1)) ** (1 / 3)) / np.linalg.norm(d, axis=1, keepdims=True)
    parts[:k, 0:3] = (centre + d).astype(np.float32)
    scene.particles = parts
    sim, o = make_pair(pkg, orc, scene)
    if wide:
        wide_walk(pkg, sim, wide)
    print(f"case: list overflow cluster, walk flag={wide}", flush=True)
    one_step_compare(sim, o, "cluster")
    # the step's sorted state and grid: the cluster overflowed its lists
    st = sim.pair_stats()
    print(f"  max in-range {st['max_in_range']}", flush=True)
    assert st["max_in_range"] > 1024
    sim.close()


def test_reorder_big_cell(pkg, orc):
    """6,000 of the 8k block's particles packed into one grid cell (a ball of
    radius 0.012 about a cell centre): the cell's segment is longer than the
    reorder's LDS stage (4,096 entries), so the rank pass stages it in turns.
    Cell counts, starts and the permutation must be the oracle's exactly,
    and the step must meet the bar."""
    scene = pkg.scenes.dam_break_8k()
    parts = scene.particles.copy()
    rng = np.random.default_rng(11)
    k = 6000
    o_grid, cell = -2.08, 0.08
    centre = parts[:, 0:3].mean(axis=0)
    centre = o_grid + (np.floor((centre - o_grid) / cell) + 0.5) * cell
    d = rng.normal(size=(k, 3))
    d *= (0.012 * rng.uniform(size=(k, 1)) ** (1 / 3)) / np.linalg.norm(d, axis=1, keepdims=True)
    parts[:k, 0:3] = (centre + d).astype(np.float32)
    scene.particles = parts
    sim, o = make_pair(pkg, orc, scene)
    print("case: reorder with a 6,000-particle cell", flush=True)
    S = pkg.abi.STAGE
    sim.step_stages(0, S["reorder"]); o.stages(0, S["reorder"])
    compare_grid(sim, o)
    assert sim.read_cell_counts().max() >= 6000
    np.testing.assert_array_equal(sim.read_sorted_to_orig(), o.sorted_to_orig())
    gs, os_ = sim.read_sorted_particles(), o.sorted_particles()
    np.testing.assert_array_equal(gs[:, 0:11], os_[:, 0:11])
    sim.step_stages(S["density"], S["integrate"]); o.stages(S["density"], S["integrate"])
    gp, op, o6 = sim.read_particles(), o.particles(), o.o64.particles()
    check_fp("big cell position", gp[:, 0:3], op[:, 0:3], o6[:, 0:3])
    check_fp("big cell velocity", gp[:, 4:7], op[:, 4:7], o6[:, 4:7])
    sim.close()


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_mouse_force_modes(pkg, orc, mode):
    scene = pkg.scenes.dam_break_8k()
    params, cont, g = pkg.scenes.make_params(scene, pkg.abi, pkg.derive_sph_params,
                                             pkg.derive_container_geometry)
    mouse = pkg.abi.default_mouse_force()
    mouse.is_active, mouse.mode, mouse.radius, mouse.strength = 1, mode, 0.3, 40.0
    mouse.position[0], mouse.position[1], mouse.position[2] = -0.6, 0.4, -0.2
    mouse.direction[0], mouse.direction[1], mouse.direction[2] = 0.3, -0.2, -1.0
    sim = pkg.make_sim(scene)
    sim.update_mouse_force(mouse)
    o = Lockstep.__new__(Lockstep)
    o.o32 = orc.from_scene(scene, pkg.abi, params, cont, g, mouse=mouse)
    o.o64 = orc.from_scene(scene, pkg.abi, params, cont, g, fp64=True, mouse=mouse)
    print(f"case: mouse mode {mode}", flush=True)
    one_step_compare(sim, o, f"mouse{mode}")
    sim.close()


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4])
def test_rigid_body_shapes(pkg, orc, shape):
    scene = pkg.scenes.dam_break_8k()
    params, cont, g = pkg.scenes.make_params(scene, pkg.abi, pkg.derive_sph_params,
                                             pkg.derive_container_geometry)
    rb = pkg.abi.default_rigid_body()
    rb.is_active, rb.shape, rb.half_extent, rb.stiffness = 1, shape, 0.12, 300.0
    rb.position[0], rb.position[1], rb.position[2] = -0.6, 0.35, -0.2
    rb.velocity[0] = 0.5
    c, s_ = np.cos(0.3), np.sin(0.3)
    rb.rot_row0[0], rb.rot_row0[1] = c, -s_
    rb.rot_row1[0], rb.rot_row1[1] = s_, c
    sdf = None
    if shape == 4:
        res = 16
        ax = (np.arange(res) + 0.5) / res * 2 - 1
        z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
        sdf = (np.sqrt(x * x + y * y + z * z) - 0.8).astype(np.float32)
    sim = pkg.SphSimulation3DGrid(scene.particles, params, cont, sdf=sdf)
    sim.set_pcisph_iterations(scene.pcisph_iterations)
    sim.update_gravity(g)
    sim.update_rigid_body(rb)
    sim.clear_rigid_body_accum()
    o = Lockstep.__new__(Lockstep)
    o.o32 = orc.from_scene(scene, pkg.abi, params, cont, g, rigid_body=rb, sdf=sdf)
    o.o64 = orc.from_scene(scene, pkg.abi, params, cont, g, fp64=True, rigid_body=rb, sdf=sdf)
    print(f"case: rigid body shape {shape}", flush=True)
    one_step_compare(sim, o, f"rb{shape}")
    acc = sim.read_rigid_body_accum()
    want = o.o32.accum()
    got = np.array([acc.force_x, acc.force_y, acc.force_z, acc.contact_count, acc.torque_x,
                    acc.torque_y, acc.torque_z, 0], dtype=np.int64)
    print(f"  accum gpu {got[:7]} oracle {want[:7]}", flush=True)
    assert got[3] == want[3] and got[3] > 0
    # each contact's x1000 value truncates to i32; f32 differences of a few ulp
    # can move a truncation by one unit per contact
    assert np.all(np.abs(got[[0, 1, 2, 4, 5, 6]] - want[[0, 1, 2, 4, 5, 6]]) <=
                  max(4, got[3] // 50))
    sim.close()


def test_spawn_then_step(pkg, orc):
    scene = pkg.scenes.small_block(10, 10, 10)
    sim = pkg.make_sim(scene, max_particles=scene.n + 300)
    got = sim.spawn_particles([0.1, 0.5, 0.05], 500, 0.1)
    assert got == 300 and sim.num_particles() == scene.n + 300
    parts = sim.read_particles()
    # golden-angle spiral, layer of 20 (sph_3d_grid.rs:1543-1556)
    i = np.arange(300, dtype=np.float32)
    ang = i * np.float32(2.39996)
    r = np.float32(0.1) * np.sqrt(i / np.float32(300))
    np.testing.assert_allclose(parts[scene.n:, 0], np.float32(0.1) + r * np.cos(ang), atol=1e-6)
    np.testing.assert_allclose(parts[scene.n:, 1],
                               np.float32(0.5) + (i // 20) * np.float32(0.1) * np.float32(0.3),
                               atol=1e-6)
    assert sim.spawn_particles([0.0, 0.5, 0.0], 10, 0.1) == 0  # at capacity
    scene.particles = parts
    params, cont, g = pkg.scenes.make_params(scene, pkg.abi, pkg.derive_sph_params,
                                             pkg.derive_container_geometry)
    o = Lockstep(orc, pkg, scene, params, cont, g)
    print("case: spawn", flush=True)
    one_step_compare(sim, o, "spawn")
    sim.close()


def test_empty_simulation(pkg):
    scene = pkg.scenes.small_block(1, 1, 1)
    scene.particles = scene.particles[:0]
    sim = pkg.make_sim(scene, max_particles=16)
    sim.step(3)
    assert sim.num_particles() == 0
    assert sim.read_cell_counts().sum() == 0
    assert sim.spawn_particles([0.0, 0.5, 0.0], 4, 0.05) == 4
    sim.step(2)
    assert sim.read_cell_counts().sum() == 4
    sim.close()


def test_block_1m_full_size(pkg, orc):
    """Full-size 1,048,576-particle block: exact grid invariants, and one
    graph step against the float32 and float64 oracles at the full bar."""
    scene = pkg.scenes.block_1m()
    print(f"case: block_1m n={scene.n}", flush=True)
    params, cont, g = pkg.scenes.make_params(scene, pkg.abi, pkg.derive_sph_params,
                                             pkg.derive_container_geometry)
    sim = pkg.make_sim(scene)
    sim.step()
    counts = sim.read_cell_counts().astype(np.int64)
    starts = sim.read_cell_starts().astype(np.int64)
    assert counts.sum() == scene.n
    np.testing.assert_array_equal(starts, np.cumsum(counts))
    s2o = sim.read_sorted_to_orig()
    assert np.array_equal(np.sort(s2o), np.arange(scene.n))
    o = Lockstep(orc, pkg, scene, params, cont, g)
    o.step()
    np.testing.assert_array_equal(sim.read_cell_counts(), o.cell_counts())
    np.testing.assert_array_equal(s2o, o.sorted_to_orig())
    gp, op, o6 = sim.read_particles(), o.particles(), o.o64.particles()
    check_fp("1m position", gp[:, 0:3], op[:, 0:3], o6[:, 0:3])
    check_fp("1m velocity", gp[:, 4:7], op[:, 4:7], o6[:, 4:7])
    check_fp("1m force", gp[:, 8:11], op[:, 8:11], o6[:, 8:11])
    gs, os_, o6s = sim.read_sorted_particles(), o.sorted_particles(), o.sorted64()
    check_fp("1m density", gs[:, 11], os_[:, 11], o6s[:, 11])
    sim.close()


def test_block_4m_full_size(pkg, orc):
    """Full-size 4,194,304-particle block (the second throughput size):
    exact grid invariants, and one graph step against the float32 and float64
    oracles at the full bar."""
    scene = pkg.scenes.block_4m()
    print(f"case: block_4m n={scene.n}", flush=True)
    params, cont, g = pkg.scenes.make_params(scene, pkg.abi, pkg.derive_sph_params,
                                             pkg.derive_container_geometry)
    sim = pkg.make_sim(scene)
    sim.step()
    counts = sim.read_cell_counts().astype(np.int64)
    starts = sim.read_cell_starts().astype(np.int64)
    assert counts.sum() == scene.n
    np.testing.assert_array_equal(starts, np.cumsum(counts))
    s2o = sim.read_sorted_to_orig()
    assert np.array_equal(np.sort(s2o), np.arange(scene.n))
    o = Lockstep(orc, pkg, scene, params, cont, g)
    o.step()
    np.testing.assert_array_equal(sim.read_cell_counts(), o.cell_counts())
    np.testing.assert_array_equal(s2o, o.sorted_to_orig())
    gp, op, o6 = sim.read_particles(), o.particles(), o.o64.particles()
    check_fp("4m position", gp[:, 0:3], op[:, 0:3], o6[:, 0:3])
    check_fp("4m velocity", gp[:, 4:7], op[:, 4:7], o6[:, 4:7])
    check_fp("4m force", gp[:, 8:11], op[:, 8:11], o6[:, 8:11])
    gs, os_, o6s = sim.read_sorted_particles(), o.sorted_particles(), o.sorted64()
    check_fp("4m density", gs[:, 11], os_[:, 11], o6s[:, 11])
    sim.close()
