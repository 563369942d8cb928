"""Pins of the oracle's whitewater passes (CPU only): spray_emit.wgsl,
spray_simulate.wgsl and the ceilings' update of src/simulation/spray.rs.

The reference holds no spray test or fixture, so these pin the oracle with
what the mathematics fixes: two-particle closed forms for the trapped-air,
divergence and crest sums (sign conventions: a converging pair scores twice
its weight, a separating one nothing; only neighbours inside the surface
curve count for the crest); exact spawn counts where the expected count is
an integer; the spawn cylinder's geometry; the ballistic, foam and bubble
updates in closed form; the wall-cap extrapolation against a numerically
integrated sphere-cap volume; and the ceilings' update rule.
"""
import numpy as np
import pytest

H = 0.08
F = np.float32


def sph_params(pkg, rest_density=8000.0):
    return pkg.derive_sph_params(H, 35.0, 0.45, 15.0, 1.0, 6.5, 0.305, 0.008, 1,
                                 rest_density=rest_density)


def container(pkg, box=(1.6, 1.2, 0.8), floor_y=0.0):
    return pkg.derive_container_geometry(box[0], box[1], box[2], floor_y)


def sorted_state(pkg, orc, pos, vel=None, dens=None, normal=None, extent=2.0):
    """Sorted particle records, counts and inclusive starts of positions
    (grid of SphSimulation3DGrid::new, within-cell order by index), and the
    sorted slot of every input particle."""
    pos = np.asarray(pos, F).reshape(-1, 3)
    n = len(pos)
    vel = np.zeros((n, 3), F) if vel is None else np.asarray(vel, F).reshape(-1, 3)
    dens = np.full(n, 8000.0, F) if dens is None else np.asarray(dens, F).reshape(-1)
    normal = np.zeros((n, 3), F) if normal is None else np.asarray(normal, F).reshape(-1, 3)
    g = orc.grid_for(H, extent, n, pkg.abi)
    o, inv, G = F(g.grid_origin_x), F(g.inv_cell_size), g.grid_size_x
    c = np.floor((pos - o) * inv).astype(np.int64)
    idx = c[:, 0] + c[:, 1] * G + c[:, 2] * G * G
    order = np.argsort(idx, kind="stable")
    counts = np.bincount(idx, minlength=g.total_cells).astype(np.uint32)
    starts = np.cumsum(counts).astype(np.uint32)
    recs = np.zeros((n, 16), F)
    recs[:, 0:3] = pos[order]
    recs[:, 4:7] = vel[order]
    recs[:, 11] = dens[order]
    recs[:, 13:16] = normal[order]
    slot = np.empty(n, np.int64)
    slot[order] = np.arange(n)
    return recs, counts, starts, g, slot


def wvol(r, rho_j=8000.0, mass=1.0):
    return mass / rho_j * 315.0 / (64.0 * np.pi * H ** 9) * (H * H - r * r) ** 3


def spray_params(pkg, **kw):
    p = pkg.abi.default_spray_params(dt=0.0078125, max_particles=kw.pop("max_particles", 64))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


# ---------------------------------------------------------------- hash
def test_hash_is_a_bijection_and_uniform(orc):
    """hash (spray_emit.wgsl:156-164) is a composition of invertible steps
    (xorshift by 16, multiplication by an odd constant), so distinct seeds
    never collide; hash_float is k / 65535 for k = the low 16 bits."""
    lib = orc.load()
    seeds = np.arange(0, 200_000, dtype=np.uint32)
    hs = np.array([lib.oracle_spray_hash(int(s)) for s in seeds[:20_000]], np.uint32)
    assert len(np.unique(hs)) == len(hs)
    # the inverse of each step undoes the hash
    inv = pow(0x45d9f3b, -1, 1 << 32)

    def unxorshift16(x):
        return x ^ (x >> 16)

    for s, x in zip(seeds[:2000], hs[:2000]):
        x = int(x)
        x = unxorshift16(x)
        x = (x * inv) & 0xFFFFFFFF
        x = unxorshift16(x)
        x = (x * inv) & 0xFFFFFFFF
        x = unxorshift16(x)
        assert x == int(s)
    fl = np.array([lib.oracle_spray_hash_float(int(s)) for s in seeds[:20_000]])
    k = np.round(fl * 65535.0)
    assert np.all(np.abs(fl - k / 65535.0) < 1e-7) and fl.min() >= 0 and fl.max() <= 1
    assert abs(fl.mean() - 0.5) < 0.01


# ---------------------------------------------------------------- emit
def test_converging_pair_trapped_air_and_divergence(pkg, orc):
    """i moves at u toward j at rest, distance r along +x: x_hat = -e_x, so
    trapped air = u (1 - (-1)) V_j W = 2 u V_j W and divergence = -u (1 - r/h);
    with j behind i, trapped air = 0 and divergence = +u (1 - r/h)."""
    sph = sph_params(pkg)
    u, r = 3.0, 0.04
    pi = np.array([0.1, 0.5, 0.1])
    for side in (+1, -1):
        pos = [pi, pi + side * np.array([r, 0, 0])]
        recs, cnt, st, g, slot = sorted_state(pkg, orc, pos, vel=[[u, 0, 0], [0, 0, 0]])
        pot, ne, fm = orc.spray_potentials(recs, cnt, st, g, sph, spray_params(pkg), 3.0, 1.0)
        got = pot[slot[0]]
        ta_want = 2 * u * wvol(r) if side > 0 else 0.0
        div_want = -side * u * (1 - r / H)
        assert got[3] == 1.0
        assert abs(got[0] - ta_want) <= 1e-5 * max(ta_want, 1e-30) + (0 if side > 0 else 0.0)
        if side < 0:
            assert got[0] == 0.0
        assert abs(got[2] - div_want) < 1e-5 * abs(div_want)
        # j is below min_speed: no neighbour work at all
        assert np.all(pot[slot[1]] == 0) and ne[slot[1]] == 0
        assert np.all(ne == 0)  # fewer than 10 neighbours: nothing spawns


def test_crest_counts_only_neighbours_inside_the_curve(pkg, orc):
    """Surface particle (density < 0.9 rho0) with normal +y: a neighbour
    below (x_hat = +y, so -x_hat . n < 0) adds (1 - n_i . n_j) V_j W; one
    above adds nothing; a particle at rest density is not eligible."""
    sph = sph_params(pkg)
    r = 0.05
    pi = np.array([0.1, 0.5, 0.1])
    pos = [pi, pi - [0, r, 0], pi + [0, r, 0]]
    vel = [[0, 3.0, 0], [0, 0, 0], [0, 0, 0]]
    for nj, factor in (([2.0, 0, 0], 1.0), ([0, 5.0, 0], 0.0), ([0, -1.0, 0], 2.0)):
        normal = [[0, 3.0, 0], nj, [1.0, 0, 0]]
        recs, cnt, st, g, slot = sorted_state(pkg, orc, pos, vel, dens=[5000, 8000, 8000],
                                              normal=normal)
        pot, _, _ = orc.spray_potentials(recs, cnt, st, g, sph, spray_params(pkg), 3.0, 1.0)
        want = factor * wvol(r)
        assert abs(pot[slot[0], 1] - want) <= 1e-5 * max(want, 1e-12), (nj, pot[slot[0]])
        recs[slot[0], 11] = 7300.0  # above 0.9 rho0: no crest
        pot, _, _ = orc.spray_potentials(recs, cnt, st, g, sph, spray_params(pkg), 3.0, 1.0)
        assert pot[slot[0], 1] == 0.0


def ring_of_neighbours(k, r, centre, seed=3):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(k, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return centre + r * d


@pytest.mark.parametrize("rate,want", [(384.0, 3), (2048.0, 16), (12800.0, 16), (0.0, 0)])
def test_spawn_count_exact_for_integer_expectation(pkg, orc, rate, want):
    """With phi_ta = phi_k = 1 and no crest weight, the expected count is
    emission_rate * dt; an integer expectation spawns exactly that many (the
    fractional draw compares against 0), capped at 16 per frame."""
    sph = sph_params(pkg)
    c = np.array([0.1, 0.5, 0.1])
    pos = np.vstack([c, ring_of_neighbours(12, 0.05, c)])
    vel = np.zeros((13, 3))
    vel[0] = [3.0, 0, 0]  # above the 2.5 m/s saturation: phi_k = 1
    recs, cnt, st, g, slot = sorted_state(pkg, orc, pos, vel)
    prm = spray_params(pkg, emission_rate=rate, k_trapped_air=1.0, k_wave_crest=0.0)
    pot, ne, fm = orc.spray_potentials(recs, cnt, st, g, sph, prm, 1e-4, 1.0)
    assert pot[slot[0], 3] == 12 and pot[slot[0], 0] > 1e-4
    assert ne[slot[0]] == want and ne.sum() == want
    assert fm[0] == pot[slot[0], 0]


def test_fractional_spawns_follow_the_hash(pkg, orc):
    """expected = 2.5 spawns 3 exactly when hash_float(seed * 13 + 5) < 0.5
    with seed = slot + frame * 7919; over many frames about half do."""
    sph = sph_params(pkg)
    lib = orc.load()
    c = np.array([0.1, 0.5, 0.1])
    pos = np.vstack([c, ring_of_neighbours(12, 0.05, c)])
    vel = np.zeros((13, 3))
    vel[0] = [3.0, 0, 0]
    recs, cnt, st, g, slot = sorted_state(pkg, orc, pos, vel)
    got = []
    for frame in range(200):
        prm = spray_params(pkg, emission_rate=320.0, k_trapped_air=1.0, k_wave_crest=0.0,
                           frame_count=frame)
        _, ne, _ = orc.spray_potentials(recs, cnt, st, g, sph, prm, 1e-4, 1.0)
        seed = (int(slot[0]) + frame * 7919) & 0xFFFFFFFF
        draw = lib.oracle_spray_hash_float((seed * 13 + 5) & 0xFFFFFFFF)
        assert ne[slot[0]] == (3 if draw < 0.5 else 2)
        got.append(int(ne[slot[0]]))
    assert 0.4 < np.mean(np.array(got) == 3) < 0.6


def test_crest_velocity_gate(pkg, orc):
    """The crest counts only while the particle moves along its normal
    (v_hat . n_hat >= 0.6); the reported maxima follow the gated values."""
    sph = sph_params(pkg)
    c = np.array([0.1, 0.5, 0.1])
    below = c + np.array([[0.01, -0.04, 0.0], [-0.02, -0.05, 0.01], [0.0, -0.03, -0.02]])
    pos = np.vstack([c, below, ring_of_neighbours(9, 0.06, c, seed=5)])
    n = len(pos)
    normal = np.zeros((n, 3))
    normal[1:] = [1.0, 0, 0]
    normal[0] = [0, 1.0, 0]
    dens = np.full(n, 8000.0)
    dens[0] = 5000.0
    for v0, kept in (([0, 3.0, 0], True), ([3.0, 0, 0], False)):
        vel = np.zeros((n, 3))
        vel[0] = v0
        recs, cnt, st, g, slot = sorted_state(pkg, orc, pos, vel, dens, normal)
        pot, _, fm = orc.spray_potentials(recs, cnt, st, g, sph, spray_params(pkg), 3.0, 1.0)
        assert pot[slot[0], 3] >= 10
        assert (pot[slot[0], 1] > 0) == kept
        assert fm[1] == (pot[slot[0], 1] if kept else 0.0)


def test_slow_particles_do_nothing(pkg, orc):
    sph = sph_params(pkg)
    c = np.array([0.1, 0.5, 0.1])
    pos = np.vstack([c, ring_of_neighbours(12, 0.05, c)])
    vel = np.zeros((13, 3))
    vel[0] = [0.79, 0, 0]  # below min_speed 0.8
    recs, cnt, st, g, slot = sorted_state(pkg, orc, pos, vel)
    pot, ne, fm = orc.spray_potentials(recs, cnt, st, g, sph, spray_params(pkg), 3.0, 1.0)
    assert not pot.any() and not ne.any() and not fm.any()


# ---------------------------------------------------------------- spawn
def test_spawn_geometry_slots_and_wrap(pkg, orc):
    """Spawns lie in the cylinder of radius h/2 and length dt*|v| along v;
    velocities are v * speed_multiplier plus a jitter of at most
    velocity_jitter per axis; lifetimes lie in [lt_min, lt_min + phi_k *
    span]; slots are handed out in sorted order from the head and wrap, the
    later spawn keeping a reused slot."""
    sph = sph_params(pkg)
    rng = np.random.default_rng(7)
    n = 40
    pos = rng.uniform(-0.5, 0.5, (n, 3))
    vel = rng.normal(size=(n, 3)) * 2.0
    vel[5] = [0, 3.0, 0]  # |v_dir.y| > 0.99 takes the x reference axis
    recs, _, _, _, _ = sorted_state(pkg, orc, pos, vel)
    ne = rng.integers(0, 5, n).astype(np.uint32)
    prm = spray_params(pkg, max_particles=1024)
    ring = np.zeros((1024, 12), np.uint32)
    head = orc.spray_spawn(recs, sph, prm, ne, 1000, ring)
    assert head == 1000 + int(ne.sum())
    f = ring.view(np.float32)
    k = 0
    for i in range(n):
        v = recs[i, 4:7].astype(np.float64)
        sp = np.linalg.norm(v)
        vd = v / sp
        phi_k = np.clip((sp - 0.8) / (2.5 - 0.8), 0, 1)
        for e in range(ne[i]):
            rec = f[(1000 + k) % 1024]
            k += 1
            off = rec[0:3] - recs[i, 0:3]
            along = off @ vd
            radial = np.linalg.norm(off - along * vd)
            assert -1e-6 <= along <= prm.dt * sp * (1 + 1e-5) + 1e-7
            assert radial <= 0.5 * H * (1 + 1e-5)
            jit = rec[4:7] - v * prm.speed_multiplier
            assert np.all(np.abs(jit) <= prm.velocity_jitter * (1 + 1e-5))
            lt_min = prm.lifetime * (1 - prm.lifetime_variation)
            span = 2 * prm.lifetime * prm.lifetime_variation
            assert lt_min * (1 - 1e-6) <= rec[3] <= lt_min + phi_k * span + 1e-5
            assert rec[3] == rec[7] and ring[(1000 + k - 1) % 1024, 8] == 0 and rec[9] == 0
    # a ring smaller than one frame's spawns keeps the last ones
    small = np.zeros((16, 12), np.uint32)
    prm16 = spray_params(pkg, max_particles=16)
    h2 = orc.spray_spawn(recs, sph, prm16, ne, 5, small)
    full = np.zeros((4096, 12), np.uint32)
    prm_big = spray_params(pkg, max_particles=4096)
    orc.spray_spawn(recs, sph, prm_big, ne, 0, full)
    total = int(ne.sum())
    assert h2 == 5 + total
    for q in range(total - 16, total):
        assert np.array_equal(small[(5 + q) % 16], full[q])


# ---------------------------------------------------------------- simulate
def one_particle_ring(pos, vel, lifetime=1.0, age=1.0, kind=0, size=4):
    ring = np.zeros((size, 12), np.uint32)
    f = ring.view(np.float32)
    f[0, 0:3], f[0, 3], f[0, 4:7], f[0, 7] = pos, lifetime, vel, lifetime
    ring[0, 8] = kind
    f[0, 9] = age
    return ring


def test_ballistic_spray_closed_form(pkg, orc):
    """No fluid within h: spray, v' = v + dt (g - drag v), x' = x + dt v',
    lifetime unchanged, age + dt; dead slots are left alone."""
    sph, cont = sph_params(pkg), container(pkg)
    recs, cnt, st, g, _ = sorted_state(pkg, orc, [[0.5, 0.2, 0.3]])
    prm = spray_params(pkg, max_particles=4)
    x0, v0 = np.array([0.0, 0.6, 0.0]), np.array([0.5, 1.0, -0.3])
    ring = one_particle_ring(x0, v0, lifetime=0.7, age=0.05)
    ring[2, 3] = np.float32(-1.0).view(np.uint32)
    before = ring.copy()
    live = orc.spray_simulate(ring, recs, cnt, st, g, sph, cont, prm)
    f = ring.view(np.float32)
    gv = np.array([0, prm.gravity_y, 0])
    v1 = v0 + prm.dt * (gv - prm.drag * v0)
    assert np.allclose(f[0, 4:7], v1, rtol=1e-6, atol=1e-7)
    assert np.allclose(f[0, 0:3], x0 + prm.dt * v1, rtol=1e-6, atol=1e-7)
    assert ring[0, 8] == 0 and f[0, 3] == np.float32(0.7)
    assert abs(f[0, 9] - (0.05 + prm.dt)) < 1e-7
    assert np.array_equal(ring[1:], before[1:])
    assert list(live) == [1, 1, 0, 0]


def _fluid_around(c, k, radius, seed=11):
    return ring_of_neighbours(k, radius, c, seed)


@pytest.mark.parametrize("k,age,kind", [(12, 1.0, 1), (12, 0.1, 0), (25, 1.0, 2), (6, 1.0, 0)])
def test_classification_and_updates(pkg, orc, k, age, kind):
    """Neighbour counts <= 6 make spray, >= 20 bubble, else foam (newborn
    foam flies as spray).  Foam takes the fluid velocity U and loses dt of
    lifetime; a bubble rises: v' = mix(v + dt (-b g), U, drag)."""
    sph, cont = sph_params(pkg), container(pkg)
    c = np.array([0.0, 0.6, 0.0])
    U = np.array([0.3, -0.2, 0.1])
    fl = _fluid_around(c, k, 0.5 * H)
    recs, cnt, st, g, _ = sorted_state(pkg, orc, fl, np.tile(U, (k, 1)))
    prm = spray_params(pkg, max_particles=4)
    v0 = np.array([1.0, 0.5, 0.0])
    ring = one_particle_ring(c, v0, lifetime=0.9, age=age)
    live = orc.spray_simulate(ring, recs, cnt, st, g, sph, cont, prm)
    f = ring.view(np.float32)
    assert ring[0, 8] == kind
    gv = np.array([0, prm.gravity_y, 0])
    if kind == 0:
        want, life = v0 + prm.dt * (gv - prm.drag * v0), 0.9
    elif kind == 1:
        want, life = U, 0.9 - prm.dt
    else:
        t = prm.bubble_drag
        want, life = (v0 + prm.dt * (-prm.bubble_buoyancy * gv)) * (1 - t) + U * t, 0.9
    assert np.allclose(f[0, 4:7], want, rtol=2e-6, atol=1e-6)
    assert abs(f[0, 3] - life) < 1e-6
    assert live[0] == 1 and live[1 + kind] == 1


def cap_fraction_numeric(t, m=200_000):
    """Volume of a unit sphere's cap of height t (midpoint rule over the
    slices z in [1 - t, 1]) over the sphere's volume."""
    z = 1 - t + (np.arange(m) + 0.5) * t / m
    return np.sum(np.pi * (1 - z * z)) * t / m / (4 / 3 * np.pi)


def test_wall_cap_extrapolation_threshold(pkg, orc):
    """A foam-range count of 12 beside the floor is extrapolated by the
    missing cap volume: the particle turns bubble exactly when 12 / (1 - cap)
    rounds to 20 or more, i.e. cap >= 1 - 12 / 19.5, with the cap volume from
    numerical integration; just outside the threshold it stays foam."""
    sph, cont = sph_params(pkg), container(pkg)
    target = 1 - 12 / 19.5
    lo, hi = 0.0, 1.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if cap_fraction_numeric(mid) < target else (lo, mid)
    d_star = H * (1 - 0.5 * (lo + hi))  # distance to the floor at the threshold
    prm = spray_params(pkg, max_particles=4)
    for d, kind in ((d_star * 0.98, 2), (d_star * 1.02, 1)):
        c = np.array([0.0, d, 0.0])  # floor at y = 0, far from the other walls
        fl = _fluid_around(c, 12, 0.5 * H, seed=13)
        recs, cnt, st, g, _ = sorted_state(pkg, orc, fl)
        ring = one_particle_ring(c, [0, 0, 0], lifetime=1.0, age=1.0)
        orc.spray_simulate(ring, recs, cnt, st, g, sph, cont, prm)
        assert ring[0, 8] == kind, (d / H, kind)


def test_leaving_the_box_kills(pkg, orc):
    sph, cont = sph_params(pkg), container(pkg)
    recs, cnt, st, g, _ = sorted_state(pkg, orc, [[0.5, 0.2, 0.3]])
    prm = spray_params(pkg, max_particles=4)
    ring = one_particle_ring([0.7999, 0.6, 0.0], [2.0, 0.0, 0.0])
    live = orc.spray_simulate(ring, recs, cnt, st, g, sph, cont, prm)
    assert ring.view(np.float32)[0, 3] == 0.0 and list(live) == [0, 0, 0, 0]


# ---------------------------------------------------------------- ceilings
def test_auto_limit_update_rule(orc):
    """spray.rs:40-51: rise 5% of the gap toward a larger maximum, decay 1%
    toward a smaller one, hold on a frame without one, never below the
    floor."""
    u = orc.spray_update_limit
    assert u(3.0, 0.0, 0.2) == 3.0 and u(3.0, 1e-6, 0.2) == 3.0
    assert abs(u(3.0, 10.0, 0.2) - (3.0 + 0.05 * 7.0)) < 1e-6
    assert abs(u(3.0, 1.0, 0.2) - (3.0 - 0.01 * 2.0)) < 1e-6
    assert u(0.2, 1e-5, 0.2) == np.float32(0.2)
    x = 1.0
    for _ in range(2000):
        x = u(x, 0.5, 0.05)
    assert abs(x - 0.5) < 1e-3
