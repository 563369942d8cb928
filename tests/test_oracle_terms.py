"""Closed-form pins of the oracle's per-term physics (CPU only).

Each pin sets up a configuration small enough to work out by hand (one, two
or three particles, one wall, one mouse, one rigid body) and compares the
float64 oracle with the value the reference's formula gives for it, computed
here in float64 numpy from the configuration alone:

  * PCISPH pressure acceleration and its CFL clamp
    (src/shaders/pcisph_solve.wgsl:150-192): a compressed pair over two
    iterations, an asymmetric triple, and momentum conservation of the
    symmetric form over a random cluster;
  * close-range separation and the index-hash direction
    (sph_force_3d_grid.wgsl:160-189), at slot differences 1 and 41;
  * the wall penalty layer (sph_integrate_3d.wgsl:105-160) on each of the
    six faces, moving in and out, upright and in a tilted container, and the
    hard backstop with its restitution (:343-386) on each face;
  * the five mouse modes and the default branch (:163-200);
  * the five rigid-body shapes, the sampled custom SDF, and the x1000
    fixed-point accumulator (:203-341).

The pins are plain functions so that tests/test_oracle_mutants.py can run
them against builds of the oracle with planted mistakes and require them to
fail.
"""
import numpy as np
import pytest

PI32 = float(np.float32(3.14159265359))


def f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def make_oracle(pkg, orc, pos, vel=None, fp64=True, mouse=None, rigid_body=None, sdf=None,
                delta_scale=1.0, **scene_kw):
    pos = np.asarray(pos, dtype=np.float64)
    sc = pkg.scenes.Scene("pin", pkg.scenes.particles_from_positions(
        pos, None if vel is None else np.asarray(vel, dtype=np.float64)), (1.6, 1.2, 0.8), 0.0)
    for k, v in scene_kw.items():
        setattr(sc, k, v)
    params, cont, g = pkg.scenes.make_params(sc, pkg.abi, pkg.derive_sph_params,
                                             pkg.derive_container_geometry)
    params.pcisph_delta *= delta_scale
    o = orc.from_scene(sc, pkg.abi, params, cont, g, fp64=fp64, mouse=mouse,
                       rigid_body=rigid_body, sdf=sdf)
    return o, params, cont, sc


def poly6(p, r2):
    return 315.0 / (64.0 * PI32 * p.kernel_radius_pow9) * (p.kernel_radius_sq - r2) ** 3


def spiky_grad(p, r):
    return 45.0 / (PI32 * p.kernel_radius_pow6) * (p.kernel_radius - r) ** 2


def near_w(p, r):
    return 15.0 / (PI32 * p.kernel_radius_pow6) * (p.kernel_radius - r) ** 3


def by_orig(o, arr):
    """Rows of a sorted-order array rearranged to original index order."""
    out = np.empty_like(arr)
    out[o.sorted_to_orig()] = arr
    return out


# ---------------------------------------------------------------- PCISPH
NO_NP = dict(near_stiffness=0.0, viscosity=0.0, surface_tension=0.0, gravity=(0.0, 0.0, 0.0))


def pcisph_two_iterations(pkg, orc, pos, rho0, delta_scale=1.0):
    """Stages 0-8 with two PCISPH iterations and no non-pressure forces: the
    first iteration has zero pressures and positions x* = x, so its density
    and pressure are closed forms of the positions, and the second one's
    pressure acceleration is the reference's symmetric form over them."""
    o, p, _, _ = make_oracle(pkg, orc, pos, rest_density=rho0, pcisph_iterations=2,
                             delta_scale=delta_scale, **NO_NP)
    o.stages(0, 8)
    A = by_orig(o, o.predicted(0))  # the second iteration wrote buffer A
    x = f32(pos)
    d = x[:, None, :] - x[None, :, :]
    r2 = (d ** 2).sum(-1)
    inr = r2 < p.kernel_radius_sq
    rho = (p.mass * np.where(inr, poly6(p, r2), 0.0)).sum(1)
    pres = p.pcisph_delta * np.maximum(0.0, rho - p.rest_density)
    return o, p, A, x, r2, inr, rho, pres


def pressure_accel(p, x, r2, inr, rho, pres):
    n = len(x)
    a = np.zeros((n, 3))
    for i in range(n):
        for j in range(n):
            if i == j or not inr[i, j] or r2[i, j] <= 1e-12:
                continue
            r = np.sqrt(r2[i, j])
            dir_ij = (x[j] - x[i]) / r  # from i toward its neighbour
            a[i] += -p.mass * (pres[i] / rho[i] ** 2 + pres[j] / rho[j] ** 2) \
                * spiky_grad(p, r) * dir_ij
    return a


def pin_pcisph_pair_repels():
    def run(pkg, orc):
        r = 0.03
        o, p, A, x, r2, inr, rho, pres = pcisph_two_iterations(
            pkg, orc, [[0.0, 0.6, 0.0], [r, 0.6, 0.0]], 4000.0)
        rr = np.sqrt(r2[0, 1])
        W0, Wr = poly6(p, 0.0), poly6(p, rr * rr)
        rho1 = p.mass * (W0 + Wr)
        p1 = p.pcisph_delta * (rho1 - p.rest_density)
        assert p1 > 0
        # a_0 = -m (p/rho^2 + p/rho^2) grad W (r) along +x (toward particle 1)
        a0 = -p.mass * 2.0 * p1 / rho1 ** 2 * spiky_grad(p, rr)
        assert abs(a0) < 0.25 * p.kernel_radius / p.dt ** 2  # below the CFL clamp
        want_v = np.array([[a0 * p.dt, 0, 0], [-a0 * p.dt, 0, 0]])
        np.testing.assert_allclose(A[:, 4:7], want_v, rtol=1e-10, atol=1e-14)
        np.testing.assert_allclose(A[:, 0:3], x + p.dt * want_v, rtol=0, atol=1e-14)
        assert A[0, 4] < 0 < A[1, 4]  # a compressed pair separates
        # pressure accumulates: p2 = p1 + delta * (rho* - rho0), rho* unchanged
        np.testing.assert_allclose(A[:, 3], [2 * p1, 2 * p1], rtol=1e-12)
        np.testing.assert_allclose(A[:, 7], [rho1, rho1], rtol=1e-12)
    return run


def pin_pcisph_triple_asymmetric():
    def run(pkg, orc):
        pos = [[0.0, 0.6, 0.0], [0.03, 0.6, 0.0], [0.07, 0.6, 0.0]]
        o, p, A, x, r2, inr, rho, pres = pcisph_two_iterations(pkg, orc, pos, 3500.0)
        assert len(set(np.round(pres, 6))) == 3 and pres.min() > 0
        a = pressure_accel(p, x, r2, inr, rho, pres)
        np.testing.assert_allclose(A[:, 4:7], p.dt * a, rtol=1e-10, atol=1e-14)
        # the symmetric form conserves momentum
        assert abs(a[:, 0].sum()) < 1e-10 * np.abs(a[:, 0]).sum()
        np.testing.assert_allclose(A[:, 7], rho, rtol=1e-12)
    return run


def pin_pcisph_cluster_momentum():
    def run(pkg, orc):
        rng = np.random.default_rng(17)
        pos = rng.uniform([-0.07, 0.53, -0.07], [0.07, 0.67, 0.07], size=(40, 3))
        x = f32(pos)
        r2 = ((x[:, None] - x[None]) ** 2).sum(-1)
        o0, p0, _, _ = make_oracle(pkg, orc, pos)
        rho = (np.where(r2 < p0.kernel_radius_sq, poly6(p0, r2), 0.0)).sum(1)
        rho0 = float(np.quantile(rho, 0.95))
        o, p, A, x, r2, inr, rho, pres = pcisph_two_iterations(pkg, orc, pos, rho0)
        a = pressure_accel(p, x, r2, inr, rho, pres)
        assert np.linalg.norm(a, axis=1).max() < 0.25 * p.kernel_radius / p.dt ** 2
        assert (pres > 0).sum() >= 2
        ap = A[:, 4:7] / p.dt
        np.testing.assert_allclose(ap, a, rtol=1e-9, atol=1e-9 * np.abs(a).max())
        tot = np.abs(ap).sum(0)
        assert np.all(np.abs(ap.sum(0)) < 1e-9 * tot)
    return run


def pin_pcisph_cfl_clamp():
    def run(pkg, orc):
        # twenty times the derived delta: the unclamped pair acceleration
        # (about 560) exceeds 0.25 h / dt^2 = 312.5, so it is cut to that
        o, p, A, x, r2, inr, rho, pres = pcisph_two_iterations(
            pkg, orc, [[0.0, 0.6, 0.0], [0.03, 0.6, 0.0]], 4000.0, delta_scale=20.0)
        amax = 0.25 * p.kernel_radius / p.dt ** 2
        a = pressure_accel(p, x, r2, inr, rho, pres)
        assert np.abs(a).max() > 1.5 * amax
        np.testing.assert_allclose(A[:, 4:7], [[-p.dt * amax, 0, 0], [p.dt * amax, 0, 0]],
                                   rtol=1e-12)
    return run


# ---------------------------------------------------------------- separation
def hash_dir(d):
    v = np.array([np.sin(d * 1.0), np.cos(d * 2.0), np.sin(d * 3.0)])
    return v / np.linalg.norm(v)


def separation_case(pkg, orc, far, near_k):
    """A pair 0.05 h apart (below 0.1 h).  `far` adds 40 particles, farther
    than h from the pair, whose cells sort between the pair's two cells, so
    the pair's sorted slots differ by 41."""
    h = 0.08
    if far:
        a, b = [0.03, 0.6, -0.002], [0.03, 0.6, 0.002]
        extra = [[-0.7 + 0.035 * k, 1.0, -0.04] for k in range(40)]
        pos = [a] + extra + [b]
    else:
        pos = [[0.0, 0.6, 0.0], [0.05 * h, 0.6, 0.0]]
    kw = dict(NO_NP)
    kw["near_stiffness"] = near_k
    o, p, _, _ = make_oracle(pkg, orc, pos, **kw)
    o.stages(0, 6)
    ia, ib = 0, len(pos) - 1
    o2s = np.empty(len(pos), dtype=np.int64)
    o2s[o.sorted_to_orig()] = np.arange(len(pos))
    sa, sb = int(o2s[ia]), int(o2s[ib])
    assert abs(sa - sb) == (41 if far else 1)
    x = f32(pos)
    r = np.linalg.norm(x[ia] - x[ib])
    assert r < 0.1 * p.kernel_radius
    # density of each pair member: self + partner (the far particles are beyond h)
    rho = p.mass * (poly6(p, 0.0) + poly6(p, r * r))
    nd = p.mass * (near_w(p, 0.0) + near_w(p, r))
    npres = p.near_stiffness * nd
    sep = 200.0 * (1.0 - r / (0.1 * p.kernel_radius))
    got = o.particles()[:, 8:11]
    for i, s, j in ((ia, sa, sb), (ib, sb, sa)):
        dirv = hash_dir(float(s) - float(j))
        mag = -p.mass * (2 * npres) / (2.0 * rho * nd) * near_w_grad(p, r) - sep
        np.testing.assert_allclose(got[i], mag * dirv, rtol=1e-10, atol=1e-12)


def near_w_grad(p, r):
    return 45.0 / (PI32 * p.kernel_radius_pow6) * (p.kernel_radius - r) ** 2


def pin_separation_adjacent_slots():
    return lambda pkg, orc: separation_case(pkg, orc, far=False, near_k=0.0)


def pin_separation_far_slots_with_near_pressure():
    return lambda pkg, orc: separation_case(pkg, orc, far=True, near_k=0.45)


# ---------------------------------------------------------------- walls
FACES = [(0, -1), (0, +1), (1, -1), (1, +1), (2, -1), (2, +1)]  # (axis, side)


def rotation_rows(cont):
    F = np.array([cont.forward_row0[:3], cont.forward_row1[:3], cont.forward_row2[:3]], float)
    I = np.array([cont.inverse_row0[:3], cont.inverse_row1[:3], cont.inverse_row2[:3]], float)
    return F, I


def wall_case(pkg, orc, axis, side, into, tilt):
    """One particle 0.2 h inside the wall layer of one face, integrate alone."""
    sc0 = pkg.scenes.Scene("w", np.zeros((1, 16), np.float32), (1.6, 1.2, 0.8), 0.0, tilt=tilt)
    params, cont, _ = pkg.scenes.make_params(sc0, pkg.abi, pkg.derive_sph_params,
                                             pkg.derive_container_geometry)
    F, I = rotation_rows(cont)
    half = np.array([cont.half_width, cont.half_height, cont.half_depth], float)
    h = params.kernel_radius
    lp = np.array([0.1, -0.05, 0.07])
    lp[axis] = side * (half[axis] - 0.2 * h)
    lv = np.array([0.15, -0.1, 0.2])
    lv[axis] = (0.5 if into else -0.5) * side  # side * +: toward the face
    wp = F @ lp + np.array([0, cont.center_y, 0])
    wv = F @ lv
    o, p, c2, _ = make_oracle(pkg, orc, [wp], [wv], tilt=tilt)
    o.stage(10)
    out = o.particles()[0]
    # expected, from the stored f32 position and velocity
    x = f32(wp)
    v = f32(wv)
    l_pos = I @ (x - np.array([0, cont.center_y, 0]))
    l_vel = I @ v
    n = np.zeros(3)
    n[axis] = -side  # inward normal of the face
    dist = half[axis] + side * (-l_pos[axis])  # distance to the face
    bl = 0.4 * h
    t = 1.0 - dist / bl
    assert 0.4 < t < 0.6
    vin = -float(l_vel @ n)  # speed into the face
    a_loc = n * (cont.wall_stiffness * t * t + (8.0 * t * vin if vin > 0 else 0.0))
    assert (vin > 0) == into
    v1 = v + p.dt * (F @ a_loc)
    x1 = x + p.dt * v1
    cy = np.array([0, cont.center_y, 0])
    l1 = I @ (x1 - cy)
    assert np.all(np.abs(l1) < half)  # the backstop stays off
    # the backstop maps the state to the container frame and back
    # (sph_integrate_3d.wgsl:348-386); with f32 rows F I differs from I by ~1e-7
    np.testing.assert_allclose(out[4:7], F @ (I @ v1), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(out[0:3], F @ l1 + cy, rtol=1e-12, atol=1e-12)


def pin_walls(tilt):
    def run(pkg, orc):
        for axis, side in FACES:
            for into in (True, False):
                wall_case(pkg, orc, axis, side, into, tilt)
    return run


def pin_wall_corner():
    def run(pkg, orc):
        # floor and the -x wall at once: the two layers add
        sc0 = pkg.scenes.Scene("w", np.zeros((1, 16), np.float32), (1.6, 1.2, 0.8), 0.0)
        params, cont, _ = pkg.scenes.make_params(sc0, pkg.abi, pkg.derive_sph_params,
                                                 pkg.derive_container_geometry)
        h = params.kernel_radius
        wp = [-0.8 + 0.1 * h, 0.25 * h, 0.0]
        wv = [-0.3, -0.2, 0.0]
        o, p, _, _ = make_oracle(pkg, orc, [wp], [wv])
        o.stage(10)
        x, v = f32(wp), f32(wv)
        bl = 0.4 * h
        tx = 1.0 - (x[0] + cont.half_width) / bl
        ty = 1.0 - x[1] / bl
        a = np.array([250.0 * tx * tx - v[0] * 8.0 * tx, 250.0 * ty * ty - v[1] * 8.0 * ty, 0.0])
        v1 = v + p.dt * a
        np.testing.assert_allclose(o.particles()[0, 4:7], v1, rtol=1e-12, atol=1e-14)
    return run


def backstop_case(pkg, orc, axis, side, tilt):
    """A particle outside the wall layer crossing one face within the step:
    no penalty acts, and the backstop puts it on the face with the normal
    velocity reflected and scaled by the damping (restitution)."""
    sc0 = pkg.scenes.Scene("w", np.zeros((1, 16), np.float32), (1.6, 1.2, 0.8), 0.0, tilt=tilt)
    params, cont, _ = pkg.scenes.make_params(sc0, pkg.abi, pkg.derive_sph_params,
                                             pkg.derive_container_geometry)
    F, I = rotation_rows(cont)
    half = np.array([cont.half_width, cont.half_height, cont.half_depth], float)
    h = params.kernel_radius
    cy = np.array([0, cont.center_y, 0])
    lp = np.array([0.1, -0.05, 0.07])
    lp[axis] = side * (half[axis] - 0.5 * h)
    lv = np.array([0.3, -0.2, 0.25])
    lv[axis] = 7.0 * side
    o, p, _, _ = make_oracle(pkg, orc, [F @ lp + cy], [F @ lv], tilt=tilt)
    o.stage(10)
    out = o.particles()[0]
    x, v = f32(F @ lp + cy), f32(F @ lv)
    l0 = I @ (x - cy)
    assert half[axis] - abs(l0[axis]) > 0.4 * h  # outside the wall layer
    l1 = I @ (x + p.dt * v - cy)
    lv1 = I @ v
    assert abs(l1[axis]) > half[axis]  # crossed the face
    l1[axis] = side * half[axis]
    lv1[axis] = -side * abs(lv1[axis]) * cont.damping
    np.testing.assert_allclose(out[0:3], F @ l1 + cy, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(out[4:7], F @ lv1, rtol=1e-12, atol=1e-12)


def pin_backstop(tilt):
    def run(pkg, orc):
        for axis, side in FACES:
            backstop_case(pkg, orc, axis, side, tilt)
    return run


# ---------------------------------------------------------------- mouse
def mouse_struct(pkg, pos, radius, strength, mode, direction, active=1):
    m = pkg.abi.GpuMouseForce()
    m.position[0], m.position[1], m.position[2] = pos
    m.radius, m.strength, m.is_active, m.mode = radius, strength, active, mode
    m.direction[0], m.direction[1], m.direction[2] = direction
    return m


def mouse_expected(mode, to_mouse, radius, strength, axis):
    dist = np.linalg.norm(to_mouse)
    if not (dist < radius and dist > 0.001):
        return np.zeros(3)
    s = strength * (1.0 - dist / radius)
    fd = to_mouse / dist
    if mode == 1:
        return fd * s
    if mode == 2:
        t = np.cross(fd, axis / np.linalg.norm(axis))
        tl = np.linalg.norm(t)
        return t / tl * s if tl > 0.001 else np.zeros(3)
    if mode == 3:
        return -fd * s * 3.0
    if mode == 4:
        d = fd + np.array([0.0, -1.0, 0.0])
        return d / np.linalg.norm(d) * s
    return -fd * s  # push, and every other mode value


def pin_mouse_modes():
    def run(pkg, orc):
        x = f32([0.1, 0.6, 0.05])
        v = f32([0.2, -0.1, 0.3])
        mp = f32([0.15, 0.68, 0.02])
        axis = f32([0.3, -0.5, 1.0])
        for mode in (0, 1, 2, 3, 4, 7):
            for radius in (0.5, 0.05):  # inside, then outside the radius
                m = mouse_struct(pkg, mp, radius, 30.0, mode, axis)
                o, p, _, _ = make_oracle(pkg, orc, [x], [v], mouse=m)
                o.stage(10)
                out = o.particles()[0]
                a = mouse_expected(mode, mp - x, f32(radius), 30.0, axis)
                if radius == 0.5:
                    assert np.linalg.norm(a) > 1.0
                v1 = v + p.dt * a
                np.testing.assert_allclose(out[4:7], v1, rtol=1e-12, atol=1e-13)
                np.testing.assert_allclose(out[0:3], x + p.dt * v1, rtol=1e-12, atol=1e-13)
        # inactive and coincident cursors do nothing
        for m in (mouse_struct(pkg, mp, 0.5, 30.0, 1, axis, active=0),
                  mouse_struct(pkg, x, 0.5, 30.0, 1, axis)):
            o, p, _, _ = make_oracle(pkg, orc, [x], [v], mouse=m)
            o.stage(10)
            np.testing.assert_allclose(o.particles()[0, 4:7], v, rtol=0, atol=0)
    return run


# ---------------------------------------------------------------- rigid body
def rot(ax, az):
    cx, sx, cz, sz = np.cos(ax), np.sin(ax), np.cos(az), np.sin(az)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return f32(Rz @ Rx)


BODY_C = f32([0.05, 0.55, -0.03])
BODY_V = f32([0.1, 0.0, -0.2])
BODY_HE = float(np.float32(0.1))
BODY_K = 200.0
SDF_COEF = np.array([0.25, 0.75, -0.5])  # planar SDF: exact in f32 at resolution 32


def body_struct(pkg, shape, R):
    b = pkg.abi.GpuRigidBody()
    b.position[0], b.position[1], b.position[2] = BODY_C
    b.velocity[0], b.velocity[1], b.velocity[2] = BODY_V
    b.half_extent, b.stiffness, b.shape, b.is_active = BODY_HE, BODY_K, shape, 1
    for k in range(3):
        b.rot_row0[k], b.rot_row1[k], b.rot_row2[k] = R[0, k], R[1, k], R[2, k]
    return b


def planar_sdf(res=32):
    """SDF texture (x fastest) of the plane c . (2 uvw - 1): trilinear
    sampling reproduces a linear function exactly away from the edges."""
    c = (2 * np.arange(res) + 1) / res - 1.0
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    v = SDF_COEF[0] * x + SDF_COEF[1] * y + SDF_COEF[2] * z
    assert np.all(v.astype(np.float32) == v)
    return v.astype(np.float32)


def shape_sdf(shape, rl, he):
    """(signed distance, outward local normal) of each shape."""
    if shape == 1:
        d = np.linalg.norm(rl)
        return d - he, rl / d
    if shape == 2:
        rd = np.hypot(rl[0], rl[2])
        d_rad, d_cap = rd - he, abs(rl[1]) - he
        if d_rad > d_cap:
            return d_rad, np.array([rl[0] / rd, 0.0, rl[2] / rd])
        return d_cap, np.array([0.0, np.sign(rl[1]), 0.0])
    if shape == 3:
        minor = he * 0.3
        xz = np.hypot(rl[0], rl[2])
        q = np.array([xz - he, rl[1]])
        ql = np.linalg.norm(q)
        return ql - minor, np.array([q[0] / ql * rl[0] / xz, q[1] / ql, q[0] / ql * rl[2] / xz])
    if shape == 4:
        return SDF_COEF @ rl, SDF_COEF / np.linalg.norm(SDF_COEF)
    d = np.abs(rl) - he
    k = int(np.argmax(d))
    nrm = np.zeros(3)
    nrm[k] = np.sign(rl[k])
    return d.max(), nrm


# local positions per shape: (each) away from every branch boundary
SHAPE_POINTS = {
    0: [[0.12, 0.03, -0.02], [0.05, -0.09, 0.02], [-0.03, 0.02, -0.14]],
    1: [[0.06, 0.07, -0.03], [-0.08, -0.02, 0.09], [0.01, -0.11, 0.02]],
    2: [[0.09, 0.02, 0.05], [0.02, 0.13, -0.03], [0.01, -0.12, 0.02]],
    3: [[0.1, 0.05, 0.04], [-0.07, -0.02, -0.06], [0.02, 0.01, 0.12]],
    4: [[0.02, 0.03, -0.01], [-0.05, -0.02, 0.03], [0.08, 0.04, 0.06]],
}
SHAPE_VELS = [[-0.4, 0.1, 0.2], [0.3, 0.2, -0.1], [0.0, -0.5, 0.4]]


def rigid_case(pkg, orc, shape):
    R = rot(-0.35, 0.6)  # rows: world -> body frame
    b = body_struct(pkg, shape, R)
    rls = np.array(SHAPE_POINTS[shape])
    wps = BODY_C + rls @ R  # R^T rl
    vels = np.array(SHAPE_VELS)
    o, p, _, _ = make_oracle(pkg, orc, wps, vels, rigid_body=b,
                             sdf=planar_sdf() if shape == 4 else None)
    o.stage(10)
    out = o.particles()
    x, v = f32(wps), f32(vels)
    acc = np.zeros(8, np.int64)
    ir = 0.7 * p.kernel_radius
    hits = 0
    for i in range(len(x)):
        wr = x[i] - BODY_C
        rl = R @ wr
        sdf, ln = shape_sdf(shape, rl, BODY_HE)
        assert sdf < ir
        hits += 1
        nw = R.T @ ln
        t = (ir - sdf) / ir
        pen = nw * BODY_K * t * t
        vn = (v[i] - BODY_V) @ nw
        damp = -nw * vn * 5.0 if vn < 0 else np.zeros(3)
        a = pen + damp
        v1 = v[i] + p.dt * a
        np.testing.assert_allclose(out[i, 4:7], v1, rtol=1e-11, atol=1e-11)
        np.testing.assert_allclose(out[i, 0:3], x[i] + p.dt * v1, rtol=1e-12, atol=1e-12)
        re = -a * p.mass
        tq = np.cross(wr, re)
        for k, val in enumerate(list(re * 1000.0) + [None] + list(tq * 1000.0)):
            if val is None:
                continue
            frac = abs(val - np.trunc(val))
            # no rounding-sensitive truncation (|val| < 1/2 truncates to 0 either way)
            assert abs(val) < 0.5 or 1e-6 < frac < 1 - 1e-6
            acc[k] += int(np.trunc(val))
        acc[3] += 1
    assert hits == len(x)
    np.testing.assert_array_equal(o.accum()[:7], acc[:7])


def pin_rigid_shapes():
    def run(pkg, orc):
        for shape in range(5):
            rigid_case(pkg, orc, shape)
    return run


def pin_rigid_out_of_range():
    def run(pkg, orc):
        R = rot(-0.35, 0.6)
        for shape in range(5):
            b = body_struct(pkg, shape, R)
            x = BODY_C + np.array([0.0, 0.2, 0.0]) @ R  # 0.1 beyond every surface
            o, p, _, _ = make_oracle(pkg, orc, [x], [[0.1, 0.0, 0.0]], rigid_body=b,
                                     sdf=planar_sdf() if shape == 4 else None)
            o.stage(10)
            np.testing.assert_allclose(o.particles()[0, 4:7], f32([0.1, 0, 0]), atol=0)
            assert np.all(o.accum() == 0)
    return run


PINS = {
    "pcisph_pair_repels": pin_pcisph_pair_repels(),
    "pcisph_triple_asymmetric": pin_pcisph_triple_asymmetric(),
    "pcisph_cluster_momentum": pin_pcisph_cluster_momentum(),
    "pcisph_cfl_clamp": pin_pcisph_cfl_clamp(),
    "separation_adjacent_slots": pin_separation_adjacent_slots(),
    "separation_far_slots": pin_separation_far_slots_with_near_pressure(),
    "walls_upright": pin_walls((0.0, 0.0)),
    "walls_tilted": pin_walls((0.3, -0.2)),
    "wall_corner": pin_wall_corner(),
    "backstop_upright": pin_backstop((0.0, 0.0)),
    "backstop_tilted": pin_backstop((0.3, -0.2)),
    "mouse_modes": pin_mouse_modes(),
    "rigid_shapes": pin_rigid_shapes(),
    "rigid_out_of_range": pin_rigid_out_of_range(),
}


@pytest.mark.parametrize("name", sorted(PINS))
def test_term_pin(pkg, orc, name):
    PINS[name](pkg, orc)
