"""Pins of the oracle's marching-cubes density field (CPU only):
mc_density.wgsl and mc_anisotropy.wgsl.

The reference holds no test or fixture for these passes, so they are pinned
by what the mathematics fixes: the poly6 splat of one particle in closed
form and its integral (the kernel is normalised, so the field integrates to
the particle count); the wall factor and the clip sentinel; the ellipsoid
fit of a cubic lattice (isotropic: G = I / h), of a single square-lattice
sheet (eigenvalues in-plane equal, normal axis clamped by k_r: axis scales
4^(1/3), 4^(1/3), 4^(-2/3) with the sheet normal on the short axis) and of a
sparse neighbourhood (anisotropy faded out); and the ellipsoid splat with
isotropic records against the poly6 splat.
"""
import numpy as np
import pytest

from test_spray_oracle import H, F, container, sorted_state

PEAK = 315.0 / (64.0 * np.pi * H ** 3)


def mc_params(pkg, grid_size=48, lo=(-0.3, 0.2, -0.3), cell=None, h=H):
    g = pkg.abi.GpuMcGridParams()
    cs = cell if cell is not None else 0.6 / grid_size
    for k in range(3):
        g.grid_min[k] = lo[k]
        g.grid_max[k] = lo[k] + grid_size * cs
    g.grid_size, g.cell_size, g.kernel_radius = grid_size, cs, h
    return g


def voxel_centres(g):
    G = g.grid_size
    idx = np.arange(G, dtype=np.float32)
    c = (idx + F(0.5)) * F(g.cell_size)
    z, y, x = np.meshgrid(c + F(g.grid_min[2]), c + F(g.grid_min[1]), c + F(g.grid_min[0]),
                          indexing="ij")
    return np.stack([x, y, z], -1).astype(np.float64)


def test_single_particle_splat_closed_form(pkg, orc):
    """One particle away from the walls: every voxel holds
    315 / (64 pi h^9) (h^2 - r^2)^3 inside the support and 0 outside."""
    p = np.array([0.013, 0.517, -0.021])
    recs, cnt, st, grid, _ = sorted_state(pkg, orc, [p])
    g = mc_params(pkg)
    a = pkg.abi.default_aniso_params(H, enabled=False)
    f = orc.mc_density(recs, cnt, st, grid, g, a, container(pkg))
    x = voxel_centres(g)
    r2 = ((x - p) ** 2).sum(-1)
    want = np.where(r2 < H * H, 315 / (64 * np.pi * H ** 9) * np.clip(H * H - r2, 0, None) ** 3, 0)
    assert np.abs(f - want).max() <= 2e-5 * want.max()
    assert (f > 0).sum() == (want > 0).sum()


def test_field_integrates_to_particle_count(pkg, orc):
    """poly6 integrates to 1 over its ball, so a block of N particles far
    from the walls integrates to N (midpoint rule, cell = h / 8)."""
    rng = np.random.default_rng(3)
    pts = rng.uniform(-0.08, 0.08, (40, 3)) + [0.0, 0.55, 0.0]
    recs, cnt, st, grid, _ = sorted_state(pkg, orc, pts)
    g = mc_params(pkg, grid_size=64, lo=(-0.2, 0.35, -0.2), cell=0.4 / 64)
    a = pkg.abi.default_aniso_params(H, enabled=False)
    f = orc.mc_density(recs, cnt, st, grid, g, a, container(pkg))
    total = f.astype(np.float64).sum() * g.cell_size ** 3
    assert abs(total / len(pts) - 1) < 5e-3


def test_wall_factor_and_clip_sentinel(pkg, orc):
    """Within h of a wall the density is divided by 0.5 + 0.5 d / h per
    wall (at least 0.15 overall); with clipping on, voxels outside the box
    hold -1."""
    cont = container(pkg, box=(1.6, 1.2, 0.8), floor_y=0.0)
    p = np.array([0.0, 0.03, 0.0])  # 0.03 above the floor
    recs, cnt, st, grid, _ = sorted_state(pkg, orc, [p])
    g = mc_params(pkg, grid_size=40, lo=(-0.1, -0.05, -0.1), cell=0.005)
    a = pkg.abi.default_aniso_params(H, enabled=False)
    f = orc.mc_density(recs, cnt, st, grid, g, a, cont)
    x = voxel_centres(g)
    r2 = ((x - p) ** 2).sum(-1)
    raw = np.where(r2 < H * H, 315 / (64 * np.pi * H ** 9) * np.clip(H * H - r2, 0, None) ** 3, 0)
    d = x[..., 1]  # distance to the floor
    gamma = np.where(d < H, 0.5 + 0.5 * np.clip(d / H, 0, 1), 1.0)
    want = raw / np.maximum(gamma, 0.15)
    assert np.abs(f - want).max() <= 3e-5 * want.max()
    cont.clip_enabled = 1
    f2 = orc.mc_density(recs, cnt, st, grid, g, a, cont)
    below = d < 0
    assert below.any() and np.all(f2[below] == -1.0) and np.array_equal(f2[~below], f[~below])


def _lattice(n, s, centre, dims=3):
    ax = (np.arange(n) - (n - 1) / 2) * s
    if dims == 3:
        g = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    else:
        g2 = np.stack(np.meshgrid(ax, ax, indexing="ij"), -1).reshape(-1, 2)
        g = np.column_stack([g2[:, 0], np.zeros(len(g2)), g2[:, 1]])
    return g + centre


def _aniso_of_centre(pkg, orc, pts, centre):
    recs, cnt, st, grid, slot = sorted_state(pkg, orc, pts)
    a = pkg.abi.default_aniso_params(H)
    out = orc.mc_aniso(recs, cnt, st, grid, a)
    k = int(np.argmin(((pts - centre) ** 2).sum(1)))
    r = out[slot[k]].astype(np.float64)
    G = np.array([[r[0], r[1], r[2]], [r[1], r[4], r[5]], [r[2], r[5], r[6]]])
    return G, r[[3, 7, 8]], r[9], r[10]


def test_cubic_lattice_is_isotropic(pkg, orc):
    """A particle inside a cubic lattice sees a covariance proportional to I
    and a zero mean offset: G = I / h, centre unmoved, reach h, amplitude
    the poly6 peak 315 / (64 pi h^3)."""
    c = np.array([0.0, 0.6, 0.0])
    pts = _lattice(9, 0.6 * H, c)
    G, cen, reach, amp = _aniso_of_centre(pkg, orc, pts, c)
    assert np.allclose(G, np.eye(3) / H, rtol=0, atol=2e-4 / H)
    assert np.allclose(cen, c, atol=1e-6)
    assert abs(reach - H) < 1e-5 * H and abs(amp - PEAK) < 1e-5 * PEAK


def test_flat_sheet_axes(pkg, orc):
    """A single square-lattice sheet (normal y): the in-plane stddevs are
    equal and the normal one is 0, clamped to s1 / k_r; volume-normalised
    the axis scales are 4^(1/3), 4^(1/3), 4^(-2/3), so G = P_plane /
    (4^(1/3) h) + n n^T / (4^(-2/3) h) and the reach is 4^(1/3) h."""
    c = np.array([0.0, 0.6, 0.0])
    pts = _lattice(11, 0.6 * H, c, dims=2)
    G, cen, reach, amp = _aniso_of_centre(pkg, orc, pts, c)
    a_in, a_n = 4 ** (1 / 3), 4 ** (-2 / 3)
    want = np.diag([1 / (a_in * H), 1 / (a_n * H), 1 / (a_in * H)])
    assert np.allclose(G, want, rtol=2e-4, atol=2e-4 / H), G * H
    assert abs(reach - a_in * H) < 1e-4 * H
    assert abs(amp - PEAK) < 1e-5 * PEAK  # a product of 1 keeps the peak
    assert np.allclose(cen, c, atol=1e-6)


def test_sparse_neighbourhood_stays_spherical(pkg, orc):
    """Fewer than 4 neighbours within 2h: the anisotropy fades to 0 (the
    smoothstep over 4..12 neighbours), so G = I / h and the centre stays."""
    c = np.array([0.0, 0.6, 0.0])
    pts = np.array([c, c + [0.05, 0, 0], c + [0, 0.03, 0.02]])
    G, cen, reach, amp = _aniso_of_centre(pkg, orc, pts, c)
    assert np.allclose(G, np.eye(3) / H, atol=1e-6 / H)
    assert np.allclose(cen, c, atol=1e-7) and abs(amp - PEAK) < 1e-6 * PEAK


def test_isotropic_ellipsoids_splat_as_poly6(pkg, orc):
    """Records with G = I / h, reach h and amplitude 315 / (64 pi h^3) give
    amplitude (1 - r^2 / h^2)^3 = 315 / (64 pi h^9) (h^2 - r^2)^3: the
    ellipsoid splat equals the poly6 splat."""
    rng = np.random.default_rng(5)
    pts = rng.uniform(-0.1, 0.1, (30, 3)) + [0.0, 0.55, 0.0]
    recs, cnt, st, grid, _ = sorted_state(pkg, orc, pts)
    g = mc_params(pkg, grid_size=40, lo=(-0.2, 0.35, -0.2), cell=0.01)
    iso = pkg.abi.default_aniso_params(H, enabled=False)
    an = pkg.abi.default_aniso_params(H)
    rec = np.zeros((len(pts), 12), np.float32)
    rec[:, 0] = rec[:, 4] = rec[:, 6] = 1.0 / H
    rec[:, 3], rec[:, 7], rec[:, 8] = recs[:, 0], recs[:, 1], recs[:, 2]
    rec[:, 9], rec[:, 10] = H, PEAK
    f_iso = orc.mc_density(recs, cnt, st, grid, g, iso, container(pkg))
    f_an = orc.mc_density(recs, cnt, st, grid, g, an, container(pkg), rec)
    assert np.abs(f_an - f_iso).max() <= 1e-4 * f_iso.max()


def test_edge_of_sheet_centre_moves_toward_neighbours(pkg, orc):
    """At the edge of a sheet the weighted mean offset mu = sum w d / sum w
    (w = 1 - (r / 2h)^3, d = x_j - x_i) points into the sheet; a flat
    neighbourhood has flatness 1, so the centre moves by lambda * mu = 0.9
    mu, capped at 0.4 h."""
    c = np.array([0.0, 0.6, 0.0])
    pts = _lattice(11, 0.6 * H, c, dims=2)
    edge = pts[np.argmin(pts[:, 0] + 1e-3 * np.abs(pts[:, 2]))]  # on the x = min edge
    G, cen, reach, amp = _aniso_of_centre(pkg, orc, pts, edge)
    d = pts - edge
    r = np.linalg.norm(d, axis=1)
    m = r < 2 * H
    w = 1 - (r[m] / (2 * H)) ** 3
    mu = (w[:, None] * d[m]).sum(0) / w.sum()
    shift = cen - edge
    want = 0.9 * mu
    if np.linalg.norm(want) > 0.4 * H:
        want *= 0.4 * H / np.linalg.norm(want)
    assert mu[0] > 0  # into the sheet
    assert np.allclose(shift, want, atol=1e-6), (shift, want)


@pytest.mark.parametrize("counts", [(13, 5, 1), (13, 1, 5), (5, 13, 1), (5, 1, 13),
                                    (1, 13, 5), (1, 5, 13)])
def test_bar_axes_follow_eigenvectors(pkg, orc, counts):
    """A one-layer strip of lattice points (spacing 0.6h; 13, 5 and 1
    points along the three axes, in every assignment to x, y, z, so the
    sort meets every input order): three distinct variances, the smallest 0
    and clamped to the largest stddev / k_r, so the result depends on which
    eigenvalue the sort puts first.  The expected G comes from numpy's eigendecomposition of
    the same weighted covariance (w = 1 - (r / 2h)^3 within 2h), stddevs
    ratio-clamped by k_r, volume-normalised and capped to [0.25, 1.6]:
    G = sum_k e_k e_k^T / (a_k h)."""
    c = np.array([0.0, 0.6, 0.0])
    ax = [(np.arange(k) - (k - 1) / 2) * 0.6 * H for k in counts]
    pts = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3) + c
    centre = pts[np.argmin(((pts - c) ** 2).sum(1))]
    G, cen, reach, amp = _aniso_of_centre(pkg, orc, pts, centre)
    d = pts - centre
    r = np.linalg.norm(d, axis=1)
    m = r < 2 * H
    w = 1 - (r[m] / (2 * H)) ** 3
    mu = (w[:, None] * d[m]).sum(0) / w.sum()
    C = (w[:, None, None] * d[m][:, :, None] * d[m][:, None, :]).sum(0) / w.sum() - np.outer(mu, mu)
    ev, vec = np.linalg.eigh(C)
    order = np.argsort(ev)[::-1]
    ev, vec = ev[order], vec[:, order]
    assert ev[0] > 1.2 * ev[1] > 1.44 * ev[2]  # well separated
    s = np.maximum(np.sqrt(np.maximum(ev, 0)), np.sqrt(ev[0]) / 4.0)
    n_cnt = m.sum()
    t = np.clip((n_cnt - 4) / 8, 0, 1)
    s_eff = t * t * (3 - 2 * t)
    a = np.clip((s / np.prod(s) ** (1 / 3)) ** s_eff, 0.25, 1.6)
    want = sum(np.outer(vec[:, k], vec[:, k]) / (a[k] * H) for k in range(3))
    assert np.allclose(G, want, rtol=2e-4, atol=2e-4 / H), (G * H, want * H)
    assert abs(reach - a.max() * H) < 2e-4 * H
