"""Dense float64 torch restatement of one SPH step, for gradient checks.

Test infrastructure only: a differentiable copy of the step the oracle
restates (fluid_oracle.cpp, which cites the WGSL line by line), written with
all-pairs tensors so torch.autograd supplies reference vector-Jacobian
products per particle.  Sized for about a thousand particles (n^2 pair
tensors).  Discrete decisions follow the oracle: cell coordinates in f32,
stable sort by cell then original index, the 27-cell adjacency of the
sorted positions, close pairs (r < 0.1 h) with the slot-difference
direction, and every branch (wall layer, approach, clamp, backstop) taken
on the float64 values.

Mouse force and rigid body are not modelled (the rollout refuses them).
"""
from __future__ import annotations

import numpy as np
import torch

PI = float(np.float32(3.14159265359))


class Uniforms:
    """Float64 copies of the f32 parameter structs, as the oracle's
    set_uniforms promotes them."""

    def __init__(self, params, grid, cont, grav, iterations: int):
        f = float
        self.h = f(params.kernel_radius)
        self.h2 = f(params.kernel_radius_sq)
        self.h6 = f(params.kernel_radius_pow6)
        self.h9 = f(params.kernel_radius_pow9)
        self.mass = f(params.mass)
        self.rho0 = f(params.rest_density)
        self.near_k = f(params.near_stiffness)
        self.visc = f(params.viscosity)
        self.dt = f(params.dt)
        self.st = f(params.surface_tension)
        self.delta = f(params.pcisph_delta)
        self.eps = f(params.xsph_epsilon)
        self.gsize = (grid.grid_size_x, grid.grid_size_y, grid.grid_size_z)
        self.inv_cell = np.float32(grid.inv_cell_size)
        self.origin = (np.float32(grid.grid_origin_x), np.float32(grid.grid_origin_y),
                       np.float32(grid.grid_origin_z))
        self.half = np.array([cont.half_width, cont.half_height, cont.half_depth], np.float64)
        self.cy = f(cont.center_y)
        self.fw = np.array([list(cont.forward_row0)[:3], list(cont.forward_row1)[:3],
                            list(cont.forward_row2)[:3]], np.float64)
        self.iv = np.array([list(cont.inverse_row0)[:3], list(cont.inverse_row1)[:3],
                            list(cont.inverse_row2)[:3]], np.float64)
        self.wall_k = f(cont.wall_stiffness)
        self.damping = f(cont.damping)
        self.gravity = np.array(list(grav.direction)[:3], np.float64)
        self.iters = int(iterations)


def sort_order(x: np.ndarray, U: Uniforms):
    """(sorted -> original permutation, clamped cell coordinates per sorted
    slot) with the oracle's f32 cell decision and within-cell order (by the
    float32 x's order key, then by index)."""
    xf = x.astype(np.float32)
    c = np.floor((xf - np.array(U.origin, np.float32)) * U.inv_cell).astype(np.int64)
    g = np.array(U.gsize)
    cc = np.clip(c, 0, g - 1)
    cell = cc[:, 0] + cc[:, 1] * g[0] + cc[:, 2] * g[0] * g[1]
    xb = np.ascontiguousarray(xf[:, 0]).view(np.uint32)
    xkey = np.where(xb >> 31, ~xb, xb | np.uint32(0x80000000))
    perm = np.lexsort((np.arange(len(x)), xkey, cell))
    return perm, c[perm], cc[perm]


def _walls_gamma(xs, U, h):
    iv = torch.as_tensor(U.iv)
    p = torch.stack([xs[:, 0], xs[:, 1] - U.cy, xs[:, 2]], 1)
    lp = p @ iv.T
    half = torch.as_tensor(U.half)
    dist = torch.cat([lp + half, half - lp], 1)  # (n, 6): lower x,y,z, upper x,y,z
    f = torch.where(dist < h, 0.5 + 0.5 * torch.clamp(dist / h, 0.0, 1.0),
                    torch.ones_like(dist))
    return torch.clamp(torch.prod(f, 1), min=0.1)


def step(x: torch.Tensor, v: torch.Tensor, U: Uniforms, visc=None, st=None, grav=None,
         trace: dict | None = None):
    """One step: canonical (x, v) -> canonical (x', v').  visc / st / grav
    may be 0-d / (3,) tensors (for parameter gradients); `trace` receives
    the intermediates (sorted order) when given."""
    n = x.shape[0]
    h, h2, m = U.h, U.h2, U.mass
    visc = torch.tensor(U.visc, dtype=torch.float64) if visc is None else visc
    st = torch.tensor(U.st, dtype=torch.float64) if st is None else st
    grav = torch.as_tensor(U.gravity) if grav is None else grav
    perm_np, cells, _ = sort_order(x.detach().numpy(), U)
    perm = torch.as_tensor(perm_np)
    xs, us = x[perm], v[perm]
    # 27-cell adjacency on the unclamped cells, clamped to [-2, g+1] as the
    # sweep does, against the particles' clamped cells
    g = np.array(U.gsize)
    ci = np.clip(cells, -2, g + 1)
    cj = np.clip(cells, 0, g - 1)
    adj = torch.as_tensor(np.all(np.abs(ci[:, None, :] - cj[None, :, :]) <= 1, axis=2))
    ps = 315.0 / (64.0 * PI * U.h9)
    ns = 15.0 / (PI * U.h6)
    gs = 45.0 / (PI * U.h6)
    aks = 32.0 / (PI * U.h9)
    h8 = U.h6 * h * h
    nscale = 945.0 / (32.0 * PI * h8)
    eye = torch.eye(n, dtype=torch.bool)

    def pairs(pts):
        rv = pts[:, None, :] - pts[None, :, :]
        r2 = (rv * rv).sum(-1)
        inr = adj & (r2 < h2)
        pair = inr & (r2 > 1e-12)
        r2s = torch.where(pair, r2, torch.ones_like(r2))
        r = torch.sqrt(r2s)
        return rv, r2, inr, pair, r

    # density (sph_density_3d_grid.wgsl:111-194)
    rv, r2, inr, pair, r = pairs(xs)
    d = torch.where(inr, h2 - r2, torch.zeros_like(r2))
    S = (m * ps * d ** 3).sum(1)
    rr = torch.where(pair, r, torch.zeros_like(r))  # self pairs: r = 0 exactly
    near = torch.where(inr, m * ns * (h - rr) ** 3, torch.zeros_like(r2)).sum(1)
    dp = torch.where(pair, d, torch.zeros_like(d))
    nrm = nscale * ((m / U.rho0) * (dp ** 2)[..., None] * rv).sum(1)
    rho = S / _walls_gamma(xs, U, h)
    # XSPH (xsph.wgsl:81-138)
    if U.eps > 0:
        W = torch.where(pair, ps * d ** 3, torch.zeros_like(d))
        f = m * W / torch.clamp(rho, min=1.0)[None, :]
        w = us + U.eps * (f[..., None] * (us[None, :, :] - us[:, None, :])).sum(1)
    else:
        w = us
    # force (sph_force_3d_grid.wgsl:105-227)
    di = torch.clamp(rho, min=1.0)
    ndi = torch.clamp(near, min=1.0)
    md = h * 0.1
    close = pair & (r < md)
    slot = torch.arange(n, dtype=torch.float64)
    idx = slot[:, None] - slot[None, :]
    cd = torch.stack([torch.sin(idx), torch.cos(2 * idx), torch.sin(3 * idx)], -1)
    cd = cd / torch.linalg.norm(cd, dim=-1, keepdim=True)
    dirv = torch.where(close[..., None], cd, -rv / r[..., None])
    grad = gs * (h - r) ** 2
    sp = -m * (U.near_k * ndi[:, None] + U.near_k * ndi[None, :]) / (2.0 * di[:, None] * ndi[None, :])
    coef = sp * grad
    coef = coef + torch.where(close, -200.0 * (1.0 - r / md), torch.zeros_like(r))
    vij = w[:, None, :] - us[None, :, :]
    vdx = (vij * rv).sum(-1)
    mu = h * vdx / (r2 + 0.01 * h2)
    avg = (di[:, None] + di[None, :]) * 0.5
    coef = coef + torch.where(vdx < 0, m * visc * mu * grad / avg, torch.zeros_like(r))
    K = 2.0 * U.rho0 / (di[:, None] + di[None, :])
    coh = torch.where(r >= h * 0.5, aks * (h - r) ** 3 * r ** 3,
                      aks * (2.0 * (h - r) ** 3 * r ** 3 - U.h6 / 64.0))
    coef = coef + st * m * K * coh
    term = coef[..., None] * dirv - (st * m * K)[..., None] * (nrm[:, None, :] - nrm[None, :, :])
    term = torch.where(pair[..., None], term, torch.zeros_like(term))
    a = term.sum(1) + grav
    if trace is not None:
        trace.update(perm=perm, rho=rho, near=near, nrm=nrm, w=w, a=a, pair=pair)
    # predict + solve (pcisph_predict.wgsl:50-73, pcisph_solve.wgsl:172-208)
    dt = U.dt
    V = w + dt * a
    X = xs + dt * V
    P = torch.zeros(n, dtype=torch.float64)
    Q = torch.zeros(n, dtype=torch.float64)
    max_a = 0.25 * h / (dt * dt)
    for _ in range(U.iters):
        rvx, r2x, inrx, pairx, rx = pairs(X)
        dx = torch.where(inrx, h2 - r2x, torch.zeros_like(r2x))
        rho_p = (m * ps * dx ** 3).sum(1)
        c = torch.where(pairx, m * (Q[:, None] + Q[None, :]) * gs * (h - rx) ** 2 / rx,
                        torch.zeros_like(rx))
        ap = (c[..., None] * rvx).sum(1)
        amag2 = (ap * ap).sum(1)
        scale = torch.where(amag2 > max_a * max_a,
                            max_a / torch.sqrt(torch.clamp(amag2, min=1e-300)),
                            torch.ones_like(amag2))
        ap = ap * scale[:, None]
        P = P + U.delta * torch.clamp(rho_p - U.rho0, min=0.0)
        V = w + dt * (a + ap)
        X = xs + dt * V
        Q = P / torch.clamp(rho_p, min=1.0) ** 2
        if trace is not None:
            trace.setdefault("iters", []).append(dict(rho_p=rho_p, ap=ap, scale=scale, P=P, V=V))
    # finalize + integrate (pcisph_finalize.wgsl:50-74, sph_integrate_3d.wgsl)
    iv = torch.as_tensor(U.iv)
    fw = torch.as_tensor(U.fw)
    half = torch.as_tensor(U.half)
    cyv = torch.tensor([0.0, U.cy, 0.0], dtype=torch.float64)
    lp = (xs - cyv) @ iv.T
    lv = V @ iv.T
    bl = h * 0.4
    wd = 8.0
    wa = torch.zeros_like(lp)
    dlo = lp + half
    t = 1.0 - dlo / bl
    wa = wa + torch.where(dlo < bl, U.wall_k * t * t - torch.where(lv < 0, lv * wd * t,
                                                                    torch.zeros_like(t)),
                          torch.zeros_like(t))
    dhi = half - lp
    t = 1.0 - dhi / bl
    wa = wa + torch.where(dhi < bl, -U.wall_k * t * t - torch.where(lv > 0, lv * wd * t,
                                                                     torch.zeros_like(t)),
                          torch.zeros_like(t))
    V2 = V + dt * (wa @ fw.T)
    Xn = xs + dt * V2
    lpn = (Xn - cyv) @ iv.T
    lvn = V2 @ iv.T
    lo, hi = lpn < -half, lpn > half
    lpn = torch.where(lo, -half.expand_as(lpn), torch.where(hi, half.expand_as(lpn), lpn))
    lvn = torch.where(lo, torch.abs(lvn) * U.damping,
                      torch.where(hi, -torch.abs(lvn) * U.damping, lvn))
    xo = lpn @ fw.T + cyv
    vo = lvn @ fw.T
    inv = torch.empty_like(perm)
    inv[perm] = torch.arange(n)
    return xo[inv], vo[inv]


def rollout(x, v, U, steps, **kw):
    for _ in range(steps):
        x, v = step(x, v, U, **kw)
    return x, v
