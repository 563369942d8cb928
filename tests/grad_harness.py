"""Gradient-check harness over the rollout C ABI.

`SphRollout` is a torch.autograd.Function whose forward is
fluid_forward_rollout and whose backward is fluid_backward_rollout; it
exists so gradient checks and the benchmark can compose the simulator with
a torch loss.  `splat_density` / `splat_mse` are that loss: the particles'
mass splatted trilinearly (cloud-in-cell) onto a cell-centred 32^3 grid over
the container, and the mean squared difference to a target field.

Inputs are the initial velocities (N x 3) and the scalars viscosity and
surface tension; the simulator's `stiffness` parameter is read by no stage
of the PCISPH step (the pressure comes from pcisph_delta), so its gradient
is exactly zero and is reported as such.
"""
from __future__ import annotations

import numpy as np
import torch


def splat_density(pos: torch.Tensor, lo, hi, res: int = 32, mass: float = 1.0) -> torch.Tensor:
    """(res, res, res) density of point masses at `pos` (N x 3) on the box
    [lo, hi], trilinear weights on cell centres, clamped at the faces."""
    lo = torch.as_tensor(lo, dtype=pos.dtype, device=pos.device)
    hi = torch.as_tensor(hi, dtype=pos.dtype, device=pos.device)
    size = (hi - lo) / res
    g = (pos - lo) / size - 0.5
    g = torch.clamp(g, 0.0, res - 1.0)
    i0 = torch.clamp(torch.floor(g).long(), 0, res - 2)
    f = g - i0.to(pos.dtype)
    out = torch.zeros(res * res * res, dtype=pos.dtype, device=pos.device)
    for dx in (0, 1):
        wx = f[:, 0] if dx else 1.0 - f[:, 0]
        for dy in (0, 1):
            wy = f[:, 1] if dy else 1.0 - f[:, 1]
            for dz in (0, 1):
                wz = f[:, 2] if dz else 1.0 - f[:, 2]
                idx = ((i0[:, 0] + dx) * res + (i0[:, 1] + dy)) * res + (i0[:, 2] + dz)
                out = out.index_add(0, idx, wx * wy * wz)
    return out.reshape(res, res, res) * (mass / torch.prod(size))


def container_box(scene):
    w, h, d = scene.box
    return (-w / 2, scene.floor_y, -d / 2), (w / 2, scene.floor_y + h, d / 2)


def splat_mse(pos, target, lo, hi, res=32):
    return torch.mean((splat_density(pos, lo, hi, res) - target) ** 2)


class SphRollout(torch.autograd.Function):
    """pos_T, vel_T = SphRollout.apply(vel0, visc, st, ctx_dict)

    ctx_dict holds: sim (SphSimulation3DGrid), particles (N x 16 float32
    initial records; their velocities are replaced by vel0), params
    (GpuSphParams3D, viscosity / surface tension replaced), steps, every."""

    @staticmethod
    def forward(ctx, vel0, visc, st, cfg):
        sim = cfg["sim"]
        parts = np.array(cfg["particles"], dtype=np.float32, copy=True)
        parts[:, 4:7] = vel0.detach().cpu().numpy()
        p = cfg["params"]
        p.viscosity = float(visc)
        p.surface_tension = float(st)
        sim.update_sph_params(p)
        sim.write_particles(parts)
        sim.forward_rollout(cfg["steps"], cfg["every"])
        out = sim.read_particles()
        ctx.cfg = cfg
        ctx.dev = vel0.device
        pos = torch.as_tensor(out[:, 0:3].copy(), device=vel0.device)
        vel = torch.as_tensor(out[:, 4:7].copy(), device=vel0.device)
        return pos, vel

    @staticmethod
    def backward(ctx, g_pos, g_vel):
        sim = ctx.cfg["sim"]
        gp = None if g_pos is None else g_pos.detach().float().cpu().numpy()
        gv = None if g_vel is None else g_vel.detach().float().cpu().numpy()
        _, dv0, dparams = sim.backward_rollout(gp, gv)
        d = torch.as_tensor(dv0, device=ctx.dev)
        return d, torch.tensor(float(dparams[0, 0])), torch.tensor(float(dparams[0, 1])), None
