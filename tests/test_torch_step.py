"""Pins of the float64 torch restatement (tests/torch_step.py) that the GPU
adjoint is compared with: its forward equals the FP64 oracle, and its
autograd gradient equals the oracle's central differences.

The finite differences use eps = 1e-8 on positions / velocities of O(1):
the benchmark's dynamics are violent (speeds reach tens of m/s within a
step), so a larger step crosses discrete events (the hard wall backstop
reflecting a particle, a pair entering the 0.1 h close range) within the
difference (at 1e-6 in the 1000-particle relaxed block, at 1e-7 in the
compressed one); the difference is also checked against eps = 1e-9 so a
crossing would show as disagreement between the two.
"""
import numpy as np
import pytest
import torch

import torch_step as ts


def _setup(pkg, orc, sc):
    p, c, g = pkg.scenes.make_params(sc, pkg.abi, pkg.derive_sph_params,
                                     pkg.derive_container_geometry)
    grid = orc.grid_for(sc.h, sc.bounds_extent, sc.n, pkg.abi)
    return p, c, g, ts.Uniforms(p, grid, c, g, sc.pcisph_iterations)


def _oracle(pkg, orc, sc, parts, steps, p, c, g):
    o = orc.from_scene(sc, pkg.abi, p, c, g, fp64=True, particles=parts)
    o.step(steps)
    out = o.particles()
    o.close()
    return out


SCENES = {
    "relaxed": lambda pkg: pkg.scenes.relaxed_block(8, 41),
    "compressed": lambda pkg: pkg.scenes.dam_break_block(8, 21),
}


@pytest.mark.parametrize("name", list(SCENES))
def test_torch_forward_equals_fp64_oracle(pkg, orc, name):
    sc = SCENES[name](pkg)
    p, c, g, U = _setup(pkg, orc, sc)
    base = sc.particles.astype(np.float64)
    ref = _oracle(pkg, orc, sc, base, 2, p, c, g)
    with torch.no_grad():
        x, v = ts.rollout(torch.tensor(base[:, 0:3]), torch.tensor(base[:, 4:7]), U, 2)
    scale = np.abs(ref[:, 4:7]).max()
    assert np.abs(x.numpy() - ref[:, 0:3]).max() < 1e-10
    assert np.abs(v.numpy() - ref[:, 4:7]).max() < 1e-10 * max(scale, 1.0)


@pytest.mark.parametrize("name", list(SCENES))
def test_torch_gradient_equals_oracle_fd(pkg, orc, name):
    sc = SCENES[name](pkg)
    p, c, g, U = _setup(pkg, orc, sc)
    base = sc.particles.astype(np.float64)
    n = sc.n
    rng = np.random.default_rng(5)
    lam_x, lam_v = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    x = torch.tensor(base[:, 0:3], requires_grad=True)
    v = torch.tensor(base[:, 4:7], requires_grad=True)
    xo, vo = ts.step(x, v, U)
    ((xo * torch.tensor(lam_x)).sum() + (vo * torch.tensor(lam_v)).sum()).backward()
    for trial in range(2):
        dx, dv = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
        ad = np.sum(x.grad.numpy() * dx) + np.sum(v.grad.numpy() * dv)
        fds = []
        for eps in (1e-8, 1e-9):
            vals = []
            for sgn in (1.0, -1.0):
                b = base.copy()
                b[:, 0:3] += sgn * eps * dx
                b[:, 4:7] += sgn * eps * dv
                out = _oracle(pkg, orc, sc, b, 1, p, c, g)
                vals.append(np.sum(lam_x * out[:, 0:3]) + np.sum(lam_v * out[:, 4:7]))
            fds.append((vals[0] - vals[1]) / (2 * eps))
        assert abs(fds[0] - fds[1]) <= 1e-5 * abs(fds[0]), (name, trial, fds)
        assert abs(ad - fds[0]) <= 1e-5 * abs(fds[0]), (name, trial, ad, fds)


def test_torch_parameter_gradients_equal_oracle_fd(pkg, orc):
    sc = SCENES["relaxed"](pkg)
    p, c, g, U = _setup(pkg, orc, sc)
    base = sc.particles.astype(np.float64)
    rng = np.random.default_rng(6)
    lam_x = rng.standard_normal((sc.n, 3))
    visc = torch.tensor(U.visc, dtype=torch.float64, requires_grad=True)
    st = torch.tensor(U.st, dtype=torch.float64, requires_grad=True)
    grav = torch.tensor(U.gravity, requires_grad=True)
    xo, _ = ts.step(torch.tensor(base[:, 0:3]), torch.tensor(base[:, 4:7]), U,
                    visc=visc, st=st, grav=grav)
    (xo * torch.tensor(lam_x)).sum().backward()
    ad = [visc.grad.item(), st.grad.item()] + grav.grad.tolist()
    # the oracle's uniforms are f32 structs: steps of 1e-3 relative keep the
    # f32 rounding of the perturbed value (6e-8 relative) far below the step
    for k in range(5):
        vals = []
        for sgn in (1.0, -1.0):
            pp, cc, gg = pkg.scenes.make_params(sc, pkg.abi, pkg.derive_sph_params,
                                                pkg.derive_container_geometry)
            if k == 0:
                pp.viscosity = np.float32(pp.viscosity * (1 + sgn * 1e-3))
                hstep = pp.viscosity
            elif k == 1:
                pp.surface_tension = np.float32(pp.surface_tension * (1 + sgn * 1e-3))
                hstep = pp.surface_tension
            else:
                gg.direction[k - 2] = np.float32(gg.direction[k - 2] + sgn * 1e-2)
                hstep = gg.direction[k - 2]
            out = _oracle(pkg, orc, sc, base, 1, pp, cc, gg)
            vals.append((np.sum(lam_x * out[:, 0:3]), float(hstep)))
        fd = (vals[0][0] - vals[1][0]) / (vals[0][1] - vals[1][1])
        assert abs(ad[k] - fd) <= 2e-3 * abs(fd) + 1e-6, (k, ad[k], fd)
