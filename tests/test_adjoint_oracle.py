"""The FP64 torch restatement used as the gradient oracle (oracle/visco2d_torch.py).

Pinned here against the C oracle's FP64 forward (every wavefield, memory
variable and trace at tau scale 1), and its autograd gradient against
central finite differences of the same FP64 forward.  The product's chain
rule (taichi_sim_amd.adjoint.coefficients) must form the same coefficients
as the oracle's own restatement.
"""
import numpy as np
import pytest
import torch

import oracle
from oracle.visco2d_torch import Visco2D
from taichi_sim_amd import workloads
from taichi_sim_amd.adjoint import coefficients
from taichi_sim_amd.backend import attenuation_from_config
from taichi_sim_amd.sim_support.simulator import Simulator


def _setup(n, steps, voids):
    cfg = workloads.visco_2d_array(n=n, steps=steps, n_elem=4, n_voids=voids)
    sim = Simulator(cfg)
    return sim, attenuation_from_config(cfg)


def _t(a):
    return torch.tensor(np.asarray(a, dtype=np.float32).astype(np.float64))


def test_torch_forward_matches_c_oracle_fp64():
    sim, att = _setup(64, 240, 1)
    o = oracle.Oracle(sim, variant="wgsl", att=att, precision="f64").run(240)
    m = Visco2D(sim, att)
    tr, st = m.forward(_t(sim._rho_grid_vx), _t(sim._cp_grid_vx), _t(sim._cs_grid_vx),
                       torch.ones(m.nx, m.ny, dtype=torch.float64), return_state=True)
    for k in ("vx", "vy", "sxx", "syy", "sxy"):
        assert oracle.rel_l2(st[k].numpy(), o.f[k]) < 1e-12, k
    assert oracle.rel_l2(st["rxx"].numpy(), o.r[0]) < 1e-12
    assert oracle.rel_l2(st["rxy"].numpy(), o.r[3]) < 1e-12
    assert np.abs(o.traces).max() > 0
    assert oracle.rel_l2(tr.numpy(), o.traces) < 1e-12


def test_coefficients_agree_with_oracle():
    sim, att = _setup(48, 10, 2)
    m = Visco2D(sim, att)
    rho, cp, cs = _t(sim._rho_grid_vx), _t(sim._cp_grid_vx), _t(sim._cs_grid_vx)
    for a, b in zip(coefficients(rho, cp, cs), m.coefficients(rho, cp, cs)):
        assert torch.equal(a, b)


@pytest.mark.parametrize("which", ["rho", "cp", "cs", "tscale"])
def test_autograd_matches_central_differences(which):
    sim, att = _setup(48, 150, 0)
    m = Visco2D(sim, att)
    base = [_t(sim._rho_grid_vx), _t(sim._cp_grid_vx), _t(sim._cs_grid_vx),
            torch.ones(m.nx, m.ny, dtype=torch.float64)]
    obs = m.forward(*base).detach()
    g = torch.Generator().manual_seed(11)
    params = [p * (1 + 0.03 * torch.randn(p.shape, generator=g, dtype=torch.float64)) for p in base]
    params = [p.clone().requires_grad_(True) for p in params]
    J = 0.5 * ((m.forward(*params) - obs) ** 2).sum()
    J.backward()
    i = ["rho", "cp", "cs", "tscale"].index(which)
    v = torch.randn(base[0].shape, generator=g, dtype=torch.float64)
    eps = 1e-6 * float(params[i].detach().abs().max())

    def Jat(s):
        ps = [p.detach().clone() for p in params]
        ps[i] = ps[i] + s * eps * v
        return float(0.5 * ((m.forward(*ps) - obs) ** 2).sum())

    fd = (Jat(1) - Jat(-1)) / (2 * eps)
    ad = float((params[i].grad * v).sum())
    assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1e-30), (ad, fd)
