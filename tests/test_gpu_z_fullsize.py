"""Full-size configurations, checked through size-independent properties
and a truncated-window comparison with the oracle."""
import numpy as np
import pytest

import oracle
from taichi_sim_amd import workloads
from taichi_sim_amd.backend import attenuation_from_config
from taichi_sim_amd.engine import VSESim
from taichi_sim_amd.sim_support.simulator import Simulator
from taichi_sim_amd.sim_support.simulator3d import Simulator3D

pytestmark = pytest.mark.gpu


def test_visco_2d_2048_window_vs_oracle():
    cfg = workloads.visco_2d_array(n=2048, steps=4000)
    sim = Simulator(cfg)
    att = attenuation_from_config(cfg)
    steps = 120
    o = oracle.Oracle(sim, variant="wgsl", att=att).run(steps)
    g = VSESim.from_simulator(sim, variant="wgsl", att=att)
    g.run(steps)
    for k in ("vx", "vy", "sxx", "syy", "sxy"):
        assert oracle.rel_l2(g.field(k), o.f[k]) <= 1e-5, k
    g.close()


def test_visco_3d_512_linearity_in_source():
    """512^3 viscoelastic block: doubling the source doubles the wavefield
    (exact on the field bulk) and the traces."""
    cfg = workloads.visco_3d_block(n=512, steps=2000)
    sim = Simulator3D(cfg)
    att = attenuation_from_config(cfg)
    g = VSESim.from_simulator(sim, variant="wgsl", att=att)
    steps = 40
    g.run(steps)
    a_f, a_t = g.field("szz"), g.traces()
    sim._source_term = (sim._source_term * np.float32(2.0)).astype(np.float32)
    g.load_sources(sim)
    g.reset()
    g.run(steps)
    b_f, b_t = g.field("szz"), g.traces()
    g.close()
    assert np.abs(a_f).max() > 0
    big = np.abs(a_f) > 1e-6 * np.abs(a_f).max()
    np.testing.assert_array_equal(b_f[big], 2 * a_f[big])
    assert oracle.rel_l2(b_t, 2 * a_t) < 1e-12


def test_visco_3d_512_fast_math_matches_reference_order():
    """The FAST arithmetic (pre-divided FD coefficients, 1/k, 1/3) against the
    reference operation order (STRICT, term-for-term equal to the FP32 oracle)
    on the full 512^3 block with the autotuned launch shapes' families: every
    field and the traces within 1e-5 relative L2 after 300 steps."""
    cfg = workloads.visco_3d_block(n=512, steps=2000)
    sim = Simulator3D(cfg)
    att = attenuation_from_config(cfg)
    g = VSESim.from_simulator(sim, variant="wgsl", att=att)
    steps = 300
    names = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    g.set_math("strict")
    g.set_tuning(0, 4, 0, 8, 1)
    g.set_tuning(1, 1, 0, 16, 1)
    g.run(steps, check_every=100)
    ref = {k: g.field(k) for k in names}
    ref_t = g.traces()
    g.set_math("fast")
    g.set_tuning(0, 4, 0, 0, 0)                 # plain kernels for the FAST run
    g.set_tuning(1, 4, 0, 0, 0)
    g.reset()
    g.run(steps, check_every=100)
    worst = 0.0
    for k in names:
        got = g.field(k)
        assert np.abs(ref[k]).max() > 0, k
        rel = oracle.rel_l2(got, ref[k])
        worst = max(worst, rel)
        assert rel <= 1e-5, (k, rel)
        del got
    rel_t = oracle.rel_l2(g.traces(), ref_t)
    print(f"FAST vs STRICT at 512^3, {steps} steps: worst field rel-L2 {worst:.2e}, traces {rel_t:.2e}")
    assert rel_t <= 1e-5
    g.close()


def test_adjoint_1024_inversion_workload():
    """The 1024^2 inversion workload (3000 steps, one FMC shot, true model with
    8 Gaussian anomalies) through the gradient engine: sqrt(N) checkpoints and
    every state stored give the same gradient bits, and a fourth-order central
    difference of J = 0.5*||traces||^2 along a gradient-weighted direction
    under the array agrees with the adjoint within 1e-3."""
    from taichi_sim_amd.adjoint import VSEGrad, chain_to_cells

    n, steps = 1024, 3000
    cfg, laws, true = workloads.adjoint_2d(n=n, steps=steps, n_shots=1)
    sim = Simulator(cfg)
    sim._emission_laws = laws
    att = attenuation_from_config(cfg)
    g = VSESim.from_simulator(sim, variant="wgsl", visco=True, att=att)
    sim.select_law(0)
    g.load_sources(sim)
    start = [np.asarray(a, np.float32).copy() for a in (sim._rho_grid_vx, sim._cp_grid_vx, sim._cs_grid_vx)]
    g.set_material(*true)
    g.reset()
    g.run(steps, check_every=500)
    obs = g.traces().copy()

    eng = VSEGrad(g)
    assert 1 < eng.seg_len < steps
    eng.set_model(*start)
    eng.forward(steps, check_every=500)
    J = eng.misfit(obs)
    c_sqrt = eng.backward()
    assert J > 0 and all(np.abs(c_sqrt[k]).max() > 0 for k in ("lambda", "mu", "rho_vx"))

    every = VSEGrad(g, seg_len=steps)
    every.set_model(*start)
    every.forward(steps, check_every=500)
    every.misfit(obs)
    c_all = every.backward()
    every.close()
    for k in c_sqrt:
        np.testing.assert_array_equal(c_all[k], c_sqrt[k], err_msg=k)

    zero = np.zeros_like(obs)

    def energy(model):
        eng.set_model(*model)
        eng.forward(steps, check_every=500)
        return eng.misfit(zero)

    energy(start)
    grads = [t.numpy() for t in chain_to_cells(eng.backward(), *start)]
    # positive Gaussian weights under the array times the gradient itself:
    # d.g = sum w g^2 cannot cancel (a random smooth direction can fall where
    # J changes by less than the FP32 forward resolves)
    rng = np.random.default_rng(5)
    y_top = int(np.min(np.nonzero(np.asarray(sim._pos_sources) != -1)[1]))
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    w = np.zeros((n, n))
    for _ in range(4):
        cx, cy, sg = rng.uniform(0.2, 0.8) * n, rng.uniform(y_top + 2, y_top + 40), rng.uniform(10.0, 40.0)
        w += rng.uniform(0.5, 1.0) * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * sg * sg))
    dirs = [w * e / np.abs(w * e).max() * float(m.max()) for e, m in zip(grads, start)]
    h = 2e-2
    models = {t: [(start[p] + t * h * dirs[p]).astype(np.float32) for p in range(3)] for t in (-2, -1, 1, 2)}
    Jt = {t: energy(m) for t, m in models.items()}
    fd = (8.0 * (Jt[1] - Jt[-1]) - (Jt[2] - Jt[-2])) / 12.0
    step = lambda p, t: models[t][p].astype(np.float64) - start[p].astype(np.float64)
    ad = sum(float((gp * (8.0 * (step(p, 1) - step(p, -1)) - (step(p, 2) - step(p, -2))) / 12.0).sum())
             for p, gp in enumerate(grads))
    err = abs(fd - ad) / max(abs(fd), 1e-30)
    print(f"FD {fd:.6e} adjoint {ad:.6e} rel {err:.1e}")
    eng.close()
    g.close()
    assert err <= 1e-3
