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
