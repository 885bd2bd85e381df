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
