"""The oracle against the reference's only numeric output.

pedro/results/result_teste_aco.npy is the B-scan the reference's Taichi 2D
elastic backend (sim_taichi_elastic.py) wrote for pedro/teste_pedro.json
(725x572 grid, 64-element array, 10000 steps).  It pins the host setup
(ROI rounding, CPML profiles, array rasterisation, source term, receiver
tables) and the Taichi-variant kernels end to end.  The reference computed it
in FP32 on an unknown GPU with Taichi's fast-math; the FP64 oracle sits 4.2e-6
rel-L2 from it over all 10000 rows, which is the golden's own rounding floor.
"""
import numpy as np
import pytest

import oracle
from taichi_sim_amd.sim_support.simulator import Simulator
from tests.conftest import pedro_config

N_STEPS = 2500   # rows 0..2499 include the direct echoes (first arrival at row 335)


@pytest.fixture(scope="module")
def pedro():
    cfg, bscan = pedro_config()
    return Simulator(cfg), bscan


def test_setup_matches_golden_geometry(pedro):
    sim, bscan = pedro
    assert (sim._nx, sim._ny) == (725, 572)
    assert bscan.shape == (sim._n_steps, sim._n_rec) == (10000, 64)
    assert sim._n_pto_rec == 256 and sim._source_term.shape == (10000, 64)


@pytest.mark.parametrize("variant", ["taichi", "wgsl"])
def test_oracle_reproduces_golden_bscan(pedro, variant):
    sim, bscan = pedro
    o = oracle.Oracle(sim, variant=variant, precision="f32").run(N_STEPS)
    tr = o.traces[:N_STEPS]
    g = bscan[:N_STEPS]
    # first arrival row identical; B-scan within the golden's FP32 floor
    assert np.nonzero(np.any(g != 0, axis=1))[0][0] == 335
    assert np.nonzero(np.any(tr[1:] != 0, axis=1))[0][0] + 1 == 335
    rel = oracle.rel_l2(tr[1:], g[1:])
    assert rel < 2e-5, rel
    if variant == "taichi":
        assert rel < 2e-6, rel
