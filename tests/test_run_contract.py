"""Simulator.run(): the per-law driver around implementation() and its outputs.

The reference's run() (sim_support/simulator.py:321-605) loops over
iterations and emission laws, applies each law's delays to the probes, calls
implementation() and reports/saves its result.  A stub backend whose result
is a function of the current source term checks the order of calls, that
each law reaches the source term, the result-dict contract, the saved file
names, and that handing laws over in groups (implementation_many) changes
nothing.
"""
import os
import zlib

import numpy as np
import pytest

from taichi_sim_amd import workloads
from taichi_sim_amd.sim_support.simulator import Simulator


class StubBackend(Simulator):
    def __init__(self, cfg):
        super().__init__(cfg, sim_type="viscoelastic")
        self._name = "stub"
        self.calls = []

    def implementation(self):
        super().implementation()
        st = np.asarray(self._source_term, dtype=np.float32)
        self.calls.append(zlib.crc32(st.tobytes()))
        field = np.full((self._nx, self._ny), st.sum(), dtype=np.float32)
        traces = st[:, :4].copy()
        return {"stress": field, "sens_stress": traces, "gpu_str": "stub", "sim_time": 0.0}


def _cfg(tmp_path, save):
    cfg, laws = workloads.multi_shot_2d(sizes=(96,), n_shots=5, steps=60, n_elem=8)[0]
    cfg["simul_configs"].update({"save_results": int(save), "save_bscan": int(save),
                                 "results_dir": str(tmp_path), "n_iter": 2})
    return cfg, laws


@pytest.mark.parametrize("per_call", [1, 2, 5, 7])
def test_laws_in_order_and_grouping_invariant(tmp_path, per_call):
    cfg, laws = _cfg(tmp_path, False)
    sim = StubBackend(cfg)
    sim._emission_laws = laws
    res = sim.run(laws_per_call=per_call)
    assert len(res) == 2 * 5
    # each law's delays reach the source term, in law order, every iteration
    ref = StubBackend(cfg)
    expect = []
    for law in range(5):
        ref._emission_laws = laws
        ref.select_law(law)
        expect.append(zlib.crc32(np.asarray(ref._source_term, np.float32).tobytes()))
    assert sim.calls == expect * 2
    assert len(set(expect)) == 5
    for r in res:
        assert set(r) >= {"stress", "sens_stress", "gpu_str", "sim_time"}
    assert sim._sim_times.shape == (10,) and sim._sim_total_times.shape == (10,)


def test_saved_outputs(tmp_path):
    cfg, laws = _cfg(tmp_path, True)
    sim = StubBackend(cfg)
    sim._save_field = True
    sim._emission_laws = laws
    res = sim.run(laws_per_call=3)
    names = sorted(os.listdir(tmp_path))
    bscans = [n for n in names if n.endswith("_bscan_stress.npy")]
    fields = [n for n in names if n.endswith("_field_stress_yy.npy")]
    assert len(bscans) == 10 and len(fields) == 10
    for it in range(2):
        for law in range(5):
            hit = [n for n in bscans if f"_96x96_60_iter_{it}_law_{law}_" in n]
            assert len(hit) == 1
            got = np.load(os.path.join(tmp_path, hit[0]))
            assert np.array_equal(got, res[it * 5 + law]["sens_stress"])


def test_without_laws_one_call_per_iteration(tmp_path):
    cfg, _ = _cfg(tmp_path, False)
    sim = StubBackend(cfg)
    res = sim.run(laws_per_call=4)
    assert len(res) == 2 and len(sim.calls) == 2
