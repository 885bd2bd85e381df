"""Simulator.run(): the per-law driver around implementation() and its outputs.

The reference's run() (sim_support/simulator.py:321-605) loops over
iterations and emission laws, applies each law's delays to the probes, calls
implementation() and reports/saves its result.  A stub backend whose result
is a function of the current source term checks the order of calls, that
each law reaches the source term, the result-dict contract, the saved file
names, and that handing laws over in groups (implementation_many) changes
nothing.  The end-of-run outputs follow simulator.py:358-599: the MSE pair
against ``result_ref_*`` files (ROI window, inf on a shape mismatch, a
message when a file is missing), the timing CSV and the description text.
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


def _ref_files(tmp_path, sim, field, traces):
    np.save(os.path.join(tmp_path, f"result_ref_{sim._sim_model}_field_pressure.npy"), field)
    np.save(os.path.join(tmp_path, f"result_ref_{sim._sim_model}_bscan_pressure.npy"), traces)


def test_mse_against_reference_files(tmp_path, capsys):
    cfg, laws = _cfg(tmp_path, False)
    sim = StubBackend(cfg)
    sim._emission_laws = laws
    # reference files: the law-0 result, with one ROI cell and one trace sample changed
    sim.select_law(0)
    r0 = sim.implementation()
    field, traces = r0["stress"].copy(), r0["sens_stress"].copy()
    roi = sim._roi
    ix, iz = roi.get_ix_min(), roi.get_iz_min()
    field[ix + 3, iz + 2] += 2.0
    field[0, 0] += 1e6               # outside the ROI window: not compared
    traces[7, 1] -= 3.0
    _ref_files(tmp_path, sim, field, traces)
    sim.calls.clear()
    sim.run()
    mse = sim._mse_values
    assert mse.shape == (10, 2)
    n_win = (roi.get_ix_max() - ix) * (roi.get_iz_max() - iz)
    assert ix > 0 and mse[0, 0] == pytest.approx(4.0 / n_win, rel=1e-6)
    assert mse[0, 1] == pytest.approx(9.0 / traces.size, rel=1e-6)
    assert mse[5, 0] == mse[0, 0]                   # law 0 of iteration 1
    assert mse[5, 1] == mse[0, 1]
    assert np.all(mse[1:5, 1] > mse[0, 1])          # other laws' traces differ from law 0's
    out = capsys.readouterr().out
    assert "MSE do campo de pressao em relacao a referencia" in out


def test_mse_shape_mismatch_is_inf_and_missing_file_is_reported(tmp_path, capsys):
    cfg, laws = _cfg(tmp_path, False)
    sim = StubBackend(cfg)
    _ref_files(tmp_path, sim, np.zeros((8, 8), np.float32), np.zeros((3, 3), np.float32))
    sim.run()
    assert sim._mse_values.shape == (2, 2) and np.all(np.isinf(sim._mse_values))
    os.remove(os.path.join(tmp_path, f"result_ref_{sim._sim_model}_bscan_pressure.npy"))
    capsys.readouterr()
    sim.run()
    assert sim._mse_values.shape == (0,)
    out = capsys.readouterr().out
    assert "nao encontrado. Nao pode ser feita a comparacao com a referencia." in out


@pytest.mark.parametrize("n_iter", [2, 7])
def test_timing_csv_and_description(tmp_path, capsys, n_iter):
    cfg, laws = _cfg(tmp_path, True)
    cfg["simul_configs"]["n_iter"] = n_iter
    sim = StubBackend(cfg)
    sim._emission_laws = laws

    times = iter(np.arange(1, 100) * 0.125)

    impl = sim.implementation

    def timed():
        r = impl()
        r["sim_time"] = float(next(times))
        r["msg_impl"] = "Implementacao: stub\n"
        return r
    sim.implementation = timed
    sim.select_law(0)
    r0 = impl()
    _ref_files(tmp_path, sim, r0["stress"], r0["sens_stress"])
    sim.run(laws_per_call=2)
    names = os.listdir(tmp_path)
    csv = [n for n in names if n.endswith("GPU_.csv")]
    desc = [n for n in names if n.endswith("__desc.txt")]
    assert len(csv) == 1 and len(desc) == 1
    assert f"_96x96_60_iter_{n_iter - 1}_GPU_.csv" in csv[0]
    got = np.loadtxt(os.path.join(tmp_path, csv[0]), delimiter=",")
    np.testing.assert_allclose(got, np.arange(1, 5 * n_iter + 1) * 0.125, atol=5e-4)
    text = open(os.path.join(tmp_path, desc[0])).read().splitlines()
    assert text[:4] == ["Parametros do ensaio", "--------------------", "", "Simulador: stub"]
    assert text[4] == f"Modelo do simulador: {sim._sim_model}"
    assert text[5:9] == ["Quantidade de iteracoes no tempo: 60", "Tamanho da ROI: 96x96", "GPU: stub",
                         f"Numero de simulacoes: {n_iter}"]
    assert text[9].startswith("Energia refletida (sensor): ") and text[10].startswith("Energia ultimo frame: ")
    assert text[11] == "Implementacao: stub"
    st = sim._sim_times
    out = capsys.readouterr().out
    if n_iter > 5:
        assert text[12] == f"Tempo medio de execucao: {st[5:].mean():.3}s"
        assert text[13] == f"Desvio padrao: {st[5:].std():.4}"
        assert text[16].startswith("MSE medio do campo de pressao: ")
        assert len(text) == 20
        assert f"Tempo medio de execucao: {st[5:].mean():.3}s (std = {st[5:].std():.4})" in out
        assert "MSE medio dos sensores de pressao" in out
    else:
        assert text[12] == f"Tempo execucao: {st[-1]:.3}s"
        assert text[13].startswith("Tempo total (inclui transferencia de dados): ")
        mse = sim._mse_values[-1]
        assert text[14:] == [f"MSE do campo de pressao: {mse[0]:.4}", f"MSE medio dos sensores de pressao: {mse[1]:.4}"]
        assert "Tempo medio de execucao" not in out


def test_display_steps_follow_the_reference_rule():
    # show_debug / show_anim act after step it when it % it_display == 0 or
    # it == 5, it = 1..n_steps (sim_webgpu_viscoelastic.py:471,519); nothing
    # when neither flag is set
    from taichi_sim_amd.backend import HIPBackend

    class Fake(HIPBackend):
        pass

    f = Fake()
    f._n_steps, f._it_display = 230, 20
    f._show_debug, f._show_anim = False, False
    assert f._display_steps() == []
    f._show_debug = True
    assert f._display_steps() == [5, 20, 40, 60, 80, 100, 120, 140, 160, 180, 200, 220]
    f._show_debug, f._show_anim = False, True
    f._n_steps, f._it_display = 4, 2
    assert f._display_steps() == [2, 4]
    f._n_steps, f._it_display = 12, 5
    assert f._display_steps() == [5, 10]
