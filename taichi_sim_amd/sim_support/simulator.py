"""2D simulator base class: configuration, setup arrays and the run() driver.

Mirrors ``sim_support/simulator.py`` of the reference: ``Simulator.__init__``
(:18-318) builds every array the time loop consumes as ``self._*`` attributes,
``run()`` (:321-605) loops over iterations and emission laws calling
``self.implementation()`` (:608-611), the backend hook a subclass overrides.

Differences, all outside the numbers of configurations the reference can run:
* the configuration may be a path (parsed with ``ast.literal_eval`` like the
  reference) or an already-parsed dict; material maps may be paths or arrays;
* probe booleans are parsed, not ``eval``'d;
* an emission-law file is read with an ``EmissionLaw`` instance (the reference
  calls the method on the class and raises TypeError) and each law rebuilds the
  source term (the reference's ``set_t0`` never reaches ``_source_term``);
* ``run()`` keeps the timing, energy and ``np.save`` outputs but no plots or
  Qt animation.
"""
import ast
import os
from datetime import datetime
from time import time

import numpy as np

from . import flt32, int32, PI, coefs_Lui, CourantError
from .emission_law import EmissionLaw
from .simul_classes import (SimulationROI, SimulationProbeLinearArray, SimulationProbePoint,
                            SimulationProbeMonoCirc)


def load_config(file_config):
    if isinstance(file_config, dict):
        return file_config
    with open(os.path.normpath(file_config), 'r') as f:
        return ast.literal_eval(f.read())


def load_map(spec):
    if isinstance(spec, np.ndarray):
        return spec.astype(flt32)
    return np.load(os.path.normpath(spec)).astype(flt32)


def fd_coefficients(acc):
    """float32 Liu coefficients for derivative accuracy ``acc`` (2..6; clamps to 6)."""
    try:
        return np.array(coefs_Lui[acc - 2], dtype=flt32)
    except IndexError:
        print(f"Acurácia das derivadas {acc} não suportada. Usando o maior valor permitido (6).")
        return np.array(coefs_Lui[-1], dtype=flt32)


def paste_map(value, m, shape, window):
    """Material grid: background value, with map ``m`` pasted into the ROI window,
    cropped, or used as is (simulator.py:88-131)."""
    grid = np.ones(shape, dtype=flt32) * value
    if m is None:
        return grid
    ms = m.shape
    if all(a < b for a, b in zip(ms, shape)):
        grid[window] = m
        return grid
    if all(a > b for a, b in zip(ms, shape)):
        return m[tuple(slice(0, n) for n in shape)]
    if tuple(ms) == tuple(shape):
        return m
    raise ValueError(f'map shape {ms} e incompativel com a ROI')


def make_probe(pcfg, dec):
    if "linear" in pcfg:
        return SimulationProbeLinearArray(**pcfg["linear"], dec=dec)
    if "point" in pcfg:
        return SimulationProbePoint(**pcfg["point"], dec=dec)
    if "mono" in pcfg:
        return SimulationProbeMonoCirc(**pcfg["mono"], dec=dec)
    raise KeyError(f"unknown probe type in {list(pcfg)}")


def receiver_tables(idx_rec):
    """offset_sensors: first table row of each receiver id; like the reference its
    first entry is the first id itself (simulator.py:289-295)."""
    numbers = list(idx_rec)
    offsets = [numbers[0]] if numbers else []
    for i in range(1, len(numbers)):
        if numbers[i] != numbers[i - 1]:
            offsets.append(int32(i))
    return np.array(offsets, dtype=int32), int32(len(numbers))


class Simulator:
    def __init__(self, file_config, sim_model="split", sim_type="acoustic"):
        self._app = None
        self._device = None
        self._name = "simulator"
        self._sim_model = sim_model
        self._sim_type = sim_type
        self._configs = load_config(file_config)
        cfg = self._configs
        sp = cfg.get("simul_params", {})
        self._deriv_acc = sp.get("acc", 2)
        self._coefs = fd_coefficients(self._deriv_acc)
        self._n_steps = sp.get("time_steps", 1000)
        self._dt = flt32(sp.get("dt", 1.0))
        self._it_display = sp.get("it_display", 10)
        self.cpml_time = sp.get("cpml_time", 1000)

        spec = cfg["specimen_params"]
        self._cp = flt32(spec.get("cp", 5.9))
        self._cs = flt32(spec.get("cs", 3.23))
        self._rho = flt32(spec.get("rho", 7800.0))
        maps = {k: load_map(spec[k + "_map"]) for k in ("cp", "cs", "rho") if k + "_map" in spec}
        if "cp" in maps:
            self._cp_map = maps["cp"]
        if "cs" in maps:
            self._cs_map = maps["cs"]
        if "rho" in maps:
            self._rho_map = maps["rho"]

        pad = self._deriv_acc - 1
        self._roi = SimulationROI(**cfg["roi"], pad=pad, rho_map=maps.get("rho"))
        roi = self._roi
        # 2D grids are (x, z); the kernels call the second axis "y".
        self._nx = roi.get_nx()
        self._ny = roi.get_nz()
        self._nz = roi.get_ny()
        self._dx = roi.get_dx()
        self._dy = roi.get_dz()
        self._dz = roi.get_dy()
        self._one_dx = flt32(1.0 / self._dx)
        self._one_dy = flt32(1.0 / self._dy)
        self._one_dz = flt32(1.0 / self._dz)

        shape = (self._nx, self._ny)
        window = (slice(roi.get_ix_min(), roi.get_ix_max()), slice(roi.get_iz_min(), roi.get_iz_max()))
        self._rho_grid_vx = paste_map(self._rho, maps.get("rho"), shape, window)
        # The reference aliases rho_grid_vy to rho_grid_vx and averages it in
        # place (simulator.py:101-105): the kernels' "rho" is the 4-point mean
        # on [:-1, :-1] with the last row and column unchanged.
        r = self._rho_grid_vx
        r[:-1, :-1] = flt32(0.25) * (r[:-1, :-1] + r[1:, :-1] + r[1:, 1:] + r[:-1, 1:])
        self._rho_grid_vy = r
        self._cp_grid_vx = paste_map(self._cp, maps.get("cp"), shape, window)
        self._cs_grid_vx = paste_map(self._cs, maps.get("cs"), shape, window)

        # Courant check (simulator.py:133-139)
        cp_max = max(self._cp_grid_vx.max(), self._cp)
        courant = flt32(cp_max * self._dt * np.sqrt(self._one_dx ** 2 + self._one_dy ** 2))
        print(f'\nNumero de Courant e {courant}')
        if courant > 1.0:
            raise CourantError("O passo de tempo e muito longo e a simulacao sera instavel", courant)
        self._cp_max = cp_max

        self._probes = [make_probe(p, roi.get_dec()) for p in cfg["probes"]]
        self._gain = flt32(0.0)
        for pr in self._probes:
            self._gain = max(self._gain, flt32(pr._gain) if hasattr(pr, "_gain") else 0.0)

        self._read_flags()
        self._build_probe_tables("2d")
        self._build_cpml(("x", "z"))
        self._build_sources("2d")

    # ------------------------------------------------------------------ setup
    def _read_flags(self):
        sc = self._configs.get("simul_configs", {})
        self._n_iter = sc.get("n_iter", 1)
        self._max_val_fields = flt32(sc.get("max_val_fields", 100.0))
        self._min_val_fields = flt32(sc.get("min_val_fields", -100.0))
        for k in ("show_anim", "show_debug", "show_figs", "plot_results", "plot_error", "plot_sensors",
                  "plot_bscan", "save_results", "save_field", "save_sensors", "save_bscan",
                  "save_sources", "source_env"):
            setattr(self, "_" + k, bool(sc.get(k, False)))
        law = sc.get("emission_laws")
        if law is not None and os.path.isfile(os.path.normpath(law)):
            self._emission_laws, _ = EmissionLaw().read_law(os.path.normpath(law))
        rd = sc.get("results_dir")
        if rd is not None and os.path.isdir(os.path.normpath(rd)):
            self._results_dir = os.path.normpath(rd)

    def _build_probe_tables(self, kind):
        """Transmit/receive points, receiver delays and element counts (simulator.py:157-181)."""
        tx, rx, delay = [], [], []
        self._n_rec = 0
        self._n_src = 0
        for pr in self._probes:
            tx += pr.get_points_roi(self._roi, simul_type=kind, dir="e")[0]
            rx += pr.get_points_roi(self._roi, simul_type=kind, dir="r")[0]
            delay += pr.get_delay_rx()
            self._n_rec += pr.receivers.count(True)
            self._n_src += pr.emitters.count(True)
        tx = np.array(tx, dtype=int32).reshape(-1, 3)
        rx = np.array(rx, dtype=int32).reshape(-1, 3)
        self._tx_points, self._rx_points = tx, rx
        if kind == "2d":
            self._ix_src, self._iy_src = tx[:, 0].astype(int32), tx[:, 2].astype(int32)
            self._ix_rec, self._iy_rec = rx[:, 0].astype(int32), rx[:, 2].astype(int32)
        else:
            self._ix_src, self._iy_src, self._iz_src = (tx[:, i].astype(int32) for i in range(3))
            self._ix_rec, self._iy_rec, self._iz_rec = (rx[:, i].astype(int32) for i in range(3))
        self._n_pto_src = tx.shape[0]
        self._delay_recv = (np.array(delay) / self._dt + 1.0).astype(int32)

    def _build_cpml(self, axes):
        """a/b/k profiles, full and half grid (simulator.py:183-209)."""
        alpha_max = flt32(2.0 * PI * (self._probes[0].get_freq() / 2.0))
        self._alpha_max_pml = alpha_max
        names = ("x", "y", "z")
        for name, ax in zip(names, axes):
            for grid, sfx in (("f", ""), ("h", "_half")):
                a, b, k = self._roi.calc_pml_array(axis=ax, grid=grid, dt=self._dt, cp=self._cp_max,
                                                   alpha_max=alpha_max)
                setattr(self, f"_a_{name}{sfx}", a.astype(flt32))
                setattr(self, f"_b_{name}{sfx}", b.astype(flt32))
                setattr(self, f"_k_{name}{sfx}", k.astype(flt32))

    def _probe_source(self, pr):
        out = 'e' if self._source_env else 'r'
        if isinstance(pr, SimulationProbePoint):
            return pr.get_source_term(samples=self._n_steps, dt=self._dt, out=out, ord_der=1)
        return pr.get_source_term(samples=self._n_steps, dt=self._dt, out=out)

    def _build_sources(self, kind):
        """Source term, source/receiver maps and receiver tables (simulator.py:235-295).

        Column and receiver-id offsets advance by points, as in the reference.
        """
        terms, idx_src, idx_rec = [], [], []
        self._idx_src_offset = 0
        self._idx_rec_offset = 0
        for pr in self._probes:
            st = self._probe_source(pr)
            _, i_src = pr.get_points_roi(sim_roi=self._roi, simul_type=kind)
            if len(i_src) > 0:
                terms.append(st)
                idx_src += [np.array(s) + self._idx_src_offset for s in i_src]
                self._idx_src_offset += len(i_src)
            i_rec = pr.get_idx_rec(sim_roi=self._roi, simul_type=kind.upper())
            if len(i_rec) > 0:
                idx_rec += [np.array(r) + self._idx_rec_offset for r in i_rec]
                self._idx_rec_offset += len(i_rec)
        self._source_term = np.concatenate(terms, axis=1) if len(terms) > 1 else terms[0]
        self._idx_src = np.array(idx_src).astype(int32).flatten()
        self._idx_rec = np.array(idx_rec).astype(np.int32).flatten()
        if kind == "2d":
            shape = (self._nx, self._ny)
            self._pos_sources = -np.ones(shape, dtype=int32)
            self._pos_sources[self._ix_src, self._iy_src] = self._idx_src
            self._pos_sensors = -np.ones(shape, dtype=int32)
            self._pos_sensors[self._ix_rec, self._iy_rec] = self._idx_rec
            self._info_rec_pt = np.column_stack((self._ix_rec, self._iy_rec, self._idx_rec)).astype(int32)
        else:
            shape = (self._nx, self._ny, self._nz)
            self._pos_sources = -np.ones(shape, dtype=int32)
            self._pos_sources[self._ix_src, self._iy_src, self._iz_src] = self._idx_src
            self._pos_sensors = -np.ones(shape, dtype=int32)
            self._pos_sensors[self._ix_rec, self._iy_rec, self._iz_rec] = self._idx_rec
            self._info_rec_pt = np.column_stack((self._ix_rec, self._iy_rec, self._iz_rec,
                                                 self._idx_rec)).astype(int32)
        self._offset_sensors, self._n_pto_rec = receiver_tables(self._idx_rec)

    def rebuild_sources(self):
        """Recompute the source term after probe delays changed (per emission law)."""
        terms = [self._probe_source(pr) for pr in self._probes
                 if len(pr.get_points_roi(sim_roi=self._roi, simul_type=self._kind)[1]) > 0]
        self._source_term = np.concatenate(terms, axis=1) if len(terms) > 1 else terms[0]

    _kind = "2d"

    # ------------------------------------------------------------------ driver
    def _field_key(self, res):
        return res["stress"] if "stress" in res else res["pressure"]

    def _trace_key(self, res):
        return res["sens_stress"] if "sens_stress" in res else res["sens_pressure"]

    def run(self, laws_per_call=1):
        """Iterations x emission laws -> implementation(); timing, energy, saves.

        ``laws_per_call`` > 1 hands that many laws at a time to
        ``implementation_many()`` (a backend may simulate them together); the
        results, their order and the saved files are the same as with 1."""
        now = datetime.now()
        sim_times, total_times, results, mse_values = [], [], [], []
        n = 0
        for n in range(self._n_iter):
            print(f'Iteracao {n}')
            n_laws = self._emission_laws.shape[0] if hasattr(self, "_emission_laws") else 1
            step = max(1, int(laws_per_call))
            for first in range(0, n_laws, step):
                laws = list(range(first, min(n_laws, first + step)))
                t0 = time()
                chunk = self.implementation_many(laws)
                sim_total = (time() - t0) / len(laws)
                for law, res in zip(laws, chunk):
                    print(f'\tLaw {law} of {n_laws}')
                    sim_times.append(res["sim_time"])
                    total_times.append(sim_total)
                    mse = self._report(res, n, law, now, sim_times[-1], total_times[-1])
                    if mse is not None:
                        mse_values.append(mse)
                    results.append(res)
        sim_times, total_times = np.array(sim_times), np.array(total_times)
        mse_values = np.array(mse_values)
        self._sim_times, self._sim_total_times, self._mse_values = sim_times, total_times, mse_values
        print(f'TEMPO - {self._n_steps} pontos de tempo')
        if self._n_iter > 5:
            print(f'Tempo medio de execucao: {sim_times[5:].mean():.3}s (std = {sim_times[5:].std():.4})')
            print(f'Tempo medio total (inclui transferencia de dados): {total_times[5:].mean():.3}s '
                  f'(std = {total_times[5:].std():.4})')
        if self._n_iter > 5 and mse_values.shape[0] > 5:
            print(f'MSE medio do campo de pressao: {mse_values[5:, 0].mean():.4} (std = {mse_values[5:, 0].std():.4})')
            print(f'MSE medio dos sensores de pressao: {mse_values[5:, 1].mean():.4} '
                  f'(std = {mse_values[5:, 1].std():.4})')
        result_dir = getattr(self, "_results_dir", ".")
        if self._save_sources:
            np.save(os.path.join(result_dir, f'sources_{self._name}_{now.strftime("%Y%m%d-%H%M%S")}'),
                    self._source_term)
        if self._save_results and results:
            self._write_run_files(result_dir, now, n, sim_times, total_times, mse_values, results[-1])
        return results

    def _write_run_files(self, result_dir, now, n, sim_times, total_times, mse_values, last):
        """The per-run outputs of simulator.py:563-599: the loop times as
        ``..._GPU_.csv`` and the run description ``..._desc.txt`` (the last
        law's GPU string, energies and implementation message)."""
        name = os.path.join(result_dir, f'result_{self._name}_{now.strftime("%Y%m%d-%H%M%S")}_'
                                        f'{self._nx}x{self._ny}_{self._n_steps}_iter_{n}_')
        np.savetxt(name + 'GPU_.csv', sim_times, '%10.3f', delimiter=',')
        lines = ['Parametros do ensaio', '--------------------', '',
                 f'Simulador: {self._name}', f'Modelo do simulador: {self._sim_model}',
                 f'Quantidade de iteracoes no tempo: {self._n_steps}', f'Tamanho da ROI: {self._nx}x{self._ny}',
                 f'GPU: {last["gpu_str"]}', f'Numero de simulacoes: {self._n_iter}',
                 f"Energia refletida (sensor): {self._reflected_energy:.2}",
                 f"Energia ultimo frame: {self._field_energy:.2}"]
        text = "\n".join(lines) + "\n" + last.get("msg_impl", "")
        if self._n_iter > 5:
            text += (f'Tempo medio de execucao: {sim_times[5:].mean():.3}s\nDesvio padrao: {sim_times[5:].std():.4}\n'
                     f'Tempo medio total (inclui transferencia de dados): {total_times[5:].mean():.3}s\n'
                     f'Desvio padrao: {total_times[5:].std():.4}\n')
            if mse_values.shape[0] > 5:
                text += (f'MSE medio do campo de pressao: {mse_values[5:, 0].mean():.4}\n'
                         f'Desvio padrao: {mse_values[5:, 0].std():.4}\n'
                         f'MSE medio dos sensores de pressao: {mse_values[5:, 1].mean():.4}\n'
                         f'Desvio padrao: {mse_values[5:, 1].std():.4}\n')
        else:
            text += (f'Tempo execucao: {sim_times[-1]:.3}s\n'
                     f'Tempo total (inclui transferencia de dados): {total_times[-1]:.3}s\n')
            if mse_values.shape[0] > 0:
                text += (f'MSE do campo de pressao: {mse_values[-1, 0]:.4}\n'
                         f'MSE medio dos sensores de pressao: {mse_values[-1, 1]:.4}\n')
        with open(name + '_desc.txt', 'w') as f:
            f.write(text)

    def select_law(self, law):
        """Apply emission law ``law`` to the probes and rebuild the source term
        (simulator.py:331-333, plus the rebuild the reference leaves out)."""
        if hasattr(self, "_emission_laws"):
            for p in self._probes:
                p.set_t0(self._emission_laws[law])
            self.rebuild_sources()

    def implementation_many(self, laws):
        """Results of ``laws``, in order: one ``implementation()`` per law."""
        out = []
        for law in laws:
            self.select_law(law)
            out.append(self.implementation())
        return out

    def _report(self, res, n, law, now, sim_time, total_time):
        """Per-law console lines, energies and np.save outputs (simulator.py:338-605)."""
        print(res["gpu_str"])
        print(f'{sim_time:.3}s')
        print(f'Tempo total (inclui transferencia de dados): {total_time:.3}s')
        result_dir = getattr(self, "_results_dir", ".")
        name = os.path.join(result_dir, f'result_{self._name}_{now.strftime("%Y%m%d-%H%M%S")}_'
                                        f'{self._nx}x{self._ny}_{self._n_steps}_iter_{n}_law_{law}')
        traces = self._trace_key(res)
        field = self._field_key(res)
        row = max(0, min(traces.shape[0] - 1, self._n_steps - self.cpml_time))
        reflected = traces[row, :]
        self._reflected_energy = (reflected ** 2).sum() * self._dt
        print(f"Energia refletida (sensor): {self._reflected_energy:.2}")
        self._field_energy = (field ** 2).sum() * self._dx * self._dy
        print(f"Energia ultimo frame: {self._field_energy:.2}")
        mse = self._compare_with_reference(result_dir, field, traces)
        tag = "pressure" if self._sim_type == "acoustic" else "stress"
        if self._save_results and self._save_field:
            np.save(name + ('_field_pressure' if tag == "pressure" else '_field_stress_yy'), field)
        if self._save_bscan:
            np.save(name + f'_bscan_{tag}', traces)
        return mse

    def _compare_with_reference(self, result_dir, field, traces):
        """MSE of the ROI window of the last field and of the B-scan against
        ``result_ref_{sim_model}_field_pressure.npy`` / ``..._bscan_pressure.npy``
        in the results folder (simulator.py:358-393): inf for a shape
        mismatch, None (and a message) when a reference file is missing."""
        roi = self._roi
        window = (slice(roi.get_ix_min(), roi.get_ix_max()), slice(roi.get_iz_min(), roi.get_iz_max()))
        try:
            ref = np.load(os.path.join(result_dir, f"result_ref_{self._sim_model}_field_pressure.npy"))[window]
            mine = field[window]
            mse_field = np.mean((ref - mine) ** 2) if ref.shape == mine.shape else np.inf
            ref_traces = np.load(os.path.join(result_dir, f"result_ref_{self._sim_model}_bscan_pressure.npy"))
            mse_traces = np.mean((ref_traces - traces) ** 2) if ref_traces.shape == traces.shape else np.inf
        except FileNotFoundError as err:
            print(f"Arquivo {err} nao encontrado. Nao pode ser feita a comparacao com a referencia.")
            return None
        print(f"MSE do campo de pressao em relacao a referencia: {mse_field:.4}")
        print(f"MSE dos sensores de pressao em relacao a referencia: {mse_traces:.4}")
        return [mse_field, mse_traces]

    def implementation(self):
        """Backend hook: returns {"stress"/"pressure", "sens_stress"/"sens_pressure",
        "gpu_str", "sim_time"} (simulator.py:608-611)."""
        print(f'Simulacao {self._name}')
