"""Drop-in backends: ``Simulator.implementation()`` on libvse.

The reference's plugin API is subclassing ``Simulator`` / ``Simulator3D`` and
overriding ``implementation(self) -> dict`` (sim_support/simulator.py:608-611,
called from run() at :337; Simulator3D.py:619-622).  ``HIPBackend`` is that
override: it reads the ``self._*`` setup attributes, runs the whole time loop
in libvse and returns the reference's result dict.  Because the 2D ``run()``
reads both key families (simulator.py:351,355 vs :476,496) the dict carries
both "stress"/"sens_stress" and "pressure"/"sens_pressure".

The mixin works on any object with the reference's attributes, so a user of
the reference can also write ``class SimulatorHIP(HIPBackend, Simulator)``
against the reference's own ``Simulator``.
"""
from time import time

import numpy as np

from .engine import VSEBatch, VSESim
from .sim_support.attenuation import AttenuationCoefficients, fit_constant_q
from .sim_support.simulator import Simulator
from .sim_support.simulator3d import Simulator3D


def attenuation_from_config(cfg):
    """The reference's hard-coded mechanisms, or a constant-Q fit when the
    configuration has an ``attenuation`` section {q_p, q_s, f_min, f_max, n_sls}
    (an extension; frequencies in MHz so tau is in microseconds)."""
    a = cfg.get("attenuation")
    if not a:
        return AttenuationCoefficients()
    return fit_constant_q(a.get("q_p", 50.0), a.get("q_s", a.get("q_p", 50.0)), a.get("f_min", 2.0),
                          a.get("f_max", 8.0), n_sls=a.get("n_sls", 3))


class HIPBackend:
    """``implementation()`` that runs the reference time loop on the GPU."""

    _variant = "wgsl"
    check_every = 100

    # Called as frame_hook(step, frame) at the reference's display steps when the
    # configuration sets show_anim: frame is the displayed slice of sigma_yy
    # (2D) or sigma_zz (3D).  None: the frame goes to the Qt windows the
    # simulator opened, if any (windows_qt.py needs PyQt6 and pyqtgraph).
    frame_hook = None

    def implementation(self):
        super().implementation()
        att = attenuation_from_config(self._configs)
        t_setup = time()
        sim = VSESim.from_simulator(self, variant=self._variant, att=att)
        self._setup_time = time() - t_setup
        try:
            ms = 0.0
            done = 0
            # the reference shows a debug line and a frame after step it when
            # it % it_display == 0 or it == 5 (sim_webgpu_viscoelastic.py:519-531,
            # sim_taichi_elastic.py:320-328, sim_webgpu_3D_viscoelastic.py:680-689)
            for it in self._display_steps():
                ms += sim.run(it - done, check_every=self.check_every, timed=True)
                done = it
                self._display(sim, it)
            ms += sim.run(self._n_steps - done, check_every=self.check_every, timed=True)
            field = sim.field("syy" if sim.ndim == 2 else "szz")
            traces = sim.traces()
            gpu = sim.device_name()
        finally:
            sim.close()
        return self._result(field, traces, gpu, ms / 1000.0, f"libvse {self._variant} variant, {sim.ndim}D, hipGraph step loop\n")

    def _display_steps(self):
        """Steps (1-based) after which show_debug / show_anim act: it % it_display
        == 0 or it == 5; none when neither is set."""
        if not (getattr(self, "_show_debug", False) or getattr(self, "_show_anim", False)):
            return []
        n, k = int(self._n_steps), max(1, int(self._it_display))
        return sorted(set(range(k, n + 1, k)) | ({5} if n >= 5 else set()))

    def _display(self, sim, it):
        """The reference's per-display-step output: two debug lines (the running
        max of |v|^2 as the WGSL backends print it, its square root; the Taichi
        backend prints the squared value) and the field frame."""
        if self._show_debug:
            v2 = sim.v2max()
            print(f'Time step # {it} out of {self._n_steps}')
            if self._variant == "taichi":
                print(f'Max norm velocity vector (m/s) = {v2}')
            else:
                print(f'Max norm velocity vector V (m/s) = {np.sqrt(v2)}')
        if self._show_anim:
            roi = self._roi
            if sim.ndim == 2:
                frame = sim.field("syy")[roi.get_ix_min():roi.get_ix_max(), roi.get_iz_min():roi.get_iz_max()]
            else:
                # the reference shows the y index gcd(ny, 16) // 2 (its workgroup width, halved)
                frame = sim.field("szz")[roi.get_ix_min():roi.get_ix_max(), int(np.gcd(int(roi.get_ny()), 16)) // 2,
                                         roi.get_iz_min():roi.get_iz_max()]
            if self.frame_hook is not None:
                self.frame_hook(it, frame)
            elif getattr(self, "_windows_gpu", None):
                self._windows_gpu[0].imv.setImage(frame, levels=[self._min_val_fields, self._max_val_fields])
                self._app.processEvents()

    @staticmethod
    def _result(field, traces, gpu, sim_time, msg):
        return {"stress": field, "sens_stress": traces, "pressure": field, "sens_pressure": traces,
                "gpu_str": gpu, "sim_time": sim_time, "msg_impl": msg}

    def implementation_many(self, laws):
        """Several emission laws of a 2D run simulated together: one libvse
        simulation per law, advanced as one batch (one launch per pass per
        step for all of them).  Each law's result equals implementation()'s;
        "sim_time" is the batch's device time divided by the number of laws."""
        if len(laws) < 2 or np.asarray(self._rho_grid_vx).ndim != 2:
            return super().implementation_many(laws)       # batches are 2D
        att = attenuation_from_config(self._configs)
        sims = []
        try:
            for law in laws:
                self.select_law(law)
                print(f'Simulacao {self._name}')
                sims.append(VSESim.from_simulator(self, variant=self._variant, att=att))
            batch = VSEBatch(sims)
            try:
                ms = batch.run(self._n_steps, check_every=self.check_every, timed=True)
            finally:
                batch.close()
            gpu = sims[0].device_name()
            msg = f"libvse {self._variant} variant, 2D, batch of {len(sims)} laws, hipGraph step loop\n"
            return [self._result(s.field("syy"), s.traces(), gpu, ms / 1000.0 / len(sims), msg) for s in sims]
        finally:
            for s in sims:
                s.close()


class SimulatorHIP(HIPBackend, Simulator):
    """2D viscoelastic backend with the WGSL semantics (sim_webgpu_viscoelastic.py:20-555)."""

    def __init__(self, file_config):
        super().__init__(file_config, sim_type="viscoelastic")
        self._name = "HIP-viscoelastic"


class SimulatorHIPTaichiElastic(HIPBackend, Simulator):
    """2D elastic backend with the Taichi semantics (sim_taichi_elastic.py:19-351)."""

    _variant = "taichi"

    def __init__(self, file_config):
        super().__init__(file_config)
        self._name = "HIP-split-elastic"


class SimulatorHIP3D(HIPBackend, Simulator3D):
    """3D viscoelastic backend with the WGSL semantics (sim_webgpu_3D_viscoelastic.py:20-713)."""

    def __init__(self, file_config):
        super().__init__(file_config, sim_type="viscoelastic")
        self._name = "HIP-3D-viscoelastic"
