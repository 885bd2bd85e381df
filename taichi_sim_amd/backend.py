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

    def implementation(self):
        super().implementation()
        att = attenuation_from_config(self._configs)
        t_setup = time()
        sim = VSESim.from_simulator(self, variant=self._variant, att=att)
        self._setup_time = time() - t_setup
        try:
            ms = sim.run(self._n_steps, check_every=self.check_every, timed=True)
            field = sim.field("syy" if sim.ndim == 2 else "szz")
            traces = sim.traces()
            gpu = sim.device_name()
        finally:
            sim.close()
        return self._result(field, traces, gpu, ms / 1000.0, f"libvse {self._variant} variant, {sim.ndim}D, hipGraph step loop\n")

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
