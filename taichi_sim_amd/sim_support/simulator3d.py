"""3D simulator base class (reference ``sim_support/Simulator3D.py:17-622``).

Same structure as the 2D ``Simulator``: grids are (x, y, z) with z fastest,
the kernel density is the host-side 8-point mean of rho (Simulator3D.py:96-103),
CPML profiles exist for all three axes and the receivers record sigma_zz.

Differences from the reference, none of which changes the numbers of a
configuration the reference can run: a ``cp_map`` is honoured (the reference
resets the grid to 2D and raises, Simulator3D.py:107-108), each probe gets the
source term of its own type (the reference dispatches on the last probe's
config, Simulator3D.py:257-277), and the Courant failure raises
``CourantError`` (a ``ValueError``) instead of a bare ``Exception``.
"""
import numpy as np

from . import flt32, CourantError
from .simulator import Simulator, load_config, load_map, fd_coefficients, paste_map, make_probe
from .simul_classes import SimulationROI


class Simulator3D(Simulator):
    _kind = "3d"

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
        for k, m in maps.items():
            setattr(self, f"_{k}_map", m)

        self._roi = SimulationROI(**cfg["roi"], pad=self._deriv_acc - 1, rho_map=maps.get("rho"))
        roi = self._roi
        self._nx, self._ny, self._nz = roi.get_nx(), roi.get_ny(), roi.get_nz()
        self._dx, self._dy, self._dz = roi.get_dx(), roi.get_dy(), roi.get_dz()
        self._one_dx = flt32(1.0 / self._dx)
        self._one_dy = flt32(1.0 / self._dy)
        self._one_dz = flt32(1.0 / self._dz)

        shape = (self._nx, self._ny, self._nz)
        window = (slice(roi.get_ix_min(), roi.get_ix_max()), slice(roi.get_iy_min(), roi.get_iy_max()),
                  slice(roi.get_iz_min(), roi.get_iz_max()))
        r = paste_map(self._rho, maps.get("rho"), shape, window)
        # In-place 8-point mean through the vx/vy alias (Simulator3D.py:96-103).
        r[:-1, :-1, :-1] = flt32(0.125) * (
            r[:-1, :-1, :-1] + r[1:, :-1, :-1] + r[:-1, 1:, :-1] + r[1:, 1:, :-1]
            + r[:-1, :-1, 1:] + r[1:, :-1, 1:] + r[:-1, 1:, 1:] + r[1:, 1:, 1:])
        self._rho_grid_vx = self._rho_grid_vy = r
        self._cp_grid_vx = paste_map(self._cp, maps.get("cp"), shape, window)
        self._cs_grid_vx = paste_map(self._cs, maps.get("cs"), shape, window)

        cp_max = max(self._cp_grid_vx.max(), self._cp)
        courant = flt32(cp_max * self._dt * np.sqrt(self._one_dx ** 2 + self._one_dy ** 2 + self._one_dz ** 2))
        print(f'\nNumero de Courant e {courant}')
        if courant > 1.0:
            raise CourantError("O passo de tempo e muito longo e a simulacao sera instavel", courant)
        self._cp_max = cp_max

        self._probes = [make_probe(p, roi.get_dec()) for p in cfg["probes"]]
        self._gain = flt32(0.0)
        for pr in self._probes:
            self._gain = max(self._gain, flt32(pr._gain) if hasattr(pr, "_gain") else 0.0)

        self._read_flags()
        self._build_probe_tables("3d")
        self._build_cpml(("x", "y", "z"))
        self._build_sources("3d")
