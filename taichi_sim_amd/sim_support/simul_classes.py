"""ROI geometry, CPML profiles and transducer models (host setup of the loop).

Mirrors the interface of the reference's ``sim_support/simul_classes.py``:
``SimulationROI`` (:30-386), ``ElementRect`` (:416-534), ``ElementCirc``
(:536-647), ``SimulationProbeLinearArray`` (:650-965), ``SimulationProbePoint``
(:967-1183) and ``SimulationProbeMonoCirc`` (:1185-1343).

The grid indices, CPML profiles and source terms these classes produce are
kernel inputs, so every rounding step (``np.round`` to the ROI's decimal
count, float32/float64 promotion order, nearest-index tie breaking) is kept
as the reference performs it; the golden B-scan depends on them.  Boolean
probe flags given as strings ("True"/"False") are parsed, never ``eval``'d.
"""
import numpy as np
from scipy.signal import gausspulse

from . import flt32
from .gauss_wavelets import gaussian_pulse, gaussian_first_dev_pulse, gaussian_second_dev_pulse


def _parse_flag(s):
    """'True'/'false'/... -> bool (the reference eval()s these strings)."""
    if isinstance(s, bool):
        return s
    v = str(s).strip().lower()
    if v in ("true", "1"):
        return True
    if v in ("false", "0", "none"):
        return False
    raise ValueError(f"invalid boolean flag {s!r}")


def _decimals(step):
    """Decimal count used for rounding coordinates of a grid of pitch ``step``."""
    return int(abs(np.log10(step))) + 2


class SimulationROI:
    """Simulation region: ROI points plus CPML layers plus ``pad`` cells per side.

    Axis naming follows the reference: ``w`` = width = x, ``d`` = depth = y,
    ``h`` = height = z.  A 2D simulation uses (x, z); the kernels call z "y".
    """

    def __init__(self, coord_ref=np.zeros((1, 3)), height=30.0, h_len=300, width=30.0, w_len=300,
                 depth=0.0, d_len=1, len_pml_xmin=10, len_pml_xmax=10, len_pml_ymin=10,
                 len_pml_ymax=10, len_pml_zmin=10, len_pml_zmax=10, rcoef_pml=0.0001,
                 npower_pml=2.0, k_max_pml=1.0, pad=1, rho_map=None):
        if isinstance(coord_ref, list):
            coord_ref = np.array(coord_ref, dtype=np.float32)
        if coord_ref.shape != (1, 3):
            raise TypeError("``coord_ref`` deve ser um vetor-linha de 3 elementos [shape = (1,3)]")
        self.coord_ref = coord_ref

        # Grid pitches, rounded to their own decimal count then stored float32
        # (simul_classes.py:141-156).  The depth pitch falls back to the width
        # pitch, whose float32 value can give a smaller decimal count.
        h = height / h_len
        self._dec_h = _decimals(h)
        self.h_step = np.round(h, decimals=self._dec_h).astype(np.float32)
        w = width / w_len
        self._dec_w = _decimals(w)
        self.w_step = np.round(w, decimals=self._dec_w).astype(np.float32)
        d = depth / d_len if (depth > 0.0 and d_len > 1) else self.w_step
        self._dec_d = _decimals(d)
        self.d_step = np.round(d, decimals=self._dec_d).astype(np.float32)

        # A material map larger than the configured ROI enlarges it (:158-190).
        w_map = h_map = d_map = 0
        if rho_map is not None:
            if rho_map.ndim == 1:
                w_map = rho_map.shape[0]
            elif rho_map.ndim == 2:
                w_map, h_map = rho_map.shape
            elif rho_map.ndim == 3:
                w_map, d_map, h_map = rho_map.shape
        self.width, self._w_len = ((np.float32(w_map * self.w_step), w_map) if w_map > w_len
                                   else (np.float32(width), w_len))
        self.height, self._h_len = ((np.float32(h_map * self.h_step), h_map) if h_map > h_len
                                    else (np.float32(height), h_len))
        self.depth, self._d_len = ((np.float32(d_map * self.d_step), d_map) if d_map > d_len
                                   else (np.float32(depth), d_len))

        def axis_points(origin, extent, n, dec):
            return np.linspace(origin, origin + extent, num=int(n), endpoint=False,
                               dtype=np.float32).round(decimals=dec)
        self.w_points = axis_points(coord_ref[0, 0], self.width, self._w_len, self._dec_w)
        self.d_points = axis_points(coord_ref[0, 1], self.depth, self._d_len, self._dec_d)
        self.h_points = axis_points(coord_ref[0, 2], self.height, self._h_len, self._dec_h)

        self._pml_xmin_len, self._pml_xmax_len = len_pml_xmin, len_pml_xmax
        self._pml_ymin_len, self._pml_ymax_len = len_pml_ymin, len_pml_ymax
        self._pml_zmin_len, self._pml_zmax_len = len_pml_zmin, len_pml_zmax
        self._rcoef_pml = rcoef_pml
        self._npower_pml = npower_pml
        self._k_max_pml = k_max_pml
        self._pad = pad

    # ---- sizes and index ranges (simul_classes.py:220-283) ----------------
    def has_pml(self):
        return any((self._pml_xmin_len, self._pml_xmax_len, self._pml_ymin_len,
                    self._pml_ymax_len, self._pml_zmin_len, self._pml_zmax_len))

    def get_dx(self):
        return flt32(self.w_step)

    def get_dy(self):
        return flt32(self.d_step)

    def get_dz(self):
        return flt32(self.h_step)

    def get_nx(self):
        return self._w_len + self._pml_xmin_len + self._pml_xmax_len + 2 * self._pad

    def get_ny(self):
        return self._d_len + self._pml_ymin_len + self._pml_ymax_len + 2 * self._pad

    def get_nz(self):
        return self._h_len + self._pml_zmin_len + self._pml_zmax_len + 2 * self._pad

    def get_len_x(self):
        return self._w_len

    def get_len_y(self):
        return self._d_len

    def get_len_z(self):
        return self._h_len

    def get_ix_min(self):
        return self._pml_xmin_len + self._pad

    def get_ix_max(self):
        return self._w_len + self._pml_xmin_len + self._pad

    def get_iy_min(self):
        return self._pml_ymin_len + self._pad

    def get_iy_max(self):
        return self._d_len + self._pml_ymin_len + self._pad

    def get_iz_min(self):
        return self._pml_zmin_len + self._pad

    def get_iz_max(self):
        return self._h_len + self._pml_zmin_len + self._pad

    def get_pml_thickness_x(self):
        return (self._pml_xmin_len + self._pml_xmax_len) * self.w_step

    def get_pml_thickness_y(self):
        return (self._pml_ymin_len + self._pml_ymax_len) * self.d_step

    def get_pml_thickness_z(self):
        return (self._pml_zmin_len + self._pml_zmax_len) * self.h_step

    def get_dec(self):
        return self._dec_w, self._dec_d, self._dec_h

    # ---- point location (simul_classes.py:285-319) -------------------------
    def is_point_in_roi(self, point):
        point = np.asarray(point, dtype=np.float32)
        return bool(self.w_points[0] <= point[0] <= self.w_points[-1]
                    and self.d_points[0] <= point[1] <= self.d_points[-1]
                    and self.h_points[0] <= point[2] <= self.h_points[-1])

    def get_nearest_grid_idx(self, point):
        """Grid indices (incl. CPML and pad) of the ROI point nearest ``point``.

        The coordinate is nudged down by step/10**(dec-1) before rounding, so
        a point exactly between two grid points resolves to the lower one;
        argmin keeps the first of equal distances.
        """
        if isinstance(point, list):
            point = np.array(point, dtype=np.float32)
        if not self.is_point_in_roi(point):
            raise IndexError(f"[{point[0]}, {point[1]}, {point[2]}] out of bounds")

        def nearest(points, coord, step, dec, first):
            target = np.round(coord - step / 10.0 ** (dec - 1), decimals=dec)
            return np.absolute(points - target).argmin() + first

        return [nearest(self.w_points, point[0], self.w_step, self._dec_w, self._pml_xmin_len + self._pad),
                nearest(self.d_points, point[1], self.d_step, self._dec_d, self._pml_ymin_len + self._pad),
                nearest(self.h_points, point[2], self.h_step, self._dec_h, self._pml_zmin_len + self._pad)]

    # ---- CPML profiles (simul_classes.py:321-386) --------------------------
    def calc_pml_array(self, axis='x', grid='f', dt=1.0, cp=1.0, alpha_max=1.0):
        """CPML coefficient profiles (a, b, k) along one axis, float32.

        Length = ROI points + both layers (the pad cells are excluded).  With
        normalised depth p in [0, 1] inside a layer:
          d = d0 p^n, d0 = -(n+1) cp ln(R) / (total thickness of both layers),
          k = 1 + (k_max - 1) p^n, alpha = alpha_max (1 - p) (0 outside layers),
          b = exp(-(d/k + alpha) dt), a = d (b - 1) / (k (d + k alpha)) where d > 1e-6.
        ``grid='h'`` evaluates at the half-cell-shifted positions.
        """
        axis = axis.lower()
        if axis == 'x':
            delta, n_min, n_len, n_max = np.float32(self.w_step), self._pml_xmin_len, self._w_len, self._pml_xmax_len
            dec, thick_total = self._dec_w, self.get_pml_thickness_x()
        elif axis == 'y':
            delta, n_min, n_len, n_max = np.float32(self.d_step), self._pml_ymin_len, self._d_len, self._pml_ymax_len
            dec, thick_total = self._dec_d, self.get_pml_thickness_y()
        elif axis == 'z':
            delta, n_min, n_len, n_max = np.float32(self.h_step), self._pml_zmin_len, self._h_len, self._pml_zmax_len
            dec, thick_total = self._dec_h, self.get_pml_thickness_z()
        else:
            raise IndexError(f"'axis' = {axis} not supported")
        npw = self._npower_pml
        n_total = n_len + n_min + n_max
        left_edge = n_min * delta
        right_edge = (n_min + n_len - 1) * delta
        thick_left = n_min * delta
        thick_right = n_max * delta
        d0 = flt32(-(npw + 1.0) * cp * np.log(self._rcoef_pml) / thick_total)

        pos = np.round(delta * np.arange(n_total), decimals=dec).astype(np.float32)
        if grid in ('f', 'F'):
            depth_left = np.round(left_edge - pos, decimals=dec)
            depth_right = np.round(pos - right_edge, decimals=dec)
        elif grid in ('h', 'H'):
            depth_left = np.round(left_edge - (pos + delta / 2.0), decimals=dec).astype(np.float32)
            depth_right = np.round((pos + delta / 2.0) - right_edge, decimals=dec).astype(np.float32)
        else:
            raise IndexError(f"'grid' = {grid} not supported")

        in_left = (depth_left >= 0.0) if thick_left > 0.0 else np.zeros(depth_left.shape, dtype=bool)
        in_right = (depth_right >= 0.0) if thick_right > 0.0 else np.zeros(depth_right.shape, dtype=bool)
        in_layer = in_left | in_right
        p = np.zeros(n_total, dtype=np.float32)
        if thick_left:
            p[in_left] = depth_left[in_left] / thick_left
        if thick_right:
            p[in_right] = depth_right[in_right] / thick_right
        d = (d0 * p ** npw).astype(np.float32)
        k = (1.0 + (self._k_max_pml - 1.0) * p ** npw).astype(np.float32)
        alpha = (alpha_max * (1.0 - np.where(in_layer, p, 1.0))).astype(np.float32)
        b = np.exp(-(d / k + alpha) * dt).astype(np.float32)
        a = np.zeros(n_total, dtype=np.float32)
        act = np.where(d > 1e-6)
        a[act] = d[act] * (b[act] - 1.0) / (k[act] * (d[act] + k[act] * alpha[act]))
        return a, b, k


class SimulationProbe:
    """Base transducer: centre coordinate and rounding decimals (:389-413)."""

    def __init__(self, coord_center=np.zeros((1, 3)), dec=(2, 2, 2)):
        self.coord_center = np.asarray(coord_center).astype(np.float32)
        self._dec = dec


def _element_excitation(t, t0, freq, bw, gain, out):
    """gain * d/dt-free Gaussian pulse, differenced: the force term of one element.

    np.diff(gain * f32(gp) / dt, append=0) with |gp| < eps(f32) flushed to 0
    (simul_classes.py:456-467, :562-573).
    """
    dt = t[1] - t[0]
    gp, _, egp = gausspulse((t - t0), fc=freq, bw=bw, retquad=True, retenv=True)
    eps = np.finfo(flt32).eps
    if out == 'e':
        egp[np.abs(egp) < eps] = 0.0
        ss = flt32(egp)
    else:
        gp[np.abs(gp) < eps] = 0.0
        ss = flt32(gp)
    return np.diff(gain * np.float32(ss) / dt, append=0.0).astype(flt32)


def _sample_times(samples, dt):
    """Time axis of a transducer source term: arange*dt rounded to dt's decimals."""
    dec = int(abs(np.log10(dt))) + 2
    return np.round(np.arange(samples, dtype=np.float32) * dt, decimals=dec)


class ElementRect:
    """Rectangular element of a linear array (simul_classes.py:416-534)."""

    def __init__(self, dim_a=0.5, dim_p=10.0, coord_center=np.zeros((1, 3)), freq=5., bw=0.5,
                 gain=1.0, t0=1.0, tx_en=True, rx_en=True, pulse_type="gaussian"):
        self.elem_dim_a = np.float32(dim_a)
        self.elem_dim_p = np.float32(dim_p)
        self.coord_center = coord_center.astype(np.float32)
        self.freq = np.float32(freq)
        self.bw = np.float32(bw)
        self.gain = np.float32(gain)
        self.t0 = np.float32(t0)
        self.tx_en = tx_en
        self.rx_en = rx_en
        self.pulse_type = pulse_type

    def get_element_exc_fn(self, t, out='r'):
        return _element_excitation(t, self.t0, self.freq, self.bw, self.gain, out)

    def _counts(self, sim_roi, rounding):
        dec_w, dec_d, _ = sim_roi.get_dec()
        dim_p = min(self.elem_dim_p, sim_roi.depth)
        n_a = int(np.round(self.elem_dim_a / sim_roi.w_step, decimals=dec_w) + rounding)
        n_p = int(np.round(dim_p / sim_roi.d_step, decimals=dec_d) + rounding) if dim_p != 0.0 else 1
        return n_a, n_p, dim_p

    def get_num_points_roi(self, sim_roi=None, simul_type="2D"):
        n_a, n_p, _ = self._counts(sim_roi, 0.0)
        return n_a * n_p if simul_type.lower() == "3d" else n_a

    def get_points_roi(self, sim_roi=None, probe_center=np.zeros((1, 3)), simul_type="2D", dir="e"):
        """Grid points of the element: a row of n_a points (times n_p in 3D).

        Points start at the grid point nearest the element's first cell centre.
        Like the reference, the 3D row index restarts at 0 rather than at the
        first point's y index (simul_classes.py:530-532).
        """
        if not isinstance(dir, str):
            raise ValueError("'dir' must be a string")
        if (dir.lower() == "e" and not self.tx_en) or (dir.lower() == "r" and not self.rx_en):
            return list()
        n_a, n_p, dim_p = self._counts(sim_roi, 0.5)
        n_coord = n_a * n_p if simul_type.lower() == "3d" else n_a
        x0 = np.float32(self.coord_center[0] - ((self.elem_dim_a - sim_roi.w_step) / 2.0 if n_a // 2
                                                else (self.elem_dim_a / 2.0)))
        if simul_type.lower() == "2d":
            y0 = np.float32(0.0)
        elif n_p // 2:
            y0 = np.float32(self.coord_center[1] - ((dim_p - sim_roi.d_step) / 2.0))
        else:
            y0 = np.float32(dim_p / 2.0)
        first = np.array([x0, y0, self.coord_center[2]], np.float32) + np.asarray(probe_center).astype(np.float32)
        p0 = sim_roi.get_nearest_grid_idx(first)
        pts = []
        for q in range(n_coord):
            row = p0[1] if simul_type == "2d" else ((q // n_a) % n_p)
            pts.append([p0[0] + (q % n_a), row, p0[2]])
        return pts


class ElementCirc:
    """Circular element of a mono transducer; radius given in inches (:536-647)."""

    def __init__(self, radius=0.25, coord_center=np.zeros((3,)), freq=5., bw=0.5, gain=1.0, t0=1.0,
                 tx_en=True, rx_en=True, pulse_type="gaussian"):
        self.radius = np.float32(radius * 25.4)
        self.coord_center = coord_center.astype(np.float32)
        self.freq = np.float32(freq)
        self.bw = np.float32(bw)
        self.gain = np.float32(gain)
        self.t0 = np.float32(t0)
        self.tx_en = tx_en
        self.rx_en = rx_en
        self.pulse_type = pulse_type

    def get_element_exc_fn(self, t, out='r'):
        return _element_excitation(t, self.t0, self.freq, self.bw, self.gain, out)

    def _candidate_points(self, sim_roi, simul_type, center):
        dec_w, dec_d, _ = sim_roi.get_dec()
        simul_type = simul_type.lower()
        n_a = int(np.round((2.0 * self.radius) / sim_roi.w_step, decimals=dec_w))
        if simul_type == "3d":
            r_d = min(self.radius, sim_roi.depth)
            n_p = int(np.round((2.0 * r_d) / sim_roi.d_step, decimals=dec_d))
        else:
            n_p, r_d = 1, 0.0
        corner = np.array([np.float32(center[0] - self.radius),
                           np.float32(0.0 if simul_type == "2d" else center[1] - r_d),
                           center[2]], dtype=flt32)
        ci = sim_roi.get_nearest_grid_idx(corner)
        pts = []
        for jj in range(n_a):
            ddx = (jj - (n_a - 1) / 2.0) * sim_roi.w_step
            for zz in range(n_p):
                ddy = (zz - (n_p - 1) / 2.0) * sim_roi.d_step if simul_type == "3d" else 0.0
                if np.sqrt(ddx ** 2 + ddy ** 2) <= self.radius:
                    pts.append([ci[0] + jj, ci[1] + (0 if simul_type == "2d" else zz), ci[2]])
        return pts

    def get_num_points_roi(self, sim_roi, simul_type="2D"):
        return len(self._candidate_points(sim_roi, simul_type, self.coord_center))

    def get_points_roi(self, sim_roi, probe_center=np.zeros((3,)), simul_type="2D", dir="e"):
        if not isinstance(dir, str):
            raise ValueError("'dir' must be a string")
        if (dir.lower() == "e" and not self.tx_en) or (dir.lower() == "r" and not self.rx_en):
            return list()
        center = self.coord_center + np.asarray(probe_center).astype(flt32).flatten()
        return self._candidate_points(sim_roi, simul_type, center)


def _t0_per_element(t0, n, what):
    if t0 is None:
        return np.zeros(n, dtype=np.float32)
    if isinstance(t0, (float, np.float32)):
        return np.ones(n, dtype=np.float32) * np.float32(t0)
    if isinstance(t0, list):
        t0 = list(t0)[:n] + [0.0] * max(0, n - len(t0))
        return np.array(t0, dtype=np.float32)
    if isinstance(t0, np.ndarray) and what == "set":
        return t0
    raise ValueError(f"{what} must be either a float [numpy.float32] or a list of floats.")


def _flags(spec, n, what):
    if isinstance(spec, str) and spec == "all":
        return [True] * n
    if isinstance(spec, str) and spec == "none":
        return [False] * n
    if isinstance(spec, list):
        f = [_parse_flag(s) for s in spec][:n]
        return f + [False] * (n - len(f))
    raise ValueError(f"{what} must be a string or a list of strings")


class SimulationProbeLinearArray(SimulationProbe):
    """Linear array of ``num_elem`` rectangular elements at pitch dim_a + inter_elem."""

    def __init__(self, coord_center=np.zeros((1, 3)), dec=(2, 2, 2), num_elem=32, dim_a=0.5, dim_p=10.0,
                 inter_elem=0.1, freq=5., bw=0.5, gain=1.0, pulse_type="gaussian", id="",
                 emitters="all", receivers="all", t0_emission=None, t0_reception=None):
        super().__init__(coord_center, dec)
        self.id = id
        self.inter_elem = np.float32(inter_elem)
        self.emitters = _flags(emitters, num_elem, "emitters")
        self.receivers = _flags(receivers, num_elem, "receivers")
        self._t0_emission = _t0_per_element(t0_emission, num_elem, "t0_emission")
        self._t0_reception = _t0_per_element(t0_reception, num_elem, "t0_reception")
        self.pitch = np.float32(dim_a + inter_elem)
        self.num_elem = num_elem
        # Element centres relative to the array centre (simul_classes.py:781-789).
        centre_shift = np.array([((num_elem - 1) * self.pitch + dim_a) / 2.0, 0.0, 0.0], dtype=np.float32)
        self.elem_list = [
            ElementRect(dim_a=dim_a, dim_p=dim_p,
                        coord_center=np.round(np.array([dim_a / 2.0 + i * self.pitch, 0.0, 0.0],
                                                       dtype=np.float32) - centre_shift,
                                              decimals=np.max(self._dec)),
                        freq=freq, bw=bw, gain=gain, pulse_type=pulse_type,
                        tx_en=self.emitters[i], rx_en=self.receivers[i],
                        t0=np.float32(self._t0_emission[i]))
            for i in range(num_elem)]
        self._dim_a, self._dim_p = dim_a, dim_p
        self._freq, self._bw, self._gain, self._pulse_type = freq, bw, gain, pulse_type

    def get_freq(self, mode='common'):
        if mode == 'mean':
            return np.array([e.freq for e in self.elem_list]).mean()
        if mode == 'max':
            return np.array([e.freq for e in self.elem_list]).max()
        return self._freq

    def _element_points(self, sim_roi, simul_type, dir):
        pts, owner = [], []
        for i, e in enumerate(self.elem_list):
            try:
                ep = e.get_points_roi(sim_roi=sim_roi, probe_center=self.coord_center,
                                      simul_type=simul_type, dir=dir)
            except IndexError:          # element outside the ROI contributes nothing
                continue
            pts += ep
            owner += [i] * len(ep)
        return pts, owner

    def get_points_roi(self, sim_roi=None, simul_type="2D", dir="e"):
        return self._element_points(sim_roi, simul_type, dir)

    def get_source_term(self, samples=1000, dt=1.0, out='r'):
        """(samples, num_elem) float32; silent columns for non-emitting elements."""
        t = _sample_times(samples, dt)
        st = np.zeros((samples, self.num_elem), dtype=np.float32)
        for i, e in enumerate(self.elem_list):
            if e.tx_en:
                st[:, i] = e.get_element_exc_fn(t, out)
        return st

    def get_idx_rec(self, sim_roi=None, simul_type="2D"):
        return self._element_points(sim_roi, simul_type, 'r')[1]

    def get_delay_rx(self):
        return [self._t0_reception[i] for i, e in enumerate(self.elem_list) if e.rx_en]

    def set_t0(self, t0_emission=None):
        """Replace the emission delays (the source term must then be rebuilt)."""
        self._t0_emission = _t0_per_element(t0_emission, self.num_elem, "set")
        for i, e in enumerate(self.elem_list):
            e.t0 = self._t0_emission[i]


class SimulationProbePoint(SimulationProbe):
    """Infinitesimal point transducer (simul_classes.py:967-1183)."""

    def __init__(self, coord_center=np.zeros((1, 3)), dec=(2, 2, 2), freq=5., bw=0.5, gain=1.0,
                 pulse_type="gaussian", id="", emitter="True", receiver="False",
                 t0_emission=None, t0_reception=None):
        super().__init__(coord_center, dec)
        self._id = id
        self.num_elem = 1
        if not isinstance(emitter, str):
            raise ValueError("emitter must be a string")
        if not isinstance(receiver, str):
            raise ValueError("receiver must be a string")
        self.emitters = [_parse_flag(emitter)]
        self.receivers = [_parse_flag(receiver)]
        self._t0_emission = np.float32(0.0 if t0_emission is None else t0_emission)
        self._t0_reception = np.float32(0.0 if t0_reception is None else t0_reception)
        self._freq, self._bw, self._gain, self._pulse_type = freq, bw, gain, pulse_type

    def get_freq(self, mode='common'):
        return self._freq

    def get_coords(self):
        return self.coord_center

    def get_points_roi(self, sim_roi=None, simul_type="2D", dir="e"):
        if not isinstance(dir, str):
            raise ValueError("'dir' must be a string")
        if (dir.lower() == "e" and not self.emitters[0]) or (dir.lower() == "r" and not self.receivers[0]):
            return list(), list()
        pt = np.array([np.float32(self.coord_center[0]),
                       np.float32(0.0 if simul_type.lower() == "2d" else self.coord_center[1]),
                       np.float32(self.coord_center[2])], np.float32)
        p0 = sim_roi.get_nearest_grid_idx(pt)
        return [[p0[0], p0[1], p0[2]]], [0]

    def get_source_term(self, samples=1000, dt=1.0, out='r', ord_der=1):
        """gain * derivative-of-Gaussian pulse on an un-rounded time axis (:1105-1141)."""
        t = np.arange(samples, dtype=flt32) * dt
        ss = np.zeros(samples, dtype=flt32)
        if self.emitters[0]:
            fn = {0: gaussian_pulse, 2: gaussian_second_dev_pulse}.get(ord_der, gaussian_first_dev_pulse)
            gp, _, egp = fn((t - self._t0_emission), fc=self._freq, bw=self._bw, retquad=True, retenv=True)
            eps = np.finfo(flt32).eps
            if out == 'e':
                egp[np.abs(egp) < eps] = 0.0
                ss = flt32(egp)
            else:
                gp[np.abs(gp) < eps] = 0.0
                ss = flt32(gp)
        return flt32(self._gain * ss)

    def get_idx_rec(self, sim_roi=None, simul_type="2D"):
        return [0] if self.receivers[0] else list()

    def get_delay_rx(self):
        return [self._t0_reception] if self.receivers[0] else list()

    def set_t0(self, t0_emission=None):
        if t0_emission is None:
            self._t0_emission = np.float32(0.0)
        elif isinstance(t0_emission, (float, np.float32)):
            self._t0_emission = t0_emission
        else:
            raise ValueError("t0_emission must be either a float [numpy.float32].")

    def get_receiver_points_count(self):
        return 1 if self.receivers[0] else 0


class SimulationProbeMonoCirc(SimulationProbe):
    """Single circular element transducer (simul_classes.py:1185-1343)."""

    def __init__(self, coord_center=np.zeros((1, 3)), dec=(2, 2, 2), radius=0.25, freq=5., bw=0.5,
                 gain=1.0, pulse_type="gaussian", id="", emitters="True", receivers="False",
                 t0_emission=None, t0_reception=None):
        super().__init__(coord_center, dec)
        self._id = id
        self.emitters = [_parse_flag(emitters)]
        self.receivers = [_parse_flag(receivers)]
        self._t0_reception = np.float32(0.0 if t0_reception is None else t0_reception)
        self._t0_emission = np.float32(0.0 if t0_emission is None else t0_emission)
        self.elem = ElementCirc(radius=radius, coord_center=np.zeros(3, dtype=np.float32), freq=freq,
                                bw=bw, gain=gain, t0=self._t0_emission, tx_en=self.emitters[0],
                                rx_en=self.receivers[0], pulse_type=pulse_type)
        self._gain = gain

    def get_freq(self, mode='common'):
        return self.elem.freq

    def get_coords(self):
        return self.coord_center

    def get_points_roi(self, sim_roi=None, simul_type="3D", dir="e"):
        try:
            pts = self.elem.get_points_roi(sim_roi=sim_roi, probe_center=self.coord_center,
                                           simul_type=simul_type, dir=dir)
        except IndexError:
            return [], []
        return pts, [0] * len(pts)

    def get_source_term(self, samples=1000, dt=1.0, out='r'):
        t = _sample_times(samples, dt)
        st = np.zeros((samples, 1), dtype=np.float32)
        if self.elem.tx_en:
            st[:, 0] = self.elem.get_element_exc_fn(t, out)
        return st

    def get_idx_rec(self, sim_roi=None, simul_type="3D"):
        return self.get_points_roi(sim_roi, simul_type, 'r')[1]

    def get_delay_rx(self):
        return [self._t0_reception] if self.receivers[0] else list()

    def set_t0(self, t0_emission=None):
        if t0_emission is None:
            self._t0_emission = np.float32(0.0)
        elif isinstance(t0_emission, (float, int, np.number)):
            self._t0_emission = np.float32(t0_emission)
        elif isinstance(t0_emission, (list, np.ndarray)):
            self._t0_emission = np.float32(t0_emission[0])
        else:
            raise ValueError("t0_emission must be either a float [numpy.float32] or a list of floats.")
        self.elem.t0 = self._t0_emission

    def get_receiver_points_count(self, simul_type="2D", sim_roi=None):
        return len(self.get_idx_rec(sim_roi=sim_roi, simul_type=simul_type))
