"""CPU oracle of the reference time loop -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this package, and only as the checker or the timed CPU
baseline; the product (``taichi_sim_amd``) never does.

``oracle.run(setup, n_steps, variant, precision)`` executes the reference's
update order (vse_oracle_body.h) on the arrays of a setup object exposing the
reference's ``Simulator`` attributes (``_nx``, ``_rho_grid_vx``, ``_a_x`` ...).

Pinning (each function is checked by at least one of these in tests/):
* 2D Taichi variant, end to end: the reference's golden B-scan
  pedro/results/result_teste_aco.npy (tests/test_oracle_golden.py);
* all variants: zero source -> identically zero fields, linearity in the
  source, FP32 vs FP64 floor, elastic limit of the Zener path with tau_eps =
  tau_sig, mirror symmetry of a centred source (tests/test_oracle_invariants.py);
* 2D WGSL, 2D Taichi and 3D, every update term: closed forms on synthetic
  inputs (tests/test_oracle_closed_form.py) -- one step from cubic velocity or
  stress fields, graded materials, random CPML profiles and random initial
  psi, r and stresses (each staggered difference equals the derivative at the
  half cell, each material mean the field at the centre, psi' = b psi + a D
  with the half profile for D+ and the full one for D-), and the geometric
  closed form of the Zener memory variables under a uniform strain rate over
  many steps; a mistake planted in this oracle (a D+/D- swap, a coefficient
  order, a half/full profile, an average, a Zener coefficient) fails them;
* 2D WGSL, 2D Taichi and 3D, many steps: plane waves along each axis, every
  FD order, against the closed form of the staggered leapfrog per Fourier
  mode with the dispersion relation sin(w dt/2) = c dt K(k)/2
  (tests/test_oracle_dispersion.py); a D+/D- swap or reversed coefficients
  planted in this oracle fail them.
  No numeric output of the WGSL backends survives (they need wgpu), so the
  expected formulas are the reference kernels' text.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libvse_oracle.so")
VARIANT = {"wgsl": 0, "taichi": 1}
PSI_COUNT = 18
MAX_ORD = 8
MAX_SLS = 8


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("ord", ctypes.c_int), ("coefs", ctypes.c_float * MAX_ORD),
        ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("dz", ctypes.c_float), ("dt", ctypes.c_float),
        ("variant", ctypes.c_int), ("visco", ctypes.c_int), ("n_sls", ctypes.c_int),
        ("tau", ctypes.c_float * (4 * MAX_SLS)), ("sum_alpha", ctypes.c_float * 2),
        ("rho", ctypes.c_void_p), ("cp", ctypes.c_void_p), ("cs", ctypes.c_void_p),
        ("prof", ctypes.c_void_p * 18),
        ("src", ctypes.c_void_p), ("n_iter", ctypes.c_int), ("src_stride", ctypes.c_int),
        ("src_map", ctypes.c_void_p),
        ("n_rec_el", ctypes.c_int), ("n_rec_pt", ctypes.c_int),
        ("info_rec_pt", ctypes.c_void_p), ("offset_sensors", ctypes.c_void_p),
        ("delay_rec", ctypes.c_void_p), ("rec_map", ctypes.c_void_p),
        ("rec_channel", ctypes.c_int), ("nthreads", ctypes.c_int),
    ]


# receiver channels, in the state's field order (vse_oracle.h ORC_CH_*)
CHANNEL = {"vx": 0, "vy": 1, "vz": 2, "sxx": 3, "syy": 4, "szz": 5, "sxy": 6, "sxz": 7, "syz": 8}


class OracleState(ctypes.Structure):
    _fields_ = [("vx", ctypes.c_void_p), ("vy", ctypes.c_void_p), ("vz", ctypes.c_void_p),
                ("sxx", ctypes.c_void_p), ("syy", ctypes.c_void_p), ("szz", ctypes.c_void_p),
                ("sxy", ctypes.c_void_p), ("sxz", ctypes.c_void_p), ("syz", ctypes.c_void_p),
                ("psi", ctypes.c_void_p * PSI_COUNT), ("r", ctypes.c_void_p * 6),
                ("traces", ctypes.c_void_p), ("v2max", ctypes.c_void_p)]


_lib = None


def build():
    """Compile the oracle with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        assert _lib.oracle_abi_version() == 2, "stale oracle build: run oracle.build()"
        _lib.oracle_params_size.restype = ctypes.c_size_t
        assert _lib.oracle_params_size() == ctypes.sizeof(OracleParams), "OracleParams layout mismatch"
        for sfx in ("f32", "f64"):
            f = getattr(_lib, "oracle_run_" + sfx)
            f.argtypes = [ctypes.POINTER(OracleParams), ctypes.POINTER(OracleState), ctypes.c_int, ctypes.c_int]
            f.restype = ctypes.c_int
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


class Oracle:
    """State + parameters of one oracle run.  ``sim`` exposes the reference's
    setup attributes; ``tau``/``sum_alpha`` override the reference's Zener
    constants (AttenuationCoefficients); ``src_stride`` defaults to the
    reference's own choice per variant."""

    def __init__(self, sim, variant="wgsl", precision="f32", visco=None, att=None, n_threads=None,
                 src_stride=None, channel=None):
        self.ndim = 3 if hasattr(sim, "_rho_grid_vx") and sim._rho_grid_vx.ndim == 3 else 2
        self.dtype = np.float32 if precision == "f32" else np.float64
        self.sfx = precision
        self.variant = variant
        self.keep = []
        p = OracleParams()
        p.ndim = self.ndim
        p.nx, p.ny = int(sim._nx), int(sim._ny)
        p.nz = int(sim._nz) if self.ndim == 3 else 1
        coefs = np.asarray(sim._coefs, dtype=np.float32)
        p.ord = len(coefs)
        for i, c in enumerate(coefs):
            p.coefs[i] = float(c)
        if self.ndim == 2:
            p.dx, p.dy, p.dz = float(sim._roi.get_dx()), float(sim._roi.get_dz()), 1.0
        else:
            p.dx, p.dy, p.dz = float(sim._roi.get_dx()), float(sim._roi.get_dy()), float(sim._roi.get_dz())
        p.dt = float(sim._dt)
        p.variant = VARIANT[variant]
        if visco is None:
            visco = bool(sim._configs.get("simul_configs", {}).get("viscoelastic_attn", 0))
        p.visco = int(bool(visco))
        if att is None:
            from taichi_sim_amd.sim_support.attenuation import AttenuationCoefficients
            att = AttenuationCoefficients()
        tau = att.tau_table()
        p.n_sls = tau.shape[1]
        for i, v in enumerate(tau.reshape(-1)):
            p.tau[i] = float(v)
        sa = att.sum_alpha()
        p.sum_alpha[0], p.sum_alpha[1] = float(sa[0]), float(sa[1])

        def keep32(a, dt=np.float32):
            a = np.ascontiguousarray(np.asarray(a), dtype=dt)
            self.keep.append(a)
            return a
        p.rho = _ptr(keep32(sim._rho_grid_vx))
        p.cp = _ptr(keep32(sim._cp_grid_vx))
        p.cs = _ptr(keep32(sim._cs_grid_vx))
        axes = ("x", "y", "z")[:self.ndim]
        for ai, ax in enumerate(axes):
            for si, nm in enumerate((f"_a_{ax}", f"_b_{ax}", f"_k_{ax}",
                                     f"_a_{ax}_half", f"_b_{ax}_half", f"_k_{ax}_half")):
                p.prof[ai * 6 + si] = _ptr(keep32(np.asarray(getattr(sim, nm)).reshape(-1)))
        src = keep32(sim._source_term)
        p.src = _ptr(src)
        p.n_iter = int(sim._n_steps)
        if src_stride is None:
            if variant == "taichi":
                src_stride = 1 if src.ndim == 1 else src.shape[1]
            else:
                src_stride = int(sim._n_src)      # get_source_term's stride (wgsl2d :67-71)
        p.src_stride = int(src_stride)
        p.src_map = _ptr(keep32(sim._pos_sources, np.int32))
        p.n_rec_el = int(sim._n_rec)
        p.n_rec_pt = int(sim._n_pto_rec)
        p.info_rec_pt = _ptr(keep32(sim._info_rec_pt, np.int32))
        p.offset_sensors = _ptr(keep32(sim._offset_sensors, np.int32))
        p.delay_rec = _ptr(keep32(sim._delay_recv, np.int32))
        p.rec_map = _ptr(keep32(sim._pos_sensors, np.int32))
        if channel is None:
            channel = "syy" if self.ndim == 2 else "szz"
        if self.ndim == 2 and channel not in ("vx", "vy", "sxx", "syy", "sxy"):
            raise ValueError(f"2D receivers cannot record {channel}")
        p.rec_channel = CHANNEL[channel]
        p.nthreads = int(n_threads or os.cpu_count() or 1)
        self.p = p

        shape = (p.nx, p.ny) if self.ndim == 2 else (p.nx, p.ny, p.nz)
        z = lambda *s: np.zeros(s if s else shape, dtype=self.dtype)
        self.f = {k: z() for k in ("vx", "vy", "sxx", "syy", "sxy")}
        if self.ndim == 3:
            for k in ("vz", "szz", "sxz", "syz"):
                self.f[k] = z()
        self.psi = [z() for _ in range(PSI_COUNT)]
        L = p.n_sls
        self.r = [z(*shape, L) for _ in range(6)]
        self.traces = z(p.n_iter, max(p.n_rec_el, 1))
        self.v2max = np.zeros(1, dtype=self.dtype)
        st = OracleState()
        for k in ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz"):
            setattr(st, k, _ptr(self.f.get(k)))
        for i in range(PSI_COUNT):
            st.psi[i] = _ptr(self.psi[i])
        for i in range(6):
            st.r[i] = _ptr(self.r[i])
        st.traces = _ptr(self.traces)
        st.v2max = _ptr(self.v2max)
        self.st = st
        self.step = 0

    def run(self, n_steps):
        fn = getattr(lib(), "oracle_run_" + self.sfx)
        rc = fn(ctypes.byref(self.p), ctypes.byref(self.st), self.step, int(n_steps))
        if rc != 0:
            raise RuntimeError(f"oracle failed rc={rc}")
        self.step += int(n_steps)
        return self

    def field(self, name):
        return self.f[name]


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))
