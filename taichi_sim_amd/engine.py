"""Python face of libvse: one device-resident simulation per ``VSESim``.

``VSESim.from_simulator(sim)`` takes the setup attributes the reference's
backends consume (``sim_webgpu_viscoelastic.py:57-191``,
``sim_taichi_elastic.py:41-132``) and hands them to libvse, converting the
reference's host tables into the library's forms:

* WGSL source indexing reads column ``e`` of row ``it`` at flat offset
  ``e + it*n_src`` where ``n_src`` counts emitting *elements* and returns 0
  for ``e >= n_src`` (wgsl2d :67-71); the source matrix handed to the library
  is that view of the flat buffer, so a partly-emitting array behaves as in
  the reference.  The Taichi variant indexes (n_steps, n_cols) directly.
* Receivers: the WGSL table walk (offset_sensors / info_rec_pt, wgsl2d
  :1125-1133) becomes a CSR per receiver; the Taichi variant uses the
  ``pos_sensors`` map in grid order, restricted to the normal-stress
  interior where the reference records (sim_taichi_elastic.py:150-190).
"""
import ctypes
import weakref

import numpy as np

from . import _lib
from .sim_support import StabilityError
from .sim_support.attenuation import AttenuationCoefficients


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def wgsl_source_matrix(source_term, n_rows, n_src):
    """(n_rows, max(n_src,1)) matrix M with M[it, e] = flat[e + it*n_src] (0 when out of range)."""
    flat = np.ascontiguousarray(source_term, dtype=np.float32).reshape(-1)
    cols = max(int(n_src), 1)
    m = np.zeros((n_rows, cols), dtype=np.float32)
    if n_src > 0:
        idx = np.arange(n_rows)[:, None] * n_src + np.arange(n_src)[None, :]
        ok = idx < flat.size
        m[:, :n_src][ok] = flat[idx[ok]]
    return m


def wgsl_receiver_csr(info_rec_pt, offset_sensors, n_rec, ndim, shape):
    """Per-receiver point lists, walking the table as the WGSL loop does:
    start at offset_sensors[r], continue while the row's id equals r."""
    info = np.asarray(info_rec_pt, dtype=np.int64).reshape(-1, ndim + 1)
    offs = np.asarray(offset_sensors, dtype=np.int64).reshape(-1)
    beg, cells = [0], []
    for r in range(n_rec):
        pt = int(offs[r]) if r < offs.size else -1
        while 0 <= pt < info.shape[0] and info[pt, ndim] == r:
            coords = info[pt, :ndim]
            if np.all(coords >= 0) and np.all(coords < np.array(shape)):
                cells.append(int(np.ravel_multi_index(tuple(coords), shape)))
            pt += 1
        beg.append(len(cells))
    return np.array(beg, dtype=np.int32), np.array(cells, dtype=np.int32)


def taichi_receiver_csr(pos_sensors, n_rec, ord_):
    """Receivers of the Taichi variant: map cells in C order, inside the
    normal-stress interior (x in [L, nx-L-1), y in [L+1, ny-L))."""
    nx, ny = pos_sensors.shape
    last = ord_ - 1
    sub = np.full_like(pos_sensors, -1)
    sub[last:nx - last - 1, last + 1:ny - last] = pos_sensors[last:nx - last - 1, last + 1:ny - last]
    flat = sub.reshape(-1)
    beg, cells = [0], []
    for r in range(n_rec):
        c = np.nonzero(flat == r)[0]
        cells += c.tolist()
        beg.append(len(cells))
    return np.array(beg, dtype=np.int32), np.array(cells, dtype=np.int32)


class VSESim:
    """A device-resident simulation.  All arrays in and out are NumPy, in the
    reference's layout."""

    def __init__(self, ndim, shape, spacing, dt, coefs, variant="wgsl", visco=False, att=None,
                 n_rows=1):
        self._lib = _lib.load()
        self.ndim = ndim
        self.shape = tuple(int(n) for n in shape)
        self.variant = variant
        self.n_rows = int(n_rows)
        att = att or AttenuationCoefficients()
        tau = att.tau_table()
        d = _lib.VseDesc()
        d.ndim = ndim
        d.nx, d.ny = self.shape[0], self.shape[1]
        d.nz = self.shape[2] if ndim == 3 else 1
        d.dx, d.dy = float(spacing[0]), float(spacing[1])
        d.dz = float(spacing[2]) if ndim == 3 else 1.0
        d.dt = float(dt)
        coefs = np.asarray(coefs, dtype=np.float32)
        d.ord = len(coefs)
        for i, c in enumerate(coefs):
            d.coefs[i] = float(c)
        d.variant = _lib.VARIANT_TAICHI if variant == "taichi" else _lib.VARIANT_WGSL
        d.visco = int(bool(visco))
        d.n_sls = tau.shape[1]
        for i, v in enumerate(tau.reshape(-1)):
            d.tau[i] = float(v)
        sa = att.sum_alpha()
        d.sum_alpha[0], d.sum_alpha[1] = float(sa[0]), float(sa[1])
        d.n_rows = self.n_rows
        self.ord = d.ord
        self.n_sls = d.n_sls
        self.visco = bool(visco)
        h = ctypes.c_void_p()
        rc = self._lib.vse_create(ctypes.byref(d), ctypes.byref(h))
        if rc != 0 or not h.value:
            raise _lib.VSEError(rc, "vse_create failed (invalid description or HIP error)")
        self._h = h
        self.n_rec = 0
        self.channel = "syy" if ndim == 2 else "szz"
        self.math = "strict"
        self.schedule = "two-pass"
        self._bands = {}
        self._dependents = weakref.WeakSet()   # open batches / gradient engines holding this handle

    # ---------------------------------------------------------------- helpers
    def _check(self, rc):
        if rc == 0:
            return
        msg = self._lib.vse_last_error(self._h).decode(errors="replace")
        if rc == _lib.VSE_ERR_UNSTABLE:
            raise StabilityError("Simulacao tornando-se instavel", msg)
        raise _lib.VSEError(rc, msg)

    def _grid(self, a, dtype=np.float32):
        a = np.ascontiguousarray(a, dtype=dtype)
        if a.shape != self.shape:
            raise ValueError(f"expected shape {self.shape}, got {a.shape}")
        return a

    # ------------------------------------------------------------------ setup
    def set_material(self, rho, cp, cs):
        rho, cp, cs = self._grid(rho), self._grid(cp), self._grid(cs)
        self._check(self._lib.vse_set_material(self._h, _fptr(rho), _fptr(cp), _fptr(cs)))

    def set_cpml(self, axis, a, b, k, a_h, b_h, k_h):
        n = self.shape[axis] - 2 * (self.ord - 1)
        arrs = [np.ascontiguousarray(np.asarray(v).reshape(-1), dtype=np.float32) for v in (a, b, k, a_h, b_h, k_h)]
        for v in arrs:
            if v.size != n:
                raise ValueError(f"CPML profile of axis {axis} must have {n} entries, got {v.size}")
        a_, _, k_, ah_, _, kh_ = arrs
        act = (a_ != 0) | (k_ != 1) | (ah_ != 0) | (kh_ != 1)
        lo = int(np.argmin(act)) if not act.all() else n
        hi = n - int(np.argmin(act[::-1])) if not act.all() else n
        if act[lo:hi].any():
            lo = hi = n
        self._bands[axis] = lo + (n - hi)
        self._check(self._lib.vse_set_cpml(self._h, axis, *[_fptr(v) for v in arrs]))

    def set_sources(self, cells, cols, src):
        cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1)
        cols = np.ascontiguousarray(cols, dtype=np.int32).reshape(-1)
        src = np.ascontiguousarray(src, dtype=np.float32)
        if src.ndim == 1:
            src = src[:, None]
        if src.shape[0] != self.n_rows or cells.size != cols.size:
            raise ValueError("source matrix must have n_rows rows and one column index per cell")
        self._check(self._lib.vse_set_sources(self._h, cells.size, _iptr(cells), _iptr(cols),
                                              src.shape[1], _fptr(src)))

    def set_receivers(self, beg, cells, delay, gate, channel=None):
        """Receiver CSR (see ``vse_set_receivers``).  channel: the field the
        receivers record ('syy', 'vy', ...; None = the reference's, sigma_yy in
        2D and sigma_zz in 3D).  In 2D the grid's y axis is the ROI's vertical
        z axis, so its vertical particle velocity is 'vy'."""
        beg = np.ascontiguousarray(beg, dtype=np.int32)
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        delay = np.ascontiguousarray(delay, dtype=np.int32)
        n_rec = beg.size - 1
        if delay.size != n_rec:
            raise ValueError("one delay per receiver")
        ch = -1 if channel is None else _lib.FIELD[channel]
        self._check(self._lib.vse_set_receivers(self._h, n_rec, _iptr(beg), _iptr(cells), _iptr(delay), int(gate),
                                                ch))
        self.n_rec = n_rec
        self.channel = channel or ("syy" if self.ndim == 2 else "szz")

    # -------------------------------------------------------------------- run
    def reset(self):
        self._check(self._lib.vse_reset(self._h))

    def run(self, n_steps, check_every=100, timed=False):
        ms = ctypes.c_float(0.0)
        self._check(self._lib.vse_run(self._h, int(n_steps), int(check_every),
                                      ctypes.byref(ms) if timed else None))
        return ms.value if timed else None

    def run_noalloc(self, n_steps, stream=None):
        """vse_run_noalloc: the same steps with no device allocation and no host
        synchronisation, on ``stream`` (a raw hipStream_t of libvse's own HIP
        runtime; None = the handle's stream).  No stability check: read
        ``v2max()`` afterwards (it synchronises)."""
        self._check(self._lib.vse_run_noalloc(self._h, int(n_steps), stream))

    def profile(self, n_steps):
        """Average (stress, velocity, sparse) kernel durations in ms over n_steps
        steps, each kernel bracketed by HIP events on the library's stream.  In
        2D the velocity kernel does the sources and receivers and the third
        value times the step-counter advance."""
        ms = (ctypes.c_float * 3)()
        self._check(self._lib.vse_profile(self._h, int(n_steps), ms))
        return tuple(float(v) for v in ms)

    def set_tuning(self, kernel, ty, nt, xchunk, remap=0):
        """Launch shape (ty, nt, xchunk, remap) of the stress (kernel 0) or
        velocity (1) pass, or of the fused 3D step (2); include/vse.h gives each
        field's meaning in 2D and 3D.  Results are identical for every shape."""
        self._check(self._lib.vse_set_tuning(self._h, int(kernel), int(ty), int(nt), int(xchunk), int(remap)))

    def set_schedule(self, sched):
        """'two-pass' (a stress and a velocity pass per step, the default) or
        'fused' (3D viscoelastic, 3 mechanisms, FD order 2: one kernel per step,
        bit-identical results)."""
        self._check(self._lib.vse_set_schedule(self._h, {"two-pass": 0, "fused": 1}[sched]))
        self.schedule = sched

    def set_math(self, mode):
        """'strict' (reference operation order, default) or 'fast'.

        'fast' multiplies by f32(coef/d) and f32(1/k) instead of dividing (one
        extra FP32 rounding per term).  In 2D it applies to the WGSL variant; the
        members of a batch must all use the same mode.
        """
        self._check(self._lib.vse_set_math(self._h, {"strict": 0, "fast": 1}[mode]))
        self.math = mode

    def band_cells(self):
        """Cells per psi slot-axis band: {axis: band width} from the CPML profiles."""
        return dict(self._bands)

    def field(self, name):
        out = np.empty(self.shape, dtype=np.float32)
        self._check(self._lib.vse_read_field(self._h, _lib.FIELD[name], _fptr(out)))
        return out

    def memory(self, comp):
        out = np.empty(self.shape + (self.n_sls,), dtype=np.float32)
        self._check(self._lib.vse_read_memory(self._h, _lib.MEMORY[comp], _fptr(out)))
        return out

    def traces(self):
        out = np.empty((self.n_rows, max(self.n_rec, 1)), dtype=np.float32)
        if self.n_rec:
            self._check(self._lib.vse_read_traces(self._h, _fptr(out)))
        return out[:, :self.n_rec]

    def v2max(self):
        v = ctypes.c_float(0.0)
        self._check(self._lib.vse_read_v2max(self._h, ctypes.byref(v)))
        return v.value

    def step(self):
        v = ctypes.c_int(0)
        self._check(self._lib.vse_get_step(self._h, ctypes.byref(v)))
        return v.value

    def device_name(self):
        return self._lib.vse_device_name(self._h).decode(errors="replace")

    def device_bytes(self):
        return int(self._lib.vse_device_bytes(self._h))

    def close(self):
        """Release the device state; a batch or gradient engine still holding
        this simulation is closed first (it would otherwise keep a dangling
        handle)."""
        for d in list(getattr(self, "_dependents", ())):
            d.close()
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.vse_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------ from a Simulator
    @classmethod
    def from_simulator(cls, sim, variant="wgsl", visco=None, att=None, channel=None):
        """Build from an object carrying the reference's Simulator attributes;
        channel: the receivers' field (None = the reference's σyy / σzz)."""
        ndim = 3 if np.asarray(sim._rho_grid_vx).ndim == 3 else 2
        if visco is None:
            visco = bool(sim._configs.get("simul_configs", {}).get("viscoelastic_attn", 0))
        if ndim == 2:
            shape = (int(sim._nx), int(sim._ny))
            spacing = (sim._roi.get_dx(), sim._roi.get_dz())
        else:
            shape = (int(sim._nx), int(sim._ny), int(sim._nz))
            spacing = (sim._roi.get_dx(), sim._roi.get_dy(), sim._roi.get_dz())
        n_rows = int(sim._n_steps)
        self = cls(ndim, shape, spacing, sim._dt, sim._coefs, variant=variant,
                   visco=visco and variant == "wgsl", att=att, n_rows=n_rows)
        self.set_material(sim._rho_grid_vx, sim._cp_grid_vx, sim._cs_grid_vx)
        for ai, ax in enumerate(("x", "y", "z")[:ndim]):
            self.set_cpml(ai, *[getattr(sim, f"_{k}_{ax}{h}") for h in ("", "_half") for k in ("a", "b", "k")])
        self.load_sources(sim)
        pos = np.asarray(sim._pos_sensors)
        if variant == "taichi":
            beg, cells = taichi_receiver_csr(pos, int(sim._n_rec), self.ord)
            gate = 1
        else:
            beg, cells = wgsl_receiver_csr(sim._info_rec_pt, sim._offset_sensors, int(sim._n_rec), ndim, shape)
            gate = 0
        delay = np.zeros(int(sim._n_rec), dtype=np.int32)
        dr = np.asarray(sim._delay_recv, dtype=np.int32).reshape(-1)
        delay[:min(dr.size, delay.size)] = dr[:delay.size]
        self.set_receivers(beg, cells, delay, gate, channel)
        self.reset()
        return self

    def load_sources(self, sim):
        """Source cells/columns from the pos_sources map and the source matrix."""
        pos = np.asarray(sim._pos_sources).reshape(-1)
        cells = np.nonzero(pos != -1)[0].astype(np.int32)
        cols = pos[cells].astype(np.int32)
        st = np.asarray(sim._source_term, dtype=np.float32)
        if self.variant == "taichi":
            src = st[:, None] if st.ndim == 1 else st
        else:
            src = wgsl_source_matrix(st, self.n_rows, int(sim._n_src))
            keep = cols < src.shape[1]          # columns >= n_src read 0 in the reference
            cells, cols = cells[keep], cols[keep]
        self.set_sources(cells, cols, src)


class VSEBatch:
    """Several 2D ``VSESim`` advanced together, one launch per pass per step
    over all of them (shots of one specimen, or specimens of different
    sizes).  Members keep their own state and results; each member ends
    exactly where ``member.run(n)`` alone would.  The batch does not own its
    members: close it before them."""

    def __init__(self, sims):
        self._lib = _lib.load()
        self.sims = list(sims)
        if not 1 <= len(self.sims) <= _lib.BATCH_MAX:
            raise ValueError(f"a batch holds 1..{_lib.BATCH_MAX} simulations, got {len(self.sims)}")
        if any(getattr(s, "_h", None) is None for s in self.sims):
            raise ValueError("batch members must be open simulations")
        arr = (ctypes.c_void_p * len(self.sims))(*[s._h.value for s in self.sims])
        h = ctypes.c_void_p()
        rc = self._lib.vse_batch_create(arr, len(self.sims), ctypes.byref(h))
        if rc != 0 or not h.value:
            raise _lib.VSEError(rc, "vse_batch_create failed: members must be open 2D simulations on one device "
                                    "with the same variant, viscoelastic flag, FD order and mechanism count")
        self._h = h
        self.schedule = "graph"
        for s in self.sims:
            s._dependents.add(self)

    def _open(self):
        if getattr(self, "_h", None) is None:
            raise ValueError("batch is closed")
        return self._h

    def run(self, n_steps, check_every=100, timed=False):
        self._open()
        ms = ctypes.c_float(0.0)
        rc = self._lib.vse_batch_run(self._h, int(n_steps), int(check_every), ctypes.byref(ms) if timed else None)
        if rc != 0:
            msg = self._lib.vse_batch_last_error(self._h).decode(errors="replace")
            if rc == _lib.VSE_ERR_UNSTABLE:
                raise StabilityError("Simulacao tornando-se instavel", msg)
            raise _lib.VSEError(rc, msg)
        return ms.value if timed else None

    def set_rows(self, rows):
        """x row-groups (of 4 rows) each workgroup walks; speed only."""
        rc = self._lib.vse_batch_set_rows(self._open(), int(rows))
        if rc != 0:
            raise _lib.VSEError(rc, self._lib.vse_batch_last_error(self._h).decode(errors="replace"))

    def set_group_bytes(self, nbytes):
        """Cache budget per group of members (0: all members in lockstep); speed only."""
        rc = self._lib.vse_batch_set_group_bytes(self._open(), int(nbytes))
        if rc != 0:
            raise _lib.VSEError(rc, self._lib.vse_batch_last_error(self._h).decode(errors="replace"))

    SCHEDULES = {"graph": 0, "persistent": 1}

    def set_schedule(self, schedule):
        """'graph' (default): hipGraphs of one stress and one velocity launch per
        step for each group.  'persistent': each group's steps as cooperative
        launches with grid-wide barriers between the passes.  Same results."""
        rc = self._lib.vse_batch_set_schedule(self._open(), self.SCHEDULES[schedule])
        if rc != 0:
            raise _lib.VSEError(rc, self._lib.vse_batch_last_error(self._h).decode(errors="replace"))
        self.schedule = schedule

    def groups(self):
        """First member index of each group."""
        n = self._lib.vse_batch_groups(self._open(), None, 0)
        out = (ctypes.c_int * max(n, 1))()
        self._lib.vse_batch_groups(self._h, out, n)
        return list(out[:n])

    def profile(self, n_steps):
        """Average (stress, velocity, step-counter advance) launch durations in ms
        of one step of all members in one launch per pass."""
        ms = (ctypes.c_float * 3)()
        rc = self._lib.vse_batch_profile(self._open(), int(n_steps), ms)
        if rc != 0:
            raise _lib.VSEError(rc, self._lib.vse_batch_last_error(self._h).decode(errors="replace"))
        return tuple(float(v) for v in ms)

    def __len__(self):
        return len(self.sims)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.vse_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
