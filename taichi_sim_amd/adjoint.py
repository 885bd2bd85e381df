"""Gradient of the receiver misfit of a 2D viscoelastic simulation.

The reference only runs the forward time loop (sim_webgpu_viscoelastic.py
:469-543).  This module differentiates it: ``VSEGrad`` drives the checkpointed
forward and the hand-written adjoint kernels of libvse (vse_grad_*), which
return FP64 gradients with respect to the coefficients the kernels use
(lambda and mu at the normal-stress point, mu at the shear point, rho at the
vx and vy points, a per-cell relaxation-time scale).  ``coefficients`` forms
those coefficients from per-cell rho, cp, cs exactly as the kernels do
(shader_webgpu_viscoelastic.wgsl:863-868, :948-951, :1019, :1035-1036), so
torch autograd through it chains the kernel-level gradient to rho, cp, cs --
and from there to any parameterisation, e.g. (rho, lambda, mu).

``ViscoTraces`` is the torch.autograd.Function: traces = f(rho, cp, cs, tscale).
``misfit_and_gradient`` is the fused path: the misfit 0.5*||traces - obs||^2
is accumulated in FP64 on the GPU and its residual drives the adjoint
without a round trip through the host.
"""
import ctypes

import numpy as np

from . import _lib

try:        # bind libvse to the ROCm runtime it was built with before torch loads its own copy
    _lib.load()
except ImportError:
    pass    # not built: VSEGrad raises when used
import torch  # noqa: E402

from .engine import VSESim  # noqa: E402
from .sim_support import StabilityError

GRAD_NAMES = ("lambda", "mu", "mu_xy", "rho_vx", "rho_vy", "tau")


def _shift(f, dx, dy):
    """g[x, y] = f[x + dx, y + dy], zero outside the grid."""
    nx, ny = f.shape
    g = torch.zeros_like(f)
    g[:nx - dx, :ny - dy] = f[dx:, dy:]
    return g


def coefficients(rho, cp, cs):
    """(lambda, mu, mu_xy, rho_vx, rho_vy) per cell from the cell maps, as the
    kernels form them: 2-point averages towards x+1 (normal stress) and y+1
    (shear), with mu = 0 where either cs is 0, and the 4-point rho of vy."""
    rx, cpx, csx = _shift(rho, 1, 0), _shift(cp, 1, 0), _shift(cs, 1, 0)
    rho_h = 0.5 * (rx + rho)
    cp_h = 0.5 * (cpx + cp)
    cs_l = 0.5 * (csx + cs)
    cs_m = torch.where(torch.minimum(csx, cs) == 0, torch.zeros_like(cs_l), cs_l)
    lam = rho_h * (cp_h * cp_h - 2.0 * cs_l * cs_l)
    mu = rho_h * (cs_m * cs_m)
    ry, csy = _shift(rho, 0, 1), _shift(cs, 0, 1)
    cs_h = torch.where(torch.minimum(csy, cs) == 0, torch.zeros_like(cs), 0.5 * (csy + cs))
    mu_xy = 0.5 * (ry + rho) * (cs_h * cs_h)
    rho_vy = 0.25 * (rho + rx + _shift(rho, 1, 1) + ry)
    return lam, mu, mu_xy, rho, rho_vy


class VSEGrad:
    """Checkpointed forward + adjoint of one 2D viscoelastic ``VSESim`` (WGSL
    variant).  ``seg_len`` = steps between checkpoints (0: ceil(sqrt(n_rows)))."""

    def __init__(self, sim, seg_len=0):
        if sim.ndim != 2 or sim.variant != "wgsl" or not sim.visco:
            raise ValueError("gradients need a 2D viscoelastic simulation with the WGSL semantics")
        if getattr(sim, "_h", None) is None:
            raise ValueError("the simulation is closed")
        self.sim = sim
        self._lib = sim._lib
        h = ctypes.c_void_p()
        rc = self._lib.vse_grad_create(sim._h, int(seg_len), ctypes.byref(h))
        if rc != 0 or not h.value:
            raise _lib.VSEError(rc, sim._lib.vse_last_error(sim._h).decode(errors="replace"))
        self._h = h
        self.n_steps = 0
        self.forward_count = 0          # forwards run so far (ViscoTraces checks it in backward)
        self.ms = {}
        sim._dependents.add(self)       # closing the simulation closes this engine first

    def _check(self, rc):
        if rc == 0:
            return
        msg = self._lib.vse_grad_last_error(self._h).decode(errors="replace")
        if rc == _lib.VSE_ERR_UNSTABLE:
            raise StabilityError("Simulacao tornando-se instavel", msg)
        raise _lib.VSEError(rc, msg)

    @property
    def seg_len(self):
        return int(self._lib.vse_grad_segment(self._h))

    def set_flush(self, every):
        """Steps per FP64 flush of the FP32 per-step gradient partials (default
        32; 1 accumulates every step's terms in FP64)."""
        self._check(self._lib.vse_grad_set_flush(self._h, int(every)))

    def set_overlap(self, on=True):
        """With sqrt(N) checkpoints, recompute the next segment on a second
        stream while the adjoint of the current one runs (default), or each
        segment before its adjoint on one stream.  The gradient is the same."""
        self._check(self._lib.vse_grad_set_overlap(self._h, 1 if on else 0))

    def device_bytes(self):
        return int(self._lib.vse_grad_device_bytes(self._h))

    def set_model(self, rho, cp, cs, tscale=None):
        """Per-cell maps (nx, ny); tscale None = 1 everywhere."""
        f = lambda a: np.ascontiguousarray(a.detach().cpu().numpy() if torch.is_tensor(a) else a, dtype=np.float32)
        self.sim.set_material(f(rho), f(cp), f(cs))
        ts = None if tscale is None else f(tscale)
        if ts is not None and ts.shape != self.sim.shape:
            raise ValueError(f"tau scale must have shape {self.sim.shape}")
        self._check(self._lib.vse_grad_set_tau_scale(self._h, None if ts is None else ts.ctypes.data_as(
            ctypes.POINTER(ctypes.c_float))))

    def forward(self, n_steps=None, check_every=100):
        """(n_steps, n_rec) float32 traces of the checkpointed forward."""
        n = int(n_steps or self.sim.n_rows)
        out = np.zeros((n, max(self.sim.n_rec, 1)), dtype=np.float32)
        ms = ctypes.c_float(0.0)
        self.forward_count += 1
        self._check(self._lib.vse_grad_forward(self._h, n, int(check_every),
                                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(ms)))
        self.n_steps = n
        self.ms["forward"] = ms.value
        return out[:, :self.sim.n_rec]

    def misfit(self, observed):
        """0.5 * sum((traces - observed)^2), accumulated in FP64 on the GPU; the
        residual becomes the adjoint source of the next ``backward()``."""
        obs = np.ascontiguousarray(observed, dtype=np.float32)
        if obs.shape != (self.n_steps, self.sim.n_rec):
            raise ValueError(f"observed traces must be {(self.n_steps, self.sim.n_rec)}")
        j = ctypes.c_double(0.0)
        self._check(self._lib.vse_grad_misfit(self._h, obs.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                              ctypes.byref(j)))
        return j.value

    def backward(self, grad_traces=None):
        """Reverse sweep; returns {name: FP64 (nx, ny) map} for GRAD_NAMES.
        grad_traces = d misfit / d traces (None: the residual of ``misfit()``)."""
        ms = ctypes.c_float(0.0)
        if grad_traces is None:
            ptr = None
        else:
            gt = np.ascontiguousarray(grad_traces, dtype=np.float32)
            if gt.shape != (self.n_steps, self.sim.n_rec):
                raise ValueError(f"grad_traces must be {(self.n_steps, self.sim.n_rec)}")
            ptr = gt.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        self._check(self._lib.vse_grad_backward(self._h, ptr, ctypes.byref(ms)))
        self.ms["backward"] = ms.value
        out = {}
        for i, name in enumerate(GRAD_NAMES):
            a = np.empty(self.sim.shape, dtype=np.float64)
            self._check(self._lib.vse_grad_read(self._h, i, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
            out[name] = a
        return out

    def state(self, field):
        """Field of the last forward's final state ('vx', 'vy', 'sxx', 'syy', 'sxy')."""
        a = np.empty(self.sim.shape, dtype=np.float32)
        self._check(self._lib.vse_grad_read_state(self._h, ("vx", "vy", "sxx", "syy", "sxy").index(field),
                                                  a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return a

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.vse_grad_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def chain_to_cells(coef_grads, rho, cp, cs):
    """Kernel-level gradients -> d/d (rho, cp, cs) per cell (FP64 torch)."""
    r, c, s = (torch.as_tensor(np.asarray(v, dtype=np.float64) if not torch.is_tensor(v) else v.detach())
               .double().clone().requires_grad_(True) for v in (rho, cp, cs))
    lam, mu, mu_xy, rho_vx, rho_vy = coefficients(r, c, s)
    G = {k: torch.as_tensor(coef_grads[k], dtype=torch.float64) for k in GRAD_NAMES}
    total = ((lam * G["lambda"]).sum() + (mu * G["mu"]).sum() + (mu_xy * G["mu_xy"]).sum()
             + (rho_vx * G["rho_vx"]).sum() + (rho_vy * G["rho_vy"]).sum())
    gr, gc, gs = torch.autograd.grad(total, (r, c, s), allow_unused=True)
    z = lambda g, like: torch.zeros_like(like) if g is None else g
    return z(gr, r), z(gc, c), z(gs, s)


class ViscoTraces(torch.autograd.Function):
    """traces (n_steps, n_rec) = forward(rho, cp, cs, tscale) through libvse;
    backward runs the adjoint.  Inputs are (nx, ny) tensors (any float dtype,
    CPU); tscale may be None (every relaxation time at scale 1).  The
    wavefields are FP32 on the GPU, the traces are returned as FP64.

    The engine keeps the checkpoints of its latest forward only, so a node's
    backward must come before the engine's next forward; backward raises
    otherwise instead of differentiating the other model."""

    @staticmethod
    def forward(ctx, eng, rho, cp, cs, tscale):
        eng.set_model(rho, cp, cs, tscale)
        tr = eng.forward()
        ctx.eng = eng
        ctx.forward_id = eng.forward_count
        ctx.save_for_backward(rho, cp, cs)
        return torch.from_numpy(tr.astype(np.float64))

    @staticmethod
    def backward(ctx, g_tr):
        eng = ctx.eng
        if eng.forward_count != ctx.forward_id:
            raise RuntimeError("ViscoTraces: the engine ran another forward since this node's; its checkpoints "
                               "are gone -- call backward before the next apply() on the same engine")
        rho, cp, cs = ctx.saved_tensors
        grads = eng.backward(g_tr.detach().cpu().numpy())
        need = ctx.needs_input_grad
        gr = gc = gs = gt = None
        if any(need[1:4]):
            with torch.enable_grad():
                gr, gc, gs = chain_to_cells(grads, rho, cp, cs)
            gr, gc, gs = (g.to(like.dtype) if n else None for g, like, n in zip((gr, gc, gs), (rho, cp, cs), need[1:4]))
        if need[4]:
            gt = torch.as_tensor(grads["tau"], dtype=torch.float64)
        return None, gr, gc, gs, gt


def velocities(rho, lam, mu):
    """(cp, cs) from (rho, lambda, mu): cp = sqrt((lambda + 2 mu)/rho), cs = sqrt(mu/rho); 0 where rho = 0."""
    safe = torch.where(rho > 0, rho, torch.ones_like(rho))
    cp = torch.where(rho > 0, torch.sqrt(torch.clamp((lam + 2.0 * mu) / safe, min=0.0)), torch.zeros_like(rho))
    cs = torch.where(rho > 0, torch.sqrt(torch.clamp(mu / safe, min=0.0)), torch.zeros_like(rho))
    return cp, cs


def misfit_and_gradient(eng, observed, rho, lam, mu, tscale=None):
    """J = 0.5 ||traces - observed||^2 (FP64 on the GPU) and dJ/d(rho, lambda,
    mu, tscale) per cell (FP64), for the model given as (rho, lambda, mu)."""
    r = torch.as_tensor(rho, dtype=torch.float64).clone().requires_grad_(True)
    l = torch.as_tensor(lam, dtype=torch.float64).clone().requires_grad_(True)
    m = torch.as_tensor(mu, dtype=torch.float64).clone().requires_grad_(True)
    with torch.no_grad():
        cp, cs = velocities(r, l, m)
    eng.set_model(r.detach(), cp, cs, tscale)
    eng.forward()
    J = eng.misfit(observed)
    grads = eng.backward()
    with torch.enable_grad():
        cp, cs = velocities(r, l, m)
        lam_k, mu_k, mu_xy, rho_vx, rho_vy = coefficients(r, cp, cs)
        G = {k: torch.as_tensor(grads[k], dtype=torch.float64) for k in GRAD_NAMES}
        total = ((lam_k * G["lambda"]).sum() + (mu_k * G["mu"]).sum() + (mu_xy * G["mu_xy"]).sum()
                 + (rho_vx * G["rho_vx"]).sum() + (rho_vy * G["rho_vy"]).sum())
        gr, gl, gm = torch.autograd.grad(total, (r, l, m), allow_unused=True)
    zero = lambda g, like: torch.zeros_like(like) if g is None else g
    return J, {"rho": zero(gr, r).numpy(), "lambda": zero(gl, l).numpy(), "mu": zero(gm, m).numpy(),
               "tau": grads["tau"], "ms": dict(eng.ms)}


def grad_engine(sim, att=None, seg_len=0):
    """VSEGrad over a new VSESim built from a Simulator (WGSL semantics, viscoelastic)."""
    g = VSESim.from_simulator(sim, variant="wgsl", visco=True, att=att)
    return VSEGrad(g, seg_len=seg_len)
