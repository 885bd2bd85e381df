"""Differentiable FP64 restatement of the 2D viscoelastic time loop -- TEST
INFRASTRUCTURE (the oracle of the gradient; only tests/ import it).

Whole-grid torch operations in float64 with the update order of the WGSL 2D
viscoelastic backend (shader_webgpu_viscoelastic.wgsl:829-1141; the same
semantics as the C oracle's wgsl variant in vse_oracle_body.h), extended by a
per-cell relaxation-time scale ``tscale``: at a cell with scale s every
relaxation time of the Zener mechanisms is multiplied by s, i.e. 1/tau_sigma
becomes (1/tau_sigma)/s in phi and h (wgsl2d :881-894); s = 1 is the
reference.  torch.autograd through this loop gives the exact gradient of the
discrete FP64 forward with respect to rho, cp, cs and tscale.

Pinned by: the FP64 C oracle (tests/test_adjoint_oracle.py compares every
wavefield and trace at s = 1), and torch's own differentiation.
"""
import numpy as np
import torch


def _shift(f, dx, dy):
    """g[x, y] = f[x + dx, y + dy], zero outside the grid."""
    nx, ny = f.shape[-2:]
    g = torch.zeros_like(f)
    xs, xd = (slice(dx, nx), slice(0, nx - dx)) if dx >= 0 else (slice(0, nx + dx), slice(-dx, nx))
    ys, yd = (slice(dy, ny), slice(0, ny - dy)) if dy >= 0 else (slice(0, ny + dy), slice(-dy, ny))
    g[..., xd, yd] = f[..., xs, ys]
    return g


class Visco2D:
    """The reference's setup attributes of ``sim`` (a Simulator), tau table
    ``att`` (AttenuationCoefficients-like: tau_table(), sum_alpha())."""

    def __init__(self, sim, att, n_steps=None):
        dt64 = torch.float64
        self.nx, self.ny = int(sim._nx), int(sim._ny)
        nx, ny = self.nx, self.ny
        self.n_steps = int(n_steps if n_steps is not None else sim._n_steps)
        self.C = [float(np.float32(c)) for c in np.asarray(sim._coefs).reshape(-1)]
        self.ord = len(self.C)
        self.last = self.ord - 1
        self.dx = float(np.float32(sim._roi.get_dx()))
        self.dy = float(np.float32(sim._roi.get_dz()))
        self.dt = float(np.float32(sim._dt))
        tau = np.asarray(att.tau_table(), dtype=np.float32).astype(np.float64)   # (4, L)
        self.L = tau.shape[1]
        sa = np.asarray(att.sum_alpha(), dtype=np.float32).astype(np.float64)
        t_ep, t_sp, t_es, t_ss = tau
        self.inv_p, self.inv_s = 1.0 / t_sp, 1.0 / t_ss
        self.A_p = self.dt * (1.0 - t_ep / t_sp) / sa[0]
        self.A_s = self.dt * (1.0 - t_es / t_ss) / sa[1]

        def prof(axis_len, name):
            # profile of index j = cell - (ord-1), padded with a=b=0, k=1 elsewhere
            v = np.asarray(getattr(sim, name), dtype=np.float32).reshape(-1).astype(np.float64)
            out = np.zeros(axis_len)
            if name.startswith("_k"):
                out[:] = 1.0
            out[self.last:self.last + v.size] = v[:axis_len - self.last]
            return out

        P = {}
        for ax, n, tag in ((0, nx, "x"), (1, ny, "y")):
            for k in ("a", "b", "k"):
                for half in ("", "_half"):
                    arr = prof(n, f"_{k}_{tag}{half}")
                    t = torch.tensor(arr, dtype=dt64)
                    P[f"{k}{'h' if half else ''}{tag}"] = t[:, None] if ax == 0 else t[None, :]
        self.P = P
        X = torch.arange(nx)[:, None]
        Y = torch.arange(ny)[None, :]
        l = self.last
        self.mN = (X >= l) & (X < nx - l - 1) & (Y >= l + 1) & (Y < ny - l)
        self.mS = (X >= l + 1) & (X < nx - l) & (Y >= l) & (Y < ny - l - 1)
        self.mX = (X >= l + 1) & (X < nx - l) & (Y >= l + 1) & (Y < ny - l)
        self.mY = (X >= l) & (X < nx - l - 1) & (Y >= l) & (Y < ny - l - 1)

        # sources: WGSL indexing src[is + it*n_src], 0 when is >= n_src (wgsl2d :67-71)
        pos = np.asarray(sim._pos_sources).reshape(nx, ny)
        n_src = int(sim._n_src)
        flat = np.asarray(sim._source_term, dtype=np.float32).reshape(-1).astype(np.float64)
        cx, cy = np.nonzero(pos != -1)
        cols = pos[cx, cy]
        keep = (cols >= 0) & (cols < n_src)
        self.src_x, self.src_y, self.src_col = cx[keep], cy[keep], cols[keep]
        rows = np.arange(self.n_steps)[:, None] * n_src + self.src_col[None, :]
        vals = np.zeros(rows.shape)
        ok = rows < flat.size
        vals[ok] = flat[rows[ok]]
        self.src_val = torch.tensor(vals, dtype=dt64)              # (n_steps, n_src_cells)

        # receivers: the WGSL table walk (wgsl2d :1117-1134)
        info = np.asarray(sim._info_rec_pt, dtype=np.int64).reshape(-1, 3)
        offs = np.asarray(sim._offset_sensors, dtype=np.int64).reshape(-1)
        self.n_rec = int(sim._n_rec)
        self.delay = np.asarray(sim._delay_recv, dtype=np.int64).reshape(-1)
        R = np.zeros((self.n_rec, nx * ny))
        for r in range(self.n_rec):
            pt = int(offs[r]) if r < offs.size else -1
            while 0 <= pt < info.shape[0] and info[pt, 2] == r:
                x, y = info[pt, 0], info[pt, 1]
                if 0 <= x < nx and 0 <= y < ny:
                    R[r, x * ny + y] += 1.0
                pt += 1
        self.R = torch.tensor(R, dtype=dt64)

    # ------------------------------------------------------------------
    def _d(self, f, ax, plus):
        """Staggered difference along ax (0: x, 1: y): D+ f(i) = sum C (f(i+c+1) - f(i-c)) / d,
        D- f(i) = sum C (f(i+c) - f(i-c-1)) / d."""
        d = self.dx if ax == 0 else self.dy
        out = torch.zeros_like(f)
        for c, C in enumerate(self.C):
            a, b = (c + 1, -c) if plus else (c, -c - 1)
            sa = _shift(f, a, 0) if ax == 0 else _shift(f, 0, a)
            sb = _shift(f, b, 0) if ax == 0 else _shift(f, 0, b)
            out = out + C * (sa - sb) / d
        return out

    def _cpml(self, psi, D, ax, half, mask):
        tag = "x" if ax == 0 else "y"
        h = "h" if half else ""
        a, b, k = self.P[f"a{h}{tag}"], self.P[f"b{h}{tag}"], self.P[f"k{h}{tag}"]
        pn = torch.where(mask, b * psi + a * D, psi)
        return pn, D / k + pn

    def coefficients(self, rho, cp, cs):
        """Lame parameters at the stress points and densities at the velocity
        points, as the kernels form them from the cell maps (wgsl2d :863-868,
        :948-951, :1019, :1035-1036)."""
        rx1, cpx1, csx1 = _shift(rho, 1, 0), _shift(cp, 1, 0), _shift(cs, 1, 0)
        rho_h = 0.5 * (rx1 + rho)
        cp_h = 0.5 * (cpx1 + cp)
        cs_l = 0.5 * (csx1 + cs)
        cs_m = torch.where(torch.minimum(csx1, cs) == 0, torch.zeros_like(cs_l), cs_l)
        lam_n = rho_h * (cp_h * cp_h - 2.0 * cs_l * cs_l)
        mu_n = rho_h * (cs_m * cs_m)
        ry1, csy1 = _shift(rho, 0, 1), _shift(cs, 0, 1)
        rho_s = 0.5 * (ry1 + rho)
        cs_h = torch.where(torch.minimum(csy1, cs) == 0, torch.zeros_like(cs), 0.5 * (csy1 + cs))
        mu_s = rho_s * (cs_h * cs_h)
        rho_y = 0.25 * (rho + rx1 + _shift(rho, 1, 1) + ry1)
        return lam_n, mu_n, mu_s, rho, rho_y

    def forward(self, rho, cp, cs, tscale, return_state=False):
        """(n_steps, n_rec) traces; rho, cp, cs, tscale are (nx, ny) float64 tensors."""
        nx, ny, L, dt = self.nx, self.ny, self.L, self.dt
        lam, mu_n, mu_s, rho_x, rho_y = self.coefficients(rho, cp, cs)
        lp2m, lpm = lam + 2.0 * mu_n, lam + mu_n
        rs = 1.0 / tscale
        z = lambda: torch.zeros((nx, ny), dtype=torch.float64)
        vx, vy, sxx, syy, sxy = z(), z(), z(), z(), z()
        rxx = [z() for _ in range(L)]
        ryy = [z() for _ in range(L)]
        rxy = [z() for _ in range(L)]
        psi = {k: z() for k in ("vxx", "vyy", "vyx", "vxy", "sxxx", "sxyy", "sxyx", "syyy")}
        phi_p = [self.A_p[l] * (self.inv_p[l] * rs) for l in range(L)]
        phi_s = [self.A_s[l] * (self.inv_s[l] * rs) for l in range(L)]
        h_p = [0.5 * dt * (self.inv_p[l] * rs) for l in range(L)]
        h_s = [0.5 * dt * (self.inv_s[l] * rs) for l in range(L)]
        m_p = [1.0 / (1.0 + h) for h in h_p]
        m_s = [1.0 / (1.0 + h) for h in h_s]
        safe_x = torch.where(rho_x > 0, rho_x, torch.ones_like(rho_x))
        safe_y = torch.where(rho_y > 0, rho_y, torch.ones_like(rho_y))
        traces = []
        sx, sy = torch.as_tensor(self.src_x), torch.as_tensor(self.src_y)
        for it in range(self.n_steps):
            # ---- stress pass (wgsl2d :829-984)
            psi["vxx"], e1 = self._cpml(psi["vxx"], self._d(vx, 0, True), 0, True, self.mN)
            psi["vyy"], e2 = self._cpml(psi["vyy"], self._d(vy, 1, False), 1, False, self.mN)
            Sxx, Syy, nxx, nyy = 0.0, 0.0, [], []
            for l in range(L):
                rn = (rxx[l] + (e1 + e2) * phi_p[l] - rxx[l] * h_p[l]) * m_p[l]
                rm = (ryy[l] + 0.5 * (e1 - e2) * phi_s[l] - ryy[l] * h_s[l]) * m_s[l]
                Sxx = Sxx + rxx[l] + rn
                Syy = Syy + ryy[l] + rm
                nxx.append(torch.where(self.mN, rn, rxx[l]))
                nyy.append(torch.where(self.mN, rm, ryy[l]))
            sxx_n = torch.where(self.mN, sxx + (lp2m * e1 + lam * e2 + 0.5 * lpm * Sxx + mu_n * Syy) * dt, sxx)
            syy_n = torch.where(self.mN, syy + (lam * e1 + lp2m * e2 + 0.5 * lpm * Sxx - mu_n * Syy) * dt, syy)
            psi["vyx"], e3 = self._cpml(psi["vyx"], self._d(vy, 0, False), 0, False, self.mS)
            psi["vxy"], e4 = self._cpml(psi["vxy"], self._d(vx, 1, True), 1, True, self.mS)
            Sxy, nxy = 0.0, []
            for l in range(L):
                rn = (rxy[l] + (e3 + e4) * phi_s[l] - rxy[l] * h_s[l]) * m_s[l]
                Sxy = Sxy + rxy[l] + rn
                nxy.append(torch.where(self.mS, rn, rxy[l]))
            sxy_n = torch.where(self.mS, sxy + (mu_s * (e3 + e4) + 0.5 * mu_s * Sxy) * dt, sxy)
            sxx, syy, sxy, rxx, ryy, rxy = sxx_n, syy_n, sxy_n, nxx, nyy, nxy
            # ---- velocity pass (wgsl2d :986-1065)
            psi["sxxx"], e5 = self._cpml(psi["sxxx"], self._d(sxx, 0, False), 0, False, self.mX)
            psi["sxyy"], e6 = self._cpml(psi["sxyy"], self._d(sxy, 1, False), 1, False, self.mX)
            vx_n = torch.where(self.mX & (rho_x > 0), (e5 + e6) * dt / safe_x + vx, vx)
            psi["sxyx"], e7 = self._cpml(psi["sxyx"], self._d(sxy, 0, True), 0, True, self.mY)
            psi["syyy"], e8 = self._cpml(psi["syyy"], self._d(syy, 1, True), 1, True, self.mY)
            vy_n = torch.where(self.mY & (rho_y > 0), (e7 + e8) * dt / safe_y + vy, vy)
            vx = torch.where(self.mY, vx_n, torch.zeros_like(vx_n))
            vy = torch.where(self.mY, vy_n, torch.zeros_like(vy_n))
            if sx.numel():
                ry = rho_y[sx, sy]
                add = torch.where(ry > 0, self.src_val[it] * dt / torch.where(ry > 0, ry, torch.ones_like(ry)),
                                  torch.zeros_like(ry))
                vy = vy.index_put((sx, sy), add, accumulate=True)
            # ---- receivers (wgsl2d :1117-1134): row it when it >= delay
            row = self.R @ syy.reshape(-1)
            gate = torch.tensor((it >= self.delay).astype(np.float64))
            traces.append(row * gate)
        out = torch.stack(traces)
        if return_state:
            return out, {"vx": vx, "vy": vy, "sxx": sxx, "syy": syy, "sxy": sxy,
                         "rxx": torch.stack(rxx, -1), "ryy": torch.stack(ryy, -1), "rxy": torch.stack(rxy, -1)}
        return out
