"""Zener standard-linear-solid relaxation times.

``AttenuationCoefficients`` carries the reference's three hard-coded mechanisms
(sim_support/attenuation.py:8-30).  Those tau values are in seconds while the
time step is in microseconds, so dt/tau_sigma is about 1e4 and the memory
variables alternate sign every step.  ``fit_constant_q`` produces mechanisms in
microseconds that give an approximately constant quality factor over a band;
the viscoelastic workloads use it.
"""
import numpy as np

from . import flt32


class AttenuationCoefficients:
    """Relaxation times in the reference's packing.

    Attributes mirror the reference: ``_n_sls``, ``_tau_epsilon_p``,
    ``_tau_sigma_p``, ``_tau_epsilon_s``, ``_tau_sigma_s`` (float32 arrays of
    length L) and ``_sum_alpha_p`` / ``_sum_alpha_s`` (float32 sums of
    tau_epsilon / tau_sigma, accumulated left to right in float32 as Python's
    ``sum`` does on float32 scalars).
    """

    # sim_support/attenuation.py:14-22 (kappa: P mechanisms, non-kappa: S)
    _REF_TAU = (
        (1.1063680886151313E-007, 3.1856366036932041E-008, 9.2197868528094345E-009),
        (1.1026580338785149E-007, 3.1830988624149110E-008, 9.1888147513272412E-009),
        (1.6007136049523215E-007, 3.1849176860846836E-008, 6.3419851645615403E-009),
        (1.5988027842097575E-007, 3.1825778973240451E-008, 6.3344086277835006E-009),
    )

    def __init__(self, tau=None):
        rows = self._REF_TAU if tau is None else tau
        t = [np.asarray(r, dtype=np.float32) for r in rows]
        self._tau_epsilon_p, self._tau_sigma_p, self._tau_epsilon_s, self._tau_sigma_s = t
        self._n_sls = int(self._tau_sigma_p.shape[0])
        alpha_p = self._tau_epsilon_p / self._tau_sigma_p
        alpha_s = self._tau_epsilon_s / self._tau_sigma_s
        self._sum_alpha_p = _f32_sum(alpha_p)
        self._sum_alpha_s = _f32_sum(alpha_s)

    def tau_table(self):
        """(4, L) float32: tau_eps_p, tau_sig_p, tau_eps_s, tau_sig_s."""
        return np.array([self._tau_epsilon_p, self._tau_sigma_p,
                         self._tau_epsilon_s, self._tau_sigma_s], dtype=np.float32)

    def sum_alpha(self):
        return np.array([self._sum_alpha_p, self._sum_alpha_s], dtype=np.float32)


def _f32_sum(a):
    s = np.float32(0.0)
    for v in a:
        s = np.float32(s + v)
    return s


def fit_constant_q(q_p, q_s, f_min, f_max, n_sls=3, n_freq=64):
    """Relaxation times that give an approximately frequency-independent
    quality factor under the reference's Zener update.

    The relaxation frequencies are spread logarithmically over [f_min, f_max]
    (MHz, so tau comes out in microseconds).  The reference's memory-variable
    update (wgsl2d :879-913, wgsl3d :1405-1445) is the Crank-Nicolson step of
    dr/dt = -r / tau_sigma + X (1 - alpha) / (tau_sigma sum_alpha), with
    alpha = tau_epsilon / tau_sigma, and the stress takes the time average of r
    (S = r + r_new, times one half).  A modulus M it relaxes (M from cp, cs, the
    unrelaxed value) therefore behaves as

        M(w) = M (1 - sum_l s_l / (1 + i w tau_sigma_l)),  s_l = (alpha_l - 1) / sum_alpha,

    so each mechanism's strength is divided by sum_alpha, about L.  The s_l
    solve Q Im M(w) = Re M(w) in the least-squares sense over the band, which
    is linear in s: sum_l s_l (w tau_l + 1/Q) / (1 + (w tau_l)^2) = 1/Q.  Then
    alpha_l = 1 + s_l L / (1 - sum s) (so that sum_alpha = L / (1 - sum s))
    and tau_epsilon = alpha tau_sigma.

    The P family (q_p) relaxes the bulk part of the reference's normal
    stresses (lambda + mu in 2D, lambda + 2 mu / 3 in 3D) and the S family (q_s)
    the shear modulus, so q_p = q_s = Q gives both body waves the quality
    factor Q (in 3D along x and y; the reference's sigma_zz Zener terms,
    wgsl3d :1440-1445, relax P waves along z more strongly).
    Returns an ``AttenuationCoefficients``.
    """
    f_relax = np.geomspace(f_min, f_max, n_sls)
    tau_s = 1.0 / (2.0 * np.pi * f_relax)
    w = 2.0 * np.pi * np.geomspace(f_min, f_max, n_freq)
    wt = w[:, None] * tau_s[None, :]
    rows = []
    for q in (q_p, q_s):
        A = (wt + 1.0 / q) / (1.0 + wt ** 2)
        s_l, *_ = np.linalg.lstsq(A, np.full(n_freq, 1.0 / q), rcond=None)
        alpha = 1.0 + s_l * n_sls / (1.0 - s_l.sum())
        rows.append((tau_s * alpha, tau_s.copy()))
    tau = (rows[0][0], rows[0][1], rows[1][0], rows[1][1])
    return AttenuationCoefficients(tau=tau)


__all__ = ["AttenuationCoefficients", "fit_constant_q", "flt32"]
