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
    """Relaxation times for an approximately frequency-independent Q.

    Standard least-squares fit (e.g. Blanch, Robertsson & Symes 1995): the
    relaxation frequencies are spread logarithmically over [f_min, f_max]
    (units: MHz, so tau comes out in microseconds) and the strengths y_l
    minimise sum_f (Q^-1(f) - sum_l y_l w tau_l / (1 + w^2 tau_l^2))^2 with
    Q^-1 ~ sum_l y_l w tau_l / (1 + w^2 tau_l^2).  Returns an
    ``AttenuationCoefficients`` whose tau_epsilon = tau_sigma (1 + y_l).
    """
    f_relax = np.geomspace(f_min, f_max, n_sls)
    tau_s = 1.0 / (2.0 * np.pi * f_relax)
    w = 2.0 * np.pi * np.geomspace(f_min, f_max, n_freq)
    A = (w[:, None] * tau_s[None, :]) / (1.0 + (w[:, None] * tau_s[None, :]) ** 2)
    rows = []
    for q in (q_p, q_s):
        y, *_ = np.linalg.lstsq(A, np.full(n_freq, 1.0 / q), rcond=None)
        rows.append((tau_s * (1.0 + y), tau_s.copy()))
    tau = (rows[0][0], rows[0][1], rows[1][0], rows[1][1])
    return AttenuationCoefficients(tau=tau)


__all__ = ["AttenuationCoefficients", "fit_constant_q", "flt32"]
