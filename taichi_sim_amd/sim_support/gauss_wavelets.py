"""Gaussian-pulse excitation signals (sim_support/gauss_wavelets.py:6-200).

``gaussian_pulse`` is SciPy's ``scipy.signal.gausspulse``.  The derivative
pulses are closed forms of d/dt and d2/dt2 of exp(-a t^2) exp(j 2 pi fc t),
where a is fixed by the fractional bandwidth ``bw`` at reference level
``bwr`` dB: exp(-a (pi fc bw)^2 / (4 a^2)) ... i.e. a = -(pi fc bw)^2 / (4 ln ref).
The dtype flow (float32 pi and phase, float64 envelope) is the reference's, so
the float32 source terms match it bit for bit.
"""
import numpy as np
from scipy.signal import gausspulse

from . import PI


def gaussian_pulse(*args, **kwargs):
    return gausspulse(*args, **kwargs)


def _bandwidth_a(fc, bw, bwr):
    if fc < 0:
        raise ValueError(f"Center frequency (fc={fc:.2f}) must be >=0.")
    if bw <= 0:
        raise ValueError(f"Fractional bandwidth (bw={bw:.2f}) must be > 0.")
    if bwr >= 0:
        raise ValueError(f"Reference level for bandwidth (bwr={bwr:.2f}) must be < 0 dB")
    ref = pow(10.0, bwr / 20.0)
    return -(PI * fc * bw) ** 2 / (4.0 * np.log(ref))


def _cutoff(a, tpr):
    if tpr >= 0:
        raise ValueError("Reference level for time cutoff must be < 0 dB")
    tref = pow(10.0, tpr / 20.0)
    return np.sqrt(-np.log(tref) / a)


def _pack(yI, yQ, yenv, retquad, retenv):
    out = (yI,)
    if retquad:
        out += (yQ,)
    if retenv:
        out += (yenv,)
    return out[0] if len(out) == 1 else out


def gaussian_first_dev_pulse(t, fc=1000, bw=0.5, bwr=-6, tpr=-60, retquad=False, retenv=False):
    """d/dt of the Gaussian-modulated sinusoid (gauss_wavelets.py:10-103).

    yI = -2 g(t) (a t cos(2 pi fc t) + pi fc sin(2 pi fc t)), g = exp(-a t^2);
    envelope -2 a t g(t); quadrature envelope * sin(2 pi fc t).
    """
    a = _bandwidth_a(fc, bw, bwr)
    if isinstance(t, str):
        if t == 'cutoff':
            return _cutoff(a, tpr)
        raise ValueError("If `t` is a string, it must be 'cutoff'")
    g = np.exp(-a * t * t)
    yenv = - 2 * g * (a * t)
    yI = - 2 * g * (a * t * np.cos(2 * PI * fc * t) + PI * fc * np.sin(2 * PI * fc * t))
    yQ = yenv * np.sin(2 * PI * fc * t)
    return _pack(yI, yQ, yenv, retquad, retenv)


def gaussian_second_dev_pulse(t, fc=1000, bw=0.5, bwr=-6, tpr=-60, retquad=False, retenv=False):
    """d2/dt2 of the Gaussian-modulated sinusoid (gauss_wavelets.py:106-200).

    With w = 2 pi fc: yI = g(t) ((4 a^2 t^2 - 2 a - w^2) cos(w t) + 4 a w t sin(w t)),
    envelope g(t)(4 a^2 t^2 - 2 a).
    """
    a = _bandwidth_a(fc, bw, bwr)
    if isinstance(t, str):
        if t == 'cutoff':
            return _cutoff(a, tpr)
        raise ValueError("If `t` is a string, it must be 'cutoff'")
    w = 2 * PI * fc
    g = np.exp(-a * t * t)
    yenv = g * (4 * a * a * t * t - 2 * a)
    yI = g * ((4 * a * a * t * t - 2 * a - w * w) * np.cos(w * t) + 4 * a * w * t * np.sin(w * t))
    yQ = yenv * np.sin(w * t)
    return _pack(yI, yQ, yenv, retquad, retenv)
