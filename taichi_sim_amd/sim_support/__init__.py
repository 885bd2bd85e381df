"""Host-side setup package with the reference's ``sim_support`` API.

It builds every array the time loop consumes: the ROI and grid sizes, material
maps, CPML profiles, FD coefficients, transducer source terms and receiver
tables.  The names follow the reference (``sim_support/__init__.py:1-90``) so a
user's configuration files and subclasses keep working.
"""
import numpy as np

flt32 = np.float32
int32 = np.int32

PI = flt32(np.pi)
DEGREES_TO_RADIANS = flt32(PI / 180.0)
ZERO = flt32(0.0)
# Divergence threshold on the velocity norm (sim_support/__init__.py:16).
STABILITY_THRESHOLD = flt32(1.0e38)
HUGEVAL = flt32(1.0e30)

# Staggered-grid FD coefficients of Liu (2009), one row per order acc = 2..6;
# row acc-2 has acc terms (sim_support/__init__.py:20-28).  Exact rationals.
_LIU_NUM_DEN = [
    [(9, 8), (-1, 24)],
    [(75, 64), (-25, 384), (3, 640)],
    [(1225, 1024), (-245, 3072), (49, 5120), (-5, 7168)],
    [(19845, 16384), (-735, 8192), (567, 40960), (-405, 229376), (35, 294912)],
    [(160083, 131072), (-12705, 131072), (22869, 1310720), (-5445, 1835008),
     (847, 2359296), (-63, 2883584)],
]
coefs_Lui = [[n / d for (n, d) in row] for row in _LIU_NUM_DEN]


def _forward_stencil(k):
    """First-derivative weights on the offsets -(k+1) .. k+2 (exact rationals):
    the unique c with sum_j c_j o_j^m = [m == 1] for m = 0 .. 2k+3."""
    from fractions import Fraction
    offs = list(range(-(k + 1), k + 3))
    n = len(offs)
    rows = [[Fraction(o) ** m for o in offs] + [Fraction(int(m == 1))] for m in range(n)]
    for i in range(n):                                   # Gauss-Jordan; the Vandermonde matrix is regular
        p = next(r for r in range(i, n) if rows[r][i] != 0)
        rows[i], rows[p] = rows[p], rows[i]
        rows[i] = [v / rows[i][i] for v in rows[i]]
        for r in range(n):
            if r != i and rows[r][i] != 0:
                f = rows[r][i]
                rows[r] = [a - f * b for a, b in zip(rows[r], rows[i])]
    return [rows[i][n] for i in range(n)]


# Forward-difference ("unsplit") first-derivative coefficients, one row per
# order: row k has 2k + 4 weights on the offsets -(k+1) .. k+2 (accuracy
# 2k + 3).  The backward derivative uses the mirrored row with the sign
# flipped (sim_support/__init__.py:32-42).
coefs_forward = [[float(c) for c in _forward_stencil(k)] for k in range(5)]


def unsplit_coefficients(row):
    """(forward, backward) float32 weights of ``coefs_forward[row]``; the
    backward weights apply to the mirrored offsets -(k+2) .. k+1."""
    fwd = np.array(coefs_forward[row], dtype=flt32)
    return fwd, -fwd[::-1]


class CourantError(ValueError):
    """Raised at setup when the Courant number exceeds 1 (sim_support/__init__.py:48-62)."""

    def __init__(self, mensagem, valor):
        super().__init__(mensagem)
        self.valor = valor

    def __str__(self):
        return f"{super().__str__()} [Valor: {self.valor}]"


class StabilityError(ValueError):
    """Raised when the velocity norm diverges (sim_support/__init__.py:65-76).

    Unlike the reference (whose ``vsn2 > 1e38`` test lets NaN through), the
    device check here also trips on NaN.
    """

    def __init__(self, mensagem, valor):
        super().__init__(mensagem)
        self.valor = valor

    def __str__(self):
        return f"{super().__str__()} [Valor: {self.valor}]"


__all__ = ['flt32', 'int32', 'PI', 'DEGREES_TO_RADIANS', 'ZERO', 'STABILITY_THRESHOLD',
           'HUGEVAL', 'coefs_Lui', 'coefs_forward', 'unsplit_coefficients', 'CourantError', 'StabilityError']
