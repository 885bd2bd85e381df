"""FD coefficient tables, pinned by the moment conditions that define them.

* Staggered (Liu 2009, sim_support/__init__.py:20-28): for acc = n terms,
  sum_c C_c (2c+1) = 1 and sum_c C_c (2c+1)^(2m+1) = 0 for m = 1 .. n-1, so
  sum_c C_c (f(i+c+1/2) - f(i-c-1/2)) is f' to order 2n.
* Forward ("unsplit", sim_support/__init__.py:32-42): row k weights the
  offsets -(k+1) .. k+2 with sum_j c_j o_j^m = [m == 1] for m = 0 .. 2k+3;
  the backward row, the mirrored weights with the sign flipped, satisfies the
  same conditions on the mirrored offsets.
"""
from fractions import Fraction

import numpy as np
import pytest

from taichi_sim_amd.sim_support import _LIU_NUM_DEN, coefs_forward, coefs_Lui, unsplit_coefficients


@pytest.mark.parametrize("row", range(5))
def test_staggered_coefficients_moment_conditions(row):
    C = [Fraction(n, d) for n, d in _LIU_NUM_DEN[row]]
    assert len(C) == row + 2
    assert sum(c * (2 * i + 1) for i, c in enumerate(C)) == 1
    for m in range(1, len(C)):
        assert sum(c * (2 * i + 1) ** (2 * m + 1) for i, c in enumerate(C)) == 0
    assert coefs_Lui[row] == [float(c) for c in C]


@pytest.mark.parametrize("row", range(5))
def test_forward_coefficients_moment_conditions(row):
    fwd = [Fraction(c).limit_denominator(10 ** 7) for c in coefs_forward[row]]
    offs = list(range(-(row + 1), row + 3))
    assert len(fwd) == len(offs) == 2 * row + 4
    for m in range(len(offs)):
        assert sum(c * Fraction(o) ** m for c, o in zip(fwd, offs)) == (1 if m == 1 else 0), m
    # the float table is the nearest double of each rational
    assert coefs_forward[row] == [float(c) for c in fwd]


def test_forward_table_spot_values_and_backward_mirror():
    # row 0 is the 4-point, third-order forward stencil -(1/3) f(-1) - (1/2) f(0) + f(1) - (1/6) f(2)
    assert coefs_forward[0] == [-2.0 / 6.0, -3.0 / 6.0, 1.0, -1.0 / 6.0]
    for row in range(5):
        fwd, bwd = unsplit_coefficients(row)
        assert fwd.dtype == np.float32 and np.array_equal(bwd, -fwd[::-1])
        # on a cubic f(x) = x^3 - 2x + 1 both give f'(0) = -2, f'(0) on the mirrored offsets
        k = row
        f = lambda x: x ** 3 - 2.0 * x + 1.0
        offs_f = np.arange(-(k + 1), k + 3, dtype=np.float64)
        offs_b = np.arange(-(k + 2), k + 2, dtype=np.float64)
        assert abs(np.dot(fwd.astype(np.float64), f(offs_f)) + 2.0) < 1e-5
        assert abs(np.dot(bwd.astype(np.float64), f(offs_b)) + 2.0) < 1e-5
