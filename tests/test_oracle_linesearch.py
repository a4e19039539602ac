"""Pins the oracle's step-selection helpers (Moré & Thuente 1994 as in
reference src/utils.h:76-230) against exact polynomial minima."""
import numpy as np
import pytest

import oracle


def cubic(c, x):
    return c[0] + c[1] * x + c[2] * x * x + c[3] * x ** 3


def dcubic(c, x):
    return c[1] + 2 * c[2] * x + 3 * c[3] * x * x


@pytest.mark.parametrize("c,a,b", [((1.0, -2.0, 0.5, 0.3), 0.0, 2.0), ((0.0, -1.0, 0.0, 1.0), 0.1, 1.5),
                                   ((2.0, 1.0, -3.0, 1.0), 1.5, 3.0), ((0.0, -1.0, 0.0, 1.0), 1.5, 0.1)])
def test_cubic_minimizer_exact(c, a, b):
    x, ok = oracle.cubic_minimizer(a, b, cubic(c, a), cubic(c, b), dcubic(c, a), dcubic(c, b))
    # local minimiser of the cubic: derivative zero and second derivative positive
    roots = np.roots([3 * c[3], 2 * c[2], c[1]])
    roots = roots[np.isreal(roots)].real
    mins = [r for r in roots if 2 * c[2] + 6 * c[3] * r > 0]
    assert ok
    assert len(mins) == 1
    assert abs(x - mins[0]) < 1e-10


def test_quadratic_minimizers_exact():
    # q(x) = 3 (x - 0.7)^2 + 1
    q = lambda x: 3 * (x - 0.7) ** 2 + 1
    dq = lambda x: 6 * (x - 0.7)
    assert abs(oracle.quadratic_minimizer_fg(0.0, 2.0, q(0.0), dq(0.0), q(2.0)) - 0.7) < 1e-12
    assert abs(oracle.quadratic_minimizer_gg(0.0, 2.0, dq(0.0), dq(2.0)) - 0.7) < 1e-12
    # a wrong sign in the interpolation would land on 2*a - 0.7
    assert abs(oracle.quadratic_minimizer_fg(1.0, 2.0, q(1.0), dq(1.0), q(2.0)) - 0.7) < 1e-12


def test_cubic_degenerates_to_quadratic():
    c = (1.0, -2.0, 1.0, 0.0)
    x, ok = oracle.cubic_minimizer(0.0, 3.0, cubic(c, 0), cubic(c, 3), dcubic(c, 0), dcubic(c, 3))
    assert ok and abs(x - 1.0) < 1e-12


def test_step_selection_cases():
    # case 1 (ft > fl): result lies between al and at
    r = oracle.step_selection(0.0, np.inf, 1.0, 0.0, np.inf, 1.0, -1.0, np.inf, 2.0)
    assert 0.0 < r < 1.0
    # non-finite trial value: midpoint
    assert oracle.step_selection(0.0, np.inf, 1.0, 0.0, np.inf, np.nan, -1.0, np.inf, 1.0) == 0.5
    # extrapolation when the upper end is infinite (case 4)
    r = oracle.step_selection(0.0, np.inf, 1.0, 0.0, np.inf, -1.0, -2.0, np.inf, -3.0)
    assert abs(r - 2.1) < 1e-12
