"""Pins the oracle's geometric cost or_cost_xy: exact on small-integer points
(every operation is exact there, so any summation order gives the same
value) and within a few ulps of an extended-precision sum elsewhere; a
transposed operand or a dropped term fails both."""
import numpy as np

import oracle


def test_exact_on_integer_points():
    rng = np.random.default_rng(0)
    X = rng.integers(-60, 60, (40, 7)).astype(float)
    Y = rng.integers(-60, 60, (33, 7)).astype(float)
    M = oracle.cost_xy(X, Y, 0.125)
    ref = ((X[:, None, :] - Y[None, :, :]) ** 2).sum(-1) * 0.125
    assert np.array_equal(M, ref)
    assert M.shape == (40, 33)


def test_close_to_extended_precision():
    rng = np.random.default_rng(1)
    for d in (1, 3, 8, 64):
        X = rng.random((25, d))
        Y = rng.random((19, d))
        M = oracle.cost_xy(X, Y, 1.0 / d)
        Xl, Yl = X.astype(np.longdouble), Y.astype(np.longdouble)
        ref = ((Xl[:, None, :] - Yl[None, :, :]) ** 2).sum(-1) / d
        rel = np.abs((M - ref) / ref).max()
        assert rel < (d + 2) * 2.3e-16, (d, rel)


def test_single_pair_by_hand():
    X = np.array([[1.0, 2.0, 3.0]])
    Y = np.array([[4.0, 6.0, 3.5], [1.0, 2.0, 3.0]])
    M = oracle.cost_xy(X, Y, 0.5)
    assert M[0, 0] == (9.0 + 16.0 + 0.25) * 0.5 and M[0, 1] == 0.0
