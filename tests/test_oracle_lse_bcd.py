"""Pins the oracle's LSE updates and BCD driver (oracle/cgot_oracle.c:or_row_lse_update,
or_col_lse_update, or_sinkhorn_bcd) against scipy and closed forms."""
import numpy as np
import pytest
from scipy.special import logsumexp

import oracle
from tests import _dense


@pytest.mark.parametrize("n,m,eta", [(5, 7, 0.1), (40, 30, 0.01), (1, 9, 0.05), (9, 1, 0.05)])
def test_lse_updates_match_scipy(n, m, eta):
    rng = np.random.default_rng(n * 100 + m)
    M = rng.random((n, m))
    a = rng.random(n) + 0.5; a /= a.sum()
    b = rng.random(m) + 0.5; b /= b.sum()
    beta = rng.normal(size=m) * 0.1
    alpha = oracle.row_lse_update(M, beta, np.log(a), eta)
    ref = eta * (np.log(a) - logsumexp((beta[None, :] - M) / eta, axis=1))
    np.testing.assert_allclose(alpha, ref, rtol=0, atol=1e-13)
    # closed form: after the alpha update the row sums of T equal a
    T = np.exp((alpha[:, None] + beta[None, :] - M) / eta)
    np.testing.assert_allclose(T.sum(1), a, rtol=1e-12)
    beta2 = oracle.col_lse_update(M, alpha, np.log(b), eta)
    ref2 = eta * (np.log(b) - logsumexp((alpha[:, None] - M) / eta, axis=0))
    np.testing.assert_allclose(beta2, ref2, rtol=0, atol=1e-13)
    T = np.exp((alpha[:, None] + beta2[None, :] - M) / eta)
    np.testing.assert_allclose(T.sum(0), b, rtol=1e-12)


def test_lse_dropped_max_term_is_caught():
    # a large-magnitude cost: without the max shift exp underflows to 0
    M = np.array([[900.0, 901.0], [902.0, 903.0]])
    a = np.array([0.5, 0.5])
    alpha = oracle.row_lse_update(M, np.zeros(2), np.log(a), 1.0)
    assert np.all(np.isfinite(alpha))
    np.testing.assert_allclose(alpha, np.log(a) - logsumexp(-M, axis=1), atol=1e-12)


@pytest.mark.parametrize("eta", [0.05, 0.5])
def test_bcd_separable_cost_gives_product_plan(eta):
    # M_ij = u_i + v_j  =>  T = a b^T exactly, for any eta
    rng = np.random.default_rng(3)
    n, m = 12, 9
    u, v = rng.random(n), rng.random(m)
    M = u[:, None] + v[None, :]
    a = rng.random(n) + 0.2; a /= a.sum()
    b = rng.random(m) + 0.2; b /= b.sum()
    out = oracle.sinkhorn_bcd(M, a, b, eta, tol=1e-12, max_iter=1000)
    np.testing.assert_allclose(out["plan"], np.outer(a, b), rtol=1e-10, atol=1e-14)


def test_bcd_matches_newton_optimum():
    rng = np.random.default_rng(7)
    n, m, eta = 30, 25, 0.05
    M = rng.random((n, m))
    a = rng.random(n) + 0.5; a /= a.sum()
    b = rng.random(m) + 0.5; b /= b.sum()
    T_ref, _, gn = _dense.newton(M, a, b, eta)
    assert gn < 1e-13
    out = oracle.sinkhorn_bcd(M, a, b, eta, tol=1e-13, max_iter=100000)
    assert np.abs(out["plan"] - T_ref).max() / T_ref.max() < 1e-10
    # stopping rule: marginal error below tol once it terminated early
    assert np.linalg.norm(out["plan"].sum(1) - a) < 1e-13


def test_bcd_max_iter_truncation():
    # reference test/test_sinkhorn_bcd.py:172-177: max_iter=10 => niter == 10
    M, a, b = np.random.default_rng(1).random((20, 15)), np.full(20, 1 / 20), np.full(15, 1 / 15)
    out = oracle.sinkhorn_bcd(M, a, b, 0.01, tol=1e-14, max_iter=10)
    assert out["niter"] == 10
