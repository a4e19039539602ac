"""Pins the oracle's Newton solve (Schur reduction + dense Cholesky) against
numpy.linalg.solve on the assembled dense matrix (reference
test/test_sparse_cholesky.py:46-61 uses scipy spsolve the same way)."""
import numpy as np
import pytest

import oracle
from tests import _dense


@pytest.mark.parametrize("n,m,K", [(10, 6, 20), (6, 10, 20), (30, 30, 120), (8, 1, 1), (1, 8, 5)])
def test_csr_solve_matches_dense(n, m, K):
    rng = np.random.default_rng(n + 7 * m)
    reg = 0.05
    gamma = rng.normal(scale=0.05, size=n + m)
    gamma[-1] = 0.0
    M = rng.random((n, m))
    a = rng.random(n) + 0.5; a /= a.sum()
    b = rng.random(m) + 0.5; b /= b.sum()
    tp = oracle.tpass(gamma, M, a, b, reg)
    si, sj = oracle.topk_pattern(gamma, M, reg, K)
    vals, col, rp = oracle.assemble_csr(gamma, M, tp["rowsum"], tp["colsum"], reg, 1e-3 * reg, si, sj)
    N = n + m - 1
    A = _dense.csr_dense_sym(vals, col, rp, N)
    rhs = rng.normal(size=(2, N))
    x = oracle.csr_solve(vals, col, rp, n, m, rhs)
    ref = np.linalg.solve(A, rhs.T).T
    np.testing.assert_allclose(x, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_dense_cholesky():
    rng = np.random.default_rng(0)
    for N in (1, 10, 100):
        L = np.tril(rng.normal(size=(N, N)))
        A = L @ L.T + 0.1 * np.eye(N)
        rhs = rng.normal(size=(3, N))
        x = oracle.dense_cholesky_solve(A, rhs)
        np.testing.assert_allclose(x, np.linalg.solve(A, rhs.T).T, rtol=1e-9, atol=1e-10)
    with pytest.raises(RuntimeError):
        oracle.dense_cholesky_solve(-np.eye(3), np.ones(3))
