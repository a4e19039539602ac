"""Pins the oracle's T pass and top-K sparsified Hessian against a dense numpy
construction, restating reference test/test_csr.py:16-157 (n=6, m=3, K=10,
eta=0.1, shift=0.01) and random small shapes."""
import numpy as np
import pytest
from scipy.sparse import csr_matrix

import oracle


def dense_reference(alpha, beta, M, a, b, reg, shift, K):
    n, m = M.shape
    T = np.exp((alpha[:, None] + beta[None, :] - M) / reg)
    rows, cols = T.sum(1), T.sum(0)
    f = reg * T.sum() - alpha @ a - beta @ b
    grad = np.concatenate([rows - a, cols[:-1] - b[:-1]])
    Tt = T.copy()
    Tt[:, -1] = 0
    flat = Tt.flatten()
    # K largest of T[:, :m-1]; ties by smaller row-major index of T[:, :m-1]
    keys = [(-flat[i * m + j], i * (m - 1) + j, i, j) for i in range(n) for j in range(m - 1)]
    keys.sort()
    Tsp = np.zeros((n, m))
    for _, _, i, j in keys[:K]:
        Tsp[i, j] = T[i, j]
    H = np.diag(np.concatenate([rows, cols[:-1]]) + shift)
    H[n:, :n] = Tsp[:, :-1].T
    return rows, cols, f, grad, csr_matrix(H)


def run_oracle(alpha, beta, M, a, b, reg, shift, K):
    gamma = np.concatenate([alpha, beta])
    tp = oracle.tpass(gamma, M, a, b, reg)
    si, sj = oracle.topk_pattern(gamma, M, reg, K)
    vals, col, rp = oracle.assemble_csr(gamma, M, tp["rowsum"], tp["colsum"], reg, shift, si, sj)
    return tp, vals, col, rp


def test_reference_csr_case():
    np.random.seed(123)   # the reference script's seed
    n, m, K, reg, shift = 6, 3, 10, 0.1, 0.01
    alpha = np.random.normal(scale=0.1, size=n)
    beta = np.random.normal(scale=0.1, size=m)
    M = np.abs(np.random.normal(scale=0.1, size=(n, m)))
    a = np.random.uniform(size=n); a /= a.sum()
    b = np.random.uniform(size=m); b /= b.sum()
    rows, cols, f, grad, H = dense_reference(alpha, beta, M, a, b, reg, shift, K)
    tp, vals, col, rp = run_oracle(alpha, beta, M, a, b, reg, shift, K)
    np.testing.assert_allclose(tp["rowsum"], rows, atol=1e-12)
    np.testing.assert_allclose(tp["colsum"], cols, atol=1e-12)
    assert abs(tp["f"] - f) < 1e-12
    np.testing.assert_allclose(tp["grad"], grad, atol=1e-12)
    assert abs(tp["gnorm"] - np.linalg.norm(grad)) < 1e-12
    np.testing.assert_allclose(vals, H.data, atol=1e-12)
    np.testing.assert_array_equal(col, H.indices)
    np.testing.assert_array_equal(rp, H.indptr)


@pytest.mark.parametrize("n,m,K", [(7, 5, 1), (7, 5, 28), (9, 12, 17), (12, 9, 40), (3, 2, 2), (5, 1, 3)])
def test_random_shapes(n, m, K):
    rng = np.random.default_rng(n * 31 + m)
    reg, shift = 0.07, 0.003
    alpha, beta = rng.normal(scale=0.1, size=n), rng.normal(scale=0.1, size=m)
    M = rng.random((n, m))
    a = rng.random(n) + 0.1; a /= a.sum()
    b = rng.random(m) + 0.1; b /= b.sum()
    Ke = min(max(K, 1), n * (m - 1))
    tp, vals, col, rp = run_oracle(alpha, beta, M, a, b, reg, shift, K)
    rows, cols, f, grad, H = dense_reference(alpha, beta, M, a, b, reg, shift, Ke)
    np.testing.assert_allclose(vals, H.data, atol=1e-12)
    np.testing.assert_array_equal(col, H.indices)
    np.testing.assert_array_equal(rp, H.indptr)


def test_ties_take_smallest_index():
    # all T equal: the K selected entries are the first K in row-major order
    n, m, K = 4, 5, 6
    M = np.zeros((n, m))
    gamma = np.zeros(n + m)
    si, sj = oracle.topk_pattern(gamma, M, 0.1, K)
    got = sorted(zip(si.tolist(), sj.tolist()))
    want = sorted([(t // (m - 1), t % (m - 1)) for t in range(K)])
    assert got == want
