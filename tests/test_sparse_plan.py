"""The symbolic analysis of the sparse Newton solve (host only): on the
Newton matrices A of real SPLR iterates (the oracle's context of solved
problems) and on synthetic patterns, the plan's schedule -- executed in numpy
exactly as the device kernel runs it -- solves A x = b to rounding against
numpy.linalg.solve, and patterns whose dense core exceeds the limit keep the
dense Schur path."""
import numpy as np
import pytest

import oracle
from curegot_backward_amd.workloads import make_problem, ragged_batch
from tests import _sparse_plan as spp


def check(n, m, rp, cj, wv, Da, Db, seed=0, rmax=196):
    h, buf = spp.make_plan(n, m, rp, cj, rmax=rmax)
    assert h is not None
    N = n + m - 1
    perm = spp.arr(buf, h.o_perm, N)
    assert sorted(perm.tolist()) == list(range(N)) and h.t + h.R == N and h.R <= rmax
    A = spp.dense_A(Da, Db, rp, cj, wv, n, m)
    rhs = np.random.default_rng(seed).normal(size=(2, N))
    x = spp.execute(h, buf, Da, Db, wv, rhs)
    xr = np.linalg.solve(A, rhs.T).T
    err = np.abs(x - xr).max() / np.abs(xr).max()
    assert err < 1e-11 * np.linalg.cond(A) ** 0.5 + 1e-13, err
    return h


@pytest.mark.parametrize("n,m,eta,seed", [(100, 120, 0.05, 3), (230, 150, 0.02, 8), (70, 300, 0.1, 9),
                                          (300, 257, 0.01, 8)])
def test_plan_on_oracle_contexts(n, m, eta, seed):
    M, a, b = make_problem(n, m, seed)
    ref = oracle.splr_bwd(M.numpy(), a.numpy(), b.numpy(), eta, tol=1e-8, max_iter=500)
    rp, cj, wv, Da, Db = spp.w_pattern(ref["csr_values"], ref["csr_colind"], ref["csr_rowptr"], n, m)
    check(n, m, rp, cj, wv, Da, Db, seed)


def test_plan_on_batch_instances():
    """Small instances of the benchmarked ragged batch at their final iterate."""
    probs = ragged_batch(128)
    done = 0
    for k, (M, a, b, eta) in enumerate(probs):
        n, m = M.shape
        if n * m > 120_000:
            continue
        ref = oracle.splr_bwd(M.numpy(), a.numpy(), b.numpy(), eta, tol=1e-8, max_iter=1000)
        rp, cj, wv, Da, Db = spp.w_pattern(ref["csr_values"], ref["csr_colind"], ref["csr_rowptr"], n, m)
        h = check(n, m, rp, cj, wv, Da, Db, k)
        print(f"instance {k} {n}x{m}: periphery {h.t}, core {h.R}, levels {h.H}")
        done += 1
        if done == 4:
            break
    assert done == 4


def random_spd_pattern(n, m, per_row, rng, dense_block=0):
    rows = []
    for i in range(n):
        cols = set(rng.choice(m - 1, size=min(per_row, m - 1), replace=False).tolist())
        if i < dense_block:
            cols |= set(range(min(dense_block, m - 1)))
        rows.append(sorted(cols))
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    cj = np.array([c for r in rows for c in r], np.int32)
    wv = rng.uniform(0.1, 1.0, size=len(cj))
    # strictly diagonally dominant: Da, Db above the row / column sums of W
    Da = np.array([wv[rp[i]:rp[i + 1]].sum() for i in range(n)]) + 0.5
    Db = np.zeros(m - 1)
    np.add.at(Db, cj, wv)
    return rp, cj, wv, Da, Db + 0.5


@pytest.mark.parametrize("n,m,per_row,dense", [(60, 50, 1, 0), (60, 50, 3, 0), (40, 90, 2, 12), (5, 4, 3, 0),
                                               (1, 6, 3, 0), (7, 2, 1, 0), (120, 110, 2, 30)])
def test_plan_on_synthetic_patterns(n, m, per_row, dense):
    rng = np.random.default_rng(n * 100 + m + per_row)
    rp, cj, wv, Da, Db = random_spd_pattern(n, m, per_row, rng, dense)
    check(n, m, rp, cj, wv, Da, Db)


def test_empty_pattern_and_no_columns():
    # no selected entries: A is diagonal, every vertex is a periphery leaf
    n, m = 9, 7
    rp = np.zeros(n + 1, np.int32)
    h = check(n, m, rp, np.zeros(0, np.int32), np.zeros(0), np.full(n, 2.0), np.full(m - 1, 3.0))
    assert h.R == 0 and h.H == 1 and h.t == n + m - 1


def test_large_dense_core_keeps_dense_path():
    rng = np.random.default_rng(5)
    rp, cj, wv, Da, Db = random_spd_pattern(80, 80, 60, rng)
    h, _ = spp.make_plan(80, 80, rp, cj, rmax=40)
    assert h is None
    h, _ = spp.make_plan(80, 80, rp, cj, rmax=196)
    assert h is not None and h.R > 40
    # a plan larger than the buffer it must fit in keeps the dense path too
    h, _ = spp.make_plan(80, 80, rp, cj, cap=1024)
    assert h is None


def test_plan_with_panel_core():
    """A top-K pattern (the oracle's selection on an n >> m problem) whose
    dense core exceeds the register-tile core's 252 vertices: the analysis
    keeps it as a core of up to 768 (factored by panels on the device) and
    the plan's schedule solves the system."""
    n, m, K = 1200, 900, 3400
    rng = np.random.default_rng(n + m)
    M = rng.random((n, m))
    gamma = np.concatenate([rng.normal(scale=0.01, size=n), rng.normal(scale=0.01, size=m)])
    si, sj = oracle.topk_pattern(gamma, M, 0.05, K)
    sel = sorted((i, j) for i, j in zip(si, sj) if j < m - 1)
    rp = np.zeros(n + 1, np.int32)
    for i, _ in sel:
        rp[i + 1] += 1
    rp = np.cumsum(rp).astype(np.int32)
    cj = np.array([j for _, j in sel], np.int32)
    wv = rng.uniform(0.1, 1.0, size=len(cj))
    Da = np.array([wv[rp[i]:rp[i + 1]].sum() for i in range(n)]) + 0.5
    Db = np.zeros(m - 1)
    np.add.at(Db, cj, wv)
    assert spp.make_plan(n, m, rp, cj, rmax=252)[0] is None
    h = check(n, m, rp, cj, wv, Da, Db + 0.5, rmax=768)
    assert 252 < h.R <= 768
