"""Pins the oracle's SPLR driver (oracle/cgot_oracle.c:or_splr_bwd, following
reference src/sinkhorn_splr.cu:1305-1611) by brute force, invariants and the
reference's own contract cases."""
import numpy as np
import pytest

import oracle
from curegot_backward_amd.workloads import readme_problem, make_problem
from tests import _dense


@pytest.mark.parametrize("n,m,eta,seed", [(64, 80, 0.05, 0), (30, 20, 0.01, 1), (20, 30, 0.1, 2), (7, 1, 0.05, 3),
                                          (1, 9, 0.05, 4), (50, 50, 0.01, 5)])
def test_splr_matches_dense_newton(n, m, eta, seed):
    """Converged at tol 1e-8, the SPLR plan is the dense-Newton optimum up to a
    small multiple of the final gradient norm."""
    M, a, b = [x.numpy() for x in make_problem(n, m, seed)]
    T_ref, dual_ref, gn = _dense.newton(M, a, b, eta)
    assert gn < 1e-13
    out = oracle.splr_bwd(M, a, b, eta, tol=1e-8, max_iter=5000)
    assert out["converged"] and out["gnorm"] < 1e-8
    err = np.abs(out["plan"] - T_ref).max() / T_ref.max()
    assert err < max(1e3 * out["gnorm"], 1e-13), err
    assert out["dual"][-1] == 0.0


def test_splr_tight_tolerance_readme_problem():
    # the README problem reaches ||g|| < 1e-13; the plan then equals the optimum to 1e-10
    M, a, b, eta = [x.numpy() if hasattr(x, "numpy") else x for x in readme_problem()]
    T_ref, dual_ref, gn = _dense.newton(M, a, b, eta)
    out = oracle.splr_bwd(M, a, b, eta, tol=1e-13, max_iter=5000)
    assert out["converged"]
    assert np.abs(out["plan"] - T_ref).max() / T_ref.max() < 1e-10
    assert np.abs(out["dual"] - dual_ref).max() < 1e-9


def test_stall_below_objective_resolution():
    # a documented property of the reference line search: once ||g||^2 is below the
    # resolution of f, the Wolfe search returns step_min and the iterate stops moving
    M, a, b = [x.numpy() for x in make_problem(64, 80, 0)]
    out = oracle.splr_bwd(M, a, b, 0.05, tol=1e-13, max_iter=60, trace=True)
    assert not out["converged"]
    assert out["trace"][-1, 2] == 1e-10


def test_splr_marginals_at_tol():
    M, a, b, eta = [x.numpy() if hasattr(x, "numpy") else x for x in readme_problem()]
    out = oracle.splr_bwd(M, a, b, eta, tol=1e-8)
    assert out["converged"] and out["gnorm"] < 1e-8
    np.testing.assert_allclose(out["plan"].sum(1), a, atol=1e-8)
    np.testing.assert_allclose(out["plan"][:, :-1].sum(0), b[:-1], atol=1e-8)


def test_plan_invariant_under_row_col_shift():
    # M + u 1^T + 1 v^T leaves the plan unchanged (reference autograd.py:271-272)
    M, a, b = [x.numpy() for x in make_problem(25, 18, 5)]
    rng = np.random.default_rng(0)
    M2 = M + rng.random(25)[:, None] + rng.random(18)[None, :]
    o1 = oracle.splr_bwd(M, a, b, 0.05, tol=1e-8, max_iter=5000)
    o2 = oracle.splr_bwd(M2, a, b, 0.05, tol=1e-8, max_iter=5000)
    assert o1["converged"] and o2["converged"]
    err = np.abs(o1["plan"] - o2["plan"]).max() / o1["plan"].max()
    assert err < 1e3 * max(o1["gnorm"], o2["gnorm"])


def test_warm_start_at_optimum_is_fallback():
    # reference stage3b_validate.py:527-537: x0 = optimum => niter == 0, fallback context
    M, a, b = [x.numpy() for x in make_problem(20, 24, 9)]
    _, dual_ref, gn = _dense.newton(M, a, b, 0.05)
    assert gn < 1e-13
    out = oracle.splr_bwd(M, a, b, 0.05, tol=1e-8, x0=dual_ref)
    assert out["niter"] == 0 and out["converged"] and out["fallback"] and not out["low_rank"]


def test_non_convergence_reported():
    M, a, b = [x.numpy() for x in make_problem(20, 24, 9)]
    out = oracle.splr_bwd(M, a, b, 0.05, tol=1e-14, max_iter=1)
    assert out["niter"] == 1 and not out["converged"]


def test_line_search_trace_decreases_objective():
    # every accepted step decreases f (sufficient decrease of the Wolfe search)
    M, a, b, eta = [x.numpy() if hasattr(x, "numpy") else x for x in readme_problem()]
    out = oracle.splr_bwd(M, a, b, eta, tol=1e-10, trace=True, candidate_sinkhorn_iter=0)
    f = out["trace"][:, 0]
    assert np.all(np.diff(f) <= 1e-15)
