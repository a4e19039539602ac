"""The adjoint workload at full size (n = m = 4096, d = 64 Gaussian points,
M = mean squared distance, uniform marginals, eta = 0.05, loss
L = sum(T^2) + <T, R>, R ~ N(0, 1)): dL/dM from the library's implicit
backward against central finite differences with h = 1e-6 on 64 random
entries, normwise relative error <= 1e-6.

Both sides are taken at the optimum: the base duals are the library's solution
polished by exact Newton steps (torch, dense Hessian through its Schur
complement), and every perturbed problem M +- h e_ij is solved by Newton steps
from there.  Each difference L(M + h e_ij) - L(M - h e_ij) is summed from the
elementwise plan difference, formed with expm1 so that it keeps its relative
accuracy, instead of as a difference of two rounded sums.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def newton(M, a, b, eta, dual, steps):
    n, m = M.shape
    for _ in range(steps):
        T = torch.exp((dual[:n, None] + dual[None, n:] - M) / eta)
        g = torch.cat([T.sum(1) - a, (T.sum(0) - b)[:-1]])
        W = T[:, :m - 1]
        Da, Db = T.sum(1), T.sum(0)[:m - 1]
        ga, gb = g[:n], g[n:]
        S = torch.diag(Db) - (W.T / Da) @ W
        xb = torch.cholesky_solve((gb - W.T @ (ga / Da))[:, None], torch.linalg.cholesky(S))[:, 0]
        xa = (ga - W @ xb) / Da
        dual = dual - eta * torch.cat([xa, xb, xa.new_zeros(1)])
    return dual


def test_adjoint_vs_central_fd_full_size():
    import curegot_backward_amd as cg
    from curegot_backward_amd.workloads import adjoint_problem
    M, a, b, eta, R = adjoint_problem()
    M, a, b, R = M.cuda(), a.cuda(), b.cuda(), R.cuda()
    n, m = M.shape
    print("adjoint FD n=m=4096 eta=0.05 h=1e-6, 64 entries", flush=True)
    out = cg.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 1000)
    assert out["converged"]
    z0 = newton(M, a, b, eta, out["dual"], 2)
    T0 = torch.exp((z0[:n, None] + z0[None, n:] - M) / eta)
    g0 = float(torch.cat([T0.sum(1) - a, (T0.sum(0) - b)[:-1]]).norm())
    # analytic gradient at the polished optimum through the library (warm start
    # at z0: the forward stops at once and its context is formed there)
    ctx_out = cg.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 10, x0=z0)
    G = 2 * ctx_out["plan"] + R
    grad = cg.solve_backward(ctx_out["plan"], G, ctx_out["context"], eta, tol=1e-14)["grad"]
    rng = np.random.default_rng(1)
    idx = [(int(rng.integers(n)), int(rng.integers(m))) for _ in range(64)]
    h = 1e-6
    fd, an = [], []
    for (i, j) in idx:
        Mp, Mm = M.clone(), M.clone()
        Mp[i, j] += h
        Mm[i, j] -= h
        zp = newton(Mp, a, b, eta, z0, 2)
        zm = newton(Mm, a, b, eta, z0, 2)
        up = (zp[:n, None] + zp[None, n:] - Mp) / eta
        um = (zm[:n, None] + zm[None, n:] - Mm) / eta
        Tm = torch.exp(um)
        D = Tm * torch.expm1(up - um)            # T(M + h e_ij) - T(M - h e_ij)
        dL = float((D * (2 * Tm + D + R)).sum())  # sum(Tp^2 - Tm^2) + <Tp - Tm, R>
        fd.append(dL / (2 * h))
        an.append(float(grad[i, j]))
        del Mp, Mm, up, um, Tm, D
    fd, an = np.array(fd), np.array(an)
    err = np.abs(fd - an).max() / np.abs(an).max()
    print(f"base |g| {g0:.2e}; max |dL/dM| on the sample {np.abs(an).max():.3e}; "
          f"max |FD - analytic| {np.abs(fd - an).max():.3e}; normwise {err:.2e}", flush=True)
    assert g0 < 1e-14
    assert err < 1e-6
