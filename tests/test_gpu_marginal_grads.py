"""dL/da and dL/db through splr_ot(..., marginal_grads=True) and
splr_ot_batch: against the oracle's adjoint lambda at the GPU's own duals
(dL/da = lambda_alpha, dL/db = (lambda_beta, 0), the identity pinned by central
differences in test_oracle_backward.py), against central differences of the
loss on the GPU (perturbed problems solved by exact Newton steps from the
polished optimum), batch results bitwise equal to single ones, and the
reference's NotImplementedError kept without the flag."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cg():
    import curegot_backward_amd
    return curegot_backward_amd


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


def problem(n, m, seed):
    from curegot_backward_amd.workloads import make_problem
    return [x.cuda() for x in make_problem(n, m, seed)]


def test_marginal_grads_vs_oracle_lambda(cg):
    n, m, eta = 40, 35, 0.1
    M, a, b = problem(n, m, 21)
    R = torch.from_numpy(np.random.default_rng(22).normal(size=(n, m))).cuda()
    a = a.clone().requires_grad_(True)
    b = b.clone().requires_grad_(True)
    Mg = M.clone().requires_grad_(True)
    T = cg.torch.splr_ot(Mg, a, b, eta=eta, tol=1e-8, marginal_grads=True, backward_tol=1e-14)
    ((T ** 2).sum() + (T * R).sum()).backward()
    info = cg.torch.splr_ot_with_info(M, a.detach(), b.detach(), eta=eta, tol=1e-8)
    G = (2 * T + R).detach().cpu().numpy()
    ref = oracle.backward(M.cpu().numpy(), info["dual"].cpu().numpy(), G, eta, preconditioner="none",
                          tol=1e-14)
    lam = ref["lambda"]
    ea = float(np.abs(a.grad.cpu().numpy() - lam[:n]).max() / np.abs(lam).max())
    eb = float(np.abs(b.grad.cpu().numpy() - np.concatenate([lam[n:], [0.0]])).max() / np.abs(lam).max())
    eM = float(np.abs(Mg.grad.cpu().numpy() - ref["grad"]).max() / np.abs(ref["grad"]).max())
    print(f"n={n} m={m}: dL/da {ea:.2e}, dL/db {eb:.2e}, dL/dM {eM:.2e} vs the oracle", flush=True)
    assert ea < 1e-9 and eb < 1e-9 and eM < 1e-9
    assert float(b.grad[-1]) == 0.0


def test_marginal_grads_vs_central_fd(cg):
    n, m, eta = 300, 260, 0.05
    M, a, b = problem(n, m, 23)
    R = torch.from_numpy(np.random.default_rng(24).normal(size=(n, m))).cuda()
    out = cg.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 500)
    assert out["converged"]
    z0 = newton(M, a, b, eta, out["dual"], 2)
    ctx_out = cg.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 10, x0=z0)
    G = 2 * ctx_out["plan"] + R
    res = cg.solve_backward(ctx_out["plan"], G, ctx_out["context"], eta, tol=1e-14)
    h = 1e-7
    rng = np.random.default_rng(25)
    picks = [("a", int(k)) for k in rng.choice(n, 8, replace=False)] + \
            [("b", int(k)) for k in rng.choice(m - 1, 8, replace=False)] + [("b", m - 1)]
    worst = 0.0
    scale = float(max(res["grad_a"].abs().max(), res["grad_b"].abs().max()))
    for name, k in picks:
        vp, vm = (a.clone(), a.clone()) if name == "a" else (b.clone(), b.clone())
        vp[k] += h
        vm[k] -= h
        ap, am = (vp, vm) if name == "a" else (a, a)
        bp, bm = (b, b) if name == "a" else (vp, vm)
        zp = newton(M, ap, bp, eta, z0, 3)
        zm = newton(M, am, bm, eta, z0, 3)
        up = (zp[:n, None] + zp[None, n:] - M) / eta
        um = (zm[:n, None] + zm[None, n:] - M) / eta
        Tm = torch.exp(um)
        D = Tm * torch.expm1(up - um)
        fd = float((D * (2 * Tm + D + R)).sum()) / (2 * h)
        an = float((res["grad_a"] if name == "a" else res["grad_b"])[k])
        worst = max(worst, abs(fd - an) / scale)
    print(f"n={n} m={m}: max |FD - analytic| / max |grad| = {worst:.2e} over {len(picks)} entries", flush=True)
    assert worst < 1e-6


def test_batch_marginal_grads_equal_single(cg):
    shapes = [(60, 50, 31), (45, 70, 32), (80, 80, 33)]
    eta = 0.05
    probs = [problem(n, m, s) for n, m, s in shapes]
    Rs = [torch.from_numpy(np.random.default_rng(s).normal(size=(n, m))).cuda() for n, m, s in shapes]
    singles = []
    for (M, a, b), R in zip(probs, Rs):
        a, b = a.clone().requires_grad_(True), b.clone().requires_grad_(True)
        T = cg.torch.splr_ot(M, a, b, eta=eta, tol=1e-8, marginal_grads=True)
        ((T ** 2).sum() + (T * R).sum()).backward()
        singles.append((a.grad.clone(), b.grad.clone()))
    As = [p[1].clone().requires_grad_(True) for p in probs]
    Bs = [p[2].clone().requires_grad_(True) for p in probs]
    Ts = cg.torch.splr_ot_batch([p[0] for p in probs], As, Bs, eta=eta, tol=1e-8, marginal_grads=True)
    sum(((T ** 2).sum() + (T * R).sum()) for T, R in zip(Ts, Rs)).backward()
    for k, (ga, gb) in enumerate(singles):
        assert torch.equal(As[k].grad, ga) and torch.equal(Bs[k].grad, gb), k


def test_marginal_grads_need_the_flag(cg):
    M, a, b = problem(8, 7, 0)
    with pytest.raises(NotImplementedError):
        cg.torch.splr_ot(M, a.clone().requires_grad_(True), b, eta=0.1)
    with pytest.raises(NotImplementedError):
        cg.torch.splr_ot_batch([M], [a], [b.clone().requires_grad_(True)], eta=0.1)
