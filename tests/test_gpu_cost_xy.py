"""The geometric cost (M formed from X, Y inside the kernels): every solve on a
GeometricCost is bitwise the solve on the stored matrix of the same formula
(built on the CPU by the oracle's or_cost_xy, the formula pinned in
test_oracle_cost.py), for the SPLR forward, the BCD (log-sum-exp) solve and
the adjoint; the forward matches the oracle's trajectory; splr_ot_xy's
gradients reach X and Y as torch's chain rule through the materialized M
gives them; ragged batches of geometric instances are bitwise their single
solves, and a batch that mixes cost kinds is refused."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cg():
    import curegot_backward_amd
    return curegot_backward_amd


def points(n, m, d, seed):
    rng = np.random.default_rng(seed)
    X = rng.random((n, d))
    Y = rng.random((m, d))
    a = rng.random(n) + 0.5
    b = rng.random(m) + 0.5
    return X, Y, a / a.sum(), b / b.sum()


def cuda(*xs):
    return [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in xs]


@pytest.mark.parametrize("n,m,d,scale,eta", [(300, 260, 8, 1 / 8, 0.05), (200, 333, 3, 0.37, 0.05),
                                              (130, 70, 1, 1.0, 0.1), (640, 520, 32, 1 / 32, 0.02)])
def test_splr_xy_bitwise_equals_stored_cost(cg, n, m, d, scale, eta):
    X, Y, a, b = points(n, m, d, n + m + d)
    M = oracle.cost_xy(X, Y, scale)
    Xg, Yg, ag, bg, Mg = cuda(X, Y, a, b, M)
    print(f"case n={n} m={m} d={d} scale={scale} eta={eta}", flush=True)
    ref = cg.sinkhorn_splr_bwd(Mg, ag, bg, eta, 1e-8, 1000)
    out = cg.sinkhorn_splr_bwd(cg.GeometricCost(Xg, Yg, scale), ag, bg, eta, 1e-8, 1000)
    assert out["converged"] and out["niter"] == ref["niter"]
    assert torch.equal(out["dual"], ref["dual"]) and torch.equal(out["plan"], ref["plan"])
    # the adjoint through either context: the same lambda and dL/dM, bitwise
    R = torch.from_numpy(np.random.default_rng(1).normal(size=(n, m))).cuda()
    G = 2 * ref["plan"] + R
    r1 = cg.solve_backward(ref["plan"], G, ref["context"], eta)
    r2 = cg.solve_backward(out["plan"], G, out["context"], eta)
    assert r1["iterations"] == r2["iterations"]
    assert torch.equal(r1["lambda"], r2["lambda"]) and torch.equal(r1["grad"], r2["grad"])


def test_splr_xy_vs_oracle(cg):
    n, m, d, eta = 240, 210, 8, 0.05
    X, Y, a, b = points(n, m, d, 5)
    M = oracle.cost_xy(X, Y, 1 / d)
    ora = oracle.splr_bwd(M, a, b, eta, tol=1e-8, max_iter=1000)
    Xg, Yg, ag, bg = cuda(X, Y, a, b)
    out = cg.sinkhorn_splr_bwd(cg.GeometricCost(Xg, Yg), ag, bg, eta, 1e-8, 1000)
    print(f"niter gpu {out['niter']} oracle {ora['niter']}", flush=True)
    assert out["niter"] == ora["niter"]
    d_ora = torch.from_numpy(ora["dual"]).cuda()
    err = float((out["dual"] - d_ora).abs().max() / d_ora.abs().max())
    assert err < 1e-10, err


@pytest.mark.parametrize("d", [2, 8])
def test_bcd_xy_bitwise_equals_stored_cost(cg, d):
    n, m = 500, 450
    X, Y, a, b = points(n, m, d, 40 + d)
    M = oracle.cost_xy(X, Y, 1 / d)
    Xg, Yg, ag, bg, Mg = cuda(X, Y, a, b, M)
    ref = cg.sinkhorn_bcd(Mg, ag, bg, 0.05, 1e-9, 500)
    out = cg.sinkhorn_bcd(cg.GeometricCost(Xg, Yg), ag, bg, 0.05, 1e-9, 500)
    assert out["niter"] == ref["niter"]
    assert torch.equal(out["dual"], ref["dual"]) and torch.equal(out["plan"], ref["plan"])
    ora = oracle.sinkhorn_bcd(M, a, b, 0.05, tol=1e-9, max_iter=500)
    assert np.abs(out["plan"].cpu().numpy() - ora["plan"]).max() < 1e-12


def test_splr_ot_xy_gradients_reach_the_points(cg):
    n, m, d, eta = 48, 48, 8, 0.05
    X, Y, a, b = points(n, m, d, 11)
    Xg, Yg, ag, bg = cuda(X, Y, a, b)
    R = torch.from_numpy(np.random.default_rng(12).normal(size=(n, m))).cuda()
    # the library path on the points
    X1, Y1 = Xg.clone().requires_grad_(True), Yg.clone().requires_grad_(True)
    T1 = cg.splr_ot_xy(X1, Y1, ag, bg, eta=eta, tol=1e-8, density=0.1)
    ((T1 ** 2).sum() + (T1 * R).sum()).backward()
    # torch's chain rule through the materialized M, solved on the stored cost
    X2, Y2 = Xg.clone().requires_grad_(True), Yg.clone().requires_grad_(True)
    M2 = ((X2[:, None, :] - Y2[None, :, :]) ** 2).sum(-1) / d
    T2 = cg.splr_ot(M2, ag, bg, eta=eta, tol=1e-8, density=0.1)
    ((T2 ** 2).sum() + (T2 * R).sum()).backward()
    eX = float((X1.grad - X2.grad).abs().max() / X2.grad.abs().max())
    eY = float((Y1.grad - Y2.grad).abs().max() / Y2.grad.abs().max())
    print(f"dL/dX {eX:.2e} dL/dY {eY:.2e} against the chain rule through M", flush=True)
    assert eX < 1e-9 and eY < 1e-9


def test_geometric_batch_bitwise_single_and_no_mixing(cg):
    shapes = [(120, 90, 2), (70, 150, 8), (200, 180, 4)]
    costs, As, Bs = [], [], []
    for k, (n, m, d) in enumerate(shapes):
        X, Y, a, b = points(n, m, d, 70 + k)
        Xg, Yg, ag, bg = cuda(X, Y, a, b)
        costs.append(cg.GeometricCost(Xg, Yg))
        As.append(ag)
        Bs.append(bg)
    outs = cg.sinkhorn_splr_bwd_batch(costs, As, Bs, 0.05, 1e-8, 1000)
    for c, a, b, o in zip(costs, As, Bs, outs):
        s = cg.sinkhorn_splr_bwd(c, a, b, 0.05, 1e-8, 1000)
        assert o["niter"] == s["niter"] and torch.equal(o["dual"], s["dual"])
    M0 = cuda(oracle.cost_xy(costs[0].X.cpu().numpy(), costs[0].Y.cpu().numpy(), costs[0].scale))[0]
    with pytest.raises(ValueError, match="mix"):
        cg.sinkhorn_splr_bwd_batch([M0, costs[1]], As[:2], Bs[:2], 0.05, 1e-8, 1000)
