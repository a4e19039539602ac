"""Ragged batching: every instance's result is bitwise the same alone and in a
batch, and splr_ot_batch equals per-slice splr_ot (reference validate_stage5_api.py:269-298)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cg():
    import curegot_backward_amd
    return curegot_backward_amd


def test_ragged_batch_bitwise_equals_single(cg):
    from curegot_backward_amd.workloads import ragged_batch
    probs = ragged_batch(B=10, seed=3, lo=40, hi=400)
    Ms = [p[0].cuda() for p in probs]
    As = [p[1].cuda() for p in probs]
    Bs = [p[2].cuda() for p in probs]
    etas = [p[3] for p in probs]
    outs = cg.sinkhorn_splr_bwd_batch(Ms, As, Bs, etas, 1e-8, 1000)
    for i in range(len(probs)):
        print("instance", i, tuple(Ms[i].shape), etas[i], flush=True)
        single = cg.sinkhorn_splr_bwd(Ms[i], As[i], Bs[i], etas[i], 1e-8, 1000)
        assert outs[i]["niter"] == single["niter"]
        assert torch.equal(outs[i]["plan"], single["plan"])
        assert torch.equal(outs[i]["dual"], single["dual"])


def test_concurrent_groups_bitwise_equal_one_group(cg):
    """The batch is split into concurrent size classes (own streams and host
    threads); the split must not change any result."""
    from curegot_backward_amd import _lib
    from curegot_backward_amd.workloads import ragged_batch
    probs = ragged_batch(B=24, seed=5, lo=64, hi=1200)
    Ms = [p[0].cuda() for p in probs]
    As = [p[1].cuda() for p in probs]
    Bs = [p[2].cuda() for p in probs]
    etas = [p[3] for p in probs]
    runs = {}
    try:
        for g in (1, 4, 2):
            _lib.set_max_groups(g)
            print("groups", g, flush=True)
            runs[g] = cg.sinkhorn_splr_bwd_batch(Ms, As, Bs, etas, 1e-8, 1000)
            torch.cuda.synchronize()
    finally:
        _lib.set_max_groups(4)
    for g in (4, 2):
        for i in range(len(probs)):
            assert runs[g][i]["niter"] == runs[1][i]["niter"]
            assert torch.equal(runs[g][i]["dual"], runs[1][i]["dual"])
            assert torch.equal(runs[g][i]["plan"], runs[1][i]["plan"])
            # the context views of each group's workspace region
            assert torch.equal(runs[g][i]["context"].inv_diag(), runs[1][i]["context"].inv_diag())
    with pytest.raises(ValueError):
        _lib.set_max_groups(0)


def test_splr_ot_batch_equals_per_slice(cg):
    from curegot_backward_amd.workloads import make_problem
    B, n, m = 4, 50, 60
    Ms, As, Bs = zip(*[make_problem(n, m, s) for s in range(B)])
    M = torch.stack(Ms).cuda().requires_grad_(True)
    a, b = torch.stack(As).cuda(), torch.stack(Bs).cuda()
    T = cg.torch.splr_ot_batch(M, a, b, eta=0.05, tol=1e-8)
    (T ** 2).sum().backward()
    for i in range(B):
        Mi = M[i].detach().clone().requires_grad_(True)
        Ti = cg.torch.splr_ot(Mi, a[i], b[i], eta=0.05, tol=1e-8)
        (Ti ** 2).sum().backward()
        assert torch.equal(T[i].detach(), Ti.detach())
        assert (M.grad[i] - Mi.grad).abs().max() <= 1e-12 * Mi.grad.abs().max()


def test_ragged_list_batch_with_backward(cg):
    from curegot_backward_amd.workloads import make_problem
    shapes = [(30, 45), (64, 20), (17, 17)]
    Ms = [make_problem(n, m, i)[0].cuda().requires_grad_(True) for i, (n, m) in enumerate(shapes)]
    As = [make_problem(n, m, i)[1].cuda() for i, (n, m) in enumerate(shapes)]
    Bs = [make_problem(n, m, i)[2].cuda() for i, (n, m) in enumerate(shapes)]
    Ts = cg.torch.splr_ot_batch(Ms, As, Bs, eta=[0.05, 0.02, 0.1], tol=1e-8)
    sum((T ** 2).sum() for T in Ts).backward()
    for i, (n, m) in enumerate(shapes):
        Mi = Ms[i].detach().clone().requires_grad_(True)
        Ti = cg.torch.splr_ot(Mi, As[i], Bs[i], eta=[0.05, 0.02, 0.1][i], tol=1e-8)
        (Ti ** 2).sum().backward()
        assert torch.equal(Ts[i].detach(), Ti.detach())
        assert torch.equal(Ms[i].grad, Mi.grad)


def test_batched_adjoint_equals_single_uneven_iterations(cg):
    """Instances whose PCG runs stop at very different iteration counts: every
    gradient of the batched adjoint equals the single solve's bitwise (the
    final assembly must cover instances that finished early)."""
    from curegot_backward_amd.backward import solve_backward_batch
    from curegot_backward_amd.workloads import serving_problem
    specs = [{"n": 395, "m": 69, "d": 16, "eta": 0.05, "alpha": 1.0, "seed": 1007},
             {"n": 64, "m": 611, "d": 4, "eta": 0.01, "alpha": 0.6, "seed": 1004},
             {"n": 300, "m": 280, "d": 64, "eta": 0.1, "alpha": 1.8, "seed": 1010}]
    g = torch.Generator().manual_seed(3)
    items, singles = [], []
    for sp in specs:
        M, a, b, eta = serving_problem(sp)
        M, a, b = M.cuda(), a.cuda(), b.cuda()
        G = torch.randn(M.shape, dtype=torch.float64, generator=g).cuda()
        o = cg.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 1000)
        items.append((G, o["context"], eta))
        singles.append(cg.solve_backward(o["plan"], G, o["context"], eta, preconditioner="none"))
    res = solve_backward_batch(items, preconditioner="none")
    iters = [r["iterations"] for r in res]
    print("iterations", iters, flush=True)
    for r, s in zip(res, singles):
        assert r["iterations"] == s["iterations"]
        assert torch.equal(r["grad"], s["grad"])
        assert torch.equal(r["lambda"], s["lambda"])


def test_full_size_ragged_batch_properties(cg):
    # The benchmarked workload at its full size: 128 instances, n, m log-uniform
    # in [128, 4096], one batched call.  Size-independent checks per instance:
    # convergence, the stop rule recomputed from the returned plan
    # (||[T1 - a; (T^T 1 - b)[:-1]]|| < tol, sinkhorn_splr.cu:1375-1382), the
    # last column's marginal through total mass, and the plan equal to
    # exp((alpha_i + beta_j - M_ij)/eta) of the returned dual.
    from curegot_backward_amd.workloads import ragged_batch
    probs = ragged_batch(128)
    Ms = [p[0].cuda() for p in probs]
    As = [p[1].cuda() for p in probs]
    Bs = [p[2].cuda() for p in probs]
    etas = [p[3] for p in probs]
    tol = 1e-8
    outs = cg.sinkhorn_splr_bwd_batch(Ms, As, Bs, etas, tol, 1000)
    print("full batch: niter max", max(o["niter"] for o in outs), flush=True)
    for k, (M, a, b, eta, o) in enumerate(zip(Ms, As, Bs, etas, outs)):
        n, m = M.shape
        assert o["converged"], k
        T, dual = o["plan"], o["dual"]
        assert torch.isfinite(T).all() and (T >= 0).all(), k
        g = torch.cat([T.sum(1) - a, (T.sum(0) - b)[:-1]])
        # the recomputed sums differ from the solver's own by summation rounding
        assert float(g.norm()) < tol * (1 + 1e-3) + 1e-14 * (n + m), (k, float(g.norm()))
        # |sum_i (T1 - a)_i - sum_{j<m} (T^T 1 - b)_j| <= sqrt(n + m) ||g|| (Cauchy-Schwarz)
        assert abs(float(T.sum(0)[-1] - b[-1])) < tol * (n + m) ** 0.5 + 1e-13, k
        Tx = torch.exp((dual[:n, None] + dual[None, n:] - M) / eta)
        assert float((T - Tx).abs().max() / Tx.abs().max()) < 1e-12, k
