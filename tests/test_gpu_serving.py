"""Serving loop: requests admitted while others run give bitwise the same
results as solving each alone (forward plan/dual and, for backward requests,
the gradient)."""
import threading
import time

import pytest
import torch

pytestmark = pytest.mark.gpu


def _problems(k, seed):
    from curegot_backward_amd.workloads import serving_problem
    reqs = []
    g = torch.Generator().manual_seed(seed)
    for i in range(k):
        n = int(torch.randint(40, 700, (1,), generator=g))
        m = int(torch.randint(40, 700, (1,), generator=g))
        spec = {"n": n, "m": m, "d": (4, 16, 64)[i % 3], "eta": (0.01, 0.05, 0.1)[i % 3],
                "alpha": 1.0, "seed": 1000 + i, "backward": i % 3 == 1}
        M, a, b, eta = serving_problem(spec)
        G = torch.randn(n, m, dtype=torch.float64, generator=g) if spec["backward"] else None
        reqs.append((M.cuda(), a.cuda(), b.cuda(), eta, None if G is None else G.cuda()))
    return reqs


@pytest.mark.parametrize("continuous", [True, False])
def test_served_results_equal_solo(continuous):
    """Both admission modes: at iteration boundaries into running groups
    (continuous) and per batch."""
    import curegot_backward_amd as cg
    from curegot_backward_amd.serving import OTServer
    reqs = _problems(18, 7)
    with OTServer(slots=3, max_batch=5, tol=1e-8, continuous=continuous) as srv:
        futs = []
        for i, (M, a, b, eta, G) in enumerate(reqs):
            futs.append(srv.submit(M, a, b, eta, G))
            if i % 4 == 3:
                time.sleep(0.01)          # later requests arrive while batches run
        outs = [f.result(timeout=120) for f in futs]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["batches"] >= 2
    for i, ((M, a, b, eta, G), o) in enumerate(zip(reqs, outs)):
        print("request", i, tuple(M.shape), eta, G is not None, flush=True)
        solo = cg.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 1000)
        assert o["niter"] == solo["niter"]
        assert torch.equal(o["plan"], solo["plan"])
        assert torch.equal(o["dual"], solo["dual"])
        if G is not None:
            ref = cg.solve_backward(solo["plan"], G, solo["context"], eta)
            assert o["bwd_converged"]
            assert torch.equal(o["grad"], ref["grad"])


def test_budget_admission_and_rejection():
    from curegot_backward_amd.serving import OTServer, splr_footprint
    reqs = _problems(6, 11)
    big = max(splr_footprint(M.shape[0], M.shape[1], backward=G is not None) for M, a, b, eta, G in reqs)
    # a budget that holds one request at a time: everything still completes
    with OTServer(slots=3, hbm_budget=big, max_batch=8) as srv:
        outs = [f.result(timeout=120) for f in [srv.submit(*r) for r in reqs]]
        assert all(o["converged"] for o in outs)
        assert srv.stats()["batches"] >= 2
    with OTServer(slots=1, hbm_budget=1024) as srv:
        M, a, b, eta, G = reqs[0]
        with pytest.raises(MemoryError):
            srv.submit(M, a, b, eta).result(timeout=10)


def test_served_trace_head_equals_solo():
    """The head of the serving workload's Poisson trace (rate 2000/s, n, m
    log-uniform in [64, 8192], d in {4, 16, 64}, Dirichlet marginals, eta in
    {0.01, 0.05, 0.1}, 30% with an upstream gradient): the first 40 requests
    (sizes up to 7015 x 7448) submitted at their trace times; every result is
    bitwise the same as solving the request alone."""
    import curegot_backward_amd as cg
    from curegot_backward_amd.serving import OTServer
    from curegot_backward_amd.workloads import serving_problem, serving_trace
    trace = serving_trace(seconds=0.05)[:40]
    g = torch.Generator().manual_seed(17)
    reqs = []
    for r in trace:
        M, a, b, eta = serving_problem(r, device="cuda")
        G = torch.randn(M.shape, dtype=torch.float64, generator=g).cuda() if r["backward"] else None
        reqs.append((M, a, b, eta, G))
    torch.cuda.synchronize()
    print("trace head: 40 requests, largest", max(M.shape for M, *_ in reqs), flush=True)
    with OTServer(slots=4, tol=1e-8) as srv:
        t0 = time.perf_counter()
        futs = []
        for r, (M, a, b, eta, G) in zip(trace, reqs):
            delay = r["t"] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            futs.append(srv.submit(M, a, b, eta, G))
        outs = [f.result(timeout=300) for f in futs]
        st = srv.stats()
    print("served:", {k: st[k] for k in ("completed", "batches", "p50_ms", "p99_ms")}, flush=True)
    assert st["completed"] == len(reqs)
    for i, ((M, a, b, eta, G), o) in enumerate(zip(reqs, outs)):
        solo = cg.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 1000)
        print("request", i, tuple(M.shape), eta, G is not None, "niter", o["niter"], flush=True)
        assert o["converged"]
        assert o["niter"] == solo["niter"]
        assert torch.equal(o["plan"], solo["plan"])
        assert torch.equal(o["dual"], solo["dual"])
        if G is not None:
            # the adjoint (default diag preconditioner, tol 1e-12, 2000 iterations)
            # is held to the same bar: bitwise equal to the solo solve, iteration
            # count and convergence flag included
            ref = cg.solve_backward(solo["plan"], G, solo["context"], eta)
            print("  backward iterations", o["bwd_iterations"], "converged", o["bwd_converged"], flush=True)
            assert o["bwd_converged"] == ref["converged"] and o["bwd_iterations"] == ref["iterations"]
            assert torch.equal(o["grad"], ref["grad"])


def test_results_outlive_slot_stream():
    """A result read on the submitter's own (non-default) stream and dropped
    while the server keeps solving: the block must not be handed to the slot's
    next batch before the submitter's stream has finished with it."""
    from curegot_backward_amd.serving import OTServer
    reqs = _problems(6, 23)
    side = torch.cuda.Stream()
    with OTServer(slots=1, max_batch=1, tol=1e-8) as srv:
        with torch.cuda.stream(side):
            M, a, b, eta, _ = reqs[0]
            side.wait_stream(torch.cuda.default_stream())
            plan = srv.submit(M, a, b, eta).result(timeout=60)["plan"]
            expect = float(plan.sum())
            acc = torch.zeros_like(plan)
            for _ in range(200):          # a long read of plan on the side stream
                acc = acc + plan
            del plan
            futs = [srv.submit(M, a, b, eta) for _ in range(3)]   # same shape: same block size
            got = float(acc.sum()) / 200.0
        outs = [f.result(timeout=60) for f in futs]
    torch.cuda.synchronize()
    assert abs(got - expect) <= 1e-12 * abs(expect)
    for o in outs:
        assert abs(float(o["plan"].sum()) - expect) <= 1e-12 * abs(expect)
