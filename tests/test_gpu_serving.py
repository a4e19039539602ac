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


def test_served_results_equal_solo():
    import curegot_backward_amd as cg
    from curegot_backward_amd.serving import OTServer
    reqs = _problems(18, 7)
    with OTServer(slots=3, max_batch=5, tol=1e-8) as srv:
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
