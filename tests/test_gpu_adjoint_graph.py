"""The adjoint's PCG loop runs its steady-state chunks (32 steps, from the second on) as a
captured HIP graph, replayed between the loop's read-backs; with the
library's profiler on it runs every step eagerly.  Both must give bitwise the
same lambda, gradient and iteration counts, for a batch whose instances stop
at different steps (the maps are repacked mid-loop, so the graph is captured
again) and for the preconditioners that allow capture."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _solves(cg, dev):
    from curegot_backward_amd.workloads import make_problem, serving_problem
    out = []
    # Dirichlet marginals with small concentration leave some masses near zero:
    # the adjoint system is ill-conditioned and PCG runs 100-600 steps, unevenly
    probs = [serving_problem({"n": n, "m": m, "d": 16, "eta": eta, "alpha": al, "seed": s})
             for (n, m, eta, al, s) in [(200, 300, 0.1, 0.1, 11), (150, 400, 0.1, 0.05, 12),
                                        (300, 250, 0.05, 0.1, 13), (120, 500, 0.1, 0.2, 14)]]
    M, a, b = make_problem(300, 400, 1)          # a well-conditioned one stops early
    probs.append((M, a, b, 0.01))
    for k, (M, a, b, eta) in enumerate(probs):
        M, a, b = M.to(dev), a.to(dev), b.to(dev)
        o = cg.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 1000)
        G = torch.randn(M.shape, generator=torch.Generator().manual_seed(k), dtype=torch.float64).to(dev)
        out.append((G, o["context"], eta))
    return out


@pytest.mark.parametrize("pc", ["none", "diag", "splr"])
def test_graph_chunks_equal_eager(pc):
    import curegot_backward_amd as cg
    from curegot_backward_amd import _lib
    from curegot_backward_amd.backward import solve_backward_batch
    dev = torch.device("cuda", 0)
    items = _solves(cg, dev)
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    runs = {}
    for mode in ("graph", "eager", "graph-side"):
        _lib.profile_enable(mode == "eager")
        try:
            if mode == "graph-side":            # capture on the solve stream itself
                side.wait_stream(torch.cuda.current_stream())
                with torch.cuda.stream(side):
                    res = solve_backward_batch(items, preconditioner=pc, max_iter=1500)
                side.synchronize()
            else:
                res = solve_backward_batch(items, preconditioner=pc, max_iter=1500)
                torch.cuda.synchronize()
        finally:
            _lib.profile_enable(False)
            _lib.profile_reset()
        runs[mode] = res
    its = [r["iterations"] for r in runs["eager"]]
    print(pc, "iterations", its, flush=True)
    assert max(its) > 28 + 3 * 32, its          # eager to 60 steps, then several captured chunks
    assert len(set(its)) > 1                     # uneven: the maps were repacked mid-loop
    for mode in ("graph", "graph-side"):
        for r, e in zip(runs[mode], runs["eager"]):
            assert r["iterations"] == e["iterations"] and r["converged"] == e["converged"]
            assert torch.equal(r["lambda"], e["lambda"])
            assert torch.equal(r["grad"], e["grad"])
