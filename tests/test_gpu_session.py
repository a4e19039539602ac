"""Sessions (continuous batching, C ABI cgot_session_*): instances added at
different iterations of a running lockstep group, with slots reused, give
bitwise the results of solving each alone (plan, dual, iteration count, and
the backward through the context the session returns)."""
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cg():
    import curegot_backward_amd
    return curegot_backward_amd


def problems():
    from curegot_backward_amd.workloads import make_problem, ragged_batch
    out = []
    for k, (n, m, eta, seed) in enumerate([(64, 80, 0.05, 0), (150, 90, 0.02, 1), (37, 210, 0.1, 2),
                                           (300, 257, 0.01, 3), (120, 120, 0.05, 4), (90, 400, 0.02, 5)]):
        M, a, b = make_problem(n, m, seed)
        out.append((M, a, b, eta))
    # a few instances of the benchmarked ragged batch (the small and mid ones)
    rb = ragged_batch(128, 0)
    for i in (112, 105, 73, 64):
        out.append(rb[i])
    return out


def test_session_matches_solo(cg):
    from curegot_backward_amd._native import SplrSession
    dev = torch.device("cuda", 0)
    probs = [(M.to(dev), a.to(dev), b.to(dev), eta) for (M, a, b, eta) in problems()]
    solo = [cg.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 1000) for (M, a, b, eta) in probs]
    torch.cuda.synchronize()
    sess = SplrSession(device=dev, capacity=4, max_sum_nm=4096, max_tiles=4096, max_nblk=16)
    got, queue, step = {}, list(range(len(probs))), 0
    print(f"session: {len(probs)} instances, capacity 4", flush=True)
    while queue or sess.live() > 0:
        # add up to two per step while there is room (later instances join a
        # group already iterating); a full group refuses without keeping anything
        for _ in range(2):
            if not queue:
                break
            i = queue[0]
            M, a, b, eta = probs[i]
            if not sess.add(M, a, b, eta, 1e-8, 1000, tag=i):
                break
            queue.pop(0)
        for tag, res in sess.step(iters=3):
            got[tag] = (res, step)
        step += 1
        assert step < 2000
    sess.close()
    assert sorted(got) == list(range(len(probs)))
    for i, (res, at) in got.items():
        ref = solo[i]
        assert res["niter"] == ref["niter"] and res["converged"] == ref["converged"], (i, res["niter"], ref["niter"])
        assert torch.equal(res["plan"], ref["plan"]), i
        assert torch.equal(res["dual"], ref["dual"]), i
    # the backward through the session's context equals the solo one, for each
    # preconditioner that reads the stored forward state
    g = torch.Generator(device="cpu").manual_seed(3)
    for i in (0, 3, 7):
        M, a, b, eta = probs[i]
        G = torch.randn(M.shape, generator=g, dtype=torch.float64).to(dev)
        for pc in ("diag", "splr"):
            r1 = cg.solve_backward(got[i][0]["plan"], G, got[i][0]["context"], eta, preconditioner=pc)
            r0 = cg.solve_backward(solo[i]["plan"], G, solo[i]["context"], eta, preconditioner=pc)
            assert r1["iterations"] == r0["iterations"], (i, pc)
            assert torch.equal(r1["grad"], r0["grad"]), (i, pc)


def test_session_limits(cg):
    from curegot_backward_amd._native import SplrSession
    from curegot_backward_amd.workloads import make_problem
    dev = torch.device("cuda", 0)
    M, a, b = make_problem(64, 80, 0)
    M, a, b = M.to(dev), a.to(dev), b.to(dev)
    sess = SplrSession(device=dev, capacity=2, max_sum_nm=1000, max_tiles=1000, max_nblk=4)
    assert sess.add(M, a, b, 0.05, 1e-8, 1000, tag="x")
    assert sess.add(M, a, b, 0.05, 1e-8, 1000, tag="y")
    assert not sess.add(M, a, b, 0.05, 1e-8, 1000, tag="z")          # capacity
    done = []
    while sess.live() > 0:
        done += sess.step(iters=50)
    assert sorted(t for t, _ in done) == ["x", "y"]
    assert torch.equal(done[0][1]["plan"], done[1][1]["plan"])
    M2, a2, b2 = make_problem(600, 500, 1)
    assert not sess.add(M2.to(dev), a2.to(dev), b2.to(dev), 0.05, 1e-8, 1000)   # sum of n + m over 1000
    sess.close()
