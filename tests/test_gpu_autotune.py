"""Tile table (what the autotuner picks): under a non-default table of columns
per lane the forward and backward still match the oracle to 1e-10 with the
oracle's iteration counts, a batch and the single solves of its instances
agree bitwise, and the tuner returns a valid table and leaves the process's
table as it found it."""
import numpy as np
import pytest
import torch

import oracle
from tests.test_gpu_parity import ctx_dict, rel

pytestmark = pytest.mark.gpu

TABLES = {"c1": [(8192, 1)], "c1-c2": [(128, 1), (8192, 2)]}
CASES = [(100, 120, 0.05, 3), (300, 257, 0.01, 8), (65, 600, 0.1, 5), (200, 1100, 0.02, 11)]


@pytest.fixture
def lib():
    from curegot_backward_amd import _lib
    _lib.set_tile_table([])
    yield _lib
    _lib.set_tile_table([])


@pytest.mark.parametrize("name", sorted(TABLES))
def test_forward_backward_vs_oracle_under_table(lib, name):
    import curegot_backward_amd as cg
    from curegot_backward_amd.workloads import make_problem
    lib.set_tile_table(TABLES[name])
    for (n, m, eta, seed) in CASES:
        want = next(c for mx, c in TABLES[name] if m <= mx)
        assert lib.tile_cols(m) == want
        M, a, b = make_problem(n, m, seed)
        ref = oracle.splr_bwd(M.numpy(), a.numpy(), b.numpy(), eta, tol=1e-8, max_iter=500)
        out = cg.sinkhorn_splr_bwd(M.cuda(), a.cuda(), b.cuda(), eta, 1e-8, 500)
        print(f"table {name} case n={n} m={m} c={want}: niter gpu {out['niter']} oracle {ref['niter']}", flush=True)
        assert out["converged"] == ref["converged"] and out["niter"] == ref["niter"]
        assert rel(out["plan"], ref["plan"]) < 1e-10
        assert rel(out["dual"], ref["dual"]) < 1e-10
        G = torch.randn(n, m, dtype=torch.float64, generator=torch.Generator().manual_seed(seed))
        ctx = out["context"]
        res = cg.solve_backward(out["plan"], G.cuda(), ctx, eta, preconditioner="diag")
        rb = oracle.backward(M.numpy(), out["dual"].cpu().numpy(), G.numpy(), eta, ctx=ctx_dict(ctx),
                             preconditioner="diag")
        assert res["converged"] == rb["converged"] and abs(res["iterations"] - rb["iterations"]) <= 1
        err = float(np.abs(res["grad"].cpu().numpy() - rb["grad"]).max())
        assert err <= 1e-10 * np.abs(rb["grad"]).max(), err


def test_batch_equals_single_under_table(lib):
    import curegot_backward_amd as cg
    from curegot_backward_amd.workloads import make_problem
    lib.set_tile_table(TABLES["c1-c2"])
    probs = [make_problem(n, m, seed) + (eta,) for (n, m, eta, seed) in
             CASES + [(40, 30, 0.05, 1), (129, 65, 0.02, 4), (500, 300, 0.05, 12), (64, 2000, 0.05, 13)]]
    Ms = [p[0].cuda() for p in probs]
    As = [p[1].cuda() for p in probs]
    Bs = [p[2].cuda() for p in probs]
    regs = [p[3] for p in probs]
    batch = cg.sinkhorn_splr_bwd_batch(Ms, As, Bs, regs, 1e-8, 500)
    for i in range(len(probs)):
        one = cg.sinkhorn_splr_bwd(Ms[i], As[i], Bs[i], regs[i], 1e-8, 500)
        assert batch[i]["niter"] == one["niter"], i
        assert torch.equal(batch[i]["plan"], one["plan"]) and torch.equal(batch[i]["dual"], one["dual"]), i


def test_tuner_returns_valid_table_and_restores(lib):
    from curegot_backward_amd import autotune
    lib.set_tile_table([(96, 1)])
    res = autotune.tune(bands=(128, 256), iters=2, reps=1, cells=1 << 20, max_batch=8, log=print)
    assert lib.get_tile_table() == [(96, 1)]
    assert [mx for mx, _ in res["table"]] == [128, 256]
    for (mx, c), band in zip(res["table"], res["bands"]):
        assert c in autotune.candidates(mx) and band["pick"] == c
        assert set(band["ms"]) == {str(k) for k in autotune.candidates(mx)}
    autotune.apply(res)
    assert lib.get_tile_table() == res["table"]
