"""End-to-end use of the drop-in: embeddings trained by Adam through the
differentiable OT layer, the workload of the reference's
examples/differentiable_splr_ot.py:25-64 (M = mean squared distance between
learnable points X and fixed points Y; the loss rewards the mass that T puts on
the pairs (i, i)).  The gradient reaching X on the first step is checked
against the oracle's implicit backward carried through the distance by hand;
that chain rule is pinned by central differences in test_oracle_backward.py."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cg():
    import curegot_backward_amd
    return curegot_backward_amd


def test_embedding_training_through_splrot(cg):
    n = m = 48
    d, eta, density = 8, 0.05, 0.10
    rng = np.random.default_rng(11)
    Y = torch.tensor(rng.random((m, d)), device="cuda")
    X = torch.nn.Parameter(torch.tensor(rng.random((n, d)), device="cuda"))
    a = torch.full((n,), 1.0 / n, dtype=torch.float64, device="cuda")
    b = torch.full((m,), 1.0 / m, dtype=torch.float64, device="cuda")
    ot = cg.torch.SPLROT(eta=eta, tol=1e-8, max_iter=1000, density=density)
    opt = torch.optim.Adam([X], lr=0.05)
    losses = []
    for step in range(40):
        opt.zero_grad()
        M = ((X[:, None, :] - Y[None, :, :]) ** 2).sum(dim=-1) / d
        T = ot(M, a, b)
        loss = -torch.diagonal(T).sum()
        loss.backward()
        assert torch.isfinite(X.grad).all()
        if step == 0:
            # dL/dX_i = (2/d) sum_j dM_ij (X_i - Y_j), with dM from the oracle's
            # solve and backward for G = dL/dT = -I
            Mh = M.detach().cpu().numpy()
            ref = oracle.splr_bwd(Mh, a.cpu().numpy(), b.cpu().numpy(), eta, tol=1e-8, max_iter=1000,
                                  density=density)
            assert ref["converged"]
            rb = oracle.backward(Mh, ref["dual"], -np.eye(n), eta, ctx=ref, preconditioner="diag")
            dM = rb["grad"]
            Xh, Yh = X.detach().cpu().numpy(), Y.cpu().numpy()
            dX = (2.0 / d) * (dM.sum(axis=1)[:, None] * Xh - dM @ Yh)
            err = np.abs(X.grad.cpu().numpy() - dX).max() / np.abs(dX).max()
            print(f"step 0: loss {loss.item():.6e}, dL/dX vs oracle {err:.2e}", flush=True)
            assert err < 1e-8
        opt.step()
        losses.append(loss.item())
    print(f"loss {losses[0]:.6e} -> {losses[-1]:.6e}", flush=True)
    assert losses[-1] < losses[0]
