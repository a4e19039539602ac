"""GPU parity on the full-size single-instance problem (n = m = 2048, eta 0.01,
density 0.01) and on every instance of the benchmarked ragged batch (B = 128,
n, m up to 4096), against the oracle's results committed in
tests/golden/batch_oracle.npz (written by tests/golden/make_batch_fixtures.py).

Per instance:
  * same iteration count as the oracle: the duals (beta[m-1] = 0 gauge) and the
    plan agree to 1e-10 relative (max norm);
  * different iteration count (the two runs took different line-search
    branches; both stop with ||g|| < tol): the instance must be in the
    committed list below, its gradient recomputed from the returned plan must
    be below tol, and one exact Newton step from its dual (dense Hessian, Schur
    complement, torch Cholesky) towards the problem's optimum z*
    (Newton-polished to ||g|| ~ 1e-16 in the fixture) must remove at least
    99.9% of the plan's distance to it (or leave less than 1e-10): the
    returned point then differs from the optimum by the first-order effect of
    its gradient, which is below tolerance.  The first iteration where the two
    trajectories part is printed from the two traces.
The split count is printed and the list asserted, so every instance is held to
one of the two stated bars.
"""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "batch_oracle.npz")

# Ragged-batch instances whose GPU iteration count differs from the oracle's;
# everything else must match.  In each, the two runs part in a strong-Wolfe
# sufficient-decrease test decided below the rounding resolution of f (the
# objectives agree to ~1e-15 relative, the required decrease c1*step*<g, d> is
# ~1e-20): one run accepts step 1, the other shrinks it.
BATCH_SPLITS = {0, 46, 76, 84, 87, 115}
SINGLE_SPLITS = False


@pytest.fixture(scope="module")
def fx():
    return np.load(FIXTURE)


@pytest.fixture(scope="module")
def cg():
    import curegot_backward_amd
    return curegot_backward_amd


def plan_of(M, dual, eta):
    n = M.shape[0]
    return torch.exp((dual[:n, None] + dual[None, n:] - M) / eta)


def maxrel(x, y):
    return float((x - y).abs().max() / y.abs().max().clamp_min(1e-300))


def reduced_grad(T, a, b):
    return torch.cat([T.sum(1) - a, (T.sum(0) - b)[:-1]])


def newton_step(M, a, b, eta, dual):
    """One exact Newton step on the reduced dual (H = Q(T)/eta, Schur
    complement onto the smaller side, torch Cholesky); returns the new dual."""
    n, m = M.shape
    T = plan_of(M, dual, eta)
    g = reduced_grad(T, a, b)
    W = T[:, :m - 1]
    Da, Db = T.sum(1), T.sum(0)[:m - 1]
    ga, gb = g[:n], g[n:]
    if m - 1 <= n:
        S = torch.diag(Db) - (W.T / Da) @ W
        xb = torch.cholesky_solve((gb - W.T @ (ga / Da))[:, None], torch.linalg.cholesky(S))[:, 0]
        xa = (ga - W @ xb) / Da
    else:
        S = torch.diag(Da) - (W / Db) @ W.T
        xa = torch.cholesky_solve((ga - W @ (gb / Db))[:, None], torch.linalg.cholesky(S))[:, 0]
        xb = (gb - W.T @ xa) / Db
    return dual - eta * torch.cat([xa, xb, xa.new_zeros(1)])


def first_divergence(tr_gpu, tr_ora):
    """First iteration whose f, ||g|| or step differ between the two traces
    (relative 1e-9 on f, 1e-6 on ||g|| and step), with the values there."""
    for it in range(min(len(tr_gpu), len(tr_ora))):
        fg, fo = tr_gpu[it][0], tr_ora[it][0]
        if (abs(fg - fo) > 1e-9 * abs(fo) or abs(tr_gpu[it][1] - tr_ora[it][1]) > 1e-6 * tr_ora[it][1]
                or abs(tr_gpu[it][2] - tr_ora[it][2]) > 1e-6 * abs(tr_ora[it][2])):
            return (f"parted at iteration {it}: |g| gpu {tr_gpu[it][1]:.6e} oracle {tr_ora[it][1]:.6e}, "
                    f"step gpu {tr_gpu[it][2]:.6e} oracle {tr_ora[it][2]:.6e}, "
                    f"f diff {abs(fg - fo) / abs(fo):.1e}")
    return "traces agree on the common iterations"


def judge(M, a, b, eta, out, dual_ora, zstar, tol, split_expected, label, tr_ora):
    """Returns (ok, split, worst error where the counts agree, message)."""
    dev = M.device
    d_ora = torch.from_numpy(dual_ora).to(dev)
    T = out["plan"]
    split = out["_split"]
    msg = f"{label}: "
    ok = bool(out["converged"])
    if split != split_expected:
        ok = False
        msg += f"SPLIT {split} NOT AS COMMITTED ({split_expected}); "
    if not split:
        e_dual = maxrel(out["dual"], d_ora)
        e_plan = maxrel(T, plan_of(M, d_ora, eta))
        ok = ok and e_dual < 1e-10 and e_plan < 1e-10
        return ok, split, max(e_dual, e_plan), msg + f"dual {e_dual:.2e} plan {e_plan:.2e}"
    gn = float(reduced_grad(T, a, b).norm())
    T_opt = plan_of(M, torch.from_numpy(zstar).to(dev), eta)
    e_gpu = maxrel(T, T_opt)
    e_newton = maxrel(plan_of(M, newton_step(M, a, b, eta, out["dual"]), eta), T_opt)
    ok = ok and gn < tol * (1 + 1e-3) and e_newton < max(1e-10, 1e-3 * e_gpu)
    msg += (f"split: |g| recomputed {gn:.2e}; plan to optimum {e_gpu:.2e}, after one Newton step {e_newton:.2e}; "
            + first_divergence(out["trace"], tr_ora))
    return ok, split, 0.0, msg


def test_single_problem_vs_oracle(cg, fx):
    """n = m = 2048, Gaussian points in R^16, eta 0.01, density 0.01 (global top-K)."""
    from curegot_backward_amd.workloads import single_problem
    M, a, b, eta = single_problem()
    M, a, b = M.cuda(), a.cuda(), b.cuda()
    print("single problem n=m=2048 eta=0.01 density=0.01", flush=True)
    out = cg.sinkhorn_splr_bwd(M, a, b, eta, float(fx["tol"]), int(fx["max_iter"]), density=0.01, trace=True)
    out["_split"] = out["niter"] != int(fx["single_niter"])
    ok, split, err, msg = judge(M, a, b, eta, out, fx["single_dual"], fx["single_zstar"], float(fx["tol"]),
                                SINGLE_SPLITS, f"single niter {out['niter']}/{int(fx['single_niter'])}",
                                fx["single_trace"])
    print(msg, flush=True)
    assert ok, msg


def test_ragged_batch_vs_oracle(cg, fx):
    """All 128 instances of the benchmarked batch in one batched call (the
    timed configuration: concurrent size classes), each against the oracle."""
    from curegot_backward_amd.workloads import ragged_batch
    probs = ragged_batch(128)
    Ms = [p[0].cuda() for p in probs]
    As = [p[1].cuda() for p in probs]
    Bs = [p[2].cuda() for p in probs]
    etas = [p[3] for p in probs]
    print("ragged batch B=128", flush=True)
    outs = cg.sinkhorn_splr_bwd_batch(Ms, As, Bs, etas, float(fx["tol"]), int(fx["max_iter"]), trace=True)
    off, toff = fx["batch_offsets"], fx["batch_trace_offsets"]
    splits, worst, bad = [], 0.0, []
    for k in range(128):
        o = outs[k]
        o["_split"] = o["niter"] != int(fx["batch_niter"][k])
        if o["_split"]:
            splits.append(k)
        ok, s, err, msg = judge(
            Ms[k], As[k], Bs[k], etas[k], o, fx["batch_dual"][off[k]:off[k + 1]],
            fx["batch_zstar"][off[k]:off[k + 1]], float(fx["tol"]), k in BATCH_SPLITS,
            f"[{k}] {tuple(Ms[k].shape)} eta {etas[k]} niter {o['niter']}/{int(fx['batch_niter'][k])}",
            fx["batch_trace"][toff[k]:toff[k + 1]])
        print(("" if ok else "FAIL ") + msg, flush=True)
        if not ok:
            bad.append(k)
        worst = max(worst, err)
    print(f"ragged batch: {len(splits)} of 128 instances split {splits}; worst error of the others {worst:.2e}",
          flush=True)
    assert not bad, f"instances failing their bar: {bad}"
    assert set(splits) == BATCH_SPLITS
