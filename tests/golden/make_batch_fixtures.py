"""Writes tests/golden/batch_oracle.npz: the oracle's SPLR results on the
single-instance problem (n = m = 2048, eta = 0.01, density 0.01, global
top-K) and on every instance of the ragged batch (B = 128, n, m log-uniform in
[128, 4096], default density), all at tol 1e-8, max_iter 1000, seed 0 of
curegot_backward_amd/workloads.py.

Per instance it stores the oracle's niter, converged flag, final ||g||, f,
its dual (beta[m-1] = 0) and its per-iteration trace (f, ||g||, step,
low_rank, Sinkhorn candidate taken, K), and the optimum z* of the same
problem, found by Newton steps with the exact Hessian Q/eta from the oracle's
dual (Schur complement onto the smaller side, scipy Cholesky) until ||g||
stops falling.  The GPU tests regenerate M, a, b from the seed and compare
per instance: duals and plan to 1e-10 where the iteration counts agree, and
the distance to z* where the two runs took different line-search branches.

    python tests/golden/make_batch_fixtures.py      (about 5 minutes on 8 cores)
"""
import os
import sys
import time

import numpy as np
import scipy.linalg as sla

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from curegot_backward_amd.workloads import ragged_batch, single_problem  # noqa: E402

TOL = 1e-8
MAX_ITER = 1000
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batch_oracle.npz")


def grad_and_plan(M, a, b, eta, dual):
    n, m = M.shape
    T = np.exp((dual[:n, None] + dual[None, n:] - M) / eta)
    g = np.concatenate([T.sum(1) - a, T.sum(0)[:-1] - b[:-1]])
    return g, T


def newton_polish(M, a, b, eta, dual, steps=4):
    """Newton on the reduced dual z = (alpha, beta[:m-1]) with H = Q(T)/eta,
    Q = [[diag(T1), W], [W^T, diag(T^T 1)[:m-1]]], W = T[:, :m-1]; the system is
    reduced onto its smaller side.  Returns the iterate of smallest ||g||."""
    n, m = M.shape
    best = dual.copy()
    g, T = grad_and_plan(M, a, b, eta, best)
    gbest = np.linalg.norm(g)
    z = best.copy()
    for _ in range(steps):
        W = T[:, :m - 1]
        Da = T.sum(1)
        Db = T.sum(0)[:m - 1]
        ga, gb = g[:n], g[n:]
        if m - 1 <= n:
            S = np.diag(Db) - (W.T / Da) @ W
            rb = gb - W.T @ (ga / Da)
            xb = sla.cho_solve(sla.cho_factor(S, lower=True), rb)
            xa = (ga - W @ xb) / Da
        else:
            S = np.diag(Da) - (W / Db) @ W.T
            ra = ga - W @ (gb / Db)
            xa = sla.cho_solve(sla.cho_factor(S, lower=True), ra)
            xb = (gb - W.T @ xa) / Db
        step = eta * np.concatenate([xa, xb, [0.0]])
        z = z - step
        g, T = grad_and_plan(M, a, b, eta, z)
        gn = np.linalg.norm(g)
        if not gn < gbest:
            break
        best, gbest = z.copy(), gn
    return best, gbest


def run(M, a, b, eta, density=None):
    r = oracle.splr_bwd(M, a, b, eta, tol=TOL, max_iter=MAX_ITER, density=density, trace=True)
    zs, gs = newton_polish(M, a, b, eta, r["dual"])
    return r, zs, gs


def main():
    t0 = time.time()
    out = {}
    M, a, b, eta = single_problem()
    r, zs, gs = run(M.numpy(), a.numpy(), b.numpy(), eta, density=0.01)
    out.update(single_niter=r["niter"], single_converged=r["converged"], single_gnorm=r["gnorm"],
               single_f=r["f"], single_dual=r["dual"], single_trace=r["trace"], single_zstar=zs,
               single_zstar_gnorm=gs)
    print(f"single: niter {r['niter']} |g| {r['gnorm']:.2e} |g*| {gs:.2e} ({time.time() - t0:.0f} s)", flush=True)
    probs = ragged_batch(128)
    niter, conv, gnorm, f, zgn, offs, traces, toffs = [], [], [], [], [], [0], [], [0]
    duals, zstars = [], []
    for k, (M, a, b, eta) in enumerate(probs):
        r, zs, gs = run(M.numpy(), a.numpy(), b.numpy(), eta)
        niter.append(r["niter"]); conv.append(r["converged"]); gnorm.append(r["gnorm"]); f.append(r["f"])
        zgn.append(gs)
        duals.append(r["dual"]); zstars.append(zs)
        offs.append(offs[-1] + len(r["dual"]))
        traces.append(r["trace"]); toffs.append(toffs[-1] + len(r["trace"]))
        print(f"batch {k:3d} {tuple(M.shape)} eta {eta}: niter {r['niter']} |g| {r['gnorm']:.2e} "
              f"|g*| {gs:.2e} ({time.time() - t0:.0f} s)", flush=True)
    out.update(batch_niter=np.array(niter), batch_converged=np.array(conv), batch_gnorm=np.array(gnorm),
               batch_f=np.array(f), batch_zstar_gnorm=np.array(zgn), batch_offsets=np.array(offs),
               batch_dual=np.concatenate(duals), batch_zstar=np.concatenate(zstars),
               batch_trace=np.concatenate(traces), batch_trace_offsets=np.array(toffs),
               tol=TOL, max_iter=MAX_ITER)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT) / 1e6:.2f} MB) in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
