"""Production backward (implicit gradient) of the SPLR solver.

Mirrors curegot.backward (reference curegot/backward.py:46-144): given the
upstream gradient G = dL/dT and the backward context of a converged forward,
solve the reduced-Hessian system

    Q(T*) lambda = mu,      mu = [(G.T*) 1_m, ((G.T*)^T 1_n)[:m-1]]

by preconditioned CG and assemble dL/dM = (1/eta) T*.(lambda_a 1^T + 1 lambda_b^T - G).

The whole PCG runs in the library's HIP kernels (one pass over M per
iteration, T* recomputed from the context's duals; no stored plan) and the
batch variant solves many instances in lockstep.

Preconditioners: "diag" (default; 1/diag(A_k) of the context matrix, the
reference's native diagonal), "splr" (B_k^{-1} = (A_k + rank-2 BFGS term)^{-1}
with the forward's stored Schur Cholesky factor of A_k, applied on the device
inside the PCG loop) and "none" (plain CG).
"""
from __future__ import annotations

import torch

from . import _native

__all__ = ["solve_backward", "solve_backward_batch", "select_preconditioner"]

_VALID_PRECONDITIONERS = ("diag", "splr", "none")


def select_preconditioner(ctx, preconditioner):
    """Return (inverse-diagonal tensor or None, factor pointer or None) for the
    native PCG.  Reference: select_preconditioner, curegot/backward.py:46-69."""
    if preconditioner == "diag":
        if not bool(ctx.has_diagonal_preconditioner):
            raise RuntimeError("SPLRBackwardContext: diagonal preconditioner not set up")
        return ctx._inv_diag, None
    if preconditioner == "none":
        return None, None
    if preconditioner == "splr":
        if not ctx.has_factor:
            raise RuntimeError("SPLRBackwardContext: preconditioner 'splr' needs the forward's stored factor "
                               "(the forward took no step); use 'diag'")
        return None, ctx.factor_ptr()
    raise ValueError(
        f"invalid preconditioner '{preconditioner}'; "
        f"expected one of: {' | '.join(_VALID_PRECONDITIONERS)}"
    )


def _check(T, G):
    if T is not None:
        if T.dim() != 2:
            raise ValueError("T must be a 2D tensor")
        if G.shape != T.shape:
            raise ValueError("G must have the same shape as T")
    elif G.dim() != 2:
        raise ValueError("G must be a 2D tensor")


def solve_backward_batch(items, *, preconditioner="diag", tol=1e-12, max_iter=2000):
    """items: list of (G, ctx, eta).  Returns one result dict per item."""
    jobs = []
    for G, ctx, eta in items:
        if G.shape != ctx.M.shape:
            raise ValueError("G must have the same shape as T")
        ctx.check_unmodified()
        inv, factor = select_preconditioner(ctx, preconditioner)
        jobs.append((ctx.M, ctx.dual, G, float(eta), inv, factor))
    lams, grads, infos, exts = _native.adjoint_batch(jobs, tol, max_iter)
    out = []
    for lam, grad, info, ext, (G, ctx, eta) in zip(lams, grads, infos, exts, items):
        if info.status == 2:
            raise RuntimeError("PCG breakdown: p^T A p <= 0 (matrix not SPD?)")
        n = ctx.M.shape[0]
        out.append({
            "lambda": lam,
            "grad": grad.to(G.dtype) if G.dtype != torch.float64 else grad,
            # gradients w.r.t. the marginals from the same lambda: the forward
            # solves row sums = a and the first m-1 column sums = b (beta_m = 0,
            # the last column takes up any imbalance), so dL/da = lambda_alpha,
            # dL/db = (lambda_beta, 0)
            "grad_a": lam[:n],
            "grad_b": ext[n:],
            "iterations": int(info.iterations),
            "converged": bool(info.converged),
            "preconditioner": preconditioner,
        })
    return out


def solve_backward(T, G, ctx, eta, *, preconditioner="diag", tol=1e-12, max_iter=2000):
    """Solve the backward pass and assemble the implicit gradient.

    Args mirror the reference (curegot/backward.py:72-81).  T is checked for
    shape only: the plan is recomputed inside the kernels from the context's
    cost matrix and converged duals (it equals the forward's plan).

    Returns dict(lambda [n+m-1], grad [n, m], grad_a [n], grad_b [m],
    iterations, converged, preconditioner); grad_a / grad_b are dL/da and
    dL/db (the reference has no counterpart for these two).
    """
    _check(T, G)
    return solve_backward_batch([(G, ctx, eta)], preconditioner=preconditioner, tol=tol,
                                max_iter=max_iter)[0]
