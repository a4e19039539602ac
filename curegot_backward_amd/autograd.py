"""Differentiable PyTorch API of the MI355X SPLR entropic-OT solver.

Same public surface and contract as the reference curegot.autograd
(curegot/autograd.py:1-423): ``splr_ot``, ``splr_ot_with_info``,
``splr_ot_batch``, ``SPLROT`` and ``SPLROTFunction``.

Differences by design:
  * the autograd graph keeps M and the converged duals, not the plan; the
    backward recomputes T inside the adjoint kernels;
  * ``splr_ot_batch`` is one native ragged-batch solve (and one batched
    adjoint), not a Python loop; it also accepts lists of differently sized
    problems.  Each slice equals ``splr_ot`` of that slice bitwise.
"""
from __future__ import annotations

import math

import torch
from torch.autograd.function import once_differentiable

from . import _native
from .cost import GeometricCost
from .backward import _VALID_PRECONDITIONERS, solve_backward_batch

__all__ = ["splr_ot", "splr_ot_with_info", "splr_ot_batch", "SPLROT", "SPLROTFunction"]


# ---------------------------------------------------------------------------
# Input contract (reference autograd.py:57-126)
# ---------------------------------------------------------------------------
def _check_inputs(M, a, b, name="splr_ot"):
    for nm, t in (("M", M), ("a", a), ("b", b)):
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name}: {nm} must be a torch.Tensor, got {type(t).__name__}")
        if not t.is_cuda:
            raise RuntimeError(f"{name}: {nm} must be a CUDA tensor ({name} supports CUDA float64 only)")
        if t.dtype != torch.float64:
            raise RuntimeError(
                f"{name}: {nm} must be float64, got {t.dtype}; the native solver is float64-only "
                "and silent conversion would make the gradient dtype misleading")
    if M.dim() != 2:
        raise ValueError(f"{name}: M must be 2D (one OT problem per call; use splr_ot_batch for "
                         f"[B, n, m]), got shape {tuple(M.shape)}")
    n, m = M.shape
    if a.dim() != 1 or a.shape[0] != n:
        raise ValueError(f"{name}: a must be 1D of length {n}, got shape {tuple(a.shape)}")
    if b.dim() != 1 or b.shape[0] != m:
        raise ValueError(f"{name}: b must be 1D of length {m}, got shape {tuple(b.shape)}")
    if not (M.device == a.device == b.device):
        raise ValueError(f"{name}: M/a/b must be on the same device, got {M.device}, {a.device}, {b.device}")


def _validate_scalars(eta, tol, max_iter, density, candidate_sinkhorn_iter, backward_tol,
                      backward_max_iter):
    if not (math.isfinite(eta) and eta > 0):
        raise ValueError(f"splr_ot: eta must be a positive finite scalar, got {eta}")
    if not (math.isfinite(tol) and tol > 0):
        raise ValueError(f"splr_ot: tol must be positive, got {tol}")
    if not int(max_iter) > 0:
        raise ValueError(f"splr_ot: max_iter must be positive, got {max_iter}")
    if density is not None and not (0.0 < float(density) <= 1.0):
        raise ValueError(f"splr_ot: density must be in (0, 1], got {density}")
    if not int(candidate_sinkhorn_iter) >= 0:
        raise ValueError(f"splr_ot: candidate_sinkhorn_iter must be >= 0, got {candidate_sinkhorn_iter}")
    if not (math.isfinite(backward_tol) and backward_tol > 0):
        raise ValueError(f"splr_ot: backward_tol must be positive, got {backward_tol}")
    if not int(backward_max_iter) > 0:
        raise ValueError(f"splr_ot: backward_max_iter must be positive, got {backward_max_iter}")


def _check_precond(p, name="splr_ot"):
    if p not in _VALID_PRECONDITIONERS:
        raise ValueError(f"{name}: invalid backward_preconditioner '{p}'; expected one of: "
                         f"{' | '.join(_VALID_PRECONDITIONERS)}")


def _native_kwargs(density, candidate_sinkhorn_iter):
    kw = {"candidate_sinkhorn_iter": int(candidate_sinkhorn_iter), "context_mode": "adopt"}
    if density is not None:
        kw["density"] = density
    return kw


def _raise_forward(out, max_iter, tol, name="splr_ot"):
    if not bool(out["converged"]):
        raise RuntimeError(f"{name}: forward solver did not converge (niter={int(out['niter'])}, "
                           f"max_iter={max_iter}, tol={tol}); relax tol or increase max_iter")


def _raise_backward(res, cfg):
    if not res["converged"]:
        raise RuntimeError(
            f"splr_ot: backward PCG did not converge (iterations={res['iterations']}, "
            f"max_iter={cfg['backward_max_iter']}, tol={cfg['backward_tol']}, "
            f"preconditioner='{cfg['backward_preconditioner']}')")


# ---------------------------------------------------------------------------
# autograd.Function (reference autograd.py:133-208)
# ---------------------------------------------------------------------------
class SPLROTFunction(torch.autograd.Function):
    """Forward: native SPLR.  Backward: device PCG adjoint with T recomputed
    from (M, duals); the plan is not saved for backward."""

    @staticmethod
    def forward(ctx, M, a, b, eta, tol, max_iter, density, candidate_sinkhorn_iter,
                backward_preconditioner, backward_tol, backward_max_iter):
        out = _native.sinkhorn_splr_bwd(M.detach(), a.detach(), b.detach(), eta, tol, max_iter, 0,
                                        **_native_kwargs(density, candidate_sinkhorn_iter))
        _raise_forward(out, max_iter, tol)
        ctx.backward_context = out["context"]
        ctx.eta = eta
        ctx.cfg = dict(backward_preconditioner=backward_preconditioner, backward_tol=backward_tol,
                       backward_max_iter=backward_max_iter)
        ctx.save_for_backward(M)
        return out["plan"]

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_output):
        ctx.saved_tensors   # version check: M must not have been modified in place
        G = grad_output if grad_output.is_contiguous() else grad_output.contiguous()
        cfg = ctx.cfg
        res = solve_backward_batch([(G, ctx.backward_context, ctx.eta)],
                                   preconditioner=cfg["backward_preconditioner"],
                                   tol=cfg["backward_tol"], max_iter=cfg["backward_max_iter"])[0]
        _raise_backward(res, cfg)
        ga = res["grad_a"] if ctx.needs_input_grad[1] else None
        gb = res["grad_b"] if ctx.needs_input_grad[2] else None
        gM = res["grad"] if ctx.needs_input_grad[0] else None
        return (gM, ga, gb, None, None, None, None, None, None, None, None)


class _SPLROTBatchFunction(torch.autograd.Function):
    """Ragged batch: inputs (cfg, *Ms, *as, *bs) -> tuple of plans."""

    @staticmethod
    def forward(ctx, cfg, *tensors):
        B = len(tensors) // 3
        Ms, As, Bs = tensors[:B], tensors[B:2 * B], tensors[2 * B:]
        outs = _native.sinkhorn_splr_bwd_batch(
            [t.detach() for t in Ms], [t.detach() for t in As], [t.detach() for t in Bs],
            cfg["etas"], cfg["tol"], cfg["max_iter"], 0,
            **_native_kwargs(cfg["density"], cfg["candidate_sinkhorn_iter"]))
        for o in outs:
            _raise_forward(o, cfg["max_iter"], cfg["tol"], "splr_ot_batch")
        ctx.contexts = [o["context"] for o in outs]
        ctx.cfg = cfg
        ctx.nb = B
        ctx.save_for_backward(*Ms)
        return tuple(o["plan"] for o in outs)

    @staticmethod
    @once_differentiable
    def backward(ctx, *grads):
        ctx.saved_tensors
        cfg = ctx.cfg
        items = []
        idx = []
        for i, g in enumerate(grads):
            if g is None:
                continue
            items.append((g.contiguous(), ctx.contexts[i], cfg["etas"][i]))
            idx.append(i)
        B = ctx.nb
        gM, gA, gB = [None] * B, [None] * B, [None] * B
        if items:
            res = solve_backward_batch(items, preconditioner=cfg["backward_preconditioner"],
                                       tol=cfg["backward_tol"], max_iter=cfg["backward_max_iter"])
            need = ctx.needs_input_grad
            for i, r in zip(idx, res):
                _raise_backward(r, cfg)
                gM[i] = r["grad"] if need[1 + i] else None
                gA[i] = r["grad_a"] if need[1 + B + i] else None
                gB[i] = r["grad_b"] if need[1 + 2 * B + i] else None
        return (None, *gM, *gA, *gB)


# ---------------------------------------------------------------------------
# Public API (reference autograd.py:214-423)
# ---------------------------------------------------------------------------
def splr_ot(M, a, b, *, eta, tol=1e-6, max_iter=1000, density=None, candidate_sinkhorn_iter=3,
            backward_preconditioner="diag", backward_tol=1e-12, backward_max_iter=2000,
            marginal_grads=False):
    """Differentiable entropic-OT plan via the SPLR solver (drop-in for
    curegot.torch.splr_ot).  M [n, m] CUDA float64 is differentiable; a, b are
    constants unless marginal_grads=True (an extension: the reference has
    only dL/dM), in which case gradients also flow to a and b:
    dL/da = lambda_alpha and dL/db = (lambda_beta, 0) from the same adjoint
    solve, for the system the forward solves (row sums a, the first m-1
    column sums b; the last column takes up any imbalance).  Raises
    RuntimeError on forward or backward non-convergence, NotImplementedError
    if a or b requires grad without marginal_grads (as the reference does)."""
    if torch.is_autocast_enabled():
        raise RuntimeError("splr_ot: AMP/autocast is unsupported (the solver is float64-only); "
                           "disable autocast for this call")
    _check_inputs(M, a, b)
    _validate_scalars(eta, tol, max_iter, density, candidate_sinkhorn_iter, backward_tol,
                      backward_max_iter)
    if (a.requires_grad or b.requires_grad) and not marginal_grads:
        raise NotImplementedError("splr_ot: gradients w.r.t. a/b are not supported (only dL/dM); "
                                  "detach them or use the lower-level API")
    _check_precond(backward_preconditioner)
    return SPLROTFunction.apply(M, a, b, float(eta), float(tol), int(max_iter), density,
                                int(candidate_sinkhorn_iter), backward_preconditioner,
                                float(backward_tol), int(backward_max_iter))


class _SPLROTXYFunction(torch.autograd.Function):
    """Forward on a geometric cost (M formed from X, Y in the kernels).
    Backward: the same adjoint as splr_ot, dL/dM carried to the points by
    dM_ij/dX_i = 2 scale (X_i - Y_j), dM_ij/dY_j = -2 scale (X_i - Y_j)."""

    @staticmethod
    def forward(ctx, X, Y, a, b, scale, eta, tol, max_iter, density, candidate_sinkhorn_iter,
                backward_preconditioner, backward_tol, backward_max_iter):
        cost = GeometricCost(X, Y, scale)
        out = _native.sinkhorn_splr_bwd(cost, a.detach(), b.detach(), eta, tol, max_iter, 0,
                                        **_native_kwargs(density, candidate_sinkhorn_iter))
        _raise_forward(out, max_iter, tol, "splr_ot_xy")
        ctx.backward_context = out["context"]
        ctx.cost = cost
        ctx.eta = eta
        ctx.cfg = dict(backward_preconditioner=backward_preconditioner, backward_tol=backward_tol,
                       backward_max_iter=backward_max_iter)
        ctx.save_for_backward(X, Y)
        return out["plan"]

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_output):
        ctx.saved_tensors   # version check: X, Y must not have been modified in place
        G = grad_output if grad_output.is_contiguous() else grad_output.contiguous()
        cfg = ctx.cfg
        res = solve_backward_batch([(G, ctx.backward_context, ctx.eta)],
                                   preconditioner=cfg["backward_preconditioner"],
                                   tol=cfg["backward_tol"], max_iter=cfg["backward_max_iter"])[0]
        _raise_backward(res, cfg)
        gX = gY = None
        if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
            gX, gY = ctx.cost.grad_points(res["grad"])
        ga = res["grad_a"] if ctx.needs_input_grad[2] else None
        gb = res["grad_b"] if ctx.needs_input_grad[3] else None
        return (gX if ctx.needs_input_grad[0] else None, gY if ctx.needs_input_grad[1] else None, ga, gb,
                None, None, None, None, None, None, None, None, None)


def splr_ot_xy(X, Y, a, b, *, eta, scale=None, tol=1e-6, max_iter=1000, density=None,
               candidate_sinkhorn_iter=3, backward_preconditioner="diag", backward_tol=1e-12,
               backward_max_iter=2000, marginal_grads=False):
    """splr_ot on the cost M_ij = scale * |X_i - Y_j|^2 (scale = 1/d by
    default: the mean squared distance of the reference's workloads,
    examples/differentiable_splr_ot.py:55) without forming M: every kernel
    computes the entries it needs from the points.  X [n, d] and Y [m, d]
    (CUDA float64) are differentiable; a, b as in splr_ot.  Same plan as
    splr_ot on the stored matrix of the same formula (GeometricCost)."""
    if torch.is_autocast_enabled():
        raise RuntimeError("splr_ot_xy: AMP/autocast is unsupported (the solver is float64-only); "
                           "disable autocast for this call")
    for name, t in (("X", X), ("Y", Y)):
        if not isinstance(t, torch.Tensor) or t.dim() != 2:
            raise ValueError(f"splr_ot_xy: {name} must be a 2D tensor")
        if t.dtype != torch.float64:
            raise RuntimeError(f"splr_ot_xy: {name} must be float64 (got {t.dtype})")
        if not t.is_cuda:
            raise ValueError(f"splr_ot_xy: {name} must be a CUDA tensor")
    if X.shape[1] != Y.shape[1]:
        raise ValueError("splr_ot_xy: X and Y must have the same number of columns")
    n, m = X.shape[0], Y.shape[0]
    if a.dim() != 1 or b.dim() != 1 or a.shape[0] != n or b.shape[0] != m:
        raise ValueError(f"splr_ot_xy: a must be [{n}] and b must be [{m}]")
    _validate_scalars(eta, tol, max_iter, density, candidate_sinkhorn_iter, backward_tol,
                      backward_max_iter)
    if (a.requires_grad or b.requires_grad) and not marginal_grads:
        raise NotImplementedError("splr_ot_xy: gradients w.r.t. a/b need marginal_grads=True")
    _check_precond(backward_preconditioner, "splr_ot_xy")
    return _SPLROTXYFunction.apply(X, Y, a, b, None if scale is None else float(scale), float(eta), float(tol),
                                   int(max_iter), density, int(candidate_sinkhorn_iter), backward_preconditioner,
                                   float(backward_tol), int(backward_max_iter))


def splr_ot_with_info(M, a, b, *, eta, tol=1e-6, max_iter=1000, density=None,
                      candidate_sinkhorn_iter=3):
    """Plan + solver metadata, not differentiable; reports non-convergence via
    "converged" instead of raising (reference autograd.py:301-337)."""
    _check_inputs(M, a, b)
    _validate_scalars(eta, tol, max_iter, density, candidate_sinkhorn_iter, 1.0, 1)
    with torch.no_grad():
        out = _native.sinkhorn_splr_bwd(M, a, b, float(eta), float(tol), int(max_iter), 0,
                                        **_native_kwargs(density, candidate_sinkhorn_iter))
    return {"plan": out["plan"].detach(), "dual": out["dual"].detach(), "niter": int(out["niter"]),
            "converged": bool(out["converged"])}


def splr_ot_batch(M, a, b, *, eta, tol=1e-6, max_iter=1000, density=None,
                  candidate_sinkhorn_iter=3, backward_preconditioner="diag", backward_tol=1e-12,
                  backward_max_iter=2000, marginal_grads=False):
    """Batched splr_ot in one native call.

    M: [B, n, m] tensor (returns a stacked [B, n, m] plan, like the reference,
    autograd.py:340-367), or a list of [n_i, m_i] tensors with lists a, b
    (returns a list of plans).  eta may be a float or one value per instance.
    Gradients flow to every M, and with marginal_grads=True to a and b too
    (as in splr_ot).
    """
    if torch.is_autocast_enabled():
        raise RuntimeError("splr_ot_batch: AMP/autocast is unsupported")
    stacked = isinstance(M, torch.Tensor)
    if stacked:
        if M.dim() != 3:
            raise ValueError(f"splr_ot_batch: M must be [B, n, m], got shape {tuple(M.shape)}")
        B = M.shape[0]
        if a.shape != (B, M.shape[1]):
            raise ValueError(f"splr_ot_batch: a must be [B, n], got shape {tuple(a.shape)}")
        if b.shape != (B, M.shape[2]):
            raise ValueError(f"splr_ot_batch: b must be [B, m], got shape {tuple(b.shape)}")
        Ms, As, Bs = list(M.unbind(0)), list(a.unbind(0)), list(b.unbind(0))
    else:
        Ms, As, Bs = list(M), list(a), list(b)
        B = len(Ms)
        if len(As) != B or len(Bs) != B:
            raise ValueError("splr_ot_batch: M, a, b lists must have the same length")
    if B == 0:
        raise ValueError("splr_ot_batch: empty batch")
    for i in range(B):
        _check_inputs(Ms[i], As[i], Bs[i], "splr_ot_batch")
        if (As[i].requires_grad or Bs[i].requires_grad) and not marginal_grads:
            raise NotImplementedError("splr_ot_batch: gradients w.r.t. a/b are not supported")
    etas = [float(e) for e in eta] if isinstance(eta, (list, tuple)) else [float(eta)] * B
    if len(etas) != B:
        raise ValueError("splr_ot_batch: eta must be a float or have one value per instance")
    for e in etas:
        _validate_scalars(e, tol, max_iter, density, candidate_sinkhorn_iter, backward_tol,
                          backward_max_iter)
    _check_precond(backward_preconditioner, "splr_ot_batch")
    cfg = dict(etas=etas, tol=float(tol), max_iter=int(max_iter), density=density,
               candidate_sinkhorn_iter=int(candidate_sinkhorn_iter),
               backward_preconditioner=backward_preconditioner, backward_tol=float(backward_tol),
               backward_max_iter=int(backward_max_iter))
    plans = _SPLROTBatchFunction.apply(cfg, *Ms, *As, *Bs)
    return torch.stack(plans) if stacked else list(plans)


class SPLROT(torch.nn.Module):
    """nn.Module holding splr_ot's hyperparameters; no parameters or buffers
    (reference autograd.py:370-423)."""

    def __init__(self, eta, *, tol=1e-6, max_iter=1000, density=None, candidate_sinkhorn_iter=3,
                 backward_preconditioner="diag", backward_tol=1e-12, backward_max_iter=2000,
                 marginal_grads=False):
        super().__init__()
        _validate_scalars(float(eta), float(tol), int(max_iter), density,
                          int(candidate_sinkhorn_iter), float(backward_tol), int(backward_max_iter))
        _check_precond(backward_preconditioner, "SPLROT")
        self.eta = float(eta)
        self.tol = float(tol)
        self.max_iter = int(max_iter)
        self.density = density
        self.candidate_sinkhorn_iter = int(candidate_sinkhorn_iter)
        self.backward_preconditioner = backward_preconditioner
        self.backward_tol = float(backward_tol)
        self.backward_max_iter = int(backward_max_iter)
        self.marginal_grads = bool(marginal_grads)

    def forward(self, M, a, b):
        return splr_ot(M, a, b, eta=self.eta, tol=self.tol, max_iter=self.max_iter,
                       density=self.density, candidate_sinkhorn_iter=self.candidate_sinkhorn_iter,
                       backward_preconditioner=self.backward_preconditioner,
                       backward_tol=self.backward_tol, backward_max_iter=self.backward_max_iter,
                       marginal_grads=self.marginal_grads)

    def extra_repr(self):
        return (f"eta={self.eta}, tol={self.tol}, density={self.density}, "
                f"backward_preconditioner='{self.backward_preconditioner}'")
