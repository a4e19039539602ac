"""Geometric cost: M_ij = scale * sum_k (X_ik - Y_jk)^2 formed on the fly by
every kernel that reads the cost, instead of a stored [n, m] matrix.

The reference's callers build M in torch from point sets
(examples/differentiable_splr_ot.py:55: ((X[:, None] - Y[None]) ** 2).sum(-1) / d;
validate_stage5_workload.py:172: cdist(X, Y)**2 / d) and pass it to splr_ot.
Passing a GeometricCost in place of M skips that matrix: the T passes, the
log-sum-exp passes, the top-K selection, the plan and the adjoint passes all
form M_ij from X and Y.  The kernels sum the d squared differences in k order
with fused multiply-adds and then multiply once by `scale`, so a solve on a
GeometricCost is bitwise the solve on the stored matrix of that same formula.
With d a power of two and scale = 1/d (the default) the multiply equals the
reference's division by d.
"""
from __future__ import annotations

import torch


class GeometricCost:
    """Stand-in for a cost tensor M [n, m]: the points X [n, d], Y [m, d]
    (float64, contiguous, on one device) and the scale.  Accepted wherever
    the native entry points take M (sinkhorn_splr_bwd / _batch, sinkhorn_bcd,
    the backward through the context).  A batch must not mix geometric and
    stored costs."""

    def __init__(self, X: torch.Tensor, Y: torch.Tensor, scale: float | None = None):
        if X.dim() != 2 or Y.dim() != 2:
            raise ValueError("GeometricCost: X and Y must be 2D [n, d] and [m, d] tensors")
        if X.shape[1] != Y.shape[1]:
            raise ValueError(f"GeometricCost: X and Y must have the same number of columns, "
                             f"got {X.shape[1]} and {Y.shape[1]}")
        if X.shape[1] < 1:
            raise ValueError("GeometricCost: d must be >= 1")
        if X.device != Y.device:
            raise ValueError("GeometricCost: X and Y must be on the same device")
        d = int(X.shape[1])
        self.scale = float(1.0 / d if scale is None else scale)
        if not torch.isfinite(torch.tensor(self.scale)):
            raise ValueError("GeometricCost: scale must be finite")
        self.X = X.detach().to(dtype=torch.float64).contiguous()
        self.Y = Y.detach().to(dtype=torch.float64).contiguous()
        self.d = d

    # the parts of the tensor interface the native layer uses
    @property
    def shape(self):
        return torch.Size((self.X.shape[0], self.Y.shape[0]))

    def dim(self) -> int:
        return 2

    @property
    def device(self):
        return self.X.device

    @property
    def dtype(self):
        return torch.float64

    @property
    def is_cuda(self) -> bool:
        return self.X.is_cuda

    @property
    def _version(self):
        return (self.X._version, self.Y._version)

    def clone(self) -> "GeometricCost":
        return GeometricCost(self.X.clone(), self.Y.clone(), self.scale)

    def to(self, device=None, dtype=None) -> "GeometricCost":
        if device is None or torch.device(device) == self.X.device:
            return self
        return GeometricCost(self.X.to(device), self.Y.to(device), self.scale)

    def contiguous(self) -> "GeometricCost":
        return self

    def materialize(self) -> torch.Tensor:
        """The [n, m] matrix of the same formula, built with torch (rounding
        may differ from the kernels' fused multiply-adds in the last bit)."""
        diff = self.X[:, None, :] - self.Y[None, :, :]
        return (diff * diff).sum(-1) * self.scale

    def grad_points(self, gM: torch.Tensor):
        """(dL/dX, dL/dY) from dL/dM: dM_ij/dX_i = 2 scale (X_i - Y_j) and
        dM_ij/dY_j = -2 scale (X_i - Y_j)."""
        g = gM.to(torch.float64)
        rs, cs = g.sum(1), g.sum(0)
        gX = 2.0 * self.scale * (rs[:, None] * self.X - g @ self.Y)
        gY = 2.0 * self.scale * (cs[:, None] * self.Y - g.T.contiguous() @ self.X)
        return gX, gY

    def __repr__(self):
        n, m = self.shape
        return f"GeometricCost(n={n}, m={m}, d={self.d}, scale={self.scale})"


def is_geometric(M) -> bool:
    return isinstance(M, GeometricCost)
