"""Seeded synthetic OT problems (CPU torch.Generator, float64), used by the
benchmark and the tests.

    readme_problem()       n=64, m=80, X,Y ~ U[0,1)^8, M = mean_k (x-y)^2, uniform
                           marginals, eta=0.05 (reference validate_stage6a_release.py:51-58,
                           README.md:91-103)
    single_problem()       n=m=2048, Gaussian points in R^16, M = |x-y|^2/max, floored
                           Dirichlet(1) marginals, eta=0.01, density=0.01
    ragged_batch()         B=128 instances, n,m log-uniform in [128, 4096], d in {2, 8, 32},
                           Dirichlet(1) marginals floored at 1e-4, eta in {0.01, 0.05}
    adjoint_problem()      n=m=4096, d=64, M = mean squared distance, uniform marginals, eta=0.05
    serving_trace()        Poisson arrivals of mixed-size requests (forward / forward+backward)
    make_problem()         random M ~ U[0,1), marginals U+0.5 normalised
                           (reference stage3b_validate.py:61-69)
"""
from __future__ import annotations

import math

import torch

f64 = torch.float64


def _gen(seed):
    return torch.Generator(device="cpu").manual_seed(int(seed))


def sqdist(X, Y, mean=False):
    D = ((X[:, None, :] - Y[None, :, :]) ** 2)
    return D.mean(-1) if mean else D.sum(-1)


def sqdist_rows(X, Y, max_elems=1 << 26):
    """sqdist(X, Y) formed a block of rows at a time (bounded temporary)."""
    n, m, d = X.shape[0], Y.shape[0], X.shape[1]
    step = max(1, max_elems // max(1, m * d))
    out = torch.empty(n, m, dtype=X.dtype, device=X.device)
    for i in range(0, n, step):
        out[i:i + step] = sqdist(X[i:i + step], Y)
    return out


def readme_problem(seed=0, n=64, m=80, d=8):
    g = _gen(seed)
    X = torch.rand(n, d, generator=g, dtype=f64)
    Y = torch.rand(m, d, generator=g, dtype=f64)
    M = sqdist(X, Y, mean=True)
    a = torch.full((n,), 1.0 / n, dtype=f64)
    b = torch.full((m,), 1.0 / m, dtype=f64)
    return M, a, b, 0.05


def make_problem(n, m, seed):
    g = _gen(seed)
    M = torch.rand(n, m, generator=g, dtype=f64)
    a = torch.rand(n, generator=g, dtype=f64) + 0.5
    b = torch.rand(m, generator=g, dtype=f64) + 0.5
    return M, a / a.sum(), b / b.sum()


def dirichlet1(k, g, floor):
    e = -torch.log(torch.rand(k, generator=g, dtype=f64).clamp_min(1e-300))   # Exp(1)
    w = e / e.sum()
    w = w.clamp_min(floor)
    return w / w.sum()


def dirichlet(k, alpha, g):
    # Gamma(alpha) by Marsaglia-Tsang on the CPU generator
    x = torch._standard_gamma(torch.full((k,), float(alpha), dtype=f64), generator=g)
    return x / x.sum()


def single_problem(seed=0, n=2048, m=2048, d=16):
    g = _gen(seed)
    X = torch.randn(n, d, generator=g, dtype=f64)
    Y = torch.randn(m, d, generator=g, dtype=f64)
    M = sqdist(X, Y)
    M = M / M.max()
    e = -torch.log(torch.rand(n, generator=g, dtype=f64))
    a = (e / e.sum()) * (1 - 1e-3) + 1e-3 / n
    e = -torch.log(torch.rand(m, generator=g, dtype=f64))
    b = (e / e.sum()) * (1 - 1e-3) + 1e-3 / m
    return M, a / a.sum(), b / b.sum(), 0.01


def ragged_sizes(B=128, lo=128, hi=4096, seed=0):
    g = _gen(seed)
    u = torch.rand(B, 2, generator=g, dtype=f64)
    nm = torch.exp(math.log(lo) + u * (math.log(hi) - math.log(lo))).round().long()
    return [(int(x), int(y)) for x, y in nm.tolist()]


def ragged_batch(B=128, seed=0, lo=128, hi=4096):
    """List of (M, a, b, eta) for the ragged batch (CPU tensors)."""
    g = _gen(seed + 1)
    out = []
    for (n, m) in ragged_sizes(B, lo, hi, seed):
        d = (2, 8, 32)[int(torch.randint(0, 3, (1,), generator=g))]
        eta = (0.01, 0.05)[int(torch.randint(0, 2, (1,), generator=g))]
        X = torch.randn(n, d, generator=g, dtype=f64)
        Y = torch.randn(m, d, generator=g, dtype=f64)
        M = sqdist(X, Y)
        M = M / M.max()
        a = dirichlet1(n, g, 1e-4)
        b = dirichlet1(m, g, 1e-4)
        out.append((M, a, b, eta))
    return out


def adjoint_problem(seed=0, n=4096, m=4096, d=64):
    g = _gen(seed)
    X = torch.randn(n, d, generator=g, dtype=f64)
    Y = torch.randn(m, d, generator=g, dtype=f64)
    M = sqdist(X, Y, mean=True)
    a = torch.full((n,), 1.0 / n, dtype=f64)
    b = torch.full((m,), 1.0 / m, dtype=f64)
    R = torch.randn(n, m, generator=g, dtype=f64)
    return M, a, b, 0.05, R


def serving_trace(rate=2000.0, seconds=60.0, seed=0, lo=64, hi=8192, backward_frac=0.3):
    """Arrival times and request specs (sizes, d, eta, alpha, backward flag)."""
    g = _gen(seed)
    t, reqs = 0.0, []
    while True:
        t += float(-math.log(max(float(torch.rand(1, generator=g, dtype=f64)), 1e-300)) / rate)
        if t >= seconds:
            break
        u = torch.rand(2, generator=g, dtype=f64)
        n = int(round(math.exp(math.log(lo) + float(u[0]) * (math.log(hi) - math.log(lo)))))
        m = int(round(math.exp(math.log(lo) + float(u[1]) * (math.log(hi) - math.log(lo)))))
        d = (4, 16, 64)[int(torch.randint(0, 3, (1,), generator=g))]
        eta = (0.01, 0.05, 0.1)[int(torch.randint(0, 3, (1,), generator=g))]
        alpha = 0.5 + 1.5 * float(torch.rand(1, generator=g, dtype=f64))
        bwd = float(torch.rand(1, generator=g, dtype=f64)) < backward_frac
        reqs.append({"t": t, "n": n, "m": m, "d": d, "eta": eta, "alpha": alpha, "backward": bwd,
                     "seed": int(torch.randint(0, 2**31 - 1, (1,), generator=g))})
    return reqs


def serving_problem(req, device=None):
    """(M, a, b, eta) of one serving request; M is formed on `device` (CPU by
    default) from the CPU-drawn points."""
    g = _gen(req["seed"])
    n, m, d = req["n"], req["m"], req["d"]
    X = torch.randn(n, d, generator=g, dtype=f64)
    Y = torch.randn(m, d, generator=g, dtype=f64)
    if device is not None:
        X, Y = X.to(device), Y.to(device)
    M = sqdist_rows(X, Y)
    M = M / M.max()
    a = dirichlet(n, req["alpha"], g).clamp_min(1e-12)
    b = dirichlet(m, req["alpha"], g).clamp_min(1e-12)
    a, b = a / a.sum(), b / b.sum()
    if device is not None:
        a, b = a.to(device), b.to(device)
    return M, a, b, req["eta"]
