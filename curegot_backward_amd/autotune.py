"""Tile-shape autotuner for the M passes.

Every kernel that streams M (the T pass, the row and column log-sum-exp
passes, the top-K selection, the adjoint's Q v) cuts an instance into row
blocks of 32 rows and strips of 64c columns, c = 1, 2 or 4 columns per lane.
The default rule takes the smallest c with 64c >= m (at most 4): one strip per
row block up to m = 256, then strips of 256 columns.  A smaller c gives more,
shorter workgroups (more parallelism for small instances, more partials to
merge); a larger one fewer, longer ones.

`tune` times, for each band of m, a short batched SPLR solve (a fixed number
of iterations, tolerance unreachable) of synthetic instances of that band
under each admissible c, best of `reps` after a warm-up, and keeps a c other
than the default only where it is faster by more than `margin`.  The table it
returns is applied with `apply` (process-wide, `cgot_set_tile_table`) and can
be kept as JSON (`save` / `load`); setting the environment variable
CUREGOT_AMD_TILE_TABLE to such a file applies it when the library loads.

c fixes the summation order of the row sums: results are bitwise reproducible
under one table, and a batch and the single solves of its instances agree
bitwise under the same table; two tables give results that differ in the last
bits (within the oracle tolerance) and, rarely, in an iteration count.
"""
from __future__ import annotations

import json
import math
import time

import torch

from . import _lib

# Upper ends of the m bands (each band is (previous, this]).
DEFAULT_BANDS = (128, 256, 512, 1024, 2048, 4096, 8192)


def default_cols(m: int) -> int:
    """The default rule: the smallest power of two c with 64c >= m, at most 4."""
    c = 1
    while 64 * c < m and c < 4:
        c *= 2
    return c


def candidates(m_hi: int):
    """Admissible c for a band ending at m_hi: every c up to the default one
    (a wider strip than 64c >= m only leaves lanes idle)."""
    return [c for c in (1, 2, 4) if c <= default_cols(m_hi)]


def _band_batch(lo: int, hi: int, cells: int, max_batch: int, seed: int, device):
    """Synthetic instances with m in (lo, hi]: n = m, uniform marginals on a
    squared-distance cost, as many as fit `cells` entries of M (1..max_batch)."""
    g = torch.Generator().manual_seed(seed)
    count = max(1, min(max_batch, cells // (hi * hi)))
    probs = []
    for k in range(count):
        m = lo + 1 + int((hi - lo - 1) * (k + 1) / count)
        n = m
        X = torch.randn(n, 8, generator=g, dtype=torch.float64)
        Y = torch.randn(m, 8, generator=g, dtype=torch.float64)
        M = torch.cdist(X, Y) ** 2
        M = (M / M.max()).to(device)
        a = torch.full((n,), 1.0 / n, dtype=torch.float64, device=device)
        b = torch.full((m,), 1.0 / m, dtype=torch.float64, device=device)
        probs.append((M, a, b))
    return probs


def _time_batch(probs, reg, iters, reps):
    from ._native import sinkhorn_splr_bwd_batch
    Ms = [p[0] for p in probs]
    As = [p[1] for p in probs]
    Bs = [p[2] for p in probs]
    best = math.inf
    for r in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sinkhorn_splr_bwd_batch(Ms, As, Bs, reg, tol=1e-300, max_iter=iters)
        torch.cuda.synchronize()
        if r > 0:                      # the first run warms streams and caches
            best = min(best, time.perf_counter() - t0)
    return best


def tune(device=None, bands=DEFAULT_BANDS, iters=4, reps=3, margin=0.03, cells=1 << 25, max_batch=64,
         reg=0.05, seed=0, log=None):
    """Time each admissible c per band and return
    {"table": [(m_max, c), ...], "bands": [...timings...], "device": name}.
    The table lists every band (the default c where no other wins); the
    process's table is restored afterwards."""
    if not torch.cuda.is_available():
        raise RuntimeError("autotune: needs a GPU")
    device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    prev = _lib.get_tile_table()
    out_bands, table = [], []
    lo = 64
    try:
        for hi in bands:
            if hi <= lo:
                raise ValueError("bands must be ascending and above 64")
            cands = candidates(hi)
            probs = _band_batch(lo, hi, cells, max_batch, seed + hi, device)
            times = {}
            for c in cands:
                # this band's candidate; every other m keeps the rows chosen so far
                _lib.set_tile_table(table + [(hi, c)])
                times[c] = _time_batch(probs, reg, iters, reps)
            c0 = default_cols(hi)
            best = min(times, key=times.get)
            pick = best if times[best] < (1.0 - margin) * times[c0] else c0
            table.append((hi, pick))
            rec = {"m_lo": lo + 1, "m_hi": hi, "instances": len(probs),
                   "ms": {str(c): 1000 * t for c, t in times.items()}, "default": c0, "pick": pick}
            out_bands.append(rec)
            if log:
                log(f"autotune m in ({lo}, {hi}] x{len(probs)}: "
                    + ", ".join(f"c={c} {1000 * t:.2f} ms" for c, t in times.items()) + f" -> c={pick}")
            lo = hi
            del probs
    finally:
        _lib.set_tile_table(prev)
    return {"table": table, "bands": out_bands, "device": torch.cuda.get_device_name(device),
            "iters": iters, "reps": reps, "margin": margin}


def apply(result_or_table):
    """Make a tuned table current (a `tune` result or a [(m_max, c), ...] list)."""
    table = result_or_table["table"] if isinstance(result_or_table, dict) else result_or_table
    _lib.set_tile_table(table)


def reset():
    """Back to the default rule for every m."""
    _lib.set_tile_table([])


def save(result, path):
    with open(path, "w") as f:
        json.dump(result, f, indent=1)


def load(path, apply_now=True):
    """Read a saved `tune` result (JSON) and, by default, apply its table."""
    with open(path) as f:
        result = json.load(f)
    result["table"] = [(int(mx), int(c)) for mx, c in result["table"]]
    if apply_now:
        apply(result)
    return result
