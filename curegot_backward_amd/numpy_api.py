"""NumPy interface: host arrays in, host arrays out, solved on the GPU.

Mirrors the reference's ``numpy`` submodule (src/bindings.cpp:112-122,
src/solvers_numpy.cpp:13-244).  ``sinkhorn_bcd`` and ``sinkhorn_splr`` take
NumPy arrays of any real dtype.  Like pybind's ``py::array_t<double>``
arguments, the arrays are converted to float64.  Each function returns
``{"niter", "plan", "dual"}`` with float64 NumPy arrays.

The arrays are copied to the current HIP device and solved by the native path
that also serves the torch interface; the results are copied back.  Shape
errors raise ``RuntimeError`` with the reference's messages before anything
touches the device.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native


def _host(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _prepare(M, a, b, kwargs):
    # checks of solvers_numpy.cpp:24-68 (sinkhorn_bcd) and 125-169 (sinkhorn_splr)
    M, a, b = _host(M), _host(a), _host(b)
    if M.ndim != 2:
        raise RuntimeError("M must be a 2D array")
    if a.ndim != 1 or b.ndim != 1:
        raise RuntimeError("a and b must be 1D arrays")
    n, m = M.shape
    if a.shape[0] != n or b.shape[0] != m:
        raise RuntimeError("Shape mismatch: M is [n x m], a should be length n, b should be length m")
    kw = dict(kwargs)
    x0 = None
    if kw.get("x0") is not None:
        x0 = _host(kw["x0"])
        if x0.ndim != 1 or x0.shape[0] != n + m:
            raise RuntimeError("x0 must be a 1D array of length n + m")
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device: the numpy interface solves on the GPU")
    device = torch.device("cuda", torch.cuda.current_device())
    dev = lambda x: torch.from_numpy(x).to(device)
    if x0 is not None:
        kw["x0"] = dev(x0)
    return dev(M), dev(a), dev(b), kw


def _to_host(res) -> dict:
    return {"niter": int(res["niter"]), "plan": res["plan"].cpu().numpy(), "dual": res["dual"].cpu().numpy()}


def sinkhorn_bcd(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs) -> dict:
    """Sinkhorn block coordinate descent.  Reference: sinkhorn_bcd,
    src/solvers_numpy.cpp:13-111 (keyword argument: x0).  The dual is not
    gauge-fixed."""
    Md, ad, bd, kw = _prepare(M, a, b, kwargs)
    bcd_kw = {"x0": kw["x0"]} if "x0" in kw else {}
    return _to_host(_native.sinkhorn_bcd(Md, ad, bd, reg, tol, max_iter, verbose, **bcd_kw))


def sinkhorn_splr(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs) -> dict:
    """Sparse-plus-low-rank quasi-Newton solve.  Reference: sinkhorn_splr,
    src/solvers_numpy.cpp:113-244 (keyword arguments: x0, density, shift,
    sparsity_pattern_cycle, candidate_sinkhorn_iter, with the defaults
    min(10/min(n, m), 1), 1e-3, 30 and 3)."""
    Md, ad, bd, kw = _prepare(M, a, b, kwargs)
    keys = ("x0", "density", "shift", "sparsity_pattern_cycle", "candidate_sinkhorn_iter")
    return _to_host(_native.sinkhorn_splr(Md, ad, bd, reg, tol, max_iter, verbose,
                                          **{k: kw[k] for k in keys if k in kw}))
