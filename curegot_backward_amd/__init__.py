"""MI355X-native entropic optimal transport (SPLR) with implicit differentiation.

Drop-in for the ``curegot.torch`` API of benben-cd/curegot_backward
(curegot/__init__.py:23-39): the same names live on the ``torch`` namespace of
this package and at top level::

    import curegot_backward_amd as curegot
    T = curegot.torch.splr_ot(M, a, b, eta=0.05)
    T.pow(2).sum().backward()

Compute runs in hand-written gfx950 HIP kernels in libcuregot_amd.so (C ABI in
include/curegot_amd.h); importing the package fails if the library is missing.
"""
from __future__ import annotations

import sys as _sys
import types as _types

from . import _lib

_lib.load()   # fail loudly if the HIP library is not built

from . import _native, autograd, backward, numpy_api  # noqa: E402
from ._native import (  # noqa: E402
    SPLRBackwardContext,
    sinkhorn_bcd,
    sinkhorn_splr,
    sinkhorn_splr_bwd,
    sinkhorn_splr_bwd_batch,
)
from .autograd import SPLROT, SPLROTFunction, splr_ot, splr_ot_batch, splr_ot_with_info, splr_ot_xy  # noqa: E402
from .cost import GeometricCost  # noqa: E402
from .backward import solve_backward, solve_backward_batch  # noqa: E402

__version__ = "0.1.0"

# Namespace mirroring the reference's native submodule plus the attached API.
torch = _types.ModuleType(__name__ + ".torch")
for _name in ("splr_ot", "splr_ot_with_info", "splr_ot_batch", "SPLROT", "solve_backward",
              "sinkhorn_splr_bwd", "sinkhorn_splr_bwd_batch", "sinkhorn_splr", "sinkhorn_bcd",
              "SPLRBackwardContext", "solve_backward_batch", "splr_ot_xy", "GeometricCost"):
    setattr(torch, _name, globals()[_name])
_sys.modules[torch.__name__] = torch

# Host-array interface, like the reference's numpy submodule
# (src/bindings.cpp:112-122): NumPy in, NumPy out, solved on the GPU.
numpy = _types.ModuleType(__name__ + ".numpy")
numpy.sinkhorn_bcd = numpy_api.sinkhorn_bcd
numpy.sinkhorn_splr = numpy_api.sinkhorn_splr
_sys.modules[numpy.__name__] = numpy

# Test hooks, like the reference's curegot.tests (src/tests.cpp:36-200).
tests = _types.ModuleType(__name__ + ".tests")
tests.test_T_computation_sparsify = _native.T_sparsify
_sys.modules[tests.__name__] = tests

__all__ = ["torch", "numpy", "tests", "splr_ot", "splr_ot_with_info", "splr_ot_batch", "SPLROT",
           "SPLROTFunction", "solve_backward", "solve_backward_batch", "sinkhorn_splr_bwd",
           "sinkhorn_splr_bwd_batch", "sinkhorn_splr", "sinkhorn_bcd", "SPLRBackwardContext", "splr_ot_xy",
           "GeometricCost"]
