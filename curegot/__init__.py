"""Drop-in import name for code written against benben-cd/curegot_backward.

``import curegot`` gives the module layout of the reference package
(curegot/__init__.py:1-39; src/bindings.cpp:83-130), backed entirely by the
MI355X library in ``curegot_backward_amd``:

    curegot.torch        the native torch submodule, named
                         ``curegot._internal_torch.torch`` as in the reference
                         (validate_stage6a_release.py:41-45), carrying
                         sinkhorn_bcd, sinkhorn_splr, sinkhorn_splr_bwd,
                         SPLRBackwardContext and the attached differentiable API
                         (splr_ot, splr_ot_with_info, splr_ot_batch, SPLROT,
                         solve_backward)
    curegot.numpy        host-array interface (sinkhorn_bcd, sinkhorn_splr)
    curegot.tests        test hooks (test_T_computation_sparsify)
    curegot.autograd, curegot.backward
                         the Python API modules
    curegot.splr_ot, ... the differentiable API at top level

Nothing here computes: every name is the object curegot_backward_amd
defines, so importing this package fails exactly when that one does (the HIP
library missing).
"""
from __future__ import annotations

import sys as _sys
import types as _types

import curegot_backward_amd as _impl
from curegot_backward_amd import autograd, backward  # noqa: F401
from curegot_backward_amd.autograd import SPLROT, splr_ot, splr_ot_batch, splr_ot_with_info  # noqa: F401
from curegot_backward_amd.backward import solve_backward  # noqa: F401

__version__ = _impl.__version__


def _submodule(name, doc, names, src):
    mod = _types.ModuleType(name, doc)
    for n in names:
        setattr(mod, n, getattr(src, n))
    _sys.modules[name] = mod
    return mod


# curegot._internal_torch with its submodule torch (the pybind layout of the
# reference's torch build), and curegot._internal_numpy with numpy and tests
_internal_torch = _types.ModuleType(__name__ + "._internal_torch", "PyTorch interface (MI355X HIP library).")
_sys.modules[_internal_torch.__name__] = _internal_torch
torch = _submodule(_internal_torch.__name__ + ".torch", "PyTorch interface.",
                   ("sinkhorn_bcd", "sinkhorn_splr", "sinkhorn_splr_bwd", "SPLRBackwardContext",
                    "sinkhorn_splr_bwd_batch", "solve_backward_batch",
                    "splr_ot", "splr_ot_with_info", "splr_ot_batch", "SPLROT", "solve_backward"),
                   _impl.torch)
_internal_torch.torch = torch

_internal_numpy = _types.ModuleType(__name__ + "._internal_numpy", "NumPy interface (MI355X HIP library).")
_sys.modules[_internal_numpy.__name__] = _internal_numpy
numpy = _submodule(_internal_numpy.__name__ + ".numpy", "Numpy interface.",
                   ("sinkhorn_bcd", "sinkhorn_splr"), _impl.numpy)
tests = _submodule(_internal_numpy.__name__ + ".tests", "Test functions.",
                   ("test_T_computation_sparsify",), _impl.tests)
_internal_numpy.numpy = numpy
_internal_numpy.tests = tests

# `import curegot.autograd` / `import curegot.backward` resolve to the same modules
_sys.modules[__name__ + ".autograd"] = autograd
_sys.modules[__name__ + ".backward"] = backward

__all__ = ["torch", "numpy", "tests", "autograd", "backward", "splr_ot", "splr_ot_with_info",
           "splr_ot_batch", "SPLROT", "solve_backward"]
