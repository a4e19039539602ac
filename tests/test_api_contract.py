"""Drop-in contract of the public API (reference autograd.py:57-126, 274-292;
validate_stage4_autograd.py:477-565; validate_stage5_api.py:158-198)."""
import math

import numpy as np
import pytest
import torch

import curegot_backward_amd as cg
from curegot_backward_amd.workloads import make_problem, readme_problem


def test_api_surface():
    for name in ("splr_ot", "splr_ot_with_info", "SPLROT", "splr_ot_batch", "solve_backward",
                 "sinkhorn_splr_bwd", "SPLRBackwardContext", "sinkhorn_splr", "sinkhorn_bcd"):
        assert callable(getattr(cg.torch, name, None)), name
    assert cg.torch.__name__ == "curegot_backward_amd.torch"
    assert callable(cg.tests.test_T_computation_sparsify)


def test_drop_in_import_name():
    """`import curegot` as the reference's release check does
    (validate_stage6a_release.py:41-45; curegot/__init__.py:23-39): the public,
    advanced and legacy names on curegot.torch, the native namespace's name,
    the numpy/tests submodules and the top-level API, all the same objects as
    this package's."""
    import importlib
    import curegot
    import curegot.autograd
    import curegot.backward
    for name in ("splr_ot", "splr_ot_with_info", "SPLROT", "splr_ot_batch", "solve_backward",
                 "sinkhorn_splr_bwd", "SPLRBackwardContext", "sinkhorn_splr", "sinkhorn_bcd"):
        assert callable(getattr(curegot.torch, name, None)), name
        assert getattr(curegot.torch, name) is getattr(cg.torch, name)
    assert curegot.torch.__name__ == "curegot._internal_torch.torch"
    assert importlib.import_module("curegot._internal_torch").torch is curegot.torch
    assert curegot.numpy.sinkhorn_bcd is cg.numpy.sinkhorn_bcd
    assert callable(curegot.tests.test_T_computation_sparsify)
    assert curegot.splr_ot is curegot.torch.splr_ot is curegot.autograd.splr_ot
    assert curegot.solve_backward is curegot.backward.solve_backward
    assert curegot.__version__ == cg.__version__


def test_cpu_input_rejected():
    M, a, b = make_problem(5, 4, 0)
    with pytest.raises(RuntimeError, match="CUDA"):
        cg.torch.splr_ot(M, a, b, eta=0.05)


@pytest.mark.parametrize("kw,exc", [({"eta": 0.0}, ValueError), ({"eta": math.inf}, ValueError),
                                    ({"eta": 0.1, "tol": -1.0}, ValueError), ({"eta": 0.1, "max_iter": 0}, ValueError),
                                    ({"eta": 0.1, "density": 1.5}, ValueError),
                                    ({"eta": 0.1, "candidate_sinkhorn_iter": -1}, ValueError),
                                    ({"eta": 0.1, "backward_tol": 0.0}, ValueError)])
def test_scalar_validation(kw, exc):
    with pytest.raises(exc):
        cg.torch.SPLROT(**kw)


def test_numpy_interface_shape_errors():
    """The numpy submodule (reference src/bindings.cpp:112-122) raises the
    reference's messages (src/solvers_numpy.cpp:28-66) before any device work."""
    assert cg.numpy.__name__ == "curegot_backward_amd.numpy"
    M, a, b = [x.numpy() for x in make_problem(6, 5, 0)]
    for f in (cg.numpy.sinkhorn_bcd, cg.numpy.sinkhorn_splr):
        with pytest.raises(RuntimeError, match="M must be a 2D array"):
            f(M.ravel(), a, b, 0.1)
        with pytest.raises(RuntimeError, match="a and b must be 1D arrays"):
            f(M, a[:, None], b, 0.1)
        with pytest.raises(RuntimeError, match="Shape mismatch"):
            f(M, a[:-1], b, 0.1)
        with pytest.raises(RuntimeError, match=r"x0 must be a 1D array of length n \+ m"):
            f(M, a, b, 0.1, x0=np.zeros(3))


def test_bad_preconditioner_name():
    with pytest.raises(ValueError):
        cg.torch.SPLROT(0.1, backward_preconditioner="chol")


@pytest.mark.gpu
class TestGPUContract:
    def test_readme_forward_backward(self):
        M, a, b, eta = readme_problem()
        M = M.cuda().requires_grad_(True)
        T = cg.torch.splr_ot(M, a.cuda(), b.cuda(), eta=eta, tol=1e-8)
        assert T.shape == M.shape and T.dtype == torch.float64 and T.device == M.device
        assert torch.allclose(T.sum(1), a.cuda(), atol=1e-6, rtol=0)
        assert torch.allclose(T.sum(0), b.cuda(), atol=1e-6, rtol=0)
        (T ** 2).sum().backward()
        assert M.grad is not None and bool(torch.isfinite(M.grad).all()) and float(M.grad.norm()) > 0

    def test_errors(self):
        M, a, b = [x.cuda() for x in make_problem(6, 5, 0)]
        with pytest.raises(RuntimeError, match="float64"):
            cg.torch.splr_ot(M.float(), a, b, eta=0.1)
        with pytest.raises(ValueError):
            cg.torch.splr_ot(M, a[:-1], b, eta=0.1)
        with pytest.raises(ValueError):
            cg.torch.splr_ot(M[None], a, b, eta=0.1)
        with pytest.raises(NotImplementedError):
            cg.torch.splr_ot(M, a.clone().requires_grad_(True), b, eta=0.1)
        with pytest.raises(ValueError):
            cg.torch.splr_ot(M, a, b, eta=0.1, backward_preconditioner="bogus")
        with torch.autocast("cuda"):
            with pytest.raises(RuntimeError, match="autocast"):
                cg.torch.splr_ot(M, a, b, eta=0.1)

    def test_forward_nonconvergence_raises(self):
        M, a, b = [x.cuda() for x in make_problem(20, 24, 9)]
        with pytest.raises(RuntimeError, match="did not converge"):
            cg.torch.splr_ot(M, a, b, eta=0.05, tol=1e-14, max_iter=1)
        info = cg.torch.splr_ot_with_info(M, a, b, eta=0.05, tol=1e-14, max_iter=1)
        assert info["niter"] == 1 and not info["converged"]

    def test_backward_nonconvergence_raises(self):
        M, a, b = [x.cuda() for x in make_problem(20, 24, 9)]
        M.requires_grad_(True)
        T = cg.torch.splr_ot(M, a, b, eta=0.05, tol=1e-8, backward_max_iter=1)
        with pytest.raises(RuntimeError, match="backward PCG did not converge"):
            (T ** 3).sum().backward()

    def test_zero_upstream_gradient(self):
        M, a, b = [x.cuda() for x in make_problem(12, 10, 2)]
        out = cg.torch.sinkhorn_splr_bwd(M, a, b, 0.05, 1e-8, 1000)
        res = cg.torch.solve_backward(out["plan"], torch.zeros_like(M), out["context"], 0.05)
        assert res["iterations"] == 0 and res["converged"]
        assert torch.count_nonzero(res["lambda"]) == 0 and torch.count_nonzero(res["grad"]) == 0

    def test_manual_path_equals_autograd(self):
        # reference validate_stage4_autograd.py:107-130 (<= 1e-10)
        M, a, b = [x.cuda() for x in make_problem(30, 25, 4)]
        Mg = M.clone().requires_grad_(True)
        T = cg.torch.splr_ot(Mg, a, b, eta=0.05, tol=1e-8)
        R = torch.randn_like(T)
        ((T ** 2).sum() + (T * R).sum()).backward()
        out = cg.torch.sinkhorn_splr_bwd(M, a, b, 0.05, 1e-8, 1000, 0, candidate_sinkhorn_iter=3)
        G = 2 * out["plan"] + R
        res = cg.torch.solve_backward(out["plan"], G, out["context"], 0.05)
        assert (res["grad"] - Mg.grad).abs().max() <= 1e-10 * Mg.grad.abs().max()

    def test_retain_graph_accumulates(self):
        M, a, b = [x.cuda() for x in make_problem(10, 12, 5)]
        M.requires_grad_(True)
        T = cg.torch.splr_ot(M, a, b, eta=0.05, tol=1e-8)
        L = (T ** 2).sum()
        L.backward(retain_graph=True)
        g1 = M.grad.clone()
        L.backward()
        assert torch.allclose(M.grad, 2 * g1, rtol=1e-12, atol=0)

    def test_module_train_eval_identical(self):
        M, a, b = [x.cuda() for x in make_problem(15, 15, 6)]
        layer = cg.torch.SPLROT(eta=0.05, tol=1e-8)
        T1 = layer(M, a, b)
        layer.eval()
        T2 = layer(M, a, b)
        assert torch.equal(T1, T2) and len(layer.state_dict()) == 0

    def test_finite_differences(self):
        # central FD (h = 1e-6) of L = sum(T^2) + <T, R> on sampled entries.  Every plan
        # is the SPLR solution polished by BCD to a 1e-14 marginal error; the adjoint is
        # taken at the polished duals through a warm-started context (niter == 0).
        n, m, eta = 12, 10, 0.1
        M, a, b = [x.cuda() for x in make_problem(n, m, 21)]
        R = torch.randn(n, m, dtype=torch.float64, generator=torch.Generator().manual_seed(0)).cuda()

        def polished(Mx):
            x0 = cg.torch.sinkhorn_splr_bwd(Mx, a, b, eta, 1e-8, 2000)["dual"]
            return cg.torch.sinkhorn_bcd(Mx, a, b, eta, 1e-14, 100000, x0=x0)

        loss = lambda T: float((T ** 2).sum() + (T * R).sum())
        p0 = polished(M)
        dual = p0["dual"] - p0["dual"][-1] * torch.cat([-torch.ones(n), torch.ones(m)]).cuda()
        out = cg.torch.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 10, x0=dual)
        assert out["niter"] == 0 and out["context"].is_fallback
        G = 2 * out["plan"] + R
        g = cg.torch.solve_backward(out["plan"], G, out["context"], eta, tol=1e-14)["grad"]
        h = 1e-6
        for (i, j) in [(0, 0), (3, 7), (11, 9), (6, 2), (5, 5)]:
            Mp, Mm = M.clone(), M.clone()
            Mp[i, j] += h
            Mm[i, j] -= h
            fd = (loss(polished(Mp)["plan"]) - loss(polished(Mm)["plan"])) / (2 * h)
            gij = float(g[i, j])
            print("fd", i, j, fd, gij, flush=True)
            assert abs(fd - gij) <= 1e-6 * max(abs(gij), 1e-2)
