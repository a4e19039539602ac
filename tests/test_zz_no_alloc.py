"""cgot_splr_precond_apply_noalloc -- the SPLR preconditioner application with
its work buffers in the library's fixed per-GPU device array -- against
cgot_splr_precond_apply (caller-provided workspace), both through the C ABI:
bitwise equal on consecutive calls, each later call on a smaller problem,
for factors of the dense Schur path and of the sparse plan; a problem too
large for the array is rejected with -2 and a message naming the sizes."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    from curegot_backward_amd import _lib
    return _lib


def context(lib, n, m, seed, eta, mode):
    import curegot_backward_amd as cg
    from curegot_backward_amd.workloads import make_problem
    M, a, b = [x.cuda() for x in make_problem(n, m, seed)]
    prev = lib.set_sparse_solve(mode)
    try:
        out = cg.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 500)
    finally:
        lib.set_sparse_solve(prev)
    assert out["converged"]
    return out["context"]


def apply_workspace(lib, ctx, r):
    L = lib.load()
    f, n, m = ctx.factor_ptr(), ctx._n, ctx._m
    nbytes = L.cgot_splr_precond_workspace_size(f, n, m)
    assert nbytes > 0
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    z = torch.empty_like(r)
    rc = L.cgot_splr_precond_apply(f, n, m, r.data_ptr(), z.data_ptr(), ws.data_ptr(), nbytes,
                                   torch.cuda.current_stream().cuda_stream)
    lib.check(rc, "cgot_splr_precond_apply")
    return z


def apply_noalloc(lib, ctx, r):
    L = lib.load()
    z = torch.full_like(r, float("nan"))
    rc = L.cgot_splr_precond_apply_noalloc(ctx.factor_ptr(), ctx._n, ctx._m, r.data_ptr(), z.data_ptr(),
                                           torch.cuda.current_stream().cuda_stream)
    lib.check(rc, "cgot_splr_precond_apply_noalloc")
    return z


def test_noalloc_apply_equals_workspace_apply(lib):
    # (n, m, seed, eta, factorization); each call after the first is smaller
    cases = [(420, 380, 1, 0.05, "dense"), (300, 260, 2, 0.05, "sparse"), (120, 90, 3, 0.1, "dense"),
             (70, 130, 4, 0.05, "sparse")]
    ctxs = [context(lib, n, m, seed, eta, mode) for n, m, seed, eta, mode in cases]
    for (n, m, seed, eta, mode), ctx in zip(cases, ctxs):
        print(f"case n={n} m={m} {mode} factor", flush=True)
        r = torch.from_numpy(np.random.default_rng(seed).normal(size=n + m - 1)).cuda()
        z_ws = apply_workspace(lib, ctx, r)
        z_na = apply_noalloc(lib, ctx, r)
        torch.cuda.synchronize()
        assert torch.isfinite(z_na).all()
        assert torch.equal(z_na, z_ws), float((z_na - z_ws).abs().max())
        # the context's own method (same kernels, caller workspace) agrees too
        assert torch.equal(z_na, ctx.apply_preconditioner(r))
    # the first (largest) problem again after the smaller ones: nothing left
    # over in the shared array changes its result
    r = torch.from_numpy(np.random.default_rng(1).normal(size=cases[0][0] + cases[0][1] - 1)).cuda()
    assert torch.equal(apply_noalloc(lib, ctxs[0], r), apply_workspace(lib, ctxs[0], r))


def test_noalloc_rejects_problem_beyond_scratch(lib):
    ctx = context(lib, 60, 50, 7, 0.1, "dense")
    L = lib.load()
    r = torch.zeros(8, dtype=torch.float64, device="cuda")
    n = m = 40000
    rc = L.cgot_splr_precond_apply_noalloc(ctx.factor_ptr(), n, m, r.data_ptr(), r.data_ptr(),
                                           torch.cuda.current_stream().cuda_stream)
    assert rc == -2
    msg = L.cgot_last_error().decode()
    print(msg, flush=True)
    assert "too large for the fixed scratch" in msg and "n=40000" in msg
    # a valid call still works after the rejection
    r = torch.from_numpy(np.random.default_rng(0).normal(size=109)).cuda()
    assert torch.equal(apply_noalloc(lib, ctx, r), apply_workspace(lib, ctx, r))
