"""Native-layer mirror: the functions the reference exports from its pybind
module ``curegot._internal_torch.torch`` (src/bindings.cpp:90-108,
src/solvers_torch.cpp), implemented over the C ABI of libcuregot_amd.so.

Tensors are plumbing: torch provides device memory (inputs, outputs and the
solver workspace, taken from torch's caching allocator) and the current HIP
stream; all compute runs in the library's HIP kernels.
"""
from __future__ import annotations

import ctypes
import math
import time
from typing import Sequence

import torch

from . import _lib
from .cost import GeometricCost

__all__ = [
    "sinkhorn_splr_bwd",
    "sinkhorn_splr_bwd_batch",
    "sinkhorn_splr",
    "sinkhorn_bcd",
    "SPLRBackwardContext",
    "SplrSession",
]


def _stream(device) -> int:
    return int(torch.cuda.current_stream(device).cuda_stream)


def _to_dev(t: torch.Tensor, device) -> torch.Tensor:
    if isinstance(t, GeometricCost):
        return t.to(device)
    return t.to(device=device, dtype=torch.float64).contiguous()


def _target_device(M: torch.Tensor):
    # reference: solvers_torch.cpp:51-65 (keep a CUDA input's device, else the default device)
    if M.is_cuda:
        return M.device
    return torch.device("cuda", torch.cuda.current_device())


def _workspace(nbytes: int, device) -> torch.Tensor:
    if nbytes == 0:
        raise RuntimeError("curegot: " + _lib.load().cgot_last_error().decode())
    return torch.empty(int(nbytes), dtype=torch.uint8, device=device)


class SPLRBackwardContext:
    """Backward context of one converged SPLR solve.

    Mirrors the reference class (src/backward_context.h:83-376;
    binding src/solvers_torch.cpp:650-704).  It keeps the converged duals and
    a view of the solver workspace holding the matrix A_k of the last completed
    iteration (its diagonal, the secant pair s, y and u = A_k^{-1} y); the plan
    is not stored: the adjoint recomputes T from M and the duals.
    """

    def __init__(self, M, dual, reg, info: _lib.SplrInfo, workspace: torch.Tensor, snapshot: bool = False):
        # The adjoint recomputes T* from M and the duals, so both must stay as
        # they were at the solve: "snapshot" keeps a private copy of M, and an
        # in-place change of the tensors the context holds is refused by
        # check_unmodified() (the reference keeps the plan T itself instead).
        self._M = M.clone() if snapshot else M
        self._dual = dual
        self._versions = (self._M._version, dual._version)
        self._ws = workspace
        self._reg = float(reg)
        self._n = int(info.n)
        self._m = int(info.m)
        self._low_rank = bool(info.low_rank)
        self._fallback = bool(info.fallback)
        self._ys = float(info.ys)
        N = self._n + self._m - 1
        self._inv_diag = self._view(info.inv_diag, N)
        self._s = self._view(info.s, N)
        self._y = self._view(info.y, N)
        self._u = self._view(info.u, N)
        self._factor = int(info.factor) if info.factor else None

    def _view(self, ptr, count) -> torch.Tensor:
        off = int(ptr) - self._ws.data_ptr()
        if off < 0 or off + 8 * count > self._ws.numel():
            raise RuntimeError("SPLRBackwardContext: context pointer outside the workspace")
        return self._ws[off:off + 8 * count].view(torch.float64)

    # ---- metadata (reference solvers_torch.cpp:650-704) ----
    @property
    def reg(self) -> float:
        return self._reg

    @property
    def dim(self) -> int:
        return self._n + self._m - 1

    @property
    def low_rank_enabled(self) -> bool:
        return self._low_rank

    @property
    def is_fallback(self) -> bool:
        return self._fallback

    @property
    def adopted_forward_state(self) -> bool:
        return not self._fallback

    @property
    def ys(self) -> float:
        return self._ys

    @property
    def has_diagonal_preconditioner(self) -> bool:
        return True

    @property
    def M(self) -> torch.Tensor:
        return self._M

    def check_unmodified(self):
        """Raise if the cost matrix or the duals this context recomputes the
        plan from were modified in place after the forward solve."""
        if (self._M._version, self._dual._version) != self._versions:
            raise RuntimeError(
                "SPLRBackwardContext: the cost matrix M or the dual returned by the forward solve was "
                "modified in place after the solve; the backward recomputes the plan from them "
                "(solve again, or use context_mode='snapshot' to keep a private copy of M)")

    @property
    def dual(self) -> torch.Tensor:
        return self._dual

    # ---- validation copies ----
    def inv_diag(self) -> torch.Tensor:
        return self._inv_diag.clone()

    def secant_s(self) -> torch.Tensor:
        """Copy of s (reference solvers_torch.cpp:611-621: unavailable without
        the rank-2 term)."""
        if not self._low_rank:
            raise RuntimeError("SPLRBackwardContext: low-rank correction is disabled, s is unavailable")
        return self._s.clone()

    def secant_y(self) -> torch.Tensor:
        """Copy of y (reference solvers_torch.cpp:623-633)."""
        if not self._low_rank:
            raise RuntimeError("SPLRBackwardContext: low-rank correction is disabled, y is unavailable")
        return self._y.clone()

    def _csr(self):
        # A of the last factorized iterate in the reference's lower CSR
        # (reference solvers_torch.cpp:583-610): copies on the device
        f = self.factor_ptr()
        L = _lib.load()
        dev = self._M.device
        nnz = ctypes.c_int64(0)
        with torch.cuda.device(dev):
            _lib.check(L.cgot_splr_context_csr(f, self._n, self._m, None, None, None, ctypes.byref(nnz),
                                               _stream(dev)), "csr")
            vals = torch.empty(nnz.value, dtype=torch.float64, device=dev)
            cols = torch.empty(nnz.value, dtype=torch.int32, device=dev)
            rowptr = torch.empty(self.dim + 1, dtype=torch.int32, device=dev)
            _lib.check(L.cgot_splr_context_csr(f, self._n, self._m, vals.data_ptr(), cols.data_ptr(),
                                               rowptr.data_ptr(), ctypes.byref(nnz), _stream(dev)), "csr")
        return vals, cols, rowptr

    def csr_values(self) -> torch.Tensor:
        return self._csr()[0]

    def csr_colind(self) -> torch.Tensor:
        return self._csr()[1]

    def csr_rowptr(self) -> torch.Tensor:
        return self._csr()[2]

    def ctx_point(self) -> torch.Tensor:
        return self._dual[: self.dim].clone()

    def _check_r(self, r: torch.Tensor):
        if not isinstance(r, torch.Tensor) or not r.is_cuda:
            raise RuntimeError("SPLRBackwardContext: r must be a CUDA tensor")
        if r.dim() != 1 or r.shape[0] != self.dim:
            raise RuntimeError(f"SPLRBackwardContext: r must have length {self.dim}")

    # ---- production methods ----
    def apply_diagonal_preconditioner(self, r: torch.Tensor) -> torch.Tensor:
        """z = diag(A_k)^{-1} r (reference backward_context.cu:642-655)."""
        self._check_r(r)
        return r.to(torch.float64) * self._inv_diag

    def inv_diag_ptr(self) -> int:
        return self._inv_diag.data_ptr()

    @property
    def has_factor(self) -> bool:
        """True when the forward's last Newton factor is stored (niter > 0)."""
        return self._factor is not None

    def factor_ptr(self) -> int:
        if self._factor is None:
            raise RuntimeError("SPLRBackwardContext: no stored factor (the forward took no step)")
        return self._factor

    def _profiled(self, r, base_only: bool, repeats: int):
        self._check_r(r)
        if int(repeats) <= 0:
            raise RuntimeError("repeats must be positive")
        f = self.factor_ptr()
        rr = r.to(torch.float64).contiguous()
        z = torch.empty_like(rr)
        prof = _lib.PrecondProfile()
        L = _lib.load()
        dev = self._M.device
        with torch.cuda.device(dev):
            nbytes = L.cgot_splr_precond_workspace_size(f, self._n, self._m)
            ws = _workspace(nbytes, dev)
            rc = L.cgot_splr_precond_apply_profiled(f, self._n, self._m, rr.data_ptr(), z.data_ptr(),
                                                    1 if base_only else 0, int(repeats), ctypes.byref(prof),
                                                    ws.data_ptr(), nbytes, _stream(dev))
            _lib.check(rc, "preconditioner profile")
        return z, prof

    # ---- debug/benchmark methods (reference solvers_torch.cpp:380-580).  The
    # dicts carry the reference's keys.  The stages differ here: there is no
    # host copy and no cuDSS, and the rank-2 term needs no host scalar or sync.
    # The copy keys hold the time to enqueue the device copies, the cuDSS keys
    # the Schur triangular solves, and the keys of stages this design does not
    # have are 0.
    def profile_preconditioner(self, r: torch.Tensor) -> dict:
        t0 = time.perf_counter()
        z, p = self._profiled(r, False, 1)
        wrapper = 1e6 * (time.perf_counter() - t0)
        return {"output": z, "host_wrapper_us": wrapper, "host_memcpy_in_us": p.host_memcpy_in_us,
                "host_solve_launch_us": p.host_solve_launch_us, "host_lowrank_total_us": p.host_lowrank_us,
                "host_final_sync_wait_us": p.host_sync_wait_us, "host_memcpy_out_us": p.host_memcpy_out_us,
                "host_total_us": p.host_total_us, "gpu_cudss_solve_us": p.gpu_solve_us,
                "low_rank_active": self._low_rank, "lowrank_setup_launch_host_us": p.host_lowrank_us,
                "lowrank_sync_wait_host_us": 0.0, "lowrank_host_scalar_us": 0.0,
                "lowrank_correction_launch_host_us": 0.0, "lowrank_reduce_gpu_us": p.gpu_lowrank_reduce_us,
                "lowrank_correction_gpu_us": p.gpu_lowrank_correction_us}

    def profile_base_solve(self, r: torch.Tensor) -> dict:
        t0 = time.perf_counter()
        z, p = self._profiled(r, True, 1)
        wrapper = 1e6 * (time.perf_counter() - t0)
        return {"output": z, "host_wrapper_us": wrapper, "host_memcpy_in_us": p.host_memcpy_in_us,
                "host_cudss_execute_us": p.host_solve_launch_us, "host_sync_wait_us": p.host_sync_wait_us,
                "host_memcpy_out_us": p.host_memcpy_out_us, "host_total_us": p.host_total_us,
                "gpu_solve_us": p.gpu_solve_us, "low_rank_enabled": self._low_rank}

    def benchmark_preconditioner_repeated(self, r: torch.Tensor, repeats: int) -> dict:
        z, p = self._profiled(r, False, repeats)
        return {"output": z, "repeats": int(repeats), "cpp_repeated_total_us": p.repeated_total_us,
                "cpp_per_call_us": p.repeated_total_us / int(repeats)}

    def benchmark_base_solve_repeated(self, r: torch.Tensor, repeats: int) -> dict:
        z, p = self._profiled(r, True, repeats)
        return {"output": z, "repeats": int(repeats), "cpp_repeated_total_us": p.repeated_total_us,
                "cpp_per_call_us": p.repeated_total_us / int(repeats)}

    def benchmark_diag_setup(self, repeats: int) -> float:
        """Median microseconds of a validated diagonal set-up (reference
        solvers_torch.cpp:380-391).  The forward already writes diag(A)^{-1};
        the set-up left to time is its validation: a device check that every
        entry is finite and positive, read back on the host."""
        if int(repeats) <= 0:
            raise RuntimeError("repeats must be positive")
        ts = []
        for _ in range(int(repeats)):
            torch.cuda.synchronize(self._M.device)
            t0 = time.perf_counter()
            ok = bool(((self._inv_diag > 0) & torch.isfinite(self._inv_diag)).all())
            ts.append(1e6 * (time.perf_counter() - t0))
            if not ok:
                raise RuntimeError("SPLRBackwardContext: non-positive or non-finite diagonal")
        return float(sorted(ts)[len(ts) // 2])

    def apply_preconditioner(self, r: torch.Tensor) -> torch.Tensor:
        """z = B^{-1} r with B = A_k + the rank-2 BFGS term of the last forward
        iteration, A_k^{-1} by the stored Schur factor (reference
        backward_context.cu:543-581)."""
        self._check_r(r)
        f = self.factor_ptr()
        rr = r.to(torch.float64).contiguous()
        z = torch.empty_like(rr)
        L = _lib.load()
        dev = self._M.device
        with torch.cuda.device(dev):
            nbytes = L.cgot_splr_precond_workspace_size(f, self._n, self._m)
            ws = _workspace(nbytes, dev)
            rc = L.cgot_splr_precond_apply(f, self._n, self._m, rr.data_ptr(), z.data_ptr(), ws.data_ptr(),
                                           nbytes, _stream(dev))
            _lib.check(rc, "apply_preconditioner")
        return z


def _kwargs_density(kwargs, n, m):
    # reference defaults: solvers_torch.cpp:783-838
    dmax = min(10.0 / min(n, m), 1.0)
    if "density" in kwargs and kwargs["density"] is not None:
        dmax = min(max(float(kwargs["density"]), 0.0), 1.0)
    return dmax


def _set_cost(p, M):
    """The cost source of a problem struct: a stored M, or a GeometricCost's
    points (then M is NULL)."""
    if isinstance(M, GeometricCost):
        p.M = None
        p.xy.X, p.xy.Y, p.xy.d, p.xy.scale = M.X.data_ptr(), M.Y.data_ptr(), M.d, M.scale
    else:
        p.M = M.data_ptr()
        p.xy.X = None


def _check_one_cost_kind(Ms, what):
    kinds = {isinstance(M, GeometricCost) for M in Ms}
    if len(kinds) > 1:
        raise ValueError(f"{what}: a batch must not mix geometric (X, Y) and stored (M) costs")


def _make_problem(M, a, b, reg, tol, max_iter, kwargs, plan, dual, x0):
    n, m = M.shape
    p = _lib.SplrProblem()
    _set_cost(p, M)
    p.a, p.b = a.data_ptr(), b.data_ptr()
    p.n, p.m = n, m
    p.reg, p.tol, p.max_iter = float(reg), float(tol), int(max_iter)
    p.density_max = _kwargs_density(kwargs, n, m)
    if p.density_max <= 0.0:
        p.density_max = 1e-300     # explicit 0 keeps K = 1, as the reference's clamp does
    p.shift_max = max(float(kwargs.get("shift", 1e-3)), 0.0)
    p.sparsity_pattern_cycle = int(kwargs.get("sparsity_pattern_cycle", 30))
    if p.sparsity_pattern_cycle <= 0:
        raise RuntimeError("sparsity_pattern_cycle must be positive")
    p.candidate_sinkhorn_iter = max(int(kwargs.get("candidate_sinkhorn_iter", 3)), 0)
    p.x0 = x0.data_ptr() if x0 is not None else None
    p.plan = plan.data_ptr()
    p.dual = dual.data_ptr()
    p.trace = None
    return p


def _check_shapes(M, a, b):
    if M.dim() != 2:
        raise RuntimeError("M must be a 2D tensor")
    if a.dim() != 1 or b.dim() != 1:
        raise RuntimeError("a and b must be 1D tensors")
    n, m = M.shape
    if a.shape[0] != n or b.shape[0] != m:
        raise RuntimeError("Shape mismatch: M is [n x m], a should be length n, b should be length m")


def sinkhorn_splr_bwd_batch(Ms: Sequence[torch.Tensor], as_: Sequence[torch.Tensor],
                            bs: Sequence[torch.Tensor], reg, tol=1e-6, max_iter=1000, verbose=0,
                            **kwargs):
    """Ragged batch of backward-capable SPLR solves in one native call.

    Ms/as_/bs are sequences of [n_i, m_i], [n_i], [m_i] tensors (shapes may
    differ); reg may be a float or one value per instance.  allow_invalid=True
    returns a non-finite instance with context None instead of raising.  Returns one result
    dict per instance, each like sinkhorn_splr_bwd's.  Every instance's result
    is bitwise the same as solving it alone.
    """
    B = len(Ms)
    if B == 0:
        return []
    if len(as_) != B or len(bs) != B:
        raise RuntimeError("Ms, as_ and bs must have the same length")
    _check_one_cost_kind(Ms, "sinkhorn_splr_bwd_batch")
    regs = [float(r) for r in reg] if isinstance(reg, (list, tuple)) else [float(reg)] * B
    if len(regs) != B:
        raise RuntimeError("reg must be a float or have one value per instance")
    ctx_mode = kwargs.get("context_mode", "adopt")
    if ctx_mode not in ("adopt", "snapshot"):
        raise RuntimeError('context_mode must be "adopt" or "snapshot"')
    device = _target_device(Ms[0])
    x0s = kwargs.get("x0", None)
    want_trace = bool(kwargs.get("trace", False))
    traces = []
    L = _lib.load()
    probs = (_lib.SplrProblem * B)()
    keep = []
    outs = []
    for i in range(B):
        _check_shapes(Ms[i], as_[i], bs[i])
        M = _to_dev(Ms[i], device)
        a = _to_dev(as_[i], device)
        b = _to_dev(bs[i], device)
        n, m = M.shape
        x0 = None
        if x0s is not None:
            x0 = x0s[i] if isinstance(x0s, (list, tuple)) else (x0s if B == 1 else x0s[i])
            if x0.dim() != 1 or x0.shape[0] != n + m:
                raise RuntimeError("x0 must be a 1D tensor of length n + m")
            x0 = _to_dev(x0, device)
        plan = torch.empty((n, m), dtype=torch.float64, device=device)
        dual = torch.empty(n + m, dtype=torch.float64, device=device)
        probs[i] = _make_problem(M, a, b, regs[i], tol, max_iter, kwargs, plan, dual, x0)
        if want_trace:
            tr = (ctypes.c_double * (8 * max(int(max_iter), 1)))()
            traces.append(tr)
            probs[i].trace = ctypes.cast(tr, ctypes.c_void_p).value
        keep.append((M, a, b, x0))
        outs.append((plan, dual))
    with torch.cuda.device(device):
        nbytes = L.cgot_splr_workspace_size(B, probs)
        ws = _workspace(nbytes, device)
        infos = (_lib.SplrInfo * B)()
        rc = L.cgot_splr_solve(B, probs, ws.data_ptr(), nbytes, _stream(device), infos)
        _lib.check(rc, "sinkhorn_splr_bwd")
    results = []
    allow_invalid = bool(kwargs.get("allow_invalid", False))
    for i in range(B):
        info = infos[i]
        plan, dual = outs[i]
        if not info.valid and not allow_invalid:
            raise RuntimeError(
                "torch_sinkhorn_splr_bwd: forward pass failed (non-finite objective value); "
                f"no backward context available (instance {i}, shape {info.n}x{info.m})")
        ctx = (SPLRBackwardContext(keep[i][0], dual, regs[i], info, ws, snapshot=(ctx_mode == "snapshot"))
               if info.valid else None)
        src = Ms[i]
        res = {
            "niter": int(info.niter),
            "converged": bool(info.converged),
            "plan": plan if (src.is_cuda and src.dtype == torch.float64) else plan.to(src.device, src.dtype),
            "dual": dual if (src.is_cuda and src.dtype == torch.float64) else dual.to(src.device, src.dtype),
            "context": ctx,
            "valid": bool(info.valid),
            "gnorm": float(info.gnorm),
            "objective": float(info.f),
        }
        if want_trace:
            import numpy as _np
            res["trace"] = _np.frombuffer(traces[i], dtype=_np.float64).reshape(-1, 8)[: int(info.niter)].copy()
        if verbose > 0:
            print(f"sinkhorn_splr_bwd (HIP) [{info.n} x {info.m}] niter={info.niter} "
                  f"converged={bool(info.converged)} gnorm={info.gnorm:.3e}")
        results.append(res)
    return results


class SplrSession:
    """Continuous batching: one persistent lockstep group of SPLR solves on a
    stream (C ABI cgot_session_*; no counterpart in the reference, whose
    splr_ot_batch is a sequential loop, curegot/autograd.py:340-367).

    add() queues a problem (returns False while the group is full); step()
    runs up to `iters` iterations of every instance in the group, newly added
    ones starting first, and returns [(tag, result)] for the instances that
    stopped, each result as sinkhorn_splr_bwd's (plan, dual, niter, converged,
    context, valid, gnorm, objective).  An instance's result is bitwise the
    same as solving it alone.  Use a session from one thread at a time; its
    work runs on `stream` (a torch.cuda.Stream; default: the current one when
    created) and is complete when step() returns.
    """

    def __init__(self, device=None, stream=None, capacity=64, max_sum_nm=1 << 20, max_tiles=1 << 20,
                 max_nblk=128):
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        self.capacity = int(capacity)
        L = _lib.load()
        with torch.cuda.device(self.device):
            nbytes = L.cgot_session_workspace_size(self.capacity, int(max_sum_nm), int(max_tiles), int(max_nblk))
            self._ws = _workspace(nbytes, self.device)
            h = ctypes.c_void_p()
            rc = L.cgot_session_create(self.capacity, int(max_sum_nm), int(max_tiles), int(max_nblk),
                                       self._ws.data_ptr(), nbytes, int(self.stream.cuda_stream), ctypes.byref(h))
            _lib.check(rc, "session_create")
        self._h = h
        self._jobs = {}
        self._next = 0

    @staticmethod
    def tiles(n, m) -> int:
        """Launch-map entries an (n, m) instance takes (for the max_tiles limit)."""
        return int(_lib.load().cgot_session_inst_tiles(int(n), int(m)))

    def live(self) -> int:
        return int(_lib.load().cgot_session_live(self._h))

    def add(self, M, a, b, reg, tol=1e-6, max_iter=1000, tag=None, **kwargs) -> bool:
        if self._h is None:
            raise RuntimeError("session is closed")
        _check_shapes(M, a, b)
        M, a, b = _to_dev(M, self.device), _to_dev(a, self.device), _to_dev(b, self.device)
        n, m = M.shape
        x0 = kwargs.get("x0", None)
        if x0 is not None:
            if x0.dim() != 1 or x0.shape[0] != n + m:
                raise RuntimeError("x0 must be a 1D tensor of length n + m")
            x0 = _to_dev(x0, self.device)
        plan = torch.empty((n, m), dtype=torch.float64, device=self.device)
        dual = torch.empty(n + m, dtype=torch.float64, device=self.device)
        p = _make_problem(M, a, b, reg, tol, max_iter, kwargs, plan, dual, x0)
        L = _lib.load()
        nbytes = L.cgot_session_inst_workspace_size(ctypes.byref(p))
        ws = _workspace(nbytes, self.device)
        key = self._next
        with torch.cuda.device(self.device):
            rc = L.cgot_session_add(self._h, ctypes.byref(p), ws.data_ptr(), nbytes, key)
        if rc == -3:
            return False
        if rc < 0:
            _lib.check(rc, "session_add")
        self._next += 1
        self._jobs[key] = (tag, M, a, b, x0, plan, dual, float(reg), ws, kwargs.get("context_mode", "adopt"))
        return True

    def step(self, iters=1):
        if self._h is None:
            raise RuntimeError("session is closed")
        L = _lib.load()
        cap = max(len(self._jobs), 1)
        tags = (ctypes.c_int64 * cap)()
        infos = (_lib.SplrInfo * cap)()
        nd = ctypes.c_int32(0)
        with torch.cuda.device(self.device):
            rc = L.cgot_session_step(self._h, int(iters), tags, infos, cap, ctypes.byref(nd))
        _lib.check(rc, "session_step")
        out = []
        for k in range(nd.value):
            tag, M, a, b, x0, plan, dual, reg, ws, mode = self._jobs.pop(int(tags[k]))
            info = infos[k]
            ctx = (SPLRBackwardContext(M, dual, reg, info, ws, snapshot=(mode == "snapshot"))
                   if info.valid else None)
            out.append((tag, {"niter": int(info.niter), "converged": bool(info.converged), "plan": plan,
                              "dual": dual, "context": ctx, "valid": bool(info.valid), "gnorm": float(info.gnorm),
                              "objective": float(info.f)}))
        return out

    def close(self):
        if self._h is not None:
            _lib.check(_lib.load().cgot_session_destroy(self._h), "session_destroy")
            self._h = None
            self._jobs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sinkhorn_splr_bwd(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """Backward-capable SPLR forward.  Reference: torch_sinkhorn_splr_bwd,
    src/solvers_torch.cpp:707-912.  Returns {niter, converged, plan, dual, context}."""
    _check_shapes(M, a, b)
    if "x0" in kwargs and kwargs["x0"] is not None:
        kwargs = dict(kwargs)
        kwargs["x0"] = [kwargs["x0"]]
    return sinkhorn_splr_bwd_batch([M], [a], [b], reg, tol, max_iter, verbose, **kwargs)[0]


def sinkhorn_splr(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """Forward-only SPLR (reference torch_sinkhorn_splr, src/solvers_torch.cpp:130-288):
    {niter, plan, dual}.  Like the reference, a non-finite objective is not an
    error here: the iterates are returned as they are (only the
    backward-capable entry needs a usable context)."""
    kwargs = dict(kwargs)
    kwargs["allow_invalid"] = True
    out = sinkhorn_splr_bwd(M, a, b, reg, tol, max_iter, verbose, **kwargs)
    return {"niter": out["niter"], "plan": out["plan"], "dual": out["dual"]}


def sinkhorn_bcd(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """Sinkhorn/BCD solve (reference torch_sinkhorn_bcd, src/solvers_torch.cpp:20-127):
    {niter, plan, dual}; the dual is not gauge-fixed."""
    _check_shapes(M, a, b)
    device = _target_device(M)
    Md, ad, bd = _to_dev(M, device), _to_dev(a, device), _to_dev(b, device)
    n, m = Md.shape
    x0 = kwargs.get("x0", None)
    if x0 is not None:
        if x0.dim() != 1 or x0.shape[0] != n + m:
            raise RuntimeError("x0 must be a 1D tensor of length n + m")
        x0 = _to_dev(x0, device)
    plan = torch.empty((n, m), dtype=torch.float64, device=device)
    dual = torch.empty(n + m, dtype=torch.float64, device=device)
    p = _lib.BcdProblem()
    _set_cost(p, Md)
    p.a, p.b = ad.data_ptr(), bd.data_ptr()
    p.n, p.m, p.reg, p.tol, p.max_iter = n, m, float(reg), float(tol), int(max_iter)
    p.x0 = x0.data_ptr() if x0 is not None else None
    p.plan, p.dual = plan.data_ptr(), dual.data_ptr()
    L = _lib.load()
    niter = ctypes.c_int32(0)
    with torch.cuda.device(device):
        nbytes = L.cgot_bcd_workspace_size(1, ctypes.byref(p))
        ws = _workspace(nbytes, device)
        rc = L.cgot_bcd_solve(1, ctypes.byref(p), ws.data_ptr(), nbytes, _stream(device), ctypes.byref(niter))
        _lib.check(rc, "sinkhorn_bcd")
    if not (M.is_cuda and M.dtype == torch.float64):
        plan, dual = plan.to(M.device, M.dtype), dual.to(M.device, M.dtype)
    if verbose > 0:
        print(f"sinkhorn_bcd (HIP) [{n} x {m}] niter={niter.value}")
    return {"niter": int(niter.value), "plan": plan, "dual": dual}


def adjoint_batch(items, tol, max_iter):
    """Batched implicit backward.  items: list of (M, dual, G, reg, inv_diag or
    None[, factor pointer or None]).  Returns (lambdas, grads, infos,
    lambda_ext): lambda_ext[i] is [n+m], lambdas[i] followed by a 0 (its first
    n+m-1 entries are lambdas[i]; all are views of one zero-filled buffer, so
    the batch costs one fill and no per-instance copies)."""
    B = len(items)
    device = items[0][0].device
    _check_one_cost_kind([it[0] for it in items], "solve_backward_batch")
    probs = (_lib.AdjointProblem * B)()
    lams, grads, keep, exts = [], [], [], []
    sizes = [it[0].shape[0] + it[0].shape[1] for it in items]
    lam_all = torch.zeros(sum(sizes), dtype=torch.float64, device=device)
    off = 0
    for i, item in enumerate(items):
        M, dual, G, reg, inv = item[:5]
        factor = item[5] if len(item) > 5 else None
        n, m = M.shape
        Gd = G.to(device=device, dtype=torch.float64).contiguous()
        ext = lam_all[off:off + n + m]
        off += n + m
        lam = ext[:n + m - 1]
        exts.append(ext)
        grad = torch.empty((n, m), dtype=torch.float64, device=device)
        p = probs[i]
        _set_cost(p, M)
        p.dual, p.G = dual.data_ptr(), Gd.data_ptr()
        p.n, p.m, p.reg = n, m, float(reg)
        p.inv_diag = inv.data_ptr() if inv is not None else None
        p.factor = factor
        p.tol, p.max_iter = float(tol), int(max_iter)
        p.lam, p.grad = lam.data_ptr(), grad.data_ptr()
        keep.append(Gd)
        lams.append(lam)
        grads.append(grad)
    L = _lib.load()
    infos = (_lib.AdjointInfo * B)()
    with torch.cuda.device(device):
        nbytes = L.cgot_adjoint_workspace_size(B, probs)
        ws = _workspace(nbytes, device)
        rc = L.cgot_adjoint_solve(B, probs, ws.data_ptr(), nbytes, _stream(device), infos)
        _lib.check(rc, "solve_backward")
    return lams, grads, [infos[i] for i in range(B)], exts


def lse_update(M, a, b, dual, reg, which, repeats=1):
    """One (or `repeats`) BCD half-sweep(s) in place on dual (device tensors)."""
    n, m = M.shape
    L = _lib.load()
    with torch.cuda.device(M.device):
        nbytes = L.cgot_lse_workspace_size(n, m)
        ws = _workspace(nbytes, M.device)
        rc = L.cgot_lse_update(M.data_ptr(), a.data_ptr(), b.data_ptr(), n, m, float(reg), int(which),
                               int(repeats), dual.data_ptr(), ws.data_ptr(), nbytes, _stream(M.device))
        _lib.check(rc, "lse_update")
    return dual


def T_sparsify(alpha, beta, M, a, b, reg, shift, K, rhs=None):
    """Device T pass + top-K sparsified Hessian (test hook; reference
    curegot.tests.test_T_computation_sparsify, src/tests.cpp:36-150)."""
    device = _target_device(M)
    Md = _to_dev(M, device)
    n, m = Md.shape
    gamma = _to_dev(torch.cat([torch.as_tensor(alpha).flatten(), torch.as_tensor(beta).flatten()]), device)
    ad, bd = _to_dev(torch.as_tensor(a), device), _to_dev(torch.as_tensor(b), device)
    N = n + m - 1
    Te = n * (m - 1)
    Kp = min(max(int(K), 1), Te) if Te > 0 else 0
    f64 = dict(dtype=torch.float64, device=device)
    rows, cols, grad = torch.empty(n, **f64), torch.empty(m, **f64), torch.empty(N, **f64)
    scal = torch.empty(2, **f64)
    vals = torch.empty(Kp + N, **f64)
    colind = torch.empty(Kp + N, dtype=torch.int32, device=device)
    rowptr = torch.empty(N + 1, dtype=torch.int32, device=device)
    rhs_d = _to_dev(torch.as_tensor(rhs), device) if rhs is not None else None
    sol = torch.empty(N, **f64) if rhs is not None else None
    L = _lib.load()
    with torch.cuda.device(device):
        nbytes = L.cgot_T_sparsify_workspace_size(n, m, int(K))
        ws = _workspace(nbytes, device)
        rc = L.cgot_T_sparsify(gamma.data_ptr(), Md.data_ptr(), ad.data_ptr(), bd.data_ptr(), n, m,
                               float(reg), float(shift), int(K), rows.data_ptr(), cols.data_ptr(),
                               grad.data_ptr(), scal.data_ptr(), vals.data_ptr(), colind.data_ptr(),
                               rowptr.data_ptr(), rhs_d.data_ptr() if rhs_d is not None else None,
                               sol.data_ptr() if sol is not None else None, ws.data_ptr(), nbytes,
                               _stream(device))
        _lib.check(rc, "T_sparsify")
    out = {"Trowsums": rows, "Tcolsums": cols, "grad": grad, "objfn": float(scal[0]),
           "gnorm": float(scal[1]), "csr_val": vals, "csr_colind": colind, "csr_rowptr": rowptr}
    if sol is not None:
        out["sol"] = sol
    return out
