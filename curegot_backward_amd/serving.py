"""Continuous-batching request loop for OT solves on one GPU.

Requests (M, a, b, eta[, G]) arrive at any time.  A fixed number of slots
(worker threads, each with its own HIP stream) take work from one FIFO queue:
a slot that becomes free opens a batch with the oldest waiting request and
adds the later waiting requests of the same size class (the log2 of the Schur
block count, which sets how long one lockstep iteration takes), in arrival
order, while their HBM footprint (solver workspace, plan, dual and, for a
backward request, the gradient and the adjoint workspace) fits in the memory
the other slots leave free.  A batch is one
ragged native solve (every instance advances in lockstep inside it), followed
by one batched adjoint solve for the requests that carry an upstream gradient
G.  Requests that arrive while other batches are on the GPU are admitted by
the next slot that frees up, so their work overlaps on the device.  Every
result is bitwise the same as solving the request alone: the kernels are
deterministic per instance.

The reference has no serving layer (its splr_ot_batch is a Python loop over
instances, curegot/autograd.py:340-367); this module is built on the same
native entry points as curegot_backward_amd.torch.
"""
from __future__ import annotations

import collections
import ctypes
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from . import _lib, _native
from .backward import solve_backward_batch


def splr_footprint(n: int, m: int, density: Optional[float] = None, backward: bool = False) -> int:
    """HBM bytes one request needs beyond its inputs: the native SPLR workspace
    of an (n, m) instance plus its plan and dual, and for a backward request
    the gradient, lambda and the adjoint workspace (an upper bound)."""
    p = _lib.SplrProblem()
    p.M = p.a = p.b = 256          # the size query does not dereference them
    p.n, p.m = int(n), int(m)
    p.reg, p.tol, p.max_iter = 1.0, 1e-8, 1
    p.density_max = -1.0 if density is None else float(density)
    p.shift_max = -1.0
    p.sparsity_pattern_cycle = 0
    p.candidate_sinkhorn_iter = -1
    ws = int(_lib.load().cgot_splr_workspace_size(1, ctypes.byref(p)))
    if ws == 0:
        raise ValueError(_lib.load().cgot_last_error().decode())
    total = ws + 8 * n * m + 8 * (n + m)
    if backward:
        total += 8 * n * m + 8 * (n + m) + 64 * (n + m) + 16 * n * m // 32 + (1 << 20)
    return total


@dataclass
class _Request:
    M: torch.Tensor
    a: torch.Tensor
    b: torch.Tensor
    eta: float
    G: Optional[torch.Tensor]
    footprint: int
    t_submit: float
    ready: torch.cuda.Event          # recorded on the submitter's stream
    stream: torch.cuda.Stream        # the submitter's stream, which will use the results
    future: Future = field(default_factory=Future)


class OTServer:
    """Continuous batching of SPLR-OT requests.

    slots:        concurrent batches (each on its own stream; the device has a
                  few hardware queues, so 2-4 is the useful range)
    hbm_budget:   bytes the in-flight batches may use together (default: 80%
                  of the device memory free at start-up)
    max_batch:    largest number of requests in one batch
    groups:       concurrent size classes inside one batch (see
                  cgot_set_max_groups; 1 keeps each batch on its slot's stream)

    The inputs of a request may still be in production on the submitting
    thread's current stream: the request waits for that stream's work up to the
    submit() call.  submit() returns a concurrent.futures.Future whose result is
    a dict with
    plan, dual, niter, converged (and grad, lambda, bwd_iterations for a request
    with G), plus queue_s (submit to start of its batch) and latency_s (submit
    to result).
    """

    def __init__(self, device=0, slots: int = 4, hbm_budget: Optional[int] = None, max_batch: int = 64,
                 tol: float = 1e-8, max_iter: int = 1000, groups: int = 1, preconditioner: str = "diag"):
        self.device = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("OTServer needs a CUDA (HIP) device")
        if slots < 1 or max_batch < 1:
            raise ValueError("slots and max_batch must be >= 1")
        if hbm_budget is None:
            free, _total = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.8 * free)
        self.hbm_budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.tol, self.max_iter = float(tol), int(max_iter)
        self.groups = int(groups)
        self.preconditioner = preconditioner
        self._queue: collections.deque = collections.deque()
        self._cv = threading.Condition()
        self._in_use = 0
        self._closed = False
        self._done: list = []          # (latency_s, queue_s, n, m, backward)
        self._log: list = []           # (slot, size class, requests, start, end) per batch
        self._batches = 0
        self._t_first: Optional[float] = None
        self._t_last: Optional[float] = None
        self._workers = [threading.Thread(target=self._run, args=(k,), daemon=True) for k in range(slots)]
        for w in self._workers:
            w.start()

    # ------------------------------------------------------------------ API
    def submit(self, M: torch.Tensor, a: torch.Tensor, b: torch.Tensor, eta: float,
               G: Optional[torch.Tensor] = None) -> Future:
        if M.dim() != 2 or a.dim() != 1 or b.dim() != 1:
            raise ValueError("M must be 2D, a and b 1D")
        n, m = M.shape
        if a.shape[0] != n or b.shape[0] != m:
            raise ValueError("shape mismatch between M, a and b")
        if G is not None and G.shape != M.shape:
            raise ValueError("G must have the shape of M")
        if not eta > 0:
            raise ValueError("eta must be positive")
        fp = splr_footprint(n, m, backward=G is not None)
        sub = torch.cuda.current_stream(self.device)
        ready = torch.cuda.Event()
        ready.record(sub)   # the inputs are complete after this point
        req = _Request(M, a, b, float(eta), G, fp, time.perf_counter(), ready, sub)
        with self._cv:
            if self._closed:
                raise RuntimeError("server is closed")
            if fp > self.hbm_budget:
                req.future.set_exception(MemoryError(
                    f"request needs {fp} bytes of HBM, more than the server budget {self.hbm_budget}"))
                return req.future
            self._queue.append(req)
            self._cv.notify_all()
        return req.future

    def close(self, wait: bool = True):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        if wait:
            for w in self._workers:
                w.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def batch_log(self) -> list:
        """(slot, size class, requests, start_s, end_s) of every finished batch,
        times relative to the first admitted request's submission."""
        with self._cv:
            t0 = self._t_first or 0.0
            return [(k, c, nb, a - t0, b - t0) for (k, c, nb, a, b) in self._log]

    def stats(self) -> dict:
        with self._cv:
            done = list(self._done)
            span = (self._t_last - self._t_first) if (self._t_first is not None and self._t_last) else 0.0
            batches = self._batches
        if not done:
            return {"completed": 0}
        lat = np.array([d[0] for d in done]) * 1000.0
        q = np.array([d[1] for d in done]) * 1000.0
        return {
            "completed": len(done),
            "batches": batches,
            "solves_per_s": len(done) / span if span > 0 else float("nan"),
            "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            "max_ms": float(lat.max()),
            "queue_p50_ms": float(np.percentile(q, 50)),
            "queue_p99_ms": float(np.percentile(q, 99)),
            "backward": int(sum(1 for d in done if d[4])),
        }

    # ------------------------------------------------------------ internals
    @staticmethod
    def _size_class(n: int, m: int) -> int:
        """log2 of the Schur block count: the length of the per-iteration
        factorization chain, which sets a lockstep batch's pace."""
        nblk = max(1, -(-max(1, min(n, m - 1)) // 64))
        return nblk.bit_length() - 1

    def _class_cap(self, cls: int) -> int:
        """Largest batch of one size class.  Long factorization chains gain
        nothing from lockstep (a batch advances at its slowest member's pace),
        so the large classes run in small batches, side by side in the slots."""
        return self.max_batch if cls <= 3 else max(1, self.max_batch >> (2 * (cls - 3)))

    def _admit(self):
        """Block until some requests fit.  The oldest waiting request opens the
        batch; later requests of the same size class join it in arrival order
        while the batch and memory limits allow, so one slow instance does not
        set the pace of many fast ones."""
        with self._cv:
            while True:
                if self._queue:
                    head = self._queue[0]
                    if self._in_use + head.footprint <= self.hbm_budget:
                        cls = self._size_class(head.M.shape[0], head.M.shape[1])
                        cap = self._class_cap(cls)
                        batch, used, keep = [], 0, collections.deque()
                        while self._queue:
                            r = self._queue.popleft()
                            if (len(batch) < cap
                                    and self._size_class(r.M.shape[0], r.M.shape[1]) == cls
                                    and self._in_use + used + r.footprint <= self.hbm_budget):
                                batch.append(r)
                                used += r.footprint
                            else:
                                keep.append(r)
                        self._queue = keep
                        self._in_use += used
                        if self._t_first is None:
                            self._t_first = min(r.t_submit for r in batch)
                        return batch, used
                elif self._closed:
                    return None, 0
                self._cv.wait()

    def _release(self, used: int):
        with self._cv:
            self._in_use -= used
            self._cv.notify_all()

    def _run(self, k: int):
        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream(self.device)          # default priority
        _lib.set_max_groups(self.groups)
        while True:
            batch, used = self._admit()
            if batch is None:
                return
            t_start = time.perf_counter()
            try:
                with torch.cuda.stream(stream):
                    for r in batch:
                        stream.wait_event(r.ready)
                    outs = _native.sinkhorn_splr_bwd_batch(
                        [r.M for r in batch], [r.a for r in batch], [r.b for r in batch],
                        [r.eta for r in batch], self.tol, self.max_iter, allow_invalid=True)
                    bwd = [(i, r) for i, r in enumerate(batch) if r.G is not None]
                    bres = {}
                    ok = [(i, r) for i, r in bwd if outs[i]["context"] is not None]
                    if ok:
                        res = solve_backward_batch([(r.G, outs[i]["context"], r.eta) for i, r in ok],
                                                   preconditioner=self.preconditioner)
                        bres = {i: x for (i, _), x in zip(ok, res)}
                    stream.synchronize()
                t_end = time.perf_counter()
                for i, r in enumerate(batch):
                    o = outs[i]
                    out = {"plan": o["plan"], "dual": o["dual"], "niter": o["niter"],
                           "converged": o["converged"], "valid": o["valid"],
                           "queue_s": t_start - r.t_submit, "latency_s": t_end - r.t_submit}
                    if r.G is not None:
                        if i in bres:
                            out.update(grad=bres[i]["grad"], **{"lambda": bres[i]["lambda"]},
                                       bwd_iterations=bres[i]["iterations"], bwd_converged=bres[i]["converged"])
                        else:
                            out.update(grad=None, bwd_converged=False)
                    # the outputs were allocated on this slot's stream: once the
                    # caller frees them, the caching allocator must not hand
                    # their blocks to this stream's next batch before the
                    # caller's stream has finished reading them
                    for key in ("plan", "dual", "grad", "lambda"):
                        t = out.get(key)
                        if torch.is_tensor(t):
                            t.record_stream(r.stream)
                    r.future.set_result(out)
                with self._cv:
                    self._batches += 1
                    self._t_last = t_end
                    self._log.append((k, self._size_class(batch[0].M.shape[0], batch[0].M.shape[1]), len(batch),
                                      t_start, t_end))
                    self._done.extend((t_end - r.t_submit, t_start - r.t_submit, r.M.shape[0], r.M.shape[1],
                                       r.G is not None) for r in batch)
            except Exception as e:                       # pragma: no cover - surfaced to the callers
                for r in batch:
                    if not r.future.done():
                        r.future.set_exception(e)
            finally:
                self._release(used)
