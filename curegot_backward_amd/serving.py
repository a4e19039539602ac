"""Continuous-batching request loop for OT solves on one GPU.

Requests (M, a, b, eta[, G]) arrive at any time.  Two admission modes:

continuous (default)
    Each forward slot (worker thread + HIP stream) owns a session
    (curegot_backward_amd._native.SplrSession): one persistent lockstep group
    into which waiting requests of the slot's size classes (the log2 of the
    Schur block count; larger classes go to lower slots, so that one slow
    instance does not set the pace of many fast ones) are admitted at every
    iteration boundary while the group's limits and the HBM budget allow, and
    from which each request leaves as soon as it has converged.  With more
    than one slot, the last one is an adjoint worker that runs the backward
    requests' implicit adjoints, batched, as their forwards finish.
batch (continuous=False)
    A free slot opens a batch with the oldest request and adds later requests
    of the same size class; the batch is one ragged native solve, followed by
    one batched adjoint for the requests with G, and holds the slot until its
    slowest member finishes.

In both, admission is by HBM footprint (solver workspace, plan, dual and, for
a backward request, the gradient and the adjoint workspace) against a budget,
requests that arrive while others run are admitted as slots and memory free
up, so their work overlaps on the device, and every result is bitwise the same
as solving the request alone: the kernels are deterministic per instance.

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
from ._native import SplrSession
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

    slots:        worker streams (a process has GPU_MAX_HW_QUEUES hardware
                  queues, 4 by default, so 2-4 is the useful range; more need
                  more queues).  Continuous mode: slots - 1 forward slots (at
                  most 7) plus one adjoint worker (one slot does both)
    hbm_budget:   bytes the in-flight requests may use together (default: 80%
                  of the device memory free at start-up)
    max_batch:    largest number of requests in one lockstep group
    continuous:   admission at iteration granularity (default) or per batch
    step_iters:   continuous mode: iterations a group runs between admissions
    groups:       batch mode: concurrent size classes inside one batch (see
                  cgot_set_max_groups; 1 keeps each batch on its slot's stream)

    The inputs of a request may still be in production on the submitting
    thread's current stream: the request waits for that stream's work up to the
    submit() call.  submit() returns a concurrent.futures.Future whose result is
    a dict with
    plan, dual, niter, converged (and grad, lambda, bwd_iterations for a request
    with G), plus queue_s (submit to start of its solve) and latency_s (submit
    to result).
    """

    NCLASS = 16

    def __init__(self, device=0, slots: int = 4, hbm_budget: Optional[int] = None, max_batch: int = 256,
                 tol: float = 1e-8, max_iter: int = 1000, groups: int = 1, preconditioner: str = "diag",
                 continuous: bool = True, step_iters: int = 3):
        self.device = torch.device("cuda", device) if isinstance(device, int) else torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("OTServer needs a CUDA (HIP) device")
        if slots < 1 or max_batch < 1 or step_iters < 1:
            raise ValueError("slots, max_batch and step_iters must be >= 1")
        if hbm_budget is None:
            free, _total = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.8 * free)
        self.hbm_budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.tol, self.max_iter = float(tol), int(max_iter)
        self.groups = int(groups)
        self.preconditioner = preconditioner
        self.continuous = bool(continuous)
        self.step_iters = int(step_iters)
        self._queue: collections.deque = collections.deque()
        self._cq = [collections.deque() for _ in range(self.NCLASS)]   # continuous mode, per size class
        self._bq: collections.deque = collections.deque()              # continuous mode, adjoint jobs
        self._cv = threading.Condition()
        self._in_use = 0
        self._closed = False
        self._fwd_open = 0
        self._done: list = []          # (latency_s, queue_s, n, m, backward)
        self._log: list = []           # (slot, size class, requests, start, end) per batch or admission
        self._batches = 0
        self._t_first: Optional[float] = None
        self._t_last: Optional[float] = None
        if self.continuous:
            self._nfwd = min(max(1, slots - 1) if slots > 1 else 1, 7)
            self._adj_worker = slots > 1
            self._fwd_open = self._nfwd
            self._workers = [threading.Thread(target=self._run_session, args=(k,), daemon=True)
                             for k in range(self._nfwd)]
            if self._adj_worker:
                self._workers.append(threading.Thread(target=self._run_adjoints, args=(self._nfwd,), daemon=True))
        else:
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
            if self.continuous:
                self._cq[min(self._size_class(n, m), self.NCLASS - 1)].append(req)
            else:
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

    # ------------------------------------------------------- continuous mode
    # Lower class bounds of the forward slots (slot 0 takes the largest
    # classes); a size class is log2 of the Schur block count, 0..7 for n, m
    # up to 8192.  The mixed-size stream puts about half of its requests in
    # classes 0-2 and a third in classes 3-4, while the work per request grows
    # with the class.
    _CLASS_BOUNDS = {1: [0], 2: [4, 0], 3: [5, 3, 0], 4: [6, 4, 2, 0], 5: [6, 5, 3, 2, 0], 6: [7, 6, 5, 3, 2, 0],
                     7: [7, 6, 5, 4, 3, 2, 0]}

    def _slot_classes(self, k: int) -> list:
        """Size classes forward slot k admits (contiguous ranges, the largest
        classes on slot 0)."""
        lo = self._CLASS_BOUNDS[self._nfwd]
        hi = [self.NCLASS] + lo[:-1]
        return list(range(lo[k], hi[k]))

    def _take_for_session(self, k: int, classes: list, room: int, empty: bool) -> list:
        """Oldest waiting requests of `classes`, in arrival order, while the
        group has room and their footprints fit the HBM budget (their memory
        is reserved here).  Blocks while the group is empty and nothing waits;
        returns None once the server is closed and nothing is left."""
        with self._cv:
            while True:
                picked = []
                while len(picked) < room:
                    best = None
                    for c in classes:
                        q = self._cq[c]
                        if q and (best is None or q[0].t_submit < best[0].t_submit):
                            best = q
                    if best is None or self._in_use + best[0].footprint > self.hbm_budget:
                        break
                    r = best.popleft()
                    self._in_use += r.footprint
                    picked.append(r)
                if picked or not empty:
                    if picked and self._t_first is None:
                        self._t_first = min(r.t_submit for r in picked)
                    return picked
                if self._closed and not any(self._cq[c] for c in classes):
                    return None
                self._cv.wait()

    def _put_back(self, r: "_Request"):
        with self._cv:
            self._cq[min(self._size_class(r.M.shape[0], r.M.shape[1]), self.NCLASS - 1)].appendleft(r)
            self._in_use -= r.footprint
            self._cv.notify_all()

    def _complete(self, r: "_Request", out: dict, t_start: float, fail: Optional[BaseException] = None):
        t_end = time.perf_counter()
        if fail is not None:
            r.future.set_exception(fail)
        else:
            out["queue_s"] = t_start - r.t_submit
            out["latency_s"] = t_end - r.t_submit
            # the outputs were allocated on a server stream: the caching
            # allocator must not hand their blocks to later work there before
            # the caller's stream has finished reading them
            for key in ("plan", "dual", "grad", "lambda"):
                t = out.get(key)
                if torch.is_tensor(t):
                    t.record_stream(r.stream)
            r.future.set_result(out)
        with self._cv:
            self._in_use -= r.footprint
            self._t_last = t_end
            if fail is None:
                self._done.append((t_end - r.t_submit, t_start - r.t_submit, r.M.shape[0], r.M.shape[1],
                                   r.G is not None))
            self._cv.notify_all()

    @staticmethod
    def _fwd_out(o: dict) -> dict:
        return {"plan": o["plan"], "dual": o["dual"], "niter": o["niter"], "converged": o["converged"],
                "valid": o["valid"]}

    def _adjoint_batch(self, jobs: list, stream):
        """Implicit adjoints of finished forwards [(request, forward result, t_start)]."""
        ok = [(r, o, t) for (r, o, t) in jobs if o["context"] is not None]
        for (r, o, t) in jobs:
            if o["context"] is None:
                self._complete(r, dict(self._fwd_out(o), grad=None, bwd_converged=False), t)
        if not ok:
            return
        with torch.cuda.stream(stream):
            for (r, _o, _t) in ok:
                stream.wait_event(r.ready)
            res = solve_backward_batch([(r.G, o["context"], r.eta) for (r, o, _t) in ok],
                                       preconditioner=self.preconditioner)
            stream.synchronize()
        for (r, o, t), x in zip(ok, res):
            self._complete(r, dict(self._fwd_out(o), grad=x["grad"], **{"lambda": x["lambda"]},
                                   bwd_iterations=x["iterations"], bwd_converged=x["converged"]), t)

    def _run_session(self, k: int):
        """Forward slot k: one session fed at every step boundary."""
        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream(self.device)          # default priority
        classes = self._slot_classes(k)
        sess = None
        inflight = {}                                     # tag -> (request, t_start)
        try:
            while True:
                if sess is None:
                    sess = SplrSession(device=self.device, stream=stream, capacity=self.max_batch,
                                       max_sum_nm=1 << 21, max_tiles=1 << 21, max_nblk=256)
                picked = self._take_for_session(k, classes, self.max_batch - len(inflight), not inflight)
                if picked is None:
                    return
                if picked:
                    t0 = time.perf_counter()
                    added = 0
                    with torch.cuda.stream(stream):
                        for i, r in enumerate(picked):
                            stream.wait_event(r.ready)
                            tag = id(r)
                            if not sess.add(r.M, r.a, r.b, r.eta, self.tol, self.max_iter, tag=tag):
                                for rr in picked[i:]:
                                    self._put_back(rr)
                                break
                            inflight[tag] = (r, t0)
                            added += 1
                    if added:
                        with self._cv:
                            self._batches += 1
                            self._log.append((k, self._size_class(picked[0].M.shape[0], picked[0].M.shape[1]),
                                              added, t0, t0))
                if not inflight:
                    continue
                try:
                    with torch.cuda.stream(stream):
                        done = sess.step(self.step_iters)
                except Exception as e:                    # pragma: no cover - surfaced to the callers
                    for (r, t) in inflight.values():
                        self._complete(r, {}, t, fail=e)
                    inflight.clear()
                    sess.close()
                    sess = None
                    continue
                bwd = []
                for tag, o in done:
                    r, t = inflight.pop(tag)
                    if r.G is None:
                        self._complete(r, self._fwd_out(o), t)
                    else:
                        bwd.append((r, o, t))
                if bwd:
                    if self._adj_worker:
                        with self._cv:
                            self._bq.extend(bwd)
                            self._cv.notify_all()
                    else:
                        self._adjoint_batch(bwd, stream)
        finally:
            if sess is not None:
                sess.close()
            with self._cv:
                self._fwd_open -= 1
                self._cv.notify_all()

    def _run_adjoints(self, k: int):
        """Adjoint worker: the backward requests' adjoints, batched as their
        forwards finish."""
        torch.cuda.set_device(self.device)
        stream = torch.cuda.Stream(self.device)          # default priority
        while True:
            with self._cv:
                while not self._bq and self._fwd_open > 0:
                    self._cv.wait()
                if not self._bq:
                    return
                jobs = [self._bq.popleft() for _ in range(min(len(self._bq), self.max_batch))]
            try:
                self._adjoint_batch(jobs, stream)
            except Exception as e:                        # pragma: no cover - surfaced to the callers
                for (r, _o, t) in jobs:
                    if not r.future.done():
                        self._complete(r, {}, t, fail=e)

    # ------------------------------------------------------------ batch mode
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
