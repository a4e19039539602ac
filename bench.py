"""Benchmark: batched SPLR-OT solves/s on one MI355X, beside one-at-a-time solves.

Workload (default): ragged batch of B=128 instances, n, m log-uniform integers
in [128, 4096], Gaussian points in R^d with d in {2, 8, 32}, M = squared
Euclidean / max, Dirichlet(1) marginals floored at 1e-4, eta in {0.01, 0.05},
FP64, tol 1e-8, default density.  A step is one batched SPLR solve of the whole
batch (forward, plans written); inputs are resident in HBM before timing.

Also measured: the same instances solved one at a time with the plain entry
(per-solve latency p50/p99); per-kernel times in one extra, single-group solve
of the batch (HIP events around each launch), which give the roofline figures
of the dominant kernel and of the HBM-bound M passes; the serving leg (the head
of the mixed-size request stream, n, m up to 8192, through the
continuous-batching server, beside one-at-a-time solves of the same
requests); and the CPU oracle on a size-stratified sample of the batch, on all
host cores and on one.

    python bench.py [--steps K] [--warmup W] [--batch B] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "batched SPLR-OT solves/s + p50/p99 latency vs sequential, FP64 tol 1e-8"
HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _omp_threads(k):
    """Set the OpenMP thread count of the oracle's parallel regions (libgomp,
    the runtime liboracle.so links)."""
    import ctypes
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(k))


def cpu_baseline(problems, budget_s, strata=4, threads=None, max_stratum=None):
    """CPU oracle (FP64 C restatement, OpenMP) on a size-stratified sample of
    the batch: the instances are split into `strata` groups of equal count by
    n*m; within each group (up to max_stratum) instances are solved in batch
    order until the group's share of the time budget is spent (at least one
    each).  The batch's CPU time is estimated as sum over groups of (group
    size x mean solve time in the group); value = B / that time."""
    import oracle
    oracle.lib()
    cores = int(threads or os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    _omp_threads(cores)
    order = sorted(range(len(problems)), key=lambda i: problems[i][0].numel())
    groups = [sorted(order[k * len(order) // strata:(k + 1) * len(order) // strata]) for k in range(strata)]
    if max_stratum is not None:
        groups = groups[:max_stratum]
    est, done, cells, t_all = 0.0, 0, 0, time.perf_counter()
    per = []
    for g in groups:
        t0 = time.perf_counter()
        k = 0
        for i in g:
            M, a, b, eta = problems[i]
            oracle.splr_bwd(M.numpy(), a.numpy(), b.numpy(), eta, tol=1e-8, max_iter=1000)
            k += 1
            cells += M.numel()
            if time.perf_counter() - t0 > budget_s / len(groups):
                break
        el = time.perf_counter() - t0
        est += len(g) * el / k
        done += k
        per.append(f"{k}/{len(g)}")
    n_inst = sum(len(g) for g in groups)
    return {"value": n_inst / est, "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"size-stratified sample of the same batch: {len(groups)} groups of equal count by n*m "
                      f"(solved/size {', '.join(per)}), {done} instances, {cells / 1e6:.1f}M cells, "
                      f"{time.perf_counter() - t_all:.1f} s; value = {n_inst} / estimated CPU time of those "
                      f"{n_inst} instances"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-sequential", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default tile rule instead of running the tile autotuner first")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--groups", type=int, default=4, help="concurrent size classes of the batched solve (1-4)")
    ap.add_argument("--newton", choices=("auto", "sparse", "dense"), default=None,
                    help="Newton solve: sparse factor by latency estimate (auto), wherever a plan fits, or dense")
    ap.add_argument("--serve", action="store_true", help="measure the serving loop on the mixed-size stream")
    ap.add_argument("--single", action="store_true", help="single n=m=2048 instance latency beside the CPU oracle")
    ap.add_argument("--lse", action="store_true",
                    help="row/column LSE and T pass in isolation on one n=m=16384 instance (2 GiB of M)")
    ap.add_argument("--adjoint", action="store_true",
                    help="n=m=4096 forward + backward times and dM vs central finite differences")
    ap.add_argument("--adjoint-sizes", action="store_true",
                    help="solve_backward wall time on the reference's published backward sizes "
                         "(100x120 .. 1000x1200, eta 0.05, density 0.10, seed 42)")
    ap.add_argument("--fd", type=int, default=64, help="entries checked by finite differences (--adjoint)")
    ap.add_argument("--fd-sweeps", type=int, default=2000, help="BCD polish sweeps per FD solve (--adjoint)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="MiB of one device-to-device copy issued first (counter calibration under rocprofv3)")
    ap.add_argument("--serve-n", type=int, default=1000,
                    help="requests replayed from the serving trace (default run: the serving leg; 0 skips it)")
    ap.add_argument("--serve-seq", type=int, default=300,
                    help="requests of the stream also solved one at a time (--serve)")
    ap.add_argument("--rate", type=float, default=2000.0, help="arrival rate of the serving trace (req/s)")
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--max-batch", type=int, default=256)
    ap.add_argument("--step-iters", type=int, default=3, help="serving: iterations between admissions")
    ap.add_argument("--serve-batch-mode", action="store_true", help="serving: admission per batch")
    args = ap.parse_args()
    if args.newton is not None:
        from curegot_backward_amd import _lib as _l
        _l.set_sparse_solve(args.newton)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (instances are independent; a multi-GPU "
              "run would shard the batch by HBM footprint). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.serve:
        serve_main(args)
        return
    if args.adjoint:
        adjoint_main(args)
        return
    if args.adjoint_sizes:
        adjoint_sizes_main(args)
        return
    if args.lse:
        lse_main(args)
        return
    if args.single:
        single_main(args)
        return

    import curegot_backward_amd as cg
    from curegot_backward_amd import _lib
    from curegot_backward_amd.workloads import ragged_batch

    dev = torch.device("cuda", 0)
    if args.copy_probe > 0:
        src = torch.ones(args.copy_probe * (1 << 17), dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe} MiB read + {args.copy_probe} MiB written")
        del src, dst
    log("generating workload")
    probs = ragged_batch(B=args.batch, seed=args.seed)
    Ms = [p[0].to(dev) for p in probs]
    As = [p[1].to(dev) for p in probs]
    Bs = [p[2].to(dev) for p in probs]
    etas = [p[3] for p in probs]
    torch.cuda.synchronize()
    cells = sum(M.numel() for M in Ms)
    log(f"B={len(Ms)} cells={cells / 1e6:.1f}M ({8 * cells / 1e9:.2f} GB of M)")

    def batched():
        return cg.sinkhorn_splr_bwd_batch(Ms, As, Bs, etas, 1e-8, 1000)

    tuned = None
    if not args.no_tune:
        # tile shapes of the M passes picked on this GPU before anything is timed
        from curegot_backward_amd import autotune
        t = time.perf_counter()
        tuned = autotune.tune(log=log)
        autotune.apply(tuned)
        tuned["tune_s"] = time.perf_counter() - t
        log(f"tile table {tuned['table']} ({tuned['tune_s']:.1f} s)")

    _lib.set_max_groups(args.groups)
    for w in range(args.warmup):
        t = time.perf_counter()
        outs = batched()
        torch.cuda.synchronize()
        log(f"warmup {w}: {time.perf_counter() - t:.3f} s")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        outs = batched()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    niters = [o["niter"] for o in outs]
    conv = sum(1 for o in outs if o["converged"])
    log(f"timed: {args.steps} steps in {el:.3f} s; converged {conv}/{len(outs)}; "
        f"niter mean {np.mean(niters):.1f} max {max(niters)}")

    # kernel figures: one more solve with HIP events around every launch, as a
    # single lockstep group so that an event interval holds only its own kernel
    _lib.set_max_groups(1)
    _lib.profile_reset()
    _lib.profile_enable(True)
    batched()
    torch.cuda.synchronize()
    _lib.profile_enable(False)
    _lib.set_max_groups(args.groups)
    prof = {k: _lib.profile_read(k) for k in _lib.PROFILE_CLASSES}
    for k, (ms, work, n) in prof.items():
        if n:
            log(f"  {k:12s} launches {n:6d}  total {ms:9.2f} ms  avg {1000 * ms / n:8.2f} us  work {work:.3e}")
    def hbm_roof(name, kernel):
        ms, byts, n = prof[name]
        ach = (byts / 1e9) / (ms / 1e3) if ms > 0 else 0.0
        return {"kernel": kernel, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": None, "avg_launch_us": 1000 * ms / max(n, 1),
                "launches": n, "bytes_per_launch": byts / max(n, 1)}

    roof_tpass = hbm_roof("tpass", "k_tpass (T pass / line-search trial over M)")
    roof_row = hbm_roof("row_lse", "k_row_lse (Sinkhorn candidate alpha half-sweep)")
    roof_col = hbm_roof("col_lse", "k_col_lse (Sinkhorn candidate beta half-sweep)")
    def mfma_roof(name, kernel, note):
        ms, fl, n = prof[name]
        ach = (fl / 1e12) / (ms / 1e3) if ms > 0 else 0.0
        return {"kernel": kernel, "bound": "mfma", "achieved": ach, "peak": FP64_MFMA_PEAK_TFS, "unit": "TFLOP/s",
                "frac": ach / FP64_MFMA_PEAK_TFS, "traffic": None, "avg_launch_us": 1000 * ms / max(n, 1),
                "launches": n, "flop_per_launch": fl / max(n, 1), "note": note}

    roof_sp = mfma_roof("sp_newton", "k_sp_newton (sparse factorization and two-RHS solve of the Newton system, "
                        "one workgroup per instance)",
                        "latency-bound: a chain of elimination-tree levels and dense-core columns per instance; "
                        "flops = periphery pair FMAs, core Schur gathers, R^3/3 core elimination, substitutions")
    roof_chol = mfma_roof("chol_update", "k_chol_update (trailing update, diagonal-block factor, next panel of the "
                          "dense Schur Cholesky)", "64x64 blocks on FP64 MFMA; 2*64^3 flops per trailing tile")
    # HBM bytes per launch measured with rocprofv3 counter passes over this
    # workload (FETCH_SIZE and WRITE_SIZE, calibrated on a copy of known size);
    # null when the measurement file is absent
    try:
        with open(os.path.join(ROOT, "profiles", "pmc_traffic_splr_ragged_b128.json")) as f:
            measured = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        measured = {}
    for roof, kname in ((roof_tpass, "k_tpass"), (roof_chol, "k_chol_update"), (roof_row, "k_row_lse"),
                        (roof_col, "k_col_lse"), (roof_sp, "k_sp_newton")):
        if kname in measured and args.batch == 128 and args.seed == 0:
            roof["traffic"] = measured[kname]["traffic_bytes_per_launch"]
            roof["traffic_source"] = "profiles/pmc_traffic_splr_ragged_b128.md"
    # the dominant kernel: the most device time in the profiled solve
    cand = [(prof["sp_newton"][0], roof_sp), (prof["chol_update"][0], roof_chol), (prof["tpass"][0], roof_tpass)]
    dominant = max(cand, key=lambda x: x[0])[1]
    dominant = dict(dominant, source="one extra solve of the same batch as a single lockstep group, HIP events "
                                     "around every launch on the solve stream (the timed steps run four "
                                     "concurrent size classes, whose kernels overlap)")

    value = len(Ms) * args.steps / el
    batch_ms = 1000 * el / args.steps
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": batch_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"ragged SPLR batch B={len(Ms)}, n,m log-uniform in [128,4096], d in {{2,8,32}}, "
                        "M=|x-y|^2/max, Dirichlet(1) marginals floored 1e-4, eta in {0.01,0.05}, tol 1e-8",
            "B": len(Ms), "tol": 1e-8, "seed": args.seed, "cells": cells, "groups": args.groups,
            "converged": conv, "niter_mean": float(np.mean(niters)), "niter_max": int(max(niters)),
        },
        "roofline": dominant,
        "roofline_hbm": {
            "tpass": roof_tpass,
            "row_lse": roof_row,
            "col_lse": roof_col,
        },
        "batched_latency_ms": {"p50": batch_ms, "p99": batch_ms,
                               "note": "one synchronous call returns every instance of the batch together"},
        "kernel_times_profiled_solve_ms": {k: v[0] for k, v in prof.items() if v[2]},
        "tile_table": ({"rows_m_max_cols": tuned["table"], "picked_by": "curegot_backward_amd.autotune.tune "
                        "(short batched solves per m band, each admissible columns-per-lane, best of 3)",
                        "bands": tuned["bands"], "tune_s": tuned["tune_s"]} if tuned else
                       {"rows_m_max_cols": [], "picked_by": "default rule (--no-tune)"}),
    }

    if not args.no_sequential:
        lat = []
        t1 = time.perf_counter()
        for i in range(len(Ms)):
            ts = time.perf_counter()
            o = cg.sinkhorn_splr_bwd(Ms[i], As[i], Bs[i], etas[i], 1e-8, 1000)
            torch.cuda.synchronize()
            lat.append(1000 * (time.perf_counter() - ts))
        seq_el = time.perf_counter() - t1
        seq = len(Ms) / seq_el
        result["sequential"] = {"solves_per_s": seq, "p50_ms": float(np.percentile(lat, 50)),
                                "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(max(lat)),
                                "total_s": seq_el}
        result["speedup_vs_sequential"] = value / seq
        log(f"sequential: {seq:.2f} solves/s, p50 {np.percentile(lat, 50):.1f} ms, "
            f"p99 {np.percentile(lat, 99):.1f} ms")

    if args.serve_n > 0:
        log(f"serving leg: {args.serve_n} requests of the mixed-size stream")
        result["served"] = serve_leg(args, seq_n=min(args.serve_n, 300))
    if args.cpu_budget > 0:
        result["cpu_baseline"] = cpu_baseline(probs, args.cpu_budget)
        log(f"cpu baseline: {result['cpu_baseline']['value']:.2f} solves/s on {result['cpu_baseline']['cores']} cores")
        one = cpu_baseline(probs, 0.5 * args.cpu_budget, threads=1, max_stratum=2)
        one["sample"] = "one thread; lower half of the batch by n*m only (two of the four groups): " + one["sample"]
        result["cpu_baseline_1core"] = one
        log(f"cpu baseline, 1 thread, lower half: {one['value']:.2f} solves/s")
    print(json.dumps(result), flush=True)


SERVE_METRIC = "served SPLR-OT solves/s + p50/p99 latency vs sequential, mixed-size Poisson stream, FP64 tol 1e-8"


def serve_leg(args, seq_n=None):
    """Replay the head of the serving trace (Poisson arrivals at args.rate
    req/s, n, m log-uniform in [64, 8192], d in {4,16,64}, Dirichlet(0.5..2)
    marginals, eta in {0.01,0.05,0.1}, 30% with an upstream gradient G) through
    the continuous-batching server, then solve the first seq_n of the same
    requests one at a time (plain entry, plus the backward where G is given)."""
    import curegot_backward_amd as cg
    from curegot_backward_amd.serving import OTServer
    from curegot_backward_amd.workloads import serving_problem, serving_trace

    dev = torch.device("cuda", 0)
    trace = serving_trace(rate=args.rate, seconds=1.5 * args.serve_n / args.rate, seed=args.seed)[: args.serve_n]
    log(f"serving sample: {len(trace)} requests over {trace[-1]['t']:.3f} s of arrivals")
    reqs = []
    g = torch.Generator().manual_seed(args.seed + 7)
    for r in trace:
        M, a, b, eta = serving_problem(r, device=dev)
        G = torch.randn(M.shape, dtype=torch.float64, generator=g).to(dev) if r["backward"] else None
        reqs.append((M, a, b, eta, G))
    torch.cuda.synchronize()
    cells = sum(q[0].numel() for q in reqs)
    log(f"cells {cells / 1e6:.1f}M, {sum(1 for q in reqs if q[4] is not None)} with backward")

    kw = dict(slots=args.slots, max_batch=args.max_batch, tol=1e-8, continuous=not args.serve_batch_mode,
              step_iters=args.step_iters)
    with OTServer(**kw) as srv:
        for q in reqs[: min(4, len(reqs))]:           # warm-up (allocator, kernels)
            srv.submit(*q).result()
    with OTServer(**kw) as srv:
        t0 = time.perf_counter()
        futs = []
        for r, q in zip(trace, reqs):
            dt = r["t"] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            futs.append(srv.submit(*q))
        outs = [f.result() for f in futs]
        el = time.perf_counter() - t0
        st = srv.stats()
        blog = srv.batch_log()
    conv = sum(1 for o in outs if o["converged"])
    log(f"served: {len(outs)} in {el:.3f} s, batches {st['batches']}, p50 {st['p50_ms']:.1f} ms, "
        f"p99 {st['p99_ms']:.1f} ms, converged {conv}")

    seq_n = len(reqs) if seq_n is None else min(seq_n, len(reqs))
    # one at a time through the public drop-in API, as a caller of the
    # reference would run them: T = splr_ot(M, a, b, eta=...) and, for a
    # request with an upstream gradient, T.backward(G) (dL/dM through the
    # implicit adjoint); splr_ot raises on a non-converged forward or backward
    # as the reference does, after the work is done
    from curegot_backward_amd.autograd import splr_ot

    def one_by_one_splr_ot():
        lat, raised = [], 0
        t1 = time.perf_counter()
        for (M, a, b, eta, G) in reqs[:seq_n]:
            ts = time.perf_counter()
            try:
                if G is None:
                    splr_ot(M, a, b, eta=eta, tol=1e-8, max_iter=1000)
                else:
                    T = splr_ot(M.detach().requires_grad_(True), a, b, eta=eta, tol=1e-8, max_iter=1000)
                    T.backward(G)
            except RuntimeError:
                raised += 1
            torch.cuda.synchronize()
            lat.append(1000 * (time.perf_counter() - ts))
        return time.perf_counter() - t1, lat, raised

    # the same requests one at a time through the lower-level native entry
    # (sinkhorn_splr_bwd + solve_backward: no autograd graph)
    def one_by_one_native():
        lat = []
        t1 = time.perf_counter()
        for (M, a, b, eta, G) in reqs[:seq_n]:
            ts = time.perf_counter()
            o = cg.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 1000, allow_invalid=True)
            if G is not None and o["context"] is not None:
                cg.solve_backward(o["plan"], G, o["context"], eta)
            torch.cuda.synchronize()
            lat.append(1000 * (time.perf_counter() - ts))
        return time.perf_counter() - t1, lat

    # order [splr_ot, native, splr_ot]: the timed splr_ot pass is the second,
    # so that both timed passes find the allocator warm for these shapes
    one_by_one_splr_ot()
    nat_el, lat_n = one_by_one_native()
    seq_el, lat, raised = one_by_one_splr_ot()
    value = len(outs) / el
    seq = seq_n / seq_el
    nat = seq_n / nat_el
    log(f"sequential splr_ot (first {seq_n}): {seq:.2f} solves/s ({raised} raised non-convergence); "
        f"served/sequential {value / seq:.2f}; native entry one at a time {nat:.2f} solves/s "
        f"(served/native {value / nat:.2f})")
    del reqs
    torch.cuda.empty_cache()
    return {
        "value": value,
        "unit": "solves/s",
        "workload": f"head of the serving trace: {len(outs)} requests, Poisson {args.rate:g} req/s, "
                    "n,m log-uniform in [64,8192], d in {4,16,64}, Dirichlet(0.5..2), eta in "
                    "{0.01,0.05,0.1}, 30% forward+backward (G ~ N(0,1)), tol 1e-8",
        "requests": len(outs), "arrival_span_s": trace[-1]["t"], "makespan_s": el, "cells": cells,
        "slots": args.slots, "max_batch": args.max_batch, "converged": conv, "batches": st["batches"],
        "admission": "per batch" if args.serve_batch_mode else f"continuous (every {args.step_iters} iterations)",
        "latency_ms": {"p50": st["p50_ms"], "p99": st["p99_ms"], "max": st["max_ms"],
                       "queue_p50": st["queue_p50_ms"], "queue_p99": st["queue_p99_ms"]},
        "sequential": {"solves_per_s": seq, "requests": seq_n, "p50_ms": float(np.percentile(lat, 50)),
                       "p99_ms": float(np.percentile(lat, 99)), "total_s": seq_el, "raised": raised,
                       "note": f"the first {seq_n} requests of the same stream, one at a time through plain "
                               "splr_ot (T.backward(G) where the request has an upstream gradient)"},
        "sequential_native": {"solves_per_s": nat, "requests": seq_n, "p50_ms": float(np.percentile(lat_n, 50)),
                              "p99_ms": float(np.percentile(lat_n, 99)), "total_s": nat_el,
                              "note": "the same requests one at a time through sinkhorn_splr_bwd + "
                                      "solve_backward (no autograd)"},
        "speedup_vs_sequential": value / seq,
        "speedup_vs_native_one_at_a_time": value / nat,
        "batch_log_tail": blog[-5:],
    }


def serve_main(args):
    r = serve_leg(args, seq_n=args.serve_seq)
    result = {
        "metric": SERVE_METRIC, "value": r["value"], "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": r["workload"], "requests": r["requests"], "slots": r["slots"],
                   "max_batch": r["max_batch"], "seed": args.seed, "converged": r["converged"]},
        "served_latency_ms": r["latency_ms"], "batches": r["batches"], "sequential": r["sequential"],
        "sequential_native": r["sequential_native"],
        "speedup_vs_sequential": r["speedup_vs_sequential"],
        "speedup_vs_native_one_at_a_time": r["speedup_vs_native_one_at_a_time"],
    }
    print(json.dumps(result), flush=True)


SINGLE_METRIC = "single-instance FP64 SPLR forward latency, n=m=2048, density 0.01, tol 1e-8"


def single_main(args):
    """n=m=2048, X, Y ~ N(0, I_16), M = squared Euclidean / max, Dirichlet(1)
    marginals mixed with 1e-3 uniform, eta = 0.01, density = 0.01 (global top-K
    maximum), tol 1e-8: solve latency on the GPU beside the CPU oracle."""
    import curegot_backward_amd as cg
    from curegot_backward_amd import _lib
    from curegot_backward_amd.workloads import single_problem

    dev = torch.device("cuda", 0)
    Mc, ac, bc, eta = single_problem(seed=args.seed)
    M, a, b = Mc.to(dev), ac.to(dev), bc.to(dev)
    solve = lambda: cg.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 1000, density=0.01)
    solve()
    torch.cuda.synchronize()
    lat = []
    for _ in range(max(args.steps, 3)):
        t0 = time.perf_counter()
        out = solve()
        torch.cuda.synchronize()
        lat.append(1000 * (time.perf_counter() - t0))
    _lib.profile_reset()
    _lib.profile_enable(True)
    solve()
    torch.cuda.synchronize()
    _lib.profile_enable(False)
    ms_t, bytes_t, n_t = _lib.profile_read("tpass")
    ach = (bytes_t / 1e9) / (ms_t / 1e3) if ms_t > 0 else 0.0
    log(f"GPU: niter {out['niter']}, p50 {np.median(lat):.1f} ms")
    cpu = None
    if args.cpu_budget > 0:
        import oracle
        oracle.lib()
        t0 = time.perf_counter()
        ref = oracle.splr_bwd(Mc.numpy(), ac.numpy(), bc.numpy(), eta, tol=1e-8, max_iter=1000, density=0.01)
        el = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
        diff = float((out["plan"].cpu() - torch.from_numpy(ref["plan"])).abs().max() / abs(ref["plan"]).max())
        cpu = {"value": 1000 * el, "unit": "ms", "cores": cores, "kind": "port", "niter": int(ref["niter"]),
               "plan_rel_diff_vs_gpu": diff, "sample": "the same instance, one solve"}
        log(f"CPU oracle: {el:.2f} s, niter {ref['niter']}, plan rel diff {diff:.2e}")
    result = {
        "metric": SINGLE_METRIC, "value": float(np.median(lat)), "unit": "ms", "n_gpus": 1,
        "higher_is_better": False, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "n=m=2048, d=16, M=|x-y|^2/max, Dirichlet(1)*(1-1e-3)+1e-3/n marginals, "
                               "eta=0.01, density=0.01 (global top-K), tol 1e-8", "seed": args.seed},
        "niter": out["niter"], "converged": out["converged"], "latency_ms": {"p50": float(np.median(lat)),
                                                                             "min": float(min(lat))},
        "tpass_hbm": {"achieved": ach, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                      "note": "M is 32 MiB and stays in the 256 MB last-level cache between passes"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)


LSE_METRIC = "row/column log-sum-exp HBM throughput in isolation, n=m=16384 FP64 (2 GiB of M)"


def adjoint_sizes_main(args):
    """solve_backward on the problems of the reference's backward benchmark
    (reference benchmark_stage3c3.py:175-230 with stage3b_validate.py:61-69):
    M ~ U[0,1), a, b ~ U[0,1) + 0.5 normalised (CPU generator, seed 42), eta
    0.05, forward tol 1e-8 at density 0.10, G ~ N(0,1) seeded per size
    ((1000003 n + 10007 m + 7) mod (2^31 - 1), benchmark_stage3c2c.py:109-112),
    PCG tol 1e-12.  Wall time of one solve_backward call (synchronised), the
    median of 20 after 3 warm-ups, per preconditioner.  The reference's table
    (STAGE3C_FINAL.md:70-75) times its Python PCG loop on an unnamed GPU."""
    import curegot_backward_amd as cg

    dev = torch.device("cuda", 0)
    eta = 0.05
    published = {(100, 120): 12.03, (300, 360): 9.57, (500, 600): 8.69, (1000, 1200): 7.30}
    rows = []
    for (n, m), ref_ms in published.items():
        gen = torch.Generator(device="cpu").manual_seed(42)
        M = torch.rand(n, m, generator=gen, dtype=torch.float64)
        a = torch.rand(n, generator=gen, dtype=torch.float64) + 0.5
        b = torch.rand(m, generator=gen, dtype=torch.float64) + 0.5
        M, a, b = M.to(dev), (a / a.sum()).to(dev), (b / b.sum()).to(dev)
        g = torch.Generator(device="cpu").manual_seed((1000003 * n + 10007 * m + 7) % (2 ** 31 - 1))
        G = torch.randn((n, m), generator=g, dtype=torch.float64).to(dev)
        out = cg.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 1000, density=0.10)
        row = {"n": n, "m": m, "forward_niter": int(out["niter"]), "reference_diag_ms": ref_ms}
        for pc in ("diag", "none", "splr"):
            ts = []
            for k in range(23):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                r = cg.solve_backward(out["plan"], G, out["context"], eta, preconditioner=pc)
                torch.cuda.synchronize()
                if k >= 3:
                    ts.append(1000 * (time.perf_counter() - t0))
            row[pc] = {"ms": float(np.median(ts)), "iterations": int(r["iterations"]),
                       "converged": bool(r["converged"])}
        log(f"{n}x{m}: forward niter {row['forward_niter']}; backward diag {row['diag']['ms']:.3f} ms "
            f"({row['diag']['iterations']} it), none {row['none']['ms']:.3f} ({row['none']['iterations']}), "
            f"splr {row['splr']['ms']:.3f} ({row['splr']['iterations']}); reference diag {ref_ms} ms")
        rows.append(row)
    print(json.dumps({"metric": "solve_backward wall ms, reference backward benchmark sizes", "unit": "ms",
                      "higher_is_better": False, "dtype": "f64", "data": "synthetic", "rows": rows}), flush=True)


def lse_main(args):
    """The library's LSE half-sweeps (k_row_lse + k_rowlse_fin, k_col_lse +
    k_collse_fin) and one T pass, each on one n = m = 16384 instance: M is
    2 GiB, far past the 256 MB MALL, so every pass streams M from HBM.
    Per-launch HIP events on the solve stream; algorithmic bytes as in the
    batch bench (LSE: 8nm + 8(n+m) + 8 max(n,m); T pass at a fixed point: 8nm + 16(n+m))."""
    import curegot_backward_amd as cg
    from curegot_backward_amd import _lib, _native

    dev = torch.device("cuda", 0)
    n = m = 16384
    g = torch.Generator(device="cpu").manual_seed(args.seed)
    X = torch.randn(n, 16, generator=g, dtype=torch.float64).to(dev)
    Y = torch.randn(m, 16, generator=g, dtype=torch.float64).to(dev)
    M = torch.cdist(X, Y).pow_(2)
    M /= M.max()
    a = torch.full((n,), 1.0 / n, dtype=torch.float64, device=dev)
    b = torch.full((m,), 1.0 / m, dtype=torch.float64, device=dev)
    reg = 0.05
    dual = torch.zeros(n + m, dtype=torch.float64, device=dev)
    reps = max(args.steps, 2)
    res = {}
    for which, name in ((0, "row_lse"), (1, "col_lse")):
        _native.lse_update(M, a, b, dual, reg, which, 2)          # warm-up
        torch.cuda.synchronize()
        _lib.profile_reset()
        _lib.profile_enable(True)
        _native.lse_update(M, a, b, dual, reg, which, reps)
        torch.cuda.synchronize()
        _lib.profile_enable(False)
        ms, work, nl = _lib.profile_read(name)
        ach = work / ms / 1e6
        res[name] = {"kernel": f"k_{name}", "launches": nl, "avg_launch_us": 1000 * ms / nl, "achieved": ach,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS}
        log(f"{name}: {nl} launches, {1000 * ms / nl:.1f} us avg, {ach:.0f} GB/s ({ach / HBM_PEAK_GBS:.3f})")
    alpha, beta = dual[:n].clone(), dual[n:].clone()
    beta -= beta[-1].item()
    _native.T_sparsify(alpha, beta, M, a, b, reg, 0.0, 1000)          # warm-up
    torch.cuda.synchronize()
    _lib.profile_reset()
    _lib.profile_enable(True)
    for _ in range(3):
        _native.T_sparsify(alpha, beta, M, a, b, reg, 0.0, 1000)
    torch.cuda.synchronize()
    _lib.profile_enable(False)
    ms, work, nl = _lib.profile_read("tpass")
    if nl:
        ach = work / ms / 1e6
        res["tpass"] = {"kernel": "k_tpass", "launches": nl, "avg_launch_us": 1000 * ms / nl, "achieved": ach,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS}
        log(f"tpass: {nl} launches, {1000 * ms / nl:.1f} us avg, {ach:.0f} GB/s ({ach / HBM_PEAK_GBS:.3f})")
    result = {
        "metric": LSE_METRIC, "value": res["row_lse"]["achieved"], "unit": "GB/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "n=m=16384, d=16 N(0,I) points, M=|x-y|^2/max, uniform marginals, reg 0.05",
                   "seed": args.seed, "repeats": reps},
        "roofline_hbm": res,
    }
    print(json.dumps(result), flush=True)


ADJ_METRIC = "FP64 splr_ot forward + implicit backward time and dM vs central FD, n=m=4096"


def adjoint_main(args):
    """n=m=4096, d=64 Gaussian points, M = mean squared distance, uniform
    marginals, eta=0.05, loss = sum(T^2) + <T,R>: forward and backward times
    through the autograd op, the backward with each preconditioner, and the
    gradient against central finite differences (h = 1e-6) on sampled entries.
    Every FD plan is the SPLR solution (warm-started at the base duals) polished
    by a fixed number of Sinkhorn/BCD sweeps; the analytic gradient is taken at
    the polished base duals.  Reported for h = 1e-6 (the specified step) and
    h = 1e-4."""
    import curegot_backward_amd as cg
    from curegot_backward_amd import _lib
    from curegot_backward_amd.workloads import adjoint_problem

    dev = torch.device("cuda", 0)
    M, a, b, eta, R = adjoint_problem(seed=args.seed)
    M, a, b, R = M.to(dev), a.to(dev), b.to(dev), R.to(dev)
    n, m = M.shape
    loss = lambda T: (T ** 2).sum() + (T * R).sum()

    def fwd_bwd():
        Mg = M.clone().requires_grad_(True)
        t0 = time.perf_counter()
        T = cg.torch.splr_ot(Mg, a, b, eta=eta, tol=1e-8)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss(T).backward()
        torch.cuda.synchronize()
        return t1 - t0, time.perf_counter() - t1, Mg.grad

    fwd_bwd()                                   # warm-up
    tf, tb, grad_autograd = fwd_bwd()
    # the autograd op's overhead over the native forward it wraps (the
    # reference's target is < 5%: benchmark_stage4_autograd.py:37-42)
    t_nat, t_ag = [], []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = cg.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 1000)
        torch.cuda.synchronize()
        t_nat.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        cg.torch.splr_ot(M, a, b, eta=eta, tol=1e-8)
        torch.cuda.synchronize()
        t_ag.append(time.perf_counter() - t0)
    ag_overhead = min(t_ag) / min(t_nat) - 1.0
    G = 2 * out["plan"] + R
    pre = {}
    for name in ("diag", "splr", "none"):
        ts = []
        for _ in range(3):          # the best of three (the first also creates streams and graphs)
            t0 = time.perf_counter()
            r = cg.solve_backward(out["plan"], G, out["context"], eta, preconditioner=name)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        pre[name] = {"ms": 1000 * min(ts), "iterations": r["iterations"], "converged": r["converged"]}
    log(f"forward {1000 * tf:.1f} ms (niter {out['niter']}), backward {1000 * tb:.1f} ms; preconditioners {pre}")
    # adjoint pass throughput: the diag PCG's k_qv_pcg launches, HIP events on the solve stream
    _lib.profile_reset()
    _lib.profile_enable(True)
    cg.solve_backward(out["plan"], G, out["context"], eta, preconditioner="diag")
    torch.cuda.synchronize()
    _lib.profile_enable(False)
    qv_ms, qv_bytes, qv_n = _lib.profile_read("qv")
    qv_roof = {"kernel": "k_qv_pcg", "launches": qv_n, "avg_launch_us": 1000 * qv_ms / max(qv_n, 1),
               "achieved": qv_bytes / qv_ms / 1e6 if qv_ms > 0 else 0.0, "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "note": "M is 128 MiB: it stays in the 256 MB MALL across PCG iterations"}
    qv_roof["frac"] = qv_roof["achieved"] / HBM_PEAK_GBS
    log(f"k_qv_pcg: {qv_n} launches, {qv_roof['avg_launch_us']:.1f} us avg, {qv_roof['achieved']:.0f} GB/s")

    def polished(Mx, x0):
        # a fixed number of BCD sweeps past the SPLR solution: the marginals are
        # ~1/n, so the plan is taken to the rounding floor (tolerance below it)
        d = cg.sinkhorn_splr_bwd(Mx, a, b, eta, 1e-8, 1000, x0=x0)["dual"]
        return cg.sinkhorn_bcd(Mx, a, b, eta, 1e-30, args.fd_sweeps, x0=d)

    base = polished(M, out["dual"])
    dual = base["dual"] - base["dual"][-1] * torch.cat([-torch.ones(n), torch.ones(m)]).to(dev)
    ctx_out = cg.sinkhorn_splr_bwd(M, a, b, eta, 1e-8, 10, x0=dual)
    g = cg.solve_backward(ctx_out["plan"], 2 * ctx_out["plan"] + R, ctx_out["context"], eta, tol=1e-14)["grad"]
    rng = np.random.default_rng(args.seed + 1)
    idx = [(int(rng.integers(n)), int(rng.integers(m))) for _ in range(args.fd)]   # --fd 0: no FD check
    gs = np.array([float(g[i, j]) for (i, j) in idx])
    fd_stats = {}
    t0 = time.perf_counter()
    for h in ((1e-6, 1e-4) if idx else ()):
        fds = []
        for k, (i, j) in enumerate(idx):
            Mp, Mm = M.clone(), M.clone()
            Mp[i, j] += h
            Mm[i, j] -= h
            fds.append((float(loss(polished(Mp, dual)["plan"])) - float(loss(polished(Mm, dual)["plan"]))) / (2 * h))
            if k % 8 == 7:
                log(f"  fd h={h:g} {k + 1}/{len(idx)} ({time.perf_counter() - t0:.1f} s)")
        err = np.abs(np.array(fds) - gs)
        fd_stats[f"h={h:g}"] = {"max_abs_err": float(err.max()),
                                "max_rel_err_normwise": float(err.max() / np.abs(gs).max()),
                                "max_rel_err_entrywise": float((err / np.maximum(np.abs(gs), 1e-300)).max())}
    result = {
        "metric": ADJ_METRIC,
        "value": 1000 * (tf + tb),
        "unit": "ms",
        "n_gpus": 1,
        "higher_is_better": False,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "n=m=4096, d=64 N(0,I) points, M=mean squared distance, uniform marginals, "
                               "eta=0.05, loss=sum(T^2)+<T,R>, R~N(0,1), tol 1e-8", "seed": args.seed},
        "forward_ms": 1000 * tf, "backward_ms": 1000 * tb, "forward_niter": out["niter"],
        "autograd_forward_overhead": ag_overhead,
        "backward_preconditioners": pre,
        "roofline_qv": qv_roof,
        "autograd_vs_manual_grad_rel": float((grad_autograd - cg.solve_backward(
            out["plan"], G, out["context"], eta)["grad"]).abs().max() / grad_autograd.abs().max()),
        "fd": {"entries": len(idx), "bcd_sweeps": args.fd_sweeps,
               "max_abs_grad": float(np.abs(gs).max()) if idx else None, **fd_stats},
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
