"""The HIP library loads and exports every symbol include/curegot_amd.h declares
(no compute calls: there is no GPU here)."""
import ctypes
import os
import re

from curegot_backward_amd import _lib
from curegot_backward_amd.workloads import ragged_sizes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "curegot_amd.h")).read()
    return sorted(set(re.findall(r"\b(cgot_[a-zA-Z0-9_]+)\s*\(", text)))


def test_exports_match_header():
    L = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
    assert set(_lib.EXPORTS) <= set(syms)
    assert L.cgot_version().decode().startswith("curegot_amd")


def test_workspace_size_is_host_only():
    L = _lib.load()
    sizes = ragged_sizes(8)
    probs = (_lib.SplrProblem * len(sizes))()
    for i, (n, m) in enumerate(sizes):
        probs[i].M = probs[i].a = probs[i].b = 16
        probs[i].n, probs[i].m, probs[i].reg, probs[i].tol, probs[i].max_iter = n, m, 0.05, 1e-8, 100
        probs[i].density_max = -1.0
        probs[i].shift_max = -1.0
    ws = L.cgot_splr_workspace_size(len(sizes), probs)
    # at least the dense Schur blocks
    assert ws > sum(8 * min(n, m - 1) ** 2 for n, m in sizes)
    probs[0].n = 0
    assert L.cgot_splr_workspace_size(len(sizes), probs) == 0
    assert b"n and m" in L.cgot_last_error()


def test_serving_footprint_host_only():
    """The serving admission estimate uses the host-only workspace query."""
    from curegot_backward_amd.serving import splr_footprint
    small = splr_footprint(100, 120)
    big = splr_footprint(1000, 1200)
    assert 0 < small < big
    assert splr_footprint(100, 120, backward=True) > small
    # dense Schur block dominates at scale: at least P^2 doubles
    assert big >= 8 * 999 * 999


def test_ctypes_structs_match_header(tmp_path):
    """Size and field offsets of every problem/result struct of the C header,
    as gcc lays them out, against the ctypes mirrors the package passes."""
    import ctypes
    import subprocess

    from curegot_backward_amd import _lib
    structs = {
        "cgot_cost_xy": (_lib.CostXY, ["X", "Y", "d", "scale"]),
        "cgot_splr_problem": (_lib.SplrProblem, ["M", "a", "b", "n", "m", "reg", "tol", "max_iter", "density_max",
                                                 "shift_max", "sparsity_pattern_cycle", "candidate_sinkhorn_iter",
                                                 "x0", "plan", "dual", "trace", "xy"]),
        "cgot_splr_info": (_lib.SplrInfo, ["niter", "converged", "valid", "low_rank", "fallback", "n", "m", "ys",
                                           "f", "gnorm", "K", "inv_diag", "s", "y", "u", "gamma", "factor"]),
        "cgot_adjoint_problem": (_lib.AdjointProblem, ["M", "dual", "G", "n", "m", "reg", "inv_diag", "tol",
                                                       "max_iter", "lambda", "grad", "factor", "xy"]),
        "cgot_adjoint_info": (_lib.AdjointInfo, ["iterations", "converged", "status"]),
        "cgot_bcd_problem": (_lib.BcdProblem, ["M", "a", "b", "n", "m", "reg", "tol", "max_iter", "x0", "plan",
                                               "dual", "xy"]),
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{os.path.join(ROOT, "include", "curegot_amd.h")}"',
             'int main(void) {']
    for cname, (_, fields) in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append('return 0; }')
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                           check=True).stdout.split("\n") if l)
    for cname, (cls, fields) in structs.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for f, cf in zip(fields, [x[0] for x in cls._fields_]):
            assert int(got[f"{cname}.{f}"]) == getattr(cls, cf).offset, (cname, f, cf)
