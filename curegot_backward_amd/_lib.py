"""ctypes binding of libcuregot_amd.so (the C ABI declared in include/curegot_amd.h).

The library is the only compute path: if it is missing or fails to load,
importing the solver raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os

# torch first: its bundled HIP runtime (soname libamdhip64.so.7) is then the
# one the library binds to.  Loading the library first would bring in the
# system runtime under that soname and torch would load a second runtime
# beside it, and the two do not share devices.
import torch  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcuregot_amd.so")

_dp = ctypes.c_void_p   # device pointers are passed as integers


class CostXY(ctypes.Structure):
    """Geometric cost M_ij = scale * sum_k (X_ik - Y_jk)^2 (X = NULL: stored M)."""
    _fields_ = [("X", _dp), ("Y", _dp), ("d", ctypes.c_int32), ("scale", ctypes.c_double)]


class SplrProblem(ctypes.Structure):
    _fields_ = [
        ("M", _dp), ("a", _dp), ("b", _dp),
        ("n", ctypes.c_int32), ("m", ctypes.c_int32),
        ("reg", ctypes.c_double), ("tol", ctypes.c_double),
        ("max_iter", ctypes.c_int32),
        ("density_max", ctypes.c_double), ("shift_max", ctypes.c_double),
        ("sparsity_pattern_cycle", ctypes.c_int32), ("candidate_sinkhorn_iter", ctypes.c_int32),
        ("x0", _dp), ("plan", _dp), ("dual", _dp), ("trace", _dp),
        ("xy", CostXY),
    ]


class SplrInfo(ctypes.Structure):
    _fields_ = [
        ("niter", ctypes.c_int32), ("converged", ctypes.c_int32), ("valid", ctypes.c_int32),
        ("low_rank", ctypes.c_int32), ("fallback", ctypes.c_int32),
        ("n", ctypes.c_int32), ("m", ctypes.c_int32),
        ("ys", ctypes.c_double), ("f", ctypes.c_double), ("gnorm", ctypes.c_double),
        ("K", ctypes.c_int64),
        ("inv_diag", _dp), ("s", _dp), ("y", _dp), ("u", _dp), ("gamma", _dp),
        ("factor", _dp),
    ]


class AdjointProblem(ctypes.Structure):
    _fields_ = [
        ("M", _dp), ("dual", _dp), ("G", _dp),
        ("n", ctypes.c_int32), ("m", ctypes.c_int32),
        ("reg", ctypes.c_double),
        ("inv_diag", _dp),
        ("tol", ctypes.c_double), ("max_iter", ctypes.c_int32),
        ("lam", _dp), ("grad", _dp),
        ("factor", _dp),
        ("xy", CostXY),
    ]


class AdjointInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("status", ctypes.c_int32)]


class BcdProblem(ctypes.Structure):
    _fields_ = [
        ("M", _dp), ("a", _dp), ("b", _dp),
        ("n", ctypes.c_int32), ("m", ctypes.c_int32),
        ("reg", ctypes.c_double), ("tol", ctypes.c_double),
        ("max_iter", ctypes.c_int32),
        ("x0", _dp), ("plan", _dp), ("dual", _dp),
        ("xy", CostXY),
    ]


class SpHdr(ctypes.Structure):
    """Header of a sparse Newton-system plan (csrc/cgot_symbolic.h)."""
    _fields_ = [(k, ctypes.c_int32) for k in ("mode", "N", "n", "m", "t", "R", "H", "nnzP", "ncore", "T")] + \
               [(k, ctypes.c_int64) for k in ("npairs", "ncpairs", "nrowpairs")] + \
               [(k, ctypes.c_int32) for k in ("nlong", "nclong", "nylong", "nblong")] + \
               [(k, ctypes.c_int64) for k in ("o_perm", "o_lvptr", "o_lcptr",
                                              "o_lcols", "o_emeta", "o_erow", "o_pairs", "o_ymeta", "o_ycore",
                                              "o_rowpairs", "o_cmeta", "o_cpairs", "o_lgptr", "o_long", "o_clong",
                                              "o_ylong", "o_blptr", "o_blong", "o_cflags", "o_meta_end", "o_lval",
                                              "o_lcore", "o_yx", "bytes")]


class PrecondProfile(ctypes.Structure):
    """cgot_precond_profile (microseconds)."""
    _fields_ = [(name, ctypes.c_double) for name in (
        "host_total_us", "host_memcpy_in_us", "host_solve_launch_us", "host_lowrank_us", "host_sync_wait_us",
        "host_memcpy_out_us", "gpu_solve_us", "gpu_lowrank_reduce_us", "gpu_lowrank_correction_us",
        "repeated_total_us")]


EXPORTS = (
    "cgot_splr_workspace_size", "cgot_splr_solve", "cgot_set_max_groups", "cgot_set_sparse_solve",
    "cgot_set_tile_table", "cgot_get_tile_table", "cgot_tile_cols",
    "cgot_adjoint_workspace_size", "cgot_adjoint_solve",
    "cgot_splr_precond_workspace_size", "cgot_splr_precond_apply", "cgot_splr_precond_apply_profiled",
    "cgot_splr_precond_apply_noalloc",
    "cgot_splr_context_csr",
    "cgot_bcd_workspace_size", "cgot_bcd_solve",
    "cgot_lse_workspace_size", "cgot_lse_update",
    "cgot_T_sparsify_workspace_size", "cgot_T_sparsify",
    "cgot_profile_enable", "cgot_profile_reset", "cgot_profile_read",
    "cgot_sparse_plan",
    "cgot_session_workspace_size", "cgot_session_inst_workspace_size", "cgot_session_inst_tiles",
    "cgot_session_create", "cgot_session_add", "cgot_session_step", "cgot_session_live", "cgot_session_destroy",
    "cgot_shutdown", "cgot_last_error", "cgot_version",
)

_lib = None


def load():
    """Load libcuregot_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"curegot_backward_amd: {LIB_PATH} not found; build it with "
            "`python curegot_backward_amd/build.py` (hipcc, gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    sz, i32, i64, dbl, vp = ctypes.c_size_t, ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
    L.cgot_splr_workspace_size.argtypes = [i32, ctypes.POINTER(SplrProblem)]
    L.cgot_splr_workspace_size.restype = sz
    L.cgot_splr_solve.argtypes = [i32, ctypes.POINTER(SplrProblem), vp, sz, vp, ctypes.POINTER(SplrInfo)]
    L.cgot_splr_solve.restype = ctypes.c_int
    L.cgot_set_sparse_solve.argtypes = [i32]
    L.cgot_set_sparse_solve.restype = ctypes.c_int
    L.cgot_set_max_groups.argtypes = [i32]
    L.cgot_set_max_groups.restype = ctypes.c_int
    L.cgot_set_tile_table.argtypes = [ctypes.POINTER(i32), ctypes.POINTER(i32), i32]
    L.cgot_set_tile_table.restype = ctypes.c_int
    L.cgot_get_tile_table.argtypes = [ctypes.POINTER(i32), ctypes.POINTER(i32), i32]
    L.cgot_get_tile_table.restype = ctypes.c_int
    L.cgot_tile_cols.argtypes = [i32]
    L.cgot_tile_cols.restype = ctypes.c_int
    L.cgot_adjoint_workspace_size.argtypes = [i32, ctypes.POINTER(AdjointProblem)]
    L.cgot_adjoint_workspace_size.restype = sz
    L.cgot_adjoint_solve.argtypes = [i32, ctypes.POINTER(AdjointProblem), vp, sz, vp,
                                     ctypes.POINTER(AdjointInfo)]
    L.cgot_adjoint_solve.restype = ctypes.c_int
    L.cgot_splr_precond_workspace_size.argtypes = [vp, i32, i32]
    L.cgot_splr_precond_workspace_size.restype = sz
    L.cgot_splr_precond_apply.argtypes = [vp, i32, i32, vp, vp, vp, sz, vp]
    L.cgot_splr_precond_apply.restype = ctypes.c_int
    L.cgot_splr_precond_apply_noalloc.argtypes = [vp, i32, i32, vp, vp, vp]
    L.cgot_splr_precond_apply_noalloc.restype = ctypes.c_int
    L.cgot_splr_precond_apply_profiled.argtypes = [vp, i32, i32, vp, vp, i32, i32, ctypes.POINTER(PrecondProfile),
                                                   vp, sz, vp]
    L.cgot_splr_precond_apply_profiled.restype = ctypes.c_int
    L.cgot_splr_context_csr.argtypes = [vp, i32, i32, vp, vp, vp, ctypes.POINTER(i64), vp]
    L.cgot_splr_context_csr.restype = ctypes.c_int
    L.cgot_bcd_workspace_size.argtypes = [i32, ctypes.POINTER(BcdProblem)]
    L.cgot_bcd_workspace_size.restype = sz
    L.cgot_bcd_solve.argtypes = [i32, ctypes.POINTER(BcdProblem), vp, sz, vp, ctypes.POINTER(i32)]
    L.cgot_bcd_solve.restype = ctypes.c_int
    L.cgot_lse_workspace_size.argtypes = [i32, i32]
    L.cgot_lse_workspace_size.restype = sz
    L.cgot_lse_update.argtypes = [vp, vp, vp, i32, i32, dbl, i32, i32, vp, vp, sz, vp]
    L.cgot_lse_update.restype = ctypes.c_int
    L.cgot_T_sparsify_workspace_size.argtypes = [i32, i32, i64]
    L.cgot_T_sparsify_workspace_size.restype = sz
    L.cgot_T_sparsify.argtypes = [vp, vp, vp, vp, i32, i32, dbl, dbl, i64, vp, vp, vp, vp, vp, vp, vp,
                                  vp, vp, vp, sz, vp]
    L.cgot_T_sparsify.restype = ctypes.c_int
    L.cgot_profile_enable.argtypes = [i32]
    L.cgot_profile_enable.restype = None
    L.cgot_profile_reset.argtypes = []
    L.cgot_profile_reset.restype = None
    L.cgot_profile_read.argtypes = [i32, ctypes.POINTER(dbl), ctypes.POINTER(dbl), ctypes.POINTER(i64)]
    L.cgot_profile_read.restype = ctypes.c_int
    L.cgot_sparse_plan.argtypes = [i32, i32, vp, vp, i64, i32, vp, vp, i64]
    L.cgot_sparse_plan.restype = i64
    L.cgot_session_workspace_size.argtypes = [i32, i64, i64, i32]
    L.cgot_session_workspace_size.restype = sz
    L.cgot_session_inst_workspace_size.argtypes = [ctypes.POINTER(SplrProblem)]
    L.cgot_session_inst_workspace_size.restype = sz
    L.cgot_session_inst_tiles.argtypes = [i32, i32]
    L.cgot_session_inst_tiles.restype = i64
    L.cgot_session_create.argtypes = [i32, i64, i64, i32, vp, sz, vp, ctypes.POINTER(vp)]
    L.cgot_session_create.restype = ctypes.c_int
    L.cgot_session_add.argtypes = [vp, ctypes.POINTER(SplrProblem), vp, sz, i64]
    L.cgot_session_add.restype = ctypes.c_int
    L.cgot_session_step.argtypes = [vp, i32, ctypes.POINTER(i64), ctypes.POINTER(SplrInfo), i32, ctypes.POINTER(i32)]
    L.cgot_session_step.restype = ctypes.c_int
    L.cgot_session_live.argtypes = [vp]
    L.cgot_session_live.restype = ctypes.c_int
    L.cgot_session_destroy.argtypes = [vp]
    L.cgot_session_destroy.restype = ctypes.c_int
    L.cgot_shutdown.argtypes = []
    L.cgot_shutdown.restype = ctypes.c_int
    L.cgot_last_error.restype = ctypes.c_char_p
    L.cgot_version.restype = ctypes.c_char_p
    _lib = L
    path = os.environ.get("CUREGOT_AMD_TILE_TABLE")
    if path:
        # a saved autotuner result (JSON with a "table" of [m_max, c] rows)
        import json
        with open(path) as f:
            set_tile_table(json.load(f)["table"])
    return L


def check(rc: int, what: str):
    if rc != 0:
        msg = load().cgot_last_error().decode(errors="replace")
        if rc == -2:
            raise ValueError(f"{what}: {msg}")
        raise RuntimeError(f"{what}: {msg}")


PROFILE_CLASSES = {"tpass": 0, "row_lse": 1, "col_lse": 2, "qv": 3, "chol_update": 4,
                   "chol_panel": 5, "trsv": 6, "sp_newton": 7}


def set_max_groups(groups: int):
    """Bound the number of concurrent size classes a batched SPLR solve issued
    from the calling thread is split into (1..4; 1 = one lockstep group on the
    caller's stream)."""
    check(load().cgot_set_max_groups(int(groups)), "set_max_groups")


def set_tile_table(rows):
    """Columns per lane of the M-pass tiles by m, process-wide: rows
    [(m_max, c), ...] with ascending m_max and c in {1, 2, 4}; row k covers the
    m <= m_max not covered before it, larger m take the default rule.  An
    empty list restores the default everywhere."""
    rows = [(int(mx), int(c)) for mx, c in rows]
    arr_m = (ctypes.c_int32 * max(1, len(rows)))(*[r[0] for r in rows])
    arr_c = (ctypes.c_int32 * max(1, len(rows)))(*[r[1] for r in rows])
    check(load().cgot_set_tile_table(arr_m, arr_c, len(rows)), "set_tile_table")


def get_tile_table():
    """The current tile table as [(m_max, c), ...] (empty: the default rule)."""
    L = load()
    cnt = L.cgot_get_tile_table(None, None, 0)
    arr_m = (ctypes.c_int32 * max(1, cnt))()
    arr_c = (ctypes.c_int32 * max(1, cnt))()
    L.cgot_get_tile_table(arr_m, arr_c, cnt)
    return [(int(arr_m[k]), int(arr_c[k])) for k in range(cnt)]


def tile_cols(m: int) -> int:
    """Columns per lane an instance with m columns is tiled with now."""
    c = load().cgot_tile_cols(int(m))
    if c < 1:
        raise ValueError("m must be >= 1")
    return c


SPARSE_MODES = {"dense": 0, "auto": 1, "sparse": 2}


def set_sparse_solve(mode) -> str:
    """How the Newton system is factored (process-wide): "dense" (Schur
    Cholesky for every instance), "sparse" (the sparse factorization wherever
    the pattern admits one) or "auto" (sparse where its latency estimate is
    the lower); True / False mean "sparse" / "dense".  Returns the previous
    mode's name."""
    if mode is True or mode is False:
        mode = "sparse" if mode else "dense"
    if mode not in SPARSE_MODES:
        raise ValueError(f"sparse solve mode {mode!r} is not one of {sorted(SPARSE_MODES)}")
    prev = load().cgot_set_sparse_solve(SPARSE_MODES[mode])
    return {v: k for k, v in SPARSE_MODES.items()}[prev]


def profile_enable(on: bool):
    load().cgot_profile_enable(1 if on else 0)


def profile_reset():
    load().cgot_profile_reset()


def profile_read(name: str):
    """(total ms, total algorithmic work, launches) of a kernel class."""
    ms, work, n = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int64(0)
    check(load().cgot_profile_read(PROFILE_CLASSES[name], ctypes.byref(ms), ctypes.byref(work),
                                   ctypes.byref(n)), "profile_read")
    return ms.value, work.value, n.value
