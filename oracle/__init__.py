"""CPU FP64 oracle for the entropic-OT SPLR path -- TEST INFRASTRUCTURE ONLY.

This package is the checker the HIP library is compared against.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it; the product package ``curegot_backward_amd`` never
does.  It wraps ``liboracle.so`` (built from ``cgot_oracle.c`` by
``oracle/Makefile``) with numpy in/out.

Every function restates a reference routine; the C file cites file:line for
each and lists what pins it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_c_int = ctypes.c_int
_c_double = ctypes.c_double
_c_i64 = ctypes.c_int64


class SplrOpts(ctypes.Structure):
    _fields_ = [
        ("density_max", ctypes.c_double),
        ("shift_max", ctypes.c_double),
        ("sparsity_pattern_cycle", ctypes.c_int),
        ("candidate_sinkhorn_iter", ctypes.c_int),
    ]


class SplrResult(ctypes.Structure):
    _fields_ = [
        ("niter", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("valid", ctypes.c_int),
        ("low_rank", ctypes.c_int),
        ("fallback", ctypes.c_int),
        ("ys", ctypes.c_double),
        ("final_f", ctypes.c_double),
        ("final_gnorm", ctypes.c_double),
        ("nnz", ctypes.c_int64),
    ]


def build() -> str:
    """Compile liboracle.so with gcc (idempotent)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_row_lse_update.argtypes = [_dp, _dp, _dp, _c_double, _c_int, _c_int, _dp]
        L.or_col_lse_update.argtypes = [_dp, _dp, _dp, _c_double, _c_int, _c_int, _dp]
        L.or_sinkhorn_bcd.argtypes = [_dp, _dp, _dp, _c_double, _c_int, _c_double, _c_int,
                                      _c_int, _dp, _dp, _dp, ctypes.POINTER(_c_int)]
        L.or_tpass.argtypes = [_dp, _dp, _dp, _c_double, _c_int, _c_int, _dp, _dp, _dp,
                               _dp, _dp]
        L.or_topk_pattern.argtypes = [_dp, _dp, _c_double, _c_int, _c_int, _c_i64, _ip, _ip]
        L.or_topk_pattern.restype = _c_i64
        L.or_assemble_csr.argtypes = [_dp, _dp, _dp, _dp, _c_double, _c_double, _c_int,
                                      _c_int, _c_i64, _ip, _ip, _dp, _ip, _ip]
        L.or_csr_solve.argtypes = [_dp, _ip, _ip, _c_int, _c_int, _c_int, _dp, _dp]
        L.or_csr_solve.restype = _c_int
        L.or_dense_cholesky_solve.argtypes = [_dp, _c_int, _c_int, _dp, _dp]
        L.or_dense_cholesky_solve.restype = _c_int
        L.or_quadratic_minimizer_fg.argtypes = [_c_double] * 5
        L.or_quadratic_minimizer_fg.restype = _c_double
        L.or_quadratic_minimizer_gg.argtypes = [_c_double] * 4
        L.or_quadratic_minimizer_gg.restype = _c_double
        L.or_cubic_minimizer.argtypes = [_c_double] * 6 + [ctypes.POINTER(_c_int)]
        L.or_cubic_minimizer.restype = _c_double
        L.or_step_selection.argtypes = [_c_double] * 9
        L.or_step_selection.restype = _c_double
        L.or_splr_bwd.argtypes = [_dp, _dp, _dp, _c_double, _c_int, _c_double, _c_int, _c_int,
                                  ctypes.POINTER(SplrOpts), _dp, _dp, _dp,
                                  ctypes.POINTER(SplrResult), _dp, _ip, _ip, _dp, _dp, _dp, _dp]
        L.or_backward.argtypes = [_dp, _dp, _dp, _c_double, _c_int, _c_int, _c_int, _dp,
                                  _dp, _ip, _ip, _dp, _dp, _c_double, _c_int, _c_double,
                                  _c_int, _dp, _dp, ctypes.POINTER(_c_int),
                                  ctypes.POINTER(_c_int)]
        L.or_backward.restype = _c_int
        L.or_q_matvec.argtypes = [_dp, _dp, _c_double, _c_int, _c_int, _dp, _dp]
        L.or_rhs_mu.argtypes = [_dp, _dp, _dp, _c_double, _c_int, _c_int, _dp]
        L.or_cost_xy.argtypes = [_dp, _dp, _c_int, _c_int, _c_int, _c_double, _dp]
        _lib = L
    return _lib


def _d(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return x, x.ctypes.data_as(_dp)


def _i(x):
    x = np.ascontiguousarray(x, dtype=np.int32)
    return x, x.ctypes.data_as(_ip)


def _nullable_d(x):
    if x is None:
        return None, None
    return _d(x)


def row_lse_update(M, beta, loga, reg):
    M, pM = _d(M)
    n, m = M.shape
    beta, pb = _d(beta)
    loga, pl = _d(loga)
    out = np.empty(n)
    lib().or_row_lse_update(pM, pb, pl, reg, n, m, out.ctypes.data_as(_dp))
    return out


def col_lse_update(M, alpha, logb, reg):
    M, pM = _d(M)
    n, m = M.shape
    alpha, pa = _d(alpha)
    logb, pl = _d(logb)
    out = np.empty(m)
    lib().or_col_lse_update(pM, pa, pl, reg, n, m, out.ctypes.data_as(_dp))
    return out


def sinkhorn_bcd(M, a, b, reg, tol=1e-6, max_iter=1000, x0=None):
    M, pM = _d(M)
    n, m = M.shape
    a, pa = _d(a)
    b, pb = _d(b)
    x0, px = _nullable_d(x0)
    plan = np.empty((n, m))
    dual = np.empty(n + m)
    it = _c_int(0)
    lib().or_sinkhorn_bcd(pM, pa, pb, reg, int(max_iter), tol, n, m, px,
                          plan.ctypes.data_as(_dp), dual.ctypes.data_as(_dp), ctypes.byref(it))
    return {"plan": plan, "dual": dual, "niter": it.value}


def tpass(gamma, M, a, b, reg):
    M, pM = _d(M)
    n, m = M.shape
    gamma, pg = _d(gamma)
    ab, pab = _d(np.concatenate([a, b]))
    rows = np.empty(n)
    cols = np.empty(m)
    grad = np.empty(n + m - 1)
    f = _c_double(0)
    gn = _c_double(0)
    lib().or_tpass(pg, pM, pab, reg, n, m, rows.ctypes.data_as(_dp), cols.ctypes.data_as(_dp),
                   grad.ctypes.data_as(_dp), ctypes.byref(f), ctypes.byref(gn))
    return {"rowsum": rows, "colsum": cols, "grad": grad, "f": f.value, "gnorm": gn.value}


def topk_pattern(gamma, M, reg, K):
    M, pM = _d(M)
    n, m = M.shape
    gamma, pg = _d(gamma)
    cap = max(int(K), 1)
    si = np.empty(cap, dtype=np.int32)
    sj = np.empty(cap, dtype=np.int32)
    k = lib().or_topk_pattern(pg, pM, reg, n, m, int(K), si.ctypes.data_as(_ip),
                              sj.ctypes.data_as(_ip))
    return si[:k].copy(), sj[:k].copy()


def assemble_csr(gamma, M, rowsum, colsum, reg, shift_reg, sel_i, sel_j):
    M, pM = _d(M)
    n, m = M.shape
    gamma, pg = _d(gamma)
    rowsum, pr = _d(rowsum)
    colsum, pc = _d(colsum)
    sel_i, psi = _i(sel_i)
    sel_j, psj = _i(sel_j)
    K = len(sel_i)
    N = n + m - 1
    vals = np.empty(K + N)
    col = np.empty(K + N, dtype=np.int32)
    rp = np.empty(N + 1, dtype=np.int32)
    lib().or_assemble_csr(pg, pM, pr, pc, reg, shift_reg, n, m, K, psi, psj,
                          vals.ctypes.data_as(_dp), col.ctypes.data_as(_ip), rp.ctypes.data_as(_ip))
    return vals, col, rp


def csr_solve(values, colind, rowptr, n, m, rhs):
    """Solve A X = rhs; rhs is [N] or [nrhs, N]."""
    values, pv = _d(values)
    colind, pci = _i(colind)
    rowptr, prp = _i(rowptr)
    rhs = np.ascontiguousarray(np.atleast_2d(rhs), dtype=np.float64)
    nrhs = rhs.shape[0]
    sol = np.empty_like(rhs)
    info = lib().or_csr_solve(pv, pci, prp, n, m, nrhs, rhs.ctypes.data_as(_dp),
                              sol.ctypes.data_as(_dp))
    if info != 0:
        raise RuntimeError(f"oracle csr_solve: Schur complement not SPD (pivot {info})")
    return sol


def dense_cholesky_solve(A, rhs):
    A, pA = _d(A)
    N = A.shape[0]
    rhs = np.ascontiguousarray(np.atleast_2d(rhs), dtype=np.float64)
    sol = np.empty_like(rhs)
    info = lib().or_dense_cholesky_solve(pA, N, rhs.shape[0], rhs.ctypes.data_as(_dp),
                                         sol.ctypes.data_as(_dp))
    if info != 0:
        raise RuntimeError(f"oracle dense cholesky: not SPD (pivot {info})")
    return sol


def cubic_minimizer(a, b, fa, fb, ga, gb):
    ex = _c_int(0)
    x = lib().or_cubic_minimizer(a, b, fa, fb, ga, gb, ctypes.byref(ex))
    return x, bool(ex.value)


def quadratic_minimizer_fg(a, b, fa, ga, fb):
    return lib().or_quadratic_minimizer_fg(a, b, fa, ga, fb)


def quadratic_minimizer_gg(a, b, ga, gb):
    return lib().or_quadratic_minimizer_gg(a, b, ga, gb)


def step_selection(al, au, at, fl, fu, ft, gl, gu, gt):
    return lib().or_step_selection(al, au, at, fl, fu, ft, gl, gu, gt)


def default_density(n, m):
    return min(10.0 / min(n, m), 1.0)


def splr_bwd(M, a, b, reg, tol=1e-6, max_iter=1000, density=None, shift=1e-3,
             sparsity_pattern_cycle=30, candidate_sinkhorn_iter=3, x0=None, trace=False):
    """SPLR forward + context export; mirrors sinkhorn_splr_bwd (src/sinkhorn_splr.cu:1305)."""
    M, pM = _d(M)
    n, m = M.shape
    a, pa = _d(a)
    b, pb = _d(b)
    x0, px = _nullable_d(x0)
    dmax = default_density(n, m) if density is None else min(max(float(density), 0.0), 1.0)
    opts = SplrOpts(dmax, max(float(shift), 0.0), int(sparsity_pattern_cycle),
                    int(candidate_sinkhorn_iter))
    N = n + m - 1
    Kmax = max(int(dmax * n * (m - 1)), 1)
    plan = np.empty((n, m))
    dual = np.empty(n + m)
    res = SplrResult()
    cv = np.zeros(Kmax + N)
    cc = np.zeros(Kmax + N, dtype=np.int32)
    cr = np.zeros(N + 1, dtype=np.int32)
    cs = np.zeros(N)
    cy = np.zeros(N)
    cd = np.zeros(N)
    tr = np.zeros(6 * max(int(max_iter), 1)) if trace else None
    lib().or_splr_bwd(pM, pa, pb, reg, int(max_iter), tol, n, m, ctypes.byref(opts), px,
                      plan.ctypes.data_as(_dp), dual.ctypes.data_as(_dp), ctypes.byref(res),
                      cv.ctypes.data_as(_dp), cc.ctypes.data_as(_ip), cr.ctypes.data_as(_ip),
                      cs.ctypes.data_as(_dp), cy.ctypes.data_as(_dp), cd.ctypes.data_as(_dp),
                      tr.ctypes.data_as(_dp) if trace else None)
    nnz = int(res.nnz)
    out = {
        "plan": plan, "dual": dual, "niter": res.niter, "converged": bool(res.converged),
        "valid": bool(res.valid), "low_rank": bool(res.low_rank), "fallback": bool(res.fallback),
        "ys": res.ys, "f": res.final_f, "gnorm": res.final_gnorm,
        "csr_values": cv[:nnz].copy(), "csr_colind": cc[:nnz].copy(), "csr_rowptr": cr,
        "s": cs, "y": cy, "inv_diag": cd,
    }
    if trace:
        out["trace"] = tr.reshape(-1, 6)[: res.niter].copy()
    return out


_PRECOND = {"none": 0, "diag": 1, "splr": 2}


def backward(M, dual, G, reg, ctx=None, preconditioner="diag", tol=1e-12, max_iter=2000):
    """Implicit backward; mirrors solve_backward (curegot/backward.py:72-144)."""
    M, pM = _d(M)
    n, m = M.shape
    dual, pdu = _d(dual)
    G, pG = _d(G)
    N = n + m - 1
    kind = _PRECOND[preconditioner]
    inv, pinv = _nullable_d(ctx["inv_diag"] if ctx is not None else None)
    if ctx is not None:
        v, pv = _d(ctx["csr_values"])
        c, pc = _i(ctx["csr_colind"])
        r, pr = _i(ctx["csr_rowptr"])
        s, ps = _d(ctx["s"])
        y, py = _d(ctx["y"])
        ys = float(ctx["ys"])
        lr = int(ctx["low_rank"])
    else:
        pv = pc = pr = ps = py = None
        ys, lr = 0.0, 0
        if kind != 0:
            raise ValueError("preconditioner needs a context")
    lam = np.empty(N)
    grad = np.empty((n, m))
    it = _c_int(0)
    conv = _c_int(0)
    st = lib().or_backward(pM, pdu, pG, reg, n, m, kind, pinv, pv, pc, pr, ps, py, ys, lr,
                           tol, int(max_iter), lam.ctypes.data_as(_dp), grad.ctypes.data_as(_dp),
                           ctypes.byref(it), ctypes.byref(conv))
    if st == 2:
        raise RuntimeError("PCG breakdown: p^T A p <= 0 (matrix not SPD?)")
    return {"lambda": lam, "grad": grad, "iterations": it.value, "converged": bool(conv.value),
            "preconditioner": preconditioner}


def q_matvec(M, dual, reg, v):
    M, pM = _d(M)
    n, m = M.shape
    dual, pd = _d(dual)
    v, pv = _d(v)
    out = np.empty(n + m - 1)
    lib().or_q_matvec(pM, pd, reg, n, m, pv, out.ctypes.data_as(_dp))
    return out


def rhs_mu(M, dual, G, reg):
    M, pM = _d(M)
    n, m = M.shape
    dual, pd = _d(dual)
    G, pG = _d(G)
    out = np.empty(n + m - 1)
    lib().or_rhs_mu(pM, pd, pG, reg, n, m, out.ctypes.data_as(_dp))
    return out


def cost_xy(X, Y, scale):
    """M_ij = scale * sum_k (X_ik - Y_jk)^2 in the library's order (fused
    multiply-adds in k order, one rounded multiply); mirrors the geometric
    cost of the kernels (the reference forms M in torch,
    examples/differentiable_splr_ot.py:55)."""
    X, pX = _d(X)
    Y, pY = _d(Y)
    n, d = X.shape
    m = Y.shape[0]
    if Y.shape[1] != d:
        raise ValueError("X and Y must have the same number of columns")
    M = np.empty((n, m))
    lib().or_cost_xy(pX, pY, n, m, d, float(scale), M.ctypes.data_as(_dp))
    return M
