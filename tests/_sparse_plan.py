"""Host-side helpers for the sparse Newton-system plan (csrc/cgot_symbolic.h):
build a plan through the library's host-only analysis hook and execute its
schedule in numpy, phase by phase as the device kernel does (periphery
levels, the core's Schur complement, the dense core, the two triangular
solves).  Test infrastructure: it checks the plan's structure against a dense
solve on the CPU; the product runs the schedule only in HIP."""
import ctypes

import numpy as np

from curegot_backward_amd import _lib


def w_pattern(values, colind, rowptr, n, m):
    """W's row CSR (pattern and values) from the reference lower CSR of A
    (rows n+j hold their W entries, then the diagonal)."""
    rows, cols, vals = [], [], []
    for j in range(m - 1):
        for p in range(rowptr[n + j], rowptr[n + j + 1] - 1):
            rows.append(int(colind[p])); cols.append(j); vals.append(values[p])
    order = np.lexsort((cols, rows))
    rows, cols, vals = np.array(rows, int)[order], np.array(cols, int)[order], np.array(vals)[order]
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    Da = np.array([values[rowptr[i]] for i in range(n)])
    Db = np.array([values[rowptr[n + j + 1] - 1] for j in range(m - 1)])
    return rp, cols.astype(np.int32), vals, Da, Db


def make_plan(n, m, rowptr, colj, cap=1 << 40, rmax=196):
    L = _lib.load()
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    colj = np.ascontiguousarray(colj if len(colj) else np.zeros(1), np.int32)
    h = _lib.SpHdr()
    nb = L.cgot_sparse_plan(n, m, rowptr.ctypes.data, colj.ctypes.data, cap, rmax, ctypes.byref(h), None, 0)
    if nb < 0:
        return None, None
    buf = np.zeros(nb, np.uint8)
    assert L.cgot_sparse_plan(n, m, rowptr.ctypes.data, colj.ctypes.data, cap, rmax, ctypes.byref(h),
                              buf.ctypes.data, nb) == nb
    return h, buf


def arr(buf, off, count, dt=np.int32):
    return np.frombuffer(buf, dtype=dt, count=count, offset=off)


def execute(h, buf, Da, Db, wval, rhs):
    """Factor and solve with the plan: returns A^{-1} rhs for rhs [k, N]."""
    N, n, t, R, H = h.N, h.n, h.t, h.R, h.H
    perm = arr(buf, h.o_perm, N)
    lvptr = arr(buf, h.o_lvptr, H + 1)
    lcptr = arr(buf, h.o_lcptr, H + 1)
    lcols = arr(buf, h.o_lcols, t + 1)
    em = arr(buf, h.o_emeta, 4 * h.nnzP).reshape(-1, 4)
    easrc, edpos = em[:, 0], em[:, 3] & ~0x40000000
    epp = np.append(em[:, 1], em[-1, 2] if h.nnzP else 0)
    erow = arr(buf, h.o_erow, h.nnzP)
    pairs = arr(buf, h.o_pairs, 2 * h.npairs).reshape(-1, 2)
    ym = arr(buf, h.o_ymeta, 4 * t).reshape(-1, 4)
    yc = arr(buf, h.o_ycore, 4 * R).reshape(-1, 4)
    rowpairs = arr(buf, h.o_rowpairs, 2 * h.nrowpairs).reshape(-1, 2)
    cm = arr(buf, h.o_cmeta, 4 * h.ncore).reshape(-1, 4)
    casrc = cm[:, 0]
    cpp = np.append(cm[:, 1], cm[-1, 2] if h.ncore else 0)
    cpairs = arr(buf, h.o_cpairs, 2 * h.ncpairs).reshape(-1, 2)
    Acat = np.concatenate([Da, Db, wval, [0.0]])      # index -1 -> the trailing 0

    def aval(src):
        return Acat[np.where(src < 0, len(Acat) - 1, src)]

    def gathered(tgt, ptr, prs, Lv):
        s = np.zeros(len(tgt))
        for k, q in enumerate(tgt):
            p0, p1 = ptr[q], ptr[q + 1]
            if p1 > p0:
                s[k] = np.sum(Lv[prs[p0:p1, 0]] * Lv[prs[p0:p1, 1]])
        return s

    k = rhs.shape[0]
    b = rhs.copy()
    y = np.zeros((k, N))
    Lv = np.zeros(h.nnzP)
    for lev in range(H):
        # the level's entries and its columns' forward substitutions (one phase)
        q = np.arange(lvptr[lev], lvptr[lev + 1])
        s = aval(easrc[q]) - gathered(q, epp, pairs, Lv)
        dg = edpos[q] == q
        assert np.all(s[dg] > 0), "periphery pivot not positive"
        Lv[q[dg]] = np.sqrt(s[dg])
        Lv[q[~dg]] = s[~dg] / Lv[edpos[q[~dg]]]
        for c in range(lcptr[lev], lcptr[lev + 1]):
            vtx, r0, r1, d = ym[c]
            d &= ~0x40000000
            j = erow[d]
            rp = rowpairs[r0:r1]
            y[:, j] = (b[:, vtx] - (Lv[rp[:, 0]] * y[:, rp[:, 1]]).sum(1)) / Lv[d]
    C = aval(casrc) - gathered(np.arange(h.ncore), cpp, cpairs, Lv)
    Cd = np.zeros((R, R))
    Cd[np.tril_indices(R)] = C
    for i in range(R):
        vtx, r0, r1, _ = yc[i]
        rp = rowpairs[r0:r1]
        y[:, t + i] = b[:, vtx] - (Lv[rp[:, 0]] * y[:, rp[:, 1]]).sum(1)
    if R:
        import scipy.linalg as sla
        Lc = np.linalg.cholesky(Cd + np.tril(Cd, -1).T)
        z = sla.solve_triangular(Lc, y[:, t:].T, lower=True)
        y[:, t:] = sla.solve_triangular(Lc.T, z, lower=False).T
    for lev in range(H - 1, -1, -1):
        for c in range(lcptr[lev], lcptr[lev + 1]):
            d, e = lcols[c], lcols[c + 1]
            j = erow[d]
            y[:, j] = (y[:, j] - (Lv[d + 1:e] * y[:, erow[d + 1:e]]).sum(1)) / Lv[d]
    x = np.zeros((k, N))
    x[:, perm] = y
    return x


def dense_A(Da, Db, rowptr, colj, wval, n, m):
    N = n + m - 1
    A = np.zeros((N, N))
    A[np.arange(n), np.arange(n)] = Da
    A[n + np.arange(m - 1), n + np.arange(m - 1)] = Db
    for i in range(n):
        for q in range(rowptr[i], rowptr[i + 1]):
            A[i, n + colj[q]] = A[n + colj[q], i] = wval[q]
    return A
