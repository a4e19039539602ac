"""Dense numpy restatements used to pin the oracle (independent of oracle/)."""
import numpy as np
from scipy.special import logsumexp


def plan(M, dual, eta):
    n, m = M.shape
    return np.exp((dual[:n, None] + dual[None, n:] - M) / eta)


def dense_Q(T):
    """Q(T) in the reduced gauge (reference stage3b_validate.py:80-90)."""
    n, m = T.shape
    Q = np.zeros((n + m - 1, n + m - 1))
    Q[:n, :n] = np.diag(T.sum(1))
    Q[n:, n:] = np.diag(T.sum(0)[:m - 1])
    Q[:n, n:] = T[:, :m - 1]
    Q[n:, :n] = T[:, :m - 1].T
    return Q


def fg(z, M, a, b, eta):
    n, m = M.shape
    al = z[:n]
    be = np.concatenate([z[n:], [0.0]])
    T = np.exp((al[:, None] + be[None, :] - M) / eta)
    f = eta * T.sum() - al @ a - be @ b
    g = np.concatenate([T.sum(1) - a, T.sum(0)[:-1] - b[:-1]])
    return f, g, T


def newton(M, a, b, eta, iters=300, gtol=1e-15):
    """Damped Newton on the dual with the exact Hessian Q/eta (brute force), run
    until ||g|| < gtol or no step reduces ||g||."""
    n, m = M.shape
    al = eta * (np.log(a) - logsumexp((-M) / eta, axis=1))
    z = np.concatenate([al, np.zeros(m - 1)])
    for _ in range(iters):
        f, g, T = fg(z, M, a, b, eta)
        if np.linalg.norm(g) < gtol:
            break
        d = -np.linalg.solve(dense_Q(T) / eta, g)
        # backtrack on the merit ||g||^2 (d is a descent direction for it), which
        # stays resolvable after f has stopped changing in double precision
        gn0 = np.linalg.norm(g)
        t = 1.0
        while t > 1e-12:
            with np.errstate(over="ignore", invalid="ignore"):
                gt = fg(z + t * d, M, a, b, eta)[1]
            if np.all(np.isfinite(gt)) and np.linalg.norm(gt) <= (1 - 1e-4 * t) * gn0:
                break
            t *= 0.5
        if t <= 1e-12:
            break
        z = z + t * d
    f, g, T = fg(z, M, a, b, eta)
    dual = np.concatenate([z, [0.0]])
    return T, dual, np.linalg.norm(g)


def csr_dense_sym(values, colind, rowptr, N):
    A = np.zeros((N, N))
    for i in range(N):
        for k in range(rowptr[i], rowptr[i + 1]):
            A[i, colind[k]] = values[k]
            A[colind[k], i] = values[k]
    return A
