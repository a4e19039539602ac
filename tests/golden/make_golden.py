"""Writes tests/golden/readme_c1.npz: the README problem (n=64, m=80, d=8,
X, Y ~ U[0,1) from torch.Generator seed 0, M = mean squared distance, uniform
marginals, eta=0.05; reference validate_stage6a_release.py:51-66) with the
oracle's SPLR outputs at tol 1e-8, its implicit gradient for L = sum(T^2), and
the dense-Newton optimum.  The reference ships no golden vectors; these are the
oracle's, pinned by tests/test_golden.py against the brute-force optimum."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from curegot_backward_amd.workloads import readme_problem  # noqa: E402
from tests import _dense  # noqa: E402


def main():
    M, a, b, eta = readme_problem()
    M, a, b = M.numpy(), a.numpy(), b.numpy()
    fw = oracle.splr_bwd(M, a, b, eta, tol=1e-8)
    G = 2.0 * fw["plan"]
    bw = oracle.backward(M, fw["dual"], G, eta, ctx=fw, preconditioner="diag")
    T_opt, dual_opt, gn = _dense.newton(M, a, b, eta)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "readme_c1.npz"),
                        M=M, a=a, b=b, eta=eta, tol=1e-8, plan=fw["plan"], dual=fw["dual"],
                        niter=fw["niter"], G=G, grad=bw["grad"], lam=bw["lambda"],
                        bwd_iters=bw["iterations"], plan_opt=T_opt, dual_opt=dual_opt)


if __name__ == "__main__":
    main()
