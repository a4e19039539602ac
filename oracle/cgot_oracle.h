/*
 * CPU FP64 oracle for the entropic-OT SPLR path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * library is compared against; it is never linked into, loaded by or called
 * from the product path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.
 *
 * It restates, in plain C, the algorithms of benben-cd/curegot_backward:
 *   - Sinkhorn/BCD log-sum-exp dual updates  (src/sinkhorn_bcd_kernel.cu:35-192, 429-634)
 *   - the dual objective / gradient T pass    (src/sinkhorn_splr_kernel.cu:214-546)
 *   - global top-K Hessian sparsification + lower CSR (src/sinkhorn_splr_kernel.cu:554-906)
 *   - the Newton solve A x = rhs (cuDSS in the reference, src/linsolve.cu:230-268)
 *   - BFGS/Woodbury search direction          (src/sinkhorn_splr_kernel_low_rank.cu:92-390)
 *   - strong-Wolfe line search                 (src/sinkhorn_splr.cu:667-920, src/utils.h:76-230)
 *   - the SPLR driver with Sinkhorn candidate  (src/sinkhorn_splr.cu:1305-1611)
 *   - the implicit backward PCG               (curegot/_backward_core.py:31-83, curegot/backward.py:72-144)
 *
 * All reductions are sequential in index order, so the oracle is deterministic.
 * Top-K ties are broken by the smaller row-major index i*(m-1)+j (the reference
 * leaves them unspecified: CUB DeviceTopK with determinism::not_guaranteed).
 *
 * Pinning status of every function is listed in the header of cgot_oracle.c.
 */
#ifndef CGOT_ORACLE_H
#define CGOT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* alpha_i = reg * (loga_i - LSE_j((beta_j - M_ij)/reg)) */
void or_row_lse_update(const double* M, const double* beta, const double* loga,
                       double reg, int n, int m, double* alpha);
/* beta_j = reg * (logb_j - LSE_i((alpha_i - M_ij)/reg)) */
void or_col_lse_update(const double* M, const double* alpha, const double* logb,
                       double reg, int n, int m, double* beta);

/* Full BCD solve (cuda_sinkhorn_bcd).  dual = (alpha, beta), not gauge-fixed. */
void or_sinkhorn_bcd(const double* M, const double* a, const double* b,
                     double reg, int max_iter, double tol, int n, int m,
                     const double* x0, double* plan, double* dual, int* niter);

/* T pass at gamma (length n+m): row/col sums, grad (n+m-1), f and ||g||. */
void or_tpass(const double* gamma, const double* M, const double* ab,
              double reg, int n, int m,
              double* rowsum, double* colsum, double* grad,
              double* f, double* gnorm);

/* Top-K selection of T[:, :m-1] at gamma.  Writes K (i, j) pairs in CSR order
 * of the lower Hessian (sorted by j, then i).  Returns the number selected. */
int64_t or_topk_pattern(const double* gamma, const double* M, double reg,
                        int n, int m, int64_t K, int32_t* sel_i, int32_t* sel_j);

/* Assemble the lower CSR of A = Q_Omega + shift_reg*I on a given pattern.
 * values/colind have K+N entries, rowptr N+1. */
void or_assemble_csr(const double* gamma, const double* M, const double* rowsum,
                     const double* colsum, double reg, double shift_reg,
                     int n, int m, int64_t K, const int32_t* sel_i, const int32_t* sel_j,
                     double* values, int32_t* colind, int32_t* rowptr);

/* Exact solve A x = rhs for nrhs right-hand sides (column-major, N each),
 * A given by its lower CSR in the layout produced by or_assemble_csr. */
int or_csr_solve(const double* values, const int32_t* colind, const int32_t* rowptr,
                 int n, int m, int nrhs, const double* rhs, double* sol);

/* Dense SPD Cholesky solve of a general N x N row-major matrix (lower used). */
int or_dense_cholesky_solve(const double* A, int N, int nrhs, const double* rhs, double* sol);

/* Moré–Thuente style step selection helpers (for unit tests). */
double or_quadratic_minimizer_fg(double a, double b, double fa, double ga, double fb);
double or_quadratic_minimizer_gg(double a, double b, double ga, double gb);
double or_cubic_minimizer(double a, double b, double fa, double fb, double ga, double gb, int* exists);
double or_step_selection(double al, double au, double at, double fl, double fu, double ft,
                         double gl, double gu, double gt);

typedef struct or_splr_opts {
    double density_max;          /* default min(10/min(n,m), 1) */
    double shift_max;            /* default 1e-3 */
    int    sparsity_pattern_cycle; /* default 30 */
    int    candidate_sinkhorn_iter; /* default 3 */
} or_splr_opts;

typedef struct or_splr_result {
    int     niter;
    int     converged;
    int     valid;             /* context valid (finite final state) */
    int     low_rank;          /* low-rank term active in the exported context */
    int     fallback;          /* niter == 0: A assembled at the final point */
    double  ys;
    double  final_f;
    double  final_gnorm;
    int64_t nnz;               /* nnz of the exported A */
} or_splr_result;

/* SPLR forward with backward-context export (cuda_sinkhorn_splr_bwd).
 * Outputs: plan [n*m], dual [n+m] (beta_{m-1}=0 gauge).
 * Context outputs (may be NULL): ctx_values/ctx_colind [Kmax+N], ctx_rowptr [N+1],
 * ctx_s, ctx_y, ctx_invdiag [N].
 * trace (may be NULL): 6 doubles per iteration up to max_iter:
 *   f, gnorm, step, low_rank, used_sinkhorn, K  (after each iteration). */
void or_splr_bwd(const double* M, const double* a, const double* b,
                 double reg, int max_iter, double tol, int n, int m,
                 const or_splr_opts* opts, const double* x0,
                 double* plan, double* dual, or_splr_result* res,
                 double* ctx_values, int32_t* ctx_colind, int32_t* ctx_rowptr,
                 double* ctx_s, double* ctx_y, double* ctx_invdiag,
                 double* trace);

/* Implicit backward: solve Q(T) lam = mu by PCG and assemble dL/dM.
 * precond: 0 none, 1 diag (uses invdiag), 2 splr (uses the exported CSR,
 * s, y, ys, low_rank).  T is recomputed from (M, dual). */
int or_backward(const double* M, const double* dual, const double* G,
                double reg, int n, int m, int precond,
                const double* invdiag,
                const double* ctx_values, const int32_t* ctx_colind, const int32_t* ctx_rowptr,
                const double* ctx_s, const double* ctx_y, double ys, int low_rank,
                double tol, int max_iter,
                double* lam, double* grad, int* iterations, int* converged);

/* Matrix-free Q(T) v and mu = [(G.T)1, ((G.T)^T 1)_{:m-1}] (for tests). */
void or_q_matvec(const double* M, const double* dual, double reg, int n, int m,
                 const double* v, double* out);
void or_cost_xy(const double* X, const double* Y, int n, int m, int d, double scale, double* M);
void or_rhs_mu(const double* M, const double* dual, const double* G, double reg,
               int n, int m, double* mu);

#ifdef __cplusplus
}
#endif
#endif
