/*
 * CPU FP64 oracle for the entropic-OT SPLR path -- TEST INFRASTRUCTURE ONLY.
 *
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg, always as the checker.  The product (curegot_backward_amd) never links,
 * loads or calls it.
 *
 * Notation (reference src/sinkhorn_splr_kernel.cu:38-100):
 *   gamma = (alpha[n], beta[m]) with the gauge beta[m-1] = 0, N = n+m-1,
 *   T_ij = exp((alpha_i + beta_j - M_ij)/reg),
 *   f = reg*sum(T) - <alpha,a> - <beta,b>,  g = (T1 - a, (T^T 1 - b)[:m-1]),
 *   A = [[diag(T1), 0], [T_Omega^T, diag((T^T 1)[:m-1])]] + shift_reg*I (lower CSR).
 *
 * Pinning of each function (tests/test_oracle_*.py):
 *   or_row/col_lse_update   scipy.special.logsumexp; marginal closed form (T1=a / T^T1=b)
 *   or_sinkhorn_bcd         closed form T = a b^T for M = u 1^T + 1 v^T; marginals; SPLR optimum
 *   or_tpass                dense numpy restatement of reference test/test_csr.py:36-58
 *   or_topk_pattern +
 *   or_assemble_csr         numpy argpartition top-K + scipy csr_matrix (test/test_csr.py:61-79)
 *   or_csr_solve,
 *   or_dense_cholesky_solve numpy.linalg.solve on the dense A (test/test_sparse_cholesky.py:46-61)
 *   step selection helpers  exact minima of cubic/quadratic polynomials
 *   or_splr_bwd             dense Newton on the dual to 1e-14 (brute force); strong-Wolfe
 *                           conditions; invariance under M + u1^T + 1v^T; warm start at optimum
 *   or_backward             dense numpy.linalg.solve(Q, mu) (stage3b_validate.py:431-457);
 *                           zero row/col sums of dL/dM; central finite differences
 *   or_cost_xy              exact on small-integer points (numpy, exact arithmetic there);
 *                           an extended-precision numpy sum within a few ulps elsewhere
 */
#include "cgot_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAX(a, b) ((a) > (b) ? (a) : (b))
#define OR_MIN(a, b) ((a) < (b) ? (a) : (b))

/* ------------------------------------------------------------------------ */
/* Geometric cost                                                            */
/* ------------------------------------------------------------------------ */

/* M_ij = scale * sum_{k<d} (X_ik - Y_jk)^2 (X [n*d], Y [m*d] row-major), the
 * terms added in k order by fused multiply-adds, then one rounded multiply.
 * The cost the library forms on the fly in place of a stored M; the
 * reference's callers form it in torch (examples/differentiable_splr_ot.py:55,
 * validate_stage5_workload.py:117, 172). */
void or_cost_xy(const double* X, const double* Y, int n, int m, int d, double scale, double* M)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < m; j++) {
            double s = 0.0;
            for (int k = 0; k < d; k++) {
                const double t = X[(size_t)i * d + k] - Y[(size_t)j * d + k];
                s = fma(t, t, s);
            }
            M[(size_t)i * m + j] = s * scale;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* BCD / LSE updates                                                         */
/* ------------------------------------------------------------------------ */

/* Reference: optimal_alpha_kernel, src/sinkhorn_bcd_kernel.cu:115-192
 * (two passes over the row: max, then sum of exp). */
void or_row_lse_update(const double* M, const double* beta, const double* loga,
                       double reg, int n, int m, double* alpha)
{
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        const double* Mi = M + (size_t)i * m;
        double mx = -INFINITY;
        for (int j = 0; j < m; j++) {
            double v = (beta[j] - Mi[j]) / reg;
            if (v > mx) mx = v;
        }
        double s = 0.0;
        for (int j = 0; j < m; j++) s += exp((beta[j] - Mi[j]) / reg - mx);
        alpha[i] = reg * (loga[i] - (mx + log(s)));
    }
}

/* Reference: optimal_beta_kernel, src/sinkhorn_bcd_kernel.cu:35-112. */
void or_col_lse_update(const double* M, const double* alpha, const double* logb,
                       double reg, int n, int m, double* beta)
{
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < m; j++) {
        double mx = -INFINITY;
        for (int i = 0; i < n; i++) {
            double v = (alpha[i] - M[(size_t)i * m + j]) / reg;
            if (v > mx) mx = v;
        }
        double s = 0.0;
        for (int i = 0; i < n; i++) s += exp((alpha[i] - M[(size_t)i * m + j]) / reg - mx);
        beta[j] = reg * (logb[j] - (mx + log(s)));
    }
}

static void or_plan(const double* M, const double* alpha, const double* beta,
                    double reg, int n, int m, double* P)
{
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++)
            P[(size_t)i * m + j] = exp((alpha[i] + beta[j] - M[(size_t)i * m + j]) / reg);
}

/* Gauge-free difference of two dual pairs.
 * Reference: compute_iter_difference, src/sinkhorn_bcd_kernel.cu:280-326. */
static double or_iter_difference(const double* al, const double* be,
                                 const double* alp, const double* bep, int n, int m)
{
    double sa = 0, sb = 0, sap = 0, sbp = 0;
    for (int i = 0; i < n; i++) { sa += al[i]; sap += alp[i]; }
    for (int j = 0; j < m; j++) { sb += be[j]; sbp += bep[j]; }
    double c = (sb - sa) / (n + m), d = (sbp - sap) / (n + m), K = c - d;
    double acc = 0;
    for (int i = 0; i < n; i++) { double t = al[i] - alp[i] + K; acc += t * t; }
    for (int j = 0; j < m; j++) { double t = be[j] - bep[j] - K; acc += t * t; }
    return sqrt(acc);
}

/* Reference: cuda_sinkhorn_bcd, src/sinkhorn_bcd_kernel.cu:429-634.
 * Stop rule: dual difference < 10*tol switches to ||T1 - a||_2 < tol. */
void or_sinkhorn_bcd(const double* M, const double* a, const double* b,
                     double reg, int max_iter, double tol, int n, int m,
                     const double* x0, double* plan, double* dual, int* niter)
{
    double* al = (double*)malloc(sizeof(double) * n);
    double* be = (double*)malloc(sizeof(double) * m);
    double* alp = (double*)malloc(sizeof(double) * n);
    double* bep = (double*)malloc(sizeof(double) * m);
    double* la = (double*)malloc(sizeof(double) * n);
    double* lb = (double*)malloc(sizeof(double) * m);
    for (int i = 0; i < n; i++) la[i] = log(a[i]);
    for (int j = 0; j < m; j++) lb[j] = log(b[j]);
    if (x0) { memcpy(al, x0, sizeof(double) * n); memcpy(be, x0 + n, sizeof(double) * m); }
    else { memset(al, 0, sizeof(double) * n); memset(be, 0, sizeof(double) * m); }
    int use_marginal = 0;
    *niter = 0;
    for (int it = 0; it < max_iter; it++) {
        memcpy(alp, al, sizeof(double) * n);
        memcpy(bep, be, sizeof(double) * m);
        or_row_lse_update(M, be, la, reg, n, m, al);
        or_col_lse_update(M, al, lb, reg, n, m, be);
        if (!use_marginal) {
            if (or_iter_difference(al, be, alp, bep, n, m) < tol * 10) use_marginal = 1;
        } else {
            double acc = 0;
            for (int i = 0; i < n; i++) {
                double s = 0;
                for (int j = 0; j < m; j++) s += exp((al[i] + be[j] - M[(size_t)i * m + j]) / reg);
                acc += (s - a[i]) * (s - a[i]);
            }
            if (sqrt(acc) < tol) { *niter = it + 1; break; }
        }
        *niter = it + 1;
    }
    if (plan) or_plan(M, al, be, reg, n, m, plan);
    if (dual) { memcpy(dual, al, sizeof(double) * n); memcpy(dual + n, be, sizeof(double) * m); }
    free(al); free(be); free(alp); free(bep); free(la); free(lb);
}

/* ------------------------------------------------------------------------ */
/* T pass: f, g, row/col sums                                                */
/* ------------------------------------------------------------------------ */

/* Reference: T_fused_kernel + obj_grad_kernel, src/sinkhorn_splr_kernel.cu:214-443.
 * f sums reg*rowsum_i - gamma.ab over all n+m entries (gamma[n+m-1] is 0). */
void or_tpass(const double* gamma, const double* M, const double* ab,
              double reg, int n, int m,
              double* rowsum, double* colsum, double* grad,
              double* f, double* gnorm)
{
    const double* al = gamma;
    const double* be = gamma + n;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double s = 0;
        for (int j = 0; j < m; j++) s += exp((al[i] + be[j] - M[(size_t)i * m + j]) / reg);
        rowsum[i] = s;
    }
    for (int j = 0; j < m; j++) colsum[j] = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++)
            colsum[j] += exp((al[i] + be[j] - M[(size_t)i * m + j]) / reg);
    double tsum = 0, dot = 0, g2 = 0;
    for (int k = 0; k < n + m; k++) dot += gamma[k] * ab[k];
    for (int i = 0; i < n; i++) {
        tsum += rowsum[i];
        double g = rowsum[i] - ab[i];
        if (grad) grad[i] = g;
        g2 += g * g;
    }
    for (int j = 0; j < m - 1; j++) {
        double g = colsum[j] - ab[n + j];
        if (grad) grad[n + j] = g;
        g2 += g * g;
    }
    if (f) *f = reg * tsum - dot;
    if (gnorm) *gnorm = sqrt(g2);
}

/* ------------------------------------------------------------------------ */
/* Top-K sparsification and CSR assembly                                     */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t key; int64_t idx; } or_kv;

static int or_kv_cmp_desc(const void* pa, const void* pb)
{
    const or_kv* x = (const or_kv*)pa;
    const or_kv* y = (const or_kv*)pb;
    if (x->key != y->key) return (x->key > y->key) ? -1 : 1;
    return (x->idx < y->idx) ? -1 : (x->idx > y->idx);
}

typedef struct { int32_t i, j; } or_ij;
static int or_ij_cmp_csr(const void* pa, const void* pb)
{
    const or_ij* x = (const or_ij*)pa;
    const or_ij* y = (const or_ij*)pb;
    if (x->j != y->j) return x->j < y->j ? -1 : 1;
    return (x->i < y->i) ? -1 : (x->i > y->i);
}

static inline uint64_t or_bits(double v)
{
    uint64_t u;
    memcpy(&u, &v, sizeof(u));
    return u;   /* T >= 0, so the IEEE bit pattern orders like the value */
}

/* Reference: launch_H_sparsification, src/sinkhorn_splr_kernel.cu:597-675
 * (global top-K of T[:, :m-1], K clamped to [1, n(m-1)]).
 * Deterministic rule: larger value first, ties by smaller i*(m-1)+j. */
int64_t or_topk_pattern(const double* gamma, const double* M, double reg,
                        int n, int m, int64_t K, int32_t* sel_i, int32_t* sel_j)
{
    const int64_t mm = m - 1;
    const int64_t Te = (int64_t)n * mm;
    if (K < 1) K = 1;
    if (K > Te) K = Te;
    if (K <= 0) return 0;
    const double* al = gamma;
    const double* be = gamma + n;
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)Te);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++)
        for (int64_t j = 0; j < mm; j++)
            keys[i * mm + j] = or_bits(exp((al[i] + be[j] - M[(size_t)i * m + j]) / reg));
    /* histogram on the top 16 bits */
    int64_t* hist = (int64_t*)calloc(65536, sizeof(int64_t));
    for (int64_t t = 0; t < Te; t++) hist[keys[t] >> 48]++;
    int64_t above = 0;
    int bstar = 65535;
    for (; bstar >= 0; bstar--) {
        if (above + hist[bstar] >= K) break;
        above += hist[bstar];
    }
    int64_t need = K - above;
    int64_t cnt = hist[bstar];
    or_kv* cand = (or_kv*)malloc(sizeof(or_kv) * (size_t)(cnt > 0 ? cnt : 1));
    or_ij* sel = (or_ij*)malloc(sizeof(or_ij) * (size_t)K);
    int64_t ns = 0, nc = 0;
    for (int64_t t = 0; t < Te; t++) {
        int b = (int)(keys[t] >> 48);
        if (b > bstar) { sel[ns].i = (int32_t)(t / mm); sel[ns].j = (int32_t)(t % mm); ns++; }
        else if (b == bstar) { cand[nc].key = keys[t]; cand[nc].idx = t; nc++; }
    }
    qsort(cand, (size_t)nc, sizeof(or_kv), or_kv_cmp_desc);
    for (int64_t c = 0; c < need; c++) {
        sel[ns].i = (int32_t)(cand[c].idx / mm);
        sel[ns].j = (int32_t)(cand[c].idx % mm);
        ns++;
    }
    qsort(sel, (size_t)ns, sizeof(or_ij), or_ij_cmp_csr);
    for (int64_t t = 0; t < ns; t++) { sel_i[t] = sel[t].i; sel_j[t] = sel[t].j; }
    free(keys); free(hist); free(cand); free(sel);
    return ns;
}

/* Reference: write_diagonal_kernel + recompute_nonzero_values_kernel +
 * launch_csr_conversion, src/sinkhorn_splr_kernel.cu:554-820.
 * Rows 0..n-1 hold only their diagonal; row n+j holds the selected T_ij of
 * column j at columns i (ascending), then its diagonal. */
void or_assemble_csr(const double* gamma, const double* M, const double* rowsum,
                     const double* colsum, double reg, double shift_reg,
                     int n, int m, int64_t K, const int32_t* sel_i, const int32_t* sel_j,
                     double* values, int32_t* colind, int32_t* rowptr)
{
    const int N = n + m - 1;
    const double* al = gamma;
    const double* be = gamma + n;
    int64_t p = 0;
    rowptr[0] = 0;
    for (int i = 0; i < n; i++) {
        values[p] = rowsum[i] + shift_reg;
        colind[p] = i;
        p++;
        rowptr[i + 1] = (int32_t)p;
    }
    int64_t k = 0;
    for (int j = 0; j < m - 1; j++) {
        while (k < K && sel_j[k] == j) {
            int i = sel_i[k];
            values[p] = exp((al[i] + be[j] - M[(size_t)i * m + j]) / reg);
            colind[p] = i;
            p++; k++;
        }
        values[p] = colsum[j] + shift_reg;
        colind[p] = n + j;
        p++;
        rowptr[n + j + 1] = (int32_t)p;
    }
    (void)N;
}

/* ------------------------------------------------------------------------ */
/* Dense Cholesky and the Schur-reduced Newton solve                        */
/* ------------------------------------------------------------------------ */

/* In-place Cholesky (Crout order) of the lower triangle of a row-major
 * P x P matrix.  Returns 0 on success, k+1 if pivot k is not positive. */
static int or_potrf(double* S, int P)
{
    for (int j = 0; j < P; j++) {
        double* Sj = S + (size_t)j * P;
        double d = Sj[j];
        for (int k = 0; k < j; k++) d -= Sj[k] * Sj[k];
        if (!(d > 0.0)) return j + 1;
        d = sqrt(d);
        Sj[j] = d;
        #pragma omp parallel for schedule(static)
        for (int i = j + 1; i < P; i++) {
            double* Si = S + (size_t)i * P;
            double v = Si[j];
            for (int k = 0; k < j; k++) v -= Si[k] * Sj[k];
            Si[j] = v / d;
        }
    }
    return 0;
}

/* Solve L L^T x = b in place. */
static void or_potrs(const double* L, int P, double* x)
{
    for (int i = 0; i < P; i++) {
        const double* Li = L + (size_t)i * P;
        double v = x[i];
        for (int k = 0; k < i; k++) v -= Li[k] * x[k];
        x[i] = v / Li[i];
    }
    for (int i = P - 1; i >= 0; i--) {
        double v = x[i];
        for (int k = i + 1; k < P; k++) v -= L[(size_t)k * P + i] * x[k];
        x[i] = v / L[(size_t)i * P + i];
    }
}

int or_dense_cholesky_solve(const double* A, int N, int nrhs, const double* rhs, double* sol)
{
    double* L = (double*)malloc(sizeof(double) * (size_t)N * N);
    memcpy(L, A, sizeof(double) * (size_t)N * N);
    int info = or_potrf(L, N);
    if (info == 0) {
        for (int r = 0; r < nrhs; r++) {
            memcpy(sol + (size_t)r * N, rhs + (size_t)r * N, sizeof(double) * N);
            or_potrs(L, N, sol + (size_t)r * N);
        }
    }
    free(L);
    return info;
}

/* Exact Newton solve with the reference's lower CSR (src/linsolve.cu solves
 * the same SPD system with cuDSS).  The two diagonal blocks are diagonal, so
 * the larger one is eliminated and the Schur complement onto the smaller one
 *   beta side : S = D_b - W^T D_a^{-1} W   (size m-1, used when m-1 <= n)
 *   alpha side: S = D_a - W D_b^{-1} W^T   (size n)
 * is factorized densely. W_ij = A[n+j, i]. */
int or_csr_solve(const double* values, const int32_t* colind, const int32_t* rowptr,
                 int n, int m, int nrhs, const double* rhs, double* sol)
{
    const int N = n + m - 1;
    const int mb = m - 1;
    double* Da = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    double* Db = (double*)malloc(sizeof(double) * (mb > 0 ? mb : 1));
    for (int i = 0; i < n; i++) Da[i] = values[rowptr[i]];
    /* column lists of W (CSR rows n+j) */
    int64_t K = 0;
    for (int j = 0; j < mb; j++) {
        int32_t b0 = rowptr[n + j], b1 = rowptr[n + j + 1];
        Db[j] = values[b1 - 1];
        K += (b1 - 1 - b0);
    }
    /* row lists of W by counting sort (ascending j inside each row) */
    int64_t* rp = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    int32_t* rj = (int32_t*)malloc(sizeof(int32_t) * (size_t)(K > 0 ? K : 1));
    double* rv = (double*)malloc(sizeof(double) * (size_t)(K > 0 ? K : 1));
    for (int j = 0; j < mb; j++)
        for (int32_t p = rowptr[n + j]; p < rowptr[n + j + 1] - 1; p++) rp[colind[p] + 1]++;
    for (int i = 0; i < n; i++) rp[i + 1] += rp[i];
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) fill[i] = rp[i];
    for (int j = 0; j < mb; j++)
        for (int32_t p = rowptr[n + j]; p < rowptr[n + j + 1] - 1; p++) {
            int i = colind[p];
            rj[fill[i]] = j; rv[fill[i]] = values[p]; fill[i]++;
        }
    const int beta_side = (mb <= n);
    const int P = beta_side ? mb : n;
    double* S = (double*)calloc((size_t)P * P + 1, sizeof(double));
    if (beta_side) {
        for (int j = 0; j < mb; j++) S[(size_t)j * P + j] = Db[j];
        for (int i = 0; i < n; i++)
            for (int64_t p = rp[i]; p < rp[i + 1]; p++)
                for (int64_t q = rp[i]; q <= p; q++)
                    S[(size_t)rj[p] * P + rj[q]] -= rv[p] * rv[q] / Da[i];
    } else {
        for (int i = 0; i < n; i++) S[(size_t)i * P + i] = Da[i];
        for (int j = 0; j < mb; j++) {
            int32_t b0 = rowptr[n + j], b1 = rowptr[n + j + 1] - 1;
            for (int32_t p = b0; p < b1; p++)
                for (int32_t q = b0; q <= p; q++)
                    S[(size_t)colind[p] * P + colind[q]] -= values[p] * values[q] / Db[j];
        }
    }
    int info = or_potrf(S, P);
    if (info == 0) {
        double* t = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
        for (int r = 0; r < nrhs; r++) {
            const double* ga = rhs + (size_t)r * N;
            const double* gb = ga + n;
            double* xa = sol + (size_t)r * N;
            double* xb = xa + n;
            if (beta_side) {
                for (int i = 0; i < n; i++) t[i] = ga[i] / Da[i];
                for (int j = 0; j < mb; j++) {
                    double v = gb[j];
                    for (int32_t p = rowptr[n + j]; p < rowptr[n + j + 1] - 1; p++)
                        v -= values[p] * t[colind[p]];
                    xb[j] = v;
                }
                or_potrs(S, P, xb);
                for (int i = 0; i < n; i++) {
                    double v = ga[i];
                    for (int64_t p = rp[i]; p < rp[i + 1]; p++) v -= rv[p] * xb[rj[p]];
                    xa[i] = v / Da[i];
                }
            } else {
                for (int j = 0; j < mb; j++) t[j] = gb[j] / Db[j];
                for (int i = 0; i < n; i++) {
                    double v = ga[i];
                    for (int64_t p = rp[i]; p < rp[i + 1]; p++) v -= rv[p] * t[rj[p]];
                    xa[i] = v;
                }
                or_potrs(S, P, xa);
                for (int j = 0; j < mb; j++) {
                    double v = gb[j];
                    for (int32_t p = rowptr[n + j]; p < rowptr[n + j + 1] - 1; p++)
                        v -= values[p] * xa[colind[p]];
                    xb[j] = v / Db[j];
                }
            }
        }
        free(t);
    }
    free(Da); free(Db); free(rp); free(rj); free(rv); free(fill); free(S);
    return info;
}

/* ------------------------------------------------------------------------ */
/* Line-search step selection (Moré & Thuente 1994, as used by the reference */
/* src/utils.h:76-230)                                                       */
/* ------------------------------------------------------------------------ */

/* Minimiser of the quadratic through (a, fa), slope ga at a, and (b, fb). */
double or_quadratic_minimizer_fg(double a, double b, double fa, double ga, double fb)
{
    const double h = b - a;
    return a + 0.5 * h * h * ga / (fa - fb + h * ga);
}

/* Minimiser of the quadratic whose derivative is ga at a and gb at b. */
double or_quadratic_minimizer_gg(double a, double b, double ga, double gb)
{
    return a + ga / (ga - gb) * (b - a);
}

/* Local minimiser of the cubic Hermite interpolant of (a, fa, ga), (b, fb, gb).
 * The cubic is written as c0 + c1 x + c2 x^2 + c3 x^3 with the coefficients
 * scaled by (b-a)^3; roots of the derivative are formed stably. */
double or_cubic_minimizer(double a, double b, double fa, double fb, double ga, double gb, int* exists)
{
    const double h = b - a, h2 = h * h, s = a + b;
    const double df = fb - fa, dg = gb - ga;
    const double c3 = (ga + gb) * h - 2.0 * df;                 /* c3 * h^3 */
    const double c2 = 0.5 * (dg * h2 - 3.0 * s * c3);           /* c2 * h^3 */
    const double c1 = df * h2 - s * c2 - (a * s + b * b) * c3;  /* c1 * h^3 */
    const double eps = 2.220446049250313e-16;
    if (fabs(c3) < eps * fabs(c2) || fabs(c3) < eps * fabs(c1)) {
        *exists = (c2 * h > 0.0);
        return *exists ? (-0.5 * c1 / c2) : b;
    }
    const double u = c2 / (3.0 * c3), v = c1 / c2, ratio = v / u;
    *exists = (ratio <= 1.0);
    if (!*exists) return b;
    double r1, r2;
    if (fabs(u) >= fabs(v)) {
        const double w = 1.0 + sqrt(1.0 - ratio);
        r1 = -u * w;
        r2 = -v / w;
    } else {
        const double sd = sqrt(fabs(u)) * sqrt(fabs(v)) * sqrt(1.0 - u / v);
        r1 = -u - sd;
        r2 = -u + sd;
    }
    return (c3 * h > 0.0) ? OR_MAX(r1, r2) : OR_MIN(r1, r2);
}

/* Trial-step selection over the bracket (al, au) given the new point at.
 * Reference: step_selection, src/utils.h:163-230. */
double or_step_selection(double al, double au, double at, double fl, double fu, double ft,
                         double gl, double gu, double gt)
{
    if (al == au) return al;
    if (!isfinite(ft) || !isfinite(gt)) return 0.5 * (al + at);
    int c_ok;
    const double ac = or_cubic_minimizer(al, at, fl, ft, gl, gt, &c_ok);
    const double aq = or_quadratic_minimizer_fg(al, at, fl, gl, ft);
    if (ft > fl) {
        if (!c_ok) return aq;
        return (fabs(ac - al) < fabs(aq - al)) ? ac : 0.5 * (aq + ac);
    }
    const double as = or_quadratic_minimizer_gg(al, at, gl, gt);
    if (gt * gl < 0.0) return (fabs(ac - at) >= fabs(as - at)) ? ac : as;
    if (fabs(gt) < fabs(gl)) {
        const int take_c = c_ok && ((ac - at) * (at - al) > 0.0) && (fabs(ac - at) < fabs(as - at));
        const double r = take_c ? ac : as;
        const double lim = at + 0.66 * (au - at);
        return (at > al) ? OR_MIN(lim, r) : OR_MAX(lim, r);
    }
    if (!isfinite(au) || !isfinite(fu) || !isfinite(gu)) return at + 1.1 * (at - al);
    int e_ok;
    const double ae = or_cubic_minimizer(at, au, ft, fu, gt, gu, &e_ok);
    const double lim = at + 0.66 * (au - at);
    return (at > al) ? OR_MIN(lim, ae) : OR_MAX(lim, ae);
}

/* ------------------------------------------------------------------------ */
/* SPLR driver                                                               */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n, m, N;
    double reg;
    const double* M;
    double* ab;
    double* logab;
    double* gamma;      /* n+m */
    double* gamma_prev; /* n+m */
    double* grad;       /* N */
    double* grad_prev;  /* N */
    double* dir;        /* N */
    double* y;          /* N */
    double* s;          /* N */
    double* u;          /* N: A^{-1} y */
    double* rowsum;     /* n */
    double* colsum;     /* m */
    int64_t Kmax, K;
    int32_t* sel_i;
    int32_t* sel_j;
    double* values;     /* Kmax+N */
    int32_t* colind;
    int32_t* rowptr;
} or_state;

static void or_eval(or_state* st, const double* gamma, double* grad, double* f, double* gnorm)
{
    or_tpass(gamma, st->M, st->ab, st->reg, st->n, st->m, st->rowsum, st->colsum, grad, f, gnorm);
}

static double or_dot(const double* x, const double* y, int N)
{
    double s = 0;
    for (int k = 0; k < N; k++) s += x[k] * y[k];
    return s;
}

/* Trial evaluation: gamma = gamma_prev + t*dir (gamma[N] = 0), f, g, <g, dir>.
 * Reference: launch_line_search_computation, src/sinkhorn_splr_kernel_line_search.cu:374-458. */
static void or_trial(or_state* st, double t, double* f, double* dg)
{
    for (int k = 0; k < st->N; k++) st->gamma[k] = st->gamma_prev[k] + t * st->dir[k];
    st->gamma[st->N] = 0.0;
    or_eval(st, st->gamma, st->grad, f, NULL);
    *dg = or_dot(st->grad, st->dir, st->N);
}

/* Strong-Wolfe line search.  Reference: SPLRSolver::line_search_wolfe,
 * src/sinkhorn_splr.cu:667-920 (c1=1e-4, c2=0.9, 20 trials, step in [1e-10, 2]). */
static double or_line_search(or_state* st, double init_step, double fx_init,
                             double* new_f, int* recompute)
{
    const double c1 = 1e-4, c2 = 0.9;
    const int max_ls = 20;
    const double step_min = 1e-10, step_max = 2.0;
    *recompute = 0;
    init_step = OR_MIN(OR_MAX(init_step, step_min), step_max);
    double step = init_step;
    double fx = fx_init;
    const double dg0 = or_dot(st->grad_prev, st->dir, st->N);
    if (dg0 > 0.0) { *recompute = 1; return init_step; }
    const double dec = c1 * dg0, curv = -c2 * dg0;
    double lo = 0.0, hi = INFINITY;
    double f_lo = 0.0, f_hi = INFINITY;
    double g_lo = (1.0 - c1) * dg0, g_hi = INFINITY;
    double psi_lo = 0.0;
    int bracketed = 0, min_guard = 1, shrink_fail = 0;
    double width = INFINITY, width_prev;
    double dg = dg0;
    for (int it = 0; it < max_ls; it++) {
        or_trial(st, step, &fx, &dg);
        *new_f = fx;
        const double psi = fx - fx_init - step * dec;
        const double dpsi = dg - dec;
        if (psi <= 0.0 && fabs(dg) <= curv) return step;
        if (step <= step_min && (psi > 0.0 || dpsi >= 0.0)) return step;
        if (step >= step_max && (psi <= 0.0 && dpsi < 0.0)) return step;
        /* the auxiliary function psi is used throughout (reference keeps f_is_psi) */
        const double ft = psi, gt = dpsi;
        if (min_guard && (psi <= 0.0 && dpsi < 0.0)) min_guard = 0;
        const int case2 = (psi <= psi_lo) && (dpsi * (lo - step) > 0.0);
        double nxt;
        if (case2) {
            nxt = OR_MIN(step_max, step + 1.1 * (step - lo));
        } else {
            nxt = or_step_selection(lo, hi, step, f_lo, f_hi, ft, g_lo, g_hi, gt);
            nxt = OR_MIN(OR_MAX(nxt, step_min), step_max);
            if (min_guard) {
                const double up = OR_MAX(step_min, (7.0 / 12.0) * step);
                nxt = OR_MIN(OR_MAX(nxt, step_min), up);
            }
        }
        if (psi > psi_lo) {
            hi = step; f_hi = ft; g_hi = gt;
        } else if (case2) {
            lo = step; f_lo = ft; g_lo = gt; psi_lo = psi;
        } else {
            hi = lo; f_hi = f_lo; g_hi = g_lo;
            lo = step; f_lo = ft; g_lo = gt; psi_lo = psi;
        }
        if (!bracketed && !case2) {
            const double l = OR_MIN(lo, hi), r = OR_MAX(lo, hi);
            bracketed = (l >= step_min && r <= step_max);
        }
        if (bracketed) {
            width_prev = width;
            width = fabs(hi - lo);
            if (width_prev < INFINITY && width > 0.66 * width_prev) shrink_fail++;
            else shrink_fail = 0;
            if (shrink_fail >= 2) { nxt = 0.5 * (lo + hi); shrink_fail = 0; }
        }
        step = nxt;
    }
    /* trials exhausted: keep the last evaluated step if it beats lo */
    {
        const double psi = fx - fx_init - step * dec;
        if (psi <= psi_lo) return step;
    }
    *recompute = 1;
    if (lo <= 0.0) return init_step;
    return lo;
}

/* Solve A [dir, u] = [g, y]; Woodbury rank-2 BFGS correction; dir <- -reg*dir.
 * Reference: SPLRSolver::compute_search_direc, src/sinkhorn_splr.cu:538-649, and
 * launch_low_rank_search_direc, src/sinkhorn_splr_kernel_low_rank.cu:296-390. */
static int or_search_dir(or_state* st, int low_rank, double ys)
{
    const int N = st->N;
    double* rhs = (double*)malloc(sizeof(double) * 2 * (size_t)N);
    double* sol = (double*)malloc(sizeof(double) * 2 * (size_t)N);
    memcpy(rhs, st->grad, sizeof(double) * N);
    memcpy(rhs + N, st->y, sizeof(double) * N);
    int info = or_csr_solve(st->values, st->colind, st->rowptr, st->n, st->m,
                            low_rank ? 2 : 1, rhs, sol);
    memcpy(st->dir, sol, sizeof(double) * N);
    if (low_rank) {
        memcpy(st->u, sol + N, sizeof(double) * N);
        const double sg = or_dot(st->s, st->grad, N);
        const double yAy = or_dot(st->y, st->u, N);
        const double yAg = or_dot(st->y, st->dir, N);
        const double sgys = sg / ys;
        const double t1 = (1.0 / st->reg + yAy / ys) * sgys - yAg / ys;
        const double t2 = sgys;
        for (int k = 0; k < N; k++) st->dir[k] += t1 * st->s[k] - t2 * st->u[k];
    }
    for (int k = 0; k < N; k++) st->dir[k] *= -st->reg;
    free(rhs); free(sol);
    return info;
}

/* Rebuild A at st->gamma on a new (update) or kept pattern.
 * Reference: SPLRSolver::dual_sphess, src/sinkhorn_splr.cu:382-407. */
static void or_sphess(or_state* st, double density, double shift_reg, int new_pattern)
{
    const int64_t Te = (int64_t)st->n * (st->m - 1);
    if (new_pattern) {
        if (density > 1.0) density = 1.0;
        if (density < 0.0) density = 0.0;
        int64_t K = (int64_t)(density * (double)Te);
        if (K > st->Kmax) K = st->Kmax;
        if (K < 1) K = 1;
        st->K = or_topk_pattern(st->gamma, st->M, st->reg, st->n, st->m, K, st->sel_i, st->sel_j);
    }
    or_assemble_csr(st->gamma, st->M, st->rowsum, st->colsum, st->reg, shift_reg,
                    st->n, st->m, st->K, st->sel_i, st->sel_j,
                    st->values, st->colind, st->rowptr);
}

/* Reference: cuda_sinkhorn_splr_bwd, src/sinkhorn_splr.cu:1305-1611. */
void or_splr_bwd(const double* M, const double* a, const double* b,
                 double reg, int max_iter, double tol, int n, int m,
                 const or_splr_opts* opts, const double* x0,
                 double* plan, double* dual, or_splr_result* res,
                 double* ctx_values, int32_t* ctx_colind, int32_t* ctx_rowptr,
                 double* ctx_s, double* ctx_y, double* ctx_invdiag,
                 double* trace)
{
    or_state st;
    const int N = n + m - 1;
    st.n = n; st.m = m; st.N = N; st.reg = reg; st.M = M;
    double dmax = opts->density_max;
    dmax = OR_MIN(OR_MAX(dmax, 0.0), 1.0);
    const double dmin = 0.01 * dmax;
    double density = 0.1 * dmax;
    const double shift_max = opts->shift_max;
    const int cycle = opts->sparsity_pattern_cycle;
    const int cand = opts->candidate_sinkhorn_iter;
    int64_t Kmax = (int64_t)(dmax * (double)n * (double)(m - 1));
    if (Kmax < 1) Kmax = 1;
    st.Kmax = Kmax;
    st.K = 0;
    const size_t nm = (size_t)n + m;
    st.ab = (double*)malloc(sizeof(double) * nm);
    st.logab = (double*)malloc(sizeof(double) * nm);
    st.gamma = (double*)malloc(sizeof(double) * nm);
    st.gamma_prev = (double*)calloc(nm, sizeof(double));
    st.grad = (double*)calloc((size_t)N, sizeof(double));
    st.grad_prev = (double*)calloc((size_t)N, sizeof(double));
    st.dir = (double*)calloc((size_t)N, sizeof(double));
    st.y = (double*)calloc((size_t)N, sizeof(double));
    st.s = (double*)calloc((size_t)N, sizeof(double));
    st.u = (double*)calloc((size_t)N, sizeof(double));
    st.rowsum = (double*)malloc(sizeof(double) * n);
    st.colsum = (double*)malloc(sizeof(double) * m);
    st.sel_i = (int32_t*)malloc(sizeof(int32_t) * (size_t)Kmax);
    st.sel_j = (int32_t*)malloc(sizeof(int32_t) * (size_t)Kmax);
    st.values = (double*)malloc(sizeof(double) * (size_t)(Kmax + N));
    st.colind = (int32_t*)malloc(sizeof(int32_t) * (size_t)(Kmax + N));
    st.rowptr = (int32_t*)malloc(sizeof(int32_t) * (size_t)(N + 1));
    double* gamma_bcd = (double*)malloc(sizeof(double) * nm);
    double* grad_bcd = (double*)malloc(sizeof(double) * (size_t)N);
    /* A_k bookkeeping for the context (the A used by the last search direction) */
    double* Ak_values = (double*)malloc(sizeof(double) * (size_t)(Kmax + N));
    int32_t* Ak_colind = (int32_t*)malloc(sizeof(int32_t) * (size_t)(Kmax + N));
    int32_t* Ak_rowptr = (int32_t*)malloc(sizeof(int32_t) * (size_t)(N + 1));
    int64_t Ak_nnz = 0;

    memcpy(st.ab, a, sizeof(double) * n);
    memcpy(st.ab + n, b, sizeof(double) * m);
    for (size_t k = 0; k < nm; k++) st.logab[k] = log(st.ab[k]);

    /* init_dual (src/sinkhorn_splr.cu:330-359) */
    if (x0) {
        memcpy(st.gamma, x0, sizeof(double) * nm);
        const double sh = st.gamma[N];
        for (int k = 0; k < N; k++) st.gamma[k] += (k < n) ? sh : -sh;
        st.gamma[N] = 0.0;
    } else {
        for (int j = 0; j < m; j++) st.gamma[n + j] = 0.0;
        or_row_lse_update(M, st.gamma + n, st.logab, reg, n, m, st.gamma);
    }

    double f, gnorm;
    or_eval(&st, st.gamma, st.grad, &f, &gnorm);
    const double gnorm_init = gnorm;
    double shift = OR_MIN(gnorm, shift_max);
    or_sphess(&st, density, shift * reg, 1);

    double step = 1.0, ys = 0.0, yy = 0.0;
    int low_rank = 0, converged = 0, invalid = 0, niter = 0;
    for (int it = 0; it < max_iter; it++) {
        if (gnorm < tol) { converged = 1; break; }
        if (!isfinite(f)) break;
        /* y = g - g_prev, s = gamma - gamma_prev (low_rank_fused_kernel) */
        ys = 0; yy = 0;
        for (int k = 0; k < N; k++) {
            st.y[k] = st.grad[k] - st.grad_prev[k];
            st.s[k] = st.gamma[k] - st.gamma_prev[k];
            ys += st.y[k] * st.s[k];
            yy += st.y[k] * st.y[k];
        }
        low_rank = (it > 0) && (ys > 1e-6 * yy);
        const int analyze = (it % cycle == 0);
        const int update = (it % cycle == cycle - 1);
        double f_bcd = 0, g_bcd = 0;
        if (analyze && cand != 0) {
            /* Sinkhorn candidate (src/sinkhorn_splr.cu:441-535): alpha from the
             * current beta, then cand (beta, alpha) sweeps, then gauge fix. */
            or_row_lse_update(M, st.gamma + n, st.logab, reg, n, m, gamma_bcd);
            for (int c = 0; c < cand; c++) {
                or_col_lse_update(M, gamma_bcd, st.logab + n, reg, n, m, gamma_bcd + n);
                or_row_lse_update(M, gamma_bcd + n, st.logab, reg, n, m, gamma_bcd);
            }
            const double sh = gamma_bcd[N];
            for (int k = 0; k < N; k++) gamma_bcd[k] += (k < n) ? sh : -sh;
            gamma_bcd[N] = 0.0;
            or_eval(&st, gamma_bcd, grad_bcd, &f_bcd, &g_bcd);
        }
        /* keep A_k for the context before it is replaced */
        memcpy(Ak_values, st.values, sizeof(double) * (size_t)(st.K + N));
        memcpy(Ak_colind, st.colind, sizeof(int32_t) * (size_t)(st.K + N));
        memcpy(Ak_rowptr, st.rowptr, sizeof(int32_t) * (size_t)(N + 1));
        Ak_nnz = st.K + N;
        or_search_dir(&st, low_rank, ys);
        memcpy(st.gamma_prev, st.gamma, sizeof(double) * (size_t)N);
        memcpy(st.grad_prev, st.grad, sizeof(double) * (size_t)N);
        int recompute = 1;
        step = or_line_search(&st, OR_MIN(1.0, 1.5 * step), f, &f, &recompute);
        const double gnorm_pre = gnorm;
        if (recompute) {
            for (int k = 0; k < N; k++) st.gamma[k] = st.gamma_prev[k] + step * st.dir[k];
            or_eval(&st, st.gamma, st.grad, &f, &gnorm);
        } else {
            gnorm = sqrt(or_dot(st.grad, st.grad, N));
        }
        int used_bcd = 0;
        if (analyze && cand != 0) {
            if ((f_bcd < f) && (g_bcd < 10 * gnorm)) {
                used_bcd = 1;
                memcpy(st.gamma, gamma_bcd, sizeof(double) * nm);
                or_eval(&st, st.gamma, st.grad, &f, &gnorm);
            }
        }
        if (update) {
            const int bad = (gnorm_pre < gnorm_init) && (gnorm > gnorm_pre);
            density *= bad ? 1.1 : 0.99;
            density = OR_MIN(dmax, OR_MAX(dmin, density));
        }
        const int conv_after = isfinite(gnorm) && (gnorm < tol);
        const int max_reached = (it + 1 >= max_iter);
        const int inv_after = !isfinite(f) || !isfinite(gnorm);
        const int stop = conv_after || max_reached || inv_after;
        if (!stop) {
            shift = OR_MIN(gnorm, shift_max);
            or_sphess(&st, density, shift * reg, update);
        }
        if (trace) {
            double* tr = trace + 6 * (size_t)it;
            tr[0] = f; tr[1] = gnorm; tr[2] = step; tr[3] = low_rank; tr[4] = used_bcd;
            tr[5] = (double)st.K;
        }
        niter = it + 1;
        if (stop) {
            if (inv_after) invalid = 1;
            if (!max_reached && conv_after) converged = 1;
            break;
        }
    }

    const int valid = !(invalid || !isfinite(f));
    int fallback = 0;
    if (niter == 0) {
        /* immediate convergence: context holds A at the final point, no low rank */
        fallback = 1;
        memcpy(Ak_values, st.values, sizeof(double) * (size_t)(st.K + N));
        memcpy(Ak_colind, st.colind, sizeof(int32_t) * (size_t)(st.K + N));
        memcpy(Ak_rowptr, st.rowptr, sizeof(int32_t) * (size_t)(N + 1));
        Ak_nnz = st.K + N;
        low_rank = 0;
        ys = 0.0;
    }
    if (ctx_values) memcpy(ctx_values, Ak_values, sizeof(double) * (size_t)Ak_nnz);
    if (ctx_colind) memcpy(ctx_colind, Ak_colind, sizeof(int32_t) * (size_t)Ak_nnz);
    if (ctx_rowptr) memcpy(ctx_rowptr, Ak_rowptr, sizeof(int32_t) * (size_t)(N + 1));
    if (ctx_s) memcpy(ctx_s, st.s, sizeof(double) * (size_t)N);
    if (ctx_y) memcpy(ctx_y, st.y, sizeof(double) * (size_t)N);
    if (ctx_invdiag) {
        for (int r = 0; r < N; r++) {
            double d = 0.0;
            for (int32_t p = Ak_rowptr[r]; p < Ak_rowptr[r + 1]; p++)
                if (Ak_colind[p] == r) d = Ak_values[p];
            ctx_invdiag[r] = (d > 0.0 && isfinite(d)) ? 1.0 / d : 0.0;
        }
    }
    if (dual) memcpy(dual, st.gamma, sizeof(double) * nm);
    if (plan) or_plan(M, st.gamma, st.gamma + n, reg, n, m, plan);
    res->niter = niter;
    res->converged = converged;
    res->valid = valid;
    res->low_rank = low_rank;
    res->fallback = fallback;
    res->ys = ys;
    res->final_f = f;
    res->final_gnorm = gnorm;
    res->nnz = Ak_nnz;

    free(st.ab); free(st.logab); free(st.gamma); free(st.gamma_prev); free(st.grad);
    free(st.grad_prev); free(st.dir); free(st.y); free(st.s); free(st.u);
    free(st.rowsum); free(st.colsum); free(st.sel_i); free(st.sel_j);
    free(st.values); free(st.colind); free(st.rowptr); free(gamma_bcd); free(grad_bcd);
    free(Ak_values); free(Ak_colind); free(Ak_rowptr);
}

/* ------------------------------------------------------------------------ */
/* Backward                                                                  */
/* ------------------------------------------------------------------------ */

/* Q(T) v in the reduced gauge.  Reference: q_matvec, curegot/_backward_core.py:31-38. */
void or_q_matvec(const double* M, const double* dual, double reg, int n, int m,
                 const double* v, double* out)
{
    const double* al = dual;
    const double* be = dual + n;
    const double* va = v;
    const double* vb = v + n;
    double* colacc = (double*)calloc((size_t)m, sizeof(double));
    double* colT = (double*)calloc((size_t)m, sizeof(double));
    for (int i = 0; i < n; i++) {
        double rs = 0, acc = 0;
        for (int j = 0; j < m; j++) {
            double t = exp((al[i] + be[j] - M[(size_t)i * m + j]) / reg);
            rs += t;
            if (j < m - 1) acc += t * vb[j];
            colT[j] += t;
            colacc[j] += t * va[i];
        }
        out[i] = rs * va[i] + acc;
    }
    for (int j = 0; j < m - 1; j++) out[n + j] = colacc[j] + colT[j] * vb[j];
    free(colacc); free(colT);
}

/* mu = [(G.T) 1_m, ((G.T)^T 1_n)[:m-1]].  Reference: rhs_Q_scale, _backward_core.py:41-45. */
void or_rhs_mu(const double* M, const double* dual, const double* G, double reg,
               int n, int m, double* mu)
{
    const double* al = dual;
    const double* be = dual + n;
    double* col = (double*)calloc((size_t)m, sizeof(double));
    for (int i = 0; i < n; i++) {
        double r = 0;
        for (int j = 0; j < m; j++) {
            double gt = G[(size_t)i * m + j] * exp((al[i] + be[j] - M[(size_t)i * m + j]) / reg);
            r += gt;
            col[j] += gt;
        }
        mu[i] = r;
    }
    for (int j = 0; j < m - 1; j++) mu[n + j] = col[j];
    free(col);
}

typedef struct {
    int kind;
    int n, m, N;
    double reg;
    const double* invdiag;
    const double* values; const int32_t* colind; const int32_t* rowptr;
    const double* s; const double* y; double ys; int low_rank;
    double* u;
} or_precond;

/* B^{-1} r with B = A + rank-2 BFGS term.  Reference: SPLRBackwardContext::
 * apply_preconditioner, src/backward_context.cu:543-581. */
static void or_apply_precond(const or_precond* P, const double* r, double* z)
{
    const int N = P->N;
    if (P->kind == 0) { memcpy(z, r, sizeof(double) * N); return; }
    if (P->kind == 1) { for (int k = 0; k < N; k++) z[k] = P->invdiag[k] * r[k]; return; }
    or_csr_solve(P->values, P->colind, P->rowptr, P->n, P->m, 1, r, z);
    if (P->low_rank) {
        const double sg = or_dot(P->s, r, N);
        const double yAy = or_dot(P->y, P->u, N);
        const double yAg = or_dot(P->y, z, N);
        const double sgys = sg / P->ys;
        const double t1 = (1.0 / P->reg + yAy / P->ys) * sgys - yAg / P->ys;
        for (int k = 0; k < N; k++) z[k] += t1 * P->s[k] - sgys * P->u[k];
    }
}

/* Reference: solve_backward + pcg + assemble_grad, curegot/backward.py:72-144,
 * curegot/_backward_core.py:47-83.  Returns 0, or 2 on PCG breakdown. */
int or_backward(const double* M, const double* dual, const double* G,
                double reg, int n, int m, int precond,
                const double* invdiag,
                const double* ctx_values, const int32_t* ctx_colind, const int32_t* ctx_rowptr,
                const double* ctx_s, const double* ctx_y, double ys, int low_rank,
                double tol, int max_iter,
                double* lam, double* grad, int* iterations, int* converged)
{
    const int N = n + m - 1;
    double* mu = (double*)malloc(sizeof(double) * N);
    or_rhs_mu(M, dual, G, reg, n, m, mu);
    int any = 0;
    for (int k = 0; k < N; k++) if (mu[k] != 0.0) { any = 1; break; }
    int status = 0;
    for (int k = 0; k < N; k++) lam[k] = 0.0;
    *iterations = 0;
    *converged = 1;
    if (any) {
        or_precond P;
        P.kind = precond; P.n = n; P.m = m; P.N = N; P.reg = reg; P.invdiag = invdiag;
        P.values = ctx_values; P.colind = ctx_colind; P.rowptr = ctx_rowptr;
        P.s = ctx_s; P.y = ctx_y; P.ys = ys; P.low_rank = low_rank; P.u = NULL;
        if (precond == 2 && low_rank) {
            P.u = (double*)malloc(sizeof(double) * N);
            or_csr_solve(ctx_values, ctx_colind, ctx_rowptr, n, m, 1, ctx_y, P.u);
        }
        double* r = (double*)malloc(sizeof(double) * N);
        double* z = (double*)malloc(sizeof(double) * N);
        double* p = (double*)malloc(sizeof(double) * N);
        double* w = (double*)malloc(sizeof(double) * N);
        memcpy(r, mu, sizeof(double) * N);   /* r = b - Q*0 */
        double bnorm = sqrt(or_dot(mu, mu, N));
        or_apply_precond(&P, r, z);
        memcpy(p, z, sizeof(double) * N);
        double rz = or_dot(r, z, N);
        *converged = 0;
        *iterations = max_iter;
        for (int k = 0; k < max_iter; k++) {
            or_q_matvec(M, dual, reg, n, m, p, w);
            const double pw = or_dot(p, w, N);
            if (pw <= 0) { status = 2; *iterations = k; break; }
            const double al = rz / pw;
            for (int q = 0; q < N; q++) { lam[q] += al * p[q]; r[q] -= al * w[q]; }
            const double rn = sqrt(or_dot(r, r, N));
            if (rn < tol * bnorm) { *iterations = k + 1; *converged = 1; break; }
            or_apply_precond(&P, r, z);
            const double rzn = or_dot(r, z, N);
            const double be = rzn / rz;
            for (int q = 0; q < N; q++) p[q] = z[q] + be * p[q];
            rz = rzn;
        }
        free(r); free(z); free(p); free(w);
        if (P.u) free(P.u);
    }
    /* dL/dM = (1/reg) T.(lam_a 1^T + 1 lam_b^T - G), lam_b[m-1] = 0 */
    const double* al = dual;
    const double* be = dual + n;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++) {
            const size_t q = (size_t)i * m + j;
            const double t = exp((al[i] + be[j] - M[q]) / reg);
            const double lb = (j < m - 1) ? lam[n + j] : 0.0;
            grad[q] = (t * (lam[i] + lb - G[q])) / reg;
        }
    free(mu);
    return status;
}
