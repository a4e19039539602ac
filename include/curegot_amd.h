/*
 * curegot_amd -- C ABI of the MI355X (gfx950) entropic-OT SPLR library.
 *
 * Plain pointers and sizes only.  All array arguments are DEVICE pointers to
 * FP64 (or int32) memory on the current HIP device; `stream` is a hipStream_t
 * passed as void* (NULL = default stream).  Every solve runs on that stream and
 * takes all of its device memory from the caller's workspace (query its size
 * first); the library allocates only a small pinned host buffer per thread.
 *
 * Functions return 0 on success and a negative code on error; the message is
 * available from cgot_last_error() (thread-local).
 *
 * Reference interfaces replaced (benben-cd/curegot_backward):
 *   cgot_splr_solve     cuda_sinkhorn_splr_bwd      src/sinkhorn.h:36-79, src/sinkhorn_splr.cu:1305-1611
 *                       (and the pybind entry torch_sinkhorn_splr_bwd, src/solvers_torch.cpp:707-912);
 *                       B > 1 replaces the sequential loop of splr_ot_batch, curegot/autograd.py:340-367
 *   cgot_adjoint_solve  solve_backward + pcg + assemble_grad, curegot/backward.py:72-144,
 *                       curegot/_backward_core.py:31-83, with the diagonal preconditioner of
 *                       SPLRBackwardContext::apply_diagonal_preconditioner, src/backward_context.cu:642-655
 *   cgot_bcd_solve      cuda_sinkhorn_bcd            src/solvers.h:4-12, src/sinkhorn_bcd_kernel.cu:429-634
 *   cgot_T_sparsify     T_computation_sparsify_host  src/sinkhorn_splr_kernel.cu:909-1032 (test hook,
 *                       exposed to Python as curegot.tests.test_T_computation_sparsify, src/tests.cpp:36-150)
 */
#ifndef CUREGOT_AMD_H
#define CUREGOT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Geometric cost in place of a stored M: with X != NULL the instance's M is
 * never read (it may be NULL) and every kernel forms
 *     M_ij = scale * sum_{k<d} (X[i*d+k] - Y[j*d+k])^2
 * on the fly (k in increasing order, each term added by a fused
 * multiply-add, then one rounded multiply by scale).  No counterpart in the
 * reference, whose callers form M = |x_i - y_j|^2 / d in torch
 * (examples/differentiable_splr_ot.py:55, validate_stage5_workload.py:172);
 * with d a power of two and scale = 1/d the division there is the same
 * multiply.  X: [n*d], Y: [m*d] row-major device arrays; d >= 1. */
typedef struct cgot_cost_xy {
    const double* X;
    const double* Y;
    int32_t d;
    double scale;
} cgot_cost_xy;

/* One OT instance of a (possibly ragged) SPLR batch. */
typedef struct cgot_splr_problem {
    const double* M;        /* [n*m] row-major cost (NULL when xy.X is set) */
    const double* a;        /* [n] source marginal (> 0) */
    const double* b;        /* [m] target marginal (> 0) */
    int32_t n, m;
    double reg;             /* entropic regularisation eta > 0 */
    double tol;             /* stop when ||g||_2 < tol */
    int32_t max_iter;
    double density_max;     /* <= 0: default min(10/min(n,m), 1) */
    double shift_max;       /* < 0: default 1e-3 */
    int32_t sparsity_pattern_cycle;   /* <= 0: default 30 */
    int32_t candidate_sinkhorn_iter;  /* < 0: default 3 */
    const double* x0;       /* [n+m] warm start or NULL */
    double* plan;           /* [n*m] output T or NULL */
    double* dual;           /* [n+m] output (alpha, beta), beta[m-1] = 0, or NULL */
    double* trace;          /* HOST [8*max_iter] per-iteration record or NULL:
                               f, ||g||, step, low_rank, sinkhorn candidate taken, K,
                               line-search trials, recomputed */
    cgot_cost_xy xy;        /* geometric cost (xy.X = NULL: M is used) */
} cgot_splr_problem;

/* Per-instance result and backward-context view (device pointers into the
 * workspace; valid while the workspace is alive and not reused). */
typedef struct cgot_splr_info {
    int32_t niter;
    int32_t converged;
    int32_t valid;          /* 0: non-finite objective, no usable context */
    int32_t low_rank;       /* rank-2 BFGS term active in the context */
    int32_t fallback;       /* niter == 0: context matrix assembled at the final point */
    int32_t n, m;
    double ys;
    double f;               /* final dual objective */
    double gnorm;           /* final ||g||_2 */
    int64_t K;              /* selected entries of T in the context matrix */
    const double* inv_diag; /* [n+m-1] 1/diag(A) of the context matrix */
    const double* s;        /* [n+m-1] secant s of the last step */
    const double* y;        /* [n+m-1] secant y of the last step */
    const double* u;        /* [n+m-1] A^{-1} y (valid if low_rank) */
    const double* gamma;    /* [n+m] final duals */
    const void* factor;     /* the instance's stored Newton factor (A of the last iteration,
                               its Schur Cholesky factor, s, y, u), for the SPLR preconditioner
                               of cgot_adjoint_solve; NULL when niter == 0 */
} cgot_splr_info;

size_t cgot_splr_workspace_size(int32_t B, const cgot_splr_problem* probs);
/* A batch of B >= 8 instances is split by Schur block count into up to 4
 * classes that run concurrently (each on its own internal HIP stream and host
 * thread, ordered after `stream` on entry and before its later work on exit).
 * Results are per instance and independent of the split.  cgot_set_max_groups
 * (1..4, default 4) bounds the number of classes for the solves issued from the
 * calling thread; 1 runs the whole batch in one lockstep group on `stream`.  No counterpart in the reference, whose
 * splr_ot_batch loops over instances (curegot/autograd.py:340-367). */
int cgot_set_max_groups(int32_t groups);
/* Tile shapes of the M passes (T pass, log-sum-exp passes, top-K selection,
 * adjoint Q v): an instance with m columns is cut into strips of 64c columns,
 * c = 1, 2 or 4 columns per lane.  Row k of the table covers the m <= m_max[k]
 * not covered by an earlier row (m_max strictly ascending, at most 64 rows);
 * m past the last row take the default c, the smallest power of two with
 * 64c >= m, at most 4.  count 0 restores the default everywhere.  Process-wide,
 * read when a solve lays out its instances.  c fixes the summation order of the
 * row sums: results are bitwise reproducible under one table, and a batch and
 * the single solves of its instances agree bitwise under the same table.  The
 * table is what the autotuner (curegot_backward_amd.autotune) picks.  No
 * counterpart in the reference, whose kernels use fixed launch shapes.
 * Returns 0, or -2 (cgot_last_error) for bad rows. */
int cgot_set_tile_table(const int32_t* m_max, const int32_t* cols, int32_t count);
/* Copies up to cap rows of the current table; returns its row count. */
int cgot_get_tile_table(int32_t* m_max, int32_t* cols, int32_t cap);
/* Columns per lane an instance with m columns takes now (-1 for m < 1). */
int cgot_tile_cols(int32_t m);
/* Newton step of the SPLR iteration: how A is factored.  mode 0: the dense
 * Cholesky of the Schur complement for every instance; mode 2: the sparse
 * factorization (host minimum-degree analysis once per pattern,
 * one-workgroup numeric factor and solves per iteration) for every instance
 * whose pattern admits one; mode 1 (the default): the sparse factorization
 * only where the plan's latency estimate is below the dense one's.  Process-wide;
 * returns the previous mode, or -1 for a mode outside 0..2.  Results agree
 * between the two factorizations to rounding.  The reference always factors
 * A with cuDSS (src/linsolve.cu:27-268). */
int cgot_set_sparse_solve(int32_t mode);
int cgot_splr_solve(int32_t B, const cgot_splr_problem* probs, void* workspace, size_t workspace_bytes,
                    void* stream, cgot_splr_info* info);

/* Sessions: continuous batching of SPLR solves.  No counterpart in the
 * reference (its splr_ot_batch is a sequential loop, curegot/autograd.py:340-367);
 * for request loops that admit new problems while others are iterating.
 * A session is one persistent lockstep group on `stream`: cgot_session_add
 * queues a problem, which starts at the next cgot_session_step; each step runs
 * up to `iters` SPLR iterations of every instance in the group and returns the
 * instances that stopped in it (their plan and dual written, tags and infos in
 * the output arrays, slots freed).  Each instance runs exactly the iterations
 * of a solo cgot_splr_solve and its results are bitwise the same.
 *   - Limits, fixed at creation: at most `capacity` instances at once, with
 *     sum of (n + m) <= max_sum_nm, sum of cgot_session_inst_tiles(n, m) <=
 *     max_tiles and ceil(min(n, m-1) / 64) <= max_nblk each; cgot_session_add
 *     returns -3 (and keeps nothing) while the group is full.
 *   - Memory: the session workspace (cgot_session_workspace_size) holds the
 *     group's shared arrays; every instance brings a workspace of its own
 *     (cgot_session_inst_workspace_size) that holds all of its state.  After
 *     the instance stops, its info.factor and the other info pointers refer to
 *     that workspace (the backward context): keep it alive while they are used.
 *   - One thread at a time per session; the session's work is ordered on
 *     `stream`, and a step returns with it complete.
 *   - A geometric-cost instance and a stored-M instance cannot be in one group. */
size_t cgot_session_workspace_size(int32_t capacity, int64_t max_sum_nm, int64_t max_tiles, int32_t max_nblk);
size_t cgot_session_inst_workspace_size(const cgot_splr_problem* p);
int64_t cgot_session_inst_tiles(int32_t n, int32_t m);
int cgot_session_create(int32_t capacity, int64_t max_sum_nm, int64_t max_tiles, int32_t max_nblk, void* workspace,
                        size_t workspace_bytes, void* stream, void** session);
/* Returns the slot (>= 0), -3 when the group has no room now, or an error. */
int cgot_session_add(void* session, const cgot_splr_problem* p, void* inst_workspace, size_t inst_workspace_bytes,
                     int64_t tag);
int cgot_session_step(void* session, int32_t iters, int64_t* tags_out, cgot_splr_info* info_out, int32_t out_cap,
                      int32_t* ndone);
/* Instances in the group (added and not yet returned by a step). */
int cgot_session_live(void* session);
int cgot_session_destroy(void* session);

/* Implicit backward of the SPLR plan w.r.t. M for one instance of a batch:
 * solve Q(T*) lambda = mu by PCG (T* recomputed from M and the duals),
 * then dL/dM = T*.(lambda_a 1^T + 1 lambda_b^T - G)/reg.  The same lambda
 * gives the gradients w.r.t. the marginals (no counterpart in the reference,
 * which raises for them): dL/da = lambda[0:n], dL/db = (lambda[n:n+m-1], 0). */
typedef struct cgot_adjoint_problem {
    const double* M;        /* [n*m] (NULL when xy.X is set) */
    const double* dual;     /* [n+m] converged duals, beta[m-1] = 0 */
    const double* G;        /* [n*m] upstream gradient dL/dT */
    int32_t n, m;
    double reg;
    const double* inv_diag; /* [n+m-1] Jacobi preconditioner, or NULL for plain CG */
    double tol;             /* relative residual tolerance */
    int32_t max_iter;
    double* lambda;         /* [n+m-1] output */
    double* grad;           /* [n*m] output dL/dM */
    const void* factor;     /* cgot_splr_info.factor of the forward solve (its workspace alive):
                               SPLR preconditioner B^{-1} r, B = A + rank-2 BFGS term
                               (SPLRBackwardContext::apply_preconditioner,
                               src/backward_context.cu:543-581); takes precedence over inv_diag;
                               NULL for none */
    cgot_cost_xy xy;        /* geometric cost of the forward solve (xy.X = NULL: M is used) */
} cgot_adjoint_problem;

typedef struct cgot_adjoint_info {
    int32_t iterations;
    int32_t converged;
    int32_t status;         /* 1 converged, 2 breakdown (p^T Q p <= 0), 3 max_iter, 4 zero rhs */
} cgot_adjoint_info;

size_t cgot_adjoint_workspace_size(int32_t B, const cgot_adjoint_problem* probs);
int cgot_adjoint_solve(int32_t B, const cgot_adjoint_problem* probs, void* workspace,
                       size_t workspace_bytes, void* stream, cgot_adjoint_info* info);

/* SPLR preconditioner of one converged forward solve applied to a vector:
 * z = B^{-1} r with B = A + rank-2 BFGS term of the last iteration, A^{-1} by
 * the stored Schur factor (factor = cgot_splr_info.factor, its workspace
 * alive; r, z: [n+m-1] device vectors).  Replaces
 * SPLRBackwardContext::apply_preconditioner, src/backward_context.cu:543-581. */
size_t cgot_splr_precond_workspace_size(const void* factor, int32_t n, int32_t m);
int cgot_splr_precond_apply(const void* factor, int32_t n, int32_t m, const double* r, double* z,
                            void* workspace, size_t workspace_bytes, void* stream);

/* cgot_splr_precond_apply without a workspace, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): no counterpart in the reference.  Its work buffers live
 * in a fixed 4 MiB device array of the library (one per GPU, held until the
 * process ends); it calls no hipMalloc, hipFree or pinned allocation.
 * Problems whose buffers exceed the array are rejected with -2 and a message
 * naming the sizes; every problem with n, m <= 24576 fits.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU race.  Same result
 * as cgot_splr_precond_apply, bit for bit. */
int cgot_splr_precond_apply_noalloc(const void* factor, int32_t n, int32_t m, const double* r, double* z,
                                    void* stream);

/* Timings of cgot_splr_precond_apply_profiled, in microseconds.  host_*: time
 * in the calling thread (enqueueing, and the final wait); gpu_*: HIP-event
 * intervals on the stream.  There is no host copy in this design, so the
 * copy fields hold the time to enqueue the device copies of r and z. */
typedef struct cgot_precond_profile {
    double host_total_us;
    double host_memcpy_in_us;       /* enqueue r -> workspace */
    double host_solve_launch_us;    /* enqueue the A^{-1} r launches */
    double host_lowrank_us;         /* enqueue the rank-2 launches */
    double host_sync_wait_us;       /* wait for the stream */
    double host_memcpy_out_us;      /* enqueue workspace -> z */
    double gpu_solve_us;            /* A^{-1} r: Schur rhs, triangular solves, back-substitution */
    double gpu_lowrank_reduce_us;   /* rank-2 dot products */
    double gpu_lowrank_correction_us;
    double repeated_total_us;       /* repeats > 1: host time of all repeats, synchronized */
} cgot_precond_profile;

/* Profiled / repeated / base-only variants of cgot_splr_precond_apply (the
 * same workspace size).  base_only != 0 applies A^{-1} only (no rank-2
 * term).  repeats >= 1 applies it that many times back to back on the same r
 * (z holds the last); with repeats == 1 the per-stage timings are filled in,
 * with repeats > 1 only repeated_total_us.  Replaces the debug/benchmark
 * methods profile_preconditioner, benchmark_preconditioner_repeated,
 * profile_base_solve and benchmark_base_solve_repeated
 * (src/solvers_torch.cpp:397-580; backward_context.cu apply_*_profiled). */
int cgot_splr_precond_apply_profiled(const void* factor, int32_t n, int32_t m, const double* r, double* z,
                                     int32_t base_only, int32_t repeats, cgot_precond_profile* prof,
                                     void* workspace, size_t workspace_bytes, void* stream);

/* The sparsified Hessian A of the stored factor (the last factorized
 * iterate) in the reference's lower CSR layout: rows 0..n-1 hold their
 * diagonal; row n+j holds its selected entries (ascending column) and then
 * its diagonal.  nnz = n + (m-1) + K.  With csr_values == NULL only *nnz is
 * written (one device read); otherwise csr_values [nnz], csr_colind [nnz]
 * and csr_rowptr [n+m] (device) are written on the stream.  Replaces the
 * validation accessors csr_values / csr_colind / csr_rowptr,
 * src/solvers_torch.cpp:583-610 (backward_context.h csr_*). */
int cgot_splr_context_csr(const void* factor, int32_t n, int32_t m, double* csr_values, int32_t* csr_colind,
                          int32_t* csr_rowptr, int64_t* nnz, void* stream);

/* Sinkhorn/BCD solve (legacy sinkhorn_bcd).  dual is not gauge-fixed. */
typedef struct cgot_bcd_problem {
    const double* M; const double* a; const double* b;
    int32_t n, m;
    double reg, tol;
    int32_t max_iter;
    const double* x0;       /* [n+m] or NULL */
    double* plan;           /* [n*m] or NULL */
    double* dual;           /* [n+m] or NULL */
    cgot_cost_xy xy;        /* geometric cost (xy.X = NULL: M is used) */
} cgot_bcd_problem;

size_t cgot_bcd_workspace_size(int32_t B, const cgot_bcd_problem* probs);
int cgot_bcd_solve(int32_t B, const cgot_bcd_problem* probs, void* workspace, size_t workspace_bytes,
                   void* stream, int32_t* niter);

/* One LSE dual update on device (BCD half-sweep), for tests and benchmarks:
 * which = 0: alpha = reg*(log a - LSE_j((beta_j - M_ij)/reg)) using dual[n:];
 * which = 1: beta  = reg*(log b - LSE_i((alpha_i - M_ij)/reg)) using dual[:n].
 * dual is updated in place. */
size_t cgot_lse_workspace_size(int32_t n, int32_t m);
int cgot_lse_update(const double* M, const double* a, const double* b, int32_t n, int32_t m,
                    double reg, int32_t which, int32_t repeats, double* dual, void* workspace,
                    size_t workspace_bytes, void* stream);

/* Test hook: T pass + top-K sparsification at gamma = (alpha, beta).
 * Outputs rowsum [n], colsum [m], grad [n+m-1], scalars {f, gnorm} and the lower
 * CSR of Hsl = [[diag(T1)+shift, 0], [T_Omega^T, diag(T^T 1)+shift]]
 * (values/colind [K'+n+m-1], rowptr [n+m]), K' = min(max(K,1), n(m-1)).
 * If rhs != NULL also solves Hsl_sym x = rhs ([n+m-1]) into sol. */
size_t cgot_T_sparsify_workspace_size(int32_t n, int32_t m, int64_t K);
int cgot_T_sparsify(const double* gamma, const double* M, const double* a, const double* b,
                    int32_t n, int32_t m, double reg, double shift, int64_t K,
                    double* rowsum, double* colsum, double* grad, double* scalars,
                    double* csr_values, int32_t* csr_colind, int32_t* csr_rowptr,
                    const double* rhs, double* sol,
                    void* workspace, size_t workspace_bytes, void* stream);

/* Test hook (host only, no device work): the symbolic analysis of the sparse
 * Newton solve for a sparsity pattern of W (rowptr [n+1], colj [K], columns
 * j < m-1): minimum-degree ordering of the bipartite graph, the dense core,
 * the elimination-tree levels and the gather lists of the factor and solves.
 * hdr_out receives the plan header (64-bit-aligned struct of 10 int32, 3 int64,
 * 4 int32 and 24 int64 fields, layout in csrc/cgot_symbolic.h); the plan image is copied
 * to plan_out when plan_cap is large enough.  Returns the plan size in bytes,
 * -1 when the instance keeps the dense Schur path (core above rmax or plan
 * above cap_bytes), -2 on bad arguments.  Replaces the reordering and
 * symbolic analysis of SparseCholeskySolver, src/linsolve.cu:171-176. */
int64_t cgot_sparse_plan(int32_t n, int32_t m, const int32_t* rowptr, const int32_t* colj, int64_t cap_bytes,
                         int32_t rmax, void* hdr_out, void* plan_out, int64_t plan_cap);

/* Kernel timing (HIP events around each launch on the stream it runs on), for
 * roofline figures.  Classes: 0 T pass (bytes), 1 row LSE (bytes), 2 column LSE
 * (bytes), 3 adjoint Q.v pass (bytes), 4 Cholesky trailing update (flops),
 * 5 Cholesky panel, 6 backward substitution.  While concurrent groups run
 * (below), a launch's event interval includes the other groups' overlap.  No
 * counterpart in the reference (its Timer, src/timer.h, prints host wall times). */
void cgot_profile_enable(int32_t on);
void cgot_profile_reset(void);
int cgot_profile_read(int32_t cls, double* ms_total, double* work_total, int64_t* launches);

/* Release the library's process-lifetime resources: join its worker threads,
 * wait for and destroy its internal size-class streams and free its pinned
 * host staging buffers.  Call only while no solve is running (returns -1
 * otherwise); a later solve sets them up again.  It also runs by itself at
 * process exit, before the HIP runtime's own teardown, when no solve is
 * running.  No counterpart in the reference, whose solver objects free their
 * state per call (src/sinkhorn_splr.cu, SPLRSolver destructor). */
int cgot_shutdown(void);

const char* cgot_last_error(void);
const char* cgot_version(void);

#ifdef __cplusplus
}
#endif
#endif
