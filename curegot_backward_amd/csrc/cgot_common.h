// Shared definitions for the MI355X (gfx950) entropic-OT kernels.
//
// Every kernel works on a ragged batch of OT instances.  An instance b owns a
// row-major cost matrix M_b [n_b x m_b] (FP64, caller memory) and a slice of
// one workspace arena (caller memory).  Kernels receive the descriptor table
// and either a tile->instance map (M-sized passes) or iterate one workgroup
// range per instance (vector work).  All reductions are fixed-order, so every
// result depends only on the instance's own data: an instance gives bitwise
// the same answer alone and inside any batch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "cgot_symbolic.h"

#define CG_WG 256          // threads per workgroup (4 wave64)
#define CG_WAVES 4
#define CG_CMAX 4          // columns per lane in an M-pass strip
#define CG_RPW 8           // rows per wave in an M-pass tile
#define CG_RB (CG_WAVES * CG_RPW)   // rows per M-pass tile
#define CG_NB 64           // dense Cholesky block size
#define CG_SCHUR_LDS 1024  // longest Schur row accumulated in LDS (longer rows update in place)

#define CG_DEV __device__ __forceinline__
#define CG_AS1 __attribute__((address_space(1)))   // global memory

// Per-instance descriptor (device resident, written once per solve).  The
// field list is shared by the host view (plain pointers) and the device view
// CgInstD, whose buffers are global-memory pointers so that loads and stores
// through them compile to global rather than flat instructions.  Both views
// have the same layout.
#define CG_INST_FIELDS(Q) \
                                                                                                 \
    const Q double* M; /* [n*m], unused for a geometric instance */                              \
    /* geometric cost (X != null): M_ij = cscale * sum_k (X_ik - Y_jk)^2 formed on the fly */    \
    const Q double* X; /* [n*d] */                                                               \
    const Q double* Y; /* [m*d] */                                                               \
    double cscale;                                                                               \
    int d;                                                                                       \
    int n, m, N; /* N = n+m-1 */                                                                 \
    int c; /* columns per lane (1..CG_CMAX) */                                                   \
    int sw; /* strip width 64*c */                                                               \
    int nstrips; /* ceil(m/sw) */                                                                \
    int nrb; /* ceil(n/CG_RB) */                                                                 \
    int tile0; /* first global tile of this instance */                                          \
    int ntiles; /* nrb*nstrips */                                                                \
    int vec0; /* first vector workgroup (ceil((n+m)/CG_WG) of them) */                           \
    int nvec;                                                                                    \
    int row0; /* first wave-map workgroup (ceil((n+m)/CG_WAVES) of them, one wave per index) */  \
    int nrow;                                                                                    \
    double reg, inv_reg;                                                                         \
    /* dual / gradient state, all FP64 */                                                        \
    Q double* ab; /* [n+m]   (a, b) */                                                           \
    Q double* logab; /* [n+m] */                                                                 \
    Q double* gamma; /* [n+m] */                                                                 \
    Q double* gamma_prev; /* [n+m] */                                                            \
    Q double* gamma_bcd; /* [n+m] */                                                             \
    Q double* grad; /* [N] */                                                                    \
    Q double* grad_prev; /* [N] */                                                               \
    Q double* grad_bcd; /* [N] */                                                                \
    Q double* dir; /* [N]  Newton / quasi-Newton direction */                                    \
    Q double* yv; /* [N]  secant y */                                                            \
    Q double* sv; /* [N]  secant s */                                                            \
    Q double* uv; /* [N]  A^{-1} y */                                                            \
    Q double* rowsum; /* [n] */                                                                  \
    Q double* colsum; /* [m] */                                                                  \
    Q double* rp0; /* [n*nstrips] row partials (sum, or LSE max) */                              \
    Q double* rp1; /* [n*nstrips] row partials (LSE sum) */                                      \
    Q double* cp0; /* [nrb*m]     column partials */                                             \
    Q double* cp1; /* [nrb*m] */                                                                 \
    Q double* wgpart; /* [4*max(nfin,nvec)] per-workgroup scalar partials */                     \
    Q unsigned int* counter; /* [4] last-workgroup counters */                                   \
    Q double* scal; /* [CG_NSCAL] per-instance scalars */                                        \
    /* sparsified Hessian (pattern + values), W_ij = T_ij on the pattern */                      \
    int64_t Kmax;                                                                                \
    Q int* w_rowptr; /* [n+1]   CSR of W by row i: column j */                                   \
    Q int* w_colj; /* [Kmax] */                                                                  \
    Q double* w_rval; /* [Kmax] */                                                               \
    Q int* w_colptr; /* [m]     CSC of W by column j (m-1 used): row i */                        \
    Q int* w_rowi; /* [Kmax] */                                                                  \
    Q double* w_cval; /* [Kmax] */                                                               \
    Q double* Da; /* [n]  diag(T1) + shift*reg */                                                \
    Q double* Db; /* [m-1] */                                                                    \
    /* top-K selection scratch */                                                                \
    Q unsigned int* hist; /* [CG_HBINS] */                                                       \
    Q int* flags; /* [nblk] backward-substitution block flags (epoch numbers) */                 \
    Q int* rcnt; /* [n]   selected per row */                                                    \
    Q int* rtie; /* [n]   ties per row */                                                        \
    Q int* rcut; /* [n]   tie cut column per row */                                              \
    Q int* ccnt; /* [nrb*m] selected per (row block, column) */                                  \
    /* dense Schur complement (row-major P x P, lower triangle used) */                          \
    int P; /* min(n, m-1) Schur size */                                                          \
    int beta_side; /* 1: Schur onto beta (P = m-1), 0: onto alpha (P = n) */                     \
    Q double* S; /* [P*P] */                                                                     \
    Q double* Winv; /* [nblk*CG_NB*CG_NB] inverses of the diagonal-block Cholesky factors */     \
    Q double* rhs; /* [2*P] Schur right-hand sides / solutions */                                \
    Q double* ltmp; /* [2*max(n,m)] Schur solutions */                                           \
    Q double* inv_diag; /* [N] 1/diag(A) of the context matrix */                                \
    /* backward with the SPLR preconditioner: device address of the forward */ \
    /* instance's descriptor (0 = none), Schur right-hand side and solution */ \
    unsigned long long fctx;                                                                     \
    int fsp; /* 1: the forward instance's factor is the sparse plan */                           \
    Q double* prhs; /* [P of the forward instance] */                                            \
    Q double* pltmp; /* [P] */                                                                   \
    /* sparse Newton solve: plan header (mode 1 = the plan in S is in use), */                   \
    /* and the solve work vectors of the backward's sparse preconditioner */                     \
    Q SpHdr* sph; /* [1] */                                                                      \
    Q double* spx; /* [2N] (backward with a sparse forward factor) */                            \
    /* outputs */                                                                                \
    Q double* plan; /* [n*m] or null */                                                          \
    Q double* dual; /* [n+m] or null */                                                         
struct CgInst {
    CG_INST_FIELDS()
};
struct CgInstD {
    CG_INST_FIELDS(CG_AS1)
};
static_assert(sizeof(CgInst) == sizeof(CgInstD), "descriptor views differ");

// A launch map holds one {instance, workgroup index within the instance}
// pair per workgroup of the launch (the maps may cover a subset of the batch).
CG_DEV int2 cg_map_entry(const int* map) { return reinterpret_cast<const int2*>(map)[blockIdx.x]; }

CG_DEV const CgInstD& cg_inst(const CgInst* insts, int b)
{
    return reinterpret_cast<const CgInstD*>(insts)[b];
}


// Cost entry M_ij of a geometric instance: cscale * sum_k (X_ik - Y_jk)^2,
// the terms added in k order by fused multiply-adds, then one rounded
// multiply.  Contraction is off here, so the final product is never fused
// into the caller's arithmetic and the value is the one a stored M of the
// same formula holds.
CG_DEV double cg_cost_xy(const CgInstD& I, int i, int j)
{
#pragma clang fp contract(off)
    const CG_AS1 double* x = I.X + (size_t)i * I.d;
    const CG_AS1 double* y = I.Y + (size_t)j * I.d;
    double s = 0.0;
    for (int k = 0; k < I.d; k++) {
        const double t = x[k] - y[k];
        s = fma(t, t, s);
    }
    return s * I.cscale;
}

// M_ij of any instance (stored or geometric); for the kernels that read a
// few entries.
CG_DEV double cg_cost(const CgInstD& I, int i, int j)
{
    return I.X ? cg_cost_xy(I, i, j) : I.M[(size_t)i * I.m + j];
}

// Row i of a stored cost at the lane's strip columns j = s0 + lane + 64 k (0
// where !ok[k]).  NT: non-temporal loads, for passes that stream M once (the
// forward's); the adjoint's Q v passes re-read M every PCG step and keep the
// ordinary loads, so that an M within the memory-side cache stays there.
template <int C, bool NT = true>
CG_DEV void cg_cost_row(const CgInstD& I, int i, int s0, int lane, const bool (&ok)[C], double (&mv)[C])
{
    const auto Mi = I.M + (size_t)i * I.m + s0 + lane;
#pragma unroll
    for (int k = 0; k < C; k++) {
        if constexpr (NT) mv[k] = ok[k] ? __builtin_nontemporal_load(&Mi[64 * k]) : 0.0;
        else mv[k] = ok[k] ? Mi[64 * k] : 0.0;
    }
}

// Rows min(i0 + r, n - 1), r < R, of a geometric cost at the lane's strip
// columns, with the rounding of cg_cost_xy.  i0 is the same for every lane of
// the wave (one wave's row group in every M-pass tile), so each X_i element
// is one wave-wide read and each Y element read serves the R rows.
template <int R, int C>
CG_DEV void cg_cost_rows_xy(const CgInstD& I, int i0, int n, int s0, int lane, const bool (&ok)[C],
                            double (&mv)[R][C])
{
#pragma clang fp contract(off)
    const int iu = __builtin_amdgcn_readfirstlane(i0);
    const CG_AS1 double* x[R];
#pragma unroll
    for (int r = 0; r < R; r++) x[r] = I.X + (size_t)min(iu + r, n - 1) * I.d;
    const CG_AS1 double* y[C];
#pragma unroll
    for (int k = 0; k < C; k++) y[k] = I.Y + (size_t)(ok[k] ? s0 + lane + 64 * k : 0) * I.d;
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int k = 0; k < C; k++) mv[r][k] = 0.0;
#pragma unroll 4
    for (int q = 0; q < I.d; q++) {
        double yq[C];
#pragma unroll
        for (int k = 0; k < C; k++) yq[k] = y[k][q];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const double xq = x[r][q];
#pragma unroll
            for (int k = 0; k < C; k++) {
                const double t = xq - yq[k];
                mv[r][k] = fma(t, t, mv[r][k]);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int k = 0; k < C; k++) mv[r][k] = ok[k] ? mv[r][k] * I.cscale : 0.0;
}

// Scalar slots in CgInst::scal
enum {
    CS_F = 0,      // objective at the last T pass
    CS_G2,         // ||g||^2 at the last T pass
    CS_DG,         // <g, dir> at the last T pass
    CS_DG0,        // <g_prev, dir>
    CS_YS, CS_YY,  // secant products
    CS_LOWRANK,    // 1.0 if the low-rank term is active
    CS_FBCD, CS_G2BCD,
    CS_STEP,       // trial step (host written)
    CS_SHIFT,      // shift*reg for A (host written)
    CS_KTH_HI,     // top-K threshold: key bits (as two 32-bit halves stored in doubles)
    CS_KTH_LO,
    CS_NEED,       // ties still to take
    CS_KSEL,       // K
    CS_PIVOT,      // Cholesky failure flag (first failing pivot + 1)
    CS_SG, CS_YAY, CS_YAG,
    CS_G2A,        // sum of (T1 - a)^2 at the last T pass
    // BCD stopping sums (share slots with the PCG block)
    BCD_SUMS = 24,
    // backward PCG
    PCG_STATUS = 24, PCG_BNORM, PCG_RZ, PCG_BETA, PCG_ALPHA, PCG_ITER,
    CG_NSCAL = 32
};

// PCG status values (PCG_STATUS)
#define PCG_RUNNING 0.0
#define PCG_CONVERGED 1.0
#define PCG_BREAKDOWN 2.0
#define PCG_MAXIT 3.0
#define PCG_ZERO_RHS 4.0

#define CG_HBITS0 12
#define CG_HBITS 13
#define CG_HBINS 8192

// ---------------------------------------------------------------------------
// wave64 / workgroup reductions (fixed order => deterministic)
// ---------------------------------------------------------------------------
// Broadcast lane `src` (wave-uniform) of a double.
CG_DEV double cg_readlane_d(double v, int src)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), src);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// A double moved across lanes by the DPP control CTRL (within 16-lane rows).
template <int CTRL>
CG_DEV double cg_dpp_d(double v)
{
    const long long b = __double_as_longlong(v);
    // every control used here is a full permutation inside each row, so no
    // lane keeps its old value and none needs initialising
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffll), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// Orders the LDS accesses of one wave: every lane's earlier LDS writes are
// visible to every lane's later LDS reads (no workgroup barrier).
CG_DEV void cg_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave reductions, all 64 lanes active; the result is the same in every lane.
// Inside each 16-lane row: DPP exchanges with lane^1 and lane^2 (quad
// permutes), then with the half-row and row mirrors; the four row results are
// combined by readlane.  No LDS traffic.
CG_DEV double cg_wave_sum(double v)
{
    v += cg_dpp_d<0xB1>(v);
    v += cg_dpp_d<0x4E>(v);
    v += cg_dpp_d<0x141>(v);
    v += cg_dpp_d<0x140>(v);
    return (cg_readlane_d(v, 0) + cg_readlane_d(v, 16)) + (cg_readlane_d(v, 32) + cg_readlane_d(v, 48));
}

CG_DEV double cg_wave_max(double v)
{
    v = fmax(v, cg_dpp_d<0xB1>(v));
    v = fmax(v, cg_dpp_d<0x4E>(v));
    v = fmax(v, cg_dpp_d<0x141>(v));
    v = fmax(v, cg_dpp_d<0x140>(v));
    return fmax(fmax(cg_readlane_d(v, 0), cg_readlane_d(v, 16)), fmax(cg_readlane_d(v, 32), cg_readlane_d(v, 48)));
}

// exp on the M passes: Tang's table-driven method (P. T. P. Tang, "Table-driven
// implementation of the exponential function in IEEE floating-point
// arithmetic", ACM TOMS 15(2), 1989) with 64 table entries.  y = (64q + j)
// ln2/64 + r with |r| <= ln2/128; exp(y) = 2^q 2^(j/64) e^r, e^r - 1 by its
// degree-5 Taylor polynomial (truncation below 4e-17 relative).  ln2/64 is
// split in two parts so that r is exact to rounding.  tab[j] = 2^(j/64) sits
// in LDS (CG_EXP_TABLE).  For finite y; exp of large negative y underflows to
// 0 through the final ldexp, of large positive y overflows to +inf.
#define CG_EXP_TABLE(tab)                                                       \
    __shared__ double tab[64];                                                  \
    if (threadIdx.x < 64) tab[threadIdx.x] = exp2((double)threadIdx.x * 0.015625); \
    __syncthreads()

CG_DEV double cg_texp(double y, const double* tab)
{
    const double kf = __builtin_rint(y * 92.332482616893656877);   // 64/ln2
    const int k = (int)kf;
    double r = fma(-kf, 0.010830424696223417, y);                  // ln2/64, leading bits
    r = fma(-kf, 2.572804622327669e-14, r);                        // ln2/64, remainder
    double p = fma(r, 1.0 / 120.0, 1.0 / 24.0);
    p = fma(r, p, 1.0 / 6.0);
    p = fma(r, p, 0.5);
    p = fma(r, p, 1.0);
    p = r * p;
    const double t = tab[k & 63];
    return __builtin_amdgcn_ldexp(fma(t, p, t), k >> 6);
}

// cg_texp for -745 <= y <= 710 only: k = y*64/ln2 rounded to nearest comes
// from one fma, as adding 1.5*2^52 leaves k in the low word (two's
// complement) and k - 1.5*2^52 exactly representable (|k| < 2^17 here).
CG_DEV double cg_texp_m(double y, const double* tab)
{
    const double sh = fma(y, 92.332482616893656877, 6755399441055744.0);   // 64/ln2 and 1.5*2^52
    const int k = __double2loint(sh);
    const double kf = sh - 6755399441055744.0;
    double r = fma(-kf, 0.010830424696223417, y);
    r = fma(-kf, 2.572804622327669e-14, r);
    double p = fma(r, 1.0 / 120.0, 1.0 / 24.0);
    p = fma(r, p, 1.0 / 6.0);
    p = fma(r, p, 0.5);
    p = fma(r, p, 1.0);
    p = r * p;
    const double t = tab[k & 63];
    return __builtin_amdgcn_ldexp(fma(t, p, t), k >> 6);
}

// exp(y) for y <= 0, -inf included: below -745 exp is under half the
// smallest subnormal, and the clamped argument gives 0 or that subnormal.
CG_DEV double cg_texp_lo(double y, const double* tab) { return cg_texp_m(fmax(y, -745.0), tab); }

// Sum over a 256-thread workgroup; result valid in every thread.
CG_DEV double cg_block_sum(double v, double* lds4)
{
    v = cg_wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds4[w] = v;
    __syncthreads();
    double r = ((lds4[0] + lds4[1]) + lds4[2]) + lds4[3];
    return r;
}

// Columns per lane for strip st of a pass over the columns j < mlim (m for
// the M passes, m - 1 for the top-K kernels): the instance's c, or fewer for
// a last strip that is narrower (its idle lanes would otherwise still issue
// the work).
CG_DEV int cg_strip_c(const CgInstD& I, int st, int mlim)
{
    const int cw = min(I.sw, mlim - st * I.sw);
    return (cw <= 64) ? 1 : (cw <= 128) ? 2 : CG_CMAX;
}

// Calls fn(tag) with tag::C the strip's columns per lane and tag::FULL true
// when all 64*C columns of the strip lie below mlim (every strip but, at
// most, an instance's last): the tile bodies then drop their per-column
// bounds tests, which otherwise put every load under its own exec-mask branch.
// tag::XY: the cost is formed from X, Y (geometric instances).  The M-pass
// kernels take XY as a template parameter and a launch covers instances of
// one kind, so the stored-M instantiations are unchanged by the geometric
// path.
template <int C_, bool FULL_, bool XY_ = false>
struct CgStrip {
    static constexpr int C = C_;
    static constexpr bool FULL = FULL_;
    static constexpr bool XY = XY_;
};

template <bool XY, class Fn>
CG_DEV void cg_strip_dispatch(const CgInstD& I, int st, int mlim, Fn&& fn)
{
    const int c = cg_strip_c(I, st, mlim);
    const bool full = mlim - st * I.sw >= 64 * c;
    if (c == 1) {
        if (full) fn(CgStrip<1, true, XY>{}); else fn(CgStrip<1, false, XY>{});
    } else if (c == 2) {
        if (full) fn(CgStrip<2, true, XY>{}); else fn(CgStrip<2, false, XY>{});
    } else {
        if (full) fn(CgStrip<CG_CMAX, true, XY>{}); else fn(CgStrip<CG_CMAX, false, XY>{});
    }
}


// Merge of two (max, sum-of-exp) log-sum-exp partials.
CG_DEV void cg_lse_merge(double& mx, double& s, double mx2, double s2)
{
    if (mx2 == -INFINITY) return;
    if (mx == -INFINITY) { mx = mx2; s = s2; return; }
    if (mx2 > mx) { s = s * exp(mx - mx2) + s2; mx = mx2; }
    else { s = s + s2 * exp(mx2 - mx); }
}

// sum_{k < count} p[k*stride] in index order (the same rounding as a plain
// loop), with CG_SUM_U loads in flight per round trip.
#define CG_SUM_U 16
template <int U = CG_SUM_U>
CG_DEV double cg_ordered_sum(const CG_AS1 double* p, int count, size_t stride)
{
    double s = 0.0;
    for (int k = 0; k < count; k += U) {
        double t[U];
#pragma unroll
        for (int u = 0; u < U; u++) t[u] = (k + u < count) ? p[(size_t)(k + u) * stride] : 0.0;
#pragma unroll
        for (int u = 0; u < U; u++)
            if (k + u < count) s += t[u];
    }
    return s;
}

// Two interleaved ordered sums (same index set, different arrays).
CG_DEV void cg_ordered_sum2(const CG_AS1 double* p, const CG_AS1 double* q, int count, size_t stride, double& sp,
                            double& sq)
{
    sp = 0.0;
    sq = 0.0;
    for (int k = 0; k < count; k += CG_SUM_U) {
        double a[CG_SUM_U], c[CG_SUM_U];
#pragma unroll
        for (int u = 0; u < CG_SUM_U; u++) {
            const bool ok = k + u < count;
            a[u] = ok ? p[(size_t)(k + u) * stride] : 0.0;
            c[u] = ok ? q[(size_t)(k + u) * stride] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < CG_SUM_U; u++)
            if (k + u < count) { sp += a[u]; sq += c[u]; }
    }
}

// Values handed between workgroups of one launch (the blocks of W = L^{-1} of
// a factorization, the blocks of a triangular solve's solution, the partials
// of a last-workgroup reduction) are written exactly once per use into storage
// that held CG_SENT before.
// The writer stores them with relaxed agent-scope atomics and the reader polls
// each value with relaxed agent-scope atomic loads until it differs from
// CG_SENT: every location is written once, so per-location coherence is all
// the hand-off needs (no flag, no release/acquire fences and their cache
// write-back and invalidation).  CG_SENT is a NaN payload that arithmetic
// never produces; a stored value with that bit pattern is replaced by the
// default NaN.  The poll is bounded: on overrun it sets scal[CS_PIVOT] = 2
// and returns 0.
#define CG_SENT 0x7FF4C0DE5E9D1E00ull

CG_DEV void cg_sent_store(CG_AS1 double* p, double v)
{
    unsigned long long u = (unsigned long long)__double_as_longlong(v);
    if (u == CG_SENT) u = 0x7FF8000000000000ull;
    __hip_atomic_store((CG_AS1 unsigned long long*)p, u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

CG_DEV unsigned long long cg_sent_peek(const CG_AS1 double* p)
{
    return __hip_atomic_load((const CG_AS1 unsigned long long*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

CG_DEV double cg_sent_wait(const CG_AS1 double* p, unsigned long long u, CG_AS1 double* scal)
{
    int spins = 0;
    while (u == CG_SENT) {
        __builtin_amdgcn_s_sleep(1);
        u = cg_sent_peek(p);
        if (++spins > (1 << 25)) {
            scal[CS_PIVOT] = 2.0;
            return 0.0;
        }
    }
    return __longlong_as_double((long long)u);
}

CG_DEV double cg_sent_load(const CG_AS1 double* p, CG_AS1 double* scal) { return cg_sent_wait(p, cg_sent_peek(p), scal); }

CG_DEV void cg_sent_reset(CG_AS1 double* p)
{
    __hip_atomic_store((CG_AS1 unsigned long long*)p, CG_SENT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Completes NQ hand-off reads of one thread: u[q] holds a first read of slot
// addr(q) (0 for slots not read); every slot still at CG_SENT is read again,
// all of them in flight together, until none is left (bounded like
// cg_sent_wait).  Values handed over by another workgroup land in a burst, so
// once the first has arrived the rest come in one more round trip instead of
// one round trip each.
template <int NQ, class Addr>
CG_DEV void cg_sent_gather(unsigned long long (&u)[NQ], Addr addr, CG_AS1 double* scal)
{
    int spins = 0;
    for (;;) {
        bool pend = false;
#pragma unroll
        for (int q = 0; q < NQ; q++) pend = pend || (u[q] == CG_SENT);
        if (!pend) return;
        if (++spins > (1 << 25)) {
            scal[CS_PIVOT] = 2.0;
#pragma unroll
            for (int q = 0; q < NQ; q++)
                if (u[q] == CG_SENT) u[q] = 0ull;
            return;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int q = 0; q < NQ; q++)
            if (u[q] == CG_SENT) u[q] = cg_sent_peek(addr(q));
    }
}

// Ordered sum of the hand-off slots p[0..count) once their values have
// arrived, every slot reset to CG_SENT for the next use; read by a whole wave
// (every lane returns the sum), one slot per lane in flight, and folded in
// index order from 0.0 through lane reads (the rounding of a plain loop).
// Call with the whole wave active.
CG_DEV double cg_ordered_take_wave(CG_AS1 double* p, int count, CG_AS1 double* scal)
{
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    for (int k = 0; k < count; k += 64) {
        const bool ok = k + lane < count;
        unsigned long long t[1] = {ok ? cg_sent_peek(p + k + lane) : 0ull};
        cg_sent_gather(t, [&](int) { return p + k + lane; }, scal);
        if (ok) cg_sent_reset(p + k + lane);
        const int lo = (int)(unsigned)t[0], hi = (int)(unsigned)(t[0] >> 32);
        const int cnt = min(64, count - k);
        for (int j = 0; j < cnt; j++) {
            const unsigned long long w = ((unsigned long long)(unsigned)__builtin_amdgcn_readlane(hi, j) << 32) |
                                         (unsigned)__builtin_amdgcn_readlane(lo, j);
            s += __longlong_as_double((long long)w);
        }
    }
    return s;
}

// Per-instance sums of NV per-thread values over all vec workgroups of the
// instance (wgi = this workgroup's index in it).  Every workgroup hands its
// block sums over through wgpart (CG_SENT-initialised) and counts itself in
// counter[slot] (relaxed); the last to arrive sums the partials in workgroup
// order, so the result does not depend on scheduling, and resets wgpart and the
// counter.  Returns true in that workgroup, where out[] holds the totals (all
// threads).
template <int NV>
CG_DEV bool cg_inst_reduce(const CgInstD& I, int wgi, double (&v)[NV], double (&out)[NV], int slot)
{
    __shared__ double lds4[4];
    __shared__ double tot[NV];
    __shared__ int is_last;
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = cg_block_sum(v[k], lds4);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) cg_sent_store(I.wgpart + (size_t)k * I.nvec + wgi, v[k]);
        const unsigned int prev = __hip_atomic_fetch_add(&I.counter[slot], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (prev == (unsigned int)(I.nvec - 1));
    }
    __syncthreads();
    if (!is_last) return false;
    for (int k = threadIdx.x >> 6; k < NV; k += CG_WAVES) {   // a wave per value
        const double t = cg_ordered_take_wave(I.wgpart + (size_t)k * I.nvec, I.nvec, I.scal);
        if ((threadIdx.x & 63) == 0) tot[k] = t;
    }
    if (threadIdx.x == 0) __hip_atomic_store(&I.counter[slot], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) out[k] = tot[k];
    return true;
}


// Log-sum-exp of count (max, sum) partials (pm[k*stride], ps[k*stride]) as
// one (mx, s) pair, partials taken in index order CG_SUM_U at a time: the
// loads of a block are in flight together, the block is reduced against its
// own maximum (independent exps), and the block's pair is merged into the
// running one (one more exp).  Empty partials (pm = -inf, ps = 0) contribute
// nothing.  tab: CG_EXP_TABLE.
CG_DEV void cg_ordered_lse(const CG_AS1 double* pm, const CG_AS1 double* ps, int count, size_t stride, double& mx,
                           double& s, const double* tab)
{
    mx = -INFINITY;
    s = 0.0;
    for (int k = 0; k < count; k += CG_SUM_U) {
        double a[CG_SUM_U], c[CG_SUM_U];
#pragma unroll
        for (int u = 0; u < CG_SUM_U; u++) {
            const bool ok = k + u < count;
            a[u] = ok ? pm[(size_t)(k + u) * stride] : -INFINITY;
            c[u] = ok ? ps[(size_t)(k + u) * stride] : 0.0;
        }
        double bm = a[0];
#pragma unroll
        for (int u = 1; u < CG_SUM_U; u++) bm = fmax(bm, a[u]);
        if (bm == -INFINITY) continue;
        double e[CG_SUM_U];
#pragma unroll
        for (int u = 0; u < CG_SUM_U; u++) e[u] = c[u] * cg_texp_lo(a[u] - bm, tab);
        double bs = e[0];
#pragma unroll
        for (int u = 1; u < CG_SUM_U; u++) bs += e[u];
        // merge (bm, bs) into (mx, s); mx = -inf gives d = +inf, f = 0, s = bs
        const double d = bm - mx;
        const double f = cg_texp_lo(-fabs(d), tab);
        s = (d > 0.0) ? fma(s, f, bs) : fma(bs, f, s);
        mx = fmax(mx, bm);
    }
}

// Order-preserving key of a non-negative double (T >= 0).
CG_DEV unsigned long long cg_key(double v) { return (unsigned long long)__double_as_longlong(v); }
