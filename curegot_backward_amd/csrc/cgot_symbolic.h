// Sparse Newton-system plan: symbolic analysis of A = [[Da, W], [W^T, Db]]
// (host) and the layout the single-workgroup factor/solve kernel reads
// (device).  Shared by the host driver and the kernels.
//
// The vertices of A are v in [0, N), N = n + m - 1: v < n is alpha_v, v >= n is
// beta_{v-n}.  A minimum-degree elimination of the bipartite graph of W orders
// them; it stops once every remaining vertex is adjacent to at least half of
// the others.  Those, and the last eliminated vertices where moving them is
// cheaper (the top of the elimination tree), form the dense core of R
// vertices.  Positions
// [0, t) are the periphery (sparse columns of L, in elimination order),
// [t, N) the core (t = N - R).
//
// Periphery columns are stored in level order of the elimination tree
// (leaves first; a column's level is one more than its deepest child's), each
// column as its diagonal followed by its sub-diagonal entries in increasing
// row order.  Every entry q of L is one gather
//     L[q] = A(q) - sum over its pair list of L[x] * L[y]
// followed by the square root (diagonal) or the division by its column's
// diagonal.  The forward substitution is the same gather
// on two extra rows (the right-hand sides): y_j = (b_j - sum L(j,k) y_k) / L_jj
// over the row-pair list of j, computed with the entries of j's level.  The
// core's Schur complement C = A22 - L21 L21^T and its right-hand sides
// b_core - L21 y are gathered the same way (C packed lower, C(i, j) at
// i(i+1)/2 + j) and factored densely in registers (4 x 4 tiles per thread).
// Each level has two phases: the gathers (diagonal entries take their square
// root), then the divisions by the column diagonals (off-diagonal entries and
// the right-hand sides).
#pragma once

#include <stdint.h>

#ifdef __cplusplus
#include <cstddef>
#include <vector>
#endif

// A-value source codes of a factor entry: -1 fill (0), v in [0, N) the
// diagonal of vertex v (Da[v] or Db[v-n]), N + q the entry q of W's row CSR
// (w_rval[q]).
struct SpHdr {
    int32_t mode;        // 1: this plan is in use (0: the dense Schur path)
    int32_t N, n, m;
    int32_t t, R, H;     // periphery size, core size, periphery levels
    int32_t nnzP;        // entries of the periphery columns of L
    int32_t ncore;       // R (R + 1) / 2
    int32_t T;           // core tiles per dimension, ceil(R / 4)
    int64_t npairs, ncpairs, nrowpairs;
    int32_t nlong, nclong, nylong, nblong;   // lists summed by eight-lane groups
    // byte offsets from the plan base (the instance's Schur buffer); the
    // integer arrays come first, in [0, o_meta_end)
    int64_t o_perm;      // int [N]        vertex at each position
    int64_t o_lvptr;     // int [H + 1]    entry range of each level
    int64_t o_lcptr;     // int [H + 1]    column range (level order) of each level
    int64_t o_lcols;     // int [t + 1]    first entry (the diagonal) of each column, level order; [t] = nnzP
    int64_t o_emeta;     // int4 [nnzP]    {A-value source, pair range begin, end, its column's diagonal entry
                         //                 (| SP_LONG_BIT when the pair list is long)}
    int64_t o_erow;      // int [nnzP]     row position
    int64_t o_pairs;     // int2 [npairs]  (entry, entry)
    int64_t o_ymeta;     // int4 [t]       per column (level order): {vertex, row-pair range begin, end,
                         //                 diagonal (| SP_LONG_BIT when long)}
    int64_t o_ycore;     // int4 [R]       per core row: {vertex, row-pair range begin, end, 1 when long}
    int64_t o_rowpairs;  // int2 [nrowpairs] (entry L(j, k), position k) of the periphery entries of each row
    int64_t o_cmeta;     // int4 [ncore]   {A-value source, pair range begin, end, 1 when long}
    int64_t o_cpairs;    // int2 [ncpairs]
    int64_t o_lgptr;     // int [H + 1]    range in `long` of each level
    int64_t o_long;      // int [nlong]    long items of the factor levels: entry q >= 0, column c as -(c + 1)
    int64_t o_clong;     // int [nclong]   long core entries (packed index)
    int64_t o_ylong;     // int [nylong]   long core rows
    int64_t o_blptr;     // int [H + 1]    range in `blong` of each level
    int64_t o_blong;     // int [nblong]   long columns of the backward substitution (level order index)
    int64_t o_cflags;    // int [t]        per column (level order): 1 when its backward sum is long
    int64_t o_meta_end;
    int64_t o_lval;      // double [nnzP]  periphery columns of L
    int64_t o_lcore;     // double [ncore] core factor (packed lower)
    int64_t o_yx;        // double [2N]    solve work vectors (two right-hand sides), by position
    int64_t bytes;       // plan size
};

// Largest dense core: its lower triangle in 4 x 4 tiles, at most two tiles
// per thread of the SP_WG-thread workgroup (T (T + 1) / 2 <= 2 SP_WG).
#define SP_WG 1024
#define SP_TILES_PER_THREAD 2
#define SP_RMAX 252
// Cores of SP_SMALL_MAX < R <= SP_RMAX_BIG (<= SP_WG: one core row per
// thread) are factored in place in the plan's packed core storage instead:
// right-looking Cholesky by panels of 8 columns (the 8 x 8 diagonal block on
// one wave, the panel's other rows by their threads, the trailing triangle by
// all waves), the right-hand sides eliminated with the panels, then the
// backward substitution by blocks of 8 rows.
#ifndef SP_SMALL_MAX
#define SP_SMALL_MAX SP_RMAX
#endif
#define SP_RMAX_BIG 768
// Pair lists longer than these are summed by eight lanes (fixed-order
// reduction) instead of one thread: SP_LONG for the periphery entries and the
// right-hand-side rows (few items per level, so latency decides), SP_CLONG
// for the core's Schur-complement entries (R (R + 1) / 2 items, throughput
// decides) and SP_BLONG for the backward substitution's columns.  A long
// periphery item carries SP_LONG_BIT in its record.
#define SP_LONG 1
#define SP_CLONG 8
#define SP_BLONG 8
#define SP_LONG_BIT 0x40000000
// Cost of one periphery level in units of one dense core column, for the
// choice of the periphery / core split.
#define SP_LEVEL_COST 4.0

// Latency estimates (microseconds per Newton solve of one instance), fitted
// to measured MI355X timings, that decide whether an instance takes its sparse
// plan.  Sparse: a fixed part, one barrier pair per periphery level, the dense
// core (its Schur gather and factorization: linear and quadratic parts, one
// fit per core kind) and the prefetch of the plan (about n + m).  Dense Schur chain: a fixed part (Schur build, right-hand
// sides, triangular solves, back-substitution) and one dependent update
// launch per CG_NB-block of the order P.
inline double sp_cost_us(const SpHdr& h)
{
    const double R = h.R;
    const double core = (h.R <= SP_SMALL_MAX) ? 0.34 * R + 8.6e-3 * R * R : 0.34 * R + 7.0e-3 * R * R;
    return 10.7 + 6.44 * h.H + core + 20.4e-3 * (double)(h.n + h.m);
}
inline double sp_dense_cost_us(int nblk) { return 150.0 + 32.0 * nblk; }

// Arithmetic of one factorization and solve with a plan (two right-hand
// sides), counting an FMA as two flops: the periphery gathers (one FMA per
// pair) and their square roots or divisions, the core's Schur complement
// gathers, the dense core's right-looking elimination (about R^3 / 3) and
// scaling, and the forward and backward substitutions of both right-hand
// sides through the periphery and the core.
inline double sp_solve_flops(const SpHdr& h)
{
    const double R = h.R, P = h.nnzP, t = h.t;
    return 2.0 * (double)h.npairs + P + 2.0 * (double)h.ncpairs + R * (R - 1.0) * (R + 1.0) / 3.0 +
           R * (R + 1.0) / 2.0 + 4.0 * (double)h.nrowpairs + 4.0 * R * (R - 1.0) + 4.0 * (P - t) + 4.0 * h.N;
}

#ifdef __cplusplus
// Host analysis.  Returns true and fills `plan` (the integer part of the
// plan image, hdr.o_meta_end bytes, offsets relative to its start; the value
// arrays up to hdr.bytes follow on the device) when a sparse plan fits in
// `cap_bytes` with a core of at most `rmax` (<= SP_RMAX_BIG); false otherwise
// (then the instance keeps the dense Schur path).  rowptr/colj: W's pattern
// by row (columns j < m - 1).
bool sp_analyze(int n, int m, const int* rowptr, const int* colj, int64_t cap_bytes, int rmax, SpHdr& hdr,
                std::vector<char>& plan);

// Doubles of the core's LDS buffers (the values of an LDS-resident launch
// follow them).  Register-tile core: two column broadcast buffers
// (4 ceil(R/4) + 2 <= R + 5 each), the pivots and two right-hand sides.
// Panel core: the panel's rows (8 R), two right-hand sides, the diagonal
// block (64), its right-hand sides (16) and two block buffers (2 x 80).
#define SP_LDS_CORE(R) ((R) <= SP_SMALL_MAX ? 5 * (R) + 16 : 10 * (R) + 240)
// LDS bytes the factor/solve kernel needs for a core of R.
inline size_t sp_lds_bytes(int R) { return sizeof(double) * (size_t)SP_LDS_CORE(R); }
// LDS of a launch that keeps the plan's values in LDS: the core's buffers, the
// nnzP periphery factor entries and the two work vectors of N
inline size_t sp_lds_vals_bytes(const SpHdr& h)
{
    return sp_lds_bytes(h.R) + sizeof(double) * ((size_t)h.nnzP + 2 * (size_t)h.N);
}
// ... and with the plan's integer arrays after the values (16-byte aligned)
inline size_t sp_lds_all_bytes(const SpHdr& h)
{
    return sp_lds_bytes(h.R) + sizeof(double) * (((size_t)h.nnzP + 2 * (size_t)h.N + 1) & ~(size_t)1) +
           (size_t)h.o_meta_end;
}
#endif
