// Sparse Newton-system plan: symbolic analysis of A = [[Da, W], [W^T, Db]]
// (host) and the layout the single-workgroup factor/solve kernel reads
// (device).  Shared by the host driver and the kernels.
//
// The vertices of A are v in [0, N), N = n + m - 1: v < n is alpha_v, v >= n is
// beta_{v-n}.  A minimum-degree elimination of the bipartite graph of W orders
// them; it stops once every remaining vertex is adjacent to at least half of
// the others, and those R remaining vertices form the dense core.  Positions
// [0, t) are the periphery (sparse columns of L, in elimination order),
// [t, N) the core (t = N - R).
//
// Periphery columns are stored in level order of the elimination tree
// (leaves first; a column's level is one more than its deepest child's), each
// column as its diagonal followed by its sub-diagonal entries in increasing
// row order.  Every entry q of L is computed by a gather:
//     L[q] = A(q) - sum over its pair list of L[x] * L[y]
// then the square root (diagonal) or the division by its column's diagonal.
// The core's Schur complement C = A22 - L21 L21^T is gathered the same way
// (packed lower triangle, C(i, j) at i(i+1)/2 + j) and factored densely.
#pragma once

#include <stdint.h>

#ifdef __cplusplus
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
    int32_t maxlev;      // largest number of entries of one level
    int64_t npairs, ncpairs;
    // byte offsets from the plan base (the instance's Schur buffer)
    int64_t o_perm;      // int [N]       vertex at each position
    int64_t o_lvptr;     // int [H + 1]   entry range of each level
    int64_t o_lcptr;     // int [H + 1]   column range (in lcols) of each level
    int64_t o_lcols;     // int [t + 1]   first entry (the diagonal) of each column, level order; [t] = nnzP
    int64_t o_emeta;     // int4 [nnzP]   {A-value source, pair range begin, end, its column's diagonal entry}
    int64_t o_erow;      // int [nnzP]    row position
    int64_t o_ecol;      // int [nnzP]    column position
    int64_t o_pairs;     // int2 [npairs] (entry, entry)
    int64_t o_cmeta;     // int4 [ncore]  {A-value source, pair range begin, end, 0}
    int64_t o_cpairs;    // int2 [ncpairs]
    int64_t o_rptr;      // int [N + 1]   per row position: its off-diagonal periphery entries
    int64_t o_rq;        // int [nnzP - t]
    int64_t o_lval;      // double [nnzP] periphery columns of L
    int64_t o_lcore;     // double [ncore] core factor (packed lower)
    int64_t o_yx;        // double [2N]   solve work vectors (two right-hand sides), by position
    int64_t bytes;       // plan size
};

// Largest core factored in one workgroup's LDS (packed lower triangle:
// 196 * 197 / 2 doubles = 151 KiB of the 160 KiB of a gfx950 workgroup).
#define SP_RMAX 196
#define SP_WG 1024

#ifdef __cplusplus
// Host analysis.  Returns true and fills `plan` (a byte image whose offsets
// are relative to its start; hdr.bytes long) when a sparse plan fits in
// `cap_bytes` with a core of at most `rmax` (<= SP_RMAX); false otherwise
// (then the instance keeps the dense Schur path).  rowptr/colj: W's pattern
// by row (columns j < m - 1).
bool sp_analyze(int n, int m, const int* rowptr, const int* colj, int64_t cap_bytes, int rmax, SpHdr& hdr,
                std::vector<char>& plan);

// LDS bytes the factor/solve kernel needs for a core of R.
inline size_t sp_lds_bytes(int R) { return sizeof(double) * ((size_t)R * (R + 1) / 2 + 2 * (size_t)R + 8); }
#endif
