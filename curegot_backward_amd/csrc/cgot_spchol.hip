// Sparse Newton solve of the SPLR step: factor A = [[Da, W], [W^T, Db]] in the
// plan's elimination order and solve A [dir, u] = [g, y], one workgroup per
// instance (plan layout and analysis: cgot_symbolic.h / .cpp).
//
// Reference: the cuDSS factorization and solve of SparseCholeskySolver
// (src/linsolve.cu:27-268) called from SPLRSolver::compute_search_direc
// (src/sinkhorn_splr.cu:538-649), which factor the same matrix every
// iteration.  Here:
//   * the plan's integer arrays are first streamed through the cache with
//     coalesced loads, so that the dependent gathers below hit L2;
//   * periphery, per elimination-tree level (workgroup barriers between
//     phases): every factor entry of the level is one gather
//     s = A(q) - sum L[x] L[y] over its pair list (diagonal entries take the
//     square root), every column of the level gathers its forward
//     substitution b_j - sum L(j,k) y_k for both right-hand sides; then the
//     divisions by the column diagonals.  Short lists are summed by one
//     thread (four items interleaved, so that their loads overlap), lists
//     longer than SP_LONG by eight lanes and a fixed-order reduction;
//   * dense core (R <= SP_RMAX): its Schur complement C = A22 - L21 L21^T and
//     right-hand sides b_core - L21 y are gathered item-parallel into scratch,
//     then held in registers, 4 x 4 tiles of the lower triangle per thread (at
//     most two) and one core row per thread; right-looking elimination
//     C(i,j) -= C(i,k) (C(j,k) / C(k,k)) with column k broadcast through LDS
//     (one barrier per column), the right-hand sides eliminated as two extra
//     rows (the forward substitution); then the scaled factor L = C D^{-1/2}
//     and the backward substitution L^T x = y by tile rows;
//   * periphery backward substitution by level in reverse (long columns by
//     eight lanes).
// Every sum runs in the fixed order of its list, so the result depends only
// on the instance's own data.
#include "cgot_kernels.h"
#include "cgot_schur_rows.h"

namespace {

// PV: the type of the value pointers (lval, yx), PI of the plan's integer
// arrays: global memory, or LDS for plans that fit beside the core's buffers.
template <class PV, class PI = const CG_AS1 int*>
struct SpViewT {
    int N, n, t, R, H, ncore, T;
    int64_t meta_ints;   // integer words in [0, o_meta_end)
    PI meta;
    PI perm;
    PI lvptr;
    PI lcptr;
    PI lcols;
    PI emeta;      // 4 ints per entry
    PI erow;
    PI pairs;      // 2 ints per pair
    PI ymeta;      // 4 ints per column
    PI ycore;      // 4 ints per core row
    PI rowpairs;   // 2 ints per pair
    PI cmeta;      // 4 ints per core entry
    PI cpairs;
    PI lgptr;
    PI longs;
    PI clong;
    PI ylong;
    PI blptr;
    PI blong;
    PI cflags;
    int nclong, nylong;
    PV lval;
    CG_AS1 double* lcore;
    PV yx;
};
using SpView = SpViewT<CG_AS1 double*>;

CG_DEV SpView sp_view(const CG_AS1 SpHdr* hp, CG_AS1 double* plan)
{
    CG_AS1 char* base = reinterpret_cast<CG_AS1 char*>(plan);
    auto ip = [&](int64_t off) { return reinterpret_cast<const CG_AS1 int*>(base + off); };
    auto dp = [&](int64_t off) { return reinterpret_cast<CG_AS1 double*>(base + off); };
    SpView v;
    v.N = hp->N; v.n = hp->n; v.t = hp->t; v.R = hp->R; v.H = hp->H; v.ncore = hp->ncore; v.T = hp->T;
    v.meta_ints = hp->o_meta_end / 4;
    v.meta = ip(0);
    v.perm = ip(hp->o_perm);
    v.lvptr = ip(hp->o_lvptr);
    v.lcptr = ip(hp->o_lcptr);
    v.lcols = ip(hp->o_lcols);
    v.emeta = ip(hp->o_emeta);
    v.erow = ip(hp->o_erow);
    v.pairs = ip(hp->o_pairs);
    v.ymeta = ip(hp->o_ymeta);
    v.ycore = ip(hp->o_ycore);
    v.rowpairs = ip(hp->o_rowpairs);
    v.cmeta = ip(hp->o_cmeta);
    v.cpairs = ip(hp->o_cpairs);
    v.lgptr = ip(hp->o_lgptr);
    v.longs = ip(hp->o_long);
    v.clong = ip(hp->o_clong);
    v.ylong = ip(hp->o_ylong);
    v.blptr = ip(hp->o_blptr);
    v.blong = ip(hp->o_blong);
    v.cflags = ip(hp->o_cflags);
    v.nclong = hp->nclong;
    v.nylong = hp->nylong;
    v.lval = dp(hp->o_lval);
    v.lcore = dp(hp->o_lcore);
    v.yx = dp(hp->o_yx);
    return v;
}

// Streams the plan's integer arrays through the cache (coalesced, 16 bytes
// per thread and load, SP_PF loads per thread in flight before any is used);
// the returned value only keeps the loads alive.
#define SP_PF 8
template <class V>
CG_DEV unsigned sp_prefetch(const V& v)
{
    unsigned acc = 0;
    const int64_t n4 = v.meta_ints & ~(int64_t)3;
    const CG_AS1 unsigned long long* m8 = reinterpret_cast<const CG_AS1 unsigned long long*>(v.meta);
    const int64_t nv = n4 / 4;   // 16-byte words
    if (nv == 0) return 0u;
    for (int64_t base = threadIdx.x; base < nv; base += (int64_t)SP_PF * SP_WG) {
        // indices past the end load the last word again (no branch, so that
        // all SP_PF loads are issued before the first wait)
        unsigned long long lo[SP_PF], hi[SP_PF];
#pragma unroll
        for (int u = 0; u < SP_PF; u++) {
            const int64_t k = min(base + (int64_t)u * SP_WG, nv - 1);
            lo[u] = m8[2 * k];
            hi[u] = m8[2 * k + 1];
        }
#pragma unroll
        for (int u = 0; u < SP_PF; u++) {
            const unsigned long long x = lo[u] ^ hi[u];
            acc ^= (unsigned)x ^ (unsigned)(x >> 32);
        }
    }
    return acc;
}

// Entry of A by source code: -1 fill, [0, n) Da, [n, N) Db, N + q W's row-CSR entry q.
CG_DEV double sp_aval(const CgInstD& I, int s, int N)
{
    if (s < 0) return 0.0;
    if (s < I.n) return I.Da[s];
    if (s < N) return I.Db[s - I.n];
    return I.w_rval[s - N];
}

// Sum of an 8-lane group (lanes 8g .. 8g+7 of the wave): exchanges with
// lane ^ 4, ^ 2, ^ 1; every lane of the group ends with the same value.  All
// 64 lanes take part.
CG_DEV double group8_sum(double v)
{
    v += __shfl_xor(v, 4, 8);
    v += __shfl_xor(v, 2, 8);
    v += __shfl_xor(v, 1, 8);
    return v;
}

// Up to four independent gathers s[k] -= sum over [p0[k], p1[k]) of
// X[x] * Y0[y] (and t[k] -= X[x] * Y1[y] with two right-hand sides), in list
// order, interleaved so that the four items' loads are in flight together.
template <int NR, class PP, class PX, class PY>
CG_DEV void gather4(PP pairs, PX X, PY Y0, PY Y1, const bool (&on)[4], const int (&p0)[4],
                    const int (&p1)[4], double (&s)[4], double (&t)[4])
{
    int len = 0;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (on[k]) len = max(len, p1[k] - p0[k]);
    for (int i = 0; i < len; i++) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int p = p0[k] + i;
            if (on[k] && p < p1[k]) {
                const int x = pairs[2 * p], y = pairs[2 * p + 1];
                const double lx = X[x];
                s[k] -= lx * Y0[y];
                if (NR > 1) t[k] -= lx * Y1[y];
            }
        }
    }
}

CG_DEV int tri(int i) { return i * (i + 1) / 2; }

// Row and column of the u-th entry of a packed lower triangle (row-major).
CG_DEV void tri_decode(int u, int& i, int& j)
{
    int r = (int)((sqrtf(8.0f * (float)u + 1.0f) - 1.0f) * 0.5f);
    while (tri(r) > u) r--;
    while (tri(r + 1) <= u) r++;
    i = r;
    j = u - tri(r);
}

// c[a][b] of a register tile with a runtime column index b (selects, no
// dynamic register indexing)
CG_DEV double tile_col(const double (&c)[4][4], int a, int b)
{
    return (b == 0) ? c[a][0] : (b == 1) ? c[a][1] : (b == 2) ? c[a][2] : c[a][3];
}

// The core tiles a thread owns: tile u = tid + s * SP_WG (s < SP_TILES_PER_THREAD)
// of the packed lower triangle of T x T tiles.
struct CoreTiles {
    bool own[SP_TILES_PER_THREAD];
    int ti[SP_TILES_PER_THREAD], tj[SP_TILES_PER_THREAD];
    double c[SP_TILES_PER_THREAD][4][4];
};

CG_DEV void core_tiles_init(CoreTiles& ct, int T)
{
    const int ntl = T * (T + 1) / 2;
#pragma unroll
    for (int s = 0; s < SP_TILES_PER_THREAD; s++) {
        const int u = threadIdx.x + s * SP_WG;
        ct.own[s] = u < ntl;
        ct.ti[s] = ct.tj[s] = 0;
        if (ct.own[s]) tri_decode(u, ct.ti[s], ct.tj[s]);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) ct.c[s][a][b] = 0.0;
    }
}

// Backward substitution L^T x = w (L the scaled core factor in the tiles, w
// in LDS, NR right-hand sides), by tile rows from the last.
template <int NR>
CG_DEV void core_bwd_tiles(const CoreTiles& ct, int R, int T, double* w0, double* w1)
{
    for (int tb = T - 1; tb >= 0; tb--) {
#pragma unroll
        for (int s = 0; s < SP_TILES_PER_THREAD; s++) {
            if (!(ct.own[s] && ct.ti[s] == tb && ct.tj[s] == tb)) continue;
            double x0[4], x1[4];
#pragma unroll
            for (int a = 3; a >= 0; a--) {
                x0[a] = x1[a] = 0.0;
                const int i = 4 * tb + a;
                if (i >= R) continue;
                double s0 = w0[i], s1 = (NR > 1) ? w1[i] : 0.0;
#pragma unroll
                for (int a2 = a + 1; a2 < 4; a2++) {
                    if (4 * tb + a2 < R) {
                        s0 -= ct.c[s][a2][a] * x0[a2];
                        if (NR > 1) s1 -= ct.c[s][a2][a] * x1[a2];
                    }
                }
                x0[a] = s0 / ct.c[s][a][a];
                w0[i] = x0[a];
                if (NR > 1) {
                    x1[a] = s1 / ct.c[s][a][a];
                    w1[i] = x1[a];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < SP_TILES_PER_THREAD; s++) {
            if (!(ct.own[s] && ct.ti[s] == tb && ct.tj[s] < tb)) continue;
            double x0[4], x1[4];
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const int i = 4 * tb + a;
                x0[a] = (i < R) ? w0[i] : 0.0;
                x1[a] = (NR > 1 && i < R) ? w1[i] : 0.0;
            }
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int j = 4 * ct.tj[s] + b;
                double s0 = w0[j], s1 = (NR > 1) ? w1[j] : 0.0;
#pragma unroll
                for (int a = 0; a < 4; a++) {
                    s0 -= ct.c[s][a][b] * x0[a];
                    if (NR > 1) s1 -= ct.c[s][a][b] * x1[a];
                }
                w0[j] = s0;
                if (NR > 1) w1[j] = s1;
            }
        }
        __syncthreads();
    }
}

// Forward substitution L y = w by tile columns (NR right-hand sides in LDS).
template <int NR>
CG_DEV void core_fwd_tiles(const CoreTiles& ct, int R, int T, double* w0, double* w1)
{
    for (int tb = 0; tb < T; tb++) {
#pragma unroll
        for (int s = 0; s < SP_TILES_PER_THREAD; s++) {
            if (!(ct.own[s] && ct.ti[s] == tb && ct.tj[s] == tb)) continue;
            double x0[4], x1[4];
#pragma unroll
            for (int b = 0; b < 4; b++) {
                x0[b] = x1[b] = 0.0;
                const int j = 4 * tb + b;
                if (j >= R) continue;
                double s0 = w0[j], s1 = (NR > 1) ? w1[j] : 0.0;
#pragma unroll
                for (int b2 = 0; b2 < b; b2++) {
                    s0 -= ct.c[s][b][b2] * x0[b2];
                    if (NR > 1) s1 -= ct.c[s][b][b2] * x1[b2];
                }
                x0[b] = s0 / ct.c[s][b][b];
                w0[j] = x0[b];
                if (NR > 1) {
                    x1[b] = s1 / ct.c[s][b][b];
                    w1[j] = x1[b];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < SP_TILES_PER_THREAD; s++) {
            if (!(ct.own[s] && ct.tj[s] == tb && ct.ti[s] > tb)) continue;
            double x0[4], x1[4];
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int j = 4 * tb + b;
                x0[b] = (j < R) ? w0[j] : 0.0;
                x1[b] = (NR > 1 && j < R) ? w1[j] : 0.0;
            }
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const int i = 4 * ct.ti[s] + a;
                if (i >= R) continue;
                double s0 = w0[i], s1 = (NR > 1) ? w1[i] : 0.0;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    s0 -= ct.c[s][a][b] * x0[b];
                    if (NR > 1) s1 -= ct.c[s][a][b] * x1[b];
                }
                w0[i] = s0;
                if (NR > 1) w1[i] = s1;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Panel core (SP_SMALL_MAX < R <= SP_RMAX_BIG): the core matrix stays in the
// plan's packed storage lcore (row-major lower triangle, entry (i, j) at
// tri(i) + j) and is factored in place; thread i < R owns core row i and its
// right-hand-side entries.  LDS (doubles): the panel's columns [0, 8R)
// (column c of row i at c R + i, so that lanes over rows read consecutive
// words), the two right-hand sides [8R, 10R), the 8 x 8 diagonal block, its
// right-hand sides, two block buffers.
// ---------------------------------------------------------------------------
// The value of x in lane `src` of the wave (a compile-time lane here), as a
// wave-uniform value: two readlane operations, no LDS round trip.
CG_DEV double lane_value(double x, int src)
{
    const unsigned long long b = __double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, src);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), src);
    return __longlong_as_double(((unsigned long long)hi << 32) | lo);
}

struct BigLds {
    double *pan, *w0, *w1, *dblk, *yb, *xb;
};
CG_DEV BigLds big_lds(double* lds, int R)
{
    BigLds b;
    b.pan = lds;
    b.w0 = lds + 8 * R;
    b.w1 = b.w0 + R;
    b.dblk = b.w1 + R;   // 64
    b.yb = b.dblk + 64;  // 16: the block's two right-hand sides
    b.xb = b.yb + 16;    // 2 x 80: block buffers of the substitutions
    return b;
}

// Backward substitution L^T x = y of the panel core, by blocks of 8 rows from
// the last: the block's rows publish their y (less the contributions of the
// blocks below) to LDS; every thread above the block solves the block's
// 8 x 8 triangle (the same arithmetic in every thread) and removes the
// block's contribution from its own row; the block's threads keep their x.
// One barrier per block (the block buffers alternate).  r0 / r1: this
// thread's entries; par: the next buffer's parity.
template <int NR>
CG_DEV void big_bwd(const CG_AS1 double* lcore, int R, const BigLds& L, double& r0, double& r1, int& par)
{
    const int tid = threadIdx.x;
    const bool own = tid < R;
    for (int kb = (R + 7) / 8 - 1; kb >= 0; kb--) {
        const int b0 = 8 * kb, nb = min(8, R - b0);
        double* xs = L.xb + 80 * par;   // [0, 16): right-hand sides, [16, 80): the block's factor
        double* lb = xs + 16;
        par ^= 1;
        const int me = tid - b0;
        if (own && me >= 0 && me < nb) {
            xs[me] = r0;
            if (NR > 1) xs[8 + me] = r1;
#pragma unroll
            for (int c = 0; c < 8; c++)
                if (c <= me) lb[8 * me + c] = lcore[tri(tid) + b0 + c];
        }
        // this row's entries in the block's rows (needed after the block's solve)
        double lr[8];
#pragma unroll
        for (int a = 0; a < 8; a++) lr[a] = (own && me < 0 && a < nb) ? lcore[tri(b0 + a) + tid] : 0.0;
        __syncthreads();
        if (own && me < nb) {
            // the thread's own entry is taken as it is computed (a compare per
            // entry: no runtime index into the arrays)
            double x0[8], x1[8];
#pragma unroll
            for (int a = 7; a >= 0; a--) {
                x0[a] = x1[a] = 0.0;
                if (a >= nb) continue;
                double s0 = xs[a], s1 = (NR > 1) ? xs[8 + a] : 0.0;
#pragma unroll
                for (int a2 = a + 1; a2 < 8; a2++) {
                    if (a2 >= nb) continue;
                    const double l = lb[8 * a2 + a];
                    s0 -= l * x0[a2];
                    if (NR > 1) s1 -= l * x1[a2];
                }
                const double d = lb[8 * a + a];
                x0[a] = s0 / d;
                if (NR > 1) x1[a] = s1 / d;
                if (me == a) {
                    r0 = x0[a];
                    if (NR > 1) r1 = x1[a];
                }
            }
            if (me < 0) {
#pragma unroll
                for (int a = 0; a < 8; a++) {
                    if (a >= nb) continue;
                    r0 -= lr[a] * x0[a];
                    if (NR > 1) r1 -= lr[a] * x1[a];
                }
            }
        }
    }
}

// Trailing update of the panel core's rows [q, R) on FP64 matrix cores:
// C(i, j) -= sum_c L(i, c) L(j, c) for q <= j <= i, by 16 x 16 tiles of the
// lower triangle (tile (ti, tj), tj <= ti, rows q + 16 ti.., columns
// q + 16 tj..), tile u = w + 16 k to wave w, two tiles per step.  A tile is
// two v_mfma_f64_16x16x4f64 (the panel's 8 columns as two k-steps of 4) with
// the tile's C entries as the accumulator and the row operand negated.
// Operand layout: lane l supplies A[l%16][l/16] and B[l/16][l%16]; the
// result holds D[4 r + l/16][l%16] in register r.  The panel's columns come
// from LDS (column-major: 16 lanes read consecutive words); the C entries of
// a row segment are 16 consecutive doubles of lcore.  Entries past R or
// above the diagonal are neither loaded nor stored.
typedef double sp_d4 __attribute__((ext_vector_type(4)));

CG_DEV void big_trailing_mfma(CG_AS1 double* lcore, const double* pan, int R, int q, int wave, int lane)
{
    const int nt = (R - q + 15) / 16;
    const int ntiles = nt * (nt + 1) / 2;
    const int li = lane & 15, lk = lane >> 4;
    for (int u0 = wave; u0 < ntiles; u0 += 2 * (SP_WG / 64)) {
        int ti[2], tj[2];
        bool on[2];
        sp_d4 acc[2];
        double av[2][2], bv[2][2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int u = u0 + t * (SP_WG / 64);
            on[t] = u < ntiles;   // uniform
            ti[t] = tj[t] = 0;
            if (on[t]) tri_decode(u, ti[t], tj[t]);
            const int i0 = q + 16 * ti[t], j0 = q + 16 * tj[t];
            const int jc = j0 + li;   // this lane's column
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int ir = i0 + 4 * r + lk;   // this lane's row of register r
                acc[t][r] = (on[t] && ir < R && jc <= ir) ? lcore[tri(ir) + jc] : 0.0;
            }
            const int ia = i0 + li, jb = j0 + li;
#pragma unroll
            for (int sk = 0; sk < 2; sk++) {
                const int c = 4 * sk + lk;
                av[t][sk] = (on[t] && ia < R) ? -pan[c * R + ia] : 0.0;
                bv[t][sk] = (on[t] && jb < R) ? pan[c * R + jb] : 0.0;
            }
        }
#pragma unroll
        for (int t = 0; t < 2; t++) {
            if (!on[t]) continue;
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t][0], bv[t][0], acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t][1], bv[t][1], acc[t], 0, 0, 0);
            const int i0 = q + 16 * ti[t], jc = q + 16 * tj[t] + li;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int ir = i0 + 4 * r + lk;
                if (ir < R && jc <= ir) lcore[tri(ir) + jc] = acc[t][r];
            }
        }
    }
}

// Factor the panel core C (in lcore, R <= SP_WG) in place into its Cholesky
// factor L (L L^T = C, the same storage and meaning as the register-tile
// core's) and solve L L^T x = w for the two right-hand sides in LDS (w0, w1
// in, x out).  Per panel of nb <= 8 columns starting at p0:
//   A  rows >= p0 read their panel entries; the diagonal block's rows
//      publish them and their right-hand-side entries;
//   B  wave 0, lane a = block row a: the block's Cholesky factor column by
//      column (pivots and multipliers exchanged by shuffles) and its forward
//      substitution; L's block to lcore;
//   C  rows below the block: their L entries (triangular solve against the
//      block) to lcore and LDS, and the block's contribution to their
//      right-hand-side entries;
//   D  the trailing triangle, C(i, j) -= sum_c L(i, c) L(j, c): a wave per
//      pair of rows, lanes over the columns.
// Returns 0 when a pivot is not positive.  Every sum runs in a fixed order.
__device__ __attribute__((noinline)) int core_big_factor(CG_AS1 double* lcore, int R, double* lds)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const BigLds L = big_lds(lds, R);
    const bool own = tid < R;
    double r0 = own ? L.w0[tid] : 0.0, r1 = own ? L.w1[tid] : 0.0;
    int ok = 1;
    for (int p0 = 0; p0 < R; p0 += 8) {
        const int nb = min(8, R - p0);
        const bool inp = own && tid >= p0;
        CG_AS1 double* row = lcore + (inp ? tri(tid) + p0 : 0);
        double x[8];
#pragma unroll
        for (int c = 0; c < 8; c++) x[c] = (inp && c < nb && p0 + c <= tid) ? row[c] : 0.0;
        if (inp && tid < p0 + nb) {
            const int a = tid - p0;
#pragma unroll
            for (int c = 0; c < 8; c++)
                if (c <= a) L.dblk[8 * a + c] = x[c];
            L.yb[a] = r0;
            L.yb[8 + a] = r1;
        }
        __syncthreads();
        if (wave == 0) {
            const int a = lane;
            const bool act = a < nb;
            double d[8], y0 = 0.0, y1 = 0.0;
#pragma unroll
            for (int c = 0; c < 8; c++) d[c] = (act && c <= a) ? L.dblk[8 * a + c] : 0.0;
            if (act) {
                y0 = L.yb[a];
                y1 = L.yb[8 + a];
            }
#pragma unroll
            for (int c = 0; c < 8; c++) {
                if (c < nb) {   // uniform; no early exit, so that the loop unrolls (readlane is convergent)
                    const double dc = lane_value(d[c], c);
                    ok &= (dc > 0.0) ? 1 : 0;
                    const double lcc = sqrt(dc);
                    const double y0c = lane_value(y0, c) / lcc, y1c = lane_value(y1, c) / lcc;
                    const double lac = (a > c) ? d[c] / lcc : lcc;   // L(a, c) for a >= c
#pragma unroll
                    for (int c2 = c + 1; c2 < 8; c2++) {
                        const double l2 = lane_value(lac, c2);   // L(c2, c)
                        if (a >= c2) d[c2] -= lac * l2;
                    }
                    if (a > c) {
                        y0 -= lac * y0c;
                        y1 -= lac * y1c;
                    } else if (a == c) {
                        y0 = y0c;
                        y1 = y1c;
                    }
                    if (a >= c) d[c] = lac;
                }
            }
            if (act) {
#pragma unroll
                for (int c = 0; c < 8; c++)
                    if (c <= a) {
                        L.dblk[8 * a + c] = d[c];
                        lcore[tri(p0 + a) + p0 + c] = d[c];
                    }
                L.yb[a] = y0;
                L.yb[8 + a] = y1;
            }
        }
        __syncthreads();
        if (inp) {
            if (tid < p0 + nb) {
                r0 = L.yb[tid - p0];
                r1 = L.yb[8 + tid - p0];
            } else {
                double l[8];
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    l[c] = 0.0;
                    if (c >= nb) continue;
                    double s0 = x[c];
#pragma unroll
                    for (int c2 = 0; c2 < c; c2++) s0 -= l[c2] * L.dblk[8 * c + c2];
                    l[c] = s0 / L.dblk[8 * c + c];
                    row[c] = l[c];
                    r0 -= l[c] * L.yb[c];
                    r1 -= l[c] * L.yb[8 + c];
                }
#pragma unroll
                for (int c = 0; c < 8; c++) L.pan[c * R + tid] = l[c];
            }
        }
        __syncthreads();
        // rows below the panel exist only when nb == 8
        const int q = p0 + nb;
        big_trailing_mfma(lcore, L.pan, R, q, wave, lane);
        __syncthreads();
    }
    int par = 0;
    big_bwd<2>(lcore, R, L, r0, r1, par);
    if (own) {
        L.w0[tid] = r0;
        L.w1[tid] = r1;
    }
    return ok;
}

// L L^T x = w with the panel core's stored factor, one right-hand side (w0 in
// LDS in, x out): the forward substitution by blocks of 8 rows from the first
// (as big_bwd, mirrored: the block's rows publish their right-hand-side
// entries and factor rows, every thread below solves the block's triangle and
// updates its own entry), then big_bwd.
__device__ __attribute__((noinline)) void core_big_solve(const CG_AS1 double* lcore, int R, double* lds)
{
    const int tid = threadIdx.x;
    const BigLds L = big_lds(lds, R);
    const bool own = tid < R;
    double r0 = own ? L.w0[tid] : 0.0, r1 = 0.0;
    int par = 0;
    for (int b0 = 0; b0 < R; b0 += 8) {
        const int nb = min(8, R - b0);
        double* xs = L.xb + 80 * par;   // [0, 8): right-hand side, [16, 80): the block's factor
        double* lb = xs + 16;
        par ^= 1;
        const int me = tid - b0;
        if (own && me >= 0 && me < nb) {
            xs[me] = r0;
#pragma unroll
            for (int c = 0; c < 8; c++)
                if (c <= me) lb[8 * me + c] = lcore[tri(tid) + b0 + c];
        }
        // this row's entries in the block's columns (needed after the block's solve)
        double lr[8];
        const CG_AS1 double* row = lcore + (own && me >= nb ? tri(tid) + b0 : 0);
#pragma unroll
        for (int a = 0; a < 8; a++) lr[a] = (own && me >= nb && a < nb) ? row[a] : 0.0;
        __syncthreads();
        if (own && me >= 0) {
            double y[8];
            double mine = r0;
#pragma unroll
            for (int a = 0; a < 8; a++) {
                y[a] = 0.0;
                if (a >= nb) continue;
                double s0 = xs[a];
#pragma unroll
                for (int a2 = 0; a2 < a; a2++) s0 -= lb[8 * a + a2] * y[a2];
                y[a] = s0 / lb[8 * a + a];
                if (me == a) mine = y[a];
            }
            if (me < nb) {
                r0 = mine;
            } else {
#pragma unroll
                for (int a = 0; a < 8; a++)
                    if (a < nb) r0 -= lr[a] * y[a];
            }
        }
    }
    big_bwd<1>(lcore, R, L, r0, r1, par);
    if (own) L.w0[tid] = r0;
}

// Forward phase of one periphery level: the gathers of the level's factor
// entries (FACTOR; diagonal entries take the square root) and of its
// columns' forward substitutions, b_j - sum L(j,k) y_k into yx[j] (divided
// by L_jj right away when the factor is already known, !FACTOR).
template <bool FACTOR, int NR, class V>
CG_DEV void level_gathers(const CgInstD& I, const V& v, int l, const CG_AS1 double* b0,
                          const CG_AS1 double* b1, decltype(V::yx) yx, bool& ok)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, N = v.N;
    const int e0 = v.lvptr[l], ne = v.lvptr[l + 1] - e0;
    const int c0 = v.lcptr[l], nc = v.lcptr[l + 1] - c0;
    // long lists: eight lanes per item (eight items per wave at a time)
    const int L0 = v.lgptr[l], L1 = v.lgptr[l + 1];
    for (int base = L0 + wave * 8; base < L1; base += SP_WG / 8) {
        const int li = base + (lane >> 3);
        const int gl = lane & 7;
        int code = 0, p0 = 0, p1 = 0;
        bool isy = false, act = li < L1;
        if (act) {
            code = v.longs[li];
            isy = code < 0;
            const auto r = isy ? v.ymeta + 4 * (-code - 1) : v.emeta + 4 * code;
            p0 = r[1];
            p1 = r[2];
            act = FACTOR || isy;
        }
        const auto pr = isy ? v.rowpairs : v.pairs;
        const auto y0 = isy ? yx : v.lval;
        double s0 = 0.0, s1 = 0.0;
        if (act)
            for (int p = p0 + gl; p < p1; p += 8) {
                const int x = pr[2 * p], y = pr[2 * p + 1];
                const double lx = v.lval[x];
                s0 += lx * y0[y];
                if (NR > 1 && isy) s1 += lx * yx[N + y];
            }
        s0 = group8_sum(s0);
        if (NR > 1) s1 = group8_sum(s1);
        if (act && gl == 0) {
            if (!isy) {
                const auto e = v.emeta + 4 * code;
                double x = sp_aval(I, e[0], N) - s0;
                if ((e[3] & ~SP_LONG_BIT) == code) {
                    ok = ok && (x > 0.0);
                    x = sqrt(x);
                }
                v.lval[code] = x;
            } else {
                const auto y = v.ymeta + 4 * (-code - 1);
                const int d = y[3] & ~SP_LONG_BIT;
                const int j = v.erow[d];
                const double div = FACTOR ? 1.0 : v.lval[d];
                yx[j] = (b0[y[0]] - s0) / div;
                if (NR > 1) yx[N + j] = (b1[y[0]] - s1) / div;
            }
        }
    }
    // short lists: a thread per item, four items at a time
    if (FACTOR) {
        for (int base = tid; base < ne; base += 4 * SP_WG) {
            bool on[4];
            int q[4], p0[4], p1[4], dp[4];
            double s[4], t[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int it = base + k * SP_WG;
                q[k] = e0 + it;
                on[k] = it < ne;
                p0[k] = p1[k] = dp[k] = 0;
                s[k] = t[k] = 0.0;
                if (on[k]) {
                    const auto e = v.emeta + 4 * q[k];
                    const int asrc = e[0];
                    p0[k] = e[1];
                    p1[k] = e[2];
                    dp[k] = e[3];
                    on[k] = !(dp[k] & SP_LONG_BIT);
                    if (on[k]) s[k] = sp_aval(I, asrc, N);
                }
            }
            gather4<1>(v.pairs, v.lval, v.lval, v.lval, on, p0, p1, s, t);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (!on[k]) continue;
                double x = s[k];
                if (dp[k] == q[k]) {
                    ok = ok && (x > 0.0);
                    x = sqrt(x);
                }
                v.lval[q[k]] = x;
            }
        }
    }
    for (int base = tid; base < nc; base += 4 * SP_WG) {
        bool on[4];
        int p0[4], p1[4], d[4];
        double s[4], t[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int it = base + k * SP_WG;
            on[k] = it < nc;
            p0[k] = p1[k] = d[k] = 0;
            s[k] = t[k] = 0.0;
            if (on[k]) {
                const auto y = v.ymeta + 4 * (c0 + it);
                const int vtx = y[0];
                p0[k] = y[1];
                p1[k] = y[2];
                d[k] = y[3];
                on[k] = !(d[k] & SP_LONG_BIT);
                if (on[k]) {
                    s[k] = b0[vtx];
                    if (NR > 1) t[k] = b1[vtx];
                }
            }
        }
        gather4<NR>(v.rowpairs, v.lval, yx, yx + N, on, p0, p1, s, t);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!on[k]) continue;
            const int j = v.erow[d[k]];
            const double div = FACTOR ? 1.0 : v.lval[d[k]];
            yx[j] = s[k] / div;
            if (NR > 1) yx[N + j] = t[k] / div;
        }
    }
}

// Second phase of a factor level: off-diagonal entries and the forward
// substitutions divided by their column's diagonal.
template <int NR, class V>
CG_DEV void level_divisions(const V& v, int l, decltype(V::yx) yx)
{
    const int tid = threadIdx.x, N = v.N;
    const int e0 = v.lvptr[l], ne = v.lvptr[l + 1] - e0;
    const int c0 = v.lcptr[l], nc = v.lcptr[l + 1] - c0;
    for (int it = tid; it < ne; it += SP_WG) {
        const int q = e0 + it;
        const int dp = v.emeta[4 * q + 3] & ~SP_LONG_BIT;
        if (dp != q) v.lval[q] = v.lval[q] / v.lval[dp];
    }
    for (int it = tid; it < nc; it += SP_WG) {
        const int d = v.ymeta[4 * (c0 + it) + 3] & ~SP_LONG_BIT;
        const int j = v.erow[d];
        const double dd = v.lval[d];
        yx[j] = yx[j] / dd;
        if (NR > 1) yx[N + j] = yx[N + j] / dd;
    }
}

// Core right-hand sides b_i - sum L(i,k) y_k of the core rows into out0/out1
// (row i at out[i]); short rows by threads, long rows by waves.
template <int NR, class V>
CG_DEV void core_rhs(const V& v, const CG_AS1 double* b0, const CG_AS1 double* b1, decltype(V::yx) yx,
                     double* out0, double* out1)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, N = v.N, R = v.R;
    for (int base = tid; base < R; base += 4 * SP_WG) {
        bool on[4];
        int p0[4], p1[4], row[4];
        double s[4], t[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            row[k] = base + k * SP_WG;
            on[k] = row[k] < R;
            p0[k] = p1[k] = 0;
            s[k] = t[k] = 0.0;
            if (on[k]) {
                const auto y = v.ycore + 4 * row[k];
                const int vtx = y[0];
                p0[k] = y[1];
                p1[k] = y[2];
                on[k] = y[3] == 0;
                if (on[k]) {
                    s[k] = b0[vtx];
                    if (NR > 1) t[k] = b1[vtx];
                }
            }
        }
        gather4<NR>(v.rowpairs, v.lval, yx, yx + N, on, p0, p1, s, t);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (!on[k]) continue;
            out0[row[k]] = s[k];
            if (NR > 1) out1[row[k]] = t[k];
        }
    }
    for (int base = wave * 8; base < v.nylong; base += SP_WG / 8) {
        const int li = base + (lane >> 3), gl = lane & 7;
        const bool act = li < v.nylong;
        const auto y = v.ycore + 4 * (act ? v.ylong[li] : 0);
        double s0 = 0.0, s1 = 0.0;
        if (act)
            for (int p = y[1] + gl; p < y[2]; p += 8) {
                const int q = v.rowpairs[2 * p], k = v.rowpairs[2 * p + 1];
                const double lq = v.lval[q];
                s0 += lq * yx[k];
                if (NR > 1) s1 += lq * yx[N + k];
            }
        s0 = group8_sum(s0);
        if (NR > 1) s1 = group8_sum(s1);
        if (act && gl == 0) {
            const int i = v.ylong[li];
            out0[i] = b0[y[0]] - s0;
            if (NR > 1) out1[i] = b1[y[0]] - s1;
        }
    }
}

// Periphery backward substitution (x overwrites y in yx), then the result by vertex.
template <int NR, class V>
CG_DEV void periphery_bwd(const V& v, CG_AS1 double* x0, CG_AS1 double* x1, decltype(V::yx) yx)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, N = v.N;
    for (int l = v.H - 1; l >= 0; l--) {
        const int B0 = v.blptr[l], B1 = v.blptr[l + 1];
        for (int base = B0 + wave * 8; base < B1; base += SP_WG / 8) {
            const int li = base + (lane >> 3), gl = lane & 7;
            const bool act = li < B1;
            const int c = act ? v.blong[li] : 0;
            const int d = v.lcols[c], e = act ? v.lcols[c + 1] : d;
            double s0 = 0.0, s1 = 0.0;
            for (int q = d + 1 + gl; q < e; q += 8) {
                const int r = v.erow[q];
                const double L = v.lval[q];
                s0 += L * yx[r];
                if (NR > 1) s1 += L * yx[N + r];
            }
            s0 = group8_sum(s0);
            if (NR > 1) s1 = group8_sum(s1);
            if (act && gl == 0) {
                const int j = v.erow[d];
                const double dd = v.lval[d];
                yx[j] = (yx[j] - s0) / dd;
                if (NR > 1) yx[N + j] = (yx[N + j] - s1) / dd;
            }
        }
        for (int c = v.lcptr[l] + tid; c < v.lcptr[l + 1]; c += SP_WG) {
            if (v.cflags[c]) continue;
            const int d = v.lcols[c], e = v.lcols[c + 1];
            const int j = v.erow[d];
            double s0 = yx[j];
            double s1 = (NR > 1) ? yx[N + j] : 0.0;
            int q = d + 1;
            for (; q + 1 < e; q += 2) {
                const int ra = v.erow[q], rb = v.erow[q + 1];
                const double la = v.lval[q], lb = v.lval[q + 1];
                s0 -= la * yx[ra];
                s0 -= lb * yx[rb];
                if (NR > 1) {
                    s1 -= la * yx[N + ra];
                    s1 -= lb * yx[N + rb];
                }
            }
            if (q < e) {
                const int ra = v.erow[q];
                const double la = v.lval[q];
                s0 -= la * yx[ra];
                if (NR > 1) s1 -= la * yx[N + ra];
            }
            const double dd = v.lval[d];
            yx[j] = s0 / dd;
            if (NR > 1) yx[N + j] = s1 / dd;
        }
        __syncthreads();
    }
    for (int p = tid; p < N; p += SP_WG) {
        const int vtx = v.perm[p];
        x0[vtx] = yx[p];
        if (NR > 1) x1[vtx] = yx[N + p];
    }
}

// Newton step of one instance (one workgroup of SP_WG threads): factor A at
// the current values (Da, Db, w_rval) and solve A [dir, uv] = [grad, yv], with
// the plan's values (periphery factor entries, work vectors) where v points.
template <bool PF, class V>
CG_DEV void sp_newton_body(const CgInstD& I, const V& v, double* sp_lds)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, N = v.N, t = v.t, R = v.R, T = v.T;
    const auto yx = v.yx;
    unsigned acc = 0u;
    if constexpr (PF) acc = sp_prefetch(v);
    bool ok = true;
    // ---- periphery: factor entries and forward substitution, by level ----
    for (int l = 0; l < v.H; l++) {
        level_gathers<true, 2>(I, v, l, I.grad, I.yv, yx, ok);
        __syncthreads();
        level_divisions<2>(v, l, yx);
        __syncthreads();
    }
    // ---- dense core ----
    if (R > 0) {
        // column broadcast buffers of 4T + 2 entries: rows [0, R), zeros for the
        // padded rows [R, 4T), the two right-hand sides at 4T and 4T + 1
        const bool big = R > SP_SMALL_MAX;   // the panel core (uniform)
        const int RS = 4 * T;
        double* cb0 = sp_lds;
        double* cb1 = sp_lds + (RS + 2);
        double* piv = sp_lds + 2 * (RS + 2);
        double* w0 = big ? big_lds(sp_lds, R).w0 : piv + R;
        double* w1 = w0 + R;
        // the Schur complement into lcore (scratch here), short lists by
        // threads, long ones by waves; the right-hand sides into w0, w1
        for (int base = tid; base < v.ncore; base += 4 * SP_WG) {
            bool on[4];
            int c[4], p0[4], p1[4];
            double s[4], tt[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                c[k] = base + k * SP_WG;
                on[k] = c[k] < v.ncore;
                p0[k] = p1[k] = 0;
                s[k] = tt[k] = 0.0;
                if (on[k]) {
                    const auto e = v.cmeta + 4 * c[k];
                    const int asrc = e[0];
                    p0[k] = e[1];
                    p1[k] = e[2];
                    on[k] = e[3] == 0;
                    if (on[k]) s[k] = sp_aval(I, asrc, N);
                }
            }
            gather4<1>(v.cpairs, v.lval, v.lval, v.lval, on, p0, p1, s, tt);
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (on[k]) v.lcore[c[k]] = s[k];
        }
        for (int base = wave * 8; base < v.nclong; base += SP_WG / 8) {
            const int li = base + (lane >> 3), gl = lane & 7;
            const bool act = li < v.nclong;
            const int c = act ? v.clong[li] : 0;
            const auto e = v.cmeta + 4 * c;
            double sum = 0.0;
            if (act)
                for (int p = e[1] + gl; p < e[2]; p += 8) sum += v.lval[v.cpairs[2 * p]] * v.lval[v.cpairs[2 * p + 1]];
            sum = group8_sum(sum);
            if (act && gl == 0) v.lcore[c] = sp_aval(I, e[0], N) - sum;
        }
        core_rhs<2>(v, I.grad, I.yv, yx, w0, w1);
        __syncthreads();
        if (big) {
            ok = core_big_factor(v.lcore, R, sp_lds) && ok;
        } else {
            CoreTiles ct;
            core_tiles_init(ct, T);
#pragma unroll
            for (int s = 0; s < SP_TILES_PER_THREAD; s++) {
                if (!ct.own[s]) continue;
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int bb = 0; bb < 4; bb++) {
                        const int i = 4 * ct.ti[s] + a, j = 4 * ct.tj[s] + bb;
                        if (i < R && j <= i) ct.c[s][a][bb] = v.lcore[tri(i) + j];
                    }
            }
            double r0 = 0.0, r1 = 0.0;   // this thread's core row tid of the two right-hand sides
            if (tid < R) {
                r0 = w0[tid];
                r1 = w1[tid];
            }
            // column 0 of C (all rows of the tiles of tile column 0) and of the right-hand sides
#pragma unroll
            for (int s = 0; s < SP_TILES_PER_THREAD; s++)
                if (ct.own[s] && ct.tj[s] == 0)
#pragma unroll
                    for (int a = 0; a < 4; a++) cb0[4 * ct.ti[s] + a] = ct.c[s][a][0];
            if (tid == 0) {
                cb0[RS] = r0;
                cb0[RS + 1] = r1;
            }
            __syncthreads();
            for (int k = 0; k < R; k++) {
                const double* cb = (k & 1) ? cb1 : cb0;
                double* nb = (k & 1) ? cb0 : cb1;
                const double dk = cb[k];
                ok = ok && (dk > 0.0);
                const double inv = 1.0 / dk;
                if (tid == 0) piv[k] = dk;
                // tiles with a column above k (the others are final); rows and
                // columns <= k and the padded rows take zero multipliers, entries
                // above the diagonal of a diagonal tile are never read
#pragma unroll
                for (int s = 0; s < SP_TILES_PER_THREAD; s++) {
                    if (!ct.own[s] || 4 * ct.tj[s] + 3 <= k) continue;
                    double ci[4], cj[4];
#pragma unroll
                    for (int a = 0; a < 4; a++) {
                        const int i = 4 * ct.ti[s] + a;
                        ci[a] = (i > k) ? cb[i] : 0.0;
                    }
#pragma unroll
                    for (int bb = 0; bb < 4; bb++) {
                        const int j = 4 * ct.tj[s] + bb;
                        cj[bb] = (j > k) ? cb[j] * inv : 0.0;
                    }
#pragma unroll
                    for (int a = 0; a < 4; a++)
#pragma unroll
                        for (int bb = 0; bb < 4; bb++) ct.c[s][a][bb] -= ci[a] * cj[bb];
                }
                if (tid > k && tid < R) {
                    const double l = cb[tid] * inv;
                    r0 -= cb[RS] * l;
                    r1 -= cb[RS + 1] * l;
                }
                // column k + 1 (after this thread's own updates of it); rows above
                // k + 1 of its tile column are written too and never read
                const int kk = k + 1;
                if (kk < R) {
#pragma unroll
                    for (int s = 0; s < SP_TILES_PER_THREAD; s++)
                        if (ct.own[s] && ct.tj[s] == (kk >> 2))
#pragma unroll
                            for (int a = 0; a < 4; a++) nb[4 * ct.ti[s] + a] = tile_col(ct.c[s], a, kk & 3);
                    if (tid == kk) {
                        nb[RS] = r0;
                        nb[RS + 1] = r1;
                    }
                }
                __syncthreads();
            }
            // scaled factor L = C D^{-1/2} (kept for the backward's preconditioner),
            // forward result y = D^{-1/2} r, then L^T x = y
#pragma unroll
            for (int s = 0; s < SP_TILES_PER_THREAD; s++) {
                if (!ct.own[s]) continue;
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int bb = 0; bb < 4; bb++) {
                        const int i = 4 * ct.ti[s] + a, j = 4 * ct.tj[s] + bb;
                        if (i < R && j <= i) {
                            const double sq = sqrt(piv[j]);
                            const double x = (i == j) ? sq : ct.c[s][a][bb] / sq;
                            ct.c[s][a][bb] = x;
                            v.lcore[tri(i) + j] = x;
                        }
                    }
            }
            if (tid < R) {
                const double sq = sqrt(piv[tid]);
                w0[tid] = r0 / sq;
                w1[tid] = r1 / sq;
            }
            __syncthreads();
            core_bwd_tiles<2>(ct, R, T, w0, w1);
        }
        if (tid < R) {
            yx[t + tid] = w0[tid];
            yx[N + t + tid] = w1[tid];
        }
        __syncthreads();
    }
    if (!ok) I.scal[CS_PIVOT] = 1.0;
    periphery_bwd<2>(v, I.dir, I.uv, yx);
    if (acc == 0x9e3779b9u && N < 0) yx[0] = 0.0;   // never true: keeps the prefetch loads
}

// The view with its values in LDS: lval at base (the plan's nnzP periphery
// entries), then the two work vectors (2N); with META also the plan's integer
// arrays, copied to LDS from sp_lds_meta_off on (same offsets as in the plan).
CG_DEV int sp_lds_meta_off(int nnzP, int N) { return (nnzP + 2 * N + 1) & ~1; }   // doubles, 16-byte aligned

template <bool META>
CG_DEV auto sp_lds_view(const SpView& g, double* base)
{
    using PI = typename std::conditional<META, const int*, const CG_AS1 int*>::type;
    SpViewT<double*, PI> v;
    v.N = g.N; v.n = g.n; v.t = g.t; v.R = g.R; v.H = g.H; v.ncore = g.ncore; v.T = g.T;
    v.meta_ints = g.meta_ints;
    v.nclong = g.nclong; v.nylong = g.nylong;
    v.lcore = g.lcore;
    const int nnzP = g.lcols[g.t];
    v.lval = base;
    v.yx = base + nnzP;
    if constexpr (META) {
        const int* m = reinterpret_cast<const int*>(base + sp_lds_meta_off(nnzP, g.N));
        auto at = [&](const CG_AS1 int* p) { return m + (p - g.meta); };
        v.meta = m; v.perm = at(g.perm); v.lvptr = at(g.lvptr); v.lcptr = at(g.lcptr); v.lcols = at(g.lcols);
        v.emeta = at(g.emeta); v.erow = at(g.erow); v.pairs = at(g.pairs); v.ymeta = at(g.ymeta);
        v.ycore = at(g.ycore); v.rowpairs = at(g.rowpairs); v.cmeta = at(g.cmeta); v.cpairs = at(g.cpairs);
        v.lgptr = at(g.lgptr); v.longs = at(g.longs); v.clong = at(g.clong); v.ylong = at(g.ylong);
        v.blptr = at(g.blptr); v.blong = at(g.blong); v.cflags = at(g.cflags);
    } else {
        v.meta = g.meta; v.perm = g.perm; v.lvptr = g.lvptr; v.lcptr = g.lcptr; v.lcols = g.lcols;
        v.emeta = g.emeta; v.erow = g.erow; v.pairs = g.pairs; v.ymeta = g.ymeta; v.ycore = g.ycore;
        v.rowpairs = g.rowpairs; v.cmeta = g.cmeta; v.cpairs = g.cpairs; v.lgptr = g.lgptr; v.longs = g.longs;
        v.clong = g.clong; v.ylong = g.ylong; v.blptr = g.blptr; v.blong = g.blong; v.cflags = g.cflags;
    }
    return v;
}

// Newton step of one instance with its plan in LDS.  LV 1: the values
// (periphery factor entries, work vectors) after the core's buffers; LV 2:
// the integer arrays copied there as well (in global memory they are
// streamed through the cache first).  The periphery factor is written to the
// plan at the end, for the backward's preconditioner.  The arithmetic is the
// same as with the plan in global memory.
template <int LV>
CG_DEV void sp_newton_in_lds(const CgInstD& I, const SpView& g, double* sp_lds)
{
    {
        double* base = sp_lds + SP_LDS_CORE(g.R);
        const int nnzP = g.lcols[g.t];
        if constexpr (LV == 2) {
            // the plan's integer arrays, 8 bytes per load (o_meta_end is a multiple of 8)
            long long* dst = reinterpret_cast<long long*>(base + sp_lds_meta_off(nnzP, g.N));
            const CG_AS1 long long* src = reinterpret_cast<const CG_AS1 long long*>(g.meta);
            const int64_t n8 = g.meta_ints / 2;
            for (int64_t x = threadIdx.x; x < n8; x += SP_WG) dst[x] = src[x];
            __syncthreads();
        }
        const auto v = sp_lds_view<LV == 2>(g, base);
        sp_newton_body<LV == 1>(I, v, sp_lds);
        __syncthreads();
        for (int q = threadIdx.x; q < nnzP; q += SP_WG) g.lval[q] = v.lval[q];
    }
}

// LV 0: the plan in global memory; LV 1: its values in LDS, and its integer
// arrays too when they fit in the launch's lds_bytes (decided per workgroup,
// uniformly).
template <int LV>
CG_DEV void sp_newton_block(const CgInst* __restrict__ insts, int b, double* sp_lds, int lds_bytes)
{
    const CgInstD& I = cg_inst(insts, b);
    const SpView g = sp_view(I.sph, I.S);
    if constexpr (LV == 0) {
        sp_newton_body<true>(I, g, sp_lds);
    } else {
        const int64_t all = 8 * (int64_t)SP_LDS_CORE(g.R) + 8 * (int64_t)sp_lds_meta_off(g.lcols[g.t], g.N) +
                            4 * g.meta_ints;
        if (all <= lds_bytes) sp_newton_in_lds<2>(I, g, sp_lds);
        else sp_newton_in_lds<1>(I, g, sp_lds);
    }
}

}  // namespace

template <int LV>
__global__ __launch_bounds__(SP_WG) void k_sp_newton(const CgInst* __restrict__ insts, const int* __restrict__ ids,
                                                     int lds_bytes)
{
    extern __shared__ double sp_lds[];
    sp_newton_block<LV>(insts, ids[blockIdx.x], sp_lds, lds_bytes);
}
template __global__ void k_sp_newton<0>(const CgInst*, const int*, int);
template __global__ void k_sp_newton<1>(const CgInst*, const int*, int);

// The Newton step of a group whose instances split between the sparse and
// the dense factorization, in one launch so that the two overlap: blocks
// [0, nsp) factor and solve the sparse instances ids[block] (as k_sp_newton);
// every later block runs four entries of the dense instances' wave map
// (CG_WAVES rows each, one 256-thread group per entry): entries [0, nsub)
// build Schur rows (as k_schur_build), entries [nsub, 2 nsub) form the Schur
// right-hand sides (as k_schur_rhs).  The parts read the iteration's A values
// and write disjoint outputs; each row is computed exactly as in its own
// kernel.  Dynamic LDS: the larger of the sparse blocks' need and SP_WG / 64
// row buffers of the largest P <= CG_SCHUR_LDS among the dense instances.
// The sparse blocks of k_newton_mixed: out of line, so that the kernel's
// Schur branch does not change the factorization's register allocation.
template <int LV>
__device__ __attribute__((noinline)) void sp_newton_call(const CgInst* __restrict__ insts, int b, double* lds,
                                                         int lds_bytes)
{
    sp_newton_block<LV>(insts, b, lds, lds_bytes);
}

// The Schur rows of one wave-map entry of k_newton_mixed.
CG_DEV void schur_rows_entry(const CgInst* __restrict__ insts, int nsp, const int* __restrict__ sub_map,
                             int nsub, const double* __restrict__ ctl, double* lds)
{
    const int e = ((int)blockIdx.x - nsp) * (SP_WG / CG_WG) + (int)(threadIdx.x / CG_WG);
    if (e >= 2 * nsub) return;
    const int2 me = reinterpret_cast<const int2*>(sub_map)[e < nsub ? e : e - nsub];
    const int b = me.x;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return;
    const CgInstD& I = cg_inst(insts, b);
    const int s = me.y * CG_WAVES + (int)((threadIdx.x >> 6) % CG_WAVES);
    if (s >= I.P) return;
    if (e < nsub) cg_schur_build_row(I, s, lds + (size_t)(threadIdx.x >> 6) * I.P);
    else cg_schur_rhs_row(I, s);
}

template <int LV>
__global__ __launch_bounds__(SP_WG) void k_newton_mixed(const CgInst* __restrict__ insts, const int* __restrict__ ids,
                                                        int nsp, const int* __restrict__ sub_map, int nsub,
                                                        const double* __restrict__ ctl, int lds_bytes)
{
    extern __shared__ double sp_lds[];
    if ((int)blockIdx.x < nsp)
        sp_newton_call<LV>(insts, ids[blockIdx.x], sp_lds, lds_bytes);
    else
        schur_rows_entry(insts, nsp, sub_map, nsub, ctl, sp_lds);
}
template __global__ void k_newton_mixed<0>(const CgInst*, const int*, int, const int*, int, const double*, int);
template __global__ void k_newton_mixed<1>(const CgInst*, const int*, int, const int*, int, const double*, int);

// Backward preconditioner of the adjoint instances ids[..] whose forward used
// the sparse factor: sv = A^{-1} yv with the forward's stored factor.
__global__ __launch_bounds__(SP_WG) void k_sp_precond(const CgInst* __restrict__ insts, const int* __restrict__ ids)
{
    extern __shared__ double sp_lds[];
    const int b = ids[blockIdx.x];
    const CgInstD& I = cg_inst(insts, b);
    if (I.scal[PCG_STATUS] != PCG_RUNNING) return;   // uniform over the workgroup
    const CgInstD& F = *reinterpret_cast<const CgInstD*>(I.fctx);
    const SpView v = sp_view(F.sph, F.S);
    const int tid = threadIdx.x, t = v.t, R = v.R, T = v.T;
    CG_AS1 double* yx = I.spx;
    const unsigned acc = sp_prefetch(v);
    bool ok = true;
    for (int l = 0; l < v.H; l++) {
        level_gathers<false, 1>(F, v, l, I.yv, nullptr, yx, ok);
        __syncthreads();
    }
    if (R > 0) {
        const bool big = R > SP_SMALL_MAX;   // the panel core (uniform)
        double* w0 = big ? big_lds(sp_lds, R).w0 : sp_lds;
        core_rhs<1>(v, I.yv, nullptr, yx, w0, nullptr);
        if (big) {
            __syncthreads();
            core_big_solve(v.lcore, R, sp_lds);
        } else {
            CoreTiles ct;
            core_tiles_init(ct, T);
#pragma unroll
            for (int s = 0; s < SP_TILES_PER_THREAD; s++) {
                if (!ct.own[s]) continue;
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int bb = 0; bb < 4; bb++) {
                        const int i = 4 * ct.ti[s] + a, j = 4 * ct.tj[s] + bb;
                        if (i < R && j <= i) ct.c[s][a][bb] = v.lcore[tri(i) + j];
                    }
            }
            __syncthreads();
            core_fwd_tiles<1>(ct, R, T, w0, nullptr);
            core_bwd_tiles<1>(ct, R, T, w0, nullptr);
        }
        if (tid < R) yx[t + tid] = w0[tid];
        __syncthreads();
    }
    periphery_bwd<1>(v, I.sv, nullptr, yx);
    if (acc == 0x9e3779b9u && v.N < 0) yx[0] = 0.0;   // never true: keeps the prefetch loads
}

// ---------------------------------------------------------------------------
// Pattern and plan transfers through pinned (coherent) host memory, one
// launch for all instances of a lockstep group instead of one or two copies
// per instance.
// ---------------------------------------------------------------------------

// W's row pattern of every instance flagged in ctl (rowptr [n + 1], then colj
// [K] with K = ctl[CTL_KSEL]) into host memory at int offset ctl[CTL_AUX];
// one workgroup per instance (grid = B).
__global__ __launch_bounds__(CG_WG) void k_pat_export(const CgInst* __restrict__ insts,
                                                      const double* __restrict__ ctl, int* __restrict__ dst)
{
    const int b = blockIdx.x;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return;
    const CgInstD& I = cg_inst(insts, b);
    int* d = dst + (long long)ctl[b * CG_NCTL + CTL_AUX];
    const int n = I.n, K = (int)ctl[b * CG_NCTL + CTL_KSEL];
    for (int x = threadIdx.x; x <= n; x += CG_WG) d[x] = I.w_rowptr[x];
    for (int x = threadIdx.x; x < K; x += CG_WG) d[n + 1 + x] = I.w_colj[x];
}

// Sparse plans of the instances flagged in ctl, from host memory at byte
// offset ctl[CTL_AUX]: the header (SP_HDR_PAD bytes reserved) into sph, then
// the plan's integer part [0, o_meta_end) into the instance's Schur buffer,
// whose value part [o_meta_end, bytes) is cleared.  A header with mode 0
// (the instance keeps the dense path) is copied alone.  SP_IMPORT_WGS
// workgroups per instance (grid = B * SP_IMPORT_WGS), 8-byte words.
__global__ __launch_bounds__(CG_WG) void k_plan_import(const CgInst* __restrict__ insts,
                                                       const double* __restrict__ ctl,
                                                       const char* __restrict__ src)
{
    const int b = blockIdx.x / SP_IMPORT_WGS, g = blockIdx.x % SP_IMPORT_WGS;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return;
    const CgInstD& I = cg_inst(insts, b);
    const char* s = src + (long long)ctl[b * CG_NCTL + CTL_AUX];
    const int* hs = reinterpret_cast<const int*>(s);
    if (g == 0)
        for (int x = threadIdx.x; x < (int)(sizeof(SpHdr) / 4); x += CG_WG) reinterpret_cast<CG_AS1 int*>(I.sph)[x] = hs[x];
    const SpHdr* hh = reinterpret_cast<const SpHdr*>(s);
    if (hh->mode == 0) return;
    const long long meta8 = hh->o_meta_end / 8, all8 = hh->bytes / 8;
    const long long* s8 = reinterpret_cast<const long long*>(s + SP_HDR_PAD);
    CG_AS1 long long* d8 = reinterpret_cast<CG_AS1 long long*>(I.S);
    const long long step = (long long)SP_IMPORT_WGS * CG_WG;
    for (long long x = (long long)g * CG_WG + threadIdx.x; x < all8; x += step) d8[x] = (x < meta8) ? s8[x] : 0LL;
}
