// Newton solve A [dir, u] = [g, y] on the Schur complement.
//
// A = [[Da, W], [W^T, Db]] has two diagonal blocks, so the larger one is
// eliminated and the Schur complement onto the smaller side is factorized:
//   beta side  (m-1 <= n): S = Db - W^T Da^{-1} W,  size P = m-1
//   alpha side (n < m-1) : S = Da - W Db^{-1} W^T,  size P = n
// S is symmetric, strictly diagonally dominant (margin >= shift*reg) with a
// non-positive off-diagonal, so an unpivoted Cholesky is stable.
//
// Dense S is row-major P x P (lower triangle used); the panels below the
// diagonal blocks are stored in S, the inverses of the diagonal-block factors
// in Winv.  The factorization is a right-looking blocked Cholesky with
// CG_NB = 64.  Every diagonal block is factored once, together with the
// inverse W of its factor, so the panel solve and both substitutions become
// small dense products:
//   k_chol_diag0:      factor/invert diagonal block 0, forward-substitute it;
//   k_chol_panel(0):   panel rows below block 0: X = A21 W_0^T;
//   k_chol_update(kb): trailing update S22 -= L21 L21^T on lower tiles, the
//                      forward-substitution update of the right-hand sides,
//                      the factor/inverse of diagonal block kb+1 (final here)
//                      and, once that is handed over, the panel of block kb+1;
//   k_trsv_flow:       backward substitution L^T x = y, one workgroup per block,
//                      chained through per-block flags within one launch.
// Grids are compact over a ragged batch: pref[b]..pref[b+1] are instance b's
// workgroups for the launch (host computed, inactive instances get none).
//
// Reference: the system is the one handed to cuDSS in src/linsolve.cu:159-268
// and src/sinkhorn_splr.cu:538-649 (full N x N SPD lower CSR there).
#include "cgot_common.h"
#include "cgot_kernels.h"
#include "cgot_schur_rows.h"

CG_DEV int cg_find_inst(const int* pref, int B, int wg)
{
    int lo = 0, hi = B;   // pref[lo] <= wg < pref[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pref[mid] <= wg) lo = mid; else hi = mid;
    }
    return lo;
}

CG_DEV void cg_tri_index(int t, int& I, int& J)
{
    // lower-triangular enumeration: t = I*(I+1)/2 + J, 0 <= J <= I
    int i = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= t) i++;
    while (i * (i + 1) / 2 > t) i--;
    I = i;
    J = t - i * (i + 1) / 2;
}

typedef double cg_d4 __attribute__((ext_vector_type(4)));

// Load a 64 x 64 tile src[r*ld + c] into LDS with all of a thread's loads in
// flight before the first LDS write; entries with r >= nrows or c >= ncols (or
// c > r when `lower`) read as 0.
CG_DEV void cg_tile_to_lds(double (*T)[CG_NB + 1], const CG_AS1 double* src, size_t ld, int nrows, int ncols,
                           bool lower)
{
    constexpr int NQ = CG_NB * CG_NB / CG_WG;
    double v[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int e = threadIdx.x + CG_WG * q, r = e / CG_NB, c = e % CG_NB;
        v[q] = (r < nrows && c < ncols && (!lower || c <= r)) ? src[(size_t)r * ld + c] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int e = threadIdx.x + CG_WG * q;
        T[e / CG_NB][e % CG_NB] = v[q];
    }
}

// Two full-width tiles at once (rows >= nrows_a / nrows_b read as 0).
CG_DEV void cg_tiles2_to_lds(double (*Ta)[CG_NB + 1], const CG_AS1 double* a, size_t lda, int nrows_a,
                             double (*Tb)[CG_NB + 1], const CG_AS1 double* b, size_t ldb, int nrows_b)
{
    constexpr int NQ = CG_NB * CG_NB / CG_WG;
    double va[NQ], vb[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int e = threadIdx.x + CG_WG * q, r = e / CG_NB, c = e % CG_NB;
        va[q] = (r < nrows_a) ? a[(size_t)r * lda + c] : 0.0;
        vb[q] = (r < nrows_b) ? b[(size_t)r * ldb + c] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int e = threadIdx.x + CG_WG * q;
        Ta[e / CG_NB][e % CG_NB] = va[q];
        Tb[e / CG_NB][e % CG_NB] = vb[q];
    }
}

// 64x64 tile product C = A B^T from LDS (A, B: 64 x 64, row stride CG_NB+1) on
// FP64 matrix cores: wave w owns rows 32*(w/2).. and columns 32*(w%2).., as
// 2 x 2 blocks of v_mfma_f64_16x16x4f64.  Operand layout of that instruction:
// lane l supplies A[l%16][l/16] and B[l/16][l%16] of the 16x4 and 4x16 blocks;
// the result holds D[4*r + l/16][l%16] in register r (cg_mfma_row below).
CG_DEV void cg_mfma_abt64(double (*A)[CG_NB + 1], double (*Bm)[CG_NB + 1], cg_d4 (&acc)[2][2])
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rw = 32 * (wv >> 1), cw = 32 * (wv & 1);
    const int i = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++) acc[a][c] = cg_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int kk = 0; kk < CG_NB / 4; kk++) {
        const int k = 4 * kk + kq;
        const double a0 = A[rw + i][k], a1 = A[rw + 16 + i][k];
        const double b0 = Bm[cw + i][k], b1 = Bm[cw + 16 + i][k];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
}

// Row/column of accumulator element (a, c, r) of this lane within the 64x64 tile.
CG_DEV int cg_mfma_row(int a, int r) { return 32 * ((threadIdx.x >> 6) >> 1) + 16 * a + 4 * r + ((threadIdx.x & 63) >> 4); }
CG_DEV int cg_mfma_col(int c) { return 32 * ((threadIdx.x >> 6) & 1) + 16 * c + (threadIdx.x & 15); }

// Zero the lower block tiles of S; the diagonal tiles also set their block of
// Winv to CG_SENT (the factor's diagonal workgroup hands W over through it).
__global__ __launch_bounds__(CG_WG) void k_chol_fill(const CgInst* __restrict__ insts, int B,
                                                     const int* __restrict__ pref)
{
    const int b = cg_find_inst(pref, B, blockIdx.x);
    const CgInstD& I = cg_inst(insts, b);
    int TI, TJ;
    cg_tri_index(blockIdx.x - pref[b], TI, TJ);
    const int P = I.P;
    for (int e = threadIdx.x; e < CG_NB * CG_NB; e += CG_WG) {
        const int r = TI * CG_NB + e / CG_NB, c = TJ * CG_NB + e % CG_NB;
        if (r < P && c < P) I.S[(size_t)r * P + c] = 0.0;
    }
    if (TI == TJ) {
        auto Wd = reinterpret_cast<CG_AS1 unsigned long long*>(I.Winv + (size_t)TI * CG_NB * CG_NB);
        for (int e = threadIdx.x; e < CG_NB * CG_NB; e += CG_WG) Wd[e] = CG_SENT;
    }
}

// Wave map: blockIdx.x -> (instance, 4 consecutive indices), one wave each.
CG_DEV bool cg_decode_wave(const CgInst* insts, const int* wave_inst, const double* ctl, int& b, int& q)
{
    const int2 e = cg_map_entry(wave_inst);
    b = e.x;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return false;
    q = e.y * CG_WAVES + (threadIdx.x >> 6);
    return true;
}

// Build row s2 of the lower Schur complement, one wave per row
// (cg_schur_build_row).
__global__ __launch_bounds__(CG_WG) void k_schur_build(const CgInst* __restrict__ insts,
                                                       const int* __restrict__ wave_inst,
                                                       const double* __restrict__ ctl)
{
    extern __shared__ double rowbuf[];   // CG_WAVES rows of the largest P <= CG_SCHUR_LDS in the launch
    int b, s2;
    if (!cg_decode_wave(insts, wave_inst, ctl, b, s2)) return;
    const CgInstD& I = cg_inst(insts, b);
    if (s2 >= I.P) return;
    cg_schur_build_row(I, s2, rowbuf + (size_t)(threadIdx.x >> 6) * I.P);
}

// Schur right-hand sides for [g, y], wave per s (cg_schur_rhs_row).
__global__ __launch_bounds__(CG_WG) void k_schur_rhs(const CgInst* __restrict__ insts,
                                                     const int* __restrict__ wave_inst,
                                                     const double* __restrict__ ctl)
{
    int b, s;
    if (!cg_decode_wave(insts, wave_inst, ctl, b, s)) return;
    const CgInstD& I = cg_inst(insts, b);
    if (s >= I.P) return;
    cg_schur_rhs_row(I, s);
}

// Back-substitute the large side and assemble [dir, u] in the (alpha, beta_t)
// layout (wave per entry of the N-vector).
__global__ __launch_bounds__(CG_WG) void k_schur_back(const CgInst* __restrict__ insts,
                                                      const int* __restrict__ wave_inst,
                                                      const double* __restrict__ ctl)
{
    int b, q;
    if (!cg_decode_wave(insts, wave_inst, ctl, b, q)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int P = I.P, n = I.n;
    if (q >= I.N) return;
    const int lane = threadIdx.x & 63;
    const auto xs0 = I.ltmp;
    const auto xs1 = I.ltmp + P;
    const bool large = I.beta_side ? (q < n) : (q >= n);
    if (!large) {
        if (lane == 0) {
            const int s = I.beta_side ? q - n : q;
            I.dir[q] = xs0[s];
            I.uv[q] = xs1[s];
        }
        return;
    }
    double acc0 = 0.0, acc1 = 0.0;
    double d;
    if (I.beta_side) {
        for (int p = I.w_rowptr[q] + lane; p < I.w_rowptr[q + 1]; p += 64) {
            const int j = I.w_colj[p];
            acc0 += I.w_rval[p] * xs0[j];
            acc1 += I.w_rval[p] * xs1[j];
        }
        d = I.Da[q];
    } else {
        const int j = q - n;
        for (int p = I.w_colptr[j] + lane; p < I.w_colptr[j + 1]; p += 64) {
            const int i = I.w_rowi[p];
            acc0 += I.w_cval[p] * xs0[i];
            acc1 += I.w_cval[p] * xs1[i];
        }
        d = I.Db[j];
    }
    acc0 = cg_wave_sum(acc0);
    acc1 = cg_wave_sum(acc1);
    if (lane == 0) {
        I.dir[q] = (I.grad[q] - acc0) / d;
        I.uv[q] = (I.yv[q] - acc1) / d;
    }
}

// ---------------------------------------------------------------------------
// Blocked Cholesky
// ---------------------------------------------------------------------------
// sqrt(x) and 1/sqrt(x) for x > 0: the hardware reciprocal-square-root
// estimate refined by two steps of the coupled Goldschmidt iteration
// (g -> sqrt(x), h -> 1/(2 sqrt(x)); each step squares the relative error).
CG_DEV void cg_sqrt_rsqrt(double x, double& sq, double& rs)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    sq = g;
    rs = 2.0 * h;
}

// Inverse of the 16x16 lower diagonal block p of L into W, lanes 0..15 of the
// calling wave (one column each, forward substitution with broadcast LDS reads).
CG_DEV void cg_diag16_inverse(double (*L)[CG_NB + 1], double (*W)[CG_NB + 1], const double* dinv, int p)
{
    const int lane = threadIdx.x & 63;
    if (lane >= 16) return;
    const int b0 = 16 * p, j = lane;
    double w[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        double v = (i == j) ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < i; k++) v = fma(-L[b0 + i][b0 + k], w[k], v);
        w[i] = (i >= j) ? v * dinv[b0 + i] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 16; i++) W[b0 + i][b0 + j] = w[i];
}

// 16x16x16 product on one wave: acc += A B with A[i][k] = pa[i*lda + k] and
// B[k][j] = pb[k*ldb + j] in LDS (v_mfma_f64_16x16x4f64, four k-steps).
CG_DEV void cg_mfma16(const double* pa, int lda, const double* pb, int ldb, cg_d4& acc)
{
    const int lane = threadIdx.x & 63, i = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int s = 0; s < 4; s++)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[i * lda + 4 * s + kq], pb[(4 * s + kq) * ldb + i], acc,
                                                   0, 0, 0);
}

// Factor the 64 x 64 lower block held in L (LDS, row stride CG_NB+1) in place
// and form the lower triangle of W = L^{-1} (the upper triangle of W is left
// undefined).  Rows/columns >= nb hold the identity (so the padded block
// factors trivially).  Blocked by 16:
//   panel p (columns 16p..16p+15, rows 16p..63): wave 0 holds one row per lane
//   in registers and runs the 16 right-looking column steps, broadcasting the
//   pivot and the column entries with readlane (the rows below the diagonal
//   block are solved in the same pass); meanwhile wave 1 inverts diagonal
//   block p-1.  The trailing lower block then takes the rank-16 update as
//   16x16 MFMA tiles spread over the four waves.
//   W: after the four 16x16 diagonal inverses, the off-diagonal blocks by
//   distance d = 1, 2, 3 on MFMA: W_ij = -W_ii sum_{k=j}^{i-1} L_ik W_kj.
CG_DEV bool cg_lds_potrf_inv(double (*L)[CG_NB + 1], double (*W)[CG_NB + 1], int nb, double* dinv)
{
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    __shared__ int bad;
    __shared__ double cb[2][16];
    if (tid == 0) bad = 0;
    if (nb < CG_NB) {
        for (int e = tid; e < CG_NB * CG_NB; e += CG_WG) {
            const int r = e / CG_NB, c = e % CG_NB;
            if (r >= nb || c >= nb) L[r][c] = (r == c) ? 1.0 : 0.0;
        }
    }
    __syncthreads();
#pragma unroll 1
    for (int p = 0; p < 4; p++) {
        const int b0 = 16 * p;
        const int R = CG_NB - b0;          // rows of this panel
        if (wv == 0) {
            const bool act = lane < R;
            double a[16];
#pragma unroll
            for (int k = 0; k < 16; k++) a[k] = act ? L[b0 + lane][b0 + k] : 0.0;
            bool badp = false;
            double myd = 0.0;
            // The pivot chain runs through readlane: the next pivot
            // a_{c+1}[c+1] - l_{c+1,c}^2 uses only lane c+1's own values.  The
            // other column entries l_kc reach the lanes through LDS (cb,
            // double-buffered by column parity), off the pivot chain.
            // Pivot chain: pivot c -> l_rc -> l_{c+1,c} (readlane) -> update of
            // column c+1 -> pivot c+1 (readlane) -> its square root.  The
            // other entries l_kc, k >= c+2, reach the lanes through LDS (cb,
            // double-buffered by column parity); they are first needed one
            // column later, so their latency overlaps the next pivot step.
            double piv = cg_readlane_d(a[0], 0);
            badp = !(piv > 0.0);
            double sq, rs;
            cg_sqrt_rsqrt(piv, sq, rs);
#pragma unroll
            for (int c = 0; c < 16; c++) {
                a[c] = (lane == c) ? sq : a[c] * rs;
                myd = (lane == c) ? rs : myd;
                if (c < 15) {
                    a[c + 1] = fma(-a[c], cg_readlane_d(a[c], c + 1), a[c + 1]);
                    const double pnext = cg_readlane_d(a[c + 1], c + 1);
                    double lk[16];
                    if (c < 14) {
                        if (lane < 16) cb[c & 1][lane] = a[c];
                        cg_wave_sync();
#pragma unroll
                        for (int k = c + 2; k < 16; k++) lk[k] = cb[c & 1][k];
                    }
                    badp = badp || !(pnext > 0.0);
                    cg_sqrt_rsqrt(pnext, sq, rs);
#pragma unroll
                    for (int k = c + 2; k < 16; k++) a[k] = fma(-a[c], lk[k], a[k]);
                }
            }
            if (act) {
#pragma unroll
                for (int k = 0; k < 16; k++)
                    if (lane >= 16 || k <= lane) L[b0 + lane][b0 + k] = a[k];
            }
            if (lane < 16) dinv[b0 + lane] = myd;
            if (badp && lane == 0) bad = 1;
        } else if (wv == 1 && p > 0) {
            cg_diag16_inverse(L, W, dinv, p - 1);
        }
        __syncthreads();
        const int T = (R - 16) / 16;       // trailing tiles per side
        if (T == 0) break;
        // lower tiles (ti >= tj) of the trailing block, wave-strided
        for (int t = wv; t < T * (T + 1) / 2; t += CG_WAVES) {
            int ti = 0;
            while ((ti + 1) * (ti + 2) / 2 <= t) ti++;
            const int tj = t - ti * (ti + 1) / 2;
            const int i0 = b0 + 16 + 16 * ti, j0 = b0 + 16 + 16 * tj;
            cg_d4 acc = {0.0, 0.0, 0.0, 0.0};
            // C -= L[i0.., b0..b0+15] L[j0.., b0..b0+15]^T: B[k][j] = L[j0 + j][b0 + k]
            const int i = lane & 15, kq = lane >> 4;
#pragma unroll
            for (int s = 0; s < 4; s++)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(L[i0 + i][b0 + 4 * s + kq], L[j0 + i][b0 + 4 * s + kq],
                                                           acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; r++) L[i0 + 4 * r + kq][j0 + i] -= acc[r];
        }
        __syncthreads();
    }
    if (wv == 0) cg_diag16_inverse(L, W, dinv, 3);
    __syncthreads();
    for (int d = 1; d < 4; d++) {
        const int jb = wv;
        if (jb < 4 - d) {
            const int ib = jb + d;
            cg_d4 t = {0.0, 0.0, 0.0, 0.0};
            for (int kb = jb; kb < ib; kb++)
                cg_mfma16(&L[16 * ib][16 * kb], CG_NB + 1, &W[16 * kb][16 * jb], CG_NB + 1, t);
            // W_ij = -W_ii T: the accumulator layout of T (row 4r + l/16, column
            // l%16 in register r) is the B-operand layout of k-step r
            const int i = lane & 15, kq = lane >> 4;
            cg_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < 4; s++)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(W[16 * ib + i][16 * ib + 4 * s + kq], t[s], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; r++) W[16 * ib + 4 * r + kq][16 * jb + i] = -acc[r];
        }
        __syncthreads();
    }
    return bad == 0;
}

// Factor + invert the (already updated) diagonal block held in L, store the
// factor and its inverse, and run the forward substitution of that block:
// y_kb = W_kb b_kb for the two right-hand sides (b already holds the updates).
// L and W are LDS blocks.
CG_DEV void cg_diag_block(const CgInstD& I, int kb, double (*L)[CG_NB + 1], double (*W)[CG_NB + 1],
                          double* dinv, double* ytmp)
{
    const int P = I.P;
    const int c0 = kb * CG_NB;
    const int nb = min(CG_NB, P - c0);
    const bool ok = cg_lds_potrf_inv(L, W, nb, dinv);
    if (!ok && threadIdx.x == 0) I.scal[CS_PIVOT] = 1.0;
    // hand W_kb to the panel workgroups of the same launch (Winv block kb
    // holds CG_SENT until then)
    auto Wd = I.Winv + (size_t)kb * CG_NB * CG_NB;
    for (int e = threadIdx.x; e < CG_NB * CG_NB; e += CG_WG) {
        const int r = e / CG_NB, c = e % CG_NB;
        cg_sent_store(Wd + e, (r < nb && c <= r) ? W[r][c] : 0.0);
    }
    if (threadIdx.x < 2 * CG_NB) {
        const int r = threadIdx.x / CG_NB, q = threadIdx.x % CG_NB;
        ytmp[threadIdx.x] = (q < nb) ? I.rhs[(size_t)r * P + c0 + q] : 0.0;
    }
    __syncthreads();
    if (threadIdx.x < 2 * CG_NB) {
        const int r = threadIdx.x / CG_NB, q = threadIdx.x % CG_NB;
        if (q < nb) {
            double v = 0.0;
            for (int k = 0; k <= q; k++) v = fma(W[q][k], ytmp[r * CG_NB + k], v);
            I.rhs[(size_t)r * P + c0 + q] = v;
        }
    }
}

// Bounded wait (thread 0) until flags[k] >= target, then a workgroup barrier
// and an acquire fence so the data published before the flag is visible.  A
// wait that exceeds the bound sets scal[CS_PIVOT] = 2 and returns.
CG_DEV void cg_wait_flag_p(const CG_AS1 int* flags, CG_AS1 double* scal, int k, int target)
{
    if (threadIdx.x == 0) {
        int spins = 0;
        while (__hip_atomic_load(&flags[k], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > (1 << 25)) {
                scal[CS_PIVOT] = 2.0;
                break;
            }
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

CG_DEV void cg_wait_flag(const CgInstD& I, int k, int target) { cg_wait_flag_p(I.flags, I.scal, k, target); }

// Publish block k: every thread's stores are made visible, then flags[k] = value.
CG_DEV void cg_publish_flag(CG_AS1 int* flags, int k, int value)
{
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(&flags[k], value, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Block 0: factor/invert the first diagonal block (one workgroup per instance).
__global__ __launch_bounds__(CG_WG) void k_chol_diag0(const CgInst* __restrict__ insts, int B,
                                                      const int* __restrict__ pref)
{
    const int b = cg_find_inst(pref, B, blockIdx.x);
    const CgInstD& I = cg_inst(insts, b);
    const int P = I.P;
    const int nb = min(CG_NB, P);
    __shared__ double L[CG_NB][CG_NB + 1];
    __shared__ double W[CG_NB][CG_NB + 1];
    __shared__ double dinv[CG_NB];
    __shared__ double ytmp[2 * CG_NB];
    cg_tile_to_lds(L, I.S, P, nb, nb, true);
    // the backward substitution hands its solution blocks over through ltmp
    auto x = reinterpret_cast<CG_AS1 unsigned long long*>(I.ltmp);
    for (int e = threadIdx.x; e < 2 * P; e += CG_WG) x[e] = CG_SENT;
    __syncthreads();
    cg_diag_block(I, 0, L, W, dinv, ytmp);
}

// Panel rows of block column kb (row tile rt + 1 below the diagonal):
// X = A21 W_kb^T, a 64x64x64 product on FP64 matrix cores.
__global__ __launch_bounds__(CG_WG) void k_chol_panel(const CgInst* __restrict__ insts, int B,
                                                      const int* __restrict__ pref, int kb)
{
    const int b = cg_find_inst(pref, B, blockIdx.x);
    const CgInstD& I = cg_inst(insts, b);
    const int rt = blockIdx.x - pref[b] + 1;
    const int P = I.P;
    const int c0 = kb * CG_NB;
    const int r0 = c0 + rt * CG_NB;
    auto S = I.S;
    const auto Wd = I.Winv + (size_t)kb * CG_NB * CG_NB;
    __shared__ double As[CG_NB][CG_NB + 1];
    __shared__ double Ws[CG_NB][CG_NB + 1];
    // block kb is full width
    cg_tiles2_to_lds(As, S + (size_t)r0 * P + c0, P, P - r0, Ws, Wd, CG_NB, CG_NB);
    __syncthreads();
    cg_d4 acc[2][2];
    cg_mfma_abt64(As, Ws, acc);
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gr = r0 + cg_mfma_row(a, r);
                if (gr < P) S[(size_t)gr * P + c0 + cg_mfma_col(c)] = acc[a][c][r];
            }
}

// Trailing update of tile (TI, TJ), kb < TJ <= TI: C -= A B^T with A, B the
// panel rows of blocks TI and TJ (C is loaded before the product).  Diagonal
// tiles also apply the forward-substitution update rhs_TI -= A y_kb.  The tile
// (kb+1, kb+1) is final after this update: it factors and inverts itself,
// hands W_{kb+1} over through Winv (see CG_SENT) and solves its rhs block.  The
// tiles (TI, kb+1) below it are the next panel: after their update they wait
// for W_{kb+1} and store X = C W_{kb+1}^T in place.
// The launch holds three segments, each with its instance prefix table
// (pref, pref + seg_stride, pref + 2 seg_stride): the diagonal tiles of all
// instances, so that every factor starts at once; the generic tiles, which
// wait on nothing; and the next-panel tiles last, which wait only on a
// diagonal tile of the first segment (dispatched before them).
__global__ __launch_bounds__(CG_WG) void k_chol_update(const CgInst* __restrict__ insts, int B,
                                                       const int* __restrict__ pref, int seg_stride, int kb)
{
    const int* prefG = pref + seg_stride;
    const int* prefP = pref + 2 * seg_stride;
    const int nD = pref[B], nG = prefG[B];
    int b, ti, tj;
    if ((int)blockIdx.x < nD) {
        b = cg_find_inst(pref, B, blockIdx.x);
        ti = tj = 0;
    } else if ((int)blockIdx.x < nD + nG) {
        const int w = blockIdx.x - nD;
        b = cg_find_inst(prefG, B, w);
        cg_tri_index(w - prefG[b], ti, tj);
        ti++;
        tj++;
    } else {
        const int w = blockIdx.x - nD - nG;
        b = cg_find_inst(prefP, B, w);
        ti = 1 + w - prefP[b];
        tj = 0;
    }
    const CgInstD& I = cg_inst(insts, b);
    const int TI = kb + 1 + ti, TJ = kb + 1 + tj;
    const int P = I.P;
    const int c0 = kb * CG_NB;
    const int ri = TI * CG_NB, rj = TJ * CG_NB;
    auto S = I.S;
    __shared__ double As[CG_NB][CG_NB + 1];
    __shared__ double Bs[CG_NB][CG_NB + 1];
    __shared__ double dinv[CG_NB];
    __shared__ double ytmp[2 * CG_NB];
    // current C in the accumulator layout (rows beyond P read as 0; the
    // diagonal tile keeps only its lower triangle)
    double cold[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gr = ri + cg_mfma_row(a, r), gc = rj + cg_mfma_col(c);
                cold[a][c][r] = (gr < P && gc < P && (TI != TJ || gc <= gr)) ? S[(size_t)gr * P + gc] : 0.0;
            }
    cg_tiles2_to_lds(As, S + (size_t)ri * P + c0, P, P - ri, Bs, S + (size_t)rj * P + c0, P, P - rj);
    if (TI == TJ && threadIdx.x < 2 * CG_NB)   // y_kb (block kb is full width)
        ytmp[threadIdx.x] = I.rhs[(size_t)(threadIdx.x / CG_NB) * P + c0 + threadIdx.x % CG_NB];
    __syncthreads();
    cg_d4 acc[2][2];
    cg_mfma_abt64(As, Bs, acc);
    if (TI != TJ && TJ != kb + 1) {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int gr = ri + cg_mfma_row(a, r), gc = rj + cg_mfma_col(c);
                    if (gr < P && gc < P) S[(size_t)gr * P + gc] = cold[a][c][r] - acc[a][c][r];
                }
        return;
    }
    if (TI == TJ) {
        // forward-substitution update of rhs block TI
        if (threadIdx.x < 2 * CG_NB) {
            const int rhsi = threadIdx.x / CG_NB, r = threadIdx.x % CG_NB;
            if (ri + r < P) {
                double v = 0.0;
#pragma unroll 8
                for (int k = 0; k < CG_NB; k++) v = fma(As[r][k], ytmp[rhsi * CG_NB + k], v);
                I.rhs[(size_t)rhsi * P + ri + r] -= v;
            }
        }
        if (TI != kb + 1) {
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int c = 0; c < 2; c++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int gr = ri + cg_mfma_row(a, r), gc = rj + cg_mfma_col(c);
                        if (gr < P && gc <= gr) S[(size_t)gr * P + gc] = cold[a][c][r] - acc[a][c][r];
                    }
            return;
        }
    }
    // tile (kb+1, kb+1) or a tile of the next panel: the updated tile goes to
    // LDS (As; padding rows/columns are 0)
    const int nb = min(CG_NB, P - ri);
    __syncthreads();   // As/Bs no longer read (product, rhs update)
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int lr = cg_mfma_row(a, r), lc = cg_mfma_col(c);
                As[lr][lc] = (lr < nb) ? cold[a][c][r] - acc[a][c][r] : 0.0;
            }
    if (TI == TJ) {
        __syncthreads();
        cg_diag_block(I, kb + 1, As, Bs, dinv, ytmp);   // Bs is free now: it holds W
        return;
    }
    // next panel: X = C W^T, W = W_{kb+1} once handed over (all of a
    // thread's loads in flight, then the ones still at CG_SENT are read again
    // together)
    {
        const auto Wd = I.Winv + (size_t)(kb + 1) * CG_NB * CG_NB;
        unsigned long long u[CG_NB * CG_NB / CG_WG];
#pragma unroll
        for (int q = 0; q < CG_NB * CG_NB / CG_WG; q++) u[q] = cg_sent_peek(Wd + threadIdx.x + CG_WG * q);
        cg_sent_gather(u, [&](int q) { return Wd + threadIdx.x + CG_WG * q; }, I.scal);
#pragma unroll
        for (int q = 0; q < CG_NB * CG_NB / CG_WG; q++) {
            const int e = threadIdx.x + CG_WG * q;
            Bs[e / CG_NB][e % CG_NB] = __longlong_as_double((long long)u[q]);
        }
    }
    __syncthreads();
    cg_mfma_abt64(As, Bs, acc);
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gr = ri + cg_mfma_row(a, r);
                if (gr < P) S[(size_t)gr * P + rj + cg_mfma_col(c)] = acc[a][c][r];
            }
}

// Operands of a blocked triangular solve with the stored factor: the panels
// of L in S (row-major P x P), the inverses of its diagonal blocks in Winv,
// right-hand sides y and solutions x (NR of each, stride P), the block flags
// and the instance scalars (for the wait-bound flag).
struct CgTriArgs {
    const CG_AS1 double* S;
    const CG_AS1 double* Winv;
    int P;
    const CG_AS1 double* y;
    CG_AS1 double* x;
    CG_AS1 int* flags;
    CG_AS1 double* scal;
};

// Backward substitution L^T x = y, block kb of a dataflow launch: accumulate
// L_{j,kb}^T x_j for j = nblk-1 .. kb+1 as each x_j arrives (x holds CG_SENT
// before the launch; the tile is loaded before the poll), then
// x_kb = W_kb^T (y_kb - sum), handed on through x.
template <int NR>
CG_DEV void cg_bwd_block(const CgTriArgs& A, int kb)
{
    const int P = A.P;
    const int nblk = (P + CG_NB - 1) / CG_NB;
    const int c0 = kb * CG_NB;
    const int nb = min(CG_NB, P - c0);
    const int tid = threadIdx.x, c = tid & 63, g = tid >> 6;
    __shared__ double Ws[CG_NB][CG_NB + 1];
    __shared__ double xs[2][CG_NB];
    __shared__ double part[2][CG_WAVES][CG_NB];
    cg_tile_to_lds(Ws, A.Winv + (size_t)kb * CG_NB * CG_NB, CG_NB, CG_NB, CG_NB, false);
    // y_kb (thread < NR*64: right-hand side tid/64, entry tid%64)
    const double yv = (tid < NR * CG_NB && (tid % CG_NB) < nb) ? A.y[(size_t)(tid / CG_NB) * P + c0 + tid % CG_NB]
                                                                : 0.0;
    double v[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) v[r] = 0.0;
    for (int j = nblk - 1; j > kb; j--) {
        // block kb is full width here (kb < nblk-1); tile rows g, g+4, ...
        const int r0 = j * CG_NB;
        const int nbj = min(CG_NB, P - r0);
        double sv[CG_NB / CG_WAVES];
#pragma unroll
        for (int q = 0; q < CG_NB / CG_WAVES; q++) {
            const int rr = g + CG_WAVES * q;
            sv[q] = (rr < nbj) ? A.S[(size_t)(r0 + rr) * P + c0 + c] : 0.0;
        }
        if (tid < NR * CG_NB) {
            const int r = tid / CG_NB, q = tid % CG_NB;
            xs[r][q] = (q < nbj) ? cg_sent_load(A.x + (size_t)r * P + r0 + q, A.scal) : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < CG_NB / CG_WAVES; q++) {
            const int rr = g + CG_WAVES * q;
#pragma unroll
            for (int r = 0; r < NR; r++) v[r] = fma(sv[q], xs[r][rr], v[r]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < NR; r++) part[r][g][c] = v[r];
    __syncthreads();
    if (tid < NR * CG_NB) {
        const int r = tid / CG_NB, q = tid % CG_NB;
        const double acc = (part[r][0][q] + part[r][1][q]) + (part[r][2][q] + part[r][3][q]);
        xs[r][q] = (q < nb) ? yv - acc : 0.0;
    }
    __syncthreads();
    {
        // x[q] = sum_{cc >= q} W[cc][q] r[cc]: thread (q, quarter g) sums cc in
        // [16g, 16g + 16); partials reduced through LDS
        double xq[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) xq[r] = 0.0;
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int cc = 16 * g + u;
            const double w = (cc >= c) ? Ws[cc][c] : 0.0;
#pragma unroll
            for (int r = 0; r < NR; r++) xq[r] = fma(w, xs[r][cc], xq[r]);
        }
#pragma unroll
        for (int r = 0; r < NR; r++) part[r][g][c] = xq[r];
    }
    __syncthreads();
    if (tid < NR * CG_NB) {
        const int r = tid / CG_NB, q = tid % CG_NB;
        const double x = (part[r][0][q] + part[r][1][q]) + (part[r][2][q] + part[r][3][q]);
        if (q < nb) cg_sent_store(A.x + (size_t)r * P + c0 + q, x);
    }
}

// Backward substitution L^T x = y of the Newton solve as a dataflow over the
// block columns: one workgroup per (instance, block kb).  Workgroups are
// numbered level by level, level t = nblk-1-kb, with the instances of each
// level in the order `order` (sorted by decreasing block count, so level t
// holds the first lvl_pref[t+1] - lvl_pref[t] of them): every instance's chain
// starts at once.  A workgroup waits only on lower levels, i.e. on workgroups
// with smaller launch indices, which were dispatched before it and themselves
// wait only on earlier ones, so every wait ends; the spin is also bounded and
// flags CS_PIVOT = 2 instead of hanging.  y lives in rhs, x goes to ltmp
// (set to CG_SENT by k_chol_diag0).
__global__ __launch_bounds__(CG_WG) void k_trsv_flow(const CgInst* __restrict__ insts, const int* __restrict__ order,
                                                     const int* __restrict__ lvl_pref, int nlvl)
{
    const int t = cg_find_inst(lvl_pref, nlvl, blockIdx.x);
    const int b = order[blockIdx.x - lvl_pref[t]];
    const CgInstD& I = cg_inst(insts, b);
    const int nblk = (I.P + CG_NB - 1) / CG_NB;
    const CgTriArgs A{I.S, I.Winv, I.P, I.rhs, I.ltmp, I.flags, I.scal};
    cg_bwd_block<2>(A, nblk - 1 - t);
}

// ---------------------------------------------------------------------------
// SPLR preconditioner of the implicit backward: q = A^{-1} r with the forward
// instance's stored Schur factor (A = the sparsified Hessian of the last
// forward iteration; its W, D and factor are left in the forward workspace).
// An adjoint instance with fctx != 0 points at that forward descriptor; these
// kernels act only while its PCG is running.  Reference: the cuDSS solve with
// the adopted factor in SPLRBackwardContext::apply_preconditioner,
// src/backward_context.cu:543-581.
// ---------------------------------------------------------------------------
CG_DEV const CgInstD& cg_fwd_of(const CgInstD& I) { return *reinterpret_cast<const CgInstD*>(I.fctx); }

CG_DEV bool cg_pc_live(const CgInstD& I) { return I.fctx != 0 && I.scal[PCG_STATUS] == PCG_RUNNING; }

// the preconditioner of this adjoint instance runs through the dense Schur
// factor (a sparse forward factor is applied by k_sp_precond instead)
CG_DEV bool cg_pc_dense_live(const CgInstD& I) { return cg_pc_live(I) && I.fsp == 0; }

// Schur right-hand side of r (the PCG residual, N layout), wave per small index.
__global__ __launch_bounds__(CG_WG) void k_pc_rhs(const CgInst* __restrict__ insts,
                                                  const int* __restrict__ wave_inst,
                                                  const double* __restrict__ ctl)
{
    int b, s;
    if (!cg_decode_wave(insts, wave_inst, ctl, b, s)) return;
    const CgInstD& I = cg_inst(insts, b);
    if (!cg_pc_dense_live(I)) return;
    const CgInstD& F = cg_fwd_of(I);
    const int P = F.P, n = F.n;
    if (s >= P) return;
    const int lane = threadIdx.x & 63;
    const auto r = I.yv;
    double acc = 0.0;
    if (F.beta_side) {
        for (int p = F.w_colptr[s] + lane; p < F.w_colptr[s + 1]; p += 64) {
            const int i = F.w_rowi[p];
            acc += F.w_cval[p] * (r[i] / F.Da[i]);
        }
    } else {
        for (int p = F.w_rowptr[s] + lane; p < F.w_rowptr[s + 1]; p += 64) {
            const int j = F.w_colj[p];
            acc += F.w_rval[p] * (r[n + j] / F.Db[j]);
        }
    }
    acc = cg_wave_sum(acc);
    if (lane == 0) {
        I.prhs[s] = r[F.beta_side ? n + s : s] - acc;
        // the backward substitution hands its solution blocks over through pltmp
        reinterpret_cast<CG_AS1 unsigned long long*>(I.pltmp)[s] = CG_SENT;
    }
}

// Forward substitution L y = rhs, block kb of a dataflow launch (y replaces rhs
// block by block): accumulate L_{kb,j} y_j for j = 0 .. kb-1 as each y_j is
// published (flags[j] >= target), then y_kb = W_kb (rhs_kb - sum).
CG_DEV void cg_fwd_block(const CgTriArgs& A, CG_AS1 double* y, int kb, int target)
{
    const int P = A.P;
    const int c0 = kb * CG_NB;
    const int nb = min(CG_NB, P - c0);
    const int tid = threadIdx.x, c = tid & 63, g = tid >> 6, lane = tid & 63;
    __shared__ double Ws[CG_NB][CG_NB + 1];
    __shared__ double xs[CG_NB];
    __shared__ double red[CG_NB];
    __shared__ double part[CG_WAVES][CG_NB];
    cg_tile_to_lds(Ws, A.Winv + (size_t)kb * CG_NB * CG_NB, CG_NB, CG_NB, CG_NB, false);
    const double rv = (tid < nb) ? y[c0 + tid] : 0.0;
    // thread (column c, row group g) accumulates rows g, g+4, ... over column c
    double acc[CG_NB / CG_WAVES];
#pragma unroll
    for (int q = 0; q < CG_NB / CG_WAVES; q++) acc[q] = 0.0;
    for (int j = 0; j < kb; j++) {
        double sv[CG_NB / CG_WAVES];
#pragma unroll
        for (int q = 0; q < CG_NB / CG_WAVES; q++) {
            const int rr = g + CG_WAVES * q;
            sv[q] = (rr < nb) ? A.S[(size_t)(c0 + rr) * P + j * CG_NB + c] : 0.0;   // block j is full
        }
        cg_wait_flag_p(A.flags, A.scal, j, target);
        if (tid < CG_NB) xs[tid] = y[j * CG_NB + tid];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < CG_NB / CG_WAVES; q++) acc[q] = fma(sv[q], xs[c], acc[q]);
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < CG_NB / CG_WAVES; q++) {
        const double t = cg_wave_sum(acc[q]);
        if (lane == 0) red[g + CG_WAVES * q] = t;
    }
    __syncthreads();
    if (tid < CG_NB) xs[tid] = (tid < nb) ? rv - red[tid] : 0.0;
    __syncthreads();
    {
        // y[q] = sum_{cc <= q} W[q][cc] r[cc]: thread (q, quarter g) sums cc in [16g, 16g + 16)
        double yq = 0.0;
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int cc = 16 * g + u;
            yq = fma((cc <= c) ? Ws[c][cc] : 0.0, xs[cc], yq);
        }
        part[g][c] = yq;
    }
    __syncthreads();
    if (tid < nb) y[c0 + tid] = (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
    cg_publish_flag(A.flags, kb, target);
}

// Dataflow launches of the two triangular solves (level tables as for
// k_trsv_flow; level t is block t forward, block nblk-1-t backward).
__global__ __launch_bounds__(CG_WG) void k_pc_fwd(const CgInst* __restrict__ insts, const int* __restrict__ order,
                                                  const int* __restrict__ lvl_pref, int nlvl, int target)
{
    const int t = cg_find_inst(lvl_pref, nlvl, blockIdx.x);
    const CgInstD& I = cg_inst(insts, order[blockIdx.x - lvl_pref[t]]);
    if (!cg_pc_dense_live(I)) return;
    const CgInstD& F = cg_fwd_of(I);
    const CgTriArgs A{F.S, F.Winv, F.P, I.prhs, I.pltmp, I.flags, I.scal};
    cg_fwd_block(A, I.prhs, t, target);
}

__global__ __launch_bounds__(CG_WG) void k_pc_bwd(const CgInst* __restrict__ insts, const int* __restrict__ order,
                                                  const int* __restrict__ lvl_pref, int nlvl)
{
    const int t = cg_find_inst(lvl_pref, nlvl, blockIdx.x);
    const CgInstD& I = cg_inst(insts, order[blockIdx.x - lvl_pref[t]]);
    if (!cg_pc_dense_live(I)) return;
    const CgInstD& F = cg_fwd_of(I);
    const int nblk = (F.P + CG_NB - 1) / CG_NB;
    const CgTriArgs A{F.S, F.Winv, F.P, I.prhs, I.pltmp, I.flags, I.scal};
    cg_bwd_block<1>(A, nblk - 1 - t);
}

// q = A^{-1} r in the N layout into z (sv): the small side is the Schur
// solution, the large side is back-substituted (wave per entry).
__global__ __launch_bounds__(CG_WG) void k_pc_back(const CgInst* __restrict__ insts,
                                                   const int* __restrict__ wave_inst,
                                                   const double* __restrict__ ctl)
{
    int b, q;
    if (!cg_decode_wave(insts, wave_inst, ctl, b, q)) return;
    const CgInstD& I = cg_inst(insts, b);
    if (!cg_pc_dense_live(I)) return;
    const CgInstD& F = cg_fwd_of(I);
    const int n = F.n;
    if (q >= F.N) return;
    const int lane = threadIdx.x & 63;
    const auto x = I.pltmp;
    const bool large = F.beta_side ? (q < n) : (q >= n);
    if (!large) {
        if (lane == 0) I.sv[q] = x[F.beta_side ? q - n : q];
        return;
    }
    double acc = 0.0, d;
    if (F.beta_side) {
        for (int p = F.w_rowptr[q] + lane; p < F.w_rowptr[q + 1]; p += 64) acc += F.w_rval[p] * x[F.w_colj[p]];
        d = F.Da[q];
    } else {
        const int j = q - n;
        for (int p = F.w_colptr[j] + lane; p < F.w_colptr[j + 1]; p += 64) acc += F.w_cval[p] * x[F.w_rowi[p]];
        d = F.Db[j];
    }
    acc = cg_wave_sum(acc);
    if (lane == 0) I.sv[q] = (I.yv[q] - acc) / d;
}
