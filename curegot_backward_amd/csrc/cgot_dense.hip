// Newton solve A [dir, u] = [g, y] on the Schur complement.
//
// A = [[Da, W], [W^T, Db]] has two diagonal blocks, so the larger one is
// eliminated and the Schur complement onto the smaller side is factorized:
//   beta side  (m-1 <= n): S = Db - W^T Da^{-1} W,  size P = m-1
//   alpha side (n < m-1) : S = Da - W Db^{-1} W^T,  size P = n
// S is symmetric, strictly diagonally dominant (margin >= shift*reg) with a
// non-positive off-diagonal, so an unpivoted Cholesky is stable.
//
// Dense S is row-major P x P (lower triangle used); the factors of the
// diagonal blocks live in Ldg, the panels below them in S.  The factorization
// is a right-looking blocked Cholesky with CG_NB = 64.  Every diagonal block
// is factored once, together with the inverse W of its factor, so the panel
// solve and both substitutions become small dense products:
//   k_chol_diag0:      factor/invert diagonal block 0, forward-substitute it;
//   k_chol_panel(kb):  panel rows below block kb: X = A21 W_kb^T;
//   k_chol_update(kb): trailing update S22 -= L21 L21^T on lower tiles, the
//                      forward-substitution update of the right-hand sides, and
//                      the factor/inverse of diagonal block kb+1 (final here);
//   k_trsv_back(kb):   backward substitution L^T x = y, one block per launch.
// Grids are compact over a ragged batch: pref[b]..pref[b+1] are instance b's
// workgroups for the launch (host computed, inactive instances get none).
//
// Reference: the system is the one handed to cuDSS in src/linsolve.cu:159-268
// and src/sinkhorn_splr.cu:538-649 (full N x N SPD lower CSR there).
#include "cgot_common.h"
#include "cgot_kernels.h"

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

// 64x64 tile product C = A B^T from LDS (A, B: 64 x 64, row stride CG_NB+1) on
// FP64 matrix cores: wave w owns rows 32*(w/2).. and columns 32*(w%2).., as
// 2 x 2 blocks of v_mfma_f64_16x16x4f64.  Operand layout of that instruction:
// lane l supplies A[l%16][l/16] and B[l/16][l%16] of the 16x4 and 4x16 blocks;
// the result holds D[4*(l/16) + r][l%16] in register r.
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

// Zero the lower block tiles of S.
__global__ __launch_bounds__(CG_WG) void k_chol_fill(const CgInst* __restrict__ insts, int B,
                                                     const int* __restrict__ pref)
{
    const int b = cg_find_inst(pref, B, blockIdx.x);
    const CgInst& I = insts[b];
    int TI, TJ;
    cg_tri_index(blockIdx.x - pref[b], TI, TJ);
    const int P = I.P;
    for (int e = threadIdx.x; e < CG_NB * CG_NB; e += CG_WG) {
        const int r = TI * CG_NB + e / CG_NB, c = TJ * CG_NB + e % CG_NB;
        if (r < P && c < P) I.S[(size_t)r * P + c] = 0.0;
    }
}

// Wave map: blockIdx.x -> (instance, 4 consecutive indices), one wave each.
CG_DEV bool cg_decode_wave(const CgInst* insts, const int* wave_inst, const double* ctl, int& b, int& q)
{
    b = wave_inst[blockIdx.x];
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return false;
    q = (blockIdx.x - insts[b].row0) * CG_WAVES + (threadIdx.x >> 6);
    return true;
}

// Build row s2 of the lower Schur complement, one wave per row:
// S[s2][s2] = D_small[s2]; then for each large index l adjacent to s2
// (ascending) the lanes cover the small indices s <= s2 adjacent to l:
// S[s2][s] -= w(s2,l) w(l,s) / D_large[l].  Each entry accumulates its terms
// in ascending l; a fence between l steps orders the lanes' updates.
__global__ __launch_bounds__(CG_WG) void k_schur_build(const CgInst* __restrict__ insts,
                                                       const int* __restrict__ wave_inst,
                                                       const double* __restrict__ ctl)
{
    int b, s2;
    if (!cg_decode_wave(insts, wave_inst, ctl, b, s2)) return;
    const CgInst& I = insts[b];
    const int P = I.P;
    if (s2 >= P) return;
    const int lane = threadIdx.x & 63;
    const int* sp; const int* sidx; const double* sval;
    const int* lp; const int* lidx; const double* lval;
    const double* Ds; const double* Dl;
    if (I.beta_side) {
        sp = I.w_colptr; sidx = I.w_rowi; sval = I.w_cval;
        lp = I.w_rowptr; lidx = I.w_colj; lval = I.w_rval;
        Ds = I.Db; Dl = I.Da;
    } else {
        sp = I.w_rowptr; sidx = I.w_colj; sval = I.w_rval;
        lp = I.w_colptr; lidx = I.w_rowi; lval = I.w_cval;
        Ds = I.Da; Dl = I.Db;
    }
    double* Srow = I.S + (size_t)s2 * P;
    if (lane == 0) Srow[s2] = Ds[s2];
    for (int p = sp[s2]; p < sp[s2 + 1]; p++) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        const int l = sidx[p];
        const double ws = sval[p];
        const double dl = Dl[l];
        for (int q = lp[l] + lane; q < lp[l + 1]; q += 64) {
            const int s = lidx[q];
            if (s <= s2) Srow[s] -= ws * lval[q] / dl;
        }
    }
}

// Schur right-hand sides for [g, y]: r_s = g_s - sum_l w(s,l) g_l / D_l (wave per s).
__global__ __launch_bounds__(CG_WG) void k_schur_rhs(const CgInst* __restrict__ insts,
                                                     const int* __restrict__ wave_inst,
                                                     const double* __restrict__ ctl)
{
    int b, s;
    if (!cg_decode_wave(insts, wave_inst, ctl, b, s)) return;
    const CgInst& I = insts[b];
    const int P = I.P, n = I.n;
    if (s >= P) return;
    const int lane = threadIdx.x & 63;
    double acc0 = 0.0, acc1 = 0.0;
    if (I.beta_side) {
        for (int p = I.w_colptr[s] + lane; p < I.w_colptr[s + 1]; p += 64) {
            const int i = I.w_rowi[p];
            acc0 += I.w_cval[p] * (I.grad[i] / I.Da[i]);
            acc1 += I.w_cval[p] * (I.yv[i] / I.Da[i]);
        }
    } else {
        for (int p = I.w_rowptr[s] + lane; p < I.w_rowptr[s + 1]; p += 64) {
            const int j = I.w_colj[p];
            acc0 += I.w_rval[p] * (I.grad[n + j] / I.Db[j]);
            acc1 += I.w_rval[p] * (I.yv[n + j] / I.Db[j]);
        }
    }
    acc0 = cg_wave_sum(acc0);
    acc1 = cg_wave_sum(acc1);
    if (lane == 0) {
        const int k = I.beta_side ? n + s : s;
        I.rhs[s] = I.grad[k] - acc0;
        I.rhs[(size_t)P + s] = I.yv[k] - acc1;
    }
}

// Back-substitute the large side and assemble [dir, u] in the (alpha, beta_t)
// layout (wave per entry of the N-vector).
__global__ __launch_bounds__(CG_WG) void k_schur_back(const CgInst* __restrict__ insts,
                                                      const int* __restrict__ wave_inst,
                                                      const double* __restrict__ ctl)
{
    int b, q;
    if (!cg_decode_wave(insts, wave_inst, ctl, b, q)) return;
    const CgInst& I = insts[b];
    const int P = I.P, n = I.n;
    if (q >= I.N) return;
    const int lane = threadIdx.x & 63;
    const double* xs0 = I.ltmp;
    const double* xs1 = I.ltmp + P;
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
// Factor the 64 x 64 lower block held in L (LDS, row stride CG_NB+1) in place
// and form W = L^{-1}.  Rows/columns >= nb hold the identity (so the padded
// block factors trivially).  Blocked by 16:
//   for each 16-column panel p: lanes 0..15 of wave 0 factor the 16x16 diagonal
//   block left-looking, one row per lane in registers (the pivot row is shared
//   through LDS within the wave, no barrier); the rows below solve against it;
//   the trailing block takes the rank-16 update;
//   W: the four 16x16 diagonal inverses, one column per lane of wave 0 in
//   registers; then the off-diagonal blocks by distance d = 1, 2, 3:
//   W_ij = -W_ii sum_{k=j}^{i-1} L_ik W_kj.
CG_DEV bool cg_lds_potrf_inv(double (*L)[CG_NB + 1], double (*W)[CG_NB + 1], int nb, double* dinv)
{
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    __shared__ int bad;
    if (tid == 0) bad = 0;
    for (int e = tid; e < CG_NB * CG_NB; e += CG_WG) {
        const int r = e / CG_NB, c = e % CG_NB;
        if (r >= nb || c >= nb) L[r][c] = (r == c) ? 1.0 : 0.0;
    }
    __syncthreads();
#pragma unroll 1
    for (int p = 0; p < 4; p++) {
        const int b0 = 16 * p;
        if (wv == 0 && lane < 16) {
            const int r = lane;
            double a[16];
#pragma unroll
            for (int k = 0; k < 16; k++) a[k] = L[b0 + r][b0 + k];
#pragma unroll
            for (int c = 0; c < 16; c++) {
                if (r == c) {
                    double d = a[c];
#pragma unroll
                    for (int k = 0; k < c; k++) d -= a[k] * a[k];
                    if (!(d > 0.0)) bad = 1;
                    const double sd = sqrt(d);
                    const double id = 1.0 / sd;
                    a[c] = sd;
                    L[b0 + c][b0 + c] = sd;
                    dinv[b0 + c] = id;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (r > c) {
                    double v = a[c];
#pragma unroll
                    for (int k = 0; k < c; k++) v -= a[k] * L[b0 + c][b0 + k];
                    v *= dinv[b0 + c];
                    a[c] = v;
                    L[b0 + r][b0 + c] = v;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        __syncthreads();
        const int rows = CG_NB - b0 - 16;
        if (rows == 0) break;
        // rows below: X L_pp^T = A  (one thread per row r, 16 sequential columns)
        if (tid < rows) {
            const int r = b0 + 16 + tid;
            double x[16];
#pragma unroll
            for (int c = 0; c < 16; c++) {
                double v = L[r][b0 + c];
#pragma unroll
                for (int q = 0; q < c; q++) v -= x[q] * L[b0 + c][b0 + q];
                x[c] = v * dinv[b0 + c];
            }
#pragma unroll
            for (int c = 0; c < 16; c++) L[r][b0 + c] = x[c];
        }
        __syncthreads();
        // trailing update of rows/cols >= b0+16 (lower part): wave wv takes rows
        // r = b0+16+wv+4t, lane takes column q = b0+16+lane
        {
            const int q = b0 + 16 + lane;
            double lq[16];
            if (q < CG_NB) {
#pragma unroll
                for (int k = 0; k < 16; k++) lq[k] = L[q][b0 + k];
            }
            for (int r = b0 + 16 + wv; r < CG_NB; r += CG_WAVES) {
                if (q < CG_NB && q <= r) {
                    double v = 0.0;
#pragma unroll
                    for (int k = 0; k < 16; k++) v += L[r][b0 + k] * lq[k];
                    L[r][q] -= v;
                }
            }
        }
        __syncthreads();
    }
    // ---- inverse: diagonal 16x16 blocks, lane = (block, column) of wave 0 ----
    for (int e = tid; e < CG_NB * CG_NB; e += CG_WG) W[e / CG_NB][e % CG_NB] = 0.0;
    __syncthreads();
    if (wv == 0) {
        const int b0 = 16 * (lane >> 4), j = lane & 15;
        double w[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            double v = (i == j) ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < i; k++) v -= L[b0 + i][b0 + k] * w[k];
            w[i] = (i >= j) ? v * dinv[b0 + i] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < 16; i++) W[b0 + i][b0 + j] = w[i];
    }
    __syncthreads();
    for (int d = 1; d < 4; d++) {
        // blocks (i = j + d, j): first T_ij = sum_{k=j}^{i-1} L_ik W_kj, stored in
        // the (i, j) block of W (not read at this level), then W_ij = -W_ii T_ij
        // with all reads before the barrier and all writes after it.
        const int nbk = 4 - d;
        for (int e = tid; e < nbk * 256; e += CG_WG) {
            const int jb = e >> 8, ib = jb + d, rr = (e >> 4) & 15, cc = e & 15;
            double v = 0.0;
            for (int kb = jb; kb < ib; kb++)
#pragma unroll
                for (int k = 0; k < 16; k++) v += L[16 * ib + rr][16 * kb + k] * W[16 * kb + k][16 * jb + cc];
            W[16 * ib + rr][16 * jb + cc] = v;
        }
        __syncthreads();
        double out[3];
        for (int e = tid, u = 0; e < nbk * 256; e += CG_WG, u++) {
            const int jb = e >> 8, ib = jb + d, rr = (e >> 4) & 15, cc = e & 15;
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < 16; k++) v += W[16 * ib + rr][16 * ib + k] * W[16 * ib + k][16 * jb + cc];
            out[u] = -v;
        }
        __syncthreads();
        for (int e = tid, u = 0; e < nbk * 256; e += CG_WG, u++) {
            const int jb = e >> 8, ib = jb + d, rr = (e >> 4) & 15, cc = e & 15;
            W[16 * ib + rr][16 * jb + cc] = out[u];
        }
        __syncthreads();
    }
    return bad == 0;
}

// Factor + invert the (already updated) diagonal block held in L, store the
// factor and its inverse, and run the forward substitution of that block:
// y_kb = W_kb b_kb for the two right-hand sides (b already holds the updates).
// L and W are LDS blocks.
CG_DEV void cg_diag_block(const CgInst& I, int kb, double (*L)[CG_NB + 1], double (*W)[CG_NB + 1],
                          double* dinv, double* ytmp)
{
    const int P = I.P;
    const int c0 = kb * CG_NB;
    const int nb = min(CG_NB, P - c0);
    const bool ok = cg_lds_potrf_inv(L, W, nb, dinv);
    if (!ok && threadIdx.x == 0) I.scal[CS_PIVOT] = 1.0;
    double* Ld = I.Ldg + (size_t)kb * CG_NB * CG_NB;
    double* Wd = I.Winv + (size_t)kb * CG_NB * CG_NB;
    for (int e = threadIdx.x; e < CG_NB * CG_NB; e += CG_WG) {
        const int r = e / CG_NB, c = e % CG_NB;
        Ld[e] = (r < nb && c <= r) ? L[r][c] : 0.0;
        Wd[e] = (r < nb && c < nb) ? W[r][c] : 0.0;
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
            for (int k = 0; k <= q; k++) v += W[q][k] * ytmp[r * CG_NB + k];
            I.rhs[(size_t)r * P + c0 + q] = v;
        }
    }
}

// Block 0: factor/invert the first diagonal block (one workgroup per instance).
__global__ __launch_bounds__(CG_WG) void k_chol_diag0(const CgInst* __restrict__ insts, int B,
                                                      const int* __restrict__ pref)
{
    const int b = cg_find_inst(pref, B, blockIdx.x);
    const CgInst& I = insts[b];
    const int P = I.P;
    const int nb = min(CG_NB, P);
    __shared__ double L[CG_NB][CG_NB + 1];
    __shared__ double W[CG_NB][CG_NB + 1];
    __shared__ double dinv[CG_NB];
    __shared__ double ytmp[2 * CG_NB];
    for (int e = threadIdx.x; e < CG_NB * CG_NB; e += CG_WG) {
        const int r = e / CG_NB, c = e % CG_NB;
        L[r][c] = (r < nb && c <= r) ? I.S[(size_t)r * P + c] : 0.0;
    }
    cg_diag_block(I, 0, L, W, dinv, ytmp);
}

// Panel rows of block column kb (row tile rt + 1 below the diagonal):
// X = A21 W_kb^T, a 64x64x64 product on FP64 matrix cores.
__global__ __launch_bounds__(CG_WG) void k_chol_panel(const CgInst* __restrict__ insts, int B,
                                                      const int* __restrict__ pref, int kb)
{
    const int b = cg_find_inst(pref, B, blockIdx.x);
    const CgInst& I = insts[b];
    const int rt = blockIdx.x - pref[b] + 1;
    const int P = I.P;
    const int c0 = kb * CG_NB;
    const int r0 = c0 + rt * CG_NB;
    double* S = I.S;
    const double* Wd = I.Winv + (size_t)kb * CG_NB * CG_NB;
    __shared__ double As[CG_NB][CG_NB + 1];
    __shared__ double Ws[CG_NB][CG_NB + 1];
    for (int e = threadIdx.x; e < CG_NB * CG_NB; e += CG_WG) {
        const int r = e / CG_NB, c = e % CG_NB;
        As[r][c] = (r0 + r < P) ? S[(size_t)(r0 + r) * P + c0 + c] : 0.0;   // block kb is full
        Ws[r][c] = Wd[e];
    }
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
// panel rows of blocks TI and TJ.  Diagonal tiles also apply the forward-
// substitution update rhs_TI -= A y_kb; the tile (kb+1, kb+1) then factors and
// inverts itself (it is final after this update) and solves its rhs block.
__global__ __launch_bounds__(CG_WG) void k_chol_update(const CgInst* __restrict__ insts, int B,
                                                       const int* __restrict__ pref, int kb)
{
    const int b = cg_find_inst(pref, B, blockIdx.x);
    const CgInst& I = insts[b];
    int ti, tj;
    cg_tri_index(blockIdx.x - pref[b], ti, tj);
    const int TI = kb + 1 + ti, TJ = kb + 1 + tj;
    const int P = I.P;
    const int c0 = kb * CG_NB;
    const int ri = TI * CG_NB, rj = TJ * CG_NB;
    double* S = I.S;
    __shared__ double As[CG_NB][CG_NB + 1];
    __shared__ double Bs[CG_NB][CG_NB + 1];
    __shared__ double dinv[CG_NB];
    __shared__ double ytmp[2 * CG_NB];
    for (int e = threadIdx.x; e < CG_NB * CG_NB; e += CG_WG) {
        const int r = e / CG_NB, c = e % CG_NB;
        As[r][c] = (ri + r < P) ? S[(size_t)(ri + r) * P + c0 + c] : 0.0;
        Bs[r][c] = (rj + r < P) ? S[(size_t)(rj + r) * P + c0 + c] : 0.0;
    }
    __syncthreads();
    cg_d4 acc[2][2];
    cg_mfma_abt64(As, Bs, acc);
    if (TI != TJ) {
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int gr = ri + cg_mfma_row(a, r), gc = rj + cg_mfma_col(c);
                    if (gr < P && gc < P) S[(size_t)gr * P + gc] -= acc[a][c][r];
                }
        return;
    }
    // diagonal tile: forward-substitution update of rhs block TI
    if (threadIdx.x < 2 * CG_NB) {
        const int rhsi = threadIdx.x / CG_NB, r = threadIdx.x % CG_NB;
        if (ri + r < P) {
            const double* y = I.rhs + (size_t)rhsi * P + c0;
            double v = 0.0;
            for (int k = 0; k < CG_NB; k++) v += As[r][k] * y[k];
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
                    if (gr < P && gc <= gr) S[(size_t)gr * P + gc] -= acc[a][c][r];
                }
        return;
    }
    // tile (kb+1, kb+1): the updated block goes to LDS (reusing As/Bs), then is
    // factored, inverted and used for its forward substitution
    const int nb = min(CG_NB, P - ri);
    __syncthreads();   // As no longer read (rhs update above)
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int lr = cg_mfma_row(a, r), lc = cg_mfma_col(c);
                As[lr][lc] = (lr < nb && lc <= lr) ? (S[(size_t)(ri + lr) * P + rj + lc] - acc[a][c][r]) : 0.0;
            }
    __threadfence_block();
    cg_diag_block(I, kb + 1, As, Bs, dinv, ytmp);   // Bs is free now: it holds W
}

// Backward substitution step kb: x_kb = W_kb^T y_kb (every workgroup computes
// it), workgroup 0 stores x_kb, workgroup t >= 1 updates block t-1:
// y_{t-1} -= L[kb, t-1]^T x_kb.  y lives in rhs, x in ltmp.
__global__ __launch_bounds__(CG_WG) void k_trsv_back(const CgInst* __restrict__ insts, int B,
                                                     const int* __restrict__ pref, int kb)
{
    const int b = cg_find_inst(pref, B, blockIdx.x);
    const CgInst& I = insts[b];
    const int t = blockIdx.x - pref[b];
    const int P = I.P;
    const int c0 = kb * CG_NB;
    const int nb = min(CG_NB, P - c0);
    const double* S = I.S;
    const double* Wd = I.Winv + (size_t)kb * CG_NB * CG_NB;
    __shared__ double ys[2][CG_NB];
    __shared__ double xs[2][CG_NB];
    if (threadIdx.x < 2 * CG_NB) {
        const int r = threadIdx.x / CG_NB, q = threadIdx.x % CG_NB;
        ys[r][q] = (q < nb) ? I.rhs[(size_t)r * P + c0 + q] : 0.0;
    }
    __syncthreads();
    if (threadIdx.x < 2 * CG_NB) {
        // x[q] = sum_{c >= q} W[c][q] y[c]
        const int r = threadIdx.x / CG_NB, q = threadIdx.x % CG_NB;
        double v = 0.0;
        for (int c = q; c < nb; c++) v += Wd[c * CG_NB + q] * ys[r][c];
        xs[r][q] = v;
    }
    __syncthreads();
    if (t == 0) {
        if (threadIdx.x < 2 * CG_NB) {
            const int r = threadIdx.x / CG_NB, q = threadIdx.x % CG_NB;
            if (q < nb) I.ltmp[(size_t)r * P + c0 + q] = xs[r][q];
        }
        return;
    }
    const int blk = t - 1;
    if (threadIdx.x < 2 * CG_NB) {
        const int r = threadIdx.x / CG_NB, q = threadIdx.x % CG_NB;
        const int col = blk * CG_NB + q;
        double v = 0.0;
        for (int c = 0; c < nb; c++) v += S[(size_t)(c0 + c) * P + col] * xs[r][c];
        I.rhs[(size_t)r * P + col] -= v;
    }
}

// Whole backward substitution L^T x = y in one workgroup of 1024 threads per
// instance (the block steps are a dependent chain, so one launch replaces one
// launch per block).  Step kb: x_kb = W_kb^T y_kb, then y_col -= sum_c
// L[kb*64 + c][col] x_kb[c] for every col < kb*64 (one thread per column,
// coalesced reads of block row kb).  y lives in rhs, x goes to ltmp.
__global__ __launch_bounds__(1024) void k_trsv_all(const CgInst* __restrict__ insts, int B,
                                                   const int* __restrict__ pref)
{
    const int b = cg_find_inst(pref, B, blockIdx.x);
    const CgInst& I = insts[b];
    const int P = I.P;
    const int nblk = (P + CG_NB - 1) / CG_NB;
    const double* S = I.S;
    double* y0 = I.rhs;
    double* y1 = I.rhs + P;
    __shared__ double xs[2][CG_NB];
    const int tid = threadIdx.x;
    for (int kb = nblk - 1; kb >= 0; kb--) {
        const int c0 = kb * CG_NB;
        const int nb = min(CG_NB, P - c0);
        const double* Wd = I.Winv + (size_t)kb * CG_NB * CG_NB;
        if (tid < 2 * CG_NB) {
            const int r = tid / CG_NB, q = tid % CG_NB;
            const double* y = r ? y1 : y0;
            double v = 0.0;
            for (int c = q; c < nb; c++) v += Wd[c * CG_NB + q] * y[c0 + c];
            xs[r][q] = v;
            if (q < nb) I.ltmp[(size_t)r * P + c0 + q] = v;
        }
        __syncthreads();
        for (int col = tid; col < c0; col += 1024) {
            double v0 = 0.0, v1 = 0.0;
            const double* Sc = S + (size_t)c0 * P + col;
#pragma unroll 4
            for (int c = 0; c < nb; c++) {
                const double sv = Sc[(size_t)c * P];
                v0 += sv * xs[0][c];
                v1 += sv * xs[1][c];
            }
            y0[col] -= v0;
            y1[col] -= v1;
        }
        __syncthreads();
    }
}
