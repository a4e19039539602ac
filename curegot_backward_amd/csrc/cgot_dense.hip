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
// is a right-looking blocked Cholesky with CG_NB = 64:
//   k_chol_panel(kb):  factor the diagonal block in LDS, solve the panel rows
//                      below it, and the forward substitution of block kb;
//   k_chol_update(kb): trailing update S22 -= L21 L21^T on lower tiles, plus the
//                      forward-substitution update of the right-hand sides;
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

// Scatter the Schur complement.  Thread s owns column s of the lower triangle:
// S[s][s] = D_small[s]; for each large index l adjacent to s (ascending) and
// each small s' >= s adjacent to l: S[s'][s] -= w(s,l) w(l,s') / D_large[l].
__global__ __launch_bounds__(CG_WG) void k_schur_build(const CgInst* __restrict__ insts,
                                                       const int* __restrict__ wg_inst,
                                                       const double* __restrict__ ctl)
{
    const int b = wg_inst[blockIdx.x];
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return;
    const CgInst& I = insts[b];
    const int s = (blockIdx.x - I.vec0) * CG_WG + threadIdx.x;
    const int P = I.P;
    if (s >= P) return;
    // small-side adjacency (s -> l) and large-side adjacency (l -> s')
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
    double* S = I.S;
    S[(size_t)s * P + s] = Ds[s];
    for (int p = sp[s]; p < sp[s + 1]; p++) {
        const int l = sidx[p];
        const double ws = sval[p];
        const double dl = Dl[l];
        for (int q = lp[l]; q < lp[l + 1]; q++) {
            const int s2 = lidx[q];
            if (s2 >= s) S[(size_t)s2 * P + s] -= lval[q] * ws / dl;
        }
    }
}

// Schur right-hand sides for [g, y]: r_s = g_s - sum_l w(s,l) g_l / D_l.
__global__ __launch_bounds__(CG_WG) void k_schur_rhs(const CgInst* __restrict__ insts,
                                                     const int* __restrict__ wg_inst,
                                                     const double* __restrict__ ctl)
{
    const int b = wg_inst[blockIdx.x];
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return;
    const CgInst& I = insts[b];
    const int s = (blockIdx.x - I.vec0) * CG_WG + threadIdx.x;
    const int P = I.P, n = I.n;
    if (s >= P) return;
    const double* src[2] = {I.grad, I.yv};
    for (int r = 0; r < 2; r++) {
        const double* g = src[r];
        double v;
        if (I.beta_side) {
            v = g[n + s];
            for (int p = I.w_colptr[s]; p < I.w_colptr[s + 1]; p++) {
                const int i = I.w_rowi[p];
                v -= I.w_cval[p] * (g[i] / I.Da[i]);
            }
        } else {
            v = g[s];
            for (int p = I.w_rowptr[s]; p < I.w_rowptr[s + 1]; p++) {
                const int j = I.w_colj[p];
                v -= I.w_rval[p] * (g[n + j] / I.Db[j]);
            }
        }
        I.rhs[(size_t)r * P + s] = v;
    }
}

// Back-substitute the large side and assemble [dir, u] in the (alpha, beta_t) layout.
__global__ __launch_bounds__(CG_WG) void k_schur_back(const CgInst* __restrict__ insts,
                                                      const int* __restrict__ wg_inst,
                                                      const double* __restrict__ ctl)
{
    const int b = wg_inst[blockIdx.x];
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return;
    const CgInst& I = insts[b];
    const int q = (blockIdx.x - I.vec0) * CG_WG + threadIdx.x;
    const int P = I.P, n = I.n, mm = I.m - 1;
    const double* src[2] = {I.grad, I.yv};
    double* dst[2] = {I.dir, I.uv};
    const double* sol = I.ltmp;   // Schur solutions [2*P]
    for (int r = 0; r < 2; r++) {
        const double* g = src[r];
        double* x = dst[r];
        const double* xs = sol + (size_t)r * P;
        if (I.beta_side) {
            if (q < n) {
                double v = g[q];
                for (int p = I.w_rowptr[q]; p < I.w_rowptr[q + 1]; p++) v -= I.w_rval[p] * xs[I.w_colj[p]];
                x[q] = v / I.Da[q];
            }
            if (q < mm) x[n + q] = xs[q];
        } else {
            if (q < mm) {
                double v = g[n + q];
                for (int p = I.w_colptr[q]; p < I.w_colptr[q + 1]; p++) v -= I.w_cval[p] * xs[I.w_rowi[p]];
                x[n + q] = v / I.Db[q];
            }
            if (q < n) x[q] = xs[q];
        }
    }
}

// ---------------------------------------------------------------------------
// Blocked Cholesky
// ---------------------------------------------------------------------------
// Factor the nb x nb lower block held in L (LDS, row stride CG_NB+1) in place.
// Right-looking with the column scaling deferred: step c subtracts
// A[r][c] A[q][c] / A[c][c] from the trailing block, one barrier per column.
CG_DEV bool cg_lds_potrf(double (*L)[CG_NB + 1], int nb, double* dsq)
{
    bool ok = true;
    for (int c = 0; c < nb; c++) {
        __syncthreads();
        const double dcc = L[c][c];
        if (!(dcc > 0.0)) ok = false;
        const int w = nb - c - 1;
        for (int e = threadIdx.x; e < w * w; e += CG_WG) {
            const int r = c + 1 + e / w, q = c + 1 + e % w;
            if (q <= r) L[r][q] -= L[r][c] * L[q][c] / dcc;
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < nb; c += CG_WG) dsq[c] = sqrt(L[c][c]);
    __syncthreads();
    for (int e = threadIdx.x; e < nb * nb; e += CG_WG) {
        const int r = e / nb, c = e % nb;
        if (r > c) L[r][c] /= dsq[c];
        else if (r == c) L[r][c] = dsq[c];
    }
    __syncthreads();
    return ok;
}

__global__ __launch_bounds__(CG_WG) void k_chol_panel(const CgInst* __restrict__ insts, int B,
                                                      const int* __restrict__ pref, int kb)
{
    const int b = cg_find_inst(pref, B, blockIdx.x);
    const CgInst& I = insts[b];
    const int rt = blockIdx.x - pref[b];
    const int P = I.P;
    const int c0 = kb * CG_NB;
    const int nb = min(CG_NB, P - c0);
    double* S = I.S;
    __shared__ double L[CG_NB][CG_NB + 1];
    __shared__ double X[CG_NB][CG_NB + 1];
    __shared__ double dsq[CG_NB];
    for (int e = threadIdx.x; e < CG_NB * CG_NB; e += CG_WG) {
        const int r = e / CG_NB, c = e % CG_NB;
        L[r][c] = (r < nb && c <= r) ? S[(size_t)(c0 + r) * P + c0 + c] : 0.0;
    }
    const bool ok = cg_lds_potrf(L, nb, dsq);
    if (rt == 0) {
        // The factor goes to a side buffer, never into S: other workgroups of
        // this launch may still be loading the unfactored block from S.
        if (!ok && threadIdx.x == 0) I.scal[CS_PIVOT] = 1.0;
        double* Ld = I.Ldg + (size_t)kb * CG_NB * CG_NB;
        for (int e = threadIdx.x; e < CG_NB * CG_NB; e += CG_WG) {
            const int r = e / CG_NB, c = e % CG_NB;
            Ld[e] = L[r][c];
        }
        // forward substitution of the block: y = L11^{-1} b (two right-hand sides)
        if (threadIdx.x < 2) {
            double* y = I.rhs + (size_t)threadIdx.x * P + c0;
            for (int r = 0; r < nb; r++) {
                double v = y[r];
                for (int q = 0; q < r; q++) v -= L[r][q] * y[q];
                y[r] = v / L[r][r];
            }
        }
        return;
    }
    // panel rows r0..r0+63: X = A21 L11^{-T}
    const int r0 = c0 + rt * CG_NB;
    const int nr = min(CG_NB, P - r0);
    for (int e = threadIdx.x; e < CG_NB * CG_NB; e += CG_WG) {
        const int r = e / CG_NB, c = e % CG_NB;
        X[r][c] = (r < nr && c < nb) ? S[(size_t)(r0 + r) * P + c0 + c] : 0.0;
    }
    __syncthreads();
    if (threadIdx.x < nr) {
        const int r = threadIdx.x;
        for (int c = 0; c < nb; c++) {
            double v = X[r][c];
            for (int q = 0; q < c; q++) v -= X[r][q] * L[c][q];
            X[r][c] = v / L[c][c];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nr * nb; e += CG_WG) {
        const int r = e / nb, c = e % nb;
        S[(size_t)(r0 + r) * P + c0 + c] = X[r][c];
    }
}

// Trailing update of tile (TI, TJ), kb < TJ <= TI: C -= A B^T with A, B the
// panel rows of blocks TI and TJ.  Diagonal tiles also apply the forward
// substitution update rhs_TI -= A y_kb.
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
    for (int e = threadIdx.x; e < CG_NB * CG_NB; e += CG_WG) {
        const int r = e / CG_NB, c = e % CG_NB;
        As[r][c] = (ri + r < P) ? S[(size_t)(ri + r) * P + c0 + c] : 0.0;
        Bs[r][c] = (rj + r < P) ? S[(size_t)(rj + r) * P + c0 + c] : 0.0;
    }
    __syncthreads();
    const int tr = (threadIdx.x >> 4) * 4, tc = (threadIdx.x & 15) * 4;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[a][c] = 0.0;
    for (int k = 0; k < CG_NB; k++) {
        double av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; a++) { av[a] = As[tr + a][k]; bv[a] = Bs[tc + a][k]; }
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int c = 0; c < 4; c++) acc[a][c] += av[a] * bv[c];
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int r = ri + tr + a, col = rj + tc + c;
            if (r < P && col < P && (TI != TJ || col <= r)) S[(size_t)r * P + col] -= acc[a][c];
        }
    if (TI == TJ && threadIdx.x < 2 * CG_NB) {
        const int rhsi = threadIdx.x / CG_NB, r = threadIdx.x % CG_NB;
        if (ri + r < P) {
            const double* y = I.rhs + (size_t)rhsi * P + c0;
            double v = 0.0;
            for (int k = 0; k < CG_NB; k++) v += As[r][k] * y[k];
            I.rhs[(size_t)rhsi * P + ri + r] -= v;
        }
    }
}

// Backward substitution step kb: x_kb = L_kk^{-T} y_kb (every workgroup
// recomputes it), workgroup 0 stores x_kb, workgroup t >= 1 updates block t-1:
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
    const double* Ld = I.Ldg + (size_t)kb * CG_NB * CG_NB;
    __shared__ double L[CG_NB][CG_NB + 1];
    __shared__ double xs[2][CG_NB];
    for (int e = threadIdx.x; e < CG_NB * CG_NB; e += CG_WG) {
        const int r = e / CG_NB, c = e % CG_NB;
        L[r][c] = (r < nb && c <= r) ? Ld[e] : 0.0;
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (w < 2) {
        // wave w solves right-hand side w; lane q holds y_q
        double yq = (lane < nb) ? I.rhs[(size_t)w * P + c0 + lane] : 0.0;
        for (int c = nb - 1; c >= 0; c--) {
            const double xc = __shfl(yq, c, 64) / L[c][c];
            if (lane == c) yq = xc;
            else if (lane < c) yq -= L[c][lane] * xc;
        }
        if (lane < nb) xs[w][lane] = yq;
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
