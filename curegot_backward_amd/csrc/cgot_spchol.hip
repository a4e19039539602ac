// Sparse Newton solve of the SPLR step: factor A = [[Da, W], [W^T, Db]] in the
// plan's elimination order and solve A [dir, u] = [g, y], one workgroup per
// instance (plan layout and analysis: cgot_symbolic.h / .cpp).
//
// Reference: the cuDSS factorization and solve of SparseCholeskySolver
// (src/linsolve.cu:27-268) called from SPLRSolver::compute_search_direc
// (src/sinkhorn_splr.cu:538-649), which factor the same matrix every
// iteration.  Here the factor has two parts:
//   * the periphery (sparse columns, a few elimination-tree levels): every
//     entry is one gather  L[q] = A(q) - sum L[x] L[y]  over its precomputed
//     pair list, then a square root or a division by its column's diagonal;
//     all entries of one level are independent, levels are separated by
//     workgroup barriers;
//   * the dense core (R <= SP_RMAX vertices adjacent to at least half of each
//     other): its Schur complement C = A22 - L21 L21^T is gathered into LDS
//     (packed lower triangle), factored right-looking in LDS (unscaled
//     C(i,j) -= C(i,k) C(j,k) / C(k,k), 4 x 4 register tiles of the trailing
//     triangle per thread, one barrier per column), then scaled to L.
// The solves run the same schedule: forward substitution by periphery level,
// the core's right-hand side, the dense core's two triangular solves in one
// wave, backward substitution by periphery level in reverse.
//
// Every sum runs in the fixed order of its list, so the result depends only
// on the instance's own data.
#include "cgot_kernels.h"

namespace {

struct SpView {
    int N, n, t, R, H, ncore;
    const CG_AS1 int* perm;
    const CG_AS1 int* lvptr;
    const CG_AS1 int* lcptr;
    const CG_AS1 int* lcols;
    const CG_AS1 int* emeta;    // 4 ints per entry
    const CG_AS1 int* erow;
    const CG_AS1 int* ecol;
    const CG_AS1 int* pairs;    // 2 ints per pair
    const CG_AS1 int* cmeta;    // 4 ints per core entry
    const CG_AS1 int* cpairs;
    const CG_AS1 int* rptr;
    const CG_AS1 int* rq;
    CG_AS1 double* lval;
    CG_AS1 double* lcore;
    CG_AS1 double* yx;
};

CG_DEV SpView sp_view(const CG_AS1 SpHdr* hp, CG_AS1 double* plan)
{
    CG_AS1 char* base = reinterpret_cast<CG_AS1 char*>(plan);
    auto ip = [&](int64_t off) { return reinterpret_cast<const CG_AS1 int*>(base + off); };
    auto dp = [&](int64_t off) { return reinterpret_cast<CG_AS1 double*>(base + off); };
    SpView v;
    v.N = hp->N; v.n = hp->n; v.t = hp->t; v.R = hp->R; v.H = hp->H; v.ncore = hp->ncore;
    v.perm = ip(hp->o_perm);
    v.lvptr = ip(hp->o_lvptr);
    v.lcptr = ip(hp->o_lcptr);
    v.lcols = ip(hp->o_lcols);
    v.emeta = ip(hp->o_emeta);
    v.erow = ip(hp->o_erow);
    v.ecol = ip(hp->o_ecol);
    v.pairs = ip(hp->o_pairs);
    v.cmeta = ip(hp->o_cmeta);
    v.cpairs = ip(hp->o_cpairs);
    v.rptr = ip(hp->o_rptr);
    v.rq = ip(hp->o_rq);
    v.lval = dp(hp->o_lval);
    v.lcore = dp(hp->o_lcore);
    v.yx = dp(hp->o_yx);
    return v;
}

// Entry of A by source code: -1 fill, [0, n) Da, [n, N) Db, N + q W's row-CSR entry q.
CG_DEV double sp_aval(const CgInstD& I, int s, int N)
{
    if (s < 0) return 0.0;
    if (s < I.n) return I.Da[s];
    if (s < N) return I.Db[s - I.n];
    return I.w_rval[s - N];
}

// A(q) - sum over the pair list [p0, p1) of L[x] L[y], in list order.
CG_DEV double sp_gather(double s, const CG_AS1 int* pairs, const CG_AS1 double* L, int p0, int p1)
{
    int p = p0;
    for (; p + 1 < p1; p += 2) {
        const int ax = pairs[2 * p], ay = pairs[2 * p + 1], bx = pairs[2 * p + 2], by = pairs[2 * p + 3];
        const double xa = L[ax], ya = L[ay], xb = L[bx], yb = L[by];
        s -= xa * ya;
        s -= xb * yb;
    }
    if (p < p1) s -= L[pairs[2 * p]] * L[pairs[2 * p + 1]];
    return s;
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

// Numeric factorization: periphery columns into lval, the core's factor into
// the LDS packed triangle C (and lcore).  Returns false on a non-positive pivot.
CG_DEV bool sp_factor(const CgInstD& I, const SpView& v, double* C)
{
    const int tid = threadIdx.x;
    bool ok = true;
    for (int l = 0; l < v.H; l++) {
        const int lo = v.lvptr[l], hi = v.lvptr[l + 1];
        for (int q = lo + tid; q < hi; q += SP_WG) {
            const CG_AS1 int* e = v.emeta + 4 * q;
            const int asrc = e[0], p0 = e[1], p1 = e[2], dpos = e[3];
            double s = sp_gather(sp_aval(I, asrc, v.N), v.pairs, v.lval, p0, p1);
            if (dpos == q) {
                ok = ok && (s > 0.0);
                s = sqrt(s);
            }
            v.lval[q] = s;
        }
        __syncthreads();
        for (int q = lo + tid; q < hi; q += SP_WG) {
            const int d = v.emeta[4 * q + 3];
            if (d != q) v.lval[q] = v.lval[q] / v.lval[d];
        }
        __syncthreads();
    }
    const int R = v.R;
    if (R == 0) return ok;
    for (int c = tid; c < v.ncore; c += SP_WG) {
        const CG_AS1 int* e = v.cmeta + 4 * c;
        const int asrc = e[0], p0 = e[1], p1 = e[2];
        C[c] = sp_gather(sp_aval(I, asrc, v.N), v.cpairs, v.lval, p0, p1);
    }
    __syncthreads();
    for (int k = 0; k < R; k++) {
        const double dk = C[tri(k) + k];
        ok = ok && (dk > 0.0);
        const double inv = 1.0 / dk;
        const int m2 = R - k - 1;
        const int nt = (m2 + 3) >> 2;
        const int ntiles = nt * (nt + 1) / 2;
        for (int u = tid; u < ntiles; u += SP_WG) {
            int ti, tj;
            tri_decode(u, ti, tj);
            const int i0 = k + 1 + 4 * ti, j0 = k + 1 + 4 * tj;
            double cj[4];
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int j = j0 + b;
                cj[b] = (j < R) ? C[tri(j) + k] * inv : 0.0;
            }
#pragma unroll
            for (int a = 0; a < 4; a++) {
                const int i = i0 + a;
                if (i < R) {
                    const double ci = C[tri(i) + k];
                    double* row = C + tri(i);
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        const int j = j0 + b;
                        if (j <= i) row[j] -= ci * cj[b];
                    }
                }
            }
        }
        __syncthreads();
    }
    // L(i, j) = C(i, j) / sqrt(C(j, j)) below the diagonal, then the diagonal
    for (int c = tid; c < v.ncore; c += SP_WG) {
        int i, j;
        tri_decode(c, i, j);
        if (i != j) {
            const double x = C[c] / sqrt(C[tri(j) + j]);
            C[c] = x;
            v.lcore[c] = x;
        }
    }
    __syncthreads();
    for (int j = tid; j < R; j += SP_WG) {
        const double x = sqrt(C[tri(j) + j]);
        C[tri(j) + j] = x;
        v.lcore[tri(j) + j] = x;
    }
    __syncthreads();
    return ok;
}

// x = A^{-1} b for NR right-hand sides (b0, b1 and x0, x1 indexed by vertex),
// with the factor in lval and the core factor in LDS C; yx [2N] holds the
// substitutions by position, rc [2R] the core's right-hand sides (LDS).
template <int NR>
CG_DEV void sp_solve(const SpView& v, const double* C, const CG_AS1 double* b0, const CG_AS1 double* b1,
                     CG_AS1 double* x0, CG_AS1 double* x1, CG_AS1 double* yx, double* rc)
{
    const int tid = threadIdx.x;
    const int N = v.N, t = v.t, R = v.R;
    for (int l = 0; l < v.H; l++) {
        for (int c = v.lcptr[l] + tid; c < v.lcptr[l + 1]; c += SP_WG) {
            const int d = v.lcols[c];
            const int j = v.erow[d];
            const int vtx = v.perm[j];
            double s0 = b0[vtx];
            double s1 = (NR > 1) ? b1[vtx] : 0.0;
            for (int p = v.rptr[j]; p < v.rptr[j + 1]; p++) {
                const int q = v.rq[p];
                const int col = v.ecol[q];
                const double L = v.lval[q];
                s0 -= L * yx[col];
                if (NR > 1) s1 -= L * yx[N + col];
            }
            const double dd = v.lval[d];
            yx[j] = s0 / dd;
            if (NR > 1) yx[N + j] = s1 / dd;
        }
        __syncthreads();
    }
    if (R > 0) {
        for (int i = t + tid; i < N; i += SP_WG) {
            const int vtx = v.perm[i];
            double s0 = b0[vtx];
            double s1 = (NR > 1) ? b1[vtx] : 0.0;
            for (int p = v.rptr[i]; p < v.rptr[i + 1]; p++) {
                const int q = v.rq[p];
                const int col = v.ecol[q];
                const double L = v.lval[q];
                s0 -= L * yx[col];
                if (NR > 1) s1 -= L * yx[N + col];
            }
            rc[i - t] = s0;
            rc[R + i - t] = s1;
        }
        __syncthreads();
        if (tid < 64) {
            // forward: column k's pivot, then the rows below it (lanes)
            for (int k = 0; k < R; k++) {
                const double lkk = C[tri(k) + k];
                const double z0 = rc[k] / lkk;
                const double z1 = rc[R + k] / lkk;
                for (int i = k + 1 + tid; i < R; i += 64) {
                    const double lik = C[tri(i) + k];
                    rc[i] -= lik * z0;
                    rc[R + i] -= lik * z1;
                }
                cg_wave_sync();
                if (tid == 0) {
                    rc[k] = z0;
                    rc[R + k] = z1;
                }
                cg_wave_sync();
            }
            // backward: x_k from row k of L^T (column k of L), then the rows above (row k of L)
            for (int k = R - 1; k >= 0; k--) {
                const double lkk = C[tri(k) + k];
                const double z0 = rc[k] / lkk;
                const double z1 = rc[R + k] / lkk;
                for (int i = tid; i < k; i += 64) {
                    const double lki = C[tri(k) + i];
                    rc[i] -= lki * z0;
                    rc[R + i] -= lki * z1;
                }
                cg_wave_sync();
                if (tid == 0) {
                    rc[k] = z0;
                    rc[R + k] = z1;
                }
                cg_wave_sync();
            }
            for (int i = tid; i < R; i += 64) {
                yx[t + i] = rc[i];
                if (NR > 1) yx[N + t + i] = rc[R + i];
            }
        }
        __syncthreads();
    }
    for (int l = v.H - 1; l >= 0; l--) {
        for (int c = v.lcptr[l] + tid; c < v.lcptr[l + 1]; c += SP_WG) {
            const int d = v.lcols[c], e = v.lcols[c + 1];
            const int j = v.erow[d];
            double s0 = yx[j];
            double s1 = (NR > 1) ? yx[N + j] : 0.0;
            for (int q = d + 1; q < e; q++) {
                const int r = v.erow[q];
                const double L = v.lval[q];
                s0 -= L * yx[r];
                if (NR > 1) s1 -= L * yx[N + r];
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

}  // namespace

// Newton step of the instances ids[0 .. gridDim.x): factor A at the current
// values (Da, Db, w_rval) and solve A [dir, uv] = [grad, yv].
__global__ __launch_bounds__(SP_WG) void k_sp_newton(const CgInst* __restrict__ insts, const int* __restrict__ ids)
{
    extern __shared__ double sp_lds[];
    const int b = ids[blockIdx.x];
    const CgInstD& I = cg_inst(insts, b);
    const SpView v = sp_view(I.sph, I.S);
    const bool ok = sp_factor(I, v, sp_lds);
    if (!ok) I.scal[CS_PIVOT] = 1.0;
    sp_solve<2>(v, sp_lds, I.grad, I.yv, I.dir, I.uv, v.yx, sp_lds + v.ncore);
}

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
    for (int c = threadIdx.x; c < v.ncore; c += SP_WG) sp_lds[c] = v.lcore[c];
    __syncthreads();
    sp_solve<1>(v, sp_lds, I.yv, nullptr, I.sv, nullptr, I.spx, sp_lds + v.ncore);
}
