// Per-wave row bodies of the dense Schur path, shared by the Schur kernels
// (k_schur_build, k_schur_rhs) and the mixed Newton launch (k_newton_mixed),
// which runs them beside the sparse factorization of the other instances.
#pragma once
#include "cgot_common.h"

// Build row s2 < P of the lower Schur complement, one wave per row:
// S[s2][s] = D_small[s2] delta - sum_{l adjacent to s2} w(s2,l) w(l,s) / D_large[l]
// for s <= s2.  Rows of up to CG_SCHUR_LDS entries accumulate in LDS: the wave
// takes the adjacent l in chunks of 8 (one group of 8 lanes per l, so the
// gathers of 8 l's are in flight together) and applies the 8 groups' updates
// one group after another, so every entry receives its terms in a fixed
// order; the finished row is written once.  Longer rows update S in place, l
// by l in ascending order with a fence between l steps.

// R: the wave's LDS row buffer of P doubles (used when P <= CG_SCHUR_LDS).
CG_DEV void cg_schur_build_row(const CgInstD& I, int s2, double* R)
{
    const int P = I.P;
    const int lane = threadIdx.x & 63;
    const CG_AS1 int* sp; const CG_AS1 int* sidx; const CG_AS1 double* sval;
    const CG_AS1 int* lp; const CG_AS1 int* lidx; const CG_AS1 double* lval;
    const CG_AS1 double* Ds; const CG_AS1 double* Dl;
    if (I.beta_side) {
        sp = I.w_colptr; sidx = I.w_rowi; sval = I.w_cval;
        lp = I.w_rowptr; lidx = I.w_colj; lval = I.w_rval;
        Ds = I.Db; Dl = I.Da;
    } else {
        sp = I.w_rowptr; sidx = I.w_colj; sval = I.w_rval;
        lp = I.w_colptr; lidx = I.w_rowi; lval = I.w_cval;
        Ds = I.Da; Dl = I.Db;
    }
    auto Srow = I.S + (size_t)s2 * P;
    const int p0 = sp[s2], p1 = sp[s2 + 1];
    if (P > CG_SCHUR_LDS) {
        if (lane == 0) Srow[s2] = Ds[s2];
        for (int p = p0; p < p1; p++) {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            const int l = sidx[p];
            const double ws = sval[p];
            const double dl = Dl[l];
            for (int q = lp[l] + lane; q < lp[l + 1]; q += 64) {
                const int s = lidx[q];
                if (s <= s2) Srow[s] -= ws * lval[q] / dl;
            }
        }
        return;
    }
    const double dss = Ds[s2];
    for (int s = lane; s <= s2; s += 64) R[s] = (s == s2) ? dss : 0.0;
    cg_wave_sync();
    const int g = lane >> 3, t = lane & 7;
    for (int pb = p0; pb < p1; pb += 8) {
        const int p = pb + g;
        const bool okp = p < p1;
        const int l = okp ? sidx[p] : 0;
        const double ws = okp ? sval[p] : 0.0;
        const double dl = okp ? Dl[l] : 1.0;
        const int q0 = okp ? lp[l] : 0;
        const int q1 = okp ? lp[l + 1] : 0;
        int len = q1 - q0;
#pragma unroll
        for (int o = 8; o < 64; o <<= 1) len = max(len, __shfl_xor(len, o));
        for (int qb = 0; qb < len; qb += 8) {
            const int q = q0 + qb + t;
            const bool okq = q < q1;
            const int s = okq ? lidx[q] : 0;
            const double v = okq ? ws * lval[q] / dl : 0.0;
            for (int gg = 0; gg < 8; gg++) {
                if (g == gg && okq && s <= s2) R[s] -= v;
                cg_wave_sync();
            }
        }
    }
    for (int s = lane; s <= s2; s += 64) Srow[s] = R[s];
}

// Schur right-hand sides for [g, y]: r_s = g_s - sum_l w(s,l) g_l / D_l (one wave, s < P).
CG_DEV void cg_schur_rhs_row(const CgInstD& I, int s)
{
    const int P = I.P, n = I.n;
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
