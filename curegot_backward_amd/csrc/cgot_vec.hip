// Finalize and vector kernels: they turn the fixed-order partials of the M
// passes into row/column sums, duals, gradients and per-instance scalars, and
// run the O(n+m) parts of the SPLR step.  Grid: the "vec" map, i.e.
// ceil((n+m)/256) workgroups per instance; the map entry of a workgroup names the
// instance.  Per-instance scalar reductions use cg_inst_reduce (last
// workgroup sums the partials in workgroup order).
#include "cgot_common.h"
#include "cgot_kernels.h"

CG_DEV bool cg_decode_vec(const CgInst* insts, const int* wg_inst, const double* ctl, int& b, int& wgi)
{
    const int2 e = cg_map_entry(wg_inst);
    b = e.x;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return false;
    wgi = e.y;
    return true;
}

// wgpart of every instance to CG_SENT (once per solve, after the descriptors
// are uploaded; cg_inst_reduce leaves it that way after each use).
__global__ __launch_bounds__(CG_WG) void k_wgpart_init(const CgInst* __restrict__ insts)
{
    const CgInstD& I = cg_inst(insts, blockIdx.x);
    if (!I.wgpart) return;
    auto p = reinterpret_cast<CG_AS1 unsigned long long*>(I.wgpart);
    for (int e = threadIdx.x; e < 8 * I.nvec; e += CG_WG) p[e] = CG_SENT;
}

// ---------------------------------------------------------------------------
// T-pass finalize: row/col sums, gradient, f = reg*sum(T) - <gamma, ab>,
// ||g||^2, <g, dir>, <g_prev, dir>.  In trial mode also stores the trial point
// into gamma.  Reference: obj_grad_kernel and line_search_fused_kernel1/3,
// src/sinkhorn_splr_kernel.cu:354-443, src/sinkhorn_splr_kernel_line_search.cu:93-358.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(CG_WG) void k_tpass_fin(const CgInst* __restrict__ insts,
                                                     const int* __restrict__ wg_inst,
                                                     const double* __restrict__ ctl, int mode)
{
    int b, wgi;
    if (!cg_decode_vec(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int n = I.n, m = I.m, N = I.N;
    const int q = wgi * CG_WG + threadIdx.x;
    const double t = ctl[b * CG_NCTL + CTL_STEP];
    auto grad = (mode == CG_PT_BCD) ? I.grad_bcd : I.grad;
    const auto gam = (mode == CG_PT_BCD) ? I.gamma_bcd : I.gamma;
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};   // tsum, dot, g2, dg, dg0, g2 over alpha
    if (q < n) {
        const double rs = cg_ordered_sum(I.rp0 + (size_t)q * I.nstrips, I.nstrips, 1);
        I.rowsum[q] = rs;
        const double g = rs - I.ab[q];
        grad[q] = g;
        v[0] = rs;
        v[2] = g * g;
        v[5] = g * g;
        if (mode == CG_PT_TRIAL) { v[3] = g * I.dir[q]; v[4] = I.grad_prev[q] * I.dir[q]; }
        const double p = (mode == CG_PT_TRIAL) ? (I.gamma_prev[q] + t * I.dir[q]) : gam[q];
        if (mode == CG_PT_TRIAL) I.gamma[q] = p;
        v[1] = p * I.ab[q];
    }
    if (q < m) {
        const double cs = cg_ordered_sum<32>(I.cp0 + q, I.nrb, m);
        I.colsum[q] = cs;
        const int k = n + q;
        if (k < N) {
            const double g = cs - I.ab[k];
            grad[k] = g;
            v[2] += g * g;
            if (mode == CG_PT_TRIAL) { v[3] += g * I.dir[k]; v[4] += I.grad_prev[k] * I.dir[k]; }
        }
        double p;
        if (mode == CG_PT_TRIAL) { p = (k < N) ? (I.gamma_prev[k] + t * I.dir[k]) : 0.0; I.gamma[k] = p; }
        else p = gam[k];
        v[1] += p * I.ab[k];
    }
    double tot[6];
    if (cg_inst_reduce<6>(I, wgi, v, tot, 0) && threadIdx.x == 0) {
        const double f = I.reg * tot[0] - tot[1];
        I.scal[CS_G2A] = tot[5];
        if (mode == CG_PT_BCD) {
            I.scal[CS_FBCD] = f;
            I.scal[CS_G2BCD] = tot[2];
        } else {
            I.scal[CS_F] = f;
            I.scal[CS_G2] = tot[2];
            I.scal[CS_DG] = tot[3];
            if (mode == CG_PT_TRIAL) I.scal[CS_DG0] = tot[4];
        }
    }
}

// alpha_i = reg*(log a_i - LSE_j) into gamma or gamma_bcd.
// Reference: optimal_alpha_kernel, src/sinkhorn_bcd_kernel.cu:181-190.
__global__ __launch_bounds__(CG_WG) void k_rowlse_fin(const CgInst* __restrict__ insts,
                                                      const int* __restrict__ wg_inst,
                                                      const double* __restrict__ ctl, int mode)
{
    CG_EXP_TABLE(tab);
    int b, wgi;
    if (!cg_decode_vec(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int q = wgi * CG_WG + threadIdx.x;
    if (q >= I.n) return;
    double mx, s;
    cg_ordered_lse(I.rp0 + (size_t)q * I.nstrips, I.rp1 + (size_t)q * I.nstrips, I.nstrips, 1, mx, s, tab);
    auto alpha = (mode == CG_PT_BCD) ? I.gamma_bcd : I.gamma;
    alpha[q] = I.reg * (I.logab[q] - (mx + log(s)));
}

// beta_j = reg*(log b_j - LSE_i).  Reference: optimal_beta_kernel,
// src/sinkhorn_bcd_kernel.cu:101-110.
__global__ __launch_bounds__(CG_WG) void k_collse_fin(const CgInst* __restrict__ insts,
                                                      const int* __restrict__ wg_inst,
                                                      const double* __restrict__ ctl, int mode)
{
    CG_EXP_TABLE(tab);
    int b, wgi;
    if (!cg_decode_vec(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int q = wgi * CG_WG + threadIdx.x;
    if (q >= I.m) return;
    double mx, s;
    cg_ordered_lse(I.cp0 + q, I.cp1 + q, I.nrb, I.m, mx, s, tab);
    auto beta = ((mode == CG_PT_BCD) ? I.gamma_bcd : I.gamma) + I.n;
    beta[q] = I.reg * (I.logab[I.n + q] - (mx + log(s)));
}

// Gauge fix alpha += beta_{m-1}, beta -= beta_{m-1}, beta_{m-1} = 0.
// Reference: shift_gamma, src/sinkhorn_splr_kernel.cu:156-197.  One workgroup
// per instance so the shift value is read before it is cleared.
__global__ __launch_bounds__(CG_WG) void k_gauge_shift(const CgInst* __restrict__ insts,
                                                       const double* __restrict__ ctl, int mode)
{
    const int b = blockIdx.x;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return;
    const CgInstD& I = cg_inst(insts, b);
    auto g = (mode == CG_PT_BCD) ? I.gamma_bcd : I.gamma;
    const double sh = g[I.N];
    __syncthreads();
    for (int k = threadIdx.x; k < I.N; k += CG_WG) g[k] += (k < I.n) ? sh : -sh;
    __syncthreads();
    if (threadIdx.x == 0) g[I.N] = 0.0;
}

// ab = (a, b), logab = log(ab), gamma_prev = 0, grad_prev = 0, and gamma = x0
// or 0 (x0 may alias gamma).  ptrs[b*3 + {0,1,2}] = a, b, x0 (x0 may be null).
// Reference: SPLRSolver constructor and init_dual, src/sinkhorn_splr.cu:182-265, 330-359.
__global__ __launch_bounds__(CG_WG) void k_init_logab(const CgInst* __restrict__ insts,
                                                      const int* __restrict__ wg_inst,
                                                      const double* __restrict__ ctl,
                                                      const double* const* __restrict__ ptrs,
                                                      const double* const* __restrict__ unused)
{
    int b, wgi;
    if (!cg_decode_vec(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int k = wgi * CG_WG + threadIdx.x;
    const int n = I.n, m = I.m;
    if (k >= n + m) return;
    const double* pa = ptrs[b * 3 + 0];
    const double* pb = ptrs[b * 3 + 1];
    const double* px = ptrs[b * 3 + 2];
    const double v = (k < n) ? pa[k] : pb[k - n];
    I.ab[k] = v;
    I.logab[k] = log(v);
    I.gamma_prev[k] = 0.0;
    if (k < I.N) I.grad_prev[k] = 0.0;
    I.gamma[k] = (px != nullptr) ? px[k] : 0.0;
}

// y = g - g_prev, s = gamma - gamma_prev, ys, yy; low-rank flag.
// Reference: low_rank_fused_kernel, src/sinkhorn_splr_kernel_low_rank.cu:92-149,
// and the rule low_rank = iter > 0 && ys > 1e-6*yy, src/sinkhorn_splr.cu:1394-1395.
__global__ __launch_bounds__(CG_WG) void k_low_rank(const CgInst* __restrict__ insts,
                                                    const int* __restrict__ wg_inst,
                                                    const double* __restrict__ ctl)
{
    int b, wgi;
    if (!cg_decode_vec(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int k = wgi * CG_WG + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (k < I.N) {
        const double yk = I.grad[k] - I.grad_prev[k];
        const double sk = I.gamma[k] - I.gamma_prev[k];
        I.yv[k] = yk;
        I.sv[k] = sk;
        v[0] = yk * sk;
        v[1] = yk * yk;
    }
    double tot[2];
    if (cg_inst_reduce<2>(I, wgi, v, tot, 1) && threadIdx.x == 0) {
        I.scal[CS_YS] = tot[0];
        I.scal[CS_YY] = tot[1];
        const bool allow = ctl[b * CG_NCTL + CTL_ALLOW_LR] != 0.0;
        I.scal[CS_LOWRANK] = (allow && tot[0] > 1e-6 * tot[1]) ? 1.0 : 0.0;
    }
}

// gamma = gamma_bcd.  Reference: update_gamma_sinkhorn, src/sinkhorn_splr.cu:929-934.
__global__ __launch_bounds__(CG_WG) void k_copy_bcd(const CgInst* __restrict__ insts,
                                                    const int* __restrict__ wg_inst,
                                                    const double* __restrict__ ctl)
{
    int b, wgi;
    if (!cg_decode_vec(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int k = wgi * CG_WG + threadIdx.x;
    if (k < I.n + I.m) I.gamma[k] = I.gamma_bcd[k];
}

// Dots for the BFGS Woodbury correction: s.g, y.u, y.dir.
// Reference: search_direc_dot_kernel, src/sinkhorn_splr_kernel_low_rank.cu:209-254.
__global__ __launch_bounds__(CG_WG) void k_woodbury_dots(const CgInst* __restrict__ insts,
                                                         const int* __restrict__ wg_inst,
                                                         const double* __restrict__ ctl)
{
    int b, wgi;
    if (!cg_decode_vec(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    if (I.scal[CS_LOWRANK] == 0.0) return;   // uniform per instance
    const int k = wgi * CG_WG + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (k < I.N) {
        v[0] = I.sv[k] * I.grad[k];
        v[1] = I.yv[k] * I.uv[k];
        v[2] = I.yv[k] * I.dir[k];
    }
    double tot[3];
    if (cg_inst_reduce<3>(I, wgi, v, tot, 2) && threadIdx.x == 0) {
        I.scal[CS_SG] = tot[0];
        I.scal[CS_YAY] = tot[1];
        I.scal[CS_YAG] = tot[2];
    }
}

// dir += t1*s - t2*u (low rank), then dir <- -reg*dir; and the history for
// the next iteration's secant pair: gamma_prev = gamma, grad_prev = grad
// (nothing reads them between here and the next k_low_rank).
// Reference: update_direc_kernel + compute_scaling_inplace, src/sinkhorn_splr_kernel_low_rank.cu:257-276,
// src/sinkhorn_splr.cu:636-638; SPLRSolver::save_history, src/sinkhorn_splr.cu:652-664.
__global__ __launch_bounds__(CG_WG) void k_woodbury_apply(const CgInst* __restrict__ insts,
                                                          const int* __restrict__ wg_inst,
                                                          const double* __restrict__ ctl)
{
    int b, wgi;
    if (!cg_decode_vec(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int k = wgi * CG_WG + threadIdx.x;
    if (k >= I.N) return;
    double d = I.dir[k];
    if (I.scal[CS_LOWRANK] != 0.0) {
        const double ys = I.scal[CS_YS];
        const double sgys = I.scal[CS_SG] / ys;
        const double t1 = (1.0 / I.reg + I.scal[CS_YAY] / ys) * sgys - I.scal[CS_YAG] / ys;
        d += (t1 * I.sv[k] - sgys * I.uv[k]);
    }
    I.dir[k] = -I.reg * d;
    I.gamma_prev[k] = I.gamma[k];
    I.grad_prev[k] = I.grad[k];
}

// dual = gamma (n+m).
__global__ __launch_bounds__(CG_WG) void k_output_dual(const CgInst* __restrict__ insts,
                                                       const int* __restrict__ wg_inst,
                                                       const double* __restrict__ ctl)
{
    int b, wgi;
    if (!cg_decode_vec(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int k = wgi * CG_WG + threadIdx.x;
    if (k < I.n + I.m && I.dual != nullptr) I.dual[k] = I.gamma[k];
}

// Vector copies: what = 0: gamma_prev = gamma (n+m);  1: beta part of gamma ->
// gamma_bcd;  2: inv_diag = 1/(Da, Db) (context preconditioner,
// reference extract_inv_diag_kernel, src/sinkhorn_splr_kernel_low_rank.cu:404-439).
__global__ __launch_bounds__(CG_WG) void k_vcopy(const CgInst* __restrict__ insts,
                                                 const int* __restrict__ wg_inst,
                                                 const double* __restrict__ ctl, int what)
{
    int b, wgi;
    if (!cg_decode_vec(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int k = wgi * CG_WG + threadIdx.x;
    if (what == 0) {
        if (k < I.n + I.m) I.gamma_prev[k] = I.gamma[k];
    } else if (what == 1) {
        if (k < I.m) I.gamma_bcd[I.n + k] = I.gamma[I.n + k];
    } else {
        if (k < I.N) {
            const double d = (k < I.n) ? I.Da[k] : I.Db[k - I.n];
            I.inv_diag[k] = (d > 0.0 && isfinite(d)) ? 1.0 / d : 0.0;
        }
    }
}

// Sums for the gauge-free BCD stopping test (reference compute_iter_difference,
// src/sinkhorn_bcd_kernel.cu:280-326): with d = alpha - alpha_prev and
// e = beta - beta_prev, sum(alpha), sum(beta), sum(alpha_prev), sum(beta_prev),
// sum(d), sum(d^2), sum(e), sum(e^2).
__global__ __launch_bounds__(CG_WG) void k_bcd_diff(const CgInst* __restrict__ insts,
                                                    const int* __restrict__ wg_inst,
                                                    const double* __restrict__ ctl)
{
    int b, wgi;
    if (!cg_decode_vec(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int q = wgi * CG_WG + threadIdx.x;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (q < I.n) {
        const double a = I.gamma[q], ap = I.gamma_prev[q], d = a - ap;
        v[0] = a; v[2] = ap; v[4] = d; v[5] = d * d;
    }
    if (q < I.m) {
        const double c = I.gamma[I.n + q], cp = I.gamma_prev[I.n + q], e = c - cp;
        v[1] = c; v[3] = cp; v[6] = e; v[7] = e * e;
    }
    double tot[8];
    if (cg_inst_reduce<8>(I, wgi, v, tot, 3) && threadIdx.x == 0)
        for (int k = 0; k < 8; k++) I.scal[BCD_SUMS + k] = tot[k];
}
