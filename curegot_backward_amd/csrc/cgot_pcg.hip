// Implicit-function adjoint: Q(T*) lambda = mu solved by preconditioned CG
// entirely on the device, T recomputed from the converged duals in every
// matrix-vector product (one pass over M per iteration, no stored plan).
//
// Reference: pcg / q_matvec / rhs_Q_scale, curegot/_backward_core.py:31-83 and
// solve_backward, curegot/backward.py:72-144 (a Python loop of torch ops with
// two host syncs per iteration there).
//
// Vector roles inside CgInst for a backward instance:
//   dir = x (lambda), yv = r, sv = z, uv = p, grad = w = Q p,
//   rowsum/colsum = T 1 / T^T 1, Da = inverse diagonal preconditioner (or null).
// Per-iteration launches: k_qv_pcg -> k_pcg_q_fin -> k_pcg_update.
// Every kernel skips an instance whose status is not "running", so the host can
// launch iterations in chunks and read the statuses once per chunk; the
// iterates are exactly those of the sequential loop.
#include "cgot_common.h"
#include "cgot_kernels.h"

CG_DEV bool cg_pcg_decode(const CgInst* insts, const int* wg_inst, const double* ctl, int& b, int& wgi)
{
    const int2 e = cg_map_entry(wg_inst);
    b = e.x;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return false;
    if (insts[b].scal[PCG_STATUS] != 0.0) return false;
    wgi = e.y;
    return true;
}

// mu, T sums, r = mu, x = 0, z = M^{-1} r, p = 0, rz, ||mu||.
__global__ __launch_bounds__(CG_WG) void k_pcg_init(const CgInst* __restrict__ insts,
                                                    const int* __restrict__ wg_inst,
                                                    const double* __restrict__ ctl)
{
    const int2 e = cg_map_entry(wg_inst);
    const int b = e.x;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return;
    const CgInstD& I = cg_inst(insts, b);
    const int wgi = e.y;
    const int q = wgi * CG_WG + threadIdx.x;
    const int n = I.n, m = I.m;
    double v[3] = {0.0, 0.0, 0.0};   // ||mu||^2, r.z, any(mu != 0)
    auto setk = [&](int k, double mu) {
        I.yv[k] = mu;
        I.dir[k] = 0.0;
        I.uv[k] = 0.0;
        const double z = (I.Da != nullptr) ? I.Da[k] * mu : mu;   // SPLR preconditioner: k_pc_* then k_pcg_rz
        I.sv[k] = z;
        v[0] += mu * mu;
        v[1] += mu * z;
        if (mu != 0.0) v[2] = 1.0;
    };
    if (q < n) {
        double a, t;
        cg_ordered_sum2(I.rp0 + (size_t)q * I.nstrips, I.rp1 + (size_t)q * I.nstrips, I.nstrips, 1, a, t);
        I.rowsum[q] = t;
        setk(q, a);
    }
    if (q < m) {
        double a, t;
        cg_ordered_sum2(I.cp0 + q, I.cp1 + q, I.nrb, m, a, t);
        I.colsum[q] = t;
        if (q < m - 1) setk(n + q, a);
    }
    double tot[3];
    if (cg_inst_reduce<3>(I, wgi, v, tot, 0) && threadIdx.x == 0) {
        I.scal[PCG_BNORM] = sqrt(tot[0]);
        I.scal[PCG_RZ] = tot[1];
        I.scal[PCG_BETA] = 0.0;
        I.scal[PCG_ITER] = 0.0;
        I.scal[PCG_STATUS] = (tot[2] != 0.0) ? 0.0 : PCG_ZERO_RHS;
    }
}

// w = Q p with p = z + beta*p_prev (the M pass already holds sum_j T_ij p_beta_j
// and sum_i T_ij p_alpha_i); store p, pw, alpha.
__global__ __launch_bounds__(CG_WG) void k_pcg_q_fin(const CgInst* __restrict__ insts,
                                                     const int* __restrict__ wg_inst,
                                                     const double* __restrict__ ctl)
{
    int b, wgi;
    if (!cg_pcg_decode(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int q = wgi * CG_WG + threadIdx.x;
    const int n = I.n, m = I.m;
    const double beta = I.scal[PCG_BETA];
    double v[1] = {0.0};
    if (q < n) {
        const double a = cg_ordered_sum(I.rp0 + (size_t)q * I.nstrips, I.nstrips, 1);
        const double p = I.sv[q] + beta * I.uv[q];
        const double w = I.rowsum[q] * p + a;
        I.uv[q] = p;
        I.grad[q] = w;
        v[0] += p * w;
    }
    if (q < m - 1) {
        const double a = cg_ordered_sum<32>(I.cp0 + q, I.nrb, m);
        const int k = n + q;
        const double p = I.sv[k] + beta * I.uv[k];
        const double w = a + I.colsum[q] * p;
        I.uv[k] = p;
        I.grad[k] = w;
        v[0] += p * w;
    }
    double tot[1];
    if (cg_inst_reduce<1>(I, wgi, v, tot, 1) && threadIdx.x == 0) {
        if (tot[0] <= 0.0) {
            I.scal[PCG_STATUS] = PCG_BREAKDOWN;
        } else {
            I.scal[PCG_ALPHA] = I.scal[PCG_RZ] / tot[0];
        }
    }
}

// x += alpha p, r -= alpha w, convergence test, z = M^{-1} r, beta.
__global__ __launch_bounds__(CG_WG) void k_pcg_update(const CgInst* __restrict__ insts,
                                                      const int* __restrict__ wg_inst,
                                                      const double* __restrict__ ctl)
{
    int b, wgi;
    if (!cg_pcg_decode(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int k = wgi * CG_WG + threadIdx.x;
    const double alpha = I.scal[PCG_ALPHA];
    double v[2] = {0.0, 0.0};
    if (k < I.N) {
        I.dir[k] += alpha * I.uv[k];
        const double r = I.yv[k] - alpha * I.grad[k];
        I.yv[k] = r;
        v[0] = r * r;
        if (I.fctx == 0) {   // with the SPLR preconditioner z and r.z come later (k_pc_*, k_pcg_rz)
            const double z = (I.Da != nullptr) ? I.Da[k] * r : r;
            I.sv[k] = z;
            v[1] = r * z;
        }
    }
    double tot[2];
    if (cg_inst_reduce<2>(I, wgi, v, tot, 2) && threadIdx.x == 0) {
        const double it = I.scal[PCG_ITER] + 1.0;
        I.scal[PCG_ITER] = it;
        const double tol = ctl[b * CG_NCTL + CTL_STEP];
        const double maxit = ctl[b * CG_NCTL + CTL_AUX];
        if (sqrt(tot[0]) < tol * I.scal[PCG_BNORM]) {
            I.scal[PCG_STATUS] = PCG_CONVERGED;
        } else {
            if (I.fctx == 0) {
                I.scal[PCG_BETA] = tot[1] / I.scal[PCG_RZ];
                I.scal[PCG_RZ] = tot[1];
            }
            if (it >= maxit) I.scal[PCG_STATUS] = PCG_MAXIT;
        }
    }
}

// ---------------------------------------------------------------------------
// SPLR preconditioner, low-rank part and the PCG scalars that follow it.
// With q = A^{-1} r in z (from k_pc_back), u = A^{-1} y and the secant pair
// (s, y) of the forward's last iteration:
//   B^{-1} r = q - (s^T r / ys) u + [(1/reg + y^T u / ys)(s^T r / ys) - y^T q / ys] s
// Reference: SPLRBackwardContext::apply_preconditioner, src/backward_context.cu:543-581,
// with launch_low_rank_search_direc, src/sinkhorn_splr_kernel_low_rank.cu:296-390.
// ---------------------------------------------------------------------------
CG_DEV const CgInstD& cg_pc_fwd(const CgInstD& I) { return *reinterpret_cast<const CgInstD*>(I.fctx); }

CG_DEV bool cg_pc_decode(const CgInst* insts, const int* wg_inst, const double* ctl, int& b, int& wgi)
{
    if (!cg_pcg_decode(insts, wg_inst, ctl, b, wgi)) return false;
    return insts[b].fctx != 0;
}

__global__ __launch_bounds__(CG_WG) void k_pc_wdots(const CgInst* __restrict__ insts,
                                                    const int* __restrict__ wg_inst,
                                                    const double* __restrict__ ctl)
{
    int b, wgi;
    if (!cg_pc_decode(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    const CgInstD& F = cg_pc_fwd(I);
    if (F.scal[CS_LOWRANK] == 0.0) return;   // uniform per instance
    const int k = wgi * CG_WG + threadIdx.x;
    double v[3] = {0.0, 0.0, 0.0};
    if (k < I.N) {
        v[0] = F.sv[k] * I.yv[k];
        v[1] = F.yv[k] * F.uv[k];
        v[2] = F.yv[k] * I.sv[k];
    }
    double tot[3];
    if (cg_inst_reduce<3>(I, wgi, v, tot, 3) && threadIdx.x == 0) {
        I.scal[CS_SG] = tot[0];
        I.scal[CS_YAY] = tot[1];
        I.scal[CS_YAG] = tot[2];
    }
}

__global__ __launch_bounds__(CG_WG) void k_pc_wapply(const CgInst* __restrict__ insts,
                                                     const int* __restrict__ wg_inst,
                                                     const double* __restrict__ ctl)
{
    int b, wgi;
    if (!cg_pc_decode(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    const CgInstD& F = cg_pc_fwd(I);
    if (F.scal[CS_LOWRANK] == 0.0) return;
    const int k = wgi * CG_WG + threadIdx.x;
    if (k >= I.N) return;
    const double ys = F.scal[CS_YS];
    const double srys = I.scal[CS_SG] / ys;
    const double t1 = (1.0 / F.reg + I.scal[CS_YAY] / ys) * srys - I.scal[CS_YAG] / ys;
    I.sv[k] += t1 * F.sv[k] - srys * F.uv[k];
}

// r.z after the SPLR preconditioner; init: rz, beta = 0; else beta = rz_new / rz.
__global__ __launch_bounds__(CG_WG) void k_pcg_rz(const CgInst* __restrict__ insts,
                                                  const int* __restrict__ wg_inst,
                                                  const double* __restrict__ ctl, int init)
{
    int b, wgi;
    if (!cg_pc_decode(insts, wg_inst, ctl, b, wgi)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int k = wgi * CG_WG + threadIdx.x;
    double v[1] = {(k < I.N) ? I.yv[k] * I.sv[k] : 0.0};
    double tot[1];
    if (cg_inst_reduce<1>(I, wgi, v, tot, 0) && threadIdx.x == 0) {
        if (init) {
            I.scal[PCG_BETA] = 0.0;
        } else {
            I.scal[PCG_BETA] = tot[0] / I.scal[PCG_RZ];
        }
        I.scal[PCG_RZ] = tot[0];
    }
}
