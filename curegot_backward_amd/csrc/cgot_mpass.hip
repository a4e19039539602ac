// M-sized passes: every kernel here streams the cost matrices of a ragged
// batch once, row-major and coalesced, and writes fixed-order partials that a
// finalize kernel combines.  These are the HBM-bound kernels of the solver.
//
// Tile geometry (per instance b): strips of sw = 64*c columns (c columns per
// lane, lane l owns columns s0 + l + 64k) by row blocks of CG_RB rows (each of
// the 4 waves owns CG_RPW consecutive rows).  Each tile writes
//   row partials  rp[i*nstrips + strip]   (one per row and strip)
//   col partials  cp[rowblock*m + j]      (one per row block and column)
// so the finalize step sums a fixed number of partials in a fixed order.
// The T pass and the log-sum-exp kernels are held to 64 VGPRs (8 waves per
// SIMD): these streams are latency-bound, and the waves in flight hide the
// HBM latency.
#include "cgot_common.h"
#include "cgot_kernels.h"

// Point (alpha, beta) the pass is evaluated at.
//   CG_PT_GAMMA : gamma
//   CG_PT_TRIAL : gamma_prev + step*dir on the first N entries, beta_{m-1} = 0
//   CG_PT_BCD   : gamma_bcd
CG_DEV double cg_point(const CgInstD& I, int mode, double t, int k)
{
    if (mode == CG_PT_GAMMA) return I.gamma[k];
    if (mode == CG_PT_BCD) return I.gamma_bcd[k];
    return (k < I.N) ? (I.gamma_prev[k] + t * I.dir[k]) : 0.0;
}

CG_DEV bool cg_decode_tile(const CgInst* insts, const int* tile_inst, const double* ctl,
                           int& b, int& rb, int& st)
{
    const int2 e = cg_map_entry(tile_inst);
    b = e.x;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return false;
    const CgInstD& I = cg_inst(insts, b);
    const int lt = e.y;
    rb = lt / I.nstrips;
    st = lt - rb * I.nstrips;
    return true;
}

// Reduction (max or sum) of R rows held as one partial per (row, lane) in
// v[r]: the partials go to this wave's LDS block red ([R][64]); lane l then
// takes row q = l / G (G = 64 / R lanes per row, gi = l % G), combines the
// row's partials gi, gi + G, gi + 2G, ... in that order, and the G lanes of
// the row finish with a DPP tree inside their 16-lane row (G = 8 or 16) and,
// for G = 32, one exchange with lane ^ 16.  Every lane of row q's group ends
// with the same value.
template <int R, bool MAX>
CG_DEV double cg_rows_reduce(const double (&v)[R], double* red)
{
    static_assert(R == 2 || R == 4 || R == 8, "32, 16 or 8 lanes per row");
    constexpr int G = 64 / R;
    const int lane = threadIdx.x & 63, q = lane / G, gi = lane % G;
#pragma unroll
    for (int r = 0; r < R; r++) red[r * 64 + lane] = v[r];
    cg_wave_sync();
    double a = red[q * 64 + gi];
#pragma unroll
    for (int j = 1; j < R; j++) {
        const double b = red[q * 64 + gi + G * j];
        a = MAX ? fmax(a, b) : a + b;
    }
    a = MAX ? fmax(a, cg_dpp_d<0xB1>(a)) : a + cg_dpp_d<0xB1>(a);     // lane ^ 1
    a = MAX ? fmax(a, cg_dpp_d<0x4E>(a)) : a + cg_dpp_d<0x4E>(a);     // lane ^ 2
    a = MAX ? fmax(a, cg_dpp_d<0x141>(a)) : a + cg_dpp_d<0x141>(a);   // half-row mirror: 8 lanes
    if (G >= 16) a = MAX ? fmax(a, cg_dpp_d<0x140>(a)) : a + cg_dpp_d<0x140>(a);   // row mirror: 16 lanes
    if (G == 32) {
        const double o = __shfl_xor(a, 16);
        a = MAX ? fmax(a, o) : a + o;
    }
    return a;
}

// ---------------------------------------------------------------------------
// T pass: T_ij = exp((alpha_i + beta_j - M_ij)/reg); row and column partial sums.
// Reference: T_fused_kernel / line_search_fused_kernel2,
// src/sinkhorn_splr_kernel.cu:214-338 (float atomics there; fixed-order here).
// Each wave takes its CG_RPW rows in groups of R and issues the loads of a
// whole group before the exponentials, so R*C loads per lane are in flight.
// ---------------------------------------------------------------------------
template <int C, bool FULL, bool XY>
CG_DEV void cg_tpass_tile(const CgInstD& I, int mode, double t, int rb, int st, double (*lds)[CG_CMAX * 64],
                          const double* tab)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s0 = st * I.sw;
    const int m = I.m, n = I.n;
    double bj[C], cs[C];
    bool ok[C];
#pragma unroll
    for (int k = 0; k < C; k++) {
        const int j = s0 + lane + 64 * k;
        ok[k] = FULL || j < m;
        bj[k] = ok[k] ? cg_point(I, mode, t, n + j) : 0.0;
        cs[k] = 0.0;
    }
    const int i0 = rb * CG_RB + w * CG_RPW;
    const double ir = I.inv_reg;
    // rows per load group: R*C loads per lane in flight; the group loop stays
    // rolled so that only one group's loads hold registers
    constexpr int R = (C >= 4) ? 2 : 4;
    constexpr int G = 64 / R;
    const int q = lane / G, gi = lane % G;
    // the row sums of a group are reduced together through lds[w], which the
    // column combine only uses after the loop (R*64 <= CG_CMAX*64)
    static_assert(R * 64 <= CG_CMAX * 64, "group partials fit the wave's column buffer");
#pragma unroll 1
    for (int r0 = 0; r0 < CG_RPW; r0 += R) {
        double mv[R][C], ai[R], part[R];
        if constexpr (XY) cg_cost_rows_xy<R, C>(I, i0 + r0, n, s0, lane, ok, mv);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = min(i0 + r0 + r, n - 1);   // clamped rows are loaded but not used
            ai[r] = cg_point(I, mode, t, i);
            if constexpr (!XY) cg_cost_row<C>(I, i, s0, lane, ok, mv[r]);
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = i0 + r0 + r;
            double rs = 0.0;
            // columns past m (mv = bj = 0 there) take a finite argument and
            // are zeroed after the exp, so no lane branches
            double v[C];
#pragma unroll
            for (int k = 0; k < C; k++) v[k] = cg_texp((ai[r] + bj[k] - mv[r][k]) * ir, tab);
#pragma unroll
            for (int k = 0; k < C; k++) {
                v[k] = ok[k] ? v[k] : 0.0;
                rs += v[k];
            }
            if (i < n) {
#pragma unroll
                for (int k = 0; k < C; k++) cs[k] += v[k];
            }
            part[r] = rs;
        }
        const double rowsum = cg_rows_reduce<R, false>(part, lds[w]);
        const int i = i0 + r0 + q;
        if (gi == 0 && i < n) I.rp0[(size_t)i * I.nstrips + st] = rowsum;
    }
    cg_wave_sync();   // the group reductions' reads of lds[w] before the column partials overwrite it
#pragma unroll
    for (int k = 0; k < C; k++) lds[w][lane + 64 * k] = cs[k];
    __syncthreads();
    for (int q = threadIdx.x; q < I.sw; q += CG_WG) {
        const int j = s0 + q;
        if (j < m) I.cp0[(size_t)rb * m + j] = ((lds[0][q] + lds[1][q]) + lds[2][q]) + lds[3][q];
    }
}

template <bool XY>
__global__ __launch_bounds__(CG_WG) __attribute__((amdgpu_waves_per_eu(XY ? 4 : 8))) void k_tpass(const CgInst* __restrict__ insts,
                                                 const int* __restrict__ tile_inst,
                                                 const double* __restrict__ ctl, int mode)
{
    CG_EXP_TABLE(tab);
    int b, rb, st;
    if (!cg_decode_tile(insts, tile_inst, ctl, b, rb, st)) return;
    const CgInstD& I = cg_inst(insts, b);
    const double t = ctl[b * CG_NCTL + CTL_STEP];
    __shared__ double lds[CG_WAVES][CG_CMAX * 64];   // one array for every instantiation
    cg_strip_dispatch<XY>(I, st, I.m, [&](auto tag) {
        using S = decltype(tag);
        cg_tpass_tile<S::C, S::FULL, S::XY>(I, mode, t, rb, st, lds, tab);
    });
}

// ---------------------------------------------------------------------------
// Row log-sum-exp partials of (beta_j - M_ij)/reg for the alpha update.
// Reference: optimal_alpha_kernel, src/sinkhorn_bcd_kernel.cu:115-192.
// Per (row, strip): max over the strip, then sum of exp(x - max).  A wave
// takes its CG_RPW rows in groups of R with the loads of a group issued
// first, and reduces the R rows together through LDS (cg_rows_reduce) instead
// of one 64-lane DPP/readlane reduction per row and quantity.
// ---------------------------------------------------------------------------

template <int C, bool FULL, bool XY>
CG_DEV void cg_rowlse_tile(const CgInstD& I, int mode, int rb, int st, const double* tab, double* red,
                           double* rmx)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s0 = st * I.sw;
    const int m = I.m, n = I.n;
    const double ir = I.inv_reg;
    const auto beta = (mode == CG_PT_BCD) ? (I.gamma_bcd + n) : (I.gamma + n);
    double bj[C];
    bool ok[C];
#pragma unroll
    for (int k = 0; k < C; k++) {
        const int j = s0 + lane + 64 * k;
        ok[k] = FULL || j < m;
        bj[k] = ok[k] ? beta[j] * ir : 0.0;
    }
    const int i0 = rb * CG_RB + w * CG_RPW;
    constexpr int R = 8 / C;   // 8 loads per lane in flight per group of R rows
    constexpr int G = 64 / R;
    const int q = lane / G, gi = lane % G;
#pragma unroll
    for (int r0 = 0; r0 < CG_RPW; r0 += R) {
        double mv[R][C];
        if constexpr (XY) cg_cost_rows_xy<R, C>(I, i0 + r0, n, s0, lane, ok, mv);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = min(i0 + r0 + r, n - 1);   // clamped rows are loaded but not used
            if constexpr (!XY) cg_cost_row<C>(I, i, s0, lane, ok, mv[r]);
        }
        // x = (beta_j - M_ij)/reg, -inf past m; formed again for the sums
        // instead of being held
        auto xat = [&](int r, int k) { return ok[k] ? fma(-mv[r][k], ir, bj[k]) : -INFINITY; };
        double part[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            double mx = xat(r, 0);
#pragma unroll
            for (int k = 1; k < C; k++) mx = fmax(mx, xat(r, k));
            part[r] = mx;
        }
        const double rowmax = cg_rows_reduce<R, true>(part, red);   // finite: column s0 is in the strip
        if (gi == 0) rmx[q] = rowmax;
        cg_wave_sync();
        // exp(-inf) past m is 0 or the smallest subnormal (cg_texp_lo), far
        // below the rounding of a sum that contains exp(0) = 1
#pragma unroll
        for (int r = 0; r < R; r++) {
            const double mr = rmx[r];
            double sm = cg_texp_lo(xat(r, 0) - mr, tab);
#pragma unroll
            for (int k = 1; k < C; k++) sm += cg_texp_lo(xat(r, k) - mr, tab);
            part[r] = sm;
        }
        const double rowsum = cg_rows_reduce<R, false>(part, red);
        const int i = i0 + r0 + q;
        if (gi == 0 && i < n) {
            I.rp0[(size_t)i * I.nstrips + st] = rowmax;
            I.rp1[(size_t)i * I.nstrips + st] = rowsum;
        }
    }
}

template <bool XY>
__global__ __launch_bounds__(CG_WG) __attribute__((amdgpu_waves_per_eu(XY ? 4 : 8))) void k_row_lse(const CgInst* __restrict__ insts,
                                                   const int* __restrict__ tile_inst,
                                                   const double* __restrict__ ctl, int mode)
{
    CG_EXP_TABLE(tab);
    int b, rb, st;
    if (!cg_decode_tile(insts, tile_inst, ctl, b, rb, st)) return;
    const CgInstD& I = cg_inst(insts, b);
    __shared__ double red[CG_WAVES][8 * 64];   // row partials of a group, per wave
    __shared__ double rmx[CG_WAVES][8];        // row maxima of a group, per wave
    const int w = threadIdx.x >> 6;
    cg_strip_dispatch<XY>(I, st, I.m, [&](auto tag) {
        using S = decltype(tag);
        cg_rowlse_tile<S::C, S::FULL, S::XY>(I, mode, rb, st, tab, red[w], rmx[w]);
    });
}

// ---------------------------------------------------------------------------
// Column log-sum-exp partials of (alpha_i - M_ij)/reg for the beta update.
// Reference: optimal_beta_kernel, src/sinkhorn_bcd_kernel.cu:35-112 (strided
// column walk there; here each lane streams its columns down the row block
// with an online max/sum, reads stay row-major and coalesced, R rows' loads
// in flight at a time).
// ---------------------------------------------------------------------------
template <int C, bool FULL, bool XY>
CG_DEV void cg_collse_tile(const CgInstD& I, int mode, int rb, int st, const double* tab,
                           double (*lmx)[CG_CMAX * 64], double (*lsm)[CG_CMAX * 64])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s0 = st * I.sw;
    const int m = I.m, n = I.n;
    const double ir = I.inv_reg;
    const auto alpha = (mode == CG_PT_BCD) ? I.gamma_bcd : I.gamma;
    double mx[C], sm[C];
    bool ok[C];
#pragma unroll
    for (int k = 0; k < C; k++) {
        mx[k] = -INFINITY;
        sm[k] = 0.0;
        ok[k] = FULL || s0 + lane + 64 * k < m;
    }
    const int i0 = rb * CG_RB + w * CG_RPW;
    constexpr int R = (C >= 4) ? 2 : 4;
#pragma unroll
    for (int r0 = 0; r0 < CG_RPW; r0 += R) {
        double mv[R][C], ai[R];
        if constexpr (XY) cg_cost_rows_xy<R, C>(I, i0 + r0, n, s0, lane, ok, mv);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = min(i0 + r0 + r, n - 1);
            ai[r] = alpha[i] * ir;
            if constexpr (!XY) cg_cost_row<C>(I, i, s0, lane, ok, mv[r]);
        }
        // online update without lane branches: e = exp(-|x - mx|) serves
        // both cases (a new maximum rescales the sum by e and adds 1, any
        // other x adds e); the first row meets mx = -inf, gets e = 0 and
        // sum 1.  Columns past m compute on mv = 0 and are never stored.
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (i0 + r0 + r >= n) break;
#pragma unroll
            for (int k = 0; k < C; k++) {
                const double x = fma(-mv[r][k], ir, ai[r]);
                const double d = x - mx[k];
                const double e = cg_texp_lo(-fabs(d), tab);
                sm[k] = (d > 0.0) ? fma(sm[k], e, 1.0) : sm[k] + e;
                mx[k] = fmax(mx[k], x);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < C; k++) { lmx[w][lane + 64 * k] = mx[k]; lsm[w][lane + 64 * k] = sm[k]; }
    __syncthreads();
    for (int q = threadIdx.x; q < I.sw; q += CG_WG) {
        const int j = s0 + q;
        if (j < m) {
            double a = lmx[0][q], s = lsm[0][q];
            cg_lse_merge(a, s, lmx[1][q], lsm[1][q]);
            cg_lse_merge(a, s, lmx[2][q], lsm[2][q]);
            cg_lse_merge(a, s, lmx[3][q], lsm[3][q]);
            I.cp0[(size_t)rb * m + j] = a;
            I.cp1[(size_t)rb * m + j] = s;
        }
    }
}

template <bool XY>
__global__ __launch_bounds__(CG_WG) __attribute__((amdgpu_waves_per_eu(XY ? 4 : 8))) void k_col_lse(const CgInst* __restrict__ insts,
                                                   const int* __restrict__ tile_inst,
                                                   const double* __restrict__ ctl, int mode)
{
    CG_EXP_TABLE(tab);
    int b, rb, st;
    if (!cg_decode_tile(insts, tile_inst, ctl, b, rb, st)) return;
    const CgInstD& I = cg_inst(insts, b);
    __shared__ double lmx[CG_WAVES][CG_CMAX * 64];
    __shared__ double lsm[CG_WAVES][CG_CMAX * 64];
    cg_strip_dispatch<XY>(I, st, I.m, [&](auto tag) {
        using S = decltype(tag);
        cg_collse_tile<S::C, S::FULL, S::XY>(I, mode, rb, st, tab, lmx, lsm);
    });
}

// ---------------------------------------------------------------------------
// Adjoint passes with T recomputed from the duals (gamma):
//   k_qv_pcg: row partial sum_j T_ij v_beta_j (j < m-1), col partial
//             sum_i T_ij v_alpha_i, with v = z + beta*p_prev (the PCG
//             direction) formed on the fly
//   k_qv_rhs: rp0/cp0 = sums of G_ij T_ij, rp1/cp1 = sums of T_ij
// Reference: q_matvec / rhs_Q_scale, curegot/_backward_core.py:31-45 (3-4 torch
// passes over a stored T there, one pass over M here).
// ---------------------------------------------------------------------------
template <int C, bool FULL, bool XY, bool RHS>
CG_DEV void cg_qv_tile(const CgInstD& I, int b, int rb, int st, const double* const* Gs, double (*lds)[CG_CMAX * 64],
                       const double* tab)
{
    constexpr bool MV = !RHS;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s0 = st * I.sw;
    const int m = I.m, n = I.n;
    const double ir = I.inv_reg;
    const double pbeta = MV ? I.scal[PCG_BETA] : 0.0;
    const CG_AS1 double* G = RHS ? (const CG_AS1 double*)Gs[b] : nullptr;
    auto vat = [&](int k) -> double { return I.sv[k] + pbeta * I.uv[k]; };
    double bj[C], vb[C], cs[C], ct[C];
    bool ok[C];
#pragma unroll
    for (int k = 0; k < C; k++) {
        const int j = s0 + lane + 64 * k;
        ok[k] = FULL || j < m;
        bj[k] = ok[k] ? I.gamma[n + j] : 0.0;
        vb[k] = (MV && ok[k] && j < m - 1) ? vat(n + j) : 0.0;   // beta_{m-1} is fixed
        cs[k] = 0.0;
        ct[k] = 0.0;
    }
    const int i0 = rb * CG_RB + w * CG_RPW;
    // R rows' loads in flight (RHS loads G beside M); the group loop stays
    // rolled (one group's loads hold registers at a time)
    constexpr int R = (C >= 4) ? 1 : 2;
#pragma unroll 1
    for (int r0 = 0; r0 < CG_RPW; r0 += R) {
        double mv[R][C], gv[RHS ? R : 1][C], ai[R], va[R];
        if constexpr (XY) cg_cost_rows_xy<R, C>(I, i0 + r0, n, s0, lane, ok, mv);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = min(i0 + r0 + r, n - 1);   // clamped rows are loaded but not used
            ai[r] = I.gamma[i];
            va[r] = MV ? vat(i) : 0.0;
            if constexpr (!XY) cg_cost_row<C, false>(I, i, s0, lane, ok, mv[r]);
            if (RHS) {
                const auto Gi = G + (size_t)i * m + s0 + lane;
#pragma unroll
                for (int k = 0; k < C; k++) gv[RHS ? r : 0][k] = ok[k] ? Gi[64 * k] : 0.0;
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = i0 + r0 + r;
            const bool live = i < n;
            // columns past m take a finite argument and are zeroed after the exp
            double tv[C];
#pragma unroll
            for (int k = 0; k < C; k++) tv[k] = cg_texp((ai[r] + bj[k] - mv[r][k]) * ir, tab);
#pragma unroll
            for (int k = 0; k < C; k++) tv[k] = ok[k] ? tv[k] : 0.0;
            double rs = 0.0, rt = 0.0;
#pragma unroll
            for (int k = 0; k < C; k++) {
                if (MV) {
                    rs += tv[k] * vb[k];
                    if (live) cs[k] += tv[k] * va[r];
                } else {
                    const double gt = gv[RHS ? r : 0][k] * tv[k];
                    rs += gt;
                    rt += tv[k];
                    if (live) {
                        cs[k] += gt;
                        ct[k] += tv[k];
                    }
                }
            }
            rs = cg_wave_sum(rs);
            if (RHS) rt = cg_wave_sum(rt);
            if (lane == 0 && live) {
                I.rp0[(size_t)i * I.nstrips + st] = rs;
                if (RHS) I.rp1[(size_t)i * I.nstrips + st] = rt;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < C; k++) lds[w][lane + 64 * k] = cs[k];
    __syncthreads();
    for (int q = threadIdx.x; q < I.sw; q += CG_WG) {
        const int j = s0 + q;
        if (j < m) I.cp0[(size_t)rb * m + j] = ((lds[0][q] + lds[1][q]) + lds[2][q]) + lds[3][q];
    }
    if (RHS) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < C; k++) lds[w][lane + 64 * k] = ct[k];
        __syncthreads();
        for (int q = threadIdx.x; q < I.sw; q += CG_WG) {
            const int j = s0 + q;
            if (j < m) I.cp1[(size_t)rb * m + j] = ((lds[0][q] + lds[1][q]) + lds[2][q]) + lds[3][q];
        }
    }
}

template <bool XY, bool RHS>
CG_DEV void cg_qv_strip(const CgInstD& I, int b, int rb, int st, const double* const* Gs, double (*lds)[CG_CMAX * 64],
                        const double* tab)
{
    cg_strip_dispatch<XY>(I, st, I.m, [&](auto tag) {
        using S = decltype(tag);
        cg_qv_tile<S::C, S::FULL, S::XY, RHS>(I, b, rb, st, Gs, lds, tab);
    });
}

// once per adjoint solve (loads G beside M; no register cap)
template <bool XY>
__global__ __launch_bounds__(CG_WG) void k_qv_rhs(const CgInst* __restrict__ insts, const int* __restrict__ tile_inst,
                                                  const double* __restrict__ ctl, const double* const* __restrict__ Gs)
{
    CG_EXP_TABLE(tab);
    int b, rb, st;
    if (!cg_decode_tile(insts, tile_inst, ctl, b, rb, st)) return;
    const CgInstD& I = cg_inst(insts, b);
    __shared__ double lds[CG_WAVES][CG_CMAX * 64];
    cg_qv_strip<XY, true>(I, b, rb, st, Gs, lds, tab);
}

// every PCG iteration; only instances whose PCG is running
template <bool XY>
__global__ __launch_bounds__(CG_WG) __attribute__((amdgpu_waves_per_eu(XY ? 4 : 8))) void k_qv_pcg(
    const CgInst* __restrict__ insts, const int* __restrict__ tile_inst, const double* __restrict__ ctl)
{
    CG_EXP_TABLE(tab);
    int b, rb, st;
    if (!cg_decode_tile(insts, tile_inst, ctl, b, rb, st)) return;
    const CgInstD& I = cg_inst(insts, b);
    if (I.scal[PCG_STATUS] != 0.0) return;
    __shared__ double lds[CG_WAVES][CG_CMAX * 64];
    cg_qv_strip<XY, false>(I, b, rb, st, nullptr, lds, tab);
}

// ---------------------------------------------------------------------------
// Elementwise writers
// ---------------------------------------------------------------------------
// P_ij = exp((alpha_i + beta_j - M_ij)/reg).  Reference: compute_transport_plan_kernel,
// src/sinkhorn_bcd_kernel.cu:383-403.
__global__ __launch_bounds__(CG_WG) void k_plan(const CgInst* __restrict__ insts,
                                                const int* __restrict__ tile_inst,
                                                const double* __restrict__ ctl)
{
    CG_EXP_TABLE(tab);
    int b, rb, st;
    if (!cg_decode_tile(insts, tile_inst, ctl, b, rb, st)) return;
    const CgInstD I = cg_inst(insts, b);
    if (I.plan == nullptr) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s0 = st * I.sw;
    const int m = I.m, n = I.n;
    const auto dual = I.gamma;
    double bj[CG_CMAX];
#pragma unroll
    for (int k = 0; k < CG_CMAX; k++) {
        const int j = s0 + lane + 64 * k;
        bj[k] = (k < I.c && j < m) ? dual[n + j] : 0.0;
    }
    const int i0 = rb * CG_RB + w * CG_RPW;
    for (int r = 0; r < CG_RPW; r++) {
        const int i = i0 + r;
        if (i >= n) break;
        const double ai = dual[i];
        auto Pi = I.plan + (size_t)i * m;
#pragma unroll
        for (int k = 0; k < CG_CMAX; k++) {
            const int j = s0 + lane + 64 * k;
            if (k < I.c && j < m) Pi[j] = cg_texp((ai + bj[k] - cg_cost(I, i, j)) * I.inv_reg, tab);
        }
    }
}

// dL/dM_ij = T_ij (lam_alpha_i + lam_beta_j - G_ij)/reg, lam_beta_{m-1} = 0.
// Reference: assemble_grad, curegot/_backward_core.py:47-53.
__global__ __launch_bounds__(CG_WG) void k_grad_assemble(const CgInst* __restrict__ insts,
                                                         const int* __restrict__ tile_inst,
                                                         const double* __restrict__ ctl,
                                                         const double* const* __restrict__ lams,
                                                         const double* const* __restrict__ Gs,
                                                         double* const* __restrict__ outs)
{
    CG_EXP_TABLE(tab);
    int b, rb, st;
    if (!cg_decode_tile(insts, tile_inst, ctl, b, rb, st)) return;
    const CgInstD I = cg_inst(insts, b);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s0 = st * I.sw;
    const int m = I.m, n = I.n;
    const CG_AS1 double* lam = (const CG_AS1 double*)lams[b];
    const CG_AS1 double* G = (const CG_AS1 double*)Gs[b];
    CG_AS1 double* out = (CG_AS1 double*)outs[b];
    double bj[CG_CMAX], lb[CG_CMAX];
#pragma unroll
    for (int k = 0; k < CG_CMAX; k++) {
        const int j = s0 + lane + 64 * k;
        const bool ok = (k < I.c && j < m);
        bj[k] = ok ? I.gamma[n + j] : 0.0;
        lb[k] = (ok && j < m - 1) ? lam[n + j] : 0.0;
    }
    const int i0 = rb * CG_RB + w * CG_RPW;
    for (int r = 0; r < CG_RPW; r++) {
        const int i = i0 + r;
        if (i >= n) break;
        const double ai = I.gamma[i];
        const double la = lam[i];
        const size_t row = (size_t)i * m;
#pragma unroll
        for (int k = 0; k < CG_CMAX; k++) {
            const int j = s0 + lane + 64 * k;
            if (k < I.c && j < m) {
                const double tv = cg_texp((ai + bj[k] - cg_cost(I, i, j)) * I.inv_reg, tab);
                out[row + j] = (tv * (la + lb[k] - G[row + j])) / I.reg;
            }
        }
    }
}

// the stored-cost and the geometric-cost instantiations of the M passes
#define CG_MPASS_INSTANTIATE(XY)                                                                        \
    template __global__ void k_tpass<XY>(const CgInst*, const int*, const double*, int);               \
    template __global__ void k_row_lse<XY>(const CgInst*, const int*, const double*, int);             \
    template __global__ void k_col_lse<XY>(const CgInst*, const int*, const double*, int);             \
    template __global__ void k_qv_rhs<XY>(const CgInst*, const int*, const double*, const double* const*); \
    template __global__ void k_qv_pcg<XY>(const CgInst*, const int*, const double*);
CG_MPASS_INSTANTIATE(false)
CG_MPASS_INSTANTIATE(true)
