// Kernel declarations and launch-time enums shared by the kernel files and the
// host driver.
#pragma once

#include "cgot_common.h"

// Host-written per-instance control block (ctl[b*CG_NCTL + field]).
enum {
    CTL_ACTIVE = 0,   // 1.0: the instance takes part in this launch
    CTL_STEP,         // trial step for CG_PT_TRIAL
    CTL_SHIFT,        // shift*reg added to the Hessian diagonal
    CTL_ALLOW_LR,     // 1.0 when the low-rank term may be used (iter > 0)
    CTL_KSEL,         // K for a new sparsity pattern
    CTL_AUX,          // kernel-specific
    CG_NCTL = 8
};

enum { CG_PT_GAMMA = 0, CG_PT_TRIAL = 1, CG_PT_BCD = 2 };

// ---- M passes (grid = total tiles; tile_inst maps tile -> instance) ----
template <bool XY>
__global__ void k_tpass(const CgInst*, const int*, const double*, int mode);
template <bool XY>
__global__ void k_row_lse(const CgInst*, const int*, const double*, int mode);
template <bool XY>
__global__ void k_col_lse(const CgInst*, const int*, const double*, int mode);
template <bool XY>
__global__ void k_qv_rhs(const CgInst*, const int*, const double*, const double* const* Gs);
template <bool XY>
__global__ void k_qv_pcg(const CgInst*, const int*, const double*);
__global__ void k_plan(const CgInst*, const int*, const double*);
__global__ void k_grad_assemble(const CgInst*, const int*, const double*, const double* const*,
                                const double* const*, double* const*);

// ---- finalize / vector kernels (grid = total fin or vec workgroups) ----
__global__ void k_tpass_fin(const CgInst*, const int*, const double*, int mode);
__global__ void k_rowlse_fin(const CgInst*, const int*, const double*, int mode);
__global__ void k_collse_fin(const CgInst*, const int*, const double*, int mode);
__global__ void k_gauge_shift(const CgInst*, const double*, int mode);      // grid = B
__global__ void k_init_logab(const CgInst*, const int*, const double*, const double* const*,
                             const double* const*);
__global__ void k_low_rank(const CgInst*, const int*, const double*);
__global__ void k_copy_bcd(const CgInst*, const int*, const double*);
__global__ void k_woodbury_dots(const CgInst*, const int*, const double*);
__global__ void k_woodbury_apply(const CgInst*, const int*, const double*);
__global__ void k_output_dual(const CgInst*, const int*, const double*);
__global__ void k_vcopy(const CgInst*, const int*, const double*, int what);
__global__ void k_bcd_diff(const CgInst*, const int*, const double*);
__global__ void k_hsl_csr(const CgInst*, const int*, const double*, double*, int*, int*);

// ---- sparsified Hessian (top-K pattern and values) ----
template <bool XY>
__global__ void k_hist(const CgInst*, const int*, const double*, int pass);
__global__ void k_hist_scan(const CgInst*, const double*, int pass);              // grid = B
__global__ void k_sel_rows_count(const CgInst*, const int* row_inst, const double*); // row grid
__global__ void k_sel_rows_scan(const CgInst*, const double*);                       // grid = B
__global__ void k_sel_rows_write(const CgInst*, const int* row_inst, const double*); // row grid
template <bool XY>
__global__ void k_sel_cols_count(const CgInst*, const int*, const double*);
__global__ void k_ctx_csr(const CgInst* fwd, double* vals, int* cols, int* rowptr);   // stored context
__global__ void k_sel_cols_prefix(const CgInst*, const int*, const double*);      // vector map
__global__ void k_sel_cols_scan(const CgInst*, const double*);                    // grid = B
template <bool XY>
__global__ void k_sel_cols_write(const CgInst*, const int*, const double*);
__global__ void k_refresh_values(const CgInst*, const int*, const double*, int wave_map);  // vector or wave grid

// ---- Newton solve on the Schur complement ----
__global__ void k_schur_build(const CgInst*, const int*, const double*);          // wave grid
__global__ void k_schur_rhs(const CgInst*, const int*, const double*);            // wave grid
__global__ void k_schur_back(const CgInst*, const int*, const double*);           // wave grid
// compact ragged grids: pref[b]..pref[b+1] are instance b's workgroups
__global__ void k_chol_fill(const CgInst*, int B, const int* pref);
__global__ void k_chol_diag0(const CgInst*, int B, const int* pref);
__global__ void k_chol_panel(const CgInst*, int B, const int* pref, int kb);
__global__ void k_wgpart_init(const CgInst*);
__global__ void k_chol_update(const CgInst*, int B, const int* pref, int seg_stride, int kb);
__global__ void k_pc_rhs(const CgInst*, const int* wave_inst, const double* ctl);
__global__ void k_pc_fwd(const CgInst*, const int* order, const int* lvl_pref, int nlvl, int target);
__global__ void k_pc_bwd(const CgInst*, const int* order, const int* lvl_pref, int nlvl);
__global__ void k_pc_back(const CgInst*, const int* wave_inst, const double* ctl);
__global__ void k_pc_wdots(const CgInst*, const int* wg_inst, const double* ctl);
__global__ void k_pc_wapply(const CgInst*, const int* wg_inst, const double* ctl);
__global__ void k_pcg_rz(const CgInst*, const int* wg_inst, const double* ctl, int init);
// sparse Newton solve (one workgroup of SP_WG threads per listed instance,
// dynamic LDS sp_lds_bytes(R) of the largest core)
// LV 0: plans in global memory (dynamic LDS sp_lds_bytes(R) of the largest
// core); 1: their values in LDS (at least the largest sp_lds_vals_bytes of
// the listed instances), and the integer arrays of each plan whose
// sp_lds_all_bytes fits the launch's lds_bytes too
template <int LV>
__global__ void k_sp_newton(const CgInst*, const int* ids, int lds_bytes);
template <int LV>
__global__ void k_newton_mixed(const CgInst*, const int* ids, int nsp, const int* sub_map, int nsub,
                               const double* ctl, int lds_bytes);   // k_sp_newton blocks, then Schur rows
__global__ void k_sp_precond(const CgInst*, const int* ids);
// pattern export / plan import through pinned host memory (grid = B, and
// B * SP_IMPORT_WGS)
#define SP_IMPORT_WGS 8
#define SP_HDR_PAD 512
__global__ void k_pat_export(const CgInst*, const double* ctl, int* dst);
__global__ void k_plan_import(const CgInst*, const double* ctl, const char* src);
__global__ void k_trsv_flow(const CgInst*, const int* order, const int* lvl_pref, int nlvl);

// ---- backward PCG (vec grid) ----
__global__ void k_pcg_init(const CgInst*, const int*, const double*);
__global__ void k_pcg_q_fin(const CgInst*, const int*, const double*);
__global__ void k_pcg_update(const CgInst*, const int*, const double*);
