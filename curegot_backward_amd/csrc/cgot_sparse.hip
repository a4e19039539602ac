// Sparsified Hessian: global top-K of T[:, :m-1] (radix select over the FP64
// bit patterns, T >= 0 so the bits order like the values), then the pattern
// in two orders: by row (CSR of W, for the large/small elimination) and by
// column (CSC of W).  Ties at the K-th value are taken in row-major index order
// i*(m-1)+j, the same rule as the oracle.
//
// Reference: launch_H_sparsification / launch_csr_conversion /
// recompute_nonzero_values_kernel, src/sinkhorn_splr_kernel.cu:597-906
// (CUB DeviceTopK + Thrust sort there).
#include "cgot_common.h"
#include "cgot_kernels.h"

CG_DEV double cg_T(const CgInstD& I, int i, int j, const double* tab)
{
    return cg_texp((I.gamma[i] + I.gamma[I.n + j] - cg_cost(I, i, j)) * I.inv_reg, tab);
}

CG_DEV unsigned long long cg_kth(const CgInstD& I) { return (unsigned long long)__double_as_longlong(I.scal[CS_KTH_HI]); }

CG_DEV bool cg_decode_tile2(const CgInst* insts, const int* tile_inst, const double* ctl,
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

// Radix histogram pass.  pass 0: top 12 bits (sign+exponent); pass p>=1: the
// next 13 bits among keys matching the prefix found so far.  A wave takes its
// CG_RPW rows in groups of R with the loads of a group issued first; the
// strip's beta and the rows' alpha are read once.  T is formed exactly as
// cg_T forms it, so the selection kernels see the same keys.
template <int C, bool FULL, bool XY>
CG_DEV void cg_hist_tile(const CgInstD& I, int rb, int st, int pass, unsigned int* h, const double* tab)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s0 = st * I.sw, mm = I.m - 1, m = I.m, n = I.n;
    const int shift = (pass == 0) ? 52 : 52 - CG_HBITS * pass;
    const unsigned long long mask = (pass == 0) ? 4095ull : (unsigned long long)(CG_HBINS - 1);
    const unsigned long long prefix = cg_kth(I);
    const double ir = I.inv_reg;
    double bj[C];
    bool ok[C];
#pragma unroll
    for (int k = 0; k < C; k++) {
        const int j = s0 + lane + 64 * k;
        ok[k] = FULL || j < mm;
        bj[k] = ok[k] ? I.gamma[n + j] : 0.0;
    }
    const int i0 = rb * CG_RB + w * CG_RPW;
    constexpr int R = 8 / C;   // 8 loads per lane in flight per group
#pragma unroll 1
    for (int r0 = 0; r0 < CG_RPW; r0 += R) {
        if (i0 + r0 >= n) break;
        double mv[R][C], ai[R];
        if constexpr (XY) cg_cost_rows_xy<R, C>(I, i0 + r0, n, s0, lane, ok, mv);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = min(i0 + r0 + r, n - 1);   // clamped rows are loaded but not counted
            ai[r] = I.gamma[i];
            if constexpr (!XY) cg_cost_row<C>(I, i, s0, lane, ok, mv[r]);
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const bool live = i0 + r0 + r < n;
#pragma unroll
            for (int k = 0; k < C; k++) {
                const unsigned long long key = cg_key(cg_texp((ai[r] + bj[k] - mv[r][k]) * ir, tab));
                if (live && ok[k] && (pass == 0 || (key >> (shift + CG_HBITS)) == prefix))
                    atomicAdd(&h[(key >> shift) & mask], 1u);
            }
        }
    }
}

template <bool XY>
__global__ __launch_bounds__(CG_WG) void k_hist(const CgInst* __restrict__ insts,
                                                const int* __restrict__ tile_inst,
                                                const double* __restrict__ ctl, int pass)
{
    CG_EXP_TABLE(tab);
    int b, rb, st;
    if (!cg_decode_tile2(insts, tile_inst, ctl, b, rb, st)) return;
    const CgInstD& I = cg_inst(insts, b);
    __shared__ unsigned int h[CG_HBINS];
    for (int q = threadIdx.x; q < CG_HBINS; q += CG_WG) h[q] = 0u;
    __syncthreads();
    cg_strip_dispatch<XY>(I, st, I.m - 1, [&](auto tag) {
        using S = decltype(tag);
        cg_hist_tile<S::C, S::FULL, S::XY>(I, rb, st, pass, h, tab);
    });
    __syncthreads();
    const int nb = (pass == 0) ? 4096 : CG_HBINS;
    for (int q = threadIdx.x; q < nb; q += CG_WG) {
        const unsigned int v = h[q];
        if (v) __hip_atomic_fetch_add(&I.hist[q], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// One workgroup per instance: find the bin holding the need-th largest key.
__global__ __launch_bounds__(CG_WG) void k_hist_scan(const CgInst* __restrict__ insts,
                                                     const double* __restrict__ ctl, int pass)
{
    const int b = blockIdx.x;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return;
    const CgInstD& I = cg_inst(insts, b);
    const int nb = (pass == 0) ? 4096 : CG_HBINS;
    const int cs = nb / CG_WG;
    __shared__ unsigned long long csum[CG_WG];
    __shared__ int sh_chunk;
    __shared__ unsigned long long sh_rem;
    unsigned long long need = (pass == 0) ? (unsigned long long)ctl[b * CG_NCTL + CTL_KSEL]
                                          : (unsigned long long)I.scal[CS_NEED];
    // chunk t covers bins [nb - (t+1)*cs, nb - t*cs), scanned from the top
    const int t = threadIdx.x;
    const int hi = nb - t * cs;
    unsigned long long s = 0;
    for (int q = hi - 1; q >= hi - cs; q--) s += I.hist[q];
    csum[t] = s;
    __syncthreads();
    if (t == 0) {
        unsigned long long acc = 0, tot = 0;
        for (int q = 0; q < CG_WG; q++) tot += csum[q];
        if (need > tot) need = tot;
        if (need == 0) need = 1;
        int ch = CG_WG - 1;
        for (int q = 0; q < CG_WG; q++) {
            if (acc + csum[q] >= need) { ch = q; break; }
            acc += csum[q];
        }
        sh_chunk = ch;
        sh_rem = need - acc;
    }
    __syncthreads();
    if (t == sh_chunk) {
        unsigned long long acc = 0, rem = sh_rem;
        int bstar = hi - cs;
        for (int q = hi - 1; q >= hi - cs; q--) {
            const unsigned long long c = I.hist[q];
            if (acc + c >= rem) { bstar = q; break; }
            acc += c;
        }
        const unsigned long long prefix = (pass == 0) ? 0ull : cg_kth(I);
        const unsigned long long np = (pass == 0) ? (unsigned long long)bstar
                                                  : ((prefix << CG_HBITS) | (unsigned long long)bstar);
        I.scal[CS_KTH_HI] = __longlong_as_double((long long)np);
        I.scal[CS_NEED] = (double)(rem - acc);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nb; q += CG_WG) I.hist[q] = 0u;
}

// Keys of U consecutive 64-column chunks of row i from column j0 (lane l
// holds column j0 + l + 64u), with every load issued before the first
// exponential.  Lanes past mm read column 0 of the row and report ok = false.
// T is formed exactly as cg_T forms it.
constexpr int CG_ROW_CHUNKS = 4;

template <int U, bool FULL>
CG_DEV void cg_row_keys(const CgInstD& I, int i, int j0, int mm, double ai, const double* tab,
                        unsigned long long (&key)[U], bool (&ok)[U])
{
    const int lane = threadIdx.x & 63;
    const auto Mi = I.M + (size_t)i * I.m;
    const auto gb = I.gamma + I.n;
    double mv[U], bv[U];
    const bool xy = I.X != nullptr;
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int j = j0 + lane + 64 * u;
        ok[u] = FULL || j < mm;
        const int jc = ok[u] ? j : 0;
        if (!xy) mv[u] = Mi[jc];
        bv[u] = gb[jc];
    }
    if (xy) {
#pragma unroll
        for (int u = 0; u < U; u++) mv[u] = cg_cost_xy(I, i, ok[u] ? j0 + lane + 64 * u : 0);
    }
#pragma unroll
    for (int u = 0; u < U; u++) key[u] = cg_key(cg_texp((ai + bv[u] - mv[u]) * I.inv_reg, tab));
}

// Row-wave kernels: one wave per row; blockIdx.x -> (instance, 4 rows) via row_inst.
// Count keys above the threshold and ties, per row.
__global__ __launch_bounds__(CG_WG) void k_sel_rows_count(const CgInst* __restrict__ insts,
                                                          const int* __restrict__ row_inst,
                                                          const double* __restrict__ ctl)
{
    CG_EXP_TABLE(tab);
    const int2 e = cg_map_entry(row_inst);
    const int b = e.x;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return;
    const CgInstD& I = cg_inst(insts, b);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = e.y * CG_WAVES + w;
    if (i >= I.n) return;
    const unsigned long long kth = cg_kth(I);
    const int mm = I.m - 1;
    const double ai = I.gamma[i];
    int gt = 0, tie = 0;
    auto tally = [&](int j0, auto tag) {
        constexpr int U = decltype(tag)::C;
        unsigned long long key[U];
        bool ok[U];
        cg_row_keys<U, decltype(tag)::FULL>(I, i, j0, mm, ai, tab, key, ok);
#pragma unroll
        for (int u = 0; u < U; u++) {
            gt += (ok[u] && key[u] > kth);
            tie += (ok[u] && key[u] == kth);
        }
    };
    int j0 = 0;
    for (; j0 + 64 * CG_ROW_CHUNKS <= mm; j0 += 64 * CG_ROW_CHUNKS) tally(j0, CgStrip<CG_ROW_CHUNKS, true>{});
    for (; j0 < mm; j0 += 64) tally(j0, CgStrip<1, false>{});
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) { gt += __shfl_xor(gt, o, 64); tie += __shfl_xor(tie, o, 64); }
    if (lane == 0) { I.rcnt[i] = gt; I.rtie[i] = tie; }
}

// Exclusive block scan of per-thread integer totals (256 threads); returns the
// exclusive prefix of this thread's total and writes the grand total.
CG_DEV long long cg_block_excl_scan(long long v, long long* lds, long long* total)
{
    lds[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long acc = 0;
        for (int q = 0; q < CG_WG; q++) { const long long x = lds[q]; lds[q] = acc; acc += x; }
        *total = acc;
    }
    __syncthreads();
    const long long r = lds[threadIdx.x];
    __syncthreads();
    return r;
}

// One workgroup per instance: ties taken per row (row-major order) and the
// CSR row pointer of the selected pattern.
__global__ __launch_bounds__(CG_WG) void k_sel_rows_scan(const CgInst* __restrict__ insts,
                                                         const double* __restrict__ ctl)
{
    const int b = blockIdx.x;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return;
    const CgInstD& I = cg_inst(insts, b);
    __shared__ long long lds[CG_WG];
    __shared__ long long total;
    const int n = I.n;
    const int per = (n + CG_WG - 1) / CG_WG;
    const int r0 = threadIdx.x * per, r1 = min(n, r0 + per);
    const long long need = (long long)I.scal[CS_NEED];
    long long tsum = 0;
    for (int r = r0; r < r1; r++) tsum += I.rtie[r];
    long long tb = cg_block_excl_scan(tsum, lds, &total);
    long long ssum = 0;
    for (int r = r0; r < r1; r++) {
        const long long tk = I.rtie[r];
        long long take = need - tb;
        take = take < 0 ? 0 : (take > tk ? tk : take);
        tb += tk;
        I.rtie[r] = (int)take;
        ssum += I.rcnt[r] + take;
    }
    long long off = cg_block_excl_scan(ssum, lds, &total);
    for (int r = r0; r < r1; r++) {
        I.w_rowptr[r] = (int)off;
        off += I.rcnt[r] + I.rtie[r];
    }
    if (threadIdx.x == 0) I.w_rowptr[n] = (int)total;
}

// One wave per row: write the selected columns of row i in ascending order and
// the tie cut column (selected ties of the row are those with j < rcut[i]).
__global__ __launch_bounds__(CG_WG) void k_sel_rows_write(const CgInst* __restrict__ insts,
                                                          const int* __restrict__ row_inst,
                                                          const double* __restrict__ ctl)
{
    CG_EXP_TABLE(tab);
    const int2 e = cg_map_entry(row_inst);
    const int b = e.x;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return;
    const CgInstD& I = cg_inst(insts, b);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = e.y * CG_WAVES + w;
    if (i >= I.n) return;
    const unsigned long long kth = cg_kth(I);
    const int mm = I.m - 1;
    const int off = I.w_rowptr[i];
    const int take = I.rtie[i];
    const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const double ai = I.gamma[i];
    int tseen = 0, wpos = 0, cut = 0;
    // chunks in column order: ties are taken, and entries placed, left to right
    auto place = [&](int j0, auto tag) {
        constexpr int U = decltype(tag)::C;
        unsigned long long key[U];
        bool ok[U];
        cg_row_keys<U, decltype(tag)::FULL>(I, i, j0, mm, ai, tab, key, ok);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int j = j0 + lane + 64 * u;
            const bool isgt = ok[u] && key[u] > kth;
            const bool istie = ok[u] && key[u] == kth;
            const unsigned long long tmask = __ballot(istie);
            const int trank = tseen + __popcll(tmask & lt);
            const bool tsel = istie && trank < take;
            const bool sel = isgt || tsel;
            const unsigned long long smask = __ballot(sel);
            const int pos = wpos + __popcll(smask & lt);
            if (sel) I.w_colj[off + pos] = j;
            if (tsel) cut = max(cut, j + 1);
            tseen += __popcll(tmask);
            wpos += __popcll(smask);
        }
    };
    int j0 = 0;
    for (; j0 + 64 * CG_ROW_CHUNKS <= mm; j0 += 64 * CG_ROW_CHUNKS) place(j0, CgStrip<CG_ROW_CHUNKS, true>{});
    for (; j0 < mm; j0 += 64) place(j0, CgStrip<1, false>{});
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cut = max(cut, __shfl_xor(cut, o, 64));
    if (lane == 0) I.rcut[i] = cut;
}

CG_DEV bool cg_selected(unsigned long long key, unsigned long long kth, int j, int cut)
{
    return key > kth || (key == kth && j < cut);
}

// Selected entries of a tile for the wave's CG_RPW rows: bit r of bits[k] is
// set when (i0 + r, s0 + lane + 64k) is selected.  Rows in groups of R with
// the loads of a group issued first; beta, alpha and the rows' tie cuts are
// read once.  T is formed exactly as cg_T forms it.
template <int C, bool FULL, bool XY>
CG_DEV void cg_sel_bits(const CgInstD& I, int rb, int st, const double* tab, unsigned int (&bits)[C])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s0 = st * I.sw, mm = I.m - 1, m = I.m, n = I.n;
    const unsigned long long kth = cg_kth(I);
    const double ir = I.inv_reg;
    double bj[C];
    bool ok[C];
#pragma unroll
    for (int k = 0; k < C; k++) {
        const int j = s0 + lane + 64 * k;
        ok[k] = FULL || j < mm;
        bj[k] = ok[k] ? I.gamma[n + j] : 0.0;
        bits[k] = 0u;
    }
    const int i0 = rb * CG_RB + w * CG_RPW;
    constexpr int R = 8 / C;
    static_assert(CG_RPW <= 32, "one bit per row of the wave");
#pragma unroll 1
    for (int r0 = 0; r0 < CG_RPW; r0 += R) {
        if (i0 + r0 >= n) break;
        double mv[R][C], ai[R];
        int cut[R];
        if constexpr (XY) cg_cost_rows_xy<R, C>(I, i0 + r0, n, s0, lane, ok, mv);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = min(i0 + r0 + r, n - 1);   // clamped rows are loaded but not taken
            ai[r] = I.gamma[i];
            cut[r] = I.rcut[i];
            if constexpr (!XY) cg_cost_row<C>(I, i, s0, lane, ok, mv[r]);
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const bool live = i0 + r0 + r < n;
#pragma unroll
            for (int k = 0; k < C; k++) {
                const unsigned long long key = cg_key(cg_texp((ai[r] + bj[k] - mv[r][k]) * ir, tab));
                if (live && ok[k] && cg_selected(key, kth, s0 + lane + 64 * k, cut[r])) bits[k] |= 1u << (r0 + r);
            }
        }
    }
}

// Tile pass: selected count per (row block, column).
template <bool XY>
__global__ __launch_bounds__(CG_WG) void k_sel_cols_count(const CgInst* __restrict__ insts,
                                                          const int* __restrict__ tile_inst,
                                                          const double* __restrict__ ctl)
{
    CG_EXP_TABLE(tab);
    int b, rb, st;
    if (!cg_decode_tile2(insts, tile_inst, ctl, b, rb, st)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s0 = st * I.sw, mm = I.m - 1;
    __shared__ int lds[CG_WAVES][CG_CMAX * 64];
    cg_strip_dispatch<XY>(I, st, mm, [&](auto tag) {
        using S = decltype(tag);
        unsigned int bits[S::C];
        cg_sel_bits<S::C, S::FULL, S::XY>(I, rb, st, tab, bits);
#pragma unroll
        for (int k = 0; k < S::C; k++) lds[w][lane + 64 * k] = __popc(bits[k]);
    });
    __syncthreads();
    // columns j < m-1 of the strip lie below 64*C of the strip's C
    for (int q = threadIdx.x; q < I.sw; q += CG_WG) {
        const int j = s0 + q;
        if (j < mm) I.ccnt[(size_t)rb * I.m + j] = lds[0][q] + lds[1][q] + lds[2][q] + lds[3][q];
    }
}

// Thread per column j < m-1 (vector-map workgroups): exclusive prefix of the
// column's counts over the row blocks, in place in ccnt (the first slot of
// every (row block, column) inside the column), and the column total into
// w_colptr[j].  Loads of CG_SUM_U row blocks are in flight at a time and
// coalesced across the threads' columns.
__global__ __launch_bounds__(CG_WG) void k_sel_cols_prefix(const CgInst* __restrict__ insts,
                                                           const int* __restrict__ wg_inst,
                                                           const double* __restrict__ ctl)
{
    const int2 e = cg_map_entry(wg_inst);
    const int b = e.x;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return;
    const CgInstD& I = cg_inst(insts, b);
    const int m = I.m, mm = m - 1, nrb = I.nrb;
    const int j = e.y * CG_WG + threadIdx.x;
    if (j >= mm) return;
    int run = 0;
    for (int r0 = 0; r0 < nrb; r0 += CG_SUM_U) {
        int c[CG_SUM_U];
#pragma unroll
        for (int u = 0; u < CG_SUM_U; u++) c[u] = (r0 + u < nrb) ? I.ccnt[(size_t)(r0 + u) * m + j] : 0;
#pragma unroll
        for (int u = 0; u < CG_SUM_U; u++)
            if (r0 + u < nrb) {
                I.ccnt[(size_t)(r0 + u) * m + j] = run;
                run += c[u];
            }
    }
    I.w_colptr[j] = run;   // column total for now
}

// One workgroup per instance: exclusive scan of the column totals into the
// CSC column pointers.  A slot of (row block rb, column j) is then
// w_colptr[j] + ccnt[rb*m + j].
__global__ __launch_bounds__(CG_WG) void k_sel_cols_scan(const CgInst* __restrict__ insts,
                                                         const double* __restrict__ ctl)
{
    const int b = blockIdx.x;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return;
    const CgInstD& I = cg_inst(insts, b);
    __shared__ long long lds[CG_WG];
    __shared__ long long total;
    const int mm = I.m - 1;
    const int per = (mm + CG_WG - 1) / CG_WG;
    const int c0 = threadIdx.x * per, c1 = min(mm, c0 + per);
    long long csum = 0;
    for (int j = c0; j < c1; j++) csum += I.w_colptr[j];
    long long off = cg_block_excl_scan(csum, lds, &total);
    for (int j = c0; j < c1; j++) {
        const int c = I.w_colptr[j];
        I.w_colptr[j] = (int)off;
        off += c;
    }
    if (threadIdx.x == 0) I.w_colptr[mm > 0 ? mm : 0] = (int)total;
}

// Tile pass: write the row indices of the CSC in ascending row order.
template <bool XY>
__global__ __launch_bounds__(CG_WG) void k_sel_cols_write(const CgInst* __restrict__ insts,
                                                          const int* __restrict__ tile_inst,
                                                          const double* __restrict__ ctl)
{
    CG_EXP_TABLE(tab);
    int b, rb, st;
    if (!cg_decode_tile2(insts, tile_inst, ctl, b, rb, st)) return;
    const CgInstD& I = cg_inst(insts, b);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int s0 = st * I.sw, mm = I.m - 1;
    const int i0 = rb * CG_RB + w * CG_RPW;
    __shared__ int lds[CG_WAVES][CG_CMAX * 64];
    cg_strip_dispatch<XY>(I, st, mm, [&](auto tag) {
        using S = decltype(tag);
        unsigned int bits[S::C];
        cg_sel_bits<S::C, S::FULL, S::XY>(I, rb, st, tab, bits);
#pragma unroll
        for (int k = 0; k < S::C; k++) lds[w][lane + 64 * k] = __popc(bits[k]);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < S::C; k++) {
            const int q = lane + 64 * k;
            const int j = s0 + q;
            if (j < mm && bits[k]) {
                int pos = I.w_colptr[j] + I.ccnt[(size_t)rb * I.m + j];
                for (int ww = 0; ww < w; ww++) pos += lds[ww][q];
                for (int r = 0; r < CG_RPW; r++)
                    if (bits[k] & (1u << r)) I.w_rowi[pos++] = i0 + r;
            }
        }
    });
}

// Refresh W values on the current pattern and the diagonal blocks
// Da = T1 + shift, Db = (T^T 1)[:m-1] + shift.  Reference:
// recompute_nonzero_values_kernel + write_diagonal_kernel,
// src/sinkhorn_splr_kernel.cu:554-734.
//
// Flat over the instance's items: N diagonal entries, then the K entries of
// W's row CSR, then the K entries of its column CSC, one thread per item
// (grid-stride over the instance's workgroups of the launch map: the vector
// map, ceil((n+m)/CG_WG) workgroups, or for dense patterns the wave map,
// ceil((n+m)/CG_WAVES); `wave_map` says which).  A row or column holds about
// K/n entries on average but the top-K concentrates on the heavy rows, so
// items rather than rows are spread over the lanes; a CSR (CSC) entry finds
// its row (column) by binary search in the pointer array.
template <class Ptr>
CG_DEV int cg_upper_index(Ptr ptr, int count, int p)
{
    // largest r in [0, count) with ptr[r] <= p (ptr non-decreasing, ptr[0] = 0)
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ptr[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(CG_WG) void k_refresh_values(const CgInst* __restrict__ insts,
                                                          const int* __restrict__ map,
                                                          const double* __restrict__ ctl, int wave_map)
{
    CG_EXP_TABLE(tab);
    const int2 e = cg_map_entry(map);
    const int b = e.x;
    if (ctl[b * CG_NCTL + CTL_ACTIVE] == 0.0) return;
    const CgInstD& I = cg_inst(insts, b);
    const double sh = ctl[b * CG_NCTL + CTL_SHIFT];
    const int n = I.n, mm = I.m - 1, N = I.N;
    const int K = I.w_rowptr[n];
    const long long items = (long long)N + 2LL * K;
    const int stride = (wave_map ? I.nrow : I.nvec) * CG_WG;
    for (long long x = (long long)e.y * CG_WG + threadIdx.x; x < items; x += stride) {
        if (x < n) {
            I.Da[x] = I.rowsum[x] + sh;
        } else if (x < N) {
            const int j = (int)x - n;
            I.Db[j] = I.colsum[j] + sh;
        } else if (x < (long long)N + K) {
            const int p = (int)(x - N);
            const int i = cg_upper_index(I.w_rowptr, n, p);
            I.w_rval[p] = cg_T(I, i, I.w_colj[p], tab);
        } else {
            const int p = (int)(x - N - K);
            const int j = cg_upper_index(I.w_colptr, mm, p);
            I.w_cval[p] = cg_T(I, I.w_rowi[p], j, tab);
        }
    }
}

// Lower CSR of Hsl in the reference layout (test hook): rows 0..n-1 hold their
// diagonal, row n+j holds the selected T_ij (ascending i) and then its diagonal.
// Reference: the CSR built by launch_sphess, src/sinkhorn_splr_kernel.cu:839-906.
CG_DEV void cg_hsl_csr_entry(const CgInstD& I, int q, double* __restrict__ vals, int* __restrict__ cols,
                              int* __restrict__ rowptr)
{
    const int n = I.n;
    if (q < n) {
        vals[q] = I.Da[q];
        cols[q] = q;
        rowptr[q] = q;
    }
    if (q < I.m - 1) {
        int p = n + I.w_colptr[q] + q;
        rowptr[n + q] = p;
        for (int e = I.w_colptr[q]; e < I.w_colptr[q + 1]; e++) { vals[p] = I.w_cval[e]; cols[p] = I.w_rowi[e]; p++; }
        vals[p] = I.Db[q];
        cols[p] = n + q;
    }
    if (q == 0) rowptr[I.N] = n + I.w_colptr[I.m - 1] + (I.m - 1);
}

__global__ __launch_bounds__(CG_WG) void k_hsl_csr(const CgInst* __restrict__ insts,
                                                   const int* __restrict__ wg_inst,
                                                   const double* __restrict__ ctl,
                                                   double* __restrict__ vals, int* __restrict__ cols,
                                                   int* __restrict__ rowptr)
{
    const int2 e = cg_map_entry(wg_inst);
    cg_hsl_csr_entry(cg_inst(insts, e.x), e.y * CG_WG + threadIdx.x, vals, cols, rowptr);
}

// The same CSR for one stored forward descriptor (the context of a finished
// solve), grid = ceil(max(n, m-1) / CG_WG).
__global__ __launch_bounds__(CG_WG) void k_ctx_csr(const CgInst* __restrict__ fwd, double* __restrict__ vals,
                                                   int* __restrict__ cols, int* __restrict__ rowptr)
{
    cg_hsl_csr_entry(cg_inst(fwd, 0), blockIdx.x * CG_WG + threadIdx.x, vals, cols, rowptr);
}

// the stored-cost and the geometric-cost instantiations of the selection passes
#define CG_SEL_INSTANTIATE(XY)                                                              \
    template __global__ void k_hist<XY>(const CgInst*, const int*, const double*, int);    \
    template __global__ void k_sel_cols_count<XY>(const CgInst*, const int*, const double*); \
    template __global__ void k_sel_cols_write<XY>(const CgInst*, const int*, const double*);
CG_SEL_INSTANTIATE(false)
CG_SEL_INSTANTIATE(true)
