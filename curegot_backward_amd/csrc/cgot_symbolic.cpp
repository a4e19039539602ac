// Symbolic analysis of the SPLR Newton system for the sparse factor/solve
// kernel (layout in cgot_symbolic.h).  Runs on the host once per sparsity
// pattern (the pattern is fixed for sparsity_pattern_cycle iterations), in
// the role of the reordering and symbolic factorization the reference hands
// to cuDSS on the CPU (src/linsolve.cu:171-176, 232-268; called once per
// pattern cycle, src/sinkhorn_splr.cu:583-591).  The numeric factorization
// and the solves run on the device every iteration.
#include "cgot_symbolic.h"

#include <algorithm>
#include <cstring>
#include <utility>

namespace {

struct Layout {
    int64_t off = 0;
    int64_t take(int64_t bytes, int64_t align = 16)
    {
        off = (off + align - 1) / align * align;
        const int64_t r = off;
        off += std::max<int64_t>(bytes, 0);
        return r;
    }
};

}  // namespace

bool sp_analyze(int n, int m, const int* rowptr, const int* colj, int64_t cap_bytes, int rmax, SpHdr& hdr,
                std::vector<char>& plan)
{
    const int N = n + m - 1;
    if (N <= 0) return false;
    rmax = std::min(rmax, SP_RMAX_BIG);
    // ---- bipartite graph of W: alpha_i -- beta_j for every selected (i, j) ----
    std::vector<std::vector<int>> adj(N);
    for (int i = 0; i < n; i++)
        for (int q = rowptr[i]; q < rowptr[i + 1]; q++) {
            const int v = n + colj[q];
            adj[i].push_back(v);
            adj[v].push_back(i);
        }
    // ---- minimum-degree elimination on the explicit elimination graph ----
    // Degree buckets are doubly linked lists (a vertex re-enters its bucket at
    // the head), so the choice among equal degrees is fixed by the pattern.
    // Stops once every remaining vertex is adjacent to at least half of the
    // others: those form the dense core.
    std::vector<int> head(N + 1, -1), nxt(N, -1), prv(N, -1), deg(N);
    auto link = [&](int v, int d) {
        deg[v] = d;
        prv[v] = -1;
        nxt[v] = head[d];
        if (head[d] >= 0) prv[head[d]] = v;
        head[d] = v;
    };
    auto unlink = [&](int v) {
        if (prv[v] >= 0) nxt[prv[v]] = nxt[v];
        else head[deg[v]] = nxt[v];
        if (nxt[v] >= 0) prv[nxt[v]] = prv[v];
    };
    for (int v = N - 1; v >= 0; v--) link(v, (int)adj[v].size());
    int mind = 0;
    std::vector<int> order;
    order.reserve(N);
    std::vector<std::vector<int>> structv(N);
    std::vector<int> mk(N, -1);
    int stamp = 0;
    int remaining = N;
    while (remaining > 0) {
        while (head[mind] < 0) mind++;
        if (2 * mind >= remaining) break;
        const int v = head[mind];
        unlink(v);
        order.push_back(v);
        std::vector<int>& nb = adj[v];
        structv[v] = nb;
        for (int u : nb) {
            std::vector<int>& au = adj[u];
            const int du = (int)au.size();
            for (size_t k = 0; k < au.size(); k++)
                if (au[k] == v) { au[k] = au.back(); au.pop_back(); break; }
            ++stamp;
            mk[u] = stamp;
            for (int w : au) mk[w] = stamp;
            for (int w : nb)
                if (mk[w] != stamp) { au.push_back(w); mk[w] = stamp; }
            if ((int)au.size() != du) {
                unlink(u);
                link(u, (int)au.size());
                mind = std::min(mind, (int)au.size());
            }
        }
        std::vector<int>().swap(nb);
        remaining--;
    }
    const int t0 = (int)order.size();
    if (N - t0 > rmax) return false;
    // core vertices in bucket order
    for (int d = 0; d <= N && (int)order.size() < N; d++)
        for (int v = head[d]; v >= 0; v = nxt[v]) order.push_back(v);
    std::vector<int> pos(N);
    for (int p = 0; p < N; p++) pos[order[p]] = p;
    // ---- elimination-tree levels of the eliminated vertices ----
    std::vector<std::vector<int>> colrows(t0);
    std::vector<int> level(t0, 0);
    for (int j = 0; j < t0; j++) {
        std::vector<int>& r = colrows[j];
        for (int w : structv[order[j]]) r.push_back(pos[w]);
        std::sort(r.begin(), r.end());
    }
    for (int j = 0; j < t0; j++)
        if (!colrows[j].empty() && colrows[j][0] < t0) {
            const int par = colrows[j][0];
            level[par] = std::max(level[par], level[j] + 1);
        }
    // ---- split: the last eliminated vertices (the top of the tree, where its
    // levels hold one or two columns each) join the dense core while that is
    // cheaper: a periphery level costs about as much as SP_LEVEL_COST core
    // columns (a level is two dependent gather phases and barriers, a core
    // column one broadcast step of the register-tile elimination) ----
    int t = t0;
    {
        std::vector<int> hpre(t0 + 1, 0);   // levels of the periphery [0, t')
        for (int j = 0; j < t0; j++) hpre[j + 1] = std::max(hpre[j], level[j] + 1);
        const int lo = std::max(0, N - rmax);
        double best = 1e300;
        for (int tc = t0; tc >= lo; tc--) {
            const double cost = SP_LEVEL_COST * hpre[tc] + (double)(N - tc);
            if (cost < best) { best = cost; t = tc; }
        }
        colrows.resize(t);
        level.resize(t);
    }
    const int R = N - t;
    int H = 0;
    for (int j = 0; j < t; j++) H = std::max(H, level[j] + 1);
    // columns in (level, position) order; entry indices: diagonal, then rows ascending
    std::vector<int> corder(t);
    for (int j = 0; j < t; j++) corder[j] = j;
    std::stable_sort(corder.begin(), corder.end(), [&](int x, int y) { return level[x] < level[y]; });
    std::vector<int> colstart(t + 1, 0);   // by position
    std::vector<int> lcols(t + 1), lvptr(H + 1, 0), lcptr(H + 1, 0);
    int64_t nnzP = 0;
    for (int c = 0; c < t; c++) {
        const int j = corder[c];
        colstart[j] = (int)nnzP;
        lcols[c] = (int)nnzP;
        nnzP += 1 + (int64_t)colrows[j].size();
    }
    lcols[t] = (int)nnzP;
    if (nnzP >= (int64_t)1 << 30) return false;
    for (int c = 0, l = 0; l <= H; l++) {
        while (c < t && level[corder[c]] < l) c++;
        lcptr[l] = c;
        lvptr[l] = lcols[c];
    }
    auto entry_of = [&](int row, int col) -> int {   // index of L(row, col), col periphery, row >= col
        if (row == col) return colstart[col];
        const std::vector<int>& r = colrows[col];
        const auto it = std::lower_bound(r.begin(), r.end(), row);
        return colstart[col] + 1 + (int)(it - r.begin());
    };
    const int ncore = R * (R + 1) / 2;
    auto core_of = [&](int a, int b) { return (a - t) * (a - t + 1) / 2 + (b - t); };   // a >= b >= t
    // per-entry data
    std::vector<int> edpos(nnzP), erow(nnzP), easrc(nnzP, -1);
    for (int j = 0; j < t; j++) {
        const int d = colstart[j];
        edpos[d] = d; erow[d] = j;
        for (size_t k = 0; k < colrows[j].size(); k++) {
            edpos[d + 1 + k] = d; erow[d + 1 + k] = colrows[j][k];
        }
    }
    std::vector<int> casrc(ncore, -1);
    // A-value sources: diagonals and W's entries (row CSR index q)
    for (int v = 0; v < N; v++) {
        const int p = pos[v];
        if (p < t) easrc[colstart[p]] = v;
        else casrc[core_of(p, p)] = v;
    }
    for (int i = 0; i < n; i++)
        for (int q = rowptr[i]; q < rowptr[i + 1]; q++) {
            const int pa = pos[i], pb = pos[n + colj[q]];
            const int hi = std::max(pa, pb), lo = std::min(pa, pb);
            if (lo < t) easrc[entry_of(hi, lo)] = N + q;
            else casrc[core_of(hi, lo)] = N + q;
        }
    // ---- gather lists: column k contributes L(a,k) L(b,k) to target (a, b), a >= b ----
    std::vector<int> ecnt(nnzP + 1, 0), ccnt(ncore + 1, 0);
    for (int k = 0; k < t; k++) {
        const std::vector<int>& r = colrows[k];
        for (size_t x = 0; x < r.size(); x++)
            for (size_t y = 0; y <= x; y++) {
                const int a = r[x], b = r[y];
                if (b < t) ecnt[entry_of(a, b) + 1]++;
                else ccnt[core_of(a, b) + 1]++;
            }
    }
    for (int64_t q = 0; q < nnzP; q++) ecnt[q + 1] += ecnt[q];
    for (int c = 0; c < ncore; c++) ccnt[c + 1] += ccnt[c];
    const int64_t npairs = ecnt[nnzP], ncpairs = ccnt[ncore];
    // ---- row-pair lists of the off-diagonal periphery entries (forward solve) ----
    std::vector<int> rptr(N + 1, 0);
    for (int64_t q = 0; q < nnzP; q++)
        if (edpos[q] != q) rptr[erow[q] + 1]++;
    for (int p = 0; p < N; p++) rptr[p + 1] += rptr[p];
    const int64_t nrowpairs = rptr[N];
    // ---- lists summed by eight-lane groups ----
    std::vector<int> lgptr(H + 1, 0), blptr(H + 1, 0), longs, blong, clong, ylong, cflags(t, 0);
    for (int l = 0; l < H; l++) {
        lgptr[l] = (int)longs.size();
        for (int q = lvptr[l]; q < lvptr[l + 1]; q++)
            if (ecnt[q + 1] - ecnt[q] > SP_LONG) longs.push_back(q);
        for (int c = lcptr[l]; c < lcptr[l + 1]; c++) {
            const int j = corder[c];
            if (rptr[j + 1] - rptr[j] > SP_LONG) longs.push_back(-(c + 1));
        }
        blptr[l] = (int)blong.size();
        for (int c = lcptr[l]; c < lcptr[l + 1]; c++)
            if ((int)colrows[corder[c]].size() > SP_BLONG) {
                blong.push_back(c);
                cflags[c] = 1;
            }
    }
    lgptr[H] = (int)longs.size();
    blptr[H] = (int)blong.size();
    for (int c = 0; c < ncore; c++)
        if (ccnt[c + 1] - ccnt[c] > SP_CLONG) clong.push_back(c);
    for (int i = 0; i < R; i++)
        if (rptr[t + i + 1] - rptr[t + i] > SP_LONG) ylong.push_back(i);
    // ---- layout: integer arrays first (prefetched together), then values ----
    Layout L;
    SpHdr h{};
    h.mode = 1;
    h.N = N; h.n = n; h.m = m;
    h.t = t; h.R = R; h.H = H;
    h.nnzP = (int)nnzP;
    h.ncore = ncore;
    h.T = (R + 3) / 4;
    h.npairs = npairs;
    h.ncpairs = ncpairs;
    h.nrowpairs = nrowpairs;
    h.nlong = (int)longs.size();
    h.nclong = (int)clong.size();
    h.nylong = (int)ylong.size();
    h.nblong = (int)blong.size();
    h.o_perm = L.take(4LL * N);
    h.o_lvptr = L.take(4LL * (H + 1));
    h.o_lcptr = L.take(4LL * (H + 1));
    h.o_lcols = L.take(4LL * (t + 1));
    h.o_emeta = L.take(16 * nnzP);
    h.o_erow = L.take(4 * nnzP);
    h.o_pairs = L.take(8 * npairs);
    h.o_ymeta = L.take(16LL * t);
    h.o_ycore = L.take(16LL * R);
    h.o_rowpairs = L.take(8 * nrowpairs);
    h.o_cmeta = L.take(16LL * ncore);
    h.o_cpairs = L.take(8 * ncpairs);
    h.o_lgptr = L.take(4LL * (H + 1));
    h.o_long = L.take(4LL * h.nlong);
    h.o_clong = L.take(4LL * h.nclong);
    h.o_ylong = L.take(4LL * h.nylong);
    h.o_blptr = L.take(4LL * (H + 1));
    h.o_blong = L.take(4LL * h.nblong);
    h.o_cflags = L.take(4LL * t);
    h.o_meta_end = L.take(0, 16);
    h.o_lval = L.take(8 * nnzP);
    h.o_lcore = L.take(8LL * ncore);
    h.o_yx = L.take(16LL * N);
    h.bytes = L.take(0, 256);
    if (h.bytes > cap_bytes || npairs >= ((int64_t)1 << 31) || ncpairs >= ((int64_t)1 << 31)) return false;
    // ---- fill the image: its integer part [0, o_meta_end); the value arrays
    // lval, lcore and yx after it are the device's ----
    plan.assign((size_t)h.o_meta_end, 0);
    char* base = plan.data();
    auto I32 = [&](int64_t off) { return reinterpret_cast<int*>(base + off); };
    std::memcpy(I32(h.o_perm), order.data(), 4LL * N);
    std::memcpy(I32(h.o_lvptr), lvptr.data(), 4LL * (H + 1));
    std::memcpy(I32(h.o_lcptr), lcptr.data(), 4LL * (H + 1));
    std::memcpy(I32(h.o_lcols), lcols.data(), 4LL * (t + 1));
    std::memcpy(I32(h.o_erow), erow.data(), 4 * nnzP);
    auto put = [&](int64_t off, const std::vector<int>& v) {
        if (!v.empty()) std::memcpy(I32(off), v.data(), 4 * v.size());
    };
    put(h.o_lgptr, lgptr);
    put(h.o_long, longs);
    put(h.o_clong, clong);
    put(h.o_ylong, ylong);
    put(h.o_blptr, blptr);
    put(h.o_blong, blong);
    put(h.o_cflags, cflags);
    {
        int* em = I32(h.o_emeta);
        for (int64_t q = 0; q < nnzP; q++) {
            em[4 * q] = easrc[q]; em[4 * q + 1] = ecnt[q]; em[4 * q + 2] = ecnt[q + 1];
            em[4 * q + 3] = edpos[q] | ((ecnt[q + 1] - ecnt[q] > SP_LONG) ? SP_LONG_BIT : 0);
        }
        int* cm = I32(h.o_cmeta);
        for (int c = 0; c < ncore; c++) {
            cm[4 * c] = casrc[c]; cm[4 * c + 1] = ccnt[c]; cm[4 * c + 2] = ccnt[c + 1];
            cm[4 * c + 3] = (ccnt[c + 1] - ccnt[c] > SP_CLONG) ? 1 : 0;
        }
        int* ym = I32(h.o_ymeta);
        for (int c = 0; c < t; c++) {
            const int j = corder[c];
            ym[4 * c] = order[j]; ym[4 * c + 1] = rptr[j]; ym[4 * c + 2] = rptr[j + 1];
            ym[4 * c + 3] = colstart[j] | ((rptr[j + 1] - rptr[j] > SP_LONG) ? SP_LONG_BIT : 0);
        }
        int* yc = I32(h.o_ycore);
        for (int i = 0; i < R; i++) {
            yc[4 * i] = order[t + i]; yc[4 * i + 1] = rptr[t + i]; yc[4 * i + 2] = rptr[t + i + 1];
            yc[4 * i + 3] = (rptr[t + i + 1] - rptr[t + i] > SP_LONG) ? 1 : 0;
        }
    }
    {
        // pairs in source-column order within each target (fixed summation order)
        int* pr = I32(h.o_pairs);
        int* cpr = I32(h.o_cpairs);
        std::vector<int> ef(ecnt.begin(), ecnt.end() - 1), cf(ccnt.begin(), ccnt.end() - 1);
        for (int k = 0; k < t; k++) {
            const std::vector<int>& r = colrows[k];
            const int d = colstart[k];
            for (size_t x = 0; x < r.size(); x++)
                for (size_t y = 0; y <= x; y++) {
                    const int a = r[x], b = r[y];
                    const int ea = d + 1 + (int)x, eb = d + 1 + (int)y;
                    if (b < t) {
                        const int64_t s = ef[entry_of(a, b)]++;
                        pr[2 * s] = ea; pr[2 * s + 1] = eb;
                    } else {
                        const int64_t s = cf[core_of(a, b)]++;
                        cpr[2 * s] = ea; cpr[2 * s + 1] = eb;
                    }
                }
        }
        // row pairs (L(j, k), k) by row j, in increasing column position k
        int* rpp = I32(h.o_rowpairs);
        std::vector<int> rf(rptr.begin(), rptr.end() - 1);
        for (int k = 0; k < t; k++) {
            const int d = colstart[k];
            for (size_t x = 0; x < colrows[k].size(); x++) {
                const int64_t s = rf[colrows[k][x]]++;
                rpp[2 * s] = d + 1 + (int)x;
                rpp[2 * s + 1] = k;
            }
        }
    }
    hdr = h;
    return true;
}
