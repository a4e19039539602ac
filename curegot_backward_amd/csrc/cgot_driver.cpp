// Host driver of the ragged-batch SPLR solver (C ABI in include/curegot_amd.h).
//
// One driver serves the plain single solve (B = 1) and the batched solve: all
// instances advance in lockstep, every launch covers every instance that takes
// part in that phase, and the host reads the per-instance scalars (objective,
// gradient norm, directional derivative) once per line-search round for the
// whole batch.  The per-instance control logic is the reference's, restated:
//   cuda_sinkhorn_splr_bwd (src/sinkhorn_splr.cu:1305-1611): density/shift
//   schedules, Sinkhorn candidate every sparsity_pattern_cycle iterations,
//   low-rank rule, stop-after-step semantics, context export;
//   SPLRSolver::line_search_wolfe (src/sinkhorn_splr.cu:667-920) as a resumable
//   state machine (one trial per round).
// Kernels are deterministic and per-instance, so a result does not depend on
// the batch it was solved in.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <cstdio>
#include <cstring>
#include <limits>
#include <array>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cgot_kernels.h"
#include "cgot_symbolic.h"
#include "../../include/curegot_amd.h"

namespace {

thread_local std::string g_last_error;

#define CG_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + \
                                     " at " __FILE__ ":" + std::to_string(__LINE__));   \
    } while (0)

// ---------------------------------------------------------------------------
// Process-lifetime resources: pinned staging buffers, the size-class streams
// and the worker threads.  They are registered here and released together by
// cgot_shutdown (called explicitly, or at exit while no solve is running);
// the generation counter tells per-thread caches that their entries are gone.
// ---------------------------------------------------------------------------
std::mutex g_res_mu;
std::vector<void*> g_res_pinned;
std::vector<hipStream_t> g_res_streams;
std::atomic<int> g_res_gen{0};
std::atomic<int> g_inflight{0};   // solves running inside the library
void register_atexit();

struct InFlight {
    InFlight() { g_inflight.fetch_add(1); }
    ~InFlight() { g_inflight.fetch_sub(1); }
};

// Pinned host staging (per thread, grown on demand).  `coherent` buffers
// (hipHostMallocCoherent) are read and written by kernels directly: device
// accesses to them go to host memory uncached.
struct Pinned {
    double* p = nullptr;
    size_t cap = 0;
    int gen = -1;
    bool coherent = false;
    Pinned() = default;
    explicit Pinned(bool coh) : coherent(coh) {}
    double* get(size_t n)
    {
        if (gen != g_res_gen.load()) { p = nullptr; cap = 0; gen = g_res_gen.load(); }
        if (n > cap) {
            // the old buffer may still be read by an in-flight copy: keep it
            // (registered, freed at shutdown).  Grow geometrically so that a
            // thread reaches its working size in a few allocations (pinning
            // host memory is slow).
            const size_t c = std::max<size_t>({n, 2 * cap, 65536});
            double* q = nullptr;
            CG_HIP(hipHostMalloc((void**)&q, c * sizeof(double),
                                 coherent ? hipHostMallocCoherent : hipHostMallocDefault));
            {
                std::lock_guard<std::mutex> lk(g_res_mu);
                g_res_pinned.push_back(q);
            }
            register_atexit();
            p = q;
            cap = c;
        }
        return p;
    }
    // the device address of the buffer (after get)
    void* dev() const
    {
        void* d = nullptr;
        CG_HIP(hipHostGetDevicePointer(&d, p, 0));
        return d;
    }
};
thread_local Pinned g_ctl_host;
thread_local Pinned g_scal_host;
thread_local Pinned g_map_host;
thread_local Pinned g_sub_host;

// ---------------------------------------------------------------------------
// Workspace layout: a bump allocator over the caller's bytes.
// ---------------------------------------------------------------------------
struct Bump {
    char* base;
    size_t off = 0;
    explicit Bump(void* b) : base((char*)b) {}
    template <class T>
    T* take(size_t count)
    {
        off = (off + 255) & ~(size_t)255;
        T* r = base ? (T*)(base + off) : nullptr;
        off += std::max<size_t>(count, 1) * sizeof(T);
        return r;
    }
};

constexpr int kCtlSlots = 16;
constexpr int kSticky = 2;   // control blocks kept at fixed addresses, uploaded only when they change

// ---------------------------------------------------------------------------
// Optional per-kernel-class timing with HIP events on the solve stream
// (enabled by cgot_profile_enable; used by the benchmark for roofline figures).
// ---------------------------------------------------------------------------
enum { PROF_TPASS = 0, PROF_ROWLSE, PROF_COLLSE, PROF_QV, PROF_CHOL_UPDATE, PROF_CHOL_PANEL,
       PROF_TRSV, PROF_SPNEWTON, PROF_NCLS = 8 };
struct ProfRec { hipEvent_t a, b; int cls; double work; };
bool g_prof = false;
thread_local std::vector<ProfRec> g_prof_pending;   // recorded by this thread, not yet read
thread_local std::vector<hipEvent_t> g_prof_pool;
std::mutex g_prof_mu;                               // guards the totals below
double g_prof_ms[PROF_NCLS] = {0};
double g_prof_work[PROF_NCLS] = {0};
long long g_prof_count[PROF_NCLS] = {0};

hipEvent_t prof_event()
{
    if (!g_prof_pool.empty()) { hipEvent_t e = g_prof_pool.back(); g_prof_pool.pop_back(); return e; }
    // timing only: without the system-scope fence that a default event
    // performs when it is recorded (an L2 write-back and invalidation that
    // would add to the interval and slow the work after it)
    hipEvent_t e;
    CG_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    return e;
}

struct ProfScope {
    hipStream_t s; ProfRec r; bool on;
    ProfScope(hipStream_t st, int cls, double work) : s(st), on(g_prof)
    {
        if (!on) return;
        r.a = prof_event(); r.b = prof_event(); r.cls = cls; r.work = work;
        CG_HIP(hipEventRecord(r.a, s));
    }
    ~ProfScope()
    {
        if (!on) return;
        (void)hipEventRecord(r.b, s);
        g_prof_pending.push_back(r);
    }
};

void prof_collect()
{
    for (auto& r : g_prof_pending) {
        float ms = 0.f;
        CG_HIP(hipEventSynchronize(r.b));
        CG_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        std::lock_guard<std::mutex> lk(g_prof_mu);
        g_prof_ms[r.cls] += ms;
        g_prof_work[r.cls] += r.work;
        g_prof_count[r.cls] += 1;
        g_prof_pool.push_back(r.a);
        g_prof_pool.push_back(r.b);
    }
    g_prof_pending.clear();
}

int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

// Columns per lane of the M-pass tiles (a strip is 64c columns wide).  The
// default: the next power of two >= ceil(m/64), at most CG_CMAX.
int default_cols(int m)
{
    const int need = std::max(1, cdiv(m, 64));
    int c = 1;
    while (c < need && c < CG_CMAX) c *= 2;
    return c;
}

// The tile table, process-wide: rows {m_max, c} by ascending m_max, row k
// covering the m <= m_max not covered by an earlier row; m past the last row
// takes the default.  Written by cgot_set_tile_table (the autotuner's choice).
std::mutex g_tile_mu;
std::vector<std::pair<int, int>> g_tile_table;

int tile_cols(int m)
{
    std::lock_guard<std::mutex> lk(g_tile_mu);
    for (const auto& r : g_tile_table)
        if (m <= r.first) return r.second;
    return default_cols(m);
}

// Geometry common to every instance kind.
void set_geometry(CgInst& I, int n, int m, double reg)
{
    I.n = n;
    I.m = m;
    I.N = n + m - 1;
    I.c = tile_cols(m);
    I.sw = 64 * I.c;
    I.nstrips = cdiv(m, I.sw);
    I.nrb = cdiv(n, CG_RB);
    I.ntiles = I.nrb * I.nstrips;
    I.nvec = cdiv(n + m, CG_WG);
    I.nrow = cdiv(n + m, CG_WAVES);
    I.reg = reg;
    I.inv_reg = 1.0 / reg;
}

// The cost source of a descriptor: a stored M, or the geometric cost (X, Y).
void set_cost(CgInst& I, const double* M, const cgot_cost_xy& xy)
{
    I.M = xy.X ? nullptr : M;
    I.X = xy.X;
    I.Y = xy.X ? xy.Y : nullptr;
    I.d = xy.X ? xy.d : 0;
    I.cscale = xy.X ? xy.scale : 0.0;
}

void check_cost(const double* M, const cgot_cost_xy& xy)
{
    if (xy.X) {
        if (!xy.Y) throw std::invalid_argument("geometric cost: X and Y must both be non-null device pointers");
        if (xy.d < 1) throw std::invalid_argument("geometric cost: d must be >= 1");
        if (!std::isfinite(xy.scale)) throw std::invalid_argument("geometric cost: scale must be finite");
    } else if (!M) {
        throw std::invalid_argument("M must be a non-null device pointer (or give the geometric cost X, Y)");
    }
}

// Maps and per-launch tables shared by all solvers.
struct Batch {
    int B = 0;
    std::vector<CgInst> h;     // host copy of descriptors
    CgInst* d_insts = nullptr;
    // Launch maps, one {instance, workgroup within the instance} pair per
    // workgroup: the current maps (and their workgroup counts) are either the
    // full maps of the batch or the packed maps of the instances still live.
    int* d_tile = nullptr;
    int* d_vec = nullptr;
    int* d_row = nullptr;
    int ntiles = 0, nvec = 0, nrow = 0;
    int* d_full = nullptr;     // tiles | vecs | rows of every instance
    int* d_live = nullptr;     // the same layout for the packed live set
    int ntiles_full = 0, nvec_full = 0, nrow_full = 0;
    long long packed_wgs = 0;  // workgroups of the current maps
    double* d_ctl = nullptr;   // kCtlSlots * B * CG_NCTL
    double* d_fixctl = nullptr;   // B * CG_NCTL, a control block at a fixed address (captured graphs)
    double* d_sticky = nullptr;   // kSticky * B * CG_NCTL (ctl_sticky)
    std::vector<double> sticky_host[kSticky];   // what each sticky block holds
    std::vector<int> ids_held, pref_held;       // what d_ids and d_pref hold (search_direction)
    double* d_scal = nullptr;  // B * CG_NSCAL
    int* d_pref = nullptr;     // dense prefix tables
    int pref_cap = 0;
    const double** d_ptrs = nullptr;   // 3*B pointers
    int* d_ids = nullptr;              // B instance ids of a sparse-solve launch
    // wave map of a subset of the instances (those on the dense Newton path),
    // so that the per-row Schur kernels dispatch no workgroups for the others
    int* d_sub = nullptr;
    int nsub = 0;
    std::vector<char> sub_key;
    std::vector<char> pc_sparse;       // adjoint: the forward factor is a sparse plan
    std::vector<int> pc_R;             // adjoint: its core size
    unsigned int* zero_base = nullptr; // counters + histograms of all instances
    size_t zero_bytes = 0;
    hipStream_t stream = nullptr;
    int slot = 0;              // next ctl slot
    double* hctl = nullptr;    // pinned, kCtlSlots * B * CG_NCTL
    double* hscal = nullptr;   // pinned, B * CG_NSCAL
    // Host staging in pageable memory instead of the pinned buffers (the
    // entry point that allocates no GPU or pinned memory).  The ring of
    // control slots and the synchronisations are the same either way.
    bool pageable = false;
    std::vector<int> pg_maps, pg_sub;
    bool xy = false;           // every instance has the geometric cost (else none has)
    // the value refresh runs on the wave map (64x the threads of the vector
    // map) when some instance's pattern may hold more than 32 items per
    // vector-map thread (N + 2 Kmax items)
    bool refresh_rows = false;
    std::vector<double> pg_ctl, pg_scal;
    std::map<const double*, const double*> hmap;   // device ctl slot -> host copy

    void layout_common(Bump& bp, int nblk_max)
    {
        d_insts = bp.take<CgInst>(B);
        const size_t nmap = 2 * ((size_t)ntiles + nvec + nrow);
        d_full = bp.take<int>(std::max<size_t>(nmap, 2));
        d_live = bp.take<int>(std::max<size_t>(nmap, 2));
        d_ctl = bp.take<double>((size_t)kCtlSlots * B * CG_NCTL);
        d_fixctl = bp.take<double>((size_t)B * CG_NCTL);
        d_sticky = bp.take<double>((size_t)kSticky * B * CG_NCTL);
        for (auto& v : sticky_host) v.clear();
        ids_held.clear();
        pref_held.clear();
        d_scal = bp.take<double>((size_t)B * CG_NSCAL);
        pref_cap = (4 * nblk_max + 3) * (B + 1) + nblk_max + 1;
        d_pref = bp.take<int>(pref_cap);
        d_ptrs = bp.take<const double*>(3 * (size_t)B);
        d_ids = bp.take<int>(B);
        d_sub = bp.take<int>(2 * (size_t)std::max(nrow, 1));
    }

    void assign_maps()
    {
        // the M passes are launched for one cost kind
        xy = !h.empty() && h[0].X != nullptr;
        for (auto& I : h)
            if ((I.X != nullptr) != xy)
                throw std::invalid_argument("a batch must not mix geometric (X, Y) and stored (M) costs");
        ntiles = nvec = nrow = 0;
        for (auto& I : h) {
            I.tile0 = ntiles; ntiles += I.ntiles;
            I.vec0 = nvec; nvec += I.nvec;
            I.row0 = nrow; nrow += I.nrow;
        }
        ntiles_full = ntiles; nvec_full = nvec; nrow_full = nrow;
    }

    // Write the maps of the instances with keep[b] into dst (device) through
    // the pinned staging buffer; returns their workgroup counts.
    void write_maps(int* dst, const std::vector<char>& keep, int& nt, int& nv, int& nr)
    {
        const size_t cap = 2 * ((size_t)ntiles_full + nvec_full + nrow_full);
        if (pageable) pg_maps.resize(cap + 1);
        int* hm = pageable ? pg_maps.data() : (int*)g_map_host.get((cap + 1) / 2);
        nt = nv = nr = 0;
        for (int b = 0; b < B; b++) if (keep[b]) { nt += h[b].ntiles; nv += h[b].nvec; nr += h[b].nrow; }
        int* pt = hm;
        int* pv = hm + 2 * (size_t)nt;
        int* pr = pv + 2 * (size_t)nv;
        for (int b = 0; b < B; b++) {
            if (!keep[b]) continue;
            const CgInst& I = h[b];
            for (int t = 0; t < I.ntiles; t++) { *pt++ = b; *pt++ = t; }
            for (int t = 0; t < I.nvec; t++) { *pv++ = b; *pv++ = t; }
            for (int t = 0; t < I.nrow; t++) { *pr++ = b; *pr++ = t; }
        }
        const size_t used = 2 * ((size_t)nt + nv + nr);
        if (used) CG_HIP(hipMemcpyAsync(dst, hm, sizeof(int) * used, hipMemcpyHostToDevice, stream));
    }

    void set_maps(int* base, int nt, int nv, int nr)
    {
        d_tile = base;
        d_vec = base + 2 * (size_t)nt;
        d_row = d_vec + 2 * (size_t)nv;
        ntiles = nt; nvec = nv; nrow = nr;
        packed_wgs = (long long)nt + nv + nr;
    }

    // Launch over every instance again (after a loop that packed the maps).
    void use_full() { set_maps(d_full, ntiles_full, nvec_full, nrow_full); }

    // Pack the maps to the instances with live[b] once they hold at most 7/8
    // of the current workgroups: the M passes and the per-row and per-vector
    // kernels then dispatch no workgroups for retired instances.  Call only
    // after a stream synchronisation since the last call (the pinned staging
    // buffer is reused) and launch nothing for instances outside `live` until
    // the next call or use_full().
    void pack_live(const std::vector<char>& live)
    {
        if (exact_maps) {
            set_exact(live);
            return;
        }
        long long w = 0;
        for (int b = 0; b < B; b++) if (live[b]) w += (long long)h[b].ntiles + h[b].nvec + h[b].nrow;
        if (8 * w > 7 * packed_wgs) return;
        int nt, nv, nr;
        write_maps(d_live, live, nt, nv, nr);
        set_maps(d_live, nt, nv, nr);
    }

    // Exact maps (a session, whose instances come and go): the maps cover
    // exactly the instances with keep[b], rewritten into d_live whenever that
    // set changes (pinned staging: at most once between two synchronisations).
    bool exact_maps = false;
    std::vector<char> map_set;
    void set_exact(const std::vector<char>& keep)
    {
        if (keep == map_set) return;
        map_set = keep;
        int nt, nv, nr;
        write_maps(d_live, keep, nt, nv, nr);
        set_maps(d_live, nt, nv, nr);
    }

    // Make d_sub the wave map of the instances with on[b] (rewritten only when
    // the set changes; the pinned staging buffer is reused, so call at most
    // once between two stream synchronisations).
    void sub_rows(const std::vector<char>& on)
    {
        if (on == sub_key) return;
        sub_key = on;
        int cnt = 0;
        for (int b = 0; b < B; b++) if (on[b]) cnt += h[b].nrow;
        nsub = cnt;
        if (!cnt) return;
        if (pageable) pg_sub.resize(2 * (size_t)cnt);
        int* hm = pageable ? pg_sub.data() : (int*)g_sub_host.get((size_t)cnt + 1);
        int* pr = hm;
        for (int b = 0; b < B; b++) {
            if (!on[b]) continue;
            for (int t = 0; t < h[b].nrow; t++) { *pr++ = b; *pr++ = t; }
        }
        CG_HIP(hipMemcpyAsync(d_sub, hm, sizeof(int) * 2 * (size_t)cnt, hipMemcpyHostToDevice, stream));
    }
    void rows_sub(void (*k)(const CgInst*, const int*, const double*), const double* c, size_t lds = 0)
    {
        if (nsub) hipLaunchKernelGGL(k, dim3(nsub), dim3(CG_WG), lds, stream, d_insts, d_sub, c);
    }

    void upload_maps(const std::vector<const double*>& ptrs)
    {
        for (int b = 0; b < B; b++) h[b].scal = d_scal + (size_t)b * CG_NSCAL;
        CG_HIP(hipMemcpyAsync(d_insts, h.data(), sizeof(CgInst) * B, hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_wgpart_init, dim3(B), dim3(CG_WG), 0, stream, d_insts);
        int nt, nv, nr;
        write_maps(d_full, std::vector<char>(B, 1), nt, nv, nr);
        set_maps(d_full, nt, nv, nr);
        if (!ptrs.empty())
            CG_HIP(hipMemcpyAsync(d_ptrs, ptrs.data(), sizeof(void*) * ptrs.size(), hipMemcpyHostToDevice, stream));
        // pageable sources: make sure the copies are done before the vectors die
        CG_HIP(hipStreamSynchronize(stream));
        if (pageable) {
            pg_ctl.assign((size_t)kCtlSlots * B * CG_NCTL, 0.0);
            pg_scal.assign((size_t)B * CG_NSCAL, 0.0);
            hctl = pg_ctl.data();
            hscal = pg_scal.data();
        } else {
            hctl = g_ctl_host.get((size_t)kCtlSlots * B * CG_NCTL);
            hscal = g_scal_host.get((size_t)B * CG_NSCAL);
        }
    }

    // Upload a control block (ctl[b*CG_NCTL + field]) and return its device copy.
    const double* ctl(const std::vector<double>& c)
    {
        if (slot == kCtlSlots) sync();
        double* hs = hctl + (size_t)slot * B * CG_NCTL;
        double* ds = d_ctl + (size_t)slot * B * CG_NCTL;
        std::memcpy(hs, c.data(), sizeof(double) * B * CG_NCTL);
        last_hctl = hs;
        hmap[ds] = hs;
        CG_HIP(hipMemcpyAsync(ds, hs, sizeof(double) * B * CG_NCTL, hipMemcpyHostToDevice, stream));
        slot++;
        return ds;
    }

    // A control block that tends to repeat from one iteration to the next
    // (the active set, the dense set): sticky block `key` is uploaded (staged
    // through the ring like ctl()) only when c differs from what it holds.
    const double* ctl_sticky(int key, const std::vector<double>& c)
    {
        double* dst = d_sticky + (size_t)key * B * CG_NCTL;
        auto& held = sticky_host[key];
        if (held != c) {
            if (slot == kCtlSlots) sync();
            double* hs = hctl + (size_t)slot * B * CG_NCTL;
            std::memcpy(hs, c.data(), sizeof(double) * B * CG_NCTL);
            CG_HIP(hipMemcpyAsync(dst, hs, sizeof(double) * B * CG_NCTL, hipMemcpyHostToDevice, stream));
            slot++;
            held = c;
        }
        hmap[dst] = held.data();
        return dst;
    }

    // Upload a control block to the fixed-address block (staged through the
    // ring like ctl()).
    const double* ctl_fixed(const std::vector<double>& c)
    {
        if (slot == kCtlSlots) sync();
        double* hs = hctl + (size_t)slot * B * CG_NCTL;
        std::memcpy(hs, c.data(), sizeof(double) * B * CG_NCTL);
        hmap[d_fixctl] = hs;
        CG_HIP(hipMemcpyAsync(d_fixctl, hs, sizeof(double) * B * CG_NCTL, hipMemcpyHostToDevice, stream));
        slot++;
        return d_fixctl;
    }

    void sync()
    {
        CG_HIP(hipStreamSynchronize(stream));
        slot = 0;
        if (!g_prof_pending.empty()) prof_collect();
    }

    // Algorithmic bytes of one M pass over the active instances of control c:
    // 8nm for M plus `vec_bytes` per (n+m) entry.
    double mpass_bytes(const double* hc, double vec_bytes) const
    {
        double t = 0.0;
        for (int b = 0; b < B; b++)
            if (hc[(size_t)b * CG_NCTL + CTL_ACTIVE] != 0.0)
                t += 8.0 * (double)h[b].n * (double)h[b].m + vec_bytes * (double)(h[b].n + h[b].m);
        return t;
    }
    const double* last_hctl = nullptr;

    // Read every instance's scalars (one copy + one sync for the batch).
    void read_scal()
    {
        CG_HIP(hipMemcpyAsync(hscal, d_scal, sizeof(double) * B * CG_NSCAL, hipMemcpyDeviceToHost, stream));
        sync();
    }
    double sc(int b, int k) const { return hscal[(size_t)b * CG_NSCAL + k]; }

    // ---- launches ----
    void mpass(void (*k)(const CgInst*, const int*, const double*, int), const double* c, int mode,
               int cls = -1, double vec_bytes = 0.0)
    {
        if (!ntiles) return;
        if (g_prof && cls >= 0) {
            ProfScope ps(stream, cls, mpass_bytes(hmap[c], vec_bytes));
            hipLaunchKernelGGL(k, dim3(ntiles), dim3(CG_WG), 0, stream, d_insts, d_tile, c, mode);
        } else {
            hipLaunchKernelGGL(k, dim3(ntiles), dim3(CG_WG), 0, stream, d_insts, d_tile, c, mode);
        }
    }
    void vec(void (*k)(const CgInst*, const int*, const double*, int), const double* c, int mode)
    {
        if (nvec) hipLaunchKernelGGL(k, dim3(nvec), dim3(CG_WG), 0, stream, d_insts, d_vec, c, mode);
    }
    void vec0(void (*k)(const CgInst*, const int*, const double*), const double* c)
    {
        if (nvec) hipLaunchKernelGGL(k, dim3(nvec), dim3(CG_WG), 0, stream, d_insts, d_vec, c);
    }
    void tiles0(void (*k)(const CgInst*, const int*, const double*), const double* c)
    {
        if (ntiles) hipLaunchKernelGGL(k, dim3(ntiles), dim3(CG_WG), 0, stream, d_insts, d_tile, c);
    }
    void rows0(void (*k)(const CgInst*, const int*, const double*), const double* c, size_t lds = 0)
    {
        if (nrow) hipLaunchKernelGGL(k, dim3(nrow), dim3(CG_WG), lds, stream, d_insts, d_row, c);
    }
    void per_inst(void (*k)(const CgInst*, const double*, int), const double* c, int arg)
    {
        hipLaunchKernelGGL(k, dim3(B), dim3(CG_WG), 0, stream, d_insts, c, arg);
    }
    void per_inst0(void (*k)(const CgInst*, const double*), const double* c)
    {
        hipLaunchKernelGGL(k, dim3(B), dim3(CG_WG), 0, stream, d_insts, c);
    }
    void check_launch() { CG_HIP(hipGetLastError()); }
};

std::vector<double> ctl_vec(int B) { return std::vector<double>((size_t)B * CG_NCTL, 0.0); }

// T pass (M pass + finalize) for the instances marked active in c.
void tpass(Batch& bt, const double* c, int mode)
{
    bt.mpass(bt.xy ? k_tpass<true> : k_tpass<false>, c, mode, PROF_TPASS, mode == CG_PT_TRIAL ? 24.0 : 16.0);
    bt.vec(k_tpass_fin, c, mode);
}

void row_lse(Batch& bt, const double* c, int mode)
{
    bt.mpass(bt.xy ? k_row_lse<true> : k_row_lse<false>, c, mode, PROF_ROWLSE, 16.0);
    bt.vec(k_rowlse_fin, c, mode);
}

void col_lse(Batch& bt, const double* c, int mode)
{
    bt.mpass(bt.xy ? k_col_lse<true> : k_col_lse<false>, c, mode, PROF_COLLSE, 16.0);
    bt.vec(k_collse_fin, c, mode);
}

// ---------------------------------------------------------------------------
// Strong-Wolfe line search as a resumable state machine.
// Reference: SPLRSolver::line_search_wolfe, src/sinkhorn_splr.cu:667-920, with
// the step selection of src/utils.h:76-230 (Moré & Thuente 1994).
// ---------------------------------------------------------------------------
double quad_fg(double a, double b, double fa, double ga, double fb)
{
    const double h = b - a;
    return a + 0.5 * h * h * ga / (fa - fb + h * ga);
}
double quad_gg(double a, double b, double ga, double gb) { return a + ga / (ga - gb) * (b - a); }
double cubic_min(double a, double b, double fa, double fb, double ga, double gb, bool& exists)
{
    const double h = b - a, h2 = h * h, s = a + b;
    const double df = fb - fa, dg = gb - ga;
    const double c3 = (ga + gb) * h - 2.0 * df;
    const double c2 = 0.5 * (dg * h2 - 3.0 * s * c3);
    const double c1 = df * h2 - s * c2 - (a * s + b * b) * c3;
    const double eps = std::numeric_limits<double>::epsilon();
    if (std::fabs(c3) < eps * std::fabs(c2) || std::fabs(c3) < eps * std::fabs(c1)) {
        exists = (c2 * h > 0.0);
        return exists ? (-0.5 * c1 / c2) : b;
    }
    const double u = c2 / (3.0 * c3), v = c1 / c2, ratio = v / u;
    exists = (ratio <= 1.0);
    if (!exists) return b;
    double r1, r2;
    if (std::fabs(u) >= std::fabs(v)) {
        const double w = 1.0 + std::sqrt(1.0 - ratio);
        r1 = -u * w;
        r2 = -v / w;
    } else {
        const double sd = std::sqrt(std::fabs(u)) * std::sqrt(std::fabs(v)) * std::sqrt(1.0 - u / v);
        r1 = -u - sd;
        r2 = -u + sd;
    }
    return (c3 * h > 0.0) ? std::max(r1, r2) : std::min(r1, r2);
}
double step_sel(double al, double au, double at, double fl, double fu, double ft, double gl, double gu, double gt)
{
    if (al == au) return al;
    if (!std::isfinite(ft) || !std::isfinite(gt)) return 0.5 * (al + at);
    bool c_ok;
    const double ac = cubic_min(al, at, fl, ft, gl, gt, c_ok);
    const double aq = quad_fg(al, at, fl, gl, ft);
    if (ft > fl) {
        if (!c_ok) return aq;
        return (std::fabs(ac - al) < std::fabs(aq - al)) ? ac : 0.5 * (aq + ac);
    }
    const double as = quad_gg(al, at, gl, gt);
    if (gt * gl < 0.0) return (std::fabs(ac - at) >= std::fabs(as - at)) ? ac : as;
    if (std::fabs(gt) < std::fabs(gl)) {
        const bool take_c = c_ok && ((ac - at) * (at - al) > 0.0) && (std::fabs(ac - at) < std::fabs(as - at));
        const double r = take_c ? ac : as;
        const double lim = at + 0.66 * (au - at);
        return (at > al) ? std::min(lim, r) : std::max(lim, r);
    }
    if (!std::isfinite(au) || !std::isfinite(fu) || !std::isfinite(gu)) return at + 1.1 * (at - al);
    bool e_ok;
    const double ae = cubic_min(at, au, ft, fu, gt, gu, e_ok);
    const double lim = at + 0.66 * (au - at);
    return (at > al) ? std::min(lim, ae) : std::max(lim, ae);
}

struct Wolfe {
    static constexpr double c1 = 1e-4, c2 = 0.9, step_min = 1e-10, step_max = 2.0;
    static constexpr int max_trials = 20;
    double init_step, step, fx_init, dg0 = 0, dec = 0, curv = 0;
    double lo = 0, hi = INFINITY, f_lo = 0, f_hi = INFINITY, g_lo = 0, g_hi = INFINITY, psi_lo = 0;
    bool bracketed = false, min_guard = true;
    int shrink_fail = 0, trials = 0;
    double width = INFINITY;
    double fx = 0, last_eval = 0;
    bool finished = false, recompute = false;
    double result = 0;

    void begin(double init, double f0)
    {
        init_step = std::min(std::max(init, step_min), step_max);
        step = init_step;
        fx_init = f0;
        fx = f0;
        trials = 0;
        finished = false;
        recompute = false;
    }

    // Feed the trial at `step` (f, <g, d>; dg0_ = <g_prev, d> on the first trial).
    // Afterwards either finished (result, recompute) or `step` is the next trial.
    void feed(double f, double dg, double dg0_)
    {
        if (trials == 0) {
            dg0 = dg0_;
            if (dg0 > 0.0) {   // not a descent direction (reference returns init_step, recompute)
                finished = true;
                recompute = true;
                result = init_step;
                last_eval = step;
                fx = f;
                return;
            }
            dec = c1 * dg0;
            curv = -c2 * dg0;
            g_lo = (1.0 - c1) * dg0;
        }
        trials++;
        fx = f;
        last_eval = step;
        const double psi = f - fx_init - step * dec;
        const double dpsi = dg - dec;
        if ((psi <= 0.0 && std::fabs(dg) <= curv) ||
            (step <= step_min && (psi > 0.0 || dpsi >= 0.0)) ||
            (step >= step_max && (psi <= 0.0 && dpsi < 0.0))) {
            finished = true;
            result = step;
            return;
        }
        const double ft = psi, gt = dpsi;
        if (min_guard && (psi <= 0.0 && dpsi < 0.0)) min_guard = false;
        const bool case2 = (psi <= psi_lo) && (dpsi * (lo - step) > 0.0);
        double nxt;
        if (case2) {
            nxt = std::min(step_max, step + 1.1 * (step - lo));
        } else {
            nxt = step_sel(lo, hi, step, f_lo, f_hi, ft, g_lo, g_hi, gt);
            nxt = std::min(std::max(nxt, step_min), step_max);
            if (min_guard) {
                const double up = std::max(step_min, (7.0 / 12.0) * step);
                nxt = std::min(std::max(nxt, step_min), up);
            }
        }
        if (psi > psi_lo) {
            hi = step; f_hi = ft; g_hi = gt;
        } else if (case2) {
            lo = step; f_lo = ft; g_lo = gt; psi_lo = psi;
        } else {
            hi = lo; f_hi = f_lo; g_hi = g_lo;
            lo = step; f_lo = ft; g_lo = gt; psi_lo = psi;
        }
        if (!bracketed && !case2) {
            const double l = std::min(lo, hi), r = std::max(lo, hi);
            bracketed = (l >= step_min && r <= step_max);
        }
        if (bracketed) {
            const double wprev = width;
            width = std::fabs(hi - lo);
            if (wprev < INFINITY && width > 0.66 * wprev) shrink_fail++;
            else shrink_fail = 0;
            if (shrink_fail >= 2) { nxt = 0.5 * (lo + hi); shrink_fail = 0; }
        }
        step = nxt;
        if (trials >= max_trials) {
            finished = true;
            const double psi2 = fx - fx_init - step * dec;
            if (psi2 <= psi_lo) { result = step; recompute = false; return; }
            recompute = true;
            result = (lo <= 0.0) ? init_step : lo;
        }
    }
};

// Largest dense core of a sparse plan on the current device (what its
// workgroup LDS holds), and parallel-for over the library's analysis pool;
// both defined with the worker pools below.
int sp_rmax();
int g_lds_max = 0;   // the device's LDS per workgroup (set with the kernels' limits in sp_rmax)
std::atomic<int> g_sparse_on{1};   // cgot_set_sparse_solve: 0 dense, 1 by cost estimate, 2 sparse where a plan fits
void parallel_for(int count, const std::function<void(int)>& fn);

// ---------------------------------------------------------------------------
// SPLR batch solver
// ---------------------------------------------------------------------------
struct HostInst {
    int n, m, N;
    double reg, tol;
    int max_iter;
    double dmax, dmin, density, shift_max;
    int cycle, cand;
    long long Te, Kmax, K;
    int P, nblk;
    bool done = false, converged = false, invalid = false;
    bool live = false;      // started (begin) and not yet finished
    bool x0 = false;        // warm start given
    int it = 0;             // iterations taken
    int niter = 0;
    double f = 0, gnorm = 0, gnorm_init = 0, step_prev = 1.0;
    bool analyze = false, update = false;
    Wolfe ls;
    bool in_ls = false;
    double* trace = nullptr;
    // sparse Newton solve: the plan of the current pattern is in use; a new
    // pattern awaits its analysis (its W pattern is being copied to pinned
    // memory at pat_off); core size
    bool sparse = false, pend = false;
    int spR = 0, spH = 0;   // sparse plan: core size, periphery levels
    double sp_flops = 0.0;  // sparse plan: flops of one factorization and solve (sp_solve_flops)
    size_t sp_vals = 0;     // sparse plan: LDS of a launch holding its values (sp_lds_vals_bytes)
    size_t sp_all = 0;      // ... its values and integer arrays (sp_lds_all_bytes)
    size_t pat_off = 0;
};

void resolve_defaults(const cgot_splr_problem& p, double& dmax, double& shift_max, int& cycle, int& cand)
{
    dmax = (p.density_max > 0.0) ? std::min(p.density_max, 1.0) : std::min(10.0 / std::min(p.n, p.m), 1.0);
    dmax = std::min(std::max(dmax, 0.0), 1.0);
    shift_max = (p.shift_max >= 0.0) ? p.shift_max : 1e-3;
    cycle = (p.sparsity_pattern_cycle > 0) ? p.sparsity_pattern_cycle : 30;
    cand = (p.candidate_sinkhorn_iter >= 0) ? p.candidate_sinkhorn_iter : 3;
}

long long kmax_of(double dmax, int n, int m)
{
    long long K = (long long)(dmax * (double)n * (double)(m - 1));
    return std::max(K, 1LL);
}

// Geometry, cost source, Kmax, Schur side and outputs of one SPLR instance.
void inst_geometry(CgInst& I, const cgot_splr_problem& p)
{
    set_geometry(I, p.n, p.m, p.reg);
    double dmax, shmax;
    int cyc, cand;
    resolve_defaults(p, dmax, shmax, cyc, cand);
    I.Kmax = kmax_of(dmax, p.n, p.m);
    I.beta_side = (p.m - 1 <= p.n) ? 1 : 0;
    I.P = I.beta_side ? (p.m - 1) : p.n;
    set_cost(I, p.M, p.xy);
    I.plan = p.plan;
    I.dual = p.dual;
}

// Zero-initialised scratch of one instance: 4 counters, the top-K histogram,
// the backward-substitution block flags.
size_t zero_words(const CgInst& I) { return 4 + CG_HBINS + (size_t)std::max(cdiv(I.P, CG_NB), 1); }
void set_zero_region(CgInst& I, unsigned int* z)
{
    I.counter = z;
    I.hist = z ? I.counter + 4 : nullptr;
    I.flags = z ? (int*)(I.hist + CG_HBINS) : nullptr;
}

// Per-instance arrays (all but the descriptor, the zero region and the scalars).
void inst_arrays(CgInst& I, Bump& bp)
{
    const int n = I.n, m = I.m, N = I.N;
    I.ab = bp.take<double>(n + m);
    I.logab = bp.take<double>(n + m);
    I.gamma = bp.take<double>(n + m);
    I.gamma_prev = bp.take<double>(n + m);
    I.gamma_bcd = bp.take<double>(n + m);
    I.grad = bp.take<double>(N);
    I.grad_prev = bp.take<double>(N);
    I.grad_bcd = bp.take<double>(N);
    I.dir = bp.take<double>(N);
    I.yv = bp.take<double>(N);
    I.sv = bp.take<double>(N);
    I.uv = bp.take<double>(N);
    I.rowsum = bp.take<double>(n);
    I.colsum = bp.take<double>(m);
    I.rp0 = bp.take<double>((size_t)n * I.nstrips);
    I.rp1 = bp.take<double>((size_t)n * I.nstrips);
    I.cp0 = bp.take<double>((size_t)I.nrb * m);
    I.cp1 = bp.take<double>((size_t)I.nrb * m);
    I.wgpart = bp.take<double>((size_t)8 * I.nvec);
    I.w_rowptr = bp.take<int>(n + 1);
    I.w_colj = bp.take<int>(I.Kmax);
    I.w_rval = bp.take<double>(I.Kmax);
    I.w_colptr = bp.take<int>(m);
    I.w_rowi = bp.take<int>(I.Kmax);
    I.w_cval = bp.take<double>(I.Kmax);
    I.Da = bp.take<double>(n);
    I.Db = bp.take<double>(m);
    I.rcnt = bp.take<int>(n);
    I.rtie = bp.take<int>(n);
    I.rcut = bp.take<int>(n);
    I.ccnt = bp.take<int>((size_t)I.nrb * m);
    I.S = bp.take<double>((size_t)I.P * I.P);
    I.sph = bp.take<SpHdr>(1);
    I.Winv = bp.take<double>((size_t)cdiv(I.P, CG_NB) * CG_NB * CG_NB);
    I.rhs = bp.take<double>(2 * (size_t)std::max(I.P, 1));
    I.ltmp = bp.take<double>(2 * (size_t)std::max(n, m));
    I.inv_diag = bp.take<double>(N);
}

// The value refresh runs on the wave map when some instance's pattern may
// hold more than 32 items per vector-map thread (N + 2 Kmax items).
bool wants_refresh_rows(const CgInst& I)
{
    return (double)I.N + 2.0 * (double)I.Kmax > 32.0 * CG_WG * (double)I.nvec;
}

// Lay out the workspace of an SPLR batch; with base == nullptr only sizes.
size_t splr_layout(Batch& bt, const cgot_splr_problem* probs, void* base)
{
    const int B = bt.B;
    bt.h.assign(B, CgInst{});
    int nblk_max = 0;
    for (int b = 0; b < B; b++) {
        inst_geometry(bt.h[b], probs[b]);
        nblk_max = std::max(nblk_max, cdiv(bt.h[b].P, CG_NB));
    }
    bt.assign_maps();
    bt.refresh_rows = false;
    for (const CgInst& I : bt.h)
        if (wants_refresh_rows(I)) bt.refresh_rows = true;
    Bump bp(base);
    bt.layout_common(bp, nblk_max);
    // scratch that must start at zero, contiguous for one clear
    const size_t zstride = 4 + CG_HBINS + (size_t)std::max(nblk_max, 1);
    bt.zero_base = bp.take<unsigned int>((size_t)B * zstride);
    bt.zero_bytes = (size_t)B * zstride * sizeof(unsigned int);
    for (int b = 0; b < B; b++) {
        CgInst& I = bt.h[b];
        set_zero_region(I, bt.zero_base ? bt.zero_base + (size_t)b * zstride : nullptr);
        inst_arrays(I, bp);
    }
    return bp.off + 256;
}

static_assert(sizeof(SpHdr) <= SP_HDR_PAD, "plan header larger than its staging slot");
thread_local Pinned g_pat_host(true);    // W patterns of new sparsity patterns (ints), written by k_pat_export
thread_local Pinned g_plan_host(true);   // sparse plans and headers, read by k_plan_import

struct SplrSolver {
    Batch& bt;
    std::vector<HostInst> hi;
    hipEvent_t pat_ev = nullptr;   // the pattern copies to g_pat_host are complete


    SplrSolver(Batch& b) : bt(b) {}
    ~SplrSolver()
    {
        if (pat_ev) (void)hipEventDestroy(pat_ev);

    }

    // After a new top-K pattern (instances in newpat): W's row pattern
    // (rowptr [n+1], colj [K]) goes to pinned host memory (one export launch);
    // the analysis waits for it just before the next Newton step, so the
    // launches queued in between (the Sinkhorn candidate) run while the host
    // analyses.
    void request_patterns(const std::vector<char>& newpat)
    {
        const int B = bt.B;
        size_t total = 0;
        auto c = ctl_vec(B);
        for (int b = 0; b < B; b++)
            if (newpat[b]) {
                hi[b].pat_off = total;
                total += (size_t)hi[b].n + 1 + (size_t)hi[b].K;
                c[(size_t)b * CG_NCTL + CTL_ACTIVE] = 1.0;
                c[(size_t)b * CG_NCTL + CTL_AUX] = (double)hi[b].pat_off;
                c[(size_t)b * CG_NCTL + CTL_KSEL] = (double)hi[b].K;
                hi[b].pend = true;
            }
        if (total == 0) return;
        g_pat_host.get((total + 1) / 2 + 1);
        hipLaunchKernelGGL(k_pat_export, dim3(B), dim3(CG_WG), 0, bt.stream, bt.d_insts, bt.ctl(c),
                           (int*)g_pat_host.dev());
        if (!pat_ev) CG_HIP(hipEventCreateWithFlags(&pat_ev, hipEventDisableTiming));
        CG_HIP(hipEventRecord(pat_ev, bt.stream));
    }

    // Analyse the pending patterns (host, in parallel) and upload each plan
    // into its instance's Schur buffer with its header (mode 0: the instance
    // keeps the dense Schur path for this pattern).
    void prepare_plans()
    {
        const int B = bt.B;
        std::vector<int> todo;
        for (int b = 0; b < B; b++)
            if (hi[b].pend) todo.push_back(b);
        if (todo.empty()) return;
        CG_HIP(hipEventSynchronize(pat_ev));
        const int* hp = (const int*)g_pat_host.p;
        const int rmax = sp_rmax();
        const int mode = g_sparse_on.load();
        std::vector<SpHdr> hdr(todo.size());
        std::vector<std::vector<char>> img(todo.size());
        std::vector<char> ok(todo.size(), 0);
        // the exported patterns must be those of the selection (row pointers
        // from 0 to K, non-decreasing, columns below m - 1) before the
        // analysis indexes with them
        for (int b : todo) {
            const HostInst& H = hi[b];
            const int* rp = hp + H.pat_off;
            const int* cj = rp + H.n + 1;
            bool good = rp[0] == 0 && rp[H.n] == H.K;
            for (int i = 0; good && i < H.n; i++) good = rp[i] <= rp[i + 1];
            for (long long q = 0; good && q < H.K; q++) good = cj[q] >= 0 && cj[q] < H.m - 1;
            if (!good) throw std::runtime_error("internal: inconsistent sparsity pattern read back for the analysis");
        }
        parallel_for((int)todo.size(), [&](int k) {
            const int b = todo[k];
            const HostInst& H = hi[b];
            const int* rp = hp + H.pat_off;
            const int64_t cap = (int64_t)sizeof(double) * bt.h[b].P * bt.h[b].P;
            ok[k] = mode != 0 && sp_analyze(H.n, H.m, rp, rp + H.n + 1, cap, rmax, hdr[k], img[k]) ? 1 : 0;
        });
        // mode 1: the sparse plan where its estimate beats the instance's own
        // dense chain; the choice depends on the instance alone, so results
        // stay independent of the rest of the batch
        if (mode == 1)
            for (size_t k = 0; k < todo.size(); k++)
                if (ok[k] && sp_cost_us(hdr[k]) > sp_dense_cost_us(hi[todo[k]].nblk)) ok[k] = 0;
        for (size_t k = 0; k < todo.size(); k++)
            if (!ok[k]) { hdr[k] = SpHdr{}; img[k].clear(); }
        // previous imports from g_plan_host have completed: every iteration
        // synchronises the stream (line-search read-backs) before the next pattern
        size_t total = 0;
        std::vector<size_t> off(todo.size());
        for (size_t k = 0; k < todo.size(); k++) {
            off[k] = total;
            total += (SP_HDR_PAD + img[k].size() + 255) & ~(size_t)255;
        }
        char* up = (char*)g_plan_host.get(total / sizeof(double) + 1);
        auto c = ctl_vec(B);
        for (size_t k = 0; k < todo.size(); k++) {
            const int b = todo[k];
            char* dst = up + off[k];
            std::memcpy(dst, &hdr[k], sizeof(SpHdr));
            if (ok[k]) std::memcpy(dst + SP_HDR_PAD, img[k].data(), img[k].size());
            c[(size_t)b * CG_NCTL + CTL_ACTIVE] = 1.0;
            c[(size_t)b * CG_NCTL + CTL_AUX] = (double)off[k];
            hi[b].sparse = ok[k] != 0;
            hi[b].spR = ok[k] ? hdr[k].R : 0;
            hi[b].spH = ok[k] ? hdr[k].H : 0;
            hi[b].sp_flops = ok[k] ? sp_solve_flops(hdr[k]) : 0.0;
            hi[b].sp_vals = ok[k] ? sp_lds_vals_bytes(hdr[k]) : 0;
            hi[b].sp_all = ok[k] ? sp_lds_all_bytes(hdr[k]) : 0;
            hi[b].pend = false;
        }
        hipLaunchKernelGGL(k_plan_import, dim3(B * SP_IMPORT_WGS), dim3(CG_WG), 0, bt.stream, bt.d_insts, bt.ctl(c),
                           (const char*)g_plan_host.dev());
    }

    std::vector<double> mask(const std::vector<char>& on)
    {
        auto c = ctl_vec(bt.B);
        for (int b = 0; b < bt.B; b++) c[(size_t)b * CG_NCTL + CTL_ACTIVE] = on[b] ? 1.0 : 0.0;
        return c;
    }

    // Zero-initialised device scratch that relies on it (counters, histograms, pivots).
    void zero_scratch()
    {
        CG_HIP(hipMemsetAsync(bt.zero_base, 0, bt.zero_bytes, bt.stream));
        CG_HIP(hipMemsetAsync(bt.d_scal, 0, sizeof(double) * bt.B * CG_NSCAL, bt.stream));
    }

    // New top-K pattern (on = instances whose pattern is rebuilt) and value
    // refresh (on2 = instances whose A is rebuilt), at the current gamma.
    void sphess(const std::vector<char>& newpat, const std::vector<char>& refresh)
    {
        const int B = bt.B;
        bool anynew = false;
        auto c = ctl_vec(B);
        for (int b = 0; b < B; b++) {
            const HostInst& H = hi[b];
            const bool on = newpat[b] && H.Te > 0;
            c[(size_t)b * CG_NCTL + CTL_ACTIVE] = on ? 1.0 : 0.0;
            if (on) {
                double density = std::min(std::max(H.density, 0.0), 1.0);
                long long K = (long long)(density * (double)H.Te);
                K = std::min(K, H.Kmax);
                K = std::max(K, 1LL);
                K = std::min(K, H.Te);
                hi[b].K = K;
                c[(size_t)b * CG_NCTL + CTL_KSEL] = (double)K;
                anynew = true;
            }
        }
        if (anynew) {
            const double* dc = bt.ctl(c);
            for (int pass = 0; pass < 5; pass++) {
                bt.mpass(bt.xy ? k_hist<true> : k_hist<false>, dc, pass);
                bt.per_inst(k_hist_scan, dc, pass);
            }
            bt.rows0(k_sel_rows_count, dc);
            bt.per_inst0(k_sel_rows_scan, dc);
            bt.rows0(k_sel_rows_write, dc);
            bt.tiles0(bt.xy ? k_sel_cols_count<true> : k_sel_cols_count<false>, dc);
            bt.vec0(k_sel_cols_prefix, dc);
            bt.per_inst0(k_sel_cols_scan, dc);
            bt.tiles0(bt.xy ? k_sel_cols_write<true> : k_sel_cols_write<false>, dc);
        }
        // instances without columns (m == 1): empty pattern
        for (int b = 0; b < B; b++) {
            if (newpat[b] && hi[b].Te == 0) {
                hi[b].K = 0;
                CG_HIP(hipMemsetAsync(bt.h[b].w_rowptr, 0, sizeof(int) * (bt.h[b].n + 1), bt.stream));
                CG_HIP(hipMemsetAsync(bt.h[b].w_colptr, 0, sizeof(int) * bt.h[b].m, bt.stream));
            }
        }
        auto r = ctl_vec(B);
        bool any = false;
        for (int b = 0; b < B; b++) {
            if (!refresh[b]) continue;
            any = true;
            r[(size_t)b * CG_NCTL + CTL_ACTIVE] = 1.0;
            const double shift = std::min(hi[b].gnorm, hi[b].shift_max);
            r[(size_t)b * CG_NCTL + CTL_SHIFT] = shift * hi[b].reg;
        }
        if (any) {
            const double* dr = bt.ctl(r);
            if (bt.refresh_rows) {
                if (bt.nrow) hipLaunchKernelGGL(k_refresh_values, dim3(bt.nrow), dim3(CG_WG), 0, bt.stream, bt.d_insts,
                                                bt.d_row, dr, 1);
            } else if (bt.nvec) {
                hipLaunchKernelGGL(k_refresh_values, dim3(bt.nvec), dim3(CG_WG), 0, bt.stream, bt.d_insts, bt.d_vec,
                                   dr, 0);
            }
        }
        request_patterns(newpat);
    }

    // Search direction for the instances in `on`: the sparse factor/solve for
    // instances whose current pattern has a plan, the dense Schur path for the
    // others.  Both write dir = A^{-1} grad and uv = A^{-1} yv.
    void search_direction(const std::vector<char>& on_all, const double*)
    {
        const int B = bt.B;
        prepare_plans();
        std::vector<char> on(B, 0);
        std::vector<int> ids;
        int rmax_used = 0;
        size_t vals_max = 0, all_max = 0;
        double sp_work = 0.0;
        for (int b = 0; b < B; b++) {
            if (!on_all[b]) continue;
            if (hi[b].sparse) {
                ids.push_back(b);
                sp_work += hi[b].sp_flops;
                rmax_used = std::max(rmax_used, hi[b].spR);
                vals_max = std::max(vals_max, hi[b].sp_vals);
                all_max = std::max(all_max, hi[b].sp_all);
            } else {
                on[b] = 1;
            }
        }
        bool anydense = false;
        for (int b = 0; b < B; b++) anydense = anydense || on[b];
        // the plans' values in LDS when every listed plan's fit, and the
        // integer arrays of each plan that fits with them (the same
        // arithmetic: only where they live).  Not when a plan has a panel
        // core (R > SP_SMALL_MAX): that combination has faulted, cause not
        // found, so such launches keep every plan in global memory.
        const int lv = (!ids.empty() && vals_max <= (size_t)g_lds_max && rmax_used <= SP_SMALL_MAX) ? 1 : 0;
        const size_t sp_lds = lv ? std::max(vals_max, all_max <= (size_t)g_lds_max ? all_max : 0) : sp_lds_bytes(rmax_used);
        // both kinds in the group: the sparse blocks share one launch with the
        // Schur rows (k_newton_mixed) instead of running before the dense
        // chain; the library profiler keeps them apart to time each
        const bool mixed = !ids.empty() && anydense && !g_prof;
        if (!ids.empty()) {
            // ids staged through a pinned buffer: the previous copy out of it
            // completed at the last iteration's read-back
            // (skipped when the list is the one the device holds)
            if (ids != bt.ids_held) {
                static thread_local Pinned pinned_ids;
                int* hid = (int*)pinned_ids.get(ids.size() / 2 + 1);
                std::memcpy(hid, ids.data(), sizeof(int) * ids.size());
                CG_HIP(hipMemcpyAsync(bt.d_ids, hid, sizeof(int) * ids.size(), hipMemcpyHostToDevice, bt.stream));
                bt.ids_held = ids;
            }
            if (!mixed) {
                ProfScope ps(bt.stream, PROF_SPNEWTON, sp_work);
                hipLaunchKernelGGL(lv ? k_sp_newton<1> : k_sp_newton<0>, dim3((unsigned)ids.size()), dim3(SP_WG),
                                   sp_lds, bt.stream, bt.d_insts, bt.d_ids, (int)sp_lds);
            }
        }
        if (!anydense) return;
        const double* dc = bt.ctl_sticky(1, mask(on));
        int nblk_max = 0;
        for (int b = 0; b < B; b++)
            if (on[b]) nblk_max = std::max(nblk_max, hi[b].nblk);
        // dense prefix tables: fill | panel[kb] | update launch kb in three
        // segments: diagonal tile[kb] | generic tiles[kb] | next-panel tiles[kb]
        // | diag0, then the backward-substitution order (instances by
        // decreasing block count) and its level prefix (level t: instances with
        // more than t blocks)
        const size_t lvl_off = (size_t)(4 * nblk_max + 3) * (B + 1);
        std::vector<int> pref(lvl_off + nblk_max + 1, 0);
        auto tab = [&](int t) { return pref.data() + (size_t)t * (B + 1); };
        for (int b = 0; b < B; b++) {
            const int nb = on[b] ? hi[b].nblk : 0;
            tab(0)[b + 1] = nb * (nb + 1) / 2;
            for (int kb = 0; kb < nblk_max; kb++) {
                const int T = nb - kb - 1;
                tab(1 + kb)[b + 1] = (kb < nb) ? (nb - kb - 1) : 0;
                tab(1 + nblk_max + kb)[b + 1] = (T > 0) ? 1 : 0;
                tab(1 + 2 * nblk_max + kb)[b + 1] = (T > 0) ? T * (T - 1) / 2 : 0;
                tab(1 + 3 * nblk_max + kb)[b + 1] = (T > 0) ? T - 1 : 0;
            }
            tab(1 + 4 * nblk_max)[b + 1] = (nb > 0) ? 1 : 0;
        }
        for (int t = 0; t < 4 * nblk_max + 2; t++)
            for (int b = 0; b < B; b++) tab(t)[b + 1] += tab(t)[b];
        int nact = 0;
        {
            int* ord = tab(2 + 4 * nblk_max);
            for (int b = 0; b < B; b++)
                if (on[b] && hi[b].nblk > 0) ord[nact++] = b;
            std::stable_sort(ord, ord + nact, [&](int x, int y) { return hi[x].nblk > hi[y].nblk; });
            int* lp = pref.data() + lvl_off;
            for (int t = 0; t < nblk_max; t++) {
                int cnt = 0;
                while (cnt < nact && hi[ord[cnt]].nblk > t) cnt++;
                lp[t + 1] = lp[t] + cnt;
            }
        }
        // The table is staged through a pinned buffer.  Every iteration reads the
        // scalars back (a stream sync) before the next search direction, so the
        // previous copy out of this buffer has completed when it is rewritten.
        static thread_local Pinned pinned_pref;
        if (nblk_max > 0 && pref != bt.pref_held) {
            if ((int)pref.size() > bt.pref_cap) throw std::runtime_error("internal: prefix table overflow");
            int* hp = (int*)pinned_pref.get((pref.size() + 1) / 2 + 1);
            std::memcpy(hp, pref.data(), pref.size() * sizeof(int));
            CG_HIP(hipMemcpyAsync(bt.d_pref, hp, pref.size() * sizeof(int), hipMemcpyHostToDevice, bt.stream));
            bt.pref_held = pref;
        }
        const int* dtab0 = bt.d_pref;
        auto dtab = [&](int t) { return dtab0 + (size_t)t * (B + 1); };
        auto total = [&](int t) { return tab(t)[B]; };
        if (nblk_max > 0 && total(0) > 0)
            hipLaunchKernelGGL(k_chol_fill, dim3(total(0)), dim3(CG_WG), 0, bt.stream, bt.d_insts, B, dtab(0));
        // LDS rows of the Schur build: CG_WAVES rows of the largest P that fits
        int pcap = 1;
        for (int b = 0; b < B; b++)
            if (on[b] && bt.h[b].P <= CG_SCHUR_LDS) pcap = std::max(pcap, bt.h[b].P);
        bt.sub_rows(on);
        if (mixed) {
            const size_t lds = std::max(sp_lds, sizeof(double) * (SP_WG / 64) * (size_t)pcap);
            const int per = SP_WG / CG_WG;   // wave-map entries per block
            const int nb = (int)ids.size() + cdiv(2LL * bt.nsub, per);
            hipLaunchKernelGGL(lv ? k_newton_mixed<1> : k_newton_mixed<0>, dim3(nb), dim3(SP_WG), lds,
                               bt.stream, bt.d_insts, bt.d_ids,
                               (int)ids.size(), bt.d_sub, bt.nsub, dc, (int)lds);
        } else {
            bt.rows_sub(k_schur_build, dc, sizeof(double) * CG_WAVES * (size_t)pcap);
            bt.rows_sub(k_schur_rhs, dc);
        }
        if (nblk_max > 0 && total(1 + 4 * nblk_max) > 0)
            hipLaunchKernelGGL(k_chol_diag0, dim3(total(1 + 4 * nblk_max)), dim3(CG_WG), 0, bt.stream, bt.d_insts, B,
                               dtab(1 + 4 * nblk_max));
        for (int kb = 0; kb < nblk_max; kb++) {
            // panel of block 0 on its own; later panels are formed by the
            // previous step's update launch
            const int tp = (kb == 0) ? total(1 + kb) : 0;
            if (tp > 0) {
                ProfScope ps(bt.stream, PROF_CHOL_PANEL, 0.0);
                hipLaunchKernelGGL(k_chol_panel, dim3(tp), dim3(CG_WG), 0, bt.stream, bt.d_insts, B, dtab(1 + kb), kb);
            }
            // diagonal tiles first (their factors start at once), then the
            // generic tiles, then the next-panel tiles (they wait for the
            // factor of their instance's diagonal tile)
            const int tu = total(1 + nblk_max + kb) + total(1 + 2 * nblk_max + kb) + total(1 + 3 * nblk_max + kb);
            if (tu > 0) {
                // algorithmic flops: 2*64^3 per full lower tile (diagonal tiles count half)
                double fl = 0.0;
                for (int b = 0; b < B; b++) {
                    const int nb = on[b] ? hi[b].nblk : 0;
                    const int T = nb - kb - 1;
                    if (T > 0) fl += 2.0 * CG_NB * CG_NB * CG_NB * (T * (T - 1) / 2 + 0.5 * T);
                }
                ProfScope ps(bt.stream, PROF_CHOL_UPDATE, fl);
                hipLaunchKernelGGL(k_chol_update, dim3(tu), dim3(CG_WG), 0, bt.stream, bt.d_insts, B,
                                   dtab(1 + nblk_max + kb), nblk_max * (B + 1), kb);
            }
        }
        const int ntrsv = (nblk_max > 0) ? pref[lvl_off + nblk_max] : 0;
        if (ntrsv > 0) {
            ProfScope ps(bt.stream, PROF_TRSV, 0.0);
            hipLaunchKernelGGL(k_trsv_flow, dim3(ntrsv), dim3(CG_WG), 0, bt.stream, bt.d_insts,
                               dtab(2 + 4 * nblk_max), dtab0 + lvl_off, nblk_max);
        }
        bt.rows_sub(k_schur_back, dc);
    }

    void woodbury(const double* dc)
    {
        bt.vec0(k_woodbury_dots, dc);
        bt.vec0(k_woodbury_apply, dc);
    }

    // Host state of instance b from its problem (its iteration counter at 0).
    void init_inst(int b, const cgot_splr_problem& p)
    {
        HostInst& H = hi[b];
        H = HostInst{};
        H.n = p.n; H.m = p.m; H.N = p.n + p.m - 1;
        H.reg = p.reg; H.tol = p.tol; H.max_iter = p.max_iter;
        resolve_defaults(p, H.dmax, H.shift_max, H.cycle, H.cand);
        H.dmin = 0.01 * H.dmax;
        H.density = 0.1 * H.dmax;
        H.Te = (long long)p.n * (p.m - 1);
        H.Kmax = kmax_of(H.dmax, p.n, p.m);
        H.K = 0;
        H.P = bt.h[b].P;
        H.nblk = cdiv(H.P, CG_NB);
        H.trace = p.trace;
        H.x0 = p.x0 != nullptr;
    }

    // Start the instances in `jn` (zeroed scratch, descriptors and input
    // pointers on the device): ab, log ab, gamma = x0 (gauge shifted) or
    // beta = 0 and alpha = row LSE, the first T pass, the first pattern.
    void begin(const std::vector<char>& jn)
    {
        const int B = bt.B;
        // patterns still awaiting their analysis are consumed first: the
        // starting instances' patterns reuse the pinned staging
        prepare_plans();
        const double* dj = bt.ctl(mask(jn));
        hipLaunchKernelGGL(k_init_logab, dim3(bt.nvec), dim3(CG_WG), 0, bt.stream, bt.d_insts, bt.d_vec, dj,
                           (const double* const*)bt.d_ptrs, (const double* const*)nullptr);
        {
            std::vector<char> wx(B, 0), nox(B, 0);
            for (int b = 0; b < B; b++)
                if (jn[b]) (hi[b].x0 ? wx : nox)[b] = 1;
            if (std::count(wx.begin(), wx.end(), 1)) bt.per_inst(k_gauge_shift, bt.ctl(mask(wx)), CG_PT_GAMMA);
            if (std::count(nox.begin(), nox.end(), 1)) row_lse(bt, bt.ctl(mask(nox)), CG_PT_GAMMA);
        }
        tpass(bt, dj, CG_PT_GAMMA);
        bt.check_launch();
        bt.read_scal();
        for (int b = 0; b < B; b++) {
            if (!jn[b]) continue;
            HostInst& H = hi[b];
            H.f = bt.sc(b, CS_F);
            H.gnorm = std::sqrt(bt.sc(b, CS_G2));
            H.gnorm_init = H.gnorm;
            H.live = true;
        }
        sphess(jn, jn);
    }

    void run(const cgot_splr_problem* probs, cgot_splr_info* info)
    {
        const int B = bt.B;
        hi.assign(B, HostInst{});
        for (int b = 0; b < B; b++) init_inst(b, probs[b]);
        zero_scratch();
        std::vector<char> all(B, 1);
        begin(all);
        while (step()) {
        }
        bt.use_full();
        finish(all, info);
    }

    // One SPLR iteration of every live instance that has not stopped; each
    // instance counts its own iterations.  Returns false (and does nothing)
    // when none is left iterating.
    bool step()
    {
        const int B = bt.B;
        {
            // loop-entry tests
            std::vector<char> act(B, 0);
            int nact = 0;
            for (int b = 0; b < B; b++) {
                HostInst& H = hi[b];
                if (!H.live || H.done) continue;
                if (H.it >= H.max_iter) { H.done = true; continue; }
                if (H.gnorm < H.tol) { H.converged = true; H.done = true; continue; }
                if (!std::isfinite(H.f)) { H.done = true; continue; }
                H.analyze = (H.it % H.cycle == 0);
                H.update = (H.it % H.cycle == H.cycle - 1);
                act[b] = 1;
                nact++;
            }
            if (nact == 0) return false;
            bt.pack_live(act);
            auto c = mask(act);
            for (int b = 0; b < B; b++) c[(size_t)b * CG_NCTL + CTL_ALLOW_LR] = hi[b].it > 0 ? 1.0 : 0.0;
            const double* dc = bt.ctl_sticky(0, c);
            bt.vec0(k_low_rank, dc);
            // Sinkhorn candidate (3 BCD sweeps from the current beta)
            std::vector<char> cand(B, 0);
            int ncand = 0;
            for (int b = 0; b < B; b++)
                if (act[b] && hi[b].analyze && hi[b].cand != 0) { cand[b] = 1; ncand++; }
            if (ncand) {
                const double* cc = bt.ctl(mask(cand));
                bt.vec(k_vcopy, cc, 1);
                row_lse(bt, cc, CG_PT_BCD);
                int maxc = 0;
                for (int b = 0; b < B; b++) if (cand[b]) maxc = std::max(maxc, hi[b].cand);
                for (int s = 0; s < maxc; s++) {
                    std::vector<char> cs(B, 0);
                    for (int b = 0; b < B; b++) cs[b] = cand[b] && s < hi[b].cand;
                    const double* c2 = bt.ctl(mask(cs));
                    col_lse(bt, c2, CG_PT_BCD);
                    row_lse(bt, c2, CG_PT_BCD);
                }
                bt.per_inst(k_gauge_shift, cc, CG_PT_BCD);
                tpass(bt, cc, CG_PT_BCD);
            }
            search_direction(act, dc);
            woodbury(dc);   // also saves gamma and grad as the next secant pair's base
            // line search rounds
            for (int b = 0; b < B; b++)
                if (act[b]) {
                    hi[b].ls = Wolfe{};
                    hi[b].ls.begin(std::min(1.0, 1.5 * hi[b].step_prev), hi[b].f);
                    hi[b].in_ls = true;
                }
            bool first = true;
            std::vector<double> g2(B, 0.0), fbcd(B, 0.0), g2bcd(B, 0.0);
            while (true) {
                auto t = ctl_vec(B);
                int ntr = 0;
                for (int b = 0; b < B; b++)
                    if (act[b] && hi[b].in_ls) {
                        t[(size_t)b * CG_NCTL + CTL_ACTIVE] = 1.0;
                        t[(size_t)b * CG_NCTL + CTL_STEP] = hi[b].ls.step;
                        ntr++;
                    }
                if (ntr == 0) break;
                tpass(bt, bt.ctl(t), CG_PT_TRIAL);
                bt.check_launch();
                bt.read_scal();
                for (int b = 0; b < B; b++) {
                    if (!(act[b] && hi[b].in_ls)) continue;
                    if (first && bt.sc(b, CS_PIVOT) >= 2.0)
                        throw std::runtime_error("internal: a block-flag wait of the backward substitution "
                                                 "exceeded its bound");
                    if (first && cand[b]) { fbcd[b] = bt.sc(b, CS_FBCD); g2bcd[b] = bt.sc(b, CS_G2BCD); }
                    g2[b] = bt.sc(b, CS_G2);
                    hi[b].ls.feed(bt.sc(b, CS_F), bt.sc(b, CS_DG), bt.sc(b, CS_DG0));
                    if (hi[b].ls.finished) hi[b].in_ls = false;
                }
                first = false;
            }
            // recompute where the returned step is not the last evaluated point
            {
                auto t = ctl_vec(B);
                int nre = 0;
                for (int b = 0; b < B; b++) {
                    if (!act[b]) continue;
                    const Wolfe& L = hi[b].ls;
                    if (L.recompute && L.result != L.last_eval) {
                        t[(size_t)b * CG_NCTL + CTL_ACTIVE] = 1.0;
                        t[(size_t)b * CG_NCTL + CTL_STEP] = L.result;
                        nre++;
                    }
                }
                if (nre) {
                    tpass(bt, bt.ctl(t), CG_PT_TRIAL);
                    bt.read_scal();
                    for (int b = 0; b < B; b++)
                        if (t[(size_t)b * CG_NCTL + CTL_ACTIVE] != 0.0) {
                            hi[b].ls.fx = bt.sc(b, CS_F);
                            g2[b] = bt.sc(b, CS_G2);
                        }
                }
            }
            std::vector<char> newpat(B, 0), refresh(B, 0), use_bcd(B, 0);
            int nbcd = 0;
            for (int b = 0; b < B; b++) {
                if (!act[b]) continue;
                HostInst& H = hi[b];
                H.step_prev = H.ls.result;
                const double gnorm_pre = H.gnorm;
                H.f = H.ls.fx;
                H.gnorm = std::sqrt(g2[b]);
                if (cand[b]) {
                    const double gb = std::sqrt(g2bcd[b]);
                    if (fbcd[b] < H.f && gb < 10.0 * H.gnorm) {
                        use_bcd[b] = 1;
                        nbcd++;
                        H.f = fbcd[b];
                        H.gnorm = gb;
                    }
                }
                if (H.update) {
                    const bool bad = (gnorm_pre < H.gnorm_init) && (H.gnorm > gnorm_pre);
                    H.density *= bad ? 1.1 : 0.99;
                    H.density = std::min(H.dmax, std::max(H.dmin, H.density));
                }
                if (H.trace) {
                    double* tr = H.trace + 8 * (size_t)H.it;
                    tr[0] = H.f; tr[1] = H.gnorm; tr[2] = H.ls.result; tr[3] = bt.sc(b, CS_LOWRANK);
                    tr[4] = use_bcd[b]; tr[5] = (double)H.K; tr[6] = H.ls.trials;
                    tr[7] = (H.ls.recompute && H.ls.result != H.ls.last_eval) ? 1.0 : 0.0;
                }
                const bool conv_after = std::isfinite(H.gnorm) && (H.gnorm < H.tol);
                const bool max_reached = (H.it + 1 >= H.max_iter);
                const bool inv_after = !std::isfinite(H.f) || !std::isfinite(H.gnorm);
                H.niter = H.it + 1;
                if (conv_after || max_reached || inv_after) {
                    H.done = true;
                    if (inv_after) H.invalid = true;
                    if (!max_reached && conv_after) H.converged = true;
                } else {
                    refresh[b] = 1;
                    newpat[b] = H.update ? 1 : 0;
                }
            }
            if (nbcd) {
                const double* cb = bt.ctl(mask(use_bcd));
                bt.vec0(k_copy_bcd, cb);
                tpass(bt, cb, CG_PT_GAMMA);   // row/col sums and grad at the accepted point
            }
            sphess(newpat, refresh);
            bt.check_launch();
            for (int b = 0; b < B; b++)
                if (act[b]) hi[b].it++;
        }
        return true;
    }

    // Outputs of the instances in `which` (the maps must cover them): the
    // gauge-fixed dual, the plan, the context's dual copy, and their infos.
    void finish(const std::vector<char>& which, cgot_splr_info* info)
    {
        const int B = bt.B;
        const double* dw = bt.ctl(mask(which));
        bt.vec0(k_output_dual, dw);
        bt.tiles0(k_plan, dw);
        bt.vec(k_vcopy, dw, 2);
        bt.check_launch();
        bt.read_scal();
        for (int b = 0; b < B; b++) {
            if (!which[b]) continue;
            const HostInst& H = hi[b];
            cgot_splr_info& o = info[b];
            const CgInst& I = bt.h[b];
            o.niter = H.niter;
            o.converged = H.converged ? 1 : 0;
            o.valid = (H.invalid || !std::isfinite(H.f)) ? 0 : 1;
            o.fallback = (H.niter == 0) ? 1 : 0;
            o.low_rank = (H.niter > 0 && bt.sc(b, CS_LOWRANK) != 0.0) ? 1 : 0;
            o.ys = (H.niter > 0) ? bt.sc(b, CS_YS) : 0.0;
            o.n = H.n;
            o.m = H.m;
            o.f = H.f;
            o.gnorm = H.gnorm;
            o.K = H.K;
            o.inv_diag = I.inv_diag;
            o.s = I.sv;
            o.y = I.yv;
            o.u = I.uv;
            o.gamma = I.gamma;
            o.factor = (H.niter > 0) ? (const void*)(bt.d_insts + b) : nullptr;
        }
    }
};

void check_problem(const cgot_splr_problem& p)
{
    if (p.n < 1 || p.m < 1) throw std::invalid_argument("n and m must be >= 1");
    if (!p.a || !p.b) throw std::invalid_argument("a and b must be non-null device pointers");
    check_cost(p.M, p.xy);
    if (!(p.reg > 0.0) || !std::isfinite(p.reg)) throw std::invalid_argument("reg must be positive and finite");
    if ((long long)p.n * (long long)p.m >= (1LL << 31) || (long long)p.n + p.m >= (1LL << 28))
        throw std::invalid_argument("instance too large for 32-bit indices");
}


// ---------------------------------------------------------------------------
// Adjoint (backward) batch: PCG on Q(T*) lambda = mu, T* recomputed from duals.
// ---------------------------------------------------------------------------
// precond_only: the buffers of a preconditioner application alone (no
// M-pass partial sums).
size_t adjoint_layout(Batch& bt, const cgot_adjoint_problem* probs, void* base, bool precond_only = false)
{
    const int B = bt.B;
    bt.h.assign(B, CgInst{});
    bt.pc_sparse.assign(B, 0);
    bt.pc_R.assign(B, 0);
    int nblk_max = 0;
    for (int b = 0; b < B; b++) {
        const auto& p = probs[b];
        CgInst& I = bt.h[b];
        set_geometry(I, p.n, p.m, p.reg);
        set_cost(I, p.M, p.xy);
        I.gamma = const_cast<double*>(p.dual);
        I.dir = p.lambda;
        I.Da = const_cast<double*>(p.inv_diag);
        I.fctx = (unsigned long long)(uintptr_t)p.factor;
        I.P = 0;
        if (p.factor) {
            // the forward descriptor (Schur size and orientation for the launch
            // geometry) and its plan header: a sparse factor is applied by
            // k_sp_precond, a dense one by the Schur kernels
            CgInst F;
            CG_HIP(hipMemcpy(&F, p.factor, sizeof(CgInst), hipMemcpyDeviceToHost));
            if (F.n != p.n || F.m != p.m) throw std::invalid_argument("factor does not match the problem shape");
            SpHdr fh{};
            if (F.sph) CG_HIP(hipMemcpy(&fh, F.sph, sizeof(SpHdr), hipMemcpyDeviceToHost));
            if (fh.mode == 1) {
                I.fsp = 1;
                bt.pc_sparse[b] = 1;
                bt.pc_R[b] = fh.R;
            } else {
                I.P = F.P;
                I.beta_side = F.beta_side;
                nblk_max = std::max(nblk_max, cdiv(F.P, CG_NB));
            }
        }
    }
    bt.assign_maps();
    bt.refresh_rows = false;
    for (const CgInst& I : bt.h)
        if ((double)I.N + 2.0 * (double)I.Kmax > 32.0 * CG_WG * (double)I.nvec) bt.refresh_rows = true;
    Bump bp(base);
    bt.layout_common(bp, nblk_max);
    const size_t zstride = 4 + (size_t)nblk_max;
    bt.zero_base = bp.take<unsigned int>((size_t)B * zstride);
    bt.zero_bytes = (size_t)B * zstride * sizeof(unsigned int);
    for (int b = 0; b < B; b++) {
        CgInst& I = bt.h[b];
        const int n = I.n, m = I.m, N = I.N;
        I.counter = bt.zero_base ? bt.zero_base + (size_t)b * zstride : nullptr;
        I.flags = bt.zero_base ? (int*)(I.counter + 4) : nullptr;
        if (I.fctx && I.fsp) {
            I.spx = bp.take<double>(2 * (size_t)N);
        } else if (I.fctx) {
            I.prhs = bp.take<double>(I.P);
            I.pltmp = bp.take<double>(I.P);
        }
        I.yv = bp.take<double>(N);
        I.sv = bp.take<double>(N);
        I.uv = bp.take<double>(N);
        I.grad = bp.take<double>(N);
        I.rowsum = bp.take<double>(n);
        I.colsum = bp.take<double>(m);
        if (!precond_only) {
            I.rp0 = bp.take<double>((size_t)n * I.nstrips);
            I.rp1 = bp.take<double>((size_t)n * I.nstrips);
            I.cp0 = bp.take<double>((size_t)I.nrb * m);
            I.cp1 = bp.take<double>((size_t)I.nrb * m);
        }
        I.wgpart = bp.take<double>((size_t)8 * I.nvec);
    }
    return bp.off + 256;
}

// SPLR preconditioner z = B^{-1} r for the adjoint instances with a forward
// factor (fctx != 0): level tables of their triangular solves (instances by
// decreasing block count; level t holds those with more than t blocks) and
// the launch sequence of one application.
struct PcSolver {
    Batch& bt;
    std::vector<int> ord, lvl;
    std::vector<int> sp_ids;   // instances whose forward factor is a sparse plan
    size_t sp_lds = 0;
    int nlvl = 0;
    int* d_ord = nullptr;
    int* d_lvl = nullptr;
    int epoch = 0;

    explicit PcSolver(Batch& b) : bt(b)
    {
        for (int i = 0; i < bt.B; i++) {
            if (!bt.h[i].fctx) continue;
            if (bt.pc_sparse[i]) {
                sp_ids.push_back(i);
                sp_lds = std::max(sp_lds, sp_lds_bytes(bt.pc_R[i]));
            } else {
                ord.push_back(i);
            }
        }
        if (!sp_ids.empty()) {
            (void)sp_rmax();   // the kernels' LDS limit is raised once per process
            CG_HIP(hipMemcpyAsync(bt.d_ids, sp_ids.data(), sizeof(int) * sp_ids.size(), hipMemcpyHostToDevice,
                                  bt.stream));
        }
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return bt.h[x].P > bt.h[y].P; });
        for (int i : ord) nlvl = std::max(nlvl, cdiv(bt.h[i].P, CG_NB));
        lvl.assign(nlvl + 1, 0);
        for (int t = 0; t < nlvl; t++) {
            int cnt = 0;
            while (cnt < (int)ord.size() && cdiv(bt.h[ord[cnt]].P, CG_NB) > t) cnt++;
            lvl[t + 1] = lvl[t] + cnt;
        }
        bt.ids_held.clear();    // d_ids and d_pref now hold this solver's lists
        bt.pref_held.clear();
        d_ord = bt.d_pref;
        d_lvl = bt.d_pref + bt.B;
        if (!ord.empty()) {
            CG_HIP(hipMemcpyAsync(d_ord, ord.data(), sizeof(int) * ord.size(), hipMemcpyHostToDevice, bt.stream));
            CG_HIP(hipMemcpyAsync(d_lvl, lvl.data(), sizeof(int) * lvl.size(), hipMemcpyHostToDevice, bt.stream));
        }
    }

    // r = yv -> z = sv for the instances still running; then r.z and beta
    // (init: beta = 0).  Flags of one application: forward blocks 2e-1 (the
    // backward blocks are handed over through pltmp).
    void apply(const double* ctlp, int init)
    {
        if (ord.empty() && sp_ids.empty()) return;
        solve(ctlp);
        lowrank_dots(ctlp);
        lowrank_apply(ctlp);
        hipLaunchKernelGGL(k_pcg_rz, dim3(bt.nvec), dim3(CG_WG), 0, bt.stream, bt.d_insts, bt.d_vec, ctlp, init);
    }
    // the stages of apply: z = A^{-1} r (Schur rhs, forward and backward
    // substitution, back-substitution into z), then the rank-2 BFGS term
    // (its dot products, then the correction of z)
    void solve(const double* ctlp)
    {
        if (!sp_ids.empty())
            hipLaunchKernelGGL(k_sp_precond, dim3((unsigned)sp_ids.size()), dim3(SP_WG), sp_lds, bt.stream,
                               bt.d_insts, bt.d_ids);
        if (ord.empty()) return;
        epoch++;
        bt.rows0(k_pc_rhs, ctlp);
        hipLaunchKernelGGL(k_pc_fwd, dim3(lvl[nlvl]), dim3(CG_WG), 0, bt.stream, bt.d_insts, d_ord, d_lvl, nlvl,
                           2 * epoch - 1);
        hipLaunchKernelGGL(k_pc_bwd, dim3(lvl[nlvl]), dim3(CG_WG), 0, bt.stream, bt.d_insts, d_ord, d_lvl, nlvl);
        bt.rows0(k_pc_back, ctlp);
    }
    void lowrank_dots(const double* ctlp) { bt.vec0(k_pc_wdots, ctlp); }
    void lowrank_apply(const double* ctlp) { bt.vec0(k_pc_wapply, ctlp); }
};

// A stream of the calling thread that only ever records captures.
hipStream_t capture_stream()
{
    thread_local hipStream_t s = nullptr;
    thread_local int gen = -1;
    if (gen != g_res_gen.load()) { s = nullptr; gen = g_res_gen.load(); }
    if (!s) {
        CG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        {
            std::lock_guard<std::mutex> lk(g_res_mu);
            g_res_streams.push_back(s);
        }
        register_atexit();
    }
    return s;
}

void adjoint_run(Batch& bt, const cgot_adjoint_problem* probs, cgot_adjoint_info* info)
{
    const int B = bt.B;
    CG_HIP(hipMemsetAsync(bt.zero_base, 0, bt.zero_bytes, bt.stream));
    CG_HIP(hipMemsetAsync(bt.d_scal, 0, sizeof(double) * B * CG_NSCAL, bt.stream));
    auto c = ctl_vec(B);
    for (int b = 0; b < B; b++) {
        c[(size_t)b * CG_NCTL + CTL_ACTIVE] = 1.0;
        c[(size_t)b * CG_NCTL + CTL_STEP] = probs[b].tol;
        c[(size_t)b * CG_NCTL + CTL_AUX] = (double)probs[b].max_iter;
    }
    const double* dc = bt.ctl(c);
    const double* const* Gs = (const double* const*)bt.d_ptrs;
    const double* const* lams = (const double* const*)(bt.d_ptrs + B);
    double* const* outs = (double* const*)(bt.d_ptrs + 2 * B);
    PcSolver pc(bt);
    auto precondition = [&](const double* ctlp, int init) { pc.apply(ctlp, init); };
    hipLaunchKernelGGL(bt.xy ? k_qv_rhs<true> : k_qv_rhs<false>, dim3(bt.ntiles), dim3(CG_WG), 0, bt.stream,
                       bt.d_insts, bt.d_tile, dc, Gs);
    bt.vec0(k_pcg_init, dc);
    precondition(dc, 1);
    bt.check_launch();
    bt.read_scal();
    std::vector<char> running(B, 0);
    int nrun = 0;
    for (int b = 0; b < B; b++) {
        running[b] = (bt.sc(b, PCG_STATUS) == PCG_RUNNING) && probs[b].max_iter > 0;
        nrun += running[b];
    }
    // One PCG step of the running instances (status words on the device stop
    // the others): Q p with T recomputed, its finish, the update, the
    // preconditioner.
    auto pcg_step = [&](const double* d2) {
        ProfScope ps(bt.stream, PROF_QV, bt.mpass_bytes(bt.hmap[d2], 24.0));
        hipLaunchKernelGGL(bt.xy ? k_qv_pcg<true> : k_qv_pcg<false>, dim3(bt.ntiles), dim3(CG_WG), 0, bt.stream,
                           bt.d_insts, bt.d_tile, d2);
        bt.vec0(k_pcg_q_fin, d2);
        bt.vec0(k_pcg_update, d2);
        precondition(d2, 0);
    };
    // Chunks of kGraphChunk steps (the loop's steady state) run as a captured
    // graph: the same launches with the same arguments every time (the
    // control at a fixed address), replayed until the maps change.  Not with
    // dense stored factors, whose preconditioner passes a per-application
    // epoch, nor under the profiler's per-launch events.
    constexpr int kGraphChunk = 32;
    const bool graphs = pc.ord.empty() && !g_prof;
    hipGraphExec_t gexec = nullptr;
    const int* g_tile = nullptr;
    int g_ntiles = -1, g_nvec = -1, g_nrow = -1;
    struct GraphGuard {
        hipGraphExec_t& e;
        ~GraphGuard() { if (e) (void)hipGraphExecDestroy(e); }
    } gguard{gexec};
    // the first full chunk runs eagerly: a loop that ends within it (most
    // well-conditioned systems) would not repay the capture
    int chunk = 4, eager32 = 0;
    while (nrun > 0) {
        bt.pack_live(running);
        auto c2 = c;
        for (int b = 0; b < B; b++) c2[(size_t)b * CG_NCTL + CTL_ACTIVE] = running[b] ? 1.0 : 0.0;
        if (graphs && chunk == kGraphChunk && eager32 > 0) {
            const double* d2 = bt.ctl_fixed(c2);
            if (gexec && (g_tile != bt.d_tile || g_ntiles != bt.ntiles || g_nvec != bt.nvec || g_nrow != bt.nrow)) {
                CG_HIP(hipGraphExecDestroy(gexec));
                gexec = nullptr;
            }
            if (!gexec) {
                bt.check_launch();
                // the null stream cannot be captured: capture on a side
                // stream (it runs nothing), launch on the solve stream
                const hipStream_t keep = bt.stream;
                const hipStream_t cs = keep ? keep : capture_stream();
                hipGraph_t g = nullptr;
                CG_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
                bt.stream = cs;
                try {
                    for (int k = 0; k < chunk; k++) pcg_step(d2);
                } catch (...) {
                    bt.stream = keep;
                    if (hipStreamEndCapture(cs, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
                    throw;
                }
                bt.stream = keep;
                const hipError_t ce = hipStreamEndCapture(cs, &g);
                CG_HIP(ce);
                const hipError_t ie = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                CG_HIP(ie);
                g_tile = bt.d_tile;
                g_ntiles = bt.ntiles;
                g_nvec = bt.nvec;
                g_nrow = bt.nrow;
            }
            CG_HIP(hipGraphLaunch(gexec, bt.stream));
        } else {
            const double* d2 = bt.ctl(c2);
            for (int k = 0; k < chunk; k++) pcg_step(d2);
        }
        bt.check_launch();
        bt.read_scal();
        nrun = 0;
        for (int b = 0; b < B; b++) {
            if (running[b] && bt.sc(b, PCG_STATUS) != PCG_RUNNING) running[b] = 0;
            nrun += running[b];
        }
        if (chunk == kGraphChunk) eager32++;
        chunk = std::min(chunk * 2, 32);
    }
    // a fresh copy of the all-active control: the slot of dc may have been
    // reused by the PCG loop's uploads (the ring holds kCtlSlots blocks)
    bt.use_full();
    const double* dfin = bt.ctl(c);
    hipLaunchKernelGGL(k_grad_assemble, dim3(bt.ntiles), dim3(CG_WG), 0, bt.stream, bt.d_insts, bt.d_tile, dfin,
                       lams, Gs, outs);
    bt.check_launch();
    bt.sync();
    for (int b = 0; b < B; b++)
        if (bt.h[b].fctx && bt.sc(b, CS_PIVOT) >= 2.0)
            throw std::runtime_error("internal: a block-flag wait of the preconditioner exceeded its bound");
    for (int b = 0; b < B; b++) {
        const double st = bt.sc(b, PCG_STATUS);
        info[b].status = (int)st;
        info[b].iterations = (st == PCG_ZERO_RHS) ? 0 : (int)bt.sc(b, PCG_ITER);
        info[b].converged = (st == PCG_CONVERGED || st == PCG_ZERO_RHS) ? 1 : 0;
        if (probs[b].max_iter <= 0 && st == PCG_RUNNING) { info[b].status = 3; info[b].converged = 0; }
    }
}

// ---------------------------------------------------------------------------
// BCD (Sinkhorn) batch.  Reference: cuda_sinkhorn_bcd, src/sinkhorn_bcd_kernel.cu:429-634.
// ---------------------------------------------------------------------------
size_t bcd_layout(Batch& bt, const cgot_bcd_problem* probs, void* base)
{
    const int B = bt.B;
    bt.h.assign(B, CgInst{});
    for (int b = 0; b < B; b++) {
        const auto& p = probs[b];
        CgInst& I = bt.h[b];
        set_geometry(I, p.n, p.m, p.reg);
        set_cost(I, p.M, p.xy);
        I.plan = p.plan;
        I.dual = p.dual;
    }
    bt.assign_maps();
    Bump bp(base);
    bt.layout_common(bp, 0);
    bt.zero_base = bp.take<unsigned int>((size_t)B * 4);
    bt.zero_bytes = (size_t)B * 4 * sizeof(unsigned int);
    for (int b = 0; b < B; b++) {
        CgInst& I = bt.h[b];
        const int n = I.n, m = I.m, N = I.N;
        I.counter = bt.zero_base ? bt.zero_base + (size_t)b * 4 : nullptr;
        I.ab = bp.take<double>(n + m);
        I.logab = bp.take<double>(n + m);
        I.gamma = bp.take<double>(n + m);
        I.gamma_prev = bp.take<double>(n + m);
        I.grad = bp.take<double>(N);
        I.grad_prev = bp.take<double>(N);
        I.rowsum = bp.take<double>(n);
        I.colsum = bp.take<double>(m);
        I.rp0 = bp.take<double>((size_t)n * I.nstrips);
        I.rp1 = bp.take<double>((size_t)n * I.nstrips);
        I.cp0 = bp.take<double>((size_t)I.nrb * m);
        I.cp1 = bp.take<double>((size_t)I.nrb * m);
        I.wgpart = bp.take<double>((size_t)8 * I.nvec);
    }
    return bp.off + 256;
}

void bcd_run(Batch& bt, const cgot_bcd_problem* probs, int32_t* niter)
{
    const int B = bt.B;
    CG_HIP(hipMemsetAsync(bt.zero_base, 0, bt.zero_bytes, bt.stream));
    CG_HIP(hipMemsetAsync(bt.d_scal, 0, sizeof(double) * B * CG_NSCAL, bt.stream));
    auto mk = [&](const std::vector<char>& on) {
        auto c = ctl_vec(B);
        for (int b = 0; b < B; b++) c[(size_t)b * CG_NCTL + CTL_ACTIVE] = on[b] ? 1.0 : 0.0;
        return bt.ctl(c);
    };
    std::vector<char> all(B, 1), act(B, 1), marg(B, 0);
    hipLaunchKernelGGL(k_init_logab, dim3(bt.nvec), dim3(CG_WG), 0, bt.stream, bt.d_insts, bt.d_vec, mk(all),
                       (const double* const*)bt.d_ptrs, (const double* const*)nullptr);
    for (int b = 0; b < B; b++) {
        niter[b] = 0;
        if (probs[b].max_iter <= 0) act[b] = 0;
    }
    for (int it = 0;; it++) {
        int na = 0;
        for (int b = 0; b < B; b++) {
            if (act[b] && it >= probs[b].max_iter) act[b] = 0;
            na += act[b];
        }
        if (na == 0) break;
        bt.pack_live(act);
        const double* dc = mk(act);
        bt.vec(k_vcopy, dc, 0);
        row_lse(bt, dc, CG_PT_GAMMA);
        col_lse(bt, dc, CG_PT_GAMMA);
        std::vector<char> p1(B, 0), p2(B, 0);
        for (int b = 0; b < B; b++) { p1[b] = act[b] && !marg[b]; p2[b] = act[b] && marg[b]; }
        if (std::count(p1.begin(), p1.end(), 1)) bt.vec0(k_bcd_diff, mk(p1));
        if (std::count(p2.begin(), p2.end(), 1)) tpass(bt, mk(p2), CG_PT_GAMMA);
        bt.check_launch();
        bt.read_scal();
        for (int b = 0; b < B; b++) {
            if (!act[b]) continue;
            niter[b] = it + 1;
            const double tol = probs[b].tol;
            if (p1[b]) {
                const double* S = &bt.hscal[(size_t)b * CG_NSCAL + BCD_SUMS];
                const int n = probs[b].n, m = probs[b].m;
                const double cc = (S[1] - S[0]) / (n + m), dd = (S[3] - S[2]) / (n + m), K = cc - dd;
                const double d2 = S[5] + 2.0 * K * S[4] + n * K * K + S[7] - 2.0 * K * S[6] + m * K * K;
                if (std::sqrt(std::max(d2, 0.0)) < tol * 10) marg[b] = 1;
            } else {
                if (std::sqrt(bt.sc(b, CS_G2A)) < tol) act[b] = 0;
            }
        }
    }
    bt.use_full();
    const double* da = mk(all);
    bt.vec0(k_output_dual, da);
    bt.tiles0(k_plan, da);
    bt.check_launch();
    bt.sync();
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
#define CG_API_BEGIN \
    InFlight cg_inflight_; \
    try {
#define CG_API_END                         \
    }                                      \
    catch (const std::invalid_argument& e) \
    {                                      \
        g_last_error = e.what();           \
        return -2;                         \
    }                                      \
    catch (const std::exception& e)        \
    {                                      \
        g_last_error = e.what();           \
        return -1;                         \
    }


// ---------------------------------------------------------------------------
// Concurrent lockstep groups of one ragged batch.
//
// The per-iteration critical path of an instance is its Schur factorization
// chain (one launch per 64-column block), so a lockstep batch advances at the
// pace of its largest instance while the GPU is mostly idle.  A batch is
// therefore split by block count into up to kMaxGroups classes (block count in
// (max/2, max], (max/4, max/2], ...); each class runs the lockstep solver on
// its own HIP stream from its own persistent host thread, so the classes
// overlap on the device.  Results are per instance and do not depend on the
// grouping.
// ---------------------------------------------------------------------------
constexpr int kMaxGroups = 4;
thread_local int g_max_groups = kMaxGroups;   // cgot_set_max_groups (per calling thread)

struct GroupPlan {
    std::vector<std::vector<int>> idx;                 // instance indices per group
    std::vector<std::vector<cgot_splr_problem>> probs;
    std::vector<size_t> off;                           // workspace offset of each group
    size_t total = 0;
};

GroupPlan plan_splr_groups(int B, const cgot_splr_problem* probs)
{
    GroupPlan g;
    std::vector<int> nblk(B);
    int mx = 1;
    for (int b = 0; b < B; b++) {
        const int P = (probs[b].m - 1 <= probs[b].n) ? probs[b].m - 1 : probs[b].n;
        nblk[b] = std::max(1, cdiv(P, CG_NB));
        mx = std::max(mx, nblk[b]);
    }
    const int G = (B >= 8) ? std::max(1, std::min(g_max_groups, kMaxGroups)) : 1;
    std::vector<std::vector<int>> cls(G);
    for (int b = 0; b < B; b++) {
        int k = 0;
        while (k + 1 < G && (long long)nblk[b] * (2LL << k) <= mx) k++;
        cls[k].push_back(b);
    }
    size_t off = 0;
    for (auto& c : cls) {
        if (c.empty()) continue;
        std::vector<cgot_splr_problem> sub;
        for (int b : c) sub.push_back(probs[b]);
        Batch tmp;
        tmp.B = (int)sub.size();
        const size_t need = splr_layout(tmp, sub.data(), nullptr);
        off = (off + 255) & ~(size_t)255;
        g.off.push_back(off);
        off += need;
        g.idx.push_back(c);
        g.probs.push_back(std::move(sub));
    }
    g.total = off + 256;
    return g;
}

// A pool of persistent host threads.  Callers may run concurrently: each
// call queues its jobs and waits for its own completion count; the pool grows
// so that every queued job has a thread, up to kMaxWorkers.  stop() (from
// cgot_shutdown) lets the idle workers finish and joins them; a later call
// starts new ones.
class GroupPool {
public:
    // the size-class workers
    static GroupPool& get()
    {
        static GroupPool* p = new GroupPool();   // outlives every caller; workers joined by stop()
        return *p;
    }
    // the sparse-plan analysis workers (separate, so that an analysis never
    // waits behind a size-class job that is waiting for it)
    static GroupPool& analysis()
    {
        static GroupPool* p = new GroupPool();
        return *p;
    }
    void run(std::vector<std::function<void()>>& jobs)
    {
        Call call;
        call.remaining = (int)jobs.size();
        std::unique_lock<std::mutex> lk(mu_);
        stop_ = false;
        for (auto& j : jobs) q_.push_back({&j, &call});
        while (idle_ < (int)q_.size() && (int)threads_.size() < kMaxWorkers) {
            threads_.emplace_back(&GroupPool::loop, this);
            idle_++;
        }
        register_atexit();
        cv_.notify_all();
        call.cv.wait(lk, [&] { return call.remaining == 0; });
    }
    void stop()
    {
        std::vector<std::thread> ts;
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            ts.swap(threads_);
            cv_.notify_all();
        }
        for (auto& t : ts) t.join();
        std::lock_guard<std::mutex> lk(mu_);
        idle_ = 0;
    }

private:
    static constexpr int kMaxWorkers = 16;
    struct Call { int remaining = 0; std::condition_variable cv; };
    struct Item { std::function<void()>* job; Call* call; };
    void loop()
    {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [&] { return !q_.empty() || stop_; });
            if (q_.empty()) return;       // stopped and nothing queued
            Item it = q_.front();
            q_.pop_front();
            idle_--;
            lk.unlock();
            (*it.job)();
            lk.lock();
            idle_++;
            if (--it.call->remaining == 0) it.call->cv.notify_all();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Item> q_;
    std::vector<std::thread> threads_;
    int idle_ = 0;
    bool stop_ = false;
};

void parallel_for(int count, const std::function<void(int)>& fn)
{
    if (count <= 0) return;
    const int hw = std::max(1, (int)std::thread::hardware_concurrency());
    const int nj = std::min({count, 8, hw});
    if (nj == 1) {
        for (int k = 0; k < count; k++) fn(k);
        return;
    }
    std::vector<std::function<void()>> jobs;
    std::vector<std::string> err(nj);
    for (int j = 0; j < nj; j++)
        jobs.push_back([&, j]() {
            try {
                for (int k = j; k < count; k += nj) fn(k);
            } catch (const std::exception& e) {
                err[j] = e.what();
            }
        });
    GroupPool::analysis().run(jobs);
    for (auto& e : err)
        if (!e.empty()) throw std::runtime_error(e);
}

int sp_rmax()
{
    // per process: the device's LDS per workgroup bounds the packed core
    static const int r = [] {
        int dev = 0, lds = 0;
        CG_HIP(hipGetDevice(&dev));
        CG_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        int R = SP_RMAX_BIG;
        while (R > 0 && sp_lds_bytes(R) > (size_t)lds) R--;
        const int bytes = (int)sp_lds_bytes(R);   // increasing in R
        // raise the kernels' dynamic-LDS limit where the runtime asks for it;
        // a launch above what the device allows fails at the launch check
        (void)hipFuncSetAttribute((const void*)k_sp_newton<0>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        (void)hipFuncSetAttribute((const void*)k_sp_newton<1>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        (void)hipFuncSetAttribute((const void*)k_sp_precond, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        (void)hipFuncSetAttribute((const void*)k_newton_mixed<0>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        (void)hipFuncSetAttribute((const void*)k_newton_mixed<1>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        g_lds_max = lds;
        (void)hipGetLastError();
        return R;
    }();
    return r;
}

// CUs kept free of size classes 1.. when a batch is split (CU-masked streams
// for classes 1..; class 0 on a stream of its own).  It paid while the
// largest class's dense factorization chain set the batch time; with the
// sparse Newton solve the smallest class (most iterations) sets it, and on
// the ragged B = 128 batch no reserve is fastest (63.4 ms against 66.4 ms
// with 80 CUs kept for class 0).  Default 0; CGOT_CU_RESERVE=<count> sets one.
int cu_reserve()
{
    static const int env = [] {
        const char* e = std::getenv("CGOT_CU_RESERVE");
        return e ? std::max(0, std::atoi(e)) : -1;
    }();
    return (env >= 0) ? env : 0;
}

// Streams of size classes 0..kMaxGroups-1 for batches issued from the calling
// thread on `device` (created once per thread and device, default priority).
// Each class index keeps its stream from call to call, so a caller uses the
// same streams (and hardware queues) every time.  With CUs reserved, classes
// 1.. run on CU-masked streams (blocking streams: they also order against
// null-stream work of other threads) and class 0 on a stream of its own
// (unmasked, non-blocking), so that no class runs on a caller's null stream;
// with no reserve, class 0 runs on the caller's stream.
hipStream_t group_stream(int device, int k)
{
    thread_local std::map<int, std::array<hipStream_t, kMaxGroups>> streams;
    thread_local int gen = -1;
    if (gen != g_res_gen.load()) { streams.clear(); gen = g_res_gen.load(); }
    auto it = streams.find(device);
    if (it == streams.end()) {
        std::array<hipStream_t, kMaxGroups> a{};
        int ncu = 0;
        CG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
        const int res = cu_reserve();
        if (res > 0 && res < ncu) {
            CG_HIP(hipStreamCreateWithFlags(&a[0], hipStreamNonBlocking));
            std::vector<uint32_t> mask((ncu + 31) / 32, 0u);
            for (int c = 0; c < ncu - res; c++) mask[c / 32] |= 1u << (c % 32);
            for (int q = 1; q < kMaxGroups; q++)
                CG_HIP(hipExtStreamCreateWithCUMask(&a[q], (uint32_t)mask.size(), mask.data()));
        } else {
            for (int q = 1; q < kMaxGroups; q++) CG_HIP(hipStreamCreateWithFlags(&a[q], hipStreamNonBlocking));
        }
        {
            std::lock_guard<std::mutex> lk(g_res_mu);
            for (hipStream_t x : a)
                if (x) g_res_streams.push_back(x);
        }
        register_atexit();
        it = streams.emplace(device, a).first;
    }
    return it->second[k];
}

// ---------------------------------------------------------------------------
// Sessions: one persistent lockstep group into which instances are added at
// iteration boundaries and from which they leave when they stop (continuous
// batching).  Each instance's arrays, its scalars and a copy of its
// descriptor live in a workspace of its own, so that its backward context
// outlives its slot; the session workspace holds the group's shared arrays
// (descriptors, maps, control blocks, scalars, prefix tables) for at most
// `cap` instances within the stated totals.
// ---------------------------------------------------------------------------
struct SessionLimits {
    int cap;
    int64_t max_sum_nm, max_tiles;
    int max_nblk;
};

int64_t inst_tiles(int n, int m)
{
    CgInst I{};
    set_geometry(I, n, m, 1.0);
    return I.ntiles;
}

// Shared arrays of a session; with base == nullptr only sizes.
size_t session_layout(Batch& bt, const SessionLimits& L, void* base)
{
    bt.B = L.cap;
    bt.h.assign(L.cap, CgInst{});
    // map capacity: tiles, vector groups (ceil((n+m)/CG_WG) each) and waves
    bt.ntiles = (int)L.max_tiles;
    bt.nvec = (int)(L.max_sum_nm / CG_WG + L.cap);
    bt.nrow = (int)(L.max_sum_nm / CG_WAVES + L.cap);
    Bump bp(base);
    bt.layout_common(bp, L.max_nblk);
    // the staging of map uploads is sized by the *_full counts: the capacity
    bt.ntiles_full = bt.ntiles;
    bt.nvec_full = bt.nvec;
    bt.nrow_full = bt.nrow;
    bt.ntiles = bt.nvec = bt.nrow = 0;
    bt.packed_wgs = 0;
    return bp.off + 256;
}

// One instance's own workspace: descriptor copy, scalars, zero region, arrays.
struct InstWs {
    CgInst* desc = nullptr;
    double* scal = nullptr;
    unsigned int* zero = nullptr;
    size_t zbytes = 0;
};
size_t inst_ws_layout(CgInst& I, const cgot_splr_problem& p, void* base, InstWs& w)
{
    inst_geometry(I, p);
    Bump bp(base);
    w.desc = bp.take<CgInst>(1);
    w.scal = bp.take<double>(CG_NSCAL);
    const size_t zw = zero_words(I);
    w.zero = bp.take<unsigned int>(zw);
    w.zbytes = zw * sizeof(unsigned int);
    set_zero_region(I, w.zero);
    inst_arrays(I, bp);
    return bp.off + 256;
}

struct Session {
    SessionLimits lim;
    int device = 0;
    hipStream_t stream = nullptr;
    Batch bt;
    SplrSolver solver;
    std::vector<char> used, pending;
    std::vector<int64_t> tag;
    std::vector<InstWs> iw;
    std::vector<const double*> ptrs;
    int64_t sum_nm = 0, tiles = 0;
    int live = 0;
    Pinned up_insts, up_ptrs, up_desc;   // staging of descriptor and pointer uploads
    Pinned ctl_host, scal_host;          // control ring and scalar read-backs
    bool xy_set = false;

    Session() : solver(bt) {}

    void start_pending()
    {
        const int B = bt.B;
        bool any = false;
        for (int b = 0; b < B; b++) any = any || pending[b];
        if (!any) return;
        for (int b = 0; b < B; b++)
            if (used[b]) bt.h[b].scal = bt.d_scal + (size_t)b * CG_NSCAL;
        // descriptors and input pointers of every slot (free slots: zero
        // descriptors, which the per-instance kernels skip); the staging was
        // last read before the previous step's final synchronisation
        CgInst* hs = (CgInst*)up_insts.get((sizeof(CgInst) * B) / sizeof(double) + 1);
        std::memcpy((void*)hs, bt.h.data(), sizeof(CgInst) * B);
        CG_HIP(hipMemcpyAsync(bt.d_insts, hs, sizeof(CgInst) * B, hipMemcpyHostToDevice, stream));
        const double** hp = (const double**)up_ptrs.get(3 * (size_t)B + 1);
        std::memcpy((void*)hp, ptrs.data(), sizeof(void*) * 3 * B);
        CG_HIP(hipMemcpyAsync(bt.d_ptrs, hp, sizeof(void*) * 3 * B, hipMemcpyHostToDevice, stream));
        for (int b = 0; b < B; b++) {
            if (!pending[b]) continue;
            CG_HIP(hipMemsetAsync(iw[b].zero, 0, iw[b].zbytes, stream));
            CG_HIP(hipMemsetAsync(bt.d_scal + (size_t)b * CG_NSCAL, 0, sizeof(double) * CG_NSCAL, stream));
        }
        hipLaunchKernelGGL(k_wgpart_init, dim3(B), dim3(CG_WG), 0, stream, bt.d_insts);
        // slots changed owners: the cached maps are keyed by slot sets only
        bt.map_set.clear();
        bt.sub_key.clear();
        std::vector<char> cover(B, 0);
        for (int b = 0; b < B; b++) cover[b] = (used[b] && solver.hi[b].live && !solver.hi[b].done) || pending[b];
        bt.set_exact(cover);
        solver.begin(pending);
        std::fill(pending.begin(), pending.end(), 0);
    }

    // Outputs of the instances that stopped; frees their slots.
    int collect(int64_t* tags_out, cgot_splr_info* info_out, int cap_out, int have)
    {
        const int B = bt.B;
        std::vector<char> fin(B, 0);
        int nf = 0;
        for (int b = 0; b < B; b++)
            if (used[b] && solver.hi[b].live && solver.hi[b].done) { fin[b] = 1; nf++; }
        if (nf == 0) return 0;
        if (have + nf > cap_out) throw std::invalid_argument("session step: output arrays too small");
        bt.set_exact(fin);
        std::vector<cgot_splr_info> info(B);
        solver.finish(fin, info.data());
        // each finished instance keeps its scalars and its descriptor in its
        // own workspace; the backward context refers to those
        CgInst* hd = (CgInst*)up_desc.get((sizeof(CgInst) * nf) / sizeof(double) + 1);
        int k = 0;
        for (int b = 0; b < B; b++) {
            if (!fin[b]) continue;
            CG_HIP(hipMemcpyAsync(iw[b].scal, bt.d_scal + (size_t)b * CG_NSCAL, sizeof(double) * CG_NSCAL,
                                  hipMemcpyDeviceToDevice, stream));
            hd[k] = bt.h[b];
            hd[k].scal = iw[b].scal;
            CG_HIP(hipMemcpyAsync(iw[b].desc, hd + k, sizeof(CgInst), hipMemcpyHostToDevice, stream));
            cgot_splr_info o = info[b];
            if (o.factor) o.factor = iw[b].desc;
            tags_out[have + k] = tag[b];
            info_out[have + k] = o;
            // free the slot (its host state too: no pattern left pending for it)
            solver.hi[b] = HostInst{};
            used[b] = 0;
            sum_nm -= bt.h[b].n + bt.h[b].m;
            tiles -= bt.h[b].ntiles;
            bt.h[b] = CgInst{};
            ptrs[3 * b] = ptrs[3 * b + 1] = ptrs[3 * b + 2] = nullptr;
            live--;
            k++;
        }
        bt.map_set.clear();
        bt.sub_key.clear();
        // the descriptor copies are read by other streams (the backward): complete them here
        CG_HIP(hipStreamSynchronize(stream));
        bt.refresh_rows = false;
        for (int b = 0; b < B; b++)
            if (used[b] && wants_refresh_rows(bt.h[b])) bt.refresh_rows = true;
        return nf;
    }
};

void splr_solve_group(const cgot_splr_problem* probs, int B, void* ws, hipStream_t stream, cgot_splr_info* info)
{
    Batch bt;
    bt.B = B;
    bt.stream = stream;
    splr_layout(bt, probs, ws);
    std::vector<const double*> ptrs(3 * (size_t)B);
    for (int b = 0; b < B; b++) {
        ptrs[3 * b] = probs[b].a;
        ptrs[3 * b + 1] = probs[b].b;
        ptrs[3 * b + 2] = probs[b].x0;
    }
    bt.upload_maps(ptrs);
    SplrSolver s(bt);
    s.run(probs, info);
}

// Release the registered resources: join the worker threads, then wait for
// and destroy the size-class streams and free the pinned staging buffers.
void shutdown_resources()
{
    GroupPool::get().stop();
    GroupPool::analysis().stop();
    std::vector<hipStream_t> streams;
    std::vector<void*> pinned;
    {
        std::lock_guard<std::mutex> lk(g_res_mu);
        streams.swap(g_res_streams);
        pinned.swap(g_res_pinned);
        g_res_gen.fetch_add(1);
    }
    for (hipStream_t s : streams) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    for (void* p : pinned) (void)hipHostFree(p);
}

// At process exit, release everything before the HIP runtime tears itself
// down (its exit handlers were registered when it loaded, before this one, so
// they run after it).  Skipped while a solve is still inside the library.
void atexit_shutdown()
{
    if (g_inflight.load() == 0) shutdown_resources();
}

void register_atexit()
{
    static std::once_flag once;
    std::call_once(once, [] { std::atexit(atexit_shutdown); });
}

}  // namespace

extern "C" {

const char* cgot_last_error(void) { return g_last_error.c_str(); }

int64_t cgot_sparse_plan(int32_t n, int32_t m, const int32_t* rowptr, const int32_t* colj, int64_t cap_bytes,
                         int32_t rmax, void* hdr_out, void* plan_out, int64_t plan_cap)
{
    try {
        if (n < 1 || m < 1 || !rowptr || (!colj && rowptr[n] > 0) || !hdr_out)
            throw std::invalid_argument("bad arguments");
        SpHdr h{};
        std::vector<char> plan;
        if (!sp_analyze(n, m, rowptr, colj, cap_bytes, rmax, h, plan)) return -1;
        std::memcpy(hdr_out, &h, sizeof(SpHdr));
        if (plan_out && plan_cap >= h.bytes) {
            // the integer part; the value arrays are the device's (zero here)
            std::memcpy(plan_out, plan.data(), plan.size());
            std::memset((char*)plan_out + plan.size(), 0, (size_t)h.bytes - plan.size());
        }
        return h.bytes;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return -2;
    }
}

int cgot_shutdown(void)
{
    if (g_inflight.load() != 0) {
        g_last_error = "cgot_shutdown: a solve is still running";
        return -1;
    }
    shutdown_resources();
    return 0;
}

void cgot_profile_enable(int32_t on)
{
    g_prof = (on != 0);
}

void cgot_profile_reset(void)
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (int k = 0; k < PROF_NCLS; k++) { g_prof_ms[k] = 0; g_prof_work[k] = 0; g_prof_count[k] = 0; }
}

int cgot_profile_read(int32_t cls, double* ms_total, double* work_total, int64_t* launches)
{
    if (cls < 0 || cls >= PROF_NCLS) return -2;
    try { prof_collect(); } catch (const std::exception& e) { g_last_error = e.what(); return -1; }
    std::lock_guard<std::mutex> lk(g_prof_mu);
    *ms_total = g_prof_ms[cls];
    *work_total = g_prof_work[cls];
    *launches = g_prof_count[cls];
    return 0;
}
const char* cgot_version(void) { return "curegot_amd 0.1.0 gfx950"; }

size_t cgot_splr_workspace_size(int32_t B, const cgot_splr_problem* probs)
{
    try {
        if (B < 1) throw std::invalid_argument("B must be >= 1");
        for (int b = 0; b < B; b++) check_problem(probs[b]);
        return plan_splr_groups(B, probs).total;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return 0;
    }
}

int cgot_splr_solve(int32_t B, const cgot_splr_problem* probs, void* workspace, size_t workspace_bytes,
                    void* stream, cgot_splr_info* info)
{
    CG_API_BEGIN
    if (B < 1) throw std::invalid_argument("B must be >= 1");
    for (int b = 0; b < B; b++) check_problem(probs[b]);
    const GroupPlan g = plan_splr_groups(B, probs);
    if (workspace == nullptr || workspace_bytes < g.total)
        throw std::invalid_argument("workspace too small: need " + std::to_string(g.total) + " bytes");
    const int G = (int)g.idx.size();
    if (G == 1) {
        splr_solve_group(probs, B, (char*)workspace + g.off[0], (hipStream_t)stream, info);
        return 0;
    }
    // group 0 on the caller's stream in this thread, the others on worker
    // streams (ordered after the caller's stream on entry and before its later
    // work on exit), so that each group has its own hardware queue
    int device = 0;
    CG_HIP(hipGetDevice(&device));
    hipEvent_t start;
    CG_HIP(hipEventCreateWithFlags(&start, hipEventDisableTiming));
    CG_HIP(hipEventRecord(start, (hipStream_t)stream));
    std::vector<std::vector<cgot_splr_info>> sub(G);
    std::vector<std::string> err(G);
    std::vector<hipEvent_t> done(G, nullptr);
    std::vector<std::function<void()>> jobs;
    std::vector<hipStream_t> gst(G, nullptr);
    for (int k = 1; k < G; k++) gst[k] = group_stream(device, k);
    // class 0: the caller's stream, or its own stream when CUs are reserved
    hipStream_t st0 = group_stream(device, 0);
    if (st0 != nullptr) CG_HIP(hipStreamWaitEvent(st0, start, 0));
    else st0 = (hipStream_t)stream;
    for (int k = 0; k < G; k++) sub[k].assign(g.idx[k].size(), cgot_splr_info{});
    for (int k = 1; k < G; k++) {
        jobs.push_back([&, k]() {
            try {
                CG_HIP(hipSetDevice(device));
                hipStream_t st = gst[k];
                CG_HIP(hipStreamWaitEvent(st, start, 0));
                splr_solve_group(g.probs[k].data(), (int)g.probs[k].size(), (char*)workspace + g.off[k], st,
                                 sub[k].data());
                CG_HIP(hipEventCreateWithFlags(&done[k], hipEventDisableTiming));
                CG_HIP(hipEventRecord(done[k], st));
            } catch (const std::exception& e) {
                err[k] = e.what();
                // the caller frees the workspace and outputs when the error
                // reaches it: let this class's queued kernels finish first
                (void)hipStreamSynchronize(gst[k]);
            }
        });
    }
    std::thread host([&]() { GroupPool::get().run(jobs); });
    try {
        splr_solve_group(g.probs[0].data(), (int)g.probs[0].size(), (char*)workspace + g.off[0], st0,
                         sub[0].data());
        if (st0 != (hipStream_t)stream) {
            CG_HIP(hipEventCreateWithFlags(&done[0], hipEventDisableTiming));
            CG_HIP(hipEventRecord(done[0], st0));
        }
    } catch (const std::exception& e) {
        err[0] = e.what();
        (void)hipStreamSynchronize(st0);
    }
    host.join();
    for (int k = 0; k < G; k++) {
        if (done[k]) {
            (void)hipStreamWaitEvent((hipStream_t)stream, done[k], 0);
            (void)hipEventDestroy(done[k]);
        }
    }
    (void)hipEventDestroy(start);
    for (int k = 0; k < G; k++)
        if (!err[k].empty()) throw std::runtime_error(err[k]);
    for (int k = 0; k < G; k++)
        for (size_t q = 0; q < g.idx[k].size(); q++) info[g.idx[k][q]] = sub[k][q];
    return 0;
    CG_API_END
}

namespace {
SessionLimits session_limits(int32_t cap, int64_t max_sum_nm, int64_t max_tiles, int32_t max_nblk)
{
    if (cap < 1 || max_sum_nm < 2 || max_tiles < 1 || max_nblk < 1)
        throw std::invalid_argument("session limits must be positive");
    if (max_tiles > (1LL << 28) || max_sum_nm > (1LL << 34))
        throw std::invalid_argument("session limits too large");
    return SessionLimits{cap, max_sum_nm, max_tiles, max_nblk};
}
}  // namespace

size_t cgot_session_workspace_size(int32_t capacity, int64_t max_sum_nm, int64_t max_tiles, int32_t max_nblk)
{
    try {
        Batch bt;
        return session_layout(bt, session_limits(capacity, max_sum_nm, max_tiles, max_nblk), nullptr);
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return 0;
    }
}

int64_t cgot_session_inst_tiles(int32_t n, int32_t m)
{
    if (n < 1 || m < 1) return -2;
    return inst_tiles(n, m);
}

size_t cgot_session_inst_workspace_size(const cgot_splr_problem* p)
{
    try {
        if (!p) throw std::invalid_argument("null problem");
        CgInst I{};
        InstWs w;
        return inst_ws_layout(I, *p, nullptr, w);
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return 0;
    }
}

int cgot_session_create(int32_t capacity, int64_t max_sum_nm, int64_t max_tiles, int32_t max_nblk, void* workspace,
                        size_t workspace_bytes, void* stream, void** session)
{
    CG_API_BEGIN
    if (!session) throw std::invalid_argument("null session handle");
    *session = nullptr;
    const SessionLimits L = session_limits(capacity, max_sum_nm, max_tiles, max_nblk);
    std::unique_ptr<Session> S(new Session());
    S->lim = L;
    S->stream = (hipStream_t)stream;
    CG_HIP(hipGetDevice(&S->device));
    Batch& bt = S->bt;
    bt.stream = S->stream;
    const size_t need = session_layout(bt, L, nullptr);
    if (!workspace || workspace_bytes < need)
        throw std::invalid_argument("session workspace too small: need " + std::to_string(need) + " bytes");
    session_layout(bt, L, workspace);
    bt.exact_maps = true;
    bt.map_set.assign(L.cap, 0);
    S->used.assign(L.cap, 0);
    S->pending.assign(L.cap, 0);
    S->tag.assign(L.cap, 0);
    S->iw.assign(L.cap, InstWs{});
    S->ptrs.assign(3 * (size_t)L.cap, nullptr);
    S->solver.hi.assign(L.cap, HostInst{});
    // control ring and scalar read-back staging of this session
    bt.hctl = S->ctl_host.get((size_t)kCtlSlots * L.cap * CG_NCTL);
    bt.hscal = S->scal_host.get((size_t)L.cap * CG_NSCAL);
    CG_HIP(hipMemsetAsync(bt.d_insts, 0, sizeof(CgInst) * L.cap, bt.stream));
    CG_HIP(hipMemsetAsync(bt.d_scal, 0, sizeof(double) * CG_NSCAL * L.cap, bt.stream));
    CG_HIP(hipStreamSynchronize(bt.stream));
    register_atexit();
    *session = S.release();
    return 0;
    CG_API_END
}

int cgot_session_add(void* session, const cgot_splr_problem* p, void* inst_workspace, size_t inst_workspace_bytes,
                     int64_t tag)
{
    CG_API_BEGIN
    Session* S = (Session*)session;
    if (!S || !p) throw std::invalid_argument("null session or problem");
    check_problem(*p);
    const bool xy = p->xy.X != nullptr;
    if (S->live > 0 && S->xy_set && xy != S->bt.xy)
        throw std::invalid_argument("a session must not mix geometric (X, Y) and stored (M) costs");
    int b = -1;
    for (int k = 0; k < S->lim.cap; k++)
        if (!S->used[k]) { b = k; break; }
    const int64_t t = inst_tiles(p->n, p->m);
    CgInst probe{};
    inst_geometry(probe, *p);
    if (b < 0 || S->sum_nm + p->n + p->m > S->lim.max_sum_nm || S->tiles + t > S->lim.max_tiles ||
        cdiv(probe.P, CG_NB) > S->lim.max_nblk)
        return -3;   // no room now
    CgInst& I = S->bt.h[b];
    I = CgInst{};
    InstWs w;
    const size_t need = inst_ws_layout(I, *p, nullptr, w);
    if (!inst_workspace || inst_workspace_bytes < need)
        throw std::invalid_argument("instance workspace too small: need " + std::to_string(need) + " bytes");
    inst_ws_layout(I, *p, inst_workspace, w);
    S->iw[b] = w;
    S->bt.xy = xy;
    S->xy_set = true;
    if (wants_refresh_rows(I)) S->bt.refresh_rows = true;
    S->ptrs[3 * b] = p->a;
    S->ptrs[3 * b + 1] = p->b;
    S->ptrs[3 * b + 2] = p->x0;
    S->solver.init_inst(b, *p);
    S->used[b] = 1;
    S->pending[b] = 1;
    S->tag[b] = tag;
    S->sum_nm += p->n + p->m;
    S->tiles += t;
    S->live++;
    return b;
    CG_API_END
}

int cgot_session_step(void* session, int32_t iters, int64_t* tags_out, cgot_splr_info* info_out, int32_t out_cap,
                      int32_t* ndone)
{
    CG_API_BEGIN
    Session* S = (Session*)session;
    if (!S || !ndone || (out_cap > 0 && (!tags_out || !info_out))) throw std::invalid_argument("bad arguments");
    *ndone = 0;
    int dev = 0;
    CG_HIP(hipGetDevice(&dev));
    if (dev != S->device) throw std::invalid_argument("session used on another device");
    S->start_pending();
    int have = 0;
    for (int k = 0; k < std::max(iters, 1); k++) {
        const bool more = S->solver.step();
        have += S->collect(tags_out, info_out, out_cap, have);
        if (!more) break;
    }
    *ndone = have;
    return 0;
    CG_API_END
}

int cgot_session_live(void* session)
{
    Session* S = (Session*)session;
    return S ? S->live : -2;
}

int cgot_session_destroy(void* session)
{
    CG_API_BEGIN
    Session* S = (Session*)session;
    if (S) {
        (void)hipStreamSynchronize(S->stream);
        delete S;
    }
    return 0;
    CG_API_END
}

int cgot_set_sparse_solve(int32_t mode)
{
    if (mode < 0 || mode > 2) return -1;
    return g_sparse_on.exchange(mode);
}

int cgot_set_tile_table(const int32_t* m_max, const int32_t* cols, int32_t count)
{
    if (count < 0 || count > 64 || (count > 0 && (!m_max || !cols))) {
        g_last_error = "tile table: count must be in [0, 64] with non-null rows";
        return -2;
    }
    std::vector<std::pair<int, int>> t;
    for (int k = 0; k < count; k++) {
        if (m_max[k] < 1 || (k > 0 && m_max[k] <= m_max[k - 1])) {
            g_last_error = "tile table: m_max must be positive and strictly ascending";
            return -2;
        }
        if (cols[k] != 1 && cols[k] != 2 && cols[k] != 4) {
            g_last_error = "tile table: columns per lane must be 1, 2 or 4";
            return -2;
        }
        t.emplace_back(m_max[k], cols[k]);
    }
    std::lock_guard<std::mutex> lk(g_tile_mu);
    g_tile_table = std::move(t);
    return 0;
}

int cgot_get_tile_table(int32_t* m_max, int32_t* cols, int32_t cap)
{
    std::lock_guard<std::mutex> lk(g_tile_mu);
    const int cnt = (int)g_tile_table.size();
    for (int k = 0; k < cnt && k < cap; k++) {
        if (m_max) m_max[k] = g_tile_table[k].first;
        if (cols) cols[k] = g_tile_table[k].second;
    }
    return cnt;
}

int cgot_tile_cols(int32_t m) { return m < 1 ? -1 : tile_cols(m); }

int cgot_set_max_groups(int32_t groups)
{
    if (groups < 1 || groups > kMaxGroups) {
        g_last_error = "max groups must be in [1, " + std::to_string(kMaxGroups) + "]";
        return -2;
    }
    g_max_groups = groups;
    return 0;
}


size_t cgot_adjoint_workspace_size(int32_t B, const cgot_adjoint_problem* probs)
{
    try {
        Batch bt;
        bt.B = B;
        return adjoint_layout(bt, probs, nullptr);
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return 0;
    }
}

int cgot_adjoint_solve(int32_t B, const cgot_adjoint_problem* probs, void* workspace, size_t workspace_bytes,
                       void* stream, cgot_adjoint_info* info)
{
    CG_API_BEGIN
    if (B < 1) throw std::invalid_argument("B must be >= 1");
    for (int b = 0; b < B; b++) {
        const auto& p = probs[b];
        if (p.n < 1 || p.m < 1) throw std::invalid_argument("n and m must be >= 1");
        if (!p.dual || !p.G || !p.lambda || !p.grad) throw std::invalid_argument("null pointer");
        check_cost(p.M, p.xy);
        if (!(p.reg > 0.0)) throw std::invalid_argument("reg must be positive");
    }
    Batch bt;
    bt.B = B;
    bt.stream = (hipStream_t)stream;
    const size_t need = adjoint_layout(bt, probs, nullptr);
    if (workspace == nullptr || workspace_bytes < need)
        throw std::invalid_argument("workspace too small: need " + std::to_string(need) + " bytes");
    adjoint_layout(bt, probs, workspace);
    std::vector<const double*> ptrs(3 * (size_t)B);
    for (int b = 0; b < B; b++) {
        ptrs[b] = probs[b].G;
        ptrs[B + b] = probs[b].lambda;
        ptrs[2 * B + b] = probs[b].grad;
    }
    bt.upload_maps(ptrs);
    adjoint_run(bt, probs, info);
    return 0;
    CG_API_END
}

size_t cgot_splr_precond_workspace_size(const void* factor, int32_t n, int32_t m)
{
    try {
        if (!factor) throw std::invalid_argument("factor must be non-null");
        Batch bt;
        bt.B = 1;
        cgot_adjoint_problem p{};
        p.n = n; p.m = m; p.reg = 1.0; p.factor = factor;
        return adjoint_layout(bt, &p, nullptr);
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return 0;
    }
}

int cgot_splr_precond_apply(const void* factor, int32_t n, int32_t m, const double* r, double* z,
                            void* workspace, size_t workspace_bytes, void* stream)
{
    return cgot_splr_precond_apply_profiled(factor, n, m, r, z, 0, 1, nullptr, workspace, workspace_bytes, stream);
}

int cgot_splr_precond_apply_profiled(const void* factor, int32_t n, int32_t m, const double* r, double* z,
                                     int32_t base_only, int32_t repeats, cgot_precond_profile* prof,
                                     void* workspace, size_t workspace_bytes, void* stream)
{
    CG_API_BEGIN
    using clk = std::chrono::steady_clock;
    const auto us = [](clk::time_point a, clk::time_point b) {
        return std::chrono::duration<double, std::micro>(b - a).count();
    };
    const auto t_start = clk::now();
    if (!factor || !r || !z) throw std::invalid_argument("factor, r and z must be non-null");
    if (n < 1 || m < 2) throw std::invalid_argument("n >= 1 and m >= 2 required");
    if (repeats < 1) throw std::invalid_argument("repeats must be positive");
    Batch bt;
    bt.B = 1;
    bt.stream = (hipStream_t)stream;
    cgot_adjoint_problem p{};
    p.n = n; p.m = m; p.reg = 1.0; p.factor = factor;
    const size_t need = adjoint_layout(bt, &p, nullptr);
    if (workspace == nullptr || workspace_bytes < need)
        throw std::invalid_argument("workspace too small: need " + std::to_string(need) + " bytes");
    adjoint_layout(bt, &p, workspace);
    bt.upload_maps(std::vector<const double*>(3, nullptr));
    CG_HIP(hipMemsetAsync(bt.zero_base, 0, bt.zero_bytes, bt.stream));
    CG_HIP(hipMemsetAsync(bt.d_scal, 0, sizeof(double) * CG_NSCAL, bt.stream));   // PCG status: running
    const CgInst& I = bt.h[0];
    const size_t N = (size_t)n + m - 1;
    auto c = ctl_vec(1);
    c[CTL_ACTIVE] = 1.0;
    PcSolver pc(bt);
    const double* dc = bt.ctl(c);
    // the production apply (prof == NULL): one application with r.z, as the
    // PCG runs it; the profiled variants time its stages and leave r.z out
    const bool stages = (prof != nullptr && repeats == 1);
    hipEvent_t ev[4] = {};
    if (stages)
        for (auto& e : ev) CG_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    cgot_precond_profile pr{};
    const auto t_loop = clk::now();
    for (int rep = 0; rep < repeats; rep++) {
        const auto a0 = clk::now();
        CG_HIP(hipMemcpyAsync(I.yv, r, sizeof(double) * N, hipMemcpyDeviceToDevice, bt.stream));
        const auto a1 = clk::now();
        if (prof == nullptr) {
            pc.apply(dc, 1);
        } else {
            if (stages) CG_HIP(hipEventRecord(ev[0], bt.stream));
            pc.solve(dc);
            if (stages) CG_HIP(hipEventRecord(ev[1], bt.stream));
            const auto a2 = clk::now();
            if (!base_only) {
                pc.lowrank_dots(dc);
                if (stages) CG_HIP(hipEventRecord(ev[2], bt.stream));
                pc.lowrank_apply(dc);
                if (stages) CG_HIP(hipEventRecord(ev[3], bt.stream));
            }
            const auto a3 = clk::now();
            pr.host_solve_launch_us += us(a1, a2);
            pr.host_lowrank_us += us(a2, a3);
        }
        const auto a4 = clk::now();
        CG_HIP(hipMemcpyAsync(z, I.sv, sizeof(double) * N, hipMemcpyDeviceToDevice, bt.stream));
        pr.host_memcpy_in_us += us(a0, a1);
        pr.host_memcpy_out_us += us(a4, clk::now());
    }
    bt.check_launch();
    const auto w0 = clk::now();
    bt.read_scal();   // stream sync
    const auto w1 = clk::now();
    if (bt.sc(0, CS_PIVOT) >= 2.0)
        throw std::runtime_error("internal: a block-flag wait of the preconditioner exceeded its bound");
    if (prof != nullptr) {
        pr.host_sync_wait_us = us(w0, w1);
        pr.host_total_us = us(t_start, w1);
        pr.repeated_total_us = us(t_loop, w1);
        if (stages) {
            float ms = 0.f;
            CG_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
            pr.gpu_solve_us = 1000.0 * ms;
            if (!base_only) {
                CG_HIP(hipEventElapsedTime(&ms, ev[1], ev[2]));
                pr.gpu_lowrank_reduce_us = 1000.0 * ms;
                CG_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
                pr.gpu_lowrank_correction_us = 1000.0 * ms;
            }
            for (auto& e : ev) (void)hipEventDestroy(e);
        }
        *prof = pr;
    }
    return 0;
    CG_API_END
}

}  // extern "C"

// ---------------------------------------------------------------------------
// cgot_splr_precond_apply_noalloc: the preconditioner application with its
// work buffers in a fixed device array instead of a caller's workspace, for
// callers that cannot allocate (code under graph capture, serving loops).
// The array lives in the library's code object: every GPU holds its own copy
// at its own address, taken once per device with hipGetSymbolAddress; the
// host only ever passes that address (the array's name is a host-side
// stand-in, not a device pointer).  Calls on one GPU share the array, so
// calls that overlap on one GPU race; each call clears the counters and
// scalars it relies on.
// ---------------------------------------------------------------------------
constexpr size_t kNoallocScratchBytes = (size_t)4 << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (taken on the first call
// there).
void* scratch_on_current_device()
{
    int dev = 0;
    CG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* ptr = nullptr;
        CG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (ptr == nullptr) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch");
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

// Upper bound of the precond-only layout for (n, m) whatever the factor: the
// layout without a factor, plus the factor's vectors (2N doubles for a sparse
// plan, 2P <= 2N for a dense one), its block tables and the alignment of the
// extra buffers.
size_t noalloc_bound(int n, int m)
{
    Batch bt;
    bt.B = 1;
    cgot_adjoint_problem p{};
    p.n = n; p.m = m; p.reg = 1.0;
    const size_t N = (size_t)n + m - 1;
    const size_t nblk = (N + CG_NB - 1) / CG_NB;
    return adjoint_layout(bt, &p, nullptr, true) + 16 * N + 4 * (9 * nblk + 8) + 3 * 256;
}
}  // namespace

extern "C" {

int cgot_splr_precond_apply_noalloc(const void* factor, int32_t n, int32_t m, const double* r, double* z,
                                    void* stream)
{
    CG_API_BEGIN
    if (!factor || !r || !z) throw std::invalid_argument("factor, r and z must be non-null");
    if (n < 1 || m < 2) throw std::invalid_argument("n >= 1 and m >= 2 required");
    const size_t bound = noalloc_bound(n, m);
    if (bound > kNoallocScratchBytes)
        throw std::invalid_argument("problem too large for the fixed scratch: n=" + std::to_string(n) + ", m=" +
                                    std::to_string(m) + " needs up to " + std::to_string(bound) +
                                    " bytes, the limit is " + std::to_string(kNoallocScratchBytes) + " bytes");
    char* base = (char*)scratch_on_current_device();
    Batch bt;
    bt.B = 1;
    bt.stream = (hipStream_t)stream;
    bt.pageable = true;
    cgot_adjoint_problem p{};
    p.n = n; p.m = m; p.reg = 1.0; p.factor = factor;
    const size_t need = adjoint_layout(bt, &p, nullptr, true);
    if (need > kNoallocScratchBytes)
        throw std::invalid_argument("problem too large for the fixed scratch: needs " + std::to_string(need) +
                                    " bytes, the limit is " + std::to_string(kNoallocScratchBytes) + " bytes");
    adjoint_layout(bt, &p, base, true);
    bt.upload_maps(std::vector<const double*>(3, nullptr));
    CG_HIP(hipMemsetAsync(bt.zero_base, 0, bt.zero_bytes, bt.stream));
    CG_HIP(hipMemsetAsync(bt.d_scal, 0, sizeof(double) * CG_NSCAL, bt.stream));   // PCG status: running
    const CgInst& I = bt.h[0];
    const size_t N = (size_t)n + m - 1;
    auto c = ctl_vec(1);
    c[CTL_ACTIVE] = 1.0;
    PcSolver pc(bt);
    const double* dc = bt.ctl(c);
    CG_HIP(hipMemcpyAsync(I.yv, r, sizeof(double) * N, hipMemcpyDeviceToDevice, bt.stream));
    pc.apply(dc, 1);
    CG_HIP(hipMemcpyAsync(z, I.sv, sizeof(double) * N, hipMemcpyDeviceToDevice, bt.stream));
    bt.check_launch();
    bt.read_scal();   // stream sync
    if (bt.sc(0, CS_PIVOT) >= 2.0)
        throw std::runtime_error("internal: a block-flag wait of the preconditioner exceeded its bound");
    return 0;
    CG_API_END
}

int cgot_splr_context_csr(const void* factor, int32_t n, int32_t m, double* csr_values, int32_t* csr_colind,
                          int32_t* csr_rowptr, int64_t* nnz, void* stream)
{
    CG_API_BEGIN
    if (!factor || !nnz) throw std::invalid_argument("factor and nnz must be non-null");
    CgInst F;
    CG_HIP(hipMemcpy(&F, factor, sizeof(CgInst), hipMemcpyDeviceToHost));
    if (F.n != n || F.m != m) throw std::invalid_argument("factor does not match the problem shape");
    int K = 0;
    if (m >= 2) CG_HIP(hipMemcpy(&K, F.w_colptr + (m - 1), sizeof(int), hipMemcpyDeviceToHost));
    *nnz = (int64_t)n + (m - 1) + K;
    if (csr_values != nullptr) {
        if (!csr_colind || !csr_rowptr) throw std::invalid_argument("csr_colind and csr_rowptr must be non-null");
        const int rows = std::max(n, m - 1);
        hipLaunchKernelGGL(k_ctx_csr, dim3(cdiv(rows, CG_WG)), dim3(CG_WG), 0, (hipStream_t)stream,
                           (const CgInst*)factor, csr_values, (int*)csr_colind, (int*)csr_rowptr);
        CG_HIP(hipGetLastError());
    }
    return 0;
    CG_API_END
}

size_t cgot_bcd_workspace_size(int32_t B, const cgot_bcd_problem* probs)
{
    try {
        Batch bt;
        bt.B = B;
        return bcd_layout(bt, probs, nullptr);
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return 0;
    }
}

int cgot_bcd_solve(int32_t B, const cgot_bcd_problem* probs, void* workspace, size_t workspace_bytes,
                   void* stream, int32_t* niter)
{
    CG_API_BEGIN
    if (B < 1) throw std::invalid_argument("B must be >= 1");
    for (int b = 0; b < B; b++) {
        const auto& p = probs[b];
        if (p.n < 1 || p.m < 1) throw std::invalid_argument("n and m must be >= 1");
        if (!p.a || !p.b) throw std::invalid_argument("null pointer");
        check_cost(p.M, p.xy);
        if (!(p.reg > 0.0)) throw std::invalid_argument("reg must be positive");
    }
    Batch bt;
    bt.B = B;
    bt.stream = (hipStream_t)stream;
    const size_t need = bcd_layout(bt, probs, nullptr);
    if (workspace == nullptr || workspace_bytes < need)
        throw std::invalid_argument("workspace too small: need " + std::to_string(need) + " bytes");
    bcd_layout(bt, probs, workspace);
    std::vector<const double*> ptrs(3 * (size_t)B);
    for (int b = 0; b < B; b++) {
        ptrs[3 * b] = probs[b].a;
        ptrs[3 * b + 1] = probs[b].b;
        ptrs[3 * b + 2] = probs[b].x0;
    }
    bt.upload_maps(ptrs);
    bcd_run(bt, probs, niter);
    return 0;
    CG_API_END
}

size_t cgot_lse_workspace_size(int32_t n, int32_t m)
{
    cgot_bcd_problem p{};
    p.n = n; p.m = m; p.reg = 1.0;
    p.M = p.a = p.b = (const double*)16;
    return cgot_bcd_workspace_size(1, &p);
}

int cgot_lse_update(const double* M, const double* a, const double* b, int32_t n, int32_t m, double reg,
                    int32_t which, int32_t repeats, double* dual, void* workspace, size_t workspace_bytes,
                    void* stream)
{
    CG_API_BEGIN
    if (n < 1 || m < 1 || !M || !a || !b || !dual) throw std::invalid_argument("bad arguments");
    if (!(reg > 0.0)) throw std::invalid_argument("reg must be positive");
    cgot_bcd_problem p{};
    p.M = M; p.a = a; p.b = b; p.n = n; p.m = m; p.reg = reg;
    Batch bt;
    bt.B = 1;
    bt.stream = (hipStream_t)stream;
    const size_t need = bcd_layout(bt, &p, nullptr);
    if (workspace == nullptr || workspace_bytes < need) throw std::invalid_argument("workspace too small");
    bcd_layout(bt, &p, workspace);
    bt.h[0].gamma = dual;   // update in place
    std::vector<const double*> ptrs = {a, b, dual};
    bt.upload_maps(ptrs);
    auto c = ctl_vec(1);
    c[CTL_ACTIVE] = 1.0;
    const double* dc = bt.ctl(c);
    hipLaunchKernelGGL(k_init_logab, dim3(bt.nvec), dim3(CG_WG), 0, bt.stream, bt.d_insts, bt.d_vec, dc,
                       (const double* const*)bt.d_ptrs, (const double* const*)nullptr);
    for (int r = 0; r < std::max(repeats, 1); r++) {
        if (which == 0) row_lse(bt, dc, CG_PT_GAMMA);
        else col_lse(bt, dc, CG_PT_GAMMA);
    }
    bt.check_launch();
    bt.sync();
    return 0;
    CG_API_END
}

size_t cgot_T_sparsify_workspace_size(int32_t n, int32_t m, int64_t K)
{
    cgot_splr_problem p{};
    p.n = n; p.m = m; p.reg = 1.0;
    p.M = p.a = p.b = (const double*)16;
    const double Te = (double)n * (double)(m - 1);
    p.density_max = (Te > 0) ? std::min(1.0, (std::max((double)K, 1.0) + 0.5) / Te) : 1.0;
    return cgot_splr_workspace_size(1, &p);
}

int cgot_T_sparsify(const double* gamma, const double* M, const double* a, const double* b, int32_t n, int32_t m,
                    double reg, double shift, int64_t K, double* rowsum, double* colsum, double* grad,
                    double* scalars, double* csr_values, int32_t* csr_colind, int32_t* csr_rowptr,
                    const double* rhs, double* sol, void* workspace, size_t workspace_bytes, void* stream)
{
    CG_API_BEGIN
    if (n < 1 || m < 1 || !gamma || !M || !a || !b) throw std::invalid_argument("bad arguments");
    cgot_splr_problem p{};
    p.M = M; p.a = a; p.b = b; p.n = n; p.m = m; p.reg = reg; p.tol = 0; p.max_iter = 1;
    const double Te = (double)n * (double)(m - 1);
    p.density_max = (Te > 0) ? std::min(1.0, (std::max((double)K, 1.0) + 0.5) / Te) : 1.0;
    p.shift_max = 0.0;
    p.sparsity_pattern_cycle = 30;
    p.candidate_sinkhorn_iter = 0;
    p.x0 = gamma;
    Batch bt;
    bt.B = 1;
    bt.stream = (hipStream_t)stream;
    const size_t need = splr_layout(bt, &p, nullptr);
    if (workspace == nullptr || workspace_bytes < need) throw std::invalid_argument("workspace too small");
    splr_layout(bt, &p, workspace);
    std::vector<const double*> ptrs = {a, b, gamma};
    bt.upload_maps(ptrs);
    SplrSolver s(bt);
    s.hi.assign(1, HostInst{});
    s.init_inst(0, p);
    s.zero_scratch();
    std::vector<char> all(1, 1);
    const double* dc = bt.ctl(s.mask(all));
    hipLaunchKernelGGL(k_init_logab, dim3(bt.nvec), dim3(CG_WG), 0, bt.stream, bt.d_insts, bt.d_vec, dc,
                       (const double* const*)bt.d_ptrs, (const double* const*)nullptr);
    tpass(bt, dc, CG_PT_GAMMA);
    // top-K with exactly K' = min(max(K,1), n(m-1)) entries; shift added as given
    HostInst& H = s.hi[0];
    H.Kmax = std::max<long long>(1, std::min<long long>(std::max<long long>(K, 1), H.Te));
    H.density = (H.Te > 0) ? ((double)H.Kmax + 0.5) / (double)H.Te : 0.0;
    H.gnorm = shift / reg;
    H.shift_max = shift / reg;
    s.sphess(all, all);
    const CgInst& I = bt.h[0];
    hipLaunchKernelGGL(k_hsl_csr, dim3(bt.nvec), dim3(CG_WG), 0, bt.stream, bt.d_insts, bt.d_vec, dc, csr_values,
                       (int*)csr_colind, (int*)csr_rowptr);
    CG_HIP(hipMemcpyAsync(rowsum, I.rowsum, sizeof(double) * n, hipMemcpyDeviceToDevice, bt.stream));
    CG_HIP(hipMemcpyAsync(colsum, I.colsum, sizeof(double) * m, hipMemcpyDeviceToDevice, bt.stream));
    CG_HIP(hipMemcpyAsync(grad, I.grad, sizeof(double) * I.N, hipMemcpyDeviceToDevice, bt.stream));
    if (rhs != nullptr && sol != nullptr) {
        CG_HIP(hipMemcpyAsync(I.grad, rhs, sizeof(double) * I.N, hipMemcpyDeviceToDevice, bt.stream));
        CG_HIP(hipMemcpyAsync(I.yv, rhs, sizeof(double) * I.N, hipMemcpyDeviceToDevice, bt.stream));
        s.search_direction(all, dc);
        CG_HIP(hipMemcpyAsync(sol, I.dir, sizeof(double) * I.N, hipMemcpyDeviceToDevice, bt.stream));
    }
    bt.check_launch();
    bt.read_scal();
    if (scalars) {
        double hs[2] = {bt.sc(0, CS_F), std::sqrt(bt.sc(0, CS_G2))};
        CG_HIP(hipMemcpyAsync(scalars, hs, sizeof(hs), hipMemcpyHostToDevice, bt.stream));
        CG_HIP(hipStreamSynchronize(bt.stream));
    }
    return 0;
    CG_API_END
}

}  // extern "C"
