// Host side of libsoapfilm.so: context, packed device layout, stage
// orchestration and the C ABI declared in include/soapfilm.h.
//
// One context = one HIP stream on one device with the packed state of any
// number of bubbles (the reference holds one bubble per process in
// __constant__ symbols, bubble.cu:6-25; here every parameter lives in a
// per-bubble descriptor array in device memory, so bubbles of different
// resolution and parameters advance together in one launch per stage).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/soapfilm.h"
#include "sf_launch.h"
#include "sf_host.h"
#include "sf_internal.h"

using namespace sfk;

namespace {

thread_local std::string g_last_error;

struct HipError {
    std::string msg;
};

#define SF_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            throw HipError{std::string(#call) + ": " + hipGetErrorString(e_)};          \
    } while (0)

struct Status {
    int code;
    std::string msg;
};

constexpr double kPi = 3.14159265358979323846;

// A caller-provided device region that the set-up carves its arrays from
// instead of allocating them (the entry point without allocation): a bump
// pointer with 256-byte alignment, each array cleared on the caller's stream.
struct Arena {
    char* base = nullptr;
    size_t size = 0, used = 0;
    hipStream_t stream = nullptr;
};
thread_local Arena* g_arena = nullptr;

struct ArenaScope {
    explicit ArenaScope(Arena* a) { g_arena = a; }
    ~ArenaScope() { g_arena = nullptr; }
};

template <typename X>
X* dalloc(size_t n, std::vector<void*>& owned) {
    void* p = nullptr;
    if (n == 0) n = 1;
    if (g_arena) {
        const size_t at = (g_arena->used + 255) & ~(size_t)255, bytes = n * sizeof(X);
        if (at + bytes > g_arena->size)
            throw Status{SF_ERR_UNSUPPORTED, "the work arrays need more than the " +
                                                 std::to_string(g_arena->size >> 20) +
                                                 " MiB of the fixed device array"};
        p = g_arena->base + at;
        g_arena->used = at + bytes;
        // cleared and finished before the set-up's synchronous copies write into it
        SF_HIP(hipMemsetAsync(p, 0, bytes, g_arena->stream));
        SF_HIP(hipStreamSynchronize(g_arena->stream));
        return static_cast<X*>(p);
    }
    hipError_t e = hipMalloc(&p, n * sizeof(X));
    if (e != hipSuccess) throw Status{SF_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)};
    owned.push_back(p);
    // the clear goes to the null stream, which does not order against the
    // context's non-blocking stream: it is finished before the array is
    // handed out (arrays allocated lazily are written by the context's
    // kernels right after)
    SF_HIP(hipMemset(p, 0, n * sizeof(X)));
    SF_HIP(hipDeviceSynchronize());
    return static_cast<X*>(p);
}

// The XCD-aware order of the work items of a launch (defined below).
template <typename X>
std::vector<X> xcd_order(const std::vector<X>& work, X pad, int nxcd);

// A schedule longer than this many inner iterations in a refinement (a solve
// far from converging, NaN input, or a stiff operator for Chebyshev) runs
// host-driven instead of as a graph of hundreds of mostly skipped launches.
constexpr int kCaptureMaxInner = 48;

// Everything that does not depend on the precision.
struct Base {
    sf_options opt{};
    int device = 0;
    int nb = 0;
    std::vector<sf_bubble_desc> desc;
    std::vector<sf_derived> der;
    std::vector<BubbleDev> bub_h;
    std::vector<BubbleTime> time_h;
    long long ncells = 0;
    int nrow_total = 0;
    long stride = 0;          // partial-sum stride (rows)
    hipStream_t stream = nullptr;
    std::vector<void*> owned;
    BubbleDev* bub_d = nullptr;
    BubbleTime* time_d = nullptr;
    SolveCtl* ctl_d = nullptr;
    SolveHist* hist_d = nullptr;
    SchedStat* sstat_d = nullptr;          // schedule statistics of the captured step
    SchedStat* sstat_h = nullptr;          // pinned host copies, two per bubble (by chunk parity)
    hipEvent_t cev[2] = {};                // end of a chunk of replays
    int chunk_sched[2] = {0, 0};           // first-refinement iterations each chunk ran with
    hipStream_t cstream = nullptr;         // capture stream (graphs are built while replays run)
    // executable graphs by (schedule, steps per launch, statistics slot),
    // with the tick of their last use (at most graph_cache_max kept)
    struct GraphEntry {
        hipGraphExec_t exec;
        unsigned long long used;
    };
    std::map<std::vector<int>, GraphEntry> graphs;
    unsigned long long graph_tick = 0;
    size_t graph_cache_max = 16;
    Tables tab{};
    LaunchGeom geom{};            // cell tiles
    std::vector<LaunchGeom> rgeo; // row blocks, one geometry per row width (multigrid)
    std::vector<int> rgeo_nph;            // the row width of each entry of rgeo
    // the row stages' launches: one per group of row widths that share a
    // cell count a thread (every width, in practice), narrow rows packed
    // several to a block; pgeo_inner the same rows for the Chebyshev
    // launches, pgeo_width the widest row of each group
    std::vector<LaunchGeom> pgeo, pgeo_inner;
    std::vector<int> pgeo_width;
    std::vector<std::vector<int>> pgeo_widths;   // the row widths of each group
    LaunchGeom sgeo{};            // one row a block, every row (row statistics)
    size_t lds_limit = 0;         // dynamic LDS a block may take
    // autotuning: the candidate cell-tile lists (widths 256, 128, 64 columns;
    // each row shorter than a width keeps its own power-of-two width) and the
    // choices made
    struct TileSet {
        int width;
        const int4* tiles;
        int ntiles;
    };
    std::vector<TileSet> tile_cands;
    bool tuned = false;
    sf_tuning tuning{};
    double* partR = nullptr;
    double* partA = nullptr;
    double* partB = nullptr;
    double* partS = nullptr;
    double* x64[2] = {nullptr, nullptr};
    double* b64 = nullptr;
    int* rsteps = nullptr;
    double* acc = nullptr;        // per-launch sums of the solve (acc_index layout)
    int nxcd = 1;
    // BIMOCQ scalar advection (allocated only when selected); its state is
    // (re)initialised from the current fields at the first step after an
    // upload of eta or Gamma
    bool bimocq = false;
    bool bimocq_dirty = true;
    BimocqPtrs bm{};
    SolveParams sp{};
    int max_outer = 3, max_inner = 500;
    bool calibrated = false;
    bool recording = false;   // host-driven solves record their inner iterations
    int calib_steps = 0;
    // extrapolated initial guess of the Gamma solve; host_hist counts the
    // steps since the state or dt last changed (the device keeps its own
    // count per bubble)
    bool guess = false;
    int host_hist = 0;
    std::vector<int> calib;   // max inner iterations per refinement seen while calibrating
    std::vector<int> sched;   // inner iterations per refinement of the fixed schedule
    int method = 1;           // inner solver: 1 PCG, 2 Chebyshev
    bool mg = false;          // PCG preconditioned by the multigrid V-cycle
    bool trace_solve = std::getenv("SF_SOLVE_TRACE") != nullptr;
    int steps_done = 0;
    int stage = 0;            // 0 idle, 1 advected, 2 assembled, 3 solved
    hipEvent_t ev[6] = {};
    sf_timing timing{};
    std::string err;
    // set up inside a caller's device array on the caller's stream: no
    // allocation, no stream, event, pinned buffer or graph of its own
    bool noalloc = false;
    virtual ~Base() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.second.exec);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : cev)
            if (e) (void)hipEventDestroy(e);
        if (sstat_h) (void)hipHostFree(sstat_h);
        if (cstream) (void)hipStreamDestroy(cstream);
        for (void* p : owned) (void)hipFree(p);
        if (stream && !noalloc) (void)hipStreamDestroy(stream);
    }
    virtual void advect() = 0;
    // applies what uploads left pending (BIMOCQ's maps and initial fields
    // restart from an uploaded eta or Gamma) as eager launches, before a
    // step replays or captures graphs that do not hold them
    virtual void apply_uploads() = 0;
    virtual void assemble() = 0;
    virtual void solve(bool host_driven, bool finalize) = 0;
    virtual void update(bool finalize) = 0;
    virtual void upload(int b, int field, const double* h) = 0;
    virtual void download(int b, int field, double* h) = 0;
    // the frame output's binary16 gray plane of bubble b (eta H 2e5)
    virtual void thickness_half(int b, uint16_t* h) = 0;
    virtual void row_stats() = 0;
    // one inner Chebyshev launch (iteration `it` of refinement `kr`)
    virtual void inner_launch(int it, int kr) = 0;
    // average time of n launches of the advection / of inner iteration 2 of
    // refinement 0 with geometry g (call between the assembly and the solve)
    virtual double time_advect_geom(const LaunchGeom& g, int n) = 0;
    virtual double time_cheb_geom(const LaunchGeom& g, int n) = 0;
    virtual bool fp64_inner() const = 0;
    // switches the inner solve to PCG with the multigrid preconditioner
    virtual void enable_mg() = 0;
    // the automatic inner method (Chebyshev) may switch to multigrid when
    // a refinement needs more inner iterations than a captured step holds
    bool auto_method() const { return opt.inner_method == 0 && !fp64_inner(); }
    // drops the stored solve corrections of bubble b (all when b < 0)
    void reset_guess(int b) {
        host_hist = 0;
        if (!guess) return;
        std::vector<BubbleTime> t(nb);
        SF_HIP(hipStreamSynchronize(stream));
        SF_HIP(hipMemcpy(t.data(), time_d, nb * sizeof(BubbleTime), hipMemcpyDeviceToHost));
        for (int k = 0; k < nb; ++k)
            if (b < 0 || b == k) t[k].ghist = t[k].gpend = 0;
        SF_HIP(hipMemcpy(time_d, t.data(), nb * sizeof(BubbleTime), hipMemcpyHostToDevice));
    }
    // the control blocks copied on the context's stream, after the launches
    // queued there (a copy on the null stream is not ordered with them: the
    // stream is non-blocking)
    void ctl_to_host(std::vector<SolveCtl>& h) {
        h.resize((size_t)nb);
        SF_HIP(hipMemcpyAsync(h.data(), ctl_d, nb * sizeof(SolveCtl), hipMemcpyDeviceToHost, stream));
        SF_HIP(hipStreamSynchronize(stream));
    }
    void ctl_to_device(const std::vector<SolveCtl>& h) {
        SF_HIP(hipMemcpyAsync(ctl_d, h.data(), nb * sizeof(SolveCtl), hipMemcpyHostToDevice, stream));
        SF_HIP(hipStreamSynchronize(stream));
    }
};

template <typename T, typename Ti>
struct Impl : Base {
    StatePtrs<T> cur{}, nxt{};
    T *ua0 = nullptr, *va0 = nullptr, *au = nullptr, *av = nullptr;
    Ti *wN = nullptr, *wW = nullptr, *dg = nullptr, *e = nullptr, *r = nullptr, *z = nullptr,
       *Ap = nullptr;
    Ti* p[2] = {nullptr, nullptr};
    T *gdel = nullptr, *gnext = nullptr;   // last solve correction, next guess

    double *vdiv = nullptr, *vfu = nullptr, *vfv = nullptr;   // VISCOUS div, forces

    // multigrid preconditioner: the arrays of every level (level 0 is the
    // assembled operator itself, its iterate z) and, per level and row
    // width, the row lists of the half-sweeps: even / odd rows of bubbles
    // with a coarser level, even / odd rows of bubbles whose coarsest level
    // this is, all rows (coarse operator)
    struct MgLaunch {
        LaunchGeom even{}, odd{}, ceven{}, codd{}, all{};
    };
    std::vector<std::vector<MgLaunch>> mgl;
    std::vector<MgPtrs<Ti>> mgp;
    int mg_levels = 0;
    AirPtrs<T> air() const { return AirPtrs<T>{ua0, va0, au, av, vfu, vfv}; }
    SysPtrs<Ti> sys(double* x_init) const {
        return SysPtrs<Ti>{b64, x_init, wN, wW, dg, e, r, z, Ap, rsteps, acc};
    }
    bool fp64_inner() const override { return sizeof(Ti) == 8; }
    void enable_mg() override {
        if (mgp.empty()) alloc_mg();
        mg = true;
        method = 1;
    }

    void alloc_fields() {
        const size_t n = (size_t)ncells;
        for (StatePtrs<T>* s : {&cur, &nxt}) {
            s->uth = dalloc<T>(n, owned);
            s->uph = dalloc<T>(n, owned);
            s->eta = dalloc<T>(n, owned);
            s->gam = dalloc<T>(n, owned);
        }
        ua0 = dalloc<T>(n, owned);
        va0 = dalloc<T>(n, owned);
        au = dalloc<T>(n, owned);
        av = dalloc<T>(n, owned);
        wN = dalloc<Ti>(n, owned);
        wW = dalloc<Ti>(n, owned);
        dg = dalloc<Ti>(n, owned);
        e = dalloc<Ti>(n, owned);
        r = dalloc<Ti>(n, owned);
        z = dalloc<Ti>(n, owned);
        Ap = dalloc<Ti>(n, owned);
        p[0] = dalloc<Ti>(n, owned);
        p[1] = dalloc<Ti>(n, owned);
        if (guess) {
            gdel = dalloc<T>(n, owned);
            gnext = dalloc<T>(n, owned);
        }
        if (opt.viscous) {
            vdiv = dalloc<double>(n, owned);
            vfu = dalloc<double>(n, owned);
            vfv = dalloc<double>(n, owned);
        }
        // reference initWithConst (solver.cu:106-110): u = 0, eta = 0.5, Gamma = 1
        // (the entry point without allocation copies the caller's state in)
        std::vector<double> h;
        for (int b = 0; b < nb && !noalloc; ++b) {
            const size_t N = (size_t)bub_h[b].nth * bub_h[b].nph;
            h.assign(N, 0.5);
            upload(b, SF_FIELD_ETA, h.data());
            h.assign(N, 1.0);
            upload(b, SF_FIELD_GAMMA, h.data());
        }
    }

    void alloc_mg() {
        // rows per level of each bubble: halved while the count is even
        std::vector<std::vector<int>> nl(nb);
        for (int b = 0; b < nb; ++b) {
            int n = bub_h[b].nth;
            nl[b].push_back(n);
            while (n % 2 == 0 && n > 1) {
                n /= 2;
                nl[b].push_back(n);
            }
            mg_levels = std::max(mg_levels, (int)nl[b].size());
        }
        mgp.assign(mg_levels, MgPtrs<Ti>{});
        mgl.assign(mg_levels, std::vector<MgLaunch>(rgeo.size()));
        for (int l = 0; l < mg_levels; ++l) {
            std::vector<MgLevelDev> lv(nb, MgLevelDev{0, 0, 0});
            long long tot = 0;
            for (int b = 0; b < nb; ++b) {
                if (l >= (int)nl[b].size()) continue;
                lv[b].n = nl[b][l];
                lv[b].off = l == 0 ? bub_h[b].off : tot;
                tot += ((long long)nl[b][l] * bub_h[b].nph + 63) / 64 * 64;
            }
            MgLevelDev* lv_d = dalloc<MgLevelDev>(nb, owned);
            SF_HIP(hipMemcpy(lv_d, lv.data(), nb * sizeof(MgLevelDev), hipMemcpyHostToDevice));
            MgPtrs<Ti>& P = mgp[l];
            P.lv = lv_d;
            if (l == 0) {
                P.wN = wN; P.wW = wW; P.dg = dg; P.x = z;
            } else {
                const size_t n = (size_t)tot;
                P.wN = dalloc<Ti>(n, owned);
                P.wW = dalloc<Ti>(n, owned);
                P.dg = dalloc<Ti>(n, owned);
                P.x = dalloc<Ti>(n, owned);
                P.U = dalloc<Ti>(n, owned);
                P.D = dalloc<Ti>(n, owned);
            }
            for (size_t w = 0; w < rgeo.size(); ++w) {
                std::vector<int2> lists[5];
                for (int b = 0; b < nb; ++b) {
                    if (bub_h[b].nph != rgeo_nph[w] || l >= (int)nl[b].size()) continue;
                    const bool coarsest = l == (int)nl[b].size() - 1;
                    for (int j = 0; j < nl[b][l]; ++j) {
                        lists[(coarsest ? 2 : 0) + (j & 1)].push_back(make_int2(b, j));
                        lists[4].push_back(make_int2(b, j));
                    }
                }
                LaunchGeom* gs[5] = {&mgl[l][w].even, &mgl[l][w].odd, &mgl[l][w].ceven,
                                     &mgl[l][w].codd, &mgl[l][w].all};
                for (int k = 0; k < 5; ++k) {
                    *gs[k] = rgeo[w];
                    gs[k]->nrows = 0;
                    gs[k]->rows = nullptr;
                    if (lists[k].empty() || (k == 4 && l == 0)) continue;
                    std::vector<int2> rows = xcd_order(lists[k], make_int2(-1, 0), nxcd);
                    int2* rows_d = dalloc<int2>(rows.size(), owned);
                    SF_HIP(hipMemcpy(rows_d, rows.data(), rows.size() * sizeof(int2),
                                     hipMemcpyHostToDevice));
                    gs[k]->rows = rows_d;
                    gs[k]->nrows = (int)rows.size();
                }
            }
        }
    }

    // the coarse operators of levels 1.. from the assembled one
    void mg_coarsen() {
        for (int l = 1; l < mg_levels; ++l)
            for (const MgLaunch& q : mgl[l])
                if (q.all.nrows) launch_mg_coarsen<Ti>(stream, q.all, bub_d, mgp[l - 1], mgp[l]);
    }

    // z <- one V(1,1) cycle applied to r (the even rows of z already hold
    // the line-Jacobi solve of r, the cycle's first half-sweep). `start`: at
    // the start of a refinement (its first launch ends a solve whose FP64
    // residual already meets the tolerance); else inside the inner PCG
    // (launches return once the inner solve is done). The r.z partials of
    // level 0 go to rz_out.
    void mg_vcycle(bool start, double* rz_out) {
        const int nw = (int)rgeo.size();
        std::vector<char> first((size_t)nw, 1);
        auto line = [&](int l, int w, const LaunchGeom& g, int flags) {
            if (g.nrows == 0) return;
            if (start && first[(size_t)w]) flags |= kMgGateStart;
            if (!start) flags |= kMgGateInner;
            first[(size_t)w] = 0;
            const MgPtrs<Ti> co = l + 1 < mg_levels ? mgp[l + 1] : MgPtrs<Ti>{};
            launch_mg_line<Ti>(stream, g, bub_d, ctl_d, mgp[l], co, r, l, flags, rsteps, sp, partR,
                               rz_out, stride);
        };
        for (int l = 0; l < mg_levels; ++l)
            for (int w = 0; w < nw; ++w) {
                const MgLaunch& q = mgl[l][w];
                const int rz = l == 0 ? kMgRz : 0;
                if (l > 0) line(l, w, q.even, kMgZeroNb);
                line(l, w, q.odd, kMgRestrict);
                // the coarsest level of some bubbles: one row is solved
                // exactly; more rows take symmetric zebra passes (even, odd,
                // even, odd, even) from zero
                if (l > 0) line(l, w, q.ceven, kMgZeroNb | (q.codd.nrows ? 0 : rz));
                if (q.codd.nrows) {
                    line(l, w, q.codd, 0);
                    line(l, w, q.ceven, 0);
                    line(l, w, q.codd, rz);
                    line(l, w, q.ceven, rz);
                }
            }
        for (int l = mg_levels - 2; l >= 0; --l)
            for (int w = 0; w < nw; ++w) {
                const MgLaunch& q = mgl[l][w];
                const int rz = l == 0 ? kMgRz : 0;
                line(l, w, q.odd, kMgProlong | rz);
                line(l, w, q.even, rz);
            }
    }

    T state_gmax(double gmax) const {
        T v = (T)gmax;
        while ((double)v <= gmax - 1e-7) v = std::nextafter(v, (T)INFINITY);
        return v;
    }

    void alloc_bimocq() {
        const size_t n = (size_t)ncells;
        for (UMap* M : {&bm.fwd, &bm.bwd, &bm.prev, &bm.nfwd, &bm.nbwd}) {
            M->x = dalloc<double>(n, owned);
            M->y = dalloc<double>(n, owned);
            M->z = dalloc<double>(n, owned);
        }
        double** arrays[] = {&bm.e_init,
                             &bm.e_initlast, &bm.e_delta, &bm.e_deltalast, &bm.g_init,
                             &bm.g_initlast, &bm.g_delta, &bm.g_deltalast, &bm.chg_e, &bm.chg_g,
                             &bm.err_e, &bm.err_g, &bm.cpy_e, &bm.cpy_g};
        for (double** a : arrays) *a = dalloc<double>(n, owned);
        bm.tab = tab;
        bm.bt = time_d;
        bm.dist_bits = dalloc<unsigned>((size_t)nb, owned);
        bm.flag = dalloc<int>((size_t)nb, owned);
        bm.dist_last = dalloc<double>((size_t)nb, owned);
        bm.reinits = dalloc<int>((size_t)nb, owned);
        SF_HIP(hipMemset(bm.dist_bits, 0, sizeof(unsigned) * nb));
        SF_HIP(hipMemset(bm.flag, 0, sizeof(int) * nb));
        SF_HIP(hipMemset(bm.dist_last, 0, sizeof(double) * nb));
        SF_HIP(hipMemset(bm.reinits, 0, sizeof(int) * nb));
    }

    // Solver::advection (bubble.cu:1171-1214); with BIMOCQ the map updates
    // of stepForward (solver.cu:265-268) come first and eta is replaced by
    // its BIMOCQ value.
    void apply_uploads() override {
        if (!bimocq || !bimocq_dirty) return;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        SF_HIP(hipStreamIsCapturing(stream, &cs));
        if (cs != hipStreamCaptureStatusNone)
            throw Status{SF_ERR_ORDER, "BIMOCQ restart pending during graph capture"};
        launch_bimocq_init<T>(stream, geom, bub_d, cur, bm);
        bimocq_dirty = false;
    }
    void advect() override {
        if (bimocq) {
            apply_uploads();
            launch_bimocq_maps<T>(stream, geom, bub_d, time_d, tab, cur, bm);
        }
        launch_advect<T>(stream, geom, bub_d, time_d, tab, cur, nxt);
        if (bimocq) launch_bimocq_eta<T>(stream, geom, bub_d, cur, nxt, bm);
    }

    // Assembly fused with the first FP64 residual and line solve; x0 is not
    // stored (the first residual launch forms it again).
    void assemble() override {
        // VISCOUS: div(u*) and the viscous forces first (Solver::bodyforce,
        // bubble.cu:2405-2423)
        if (vfu) launch_viscous<T>(stream, geom, bub_d, tab, nxt, vdiv, vfu, vfv);
        // the broken-cell value of Gamma* is the state-precision gamma_max,
        // one value for a context whose bubbles share Gamma_m and T
        for (const LaunchGeom& rg : pgeo)
            launch_assemble<T, Ti>(stream, rg, bub_d, time_d, tab, nxt, air(), sys(x64[0]),
                                   (double)state_gmax(bub_h[0].gamma_max), ctl_d, partR, stride,
                                   gnext, sp);
        if (mg) {
            mg_coarsen();
            mg_vcycle(true, partR + 2 * stride);
        }
    }

    // FP64 residual evaluation k >= 1: x64[k&1] = x64[(k-1)&1] + ecorr (k = 1:
    // x_0 formed again from Gamma* and the guess, as the assembly formed it); sh
    // (captured steps) collects the schedule statistics. `last`: no line
    // solve of the new residual (no refinement follows, or only one whose
    // residual launch forms its correction itself); `check_prev`: refinement
    // k - 1 ran no inner iteration, so this launch tests its start residual;
    // ecorr = nullptr: the correction is M^-1 r of the stored residual,
    // formed by this launch.
    void resid(int k, Ti* ecorr, SchedStat* sh, bool last = false, bool check_prev = false) {
        const double* xo = x64[(k + 1) & 1];
        double* xn = x64[k & 1];
        SysPtrs<Ti> s = sys(x64[0]);
        s.e = ecorr;
        if (last) s.z = nullptr;
        for (const LaunchGeom& rg : pgeo)
            launch_resid<T, Ti>(stream, rg, bub_d, time_d, tab, nxt, ctl_d, s, gnext, xo, xn, k,
                                partR, stride, partB, sp, method == 2 ? sh : nullptr,
                                check_prev ? 1 : 0);
        if (mg && !last) mg_vcycle(true, partR + 2 * stride);
    }
    // Chebyshev iterates: e_1 is z; e_k for k >= 2 rotates over e, Ap, p[0]
    // (k mod 3), so e_{k-1}, e_k and e_{k+1} are distinct.
    Ti* cheb_e(int k) {
        if (k <= 1) return z;
        Ti* const bufs[3] = {e, Ap, p[0]};
        return bufs[k % 3];
    }
    void cheb(int it, int kr) {
        for (const LaunchGeom& rg : pgeo_inner)
            launch_cheb<Ti>(stream, rg, bub_d, ctl_d, sys(x64[0]), cheb_e(it), cheb_e(it - 1),
                            cheb_e(it + 1), it, kr, sp, partB, stride);
    }
    void inner_launch(int it, int kr) override { cheb(it, kr); }
    double time_advect_geom(const LaunchGeom& g, int n) override {
        launch_advect<T>(stream, g, bub_d, time_d, tab, cur, nxt);
        SF_HIP(hipEventRecord(ev[0], stream));
        for (int k = 0; k < n; ++k) launch_advect<T>(stream, g, bub_d, time_d, tab, cur, nxt);
        SF_HIP(hipEventRecord(ev[5], stream));
        SF_HIP(hipEventSynchronize(ev[5]));
        float ms = 0.f;
        SF_HIP(hipEventElapsedTime(&ms, ev[0], ev[5]));
        return (double)ms / n;
    }
    double time_cheb_geom(const LaunchGeom& g, int n) override {
        auto one = [&] {
            launch_cheb<Ti>(stream, g, bub_d, ctl_d, sys(x64[0]), cheb_e(2), cheb_e(1), cheb_e(3), 2,
                            0, sp, partB, stride);
        };
        one();
        SF_HIP(hipEventRecord(ev[0], stream));
        for (int k = 0; k < n; ++k) one();
        SF_HIP(hipEventRecord(ev[5], stream));
        SF_HIP(hipEventSynchronize(ev[5]));
        float ms = 0.f;
        SF_HIP(hipEventElapsedTime(&ms, ev[0], ev[5]));
        return (double)ms / n;
    }
    // one accumulator slot of every bubble, copied to the host
    std::vector<double> slot_h;
    void fetch_slot(int slot) {
        const size_t w = sizeof(double) * kAccVals * kAccSub;
        slot_h.resize((size_t)nb * kAccVals * kAccSub);
        SF_HIP(hipMemcpy2DAsync(slot_h.data(), w, acc + acc_index(0, slot, 0, 0),
                                sizeof(double) * kAccBubble, w, nb, hipMemcpyDeviceToHost, stream));
        SF_HIP(hipStreamSynchronize(stream));
    }
    // the largest sub-accumulator of value v of bubble b (a maximum slot)
    double slot_max(int b, int v) const {
        double m = 0.0;
        for (int q = 0; q < kAccSub; ++q) m = std::max(m, slot_h[((size_t)b * kAccVals + v) * kAccSub + q]);
        return m;
    }
    // Chebyshev on the line-Jacobi interval [1 - mu, 1 + mu] reduces the
    // residual by about mu / (1 + sqrt(1 - mu^2)) an iteration: above mu =
    // 0.96 the FP32 inner solve's 1e-6 takes more than the captured
    // schedule holds (kCaptureMaxInner). The bound of the assembled system.
    bool stiff_for_chebyshev() {
        fetch_slot(acc_slot_res(0));
        for (int b = 0; b < nb; ++b)
            if (slot_max(b, 4) > 0.96) return true;
        return false;
    }
    double slot_at(int b, int v) const {
        double s = 0.0;
        for (int q = 0; q < kAccSub; ++q) s += slot_h[((size_t)b * kAccVals + v) * kAccSub + q];
        return s;
    }
    // per bubble: |r|^2, |b|^2 and |r_in|^2 at the start of refinement k,
    // and whether that residual met the tolerance
    struct Start { double rr, bb, rin; bool done; };
    std::vector<Start> refinement_start(int k) {
        std::vector<Start> st(nb);
        fetch_slot(acc_slot_res(k));
        for (int b = 0; b < nb; ++b) {
            st[b].rr = slot_at(b, 0);
            st[b].rin = slot_at(b, 3);
            const double bb = slot_at(b, 1);
            const double res = sp.absolute ? std::sqrt(st[b].rr)
                                           : std::sqrt(st[b].rr) / (bb > 0.0 ? std::sqrt(bb) : 1.0);
            st[b].done = res <= sp.tol;
        }
        fetch_slot(acc_slot_res(0));
        for (int b = 0; b < nb; ++b) st[b].bb = slot_at(b, 1);
        return st;
    }
    // The inner solve of a refinement is done for a bubble when its inner
    // residual meets the final tolerance (times inner_frac) or has fallen by
    // the factor inner_rho from the refinement's start.
    // ... or, for a bubble whose line-Jacobi bound mu is at most 0.9, when
    // it stagnates at the FP32 rounding floor (a target the FP32 iteration
    // cannot reach, such as an absolute tolerance far below the rounding of
    // |b|): |r|^2 has fallen by less than a factor 2 over the last eight
    // iterations, where Chebyshev with mu <= 0.9 reduces it by more than
    // 1e4 (the FP64 refinement continues from there; closer to mu = 1 the
    // Chebyshev residual is not monotone enough for this test). `hist`
    // holds each bubble's |r|^2 of the iterations so far, `mu` the bounds.
    bool cheb_done(const std::vector<Start>& st, std::vector<double>& pb,
                   std::vector<std::vector<double>>& hist, const std::vector<double>& mu) {
        SF_HIP(hipMemcpyAsync(pb.data(), partB + stride, sizeof(double) * stride,
                              hipMemcpyDeviceToHost, stream));
        SF_HIP(hipStreamSynchronize(stream));
        bool all = true;
        for (int b = 0; b < nb; ++b) {
            if (st[b].done) continue;
            double rr = 0.0;
            for (int j = 0; j < bub_h[b].nth; ++j) rr += pb[bub_h[b].pr0 + j];
            hist[(size_t)b].push_back(rr);
            const std::vector<double>& h = hist[(size_t)b];
            const double tgt = sp.inner_frac * sp.tol;
            const bool stalled = mu[(size_t)b] <= 0.9 && h.size() > 8 && !(rr < 0.5 * h[h.size() - 9]);
            if (!(rr <= tgt * tgt * (sp.absolute ? 1.0 : st[b].bb) || rr <= sp.inner_rho2 * st[b].rin ||
                  stalled))
                all = false;
        }
        return all;
    }
    void inner(int it) {
        for (const LaunchGeom& rg : pgeo)
            launch_spmv<Ti>(stream, rg, bub_d, ctl_d, sys(x64[0]), p[(it + 1) & 1], p[it & 1], it, sp,
                            partR, partB, partA, stride);
        cgstep(it);
    }
    // x, r update and z = M^-1 r of inner PCG iteration `it` (with the
    // multigrid preconditioner: the line-Jacobi z is the V-cycle's first
    // half-sweep)
    void cgstep(int it) {
        for (const LaunchGeom& rg : pgeo)
            launch_cgstep<Ti>(stream, rg, bub_d, ctl_d, sys(x64[0]), p[it & 1], it, partA, partB,
                              stride);
        if (mg) mg_vcycle(false, partB);
    }

    bool all_flags(std::vector<SolveCtl>& c, bool inner_flag) {
        SF_HIP(hipMemcpyAsync(c.data(), ctl_d, sizeof(SolveCtl) * nb, hipMemcpyDeviceToHost, stream));
        SF_HIP(hipStreamSynchronize(stream));
        for (int b = 0; b < nb; ++b) {
            if (c[b].outer_done) continue;
            if (!inner_flag || !c[b].inner_done) return false;
        }
        return true;
    }

    // Refinement loop after the assembly. Host-driven: convergence flags are
    // read back after each residual and every four inner iterations (the
    // plain path, and the calibration steps of the graph schedule). Fixed:
    // the calibrated schedule, every kernel launched, converged bubbles
    // skip the work (graph capture).
    void solve(bool host_driven, bool finalize) override {
        if (host_driven && method == 2 && auto_method() && stiff_for_chebyshev()) {
            // the automatic inner method continues with multigrid: its
            // coarse levels and first cycle for this solve, then the PCG
            // branch below
            if (trace_solve) std::fprintf(stderr, "[sf solve] stiff operator: multigrid from here on\n");
            enable_mg();
            mg_coarsen();
            mg_vcycle(true, partR + 2 * stride);
        }
        std::vector<int> used;
        if (method == 2) {
            // Chebyshev: kernel `it` measures the residual of e_it and
            // produces e_{it+1}; a refinement of n kernels hands e_{n+1} on
            if (host_driven) {
                std::vector<double> pb(stride);
                std::vector<double> mu((size_t)nb);
                fetch_slot(acc_slot_res(0));
                for (int b = 0; b < nb; ++b) mu[(size_t)b] = slot_max(b, 4);
                for (int k = 0; k < max_outer; ++k) {
                    const std::vector<Start> st = refinement_start(k);
                    bool done = true;
                    for (int b = 0; b < nb; ++b) done = done && st[b].done;
                    if (trace_solve)
                        std::fprintf(stderr, "[sf solve] outer %d residual %.3e%s\n", k,
                                     std::sqrt(st[0].rr / (st[0].bb > 0 ? st[0].bb : 1.0)),
                                     done ? " converged" : "");
                    if (done) break;
                    int n = max_inner;
                    std::vector<std::vector<double>> rhist((size_t)nb);
                    for (int it = 1; it <= max_inner; ++it) {
                        cheb(it, k);
                        if (cheb_done(st, pb, rhist, mu)) { n = it; break; }
                    }
                    used.push_back(n);
                    if (trace_solve) {
                        std::fprintf(stderr, "[sf solve]   chebyshev iterations %d, |r|/|b|:", n);
                        for (double rr : rhist[0])
                            std::fprintf(stderr, " %.2e", std::sqrt(rr / (st[0].bb > 0 ? st[0].bb : 1.0)));
                        std::fprintf(stderr, "\n");
                    }
                    resid(k + 1, cheb_e(n + 1), nullptr);
                }

            } else {
                // a refinement of no inner iteration (only ever the last one)
                // takes z = M^-1 r of the residual before as its correction;
                // its residual launch forms that z itself (on its rows and
                // one row either side), so the residual launch before it
                // solves no line system: the spare refinement usually finds
                // the tolerance met and returns at once
                for (size_t k = 0; k < sched.size(); ++k) {
                    for (int it = 1; it <= sched[k]; ++it) cheb(it, (int)k);
                    const bool zp = sched[k] == 0 && k > 0;
                    const bool zp_next = k + 2 == sched.size() && sched[k + 1] == 0;
                    resid((int)k + 1, zp ? nullptr : cheb_e(sched[k] + 1), sstat_d,
                          k + 1 == sched.size() || zp_next, zp);
                }
            }
        } else if (host_driven) {
            std::vector<SolveCtl> c(nb);
            std::vector<int> before(nb);
            for (int k = 0; k < max_outer; ++k) {
                for (const LaunchGeom& rg : pgeo)
                    launch_spmv<Ti>(stream, rg, bub_d, ctl_d, sys(x64[0]), p[1], p[0], 0, sp, partR,
                                    partB, partA, stride);
                const bool done = all_flags(c, false);
                if (trace_solve)
                    std::fprintf(stderr, "[sf solve] outer %d residual %.3e (|b|^2 %.3e)%s\n", k,
                                 c[0].rel_res, c[0].bb, done ? " converged" : "");
                if (done) break;
                for (int b = 0; b < nb; ++b) before[b] = c[b].iters;
                cgstep(0);
                for (int it = 1; it < max_inner; ++it) {
                    inner(it);
                    if ((mg || (it & 3) == 3) && all_flags(c, true)) break;
                }
                all_flags(c, true);
                int mx = 0;
                for (int b = 0; b < nb; ++b) mx = std::max(mx, c[b].iters - before[b]);
                used.push_back(mx);
                if (trace_solve)
                    std::fprintf(stderr, "[sf solve]   inner iterations %d\n", c[0].iters - before[0]);
                resid(k + 1, e, nullptr);
            }
        } else {
            for (size_t k = 0; k < sched.size(); ++k) {
                for (int it = 0; it < sched[k]; ++it) inner(it);
                resid((int)k + 1, e, nullptr);
            }
        }
        if (host_driven && recording) {
            if (calib.size() < used.size()) calib.resize(used.size(), 0);
            for (size_t k = 0; k < used.size(); ++k) calib[k] = std::max(calib[k], used[k]);
        }
        if (finalize) launch_finalize(stream, nb, bub_d, ctl_d, hist_d, sp, partR, stride);
    }

    void update(bool finalize) override {
        launch_update<T>(stream, geom, bub_d, time_d, ctl_d, tab, nxt, air(), x64[0], x64[1], cur,
                         finalize ? 1 : 0, hist_d, sp, partR, stride,
                         bimocq ? bm.chg_e : nullptr, bimocq ? bm.chg_g : nullptr, gdel, gnext,
                         acc);
        if (bimocq) {
            // the change accumulation of bodyforce (bubble.cu:2486-2502), then
            // estimateDistortion and reInitializeMapping (solver.cu:279-286)
            launch_bimocq_accumulate(stream, geom, bub_d, nb, bm);
            launch_bimocq_reinit<T>(stream, geom, bub_d, cur, bm);
        }
    }

    // ---- host <-> device copies
    void upload(int b, int field, const double* h) override {
        const BubbleDev& B = bub_h[b];
        const size_t N = (size_t)B.nth * B.nph;
        size_t n = N;
        T* dst = nullptr;
        switch (field) {
            case SF_FIELD_UTHETA: dst = cur.uth; n = N - B.nph; break;
            case SF_FIELD_UPHI: dst = cur.uph; break;
            case SF_FIELD_ETA: dst = cur.eta; break;
            case SF_FIELD_GAMMA: dst = cur.gam; break;
            case SF_FIELD_UAIR0: dst = ua0; break;
            case SF_FIELD_VAIR0: dst = va0; n = N - B.nph; break;
            default: throw Status{SF_ERR_INVALID, "field cannot be uploaded"};
        }
        std::vector<T> tmp(N, T(0));
        bool nan = false;
        for (size_t k = 0; k < n; ++k) {
            tmp[k] = (T)h[k];
            nan |= std::isnan(h[k]);
        }
        SF_HIP(hipMemcpyAsync(dst + B.off, tmp.data(), N * sizeof(T), hipMemcpyHostToDevice, stream));
        if (field == SF_FIELD_ETA || field == SF_FIELD_GAMMA) bimocq_dirty = true;
        if (field <= SF_FIELD_GAMMA) reset_guess(b);
        if (field <= SF_FIELD_GAMMA) {
            // a new state steps again after a failed solve; the advection
            // samples a state with NaNs by NaN-skipping interpolation
            BubbleTime bt;
            SF_HIP(hipMemcpyAsync(&bt, time_d + b, sizeof(bt), hipMemcpyDeviceToHost, stream));
            SF_HIP(hipStreamSynchronize(stream));
            bt.stop &= 1;
            if (nan) bt.nans = 1;
            SF_HIP(hipMemcpyAsync(time_d + b, &bt, sizeof(bt), hipMemcpyHostToDevice, stream));
        }
        SF_HIP(hipStreamSynchronize(stream));
    }

    template <typename X>
    void fetch(const X* src, size_t off, size_t n, double* h) {
        std::vector<X> tmp(n);
        SF_HIP(hipMemcpyAsync(tmp.data(), src + off, n * sizeof(X), hipMemcpyDeviceToHost, stream));
        SF_HIP(hipStreamSynchronize(stream));
        for (size_t k = 0; k < n; ++k) h[k] = (double)tmp[k];
    }

    unsigned short* half_d = nullptr;   // frame output plane (allocated on first use)
    size_t half_cap = 0;
    void thickness_half(int b, uint16_t* h) override {
        const BubbleDev& B = bub_h[b];
        const size_t N = (size_t)B.nth * B.nph;
        if (half_cap < N) {
            size_t most = 0;
            for (int k = 0; k < nb; ++k) most = std::max(most, (size_t)bub_h[k].nth * bub_h[k].nph);
            half_d = dalloc<unsigned short>(most, owned);
            half_cap = most;
        }
        launch_thickness_half<T>(stream, cur.eta + B.off, (long)N, desc[b].H, half_d);
        SF_HIP(hipMemcpyAsync(h, half_d, N * sizeof(uint16_t), hipMemcpyDeviceToHost, stream));
        SF_HIP(hipStreamSynchronize(stream));
    }
    void download(int b, int field, double* h) override {
        const BubbleDev& B = bub_h[b];
        const size_t N = (size_t)B.nth * B.nph, Nt = N - B.nph, o = (size_t)B.off;
        switch (field) {
            case SF_FIELD_UTHETA: fetch(cur.uth, o, Nt, h); break;
            case SF_FIELD_UPHI: fetch(cur.uph, o, N, h); break;
            case SF_FIELD_ETA: fetch(cur.eta, o, N, h); break;
            case SF_FIELD_GAMMA: fetch(cur.gam, o, N, h); break;
            case SF_FIELD_UAIR0: fetch(ua0, o, N, h); break;
            case SF_FIELD_VAIR0: fetch(va0, o, Nt, h); break;
            case SF_FIELD_UTHETA_STAR: fetch(nxt.uth, o, Nt, h); break;
            case SF_FIELD_UPHI_STAR: fetch(nxt.uph, o, N, h); break;
            case SF_FIELD_ETA_STAR: fetch(nxt.eta, o, N, h); break;
            case SF_FIELD_GAMMA_STAR: fetch(nxt.gam, o, N, h); break;
            case SF_FIELD_RHS: fetch(b64, o, N, h); break;
            case SF_FIELD_DIAG: fetch(dg, o, N, h); break;
            case SF_FIELD_WNORTH: fetch(wN, o, N, h); break;
            case SF_FIELD_WWEST: fetch(wW, o, N, h); break;
            case SF_FIELD_SOLUTION: {
                SolveCtl c;
                SF_HIP(hipMemcpyAsync(&c, ctl_d + b, sizeof(c), hipMemcpyDeviceToHost, stream));
                SF_HIP(hipStreamSynchronize(stream));
                fetch(x64[c.xbuf & 1], o, N, h);
                break;
            }
            default: throw Status{SF_ERR_INVALID, "unknown field"};
        }
    }

    void row_stats() override {
        launch_row_stats<T>(stream, sgeo, bub_d, tab, cur.eta, partS, stride);
    }
};

// ------------------------------------------------------------------ setup --
bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

// Cells per thread of the row blocks of a row of nph cells: C in {2, 4, 8,
// 16} dividing nph, so that the row's nph / C threads (at least 4, at most
// kRowThreadsMax) each hold C consecutive cells. The first choice is 2 for
// rows of at most 128 cells, else 4, doubled until the row fits one block;
// a row that no doubling fits (nph / C not a whole number) takes the next
// smaller C that divides it. 0: no C fits.
int row_cpt(int nph) {
    const int pref = nph <= 128 ? 2 : 4;
    int cands[8], n = 0;
    for (int c = pref; c <= 16; c *= 2) cands[n++] = c;
    for (int c = pref / 2; c >= 2; c /= 2) cands[n++] = c;
    for (int k = 0; k < n; ++k) {
        const int c = cands[k];
        if (nph % c == 0 && nph / c <= kRowThreadsMax && nph / c >= 4) return c;
    }
    return 0;
}

int next_pow2(int v) {
    int p = 1;
    while (p < v) p *= 2;
    return p;
}

// Workgroup k of a launch is dispatched to XCD k mod nxcd, and each XCD has
// its own L2. Work item l (a row or a tile, in spatial order) is therefore
// placed at launch slot (l mod chunk) * nxcd + l / chunk with chunk =
// ceil(n / nxcd): every XCD receives one contiguous band of rows, so the
// neighbour rows that stencils and advection gathers read are mostly in the
// same L2. Slots past the end hold `pad` (bubble index -1, skipped).
template <typename X>
std::vector<X> xcd_order(const std::vector<X>& work, X pad, int nxcd) {
    const size_t n = work.size(), chunk = (n + nxcd - 1) / nxcd;
    std::vector<X> out(chunk * nxcd, pad);
    for (size_t l = 0; l < n; ++l) out[(l % chunk) * nxcd + l / chunk] = work[l];
    return out;
}

void derive(const sf_bubble_desc& d, sf_derived& o) { sfh::derive_params(d, o); }

void build_tables(Base& c, std::vector<void*>& owned) {
    std::vector<double> sTc, cTc, sTf, cTf, sF, cscTc, sPc, cPc, sPh, cPh;
    for (int b = 0; b < c.nb; ++b) {
        const BubbleDev& B = c.bub_h[b];
        const double h = B.h;
        for (int j = 0; j <= B.nth; ++j) {
            // centres (j + 1/2) h, as gTheta = (thetaId + 0.5) * gridLen
            sTc.push_back(std::sin(((double)j + 0.5) * h));
            cTc.push_back(std::cos(((double)j + 0.5) * h));
            cscTc.push_back(1.0 / std::sin(((double)j + 0.5) * h));
            // u_theta positions (thetaId + 1.0) * gridLen, indexed by face j
            sTf.push_back(std::sin((double)j * h));
            cTf.push_back(std::cos((double)j * h));
            // assembly face sines sin(gTheta - halfStep); both pole faces
            // carry no flux
            sF.push_back((j == 0 || j == B.nth) ? 0.0 : std::sin(((double)j + 0.5) * h - 0.5 * h));
        }
        for (int i = 0; i < B.nph; ++i) {
            sPc.push_back(std::sin((double)i * h));
            cPc.push_back(std::cos((double)i * h));
            sPh.push_back(std::sin(((double)i - 0.5) * h));
            cPh.push_back(std::cos(((double)i - 0.5) * h));
        }
    }
    auto up64 = [&](const std::vector<double>& v) {
        double* d = dalloc<double>(v.size(), owned);
        SF_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
        return d;
    };
    auto up32 = [&](const std::vector<double>& v) {
        std::vector<float> f(v.begin(), v.end());
        float* d = dalloc<float>(f.size(), owned);
        SF_HIP(hipMemcpy(d, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
        return d;
    };
    Tables& t = c.tab;
    t.sinTc = up64(sTc); t.cosTc = up64(cTc); t.sinTf = up64(sTf); t.cosTf = up64(cTf);
    t.sinFace = up64(sF); t.cscTc = up64(cscTc);
    t.sinTc32 = up32(sTc); t.cosTc32 = up32(cTc); t.sinTf32 = up32(sTf); t.cosTf32 = up32(cTf);
    t.sinPc = up64(sPc); t.cosPc = up64(cPc); t.sinPh = up64(sPh); t.cosPh = up64(cPh);
    t.sinPc32 = up32(sPc); t.cosPc32 = up32(cPc); t.sinPh32 = up32(sPh); t.cosPh32 = up32(cPh);
}

template <typename T, typename Ti>
void upload_air_shapes(Impl<T, Ti>& c) {
    for (int b = 0; b < c.nb; ++b) {
        const BubbleDev& B = c.bub_h[b];
        const size_t N = (size_t)B.nth * B.nph;
        std::vector<double> su(N, 0.0), sv(N, 0.0);
        sfh::air_shapes(B.nth, c.der[b].radius, c.der[b].U, su.data(), sv.data());
        std::vector<T> tu(su.begin(), su.end()), tv(sv.begin(), sv.end());
        SF_HIP(hipMemcpy(c.au + B.off, tu.data(), N * sizeof(T), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(c.av + B.off, tv.data(), N * sizeof(T), hipMemcpyHostToDevice));
    }
}

std::unique_ptr<Base> make_impl(int prec) {
    switch (prec) {
        case SF_PREC_FP64: return std::unique_ptr<Base>(new Impl<double, double>());
        case SF_PREC_MIXED: return std::unique_ptr<Base>(new Impl<float, float>());
        case SF_PREC_FP32_ADV: return std::unique_ptr<Base>(new Impl<float, double>());
        default: return nullptr;
    }
}

template <typename T, typename Ti>
void finish_setup(Impl<T, Ti>& c) {
    c.alloc_fields();
    upload_air_shapes(c);
    if (c.bimocq) c.alloc_bimocq();
    if (c.mg) c.alloc_mg();
}

// The row blocks of one launch over the rows of the widths `ws` (C cells a
// thread): a block has the threads of the widest row (whole waves); a row of
// R = nph / C threads rounded up to w whole waves shares a block with
// blockDim / w - 1 further rows of its bubble when two or more fit (the
// kernels' row_seg), and the block's LDS holds each of its rows' reduced
// line system (6 R values, 8 R when R is no power of two). Items are
// (bubble, first row), in the XCD-aware order.
LaunchGeom pack_rows(Base& c, const std::vector<int>& ws, int C, size_t esz, size_t lds_limit) {
    int T = 64;
    for (int w : ws) T = std::max(T, (w / C + 63) / 64 * 64);
    size_t lds = 0;
    bool packed = false;
    std::vector<int2> items;
    for (int b = 0; b < c.nb; ++b) {
        const int nph = c.bub_h[b].nph;
        if (std::find(ws.begin(), ws.end(), nph) == ws.end()) continue;
        const int R = nph / C, w = (R + 63) / 64 * 64;
        const int k = 2 * w > T ? 1 : T / w;
        packed = packed || k > 1;
        lds = std::max(lds, (size_t)k * (size_t)(is_pow2(R) ? 6 : 8) * (size_t)R * esz);
        for (int j = 0; j < c.bub_h[b].nth; j += k) items.push_back(make_int2(b, j));
    }
    if (lds > lds_limit)
        throw Status{SF_ERR_UNSUPPORTED, "line solve of rows of " + std::to_string(ws.back()) +
                                             " cells needs " + std::to_string(lds) +
                                             " B of LDS per block, device allows " +
                                             std::to_string(lds_limit) +
                                             " (use SF_PREC_MIXED for this size)"};
    items = xcd_order(items, make_int2(-1, 0), c.nxcd);
    int2* rows_d = dalloc<int2>(items.size(), c.owned);
    SF_HIP(hipMemcpy(rows_d, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice));
    LaunchGeom g = c.geom;
    g.rows = rows_d;
    g.nrows = (int)items.size();
    g.cpt = C;
    g.row_threads = T;
    g.lds_bytes = lds;
    g.packed = packed;
    g.pow2_rows = true;
    for (int w : ws) g.pow2_rows = g.pow2_rows && is_pow2(w);
    return g;
}

void setup(Base& c, int prec) {
    const int nb = c.nb;
    SF_HIP(hipSetDevice(c.device));
    if (!c.noalloc) {
        SF_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        for (auto& e : c.ev) SF_HIP(hipEventCreate(&e));
    }
    hipDeviceProp_t prop;
    SF_HIP(hipGetDeviceProperties(&prop, c.device));

    // bubble descriptors, offsets, row/column table bases
    long long off = 0;
    int rt = 0, ct = 0, pr = 0, nph_max = 0;
    c.bub_h.resize(nb);
    c.der.resize(nb);
    c.time_h.resize(nb);
    for (int b = 0; b < nb; ++b) {
        const sf_bubble_desc& d = c.desc[b];
        derive(d, c.der[b]);
        BubbleDev& B = c.bub_h[b];
        B.nth = d.n_theta;
        B.nph = 2 * d.n_theta;
        // power-of-two rows wrap phi by a mask; other rows by a comparison
        B.mask = is_pow2(B.nph) ? B.nph - 1 : -1;
        B.rt0 = rt;
        B.ct0 = ct;
        B.pr0 = pr;
        B.off = off;
        B.h = kPi / d.n_theta;
        B.invh = d.n_theta / kPi;
        B.dt = c.der[b].dt;
        B.M = c.der[b].M;
        B.Cr = c.der[b].Cr;
        B.gs = c.der[b].gs;
        B.gamma_max = c.der[b].gamma_max;
        B.eva = c.der[b].eva;
        B.gravity = d.gravity;
        B.evaporation = d.evaporation;
        B.airflow = d.airflow;
        B.viscous = c.opt.viscous ? 1 : 0;
        B.re = c.der[b].re;
        // every bubble starts on a 256-byte boundary (the row kernels load
        // C consecutive cells as one aligned vector)
        off += ((long long)B.nth * B.nph + 63) / 64 * 64;
        rt += B.nth + 1;
        ct += B.nph;
        pr += B.nth;
        nph_max = std::max(nph_max, B.nph);
        c.time_h[b] = BubbleTime{0.0, d.dt * 1.0, c.opt.advance_time, 0, 0, 0, 0,
                                 c.opt.halt_on_failure ? 1 : 0};
    }
    c.ncells = off;
    c.nrow_total = pr;
    c.stride = pr;

    c.bub_d = dalloc<BubbleDev>(nb, c.owned);
    SF_HIP(hipMemcpy(c.bub_d, c.bub_h.data(), nb * sizeof(BubbleDev), hipMemcpyHostToDevice));
    c.time_d = dalloc<BubbleTime>(nb, c.owned);
    SF_HIP(hipMemcpy(c.time_d, c.time_h.data(), nb * sizeof(BubbleTime), hipMemcpyHostToDevice));
    c.ctl_d = dalloc<SolveCtl>(nb, c.owned);
    c.hist_d = dalloc<SolveHist>(nb, c.owned);
    c.sstat_d = dalloc<SchedStat>(nb, c.owned);
    if (!c.noalloc) {
        SF_HIP(hipHostMalloc((void**)&c.sstat_h, 2 * nb * sizeof(SchedStat)));
        std::memset(c.sstat_h, 0, 2 * nb * sizeof(SchedStat));
        for (auto& e : c.cev) SF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        SF_HIP(hipStreamCreateWithFlags(&c.cstream, hipStreamNonBlocking));
    }
    build_tables(c, c.owned);

    int nxcd = 1;
    if (hipDeviceGetAttribute(&nxcd, hipDeviceAttributeNumberOfXccs, c.device) != hipSuccess || nxcd < 1)
        nxcd = 1;
    if (const char* e = std::getenv("SF_XCD_ORDER")) nxcd = std::max(1, std::atoi(e));
    c.nxcd = nxcd;
    // cell tiles of 256 threads: a row chunk of `width` columns times
    // 256 / width rows (a shorter row: its power-of-two width, several rows);
    // width 256 by default, 128 and 64 as the autotuner's candidates
    auto build_tiles = [&](int width) {
        std::vector<int4> tiles;
        for (int b = 0; b < nb; ++b) {
            const BubbleDev& B = c.bub_h[b];
            const int w = std::min(width, next_pow2(B.nph)), rpt = 256 / w;
            for (int j = 0; j < B.nth; j += rpt)
                for (int i = 0; i < B.nph; i += w) tiles.push_back(make_int4(b, j, i, w));
        }
        tiles = xcd_order(tiles, make_int4(-1, 0, 0, 1), nxcd);
        int4* tiles_d = dalloc<int4>(tiles.size(), c.owned);
        SF_HIP(hipMemcpy(tiles_d, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice));
        return Base::TileSet{width, tiles_d, (int)tiles.size()};
    };
    c.tile_cands.clear();
    c.tile_cands.push_back(build_tiles(256));
    if (c.opt.autotune && !c.noalloc)
        for (int w : {128, 64}) c.tile_cands.push_back(build_tiles(w));
    c.geom.tiles = c.tile_cands[0].tiles;
    c.geom.ntiles = c.tile_cands[0].ntiles;
    c.geom.pow2_rows = true;
    for (int b = 0; b < nb; ++b) c.geom.pow2_rows = c.geom.pow2_rows && c.bub_h[b].mask >= 0;
    // Row blocks, one launch per row width: C consecutive cells per thread
    // (4, or 2 for rows of at most 128 cells, doubled until the row needs at
    // most kRowThreadsMax threads), one thread per C cells; a row holds at
    // least 4 C cells for the reduced line system. Grouping the rows by
    // width keeps a narrow row from occupying a wide row's block.
    const size_t esz = c.fp64_inner() ? 8 : 4;
    const size_t lds_limit = (size_t)prop.sharedMemPerBlock - 1024;
    c.lds_limit = lds_limit;
    std::vector<int> widths;
    for (int b = 0; b < nb; ++b) widths.push_back(c.bub_h[b].nph);
    std::sort(widths.begin(), widths.end());
    widths.erase(std::unique(widths.begin(), widths.end()), widths.end());
    c.rgeo.clear();
    c.rgeo_nph.clear();
    for (int nph : widths) {
        std::vector<int2> rows;
        for (int b = 0; b < nb; ++b)
            if (c.bub_h[b].nph == nph)
                for (int j = 0; j < c.bub_h[b].nth; ++j) rows.push_back(make_int2(b, j));
        rows = xcd_order(rows, make_int2(-1, 0), nxcd);
        int2* rows_d = dalloc<int2>(rows.size(), c.owned);
        SF_HIP(hipMemcpy(rows_d, rows.data(), rows.size() * sizeof(int2), hipMemcpyHostToDevice));
        const int cpt = row_cpt(nph);
        if (cpt == 0)
            throw Status{SF_ERR_UNSUPPORTED, "rows of " + std::to_string(nph) + " cells are not supported"};
        LaunchGeom g = c.geom;
        g.rows = rows_d;
        g.nrows = (int)rows.size();
        g.cpt = cpt;
        g.pow2_rows = is_pow2(nph);
        g.row_threads = std::max(64, (nph / cpt + 63) / 64 * 64);
        // the reduced system of R = nph / C interface unknowns: 6 R values
        // (8 R when R is no power of two: the correction's second right-hand
        // side)
        auto lds_of = [&](int R) { return (size_t)(is_pow2(R) ? 6 : 8) * (size_t)R * esz; };
        g.lds_bytes = lds_of(nph / cpt);
        if (g.lds_bytes > lds_limit) {
            throw Status{SF_ERR_UNSUPPORTED,
                         "line solve of nPhi = " + std::to_string(nph) + " needs " +
                             std::to_string(g.lds_bytes) + " B of LDS per block, device allows " +
                             std::to_string(lds_limit) + " (use SF_PREC_MIXED for this size)"};
        }
        c.rgeo.push_back(g);
        c.rgeo_nph.push_back(nph);
    }
    // The row stages' launches: widths grouped (in increasing order) while
    // one cell count a thread C suits all of them (C divides the row, 4 to
    // 1024 threads a row), C the largest any of them prefers; in each
    // group's launch a block has the threads of the widest row, and a
    // narrower row takes its thread count rounded up to whole waves, as
    // many rows of one bubble to a block as fit (row_seg). The Chebyshev
    // launches take 8 cells a thread where the assembly takes 4 and the
    // widest row has at least 2048 cells.
    auto fits = [](int nph, int C) { return nph % C == 0 && nph / C >= 4 && nph / C <= kRowThreadsMax; };
    c.pgeo.clear();
    c.pgeo_inner.clear();
    c.pgeo_width.clear();
    c.pgeo_widths.clear();
    std::vector<std::vector<int>> groups;
    std::vector<int> gcpt;
    for (int nph : widths) {
        if (!groups.empty()) {
            // a shared C above 4 only where every member takes it anyway (8
            // or more cells a thread spill the assembly's registers), and
            // widths within a factor 4 of the group's narrowest: a launch
            // whose widest rows fill whole blocks runs them on the packed
            // code, which holds more registers (measured on 68 bubbles of
            // 4096 and 256 cells a row in one launch: 7 % slower than two)
            const int C = std::max(gcpt.back(), row_cpt(nph));
            bool ok = fits(nph, C) && (C <= 4 || row_cpt(nph) == C) && nph <= 4 * groups.back().front();
            for (int w : groups.back()) ok = ok && fits(w, C) && (C <= 4 || row_cpt(w) == C);
            if (ok) {
                groups.back().push_back(nph);
                gcpt.back() = C;
                continue;
            }
        }
        groups.push_back({nph});
        gcpt.push_back(row_cpt(nph));
    }
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        const std::vector<int>& ws = groups[gi];
        const int C = gcpt[gi];
        const int wmax = ws.back();
        int Ci = C;
        if (!c.fp64_inner() && C == 4 && wmax / 8 >= 256) {
            bool ok8 = true;
            for (int w : ws) ok8 = ok8 && fits(w, 8);
            if (ok8) Ci = 8;
        }
        c.pgeo.push_back(pack_rows(c, ws, C, esz, lds_limit));
        c.pgeo_inner.push_back(pack_rows(c, ws, Ci, esz, lds_limit));
        c.pgeo_width.push_back(wmax);
        c.pgeo_widths.push_back(ws);
    }
    {
        // row statistics: one row a block, all rows, 256 threads striding
        std::vector<int2> rows;
        for (int b = 0; b < nb; ++b)
            for (int j = 0; j < c.bub_h[b].nth; ++j) rows.push_back(make_int2(b, j));
        rows = xcd_order(rows, make_int2(-1, 0), nxcd);
        int2* rows_d = dalloc<int2>(rows.size(), c.owned);
        SF_HIP(hipMemcpy(rows_d, rows.data(), rows.size() * sizeof(int2), hipMemcpyHostToDevice));
        c.sgeo = c.geom;
        c.sgeo.rows = rows_d;
        c.sgeo.nrows = (int)rows.size();
        c.sgeo.row_threads = 256;
        c.sgeo.cpt = 1;
        c.sgeo.lds_bytes = 0;
    }

    c.partR = dalloc<double>(5 * (size_t)c.stride, c.owned);
    c.partA = dalloc<double>((size_t)c.stride, c.owned);
    c.partB = dalloc<double>(2 * (size_t)c.stride, c.owned);
    c.partS = dalloc<double>(2 * (size_t)c.stride, c.owned);
    const size_t n = (size_t)c.ncells;
    c.x64[0] = dalloc<double>(n, c.owned);
    c.x64[1] = dalloc<double>(n, c.owned);
    c.b64 = dalloc<double>(n, c.owned);
    c.rsteps = dalloc<int>((size_t)c.stride, c.owned);
    c.acc = dalloc<double>((size_t)nb * kAccBubble, c.owned);

    // solver schedule; the inner method: PCG for an FP64 inner solve (the
    // plain path), Chebyshev for FP32 unless asked otherwise
    c.method = c.opt.inner_method > 0 ? c.opt.inner_method : (c.fp64_inner() ? 1 : 2);
    if (c.method == 3) {
        // PCG with the multigrid preconditioner
        c.method = 1;
        c.mg = true;
    }
    if (c.method != 1 && c.method != 2) throw Status{SF_ERR_INVALID, "unknown inner method"};
    c.sp.tol = c.opt.tol > 0 ? c.opt.tol : 1e-10;
    c.sp.absolute = c.opt.tol_absolute;
    c.sp.inner_frac = 0.5;
    // test hook: a smaller executable-graph cache (1 evicts on every capture)
    if (const char* e = std::getenv("SF_GRAPH_CACHE_MAX")) c.graph_cache_max = (size_t)std::max(1, std::atoi(e));
    if (c.opt.scalar_advection < 0 || c.opt.scalar_advection > 1)
        throw Status{SF_ERR_INVALID, "scalar_advection must be 0 (plain) or 1 (BIMOCQ)"};
    c.bimocq = c.opt.scalar_advection == 1;
    c.bm.blend = c.opt.blend;
    c.bm.threshold = c.opt.reinit_threshold != 0.0 ? c.opt.reinit_threshold : 4.0;
    c.bm.multisample = c.opt.multisample ? 1 : 0;
    if (c.opt.initial_guess < 0 || c.opt.initial_guess > 2)
        throw Status{SF_ERR_INVALID, "initial_guess must be 0, 1 or 2"};
    c.guess = c.opt.initial_guess == 2 || (c.opt.initial_guess == 0 && prec != SF_PREC_FP64);
    if (c.fp64_inner()) {
        c.sp.inner_rho2 = 0.0;
        c.max_outer = c.opt.max_outer > 0 ? c.opt.max_outer : 3;
        if (c.max_outer > kAccRefine) throw Status{SF_ERR_INVALID, "max_outer is at most 16"};
        c.max_inner = c.opt.max_inner > 0 ? c.opt.max_inner : 1000;
    } else {
        // an FP32 inner solve reduces its residual by about 1e-6 before its
        // rounding floor; the FP64 refinement takes it from there
        const double rho = c.opt.inner_rho > 0 ? c.opt.inner_rho : 1e-6;
        c.sp.inner_rho2 = rho * rho;
        c.max_outer = c.opt.max_outer > 0 ? c.opt.max_outer : 4;
        if (c.max_outer > kAccRefine) throw Status{SF_ERR_INVALID, "max_outer is at most 16"};
        c.max_inner = c.opt.max_inner > 0 ? c.opt.max_inner : 200;
    }

    switch (prec) {
        case SF_PREC_FP64: finish_setup(static_cast<Impl<double, double>&>(c)); break;
        case SF_PREC_MIXED: finish_setup(static_cast<Impl<float, float>&>(c)); break;
        default: finish_setup(static_cast<Impl<float, double>&>(c)); break;
    }
    SF_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace

struct sf_ctx {
    std::unique_ptr<Base> impl;
    int prec = 0;
};

namespace {

int fail(sf_ctx* ctx, int code, const std::string& msg) {
    if (ctx && ctx->impl) ctx->impl->err = msg;
    g_last_error = msg;
    return code;
}

template <typename F>
int guarded(sf_ctx* ctx, F&& f) {
    try {
        return f();
    } catch (const HipError& e) {
        return fail(ctx, SF_ERR_HIP, e.msg);
    } catch (const Status& s) {
        return fail(ctx, s.code, s.msg);
    } catch (const std::exception& e) {
        return fail(ctx, SF_ERR_INVALID, e.what());
    }
}

void time_stage(Base& c, int e0, int e1, double& acc) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c.ev[e0], c.ev[e1]) == hipSuccess) acc += ms;
}

// One eager step with per-stage events.
void eager_step(Base& c, bool host_driven) {
    SF_HIP(hipEventRecord(c.ev[0], c.stream));
    c.advect();
    SF_HIP(hipEventRecord(c.ev[1], c.stream));
    c.assemble();
    SF_HIP(hipEventRecord(c.ev[2], c.stream));
    c.solve(host_driven, false);
    SF_HIP(hipEventRecord(c.ev[3], c.stream));
    c.update(true);
    SF_HIP(hipEventRecord(c.ev[4], c.stream));
    SF_HIP(hipEventSynchronize(c.ev[4]));
    time_stage(c, 0, 1, c.timing.advect_ms);
    time_stage(c, 1, 2, c.timing.assemble_ms);
    time_stage(c, 2, 3, c.timing.solve_ms);
    time_stage(c, 3, 4, c.timing.update_ms);
    time_stage(c, 0, 4, c.timing.step_ms);
    c.timing.steps += 1;
    c.host_hist += 1;
}

// Autotuner (sf_options.autotune), at the start of the first sf_step call
// of a captured context after kTuneAfterSteps steps (the advection's cost
// depends on the flow: a film started from rest has barely moved before):
// times the advection with each candidate cell-tile list and, per
// row width, the Chebyshev launch at C, 2C and 4C cells a thread (C the
// assembly's, so the reduction depth it records for the line solve stays
// deep enough), on the context's own state, and keeps the fastest of each.
// The launches repeat the step's advection and the second inner iteration
// of the first refinement (whose output the solve rewrites before reading
// it); the control blocks are restored, and the step that follows computes
// everything again.
constexpr int kTuneAfterSteps = 16;

void autotune(Base& c) {
    const int n = 20;
    sf_tuning& t = c.tuning;
    std::memset(&t, 0, sizeof(t));
    double best = 1e300;
    size_t bi = 0;
    for (size_t k = 0; k < c.tile_cands.size() && k < 3; ++k) {
        LaunchGeom g = c.geom;
        g.tiles = c.tile_cands[k].tiles;
        g.ntiles = c.tile_cands[k].ntiles;
        (void)c.time_advect_geom(g, 3);
        const double ms = c.time_advect_geom(g, n);
        t.advect_ms[k] = ms;
        if (ms < best) { best = ms; bi = k; }
    }
    c.geom.tiles = c.tile_cands[bi].tiles;
    c.geom.ntiles = c.tile_cands[bi].ntiles;
    t.advect_tile_width = c.tile_cands[bi].width;
    t.nwidths = (int)std::min<size_t>(c.pgeo_inner.size(), 8);
    for (int w = 0; w < t.nwidths; ++w) {
        t.row_width[w] = c.pgeo_width[(size_t)w];
        t.inner_cells[w] = c.pgeo_inner[(size_t)w].cpt;
    }
    if (c.method == 2 && !c.fp64_inner()) {
        c.advect();
        c.assemble();
        std::vector<SolveCtl> keep;
        c.ctl_to_host(keep);
        for (int w = 0; w < t.nwidths; ++w) {
            const std::vector<int>& ws = c.pgeo_widths[(size_t)w];
            const int C0 = c.pgeo[(size_t)w].cpt;
            double bt = 1e300;
            LaunchGeom bg = c.pgeo_inner[(size_t)w];
            for (int q = 0; q < 3; ++q) {
                const int C = C0 << q;
                bool ok = C <= 16;
                for (int nph : ws) ok = ok && nph % C == 0 && nph / C >= 4 && nph / C <= kRowThreadsMax;
                if (!ok) continue;
                LaunchGeom gi;
                try {
                    gi = pack_rows(c, ws, C, sizeof(float), c.lds_limit);
                } catch (const Status&) {
                    continue;   // more LDS than a block has
                }
                (void)c.time_cheb_geom(gi, 5);
                const double ms = c.time_cheb_geom(gi, n);
                t.inner_ms[w][q] = ms;
                if (ms < bt) { bt = ms; bg = gi; }
            }
            c.pgeo_inner[(size_t)w] = bg;
            t.inner_cells[w] = bg.cpt;
        }
        c.ctl_to_device(keep);
    }
    t.tuned = 1;
    c.tuned = true;
    if (c.trace_solve) {
        std::fprintf(stderr, "[sf tune] advection tiles %d columns (%.4f %.4f %.4f ms)\n",
                     t.advect_tile_width, t.advect_ms[0], t.advect_ms[1], t.advect_ms[2]);
        for (int w = 0; w < t.nwidths; ++w)
            std::fprintf(stderr, "[sf tune] rows of %d cells: Chebyshev %d cells a thread (%.4f %.4f %.4f ms)\n",
                         t.row_width[w], t.inner_cells[w], t.inner_ms[w][0], t.inner_ms[w][1],
                         t.inner_ms[w][2]);
    }
}

// The executable graph of `nsteps` steps for the current schedule, followed
// (slot >= 0) by the hand-off of the chunk's schedule statistics to pinned
// slot `slot`; captured on the capture stream (replays already queued on the
// context's stream keep running) and kept for reuse. One graph launch per
// chunk: a launch costs about as much as a short kernel. The cache holds at
// most graph_cache_max executables (the least recently used goes first):
// unbounded, a schedule that drifts over many lengths kept three
// executables of four steps per length, e.g. about 260 000 instantiated
// kernel nodes for lengths up to 200 inner launches.
hipGraphExec_t step_graph(Base& c, int nsteps, int slot) {
    std::vector<int> key = c.sched;
    key.push_back(nsteps);
    key.push_back(slot);
    auto it = c.graphs.find(key);
    if (it != c.graphs.end()) {
        it->second.used = ++c.graph_tick;
        return it->second.exec;
    }
    if (c.graphs.size() >= c.graph_cache_max) {
        // the least recently used executable goes, once no replay of it can
        // still be queued
        SF_HIP(hipStreamSynchronize(c.stream));
        auto lru = c.graphs.begin();
        for (auto q = c.graphs.begin(); q != c.graphs.end(); ++q)
            if (q->second.used < lru->second.used) lru = q;
        (void)hipGraphExecDestroy(lru->second.exec);
        c.graphs.erase(lru);
    }
    hipGraph_t g;
    std::swap(c.stream, c.cstream);
    hipError_t e = hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        for (int q = 0; q < nsteps; ++q) {
            c.advect();
            c.assemble();
            c.solve(false, false);
            c.update(true);
        }
        if (slot >= 0) launch_sched_stats(c.stream, c.sstat_d, c.sstat_h + slot * c.nb, c.nb);
        e = hipStreamEndCapture(c.stream, &g);
    }
    std::swap(c.stream, c.cstream);
    SF_HIP(e);
    if (c.trace_solve) {
        size_t nodes = 0;
        (void)hipGraphGetNodes(g, nullptr, &nodes);
        std::fprintf(stderr, "[sf solve] captured %d steps, schedule %d: %zu graph nodes\n", nsteps,
                     c.sched.empty() ? 0 : c.sched[0], nodes);
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    SF_HIP(e);
    c.graphs[key] = Base::GraphEntry{x, ++c.graph_tick};
    return x;
}

constexpr int kAdaptChunk = 4;
// Steps captured in one graph launch (a whole chunk: measured 1, 2 and 4
// steps per launch, 4 the fastest).
constexpr int kGraphSteps = 4;
// The cap on a captured schedule (kCaptureMaxInner); the next sf_step after
// a longer one calibrates again.
constexpr int kGraphMaxInner = kCaptureMaxInner;

void drop_calibration(Base& c) {
    c.calibrated = false;
    c.calib_steps = 0;
    c.calib.clear();
}

// Follows the schedule statistics of a finished chunk of replays that ran
// with `ran` first-refinement iterations: one more when a step needed a
// further refinement, one fewer when every step's iterate before the last
// already met the inner target. Statistics of a chunk that ran with another
// schedule than the current one only ever raise it.
void adapt_schedule(Base& c, const SchedStat* st, int ran) {
    double ratio = 0.0;
    int extra = 0;
    for (int b = 0; b < c.nb; ++b) {
        ratio = std::max(ratio, st[b].tail_ratio);
        extra += st[b].extra;
    }
    int n = c.sched[0];
    if (extra > 0) n = std::max(n, ran + 1);
    else if (ran == n && ratio <= 1.0 && n > 1) n -= 1;
    n = std::min(n, c.max_inner);
    if (n == c.sched[0]) return;
    if (c.trace_solve)
        std::fprintf(stderr, "[sf solve] schedule %d -> %d (tail ratio %.3g, extra %d)\n", c.sched[0], n,
                     ratio, extra);
    c.sched[0] = n;
}

void check_index(sf_ctx* ctx, int b) {
    if (!ctx || !ctx->impl) throw Status{SF_ERR_INVALID, "null context"};
    if (b < 0 || b >= ctx->impl->nb) throw Status{SF_ERR_INVALID, "bubble index out of range"};
}

// Solver statistics of bubble b (sf_get_stats): the row statistics launch,
// then the control blocks and partial sums read back.
void fill_stats(Base& c, int b, sf_bubble_stats* out) {
    c.row_stats();
    const BubbleDev& B = c.bub_h[b];
    std::vector<double> ps(2 * (size_t)c.stride);
    SF_HIP(hipMemcpyAsync(ps.data(), c.partS, ps.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c.stream));
    SolveCtl sc;
    SF_HIP(hipMemcpyAsync(&sc, c.ctl_d + b, sizeof(sc), hipMemcpyDeviceToHost, c.stream));
    SolveHist hs;
    SF_HIP(hipMemcpyAsync(&hs, c.hist_d + b, sizeof(hs), hipMemcpyDeviceToHost, c.stream));
    BubbleTime bt;
    SF_HIP(hipMemcpyAsync(&bt, c.time_d + b, sizeof(bt), hipMemcpyDeviceToHost, c.stream));
    SF_HIP(hipStreamSynchronize(c.stream));
    double vol = 0.0, pos = 0.0;
    for (int j = 0; j < B.nth; ++j) {
        vol += ps[B.pr0 + j];
        pos += ps[c.stride + B.pr0 + j];
    }
    std::memset(out, 0, sizeof(*out));
    out->iterations = sc.iters;
    out->outer = sc.outers;
    out->converged = sc.converged;
    out->residual = sc.rel_res;
    out->volume = vol * B.h * B.h;
    out->broken = (pos == 0.0) || (c.steps_done > 0 && !sc.converged);
    out->no_film = pos == 0.0 ? 1 : 0;
    out->steps = c.steps_done;
    out->time = bt.t + (bt.pending ? bt.dt_s : 0.0);
    out->failed_solves = hs.fails;
    out->max_residual = hs.max_res;
    if (c.bimocq) {
        double dl = 0.0;
        int ri = 0;
        SF_HIP(hipMemcpy(&dl, c.bm.dist_last + b, sizeof(double), hipMemcpyDeviceToHost));
        SF_HIP(hipMemcpy(&ri, c.bm.reinits + b, sizeof(int), hipMemcpyDeviceToHost));
        out->distortion = dl;
        out->reinits = ri;
    }
}

// The argument checks of sf_create (a bubble's size and physical
// parameters, the options), shared by the entry point without allocation.
int validate_create(const sf_options* opt, int32_t nbubbles, const sf_bubble_desc* bubbles) {
    if (!opt || !bubbles || nbubbles <= 0) return fail(nullptr, SF_ERR_INVALID, "no bubbles");
    for (int b = 0; b < nbubbles; ++b) {
        const sf_bubble_desc& d = bubbles[b];
        // a row of 2 n_theta cells is one block of at most kRowThreadsMax
        // threads holding 2, 4, 8 or 16 cells each
        if (d.n_theta < 4 || d.n_theta > 8 * kRowThreadsMax || row_cpt(2 * d.n_theta) == 0)
            return fail(nullptr, SF_ERR_UNSUPPORTED,
                        "n_theta must be from 4 to " + std::to_string(8 * kRowThreadsMax) +
                            ", with its row of 2 n_theta cells splitting into at most " +
                            std::to_string(kRowThreadsMax) +
                            " equal parts of 2, 4, 8 or 16 cells (bubble " + std::to_string(b) + ")");
        if (!(d.radius > 0) || !(d.H > 0) || !(d.U > 0) || !(d.Gamma_m > 0) || !(d.T > 0) ||
            !(d.dt > 0) || !(d.M_scale > 0))
            return fail(nullptr, SF_ERR_INVALID, "non-positive physical parameter");
        if (d.Gamma_m != bubbles[0].Gamma_m || d.T != bubbles[0].T)
            return fail(nullptr, SF_ERR_UNSUPPORTED, "bubbles of one context share Gamma_m and T");
    }
    if (opt->precision < 0 || opt->precision > 2) return fail(nullptr, SF_ERR_INVALID, "bad precision");
    if (opt->initial_guess < 0 || opt->initial_guess > 2)
        return fail(nullptr, SF_ERR_INVALID, "initial_guess must be 0, 1 or 2");
    if (opt->scalar_advection < 0 || opt->scalar_advection > 1)
        return fail(nullptr, SF_ERR_INVALID, "scalar_advection must be 0 (plain) or 1 (BIMOCQ)");
    if (opt->max_outer > kAccRefine) return fail(nullptr, SF_ERR_INVALID, "max_outer is at most 16");
    if (opt->viscous < 0 || opt->viscous > 1) return fail(nullptr, SF_ERR_INVALID, "viscous must be 0 or 1");
    if (opt->halt_on_failure < 0 || opt->halt_on_failure > 1)
        return fail(nullptr, SF_ERR_INVALID, "halt_on_failure must be 0 or 1");
    if (opt->autotune < 0 || opt->autotune > 1)
        return fail(nullptr, SF_ERR_INVALID, "autotune must be 0 or 1");
    return SF_OK;
}

}  // namespace

namespace sfapi {

int record_error(int code, const char* msg) { return fail(nullptr, code, msg); }

// Bubbles up to this many cells: the work arrays of one at 2^18 cells take
// about 38 MiB of the fixed array (mixed precision, VISCOUS on).
constexpr long long kNoAllocMaxCells = 1LL << 18;

int check_noalloc(const sf_options* opt, const sf_bubble_desc* bubble, int nsteps,
                  const float* const in[4], const float* uair0, const float* vair0,
                  float* const out[4]) {
    const int rc0 = validate_create(opt, 1, bubble);
    if (rc0 != SF_OK) return rc0;
    return guarded(nullptr, [&] {
        if (opt->precision != SF_PREC_MIXED)
            throw Status{SF_ERR_UNSUPPORTED, "sf_step_noalloc takes SF_PREC_MIXED (FP32 fields)"};
        if (opt->scalar_advection != 0)
            throw Status{SF_ERR_UNSUPPORTED, "sf_step_noalloc runs the plain scalar advection"};
        if (nsteps < 0) throw Status{SF_ERR_INVALID, "negative step count"};
        for (int k = 0; k < 4; ++k)
            if (!in[k] || !out[k]) throw Status{SF_ERR_INVALID, "null field pointer"};
        if (!uair0 || !vair0) throw Status{SF_ERR_INVALID, "null air-flow pointer"};
        const long long cells = 2LL * bubble->n_theta * bubble->n_theta;
        if (cells > kNoAllocMaxCells)
            throw Status{SF_ERR_UNSUPPORTED,
                         "sf_step_noalloc takes bubbles of at most " + std::to_string(kNoAllocMaxCells) +
                             " cells (n_theta " + std::to_string(bubble->n_theta) + " has " +
                             std::to_string(cells) + ")"};
        return (int)SF_OK;
    });
}

int step_in_scratch(void* scratch, size_t bytes, const sf_options* opt,
                    const sf_bubble_desc* bubble, double t_seconds, int nsteps,
                    hipStream_t stream, const float* const in[4], const float* uair0,
                    const float* vair0, float* const out[4], sf_bubble_stats* stats) {
    const int rc0 = check_noalloc(opt, bubble, nsteps, in, uair0, vair0, out);
    if (rc0 != SF_OK) return rc0;
    return guarded(nullptr, [&] {
        std::unique_ptr<Impl<float, float>> impl(new Impl<float, float>());
        Impl<float, float>& c = *impl;
        c.opt = *opt;
        c.device = opt->device;
        c.nb = 1;
        c.desc.assign(bubble, bubble + 1);
        c.noalloc = true;
        c.stream = stream;
        {
            Arena a;
            a.base = static_cast<char*>(scratch);
            a.size = bytes;
            a.stream = stream;
            ArenaScope scope(&a);
            setup(c, SF_PREC_MIXED);
        }
        const BubbleDev& B = c.bub_h[0];
        const size_t N = (size_t)B.nth * B.nph, Nt = N - B.nph;
        c.time_h[0].t = t_seconds;
        SF_HIP(hipMemcpy(c.time_d, c.time_h.data(), sizeof(BubbleTime), hipMemcpyHostToDevice));
        // the caller's state and air flow into the work arrays (the u_theta
        // array keeps its zero padding row)
        SF_HIP(hipMemcpyAsync(c.cur.uth, in[0], Nt * sizeof(float), hipMemcpyDeviceToDevice, stream));
        SF_HIP(hipMemcpyAsync(c.cur.uph, in[1], N * sizeof(float), hipMemcpyDeviceToDevice, stream));
        SF_HIP(hipMemcpyAsync(c.cur.eta, in[2], N * sizeof(float), hipMemcpyDeviceToDevice, stream));
        SF_HIP(hipMemcpyAsync(c.cur.gam, in[3], N * sizeof(float), hipMemcpyDeviceToDevice, stream));
        SF_HIP(hipMemcpyAsync(c.ua0, uair0, N * sizeof(float), hipMemcpyDeviceToDevice, stream));
        SF_HIP(hipMemcpyAsync(c.va0, vair0, Nt * sizeof(float), hipMemcpyDeviceToDevice, stream));
        launch_nan_scan<float>(stream, c.geom, c.bub_d, c.time_d, c.cur);
        // a fixed schedule (no host round trip inside the steps): max_outer
        // refinements of max_inner inner iterations, then the spare
        // refinement whose correction is the last line solve; a bubble that
        // has converged skips the remaining kernels
        const int K = opt->max_inner > 0 ? opt->max_inner : 16;
        const int R = opt->max_outer > 0 ? std::min(opt->max_outer, kAccRefine - 1) : 2;
        c.sched.assign((size_t)R, K);
        c.sched.push_back(0);
        for (int k = 0; k < nsteps; ++k) {
            c.advect();
            c.assemble();
            c.solve(false, false);
            c.update(true);
        }
        c.steps_done = nsteps;
        SF_HIP(hipMemcpyAsync(out[0], c.cur.uth, Nt * sizeof(float), hipMemcpyDeviceToDevice, stream));
        SF_HIP(hipMemcpyAsync(out[1], c.cur.uph, N * sizeof(float), hipMemcpyDeviceToDevice, stream));
        SF_HIP(hipMemcpyAsync(out[2], c.cur.eta, N * sizeof(float), hipMemcpyDeviceToDevice, stream));
        SF_HIP(hipMemcpyAsync(out[3], c.cur.gam, N * sizeof(float), hipMemcpyDeviceToDevice, stream));
        if (stats) fill_stats(c, 0, stats);
        SF_HIP(hipStreamSynchronize(stream));
        return (int)SF_OK;
    });
}

}  // namespace sfapi

namespace {
}  // namespace

extern "C" {

int sf_derive_params(const sf_bubble_desc* desc, sf_derived* out) {
    if (!desc || !out) return fail(nullptr, SF_ERR_INVALID, "null argument");
    sfh::derive_params(*desc, *out);
    return SF_OK;
}

int sf_create(const sf_options* opt, int32_t nbubbles, const sf_bubble_desc* bubbles,
              sf_ctx** out) {
    if (!out) return fail(nullptr, SF_ERR_INVALID, "null output pointer");
    *out = nullptr;
    const int rc0 = validate_create(opt, nbubbles, bubbles);
    if (rc0 != SF_OK) return rc0;
    auto ctx = std::unique_ptr<sf_ctx>(new sf_ctx());
    ctx->prec = opt->precision;
    ctx->impl = make_impl(opt->precision);
    Base& c = *ctx->impl;
    c.opt = *opt;
    c.device = opt->device;
    c.nb = nbubbles;
    c.desc.assign(bubbles, bubbles + nbubbles);
    int rc = guarded(nullptr, [&] {
        setup(c, opt->precision);
        return (int)SF_OK;
    });
    if (rc != SF_OK) return rc;
    *out = ctx.release();
    return SF_OK;
}

int sf_destroy(sf_ctx* ctx) {
    if (!ctx) return SF_OK;
    delete ctx;
    return SF_OK;
}

int sf_upload_field(sf_ctx* ctx, int32_t b, int32_t field, const double* host) {
    return guarded(ctx, [&] {
        check_index(ctx, b);
        if (!host) throw Status{SF_ERR_INVALID, "null host buffer"};
        ctx->impl->upload(b, field, host);
        return (int)SF_OK;
    });
}

int sf_download_field(sf_ctx* ctx, int32_t b, int32_t field, double* host) {
    return guarded(ctx, [&] {
        check_index(ctx, b);
        if (!host) throw Status{SF_ERR_INVALID, "null host buffer"};
        ctx->impl->download(b, field, host);
        return (int)SF_OK;
    });
}

int sf_set_time(sf_ctx* ctx, int32_t b, double t) {
    return guarded(ctx, [&] {
        if (!ctx || !ctx->impl) throw Status{SF_ERR_INVALID, "null context"};
        Base& c = *ctx->impl;
        if (b >= c.nb) throw Status{SF_ERR_INVALID, "bubble index out of range"};
        // in stream order, after the steps already queued (no host wait)
        launch_set_time(c.stream, c.time_d, c.nb, b < 0 ? -1 : b, t);
        SF_HIP(hipGetLastError());
        return (int)SF_OK;
    });
}

int sf_set_dt(sf_ctx* ctx, int32_t b, double dt_seconds) {
    return guarded(ctx, [&] {
        if (!ctx || !ctx->impl) throw Status{SF_ERR_INVALID, "null context"};
        if (!(dt_seconds > 0)) throw Status{SF_ERR_INVALID, "dt must be positive"};
        Base& c = *ctx->impl;
        if (b >= c.nb) throw Status{SF_ERR_INVALID, "bubble index out of range"};
        SF_HIP(hipStreamSynchronize(c.stream));
        SF_HIP(hipMemcpy(c.time_h.data(), c.time_d, c.nb * sizeof(BubbleTime), hipMemcpyDeviceToHost));
        for (int k = 0; k < c.nb; ++k) {
            if (b >= 0 && b != k) continue;
            // the last step's time advance, still pending for the next
            // advection, is by the step size it was taken with
            if (c.time_h[k].pending) {
                c.time_h[k].t += c.time_h[k].dt_s;
                c.time_h[k].pending = 0;
            }
            // timeStepGlobal = dt U / radius with the member U = 1 (bubble.cu:2675)
            const double dtn = dt_seconds * 1.0 / c.der[k].radius;
            // a change of more than a relative 1e-9 (not the rounding of a
            // frame's last step that lands on the frame's end)
            if (std::fabs(dtn - c.bub_h[k].dt) > 1e-9 * c.bub_h[k].dt) {
                // the stored solve corrections and the inner schedule belong
                // to the old step
                c.time_h[k].ghist = c.time_h[k].gpend = 0;
                c.host_hist = 0;
                drop_calibration(c);
            }
            c.bub_h[k].dt = dtn;
            c.time_h[k].dt_s = dt_seconds;
        }
        SF_HIP(hipMemcpy(c.bub_d, c.bub_h.data(), c.nb * sizeof(BubbleDev), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(c.time_d, c.time_h.data(), c.nb * sizeof(BubbleTime), hipMemcpyHostToDevice));
        return (int)SF_OK;
    });
}

int sf_step(sf_ctx* ctx, int32_t nsteps) {
    return guarded(ctx, [&] {
        if (!ctx || !ctx->impl) throw Status{SF_ERR_INVALID, "null context"};
        if (nsteps < 0) throw Status{SF_ERR_INVALID, "negative step count"};
        Base& c = *ctx->impl;
        if (c.stage != 0) throw Status{SF_ERR_ORDER, "a stage-level step is in progress"};
        const bool want_graph = c.opt.use_graph && !c.fp64_inner();
        // captured steps hold no BIMOCQ restart: an upload's takes effect here
        c.apply_uploads();
        // the extrapolated initial guess is in full effect from the third
        // solve after a reset; steps before that run host-driven
        const int warm = c.guess ? 2 : 0;
        int k = 0;
        if (want_graph && c.opt.autotune && !c.tuned && c.steps_done >= kTuneAfterSteps) {
            autotune(c);
            // captured graphs hold the old launch geometry
            for (auto& g : c.graphs) (void)hipGraphExecDestroy(g.second.exec);
            c.graphs.clear();
            drop_calibration(c);
        }
        for (int pass = 0; pass < 2 && want_graph && !c.calibrated; ++pass) {
            const bool was_mg = c.mg;
            // calibration: one host-driven step with the guess in effect sets
            // the first schedule
            for (; k < nsteps && c.calib_steps < 1; ++k) {
                c.recording = c.host_hist >= warm;
                eager_step(c, true);
                if (c.recording) ++c.calib_steps;
                c.recording = false;
            }
            if (c.calib_steps == 1) {
                // schedule: per refinement one inner iteration fewer than the
                // host-driven step ran (it stopped at the first n whose
                // residual met the target, and launch n - 1 already produces
                // that iterate e_n), then one spare refinement of one
                // iteration (a further factor of about ten) that catches a
                // step needing more; a bubble that converged skips the
                // remaining kernels
                c.sched.clear();
                for (int n : c.calib)
                    c.sched.push_back(c.opt.max_inner > 0 ? c.opt.max_inner
                                                          : std::max(c.method == 2 ? n - 1 : n, 1));
                if (c.sched.empty()) c.sched.push_back(1);
                // a solve that misses its tolerance stops the bubble when
                // halt_on_failure is set: one more refinement of margin
                if (c.opt.halt_on_failure) c.sched.push_back(1);
                c.sched.push_back(1);
                if (c.opt.max_outer > 0) c.sched.resize((size_t)c.opt.max_outer, 1);
                if (c.sched.size() > (size_t)kAccRefine) c.sched.resize(kAccRefine);
                // Chebyshev: the last refinement (run only when those before
                // missed the tolerance) takes the line solve z = M^-1 r that
                // the residual before it produced as its correction, with no
                // inner launch (the start-residual test moves into its
                // residual launch): one skipped launch per step fewer
                if (c.method == 2 && c.sched.size() >= 2) c.sched.back() = 0;
                c.calibrated = true;
                if (c.sched[0] > kGraphMaxInner) {
                    drop_calibration(c);
                    // the calibration step switched to multigrid: calibrate
                    // that schedule
                    if (c.mg && !was_mg) continue;
                }
                if (c.trace_solve) {
                    std::fprintf(stderr, "[sf solve] graph schedule:");
                    for (int n : c.sched) std::fprintf(stderr, " %d", n);
                    std::fprintf(stderr, "\n");
                }
                if (c.calibrated) {
                    const bool adapt = c.method == 2 && c.opt.max_inner <= 0;
                    (void)step_graph(c, kGraphSteps, -1);
                    if (adapt) {
                        // the graphs of the calibrated schedule and of one
                        // inner iteration fewer (where the schedule usually
                        // settles as the extrapolated guess improves), so
                        // that no instantiation falls between replays
                        const int n0 = c.sched[0];
                        for (int n : {n0, n0 - 1}) {
                            if (n < 1) continue;
                            c.sched[0] = n;
                            (void)step_graph(c, kGraphSteps, 0);
                            (void)step_graph(c, kGraphSteps, 1);
                        }
                        c.sched[0] = n0;
                    }
                }
            }
        }
        if (want_graph && c.calibrated) {
            for (; k < nsteps && c.host_hist < warm; ++k) eager_step(c, true);
            // replays in chunks of kAdaptChunk steps, one graph launch each;
            // with the automatic schedule each chunk's graph ends by handing
            // its statistics to the host, which reads them one chunk later,
            // so the device never waits for the host
            const bool adapt = c.method == 2 && c.opt.max_inner <= 0;
            if (k < nsteps) {
                SF_HIP(hipEventRecord(c.ev[0], c.stream));
                int chunk = 0;
                while (k < nsteps && c.sched[0] <= kGraphMaxInner) {
                    const int n = std::min(nsteps - k, kAdaptChunk);
                    const int p = chunk & 1;
                    // kGraphSteps steps per launch, the chunk's last launch
                    // ending with the statistics hand-off
                    for (int q = 0; q < n; q += kGraphSteps) {
                        const int m = std::min(kGraphSteps, n - q);
                        const bool last = q + m == n;
                        SF_HIP(hipGraphLaunch(step_graph(c, m, adapt && last ? p : -1), c.stream));
                    }
                    k += n;
                    c.host_hist += n;
                    if (!adapt) continue;
                    SF_HIP(hipEventRecord(c.cev[p], c.stream));
                    c.chunk_sched[p] = c.sched[0];
                    if (chunk > 0) {
                        SF_HIP(hipEventSynchronize(c.cev[p ^ 1]));
                        adapt_schedule(c, c.sstat_h + (p ^ 1) * c.nb, c.chunk_sched[p ^ 1]);
                    }
                    ++chunk;
                }
                SF_HIP(hipEventRecord(c.ev[5], c.stream));
                SF_HIP(hipEventSynchronize(c.ev[5]));
                if (adapt && chunk > 0) {
                    const int p = (chunk - 1) & 1;
                    adapt_schedule(c, c.sstat_h + p * c.nb, c.chunk_sched[p]);
                }
                float ms = 0.f;
                SF_HIP(hipEventElapsedTime(&ms, c.ev[0], c.ev[5]));
                c.timing.step_ms += ms;
            }
            if (c.sched[0] > kGraphMaxInner) drop_calibration(c);
        }
        for (; k < nsteps; ++k) eager_step(c, true);
        c.steps_done += nsteps;
        return (int)SF_OK;
    });
}

int sf_time_advection(sf_ctx* ctx, int32_t nlaunch, double* avg_ms) {
    return guarded(ctx, [&] {
        if (!ctx || !ctx->impl || !avg_ms) throw Status{SF_ERR_INVALID, "null argument"};
        if (nlaunch <= 0) throw Status{SF_ERR_INVALID, "launch count must be positive"};
        Base& c = *ctx->impl;
        if (c.stage != 0) throw Status{SF_ERR_ORDER, "sf_time_advection must start a step"};
        // every launch recomputes the advected fields from the same state
        c.advect();
        SF_HIP(hipEventRecord(c.ev[0], c.stream));
        for (int k = 0; k < nlaunch; ++k) c.advect();
        SF_HIP(hipEventRecord(c.ev[5], c.stream));
        SF_HIP(hipEventSynchronize(c.ev[5]));
        float ms = 0.f;
        SF_HIP(hipEventElapsedTime(&ms, c.ev[0], c.ev[5]));
        *avg_ms = (double)ms / nlaunch;
        return (int)SF_OK;
    });
}

int sf_time_inner(sf_ctx* ctx, int32_t nlaunch, double* avg_ms) {
    return guarded(ctx, [&] {
        if (!ctx || !ctx->impl || !avg_ms) throw Status{SF_ERR_INVALID, "null argument"};
        if (nlaunch <= 0) throw Status{SF_ERR_INVALID, "launch count must be positive"};
        Base& c = *ctx->impl;
        if (c.stage != 2) throw Status{SF_ERR_ORDER, "sf_time_inner must follow sf_assemble"};
        if (c.method != 2) throw Status{SF_ERR_INVALID, "the inner method is not Chebyshev"};
        // the launches are the second inner iteration of the first
        // refinement (no start test; e_1 = z from the assembly, e_2 whatever
        // its buffer holds): they write e_3, which the solve writes before
        // it reads it, so the solve that follows is unchanged; the control
        // blocks (their iteration counts) are restored
        std::vector<SolveCtl> keep;
        c.ctl_to_host(keep);
        c.inner_launch(2, 0);
        SF_HIP(hipEventRecord(c.ev[0], c.stream));
        for (int k = 0; k < nlaunch; ++k) c.inner_launch(2, 0);
        SF_HIP(hipEventRecord(c.ev[5], c.stream));
        SF_HIP(hipEventSynchronize(c.ev[5]));
        float ms = 0.f;
        SF_HIP(hipEventElapsedTime(&ms, c.ev[0], c.ev[5]));
        c.ctl_to_device(keep);
        *avg_ms = (double)ms / nlaunch;
        return (int)SF_OK;
    });
}

int sf_advect(sf_ctx* ctx) {
    return guarded(ctx, [&] {
        if (!ctx || !ctx->impl) throw Status{SF_ERR_INVALID, "null context"};
        Base& c = *ctx->impl;
        if (c.stage != 0) throw Status{SF_ERR_ORDER, "sf_advect must start a step"};
        c.advect();
        c.stage = 1;
        return (int)SF_OK;
    });
}

int sf_assemble(sf_ctx* ctx) {
    return guarded(ctx, [&] {
        if (!ctx || !ctx->impl) throw Status{SF_ERR_INVALID, "null context"};
        Base& c = *ctx->impl;
        if (c.stage != 1) throw Status{SF_ERR_ORDER, "sf_assemble must follow sf_advect"};
        c.assemble();
        c.stage = 2;
        return (int)SF_OK;
    });
}

int sf_solve(sf_ctx* ctx) {
    return guarded(ctx, [&] {
        if (!ctx || !ctx->impl) throw Status{SF_ERR_INVALID, "null context"};
        Base& c = *ctx->impl;
        if (c.stage != 2) throw Status{SF_ERR_ORDER, "sf_solve must follow sf_assemble"};
        c.solve(true, true);
        c.stage = 3;
        SF_HIP(hipStreamSynchronize(c.stream));
        std::vector<SolveCtl> h(c.nb);
        SF_HIP(hipMemcpy(h.data(), c.ctl_d, c.nb * sizeof(SolveCtl), hipMemcpyDeviceToHost));
        for (auto& s : h)
            if (!s.converged) throw Status{SF_ERR_NOT_CONVERGED, "solve missed its tolerance"};
        return (int)SF_OK;
    });
}

int sf_update(sf_ctx* ctx) {
    return guarded(ctx, [&] {
        if (!ctx || !ctx->impl) throw Status{SF_ERR_INVALID, "null context"};
        Base& c = *ctx->impl;
        if (c.stage != 3) throw Status{SF_ERR_ORDER, "sf_update must follow sf_solve"};
        c.update(false);
        c.stage = 0;
        c.steps_done += 1;
        c.host_hist += 1;
        SF_HIP(hipStreamSynchronize(c.stream));
        return (int)SF_OK;
    });
}

int sf_synchronize(sf_ctx* ctx) {
    return guarded(ctx, [&] {
        if (!ctx || !ctx->impl) throw Status{SF_ERR_INVALID, "null context"};
        SF_HIP(hipStreamSynchronize(ctx->impl->stream));
        return (int)SF_OK;
    });
}

int sf_get_stats(sf_ctx* ctx, int32_t b, sf_bubble_stats* out) {
    return guarded(ctx, [&] {
        check_index(ctx, b);
        if (!out) throw Status{SF_ERR_INVALID, "null output"};
        fill_stats(*ctx->impl, b, out);
        return (int)SF_OK;
    });
}

int sf_get_tuning(sf_ctx* ctx, sf_tuning* out) {
    if (!ctx || !ctx->impl || !out) return fail(ctx, SF_ERR_INVALID, "null argument");
    *out = ctx->impl->tuning;
    return SF_OK;
}

int sf_get_timing(sf_ctx* ctx, sf_timing* out) {
    if (!ctx || !ctx->impl || !out) return fail(ctx, SF_ERR_INVALID, "null argument");
    *out = ctx->impl->timing;
    return SF_OK;
}

int sf_reset_timing(sf_ctx* ctx) {
    if (!ctx || !ctx->impl) return fail(ctx, SF_ERR_INVALID, "null argument");
    std::memset(&ctx->impl->timing, 0, sizeof(sf_timing));
    return SF_OK;
}

const char* sf_last_error(const sf_ctx* ctx) {
    if (ctx && ctx->impl) return ctx->impl->err.c_str();
    return g_last_error.c_str();
}

int sf_thickness_half(sf_ctx* ctx, int32_t b, uint16_t* host) {
    return guarded(ctx, [&] {
        check_index(ctx, b);
        if (!host) throw Status{SF_ERR_INVALID, "null buffer"};
        Base& c = *ctx->impl;
        if (c.stage != 0) throw Status{SF_ERR_ORDER, "a stage-level step is in progress"};
        c.thickness_half(b, host);
        return (int)SF_OK;
    });
}

int sf_write_thickness_exr(sf_ctx* ctx, int32_t b, const char* path) {
    return guarded(ctx, [&] {
        check_index(ctx, b);
        if (!path) throw Status{SF_ERR_INVALID, "null path"};
        Base& c = *ctx->impl;
        const BubbleDev& B = c.bub_h[b];
        std::vector<double> eta((size_t)B.nth * B.nph);
        c.download(b, SF_FIELD_ETA, eta.data());
        if (!sfh::write_thickness_exr(path, B.nth, B.nph, eta.data(), c.desc[b].H))
            throw Status{SF_ERR_INVALID, std::string("cannot write ") + path};
        return (int)SF_OK;
    });
}

}  // extern "C"
