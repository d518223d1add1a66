// Host-side launch interface of the soap-film kernels (sf_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "sf_device.h"

namespace sfk {

// BIMOCQ state of all bubbles (packed like the fields, FP64): maps (unit
// vectors of the mapped positions), the new maps of the current step, and for eta and Gamma the
// initial fields and accumulated changes of the current and the previous
// reinitialisation interval; per-step changes, errors and copies; per-bubble
// distortion, reinitialisation flag and counters.
struct BimocqPtrs {
    UMap fwd, bwd, prev, nfwd, nbwd;   // maps as unit vectors (x, y, z arrays)
    double *e_init, *e_initlast, *e_delta, *e_deltalast;
    double *g_init, *g_initlast, *g_delta, *g_deltalast;
    double *chg_e, *chg_g, *err_e, *err_g, *cpy_e, *cpy_g;
    unsigned* dist_bits;   // per bubble: running max distortion (float bits)
    int* flag;             // per bubble: reinitialise at this step
    double* dist_last;     // per bubble: distortion of the last step
    int* reinits;          // per bubble: reinitialisations so far
    double blend, threshold;
    int multisample;
    Tables tab;            // the grid's sin/cos tables (cell-relative map positions)
    const BubbleTime* bt;  // per bubble: the NaN flag selects the NaN-skipping samplers
};

template <typename T>
struct StatePtrs {
    T* uth;
    T* uph;
    T* eta;
    T* gam;
};

template <typename T>
struct AirPtrs {
    const T* ua0;  // u_phi-face air velocity at t = 0 (uair_init)
    const T* va0;  // u_theta-face air velocity at t = 0 (vair_init)
    const T* au;   // corridor shape projected on e_phi (per unit time factor)
    const T* av;   // corridor shape projected on e_theta
    // VISCOUS forces (null without): fu on the u_phi faces, fv on the
    // u_theta faces, in the cell layout (FP64)
    const double* fu;
    const double* fv;
};

template <typename Ti>
struct SysPtrs {
    double* b64;     // right-hand side
    double* x64;     // initial guess written by the assembly
    Ti* wN;          // north-face weight of each cell (row 0: 0)
    Ti* wW;
    Ti* diag;
    Ti* e;           // inner correction
    Ti* r;           // inner residual
    Ti* z;           // preconditioned residual
    Ti* Ap;          // PCG: A p; Chebyshev: the second inner iterate buffer
    int* rsteps;     // line-solve reductions per row (set by the assembly)
    double* acc;     // per-launch sums (acc_index layout)
};

struct SolveParams {
    double tol;         // outer tolerance on |r| / |b| (or |r|)
    double inner_rho2;  // inner stop: |r|^2 <= inner_rho2 * |r_0|^2
    double inner_frac;  // ... or |r| <= inner_frac * tol * |b|
    int absolute;       // 1: tolerance on |r| (AmgX ABSOLUTE)
    int pad_;
};

// One level of the multigrid preconditioner, per bubble: its row count (0:
// the bubble has no such level) and the first cell of its rows in the
// level's packed arrays.
struct MgLevelDev {
    int n;
    int pad_;
    long long off;
};

// The arrays of one multigrid level, packed over the bubbles: operator
// (north and west face weights, diagonal), the iterate, and the two terms of
// the right-hand side restricted from the level below (coarse row J: U the
// term of its own odd row 2J + 1, D the term of the odd row 2J - 1 above).
template <typename Ti>
struct MgPtrs {
    Ti* wN;
    Ti* wW;
    Ti* dg;
    Ti* x;
    Ti* U;
    Ti* D;
    const MgLevelDev* lv;
};

// Options of a multigrid half-sweep (k_mg_line).
constexpr int kMgZeroNb = 1, kMgProlong = 2, kMgRestrict = 4, kMgRz = 8, kMgGateInner = 16,
              kMgGateStart = 32;

// Threads of a row block at most: one per C consecutive cells of the widest
// row, C in {2, 4, 8, 16}.
constexpr int kRowThreadsMax = 1024;

struct LaunchGeom {
    const int4* tiles;   // cell tiles: (bubble, row, first column, width)
    int ntiles;
    const int2* rows;    // row blocks: (bubble, row)
    int nrows;
    int row_threads;     // threads of a row block
    int cpt;             // consecutive cells per thread of a row block (2, 4, 8 or 16)
    size_t lds_bytes;    // dynamic LDS of a row block (the reduced line system)
    bool pow2_rows;      // every bubble's row length is a power of two
    bool packed;         // some block holds several rows (RowSeg)
};

template <typename T>
void launch_advect(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, BubbleTime* bt,
                   const Tables& tab, StatePtrs<T> in, StatePtrs<T> out);
template <typename T, typename Ti>
void launch_assemble(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub,
                     BubbleTime* bt, const Tables& tab, StatePtrs<T> st, AirPtrs<T> air,
                     SysPtrs<Ti> sys, double gmax_state, SolveCtl* ctl, double* part, long stride,
                     const T* guess, const SolveParams& sp);
template <typename T, typename Ti>
void launch_resid(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, const BubbleTime* bt,
                  const Tables& tab, StatePtrs<T> st, SolveCtl* ctl, SysPtrs<Ti> sys,
                  const T* guess, const double* xo, double* xn, int k, double* part, long stride,
                  const double* partB, const SolveParams& sp, SchedStat* sh, int check_prev);
template <typename Ti>
void launch_spmv(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                 SysPtrs<Ti> sys, const Ti* pold, Ti* pnew, int it, const SolveParams& sp,
                 const double* partR, const double* partB, double* partA, long stride);
template <typename Ti>
void launch_cgstep(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                   SysPtrs<Ti> sys, const Ti* p, int it, const double* partA, double* partB,
                   long stride);
template <typename Ti>
void launch_cheb(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                 SysPtrs<Ti> sys, const Ti* ek, const Ti* ekm1, Ti* enext, int it, int kr,
                 const SolveParams& sp, double* partB, long stride);
// VISCOUS: div(u) at the centres, then fu, fv (cell layout)
template <typename T>
void launch_viscous(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, const Tables& tab,
                    StatePtrs<T> st, double* div, double* fu, double* fv);
// multigrid preconditioner: the coarse operator of a level's rows from the
// level below; one half-sweep of zebra phi-line relaxation (flags as
// k_mg_line)
template <typename Ti>
void launch_mg_coarsen(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, MgPtrs<Ti> fine,
                       MgPtrs<Ti> coarse);
template <typename Ti>
void launch_mg_line(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                    MgPtrs<Ti> lvl, MgPtrs<Ti> coarse, const Ti* r0, int level, int flags,
                    const int* rsteps, const SolveParams& sp, const double* partR,
                    double* rz_out, long stride);
void launch_finalize(hipStream_t s, int nb, const BubbleDev* bub, SolveCtl* ctl, SolveHist* hist,
                     const SolveParams& sp, const double* partR, long stride);
void launch_sched_stats(hipStream_t s, SchedStat* d, SchedStat* h, int nb);
template <typename T>
void launch_update(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, BubbleTime* bt,
                   SolveCtl* ctl, const Tables& tab, StatePtrs<T> st, AirPtrs<T> air,
                   const double* x0, const double* x1, StatePtrs<T> out, int fin, SolveHist* hist,
                   const SolveParams& sp, const double* partR, long stride, double* chg_e,
                   double* chg_g, T* gdel, T* gnext, double* acc);
// the air-flow time of bubble b (all when b < 0), no pending advance
void launch_set_time(hipStream_t s, BubbleTime* bt, int nb, int b, double t);
// the frame output's half-precision gray plane of n cells (k_thickness_half)
template <typename T>
void launch_thickness_half(hipStream_t s, const T* eta, long n, double H, unsigned short* out);
// BIMOCQ stages (packed over all bubbles)
template <typename T>
void launch_bimocq_init(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub,
                        StatePtrs<T> cur, const BimocqPtrs& P);
template <typename T>
void launch_bimocq_maps(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub,
                        BubbleTime* bt, const Tables& tab, StatePtrs<T> cur, const BimocqPtrs& P);
template <typename T>
void launch_bimocq_eta(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub,
                       StatePtrs<T> cur, StatePtrs<T> out, const BimocqPtrs& P);
void launch_bimocq_accumulate(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, int nb,
                              const BimocqPtrs& P);
template <typename T>
void launch_bimocq_reinit(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub,
                          StatePtrs<T> cur, const BimocqPtrs& P);
// sets the per-bubble NaN flag of bubbles whose state holds a NaN
template <typename T>
void launch_nan_scan(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, BubbleTime* bt,
                     StatePtrs<T> st);
template <typename T>
void launch_row_stats(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, const Tables& tab,
                      const T* eta, double* out, long stride);

}  // namespace sfk
