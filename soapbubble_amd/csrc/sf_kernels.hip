// HIP kernels of the soap-film step for gfx950 (MI355X).
//
// Stage kernels (one launch each per step, packed over all bubbles):
//   k_advect    great-circle semi-Lagrangian advection of u_theta, u_phi,
//               eta and Gamma (advectionVSphere{Theta,Phi}Kernel and
//               advectionAllCentered, bubble.cu:639-961) fused per cell
//   k_assemble  the sinTheta-scaled implicit Gamma system
//               (concentrationLinearSystemKernel, bubble.cu:1301-1473) as a
//               matrix-free 5-point operator: face weights, diagonal, FP64 rhs
//   k_resid     FP64 residual of the refinement loop + phi-line solve
//   k_spmv      inner PCG: p update, A p, p.Ap
//   k_cgstep    inner PCG: x, r update, phi-line solve, r.z
//   k_finalize  convergence verdict of the solve
//   k_update    divergenceKernel_fromGamma, applyforcevel{theta,phi} and
//               applyforceThickness (bubble.cu:1276-1295, 1819-1935,
//               2041-2063) fused per cell
//
// Template parameter T is the state precision (u, eta, Gamma, air), Ti the
// precision of the inner solve. The refinement vectors (x, b, residual) and
// the assembled coefficients are FP64.
#include "sf_device.h"
#include "sf_launch.h"

namespace sfk {

constexpr double kEps = 1e-7;   // bubble.cu:28
constexpr double kThresh = 0.0; // bubble.cu:27

// ------------------------------------------------------------ advection ----
template <typename T>
__global__ __launch_bounds__(256) void k_advect(const int4* __restrict__ tiles,
                                                const BubbleDev* __restrict__ bub,
                                                BubbleTime* __restrict__ btime, Tables tab,
                                                StatePtrs<T> in, StatePtrs<T> out) {
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0) return;
    const BubbleDev B = bub[tl.x];
    // the previous step's time advance lands here, before any kernel of this
    // step reads the time
    if (tl.y == 0 && tl.z == 0 && threadIdx.x == 0 && btime[tl.x].pending) {
        btime[tl.x].t += btime[tl.x].dt_s;
        btime[tl.x].pending = 0;
    }
    const int j = tl.y + (int)threadIdx.x / tl.w;
    const int i = tl.z + (int)threadIdx.x % tl.w;
    if (j >= B.nth || i >= B.nph) return;
    const size_t c = (size_t)j * B.nph + i;

    VelCtx<T> v;
    v.uth = Field<T>{in.uth + B.off, B.nph, B.mask, B.nth - 2};
    v.uph = Field<T>{in.uph + B.off, B.nph, B.mask, B.nth - 1};
    v.nth = B.nth;
    v.invh = (T)B.invh;
    v.dt = (T)B.dt;
    const Field<T> feta{in.eta + B.off, B.nph, B.mask, B.nth - 1};
    const Field<T> fgam{in.gam + B.off, B.nph, B.mask, B.nth - 1};

    const T sinTc = TabSel<T>::sTc(tab)[B.rt0 + j], cosTc = TabSel<T>::cTc(tab)[B.rt0 + j];
    const T sinPc = TabSel<T>::sPc(tab)[B.ct0 + i], cosPc = TabSel<T>::cPc(tab)[B.ct0 + i];

    // Start-point samples. Away from the two polar rows the sample positions
    // are grid-aligned, and the samplers reduce to their interior branch with
    // weights 0 or 1/2 (exactly the values the general path computes); the
    // polar rows go through the general samplers.
    const bool interior = j > 0 && j < B.nth - 1;

    // eta, Gamma at the cell centre (j+1/2, i): traceGreatCircle + two samples
    {
        const T gx = (T)j + T(0.5), gy = (T)i;
        T guT, guP;
        if (interior) {
            guT = lerp_nan(v.uth.belt(j - 1, i, T(0)), v.uth.belt(j, i, T(0)), T(0.5));
            guP = lerp_nan(v.uph.belt(j, i, T(0.5)), v.uph.belt(j + 1, i, T(0.5)), T(0));
        } else {
            guT = sample_vtheta(v.uth, B.nth, gx, gy);
            guP = sample_vphi(v.uph, B.nth, gx, gy);
        }
        Circle<T> o;
        T tx = gx, ty = gy;
        if (circle_mid(v, sinTc, cosTc, sinPc, cosPc, guT, guP, o)) {
            T s, co, px, py, pz, rho;
            circle_depart(v, o, s, co, px, py, pz, tx, ty, rho);
        }
        T en, gn;
        sample_centered2(feta, fgam, B.nth, tx, ty, en, gn);
        out.eta[B.off + c] = en;
        out.gam[B.off + c] = gn;
    }
    // u_phi at the west face (j+1/2, i-1/2)
    {
        const T gx = (T)j + T(0.5), gy = (T)i - T(0.5);
        const T sinP = TabSel<T>::sPh(tab)[B.ct0 + i], cosP = TabSel<T>::cPh(tab)[B.ct0 + i];
        T guT = interior
                    ? lerp_nan(v.uth.belt(j - 1, i - 1, T(0.5)), v.uth.belt(j, i - 1, T(0.5)), T(0.5))
                    : sample_vtheta(v.uth, B.nth, gx, gy);
        T guP = in.uph[B.off + c];
        Circle<T> o;
        T res;
        if (!circle_mid(v, sinTc, cosTc, sinP, cosP, guT, guP, o)) {
            res = o.unorm;
        } else {
            T r0, r1, r2;
            transport_back(v, o, r0, r1, r2);
            res = -sinP * r0 + cosP * r1;
        }
        out.uph[B.off + c] = res;
    }
    // u_theta at the south face (j+1, i); sampleVPhi there is always the
    // interior branch with weights 1/2
    if (j < B.nth - 1) {
        const T sinT = TabSel<T>::sTf(tab)[B.rt0 + j + 1], cosT = TabSel<T>::cTf(tab)[B.rt0 + j + 1];
        T guT = in.uth[B.off + c];
        T guP = lerp_nan(v.uph.belt(j, i, T(0.5)), v.uph.belt(j + 1, i, T(0.5)), T(0.5));
        Circle<T> o;
        T res;
        if (!circle_mid(v, sinT, cosT, sinPc, cosPc, guT, guP, o)) {
            res = o.unorm;
        } else {
            T r0, r1, r2;
            transport_back(v, o, r0, r1, r2);
            res = r0 * (cosT * cosPc) + r1 * (cosT * sinPc) - r2 * sinT;
        }
        out.uth[B.off + c] = res;
    }
}

// ---------------------------------------------------------- air flow ----
// u_air(t) = u_air_init + (3 - |t - 3|) * corridor, NaN -> 0 (airFlowU/V,
// bubble.cu:2301-2380). The corridor shape is time-independent and is
// precomputed on the host.
template <typename T>
__device__ __forceinline__ double air_at(const T* __restrict__ a0, const T* __restrict__ shape,
                                         size_t k, double kt) {
    double v = (double)a0[k] + (double)shape[k] * kt;
    return __builtin_isnan(v) ? 0.0 : v;
}

__device__ __forceinline__ bool is_broken(double gamma, double eta, double gmax) {
    return gamma > gmax - kEps || eta == kThresh;
}

// ---------------------------------------------------------- assembly ----
// Face weights are the negated off-diagonal entries of the reference matrix:
// w_N = s2 sin(theta_N) M dt / (eta_N,f + Cr dt), w_W = s2 csc(theta) M dt /
// (eta_W,f + Cr dt). A broken cell keeps the reference's identity row with
// rhs gamma_max; its coupling to unbroken neighbours is moved to their
// right-hand sides (x_broken = gamma_max is known), which leaves the
// solution unchanged and keeps the operator symmetric.
struct CellSys {
    double wN, wS, wW, wE;   // effective face weights (0 across broken cells)
    double diag, rhs;
    double x0, x0N, x0S, x0W, x0E;  // initial guess Gamma* at the cell and neighbours
    bool brk;
};

template <typename T>
__device__ __forceinline__ CellSys assemble_cell(const BubbleDev& B, double t, const Tables& tab,
                                                 const StatePtrs<T>& s, const AirPtrs<T>& air,
                                                 int j, int i) {
    const int nph = B.nph, m = B.mask, nth = B.nth;
    const size_t o = (size_t)B.off;
    const size_t c = o + (size_t)j * nph + i;
    const int iE = (i + 1) & m, iW = (i - 1) & m;
    const double gmax = B.gamma_max;
    const T* eta = s.eta;
    const T* gam = s.gam;
    const double e = (double)eta[c];
    const double gamma = (double)gam[c];
    CellSys r;
    r.brk = is_broken(gamma, e, gmax);

    const size_t cW = o + (size_t)j * nph + iW, cE = o + (size_t)j * nph + iE;
    const size_t cN = j > 0 ? c - nph : o + (size_t)((i + nth) & m);
    const size_t cS = j < nth - 1 ? c + nph : o + (size_t)j * nph + ((i + nth) & m);
    const double etaW = (double)eta[cW], etaE = (double)eta[cE];
    const double etaN = (double)eta[cN], etaS = (double)eta[cS];
    const double gW = (double)gam[cW], gE = (double)gam[cE];
    const double gN = (double)gam[cN], gS = (double)gam[cS];
    const bool bW = is_broken(gW, etaW, gmax);
    const bool bE = is_broken(gE, etaE, gmax);
    const bool bN = j > 0 && is_broken(gN, etaN, gmax);
    const bool bS = j < nth - 1 && is_broken(gS, etaS, gmax);

    const double dt = B.dt, invDt = 1.0 / dt, Cr = B.Cr;
    const double CrDt = Cr * dt, MDt = B.M * dt, s2 = B.invh * B.invh;
    const double sinN = tab.sinFace[B.rt0 + j], sinS = tab.sinFace[B.rt0 + j + 1];
    const double sinTheta = tab.sinTc[B.rt0 + j], csc = tab.cscTc[B.rt0 + j];
    const double etaUp = (etaN + e) / 2.0, etaLeft = (etaW + e) / 2.0;
    const double etaRight = (etaE + e) / 2.0, etaDown = (etaS + e) / 2.0;
    const double wN = s2 * sinN * MDt / (etaUp + CrDt);
    const double wW = s2 * csc * MDt / (etaLeft + CrDt);
    const double wE = s2 * csc * MDt / (etaRight + CrDt);
    const double wS = s2 * sinS * MDt / (etaDown + CrDt);

    if (r.brk) {
        r.diag = 1.0;
        r.rhs = gmax;
    } else {
        r.diag = sinTheta * invDt / gamma - (((-wN) + (-wW)) + (-wE) + (-wS));
        double uW = (double)s.uph[c], uE = (double)s.uph[cE];
        double vN = j > 0 ? (double)s.uth[c - nph] : 0.0;
        double vS = j < nth - 1 ? (double)s.uth[c] : 0.0;
        if (__builtin_isnan(vN)) vN = 0.0;
        if (__builtin_isnan(vS)) vS = 0.0;
        if (__builtin_isnan(uW)) uW = 0.0;
        if (__builtin_isnan(uE)) uE = 0.0;
        double div = B.invh * (uE * etaRight / (etaRight + Cr * dt) -
                               uW * etaLeft / (etaLeft + Cr * dt) +
                               (vS * sinS * etaDown / (etaDown + Cr * dt) -
                                vN * sinN * etaUp / (etaUp + Cr * dt)));
        if (B.airflow) {
            const double kt = 3.0 - fabs(t - 3.0);
            const double bl = t / 12.0;
            const size_t row = o + (size_t)j * nph;
            const int opp = (nph - i) & m, oppE = (nph - iE) & m;
            const double uAirW = 0.5 * (air_at(air.ua0, air.au, row + i, kt) * bl +
                                        air_at(air.ua0, air.au, row + opp, kt) * (1 - bl));
            const double uAirE = 0.5 * (air_at(air.ua0, air.au, row + iE, kt) * bl +
                                        air_at(air.ua0, air.au, row + oppE, kt) * (1 - bl));
            double vAirN = 0.0, vAirS = 0.0;
            if (j > 0) {
                const size_t rn = row - nph;
                vAirN = 0.5 * (air_at(air.va0, air.av, rn + i, kt) * bl -
                               air_at(air.va0, air.av, rn + opp, kt) * (1 - bl));
            }
            if (j < nth - 1) {
                vAirS = 0.5 * (air_at(air.va0, air.av, row + i, kt) * bl -
                               air_at(air.va0, air.av, row + opp, kt) * (1 - bl));
            }
            div += B.invh * (uAirE / (etaRight / Cr * invDt + 1.0) -
                             uAirW / (etaLeft / Cr * invDt + 1.0) +
                             vAirS * sinS / (etaDown / Cr * invDt + 1.0) -
                             vAirN * sinN / (etaUp / Cr * invDt + 1.0));
        }
        if (B.gravity) {
            div += B.gs * B.invh *
                   (sinS * sinS * etaDown / (invDt * etaDown + Cr) -
                    sinN * sinN * etaUp / (invDt * etaUp + Cr));
        }
        r.rhs = sinTheta * invDt - div;
        if (bN) r.rhs += wN * gmax;
        if (bS) r.rhs += wS * gmax;
        if (bW) r.rhs += wW * gmax;
        if (bE) r.rhs += wE * gmax;
    }
    r.wN = (r.brk || bN) ? 0.0 : wN;
    r.wS = (r.brk || bS) ? 0.0 : wS;
    r.wW = (r.brk || bW) ? 0.0 : wW;
    r.wE = (r.brk || bE) ? 0.0 : wE;
    // x0 = Gamma* after the broken-cell override (bubble.cu:1325, 1480-1486)
    r.x0 = r.brk ? gmax : gamma;
    r.x0N = bN ? gmax : gN;
    r.x0S = bS ? gmax : gS;
    r.x0W = bW ? gmax : gW;
    r.x0E = bE ? gmax : gE;
    return r;
}

// Row-block assembly: coefficients and rhs of one theta row, the initial
// FP64 residual r0 = b - A Gamma*, r_in = (Ti) r0 and the first line solve
// z = M^-1 r_in, whose reduction depth is recorded for the later line solves
// of the same matrix. Partials as k_resid. The block of row 0 resets the
// bubble's solver control block; the broken-cell override of Gamma* is
// written back to the state as the reference does.
template <typename T, typename Ti, int CPT>
__global__ void k_assemble(const int2* __restrict__ rows, const BubbleDev* __restrict__ bub,
                           const BubbleTime* __restrict__ btime, Tables tab, StatePtrs<T> s,
                           AirPtrs<T> air, SysPtrs<Ti> sys, T gmax_state,
                           SolveCtl* __restrict__ ctl, double* __restrict__ part, long stride) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double red[16 * 4];
    __shared__ float redf[16];
    const int2 rw = rows[blockIdx.x];
    if (rw.x < 0) return;
    const BubbleDev B = bub[rw.x];
    const int j = rw.y, nph = B.nph;
    const double t = btime[rw.x].t;
    Ti* sa = reinterpret_cast<Ti*>(lds_raw);
    Ti* sc = sa + nph;
    Ti* sd = sc + nph;
    const size_t row = (size_t)B.off + (size_t)j * nph;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    double mu = 0.0;
    for (int i = threadIdx.x; i < nph; i += blockDim.x) {
        const CellSys q = assemble_cell<T>(B, t, tab, s, air, j, i);
        // x^T N x <= sum (wN + wS) x_i^2 and x^T M x >= sum (diag - wW - wE) x_i^2
        // for the theta couplings N and the phi-line blocks M, so the
        // eigenvalues of M^-1 A lie in [1 - mu, 1 + mu] with mu the largest
        // ratio below
        if (!q.brk) mu = fmax(mu, (q.wN + q.wS) / (q.diag - q.wW - q.wE));
        const size_t c = row + i;
        const double ax = q.diag * q.x0 - q.wN * q.x0N - q.wS * q.x0S - q.wW * q.x0W -
                          q.wE * q.x0E;
        const double r0 = q.rhs - ax;
        const Ti ri = (Ti)r0;
        sys.wN64[c] = q.wN;
        sys.wW64[c] = q.wW;
        sys.diag64[c] = q.diag;
        sys.b64[c] = q.rhs;
        sys.x64[c] = q.x0;
        sys.wN[c] = (Ti)q.wN;
        sys.wW[c] = (Ti)q.wW;
        sys.diag[c] = (Ti)q.diag;
        sys.r[c] = ri;
        const Ti inv = pre_rcp((Ti)q.diag);
        sa[i] = -(Ti)q.wW * inv;
        sc[i] = -(Ti)q.wE * inv;
        sd[i] = ri * inv;
        acc[0] += r0 * r0;
        acc[1] += q.rhs * q.rhs;
        acc[3] += (double)ri * (double)ri;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nph; i += blockDim.x) {
        const size_t c = row + i;
        if (is_broken((double)s.gam[c], (double)s.eta[c], B.gamma_max)) s.gam[c] = gmax_state;
    }
    const int steps = pcr_periodic<Ti, CPT>(sa, sc, sd, nph, -1, redf);
    for (int i = threadIdx.x; i < nph; i += blockDim.x) {
        const Ti z = sd[i];
        sys.z[row + i] = z;
        acc[2] += (double)sys.r[row + i] * (double)z;
    }
    block_sum<4>(acc, red);
    mu = block_max(mu, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) part[q * stride + B.pr0 + j] = acc[q];
        part[4 * stride + B.pr0 + j] = mu;
        sys.rsteps[B.pr0 + j] = steps;
        if (j == 0) {
            SolveCtl z = {};
            z.xbuf = 0;
            z.outers = 1;
            ctl[rw.x] = z;
        }
    }
}

// ----------------------------------------------------- FP64 residual ----
// x_new = x_old + e (the inner correction), r = b - A x_new in FP64,
// r_in = (Ti) r, z = M^-1 r_in. Partials: |r|^2, |b|^2, r_in.z, |r_in|^2.
template <typename Ti, int CPT>
__global__ void k_resid(const int2* __restrict__ rows, const BubbleDev* __restrict__ bub,
                        SolveCtl* __restrict__ ctl, SysPtrs<Ti> sys, const double* __restrict__ xo,
                        double* __restrict__ xn, int k, double* __restrict__ part, long stride) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double red[16 * 4];
    __shared__ float redf[16];
    const int2 rw = rows[blockIdx.x];
    if (rw.x < 0 || ctl[rw.x].outer_done) return;
    const BubbleDev B = bub[rw.x];
    const int j = rw.y, nph = B.nph, m = B.mask, nth = B.nth;
    const int steps = sys.rsteps[B.pr0 + j];
    Ti* sa = reinterpret_cast<Ti*>(lds_raw);
    Ti* sc = sa + nph;
    Ti* sd = sc + nph;
    const size_t row = (size_t)B.off + (size_t)j * nph;
    const Ti* __restrict__ e = sys.e;
    const bool hasN = j > 0, hasS = j < nth - 1;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    Ti rin[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int i = threadIdx.x + q * blockDim.x;
        rin[q] = Ti(0);
        if (i < nph) {
            const size_t c = row + i;
            const size_t cW = row + ((i - 1) & m), cE = row + ((i + 1) & m);
            const double xc = xo[c] + (double)e[c];
            double ax = sys.diag64[c] * xc - sys.wW64[c] * (xo[cW] + (double)e[cW]) -
                        sys.wW64[cE] * (xo[cE] + (double)e[cE]);
            if (hasN) ax -= sys.wN64[c] * (xo[c - nph] + (double)e[c - nph]);
            if (hasS) ax -= sys.wN64[c + nph] * (xo[c + nph] + (double)e[c + nph]);
            xn[c] = xc;
            const double b = sys.b64[c];
            const double r = b - ax;
            const Ti ri = (Ti)r;
            rin[q] = ri;
            sys.r[c] = ri;
            const Ti inv = pre_rcp(sys.diag[c]);
            sa[i] = -sys.wW[c] * inv;
            sc[i] = -sys.wW[cE] * inv;
            sd[i] = ri * inv;
            acc[0] += r * r;
            acc[1] += b * b;
            acc[3] += (double)ri * (double)ri;
        }
    }
    __syncthreads();
    pcr_periodic<Ti, CPT>(sa, sc, sd, nph, steps, redf);
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int i = threadIdx.x + q * blockDim.x;
        if (i < nph) {
            const Ti z = sd[i];
            sys.z[row + i] = z;
            acc[2] += (double)rin[q] * (double)z;
        }
    }
    block_sum<4>(acc, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) part[q * stride + B.pr0 + j] = acc[q];
        if (j == 0) {
            ctl[rw.x].xbuf = k & 1;
            ctl[rw.x].outers = k + 1;
        }
    }
}

// ------------------------------------------------------ inner PCG: A p ----
// p = z + beta p_old (beta from the r.z partials of the previous step),
// Ap, partial p.Ap. Iteration 0 of a refinement also decides whether the
// FP64 residual already meets the tolerance.
template <typename Ti, int CPT>
__global__ void k_spmv(const int2* __restrict__ rows, const BubbleDev* __restrict__ bub,
                       SolveCtl* __restrict__ ctl, SysPtrs<Ti> sys, const Ti* __restrict__ pold,
                       Ti* __restrict__ pnew, int it, SolveParams sp, const double* __restrict__ partR,
                       const double* __restrict__ partB, double* __restrict__ partA, long stride) {
    __shared__ double red[16 * 4];
    const int2 rw = rows[blockIdx.x];
    if (rw.x < 0) return;
    SolveCtl* C = ctl + rw.x;
    if (C->outer_done || (it > 0 && C->inner_done)) return;
    const BubbleDev B = bub[rw.x];
    const int j = rw.y, nph = B.nph, m = B.mask, nth = B.nth;
    const bool leader = (j == 0 && threadIdx.x == 0);
    const bool hasN = j > 0, hasS = j < nth - 1;
    const size_t row = (size_t)B.off + (size_t)j * nph;
    const Ti* __restrict__ z = sys.z;
    double beta = 0.0;
    if (it == 0) {
        double sm[4];
        sum_partials<4>(partR, stride, B.pr0, nth, sm, red);
        const double rr = sm[0], bb = sm[1], rz = sm[2], rrin = sm[3];
        const double res = sp.absolute ? sqrt(rr) : sqrt(rr) / (bb > 0.0 ? sqrt(bb) : 1.0);
        if (res <= sp.tol) {
            if (leader) { C->outer_done = 1; C->rel_res = res; }
            return;
        }
        if (leader) {
            C->rr_in0 = rrin;
            C->inner_done = 0;
            C->rz[0] = rz;
            C->rel_res = res;
            C->bb = bb;
        }
    } else {
        double sm[2];
        sum_partials<2>(partB, stride, B.pr0, nth, sm, red);
        const double rz = sm[0], rr = sm[1];
        const double tgt = sp.inner_frac * sp.tol;
        const double bbe = sp.absolute ? 1.0 : C->bb;
        if (rr <= sp.inner_rho2 * C->rr_in0 || rr <= tgt * tgt * bbe) {
            if (leader) C->inner_done = 1;
            return;
        }
        beta = rz / C->rz[(it - 1) & 1];
        if (leader) C->rz[it & 1] = rz;
    }
    // operands (the decision above comes first, so a converged solve reads none)
    Ti zc[CPT], zw[CPT], ze[CPT], zn[CPT], zs[CPT];
    Ti pc[CPT], pw[CPT], pe[CPT], pn[CPT], ps[CPT];
    Ti dg[CPT], wn[CPT], wsv[CPT], ww[CPT], we[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int i = threadIdx.x + q * blockDim.x;
        zc[q] = zw[q] = ze[q] = zn[q] = zs[q] = Ti(0);
        pc[q] = pw[q] = pe[q] = pn[q] = ps[q] = Ti(0);
        dg[q] = wn[q] = wsv[q] = ww[q] = we[q] = Ti(0);
        if (i < nph) {
            const size_t c = row + i, cW = row + ((i - 1) & m), cE = row + ((i + 1) & m);
            zc[q] = z[c]; zw[q] = z[cW]; ze[q] = z[cE];
            dg[q] = sys.diag[c]; ww[q] = sys.wW[c]; we[q] = sys.wW[cE];
            if (hasN) { zn[q] = z[c - nph]; wn[q] = sys.wN[c]; }
            if (hasS) { zs[q] = z[c + nph]; wsv[q] = sys.wN[c + nph]; }
            if (it > 0) {
                pc[q] = pold[c]; pw[q] = pold[cW]; pe[q] = pold[cE];
                if (hasN) pn[q] = pold[c - nph];
                if (hasS) ps[q] = pold[c + nph];
            }
        }
    }
    const Ti bt = (Ti)beta;
    double acc[1] = {0.0};
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int i = threadIdx.x + q * blockDim.x;
        if (i < nph) {
            const Ti p0 = zc[q] + bt * pc[q];
            const Ti ap = dg[q] * p0 - ww[q] * (zw[q] + bt * pw[q]) - we[q] * (ze[q] + bt * pe[q]) -
                          wn[q] * (zn[q] + bt * pn[q]) - wsv[q] * (zs[q] + bt * ps[q]);
            pnew[row + i] = p0;
            sys.Ap[row + i] = ap;
            acc[0] += (double)p0 * (double)ap;
        }
    }
    block_sum<1>(acc, red);
    if (threadIdx.x == 0) partA[B.pr0 + j] = acc[0];
}

// -------------------------------------------- inner PCG: update + M^-1 ----
template <typename Ti, int CPT>
__global__ void k_cgstep(const int2* __restrict__ rows, const BubbleDev* __restrict__ bub,
                         SolveCtl* __restrict__ ctl, SysPtrs<Ti> sys, const Ti* __restrict__ p,
                         int it, const double* __restrict__ partA, double* __restrict__ partB,
                         long stride) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double red[16 * 4];
    __shared__ float redf[16];
    const int2 rw = rows[blockIdx.x];
    if (rw.x < 0) return;
    SolveCtl* C = ctl + rw.x;
    if (C->outer_done || C->inner_done) return;
    const BubbleDev B = bub[rw.x];
    const int j = rw.y, nph = B.nph, m = B.mask, nth = B.nth;
    const int steps = sys.rsteps[B.pr0 + j];
    Ti* sa = reinterpret_cast<Ti*>(lds_raw);
    Ti* sc = sa + nph;
    Ti* sd = sc + nph;
    const size_t row = (size_t)B.off + (size_t)j * nph;
    Ti pv[CPT], apv[CPT], ev[CPT], rv[CPT], ww[CPT], we[CPT], dg[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int i = threadIdx.x + q * blockDim.x;
        pv[q] = apv[q] = ev[q] = rv[q] = ww[q] = we[q] = Ti(0);
        dg[q] = Ti(1);
        if (i < nph) {
            const size_t c = row + i;
            pv[q] = p[c];
            apv[q] = sys.Ap[c];
            if (it > 0) ev[q] = sys.e[c];
            rv[q] = sys.r[c];
            ww[q] = sys.wW[c];
            we[q] = sys.wW[row + ((i + 1) & m)];
            dg[q] = sys.diag[c];
        }
    }
    double sm[1];
    sum_partials<1>(partA, stride, B.pr0, nth, sm, red);
    const Ti al = (Ti)(C->rz[it & 1] / sm[0]);
    double acc[2] = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int i = threadIdx.x + q * blockDim.x;
        if (i < nph) {
            const size_t c = row + i;
            sys.e[c] = ev[q] + al * pv[q];
            const Ti r = rv[q] - al * apv[q];
            rv[q] = r;
            sys.r[c] = r;
            const Ti inv = pre_rcp(dg[q]);
            sa[i] = -ww[q] * inv;
            sc[i] = -we[q] * inv;
            sd[i] = r * inv;
            acc[1] += (double)r * (double)r;
        }
    }
    __syncthreads();
    pcr_periodic<Ti, CPT>(sa, sc, sd, nph, steps, redf);
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int i = threadIdx.x + q * blockDim.x;
        if (i < nph) {
            const Ti z = sd[i];
            sys.z[row + i] = z;
            acc[0] += (double)rv[q] * (double)z;
        }
    }
    block_sum<2>(acc, red);
    if (threadIdx.x == 0) {
        partB[B.pr0 + j] = acc[0];
        partB[stride + B.pr0 + j] = acc[1];
        if (j == 0) C->iters += 1;
    }
}

// ------------------------------------------- inner Chebyshev iteration ----
// Chebyshev acceleration of the phi-line Jacobi iteration for A e = r_in
// (Saad, Iterative Methods for Sparse Linear Systems, Alg. 12.1), with the
// spectrum of M^-1 A inside [1 - mu, 1 + mu] (mu from the assembly, widened
// slightly). No inner products: one kernel per iteration. Iteration k >= 1
// reads e_k (k = 1: e_1 = d_0 = z_0 = M^-1 r_in) and d_{k-1}, forms
// r_k = r_in - A e_k, z_k = M^-1 r_k, rho_k = 1 / (2 sigma - rho_{k-1}),
// d_k = rho_k rho_{k-1} d_{k-1} + (2 rho_k / delta) z_k, e_{k+1} = e_k + d_k.
// Partial |r_k|^2 goes to partB slot 1 (calibration and tracing).
template <typename Ti, int CPT>
__global__ void k_cheb(const int2* __restrict__ rows, const BubbleDev* __restrict__ bub,
                       SolveCtl* __restrict__ ctl, SysPtrs<Ti> sys, const Ti* __restrict__ ek,
                       Ti* __restrict__ enext, Ti* __restrict__ dv, int it, SolveParams sp,
                       const double* __restrict__ partR, double* __restrict__ partB, long stride) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double red[16 * 4];
    __shared__ float redf[16];
    const int2 rw = rows[blockIdx.x];
    if (rw.x < 0) return;
    SolveCtl* C = ctl + rw.x;
    if (C->outer_done) return;
    const BubbleDev B = bub[rw.x];
    const int j = rw.y, nph = B.nph, m = B.mask, nth = B.nth;
    const bool leader = (j == 0 && threadIdx.x == 0);
    double mu;
    if (it == 1) {
        double sm[2];
        sum_partials<2>(partR, stride, B.pr0, nth, sm, red);
        const double res = sp.absolute ? sqrt(sm[0]) : sqrt(sm[0]) / (sm[1] > 0.0 ? sqrt(sm[1]) : 1.0);
        if (res <= sp.tol) {
            if (leader) { C->outer_done = 1; C->rel_res = res; }
            return;
        }
        mu = max_partials(partR + 4 * stride, B.pr0, nth, red);
        if (leader) {
            C->mu = mu;
            C->rel_res = res;
            C->bb = sm[1];
        }
    } else {
        mu = C->mu;
    }
    const double delta = fmax(mu * 1.02 + 1e-3, 1e-4);
    const double sigma = 1.0 / delta;
    double rho_prev = delta, rho = delta;      // rho_0 = 1 / sigma
    for (int q = 1; q <= it; ++q) {
        rho_prev = rho;
        rho = 1.0 / (2.0 * sigma - rho_prev);
    }
    const Ti cd = (Ti)(rho * rho_prev), cz = (Ti)(2.0 * rho / delta);
    const int steps = sys.rsteps[B.pr0 + j];
    Ti* sa = reinterpret_cast<Ti*>(lds_raw);
    Ti* sc = sa + nph;
    Ti* sd = sc + nph;
    const size_t row = (size_t)B.off + (size_t)j * nph;
    const bool hasN = j > 0, hasS = j < nth - 1;
    const Ti* __restrict__ e = it == 1 ? sys.z : ek;
    const Ti* __restrict__ dprev = it == 1 ? sys.z : dv;
    Ti ev[CPT], dvv[CPT];
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int i = threadIdx.x + q * blockDim.x;
        ev[q] = dvv[q] = Ti(0);
        if (i < nph) {
            const size_t c = row + i, cW = row + ((i - 1) & m), cE = row + ((i + 1) & m);
            const Ti ec = e[c];
            Ti ax = sys.diag[c] * ec - sys.wW[c] * e[cW] - sys.wW[cE] * e[cE];
            if (hasN) ax -= sys.wN[c] * e[c - nph];
            if (hasS) ax -= sys.wN[c + nph] * e[c + nph];
            const Ti r = sys.r[c] - ax;
            ev[q] = ec;
            dvv[q] = dprev[c];
            const Ti inv = pre_rcp(sys.diag[c]);
            sa[i] = -sys.wW[c] * inv;
            sc[i] = -sys.wW[cE] * inv;
            sd[i] = r * inv;
            acc += (double)r * (double)r;
        }
    }
    __syncthreads();
    pcr_periodic<Ti, CPT>(sa, sc, sd, nph, steps, redf);
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int i = threadIdx.x + q * blockDim.x;
        if (i < nph) {
            const Ti dn = cd * dvv[q] + cz * sd[i];
            dv[row + i] = dn;
            enext[row + i] = ev[q] + dn;
        }
    }
    double a1[1] = {acc};
    block_sum<1>(a1, red);
    if (threadIdx.x == 0) {
        partB[stride + B.pr0 + j] = a1[0];
        if (j == 0) C->iters += 1;
    }
}

// ----------------------------------------------------------- finalize ----
// The verdict from the last FP64 residual evaluated (block-wide; every
// thread of the block must call it).
__device__ __forceinline__ void finalize_block(const BubbleDev& B, SolveCtl* __restrict__ C,
                                               SolveHist* __restrict__ Hh, const SolveParams& sp,
                                               const double* __restrict__ partR, long stride,
                                               double* red) {
    double s[2];
    sum_partials<2>(partR, stride, B.pr0, B.nth, s, red);
    if (threadIdx.x == 0) {
        const double res = sp.absolute ? sqrt(s[0]) : sqrt(s[0]) / (s[1] > 0.0 ? sqrt(s[1]) : 1.0);
        C->rel_res = res;
        C->bb = s[1];
        C->converged = res <= sp.tol ? 1 : 0;
        Hh->solves += 1;
        if (!(res <= sp.tol)) Hh->fails += 1;
        if (!(res <= Hh->max_res)) Hh->max_res = res;
    }
}

__global__ void k_finalize(const BubbleDev* __restrict__ bub, SolveCtl* __restrict__ ctl,
                           SolveHist* __restrict__ hist, SolveParams sp,
                           const double* __restrict__ partR, long stride) {
    __shared__ double red[16 * 4];
    const BubbleDev B = bub[blockIdx.x];
    finalize_block(B, ctl + blockIdx.x, hist + blockIdx.x, sp, partR, stride, red);
}

// ------------------------------------------------------------- update ----
__device__ __forceinline__ double clamp_gamma(double g, double gmax) {
    return (g <= 0.0 || g > gmax - kEps) ? gmax : g;
}

template <typename T>
__global__ __launch_bounds__(256) void k_update(const int4* __restrict__ tiles,
                                                const BubbleDev* __restrict__ bub,
                                                BubbleTime* __restrict__ btime,
                                                SolveCtl* __restrict__ ctl, Tables tab,
                                                StatePtrs<T> st, AirPtrs<T> air,
                                                const double* __restrict__ x0,
                                                const double* __restrict__ x1, StatePtrs<T> out,
                                                int fin, SolveHist* __restrict__ hist,
                                                SolveParams sp, const double* __restrict__ partR,
                                                long stride) {
    __shared__ double red[16 * 4];
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0) return;
    const BubbleDev B = bub[tl.x];
    if (tl.y == 0 && tl.z == 0) {
        // first tile of the bubble: solve verdict, and the time advance that
        // the next step's advection applies
        if (fin) finalize_block(B, ctl + tl.x, hist + tl.x, sp, partR, stride, red);
        if (threadIdx.x == 0 && btime[tl.x].advance) btime[tl.x].pending = 1;
    }
    const int j = tl.y + (int)threadIdx.x / tl.w;
    const int i = tl.z + (int)threadIdx.x % tl.w;
    if (j >= B.nth || i >= B.nph) return;
    const int nph = B.nph, m = B.mask, nth = B.nth;
    const double* __restrict__ x = ctl[tl.x].xbuf ? x1 : x0;
    const size_t c = (size_t)B.off + (size_t)j * nph + i;
    const size_t cW = (size_t)B.off + (size_t)j * nph + ((i - 1) & m);
    const double gmax = B.gamma_max, dt = B.dt, invh = B.invh, Cr = B.Cr, M = B.M;
    double kt = 0.0;
    if (B.airflow) { const double t = btime[tl.x].t; kt = 3.0 - fabs(t - 3.0); }

    // divergenceKernel_fromGamma (bubble.cu:1276-1295)
    const double ga = (double)st.gam[c];
    const double g = x[c];
    double gnew, div;
    if (g <= 0.0 || g > gmax - kEps) {
        gnew = gmax;
        div = 1.0 / dt;
    } else {
        gnew = g;
        div = (1.0 - fmin(g, 2 * ga) / ga) / dt;
    }
    const double eta = (double)st.eta[c];
    // applyforceThickness (bubble.cu:2041-2063)
    {
        double v = eta * (1 - dt * div);
        if (B.evaporation) v += B.eva * dt;
        out.eta[c] = (T)fmax(kThresh, v);
        out.gam[c] = (T)gnew;
    }
    // applyforcevelphiKernel (bubble.cu:1879-1935)
    {
        const double etaW = (double)st.eta[cW];
        const double u1 = (double)st.uph[c];
        const double invEta = 2. / (etaW + eta);
        double res;
        if (__builtin_isinf(invEta)) {
            res = __builtin_nan("");
        } else {
            const double gW = clamp_gamma(x[cW], gmax);
            const double pGpx = invh * (gnew - gW) * tab.cscTc[B.rt0 + j];
            const double f1 = -M * invEta * pGpx;
            const double uAir = B.airflow ? air_at(air.ua0, air.au, c, kt) : 0.0;
            const double f2 = Cr * invEta * uAir;
            res = (u1 / dt + (f1 + f2 + 0.0 + 0.0)) / (1. / dt + Cr * invEta);
        }
        out.uph[c] = (T)res;
    }
    // applyforcevelthetaKernel (bubble.cu:1819-1876)
    if (j < nth - 1) {
        const double etaS = (double)st.eta[c + nph];
        const double v1 = (double)st.uth[c];
        const double invEta = 2. / (eta + etaS);
        double res;
        if (__builtin_isinf(invEta)) {
            res = __builtin_nan("");
        } else {
            const double gS = clamp_gamma(x[c + nph], gmax);
            const double pGpy = invh * (gS - gnew);
            const double f1 = -M * invEta * pGpy;
            const double vAir = B.airflow ? air_at(air.va0, air.av, c, kt) : 0.0;
            const double f2 = Cr * invEta * vAir;
            const double f3 = B.gravity ? B.gs * tab.sinTf[B.rt0 + j + 1] : 0.0;
            res = (v1 / dt + f1 + f2 + f3 + 0.0) / (1. / dt + Cr * invEta);
        }
        out.uth[c] = (T)res;
    }
}

// Per-row film volume (sum eta sin(theta)) and count of cells with eta > 0.
template <typename T>
__global__ void k_row_stats(const int2* __restrict__ rows, const BubbleDev* __restrict__ bub,
                            Tables tab, const T* __restrict__ eta, double* __restrict__ out,
                            long stride) {
    __shared__ double red[16 * 4];
    const int2 rw = rows[blockIdx.x];
    if (rw.x < 0) return;
    const BubbleDev B = bub[rw.x];
    const size_t row = (size_t)B.off + (size_t)rw.y * B.nph;
    double acc[2] = {0.0, 0.0};
    for (int i = threadIdx.x; i < B.nph; i += blockDim.x) {
        const double e = (double)eta[row + i];
        acc[0] += e;
        acc[1] += e > 0.0 ? 1.0 : 0.0;
    }
    block_sum<2>(acc, red);
    if (threadIdx.x == 0) {
        out[B.pr0 + rw.y] = acc[0] * tab.sinTc[B.rt0 + rw.y];
        out[stride + B.pr0 + rw.y] = acc[1];
    }
}

// ------------------------------------------------------------ launchers ----
template <typename T>
void launch_advect(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, BubbleTime* bt,
                   const Tables& tab, StatePtrs<T> in, StatePtrs<T> out) {
    hipLaunchKernelGGL(k_advect<T>, dim3(g.ntiles), dim3(256), 0, s, g.tiles, bub, bt, tab, in,
                       out);
}

template <typename T, typename Ti, int CPT>
static void assemble_cpt(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub,
                         const BubbleTime* bt, const Tables& tab, StatePtrs<T> st, AirPtrs<T> air,
                         SysPtrs<Ti> sys, double gmax_state, SolveCtl* ctl, double* part,
                         long stride) {
    hipLaunchKernelGGL((k_assemble<T, Ti, CPT>), dim3(g.nrows), dim3(g.row_threads), g.lds_bytes,
                       s, g.rows, bub, bt, tab, st, air, sys, (T)gmax_state, ctl, part, stride);
}
template <typename T, typename Ti>
void launch_assemble(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub,
                     const BubbleTime* bt, const Tables& tab, StatePtrs<T> st, AirPtrs<T> air,
                     SysPtrs<Ti> sys, double gmax_state, SolveCtl* ctl, double* part, long stride) {
    switch (g.cpt) {
        case 1: assemble_cpt<T, Ti, 1>(s, g, bub, bt, tab, st, air, sys, gmax_state, ctl, part, stride); break;
        case 2: assemble_cpt<T, Ti, 2>(s, g, bub, bt, tab, st, air, sys, gmax_state, ctl, part, stride); break;
        case 4: assemble_cpt<T, Ti, 4>(s, g, bub, bt, tab, st, air, sys, gmax_state, ctl, part, stride); break;
        default: assemble_cpt<T, Ti, 8>(s, g, bub, bt, tab, st, air, sys, gmax_state, ctl, part, stride); break;
    }
}

template <typename Ti, int CPT>
static void resid_cpt(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                      SysPtrs<Ti> sys, const double* xo, double* xn, int k, double* part,
                      long stride) {
    hipLaunchKernelGGL((k_resid<Ti, CPT>), dim3(g.nrows), dim3(g.row_threads), g.lds_bytes, s,
                       g.rows, bub, ctl, sys, xo, xn, k, part, stride);
}
template <typename Ti>
void launch_resid(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                  SysPtrs<Ti> sys, const double* xo, double* xn, int k, double* part,
                  long stride) {
    switch (g.cpt) {
        case 1: resid_cpt<Ti, 1>(s, g, bub, ctl, sys, xo, xn, k, part, stride); break;
        case 2: resid_cpt<Ti, 2>(s, g, bub, ctl, sys, xo, xn, k, part, stride); break;
        case 4: resid_cpt<Ti, 4>(s, g, bub, ctl, sys, xo, xn, k, part, stride); break;
        default: resid_cpt<Ti, 8>(s, g, bub, ctl, sys, xo, xn, k, part, stride); break;
    }
}

template <typename Ti, int CPT>
static void spmv_cpt(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                     SysPtrs<Ti> sys, const Ti* pold, Ti* pnew, int it, const SolveParams& sp,
                     const double* partR, const double* partB, double* partA, long stride) {
    hipLaunchKernelGGL((k_spmv<Ti, CPT>), dim3(g.nrows), dim3(g.row_threads), 0, s, g.rows, bub,
                       ctl, sys, pold, pnew, it, sp, partR, partB, partA, stride);
}
template <typename Ti>
void launch_spmv(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                 SysPtrs<Ti> sys, const Ti* pold, Ti* pnew, int it, const SolveParams& sp,
                 const double* partR, const double* partB, double* partA, long stride) {
    switch (g.cpt) {
        case 1: spmv_cpt<Ti, 1>(s, g, bub, ctl, sys, pold, pnew, it, sp, partR, partB, partA, stride); break;
        case 2: spmv_cpt<Ti, 2>(s, g, bub, ctl, sys, pold, pnew, it, sp, partR, partB, partA, stride); break;
        case 4: spmv_cpt<Ti, 4>(s, g, bub, ctl, sys, pold, pnew, it, sp, partR, partB, partA, stride); break;
        default: spmv_cpt<Ti, 8>(s, g, bub, ctl, sys, pold, pnew, it, sp, partR, partB, partA, stride); break;
    }
}

template <typename Ti, int CPT>
static void cgstep_cpt(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                       SysPtrs<Ti> sys, const Ti* p, int it, const double* partA, double* partB,
                       long stride) {
    hipLaunchKernelGGL((k_cgstep<Ti, CPT>), dim3(g.nrows), dim3(g.row_threads), g.lds_bytes, s,
                       g.rows, bub, ctl, sys, p, it, partA, partB, stride);
}
template <typename Ti>
void launch_cgstep(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                   SysPtrs<Ti> sys, const Ti* p, int it, const double* partA, double* partB,
                   long stride) {
    switch (g.cpt) {
        case 1: cgstep_cpt<Ti, 1>(s, g, bub, ctl, sys, p, it, partA, partB, stride); break;
        case 2: cgstep_cpt<Ti, 2>(s, g, bub, ctl, sys, p, it, partA, partB, stride); break;
        case 4: cgstep_cpt<Ti, 4>(s, g, bub, ctl, sys, p, it, partA, partB, stride); break;
        default: cgstep_cpt<Ti, 8>(s, g, bub, ctl, sys, p, it, partA, partB, stride); break;
    }
}

template <typename Ti, int CPT>
static void cheb_cpt(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                     SysPtrs<Ti> sys, const Ti* ek, Ti* enext, Ti* dv, int it, const SolveParams& sp,
                     const double* partR, double* partB, long stride) {
    hipLaunchKernelGGL((k_cheb<Ti, CPT>), dim3(g.nrows), dim3(g.row_threads), g.lds_bytes, s,
                       g.rows, bub, ctl, sys, ek, enext, dv, it, sp, partR, partB, stride);
}
template <typename Ti>
void launch_cheb(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                 SysPtrs<Ti> sys, const Ti* ek, Ti* enext, Ti* dv, int it, const SolveParams& sp,
                 const double* partR, double* partB, long stride) {
    switch (g.cpt) {
        case 1: cheb_cpt<Ti, 1>(s, g, bub, ctl, sys, ek, enext, dv, it, sp, partR, partB, stride); break;
        case 2: cheb_cpt<Ti, 2>(s, g, bub, ctl, sys, ek, enext, dv, it, sp, partR, partB, stride); break;
        case 4: cheb_cpt<Ti, 4>(s, g, bub, ctl, sys, ek, enext, dv, it, sp, partR, partB, stride); break;
        default: cheb_cpt<Ti, 8>(s, g, bub, ctl, sys, ek, enext, dv, it, sp, partR, partB, stride); break;
    }
}

void launch_finalize(hipStream_t s, int nb, const BubbleDev* bub, SolveCtl* ctl, SolveHist* hist,
                     const SolveParams& sp, const double* partR, long stride) {
    hipLaunchKernelGGL(k_finalize, dim3(nb), dim3(256), 0, s, bub, ctl, hist, sp, partR, stride);
}

template <typename T>
void launch_update(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, BubbleTime* bt,
                   SolveCtl* ctl, const Tables& tab, StatePtrs<T> st, AirPtrs<T> air,
                   const double* x0, const double* x1, StatePtrs<T> out, int fin, SolveHist* hist,
                   const SolveParams& sp, const double* partR, long stride) {
    hipLaunchKernelGGL(k_update<T>, dim3(g.ntiles), dim3(256), 0, s, g.tiles, bub, bt, ctl, tab, st,
                       air, x0, x1, out, fin, hist, sp, partR, stride);
}

template <typename T>
void launch_row_stats(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, const Tables& tab,
                      const T* eta, double* out, long stride) {
    hipLaunchKernelGGL(k_row_stats<T>, dim3(g.nrows), dim3(256), 0, s, g.rows, bub, tab, eta, out,
                       stride);
}

// explicit instantiations
template void launch_advect<float>(hipStream_t, const LaunchGeom&, const BubbleDev*, BubbleTime*,
                                   const Tables&, StatePtrs<float>, StatePtrs<float>);
template void launch_advect<double>(hipStream_t, const LaunchGeom&, const BubbleDev*, BubbleTime*,
                                    const Tables&, StatePtrs<double>, StatePtrs<double>);
#define SF_INST_SYS(T, Ti)                                                                        \
    template void launch_assemble<T, Ti>(hipStream_t, const LaunchGeom&, const BubbleDev*,        \
                                         const BubbleTime*, const Tables&, StatePtrs<T>,          \
                                         AirPtrs<T>, SysPtrs<Ti>, double, SolveCtl*, double*,     \
                                         long);
SF_INST_SYS(float, float)
SF_INST_SYS(float, double)
SF_INST_SYS(double, double)
#define SF_INST_TI(Ti)                                                                            \
    template void launch_resid<Ti>(hipStream_t, const LaunchGeom&, const BubbleDev*, SolveCtl*,   \
                                   SysPtrs<Ti>, const double*, double*, int, double*, long);      \
    template void launch_spmv<Ti>(hipStream_t, const LaunchGeom&, const BubbleDev*, SolveCtl*,    \
                                  SysPtrs<Ti>, const Ti*, Ti*, int, const SolveParams&,           \
                                  const double*, const double*, double*, long);                   \
    template void launch_cgstep<Ti>(hipStream_t, const LaunchGeom&, const BubbleDev*, SolveCtl*,  \
                                    SysPtrs<Ti>, const Ti*, int, const double*, double*, long);   \
    template void launch_cheb<Ti>(hipStream_t, const LaunchGeom&, const BubbleDev*, SolveCtl*,    \
                                  SysPtrs<Ti>, const Ti*, Ti*, Ti*, int, const SolveParams&,      \
                                  const double*, double*, long);
SF_INST_TI(float)
SF_INST_TI(double)
template void launch_update<float>(hipStream_t, const LaunchGeom&, const BubbleDev*, BubbleTime*,
                                   SolveCtl*, const Tables&, StatePtrs<float>, AirPtrs<float>,
                                   const double*, const double*, StatePtrs<float>, int,
                                   SolveHist*, const SolveParams&, const double*, long);
template void launch_update<double>(hipStream_t, const LaunchGeom&, const BubbleDev*, BubbleTime*,
                                    SolveCtl*, const Tables&, StatePtrs<double>, AirPtrs<double>,
                                    const double*, const double*, StatePtrs<double>, int,
                                    SolveHist*, const SolveParams&, const double*, long);
template void launch_row_stats<float>(hipStream_t, const LaunchGeom&, const BubbleDev*,
                                      const Tables&, const float*, double*, long);
template void launch_row_stats<double>(hipStream_t, const LaunchGeom&, const BubbleDev*,
                                       const Tables&, const double*, double*, long);

}  // namespace sfk
