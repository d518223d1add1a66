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
// One cell: eta and Gamma at the centre (j+1/2, i), u_phi at the west face
// (j+1/2, i-1/2), u_theta at the south face (j+1, i). NANS selects the
// reference's NaN-skipping interpolation; without it the fields hold no NaN.
// P2: every bubble of the launch has rows whose length is a power of two.
template <typename T, bool NANS, bool P2>
__device__ __forceinline__ void advect_cell(const BubbleDev& B, const Tables& tab,
                                            const StatePtrs<T>& in, const StatePtrs<T>& out, int j,
                                            int i) {
    const size_t c = (size_t)j * B.nph + i;
    VelCtx<T, P2> v;
    v.uth = Field<T, P2>{in.uth + B.off, B.nph, B.mask, B.nth - 2};
    v.uph = Field<T, P2>{in.uph + B.off, B.nph, B.mask, B.nth - 1};
    v.nth = B.nth;
    v.invh = (T)B.invh;
    v.dt = (T)B.dt;
    const Field<T, P2> feta{in.eta + B.off, B.nph, B.mask, B.nth - 1};
    const Field<T, P2> fgam{in.gam + B.off, B.nph, B.mask, B.nth - 1};

    const T sinTc = TabSel<T>::sTc(tab)[B.rt0 + j], cosTc = TabSel<T>::cTc(tab)[B.rt0 + j];
    const T sinPc = TabSel<T>::sPc(tab)[B.ct0 + i], cosPc = TabSel<T>::cPc(tab)[B.ct0 + i];

    // Start-point samples. Away from the two polar rows the sample positions
    // are grid-aligned, and the samplers reduce to their interior branch with
    // weights 0 or 1/2 (exactly the values the general path computes); the
    // polar rows go through the general samplers.
    const bool interior = j > 0 && j < B.nth - 1;

    // eta, Gamma at the cell centre: traceGreatCircle + two samples
    {
        const T gx = (T)j + T(0.5), gy = (T)i;
        T guT, guP;
        if (interior && !NANS) {
            // the same samples without their zero-weight terms (which add
            // 0 x finite: nothing, in fields that hold no NaN)
            guT = lerp_s<T, false>(v.uth.at(j - 1, i), v.uth.at(j, i), T(0.5));
            guP = v.uph.template belt<false>(j, i, T(0.5));
        } else if (interior) {
            guT = lerp_s<T, NANS>(v.uth.template belt<NANS>(j - 1, i, T(0)),
                                  v.uth.template belt<NANS>(j, i, T(0)), T(0.5));
            guP = lerp_s<T, NANS>(v.uph.template belt<NANS>(j, i, T(0.5)),
                                  v.uph.template belt<NANS>(j + 1, i, T(0.5)), T(0));
        } else {
            guT = sample_vtheta<T, NANS>(v.uth, B.nth, gx, gy);
            guP = sample_vphi<T, NANS>(v.uph, B.nth, gx, gy);
        }
        Circle<T> o;
        TPos<T> tp;
        if (circle_mid<T, NANS>(v, gx, gy, sinTc, cosTc, sinPc, cosPc, guT, guP, o)) {
            T s, co, px, py, pz, rho, ir;
            circle_depart(v, o, s, co, px, py, pz, tp, rho, ir);
        } else {
            tp = o.start;
        }
        T en, gn;
        sample_centered2<T, NANS>(feta, fgam, B.nth, tp, en, gn);
        out.eta[B.off + c] = en;
        out.gam[B.off + c] = gn;
    }
    // u_phi at the west face
    {
        const T gx = (T)j + T(0.5), gy = (T)i - T(0.5);
        const T sinP = TabSel<T>::sPh(tab)[B.ct0 + i], cosP = TabSel<T>::cPh(tab)[B.ct0 + i];
        T guT = interior ? lerp_s<T, NANS>(v.uth.template belt<NANS>(j - 1, i - 1, T(0.5)),
                                           v.uth.template belt<NANS>(j, i - 1, T(0.5)), T(0.5))
                         : sample_vtheta<T, NANS>(v.uth, B.nth, gx, gy);
        T guP = in.uph[B.off + c];
        Circle<T> o;
        T res;
        if (!circle_mid<T, NANS>(v, gx, gy, sinTc, cosTc, sinP, cosP, guT, guP, o)) {
            res = o.unorm;
        } else {
            T r0, r1, r2;
            transport_back<T, NANS>(v, o, r0, r1, r2);
            res = -sinP * r0 + cosP * r1;
        }
        out.uph[B.off + c] = res;
    }
    // u_theta at the south face; sampleVPhi there is always the interior
    // branch with weights 1/2
    if (j < B.nth - 1) {
        const T gx = (T)j + T(1), gy = (T)i;
        const T sinT = TabSel<T>::sTf(tab)[B.rt0 + j + 1], cosT = TabSel<T>::cTf(tab)[B.rt0 + j + 1];
        T guT = in.uth[B.off + c];
        T guP = lerp_s<T, NANS>(v.uph.template belt<NANS>(j, i, T(0.5)),
                                v.uph.template belt<NANS>(j + 1, i, T(0.5)), T(0.5));
        Circle<T> o;
        T res;
        if (!circle_mid<T, NANS>(v, gx, gy, sinT, cosT, sinPc, cosPc, guT, guP, o)) {
            res = o.unorm;
        } else {
            T r0, r1, r2;
            transport_back<T, NANS>(v, o, r0, r1, r2);
            res = r0 * (cosT * cosPc) + r1 * (cosT * sinPc) - r2 * sinT;
        }
        out.uth[B.off + c] = res;
    }
}

template <typename T, bool P2>
__global__ __launch_bounds__(256) void k_advect(const int4* __restrict__ tiles,
                                                const BubbleDev* __restrict__ bub,
                                                BubbleTime* __restrict__ btime, Tables tab,
                                                StatePtrs<T> in, StatePtrs<T> out) {
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0 || stopped_early(btime[tl.x])) return;
    const BubbleDev B = bub[tl.x];
    // the previous step's time advance lands here, before any kernel of this
    // step reads the time
    if (tl.y == 0 && tl.z == 0 && threadIdx.x == 0) {
        if (btime[tl.x].pending) {
            btime[tl.x].t += btime[tl.x].dt_s;
            btime[tl.x].pending = 0;
        }
        if (btime[tl.x].gpend) {
            btime[tl.x].ghist = min(btime[tl.x].ghist + 1, 2);
            btime[tl.x].gpend = 0;
        }
    }
    int j, i;
    tile_cell(tl, j, i);
    if (j >= B.nth || i >= B.nph) return;
    if (btime[tl.x].nans)
        advect_cell<T, true, P2>(B, tab, in, out, j, i);
    else
        advect_cell<T, false, P2>(B, tab, in, out, j, i);
}

// ---------------------------------------------------------- air flow ----
// u_air(t) = u_air_init + (3 - |t - 3|) * corridor, NaN -> 0 (airFlowU/V,
// bubble.cu:2301-2380). The corridor shape is time-independent and is
// precomputed on the host.
// FP32 fields (the mixed paths) evaluate it in FP32, as the reference's
// float kernels do; FP64 fields in FP64.
template <typename T>
__device__ __forceinline__ double air_val(T a0, T shape, double kt) {
    if constexpr (sizeof(T) == 4) {
        float v = fmaf(shape, (float)kt, a0);
        v = __builtin_isnan(v) ? 0.0f : v;
        return (double)v;
    } else {
        double v = (double)a0 + (double)shape * kt;
        return __builtin_isnan(v) ? 0.0 : v;
    }
}
template <typename T>
__device__ __forceinline__ double air_at(const T* __restrict__ a0, const T* __restrict__ shape,
                                         size_t k, double kt) {
    return air_val(a0[k], shape[k], kt);
}



// ----------------------------------------------------------- VISCOUS ----
// The VISCOUS stress path (header.cuh:99; off in the reference): div(u) at
// the centres, then the viscous forces on the u_phi and u_theta faces from
// the advected velocity and thickness; they enter the assembly (with
// gravity) and the velocity updates. FP64 arithmetic in every precision
// mode (the metric terms csc, cot grow at the poles).

// helper_math.h's clamp(float, float, float) = fmaxf(a, fminf(f, b)): the
// value is rounded to float and NaN gives the upper bound.
__device__ __forceinline__ double clamp_ref(double x, float a, float b) {
    return (double)fmaxf(a, fminf((float)x, b));
}

// divergenceKernel (bubble.cu:1217-1269) of cell (j, i).
template <typename T>
__global__ __launch_bounds__(256) void k_visc_div(const int4* __restrict__ tiles,
                                                  const BubbleDev* __restrict__ bub, Tables tab,
                                                  StatePtrs<T> st, double* __restrict__ div) {
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0) return;
    const BubbleDev B = bub[tl.x];
    int j, i;
    tile_cell(tl, j, i);
    if (j >= B.nth || i >= B.nph) return;
    const int nph = B.nph, m = B.mask;
    const size_t row = (size_t)B.off + (size_t)j * nph;
    const double uW = (double)st.uph[row + i], uE = (double)st.uph[row + wrap_col(i + 1, nph, m)];
    const double vN = j != 0 ? (double)st.uth[row - nph + i] : 0.0;
    const double vS = j != B.nth - 1 ? (double)st.uth[row + i] : 0.0;
    const double sinN = tab.sinFace[B.rt0 + j], sinS = tab.sinFace[B.rt0 + j + 1];
    const double factor = tab.cscTc[B.rt0 + j] * B.invh;
    const double termTheta = factor * (vS * sinS - vN * sinN);
    const double termPhi = B.invh * (uE - uW);
    div[row + i] = termTheta + termPhi;
}

// computeforcevelthetaKernel (bubble.cu:1511-1646) on the u_theta face below
// cell (j, i) and computeforcevelphiKernel_viscous (bubble.cu:1649-1816) on
// its u_phi face, from the advected state and div.
template <typename T>
__global__ __launch_bounds__(256) void k_visc_force(const int4* __restrict__ tiles,
                                                    const BubbleDev* __restrict__ bub, Tables tab,
                                                    StatePtrs<T> st, const double* __restrict__ dv,
                                                    double* __restrict__ fu,
                                                    double* __restrict__ fv) {
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0) return;
    const BubbleDev B = bub[tl.x];
    int j, i;
    tile_cell(tl, j, i);
    if (j >= B.nth || i >= B.nph) return;
    const int nph = B.nph, m = B.mask, nth = B.nth, hn = nph / 2;
    const size_t base = (size_t)B.off;
    const double ih = B.invh, re = B.re;
    auto U = [&](int jj, int ii) { return (double)st.uph[base + (size_t)jj * nph + ii]; };
    auto V = [&](int jj, int ii) { return (double)st.uth[base + (size_t)jj * nph + ii]; };
    auto E = [&](int jj, int ii) { return (double)st.eta[base + (size_t)jj * nph + ii]; };
    auto D = [&](int jj, int ii) { return dv[base + (size_t)jj * nph + ii]; };
    const int iW = wrap_col(i - 1, nph, m), iE = wrap_col(i + 1, nph, m);
    if (j < nth - 1) {
        const int opp = wrap_col(i + hn, nph, m);
        const double u0 = U(j, i), u1 = U(j, iE), u2 = U(j + 1, i), u3 = U(j + 1, iE);
        const double divN = D(j, i), divS = D(j + 1, i);
        const double etaN = E(j, i), etaS = E(j + 1, i);
        double out;
        const double invEta = 2. / (etaN + etaS);
        if (__builtin_isinf(invEta)) {
            out = __builtin_nan("");
        } else {
            const double dvm = 0.5 * (divN + divS);
            const double uPhi = 0.25 * (u0 + u1 + u2 + u3);
            const double d0 = E(j, iW), d1 = E(j, iE), d2 = E(j + 1, iW), d3 = E(j + 1, iE);
            const double pDpx = 0.25 * ih * (d1 + d3 - d0 - d2);
            const double pDpy = ih * (etaS - etaN);
            const double v1 = V(j, i), v3 = V(j, iW), v4 = V(j, iE);
            const double v0 = j != 0 ? V(j - 1, i) : 0.5 * (V(j, i) - V(j, opp));
            const double v2 = j != nth - 2 ? V(j + 1, i) : 0.5 * (V(j, i) - V(j, opp));
            const double pvpy = 0.5 * ih * (v2 - v0);
            const double pvpyNorth = ih * (v1 - v0), pvpySouth = ih * (v2 - v1);
            const double pvpx = 0.5 * ih * (v4 - v3);
            const double pupy = 0.5 * ih * (u2 + u3 - u0 - u1);
            const double pupx = 0.5 * ih * (u1 + u3 - u0 - u2);
            const double pupxy = ih * ih * (u0 + u3 - u1 - u2);
            const double pvpxx = ih * ih * (v3 + v4 - 2 * v1);
            const double sinT = tab.sinTf[B.rt0 + j + 1], cosT = tab.cosTf[B.rt0 + j + 1];
            const double cscT = 1. / sinT, cotT = cosT * cscT;
            const double s11N = 0. + 2. * (pvpyNorth + divN);
            const double s11S = 0. + 2. * (pvpySouth + divS);
            const double s22 = -2. * (pvpy - 2. * dvm);
            const double s12 = cscT * pvpx + pupy - uPhi * cotT;
            // sin(theta -+ h/2) at the neighbouring centres
            const double sinN = tab.sinTc[B.rt0 + j], sinS = tab.sinTc[B.rt0 + j + 1];
            const double psspy = ih * (s11S * sinS - s11N * sinN);
            const double pspx = cscT * pvpxx + pupxy - cotT * pupx;
            const double f2 = re * cscT * invEta * pDpx * s12;
            const double f4 = re * invEta * pDpy * 2 * (dvm + pvpy);
            const double f5 = re * cscT * (psspy + pspx - cosT * s22);
            const double f7 = B.gs * sinT;
            out = sinT < 0.1 ? f7 : clamp_ref(f2 + f5 + f4, -0.1f, 0.1f) + f7;
        }
        fv[base + (size_t)j * nph + i] = out;
    }
    {
        const int oW = wrap_col(iW + hn, nph, m), oC = wrap_col(i + hn, nph, m);
        const double divW = D(j, iW), divE = D(j, i);
        const double etaW = E(j, iW), etaE = E(j, i);
        double out;
        const double invEta = 2. / (etaW + etaE);
        if (__builtin_isinf(invEta)) {
            out = __builtin_nan("");
        } else {
            double v0, v1, v2, v3;
            if (j != 0) { v0 = V(j - 1, iW); v1 = V(j - 1, i); }
            else { v0 = 0.5 * (V(j, iW) - V(j, oW)); v1 = 0.5 * (V(j, i) - V(j, oC)); }
            if (j != nth - 1) { v2 = V(j, iW); v3 = V(j, i); }
            else { v2 = 0.5 * (V(j - 1, iW) - V(j - 1, oW)); v3 = 0.5 * (V(j - 1, i) - V(j - 1, oC)); }
            const double dvm = 0.5 * (divW + divE);
            const double pvpx = 0.5 * ih * (v1 + v3 - v0 - v2);
            const double pvpyWest = ih * (v2 - v0), pvpyEast = ih * (v3 - v1);
            const double pvpy = 0.5 * ih * (v2 + v3 - v0 - v1);
            double pupyNorth = 0.0, pupySouth = 0.0, u3, u4;
            const double u1 = U(j, i);
            if (j != 0) { u3 = U(j - 1, i); pupyNorth = ih * (u1 - u3); }
            else u3 = -U(j, oC);
            if (j != nth - 1) { u4 = U(j + 1, i); pupySouth = ih * (u4 - u1); }
            else u4 = -U(j, oC);
            const double pupy = 0.5 * ih * (u4 - u3);
            const double sinT = tab.sinTc[B.rt0 + j], cosT = tab.cosTc[B.rt0 + j];
            const double cscT = 1. / sinT, cotT = cosT * cscT;
            const double s12 = cscT * pvpx + pupy - u1 * cotT;
            const double pDpx = ih * (etaE - etaW);
            double d0, d1, d2, d3;
            if (j != 0) { d0 = E(j - 1, iW); d2 = E(j - 1, i); }
            else { d0 = E(j, oW); d2 = E(j, oC); }
            if (j != nth - 1) { d1 = E(j + 1, iW); d3 = E(j + 1, i); }
            else { d1 = E(j, oW); d3 = E(j, oC); }
            const double pDpy = 0.25 * ih * (d1 + d3 - d0 - d2);
            // sin, cos at theta -+ h/2: the faces j and j + 1
            const double sinN = tab.sinTf[B.rt0 + j], sinS = tab.sinTf[B.rt0 + j + 1];
            const double cosN = tab.cosTf[B.rt0 + j], cosS = tab.cosTf[B.rt0 + j + 1];
            double uNorth = 0.5 * (u3 + u1), uSouth = 0.5 * (u1 + u4);
            if (j == 0) uNorth = 0.5 * (u1 - u3);
            if (j == nth - 1) uSouth = 0.5 * (u1 - u4);
            const double psspy = ih * (ih * (v0 + v3 - v1 - v2) + sinS * pupySouth -
                                       sinN * pupyNorth - cosS * uSouth + cosN * uNorth);
            const double s22W = 2 * (2 * divW - pvpyWest), s22E = 2 * (2 * divE - pvpyEast);
            const double pspx = ih * (s22E - s22W);
            const double f2 = re * invEta * pDpy * s12;
            const double f4 = re * invEta * cscT * pDpx * 2 * (2 * dvm - pvpy);
            const double f5 = re * cscT * (psspy + pspx + cosT * s12);
            out = sinT < 0.1 ? 0. : clamp_ref(f2 + f4 + f5, -0.1f, 0.1f);
        }
        fu[base + (size_t)j * nph + i] = out;
    }
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
// ------------------------------------------------- blocked row layout ----
// A row block gives thread t the C consecutive cells [C t, C t + C) of its
// theta row: row loads and stores are one contiguous vector per thread, and
// the phi-line solve eliminates each thread's interior cells locally.
template <int C, typename X>
__device__ __forceinline__ void ld_row(const X* __restrict__ p, X (&v)[C]) {
    const X* q = (const X*)__builtin_assume_aligned(p, sizeof(X) * C);
#pragma unroll
    for (int k = 0; k < C; ++k) v[k] = q[k];
}
// A phi column wrapped into the row: for rows known at compile time to be
// of power-of-two length one AND (P2), else wrap_col.
template <bool P2>
__device__ __forceinline__ int wrap_t(int x, int nph, int m) {
    if constexpr (P2) return x & m;
    else return wrap_col(x, nph, m);
}

// C consecutive cells of a row from column `start` on, wrapping past the
// row's end: the phi + pi cells across a pole. One aligned vector when the
// block stays inside the row on a multiple of C (always for power-of-two
// rows, where n_theta is a multiple of C), else cell by cell.
template <int C, bool P2, typename X>
__device__ __forceinline__ void ld_row_wrap(const X* __restrict__ rowp, int start, int nph, int m,
                                            X (&v)[C]) {
    if (P2 || (start % C == 0 && start + C <= nph)) {
        ld_row<C>(rowp + start, v);
        return;
    }
#pragma unroll
    for (int k = 0; k < C; ++k) v[k] = rowp[wrap_col(start + k, nph, m)];
}

template <int C, typename X>
__device__ __forceinline__ void st_row(X* __restrict__ p, const X (&v)[C]) {
    X* q = (X*)__builtin_assume_aligned(p, sizeof(X) * C);
#pragma unroll
    for (int k = 0; k < C; ++k) q[k] = v[k];
}

// The t/12 blend of the air with its phi mirror (bubble.cu:1364-1379) at the
// N consecutive columns i0 + k of a row (N = C: cells, N = C + 1: the phi
// faces of C cells, the last one wrapping): 0.5 (a(i) bl +- a(nph - i) (1 -
// bl)), + for u_phi, - for u_theta. The mirror columns nph - i0 - k, k >= 1,
// are the C-aligned block [nph - i0 - C, nph - i0) read backwards, so both
// sides come as row vectors; column (nph - i0) mod nph is the one extra.
// FP32 fields only; the blend in FP32 too (the reference's float
// arithmetic), which keeps the 3 C + 1 results in single registers.
template <int C, int N, bool MINUS, bool P2>
__device__ __forceinline__ void air_blend(const float* __restrict__ a0,
                                          const float* __restrict__ shape, size_t row, int i0,
                                          int nph, int m, float kt, float bl, float (&out)[N]) {
    float a[C], sh[C], am[C], sm[C];
    ld_row<C>(a0 + row + i0, a);
    ld_row<C>(shape + row + i0, sh);
    const size_t mb = row + (size_t)(nph - i0 - C), cm = row + wrap_t<P2>(nph - i0, nph, m);
    ld_row<C>(a0 + mb, am);
    ld_row<C>(shape + mb, sm);
    auto val = [&](float x, float y) {
        const float v = fmaf(y, kt, x);
        return __builtin_isnan(v) ? 0.0f : v;
    };
    const size_t ce = row + wrap_t<P2>(i0 + C, nph, m);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const float own = k < C ? val(a[k], sh[k]) : val(a0[ce], shape[ce]);
        const float mir = k == 0 ? val(a0[cm], shape[cm]) : val(am[C - k], sm[C - k]);
        out[k] = MINUS ? 0.5f * (own * bl - mir * (1.0f - bl)) : 0.5f * (own * bl + mir * (1.0f - bl));
    }
}

// Operator (and optionally right-hand side) of the C consecutive cells
// [i0, i0 + C) of row j, computed from row-vector loads. Every phi face is
// evaluated once and shared by its two cells; each expression is the one of
// concentrationLinearSystemKernel (bubble.cu:1330-1472), so a face weight
// has the same value whichever cell computes it.
template <int C, typename T>
struct RowSys {
    double wN[C], wS[C], wF[C + 1];   // effective weights; wF[k] is the face west of cell k
    double diag[C], rhs[C];
    bool brk[C];
    // the initial iterate on the cells and their neighbours: a value of the
    // state precision, or gamma_max on a broken cell; kept as the state's
    // value and the broken flags (an FP32 state then takes half the
    // registers of FP64 copies)
    T x0v[C], x0Nv[C], x0Sv[C], x0Wv, x0Ev;
    bool bN[C], bS[C], bW, bE;
    double gmax;
    __device__ __forceinline__ double x0(int k) const { return brk[k] ? gmax : (double)x0v[k]; }
    __device__ __forceinline__ double x0N(int k) const { return bN[k] ? gmax : (double)x0Nv[k]; }
    __device__ __forceinline__ double x0S(int k) const { return bS[k] ? gmax : (double)x0Sv[k]; }
    __device__ __forceinline__ double x0W() const { return bW ? gmax : (double)x0Wv; }
    __device__ __forceinline__ double x0E() const { return bE ? gmax : (double)x0Ev; }
};

// With `guess` (assembly only) the initial iterate is Gamma* + guess instead
// of Gamma* on cells that are not broken.
template <typename T, int C, bool RHS, bool VISC, bool P2>
__device__ __forceinline__ void row_system(const BubbleDev& B, double t, const Tables& tab,
                                           const StatePtrs<T>& s, const AirPtrs<T>& air, int j,
                                           int i0, RowSys<C, T>& q,
                                           const T* __restrict__ guess = nullptr) {
    const int nph = B.nph, m = B.mask, nth = B.nth;
    const size_t row = (size_t)B.off + (size_t)j * nph;
    const double gmax = B.gamma_max;
    const bool hasN = j > 0, hasS = j < nth - 1;
    // neighbours across the poles are the phi + pi cells of the same row
    // (bubble.cu:1384-1401); their face weights vanish with sin(theta_face)
    const size_t rowN = hasN ? row - nph : row, rowS = hasS ? row + nph : row;
    const int iNS = hasN ? i0 : wrap_t<P2>(i0 + nth, nph, m);
    const int iSS = hasS ? i0 : wrap_t<P2>(i0 + nth, nph, m);
    T ev[C], eNv[C], eSv[C], gv[C], gNv[C], gSv[C];
    ld_row<C>(s.eta + row + i0, ev);
    ld_row<C>(s.gam + row + i0, gv);
    ld_row_wrap<C, P2>(s.eta + rowN, iNS, nph, m, eNv);
    ld_row_wrap<C, P2>(s.gam + rowN, iNS, nph, m, gNv);
    ld_row_wrap<C, P2>(s.eta + rowS, iSS, nph, m, eSv);
    ld_row_wrap<C, P2>(s.gam + rowS, iSS, nph, m, gSv);
    const int iWc = wrap_t<P2>(i0 - 1, nph, m), iEc = wrap_t<P2>(i0 + C, nph, m);
    const double eW = (double)s.eta[row + iWc], eE = (double)s.eta[row + iEc];
    double gW = (double)s.gam[row + iWc], gE = (double)s.gam[row + iEc];

    const double dt = B.dt, invDt = 1.0 / dt, Cr = B.Cr;
    const double CrDt = Cr * dt, MDt = B.M * dt, s2 = B.invh * B.invh;
    const double sinN = tab.sinFace[B.rt0 + j], sinS = tab.sinFace[B.rt0 + j + 1];
    const double sinTheta = tab.sinTc[B.rt0 + j], csc = tab.cscTc[B.rt0 + j];
    const double cW = s2 * csc * MDt, cN = s2 * sinN * MDt, cS = s2 * sinS * MDt;

    bool bW = is_broken(gW, eW, gmax), bE = is_broken(gE, eE, gmax);
    bool bN[C], bS[C];
    double e[C], g[C];
#pragma unroll
    for (int k = 0; k < C; ++k) {
        e[k] = (double)ev[k];
        g[k] = (double)gv[k];
        q.brk[k] = is_broken(g[k], e[k], gmax);
        bN[k] = hasN && is_broken((double)gNv[k], (double)eNv[k], gmax);
        bS[k] = hasS && is_broken((double)gSv[k], (double)eSv[k], gmax);
    }
    // the initial iterate: Gamma* plus the guess, rounded to the state
    // precision (its residual is evaluated exactly below, so the rounding
    // only perturbs the guess)
    T gxv[C];
#pragma unroll
    for (int k = 0; k < C; ++k) gxv[k] = gv[k];
    if (guess) {
        T gc[C];
        ld_row<C>(guess + row + i0, gc);
#pragma unroll
        for (int k = 0; k < C; ++k) gxv[k] = gv[k] + gc[k];
        ld_row_wrap<C, P2>(guess + rowN, iNS, nph, m, gc);
#pragma unroll
        for (int k = 0; k < C; ++k) gNv[k] = gNv[k] + gc[k];
        ld_row_wrap<C, P2>(guess + rowS, iSS, nph, m, gc);
#pragma unroll
        for (int k = 0; k < C; ++k) gSv[k] = gSv[k] + gc[k];
        gW = (double)(T)(gW + (double)guess[row + iWc]);
        gE = (double)(T)(gE + (double)guess[row + iEc]);
    }
    // Every face term of concentrationLinearSystemKernel has the factor
    // 1 / (eta_f + Cr dt): the weight s2 sin M dt / (eta_f + Cr dt), the
    // film term u eta_f / (eta_f + Cr dt), the air term u_air / (eta_f / Cr
    // / dt + 1) = u_air Cr dt / (eta_f + Cr dt) and the gravity term
    // eta_f / (eta_f / dt + Cr) = eta_f dt / (eta_f + Cr dt). One reciprocal
    // per face serves all of them (a few ulp from the reference's divisions).
    // phi faces: face k lies between cells k-1 and k
    double wFr[C + 1], etaF[C + 1], invF[C + 1];
#pragma unroll
    for (int k = 0; k <= C; ++k) {
        const double l = k == 0 ? eW : e[k - 1];
        const double r = k == C ? eE : e[k];
        etaF[k] = (l + r) / 2.0;
        invF[k] = rcp64(etaF[k] + CrDt);
        wFr[k] = cW * invF[k];
        const bool bl = k == 0 ? bW : q.brk[k - 1];
        const bool br = k == C ? bE : q.brk[k];
        q.wF[k] = (bl || br) ? 0.0 : wFr[k];
    }
    T uv[C + 1], vNv[C], vSv[C];            // the film velocities (state precision)
    T uAirF[C + 1], vAirNv[C], vAirSv[C];   // the blended air (state precision)
#pragma unroll
    for (int k = 0; k < C; ++k) vAirNv[k] = vAirSv[k] = T(0);
    const double kt = 3.0 - fabs(t - 3.0);
    const double bl = t / 12.0;
    if (RHS) {
        T u[C], vn[C], vs[C];
        ld_row<C>(s.uph + row + i0, u);
#pragma unroll
        for (int k = 0; k < C; ++k) { uv[k] = u[k]; vNv[k] = vSv[k] = T(0); }
        uv[C] = s.uph[row + iEc];
        if (hasN) {
            ld_row<C>(s.uth + row - nph + i0, vn);
#pragma unroll
            for (int k = 0; k < C; ++k) vNv[k] = vn[k];
        }
        if (hasS) {
            ld_row<C>(s.uth + row + i0, vs);
#pragma unroll
            for (int k = 0; k < C; ++k) vSv[k] = vs[k];
        }
        if (B.airflow) {
            // per phi face, shared by the two cells of the face; v at the
            // north and south faces of the cells. FP32 fields: row vectors;
            // FP64 fields (twice the registers) per column.
            if constexpr (sizeof(T) == 4) {
                const float ktf = (float)kt, blf = (float)bl;
                air_blend<C, C + 1, false, P2>(air.ua0, air.au, row, i0, nph, m, ktf, blf, uAirF);
                if (hasN)
                    air_blend<C, C, true, P2>(air.va0, air.av, row - nph, i0, nph, m, ktf, blf, vAirNv);
                if (hasS) air_blend<C, C, true, P2>(air.va0, air.av, row, i0, nph, m, ktf, blf, vAirSv);
            } else {
#pragma unroll
                for (int k = 0; k <= C; ++k) {
                    const int i = wrap_t<P2>(i0 + k, nph, m), opp = wrap_t<P2>(nph - i, nph, m);
                    uAirF[k] = 0.5 * (air_at(air.ua0, air.au, row + i, kt) * bl +
                                      air_at(air.ua0, air.au, row + opp, kt) * (1 - bl));
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < C; ++k) {
        const double eN = (double)eNv[k], eS = (double)eSv[k];
        const double etaUp = (eN + e[k]) / 2.0, etaDown = (eS + e[k]) / 2.0;
        const double etaLeft = etaF[k], etaRight = etaF[k + 1];
        const double invU = rcp64(etaUp + CrDt), invD = rcp64(etaDown + CrDt);
        const double invL = invF[k], invR = invF[k + 1];
        const double wN = cN * invU;
        const double wS = cS * invD;
        const double wW = wFr[k], wE = wFr[k + 1];
        const bool brk = q.brk[k];
        q.wN[k] = (brk || bN[k]) ? 0.0 : wN;
        q.wS[k] = (brk || bS[k]) ? 0.0 : wS;
        q.x0v[k] = gxv[k];
        q.x0Nv[k] = gNv[k];
        q.x0Sv[k] = gSv[k];
        q.bN[k] = bN[k];
        q.bS[k] = bS[k];
        q.rhs[k] = 0.0;
        if (brk) {
            q.diag[k] = 1.0;
            q.rhs[k] = gmax;
            continue;
        }
        q.diag[k] = qdiv(sinTheta * invDt, g[k]) - (((-wN) + (-wW)) + (-wE) + (-wS));
        if (!RHS) continue;
        const int i = i0 + k;
        double uW = (double)uv[k], uE = (double)uv[k + 1], vN = (double)vNv[k], vS = (double)vSv[k];
        if (__builtin_isnan(vN)) vN = 0.0;
        if (__builtin_isnan(vS)) vS = 0.0;
        if (__builtin_isnan(uW)) uW = 0.0;
        if (__builtin_isnan(uE)) uE = 0.0;
        double div = B.invh * (uE * etaRight * invR - uW * etaLeft * invL +
                               (vS * sinS * etaDown * invD - vN * sinN * etaUp * invU));
        if (B.airflow) {
            double vAirN = (double)vAirNv[k], vAirS = (double)vAirSv[k];
            if constexpr (sizeof(T) != 4) {
                const int opp = wrap_t<P2>(nph - i, nph, m);
                if (hasN) {
                    const size_t rn = row - nph;
                    vAirN = 0.5 * (air_at(air.va0, air.av, rn + i, kt) * bl -
                                   air_at(air.va0, air.av, rn + opp, kt) * (1 - bl));
                }
                if (hasS) {
                    vAirS = 0.5 * (air_at(air.va0, air.av, row + i, kt) * bl -
                                   air_at(air.va0, air.av, row + opp, kt) * (1 - bl));
                }
            }
            div += B.invh * CrDt *
                   ((double)uAirF[k + 1] * invR - (double)uAirF[k] * invL + vAirS * sinS * invD -
                    vAirN * sinN * invU);
        }
        if (VISC && B.gravity) {
            // `gravity` with `VISCOUS` (bubble.cu:1458-1464): the viscous
            // forces in place of the gravity term
            const size_t cc = row + i;
            const double fuW = air.fu[cc], fuE = air.fu[row + wrap_t<P2>(i + 1, nph, m)];
            const double fvN = hasN ? air.fv[cc - nph] : 0.0;
            const double fvS = hasS ? air.fv[cc] : 0.0;
            div += B.invh * dt *
                   (sinS * fvS * etaDown * invD - sinN * fvN * etaUp * invU +
                    fuE * etaRight * invR - fuW * etaLeft * invL);
        } else if (B.gravity) {
            div += B.gs * B.invh * dt *
                   (sinS * sinS * etaDown * invD - sinN * sinN * etaUp * invU);
        }
        double rr = sinTheta * invDt - div;
        if (bN[k]) rr += wN * gmax;
        if (bS[k]) rr += wS * gmax;
        if (k == 0 ? bW : q.brk[k - 1]) rr += wW * gmax;
        if (k == C - 1 ? bE : q.brk[k + 1]) rr += wE * gmax;
        q.rhs[k] = rr;
    }
    q.x0Wv = (T)gW;
    q.x0Ev = (T)gE;
    q.bW = bW;
    q.bE = bE;
    q.gmax = gmax;
}

// Periodic phi-line solve of one row, normalised to unit diagonal
// (x_i + a_i x_{i-1} + c_i x_{i+1} = d_i), C dividing n, thread t owning
// cells [C t, C t + C) in registers (threads t >= n / C hold nothing but
// take part in the barriers). Partition method: each thread expresses its
// interior cells k < C-1 as x_k = al_k + be_k X_{t-1} + ga_k X_t, where X_t is
// its last cell, by a Thomas sweep with three right-hand sides; substituting
// the neighbours' interior cells into the last-cell equations leaves a
// periodic tridiagonal system in the R = n / C unknowns X_t, solved by
// parallel cyclic reduction in LDS (steps as pcr_periodic) when R is a power
// of two, else by pcr_general; back-substitution recovers the interior
// cells. On exit d holds x. `lds` holds 6 R values (8 R when R is not a
// power of two).
template <typename Ti, int C>
__device__ __forceinline__ int line_solve(Ti (&a)[C], Ti (&c)[C], Ti (&d)[C], int n, int steps,
                                          Ti* lds, float* redf, const RowSeg& sg) {
    const int R = n / C;
    // threads past the block's last segment take no part (an index past the
    // row keeps them out of the LDS)
    const int tid = sg.seg < sg.nseg ? sg.lt : R;
    const bool act = tid < R;
    // this row's share of the block's LDS
    Ti* ra = lds + (size_t)(sg.seg < sg.nseg ? sg.seg : 0) * ((R & (R - 1)) == 0 ? 6 : 8) * R;
    Ti* rc = ra + R;
    Ti* rd = rc + R;
    Ti* xa = rd + R;
    Ti* xb = xa + R;
    Ti* xg = xb + R;
    Ti al[C], be[C], ga[C];
#pragma unroll
    for (int k = 0; k < C; ++k) al[k] = be[k] = ga[k] = Ti(0);
    if (C > 1 && act) {
        constexpr int NI = C - 1;
        Ti cp[C], pa[C], pb[C], pg[C];
        cp[0] = NI > 1 ? c[0] : Ti(0);
        pa[0] = d[0];
        pb[0] = -a[0];
        pg[0] = NI == 1 ? -c[0] : Ti(0);
#pragma unroll
        for (int k = 1; k < NI; ++k) {
            const Ti im = pre_rcp(Ti(1) - a[k] * cp[k - 1]);
            cp[k] = (k < NI - 1 ? c[k] : Ti(0)) * im;
            pa[k] = (d[k] - a[k] * pa[k - 1]) * im;
            pb[k] = (-a[k] * pb[k - 1]) * im;
            pg[k] = ((k == NI - 1 ? -c[k] : Ti(0)) - a[k] * pg[k - 1]) * im;
        }
        al[NI - 1] = pa[NI - 1];
        be[NI - 1] = pb[NI - 1];
        ga[NI - 1] = pg[NI - 1];
#pragma unroll
        for (int k = NI - 2; k >= 0; --k) {
            al[k] = pa[k] - cp[k] * al[k + 1];
            be[k] = pb[k] - cp[k] * be[k + 1];
            ga[k] = pg[k] - cp[k] * ga[k + 1];
        }
        xa[tid] = al[0];
        xb[tid] = be[0];
        xg[tid] = ga[0];
    }
    __syncthreads();
    if (act) {
        constexpr int L = C - 1;
        Ti A, Bd, Cc, D;
        if (C > 1) {
            const int tn = tid + 1 == R ? 0 : tid + 1;
            const Ti a0n = xa[tn], b0n = xb[tn], g0n = xg[tn];
            A = a[L] * be[L > 0 ? L - 1 : 0];
            Bd = Ti(1) + a[L] * ga[L > 0 ? L - 1 : 0] + c[L] * b0n;
            Cc = c[L] * g0n;
            D = d[L] - a[L] * al[L > 0 ? L - 1 : 0] - c[L] * a0n;
        } else {
            A = a[0];
            Bd = Ti(1);
            Cc = c[0];
            D = d[0];
        }
        const Ti ib = pre_rcp(Bd);
        ra[tid] = A * ib;
        rc[tid] = Cc * ib;
        rd[tid] = D * ib;
    }
    __syncthreads();
    // the interface unknowns' arrays xa, xb, xg are free from here on: the
    // second buffer of the reduction
    Ti* xs;
    const int done = (R & (R - 1)) == 0
                         ? pcr_periodic<Ti, 1>(ra, rc, rd, xa, xb, xg, R, steps, redf, &xs, tid, sg.nt)
                         : pcr_general<Ti>(ra, rc, rd, xg + R, xa, xb, xg, xg + 2 * R, R, &xs, tid);
    if (act) {
        const Ti Xt = xs[tid], Xp = xs[tid == 0 ? R - 1 : tid - 1];
#pragma unroll
        for (int k = 0; k < C - 1; ++k) d[k] = al[k] + be[k] * Xp + ga[k] * Xt;
        d[C - 1] = Xt;
    }
    __syncthreads();
    return done;
}

// Row-block assembly: coefficients and rhs of one theta row, the initial
// FP64 residual r0 = b - A Gamma*, r_in = (Ti) r0, the spectral bound mu of
// the line-preconditioned operator and the first line solve z = M^-1 r_in,
// whose reduction depth is recorded for the later line solves of the same
// matrix. Partials as k_resid. The block of row 0 resets the bubble's
// solver control block; the broken-cell override of Gamma* is written back
// to the state as the reference does.
template <typename T, typename Ti, int C, bool VISC, bool P2, bool PK>
__global__ void k_assemble(const int2* __restrict__ rows, const BubbleDev* __restrict__ bub,
                           BubbleTime* __restrict__ btime, Tables tab, StatePtrs<T> s,
                           AirPtrs<T> air, SysPtrs<Ti> sys, T gmax_state,
                           SolveCtl* __restrict__ ctl, double* __restrict__ part, long stride,
                           const T* __restrict__ guess, SolveParams sp) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double red[16 * 4];
    __shared__ float redf[16];
    const int2 rw = rows[blockIdx.x];
    if (rw.x < 0) return;
    if (stopped_early(btime[rw.x])) {
        // a bubble stopped by a failed solve: its solve kernels return at
        // once, its verdict stays the failed one
        if (rw.y == 0 && threadIdx.x == 0) {
            flag_store(&ctl[rw.x].outer_done, 1);
            btime[rw.x].stop |= 4;
        }
        return;
    }
    const BubbleDev B = bub[rw.x];
    const int nph = B.nph;
    // this thread's row of the block (several rows of a narrow width share one)
    const RowSeg sg = row_seg<PK>(nph / C);
    const int j = rw.y + sg.seg;
    const bool valid = !PK || (sg.seg < sg.nseg && j < B.nth);
    const bool act = valid && sg.lt < nph / C;
    const int i0 = C * sg.lt;
    const double t = btime[rw.x].t;
    const size_t row = (size_t)B.off + (size_t)j * nph;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    double mu = 0.0;
    float muf = 0.0f;
    Ti a[C], cc[C], d[C], rin[C];
#pragma unroll
    for (int k = 0; k < C; ++k) a[k] = cc[k] = d[k] = rin[k] = Ti(0);
    if (act) {
        RowSys<C, T> q;
        // the extrapolated correction once a previous solve stored one
        row_system<T, C, true, VISC, P2>(B, t, tab, s, air, j, i0, q,
                                     btime[rw.x].ghist > 0 ? guess : nullptr);
        Ti wn[C], ww[C], dg[C];
#pragma unroll
        for (int k = 0; k < C; ++k) {
            const double wW = q.wF[k], wE = q.wF[k + 1];
            const double xl = k == 0 ? q.x0W() : q.x0(k - 1);
            const double xr = k == C - 1 ? q.x0E() : q.x0(k + 1);
            const double ax = q.diag[k] * q.x0(k) - q.wN[k] * q.x0N(k) - q.wS[k] * q.x0S(k) -
                              wW * xl - wE * xr;
            const double r0 = q.rhs[k] - ax;
            // x^T N x <= sum (wN + wS) x_i^2 and x^T M x >= sum (diag - wW - wE)
            // x_i^2 for the theta couplings N and the phi-line blocks M, so
            // the eigenvalues of M^-1 A lie in [1 - mu, 1 + mu] with mu the
            // largest ratio below
            // (a bound with a 2 % margin downstream: FP32 is ample)
            if (!q.brk[k])
                muf = fmaxf(muf, (float)(q.wN[k] + q.wS[k]) *
                                     __builtin_amdgcn_rcpf((float)(q.diag[k] - wW - wE)));
            rin[k] = (Ti)r0;
            wn[k] = (Ti)q.wN[k];
            ww[k] = (Ti)wW;
            dg[k] = (Ti)q.diag[k];
            const Ti inv = pre_rcp(dg[k]);
            a[k] = -ww[k] * inv;
            cc[k] = -(Ti)wE * inv;
            d[k] = rin[k] * inv;
            acc[0] += r0 * r0;
            acc[1] += q.rhs[k] * q.rhs[k];
            acc[3] += (double)rin[k] * (double)rin[k];
        }
        st_row<C>(sys.b64 + row + i0, q.rhs);
        st_row<C>(sys.wN + row + i0, wn);
        st_row<C>(sys.wW + row + i0, ww);
        st_row<C>(sys.diag + row + i0, dg);
        st_row<C>(sys.r + row + i0, rin);
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int k = 0; k < C; ++k) {
            const size_t c = row + i0 + k;
            if (is_broken((double)s.gam[c], (double)s.eta[c], B.gamma_max)) s.gam[c] = gmax_state;
        }
    }
    const int steps = line_solve<Ti, C>(a, cc, d, nph, -1, reinterpret_cast<Ti*>(lds_raw), redf, sg);
    if (act) {
        st_row<C>(sys.z + row + i0, d);
#pragma unroll
        for (int k = 0; k < C; ++k) acc[2] += (double)rin[k] * (double)d[k];
    }
    seg_sum<4>(acc, red, sg);
    mu = seg_max((double)muf, red, sg);
    if (!valid) return;
    if (sg.lt == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) part[q * stride + B.pr0 + j] = acc[q];
        part[4 * stride + B.pr0 + j] = mu;
        sys.rsteps[B.pr0 + j] = steps;
    }
    if (sg.lt == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) acc_add(sys.acc, rw.x, acc_slot_res(0), q, j, acc[q]);
        acc_put_max(sys.acc, rw.x, acc_slot_res(0), 4, j, mu);
        if (j == 0) {
            SolveCtl z = {};
            z.xbuf = 0;
            z.outers = 1;
            ctl[rw.x] = z;
        }
    }
    // the first refinement's sums start from zero
    if (j == 0 && sg.lt < 64) acc_zero(sys.acc, rw.x, acc_slot_res(1));
}

// The five-point operator of the inner solve applied to v on this
// thread's cells: out_k = diag v_k - wW v_{k-1} - wE v_{k+1} - wN vN_k - wS vS_k.
template <typename Ti, int C>
struct RowOp {
    Ti dg[C], ww[C], wwE, wn[C], wsv[C];
    __device__ __forceinline__ void load(const SysPtrs<Ti>& sys, size_t row, int i0, int m,
                                         int nph, bool hasN, bool hasS) {
        ld_row<C>(sys.diag + row + i0, dg);
        ld_row<C>(sys.wW + row + i0, ww);
        wwE = sys.wW[row + wrap_col(i0 + C, nph, m)];
#pragma unroll
        for (int k = 0; k < C; ++k) wn[k] = wsv[k] = Ti(0);
        if (hasN) ld_row<C>(sys.wN + row + i0, wn);
        if (hasS) ld_row<C>(sys.wN + row + nph + i0, wsv);
    }
    __device__ __forceinline__ Ti wE(int k) const { return k == C - 1 ? wwE : ww[k + 1]; }
    __device__ __forceinline__ void apply(const Ti (&v)[C], Ti vW, Ti vE, const Ti (&vN)[C],
                                          const Ti (&vS)[C], Ti (&out)[C]) const {
#pragma unroll
        for (int k = 0; k < C; ++k) {
            const Ti l = k == 0 ? vW : v[k - 1];
            const Ti r = k == C - 1 ? vE : v[k + 1];
            out[k] = dg[k] * v[k] - ww[k] * l - wE(k) * r - wn[k] * vN[k] - wsv[k] * vS[k];
        }
    }
    __device__ __forceinline__ void line(const Ti (&r)[C], Ti (&a)[C], Ti (&c)[C], Ti (&d)[C]) const {
#pragma unroll
        for (int k = 0; k < C; ++k) {
            const Ti inv = pre_rcp(dg[k]);
            a[k] = -ww[k] * inv;
            c[k] = -wE(k) * inv;
            d[k] = r[k] * inv;
        }
    }
};

// A row vector of this thread's cells with its west/east and north/south
// neighbours (zero where the row has no neighbour).
template <typename Ti, int C>
__device__ __forceinline__ void ld_stencil(const Ti* __restrict__ v, size_t row, int i0, int m,
                                           int nph, bool hasN, bool hasS, Ti (&c)[C], Ti& w, Ti& e,
                                           Ti (&n)[C], Ti (&s)[C]) {
    ld_row<C>(v + row + i0, c);
    w = v[row + wrap_col(i0 - 1, nph, m)];
    e = v[row + wrap_col(i0 + C, nph, m)];
#pragma unroll
    for (int k = 0; k < C; ++k) n[k] = s[k] = Ti(0);
    if (hasN) ld_row<C>(v + row - nph + i0, n);
    if (hasS) ld_row<C>(v + row + nph + i0, s);
}

// z = M^-1 r on row jr of this thread's segment (zeros where the row does
// not exist or the thread holds no cells), from the stored FP32 residual,
// diagonal and west-face weights: the line system the residual launch that
// stored r formed from the same roundings. Every thread of the block calls
// it (the line solve synchronises the block).
template <typename Ti, int C>
__device__ __forceinline__ void z_of_stored_residual(const SysPtrs<Ti>& sys, const BubbleDev& B,
                                                     int jr, int i0, bool act, int steps,
                                                     Ti* lds, float* redf, const RowSeg& sg,
                                                     Ti (&z)[C]) {
    Ti a[C], c[C];
#pragma unroll
    for (int q = 0; q < C; ++q) a[q] = c[q] = z[q] = Ti(0);
    if (act && jr >= 0 && jr < B.nth) {
        const size_t row = (size_t)B.off + (size_t)jr * B.nph;
        Ti rr[C], dg[C], ww[C];
        ld_row<C>(sys.r + row + i0, rr);
        ld_row<C>(sys.diag + row + i0, dg);
        ld_row<C>(sys.wW + row + i0, ww);
        const Ti wwE = sys.wW[row + wrap_col(i0 + C, B.nph, B.mask)];
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const Ti inv = pre_rcp(dg[q]);
            a[q] = -ww[q] * inv;
            c[q] = -(q + 1 < C ? ww[q + 1] : wwE) * inv;
            z[q] = rr[q] * inv;
        }
    }
    (void)line_solve<Ti, C>(a, c, z, B.nph, steps, lds, redf, sg);
}

// ----------------------------------------------------- FP64 residual ----
// x_new = x_old + e (the inner correction of refinement k-1), r = b - A x_new
// in FP64 with the operator recomputed from eta and Gamma* (bitwise the
// assembly's), r_in = (Ti) r, z = M^-1 r_in. Partials |r|^2, |b|^2, r_in.z,
// |r_in|^2 go to the per-row slots and the refinement's sums. With `sh`
// (Chebyshev) the block of row 0 records for the schedule of the inner
// iterations: after the first refinement, |r_n|^2 of its last inner
// iteration n over (inner_frac tol)^2 |b|^2 (at most 1: n - 1 iterations
// would have done), and whether a further refinement ran. Without sys.z (the
// last refinement of a step) no line solve follows. With `check_prev` (the
// refinement had no inner iteration: its correction is the z of the residual
// before) the start-of-refinement test of the Chebyshev launches is done
// here: a start residual (slot k - 1) that met the tolerance ends the solve.
// FIRST (k = 1): x_0 is formed again here, bitwise as the assembly formed it
// (Gamma* plus the stored guess, gamma_max on broken cells), instead of being
// stored by the assembly and read back. ZP (the last refinement, with no
// inner iteration): its correction e = M^-1 r of the residual before is
// formed here, on the block's rows and one row either side, so that the
// residual launch before needs no line solve when no further refinement
// runs (the usual case: this launch then returns at once); r is not stored
// (neighbouring blocks read the stored one).
template <typename T, typename Ti, int C, bool FIRST, bool P2, bool PK, bool ZP>
__global__ void k_resid(const int2* __restrict__ rows, const BubbleDev* __restrict__ bub,
                        const BubbleTime* __restrict__ btime, Tables tab, StatePtrs<T> s,
                        SolveCtl* __restrict__ ctl, SysPtrs<Ti> sys, const T* __restrict__ guess,
                        const double* __restrict__ xo, double* __restrict__ xn, int k,
                        double* __restrict__ part, long stride, const double* __restrict__ partB,
                        SolveParams sp, SchedStat* __restrict__ sh, int check_prev) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double red[16 * 4];
    __shared__ float redf[16];
    const int2 rw = rows[blockIdx.x];
    if (rw.x < 0 || flag_load(&ctl[rw.x].outer_done)) return;
    if (check_prev) {
        const double rr = acc_sum(sys.acc, rw.x, acc_slot_res(k - 1), 0);
        const double bb = acc_sum(sys.acc, rw.x, acc_slot_res(k - 1), 1);
        if (residual_norm(rr, bb, sp.absolute) <= sp.tol) {
            if (rw.y == 0 && threadIdx.x == 0) flag_store(&ctl[rw.x].outer_done, 1);
            return;
        }
    }
    const bool next = sys.z != nullptr;
    const Ti* __restrict__ ecorr = sys.e;
    (void)ecorr;
    const BubbleDev B = bub[rw.x];
    const int nph = B.nph, m = B.mask, nth = B.nth;
    const RowSeg sg = row_seg<PK>(nph / C);
    const int j = rw.y + sg.seg;
    const bool valid = !PK || (sg.seg < sg.nseg && j < nth);
    const bool act = valid && sg.lt < nph / C;
    const int i0 = C * sg.lt;
    const int steps = next ? block_steps<PK>(sys.rsteps, B.pr0, rw.y, nth - rw.y, sg, red) : 0;
    const size_t row = (size_t)B.off + (size_t)j * nph;
    const bool hasN = j > 0, hasS = j < nth - 1;
    Ti ec[C], eN[C], eS[C], eW = Ti(0), eE = Ti(0);
    if constexpr (ZP) {
        // the depth of the block's rows and the two beside them
        const int jlo = max(rw.y - 1, 0);
        const int jhi = min(rw.y + (PK ? sg.nseg : 1), nth - 1);
        int zsteps = -1;
        if (threadIdx.x < 64)
            for (int r = jlo + (int)threadIdx.x; r <= jhi; r += 64)
                zsteps = max(zsteps, sys.rsteps[B.pr0 + r]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) zsteps = max(zsteps, __shfl_xor(zsteps, o, 64));
        if (threadIdx.x == 0) red[0] = (double)zsteps;
        __syncthreads();
        zsteps = (int)red[0];
        __syncthreads();
        Ti* zl = reinterpret_cast<Ti*>(lds_raw);
        z_of_stored_residual<Ti, C>(sys, B, j - 1, i0, act && hasN, zsteps, zl, redf, sg, eN);
        z_of_stored_residual<Ti, C>(sys, B, j + 1, i0, act && hasS, zsteps, zl, redf, sg, eS);
        z_of_stored_residual<Ti, C>(sys, B, j, i0, act, zsteps, zl, redf, sg, ec);
        // the west and east neighbours of the thread's cells on row j: the
        // last cell of the thread before and the first of the thread after
        // (periodic in phi); the line solve left the LDS free
        const int R = nph / C;
        Ti* first = zl + (size_t)(sg.seg < sg.nseg ? sg.seg : 0) * 2 * R;
        Ti* last = first + R;
        if (act) {
            first[sg.lt] = ec[0];
            last[sg.lt] = ec[C - 1];
        }
        __syncthreads();
        if (act) {
            eW = last[sg.lt == 0 ? R - 1 : sg.lt - 1];
            eE = first[sg.lt + 1 == R ? 0 : sg.lt + 1];
        }
        __syncthreads();
    }
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    Ti a[C], cc[C], d[C], rin[C];
#pragma unroll
    for (int q = 0; q < C; ++q) a[q] = cc[q] = d[q] = rin[q] = Ti(0);
    if (act) {
        double xc[C], xN[C], xS[C], xW, xE;
        RowSys<C, T> w;
        if constexpr (FIRST) {
            row_system<T, C, false, false, P2>(B, 0.0, tab, s, AirPtrs<T>{}, j, i0, w,
                                    btime[rw.x].ghist > 0 ? guess : nullptr);
#pragma unroll
            for (int q = 0; q < C; ++q) {
                xc[q] = w.x0(q);
                xN[q] = w.x0N(q);
                xS[q] = w.x0S(q);
            }
            xW = w.x0W();
            xE = w.x0E();
        } else {
            ld_stencil<double, C>(xo, row, i0, m, nph, hasN, hasS, xc, xW, xE, xN, xS);
            row_system<T, C, false, false, P2>(B, 0.0, tab, s, AirPtrs<T>{}, j, i0, w);
        }
        if constexpr (!ZP) ld_stencil<Ti, C>(ecorr, row, i0, m, nph, hasN, hasS, ec, eW, eE, eN, eS);
        double bv[C];
        ld_row<C>(sys.b64 + row + i0, bv);
        double xnew[C];
#pragma unroll
        for (int q = 0; q < C; ++q) xnew[q] = xc[q] + (double)ec[q];
        const double xWn = xW + (double)eW, xEn = xE + (double)eE;
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const double l = q == 0 ? xWn : xnew[q - 1];
            const double r = q == C - 1 ? xEn : xnew[q + 1];
            double ax = w.diag[q] * xnew[q] - w.wF[q] * l - w.wF[q + 1] * r;
            if (hasN) ax -= w.wN[q] * (xN[q] + (double)eN[q]);
            if (hasS) ax -= w.wS[q] * (xS[q] + (double)eS[q]);
            const double res = bv[q] - ax;
            rin[q] = (Ti)res;
            acc[0] += res * res;
            acc[1] += bv[q] * bv[q];
            acc[3] += (double)rin[q] * (double)rin[q];
        }
        st_row<C>(xn + row + i0, xnew);
        if constexpr (!ZP) st_row<C>(sys.r + row + i0, rin);
        // the line system of the inner operator: the assembly stored exactly
        // these roundings of the same FP64 weights
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const Ti inv = pre_rcp((Ti)w.diag[q]);
            a[q] = -(Ti)w.wF[q] * inv;
            cc[q] = -(Ti)w.wF[q + 1] * inv;
            d[q] = rin[q] * inv;
        }
    }
    if (next) {
        line_solve<Ti, C>(a, cc, d, nph, steps, reinterpret_cast<Ti*>(lds_raw), redf, sg);
        if (act) {
            st_row<C>(sys.z + row + i0, d);
#pragma unroll
            for (int q = 0; q < C; ++q) acc[2] += (double)rin[q] * (double)d[q];
        }
    }
    seg_sum<4>(acc, red, sg);
    if (!valid) return;
    if (sg.lt == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) part[q * stride + B.pr0 + j] = acc[q];
    }
    if (sg.lt == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) acc_add(sys.acc, rw.x, acc_slot_res(k), q, j, acc[q]);
        if (j == 0) {
            ctl[rw.x].xbuf = k & 1;
            ctl[rw.x].outers = k + 1;
        }
    }
    // the next refinement's sums start from zero
    if (j == 0 && sg.lt < 64 && k < kAccRefine) acc_zero(sys.acc, rw.x, acc_slot_res(k + 1));
    if (sh && j == 0 && sg.lt < 64) {
        if (k == 1) {
            double rr = 0.0;
            for (int r = sg.lt; r < nth; r += 64) rr += partB[stride + B.pr0 + r];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) rr += __shfl_xor(rr, o, 64);
            const double tgt = sp.inner_frac * sp.tol;
            const double bb = acc_sum(sys.acc, rw.x, acc_slot_res(0), 1);
            const double ratio = rr / (tgt * tgt * (sp.absolute ? 1.0 : bb));
            if (sg.lt == 0 && !(ratio <= sh[rw.x].tail_ratio)) sh[rw.x].tail_ratio = ratio;
        } else if (sg.lt == 0) {
            sh[rw.x].extra += 1;
        }
    }
}

// ------------------------------------------------------ inner PCG: A p ----
// p = z + beta p_old (beta from the r.z partials of the previous step),
// Ap, partial p.Ap. Iteration 0 of a refinement also decides whether the
// FP64 residual already meets the tolerance.
template <typename Ti, int C, bool PK>
__global__ void k_spmv(const int2* __restrict__ rows, const BubbleDev* __restrict__ bub,
                       SolveCtl* __restrict__ ctl, SysPtrs<Ti> sys, const Ti* __restrict__ pold,
                       Ti* __restrict__ pnew, int it, SolveParams sp, const double* __restrict__ partR,
                       const double* __restrict__ partB, double* __restrict__ partA, long stride) {
    __shared__ double red[16 * 4];
    const int2 rw = rows[blockIdx.x];
    if (rw.x < 0) return;
    SolveCtl* Cb = ctl + rw.x;
    if (Cb->outer_done || (it > 0 && Cb->inner_done)) return;
    const BubbleDev B = bub[rw.x];
    const int nph = B.nph, m = B.mask, nth = B.nth;
    const RowSeg sg = row_seg<PK>(nph / C);
    const int j = rw.y + sg.seg;
    const bool valid = !PK || (sg.seg < sg.nseg && j < nth);
    // the block-level decisions below are the same in every block
    const bool leader = (rw.y == 0 && threadIdx.x == 0);
    const bool hasN = j > 0, hasS = j < nth - 1;
    const bool act = valid && sg.lt < nph / C;
    const int i0 = C * sg.lt;
    const size_t row = (size_t)B.off + (size_t)j * nph;
    double beta = 0.0;
    if (it == 0) {
        double sm[4];
        sum_partials<4>(partR, stride, B.pr0, nth, sm, red);
        const double rr = sm[0], bb = sm[1], rz = sm[2], rrin = sm[3];
        const double res = sp.absolute ? sqrt(rr) : sqrt(rr) / (bb > 0.0 ? sqrt(bb) : 1.0);
        if (res <= sp.tol) {
            if (leader) { Cb->outer_done = 1; Cb->rel_res = res; }
            return;
        }
        if (leader) {
            Cb->rr_in0 = rrin;
            Cb->inner_done = 0;
            Cb->rz[0] = rz;
            Cb->rel_res = res;
            Cb->bb = bb;
        }
    } else {
        double sm[2];
        sum_partials<2>(partB, stride, B.pr0, nth, sm, red);
        const double rz = sm[0], rr = sm[1];
        const double tgt = sp.inner_frac * sp.tol;
        const double bbe = sp.absolute ? 1.0 : Cb->bb;
        if (rr <= sp.inner_rho2 * Cb->rr_in0 || rr <= tgt * tgt * bbe) {
            if (leader) Cb->inner_done = 1;
            return;
        }
        beta = rz / Cb->rz[(it - 1) & 1];
        if (leader) Cb->rz[it & 1] = rz;
    }
    double acc[1] = {0.0};
    if (act) {
        const Ti bt = (Ti)beta;
        Ti zc[C], zN[C], zS[C], zW, zE;
        ld_stencil<Ti, C>(sys.z, row, i0, m, nph, hasN, hasS, zc, zW, zE, zN, zS);
        if (it > 0) {
            Ti pc[C], pN[C], pS[C], pW, pE;
            ld_stencil<Ti, C>(pold, row, i0, m, nph, hasN, hasS, pc, pW, pE, pN, pS);
#pragma unroll
            for (int q = 0; q < C; ++q) {
                zc[q] += bt * pc[q];
                zN[q] += bt * pN[q];
                zS[q] += bt * pS[q];
            }
            zW += bt * pW;
            zE += bt * pE;
        }
        RowOp<Ti, C> op;
        op.load(sys, row, i0, m, nph, hasN, hasS);
        Ti ap[C];
        op.apply(zc, zW, zE, zN, zS, ap);
        st_row<C>(pnew + row + i0, zc);
        st_row<C>(sys.Ap + row + i0, ap);
#pragma unroll
        for (int q = 0; q < C; ++q) acc[0] += (double)zc[q] * (double)ap[q];
    }
    seg_sum<1>(acc, red, sg);
    if (valid && sg.lt == 0) partA[B.pr0 + j] = acc[0];
}

// -------------------------------------------- inner PCG: update + M^-1 ----
template <typename Ti, int C, bool PK>
__global__ void k_cgstep(const int2* __restrict__ rows, const BubbleDev* __restrict__ bub,
                         SolveCtl* __restrict__ ctl, SysPtrs<Ti> sys, const Ti* __restrict__ p,
                         int it, const double* __restrict__ partA, double* __restrict__ partB,
                         long stride) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double red[16 * 4];
    __shared__ float redf[16];
    const int2 rw = rows[blockIdx.x];
    if (rw.x < 0) return;
    SolveCtl* Cb = ctl + rw.x;
    if (Cb->outer_done || Cb->inner_done) return;
    const BubbleDev B = bub[rw.x];
    const int nph = B.nph, m = B.mask, nth = B.nth;
    const RowSeg sg = row_seg<PK>(nph / C);
    const int j = rw.y + sg.seg;
    const bool valid = !PK || (sg.seg < sg.nseg && j < nth);
    const bool act = valid && sg.lt < nph / C;
    const int i0 = C * sg.lt;
    const int steps = block_steps<PK>(sys.rsteps, B.pr0, rw.y, nth - rw.y, sg, red);
    const size_t row = (size_t)B.off + (size_t)j * nph;
    double sm[1];
    sum_partials<1>(partA, stride, B.pr0, nth, sm, red);
    const Ti al = (Ti)(Cb->rz[it & 1] / sm[0]);
    double acc[2] = {0.0, 0.0};
    Ti a[C], cc[C], d[C], rv[C];
#pragma unroll
    for (int q = 0; q < C; ++q) a[q] = cc[q] = d[q] = rv[q] = Ti(0);
    if (act) {
        Ti pv[C], apv[C], ev[C];
        ld_row<C>(p + row + i0, pv);
        ld_row<C>(sys.Ap + row + i0, apv);
        ld_row<C>(sys.r + row + i0, rv);
        if (it > 0) ld_row<C>(sys.e + row + i0, ev);
#pragma unroll
        for (int q = 0; q < C; ++q) {
            ev[q] = (it > 0 ? ev[q] : Ti(0)) + al * pv[q];
            rv[q] = rv[q] - al * apv[q];
            acc[1] += (double)rv[q] * (double)rv[q];
        }
        st_row<C>(sys.e + row + i0, ev);
        st_row<C>(sys.r + row + i0, rv);
        RowOp<Ti, C> op;
        op.load(sys, row, i0, m, nph, j > 0, j < nth - 1);
        op.line(rv, a, cc, d);
    }
    line_solve<Ti, C>(a, cc, d, nph, steps, reinterpret_cast<Ti*>(lds_raw), redf, sg);
    if (act) {
        st_row<C>(sys.z + row + i0, d);
#pragma unroll
        for (int q = 0; q < C; ++q) acc[0] += (double)rv[q] * (double)d[q];
    }
    seg_sum<2>(acc, red, sg);
    if (valid && sg.lt == 0) {
        partB[B.pr0 + j] = acc[0];
        partB[stride + B.pr0 + j] = acc[1];
        if (j == 0) Cb->iters += 1;
    }
}

// ------------------------------------------ multigrid preconditioner ----
// z = B r for the inner PCG: one V(1,1) cycle of geometric multigrid built on
// the phi lines. Level l + 1 merges the row pairs (2J, 2J + 1) of level l
// (semi-coarsening in theta; phi keeps every column, so the phi-line
// relaxation stays exact for the csc(theta) anisotropy at the poles). Its
// operator is the sum of the two rows' equations with the theta couplings
// halved (the re-discretisation at twice the spacing: a face weight
// s2 sin(theta_f) K_f of spacing h becomes s2 sin(theta_f) K_f / 2 for 2h);
// restriction sums the two rows' residuals, prolongation copies a coarse
// value to both rows (each the transpose of the other, so the cycle is a
// symmetric preconditioner). Smoothing: zebra phi-line Gauss-Seidel, even
// then odd rows before the coarse correction and odd then even after it.
// From a zero start the even rows' line solve has no theta terms, which
// makes it the line-Jacobi solve the PCG kernels already produce at level
// 0, and leaves residuals only on the even rows, r(2J) = wN(2J) x(2J - 1) +
// wS(2J) x(2J + 1): the odd rows' launch writes those terms into the coarse
// level's right-hand side (U, D). After the coarse correction an odd row's
// new value depends on its even neighbours only (the line solve replaces
// the old value), so the prolongation is folded into the neighbours it
// reads. The coarsest level is one row (solved exactly) or, for a row
// count that cannot be halved, symmetric zebra passes.

// Coarse operator of level l + 1 (row J) from level l (rows 2J, 2J + 1).
template <typename Ti>
__global__ __launch_bounds__(256) void k_mg_coarsen(const int2* __restrict__ rows,
                                                    const BubbleDev* __restrict__ bub,
                                                    MgPtrs<Ti> F, MgPtrs<Ti> Cn) {
    const int2 rw = rows[blockIdx.x];
    if (rw.x < 0) return;
    const BubbleDev B = bub[rw.x];
    const MgLevelDev lf = F.lv[rw.x], lc = Cn.lv[rw.x];
    const int J = rw.y, nph = B.nph;
    const size_t f0 = (size_t)lf.off + (size_t)(2 * J) * nph, f1 = f0 + nph;
    const size_t c0 = (size_t)lc.off + (size_t)J * nph;
    const bool south = 2 * J + 2 < lf.n;
    for (int i = threadIdx.x; i < nph; i += blockDim.x) {
        const double wN0 = (double)F.wN[f0 + i], wN1 = (double)F.wN[f1 + i];
        const double wN2 = south ? (double)F.wN[f1 + nph + i] : 0.0;
        Cn.wW[c0 + i] = (Ti)((double)F.wW[f0 + i] + (double)F.wW[f1 + i]);
        Cn.wN[c0 + i] = (Ti)(0.5 * wN0);
        Cn.dg[c0 + i] =
            (Ti)((double)F.dg[f0 + i] + (double)F.dg[f1 + i] - 2.0 * wN1 - 0.5 * (wN0 + wN2));
    }
}

// One half-sweep of zebra line relaxation on the rows of one parity of
// level `level` (the launch's row list), with the options of `flags`:
// kMgZeroNb (the first half-sweep from a zero start: no theta terms),
// kMgProlong (the neighbours carry the coarse level's correction),
// kMgRestrict (write this odd row's residual terms into the coarse
// right-hand side), kMgRz (level 0: this row's r.z partial into rz_out).
// Gates: every launch returns for a converged bubble; kMgGateInner also
// for a finished inner solve; kMgGateStart (the first launch of the cycle at
// the start of a refinement) ends the solve when the FP64 residual in partR
// already meets the tolerance. The right-hand side is r0 at level 0 and U +
// D above it.
template <typename Ti, int C>
__global__ void k_mg_line(const int2* __restrict__ rows, const BubbleDev* __restrict__ bub,
                          SolveCtl* __restrict__ ctl, MgPtrs<Ti> L, MgPtrs<Ti> Cn,
                          const Ti* __restrict__ r0, int level, int flags,
                          const int* __restrict__ rsteps, SolveParams sp,
                          const double* __restrict__ partR, double* __restrict__ rz_out,
                          long stride) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double red[16 * 4];
    __shared__ float redf[16];
    const int2 rw = rows[blockIdx.x];
    if (rw.x < 0) return;
    SolveCtl* Cb = ctl + rw.x;
    if (flag_load(&Cb->outer_done)) return;
    if ((flags & kMgGateInner) && Cb->inner_done) return;
    const BubbleDev B = bub[rw.x];
    if (flags & kMgGateStart) {
        double sm[2];
        sum_partials<2>(partR, stride, B.pr0, B.nth, sm, red);
        if (residual_norm(sm[0], sm[1], sp.absolute) <= sp.tol) {
            if (rw.y == 0 && threadIdx.x == 0) flag_store(&Cb->outer_done, 1);
            return;
        }
    }
    const MgLevelDev lv = L.lv[rw.x];
    const int j = rw.y, nph = B.nph, m = B.mask, n = lv.n;
    const size_t row = (size_t)lv.off + (size_t)j * nph;
    // one row a block (the half-sweeps' row lists are per row width)
    const RowSeg sg = row_seg<false>(nph / C);
    const bool act = (int)threadIdx.x < nph / C;
    const int i0 = C * threadIdx.x;
    Ti a[C], cc[C], d[C], rv[C];
#pragma unroll
    for (int q = 0; q < C; ++q) a[q] = cc[q] = d[q] = rv[q] = Ti(0);
    if (act) {
        if (level == 0) {
            ld_row<C>(r0 + row + i0, rv);
        } else {
            Ti u[C], dd[C];
            ld_row<C>(L.U + row + i0, u);
            ld_row<C>(L.D + row + i0, dd);
#pragma unroll
            for (int q = 0; q < C; ++q) rv[q] = u[q] + dd[q];
        }
        Ti rhs[C];
#pragma unroll
        for (int q = 0; q < C; ++q) rhs[q] = rv[q];
        if (!(flags & kMgZeroNb)) {
            const size_t crow = Cn.lv ? (size_t)Cn.lv[rw.x].off : 0;
            if (j > 0) {
                Ti xn[C], w[C];
                ld_row<C>(L.x + row - nph + i0, xn);
                if (flags & kMgProlong) {
                    Ti e[C];
                    ld_row<C>(Cn.x + crow + (size_t)((j - 1) / 2) * nph + i0, e);
#pragma unroll
                    for (int q = 0; q < C; ++q) xn[q] += e[q];
                }
                ld_row<C>(L.wN + row + i0, w);
#pragma unroll
                for (int q = 0; q < C; ++q) rhs[q] += w[q] * xn[q];
            }
            if (j < n - 1) {
                Ti xs[C], w[C];
                ld_row<C>(L.x + row + nph + i0, xs);
                if (flags & kMgProlong) {
                    Ti e[C];
                    ld_row<C>(Cn.x + crow + (size_t)((j + 1) / 2) * nph + i0, e);
#pragma unroll
                    for (int q = 0; q < C; ++q) xs[q] += e[q];
                }
                ld_row<C>(L.wN + row + nph + i0, w);
#pragma unroll
                for (int q = 0; q < C; ++q) rhs[q] += w[q] * xs[q];
            }
        }
        Ti dg[C], ww[C];
        ld_row<C>(L.dg + row + i0, dg);
        ld_row<C>(L.wW + row + i0, ww);
        const Ti wwE = L.wW[row + wrap_col(i0 + C, nph, m)];
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const Ti inv = pre_rcp(dg[q]);
            a[q] = -ww[q] * inv;
            cc[q] = -(q == C - 1 ? wwE : ww[q + 1]) * inv;
            d[q] = rhs[q] * inv;
        }
    }
    const int steps = level == 0 ? rsteps[B.pr0 + j] : -1;
    line_solve<Ti, C>(a, cc, d, nph, steps, reinterpret_cast<Ti*>(lds_raw), redf, sg);
    double acc[1] = {0.0};
    if (act) {
        st_row<C>(L.x + row + i0, d);
        if (flags & kMgRestrict) {
            // odd row j = 2J + 1: wS(2J) x(2J + 1) into coarse row J, wN(2J + 2)
            // x(2J + 1) into coarse row J + 1
            const size_t cb = (size_t)Cn.lv[rw.x].off + (size_t)(j / 2) * nph + i0;
            Ti w[C], o[C];
            ld_row<C>(L.wN + row + i0, w);
#pragma unroll
            for (int q = 0; q < C; ++q) o[q] = w[q] * d[q];
            st_row<C>(Cn.U + cb, o);
            if (j + 1 < n) {
                ld_row<C>(L.wN + row + nph + i0, w);
#pragma unroll
                for (int q = 0; q < C; ++q) o[q] = w[q] * d[q];
                st_row<C>(Cn.D + cb + nph, o);
            }
        }
        if (flags & kMgRz) {
#pragma unroll
            for (int q = 0; q < C; ++q) acc[0] += (double)rv[q] * (double)d[q];
        }
    }
    if (flags & kMgRz) {
        block_sum<1>(acc, red);
        if (threadIdx.x == 0) rz_out[B.pr0 + j] = acc[0];
    }
}

// ------------------------------------------- inner Chebyshev iteration ----
// Line-Jacobi bound up to which the Chebyshev recurrence starts after one
// plain line-Jacobi step (see k_cheb).
constexpr double kChebLeadMu = 0.5;

// Chebyshev acceleration of the phi-line Jacobi iteration for A e = r_in
// (Saad, Iterative Methods for Sparse Linear Systems, Alg. 12.1), with the
// spectrum of M^-1 A inside [1 - mu, 1 + mu] (mu from the assembly, widened
// slightly). No inner products: one kernel per iteration. Iteration k >= 1
// reads e_k (k = 1: e_1 = d_0 = z_0 = M^-1 r_in) and e_{k-1} (d_{k-1} =
// e_k - e_{k-1}, so one vector is written per iteration), forms
// r_k = r_in - A e_k, z_k = M^-1 r_k, rho_k = 1 / (2 sigma - rho_{k-1}),
// d_k = rho_k rho_{k-1} d_{k-1} + (2 rho_k / delta) z_k, e_{k+1} = e_k + d_k.
// Launch 1 of refinement `kr` ends the solve (outer_done) when the FP64
// residual at the refinement's start met the tolerance; later launches of
// the step then return at once. Partial |r_k|^2 goes to partB slot 1.
template <typename Ti, int C, bool PK>
__global__ void k_cheb(const int2* __restrict__ rows, const BubbleDev* __restrict__ bub,
                       SolveCtl* __restrict__ ctl, SysPtrs<Ti> sys, const Ti* __restrict__ ek,
                       const Ti* __restrict__ ekm1, Ti* __restrict__ enext, int it, int kr,
                       SolveParams sp, double* __restrict__ partB, long stride) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double red[16 * 4];
    __shared__ float redf[16];
    const int2 rw = rows[blockIdx.x];
    if (rw.x < 0) return;
    const int b = rw.x;
    SolveCtl* Cb = ctl + b;
    const int od = flag_load(&Cb->outer_done);
    const double* __restrict__ acc0 = sys.acc;
    const double mu = acc_max(acc0, b, acc_slot_res(0), 4);
    const bool converged =
        it == 1 && residual_norm(acc_sum(acc0, b, acc_slot_res(kr), 0),
                                 acc_sum(acc0, b, acc_slot_res(kr), 1), sp.absolute) <= sp.tol;
    const BubbleDev B = bub[b];
    const int nph = B.nph, m = B.mask, nth = B.nth;
    const RowSeg sg = row_seg<PK>(nph / C);
    const int j = rw.y + sg.seg;
    const bool valid = !PK || (sg.seg < sg.nseg && j < nth);
    const bool leader = (rw.y == 0 && threadIdx.x == 0);
    if (od || converged) {
        if (leader && converged) flag_store(&Cb->outer_done, 1);
        return;
    }
    if (leader) Cb->iters += 1;
    // the interval [1 - delta, 1 + delta] widened a little, and kept clear of
    // zero (a lower end at or below zero makes the iteration diverge)
    const double delta = fmin(fmax(mu * 1.02 + 1e-3, 1e-4), 0.9999);
    const double sigma = 1.0 / delta;
    // Jacobi lead: on a weakly coupled system (mu <= kChebLeadMu) the
    // recurrence restarts after the first iterate e_1 = z_0 (a plain line-
    // Jacobi step), so the residual polynomial after k line solves is
    // nu T_{k-1}(nu / delta) / T_{k-1}(1 / delta) (nu the eigenvalues of the
    // line-Jacobi iteration matrix, in [-delta, delta]) instead of
    // T_k(nu / delta) / T_k(1 / delta): a zero at nu = 0 for every k, where
    // the line solves leave most of the residual of such a system (even-
    // degree Chebyshev polynomials are +-1 / T_k(1 / delta) there).
    const int kk = it - (mu <= kChebLeadMu ? 1 : 0);
    double rho_prev = delta, rho = delta;      // rho_0 = 1 / sigma
    for (int q = 1; q <= kk; ++q) {
        rho_prev = rho;
        rho = 1.0 / (2.0 * sigma - rho_prev);
    }
    // the restart's first step: e_{k+1} = e_k + z_k
    const Ti cd = kk == 0 ? Ti(0) : (Ti)(rho * rho_prev),
             cz = kk == 0 ? Ti(1) : (Ti)(2.0 * rho / delta);
    const int steps = block_steps<PK>(sys.rsteps, B.pr0, rw.y, nth - rw.y, sg, red);
    const size_t row = (size_t)B.off + (size_t)j * nph;
    const bool hasN = j > 0, hasS = j < nth - 1;
    const bool act = valid && sg.lt < nph / C;
    const int i0 = C * sg.lt;
    // e_1 = z_0 and e_0 = 0, so d_0 = e_1 - e_0 = z_0
    const Ti* __restrict__ e = it == 1 ? sys.z : ek;
    Ti ev[C], dvv[C], a[C], cc[C], d[C];
#pragma unroll
    for (int q = 0; q < C; ++q) ev[q] = dvv[q] = a[q] = cc[q] = d[q] = Ti(0);
    double acc = 0.0;
    if (act) {
        Ti eN[C], eS[C], eW, eE;
        ld_stencil<Ti, C>(e, row, i0, m, nph, hasN, hasS, ev, eW, eE, eN, eS);
        if (it > 1) {
            // d_{k-1} = e_k - e_{k-1}
            Ti em[C];
            ld_row<C>(ekm1 + row + i0, em);
#pragma unroll
            for (int q = 0; q < C; ++q) dvv[q] = ev[q] - em[q];
        } else {
#pragma unroll
            for (int q = 0; q < C; ++q) dvv[q] = ev[q];
        }
        Ti rin[C];
        ld_row<C>(sys.r + row + i0, rin);
        RowOp<Ti, C> op;
        op.load(sys, row, i0, m, nph, hasN, hasS);
        Ti ax[C], r[C];
        op.apply(ev, eW, eE, eN, eS, ax);
#pragma unroll
        for (int q = 0; q < C; ++q) {
            r[q] = rin[q] - ax[q];
            acc += (double)r[q] * (double)r[q];
        }
        op.line(r, a, cc, d);
    }
    line_solve<Ti, C>(a, cc, d, nph, steps, reinterpret_cast<Ti*>(lds_raw), redf, sg);
    if (act) {
        Ti en[C];
#pragma unroll
        for (int q = 0; q < C; ++q) en[q] = ev[q] + (cd * dvv[q] + cz * d[q]);
        st_row<C>(enext + row + i0, en);
    }
    double a1[1] = {acc};
    seg_sum<1>(a1, red, sg);
    if (valid && sg.lt == 0) partB[stride + B.pr0 + j] = a1[0];
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

// The main-path loads of one cell of k_update (the neighbours west and
// south as the reference reads them).
template <typename T>
struct UpdIn {
    double ga, g, gW, gS, eta, etaW, etaS, u1, v1;
    T ua0, uas, va0, vas, gd;
};

// One cell of the post-solve updates; writes u_phi, u_theta, eta, Gamma,
// the guess pair and the BIMOCQ changes of cell c; returns whether a NaN
// was written.
template <typename T>
__device__ __forceinline__ bool update_cell(const BubbleDev& B, int ghist,
                                            const Tables& tab, const AirPtrs<T>& air,
                                            const double* __restrict__ x, int j, int i, size_t c,
                                            const UpdIn<T>& in, double kt, double tnow,
                                            double invDt, T& o_eta, T& o_gam, T& o_uph,
                                            T& o_uth, T& o_gdel, T& o_gnext, bool guess,
                                            double* __restrict__ chg_e,
                                            double* __restrict__ chg_g) {
    const int nph = B.nph, m = B.mask, nth = B.nth;
    const double gmax = B.gamma_max, dt = B.dt, invh = B.invh, Cr = B.Cr, M = B.M;
    // VISCOUS (applyforcevel{theta,phi}Kernel_viscous, bubble.cu:1938-2038):
    // the viscous force, the t/12 mirror blend of the air and no separate
    // gravity term
    const bool visc = air.fu != nullptr;
    const int opp = wrap_col(nph - i, nph, m);

    // divergenceKernel_fromGamma (bubble.cu:1276-1295)
    const double ga = in.ga;
    const double g = in.g;
    double gnew, div;
    if (g <= 0.0 || g > gmax - kEps) {
        gnew = gmax;
        div = invDt;
    } else {
        gnew = g;
        div = (1.0 - qdiv(fmin(g, 2 * ga), ga)) * invDt;
    }
    if (guess) {
        // initial guess of the next solve: its correction Gamma^{n+1} - Gamma*
        // extrapolated linearly from this step's and the previous step's
        // (this step's alone after a reset); a non-finite value is dropped
        double dl = g - ga;
        if (!__builtin_isfinite(dl)) dl = 0.0;
        const double gn = ghist > 0 ? 2.0 * dl - (double)in.gd : dl;
        o_gdel = (T)dl;
        o_gnext = __builtin_isfinite(gn) ? (T)gn : T(0);
    }
    const double eta = in.eta;
    bool nan_out;
    // applyforceThickness (bubble.cu:2041-2063)
    {
        double v = eta * (1 - dt * div);
        if (B.evaporation) v += B.eva * dt;
        o_eta = (T)fmax(kThresh, v);
        o_gam = (T)gnew;
        nan_out = __builtin_isnan(gnew);
        if (chg_e) {
            // the changes of this step for BIMOCQ (bubble.cu:2062, 2486-2490),
            // from the stored values
            chg_e[c] = (double)(T)fmax(kThresh, v) - eta;
            chg_g[c] = (double)(T)gnew - ga;
        }
    }
    // applyforcevelphiKernel (bubble.cu:1879-1935)
    {
        const double etaW = in.etaW;
        const double u1 = in.u1;
        const double invEta = qdiv(2., etaW + eta);
        double res;
        if (__builtin_isinf(invEta)) {
            res = __builtin_nan("");
        } else if (visc) {
            const double gW = clamp_gamma(in.gW, gmax);
            const double pGpx = invh * (gnew - gW);
            const double f3 = -M * invEta * tab.cscTc[B.rt0 + j] * pGpx;
            double uAir = 0.0;
            if (B.airflow) {
                const size_t rowc = (size_t)B.off + (size_t)j * nph;
                uAir = 0.5 * (air_at(air.ua0, air.au, c, kt) * tnow / 12.0 +
                              air_at(air.ua0, air.au, rowc + opp, kt) * (1 - tnow / 12.0));
            }
            const double f6 = Cr * invEta * uAir;
            res = qdiv(u1 * invDt + f3 + f6 + air.fu[c], invDt + Cr * invEta);
        } else {
            const double gW = clamp_gamma(in.gW, gmax);
            const double pGpx = invh * (gnew - gW) * tab.cscTc[B.rt0 + j];
            const double f1 = -M * invEta * pGpx;
            const double uAir = B.airflow ? air_val(in.ua0, in.uas, kt) : 0.0;
            const double f2 = Cr * invEta * uAir;
            res = qdiv(u1 * invDt + (f1 + f2 + 0.0 + 0.0), invDt + Cr * invEta);
        }
        o_uph = (T)res;
        nan_out |= __builtin_isnan(res);
    }
    // applyforcevelthetaKernel (bubble.cu:1819-1876)
    if (j < nth - 1) {
        const double etaS = in.etaS;
        const double v1 = in.v1;
        const double invEta = qdiv(2., eta + etaS);
        double res;
        if (__builtin_isinf(invEta)) {
            res = __builtin_nan("");
        } else if (visc) {
            const double gS = clamp_gamma(in.gS, gmax);
            const double pGpy = invh * (gS - gnew);
            const double f3 = -M * invEta * pGpy;
            double vAir = 0.0;
            if (B.airflow) {
                const size_t rowc = (size_t)B.off + (size_t)j * nph;
                vAir = 0.5 * (air_at(air.va0, air.av, c, kt) * tnow / 12.0 -
                              air_at(air.va0, air.av, rowc + opp, kt) * (1 - tnow / 12.0));
            }
            const double f6 = Cr * invEta * vAir;
            res = qdiv(v1 * invDt + f3 + f6 + air.fv[c], invDt + Cr * invEta);
        } else {
            const double gS = clamp_gamma(in.gS, gmax);
            const double pGpy = invh * (gS - gnew);
            const double f1 = -M * invEta * pGpy;
            const double vAir = B.airflow ? air_val(in.va0, in.vas, kt) : 0.0;
            const double f2 = Cr * invEta * vAir;
            const double f3 = B.gravity ? B.gs * tab.sinTf[B.rt0 + j + 1] : 0.0;
            res = qdiv(v1 * invDt + f1 + f2 + f3 + 0.0, invDt + Cr * invEta);
        }
        o_uth = (T)res;
        nan_out |= __builtin_isnan(res);
    }
    return nan_out;
}

// Two consecutive cells of a row per thread (cells 2t, 2t + 1 of the tile;
// tile widths are even), so that each field is one 2-vector load and store.
template <typename T>
struct alignas(2 * sizeof(T)) Pair2 {
    T a, b;
};
template <typename T>
__device__ __forceinline__ Pair2<T> ld2(const T* __restrict__ p) {
    return *reinterpret_cast<const Pair2<T>*>(p);
}
template <typename T>
__device__ __forceinline__ void st2(T* __restrict__ p, T a, T b) {
    *reinterpret_cast<Pair2<T>*>(p) = Pair2<T>{a, b};
}

template <typename T>
__global__ __launch_bounds__(128) void k_update(const int4* __restrict__ tiles,
                                                const BubbleDev* __restrict__ bub,
                                                BubbleTime* __restrict__ btime,
                                                SolveCtl* __restrict__ ctl, Tables tab,
                                                StatePtrs<T> st, AirPtrs<T> air,
                                                const double* __restrict__ x0,
                                                const double* __restrict__ x1, StatePtrs<T> out,
                                                int fin, SolveHist* __restrict__ hist,
                                                SolveParams sp, const double* __restrict__ partR,
                                                long stride, double* __restrict__ chg_e,
                                                double* __restrict__ chg_g,
                                                T* __restrict__ gdel, T* __restrict__ gnext,
                                                double* __restrict__ acc) {
    __shared__ double red[16 * 4];
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0 || stopped(btime[tl.x])) return;
    const BubbleDev B = bub[tl.x];
    if (tl.y == 0 && tl.z == 0) {
        // first tile of the bubble: solve verdict, and the time advance that
        // the next step's advection applies
        if (fin) finalize_block(B, ctl + tl.x, hist + tl.x, sp, partR, stride, red);
        if (fin && threadIdx.x == 0 && (btime[tl.x].stop & 1) && !ctl[tl.x].converged)
            btime[tl.x].stop |= 2;
        if (threadIdx.x == 0 && btime[tl.x].advance) btime[tl.x].pending = 1;
        if (threadIdx.x == 0 && gdel) btime[tl.x].gpend = 1;
        // the next assembly adds its sums into a cleared slot
        if (threadIdx.x < 64) acc_zero(acc, tl.x, acc_slot_res(0));
    }
    const int sh = __builtin_ctz((unsigned)tl.w) - 1;   // tile width / 2 = 2^sh
    const int j = tl.y + ((int)threadIdx.x >> sh);
    const int i = tl.z + 2 * ((int)threadIdx.x & ((tl.w >> 1) - 1));
    if (j >= B.nth || i >= B.nph) return;
    const int nph = B.nph, m = B.mask, nth = B.nth;
    const bool south = j < nth - 1;
    const double* __restrict__ x = ctl[tl.x].xbuf ? x1 : x0;
    const size_t row = (size_t)B.off + (size_t)j * nph;
    const size_t c = row + i, cW = row + wrap_col(i - 1, nph, m);
    // the reference's divisions by dt as products with 1 / dt (an ulp)
    const double invDt = 1.0 / B.dt;
    const double tnow = btime[tl.x].t;
    const int ghist = btime[tl.x].ghist;
    const double kt = B.airflow ? 3.0 - fabs(tnow - 3.0) : 0.0;

    const Pair2<T> gam = ld2(st.gam + c), eta = ld2(st.eta + c), uph = ld2(st.uph + c);
    const Pair2<double> xc = ld2(x + c);
    UpdIn<T> in0{}, in1{};
    in0.ga = (double)gam.a; in1.ga = (double)gam.b;
    in0.g = xc.a; in1.g = xc.b;
    in0.gW = x[cW]; in1.gW = xc.a;
    in0.eta = (double)eta.a; in1.eta = (double)eta.b;
    in0.etaW = (double)st.eta[cW]; in1.etaW = (double)eta.a;
    in0.u1 = (double)uph.a; in1.u1 = (double)uph.b;
    if (south) {
        const Pair2<T> etaS = ld2(st.eta + c + nph), uth = ld2(st.uth + c);
        const Pair2<double> xS = ld2(x + c + nph);
        in0.gS = xS.a; in1.gS = xS.b;
        in0.etaS = (double)etaS.a; in1.etaS = (double)etaS.b;
        in0.v1 = (double)uth.a; in1.v1 = (double)uth.b;
    }
    if (B.airflow) {
        const Pair2<T> ua = ld2(air.ua0 + c), us = ld2(air.au + c);
        in0.ua0 = ua.a; in1.ua0 = ua.b; in0.uas = us.a; in1.uas = us.b;
        if (south) {
            const Pair2<T> va = ld2(air.va0 + c), vs = ld2(air.av + c);
            in0.va0 = va.a; in1.va0 = va.b; in0.vas = vs.a; in1.vas = vs.b;
        }
    }
    const bool guess = gdel != nullptr;
    if (guess) {
        const Pair2<T> gd = ld2(gdel + c);
        in0.gd = gd.a; in1.gd = gd.b;
    }
    T e0, e1, g0, g1, u0, u1, v0 = T(0), v1 = T(0), d0 = T(0), d1 = T(0), n0 = T(0), n1 = T(0);
    bool nan_out = update_cell(B, ghist, tab, air, x, j, i, c, in0, kt, tnow, invDt, e0, g0, u0,
                               v0, d0, n0, guess, chg_e, chg_g);
    nan_out |= update_cell(B, ghist, tab, air, x, j, i + 1, c + 1, in1, kt, tnow, invDt, e1, g1,
                           u1, v1, d1, n1, guess, chg_e, chg_g);
    st2(out.eta + c, e0, e1);
    st2(out.gam + c, g0, g1);
    st2(out.uph + c, u0, u1);
    if (south) st2(out.uth + c, v0, v1);
    if (guess) {
        st2(gdel + c, d0, d1);
        st2(gnext + c, n0, n1);
    }
    // the advection of the next step takes its NaN-skipping path from here on
    if (nan_out) btime[tl.x].nans = 1;
}

// currentTimeGlobal set by the host (bubble.cu:2662): the air-flow time of
// bubble b (all when b < 0), with no pending advance, in stream order.
__global__ void k_set_time(BubbleTime* __restrict__ bt, int nb, int b, double t) {
    for (int k = threadIdx.x; k < nb; k += blockDim.x)
        if (b < 0 || b == k) {
            bt[k].t = t;
            bt[k].pending = 0;
        }
}

void launch_set_time(hipStream_t s, BubbleTime* bt, int nb, int b, double t) {
    hipLaunchKernelGGL(k_set_time, dim3(1), dim3(64), 0, s, bt, nb, b, t);
}

// ------------------------------------------------------ frame output ----
// write_thickness_img (solver.cu:560-582) up to the file: the gray value
// eta H 2e5 of every cell, rounded to binary32 as the host writer rounds it
// (the products in FP64, then one rounding), then to binary16 as the
// reference's EXR writer converts (tinyexr float_to_half_full: truncate the
// significand, add one unit when the highest dropped bit is set; binary32
// subnormals to signed zero, overflow to infinity, NaN to a quiet NaN).
// Cell c of the bubble is pixel (x = i, y = j) of the nPhi x nTheta image.
__device__ __forceinline__ unsigned short half_of(float f) {
    const unsigned int x = __float_as_uint(f);
    const unsigned int sign = (x >> 16) & 0x8000u;
    const int e = (int)((x >> 23) & 0xFFu);
    const unsigned int mant = x & 0x7FFFFFu;
    if (e == 0) return (unsigned short)sign;
    if (e == 255) return (unsigned short)(sign | 0x7C00u | (mant ? 0x200u : 0u));
    const int ne = e - 127 + 15;
    if (ne >= 31) return (unsigned short)(sign | 0x7C00u);
    if (ne <= 0) {
        const int shift = 14 - ne;
        if (shift > 24) return (unsigned short)sign;
        const unsigned int m = mant | 0x800000u;
        unsigned int r = m >> shift;
        if ((m >> (shift - 1)) & 1u) ++r;
        return (unsigned short)(sign | r);
    }
    unsigned int r = ((unsigned int)ne << 10) | (mant >> 13);
    if (mant & 0x1000u) ++r;
    return (unsigned short)(sign | r);
}

template <typename T>
__global__ __launch_bounds__(256) void k_thickness_half(const T* __restrict__ eta, long n,
                                                        double H, unsigned short* __restrict__ out) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    out[k] = half_of((float)((double)eta[k] * H * 2e5));
}

// Marks a bubble whose state fields hold a NaN (the sticky flag that selects
// the NaN-skipping samplers), for state that reached the device without
// passing through the host upload's scan.
template <typename T>
__global__ __launch_bounds__(256) void k_nan_scan(const int4* __restrict__ tiles,
                                                  const BubbleDev* __restrict__ bub,
                                                  BubbleTime* __restrict__ btime, StatePtrs<T> st) {
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0) return;
    const BubbleDev B = bub[tl.x];
    int j, i;
    tile_cell(tl, j, i);
    if (j >= B.nth || i >= B.nph) return;
    const size_t c = (size_t)B.off + (size_t)j * B.nph + i;
    bool nan = tisnan(st.uph[c]) || tisnan(st.eta[c]) || tisnan(st.gam[c]);
    if (j < B.nth - 1) nan = nan || tisnan(st.uth[c]);
    if (nan) btime[tl.x].nans = 1;
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

// ------------------------------------------------------------ BIMOCQ ----
// The BIMOCQ scalar advection of eta (header.cuh:97-98 BIMOCQ and
// MULTISAMPLE): backward and forward maps between the current grid and the
// positions at the last reinitialisation, eta sampled through them from its
// initial field plus the changes accumulated since (bubble.cu:967-1013),
// error-corrected reinitialisation when the maps drift apart
// (bubble.cu:2509-2564). Gamma keeps the plain advection ("do not apply
// bimocq on concentration for stability", bubble.cu:1195) but carries its
// init/changes for the correction. One thread per cell, packed tiles.

// initMapping (solver.cu:12-22) and Init/InitLast = current, changes = 0
__device__ __forceinline__ void set_identity(const UMap& M, size_t c, double h, int j, int i) {
    const U3 u = unit_of_index(h, j + 0.5, (double)i);
    M.x[c] = u.x;
    M.y[c] = u.y;
    M.z[c] = u.z;
}

template <typename T>
__global__ __launch_bounds__(256) void k_bm_init(const int4* __restrict__ tiles,
                                                 const BubbleDev* __restrict__ bub,
                                                 StatePtrs<T> cur, BimocqPtrs P) {
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0) return;
    const BubbleDev B = bub[tl.x];
    int j, i;
    tile_cell(tl, j, i);
    if (j >= B.nth || i >= B.nph) return;
    const size_t c = (size_t)B.off + (size_t)j * B.nph + i;
    set_identity(P.fwd, c, B.h, j, i);
    set_identity(P.bwd, c, B.h, j, i);
    set_identity(P.prev, c, B.h, j, i);
    P.e_init[c] = P.e_initlast[c] = (double)cur.eta[c];
    P.g_init[c] = P.g_initlast[c] = (double)cur.gam[c];
    P.e_delta[c] = P.e_deltalast[c] = P.g_delta[c] = P.g_deltalast[c] = 0.0;
}

// traceGreatCircle from the cell centre (j + 1/2, i) over dt (either sign)
template <typename T, bool NANS>
__device__ __forceinline__ void trace_centre(const VelCtx<T>& v, const Tables& tab,
                                             const BubbleDev& B, int j, int i, double& tx,
                                             double& ty) {
    const T gx = (T)j + T(0.5), gy = (T)i;
    const T sinTc = TabSel<T>::sTc(tab)[B.rt0 + j], cosTc = TabSel<T>::cTc(tab)[B.rt0 + j];
    const T sinPc = TabSel<T>::sPc(tab)[B.ct0 + i], cosPc = TabSel<T>::cPc(tab)[B.ct0 + i];
    const T guT = sample_vtheta<T, NANS>(v.uth, B.nth, gx, gy);
    const T guP = sample_vphi<T, NANS>(v.uph, B.nth, gx, gy);
    Circle<T> o;
    TPos<T> tp;
    if (circle_mid<T, NANS>(v, gx, gy, sinTc, cosTc, sinPc, cosPc, guT, guP, o)) {
        T s, co, px, py, pz, rho, ir;
        circle_depart(v, o, s, co, px, py, pz, tp, rho, ir);
    } else {
        tp = o.start;
    }
    // the split position summed in FP64 (exact)
    tx = (double)tp.ix + (double)tp.fx;
    ty = (double)tp.iy + (double)tp.fy;
}

// updateForward / updateBackward (bubble.cu:1049-1118 with
// updateMappingKernel, bubble.cu:606-633): new map(x) = map(trace(x, -dt))
// for the forward map and map(trace(x, +dt)) for the backward map.
template <typename T, bool NANS>
__device__ __forceinline__ void bm_maps_cell(const BubbleDev& B, const Tables& tab,
                                             const StatePtrs<T>& cur, const BimocqPtrs& P, int j,
                                             int i) {
    VelCtx<T> v;
    v.uth = Field<T>{cur.uth + B.off, B.nph, B.mask, B.nth - 2};
    v.uph = Field<T>{cur.uph + B.off, B.nph, B.mask, B.nth - 1};
    v.nth = B.nth;
    v.invh = (T)B.invh;
    const size_t o = (size_t)B.off, c = o + (size_t)j * B.nph + i;
    for (int dir = 0; dir < 2; ++dir) {
        v.dt = dir == 0 ? (T)(-B.dt) : (T)B.dt;
        double tx, ty;
        trace_centre<T, NANS>(v, tab, B, j, i, tx, ty);
        const U3 s = sample_map_u(dir == 0 ? P.fwd : P.bwd, o, B.nth, B.nph, tx, ty);
        const UMap& out = dir == 0 ? P.nfwd : P.nbwd;
        out.x[c] = s.x;
        out.y[c] = s.y;
        out.z[c] = s.z;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bm_maps(const int4* __restrict__ tiles,
                                                 const BubbleDev* __restrict__ bub,
                                                 const BubbleTime* __restrict__ btime, Tables tab,
                                                 StatePtrs<T> cur, BimocqPtrs P) {
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0 || stopped_early(btime[tl.x])) return;
    const BubbleDev B = bub[tl.x];
    int j, i;
    tile_cell(tl, j, i);
    if (j >= B.nth || i >= B.nph) return;
    if (btime[tl.x].nans)
        bm_maps_cell<T, true>(B, tab, cur, P, j, i);
    else
        bm_maps_cell<T, false>(B, tab, cur, P, j, i);
}

// advectionCenteredBimocq (bubble.cu:967-1013) through the new backward map
// and the previous interval's backward map; overwrites the plain eta*.
template <typename T, bool NANS>
__device__ __forceinline__ void bm_eta_cell(const BubbleDev& B, StatePtrs<T> cur, StatePtrs<T> out,
                                            const BimocqPtrs& P, int j, int i) {
    const size_t o = (size_t)B.off, c = o + (size_t)j * B.nph + i;
    if ((double)cur.eta[c] == kThresh) {
        out.eta[c] = (T)kThresh;
        return;
    }
    auto F = [&](const double* a) { return Field<double>{a + o, B.nph, B.mask, B.nth - 1}; };
    const Field<double> fil = F(P.e_initlast), fin = F(P.e_init), fd = F(P.e_delta),
                        fdl = F(P.e_deltalast);
    const int ns = P.multisample ? 5 : 1;
    const double bl = P.blend;
    const CellRef cr = cell_ref(P.tab, B, j, i);
    double a = 0.0;
    for (int s = 0; s < ns; ++s) {
        const int q = ns == 5 ? s : 4;
        const double w = ns == 5 ? kMsW[q] : 1.0;
        const P2d ip = index_near(B.invh, sample_map_u(P.nbwd, o, B.nth, B.nph,
                                                        j + 0.5 + kMsDx[q], i + kMsDy[q]), cr);
        const P2d lp = index_near(B.invh, sample_map_u(P.prev, o, B.nth, B.nph, ip.x, ip.y), cr);
        a += (1.0 - bl) * w *
             (sample_centered<double, NANS>(fil, B.nth, lp.x, lp.y) +
              sample_centered<double, NANS>(fd, B.nth, ip.x, ip.y) +
              sample_centered<double, NANS>(fdl, B.nth, lp.x, lp.y));
        a += bl * w * (sample_centered<double, NANS>(fin, B.nth, ip.x, ip.y) +
                       sample_centered<double, NANS>(fd, B.nth, ip.x, ip.y));
    }
    out.eta[c] = (T)fmax(kThresh, a);
}

// The NaN-skipping samplers only once a NaN exists in the bubble (its
// sticky flag; the BIMOCQ fields derive from eta, which holds none before).
template <typename T>
__global__ __launch_bounds__(256) void k_bm_eta(const int4* __restrict__ tiles,
                                                const BubbleDev* __restrict__ bub,
                                                StatePtrs<T> cur, StatePtrs<T> out, BimocqPtrs P) {
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0) return;
    const BubbleDev B = bub[tl.x];
    int j, i;
    tile_cell(tl, j, i);
    if (j >= B.nth || i >= B.nph) return;
    if (P.bt[tl.x].nans)
        bm_eta_cell<T, true>(B, cur, out, P, j, i);
    else
        bm_eta_cell<T, false>(B, cur, out, P, j, i);
}

// accumulateChangesCentered (bubble.cu:2081-2114) for eta and Gamma through
// the new forward map, the new maps made current, and the distortion
// (estimateDistortion, bubble.cu:1133-1168) of the new maps: its block
// maximum goes to the bubble's running maximum.
__global__ __launch_bounds__(256) void k_bm_accum(const int4* __restrict__ tiles,
                                                  const BubbleDev* __restrict__ bub,
                                                  BimocqPtrs P) {
    __shared__ float red[8];
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0 || stopped(P.bt[tl.x])) return;
    const BubbleDev B = bub[tl.x];
    int j, i;
    tile_cell(tl, j, i);
    float dmax = 0.0f;
    if (j < B.nth && i < B.nph) {
        const size_t o = (size_t)B.off, c = o + (size_t)j * B.nph + i;
        const Field<double> ce{P.chg_e + o, B.nph, B.mask, B.nth - 1};
        const Field<double> cg{P.chg_g + o, B.nph, B.mask, B.nth - 1};
        const int ns = P.multisample ? 5 : 1;
        const CellRef cr = cell_ref(P.tab, B, j, i);
        // the changes derive from eta and Gamma: NaN-free while the bubble's
        // flag is clear (k_update sets it before this launch)
        const bool nans = P.bt[tl.x].nans != 0;
        double de = 0.0, dg = 0.0;
        for (int s = 0; s < ns; ++s) {
            const int q = ns == 5 ? s : 4;
            const double w = ns == 5 ? kMsW[q] : 1.0;
            const P2d ip = index_near(B.invh, sample_map_u(P.nfwd, o, B.nth, B.nph,
                                                            j + 0.5 + kMsDx[q], i + kMsDy[q]), cr);
            if (nans) {
                de += w * sample_centered<double, true>(ce, B.nth, ip.x, ip.y);
                dg += w * sample_centered<double, true>(cg, B.nth, ip.x, ip.y);
            } else {
                de += w * sample_centered<double, false>(ce, B.nth, ip.x, ip.y);
                dg += w * sample_centered<double, false>(cg, B.nth, ip.x, ip.y);
            }
        }
        P.e_delta[c] += de;
        P.g_delta[c] += dg;
        const U3 f{P.nfwd.x[c], P.nfwd.y[c], P.nfwd.z[c]};
        const U3 b{P.nbwd.x[c], P.nbwd.y[c], P.nbwd.z[c]};
        const P2d fi = index_near(B.invh, f, cr), bi = index_near(B.invh, b, cr);
        // the cell centre from the sin/cos tables
        const U3 x{cr.cP * cr.sT, cr.sP * cr.sT, cr.cT};
        const U3 a = sample_map_u(P.nbwd, o, B.nth, B.nph, fi.x, fi.y);
        const U3 e = sample_map_u(P.nfwd, o, B.nth, B.nph, bi.x, bi.y);
        // dist (bubble.cu:60-79): great-circle distance in cells
        const double da = B.invh * safe_acos_d(x.x * a.x + x.y * a.y + x.z * a.z);
        const double dd = B.invh * safe_acos_d(x.x * e.x + x.y * e.y + x.z * e.z);
        dmax = fmaxf((float)da, (float)dd);
        P.fwd.x[c] = f.x; P.fwd.y[c] = f.y; P.fwd.z[c] = f.z;
        P.bwd.x[c] = b.x; P.bwd.y[c] = b.y; P.bwd.z[c] = b.z;
    }
    dmax = wave_max(dmax);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = 0.0f;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = fmaxf(m, red[w]);
        // non-negative floats order as their bit patterns
        atomicMax(P.dist_bits + tl.x, __float_as_uint(m));
    }
}

// The reinitialisation decision of each bubble; clears the running maximum.
__global__ void k_bm_decide(int nb, BimocqPtrs P) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    if (stopped(P.bt[b])) {
        P.flag[b] = 0;
        return;
    }
    const float d = __uint_as_float(P.dist_bits[b]);
    P.dist_bits[b] = 0u;
    P.dist_last[b] = (double)d;
    const int f = (double)d > P.threshold ? 1 : 0;
    P.flag[b] = f;
    P.reinits[b] += f;
}

// correctCentered1 (bubble.cu:2180-2220): the error of the forward-mapped
// current fields against init + changes; keeps copies of the fields
template <typename T>
__global__ __launch_bounds__(256) void k_bm_correct1(const int4* __restrict__ tiles,
                                                     const BubbleDev* __restrict__ bub,
                                                     StatePtrs<T> cur, BimocqPtrs P) {
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0 || !P.flag[tl.x]) return;
    const BubbleDev B = bub[tl.x];
    int j, i;
    tile_cell(tl, j, i);
    if (j >= B.nth || i >= B.nph) return;
    const size_t o = (size_t)B.off, c = o + (size_t)j * B.nph + i;
    const Field<T> fe{cur.eta + o, B.nph, B.mask, B.nth - 1};
    const Field<T> fg{cur.gam + o, B.nph, B.mask, B.nth - 1};
    double te = 0.0, tg = 0.0;
    const CellRef cr = cell_ref(P.tab, B, j, i);
    for (int q = 0; q < 5; ++q) {
        const P2d ip = index_near(B.invh, sample_map_u(P.fwd, o, B.nth, B.nph,
                                                        j + 0.5 + kMsDx[q], i + kMsDy[q]), cr);
        te += kMsW[q] * (double)sample_centered<T, true>(fe, B.nth, (T)ip.x, (T)ip.y);
        tg += kMsW[q] * (double)sample_centered<T, true>(fg, B.nth, (T)ip.x, (T)ip.y);
    }
    P.err_e[c] = (te - P.e_delta[c] - P.e_init[c]) * 0.5;
    P.err_g[c] = (tg - P.g_delta[c] - P.g_init[c]) * 0.5;
    P.cpy_e[c] = (double)cur.eta[c];
    P.cpy_g[c] = (double)cur.gam[c];
}

// Bounds of the 3 x 3 neighbourhood of a cell in float, as the reference's
// fminf/fmaxf take them (bubble.cu:2248-2262); theta = -1 reflects across the
// pole like theta = nTheta (the reference's validateId leaves -1 negative).
__device__ __forceinline__ void local_bounds(const double* __restrict__ a, int nth, int nph, int j,
                                             int i, float& lo, float& hi) {
    lo = hi = (float)a[(size_t)j * nph + i];
    for (int t = -1; t <= 1; ++t)
        for (int s = -1; s <= 1; ++s) {
            int jj = j + t, ii = i + s;
            if (jj < 0 || jj >= nth) {
                jj = jj < 0 ? 0 : nth - 1;
                ii += nth;
            }
            const float v = (float)a[(size_t)jj * nph + wrap_col(ii, nph, (nph & (nph - 1)) ? -1 : nph - 1)];
            lo = fminf(lo, v);
            hi = fmaxf(hi, v);
        }
}

// correctCentered2 (bubble.cu:2223-2280): subtract the backward-mapped error,
// clamp to the local extrema of the uncorrected field (in float)
template <typename T>
__global__ __launch_bounds__(256) void k_bm_correct2(const int4* __restrict__ tiles,
                                                     const BubbleDev* __restrict__ bub,
                                                     StatePtrs<T> cur, BimocqPtrs P) {
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0 || !P.flag[tl.x]) return;
    const BubbleDev B = bub[tl.x];
    int j, i;
    tile_cell(tl, j, i);
    if (j >= B.nth || i >= B.nph) return;
    const size_t o = (size_t)B.off, c = o + (size_t)j * B.nph + i;
    const Field<double> ee{P.err_e + o, B.nph, B.mask, B.nth - 1};
    const Field<double> eg{P.err_g + o, B.nph, B.mask, B.nth - 1};
    double ve = P.cpy_e[c], vg = P.cpy_g[c];
    const CellRef cr = cell_ref(P.tab, B, j, i);
    for (int q = 0; q < 5; ++q) {
        const P2d ip = index_near(B.invh, sample_map_u(P.bwd, o, B.nth, B.nph,
                                                        j + 0.5 + kMsDx[q], i + kMsDy[q]), cr);
        ve -= kMsW[q] * sample_centered<double, true>(ee, B.nth, ip.x, ip.y);
        vg -= kMsW[q] * sample_centered<double, true>(eg, B.nth, ip.x, ip.y);
    }
    float lo, hi;
    local_bounds(P.cpy_e + o, B.nth, B.nph, j, i, lo, hi);
    cur.eta[c] = (T)fmax(kThresh, (double)fmaxf(lo, fminf((float)ve, hi)));
    local_bounds(P.cpy_g + o, B.nth, B.nph, j, i, lo, hi);
    cur.gam[c] = (T)fmax(kThresh, (double)fmaxf(lo, fminf((float)vg, hi)));
}

// The rest of reInitializeMapping (bubble.cu:2538-2563): the previous
// interval takes the current init, changes and backward map; init = the
// corrected fields, changes = 0, maps = identity.
template <typename T>
__global__ __launch_bounds__(256) void k_bm_finish(const int4* __restrict__ tiles,
                                                   const BubbleDev* __restrict__ bub,
                                                   StatePtrs<T> cur, BimocqPtrs P) {
    const int4 tl = tiles[blockIdx.x];
    if (tl.x < 0 || !P.flag[tl.x]) return;
    const BubbleDev B = bub[tl.x];
    int j, i;
    tile_cell(tl, j, i);
    if (j >= B.nth || i >= B.nph) return;
    const size_t c = (size_t)B.off + (size_t)j * B.nph + i;
    P.e_initlast[c] = P.e_init[c];
    P.e_deltalast[c] = P.e_delta[c];
    P.g_initlast[c] = P.g_init[c];
    P.g_deltalast[c] = P.g_delta[c];
    P.prev.x[c] = P.bwd.x[c];
    P.prev.y[c] = P.bwd.y[c];
    P.prev.z[c] = P.bwd.z[c];
    P.e_init[c] = (double)cur.eta[c];
    P.g_init[c] = (double)cur.gam[c];
    P.e_delta[c] = P.g_delta[c] = 0.0;
    set_identity(P.fwd, c, B.h, j, i);
    set_identity(P.bwd, c, B.h, j, i);
}

// ------------------------------------------------------------ launchers ----
template <typename T>
void launch_advect(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, BubbleTime* bt,
                   const Tables& tab, StatePtrs<T> in, StatePtrs<T> out) {
    if (g.pow2_rows)
        hipLaunchKernelGGL((k_advect<T, true>), dim3(g.ntiles), dim3(256), 0, s, g.tiles, bub, bt,
                           tab, in, out);
    else
        hipLaunchKernelGGL((k_advect<T, false>), dim3(g.ntiles), dim3(256), 0, s, g.tiles, bub, bt,
                           tab, in, out);
}

#define SF_CASES(CALL)                                 \
    switch (g.cpt) {                                   \
        case 2: CALL(2); break;                        \
        case 4: CALL(4); break;                        \
        case 8: CALL(8); break;                        \
        default: CALL(16); break;                      \
    }

template <typename T, typename Ti>
void launch_assemble(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub,
                     BubbleTime* bt, const Tables& tab, StatePtrs<T> st, AirPtrs<T> air,
                     SysPtrs<Ti> sys, double gmax_state, SolveCtl* ctl, double* part, long stride,
                     const T* guess, const SolveParams& sp) {
#define SF_ASM1(C, V, P, K)                                                                        \
    hipLaunchKernelGGL((k_assemble<T, Ti, C, V, P, K>), dim3(g.nrows), dim3(g.row_threads),        \
                       g.lds_bytes, s, g.rows, bub, bt, tab, st, air, sys, (T)gmax_state, ctl,     \
                       part, stride, guess, sp)
#define SF_ASM(C, V, P)                                                                            \
    do {                                                                                           \
        if (g.packed) SF_ASM1(C, V, P, true); else SF_ASM1(C, V, P, false);                        \
    } while (0)
#define SF_CALL(C)                                                                                 \
    do {                                                                                           \
        if (air.fu) {                                                                              \
            if (g.pow2_rows) SF_ASM(C, true, true); else SF_ASM(C, true, false);                   \
        } else {                                                                                   \
            if (g.pow2_rows) SF_ASM(C, false, true); else SF_ASM(C, false, false);                 \
        }                                                                                          \
    } while (0)
    SF_CASES(SF_CALL)
#undef SF_CALL
#undef SF_ASM
#undef SF_ASM1
}

template <typename T, typename Ti>
void launch_resid(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, const BubbleTime* bt,
                  const Tables& tab, StatePtrs<T> st, SolveCtl* ctl, SysPtrs<Ti> sys,
                  const T* guess, const double* xo, double* xn, int k, double* part, long stride,
                  const double* partB, const SolveParams& sp, SchedStat* sh, int check_prev) {
#define SF_RES1(C, F, P, K, Z)                                                                   \
    hipLaunchKernelGGL((k_resid<T, Ti, C, F, P, K, Z>), dim3(g.nrows), dim3(g.row_threads),       \
                       g.lds_bytes, s, g.rows, bub, bt, tab, st, ctl, sys, guess, xo, xn, k, part, \
                       stride, partB, sp, sh, check_prev)
#define SF_RES(C, F, P, Z)                                                                        \
    do {                                                                                          \
        if (g.packed) SF_RES1(C, F, P, true, Z); else SF_RES1(C, F, P, false, Z);                 \
    } while (0)
#define SF_CALL(C)                                                                                \
    do {                                                                                          \
        if (k == 1) {                                                                             \
            if (g.pow2_rows) SF_RES(C, true, true, false); else SF_RES(C, true, false, false);    \
        } else if (sys.e) {                                                                       \
            if (g.pow2_rows) SF_RES(C, false, true, false); else SF_RES(C, false, false, false);  \
        } else {                                                                                  \
            if (g.pow2_rows) SF_RES(C, false, true, true); else SF_RES(C, false, false, true);    \
        }                                                                                         \
    } while (0)
    SF_CASES(SF_CALL)
#undef SF_CALL
#undef SF_RES
#undef SF_RES1
}

template <typename Ti>
void launch_spmv(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                 SysPtrs<Ti> sys, const Ti* pold, Ti* pnew, int it, const SolveParams& sp,
                 const double* partR, const double* partB, double* partA, long stride) {
#define SF_CALL1(C, K)                                                                       \
    hipLaunchKernelGGL((k_spmv<Ti, C, K>), dim3(g.nrows), dim3(g.row_threads), 0, s, g.rows, \
                       bub, ctl, sys, pold, pnew, it, sp, partR, partB, partA, stride)
#define SF_CALL(C)                                                                           \
    do {                                                                                     \
        if (g.packed) SF_CALL1(C, true); else SF_CALL1(C, false);                            \
    } while (0)
    SF_CASES(SF_CALL)
#undef SF_CALL
#undef SF_CALL1
}

template <typename Ti>
void launch_cgstep(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                   SysPtrs<Ti> sys, const Ti* p, int it, const double* partA, double* partB,
                   long stride) {
#define SF_CALL1(C, K)                                                                            \
    hipLaunchKernelGGL((k_cgstep<Ti, C, K>), dim3(g.nrows), dim3(g.row_threads), g.lds_bytes, s, \
                       g.rows, bub, ctl, sys, p, it, partA, partB, stride)
#define SF_CALL(C)                                                                                \
    do {                                                                                          \
        if (g.packed) SF_CALL1(C, true); else SF_CALL1(C, false);                                 \
    } while (0)
    SF_CASES(SF_CALL)
#undef SF_CALL
#undef SF_CALL1
}

template <typename Ti>
void launch_cheb(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                 SysPtrs<Ti> sys, const Ti* ek, const Ti* ekm1, Ti* enext, int it, int kr,
                 const SolveParams& sp, double* partB, long stride) {
#define SF_CALL1(C, K)                                                                          \
    hipLaunchKernelGGL((k_cheb<Ti, C, K>), dim3(g.nrows), dim3(g.row_threads), g.lds_bytes, s, \
                       g.rows, bub, ctl, sys, ek, ekm1, enext, it, kr, sp, partB, stride)
#define SF_CALL(C)                                                                              \
    do {                                                                                        \
        if (g.packed) SF_CALL1(C, true); else SF_CALL1(C, false);                               \
    } while (0)
    SF_CASES(SF_CALL)
#undef SF_CALL
#undef SF_CALL1
}

template <typename T>
void launch_viscous(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, const Tables& tab,
                    StatePtrs<T> st, double* div, double* fu, double* fv) {
    hipLaunchKernelGGL(k_visc_div<T>, dim3(g.ntiles), dim3(256), 0, s, g.tiles, bub, tab, st, div);
    hipLaunchKernelGGL(k_visc_force<T>, dim3(g.ntiles), dim3(256), 0, s, g.tiles, bub, tab, st, div,
                       fu, fv);
}
template void launch_viscous<float>(hipStream_t, const LaunchGeom&, const BubbleDev*,
                                    const Tables&, StatePtrs<float>, double*, double*, double*);
template void launch_viscous<double>(hipStream_t, const LaunchGeom&, const BubbleDev*,
                                     const Tables&, StatePtrs<double>, double*, double*, double*);

template <typename Ti>
void launch_mg_coarsen(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, MgPtrs<Ti> fine,
                       MgPtrs<Ti> coarse) {
    hipLaunchKernelGGL(k_mg_coarsen<Ti>, dim3(g.nrows), dim3(256), 0, s, g.rows, bub, fine, coarse);
}

template <typename Ti>
void launch_mg_line(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, SolveCtl* ctl,
                    MgPtrs<Ti> lvl, MgPtrs<Ti> coarse, const Ti* r0, int level, int flags,
                    const int* rsteps, const SolveParams& sp, const double* partR,
                    double* rz_out, long stride) {
#define SF_CALL(C)                                                                                 \
    hipLaunchKernelGGL((k_mg_line<Ti, C>), dim3(g.nrows), dim3(g.row_threads), g.lds_bytes, s,     \
                       g.rows, bub, ctl, lvl, coarse, r0, level, flags, rsteps, sp, partR, rz_out, \
                       stride)
    SF_CASES(SF_CALL)
#undef SF_CALL
}

void launch_finalize(hipStream_t s, int nb, const BubbleDev* bub, SolveCtl* ctl, SolveHist* hist,
                     const SolveParams& sp, const double* partR, long stride) {
    hipLaunchKernelGGL(k_finalize, dim3(nb), dim3(256), 0, s, bub, ctl, hist, sp, partR, stride);
}

// The schedule statistics of a chunk of captured steps handed to the host:
// copied into the host's pinned slot and cleared for the next chunk (a
// kernel, so that the chunk's graph holds it).
__global__ __launch_bounds__(64) void k_sched_stats(SchedStat* __restrict__ d,
                                                    SchedStat* __restrict__ h, int nb) {
    for (int b = threadIdx.x; b < nb; b += 64) {
        h[b] = d[b];
        d[b] = SchedStat{0.0, 0, 0};
    }
}
void launch_sched_stats(hipStream_t s, SchedStat* d, SchedStat* h, int nb) {
    hipLaunchKernelGGL(k_sched_stats, dim3(1), dim3(64), 0, s, d, h, nb);
}

template <typename T>
void launch_update(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, BubbleTime* bt,
                   SolveCtl* ctl, const Tables& tab, StatePtrs<T> st, AirPtrs<T> air,
                   const double* x0, const double* x1, StatePtrs<T> out, int fin, SolveHist* hist,
                   const SolveParams& sp, const double* partR, long stride, double* chg_e,
                   double* chg_g, T* gdel, T* gnext, double* acc) {
    hipLaunchKernelGGL(k_update<T>, dim3(g.ntiles), dim3(128), 0, s, g.tiles, bub, bt, ctl, tab, st,
                       air, x0, x1, out, fin, hist, sp, partR, stride, chg_e, chg_g, gdel, gnext, acc);
}

template <typename T>
void launch_thickness_half(hipStream_t s, const T* eta, long n, double H, unsigned short* out) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_thickness_half<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, eta,
                       n, H, out);
}

template <typename T>
void launch_nan_scan(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, BubbleTime* bt,
                     StatePtrs<T> st) {
    hipLaunchKernelGGL(k_nan_scan<T>, dim3(g.ntiles), dim3(256), 0, s, g.tiles, bub, bt, st);
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
                                         BubbleTime*, const Tables&, StatePtrs<T>,                \
                                         AirPtrs<T>, SysPtrs<Ti>, double, SolveCtl*, double*,     \
                                         long, const T*, const SolveParams&);                     \
    template void launch_resid<T, Ti>(hipStream_t, const LaunchGeom&, const BubbleDev*,           \
                                      const BubbleTime*, const Tables&, StatePtrs<T>, SolveCtl*,  \
                                      SysPtrs<Ti>, const T*, const double*, double*, int,         \
                                      double*, long, const double*, const SolveParams&,           \
                                      SchedStat*, int);
SF_INST_SYS(float, float)
SF_INST_SYS(float, double)
SF_INST_SYS(double, double)
#define SF_INST_TI(Ti)                                                                            \
    template void launch_spmv<Ti>(hipStream_t, const LaunchGeom&, const BubbleDev*, SolveCtl*,    \
                                  SysPtrs<Ti>, const Ti*, Ti*, int, const SolveParams&,           \
                                  const double*, const double*, double*, long);                   \
    template void launch_cgstep<Ti>(hipStream_t, const LaunchGeom&, const BubbleDev*, SolveCtl*,  \
                                    SysPtrs<Ti>, const Ti*, int, const double*, double*, long);   \
    template void launch_cheb<Ti>(hipStream_t, const LaunchGeom&, const BubbleDev*, SolveCtl*,    \
                                  SysPtrs<Ti>, const Ti*, const Ti*, Ti*, int, int,               \
                                  const SolveParams&,                                             \
                                  double*, long);
SF_INST_TI(float)
SF_INST_TI(double)
#define SF_INST_MG(Ti)                                                                            \
    template void launch_mg_coarsen<Ti>(hipStream_t, const LaunchGeom&, const BubbleDev*,         \
                                        MgPtrs<Ti>, MgPtrs<Ti>);                                  \
    template void launch_mg_line<Ti>(hipStream_t, const LaunchGeom&, const BubbleDev*, SolveCtl*, \
                                     MgPtrs<Ti>, MgPtrs<Ti>, const Ti*, int, int, const int*,     \
                                     const SolveParams&, const double*, double*, long);
SF_INST_MG(float)
SF_INST_MG(double)
template void launch_update<float>(hipStream_t, const LaunchGeom&, const BubbleDev*, BubbleTime*,
                                   SolveCtl*, const Tables&, StatePtrs<float>, AirPtrs<float>,
                                   const double*, const double*, StatePtrs<float>, int,
                                   SolveHist*, const SolveParams&, const double*, long, double*,
                                   double*, float*, float*, double*);
template void launch_update<double>(hipStream_t, const LaunchGeom&, const BubbleDev*, BubbleTime*,
                                    SolveCtl*, const Tables&, StatePtrs<double>, AirPtrs<double>,
                                    const double*, const double*, StatePtrs<double>, int,
                                    SolveHist*, const SolveParams&, const double*, long, double*,
                                    double*, double*, double*, double*);
template void launch_thickness_half<float>(hipStream_t, const float*, long, double, unsigned short*);
template void launch_thickness_half<double>(hipStream_t, const double*, long, double,
                                            unsigned short*);
template void launch_nan_scan<float>(hipStream_t, const LaunchGeom&, const BubbleDev*,
                                     BubbleTime*, StatePtrs<float>);
template void launch_nan_scan<double>(hipStream_t, const LaunchGeom&, const BubbleDev*,
                                      BubbleTime*, StatePtrs<double>);
template void launch_row_stats<float>(hipStream_t, const LaunchGeom&, const BubbleDev*,
                                      const Tables&, const float*, double*, long);
template void launch_row_stats<double>(hipStream_t, const LaunchGeom&, const BubbleDev*,
                                       const Tables&, const double*, double*, long);


template <typename T>
void launch_bimocq_init(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub,
                        StatePtrs<T> cur, const BimocqPtrs& P) {
    hipLaunchKernelGGL(k_bm_init<T>, dim3(g.ntiles), dim3(256), 0, s, g.tiles, bub, cur, P);
}
template <typename T>
void launch_bimocq_maps(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub,
                        BubbleTime* bt, const Tables& tab, StatePtrs<T> cur, const BimocqPtrs& P) {
    hipLaunchKernelGGL(k_bm_maps<T>, dim3(g.ntiles), dim3(256), 0, s, g.tiles, bub, bt, tab, cur, P);
}
template <typename T>
void launch_bimocq_eta(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub,
                       StatePtrs<T> cur, StatePtrs<T> out, const BimocqPtrs& P) {
    hipLaunchKernelGGL(k_bm_eta<T>, dim3(g.ntiles), dim3(256), 0, s, g.tiles, bub, cur, out, P);
}
void launch_bimocq_accumulate(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub, int nb,
                              const BimocqPtrs& P) {
    hipLaunchKernelGGL(k_bm_accum, dim3(g.ntiles), dim3(256), 0, s, g.tiles, bub, P);
    hipLaunchKernelGGL(k_bm_decide, dim3((nb + 63) / 64), dim3(64), 0, s, nb, P);
}
template <typename T>
void launch_bimocq_reinit(hipStream_t s, const LaunchGeom& g, const BubbleDev* bub,
                          StatePtrs<T> cur, const BimocqPtrs& P) {
    hipLaunchKernelGGL(k_bm_correct1<T>, dim3(g.ntiles), dim3(256), 0, s, g.tiles, bub, cur, P);
    hipLaunchKernelGGL(k_bm_correct2<T>, dim3(g.ntiles), dim3(256), 0, s, g.tiles, bub, cur, P);
    hipLaunchKernelGGL(k_bm_finish<T>, dim3(g.ntiles), dim3(256), 0, s, g.tiles, bub, cur, P);
}
#define SF_INST_BM(T)                                                                             \
    template void launch_bimocq_init<T>(hipStream_t, const LaunchGeom&, const BubbleDev*,         \
                                        StatePtrs<T>, const BimocqPtrs&);                         \
    template void launch_bimocq_maps<T>(hipStream_t, const LaunchGeom&, const BubbleDev*,         \
                                        BubbleTime*, const Tables&, StatePtrs<T>,                 \
                                        const BimocqPtrs&);                                       \
    template void launch_bimocq_eta<T>(hipStream_t, const LaunchGeom&, const BubbleDev*,          \
                                       StatePtrs<T>, StatePtrs<T>, const BimocqPtrs&);            \
    template void launch_bimocq_reinit<T>(hipStream_t, const LaunchGeom&, const BubbleDev*,       \
                                          StatePtrs<T>, const BimocqPtrs&);
SF_INST_BM(float)
SF_INST_BM(double)

}  // namespace sfk
