// Device-side data structures and helpers of the soap-film step for gfx950.
//
// Layout in HBM (packed over all bubbles of a context):
//   every cell field is one array; bubble b owns cells [off, off + nth*nph)
//   row-major by theta with phi contiguous, as the reference stores them
//   (quantity.cu:20-29). u_theta uses rows 0..nth-2 of its block; row nth-1
//   is padding that stays zero.
//   Per-row tables (nth+1 slots per bubble, base rt0) and per-column tables
//   (nph slots, base ct0) hold sin/cos of the cell coordinates in FP64 and
//   FP32, so no kernel evaluates the trigonometry of its own grid point.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfk {

struct BubbleDev {
    int nth, nph, mask;          // nph = 2 nth; mask = nph - 1 when nph is a power of two, else -1
    int rt0, ct0, pr0;           // row-table, column-table and partial-sum bases
    long long off;               // first cell of this bubble in every cell field
    double h, invh, dt;          // grid spacing, 1/h, non-dimensional step
    double M, Cr, gs, gamma_max, eva;
    double re;                   // Re^-1 of the VISCOUS path
    int gravity, evaporation, airflow, viscous;
};

// Air-flow time of one bubble (seconds) and its step in seconds; kernels
// derive the corridor factor 3 - |t - 3| and the mirror blend t/12. The
// update kernel marks a pending advance (when `advance`), which the next
// step's advection kernel applies before the assembly reads the time.
struct BubbleTime {
    double t;
    double dt_s;
    int advance;
    int pending;
    int nans;   // sticky: set once a NaN has been written into the state fields
    // initial guess of the Gamma solve: solves whose correction Gamma^{n+1} -
    // Gamma* is stored (0..2; the host resets it when the state or dt
    // changes), and the increment the update kernel leaves for the next
    // advection to apply
    int ghist;
    int gpend;
    // stop on a failed solve (the reference's `broken`, bubble.cu:1500-1507,
    // after which its frame loop takes no further step): bit 0 the option,
    // bit 1 set by the update kernel when this bubble's solve missed its
    // tolerance, bit 2 set by the next step's assembly. Kernels before the
    // assembly skip a bubble with bits 0 and 1, those after it one with bit
    // 2, so each launch reads a bit that an earlier launch wrote.
    int stop;
};

__device__ __forceinline__ bool stopped_early(const BubbleTime& t) { return (t.stop & 3) == 3; }
__device__ __forceinline__ bool stopped(const BubbleTime& t) { return (t.stop & 4) != 0; }

// Per-bubble control block of the iterative solve. Every field is written by
// one kernel (the block of a row launch that arrives last, see arrive_last)
// and read only by later kernels of the same stream.
struct SolveCtl {
    double rz[2];        // r.z of inner iterations, slot = iteration parity
    double rr_in0;       // |r_in|^2 at the start of the current inner solve
    double rel_res;      // last relative (or absolute) residual, FP64
    double bb;           // |b|^2
    double mu;           // spectral bound of the line-preconditioned operator
    int outer_done;      // 1 once the FP64 residual met the tolerance
    int inner_done;      // 1 once the inner solve met its reduction
    int iters;           // inner iterations executed
    int outers;          // FP64 residual evaluations
    int xbuf;            // which x buffer holds the current iterate
    int converged;       // final verdict written by the finalize kernel
};

// Per-bubble history of the solves since creation (written by k_finalize).
struct SolveHist {
    int fails;
    int solves;
    double max_res;
};

// Statistics of the captured step for its inner-iteration schedule, since
// the host last cleared them: the largest tail ratio of the first
// refinement and the number of further refinements that ran.
struct SchedStat {
    double tail_ratio;
    int extra;
    int pad_;
};

struct Tables {
    // per-row, FP64: centres (j+1/2)h, faces j h (advection positions),
    // face sines of the assembly (pole faces exactly 0), csc of centres
    const double *sinTc, *cosTc, *sinTf, *cosTf, *sinFace, *cscTc;
    // per-row, FP32 copies
    const float *sinTc32, *cosTc32, *sinTf32, *cosTf32;
    // per-column: phi = i h and phi = (i - 1/2) h
    const double *sinPc, *cosPc, *sinPh, *cosPh;
    const float *sinPc32, *cosPc32, *sinPh32, *cosPh32;
};

template <typename T> struct TabSel;
template <> struct TabSel<double> {
    static __device__ __forceinline__ const double* sTc(const Tables& t) { return t.sinTc; }
    static __device__ __forceinline__ const double* cTc(const Tables& t) { return t.cosTc; }
    static __device__ __forceinline__ const double* sTf(const Tables& t) { return t.sinTf; }
    static __device__ __forceinline__ const double* cTf(const Tables& t) { return t.cosTf; }
    static __device__ __forceinline__ const double* sPc(const Tables& t) { return t.sinPc; }
    static __device__ __forceinline__ const double* cPc(const Tables& t) { return t.cosPc; }
    static __device__ __forceinline__ const double* sPh(const Tables& t) { return t.sinPh; }
    static __device__ __forceinline__ const double* cPh(const Tables& t) { return t.cosPh; }
};
template <> struct TabSel<float> {
    static __device__ __forceinline__ const float* sTc(const Tables& t) { return t.sinTc32; }
    static __device__ __forceinline__ const float* cTc(const Tables& t) { return t.cosTc32; }
    static __device__ __forceinline__ const float* sTf(const Tables& t) { return t.sinTf32; }
    static __device__ __forceinline__ const float* cTf(const Tables& t) { return t.cosTf32; }
    static __device__ __forceinline__ const float* sPc(const Tables& t) { return t.sinPc32; }
    static __device__ __forceinline__ const float* cPc(const Tables& t) { return t.cosPc32; }
    static __device__ __forceinline__ const float* sPh(const Tables& t) { return t.sinPh32; }
    static __device__ __forceinline__ const float* cPh(const Tables& t) { return t.cosPh32; }
};

// ---------------------------------------------------------------- math ----
__device__ __forceinline__ float tfloor(float x) { return floorf(x); }
__device__ __forceinline__ double tfloor(double x) { return floor(x); }
// FP32: the hardware square root (within an ulp) without the
// correctly-rounded refinement sequence sqrtf expands to
__device__ __forceinline__ float tsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double tsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ bool tisnan(float x) { return __builtin_isnan(x); }
__device__ __forceinline__ bool tisnan(double x) { return __builtin_isnan(x); }

// FP32 atan2 for the traced points: octant reduction to a = min/max in
// [0, 1] and atan(a) = a P(a^2), P of degree 8 fitted to atan(a)/a on [0, 1]
// by iteratively reweighted least squares on the relative error (Lawson's
// algorithm); maximum absolute error 1.2e-7 rad evaluated in FP32, i.e. a
// few 1e-5 of a cell at nTheta = 1024, below the FP32 resolution of the
// index-space position. FP64 keeps the library atan2.
__device__ __forceinline__ float tatan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float a = mx > 0.0f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
    const float s = a * a;
    float p = 0.002850157907232642f;
    p = p * s - 0.01606966182589531f;
    p = p * s + 0.0426931269466877f;
    p = p * s - 0.07504424452781677f;
    p = p * s + 0.10640992224216461f;
    p = p * s - 0.1420365869998932f;
    p = p * s + 0.1999262124300003f;
    p = p * s - 0.3333307206630707f;
    p = p * s + 1.0f;
    float r = a * p;
    if (ay > ax) r = 1.57079632679489662f - r;
    if (x < 0.0f) r = 3.14159265358979324f - r;
    return copysignf(r, y);
}
__device__ __forceinline__ double tatan2(double y, double x) { return atan2(y, x); }

// Normalisation of a direction: returns |v| and 1/|v| from |v|^2. FP32 uses
// the hardware reciprocal square root; FP64 the correctly rounded sqrt and
// division.
__device__ __forceinline__ void norm_inv(float s, float& n, float& inv) {
    inv = __builtin_amdgcn_rsqf(s);
    n = s > 0.0f ? s * inv : s;   // |0| = 0 as length() gives, not 0 * inf
}
__device__ __forceinline__ void norm_inv(double s, double& n, double& inv) {
    n = sqrt(s);
    inv = 1.0 / n;
}
// FP64 reciprocal of a positive finite y: the hardware estimate refined by
// two Newton steps r += r (1 - y r), each of which squares the relative
// error; the result is within an ulp or two of 1 / y.
__device__ __forceinline__ double rcp64(double y) {
    double r = __builtin_amdgcn_rcp(y);
    double e = fma(-y, r, 1.0);
    r = fma(r, e, r);
    e = fma(-y, r, 1.0);
    return fma(r, e, r);
}
// a / y through the refined reciprocal (within an ulp or two of the
// quotient, a quarter of the division's cost); a divisor of magnitude
// outside [1e-300, 1e300] (zero, denormal, huge, infinite or NaN) takes the
// division itself, so that 2 / 0 = inf and 0 / 0 = NaN come out as the
// reference's.
__device__ __forceinline__ double qdiv(double a, double y) {
    const double ay = fabs(y);
    if (ay >= 1e-300 && ay <= 1e300) return a * rcp64(y);
    return a / y;
}
__device__ __forceinline__ float trcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double trcp(double x) { return 1.0 / x; }

// sin/cos of the great-circle step. The step is a small angle (|u| dt), so
// FP32 uses the Taylor series to x^7 / x^6 below 1/16 rad (truncation below
// 2e-16 / 6e-15, under FP32 rounding); FP64 and large steps use sincos.
__device__ __forceinline__ void sincos_step(float x, float* s, float* c) {
    if (fabsf(x) < 0.0625f) {
        float x2 = x * x;
        *s = x * (1.0f + x2 * (-1.0f / 6.0f + x2 * (1.0f / 120.0f + x2 * (-1.0f / 5040.0f))));
        *c = 1.0f + x2 * (-0.5f + x2 * (1.0f / 24.0f + x2 * (-1.0f / 720.0f)));
    } else {
        sincosf(x, s, c);
    }
}
__device__ __forceinline__ void sincos_step(double x, double* s, double* c) { sincos(x, s, c); }

// bubbleLerp (bubble.cu:153-160): NaN operands are skipped. With NANS =
// false the fields are known to hold no NaN (a per-bubble flag that every
// writer of a NaN sets) and the same expression is evaluated without the
// checks.
template <typename T>
__device__ __forceinline__ T lerp_nan(T from, T to, T alpha) {
    if (tisnan(from)) return to;
    if (tisnan(to)) return from;
    return (T(1) - alpha) * from + alpha * to;
}
template <typename T, bool NANS>
__device__ __forceinline__ T lerp_s(T from, T to, T alpha) {
    if (NANS) return lerp_nan(from, to, alpha);
    return (T(1) - alpha) * from + alpha * to;
}

// validateCoord (bubble.cu:116-137). The phi wrap fmod(y + nph, nph) is done
// by exact conditional subtraction (y + nph lies in [0, 3 nph) for every
// point the trace produces), which gives the same value as fmod.
template <typename T>
__device__ __forceinline__ bool validate_coord(T& x, T& y, int nth, int nph) {
    bool ret = false;
    const T tn = (T)nth, tp = (T)nph;
    if (x >= tn) { x = tp - x; y += tn; ret = !ret; }
    if (x < T(0)) { x = -x; y += tn; ret = !ret; }
    T a = y + tp;
    if (a >= T(2) * tp) a -= T(2) * tp;
    else if (a >= tp) a -= tp;
    y = a;
    return ret;
}

// A phi column x in [-nph, 2 nph) wrapped into the row: one AND for a row
// whose length is a power of two (mask = nph - 1); otherwise (mask = -1) a
// conditional add or subtract, clamped into the row so that no column,
// however malformed, addresses memory outside it.
__device__ __forceinline__ int wrap_col(int x, int nph, int mask) {
    if (mask >= 0) return x & mask;
    const int r = x < 0 ? x + nph : (x >= nph ? x - nph : x);
    return min(max(r, 0), nph - 1);
}

// Cell (j, i) of thread threadIdx.x in a packed tile {bubble, first row,
// first column, width}; tile widths are powers of two (a row shorter than
// its tile leaves the tile's last columns idle).
__device__ __forceinline__ void tile_cell(const int4& tl, int& j, int& i) {
    const int sh = __builtin_ctz((unsigned)tl.w);
    j = tl.y + ((int)threadIdx.x >> sh);
    i = tl.z + ((int)threadIdx.x & (tl.w - 1));
}

// A field of one bubble. P2 = true: the row length is known to be a power
// of two (mask = nph - 1), so the phi wrap is one AND; otherwise wrap_col
// takes the row's mask at run time.
template <typename T, bool P2 = false>
struct Field {
    const T* __restrict__ p;   // bubble base
    int nph, mask, maxrow;
    __device__ __forceinline__ int col(int c) const {
        if constexpr (P2) return c & mask;
        else return wrap_col(c, nph, mask);
    }
    __device__ __forceinline__ bool pow2() const { return P2 || mask >= 0; }
    // rows are clamped into the field so that no position, however
    // malformed, addresses memory outside it
    __device__ __forceinline__ T at(int row, int c) const {
        row = min(max(row, 0), maxrow);
        return p[(unsigned)(row * nph + col(c))];
    }
    template <bool NANS>
    __device__ __forceinline__ T belt(int row, int lo, T a) const {
        row = min(max(row, 0), maxrow);
        const T* r = p + (unsigned)(row * nph);
        return lerp_s<T, NANS>(r[col(lo)], r[col(lo + 1)], a);
    }
    __device__ __forceinline__ T belt(int row, int lo, T a) const { return belt<true>(row, lo, a); }
};

// ------------------------------------------- FP32 interior fast path ----
// Away from the polar rows every sampler reduces to one bilinear
// interpolation between two adjacent rows with no reflection (validateCoord
// leaves theta alone, no pole branch applies). For FP32 fields without NaNs
// that case is evaluated directly: the phi wrap by one conditional add or
// subtract (y + nPhi stays unrounded where the reference's fmod(y + nPhi,
// nPhi) would round it at the magnitude of nPhi), byte offsets into a
// buffer resource, and the next row through the scalar offset. A buffer
// load outside the field returns 0, so no position, however malformed, can
// read outside it.
struct FastIdx {
    int off, off1;   // byte offsets of the lower row's two phi neighbours
    float aT, aP;
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const float* p, int rows, int nph) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, rows * nph * 4, 0x00020000);
}

__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, int off, int soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
}

// x, y: sampler-relative index coordinates, x in [0, rows - 1). The phi
// wrap acts on the integer column (the fast path serves rows whose length
// is a power of two), so the weight y - floor(y) keeps every bit of y, also
// for y < 0.
__device__ __forceinline__ FastIdx fast_index(float x, float y, int nph, int mask) {
    const float fx = floorf(x), fy = floorf(y);
    FastIdx q;
    q.aT = x - fx;
    q.aP = y - fy;
    const int row = (int)fx * nph, pI = (int)fy;
    q.off = (row + (pI & mask)) * 4;
    q.off1 = (row + ((pI + 1) & mask)) * 4;
    return q;
}

// The three interpolations as v0 + a (v1 - v0): two operations each instead
// of the reference's three for (1 - a) v0 + a v1, the same value within an
// FP32 rounding.
__device__ __forceinline__ float fast_bilerp(__amdgpu_buffer_rsrc_t r, const FastIdx& q, int row_bytes) {
    const float v00 = bload(r, q.off, 0), v01 = bload(r, q.off1, 0);
    const float v10 = bload(r, q.off, row_bytes), v11 = bload(r, q.off1, row_bytes);
    const float lo = fmaf(q.aP, v01 - v00, v00);
    const float hi = fmaf(q.aP, v11 - v10, v10);
    return fmaf(q.aT, hi - lo, lo);
}

// A traced position in index coordinates, x = ix + fx, y = iy + fy. FP32
// traces keep the start point's integer index apart from the fraction,
// which holds the start point's half index and the increment turned along
// the circle: the interpolation weights then carry the increment's
// precision instead of the ulp of the absolute index (6e-5 of a cell at
// index 1000). FP64 traces hold the absolute position in fx, fy (ix = iy =
// 0).
template <typename T>
struct TPos {
    int ix, iy;
    T fx, fy;
    __device__ __forceinline__ T x() const { return (T)ix + fx; }
    __device__ __forceinline__ T y() const { return (T)iy + fy; }
};

// fast_index of the point p + (ox, oy) (ox, oy half-integers, added to the
// fraction exactly): weights from the fraction, the row and column from the
// integer parts. True when the lower row lies in [0, rows - 2], the case
// the fast path serves (rows: the field's row count).
__device__ __forceinline__ bool fast_index_rel(const TPos<float>& p, float ox, float oy, int rows,
                                               int nph, int mask, FastIdx& q) {
    const float rx = p.fx + ox, ry = p.fy + oy;
    const float fx = floorf(rx), fy = floorf(ry);
    q.aT = rx - fx;
    q.aP = ry - fy;
    const int tI = p.ix + (int)fx, pI = p.iy + (int)fy;
    const int row = tI * nph;
    q.off = (row + (pI & mask)) * 4;
    q.off1 = (row + ((pI + 1) & mask)) * 4;
    return (unsigned)tI < (unsigned)(rows - 1);
}

template <typename T, bool NANS>
struct FastPath { static constexpr bool on = false; };
template <>
struct FastPath<float, false> { static constexpr bool on = true; };

// Split a sample position into indices and weights; the phi index is
// clamped into range so a NaN position cannot address memory outside the
// field.
template <typename T>
__device__ __forceinline__ void split_index(T x, T y, int nph, int& tI, int& pI, T& aT, T& aP) {
    T fy = tfloor(y), fx = tfloor(x);
    aP = y - fy;
    aT = x - fx;
    pI = (int)fy;
    tI = (int)fx;
    if ((unsigned)pI >= (unsigned)nph) pI = pI < 0 ? 0 : nph - 1;
}

// sampleCentered (bubble.cu:320-376)
template <typename T, bool NANS = true, bool P2 = false, bool FAST = true>
__device__ T sample_centered(const Field<T, P2>& f, int nth, T rx, T ry) {
    if constexpr (FAST && FastPath<T, NANS>::on) {
        if (rx >= 0.5f && rx < (float)nth - 0.5f && f.pow2())
            return fast_bilerp(field_rsrc(f.p, nth, f.nph), fast_index(rx - 0.5f, ry, f.nph, f.mask),
                               f.nph * 4);
    }
    if constexpr (sizeof(T) == 8 && !NANS) {
        // FP64 fields without NaNs, away from the polar rows: the bilinear
        // interpolation of rows floor(x) and floor(x) + 1 that the general
        // branch below reduces to, the phi wrap on the integer column
        if (rx >= 0.5 && rx < (double)nth - 0.5 && f.pow2()) {
            const T x = rx - T(0.5), fx = tfloor(x), fy = tfloor(ry);
            const T aT = x - fx, aP = ry - fy;
            const int pI = (int)fy;
            const T* r0 = f.p + (size_t)(int)fx * f.nph;
            const T* r1 = r0 + f.nph;
            const int c0 = f.col(pI), c1 = f.col(pI + 1);
            const T lower = (T(1) - aP) * r0[c0] + aP * r0[c1];
            const T higher = (T(1) - aP) * r1[c0] + aP * r1[c1];
            return (T(1) - aT) * lower + aT * higher;
        }
    }
    T x = rx - T(0.5), y = ry;
    const int nph = f.nph;
    bool fl = validate_coord(x, y, nth, nph);
    int tI, pI; T aT, aP;
    split_index(x, y, nph, tI, pI, aT, aP);
    if (tI == 0 && fl) {
        T higher = f.template belt<NANS>(0, pI, aP);
        T lower = f.template belt<NANS>(0, pI + nph / 2, aP);
        return lerp_s<T, NANS>(lower, higher, aT);
    }
    if (fl) tI -= 1;
    if (tI == nth - 1) {
        T lower = f.template belt<NANS>(tI, pI, aP);
        T higher = f.template belt<NANS>(tI, pI + nph / 2, aP);
        return lerp_s<T, NANS>(lower, higher, aT);
    }
    T lower = f.template belt<NANS>(tI, pI, aP);
    T higher = f.template belt<NANS>(tI + 1, pI, aP);
    return lerp_s<T, NANS>(lower, higher, aT);
}

// sampleCentered of two fields at one point (the eta and Gamma samples of
// advectionAllCentered, bubble.cu:957-960) sharing the index computation.
template <typename T, bool NANS = true, bool P2 = false, bool FAST = true>
__device__ void sample_centered2(const Field<T, P2>& f, const Field<T, P2>& g, int nth, T rx, T ry, T& of,
                                 T& og) {
    if constexpr (FAST && FastPath<T, NANS>::on) {
        if (rx >= 0.5f && rx < (float)nth - 0.5f && f.pow2()) {
            const FastIdx q = fast_index(rx - 0.5f, ry, f.nph, f.mask);
            of = fast_bilerp(field_rsrc(f.p, nth, f.nph), q, f.nph * 4);
            og = fast_bilerp(field_rsrc(g.p, nth, g.nph), q, g.nph * 4);
            return;
        }
    }
    T x = rx - T(0.5), y = ry;
    const int nph = f.nph;
    bool fl = validate_coord(x, y, nth, nph);
    int tI, pI; T aT, aP;
    split_index(x, y, nph, tI, pI, aT, aP);
    int r0 = tI, r1 = tI + 1, c0 = pI, c1 = pI;
    if (tI == 0 && fl) {
        r0 = 0; r1 = 0; c0 = pI + nph / 2;
    } else {
        if (fl) tI -= 1;
        if (tI == nth - 1) { r0 = tI; r1 = tI; c1 = pI + nph / 2; }
        else { r0 = tI; r1 = tI + 1; }
    }
    of = lerp_s<T, NANS>(f.template belt<NANS>(r0, c0, aP), f.template belt<NANS>(r1, c1, aP), aT);
    og = lerp_s<T, NANS>(g.template belt<NANS>(r0, c0, aP), g.template belt<NANS>(r1, c1, aP), aT);
}

// sampleVPhi (bubble.cu:178-240): values taken across the pole flip sign.
template <typename T, bool NANS = true, bool P2 = false, bool FAST = true>
__device__ T sample_vphi(const Field<T, P2>& f, int nth, T rx, T ry) {
    if constexpr (FAST && FastPath<T, NANS>::on) {
        if (rx >= 0.5f && rx < (float)nth - 0.5f && f.pow2())
            return fast_bilerp(field_rsrc(f.p, nth, f.nph),
                               fast_index(rx - 0.5f, ry + 0.5f, f.nph, f.mask), f.nph * 4);
    }
    T x = rx - T(0.5), y = ry + T(0.5);
    const int nph = f.nph;
    bool fl = validate_coord(x, y, nth, nph);
    int tI, pI; T aT, aP;
    split_index(x, y, nph, tI, pI, aT, aP);
    if (tI == 0 && fl) {
        T higher = -f.template belt<NANS>(0, pI, aP);
        T lower = f.template belt<NANS>(0, pI + nph / 2, aP);
        return lerp_s<T, NANS>(lower, higher, aT);
    }
    if (fl) tI -= 1;
    if (tI == nth - 1) {
        T lower = f.template belt<NANS>(tI, pI, aP);
        T higher = -f.template belt<NANS>(tI, pI + nph / 2, aP);
        T v = lerp_s<T, NANS>(lower, higher, aT);
        return fl ? -v : v;
    }
    T lower = f.template belt<NANS>(tI, pI, aP);
    T higher = f.template belt<NANS>(tI + 1, pI, aP);
    T v = lerp_s<T, NANS>(lower, higher, aT);
    return fl ? -v : v;
}

// sampleVTheta (bubble.cu:247-313): u_theta has nth-1 rows; the polar caps
// interpolate between the first/last row and its phi+pi mirror with half
// weight. A sample exactly at theta = pi (which the reference reads past
// the array for) takes the south-cap branch.
template <typename T, bool NANS = true, bool P2 = false, bool FAST = true>
__device__ T sample_vtheta(const Field<T, P2>& f, int nth, T rx, T ry) {
    if constexpr (FAST && FastPath<T, NANS>::on) {
        if (rx >= 1.0f && rx < (float)nth - 1.0f && f.pow2())
            return fast_bilerp(field_rsrc(f.p, nth - 1, f.nph),
                               fast_index(rx - 1.0f, ry, f.nph, f.mask), f.nph * 4);
    }
    T x = rx - T(1), y = ry;
    const int nph = f.nph;
    bool fl = validate_coord(x, y, nth, nph);
    int tI, pI; T aT, aP;
    split_index(x, y, nph, tI, pI, aT, aP);
    const T tn = (T)nth;
    if ((rx < T(0) && rx > T(-1)) || (rx > tn && rx < tn + T(1))) {
        tI -= 1;
        aT += T(1);
    } else if (rx >= tn + T(1) || rx <= T(-1)) {
        tI -= 2;
    }
    if (tI == 0 && fl && rx > T(-1)) {
        T higher = -f.template belt<NANS>(0, pI, aP);
        T lower = f.template belt<NANS>(0, pI + nph / 2, aP);
        return lerp_s<T, NANS>(lower, higher, T(0.5) * aT);
    }
    if (tI >= nth - 1 && !fl) { tI = nth - 2; aT = T(1); }
    if (tI == nth - 2) {
        T lower = f.template belt<NANS>(tI, pI, aP);
        T higher = -f.template belt<NANS>(tI, pI + nph / 2, aP);
        T v = lerp_s<T, NANS>(lower, higher, T(0.5) * aT);
        return fl ? -v : v;
    }
    T lower = f.template belt<NANS>(tI, pI, aP);
    T higher = f.template belt<NANS>(tI + 1, pI, aP);
    T v = lerp_s<T, NANS>(lower, higher, aT);
    return fl ? -v : v;
}

// The samplers at a traced position: the FP32 interior fast path from the
// split position, else the general sampler at the absolute one.
template <typename T, bool NANS = true, bool P2 = false>
__device__ __forceinline__ T sample_vtheta(const Field<T, P2>& f, int nth, const TPos<T>& p) {
    if constexpr (FastPath<T, NANS>::on) {
        FastIdx q;
        if (f.pow2() && fast_index_rel(p, -1.0f, 0.0f, nth - 1, f.nph, f.mask, q))
            return fast_bilerp(field_rsrc(f.p, nth - 1, f.nph), q, f.nph * 4);
    }
    return sample_vtheta<T, NANS, P2, false>(f, nth, p.x(), p.y());
}
template <typename T, bool NANS = true, bool P2 = false>
__device__ __forceinline__ T sample_vphi(const Field<T, P2>& f, int nth, const TPos<T>& p) {
    if constexpr (FastPath<T, NANS>::on) {
        FastIdx q;
        if (f.pow2() && fast_index_rel(p, -0.5f, 0.5f, nth, f.nph, f.mask, q))
            return fast_bilerp(field_rsrc(f.p, nth, f.nph), q, f.nph * 4);
    }
    return sample_vphi<T, NANS, P2, false>(f, nth, p.x(), p.y());
}
template <typename T, bool NANS = true, bool P2 = false>
__device__ __forceinline__ void sample_centered2(const Field<T, P2>& f, const Field<T, P2>& g, int nth,
                                                 const TPos<T>& p, T& of, T& og) {
    if constexpr (FastPath<T, NANS>::on) {
        FastIdx q;
        if (f.pow2() && fast_index_rel(p, -0.5f, 0.0f, nth, f.nph, f.mask, q)) {
            of = fast_bilerp(field_rsrc(f.p, nth, f.nph), q, f.nph * 4);
            og = fast_bilerp(field_rsrc(g.p, nth, g.nph), q, g.nph * 4);
            return;
        }
    }
    sample_centered2<T, NANS, P2, false>(f, g, nth, p.x(), p.y(), of, og);
}

// ------------------------------------------------------ great circle ----
// The start frame (w, u_, v_) and speed after the transported midpoint
// velocity, i.e. the first two stages shared by traceGreatCircle
// (bubble.cu:416-496) and the vector advection kernels (bubble.cu:651-731,
// 787-867). Positions on the sphere are kept Cartesian: the sin/cos of a
// traced point's theta and phi are its z, rho, x/rho and y/rho, which is
// the reference's cos(acos(z)), sin(acos(z)), cos(atan2(y,x)), sin(atan2(y,x))
// without the round trip.
template <typename T>
struct Circle {
    T w0, w1, w2, u0, u1, u2, v0, v1, v2, unorm;
    T gx, gy, sT, cT, sP, cP;   // start point: index coordinates and trig
    TPos<T> start;              // start point, split as the traced positions
};

template <typename T, bool P2 = false>
struct VelCtx {
    Field<T, P2> uth, uph;
    int nth;
    T invh, dt;
};

// The small angle a = atan2(n, d) between a traced point and its start
// point, given (n, d) = k (sin a, cos a) and k_inv = 1 / k: with s = n k_inv
// = sin a, the odd series s + s^3/6 + 3 s^5/40 of asin(s) below |s| = 1/10
// (truncation under 5e-9 rad; no division, the scale k is 1 or the 1/rho
// that the tangent frame needs anyway), the general atan2 otherwise (a
// large step, or the phi change of a trace that passes close to a pole).
__device__ __forceinline__ float small_angle(float n, float d, float k_inv) {
    const float sn = n * k_inv, cs = d * k_inv;
    if (cs > 0.0f && fabsf(sn) < 0.1f) {
        const float s2 = sn * sn;
        return sn * (1.0f + s2 * (1.0f / 6.0f + s2 * (3.0f / 40.0f)));
    }
    return tatan2(n, d);
}

// Index-space coordinates of a traced point (x, y, z), rho = sin(theta) and
// ir = 1 / rho (used where rho > 0). FP64: theta = atan2(rho, z), phi =
// atan2(y, x). FP32: theta and phi as the start point's plus the angles
// turned from it, theta - theta0 = atan2(rho z0 - z rho0, z z0 + rho rho0) in
// the meridian half-plane (a unit vector: scale 1) and phi - phi0 =
// atan2(y x0 - x y0, x x0 + y y0) (x0, y0 the start point's cos phi0,
// sin phi0; scale rho): the same angles, with the grid-exact start index
// carrying the magnitude so FP32 keeps the small increment's precision.
template <typename T>
__device__ __forceinline__ void point_index(const Circle<T>& o, T x, T y, T z, T invh, TPos<T>& p,
                                            T& rho, T& ir);
template <>
__device__ __forceinline__ void point_index(const Circle<double>& o, double x, double y, double z,
                                            double invh, TPos<double>& p, double& rho, double& ir) {
    rho = sqrt(x * x + y * y);
    ir = 1.0 / rho;
    p.ix = p.iy = 0;
    p.fx = atan2(rho, z) * invh;
    p.fy = atan2(y, x) * invh;
}
template <>
__device__ __forceinline__ void point_index(const Circle<float>& o, float x, float y, float z,
                                            float invh, TPos<float>& p, float& rho, float& ir) {
    const float r2 = x * x + y * y;
    ir = __builtin_amdgcn_rsqf(r2);
    rho = r2 > 0.0f ? r2 * ir : 0.0f;
    p.ix = o.start.ix;
    p.iy = o.start.iy;
    p.fx = o.start.fx + small_angle(rho * o.cT - z * o.sT, z * o.cT + rho * o.sT, 1.0f) * invh;
    p.fy = o.start.fy + small_angle(y * o.cP - x * o.sP, x * o.cP + y * o.sP, ir) * invh;
}

// The start point (index coordinates gx, gy: multiples of 1/2) in the split
// form of the traced positions.
template <typename T>
__device__ __forceinline__ TPos<T> start_pos(T gx, T gy) {
    TPos<T> p;
    if constexpr (sizeof(T) == 4) {
        const T fx = tfloor(gx), fy = tfloor(gy);
        p.ix = (int)fx;
        p.iy = (int)fy;
        p.fx = gx - fx;
        p.fy = gy - fy;
    } else {
        p.ix = p.iy = 0;
        p.fx = gx;
        p.fy = gy;
    }
    return p;
}

// The components along uc and along vc = q x uc of the tangent vector with
// (e_theta, e_phi) components (vT, vP) at the unit point q = (x, y, z)
// (rho = sin theta, ir = 1 / rho), uc a unit tangent at q. (e_theta,
// e_phi, q) is right-handed, so vc = (e_theta . uc) e_phi - (e_phi . uc)
// e_theta and the two components are a plane rotation of (vT, vP): the
// reference's dot products with the Cartesian vector (bubble.cu:476-487,
// 754-765) without forming it. At a pole (rho = 0) the frame of phi = 0.
template <typename T>
__device__ __forceinline__ void circle_components(T x, T y, T z, T rho, T ir, T vT, T vP, T uc0,
                                                  T uc1, T uc2, T& along, T& across) {
    T cp = T(1), sp = T(0);
    if (rho > T(0)) { cp = x * ir; sp = y * ir; }
    const T ct = z * (cp * uc0 + sp * uc1) - rho * uc2;   // e_theta . uc
    const T cf = cp * uc1 - sp * uc0;                       // e_phi . uc
    along = vT * ct + vP * cf;
    across = vP * ct - vT * cf;
}

template <typename T, bool NANS = true, bool P2 = false>
__device__ __forceinline__ bool circle_mid(const VelCtx<T, P2>& c, T gx, T gy, T sinT, T cosT, T sinP,
                                           T cosP, T guT, T guP, Circle<T>& o) {
    o.gx = gx; o.gy = gy; o.sT = sinT; o.cT = cosT; o.sP = sinP; o.cP = cosP;
    o.start = start_pos(gx, gy);
    T u0 = guT * (cosT * cosP) - guP * sinP;
    T u1 = guT * (cosT * sinP) + guP * cosP;
    T u2 = -guT * sinT;
    o.w0 = cosP * sinT; o.w1 = sinP * sinT; o.w2 = cosT;
    T inv;
    norm_inv(u0 * u0 + u1 * u1 + u2 * u2, o.unorm, inv);
    if (!(o.unorm > T(0))) return false;
    o.u0 = u0 * inv; o.u1 = u1 * inv; o.u2 = u2 * inv;
    T s, co;
    sincos_step(T(-0.5) * o.unorm * c.dt, &s, &co);
    T mx = o.u0 * s + o.w0 * co, my = o.u1 * s + o.w1 * co, mz = o.u2 * s + o.w2 * co;
    TPos<T> mp;
    T rho, ir;
    point_index(o, mx, my, mz, c.invh, mp, rho, ir);
    T muT = sample_vtheta<T, NANS>(c.uth, c.nth, mp);
    T muP = sample_vphi<T, NANS>(c.uph, c.nth, mp);
    // the midpoint velocity along the circle (mguP) and across it (mguT)
    T uc0 = o.u0 * co - o.w0 * s, uc1 = o.u1 * co - o.w1 * s, uc2 = o.u2 * co - o.w2 * s;
    T mguP, mguT;
    circle_components(mx, my, mz, rho, ir, muT, muP, uc0, uc1, uc2, mguP, mguT);
    // v_ of the start frame = w x u_
    T v0 = o.w1 * o.u2 - o.w2 * o.u1, v1 = o.w2 * o.u0 - o.w0 * o.u2, v2 = o.w0 * o.u1 - o.w1 * o.u0;
    u0 = mguP * o.u0 + mguT * v0;
    u1 = mguP * o.u1 + mguT * v1;
    u2 = mguP * o.u2 + mguT * v2;
    norm_inv(u0 * u0 + u1 * u1 + u2 * u2, o.unorm, inv);
    if (!(o.unorm > T(0))) return false;
    o.u0 = u0 * inv; o.u1 = u1 * inv; o.u2 = u2 * inv;
    o.v0 = o.w1 * o.u2 - o.w2 * o.u1;
    o.v1 = o.w2 * o.u0 - o.w0 * o.u2;
    o.v2 = o.w0 * o.u1 - o.w1 * o.u0;
    return true;
}

// Departure point of the full step: returns its index-space coordinates and
// the Cartesian point (bubble.cu:498-505, 734-743).
template <typename T, bool P2>
__device__ __forceinline__ void circle_depart(const VelCtx<T, P2>& c, const Circle<T>& o, T& s, T& co,
                                              T& px, T& py, T& pz, TPos<T>& pos, T& rho, T& ir) {
    sincos_step(-o.unorm * c.dt, &s, &co);
    px = o.u0 * s + o.w0 * co;
    py = o.u1 * s + o.w1 * co;
    pz = o.u2 * s + o.w2 * co;
    point_index(o, px, py, pz, c.invh, pos, rho, ir);
}

// Velocity at the departure point transported back along the circle into
// the start frame (bubble.cu:745-766, 881-902).
template <typename T, bool NANS = true, bool P2 = false>
__device__ __forceinline__ void transport_back(const VelCtx<T, P2>& c, const Circle<T>& o, T& r0, T& r1,
                                               T& r2) {
    T s, co, px, py, pz, rho, ir;
    TPos<T> pp;
    circle_depart(c, o, s, co, px, py, pz, pp, rho, ir);
    T puT = sample_vtheta<T, NANS>(c.uth, c.nth, pp);
    T puP = sample_vphi<T, NANS>(c.uph, c.nth, pp);
    T uc0 = o.u0 * co - o.w0 * s, uc1 = o.u1 * co - o.w1 * s, uc2 = o.u2 * co - o.w2 * s;
    T b, a;
    circle_components(px, py, pz, rho, ir, puT, puP, uc0, uc1, uc2, b, a);
    r0 = b * o.u0 + a * o.v0;
    r1 = b * o.u1 + a * o.v1;
    r2 = b * o.u2 + a * o.v2;
}

// ------------------------------------------------------------ BIMOCQ ----
// Mapping arithmetic of the BIMOCQ scalar advection, FP64 in every precision
// mode. The reference keeps a map as index-space (theta, phi) pairs and
// converts every corner to a unit vector inside each spherical
// interpolation (lerpCoords, bubble.cu:512-536), and the result back. Here a
// map holds the unit vectors themselves: the interpolations chain on unit
// vectors, and only a position that samples a field is converted to index
// coordinates (acos of z, atan2 of y, x, as lerpCoords ends). The same
// arithmetic up to rounding, without two sin/cos pairs and an acos/atan2 per
// interpolation.
struct U3 {
    double x, y, z;
};
struct UMap {
    double *x, *y, *z;
};

__device__ __forceinline__ double safe_acos_d(double x) { return acos(fmax(-1.0, fmin(1.0, x))); }

__device__ __forceinline__ U3 unit_of_index(double h, double t, double f) {
    double st, ct, sf, cf;
    sincos(t * h, &st, &ct);
    sincos(f * h, &sf, &cf);
    return U3{cf * st, sf * st, ct};
}

struct P2d {
    double x, y;   // (theta, phi) index coordinates
};

// Index coordinates (acos z, atan2(y, x)) / h of a unit vector a few cells
// from the centre of cell (j, i): the
// centre's index coordinates (j + 1/2, i) plus the angles turned from it,
// each the asin of a sine that the (sin, cos) pair of the meridian half-plane
// (a unit vector) or the equatorial plane (scale rho) gives, by its series
// below 0.02 (truncation under 1e-22 rad); farther points take acos and
// atan2. The same angles up to rounding, phi in the centre's period rather
// than atan2's (-pi, pi] (the samplers wrap phi either way).
struct CellRef {
    double gx, gy, sT, cT, sP, cP;
};
__device__ __forceinline__ CellRef cell_ref(const Tables& tab, const BubbleDev& B, int j, int i) {
    return CellRef{j + 0.5, (double)i, tab.sinTc[B.rt0 + j], tab.cosTc[B.rt0 + j],
                   tab.sinPc[B.ct0 + i], tab.cosPc[B.ct0 + i]};
}
__device__ __forceinline__ double asin_small(double s) {
    const double s2 = s * s;
    return s * (1.0 + s2 * (1.0 / 6 + s2 * (3.0 / 40 + s2 * (5.0 / 112 + s2 * (35.0 / 1152 +
                                                                              s2 * (63.0 / 2816))))));
}
__device__ __forceinline__ P2d index_near(double invh, const U3& r, const CellRef& o) {
    const double rho = sqrt(r.x * r.x + r.y * r.y);
    const double nt = rho * o.cT - r.z * o.sT, ct = r.z * o.cT + rho * o.sT;
    const double np = r.y * o.cP - r.x * o.sP, cp = r.x * o.cP + r.y * o.sP;
    P2d q;
    if (ct > 0.0 && fabs(nt) < 0.02) {
        q.x = o.gx + asin_small(nt) * invh;
    } else {
        q.x = safe_acos_d(r.z) * invh;
    }
    if (cp > 0.0 && fabs(np) < 0.02 * rho) {
        q.y = o.gy + asin_small(np * rcp64(rho)) * invh;
    } else {
        q.y = atan2(r.y, r.x) * invh;
    }
    return q;
}

// lerpCoords on unit vectors: the point a fraction alpha along the great
// circle from f to t (Rodrigues rotation about k = f x t / |f x t|); f when
// the two coincide (k undefined), as the reference returns `from`.
// Map values at neighbouring grid points are a few cells apart, so the angle
// theta between f and t is small: below |t - f| = 0.02 the rotation is the
// same point written as a f + b t with a = sin((1 - alpha) theta) / sin theta,
// b = sin(alpha theta) / sin theta, both by their series in theta^2, and
// theta^2 = (2 asin(|t - f| / 2))^2 by its series in |t - f|^2 (every
// truncation below 1e-16 relative there): no square root, division or
// normalisation. Larger angles rotate with acos and sincos.
__device__ __forceinline__ double slerp_coef(double beta, double u) {
    const double b2 = beta * beta;
    return beta * (1.0 + u * (1.0 - b2) *
                             (1.0 / 6 + u * ((7.0 - 3.0 * b2) / 360 +
                                             u * (31.0 + b2 * (-18.0 + 3.0 * b2)) / 15120)));
}
__device__ __forceinline__ U3 slerp_u(const U3& f, const U3& t, double alpha) {
    const double dx = t.x - f.x, dy = t.y - f.y, dz = t.z - f.z;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < 4e-4) {
        if (d2 == 0.0) return f;
        const double u = d2 * (1.0 + d2 * (1.0 / 12 + d2 * (1.0 / 90 + d2 * (1.0 / 560))));
        const double a = slerp_coef(1.0 - alpha, u), b = slerp_coef(alpha, u);
        return U3{a * f.x + b * t.x, a * f.y + b * t.y, a * f.z + b * t.z};
    }
    double kx = f.y * t.z - f.z * t.y, ky = f.z * t.x - f.x * t.z, kz = f.x * t.y - f.y * t.x;
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double kin = 1.0 / sqrt(k2);
    kx *= kin; ky *= kin; kz *= kin;
    if (__builtin_isnan(kx)) return f;
    const double cth = f.x * t.x + f.y * t.y + f.z * t.z;
    double sn, cs;
    const double span = safe_acos_d(cth) * alpha;
    sincos(span, &sn, &cs);
    const double kf = kx * f.x + ky * f.y + kz * f.z;
    const double cx = ky * f.z - kz * f.y, cy = kz * f.x - kx * f.z, cz = kx * f.y - ky * f.x;
    return U3{f.x * cs + cx * sn + kx * kf * (1 - cs), f.y * cs + cy * sn + ky * kf * (1 - cs),
              f.z * cs + cz * sn + kz * kf * (1 - cs)};
}

// sampleMapping (bubble.cu:539-603) on a unit-vector map: the corner
// selection of the reference, three spherical interpolations.
__device__ U3 sample_map_u(const UMap& M, size_t off, int nth, int nph, double rx, double ry) {
    const int m = (nph & (nph - 1)) == 0 ? nph - 1 : -1;
    if (rx >= 0.5 && rx < (double)nth - 0.5 && m >= 0) {
        // away from the polar rows: the corners of rows floor(x), floor(x) + 1
        // that the general branch below selects there, the phi wrap on the
        // integer column
        const double x = rx - 0.5, fx = floor(x), fy = floor(ry);
        const double aT = x - fx, aP = ry - fy;
        const int pI = (int)fy;
        const size_t r0 = off + (size_t)(int)fx * nph, r1 = r0 + nph;
        const size_t c0 = (size_t)(pI & m), c1 = (size_t)((pI + 1) & m);
        auto at = [&](size_t c) { return U3{M.x[c], M.y[c], M.z[c]}; };
        return slerp_u(slerp_u(at(r0 + c0), at(r0 + c1), aP), slerp_u(at(r1 + c0), at(r1 + c1), aP),
                       aT);
    }
    double x = rx - 0.5, y = ry;
    const bool fl = validate_coord(x, y, nth, nph);
    int pI = (int)floor(y), tI = (int)floor(x);
    const double aP = y - pI, aT = x - tI;
    if ((unsigned)pI >= (unsigned)nph) pI = pI < 0 ? 0 : nph - 1;
    auto at = [&](int row, int col) {
        row = min(max(row, 0), nth - 1);
        const size_t c = off + (size_t)row * nph + wrap_col(col, nph, m);
        return U3{M.x[c], M.y[c], M.z[c]};
    };
    U3 ll{0, 0, 1}, lr{0, 0, 1}, hl{0, 0, 1}, hr{0, 0, 1};
    if (fl) {
        if (tI == 0) {
            hl = at(0, pI); hr = at(0, pI + 1);
            ll = at(0, pI + nph / 2); lr = at(0, pI + 1 + nph / 2);
        } else {
            tI -= 1;
        }
    }
    if (tI == nth - 1) {
        ll = at(tI, pI); lr = at(tI, pI + 1);
        hl = at(tI, pI + nph / 2); hr = at(tI, pI + 1 + nph / 2);
    } else if (tI != 0 || !fl) {
        ll = at(tI, pI); lr = at(tI, pI + 1);
        hl = at(tI + 1, pI); hr = at(tI + 1, pI + 1);
    }
    return slerp_u(slerp_u(ll, lr, aP), slerp_u(hl, hr, aP), aT);
}

// MULTISAMPLE (header.cuh:98): five samples around a point with these
// weights and offsets (bubble.cu:983-995)
__constant__ const double kMsW[5] = {0.125, 0.125, 0.125, 0.125, 0.5};
__constant__ const double kMsDx[5] = {-0.25, 0.25, -0.25, 0.25, 0.0};
__constant__ const double kMsDy[5] = {-0.25, -0.25, 0.25, 0.25, 0.0};

// -------------------------------------------------------- reductions ----
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Deterministic block sum of NV values; every thread receives the sums.
// `red` must hold 16 * NV doubles of LDS.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) red[k * 16 + wid] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += red[k * 16 + w];
        v[k] = s;
    }
    __syncthreads();
}

// Deterministic block maximum of a non-negative value; every thread receives
// it. `red` must hold 16 doubles of LDS.
__device__ __forceinline__ double block_max(double v, double* red) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double m = 0.0;
    for (int w = 0; w < nw; ++w) m = fmax(m, red[w]);
    __syncthreads();
    return m;
}

// The rows of a row block. A row of R threads (C cells each) takes R rounded
// up to whole waves; a block of at least twice that holds blockDim.x / w
// rows j0, j0 + 1, ... (packed: rows of several widths share one launch),
// one segment of w threads each; otherwise its one row has the whole block.
// seg: this thread's row (j0 + seg), lt: its place in the row, nt: threads a
// segment, nseg: segments of the block.
struct RowSeg {
    int seg, lt, nt, nseg;
};
// PACKED = false: the launch has one row a block (compiled to constants).
template <bool PACKED = true>
__device__ __forceinline__ RowSeg row_seg(int R) {
    const int w = (R + 63) & ~63;
    RowSeg s;
    if (!PACKED || 2 * w > (int)blockDim.x) {
        s.nt = blockDim.x;
        s.seg = 0;
        s.lt = threadIdx.x;
        s.nseg = 1;
    } else {
        s.nt = w;
        // segments are whole waves, so the segment (and the row it gives)
        // is wave-uniform: said to the compiler, per-row values (the row's
        // table entries, offsets) load into scalar registers
        s.seg = __builtin_amdgcn_readfirstlane((int)threadIdx.x / w);
        s.lt = (int)threadIdx.x - s.seg * w;
        s.nseg = (int)blockDim.x / w;
    }
    return s;
}

// Deterministic sums of NV values over each segment's threads (the block
// sum when the block holds one row); every thread receives its segment's.
// Segments are whole waves. `red` must hold 16 * NV doubles of LDS.
template <int NV>
__device__ __forceinline__ void seg_sum(double (&v)[NV], double* red, const RowSeg& s) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) red[k * 16 + wid] = v[k];
    }
    __syncthreads();
    const int w0 = (s.seg * s.nt) >> 6, nw = (s.nt + 63) >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double t = 0.0;
        for (int w = 0; w < nw; ++w) t += red[k * 16 + w0 + w];
        v[k] = t;
    }
    __syncthreads();
}
// The segment's maximum of a non-negative value.
__device__ __forceinline__ double seg_max(double v, double* red, const RowSeg& s) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    const int w0 = (s.seg * s.nt) >> 6, nw = (s.nt + 63) >> 6;
    double m = 0.0;
    for (int w = 0; w < nw; ++w) m = fmax(m, red[w0 + w]);
    __syncthreads();
    return m;
}
// The line-solve depth of a block's rows: the recorded depth of its one row,
// or the largest over a packed block's rows (every row of a block runs the
// same reduction levels; a deeper reduction than recorded only continues
// an already decoupled one). `red` must hold 16 doubles of LDS.
template <bool PACKED = true>
__device__ __forceinline__ int block_steps(const int* __restrict__ rsteps, int base, int j0,
                                           int nrows, const RowSeg& s, double* red) {
    if (!PACKED || s.nseg == 1) return rsteps[base + j0];
    int mx = -1;
    if (threadIdx.x < 64)
        for (int r = threadIdx.x; r < min(s.nseg, nrows); r += 64) mx = max(mx, rsteps[base + j0 + r]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    __syncthreads();
    if (threadIdx.x == 0) red[0] = (double)mx;
    __syncthreads();
    const int r = (int)red[0];
    __syncthreads();
    return r;
}

// Maximum of the per-row values [base, base+n) of one slot.
__device__ __forceinline__ double max_partials(const double* __restrict__ part, int base, int n,
                                               double* red) {
    double m = 0.0;
    for (int r = threadIdx.x; r < n; r += blockDim.x) m = fmax(m, part[base + r]);
    return block_max(m, red);
}

// Sum the per-row partials [base, base+n) of NV interleaved slots
// (layout part[slot * stride + row]).
template <int NV>
__device__ __forceinline__ void sum_partials(const double* __restrict__ part, long stride, int base,
                                             int n, double (&out)[NV], double* red) {
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = 0.0;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
#pragma unroll
        for (int k = 0; k < NV; ++k) out[k] += part[k * stride + base + r];
    }
    block_sum<NV>(out, red);
}

// Per-launch sums of the solve, accumulated without fences: every row block
// adds its block sums with device-scope atomics into one of kAccSub = 64
// sub-accumulators (row j uses j mod 64, so at most nth / 64 atomics meet
// on one address), and kernels of later launches read them (stream order
// makes them visible); a wave reads one slot with one load per lane. Slots
// per bubble: the assembly (|r0|^2, |b|^2, r_in.z, |r_in|^2, mu as a
// maximum) and the FP64 residual of each refinement k = 1..kAccRefine (the
// same first four). A slot is zeroed by a kernel that neither reads nor adds
// to it, before the launch that adds to it: the assembly's by the update
// kernel, refinement 1's by the assembly, refinement k+1's by refinement k.
constexpr int kAccSub = 64, kAccVals = 5, kAccRefine = 16;
constexpr int kAccSlots = 1 + kAccRefine;
constexpr int kAccBubble = kAccSlots * kAccVals * kAccSub;
__host__ __device__ constexpr int acc_slot_res(int k) { return k; }          // k = 0: assembly
__host__ __device__ inline size_t acc_index(int b, int slot, int v, int sub) {
    return (((size_t)b * kAccSlots + slot) * kAccVals + v) * kAccSub + sub;
}
// The sum (maximum) of one value of a slot, in every lane of the calling
// wave; the same in every wave (fixed reduction order).
__device__ __forceinline__ double acc_sum(const double* __restrict__ acc, int b, int slot, int v) {
    double x = acc[acc_index(b, slot, v, threadIdx.x & 63)];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}
__device__ __forceinline__ double acc_max(const double* __restrict__ acc, int b, int slot, int v) {
    double x = acc[acc_index(b, slot, v, threadIdx.x & 63)];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
    return x;
}
__device__ __forceinline__ void acc_add(double* acc, int b, int slot, int v, int sub, double x) {
    unsafeAtomicAdd(acc + acc_index(b, slot, v, sub & (kAccSub - 1)), x);
}
// maximum of non-negative values through their bit patterns
__device__ __forceinline__ void acc_put_max(double* acc, int b, int slot, int v, int sub, double x) {
    atomicMax(reinterpret_cast<unsigned long long*>(acc + acc_index(b, slot, v, sub & (kAccSub - 1))),
              (unsigned long long)__double_as_longlong(x));
}
// zeroes one slot; called by every lane of one wave
__device__ __forceinline__ void acc_zero(double* acc, int b, int slot) {
#pragma unroll
    for (int v = 0; v < kAccVals; ++v) acc[acc_index(b, slot, v, threadIdx.x & 63)] = 0.0;
}

// Flags of SolveCtl that blocks of one launch may set while other blocks of
// the same launch read them: each flag caches a decision every block makes
// from the same sums, so a block sees either value and acts alike.
__device__ __forceinline__ int flag_load(const int* f) {
    return __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void flag_store(int* f, int v) {
    __hip_atomic_store(f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The FP64 residual norm the tolerance applies to: relative to |b| unless
// the tolerance is absolute.
__device__ __forceinline__ double residual_norm(double rr, double bb, bool absolute) {
    return absolute ? sqrt(rr) : sqrt(rr) / (bb > 0.0 ? sqrt(bb) : 1.0);
}

// --------------------------------------------- periodic line solve ----
// Reciprocal for the preconditioner: FP32 uses the hardware v_rcp_f32
// (1 ulp; the line solve is a preconditioner, not the residual), FP64 the
// correctly rounded division.
__device__ __forceinline__ float pre_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double pre_rcp(double x) { return 1.0 / x; }

// Coupling below which a reduced row is decoupled: below the rounding of the
// precision, so stopping the reduction there changes z by less than one ulp.
template <typename Ti> struct PcrTol;
template <> struct PcrTol<float> { static constexpr float v = 2e-9f; };
template <> struct PcrTol<double> { static constexpr double v = 1e-18; };

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_down(v, o, 64));
    return __shfl(v, 0, 64);
}

// Parallel cyclic reduction of one periodic tridiagonal row held in LDS,
// normalised to unit diagonal: x_i + a_i x_{i-1} + c_i x_{i+1} = d_i, n a
// power of two >= 4. Each reduction doubles the coupling stride and, for a
// diagonally dominant row, squares the size of the couplings; after
// log2(n)-1 reductions the stride is n/2, where x_{i-n/2} = x_{i+n/2} and
// the rows pair up into 2x2 systems. A row whose couplings have all fallen
// below PcrTol is decoupled early (x_i = d_i).
// steps < 0: reduce adaptively (a block-wide max per reduction) and return
// the number of reductions done; steps >= 0: do exactly that many (the
// count returned by the adaptive solve of the same matrix).
// The reductions alternate between the arrays (sa, sc, sd) and the free
// arrays (ta, tc, td), so one barrier per reduction orders them; on exit
// *x points to the array that holds the solution. CPT bounds the cells per
// thread.
template <typename Ti, int CPT>
__device__ __forceinline__ int pcr_periodic(Ti* sa, Ti* sc, Ti* sd, Ti* ta, Ti* tc, Ti* td, int n,
                                            int steps, float* red, Ti** x, int tid, int nt) {
    const int m = n - 1;
    const int full = 31 - __clz(n) - 1;     // log2(n) - 1
    Ti *a = sa, *c = sc, *d = sd, *na = ta, *nc = tc, *nd = td;
    int done = 0;
    for (int s = 1; s < n / 2; s <<= 1) {
        if (steps >= 0 && done >= steps) break;
        float mx = 0.0f;
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            int i = tid + k * nt;
            if (i < n) {
                int im = (i - s) & m, ip = (i + s) & m;
                Ti ai = a[i], ci = c[i], di = d[i];
                Ti am = a[im], cm = c[im], dm = d[im];
                Ti ap = a[ip], cp = c[ip], dp = d[ip];
                Ti iD = pre_rcp(Ti(1) - ai * cm - ci * ap);
                const Ti a2 = -ai * am * iD, c2 = -ci * cp * iD;
                na[i] = a2;
                nc[i] = c2;
                nd[i] = (di - ai * dm - ci * dp) * iD;
                mx = fmaxf(mx, (float)(fabs(a2) + fabs(c2)));
            }
        }
        ++done;
        Ti* t;
        t = a; a = na; na = t;
        t = c; c = nc; nc = t;
        t = d; d = nd; nd = t;
        if (steps < 0) {
            // block-wide (every row of a packed block takes the same depth)
            mx = wave_max(mx);
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
            __syncthreads();
            float bm = 0.0f;
            for (int w = 0; w < ((int)blockDim.x + 63) >> 6; ++w) bm = fmaxf(bm, red[w]);
            __syncthreads();
            if (bm < (float)PcrTol<Ti>::v) {   // d already holds x
                *x = d;
                return done;
            }
        } else {
            __syncthreads();
        }
    }
    if (done < full) {   // truncated: x_i = d_i
        *x = d;
        return done;
    }
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        int i = tid + k * nt;
        if (i < n) {
            int j = (i + n / 2) & m;
            Ti ei = a[i] + c[i], ej = a[j] + c[j];
            nd[i] = (d[i] - ei * d[j]) * pre_rcp(Ti(1) - ei * ej);
        }
    }
    __syncthreads();
    *x = nd;
    return done;
}

// The periodic row of a length n that is not a power of two (n >= 2, one
// row per thread: threads t < n hold row t), by the rank-one correction of
// the periodic system into a non-periodic one (Sherman-Morrison; the form of
// Numerical Recipes' `cyclic`): with the corner couplings alpha = a_0 and
// beta = c_{n-1} and gamma = -1, T = T' + u v^T where T' drops the corners
// and has the diagonals 1 - gamma = 2 (row 0) and 1 - alpha beta / gamma =
// 1 + alpha beta (row n-1), u = (gamma, 0, ..., 0, beta), v = (1, 0, ..., 0,
// alpha / gamma). T' y = d and T' z = u are reduced together by
// non-periodic cyclic reduction (a coupling past an end of the row is zero,
// so ceil(log2 n) reductions decouple every row, exactly), and x = y -
// (v . y) / (1 + v . z) z. The arrays (sa, sc, sd, su) and (ta, tc, td, tu)
// alternate as in pcr_periodic; on exit *x points to the solution. Returns
// the number of reductions.
template <typename Ti>
__device__ __forceinline__ int pcr_general(Ti* sa, Ti* sc, Ti* sd, Ti* su, Ti* ta, Ti* tc, Ti* td,
                                           Ti* tu, int n, Ti** x, int i) {
    const bool act = i < n;
    // the corner couplings, read by every thread before rows 0 and n-1 change
    const Ti alpha = sa[0], beta = sc[n - 1];
    __syncthreads();
    if (act) {
        Ti u = Ti(0);
        if (i == 0) {
            sa[0] = Ti(0);
            sc[0] = sc[0] * Ti(0.5);
            sd[0] = sd[0] * Ti(0.5);
            u = Ti(-0.5);
        }
        if (i == n - 1) {
            const Ti inv = Ti(1) / (Ti(1) + alpha * beta);
            sa[i] = sa[i] * inv;
            sc[i] = Ti(0);
            sd[i] = sd[i] * inv;
            u = beta * inv;
        }
        su[i] = u;
    }
    __syncthreads();
    Ti *a = sa, *c = sc, *d = sd, *w = su, *na = ta, *nc = tc, *nd = td, *nw = tu;
    int done = 0;
    for (int s = 1; s < n; s <<= 1) {
        if (act) {
            const int im = i - s, ip = i + s;
            const Ti ai = a[i], ci = c[i];
            Ti am = Ti(0), cm = Ti(0), dm = Ti(0), wm = Ti(0);
            Ti ap = Ti(0), cp = Ti(0), dp = Ti(0), wp = Ti(0);
            if (im >= 0) { am = a[im]; cm = c[im]; dm = d[im]; wm = w[im]; }
            if (ip < n) { ap = a[ip]; cp = c[ip]; dp = d[ip]; wp = w[ip]; }
            const Ti iD = Ti(1) / (Ti(1) - ai * cm - ci * ap);
            na[i] = -ai * am * iD;
            nc[i] = -ci * cp * iD;
            nd[i] = (d[i] - ai * dm - ci * dp) * iD;
            nw[i] = (w[i] - ai * wm - ci * wp) * iD;
        }
        ++done;
        Ti* t;
        t = a; a = na; na = t;
        t = c; c = nc; nc = t;
        t = d; d = nd; nd = t;
        t = w; w = nw; nw = t;
        __syncthreads();
    }
    // d holds y, w holds z
    const Ti f = (d[0] - alpha * d[n - 1]) / (Ti(1) + w[0] - alpha * w[n - 1]);
    if (act) nd[i] = d[i] - f * w[i];
    __syncthreads();
    *x = nd;
    return done;
}

}  // namespace sfk
