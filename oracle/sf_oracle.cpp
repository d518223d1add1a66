/*
 * FP64 OpenMP oracle for the plain soap-film step of RiverIntheSky/SoapBubble.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product path never does.
 *
 * It restates, in plain C++ over dense host arrays, the CUDA code of
 *   kernel/bubble.cu   (samplers, great-circle trace, advection kernels,
 *                       air flow, linear-system assembly, force updates)
 *   kernel/solver.cu   (CSR pattern, stepForward orchestration)
 * with one OpenMP thread team over theta rows in place of one CUDA thread
 * per cell. Each function names the reference lines it follows.
 *
 * Declared deviations from the reference text (all at rounding level or in
 * code the reference never reaches with valid inputs):
 *  D1 double sin/cos replace the sinf/cosf of the FP64 assembly
 *     (bubble.cu:1333-1337); smoothstep stays float as in helper_math.h.
 *  D2 sample indices use floor() on doubles; floorf_mode=1 reproduces the
 *     reference's float rounding (bubble.cu:183-184, 252-253, 325-326).
 *  D3 the air-flow mirror index is (nPhi - i) mod nPhi; the reference reads
 *     index nPhi (one past the row) at i = 0 (bubble.cu:1364-1392).
 *  D4 a u_theta sample exactly at theta = pi would read rows nTheta-1 and
 *     nTheta of an (nTheta-1)-row array (bubble.cu:299-308); it is clamped
 *     to the south-cap branch instead.
 *  D5 the linear solve is PCG (phi-line block-Jacobi) to a stated residual
 *     instead of AmgX PCGF (bubble.cu:1479-1508). Rows of broken cells are
 *     identity rows as assembled; their known value gamma_max is moved to
 *     the right-hand side of the neighbouring rows, which leaves the
 *     solution unchanged and makes the system symmetric.
 *
 * How each function is pinned (tests/test_oracle_samplers.py,
 * tests/test_oracle_step.py, tests/test_golden.py):
 *  samplers     a ghost-row bilinear restatement on 8x16 and 16x32 grids,
 *               constants, exact node values, periodicity, NaN skipping
 *  trace        zero velocity, uniform flow along the equator and a
 *               meridian (exact), O(h^2) convergence for solid-body rotation
 *  advect       zero-velocity identity, constants, the two grid reflections,
 *               a geodesic flow along a grid meridian (exact); the
 *               velocity transport (tests/test_oracle_transport.py): a
 *               restatement through rotation matrices on random and
 *               tilted-rotation fields (1e-12), the departure point of a
 *               tilted solid-body rotation at third order in dt, and the
 *               advected velocity against the sampled field transported
 *               along the great circle from the exact departure point,
 *               third order (a flipped frame cross product gives 2 or 1)
 *  airflow      closed form: 0.5 (3 - |t - 3|) taper(beta) (n x r), t = 0
 *  col_ind      the neighbour pattern
 *  assemble     symmetry, row sums, Cholesky, known answer, O(h^2) Y_lm
 *               consistency, divergence-free cancellation, closed forms of
 *               the film, air and gravity terms, broken rows
 *  solve        dense direct solves (with and without broken cells)
 *  update       per-cell closed forms (friction, div, broken, evaporation)
 *  derive       closed forms and the published parameter values
 *  volume       midpoint-rule area of the sphere, O(h^2)
 *  air blend    the t/12 mirror blend of the assembly by its symmetries:
 *               mirror-symmetric u with antisymmetric v is t-independent,
 *               antisymmetric u with symmetric v scales by t/6 - 1
 *  BIMOCQ       (tests/test_oracle_bimocq.py) lerpCoords closed forms,
 *               identity maps, zero-velocity map updates, the still film
 *               equal to the plain step, reinitialisation bounds; on a
 *               meridian flow the closed forms bwd = x - k s, fwd = x + k s
 *               of k map updates, the change accumulated through the
 *               forward map, the BIMOCQ eta through both backward maps,
 *               and the error correction of the reinitialisation
 *               (correctCentered1/2) through both maps (linear fields:
 *               exact values)
 * Parity unpinned (reference text only): the u_theta sample exactly at the
 * north pole; floorf_mode (D2) and the float sinf/cosf of D1, which the
 * oracle does not reproduce.
 */
#include "sf_oracle.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* header.cuh:76-83 physical constants */
const double kR = 8.3144598;
const double kSigmaA = 0.07275;
const double kRho = 1e3;
const double kG = 9.8;
const double kRhoA = 1.184;
const double kNuA = 1.562e-5;
const double kNu = 1.25e-4;                /* soap solution viscosity */
const double kEvaporation = -0.00000003; /* header.cuh:95 */
const double kEps = 1e-7;                /* bubble.cu:28 */
const double kThresh = 0.0;              /* bubble.cu:27 */
const double kPi = 3.14159265358979323846;
const double kHalfPi = 1.57079632679489661923;

struct V3 { double x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(double s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double length(V3 v) { return std::sqrt(dot(v, v)); }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
/* utils.h:88-91: v * rsqrt(dot(v,v)) */
inline V3 normalize(V3 v) { return v * (1.0 / std::sqrt(dot(v, v))); }
/* utils.h:109-111 */
inline double safe_acos(double x) { return std::acos(std::max(-1.0, std::min(1.0, x))); }

/* helper_math.h:1432-1436, float arithmetic */
inline float smoothstep_f(float a, float b, float x) {
    float y = (x - a) / (b - a);
    y = y < 0.0f ? 0.0f : (y > 1.0f ? 1.0f : y);
    return y * y * (3.0f - (2.0f * y));
}

struct Grid {
    int nth, nph;
    double h, invh;
    int floorf_mode;
    explicit Grid(const sfo_params* p)
        : nth(p->n_theta), nph(2 * p->n_theta), h(kPi / p->n_theta),
          invh(p->n_theta / kPi), floorf_mode(p->floorf_mode) {}
    double flr(double x) const {
        return floorf_mode ? (double)std::floor((float)x) : std::floor(x);
    }
};

/* bubble.cu:153-160 */
inline double bubble_lerp(double from, double to, double alpha) {
    if (std::isnan(from)) return to;
    if (std::isnan(to)) return from;
    return (1.0 - alpha) * from + alpha * to;
}

/* bubble.cu:116-137. Id.x = theta index, Id.y = phi index. */
inline bool validate_coord(const Grid& g, double& x, double& y) {
    bool ret = false;
    if (x >= g.nth) { x = g.nph - x; y += g.nth; ret = !ret; }
    if (x < 0) { x = -x; y += g.nth; ret = !ret; }
    y = std::fmod(y + g.nph, (double)g.nph);
    return ret;
}

inline size_t umod(long v, long n) { return (size_t)(((v % n) + n) % n); }

/* sampleVPhi, bubble.cu:178-240 */
double sample_vphi(const Grid& g, const double* in, double rx, double ry) {
    double x = rx - 0.5, y = ry + 0.5;
    bool flipped = validate_coord(g, x, y);
    int phiIndex = (int)g.flr(y);
    int thetaIndex = (int)g.flr(x);
    double aPhi = y - phiIndex, aTheta = x - thetaIndex;
    const long n = g.nph;
    if (thetaIndex == 0 && flipped) {
        size_t lo = umod(phiIndex, n), hi = (lo + 1) % n;
        double higher = -bubble_lerp(in[lo], in[hi], aPhi);
        lo = umod(phiIndex + n / 2, n); hi = (lo + 1) % n;
        double lower = bubble_lerp(in[lo], in[hi], aPhi);
        return bubble_lerp(lower, higher, aTheta);
    }
    if (flipped) thetaIndex -= 1;
    if (thetaIndex == g.nth - 1) {
        const double* row = in + (size_t)thetaIndex * n;
        size_t lo = umod(phiIndex, n), hi = (lo + 1) % n;
        double lower = bubble_lerp(row[lo], row[hi], aPhi);
        lo = umod(phiIndex + n / 2, n); hi = (lo + 1) % n;
        double higher = -bubble_lerp(row[lo], row[hi], aPhi);
        double v = bubble_lerp(lower, higher, aTheta);
        return flipped ? -v : v;
    }
    size_t lo = umod(phiIndex, n), hi = (lo + 1) % n;
    const double* r0 = in + (size_t)thetaIndex * n;
    const double* r1 = r0 + n;
    double lower = bubble_lerp(r0[lo], r0[hi], aPhi);
    double higher = bubble_lerp(r1[lo], r1[hi], aPhi);
    double v = bubble_lerp(lower, higher, aTheta);
    return flipped ? -v : v;
}

/* sampleVTheta, bubble.cu:247-313 (input has nTheta-1 rows) */
double sample_vtheta(const Grid& g, const double* in, double rx, double ry) {
    double x = rx - 1.0, y = ry;
    bool flipped = validate_coord(g, x, y);
    int phiIndex = (int)g.flr(y);
    int thetaIndex = (int)g.flr(x);
    double aPhi = y - phiIndex, aTheta = x - thetaIndex;
    const long n = g.nph;
    if ((rx < 0 && rx > -1) || (rx > g.nth && rx < g.nth + 1)) {
        thetaIndex -= 1;
        aTheta += 1;
    } else if (rx >= g.nth + 1 || rx <= -1) {
        thetaIndex -= 2;
    }
    if (thetaIndex == 0 && flipped && rx > -1) {
        size_t lo = umod(phiIndex, n), hi = (lo + 1) % n;
        double higher = -bubble_lerp(in[lo], in[hi], aPhi);
        lo = (lo + n / 2) % n; hi = (hi + n / 2) % n;
        double lower = bubble_lerp(in[lo], in[hi], aPhi);
        aTheta = 0.5 * aTheta;
        return bubble_lerp(lower, higher, aTheta);
    }
    if (thetaIndex >= g.nth - 1 && !flipped) { /* D4: exactly theta = pi */
        thetaIndex = g.nth - 2;
        aTheta = 1.0;
    }
    if (thetaIndex == g.nth - 2) {
        const double* row = in + (size_t)thetaIndex * n;
        size_t lo = umod(phiIndex, n), hi = (lo + 1) % n;
        double lower = bubble_lerp(row[lo], row[hi], aPhi);
        lo = (lo + n / 2) % n; hi = (hi + n / 2) % n;
        double higher = -bubble_lerp(row[lo], row[hi], aPhi);
        aTheta = 0.5 * aTheta;
        double v = bubble_lerp(lower, higher, aTheta);
        return flipped ? -v : v;
    }
    size_t lo = umod(phiIndex, n), hi = (lo + 1) % n;
    const double* r0 = in + (size_t)thetaIndex * n;
    const double* r1 = r0 + n;
    double lower = bubble_lerp(r0[lo], r0[hi], aPhi);
    double higher = bubble_lerp(r1[lo], r1[hi], aPhi);
    double v = bubble_lerp(lower, higher, aTheta);
    return flipped ? -v : v;
}

/* sampleCentered, bubble.cu:320-376 */
double sample_centered(const Grid& g, const double* in, double rx, double ry) {
    double x = rx - 0.5, y = ry;
    bool flipped = validate_coord(g, x, y);
    int phiIndex = (int)g.flr(y);
    int thetaIndex = (int)g.flr(x);
    double aPhi = y - phiIndex, aTheta = x - thetaIndex;
    const long n = g.nph;
    if (thetaIndex == 0 && flipped) {
        size_t lo = umod(phiIndex, n), hi = (lo + 1) % n;
        double higher = bubble_lerp(in[lo], in[hi], aPhi);
        lo = (lo + n / 2) % n; hi = (hi + n / 2) % n;
        double lower = bubble_lerp(in[lo], in[hi], aPhi);
        return bubble_lerp(lower, higher, aTheta);
    }
    if (flipped) thetaIndex -= 1;
    if (thetaIndex == g.nth - 1) {
        const double* row = in + (size_t)thetaIndex * n;
        size_t lo = umod(phiIndex, n), hi = (lo + 1) % n;
        double lower = bubble_lerp(row[lo], row[hi], aPhi);
        lo = (lo + n / 2) % n; hi = (hi + n / 2) % n;
        double higher = bubble_lerp(row[lo], row[hi], aPhi);
        return bubble_lerp(lower, higher, aTheta);
    }
    size_t lo = umod(phiIndex, n), hi = (lo + 1) % n;
    const double* r0 = in + (size_t)thetaIndex * n;
    const double* r1 = r0 + n;
    double lower = bubble_lerp(r0[lo], r0[hi], aPhi);
    double higher = bubble_lerp(r1[lo], r1[hi], aPhi);
    return bubble_lerp(lower, higher, aTheta);
}

/* Shared first two stages of the great-circle method: the start frame, the
 * half step and the parallel-transported midpoint velocity
 * (bubble.cu:427-496 / 651-731 / 787-867). Returns false where the
 * reference returns early (|u| = 0 or NaN); `unorm` then holds |u|. */
struct Circle {
    V3 w, u_, v_;
    double unorm;
};

bool circle_midpoint(const Grid& g, const double* uth, const double* uph, double dt,
                     double sinT, double cosT, double sinP, double cosP,
                     double guTheta, double guPhi, Circle& c) {
    V3 eTheta = {cosT * cosP, cosT * sinP, -sinT};
    V3 ePhi = {-sinP, cosP, 0.0};
    c.w = {cosP * sinT, sinP * sinT, cosT};
    V3 u = guTheta * eTheta + guPhi * ePhi;
    c.unorm = length(u);
    if (!(c.unorm > 0)) return false;
    c.u_ = normalize(u);
    c.v_ = cross(c.w, c.u_);
    /* RUNGE_KUTTA: half step, bubble.cu:456-496 */
    double dS = -0.5 * c.unorm * dt;
    V3 midx = c.u_ * std::sin(dS) + c.w * std::cos(dS);
    double mct = safe_acos(midx.z), mcp = std::atan2(midx.y, midx.x);
    double muTheta = sample_vtheta(g, uth, mct * g.invh, mcp * g.invh);
    double muPhi = sample_vphi(g, uph, mct * g.invh, mcp * g.invh);
    V3 mu = {muTheta * std::cos(mct) * std::cos(mcp) - muPhi * std::sin(mcp),
             muTheta * std::cos(mct) * std::sin(mcp) + muPhi * std::cos(mcp),
             -muTheta * std::sin(mct)};
    V3 uCircleMid = c.u_ * std::cos(dS) - c.w * std::sin(dS);
    V3 vCircleMid = cross(midx, uCircleMid);
    double mguTheta = dot(mu, vCircleMid);
    double mguPhi = dot(mu, uCircleMid);
    u = mguPhi * c.u_ + mguTheta * c.v_;
    c.unorm = length(u);
    if (!(c.unorm > 0)) return false;
    c.u_ = normalize(u);
    c.v_ = cross(c.w, c.u_);
    return true;
}

/* traceGreatCircle, bubble.cu:416-506 */
void trace_great_circle(const Grid& g, const double* uth, const double* uph, double dt,
                        double gx, double gy, double& ox, double& oy) {
    double ct = gx * g.h, cp = gy * g.h;
    double sinT = std::sin(ct), cosT = std::cos(ct), sinP = std::sin(cp), cosP = std::cos(cp);
    double guTheta = sample_vtheta(g, uth, gx, gy);
    double guPhi = sample_vphi(g, uph, gx, gy);
    Circle c;
    if (!circle_midpoint(g, uth, uph, dt, sinT, cosT, sinP, cosP, guTheta, guPhi, c)) {
        ox = gx; oy = gy;
        return;
    }
    double dS = -c.unorm * dt;
    V3 px = c.u_ * std::sin(dS) + c.w * std::cos(dS);
    ox = safe_acos(px.z) * g.invh;
    oy = std::atan2(px.y, px.x) * g.invh;
}

/* Third stage of advectionVSphere{Phi,Theta}Kernel (bubble.cu:734-767 /
 * 870-903): trace to the departure point, sample there and transport back.
 * Returns the transported vector in the start frame. */
V3 circle_departure(const Grid& g, const double* uth, const double* uph, double dt,
                    const Circle& c) {
    double dS = -c.unorm * dt;
    V3 px = c.u_ * std::sin(dS) + c.w * std::cos(dS);
    double pct = safe_acos(px.z), pcp = std::atan2(px.y, px.x);
    double puTheta = sample_vtheta(g, uth, pct * g.invh, pcp * g.invh);
    double puPhi = sample_vphi(g, uph, pct * g.invh, pcp * g.invh);
    V3 pu = {puTheta * std::cos(pct) * std::cos(pcp) - puPhi * std::sin(pcp),
             puTheta * std::cos(pct) * std::sin(pcp) + puPhi * std::cos(pcp),
             -puTheta * std::sin(pct)};
    V3 uCircleP = c.u_ * std::cos(dS) - c.w * std::sin(dS);
    V3 vCircleP = cross(px, uCircleP);
    double a = dot(pu, vCircleP);
    double b = dot(pu, uCircleP);
    return b * c.u_ + a * c.v_;
}

/* The air-flow corridor vector of airFlowU/airFlowV (bubble.cu:2317-2335),
 * before projection on e_phi / e_theta. */
V3 air_corridor(double sinT, double cosT, double sinP, double cosP, double U,
                double radius, double t) {
    double angle = kPi / 2.;
    V3 wind = normalize(V3{0., std::cos(angle), std::sin(angle)});
    V3 position = {1., 0, 0};
    V3 normal = normalize(cross(position, wind));
    V3 cur = {sinT * cosP, sinT * sinP, cosT};
    double beta = safe_acos(dot(normal, cur));
    V3 projected = normalize(cur - dot(normal, cur) * normal);
    V3 projectedWind = cross(normal, projected);
    double ss = (double)smoothstep_f(0.2f, 0.4f, (float)std::fabs(beta - kHalfPi));
    V3 air = 0.5 * std::sin(beta) / U * radius * projectedWind;
    air = air * (1. - ss);
    air = air * (3. - std::fabs(t - 3.));
    return air;
}

/* Periodic tridiagonal solve by the Sherman-Morrison correction of the
 * Thomas algorithm (Numerical Recipes section 2.7, `cyclic`).
 * a: coefficient of x[i-1], b: diagonal, c: coefficient of x[i+1];
 * a[0] couples x[n-1] and c[n-1] couples x[0]. */
void cyclic_tridiag(int n, const double* a, const double* b, const double* c,
                    const double* d, double* x, std::vector<double>& work) {
    work.resize(4 * (size_t)n);
    double* bb = work.data();
    double* u = bb + n;
    double* z = u + n;
    double* cp = z + n;
    double gamma = -b[0];
    double alpha = a[0], beta = c[n - 1];
    for (int i = 0; i < n; ++i) bb[i] = b[i];
    bb[0] = b[0] - gamma;
    bb[n - 1] = b[n - 1] - alpha * beta / gamma;
    auto thomas = [&](const double* rhs, double* out) {
        double den = bb[0];
        cp[0] = c[0] / den;
        out[0] = rhs[0] / den;
        for (int i = 1; i < n; ++i) {
            den = bb[i] - a[i] * cp[i - 1];
            cp[i] = (i < n - 1 ? c[i] : 0.0) / den;
            out[i] = (rhs[i] - a[i] * out[i - 1]) / den;
        }
        for (int i = n - 2; i >= 0; --i) out[i] -= cp[i] * out[i + 1];
    };
    thomas(d, x);
    for (int i = 0; i < n; ++i) u[i] = 0.0;
    u[0] = gamma;
    u[n - 1] = beta;
    thomas(u, z);
    double fact = (x[0] + alpha * x[n - 1] / gamma) / (1.0 + z[0] + alpha * z[n - 1] / gamma);
    for (int i = 0; i < n; ++i) x[i] -= fact * z[i];
}


/* ------------------------------------------------------------ BIMOCQ ---- */
struct P2 { double x, y; };   /* (theta, phi) index coordinates */

/* lerpCoords, bubble.cu:512-536: spherical interpolation of two index-space
 * points along their great circle (Rodrigues rotation about k). */
P2 lerp_coords(const Grid& g, P2 from, P2 to, double alpha) {
    const double h = kPi / g.nth;
    const double ft = from.x * h, fp = from.y * h, tt = to.x * h, tp = to.y * h;
    V3 f3 = normalize(V3{std::cos(fp) * std::sin(ft), std::sin(fp) * std::sin(ft), std::cos(ft)});
    V3 t3 = normalize(V3{std::cos(tp) * std::sin(tt), std::sin(tp) * std::sin(tt), std::cos(tt)});
    V3 k = normalize(cross(f3, t3));
    if (std::isnan(k.x)) return from;
    double span = safe_acos(dot(f3, t3)) * alpha;
    V3 r = f3 * std::cos(span) + cross(k, f3) * std::sin(span) + k * dot(k, f3) * (1 - std::cos(span));
    return P2{safe_acos(r.z) / h, std::atan2(r.y, r.x) / h};
}

/* sampleMapping, bubble.cu:539-603. The corner selection follows the
 * reference's branches; the phi indices wrap modulo nPhi. */
P2 sample_mapping(const Grid& g, const double* mt, const double* mp, P2 raw) {
    double x = raw.x - 0.5, y = raw.y;
    const bool fl = validate_coord(g, x, y);
    const long n = g.nph;
    int pI = (int)g.flr(y), tI = (int)g.flr(x);
    const double aP = y - pI, aT = x - tI;
    P2 ll{0, 0}, lr{0, 0}, hl{0, 0}, hr{0, 0};
    auto at = [&](int row, long col) { size_t c = (size_t)row * n + umod(col, n); return P2{mt[c], mp[c]}; };
    if (fl) {
        if (tI == 0) {
            hl = at(0, pI); hr = at(0, pI + 1);
            ll = at(0, pI + n / 2); lr = at(0, pI + 1 + n / 2);
        } else {
            tI -= 1;
        }
    }
    if (tI == g.nth - 1) {
        ll = at(tI, pI); lr = at(tI, pI + 1);
        hl = at(tI, pI + n / 2); hr = at(tI, pI + 1 + n / 2);
    } else if (tI != 0 || !fl) {
        ll = at(tI, pI); lr = at(tI, pI + 1);
        hl = at(tI + 1, pI); hr = at(tI + 1, pI + 1);
    }
    return lerp_coords(g, lerp_coords(g, ll, lr, aP), lerp_coords(g, hl, hr, aP), aT);
}

/* dist, bubble.cu:60-79: great-circle distance in grid cells */
double grid_dist(const Grid& g, P2 a, P2 b) {
    const double h = kPi / g.nth;
    V3 a3 = normalize(V3{std::cos(a.y * h) * std::sin(a.x * h), std::sin(a.y * h) * std::sin(a.x * h), std::cos(a.x * h)});
    V3 b3 = normalize(V3{std::cos(b.y * h) * std::sin(b.x * h), std::sin(b.y * h) * std::sin(b.x * h), std::cos(b.x * h)});
    return g.invh * safe_acos(dot(a3, b3));
}

/* MULTISAMPLE (header.cuh:98): five samples around the point
 * (bubble.cu:983-995) */
const double kMsW[5] = {0.125, 0.125, 0.125, 0.125, 0.5};
const double kMsDx[5] = {-0.25, 0.25, -0.25, 0.25, 0.0};
const double kMsDy[5] = {-0.25, -0.25, 0.25, 0.25, 0.0};

}  // namespace

extern "C" {

int sfo_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Bubble::Bubble initialiser list (bubble.cu:2573-2580) and the constants
 * uploaded by Bubble::run (bubble.cu:2627-2638). Inside the initialiser list
 * the parameter U shadows the member (which is forced to 1.0); run() uses
 * the member, so UGlobal = 1 and eva = evaporation * radius / H. */
void sfo_derive(double radius, double H, double U, double Gamma_m, double T,
                double dt_seconds, int gravity, int evaporation, sfo_params* out) {
    double sigma_r = kR * T;
    out->M = Gamma_m * sigma_r / (3 * kRho * H * U * U);
    out->Cr = kRhoA * std::sqrt(kNuA * radius / U) / (kRho * H);
    out->gs = kG * radius / (U * U);
    out->gamma_max = kSigmaA / sigma_r / Gamma_m;
    out->dt = dt_seconds * U / radius;
    out->radius = radius;
    out->U = 1.0;
    out->eva = evaporation ? kEvaporation * radius / (H * 1.0) : 0.0;
    out->gravity = gravity;
    out->evaporation = evaporation;
    out->re = kNu / (U * radius);   /* re(nu / (U radius)) with the U argument */
}

double sfo_sample(const sfo_params* p, int kind, const double* field,
                  double theta_id, double phi_id) {
    Grid g(p);
    if (kind == 0) return sample_centered(g, field, theta_id, phi_id);
    if (kind == 1) return sample_vphi(g, field, theta_id, phi_id);
    return sample_vtheta(g, field, theta_id, phi_id);
}

void sfo_trace(const sfo_params* p, const double* uth, const double* uph,
               double theta_id, double phi_id, double* out2) {
    Grid g(p);
    trace_great_circle(g, uth, uph, p->dt, theta_id, phi_id, out2[0], out2[1]);
}

/* Solver::advection, plain branch (bubble.cu:1171-1214): u_theta, u_phi and
 * the scalars all read the old velocity. */
void sfo_advect(const sfo_params* p, const double* uth, const double* uph,
                const double* eta, const double* gam, double* uth_out,
                double* uph_out, double* eta_out, double* gam_out) {
    Grid g(p);
    const int nth = g.nth, nph = g.nph;
    const double dt = p->dt;
    /* advectionVSphereThetaKernel, bubble.cu:775-904 */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth - 1; ++j) {
        for (int i = 0; i < nph; ++i) {
            double gx = j + 1.0, gy = (double)i;
            double ct = gx * g.h, cp = gy * g.h;
            double sinT = std::sin(ct), cosT = std::cos(ct);
            double sinP = std::sin(cp), cosP = std::cos(cp);
            double guTheta = uth[(size_t)j * nph + i];
            double guPhi = sample_vphi(g, uph, gx, gy);
            Circle c;
            double out;
            if (!circle_midpoint(g, uth, uph, dt, sinT, cosT, sinP, cosP, guTheta, guPhi, c)) {
                out = c.unorm;
            } else {
                V3 pu = circle_departure(g, uth, uph, dt, c);
                V3 eTheta = {cosT * cosP, cosT * sinP, -sinT};
                out = dot(pu, eTheta);
            }
            uth_out[(size_t)j * nph + i] = out;
        }
    }
    /* advectionVSpherePhiKernel, bubble.cu:639-768 */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth; ++j) {
        for (int i = 0; i < nph; ++i) {
            double gx = j + 0.5, gy = i - 0.5;
            double ct = gx * g.h, cp = gy * g.h;
            double sinT = std::sin(ct), cosT = std::cos(ct);
            double sinP = std::sin(cp), cosP = std::cos(cp);
            double guTheta = sample_vtheta(g, uth, gx, gy);
            double guPhi = uph[(size_t)j * nph + i];
            Circle c;
            double out;
            if (!circle_midpoint(g, uth, uph, dt, sinT, cosT, sinP, cosP, guTheta, guPhi, c)) {
                out = c.unorm;
            } else {
                V3 pu = circle_departure(g, uth, uph, dt, c);
                V3 ePhi = {-sinP, cosP, 0.0};
                out = dot(pu, ePhi);
            }
            uph_out[(size_t)j * nph + i] = out;
        }
    }
    /* advectionAllCentered, bubble.cu:937-961 */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth; ++j) {
        for (int i = 0; i < nph; ++i) {
            double tx, ty;
            trace_great_circle(g, uth, uph, dt, j + 0.5, (double)i, tx, ty);
            eta_out[(size_t)j * nph + i] = sample_centered(g, eta, tx, ty);
            gam_out[(size_t)j * nph + i] = sample_centered(g, gam, tx, ty);
        }
    }
}

/* airFlowU (bubble.cu:2301-2339) on u_phi faces and airFlowV
 * (bubble.cu:2342-2380) on u_theta faces. */
void sfo_airflow(const sfo_params* p, const double* uair0, const double* vair0,
                 double* uair, double* vair) {
    Grid g(p);
    const int nth = g.nth, nph = g.nph;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth; ++j) {
        for (int i = 0; i < nph; ++i) {
            double gPhi = ((double)i - 0.5) * g.h;
            double gTheta = ((double)j + 0.5) * g.h;
            double sinT = std::sin(gTheta), cosT = std::cos(gTheta);
            double sinP = std::sin(gPhi), cosP = std::cos(gPhi);
            V3 air = air_corridor(sinT, cosT, sinP, cosP, p->U, p->radius, p->time_s);
            V3 ePhi = {-sinP, cosP, 0.};
            double v = uair0[(size_t)j * nph + i] + dot(air, ePhi);
            uair[(size_t)j * nph + i] = std::isnan(v) ? 0.0 : v;
        }
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth - 1; ++j) {
        for (int i = 0; i < nph; ++i) {
            double gTheta = ((double)j + 1.0) * g.h;
            double gPhi = ((double)i + 0.0) * g.h;
            double sinT = std::sin(gTheta), cosT = std::cos(gTheta);
            double sinP = std::sin(gPhi), cosP = std::cos(gPhi);
            V3 air = air_corridor(sinT, cosT, sinP, cosP, p->U, p->radius, p->time_s);
            V3 eTheta = {cosT * cosP, cosT * sinP, -sinT};
            double v = vair0[(size_t)j * nph + i] + dot(air, eTheta);
            vair[(size_t)j * nph + i] = std::isnan(v) ? 0.0 : v;
        }
    }
}

/* initLinearSystem, solver.cu:25-60 */
void sfo_col_ind(const sfo_params* p, int32_t* col) {
    const long nth = p->n_theta, nph = 2 * nth, N = nth * nph;
    for (long idx = 0; idx < N; ++idx) {
        long i5 = 5 * idx;
        col[i5] = (int32_t)(idx < nph ? (idx + nth) % nph : idx - nph);
        col[i5 + 1] = (int32_t)((idx % nph) == 0 ? idx + nph - 1 : idx - 1);
        col[i5 + 2] = (int32_t)idx;
        col[i5 + 3] = (int32_t)((idx % nph) == nph - 1 ? idx - nph + 1 : idx + 1);
        if (idx >= (nth - 1) * nph + nth) col[i5 + 4] = (int32_t)(idx - nth);
        else if (idx >= (nth - 1) * nph) col[i5 + 4] = (int32_t)(idx + nth);
        else col[i5 + 4] = (int32_t)(idx + nph);
    }
}

/* concentrationLinearSystemKernel, bubble.cu:1301-1473 (W = 0 since
 * W1 = W2 = 0, bubble.cu:2633-2638; fu/fv only enter with VISCOUS). */

/* ----------------------------------------------------------- VISCOUS ---- */
/* helper_math.h's clamp(float, float, float): fmaxf(a, fminf(f, b)), so a
 * double argument is rounded to float and NaN gives the upper bound. */
static double clampf_ref(double x, double a, double b) {
    return (double)std::fmax((float)a, std::fmin((float)x, (float)b));
}

/* divergenceKernel, bubble.cu:1217-1269 (D1: double sin for sinf). */
void sfo_visc_div(const sfo_params* p, const double* uth, const double* uph, double* div) {
    Grid g(p);
    const int nth = g.nth, nph = g.nph;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth; ++j)
        for (int i = 0; i < nph; ++i) {
            const double th = ((double)j + 0.5) * g.h, half = 0.5 * g.h;
            const double uW = uph[(size_t)j * nph + i], uE = uph[(size_t)j * nph + (i + 1) % nph];
            const double vN = j != 0 ? uth[(size_t)(j - 1) * nph + i] : 0.0;
            const double vS = j != nth - 1 ? uth[(size_t)j * nph + i] : 0.0;
            const double invSin = 1.0 / std::sin(th);
            const double sinN = std::sin(th - half), sinS = std::sin(th + half);
            const double factor = invSin * g.invh;
            const double termTheta = factor * (vS * sinS - vN * sinN);
            const double termPhi = g.invh * (uE - uW);
            div[(size_t)j * nph + i] = termTheta + termPhi;
        }
}

/* computeforcevelthetaKernel (bubble.cu:1511-1646) into fv[(nth-1) nph] and
 * computeforcevelphiKernel_viscous (bubble.cu:1649-1816) into fu[nth nph]
 * (D1: double sin/cos for the phi kernel's sinf/cosf). */
void sfo_visc_forces(const sfo_params* p, const double* uth, const double* uph,
                     const double* eta, const double* div, double* fu, double* fv) {
    Grid g(p);
    const int nth = g.nth, nph = g.nph, half_n = nph / 2;
    const double ih = g.invh, re = p->re;
    auto U = [&](int j, int i) { return uph[(size_t)j * nph + i]; };
    auto V = [&](int j, int i) { return uth[(size_t)j * nph + i]; };
    auto E = [&](int j, int i) { return eta[(size_t)j * nph + i]; };
    auto D = [&](int j, int i) { return div[(size_t)j * nph + i]; };
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth - 1; ++j)
        for (int i = 0; i < nph; ++i) {
            const double gTheta = ((double)j + 1.0) * g.h;
            const int iW = (i - 1 + nph) % nph, iE = (i + 1) % nph, opp = (i + half_n) % nph;
            const double u0 = U(j, i), u1 = U(j, iE), u2 = U(j + 1, i), u3 = U(j + 1, iE);
            const double divN = D(j, i), divS = D(j + 1, i);
            const double etaN = E(j, i), etaS = E(j + 1, i);
            double& out = fv[(size_t)j * nph + i];
            const double invEta = 2. / (etaN + etaS);
            if (std::isinf(invEta)) { out = NAN; continue; }
            const double dv = 0.5 * (divN + divS);
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
            const double sinT = std::sin(gTheta), cscT = 1. / sinT, cosT = std::cos(gTheta);
            const double cotT = cosT * cscT;
            const double s11N = 0. + 2. * (pvpyNorth + divN);
            const double s11S = 0. + 2. * (pvpySouth + divS);
            const double s22 = -2. * (pvpy - 2. * dv);
            const double s12 = cscT * pvpx + pupy - uPhi * cotT;
            const double halfStep = 0.5 * g.h;
            const double sinN = std::sin(gTheta - halfStep), sinS = std::sin(gTheta + halfStep);
            const double psspy = ih * (s11S * sinS - s11N * sinN);
            const double pspx = cscT * pvpxx + pupxy - cotT * pupx;
            const double f2 = re * cscT * invEta * pDpx * s12;
            const double f4 = re * invEta * pDpy * 2 * (dv + pvpy);
            const double f5 = re * cscT * (psspy + pspx - cosT * s22);
            const double f7 = p->gs * sinT;
            out = clampf_ref(f2 + f5 + f4, -0.1, 0.1) + f7;
            if (sinT < 0.1) out = f7;
        }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth; ++j)
        for (int i = 0; i < nph; ++i) {
            const double gTheta = ((double)j + 0.5) * g.h;
            const int iW = (i - 1 + nph) % nph;
            const int oW = (iW + half_n) % nph, oC = (i + half_n) % nph;
            const double divW = D(j, iW), divE = D(j, i);
            const double etaW = E(j, iW), etaE = E(j, i);
            double v0, v1, v2, v3;
            if (j != 0) { v0 = V(j - 1, iW); v1 = V(j - 1, i); }
            else { v0 = 0.5 * (V(j, iW) - V(j, oW)); v1 = 0.5 * (V(j, i) - V(j, oC)); }
            if (j != nth - 1) { v2 = V(j, iW); v3 = V(j, i); }
            else { v2 = 0.5 * (V(j - 1, iW) - V(j - 1, oW)); v3 = 0.5 * (V(j - 1, i) - V(j - 1, oC)); }
            double& out = fu[(size_t)j * nph + i];
            const double invEta = 2. / (etaW + etaE);
            if (std::isinf(invEta)) { out = NAN; continue; }
            const double dv = 0.5 * (divW + divE);
            const double pvpx = 0.5 * ih * (v1 + v3 - v0 - v2);
            const double pvpyWest = ih * (v2 - v0), pvpyEast = ih * (v3 - v1);
            const double pvpy = 0.5 * ih * (v2 + v3 - v0 - v1);
            double pupyNorth = 0.0, pupySouth = 0.0;
            const double u1 = U(j, i);
            double u3, u4;
            if (j != 0) { u3 = U(j - 1, i); pupyNorth = ih * (u1 - u3); }
            else u3 = -U(j, oC);
            if (j != nth - 1) { u4 = U(j + 1, i); pupySouth = ih * (u4 - u1); }
            else u4 = -U(j, oC);
            const double pupy = 0.5 * ih * (u4 - u3);
            const double sinT = std::sin(gTheta), cscT = 1. / sinT, cosT = std::cos(gTheta);
            const double cotT = cosT * cscT;
            const double s12 = cscT * pvpx + pupy - u1 * cotT;
            const double pDpx = ih * (etaE - etaW);
            double d0, d1, d2, d3;
            if (j != 0) { d0 = E(j - 1, iW); d2 = E(j - 1, i); }
            else { d0 = E(j, oW); d2 = E(j, oC); }
            if (j != nth - 1) { d1 = E(j + 1, iW); d3 = E(j + 1, i); }
            else { d1 = E(j, oW); d3 = E(j, oC); }
            const double pDpy = 0.25 * ih * (d1 + d3 - d0 - d2);
            const double halfStep = 0.5 * g.h;
            const double tS = gTheta + halfStep, tN = gTheta - halfStep;
            const double sinN = std::sin(tN), sinS = std::sin(tS);
            const double cosN = std::cos(tN), cosS = std::cos(tS);
            double uNorth = 0.5 * (u3 + u1), uSouth = 0.5 * (u1 + u4);
            if (j == 0) uNorth = 0.5 * (u1 - u3);
            if (j == nth - 1) uSouth = 0.5 * (u1 - u4);
            const double psspy = ih * (ih * (v0 + v3 - v1 - v2) + sinS * pupySouth -
                                       sinN * pupyNorth - cosS * uSouth + cosN * uNorth);
            const double s22W = 2 * (2 * divW - pvpyWest), s22E = 2 * (2 * divE - pvpyEast);
            const double pspx = ih * (s22E - s22W);
            const double f2 = re * invEta * pDpy * s12;
            const double f4 = re * invEta * cscT * pDpx * 2 * (2 * dv - pvpy);
            const double f5 = re * cscT * (psspy + pspx + cosT * s12);
            out = clampf_ref(f2 + f4 + f5, -0.1, 0.1);
            if (sinT < 0.1) out = 0.;
        }
}

/* fu, fv of the VISCOUS stages from the (advected) velocity and thickness */
static void visc_forces(const sfo_params* p, const double* uth, const double* uph,
                        const double* eta, std::vector<double>& fu, std::vector<double>& fv) {
    const size_t nth = p->n_theta, nph = 2 * nth, N = nth * nph;
    std::vector<double> div(N);
    fu.assign(N, 0.0);
    fv.assign(N - nph, 0.0);
    sfo_visc_div(p, uth, uph, div.data());
    sfo_visc_forces(p, uth, uph, eta, div.data(), fu.data(), fv.data());
}

void sfo_assemble(const sfo_params* p, const double* uth, const double* uph,
                  double* gam, const double* eta, const double* uair,
                  const double* vair, double* val, double* rhs) {
    Grid g(p);
    const int nth = g.nth, nph = g.nph;
    const double dt = p->dt, t = p->time_s;
    std::vector<double> fu, fv;
    const bool visc = p->viscous && p->gravity;
    if (visc) visc_forces(p, uth, uph, eta, fu, fv);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth; ++j) {
        for (int i = 0; i < nph; ++i) {
            size_t idx = (size_t)j * nph + i, idx5 = 5 * idx;
            double gamma = gam[idx];
            double e = eta[idx];
            if (gamma > (p->gamma_max - kEps) || e == kThresh) {
                val[idx5] = 0.0; val[idx5 + 1] = 0.0; val[idx5 + 3] = 0.0; val[idx5 + 4] = 0.0;
                val[idx5 + 2] = 1.0;
                rhs[idx] = p->gamma_max;
                gam[idx] = p->gamma_max;
                continue;
            }
            double invDt = 1.0 / dt;
            double gTheta = ((double)j + 0.5) * g.h;
            double halfStep = 0.5 * g.h;
            double sinS = std::sin(gTheta + halfStep);   /* D1 */
            double sinN = std::sin(gTheta - halfStep);
            double sinTheta = std::sin(gTheta);
            double cscTheta = 1. / sinTheta;
            int iE = (i + 1) % nph, iW = (i - 1 + nph) % nph;
            double etaW = eta[(size_t)j * nph + iW], etaE = eta[(size_t)j * nph + iE];
            double etaN = 0.0, etaS = 0.0;
            double uW = uph[(size_t)j * nph + i], uE = uph[(size_t)j * nph + iE];
            double vN = 0.0, vS = 0.0;
            int opp = (nph - i) % nph;          /* D3 */
            int oppE = (nph - iE) % nph;
            double uAirW = 0.5 * (uair[(size_t)j * nph + i] * t / 12.0 +
                                  uair[(size_t)j * nph + opp] * (1 - t / 12.0));
            double uAirE = 0.5 * (uair[(size_t)j * nph + iE] * t / 12.0 +
                                  uair[(size_t)j * nph + oppE] * (1 - t / 12.0));
            double vAirN = 0.0, vAirS = 0.0;
            if (j != 0) {
                size_t jn = (size_t)(j - 1) * nph;
                vAirN = 0.5 * (vair[jn + i] * t / 12.0 - vair[jn + opp] * (1 - t / 12.0));
                vN = uth[jn + i];
                etaN = eta[jn + i];
            } else {
                etaN = eta[(i + nth) % nph];
            }
            if (j != nth - 1) {
                size_t js = (size_t)j * nph;
                vAirS = 0.5 * (vair[js + i] * t / 12.0 - vair[js + opp] * (1 - t / 12.0));
                vS = uth[js + i];
                etaS = eta[(size_t)(j + 1) * nph + i];
            } else {
                etaS = eta[(size_t)j * nph + (i + nth) % nph];
            }
            if (std::isnan(vN)) vN = 0.0;
            if (std::isnan(vS)) vS = 0.0;
            if (std::isnan(uW)) uW = 0.0;
            if (std::isnan(uE)) uE = 0.0;
            double CrDt = p->Cr * dt, MDt = p->M * dt, s2 = g.invh * g.invh;
            double etaUp = (etaN + e) / 2.0;
            val[idx5] = -s2 * sinN * MDt / (etaUp + CrDt);
            double etaLeft = (etaW + e) / 2.0;
            val[idx5 + 1] = -s2 * cscTheta * MDt / (etaLeft + CrDt);
            double etaRight = (etaE + e) / 2.0;
            val[idx5 + 3] = -s2 * cscTheta * MDt / (etaRight + CrDt);
            double etaDown = (etaS + e) / 2.0;
            val[idx5 + 4] = -s2 * sinS * MDt / (etaDown + CrDt);
            val[idx5 + 2] = sinTheta * invDt / gamma -
                            (val[idx5] + val[idx5 + 1] + val[idx5 + 3] + val[idx5 + 4]);
            double div = g.invh * (uE * etaRight / (etaRight + p->Cr * dt) -
                                   uW * etaLeft / (etaLeft + p->Cr * dt) +
                                   (vS * sinS * etaDown / (etaDown + p->Cr * dt) -
                                    vN * sinN * etaUp / (etaUp + p->Cr * dt)));
            if (p->airflow) {
                div += g.invh * (uAirE / (etaRight / p->Cr * invDt + 1.0) -
                                 uAirW / (etaLeft / p->Cr * invDt + 1.0) +
                                 vAirS * sinS / (etaDown / p->Cr * invDt + 1.0) -
                                 vAirN * sinN / (etaUp / p->Cr * invDt + 1.0));
            }
            if (visc) {
                /* `gravity` with `VISCOUS`, bubble.cu:1458-1464 */
                const double fuW = fu[(size_t)j * nph + i], fuE = fu[(size_t)j * nph + iE];
                const double fvN = j != 0 ? fv[(size_t)(j - 1) * nph + i] : 0.0;
                const double fvS = j != nth - 1 ? fv[(size_t)j * nph + i] : 0.0;
                div += g.invh * (sinS * fvS * etaDown / (invDt * etaDown + p->Cr) -
                                 sinN * fvN * etaUp / (invDt * etaUp + p->Cr) +
                                 fuE * etaRight / (invDt * etaRight + p->Cr) -
                                 fuW * etaLeft / (invDt * etaLeft + p->Cr));
            } else if (p->gravity) {
                div += p->gs * g.invh *
                       (sinS * sinS * etaDown / (invDt * etaDown + p->Cr) -
                        sinN * sinN * etaUp / (invDt * etaUp + p->Cr));
            }
            rhs[idx] = sinTheta * invDt - div;
        }
    }
}

void sfo_spmv(const sfo_params* p, const double* val, const double* x, double* y) {
    const long nth = p->n_theta, nph = 2 * nth, N = nth * nph;
    std::vector<int32_t> col(5 * N);
    sfo_col_ind(p, col.data());
#pragma omp parallel for schedule(static)
    for (long r = 0; r < N; ++r) {
        double s = 0.0;
        for (int k = 0; k < 5; ++k) s += val[5 * r + k] * x[col[5 * r + k]];
        y[r] = s;
    }
}

/* D5: PCG with phi-line block-Jacobi preconditioning on the symmetrised
 * system. */
int sfo_solve(const sfo_params* p, const double* val, const double* rhs,
              double* x, sfo_stats* st) {
    const int nth = p->n_theta, nph = 2 * nth;
    const size_t N = (size_t)nth * nph;
    std::vector<int32_t> col(5 * N);
    sfo_col_ind(p, col.data());
    /* identity rows = cells whose row is e_i with rhs gamma_max */
    std::vector<char> ident(N);
    for (size_t r = 0; r < N; ++r)
        ident[r] = val[5 * r] == 0.0 && val[5 * r + 1] == 0.0 && val[5 * r + 3] == 0.0 &&
                   val[5 * r + 4] == 0.0 && val[5 * r + 2] == 1.0 && rhs[r] == p->gamma_max;
    std::vector<double> A(5 * N), b(N);
    for (size_t r = 0; r < N; ++r) {
        b[r] = rhs[r];
        for (int k = 0; k < 5; ++k) A[5 * r + k] = val[5 * r + k];
        if (ident[r]) { x[r] = rhs[r]; continue; }
        for (int k = 0; k < 5; ++k) {
            if (k == 2) continue;
            size_t c = (size_t)col[5 * r + k];
            if (ident[c]) { b[r] -= A[5 * r + k] * rhs[c]; A[5 * r + k] = 0.0; }
        }
    }
    auto spmv = [&](const std::vector<double>& v, std::vector<double>& out) {
#pragma omp parallel for schedule(static)
        for (long r = 0; r < (long)N; ++r) {
            double s = 0.0;
            for (int k = 0; k < 5; ++k) s += A[5 * r + k] * v[col[5 * r + k]];
            out[r] = s;
        }
    };
    /* deterministic dot: per-row partials summed in row order */
    std::vector<double> part(nth);
    auto dotp = [&](const std::vector<double>& u, const std::vector<double>& v) {
#pragma omp parallel for schedule(static)
        for (int j = 0; j < nth; ++j) {
            double s = 0.0;
            for (int i = 0; i < nph; ++i) s += u[(size_t)j * nph + i] * v[(size_t)j * nph + i];
            part[j] = s;
        }
        double s = 0.0;
        for (int j = 0; j < nth; ++j) s += part[j];
        return s;
    };
    auto precond = [&](const std::vector<double>& r, std::vector<double>& z) {
#pragma omp parallel
        {
            std::vector<double> a(nph), bd(nph), c(nph), work;
#pragma omp for schedule(static)
            for (int j = 0; j < nth; ++j) {
                size_t o = (size_t)j * nph;
                for (int i = 0; i < nph; ++i) {
                    a[i] = A[5 * (o + i) + 1];
                    bd[i] = A[5 * (o + i) + 2];
                    c[i] = A[5 * (o + i) + 3];
                }
                cyclic_tridiag(nph, a.data(), bd.data(), c.data(), r.data() + o, z.data() + o, work);
            }
        }
    };
    std::vector<double> xv(x, x + N), r(N), z(N), pv(N), q(N);
    double bnorm = std::sqrt(dotp(b, b));
    if (bnorm == 0.0) bnorm = 1.0;
    double target = p->tol_absolute ? p->tol : p->tol * bnorm;
    int maxit = p->max_iters > 0 ? p->max_iters : 10000;
    int it = 0;
    /* CG on the recursive residual; restart from the true residual when the
       two have drifted apart (at most three restarts). */
    for (int restart = 0; restart < 4 && it < maxit; ++restart) {
        spmv(xv, q);
        for (size_t i = 0; i < N; ++i) r[i] = b[i] - q[i];
        double rnorm = std::sqrt(dotp(r, r));
        if (rnorm <= target) break;
        precond(r, z);
        pv = z;
        double rz = dotp(r, z);
        while (it < maxit) {
            ++it;
            spmv(pv, q);
            double alpha = rz / dotp(pv, q);
            for (size_t i = 0; i < N; ++i) { xv[i] += alpha * pv[i]; r[i] -= alpha * q[i]; }
            rnorm = std::sqrt(dotp(r, r));
            if (rnorm <= 0.5 * target) break;
            precond(r, z);
            double rz2 = dotp(r, z);
            double beta = rz2 / rz;
            rz = rz2;
            for (size_t i = 0; i < N; ++i) pv[i] = z[i] + beta * pv[i];
        }
    }
    /* recompute the true residual of the assembled system in FP64 */
    std::vector<double> y(N);
    sfo_spmv(p, val, xv.data(), y.data());
    double rr = 0.0, bb = 0.0;
    for (size_t i = 0; i < N; ++i) { double d = rhs[i] - y[i]; rr += d * d; bb += rhs[i] * rhs[i]; }
    std::copy(xv.begin(), xv.end(), x);
    if (st) {
        st->iters = it;
        st->abs_residual = std::sqrt(rr);
        st->rel_residual = std::sqrt(rr) / (bb > 0 ? std::sqrt(bb) : 1.0);
        st->converged = (p->tol_absolute ? st->abs_residual : st->rel_residual) <= p->tol * 1.0001;
    }
    double fin = p->tol_absolute ? std::sqrt(rr) : std::sqrt(rr) / (bb > 0 ? std::sqrt(bb) : 1.0);
    return fin <= p->tol * 1.0001 ? 0 : 1;
}

void sfo_update(const sfo_params* p, const double* uth, const double* uph,
                const double* eta, const double* gam, double* gam_next,
                const double* uair, const double* vair, double* uth_out,
                double* uph_out, double* eta_out) {
    Grid g(p);
    const int nth = g.nth, nph = g.nph;
    const double dt = p->dt;
    const size_t N = (size_t)nth * nph;
    std::vector<double> div(N);
    /* divergenceKernel_fromGamma, bubble.cu:1276-1295 */
#pragma omp parallel for schedule(static)
    for (long k = 0; k < (long)N; ++k) {
        double gamma_a = gam[k];
        double gamma = gam_next[k];
        if (gamma <= 0.0 || gamma > p->gamma_max - kEps) {
            gam_next[k] = p->gamma_max;
            div[k] = 1.0 / dt;
        } else {
            gamma = std::fmin(gamma, 2 * gamma_a);
            div[k] = (1.0 - gamma / gamma_a) / dt;
        }
    }
    if (p->viscous) {
        /* applyforcevelthetaKernel_viscous and applyforcevelphiKernel_viscous
         * (bubble.cu:1938-2038): the viscous forces, the t/12 mirror blend of
         * the air (D3 index), no separate gravity term (it is in fv) */
        std::vector<double> fu, fv;
        visc_forces(p, uth, uph, eta, fu, fv);
        const double t = p->time_s;
#pragma omp parallel for schedule(static)
        for (int j = 0; j < nth - 1; ++j)
            for (int i = 0; i < nph; ++i) {
                size_t k = (size_t)j * nph + i, ks = k + nph;
                double invEta = 2. / (eta[k] + eta[ks]);
                if (std::isinf(invEta)) { uth_out[k] = NAN; continue; }
                double v1 = uth[k];
                double pGpy = g.invh * (gam_next[ks] - gam_next[k]);
                double f3 = -p->M * invEta * pGpy;
                int opp = (nph - i) % nph;   /* D3 */
                double vAir = 0.5 * (vair[(size_t)j * nph + i] * t / 12.0 -
                                     vair[(size_t)j * nph + opp] * (1 - t / 12.0));
                double f6 = p->Cr * invEta * vAir;
                uth_out[k] = (v1 / dt + f3 + f6 + fv[k]) / (1. / dt + p->Cr * invEta);
            }
#pragma omp parallel for schedule(static)
        for (int j = 0; j < nth; ++j) {
            double gTheta = ((double)j + 0.5) * g.h;
            double cscTheta = 1. / std::sin(gTheta);
            for (int i = 0; i < nph; ++i) {
                size_t k = (size_t)j * nph + i, kw = (size_t)j * nph + (i - 1 + nph) % nph;
                double invEta = 2. / (eta[kw] + eta[k]);
                if (std::isinf(invEta)) { uph_out[k] = NAN; continue; }
                double u1 = uph[k];
                double pGpx = g.invh * (gam_next[k] - gam_next[kw]);
                double f3 = -p->M * invEta * cscTheta * pGpx;
                int opp = (nph - i) % nph;   /* D3 */
                double uAir = 0.5 * (uair[(size_t)j * nph + i] * t / 12.0 +
                                     uair[(size_t)j * nph + opp] * (1 - t / 12.0));
                double f6 = p->Cr * invEta * uAir;
                uph_out[k] = (u1 / dt + f3 + f6 + fu[k]) / (1. / dt + p->Cr * invEta);
            }
        }
    } else {
    /* applyforcevelthetaKernel, bubble.cu:1819-1876 */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth - 1; ++j) {
        for (int i = 0; i < nph; ++i) {
            size_t k = (size_t)j * nph + i, ks = k + nph;
            double gTheta = ((double)j + 1.0) * g.h;
            double v1 = uth[k];
            double invEta = 2. / (eta[k] + eta[ks]);
            if (std::isinf(invEta)) { uth_out[k] = NAN; continue; }
            double pGpy = g.invh * (gam_next[ks] - gam_next[k]);
            double f1 = -p->M * invEta * pGpy;
            double f2 = p->Cr * invEta * vair[k];
            double f3 = p->gravity ? p->gs * std::sin(gTheta) : 0.0;
            double f4 = 0.0;
            uth_out[k] = (v1 / dt + f1 + f2 + f3 + f4) / (1. / dt + p->Cr * invEta);
        }
    }
    /* applyforcevelphiKernel, bubble.cu:1879-1935 */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth; ++j) {
        double gTheta = ((double)j + 0.5) * g.h;
        double cscTheta = 1.0 / std::sin(gTheta);
        for (int i = 0; i < nph; ++i) {
            size_t k = (size_t)j * nph + i, kw = (size_t)j * nph + (i - 1 + nph) % nph;
            double u1 = uph[k];
            double invEta = 2. / (eta[kw] + eta[k]);
            if (std::isinf(invEta)) { uph_out[k] = NAN; continue; }
            double pGpx = g.invh * (gam_next[k] - gam_next[kw]) * cscTheta;
            double f1 = -p->M * invEta * pGpx;
            double f2 = p->Cr * invEta * uair[k];
            double f3 = 0.0, f4 = 0.0;
            uph_out[k] = (u1 / dt + (f1 + f2 + f3 + f4)) / (1. / dt + p->Cr * invEta);
        }
    }
    }
    /* applyforceThickness, bubble.cu:2041-2063 */
#pragma omp parallel for schedule(static)
    for (long k = 0; k < (long)N; ++k) {
        double v = eta[k] * (1 - dt * div[k]);
        if (p->evaporation) v += p->eva * dt;
        eta_out[k] = std::fmax(kThresh, v);
    }
}

double sfo_volume(const sfo_params* p, const double* eta) {
    Grid g(p);
    double v = 0.0;
    for (int j = 0; j < g.nth; ++j) {
        double s = 0.0;
        for (int i = 0; i < g.nph; ++i) s += eta[(size_t)j * g.nph + i];
        v += s * std::sin((j + 0.5) * g.h);
    }
    return v * g.h * g.h;
}

/* Solver::stepForward (solver.cu:260-292) with Solver::advection
 * (bubble.cu:1171-1214) and Solver::bodyforce (bubble.cu:2383-2506). */
int sfo_step(const sfo_params* p, double* uth, double* uph, double* eta,
             double* gam, const double* uair0, const double* vair0, sfo_stats* st) {
    const int nth = p->n_theta, nph = 2 * nth;
    const size_t N = (size_t)nth * nph, Nt = (size_t)(nth - 1) * nph;
    std::vector<double> uth1(Nt), uph1(N), eta1(N), gam1(N);
    sfo_advect(p, uth, uph, eta, gam, uth1.data(), uph1.data(), eta1.data(), gam1.data());
    std::vector<double> uair(N), vair(Nt);
    if (p->airflow) {
        sfo_airflow(p, uair0, vair0, uair.data(), vair.data());
    } /* else: without `airflow` the reference never writes uair/vair
         (bubble.cu:2389-2396); they are taken as zero here. */
    std::vector<double> val(5 * N), rhs(N);
    sfo_assemble(p, uth1.data(), uph1.data(), gam1.data(), eta1.data(), uair.data(),
                 vair.data(), val.data(), rhs.data());
    std::vector<double> x(gam1);
    sfo_stats local;
    sfo_stats* s = st ? st : &local;
    int rc = sfo_solve(p, val.data(), rhs.data(), x.data(), s);
    sfo_update(p, uth1.data(), uph1.data(), eta1.data(), gam1.data(), x.data(),
               uair.data(), vair.data(), uth, uph, eta);
    std::copy(x.begin(), x.end(), gam);
    s->volume = sfo_volume(p, eta);
    int any = 0;
    for (size_t k = 0; k < N; ++k) if (eta[k] > 0) { any = 1; break; }
    s->broken = !any || rc != 0;
    return rc;
}


/* ------------------------------------------------------------ BIMOCQ ---- */
void sfo_lerp_coords(const sfo_params* p, const double* from2, const double* to2, double alpha,
                     double* out2) {
    Grid g(p);
    P2 r = lerp_coords(g, P2{from2[0], from2[1]}, P2{to2[0], to2[1]}, alpha);
    out2[0] = r.x;
    out2[1] = r.y;
}

void sfo_sample_mapping(const sfo_params* p, const double* mt, const double* mp, double rx,
                        double ry, double* out2) {
    Grid g(p);
    P2 r = sample_mapping(g, mt, mp, P2{rx, ry});
    out2[0] = r.x;
    out2[1] = r.y;
}

/* updateMappingKernel, bubble.cu:606-633: out(x) = map(trace(x, dt)), the
 * result through validateCoord. */
void sfo_map_update(const sfo_params* p, const double* uth, const double* uph, double dt,
                    const double* mt, const double* mp, double* out_t, double* out_p) {
    Grid g(p);
    const int nth = g.nth, nph = g.nph;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth; ++j)
        for (int i = 0; i < nph; ++i) {
            double bx, by;
            trace_great_circle(g, uth, uph, dt, j + 0.5, (double)i, bx, by);
            P2 s = sample_mapping(g, mt, mp, P2{bx, by});
            validate_coord(g, s.x, s.y);
            out_t[(size_t)j * nph + i] = s.x;
            out_p[(size_t)j * nph + i] = s.y;
        }
}

/* initMapping, solver.cu:12-22 */
static void init_mapping(const Grid& g, double* mt, double* mp) {
    for (int j = 0; j < g.nth; ++j)
        for (int i = 0; i < g.nph; ++i) {
            mt[(size_t)j * g.nph + i] = j + 0.5;
            mp[(size_t)j * g.nph + i] = (double)i;
        }
}

void sfo_bimocq_init(const sfo_params* p, const double* eta, const double* gam, sfo_bimocq* b) {
    Grid g(p);
    const size_t N = (size_t)g.nth * g.nph;
    init_mapping(g, b->fwd_t, b->fwd_p);
    init_mapping(g, b->bwd_t, b->bwd_p);
    init_mapping(g, b->bwdprev_t, b->bwdprev_p);
    for (size_t k = 0; k < N; ++k) {
        b->eta_init[k] = b->eta_initlast[k] = eta[k];
        b->gam_init[k] = b->gam_initlast[k] = gam[k];
        b->eta_delta[k] = b->eta_deltalast[k] = 0.0;
        b->gam_delta[k] = b->gam_deltalast[k] = 0.0;
    }
    b->reinit_count = 0;
    b->last_distortion = 0.0;
}

/* advectionCenteredBimocq, bubble.cu:967-1013 */
void sfo_advect_bimocq(const sfo_params* p, const double* eta, const sfo_bimocq* b,
                       double* eta_out) {
    Grid g(p);
    const int nth = g.nth, nph = g.nph;
    const int ns = b->multisample ? 5 : 1;
    const double bl = b->blend;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth; ++j)
        for (int i = 0; i < nph; ++i) {
            const size_t c = (size_t)j * nph + i;
            if (eta[c] == kThresh) { eta_out[c] = kThresh; continue; }
            double a = 0.0;
            for (int s = 0; s < ns; ++s) {
                const int q = ns == 5 ? s : 4;
                const double w = ns == 5 ? kMsW[q] : 1.0;
                P2 pos{j + 0.5 + kMsDx[q], i + kMsDy[q]};
                P2 ip = sample_mapping(g, b->bwd_t, b->bwd_p, pos);
                P2 lp = sample_mapping(g, b->bwdprev_t, b->bwdprev_p, ip);
                a += (1.0 - bl) * w * (sample_centered(g, b->eta_initlast, lp.x, lp.y) +
                                       sample_centered(g, b->eta_delta, ip.x, ip.y) +
                                       sample_centered(g, b->eta_deltalast, lp.x, lp.y));
                a += bl * w * (sample_centered(g, b->eta_init, ip.x, ip.y) +
                               sample_centered(g, b->eta_delta, ip.x, ip.y));
            }
            eta_out[c] = std::fmax(kThresh, a);
        }
}

/* accumulateChangesCentered, bubble.cu:2081-2114: delta(x) += sum_s w_s
 * change(fwd(x + d_s)) */
void sfo_accumulate(const sfo_params* p, const double* fwd_t, const double* fwd_p,
                    const double* change, double* delta, int multisample) {
    Grid g(p);
    const int nth = g.nth, nph = g.nph;
    const int ns = multisample ? 5 : 1;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth; ++j)
        for (int i = 0; i < nph; ++i) {
            const size_t c = (size_t)j * nph + i;
            for (int s = 0; s < ns; ++s) {
                const int q = ns == 5 ? s : 4;
                const double w = ns == 5 ? kMsW[q] : 1.0;
                P2 ip = sample_mapping(g, fwd_t, fwd_p, P2{j + 0.5 + kMsDx[q], i + kMsDy[q]});
                delta[c] += w * sample_centered(g, change, ip.x, ip.y);
            }
        }
}

/* estimateDistortion, bubble.cu:1133-1168: the larger of max |x - bwd(fwd(x))|
 * and max |x - fwd(bwd(x))| in grid cells */
double sfo_distortion(const sfo_params* p, const sfo_bimocq* b) {
    Grid g(p);
    const int nth = g.nth, nph = g.nph;
    double mx = 0.0;
#pragma omp parallel for schedule(static) reduction(max : mx)
    for (int j = 0; j < nth; ++j)
        for (int i = 0; i < nph; ++i) {
            const size_t c = (size_t)j * nph + i;
            P2 x{j + 0.5, (double)i};
            P2 a = sample_mapping(g, b->bwd_t, b->bwd_p, P2{b->fwd_t[c], b->fwd_p[c]});
            P2 e = sample_mapping(g, b->fwd_t, b->fwd_p, P2{b->bwd_t[c], b->bwd_p[c]});
            mx = std::max(mx, std::max((double)(float)grid_dist(g, x, a),
                                       (double)(float)grid_dist(g, x, e)));
        }
    return mx;
}

/* reInitializeMapping with its error correction (bubble.cu:2509-2564,
 * correctCentered1/2 at bubble.cu:2180-2280). The neighbourhood bounds and
 * the clamp are evaluated in float as the reference's fminf/fmaxf/clamp
 * do. Deviation D6: the 3x3 neighbourhood wraps phi and reflects theta = -1
 * across the pole like theta = nTheta (the reference's validateId leaves
 * index -1 negative and reads before the row). */
void sfo_bimocq_reinit(const sfo_params* p, double* eta, double* gam, sfo_bimocq* b) {
    Grid g(p);
    const int nth = g.nth, nph = g.nph;
    const size_t N = (size_t)nth * nph;
    std::vector<double> eerr(N), gerr(N), eta0(eta, eta + N), gam0(gam, gam + N);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth; ++j)
        for (int i = 0; i < nph; ++i) {
            const size_t c = (size_t)j * nph + i;
            double te = 0.0, tg = 0.0;
            for (int q = 0; q < 5; ++q) {
                P2 ip = sample_mapping(g, b->fwd_t, b->fwd_p, P2{j + 0.5 + kMsDx[q], i + kMsDy[q]});
                te += kMsW[q] * sample_centered(g, eta0.data(), ip.x, ip.y);
                tg += kMsW[q] * sample_centered(g, gam0.data(), ip.x, ip.y);
            }
            eerr[c] = (te - b->eta_delta[c] - b->eta_init[c]) * 0.5;
            gerr[c] = (tg - b->gam_delta[c] - b->gam_init[c]) * 0.5;
        }
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nth; ++j)
        for (int i = 0; i < nph; ++i) {
            const size_t c = (size_t)j * nph + i;
            double ve = eta0[c], vg = gam0[c];
            for (int q = 0; q < 5; ++q) {
                P2 ip = sample_mapping(g, b->bwd_t, b->bwd_p, P2{j + 0.5 + kMsDx[q], i + kMsDy[q]});
                ve -= kMsW[q] * sample_centered(g, eerr.data(), ip.x, ip.y);
                vg -= kMsW[q] * sample_centered(g, gerr.data(), ip.x, ip.y);
            }
            auto clamp_local = [&](const std::vector<double>& in, double v) {
                float lo = (float)in[c], hi = lo;
                for (int t = -1; t <= 1; ++t)
                    for (int s = -1; s <= 1; ++s) {
                        int jj = j + t, ii = i + s;
                        if (jj < 0 || jj >= nth) { jj = jj < 0 ? 0 : nth - 1; ii += nth; }
                        const float cv = (float)in[(size_t)jj * nph + umod(ii, nph)];
                        lo = std::fmin(lo, cv);
                        hi = std::fmax(hi, cv);
                    }
                const float cl = std::fmax(lo, std::fmin((float)v, hi));
                return std::fmax(kThresh, (double)cl);
            };
            eta[c] = clamp_local(eta0, ve);
            gam[c] = clamp_local(gam0, vg);
        }
    std::swap(b->eta_initlast, b->eta_init);
    std::swap(b->eta_deltalast, b->eta_delta);
    std::swap(b->gam_initlast, b->gam_init);
    std::swap(b->gam_deltalast, b->gam_delta);
    std::swap(b->bwdprev_t, b->bwd_t);
    std::swap(b->bwdprev_p, b->bwd_p);
    for (size_t k = 0; k < N; ++k) {
        b->eta_init[k] = eta[k];
        b->gam_init[k] = gam[k];
        b->eta_delta[k] = 0.0;
        b->gam_delta[k] = 0.0;
    }
    init_mapping(g, b->fwd_t, b->fwd_p);
    init_mapping(g, b->bwd_t, b->bwd_p);
    b->reinit_count += 1;
}

/* Solver::stepForward with BIMOCQ (solver.cu:260-292): mapping updates,
 * advection (eta by BIMOCQ, Gamma plain, bubble.cu:1188-1205), bodyforce
 * with the change accumulation (bubble.cu:2481-2502), distortion check and
 * reinitialisation. */
int sfo_bimocq_step(const sfo_params* p, double* uth, double* uph, double* eta, double* gam,
                    const double* uair0, const double* vair0, sfo_bimocq* b, sfo_stats* st) {
    Grid g(p);
    const int nth = g.nth, nph = g.nph;
    const size_t N = (size_t)nth * nph, Nt = (size_t)(nth - 1) * nph;
    std::vector<double> tt(N), tp(N);
    sfo_map_update(p, uth, uph, -p->dt, b->fwd_t, b->fwd_p, tt.data(), tp.data());
    std::copy(tt.begin(), tt.end(), b->fwd_t);
    std::copy(tp.begin(), tp.end(), b->fwd_p);
    sfo_map_update(p, uth, uph, p->dt, b->bwd_t, b->bwd_p, tt.data(), tp.data());
    std::copy(tt.begin(), tt.end(), b->bwd_t);
    std::copy(tp.begin(), tp.end(), b->bwd_p);

    std::vector<double> uth1(Nt), uph1(N), eta1(N), gam1(N), eta_plain(N);
    sfo_advect(p, uth, uph, eta, gam, uth1.data(), uph1.data(), eta_plain.data(), gam1.data());
    sfo_advect_bimocq(p, eta, b, eta1.data());

    std::vector<double> uair(N), vair(Nt);
    if (p->airflow) sfo_airflow(p, uair0, vair0, uair.data(), vair.data());
    std::vector<double> val(5 * N), rhs(N);
    sfo_assemble(p, uth1.data(), uph1.data(), gam1.data(), eta1.data(), uair.data(), vair.data(),
                 val.data(), rhs.data());
    std::vector<double> x(gam1);
    sfo_stats local;
    sfo_stats* s = st ? st : &local;
    int rc = sfo_solve(p, val.data(), rhs.data(), x.data(), s);
    sfo_update(p, uth1.data(), uph1.data(), eta1.data(), gam1.data(), x.data(), uair.data(),
               vair.data(), uth, uph, eta);
    std::copy(x.begin(), x.end(), gam);
    std::vector<double> de(N), dg(N);
    for (size_t k = 0; k < N; ++k) {
        de[k] = eta[k] - eta1[k];
        dg[k] = gam[k] - gam1[k];
    }
    sfo_accumulate(p, b->fwd_t, b->fwd_p, de.data(), b->eta_delta, b->multisample);
    sfo_accumulate(p, b->fwd_t, b->fwd_p, dg.data(), b->gam_delta, b->multisample);

    b->last_distortion = sfo_distortion(p, b);
    if (b->last_distortion > b->threshold) sfo_bimocq_reinit(p, eta, gam, b);

    s->volume = sfo_volume(p, eta);
    int any = 0;
    for (size_t k = 0; k < N; ++k) if (eta[k] > 0) { any = 1; break; }
    s->broken = !any || rc != 0;
    return rc;
}

}  // extern "C"
