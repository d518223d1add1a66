// Host-only helpers: parameter derivation, air-flow corridor shape, seeded
// synthetic initial fields and the OpenEXR thickness writer.
#include "sf_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <thread>

#include <cerrno>
#include <climits>
#include <fcntl.h>
#include <sys/uio.h>
#include <unistd.h>

namespace sfh {

namespace {
// header.cuh:76-83 physical constants
const double kR = 8.3144598;
const double kSigmaA = 0.07275;
const double kRho = 1e3;
const double kG = 9.8;
const double kRhoA = 1.184;
const double kNuA = 1.562e-5;
const double kNu = 1.25e-4;   // soap solution kinematic viscosity, header.cuh:83
const double kEvaporation = -0.00000003;  // header.cuh:95
const double kPi = 3.14159265358979323846;
const double kHalfPi = 1.57079632679489661923;

struct V3 { double x, y, z; };
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(double s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 normalize(V3 v) { return v * (1.0 / std::sqrt(dot(v, v))); }

// smoothstep of helper_math.h (float arithmetic)
float smoothstepf(float a, float b, float x) {
    float y = (x - a) / (b - a);
    y = std::min(std::max(y, 0.0f), 1.0f);
    return y * y * (3.0f - 2.0f * y);
}

// Corridor vector without its time factor: a great circle through the x
// axis normal to the wind (0, cos(pi/2), sin(pi/2)), speed 0.5 sin(beta) r/U
// faded out by smoothstep(0.2, 0.4, |beta - pi/2|).
V3 corridor(double sinT, double cosT, double sinP, double cosP, double U, double radius) {
    const double angle = kPi / 2.;
    V3 wind = normalize(V3{0., std::cos(angle), std::sin(angle)});
    V3 position{1., 0., 0.};
    V3 normal = normalize(cross(position, wind));
    V3 cur{sinT * cosP, sinT * sinP, cosT};
    double c = std::max(-1.0, std::min(1.0, dot(normal, cur)));
    double beta = std::acos(c);
    V3 projected = normalize(cur - dot(normal, cur) * normal);
    V3 pw = cross(normal, projected);
    double ss = (double)smoothstepf(0.2f, 0.4f, (float)std::fabs(beta - kHalfPi));
    V3 a = 0.5 * std::sin(beta) / U * radius * pw;
    return a * (1. - ss);
}

// SplitMix64 (Steele, Lea, Flood 2014) seeding xoshiro256** (Blackman and
// Vigna 2018).
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (auto& w : s) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
            x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
            w = x ^ (x >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
};

struct Mode { double a; int l, m; };

std::vector<Mode> draw_modes(uint64_t seed, int lmax) {
    Rng r(seed);
    std::vector<Mode> modes(16);
    for (auto& md : modes) {
        md.a = 2.0 * r.uniform() - 1.0;
        md.l = 1 + (int)std::floor(r.uniform() * lmax);
        md.m = -md.l + (int)std::floor(r.uniform() * (2 * md.l + 1));
    }
    return modes;
}

// Fully normalised associated Legendre function Pbar_l^m(cos theta) (no
// Condon-Shortley phase) by the standard three-term recurrence in l.
double pbar(int l, int m, double theta) {
    const double x = std::cos(theta), s = std::sin(theta);
    double pmm = (2 * m + 1) / (4.0 * kPi);
    for (int k = 1; k <= m; ++k) pmm *= (2.0 * k - 1.0) / (2.0 * k);
    pmm = std::sqrt(pmm) * std::pow(std::fabs(s), m);
    if (l == m) return pmm;
    double p1 = x * std::sqrt(2.0 * m + 3.0) * pmm;
    if (l == m + 1) return p1;
    double p0 = pmm;
    double a_prev = std::sqrt(2.0 * m + 3.0);  // a_{m+1,m}
    for (int ll = m + 2; ll <= l; ++ll) {
        double a = std::sqrt((4.0 * ll * ll - 1.0) / ((double)ll * ll - (double)m * m));
        double p2 = a * (x * p1 - p0 / a_prev);
        p0 = p1;
        p1 = p2;
        a_prev = a;
    }
    return p1;
}

// f = sum of the 16 modes of `seed` at points (theta_j, phi_i), returned on
// an (nrow x ncol) grid given by the coordinate lists.
std::vector<double> mode_sum(uint64_t seed, int lmax, const std::vector<double>& th,
                             const std::vector<double>& ph) {
    const size_t nr = th.size(), nc = ph.size();
    std::vector<double> f(nr * nc, 0.0), leg(nr), trig(nc);
    for (const Mode& md : draw_modes(seed, lmax)) {
        const int am = std::abs(md.m);
        for (size_t j = 0; j < nr; ++j) leg[j] = pbar(md.l, am, th[j]);
        const double w = md.m == 0 ? 1.0 : std::sqrt(2.0);
        for (size_t i = 0; i < nc; ++i)
            trig[i] = md.m > 0 ? std::cos(am * ph[i]) : (md.m < 0 ? std::sin(am * ph[i]) : 1.0);
        for (size_t j = 0; j < nr; ++j) {
            const double cj = md.a * w * leg[j];
            double* row = &f[j * nc];
            for (size_t i = 0; i < nc; ++i) row[i] += cj * trig[i];
        }
    }
    return f;
}

}  // namespace

void derive_params(const sf_bubble_desc& d, sf_derived& o) {
    // Inside Bubble's initialiser list the parameter U shadows the member
    // (forced to 1.0); run() uses the member, so UGlobal = 1 and the
    // evaporation rate is evaporation * radius / (H * 1).
    const double sigma_r = kR * d.T;
    o.M = d.Gamma_m * sigma_r / (3 * kRho * d.H * d.U * d.U) * d.M_scale;
    o.Cr = kRhoA * std::sqrt(kNuA * d.radius / d.U) / (kRho * d.H);
    o.gs = kG * d.radius / (d.U * d.U);
    o.gamma_max = kSigmaA / sigma_r / d.Gamma_m;
    o.dt = d.dt * d.U / d.radius;
    o.radius = d.radius;
    o.U = 1.0;
    o.eva = d.evaporation ? kEvaporation * d.radius / (d.H * 1.0) : 0.0;
    o.h = kPi / d.n_theta;
    o.re = kNu / (d.U * d.radius);   // re(nu / (U radius)) with the U argument
}

void air_shapes(int nth, double radius, double U, double* su, double* sv) {
    const int nph = 2 * nth;
    const double h = kPi / nth;
    for (int j = 0; j < nth; ++j) {
        const double th = ((double)j + 0.5) * h;
        const double sT = std::sin(th), cT = std::cos(th);
        for (int i = 0; i < nph; ++i) {
            const double ph = ((double)i - 0.5) * h;
            const double sP = std::sin(ph), cP = std::cos(ph);
            V3 a = corridor(sT, cT, sP, cP, U, radius);
            su[(size_t)j * nph + i] = dot(a, V3{-sP, cP, 0.});
        }
    }
    for (int j = 0; j < nth - 1; ++j) {
        const double th = ((double)j + 1.0) * h;
        const double sT = std::sin(th), cT = std::cos(th);
        for (int i = 0; i < nph; ++i) {
            const double ph = (double)i * h;
            const double sP = std::sin(ph), cP = std::cos(ph);
            V3 a = corridor(sT, cT, sP, cP, U, radius);
            sv[(size_t)j * nph + i] = dot(a, V3{cT * cP, cT * sP, -sT});
        }
    }
}

double real_ylm(int l, int m, double theta, double phi) {
    const int am = std::abs(m);
    if (am > l || l < 0) return 0.0;
    const double p = pbar(l, am, theta);
    if (m == 0) return p;
    return std::sqrt(2.0) * p * (m > 0 ? std::cos(am * phi) : std::sin(am * phi));
}

void synth_fields(int nth, uint64_t seed, int lmax, int lair, double noise, double air_amp,
                  double* eta, double* gam, double* uair0, double* vair0) {
    const int nph = 2 * nth;
    const double h = kPi / nth;
    const size_t N = (size_t)nth * nph;
    std::vector<double> thc(nth), phc(nph), thn(nth + 1), phw(nph);
    for (int j = 0; j < nth; ++j) thc[j] = ((double)j + 0.5) * h;
    for (int j = 0; j <= nth; ++j) thn[j] = (double)j * h;
    for (int i = 0; i < nph; ++i) {
        phc[i] = (double)i * h;
        phw[i] = ((double)i - 0.5) * h;
    }
    auto normalised = [&](uint64_t s) {
        std::vector<double> f = mode_sum(s, lmax, thc, phc);
        double mx = 0.0;
        for (double v : f) mx = std::max(mx, std::fabs(v));
        if (mx > 0)
            for (double& v : f) v /= mx;
        return f;
    };
    if (eta) {
        std::vector<double> f = normalised(seed);
        Rng r(seed + 3);
        for (size_t k = 0; k < N; ++k) {
            double v = 1.0 + 0.1 * f[k];
            if (noise != 0.0) v += noise * (2.0 * r.uniform() - 1.0);
            eta[k] = v;
        }
    }
    if (gam) {
        std::vector<double> f = normalised(seed + 1);
        for (size_t k = 0; k < N; ++k) gam[k] = 1.0 + 0.05 * f[k];
    }
    if (uair0 || vair0) {
        // stream function at cell corners (theta = j h, phi = (i - 1/2) h);
        // u_phi = d(psi)/d(theta), u_theta = -d(psi)/d(phi) / sin(theta) as
        // differences across the face, so the discrete MAC divergence of
        // (u_theta, u_phi) vanishes identically.
        std::vector<double> psi = mode_sum(seed + 2, lair, thn, phw);
        std::vector<double> u(N, 0.0), v(N - nph, 0.0);
        for (int j = 0; j < nth; ++j)
            for (int i = 0; i < nph; ++i)
                u[(size_t)j * nph + i] = (psi[(size_t)(j + 1) * nph + i] - psi[(size_t)j * nph + i]) / h;
        for (int j = 0; j < nth - 1; ++j) {
            const double s = std::sin((double)(j + 1) * h);
            for (int i = 0; i < nph; ++i) {
                const int ie = (i + 1) % nph;
                v[(size_t)j * nph + i] =
                    -(psi[(size_t)(j + 1) * nph + ie] - psi[(size_t)(j + 1) * nph + i]) / (h * s);
            }
        }
        double mx = 0.0;
        for (int j = 0; j < nth; ++j)
            for (int i = 0; i < nph; ++i) {
                const double up = 0.5 * (u[(size_t)j * nph + i] + u[(size_t)j * nph + (i + 1) % nph]);
                const double vn = j > 0 ? v[(size_t)(j - 1) * nph + i] : 0.0;
                const double vs = j < nth - 1 ? v[(size_t)j * nph + i] : 0.0;
                const double vt = 0.5 * (vn + vs);
                mx = std::max(mx, std::sqrt(up * up + vt * vt));
            }
        const double sc = mx > 0 ? air_amp / mx : 0.0;
        if (uair0)
            for (size_t k = 0; k < N; ++k) uair0[k] = u[k] * sc;
        if (vair0)
            for (size_t k = 0; k < N - nph; ++k) vair0[k] = v[k] * sc;
    }
}

// ------------------------------------------------------------------ EXR ----
namespace {

// binary32 -> binary16 as the reference's EXR writer converts (tinyexr,
// float_to_half_full): truncate the significand and add one unit whenever
// the highest dropped bit is set (round to nearest, ties away from zero);
// binary32 subnormals become signed zero, overflow becomes infinity, NaN a
// quiet NaN.
uint16_t float_to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
    const int e = (int)((x >> 23) & 0xFFu);
    const uint32_t mant = x & 0x7FFFFFu;
    if (e == 0) return sign;
    if (e == 255) return (uint16_t)(sign | 0x7C00u | (mant ? 0x200u : 0u));
    const int ne = e - 127 + 15;
    if (ne >= 31) return (uint16_t)(sign | 0x7C00u);
    if (ne <= 0) {
        const int shift = 14 - ne;   // hidden bit included
        if (shift > 24) return sign;
        const uint32_t m = mant | 0x800000u;
        uint32_t r = m >> shift;
        if ((m >> (shift - 1)) & 1u) ++r;
        return (uint16_t)(sign | r);
    }
    uint32_t r = ((uint32_t)ne << 10) | (mant >> 13);
    if (mant & 0x1000u) ++r;   // may carry into the exponent, up to infinity
    return (uint16_t)(sign | r);
}

float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    uint32_t x;
    if (e == 0) {
        if (m == 0) {
            x = sign;
        } else {
            int k = -1;
            uint32_t mm = m;
            do { ++k; mm <<= 1; } while (!(mm & 0x400u));
            x = sign | ((uint32_t)(127 - 15 - k) << 23) | ((mm & 0x3FFu) << 13);
        }
    } else if (e == 31) {
        x = sign | 0x7F800000u | (m << 13);
    } else {
        x = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

void put32(std::string& s, int32_t v) { s.append(reinterpret_cast<const char*>(&v), 4); }
void putf(std::string& s, float v) { s.append(reinterpret_cast<const char*>(&v), 4); }
void attr(std::string& s, const char* name, const char* type, const std::string& val) {
    s += name;
    s += '\0';
    s += type;
    s += '\0';
    put32(s, (int32_t)val.size());
    s += val;
}

}  // namespace

namespace {

// The header of write_image's file (solver.cu:338-385): channels B, G, R,
// HALF, uncompressed, w x hgt, the attributes in the reference writer's
// order, then the end-of-header byte.
std::string exr_header(int w, int hgt) {
    std::string hdr;
    const uint32_t magic = 20000630u;
    hdr.append(reinterpret_cast<const char*>(&magic), 4);
    put32(hdr, 2);
    std::string ch;
    for (const char* name : {"B", "G", "R"}) {
        ch += name;
        ch += '\0';
        put32(ch, 1);  // HALF
        ch += std::string(4, '\0');
        put32(ch, 1);
        put32(ch, 1);
    }
    ch += '\0';
    attr(hdr, "channels", "chlist", ch);
    attr(hdr, "compression", "compression", std::string(1, '\0'));
    std::string box;
    put32(box, 0); put32(box, 0); put32(box, w - 1); put32(box, hgt - 1);
    attr(hdr, "dataWindow", "box2i", box);
    attr(hdr, "displayWindow", "box2i", box);
    attr(hdr, "lineOrder", "lineOrder", std::string(1, '\0'));
    std::string f1;
    putf(f1, 1.0f);
    attr(hdr, "pixelAspectRatio", "float", f1);
    std::string v2;
    putf(v2, 0.0f); putf(v2, 0.0f);
    attr(hdr, "screenWindowCenter", "v2f", v2);
    std::string fw;
    putf(fw, (float)w);   // the image width, as the reference's writer stores it
    attr(hdr, "screenWindowWidth", "float", fw);
    hdr += '\0';
    return hdr;
}

}  // namespace

bool write_image_exr(const std::string& path, int w, int hgt, const float* r, const float* g,
                     const float* b) {
    // write_image (solver.cu:338-385): channels B, G, R, HALF, uncompressed
    const std::string hdr = exr_header(w, hgt);
    // offset table, then per scanline: y, byte count, the B, G, R halves
    const uint64_t block = 8 + (uint64_t)w * 3 * 2;
    const size_t head = hdr.size() + 8ull * hgt;
    std::vector<char> out(head + block * (uint64_t)hgt);
    std::memcpy(out.data(), hdr.data(), hdr.size());
    for (int y = 0; y < hgt; ++y) {
        const uint64_t pos = head + block * (uint64_t)y;
        std::memcpy(out.data() + hdr.size() + 8ull * y, &pos, 8);
    }
    // one plane given for all three channels (a gray image) is converted once
    const bool gray = r == g && g == b;
    auto rows = [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y) {
            char* q = out.data() + head + block * (uint64_t)y;
            const int32_t yy = y, nbytes = w * 3 * 2;
            std::memcpy(q, &yy, 4);
            std::memcpy(q + 4, &nbytes, 4);
            uint16_t* hb = reinterpret_cast<uint16_t*>(q + 8);
            const float* planes[3] = {b, g, r};
            for (int k = 0; k < 3; ++k) {
                uint16_t* dst = hb + (size_t)k * w;
                if (gray && k > 0) {
                    std::memcpy(dst, hb, 2 * (size_t)w);
                    continue;
                }
                const float* src = planes[k] + (size_t)y * w;
                for (int i = 0; i < w; ++i) dst[i] = float_to_half(src[i]);
            }
        }
    };
    // large images: scanlines split over a few host threads
    const int nt = (int)std::min<long>(
        std::max(1u, std::min(8u, std::thread::hardware_concurrency())),
        std::max<long>(1, (long)w * hgt / (1L << 18)));
    if (nt <= 1) {
        rows(0, hgt);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t)
            th.emplace_back(rows, (int)((long)hgt * t / nt), (int)((long)hgt * (t + 1) / nt));
        for (auto& x : th) x.join();
    }
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f.write(out.data(), (std::streamsize)out.size());
    return (bool)f;
}

// The file write_image_exr writes for a gray image whose halves are
// `gray` (row-major, w x hgt): the header and offset table, then each
// scanline's y and byte count and its row three times (B, G, R), gathered
// straight from `gray` by writev.
bool write_half_exr(const std::string& path, int w, int hgt, const uint16_t* gray) {
    std::string head = exr_header(w, hgt);
    const uint64_t block = 8 + (uint64_t)w * 3 * 2;
    const uint64_t start = head.size() + 8ull * hgt;
    for (int y = 0; y < hgt; ++y) {
        const uint64_t pos = start + block * (uint64_t)y;
        head.append(reinterpret_cast<const char*>(&pos), 8);
    }
    std::vector<int32_t> lh(2 * (size_t)hgt);
    for (int y = 0; y < hgt; ++y) {
        lh[2 * (size_t)y] = y;
        lh[2 * (size_t)y + 1] = w * 3 * 2;
    }
    const int fd = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) return false;
    std::vector<iovec> iov;
    iov.reserve(1 + 4 * (size_t)hgt);
    iov.push_back(iovec{head.data(), head.size()});
    for (int y = 0; y < hgt; ++y) {
        void* row = const_cast<uint16_t*>(gray + (size_t)y * w);
        iov.push_back(iovec{&lh[2 * (size_t)y], 8});
        for (int k = 0; k < 3; ++k) iov.push_back(iovec{row, 2 * (size_t)w});
    }
    // writev in batches of at most IOV_MAX, resuming after a short write
    bool ok = true;
    size_t at = 0;
    while (ok && at < iov.size()) {
        const int cnt = (int)std::min<size_t>(iov.size() - at, IOV_MAX);
        const ssize_t got = ::writev(fd, iov.data() + at, cnt);
        if (got < 0) {
            if (errno == EINTR) continue;
            ok = false;
            break;
        }
        size_t left = (size_t)got;
        while (left > 0 && at < iov.size()) {
            if (left >= iov[at].iov_len) {
                left -= iov[at].iov_len;
                ++at;
            } else {
                iov[at].iov_base = static_cast<char*>(iov[at].iov_base) + left;
                iov[at].iov_len -= left;
                left = 0;
            }
        }
        if (got == 0 && cnt > 0) ok = false;
    }
    if (::close(fd) != 0) ok = false;
    return ok;
}

bool write_thickness_exr(const std::string& path, int nth, int nph, const double* eta, double H) {
    // write_thickness_img (solver.cu:560-582): eta * H * 2e5 in all three
    // channels
    std::vector<float> gray((size_t)nth * nph);
    for (size_t k = 0; k < gray.size(); ++k) gray[k] = (float)(eta[k] * H * 2e5);
    return write_image_exr(path, nph, nth, gray.data(), gray.data(), gray.data());
}

// Every read is checked against the file (and an attribute's reads against
// the attribute's own size): a truncated or corrupted file returns false,
// never a read outside the buffer or a loop that does not advance.
bool read_exr_channel(const std::string& path, const std::string& channel, int* width,
                      int* height, std::vector<float>* data) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::string buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    const size_t n = buf.size();
    size_t p = 0;
    auto rd32 = [&](int32_t& v) {
        if (p > n || n - p < 4) return false;
        std::memcpy(&v, buf.data() + p, 4);
        p += 4;
        return true;
    };
    // a NUL-terminated string that ends before `lim`
    auto rdstr = [&](std::string& s, size_t lim) {
        if (p >= lim) return false;
        size_t e = buf.find('\0', p);
        if (e == std::string::npos || e >= lim) return false;
        s = buf.substr(p, e - p);
        p = e + 1;
        return true;
    };
    int32_t magic, version;
    if (!rd32(magic) || magic != 20000630 || !rd32(version) || (version & 0xFF) != 2) return false;
    if (version & 0x1E00) return false;  // tiled / long names / deep / multipart unsupported
    struct Chan { std::string name; int type; };
    std::vector<Chan> chans;
    int comp = -1, x0 = 0, y0 = 0, x1 = -1, y1 = -1;
    bool have_window = false;
    for (;;) {
        std::string name, type;
        if (!rdstr(name, n)) return false;
        if (name.empty()) break;
        int32_t size;
        if (!rdstr(type, n) || !rd32(size) || size < 0 || (size_t)size > n - p) return false;
        const size_t end = p + (size_t)size;
        if (name == "channels") {
            // name, pixel type (0 uint, 1 half, 2 float), pLinear, reserved,
            // xSampling, ySampling; a NUL ends the list
            while (p < end && buf[p] != '\0') {
                Chan c;
                if (!rdstr(c.name, end) || end - p < 16) return false;
                int32_t t, xs, ys;
                std::memcpy(&t, buf.data() + p, 4);
                std::memcpy(&xs, buf.data() + p + 8, 4);
                std::memcpy(&ys, buf.data() + p + 12, 4);
                if (t < 0 || t > 2 || xs != 1 || ys != 1) return false;
                c.type = t;
                p += 16;
                chans.push_back(c);
            }
        } else if (name == "compression") {
            if (size < 1) return false;
            comp = (unsigned char)buf[p];
        } else if (name == "dataWindow") {
            if (size < 16) return false;
            std::memcpy(&x0, buf.data() + p, 4);
            std::memcpy(&y0, buf.data() + p + 4, 4);
            std::memcpy(&x1, buf.data() + p + 8, 4);
            std::memcpy(&y1, buf.data() + p + 12, 4);
            have_window = true;
        }
        p = end;
    }
    if (comp != 0 || chans.empty() || !have_window) return false;
    const int64_t w64 = (int64_t)x1 - x0 + 1, h64 = (int64_t)y1 - y0 + 1;
    if (w64 <= 0 || h64 <= 0) return false;
    std::sort(chans.begin(), chans.end(), [](const Chan& a, const Chan& b) { return a.name < b.name; });
    int ci = -1;
    uint64_t chan_off = 0, pix_bytes = 0;
    for (size_t k = 0; k < chans.size(); ++k) {
        if (chans[k].name == channel) { ci = (int)k; chan_off = pix_bytes; }
        pix_bytes += chans[k].type == 1 ? 2 : 4;
    }
    if (ci < 0) return false;
    // every scanline holds its 8-byte offset, its 8-byte header and
    // pix_bytes per pixel: the image must fit in the file
    if ((uint64_t)w64 > n / pix_bytes || (uint64_t)h64 > n / 16 ||
        (uint64_t)h64 * (16 + (uint64_t)w64 * pix_bytes) > n)
        return false;
    const int w = (int)w64, h = (int)h64;
    const size_t line_bytes = (size_t)pix_bytes * (size_t)w;
    chan_off *= (uint64_t)w;
    std::vector<uint64_t> offs((size_t)h);
    if (p > n || (n - p) / 8 < (size_t)h) return false;
    std::memcpy(offs.data(), buf.data() + p, 8 * (size_t)h);
    data->assign((size_t)w * h, 0.0f);
    for (int k = 0; k < h; ++k) {
        const uint64_t q0 = offs[(size_t)k];
        int32_t y, sz;
        if (n < 8 || q0 > n - 8) return false;
        size_t q = (size_t)q0;
        std::memcpy(&y, buf.data() + q, 4);
        std::memcpy(&sz, buf.data() + q + 4, 4);
        q += 8;
        if (sz < 0 || (size_t)sz != line_bytes || line_bytes > n - q || y < y0 || y > y1)
            return false;
        const char* src = buf.data() + q + chan_off;
        float* dst = data->data() + (size_t)((int64_t)y - y0) * w;
        for (int i = 0; i < w; ++i) {
            if (chans[ci].type == 1) {
                uint16_t hv;
                std::memcpy(&hv, src + 2 * i, 2);
                dst[i] = half_to_float(hv);
            } else if (chans[ci].type == 2) {
                std::memcpy(&dst[i], src + 4 * i, 4);
            } else {
                uint32_t u;
                std::memcpy(&u, src + 4 * i, 4);
                dst[i] = (float)u;
            }
        }
    }
    *width = w;
    *height = h;
    return true;
}

}  // namespace sfh

namespace sfh {

namespace {

// Bilinear resize of one float channel, W x H -> w x h: destination pixel
// centre x + 1/2 maps to source x_s = (x + 1/2) W / w - 1/2; taps outside
// the image are clamped to the border pixel.
std::vector<float> resize_linear(const std::vector<float>& s, int W, int H, int w, int h) {
    std::vector<float> out((size_t)w * h);
    auto tap = [](double f, int n, int& i0, int& i1, double& a) {
        double fl = std::floor(f);
        a = f - fl;
        i0 = (int)fl;
        i1 = i0 + 1;
        if (i0 < 0) { i0 = i1 = 0; a = 0.0; }
        if (i1 > n - 1) { i1 = n - 1; if (i0 > n - 1) i0 = n - 1; }
        if (i0 == i1) a = 0.0;
    };
    for (int y = 0; y < h; ++y) {
        int y0, y1;
        double ay;
        tap((y + 0.5) * (double)H / h - 0.5, H, y0, y1, ay);
        for (int x = 0; x < w; ++x) {
            int x0, x1;
            double ax;
            tap((x + 0.5) * (double)W / w - 0.5, W, x0, x1, ax);
            const double lo = (1 - ax) * s[(size_t)y0 * W + x0] + ax * s[(size_t)y0 * W + x1];
            const double hi = (1 - ax) * s[(size_t)y1 * W + x0] + ax * s[(size_t)y1 * W + x1];
            out[(size_t)y * w + x] = (float)((1 - ay) * lo + ay * hi);
        }
    }
    return out;
}

bool load_channel(const char* path, const char* ch, int nph, int nth, std::vector<float>& out) {
    int W = 0, H = 0;
    std::vector<float> raw;
    if (!read_exr_channel(path, ch, &W, &H, &raw)) return false;
    out = resize_linear(raw, W, H, nph, nth);
    return true;
}

}  // namespace

void texture_fields(const char* thickness_path, const char* air_path, int nth, double radius,
                    double* eta, double* uair0, double* vair0, int* loaded) {
    const int nph = 2 * nth;
    int got = 0;
    std::vector<float> B, G, R;
    if (eta && thickness_path && *thickness_path && load_channel(thickness_path, "B", nph, nth, B)) {
        for (size_t k = 0; k < B.size(); ++k) {
            double C = (double)B[k];
            eta[k] = (C - 0.5) * 0.3 + 0.5;
        }
        got |= 1;
    }
    if ((uair0 || vair0) && air_path && *air_path && load_channel(air_path, "G", nph, nth, G) &&
        load_channel(air_path, "R", nph, nth, R)) {
        for (int j = 0; j < nth; ++j)
            for (int i = 0; i < nph; ++i) {
                const size_t c = (size_t)j * nph + i, w = (size_t)j * nph + (i + nph - 1) % nph;
                if (uair0) uair0[c] = 0.5 * (double)(G[c] + G[w]) * radius;
                if (vair0 && j < nth - 1) vair0[c] = -0.5 * (double)(R[c] + R[c + nph]) * radius;
            }
        got |= 2;
    }
    if (loaded) *loaded = got;
}

}  // namespace sfh

extern "C" {

int sf_write_half_exr(const char* path, int32_t width, int32_t height, const uint16_t* gray) {
    if (!path || !gray || width <= 0 || height <= 0) return SF_ERR_INVALID;
    return sfh::write_half_exr(path, width, height, gray) ? SF_OK : SF_ERR_INVALID;
}

int sf_write_image_exr(const char* path, int32_t width, int32_t height, const float* r,
                       const float* g, const float* b) {
    if (!path || !r || !g || !b || width <= 0 || height <= 0) return SF_ERR_INVALID;
    return sfh::write_image_exr(path, width, height, r, g, b) ? SF_OK : SF_ERR_INVALID;
}

int sf_texture_fields(const char* thickness_exr, const char* air_exr, int32_t n_theta,
                      double radius, double* eta, double* uair0, double* vair0, int32_t* loaded) {
    if (n_theta < 4) return SF_ERR_UNSUPPORTED;
    int got = 0;
    sfh::texture_fields(thickness_exr, air_exr, n_theta, radius, eta, uair0, vair0, &got);
    if (loaded) *loaded = got;
    return SF_OK;
}

int sf_synth_fields(int32_t n_theta, uint64_t seed, int32_t l_max, int32_t l_air, double noise,
                    double air_amp, double* eta, double* gamma, double* uair0, double* vair0) {
    if (n_theta < 2 || l_max < 1 || l_air < 1) return SF_ERR_INVALID;
    sfh::synth_fields(n_theta, seed, l_max, l_air, noise, air_amp, eta, gamma, uair0, vair0);
    return SF_OK;
}

double sf_real_ylm(int32_t l, int32_t m, double theta, double phi) {
    return sfh::real_ylm(l, m, theta, phi);
}

}  // extern "C"
