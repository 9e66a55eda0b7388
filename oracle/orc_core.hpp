// oracle/orc_core.hpp — CPU oracle, core math.
//
// TEST INFRASTRUCTURE ONLY.  The oracle is the checker for libaurora's GPU
// path: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
// may load it.  Product code never includes or links anything under oracle/.
//
// This is an independent CPU restatement of the path-tracing hot path, in
// plain C++ templated on the real type (float: parity with the GPU; double:
// convergence reference).  It is compiled with -ffp-contract=off so every
// float expression below rounds exactly once per operation, in the order
// written, as IEEE-754 requires; the GPU kernels are written to the same
// operation order, which is what makes bit-exact parity possible.
//
// Reference anchors (the reference has no path tracer; see each function):
//   vector algebra, reflect, refract: src/basic/vec3.cpp:6-199
//   normalised ray directions:       src/basic/ray.cpp:5
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

// ---------------------------------------------------------------- vec3 ---
// Component-wise algebra following src/basic/vec3.cpp:6-179 (the reference's
// double Vec3); dot/cross evaluate left to right as written there
// (vec3.cpp:149-158).
template <class R>
struct V3 {
    R x, y, z;
};
template <class R> inline V3<R> mk(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
// unit vector: each component divided by the length (ray.cpp:5 normalises
// every ray direction the same way; vec3.cpp:125-131 is the reference form).
template <class R> inline V3<R> unit(V3<R> a) { R l = length(a); return a / l; }
// near_zero: all |components| < 1e-8 (vec3.cpp:143-146).
template <class R> inline bool near_zero(V3<R> a) {
    const R s = R(1e-8);
    return std::fabs(a.x) < s && std::fabs(a.y) < s && std::fabs(a.z) < s;
}
// reflect(v, n) = v - 2 (v.n) n  (vec3.cpp:182-184)
template <class R> inline V3<R> reflect(V3<R> v, V3<R> n) {
    R k = R(2) * dot(v, n);
    return v - k * n;
}
// refract(uv, n, eta) (vec3.cpp:188-199 without its dead TIR branch; the
// caller decides total internal reflection before calling, as a dielectric
// must).
template <class R> inline V3<R> refract(V3<R> uv, V3<R> n, R eta) {
    R cos_t = std::fmin(dot(-uv, n), R(1));
    V3<R> perp = eta * (uv + cos_t * n);
    R par_len = -std::sqrt(std::fabs(R(1) - dot(perp, perp)));
    V3<R> par = par_len * n;
    return perp + par;
}
template <class R, class S> inline V3<R> cvt(V3<S> a) { return {R(a.x), R(a.y), R(a.z)}; }

// -------------------------------------------------------------- Philox ---
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with the
// published multipliers and Weyl key increments.  Checked against the
// Random123 known-answer vectors in tests/test_oracle_philox.py.
struct U4 { uint32_t x, y, z, w; };
inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = uint64_t(M0) * c.x;
        uint64_t p1 = uint64_t(M1) * c.z;
        uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}
// 24-bit uniform in [0,1): exactly representable in float and double.
inline float u01f(uint32_t x) { return float(x >> 8) * (1.0f / 16777216.0f); }

// Stream ids: key word 1 of every Philox draw.
enum : uint32_t {
    STREAM_RENDER = 0u,
    STREAM_SCENE = 1u,
    STREAM_PERLIN = 2u,
};

// ------------------------------------------------ deterministic sin/atan ---
// Transcendentals outside the parity-critical arithmetic would differ between
// the host libm and the device libraries, so the marble texture's sine and the
// sphere (u,v) angles use these double-precision versions written from the
// textbook series: Cody–Waite style reduction by pi/2 and Taylor polynomials
// (sin, cos), argument halving tan(x/2) = x/(1+sqrt(1+x^2)) plus the Taylor
// series of atan.  Every step is +,-,*,/,sqrt, so host and device agree
// bit-for-bit.
inline double orc_sin_d(double x) {
    const double TWO_OVER_PI = 0.63661977236758134308;
    const double PIO2_HI = 1.5707963267948966192;     // double nearest pi/2
    const double PIO2_LO = 6.123233995736766036e-17;  // pi/2 - PIO2_HI
    double k = std::nearbyint(x * TWO_OVER_PI);
    double r = (x - k * PIO2_HI) - k * PIO2_LO;
    long long q = (long long)k;
    double r2 = r * r;
    // sin r = r - r^3/3! + ... + r^21/21!  (|r| <= pi/4 + tiny)
    double s = 1.0 / 51090942171709440000.0;          // 1/21!
    s = s * r2 - 1.0 / 121645100408832000.0;          // 1/19!
    s = s * r2 + 1.0 / 355687428096000.0;             // 1/17!
    s = s * r2 - 1.0 / 1307674368000.0;               // 1/15!
    s = s * r2 + 1.0 / 6227020800.0;                  // 1/13!
    s = s * r2 - 1.0 / 39916800.0;                    // 1/11!
    s = s * r2 + 1.0 / 362880.0;                      // 1/9!
    s = s * r2 - 1.0 / 5040.0;                        // 1/7!
    s = s * r2 + 1.0 / 120.0;                         // 1/5!
    s = s * r2 - 1.0 / 6.0;                           // 1/3!
    s = s * r2 + 1.0;
    s = s * r;
    // cos r = 1 - r^2/2! + ... + r^20/20!
    double c = 1.0 / 2432902008176640000.0;           // 1/20!
    c = c * r2 - 1.0 / 6402373705728000.0;            // 1/18!
    c = c * r2 + 1.0 / 20922789888000.0;              // 1/16!
    c = c * r2 - 1.0 / 87178291200.0;                 // 1/14!
    c = c * r2 + 1.0 / 479001600.0;                   // 1/12!
    c = c * r2 - 1.0 / 3628800.0;                     // 1/10!
    c = c * r2 + 1.0 / 40320.0;                       // 1/8!
    c = c * r2 - 1.0 / 720.0;                         // 1/6!
    c = c * r2 + 1.0 / 24.0;                          // 1/4!
    c = c * r2 - 1.0 / 2.0;                           // 1/2!
    c = c * r2 + 1.0;
    switch (q & 3) {
        case 0: return s;
        case 1: return c;
        case 2: return -s;
        default: return -c;
    }
}

// atan for x >= 0
inline double orc_atan_pos(double x) {
    const double PIO2 = 1.5707963267948966192;
    bool inv = x > 1.0;
    double a = inv ? 1.0 / x : x;
    // two half-angle steps: |a| <= tan(pi/16) ~ 0.199
    a = a / (1.0 + std::sqrt(1.0 + a * a));
    a = a / (1.0 + std::sqrt(1.0 + a * a));
    double a2 = a * a;
    // atan a = a - a^3/3 + a^5/5 - ... - a^31/31   (Horner from the top)
    double p = -1.0 / 31.0;
    p = p * a2 + 1.0 / 29.0;
    p = p * a2 - 1.0 / 27.0;
    p = p * a2 + 1.0 / 25.0;
    p = p * a2 - 1.0 / 23.0;
    p = p * a2 + 1.0 / 21.0;
    p = p * a2 - 1.0 / 19.0;
    p = p * a2 + 1.0 / 17.0;
    p = p * a2 - 1.0 / 15.0;
    p = p * a2 + 1.0 / 13.0;
    p = p * a2 - 1.0 / 11.0;
    p = p * a2 + 1.0 / 9.0;
    p = p * a2 - 1.0 / 7.0;
    p = p * a2 + 1.0 / 5.0;
    p = p * a2 - 1.0 / 3.0;
    p = p * a2 + 1.0;
    double r = 4.0 * (p * a);
    return inv ? PIO2 - r : r;
}
inline double orc_atan2_d(double y, double x) {
    const double PI = 3.14159265358979323846;
    if (x == 0.0 && y == 0.0) return 0.0;
    double ay = std::fabs(y), ax = std::fabs(x);
    double r;
    if (ax >= ay) r = orc_atan_pos(ay / ax);
    else r = PI / 2.0 - orc_atan_pos(ax / ay);
    if (x < 0.0) r = PI - r;
    return y < 0.0 ? -r : r;
}
inline double orc_acos_d(double y) {
    double s = std::sqrt((1.0 - y) * (1.0 + y));
    return orc_atan2_d(s, y);
}

}  // namespace orc
