// dev_common.hpp — device-side math shared by libaurora's gfx950 kernels:
// float3 algebra, Philox4x32-10, primitive intersection, materials and
// textures.
//
// Floating-point contract: every kernel is built with -ffp-contract=off and
// correctly rounded f32 division and square root, so each expression below
// rounds once per operation in the written order.  The CPU oracle states the
// same algorithm in the same order, which is what makes per-ray and
// per-pixel results bit-identical between host and device.
//
// Reference anchors (the reference renderer has no GPU code and no path
// tracer; these are the nearest analogues):
//   vector algebra, reflect, refract  src/basic/vec3.cpp:149-199
//   unit ray directions               src/basic/ray.cpp:5
//   Möller–Trumbore (Q,u,v) triangle  src/object/triangle.cpp:82-121
//   bilinear texel sampling           src/texture.cpp:310-330
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/aurora.h"

namespace aur {

// ---------------------------------------------------------------- float3 ---
struct v3 {
    float x, y, z;
};
__host__ __device__ __forceinline__ v3 V(float x, float y, float z) { return v3{x, y, z}; }
__host__ __device__ __forceinline__ v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
__host__ __device__ __forceinline__ v3 operator*(v3 a, v3 b) { return v3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__host__ __device__ __forceinline__ v3 operator*(float s, v3 a) { return v3{s * a.x, s * a.y, s * a.z}; }
__host__ __device__ __forceinline__ v3 operator/(v3 a, float s) { return v3{a.x / s, a.y / s, a.z / s}; }
__host__ __device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ v3 cross(v3 a, v3 b) {
    return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ v3 unitv(v3 a) {
    float l = sqrtf(dot(a, a));
    return a / l;
}
__device__ __forceinline__ bool near_zero(v3 a) {
    const float s = 1e-8f;
    return fabsf(a.x) < s && fabsf(a.y) < s && fabsf(a.z) < s;
}
__device__ __forceinline__ v3 reflect(v3 v, v3 n) {
    float k = 2.0f * dot(v, n);
    return v - k * n;
}
__device__ __forceinline__ v3 refract(v3 uv, v3 n, float eta) {
    float c = fminf(dot(-uv, n), 1.0f);
    v3 perp = eta * (uv + c * n);
    float pl = -sqrtf(fabsf(1.0f - dot(perp, perp)));
    return perp + pl * n;
}
__device__ __forceinline__ v3 xyz(float4 a) { return v3{a.x, a.y, a.z}; }
__device__ __forceinline__ v3 clamp01(v3 a) {
    return v3{fminf(fmaxf(a.x, 0.0f), 1.0f), fminf(fmaxf(a.y, 0.0f), 1.0f), fminf(fmaxf(a.z, 0.0f), 1.0f)};
}

// ---------------------------------------------------------------- Philox ---
// Philox4x32-10 (Salmon et al., SC'11).  Counter = (pixel, sample, bounce,
// draw), key = (seed, stream).
struct u4 {
    uint32_t x, y, z, w;
};
__host__ __device__ __forceinline__ u4 philox(u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = u4{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}
__host__ __device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

enum : uint32_t { STREAM_RENDER = 0u, STREAM_SCENE = 1u, STREAM_PERLIN = 2u };

// ------------------------------------------------------- 8-bit resolve ---
// One channel of a radiance sum over spp samples to a byte.  The mean is
// sum / spp (one correctly rounded division); NaN counts as 0.
//   GAMMA2  : "Ray Tracing in One Weekend" write_color: sqrt (0 for x <= 0),
//             clamped to [0, 0.999], times 256, truncated.
//   GAMMA22 : the writer of experiments/rt.cpp:377-392: clamp to [0, 1],
//             pow(c, 1/2.2f) times 255 in float, truncated.  The power is
//             evaluated in double and rounded once to float (the correctly
//             rounded powf except within 2^-29 of a float midpoint), so host
//             and device agree.
//   LINEAR  : are::Texture::save_texture (src/texture.cpp:378-391):
//             clamp(c * 255.0, 0, 255) in double, truncated.
__host__ __device__ __forceinline__ uint8_t resolve_byte(float sum, float spp, uint32_t transfer) {
    float x = sum / spp;
    if (!(x == x)) x = 0.0f;
    if (transfer == AURORA_TRANSFER_GAMMA22) {
        const float c = fminf(fmaxf(x, 0.0f), 1.0f);
        const float g = (float)pow((double)c, (double)(1.0f / 2.2f));
        return (uint8_t)(g * 255.0f);
    }
    if (transfer == AURORA_TRANSFER_LINEAR) {
        const double y = (double)x * 255.0;
        return (uint8_t)(y < 0.0 ? 0.0 : y > 255.0 ? 255.0 : y);
    }
    const float g = x > 0.0f ? sqrtf(x) : 0.0f;
    return (uint8_t)(256.0f * fminf(fmaxf(g, 0.0f), 0.999f));
}

// ------------------------------------------------ sin / atan2 / acos (f64) ---
// Written from the series definitions (reduction by pi/2 with a two-part
// constant, Taylor polynomials; two half-angle steps then the atan series),
// in double so that host and device produce identical bits.
__device__ __forceinline__ double dsin(double x) {
    const double k = rint(x * 0.63661977236758134308);
    const double r = (x - k * 1.5707963267948966192) - k * 6.123233995736766036e-17;
    const long long q = (long long)k;
    const double r2 = r * r;
    double s = 1.0 / 51090942171709440000.0;
    s = s * r2 - 1.0 / 121645100408832000.0;
    s = s * r2 + 1.0 / 355687428096000.0;
    s = s * r2 - 1.0 / 1307674368000.0;
    s = s * r2 + 1.0 / 6227020800.0;
    s = s * r2 - 1.0 / 39916800.0;
    s = s * r2 + 1.0 / 362880.0;
    s = s * r2 - 1.0 / 5040.0;
    s = s * r2 + 1.0 / 120.0;
    s = s * r2 - 1.0 / 6.0;
    s = s * r2 + 1.0;
    s = s * r;
    double c = 1.0 / 2432902008176640000.0;
    c = c * r2 - 1.0 / 6402373705728000.0;
    c = c * r2 + 1.0 / 20922789888000.0;
    c = c * r2 - 1.0 / 87178291200.0;
    c = c * r2 + 1.0 / 479001600.0;
    c = c * r2 - 1.0 / 3628800.0;
    c = c * r2 + 1.0 / 40320.0;
    c = c * r2 - 1.0 / 720.0;
    c = c * r2 + 1.0 / 24.0;
    c = c * r2 - 1.0 / 2.0;
    c = c * r2 + 1.0;
    switch (q & 3) {
        case 0: return s;
        case 1: return c;
        case 2: return -s;
        default: return -c;
    }
}
__device__ __forceinline__ double datan_pos(double x) {
    const bool inv = x > 1.0;
    double a = inv ? 1.0 / x : x;
    a = a / (1.0 + sqrt(1.0 + a * a));
    a = a / (1.0 + sqrt(1.0 + a * a));
    const double a2 = a * a;
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
    const double r = 4.0 * (p * a);
    return inv ? 1.5707963267948966192 - r : r;
}
__device__ __forceinline__ double datan2(double y, double x) {
    if (x == 0.0 && y == 0.0) return 0.0;
    const double ay = fabs(y), ax = fabs(x);
    double r = ax >= ay ? datan_pos(ay / ax) : 3.14159265358979323846 / 2.0 - datan_pos(ax / ay);
    if (x < 0.0) r = 3.14159265358979323846 - r;
    return y < 0.0 ? -r : r;
}
__device__ __forceinline__ double dacos(double y) { return datan2(sqrt((1.0 - y) * (1.0 + y)), y); }

// ----------------------------------------------------------- scene view ---
// Device texture record: aurora_texture with `even` reused as the Perlin
// table slot of a noise texture.
// Loads through the global address space.  Scene arrays reach the kernels
// through tables in device memory, so the compiler cannot tell their
// pointers from LDS pointers and would emit flat loads, which are slower to
// issue and make every wait also wait for the LDS stack traffic.
typedef float aur_f4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const aur_f4v* aur_g4p;
typedef __attribute__((address_space(1))) const uint32_t* aur_g1p;
__device__ __forceinline__ float4 ldg4(const float4* p, size_t i) {
    const aur_f4v v = ((aur_g4p)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint32_t ldg1(const uint32_t* p, size_t i) { return ((aur_g1p)p)[i]; }
// Loads at a wave-uniform address from an array that no kernel writes while
// the reading kernel runs.  Through the constant address space the compiler
// issues them as scalar loads: no vector memory instruction, no wait on the
// vector memory counter, and the values land in scalar registers that the
// vector instructions read directly.  The pointer and index must be
// wave-uniform (uniform_ptr, readfirstlane).
typedef __attribute__((address_space(4))) const aur_f4v* aur_c4p;
__device__ __forceinline__ float4 ldu4(const float4* p, size_t i) {
    const aur_f4v v = ((aur_c4p)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
template <class T>
__device__ __forceinline__ const T* uniform_ptr(const T* p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (const T*)(((uint64_t)hi << 32) | lo);
}

struct DevScene {
    const float4* spheres;        // (centre, radius)
    const uint32_t* sphere_mat;
    const float4* quads;          // 3 per quad: (Q, mat bits) (u, kind bits) (v, 0)
    const float4* quad_n;         // per quad: (unitv(u x v), 0), computed once at upload
    const aurora_material* mats;
    const aurora_texture* texs;
    const float* texels;
    const float* perlin_vec;      // per slot: 256 x 3 floats
    const int* perlin_perm;       // per slot: 3 x 256 ints
    const float4* nodes;          // 2 float4 per node: (lo, child) (hi, parent)
    const float4* wide;           // four-wide traversal nodes: 4 slots of (lo, ref) (hi, 0)
    const float4* flat;           // small scenes: the leaf records in Morton order, (lo, leaf) (hi, parent)
    uint32_t n_spheres, n_quads, n_nodes, n_wide, n_flat;
};

// 1/x rounded as IEEE division rounds it, in 5 VALU instructions instead of
// the 11 of the division sequence: the hardware reciprocal (within one ulp)
// refined by one Newton-Raphson step on fused multiply-adds, e = 1 - x r,
// r' = r + r e (P. Markstein, IA-64 and Elementary Functions, 2000).  For
// every x with 2^-126 <= |x| < 2^126 the fast form equals 1.0f / x bit for
// bit: aurora_selftest_rcp compares all of them on the device (the sign is
// symmetric).  Other x (zero, denormal, huge, infinite or NaN), whose
// quotient may be denormal or special, take the division.
__device__ __forceinline__ float rcp_fast(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(__builtin_fmaf(-x, r, 1.0f), r, r);
}
__device__ __forceinline__ float rcp_rn(float x) {
    float y = rcp_fast(x);
    const float ax = fabsf(x);
    if (!(ax >= 0x1p-126f && ax < 0x1p126f)) {
        // an empty volatile statement keeps the division behind the branch
        // (lanes with a normal x skip it); otherwise the compiler runs it for
        // every lane and selects
        __asm__ volatile("");
        y = 1.0f / x;
    }
    return y;
}
// rcp_rn for a caller that discards the quotient when |x| < 1e-20 (the quad
// test's determinant): below 2^-126 any value will do, so only the top of the
// verified range is checked.  Equal to rcp_rn for every |x| >= 2^-126 and NaN.
__device__ __forceinline__ float rcp_rn_unless_tiny(float x) {
    float y = rcp_fast(x);
    if (!(fabsf(x) < 0x1p126f)) {
        __asm__ volatile("");
        y = 1.0f / x;
    }
    return y;
}

// Ray/sphere roots without the cancellations of the textbook quadratic
// (Haines, Guenther and Akenine-Moeller, "Precision Improvements for
// Ray/Sphere Intersection", Ray Tracing Gems, 2019): the discriminant
// h^2 - a*c is formed as a*r^2 - |d x oc|^2 (Lagrange's identity), whose
// terms are of the size of r^2 instead of |oc|^2, and the roots as c/q and
// q/a with q = h + sign(h)*sqrt(disc), which adds terms of one sign.  The
// textbook form loses every digit of the discriminant for a small sphere
// seen from afar and then reports hits outside the sphere.  The near root
// (c/q when q > 0, else q/a) is taken when it lies in (tmin, tmax), else the
// far one; each is one division of a selected numerator and denominator.
__device__ __forceinline__ bool hit_sphere(float4 sp, v3 o, v3 d, float tmin, float tmax, float& t) {
    const v3 oc = V(sp.x, sp.y, sp.z) - o;
    const float a = dot(d, d);
    const float h = dot(d, oc);
    const float rr = sp.w * sp.w;
    const float c = dot(oc, oc) - rr;
    const v3 x = cross(d, oc);
    const float disc = a * rr - dot(x, x);
    if (!(disc >= 0.0f)) return false;
    const float q = h + copysignf(sqrtf(disc), h);
    if (q == 0.0f) return false;  // tangent at the origin
    const bool pos = q > 0.0f;
    float root = (pos ? c : q) / (pos ? q : a);
    if (!(root > tmin && root < tmax)) {
        root = (pos ? q : c) / (pos ? a : q);
        if (!(root > tmin && root < tmax)) return false;
    }
    t = root;
    return true;
}

// hit_sphere for the traversals with its early returns folded into one
// predicate (same operations, same decisions); the far root is still
// computed only by lanes that need it, which most waves skip entirely.
__device__ __forceinline__ bool hit_sphere_t(float4 sp, v3 o, v3 d, float tmin, float tmax, float& t) {
    const v3 oc = V(sp.x, sp.y, sp.z) - o;
    const float a = dot(d, d);
    const float h = dot(d, oc);
    const float rr = sp.w * sp.w;
    const float c = dot(oc, oc) - rr;
    const v3 x = cross(d, oc);
    const float disc = a * rr - dot(x, x);
    const float q = h + copysignf(sqrtf(disc), h);
    const bool cand = (disc >= 0.0f) & !(q == 0.0f);
    const bool pos = q > 0.0f;
    float root = (pos ? c : q) / (pos ? q : a);
    bool in = (root > tmin) & (root < tmax);
    if (cand & !in) {
        root = (pos ? q : c) / (pos ? a : q);
        in = (root > tmin) & (root < tmax);
    }
    t = root;
    return cand & in;
}

// Möller–Trumbore on the (Q, u, v) record: alpha multiplies u, beta v;
// inclusive edges as in src/object/triangle.cpp:98-111.
__device__ __forceinline__ bool hit_quad(float4 q0, float4 q1, float4 q2, v3 o, v3 d, float tmin, float tmax,
                                         float& t, float& al, float& be) {
    v3 u = xyz(q1), v = xyz(q2);
    v3 h = cross(d, v);
    float a = dot(u, h);
    if (fabsf(a) < 1e-20f) return false;
    float f = rcp_rn(a);
    v3 s = o - xyz(q0);
    float alpha = f * dot(s, h);
    if (alpha < 0.0f || alpha > 1.0f) return false;
    v3 qv = cross(s, u);
    float beta = f * dot(d, qv);
    if (beta < 0.0f) return false;
    const bool tri = __float_as_uint(q1.w) == AURORA_QUAD_TRIANGLE;
    if (tri ? (alpha + beta > 1.0f) : (beta > 1.0f)) return false;
    float tt = f * dot(v, qv);
    if (!(tt > tmin && tt < tmax)) return false;
    t = tt;
    al = alpha;
    be = beta;
    return true;
}

// hit_quad's test for the traversals, without branches: the same operations
// and the same accept/reject decisions (every comparison written as in
// hit_quad, so NaNs decide alike), all evaluated.  A branch per early return
// costs exec-mask saves and restores on every lane and saves VALU work only
// when no lane of the wave goes on.
__device__ __forceinline__ bool hit_quad_t(float4 q0, float4 q1, float4 q2, v3 o, v3 d, float tmin, float tmax,
                                           float& t) {
    const v3 u = xyz(q1), v = xyz(q2);
    const v3 h = cross(d, v);
    const float a = dot(u, h);
    // a quotient for |a| < 1e-20 is never used: the test rejects those and t
    // counts only on a hit
    const float f = rcp_rn_unless_tiny(a);
    const v3 s = o - xyz(q0);
    const float alpha = f * dot(s, h);
    const v3 qv = cross(s, u);
    const float beta = f * dot(d, qv);
    const bool tri = __float_as_uint(q1.w) == AURORA_QUAD_TRIANGLE;
    const float tt = f * dot(v, qv);
    const bool edge = tri ? (alpha + beta > 1.0f) : (beta > 1.0f);
    t = tt;
    return !(fabsf(a) < 1e-20f) & !(alpha < 0.0f) & !(alpha > 1.0f) & !(beta < 0.0f) & !edge & (tt > tmin) &
           (tt < tmax);
}

// slab test against (tmin, tcur]; fminf/fmaxf are IEEE minNum/maxNum.
__device__ __forceinline__ bool slab(float4 lo, float4 hi, v3 o, v3 inv, float tmin, float tcur, float& tnear) {
    float tx0 = (lo.x - o.x) * inv.x, tx1 = (hi.x - o.x) * inv.x;
    float ty0 = (lo.y - o.y) * inv.y, ty1 = (hi.y - o.y) * inv.y;
    float tz0 = (lo.z - o.z) * inv.z, tz1 = (hi.z - o.z) * inv.z;
    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tcur));
    tnear = tn;
    return tn <= tf;
}

// slab() for a walk whose tcur is never NaN (it starts at a tmax that is not
// NaN and only takes hit distances): the same decision and tnear.  tcur
// enters by a compare instead of a min, so no NaN-quieting max is issued on
// it per box.  For tn not NaN, tn <= fminf(a, tcur) is !(tn > a) & (tn <= tcur)
// whether or not a is NaN; for tn NaN both forms are false.  tmin must be
// wave-uniform (it is read from a scalar register) and not a signalling NaN.
__device__ __forceinline__ bool slab_tcur(float4 lo, float4 hi, v3 o, v3 inv, float tmin, float tcur, float& tnear) {
    float tx0 = (lo.x - o.x) * inv.x, tx1 = (hi.x - o.x) * inv.x;
    float ty0 = (lo.y - o.y) * inv.y, ty1 = (hi.y - o.y) * inv.y;
    float tz0 = (lo.z - o.z) * inv.z, tz1 = (hi.z - o.z) * inv.z;
    // fmaxf(min z, tmin) by one v_max_f32 with tmin from a scalar register: no
    // input here is a signalling NaN (products, and a tmin the host made), so
    // the instruction's IEEE-mode result is fmaxf's
    float mzt;
    __asm__("v_max_f32 %0, %1, %2" : "=v"(mzt) : "s"(tmin), "v"(fminf(tz0, tz1)));
    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), mzt);
    float a = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    tnear = tn;
    return !(tn > a) & (tn <= tcur);
}

// the box of primitive id as the LBVH stores it in the primitive's leaf
__device__ __forceinline__ void prim_box(const DevScene& sc, uint32_t id, v3& lo, v3& hi) {
    if (id < sc.n_spheres) {
        float4 s = sc.spheres[id];
        lo = V(s.x - s.w, s.y - s.w, s.z - s.w);
        hi = V(s.x + s.w, s.y + s.w, s.z + s.w);
    } else {
        const float4* q = sc.quads + 3 * (size_t)(id - sc.n_spheres);
        float4 q0 = q[0], q1 = q[1], q2 = q[2];
        const bool para = __float_as_uint(q1.w) == AURORA_QUAD_PARALLELOGRAM;
        float a[3] = {q0.x, q0.y, q0.z}, u[3] = {q1.x, q1.y, q1.z}, v[3] = {q2.x, q2.y, q2.z};
        float l[3], h[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            float p0 = a[k], p1 = a[k] + u[k], p2 = a[k] + v[k];
            float lk = fminf(p0, fminf(p1, p2)), hk = fmaxf(p0, fmaxf(p1, p2));
            if (para) {
                float p3 = (a[k] + u[k]) + v[k];
                lk = fminf(lk, p3);
                hk = fmaxf(hk, p3);
            }
            l[k] = lk;
            h[k] = hk;
        }
        lo = V(l[0], l[1], l[2]);
        hi = V(h[0], h[1], h[2]);
    }
    // pad by 2^-16 of the largest coordinate magnitude (planar quads get
    // thickness; slab-test rounding cannot cull a true hit)
    float m = fmaxf(fmaxf(fabsf(lo.x), fabsf(hi.x)), 0.0f);
    m = fmaxf(m, fmaxf(fabsf(lo.y), fabsf(hi.y)));
    m = fmaxf(m, fmaxf(fabsf(lo.z), fabsf(hi.z)));
    float e = m * (1.0f / 65536.0f) + 1e-30f;
    lo = V(lo.x - e, lo.y - e, lo.z - e);
    hi = V(hi.x + e, hi.y + e, hi.z + e);
}

// ------------------------------------------------------------- textures ---
__device__ __forceinline__ float perlin_noise(const DevScene& sc, uint32_t slot, v3 p) {
    const float* rv = sc.perlin_vec + (size_t)slot * 768;
    const int* pm = sc.perlin_perm + (size_t)slot * 768;
    float fx = floorf(p.x), fy = floorf(p.y), fz = floorf(p.z);
    float u = p.x - fx, v = p.y - fy, w = p.z - fz;
    int i = (int)fminf(fmaxf(fx, -1073741824.0f), 1073741824.0f);
    int j = (int)fminf(fmaxf(fy, -1073741824.0f), 1073741824.0f);
    int k = (int)fminf(fmaxf(fz, -1073741824.0f), 1073741824.0f);
    float uu = (u * u) * (3.0f - 2.0f * u);
    float vv = (v * v) * (3.0f - 2.0f * v);
    float ww = (w * w) * (3.0f - 2.0f * w);
    float acc = 0.0f;
#pragma unroll
    for (int di = 0; di < 2; ++di)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int dk = 0; dk < 2; ++dk) {
                int idx = pm[(i + di) & 255] ^ pm[256 + ((j + dj) & 255)] ^ pm[512 + ((k + dk) & 255)];
                v3 g = V(rv[3 * idx], rv[3 * idx + 1], rv[3 * idx + 2]);
                v3 wv = V(u - (float)di, v - (float)dj, w - (float)dk);
                float wx = di ? uu : 1.0f - uu;
                float wy = dj ? vv : 1.0f - vv;
                float wz = dk ? ww : 1.0f - ww;
                acc = acc + ((wx * wy) * wz) * dot(g, wv);
            }
    return acc;
}

__device__ __forceinline__ float perlin_turb(const DevScene& sc, uint32_t slot, v3 p) {
    float acc = 0.0f, wt = 1.0f;
    v3 tp = p;
    for (int i = 0; i < 7; ++i) {
        acc = acc + wt * perlin_noise(sc, slot, tp);
        wt = wt * 0.5f;
        tp = 2.0f * tp;
    }
    return fabsf(acc);
}

__device__ __forceinline__ v3 tex_leaf(const DevScene& sc, const aurora_texture& t, float u, float v, v3 p) {
    if (t.type == AURORA_TEX_SOLID) return V(t.color[0], t.color[1], t.color[2]);
    if (t.type == AURORA_TEX_NOISE) {
        float arg = t.scale * p.z + 10.0f * perlin_turb(sc, t.even, p);
        float sn = (float)dsin((double)arg);
        float g = 0.5f * (1.0f + sn);
        return V(g, g, g);
    }
    // image: bilinear, (u, v) -> x = u (w-1), y = (1-v) (h-1)
    float cu = fminf(fmaxf(u, 0.0f), 1.0f);
    float cv = 1.0f - fminf(fmaxf(v, 0.0f), 1.0f);
    float x = cu * (float)(t.width - 1);
    float y = cv * (float)(t.height - 1);
    int ix = (int)floorf(x), iy = (int)floorf(y);
    int ix1 = min(ix + 1, (int)t.width - 1), iy1 = min(iy + 1, (int)t.height - 1);
    float tx = x - (float)ix, ty = y - (float)iy;
    const float* base = sc.texels + 3 * t.texel_offset;
    auto tex = [&](int xx, int yy) {
        const float* c = base + 3 * ((size_t)yy * t.width + (size_t)xx);
        return V(c[0], c[1], c[2]);
    };
    v3 c00 = tex(ix, iy), c10 = tex(ix1, iy), c01 = tex(ix, iy1), c11 = tex(ix1, iy1);
    float w00 = (1.0f - tx) * (1.0f - ty), w10 = tx * (1.0f - ty), w01 = (1.0f - tx) * ty, w11 = tx * ty;
    return ((w00 * c00 + w10 * c10) + w01 * c01) + w11 * c11;
}

__device__ __forceinline__ v3 texture_value(const DevScene& sc, uint32_t ti, float u, float v, v3 p) {
    aurora_texture t = sc.texs[ti];
    if (t.type == AURORA_TEX_CHECKER) {
        float inv = rcp_rn(t.scale);
        int cx = (int)fminf(fmaxf(floorf(inv * p.x), -1073741824.0f), 1073741824.0f);
        int cy = (int)fminf(fmaxf(floorf(inv * p.y), -1073741824.0f), 1073741824.0f);
        int cz = (int)fminf(fmaxf(floorf(inv * p.z), -1073741824.0f), 1073741824.0f);
        int sum = cx + cy + cz;
        t = sc.texs[(sum & 1) == 0 ? t.even : t.odd];
    }
    return tex_leaf(sc, t, u, v, p);
}

__device__ __forceinline__ bool tex_needs_uv(const DevScene& sc, uint32_t ti) {
    const aurora_texture& t = sc.texs[ti];
    if (t.type == AURORA_TEX_IMAGE) return true;
    if (t.type == AURORA_TEX_CHECKER)
        return sc.texs[t.even].type == AURORA_TEX_IMAGE || sc.texs[t.odd].type == AURORA_TEX_IMAGE;
    return false;
}

__device__ __forceinline__ void sphere_uv(v3 n, float& u, float& v) {
    const double PI = 3.14159265358979323846;
    double theta = dacos(-(double)n.y);
    double phi = datan2(-(double)n.z, (double)n.x) + PI;
    u = (float)(phi / (2.0 * PI));
    v = (float)(theta / PI);
}

}  // namespace aur
