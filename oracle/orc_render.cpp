// oracle/orc_render.cpp — CPU oracle: camera, path integrator, thread pool
// and the oracle's C ABI (oracle.h).  TEST INFRASTRUCTURE ONLY (see
// orc_core.hpp): loaded by tests/, __graft_entry__.smoke() and bench.py's
// CPU-baseline leg, never by libaurora.
//
// Reference anchors:
//   camera basis and pixel mapping: experiments/rt.cpp:339-343, 364-366
//     (forward/right/up from target - position and vup; pixel centre x+0.5),
//     restated in the "Ray Tracing in One Weekend" camera form (lookfrom,
//     lookat, vup, vfov, defocus disk) that the host API keeps.
//   per-pixel loop: experiments/rt.cpp:337-374; trace: rt.cpp:251-334 (the
//     reference's estimator is AO + cosine bounces + Russian roulette; this
//     restates the hit -> emitted -> scatter -> attenuate loop instead).
//   reflect/refract: src/basic/vec3.cpp:182-199.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "oracle.h"
#include "orc_scene.hpp"

namespace orc {

struct Cam {
    V3<float> center, p00, du, dv, disk_u, disk_v;
    bool defocus;
    uint32_t W, H;
};

// the tree the renderer walks: 0 the binary LBVH, 1 the four-wide tree (as
// aurora_set_render_tree selects on the GPU)
static std::atomic<int> g_render_wide{0};

// when set (on this thread), trace_path appends per segment: origin (3),
// direction (3), hit primitive (int32 bits) and hit t
static thread_local std::vector<float>* g_seg_log = nullptr;

static std::string camera_setup(const aurora_camera_desc& c, Cam& out) {
    if (c.image_width == 0 || c.image_height == 0) return "image width and height must be positive";
    if (!(c.vfov > 0.0f && c.vfov < 180.0f)) return "vfov must be in (0, 180) degrees";
    if (!(c.focus_dist > 0.0f)) return "focus_dist must be positive";
    const double PI = 3.1415926535897932385;
    V3<double> lf = mk<double>(c.lookfrom[0], c.lookfrom[1], c.lookfrom[2]);
    V3<double> la = mk<double>(c.lookat[0], c.lookat[1], c.lookat[2]);
    V3<double> up = mk<double>(c.vup[0], c.vup[1], c.vup[2]);
    if (length(lf - la) == 0.0) return "lookfrom and lookat coincide";
    double theta = double(c.vfov) * PI / 180.0;
    double h = std::tan(theta / 2.0);
    double fd = c.focus_dist;
    double vh = 2.0 * h * fd;
    double vw = vh * (double(c.image_width) / double(c.image_height));
    V3<double> w = unit(lf - la);
    V3<double> cu = cross(up, w);
    if (length(cu) == 0.0) return "vup is parallel to the view direction";
    V3<double> u = unit(cu);
    V3<double> v = cross(w, u);
    V3<double> vu = vw * u, vv = vh * (-v);
    V3<double> du = vu / double(c.image_width), dv = vv / double(c.image_height);
    V3<double> ul = ((lf - fd * w) - vu / 2.0) - vv / 2.0;
    V3<double> p00 = ul + 0.5 * (du + dv);
    double rad = fd * std::tan((double(c.defocus_angle) / 2.0) * PI / 180.0);
    out.center = cvt<float>(lf);
    out.p00 = cvt<float>(p00);
    out.du = cvt<float>(du);
    out.dv = cvt<float>(dv);
    out.disk_u = cvt<float>(rad * u);
    out.disk_v = cvt<float>(rad * v);
    out.defocus = c.defocus_angle > 0.0f;
    out.W = c.image_width;
    out.H = c.image_height;
    return "";
}

// sin and cos of 2 pi u for u in [0, 1), u a multiple of 2^-24: the
// reduction to an eighth of a turn is exact (8u, its parts, 1 - f), then
// Taylor polynomials on [0, pi/4] (truncation error below 2e-9) evaluated
// in a fixed order, and octant symmetry.  Not libm's sin: the same float
// operations run on the GPU, so both produce the same bits.
template <class R>
static void sincos_turn(R u, R& s, R& c) {
    const R t = R(8) * u;
    const R fk = std::floor(t);
    const int k = int(fk);
    const R f = t - fk;
    const bool odd = (k & 1) != 0;
    const R a = (odd ? R(1) - f : f) * R(0.785398163f);
    const R a2 = a * a;
    const R sp = a + (a * a2) * (R(-0.166666672f) + a2 * (R(0.00833333377f) + a2 * (R(-0.000198412701f) +
                                                                                       a2 * R(2.75573188e-6f))));
    const R cp = R(1) + a2 * (R(-0.5f) + a2 * (R(0.0416666679f) + a2 * (R(-0.00138888892f) +
                                                                          a2 * (R(2.48015876e-5f) + a2 * R(-2.75573188e-7f)))));
    const R sa = odd ? cp : sp, ca = odd ? sp : cp;
    const int q = k >> 1;
    s = q == 0 ? sa : q == 1 ? ca : q == 2 ? -sa : -ca;
    c = q == 0 ? ca : q == 1 ? -sa : q == 2 ? -ca : sa;
}

template <class R>
static V3<R> random_unit(uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t seed) {
    // uniform on the unit sphere from one draw: z uniform in (-1, 1] (exact),
    // azimuth uniform (Archimedes)
    U4 r = philox4x32_10(U4{pixel, sample, bounce, 0u}, seed, STREAM_RENDER);
    R z = R(1) - R(2) * R(u01f(r.x));
    R rho = std::sqrt(std::fmax(R(0), R(1) - z * z));
    R sn, cs;
    sincos_turn<R>(R(u01f(r.y)), sn, cs);
    return mk<R>(rho * cs, rho * sn, z);
}

template <class R>
static R schlick(R cosine, R ri) {
    R r0 = (R(1) - ri) / (R(1) + ri);
    r0 = r0 * r0;
    R x = R(1) - cosine;
    R x5 = ((x * x) * (x * x)) * x;
    return r0 + (R(1) - r0) * x5;
}

template <class R>
static V3<R> clamp01(V3<R> a) {
    auto c = [](R x) { return std::fmin(std::fmax(x, R(0)), R(1)); };
    return mk<R>(c(a.x), c(a.y), c(a.z));
}

// (u,v) of a unit-sphere point: theta = acos(-y), phi = atan2(-z, x) + pi
template <class R>
static void sphere_uv(V3<R> n, R& u, R& v) {
    const double PI = 3.14159265358979323846;
    double theta = orc_acos_d(-double(n.y));
    double phi = orc_atan2_d(-double(n.z), double(n.x)) + PI;
    u = R(phi / (2.0 * PI));
    v = R(theta / PI);
}

static bool tex_needs_uv(const Scene& s, uint32_t ti) {
    const aurora_texture& t = s.texs[ti];
    if (t.type == AURORA_TEX_IMAGE) return true;
    if (t.type == AURORA_TEX_CHECKER)
        return s.texs[t.even].type == AURORA_TEX_IMAGE || s.texs[t.odd].type == AURORA_TEX_IMAGE;
    return false;
}

struct PathOut {
    uint32_t segments;
};

// First ray of one sample through pixel (i, j): the pixel centre
// p00 + i*du + j*dv jittered by U[-1/2, 1/2)^2 from Philox(pixel, sample, 0,
// 0); with depth of field the origin is a lens-disk point found by rejection
// among up to 32 draws (pixel, sample, 0, 1+k).  Follows the pixel mapping of
// experiments/rt.cpp:364-366 on the viewport of camera_setup.
template <class R>
static void camera_ray(const Cam& cam, uint32_t pixel, uint32_t i, uint32_t j, uint32_t sample, uint32_t seed,
                       V3<R>& org, V3<R>& dir) {
    U4 r = philox4x32_10(U4{pixel, sample, 0u, 0u}, seed, STREAM_RENDER);
    R ox = R(u01f(r.x)) - R(0.5), oy = R(u01f(r.y)) - R(0.5);
    V3<R> p00 = cvt<R>(cam.p00), du = cvt<R>(cam.du), dv = cvt<R>(cam.dv);
    V3<R> ps = (p00 + (R(float(i)) + ox) * du) + (R(float(j)) + oy) * dv;
    org = cvt<R>(cam.center);
    if (cam.defocus) {
        R px = R(0), py = R(0);
        for (uint32_t k = 0; k < 32; ++k) {
            U4 q = philox4x32_10(U4{pixel, sample, 0u, 1u + k}, seed, STREAM_RENDER);
            R ax = R(2) * R(u01f(q.x)) - R(1), ay = R(2) * R(u01f(q.y)) - R(1);
            if (ax * ax + ay * ay < R(1)) { px = ax; py = ay; break; }
        }
        org = (org + px * cvt<R>(cam.disk_u)) + py * cvt<R>(cam.disk_v);
    }
    dir = unit(ps - org);
}

// One camera sample through pixel (i, j): the wavefront loop of libaurora,
// restated as a single sequential loop.
template <class R>
static V3<R> trace_path(const Scene& s, const Cam& cam, const aurora_render_desc& rd, uint32_t i, uint32_t j,
                        uint32_t sample, V3<R>* alb, V3<R>* nrm, uint32_t& segments) {
    const uint32_t pixel = j * cam.W + i;
    const uint32_t seed = rd.seed;
    const R tmin = rd.tmin > 0.0f ? R(rd.tmin) : R(0.001f);
    V3<R> org, dir;
    camera_ray<R>(cam, pixel, i, j, sample, seed, org, dir);
    V3<R> L = mk<R>(0, 0, 0), beta = mk<R>(1, 1, 1);
    const V3<R> bg = mk<R>(rd.background[0], rd.background[1], rd.background[2]);
    const uint32_t ns = uint32_t(s.spheres.size());
    for (uint32_t depth = 0; depth < rd.max_depth; ++depth) {
        segments++;
        Hit<R> h = g_render_wide.load(std::memory_order_relaxed)
                       ? closest_bvh4<R>(s, org, dir, tmin, R(INFINITY), nullptr)
                       : closest_bvh<R>(s, org, dir, tmin, R(INFINITY), nullptr);
        if (g_seg_log) {
            float pf;
            const int32_t pi = h.prim;
            std::memcpy(&pf, &pi, 4);
            for (float v : {float(org.x), float(org.y), float(org.z), float(dir.x), float(dir.y), float(dir.z), pf,
                            float(h.t)})
                g_seg_log->push_back(v);
        }
        if (h.prim < 0) {
            L = L + beta * bg;
            if (depth == 0 && alb) { *alb = clamp01(bg); *nrm = mk<R>(0, 0, 0); }
            break;
        }
        uint32_t id = uint32_t(h.prim);
        V3<R> p = org + h.t * dir;
        V3<R> outward;
        uint32_t mi;
        R tu = R(0), tv = R(0);
        bool is_sphere = id < ns;
        if (is_sphere) {
            const Sphere& sp = s.spheres[id];
            outward = (p - mk<R>(sp.c[0], sp.c[1], sp.c[2])) / R(sp.r);
            mi = s.sphere_mat[id];
        } else {
            const Quad& q = s.quads[id - ns];
            outward = unit(cross(mk<R>(q.u[0], q.u[1], q.u[2]), mk<R>(q.v[0], q.v[1], q.v[2])));
            mi = s.quad_mat[id - ns];
        }
        const aurora_material& m = s.mats[mi];
        if (m.type != AURORA_MAT_DIELECTRIC && tex_needs_uv(s, m.texture)) {
            if (is_sphere) {
                sphere_uv(outward, tu, tv);
            } else {
                R tt;
                hit_quad<R>(s.quads[id - ns], org, dir, tmin, R(INFINITY), tt, tu, tv);
            }
        }
        bool front = dot(dir, outward) < R(0);
        V3<R> n = front ? outward : -outward;
        V3<R> att, nd;
        if (m.type == AURORA_MAT_DIFFUSE_LIGHT) {
            V3<R> e = texture_value<R>(s, m.texture, tu, tv, p);
            L = L + beta * e;
            if (depth == 0 && alb) { *alb = clamp01(e); *nrm = n; }
            break;
        } else if (m.type == AURORA_MAT_LAMBERTIAN) {
            att = texture_value<R>(s, m.texture, tu, tv, p);
            V3<R> sd = n + random_unit<R>(pixel, sample, depth + 1, seed);
            if (near_zero(sd)) sd = n;
            nd = unit(sd);
        } else if (m.type == AURORA_MAT_METAL) {
            att = texture_value<R>(s, m.texture, tu, tv, p);
            V3<R> rf = unit(reflect(dir, n));
            if (m.fuzz > 0.0f) rf = rf + R(m.fuzz) * random_unit<R>(pixel, sample, depth + 1, seed);
            if (depth == 0 && alb) { *alb = clamp01(att); *nrm = n; }
            if (dot(rf, n) <= R(0)) break;  // absorbed
            nd = unit(rf);
        } else {
            att = mk<R>(1, 1, 1);
            R ri = front ? R(1) / R(m.ior) : R(m.ior);
            R cos_t = std::fmin(dot(-dir, n), R(1));
            R sin_t = std::sqrt(R(1) - cos_t * cos_t);
            bool cannot = ri * sin_t > R(1);
            U4 q = philox4x32_10(U4{pixel, sample, depth + 1, 0u}, seed, STREAM_RENDER);
            R uu = R(u01f(q.x));
            V3<R> d2 = (cannot || schlick(cos_t, ri) > uu) ? reflect(dir, n) : refract(dir, n, ri);
            nd = unit(d2);
        }
        if (depth == 0 && alb) { *alb = clamp01(att); *nrm = n; }
        beta = beta * att;
        org = p;
        dir = nd;
    }
    return L;
}

}  // namespace orc

using namespace orc;

struct orc_scene {
    Scene s;
};

static thread_local std::string g_err;

template <class F>
static void parallel_for(uint64_t n, int threads, uint64_t chunk, F f) {
    if (threads <= 0) threads = int(std::thread::hardware_concurrency());
    if (threads < 1) threads = 1;
    std::atomic<uint64_t> next{0};
    auto work = [&]() {
        for (;;) {
            uint64_t b = next.fetch_add(chunk);
            if (b >= n) break;
            uint64_t e = std::min(n, b + chunk);
            for (uint64_t i = b; i < e; ++i) f(i);
        }
    };
    if (threads == 1) { work(); return; }
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t) ts.emplace_back(work);
    for (auto& t : ts) t.join();
}

template <class R>
static int render_pixels(orc_scene* sc, const aurora_camera_desc* c, const aurora_render_desc* rd,
                         const uint32_t* pixels, uint32_t n, int threads, R* rgb, R* albedo, R* normal,
                         uint64_t* segments) {
    Cam cam;
    std::string e = camera_setup(*c, cam);
    if (!e.empty()) { g_err = e; return AURORA_ERR_INVALID; }
    if (rd->max_depth == 0) { g_err = "max_depth must be >= 1"; return AURORA_ERR_INVALID; }
    if (!sc->s.built) { g_err = "BVH not built"; return AURORA_ERR_STATE; }
    uint64_t count = pixels ? n : uint64_t(cam.W) * cam.H;
    std::atomic<uint64_t> segs{0};
    parallel_for(count, threads, 64, [&](uint64_t k) {
        uint32_t pix = pixels ? pixels[k] : uint32_t(k);
        uint32_t i = pix % cam.W, j = pix / cam.W;
        V3<R> acc = mk<R>(rgb[3 * k], rgb[3 * k + 1], rgb[3 * k + 2]);
        V3<R> aacc{}, nacc{};
        if (albedo) {
            aacc = mk<R>(albedo[3 * k], albedo[3 * k + 1], albedo[3 * k + 2]);
            nacc = mk<R>(normal[3 * k], normal[3 * k + 1], normal[3 * k + 2]);
        }
        uint32_t sg = 0;
        for (uint32_t s = 0; s < rd->spp; ++s) {
            V3<R> a = mk<R>(0, 0, 0), nn = mk<R>(0, 0, 0);
            V3<R> L = trace_path<R>(sc->s, cam, *rd, i, j, rd->sample_offset + s, albedo ? &a : nullptr,
                                    albedo ? &nn : nullptr, sg);
            acc = acc + L;
            if (albedo) { aacc = aacc + a; nacc = nacc + nn; }
        }
        segs += sg;
        rgb[3 * k] = acc.x; rgb[3 * k + 1] = acc.y; rgb[3 * k + 2] = acc.z;
        if (albedo) {
            albedo[3 * k] = aacc.x; albedo[3 * k + 1] = aacc.y; albedo[3 * k + 2] = aacc.z;
            normal[3 * k] = nacc.x; normal[3 * k + 1] = nacc.y; normal[3 * k + 2] = nacc.z;
        }
    });
    if (segments) *segments = segs.load();
    return AURORA_OK;
}


extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    U4 r = philox4x32_10(U4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}

void orc_philox_batch(const uint32_t* ctr, uint32_t n, uint32_t k0, uint32_t k1, uint32_t* out) {
    for (uint32_t i = 0; i < n; ++i) {
        U4 r = philox4x32_10(U4{ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]}, k0, k1);
        out[4 * i] = r.x; out[4 * i + 1] = r.y; out[4 * i + 2] = r.z; out[4 * i + 3] = r.w;
    }
}

void orc_perlin(uint32_t seed, float* ranvec, int32_t* perm) {
    Perlin p;
    perlin_tables(seed, p);
    std::memcpy(ranvec, p.ranvec, sizeof p.ranvec);
    for (int a = 0; a < 3; ++a)
        for (int i = 0; i < 256; ++i) perm[a * 256 + i] = p.perm[a][i];
}

double orc_sin(double x) { return orc_sin_d(x); }

void orc_reflect(const float* v, const float* n, float* out) {
    V3<float> r = reflect(mk(v[0], v[1], v[2]), mk(n[0], n[1], n[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void orc_refract(const float* v, const float* n, float eta, float* out) {
    V3<float> r = refract(mk(v[0], v[1], v[2]), mk(n[0], n[1], n[2]), eta);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

float orc_schlick(float cosine, float ri) { return schlick<float>(cosine, ri); }

void orc_random_unit(uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t seed, float* out) {
    V3<float> r = random_unit<float>(pixel, sample, bounce, seed);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
double orc_atan2(double y, double x) { return orc_atan2_d(y, x); }
double orc_acos(double x) { return orc_acos_d(x); }

int orc_scene_create(const aurora_scene_desc* d, orc_scene** out) {
    try {
        if (!d || !out) throw std::invalid_argument("null argument");
        auto* sc = new orc_scene();
        try {
            load(sc->s, *d);
        } catch (...) {
            delete sc;
            throw;
        }
        *out = sc;
        return AURORA_OK;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return AURORA_ERR_INVALID;
    } catch (const std::exception& e) {
        g_err = e.what();
        return AURORA_ERR_RUNTIME;
    }
}

void orc_scene_destroy(orc_scene* s) { delete s; }

int orc_bvh_build(orc_scene* sc, aurora_bvh_stats* st) {
    try {
        build_lbvh(sc->s);
    } catch (const std::exception& e) {
        sc->s.built = false;
        g_err = e.what();
        return AURORA_ERR_RUNTIME;
    }
    if (st) *st = sc->s.stats;
    return AURORA_OK;
}

int orc_bvh_nodes(orc_scene* sc, aurora_bvh_node* nodes, uint32_t cap) {
    if (cap < sc->s.nodes.size()) { g_err = "capacity too small"; return AURORA_ERR_INVALID; }
    std::memcpy(nodes, sc->s.nodes.data(), sc->s.nodes.size() * sizeof(aurora_bvh_node));
    return AURORA_OK;
}

int orc_bvh_wide(orc_scene* sc, aurora_bvh4_node* nodes, uint32_t cap) {
    if (cap < sc->s.wide.size()) { g_err = "capacity too small"; return AURORA_ERR_INVALID; }
    std::memcpy(nodes, sc->s.wide.data(), sc->s.wide.size() * sizeof(aurora_bvh4_node));
    return AURORA_OK;
}

int orc_bvh_order(orc_scene* sc, uint32_t* morton, uint32_t* order, uint32_t cap) {
    if (cap < sc->s.order.size()) { g_err = "capacity too small"; return AURORA_ERR_INVALID; }
    std::memcpy(morton, sc->s.sorted_morton.data(), sc->s.order.size() * 4);
    std::memcpy(order, sc->s.order.data(), sc->s.order.size() * 4);
    return AURORA_OK;
}

void orc_prim_box(orc_scene* sc, uint32_t id, float* lo, float* hi) { prim_box(sc->s, id, lo, hi); }

int orc_trace(orc_scene* sc, const float* o, const float* d, uint32_t n, float tmin, float tmax, int brute,
              int32_t* prim, float* t, uint32_t* counts, int threads) {
    if (!sc->s.built && !brute) { g_err = "BVH not built"; return AURORA_ERR_STATE; }
    parallel_for(n, threads, 4096, [&](uint64_t i) {
        V3<float> oo = mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
        V3<float> dd = mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        Hit<float> h;
        if (brute == 1) {
            h = closest_brute<float>(sc->s, oo, dd, tmin, tmax);
        } else {
            Counts c;
            h = brute == 2 ? closest_bvh<float>(sc->s, oo, dd, tmin, tmax, &c)
                           : closest_bvh4<float>(sc->s, oo, dd, tmin, tmax, &c);
            if (counts) { counts[3 * i] = c.nodes; counts[3 * i + 1] = c.spheres; counts[3 * i + 2] = c.quads; }
        }
        prim[i] = h.prim;
        t[i] = h.t;
    });
    return AURORA_OK;
}

int orc_path_segments(orc_scene* sc, const aurora_camera_desc* c, const aurora_render_desc* rd, uint32_t pixel,
                      uint32_t sample, float* out, uint32_t cap) {
    aurora_render_desc one = *rd;
    one.spp = 1;
    one.sample_offset = sample;
    std::vector<float> log;
    float rgb[3] = {0, 0, 0};
    g_seg_log = &log;
    const int rc = render_pixels<float>(sc, c, &one, &pixel, 1, 1, rgb, nullptr, nullptr, nullptr);
    g_seg_log = nullptr;
    if (rc) return -rc;
    const uint32_t n = uint32_t(log.size() / 8);
    std::memcpy(out, log.data(), std::min(n, cap) * 8 * sizeof(float));
    return int(n);
}

void orc_set_render_tree(int wide) { g_render_wide.store(wide ? 1 : 0); }

int orc_wide_order(const float* t, const int32_t* hit, int32_t* ord) {
    bool h[4];
    int o[4];
    for (int k = 0; k < 4; ++k) h[k] = hit[k] != 0;
    const int m = wide_order<float>(h, t, o);
    for (int k = 0; k < 4; ++k) ord[k] = o[k];
    return m;
}

int orc_hit_sphere(const float* sphere, const float* o, const float* d, float tmin, float tmax, float* t) {
    Sphere sp{{sphere[0], sphere[1], sphere[2]}, sphere[3]};
    return hit_sphere<float>(sp, mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]), tmin, tmax, *t) ? 1 : 0;
}

int orc_hit_quad(const float* quad, uint32_t kind, const float* o, const float* d, float tmin, float tmax,
                 float* tab) {
    Quad q;
    for (int k = 0; k < 3; ++k) { q.Q[k] = quad[k]; q.u[k] = quad[3 + k]; q.v[k] = quad[6 + k]; }
    q.kind = kind;
    return hit_quad<float>(q, mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]), tmin, tmax, tab[0], tab[1], tab[2]) ? 1
                                                                                                                : 0;
}

void orc_sphere_uv(const float* n, float* uv) {
    float u, v;
    sphere_uv<float>(mk(n[0], n[1], n[2]), u, v);
    uv[0] = u;
    uv[1] = v;
}

void orc_texture_value(orc_scene* sc, uint32_t tex, float u, float v, const float* p, float* out) {
    V3<float> c = texture_value<float>(sc->s, tex, u, v, mk(p[0], p[1], p[2]));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}

int orc_camera(const aurora_camera_desc* c, float* out18) {
    Cam cam;
    std::string e = camera_setup(*c, cam);
    if (!e.empty()) { g_err = e; return AURORA_ERR_INVALID; }
    V3<float> v[6] = {cam.center, cam.p00, cam.du, cam.dv, cam.disk_u, cam.disk_v};
    for (int i = 0; i < 6; ++i) { out18[3 * i] = v[i].x; out18[3 * i + 1] = v[i].y; out18[3 * i + 2] = v[i].z; }
    return AURORA_OK;
}

int orc_camera_rays(const aurora_camera_desc* c, uint32_t seed, const uint32_t* pixels, const uint32_t* samples,
                    uint32_t n, float* org, float* dir) {
    Cam cam;
    std::string e = camera_setup(*c, cam);
    if (!e.empty()) { g_err = e; return AURORA_ERR_INVALID; }
    for (uint32_t k = 0; k < n; ++k) {
        V3<float> o, d;
        camera_ray<float>(cam, pixels[k], pixels[k] % cam.W, pixels[k] / cam.W, samples[k], seed, o, d);
        org[3 * k] = o.x; org[3 * k + 1] = o.y; org[3 * k + 2] = o.z;
        dir[3 * k] = d.x; dir[3 * k + 1] = d.y; dir[3 * k + 2] = d.z;
    }
    return AURORA_OK;
}

// 8-bit resolve of one channel (transfer curves as include/aurora.h states):
// GAMMA2 after "Ray Tracing in One Weekend" write_color, GAMMA22 after the
// writer of experiments/rt.cpp:377-392 (clamp, pow(c, 1/2.2f), x255),
// LINEAR after are::Texture::save_texture (src/texture.cpp:378-391).
static uint8_t resolve_byte(float sum, float spp, uint32_t transfer) {
    float x = sum / spp;
    if (std::isnan(x)) x = 0.0f;
    if (transfer == AURORA_TRANSFER_GAMMA22) {
        const float c = std::min(std::max(x, 0.0f), 1.0f);
        const float g = float(std::pow(double(c), double(1.0f / 2.2f)));
        return uint8_t(g * 255.0f);
    }
    if (transfer == AURORA_TRANSFER_LINEAR) {
        const double y = double(x) * 255.0;
        return uint8_t(std::clamp(y, 0.0, 255.0));
    }
    const float g = x > 0.0f ? std::sqrt(x) : 0.0f;
    return uint8_t(256.0f * std::min(std::max(g, 0.0f), 0.999f));
}

int orc_resolve_rgb8(const float* sum, uint64_t n, uint32_t spp, uint32_t transfer, uint8_t* out) {
    if (spp == 0 || transfer > AURORA_TRANSFER_LINEAR) { g_err = "spp == 0 or unknown transfer"; return AURORA_ERR_INVALID; }
    for (uint64_t i = 0; i < n; ++i) out[i] = resolve_byte(sum[i], float(spp), transfer);
    return AURORA_OK;
}

int orc_render_pixels(orc_scene* sc, const aurora_camera_desc* c, const aurora_render_desc* rd,
                      const uint32_t* pixels, uint32_t n, int threads, float* rgb, float* albedo, float* normal,
                      uint64_t* segments) {
    return render_pixels<float>(sc, c, rd, pixels, n, threads, rgb, albedo, normal, segments);
}

int orc_render_pixels_f64(orc_scene* sc, const aurora_camera_desc* c, const aurora_render_desc* rd,
                          const uint32_t* pixels, uint32_t n, int threads, double* rgb, double* albedo,
                          double* normal, uint64_t* segments) {
    return render_pixels<double>(sc, c, rd, pixels, n, threads, rgb, albedo, normal, segments);
}

}  // extern "C"
