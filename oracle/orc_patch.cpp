// orc_patch.cpp — CPU restatement of the reference's recursive mirror
// "patch-projection" renderer (experiments/rt10.cpp), in double, as a
// sequential painter over whole images exactly as the reference composes
// them.  TEST INFRASTRUCTURE ONLY: the GPU renderer of libaurora
// (aurora_patch_render) is checked against it, and it is checked against
// rt10.cpp itself compiled from its source (oracle/ref, oracle/_ref).
//
// Every floating operation follows the reference's order (compiled with
// -ffp-contract=off), so images agree bit for bit:
//   normal / centroid / barycentric3D       rt10.cpp:160-168, 175-193
//   pointInTriangle3D / pointInTriangle2D   rt10.cpp:195-219
//   mirror point across a plane             rt10.cpp:268-273
//   projection onto the viewport plane      rt10.cpp:279-304
//   shoelace area, Sutherland-Hodgman clip  rt10.cpp:311-381
//   warp raster of one fan triangle         rt10.cpp:384-435 (pixel loop :414-434)
//   fan of a clipped polygon                rt10.cpp:437-456
//   candidate collection                    rt10.cpp:462-484
//   recursion                               rt10.cpp:551-664
//   viewport render, mask, composition      rt10.cpp:677-787
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "oracle.h"

namespace {

struct V2 {
    double x, y;
};
struct V3 {
    double x, y, z;
};

V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
V2 operator-(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
V2 operator*(V2 a, double s) { return {a.x * s, a.y * s}; }
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }

double cross2(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// the reference clamps with std::max(lo, std::min(hi, x))
double clamp(double x, double lo, double hi) { return std::max(lo, std::min(hi, x)); }
int clampi(int x, int lo, int hi) { return std::max(lo, std::min(hi, x)); }
V3 lerp3(V3 a, V3 b, double t) { return a * (1.0 - t) + b * t; }

constexpr double kTiny = 1e-9;

struct Tri {
    V3 p[3];
    V2 uv[3];
    int id;
    const aurora_patch_material* mat;
    V3 normal() const {
        V3 c = cross(p[1] - p[0], p[2] - p[0]);
        double len = std::sqrt(dot(c, c));
        if (len < kTiny) return {0, 0, 0};
        return c / len;
    }
    V3 centroid() const { return (p[0] + p[1] + p[2]) / 3.0; }
};

Tri to_tri(const aurora_patch_triangle& t, const aurora_patch_material* mats) {
    Tri r;
    for (int i = 0; i < 3; ++i) {
        r.p[i] = {t.p[i][0], t.p[i][1], t.p[i][2]};
        r.uv[i] = {t.uv[i][0], t.uv[i][1]};
    }
    r.id = t.id;
    r.mat = (mats && t.material != AURORA_PATCH_NO_MATERIAL) ? &mats[t.material] : nullptr;
    return r;
}

struct Img {
    int w = 0, h = 0;
    std::vector<V3> px;
    Img(int w_, int h_, V3 fill) : w(w_), h(h_), px(size_t(w_) * size_t(h_), fill) {}
    V3 at(int x, int y) const { return px[size_t(clampi(y, 0, h - 1)) * size_t(w) + size_t(clampi(x, 0, w - 1))]; }
    void put(int x, int y, V3 c) {
        if (x >= 0 && x < w && y >= 0 && y < h) px[size_t(y) * size_t(w) + size_t(x)] = c;
    }
    V3 bilinear(double u, double v) const {
        u = clamp(u, 0.0, 1.0);
        v = clamp(v, 0.0, 1.0);
        const double fx = u * (w - 1), fy = v * (h - 1);
        const int x0 = int(std::floor(fx)), y0 = int(std::floor(fy));
        const int x1 = std::min(x0 + 1, w - 1), y1 = std::min(y0 + 1, h - 1);
        const double tx = fx - x0, ty = fy - y0;
        return lerp3(lerp3(at(x0, y0), at(x1, y0), tx), lerp3(at(x0, y1), at(x1, y1), tx), ty);
    }
};

// barycentric coordinates (u, v, w) of P in T's plane; (-1,-1,-1) when T is degenerate
void bary3(V3 P, const Tri& T, double out[3]) {
    const V3 e0 = T.p[1] - T.p[0], e1 = T.p[2] - T.p[0], q = P - T.p[0];
    const double d00 = dot(e0, e0), d01 = dot(e0, e1), d11 = dot(e1, e1);
    const double d20 = dot(q, e0), d21 = dot(q, e1);
    const double den = d00 * d11 - d01 * d01;
    if (std::fabs(den) < kTiny) {
        out[0] = out[1] = out[2] = -1.0;
        return;
    }
    const double b = (d11 * d20 - d01 * d21) / den;
    const double c = (d00 * d21 - d01 * d20) / den;
    out[0] = 1.0 - b - c;
    out[1] = b;
    out[2] = c;
}

bool inside3(V3 P, const Tri& T) {
    double b[3];
    bary3(P, T, b);
    return b[0] >= -1e-8 && b[1] >= -1e-8 && b[2] >= -1e-8;
}

bool inside2(V2 P, V2 A, V2 B, V2 C) {
    const double eps = 1e-10;
    const double c0 = cross2(B - A, P - A), c1 = cross2(C - B, P - B), c2 = cross2(A - C, P - C);
    const bool neg = c0 < -eps || c1 < -eps || c2 < -eps;
    const bool pos = c0 > eps || c1 > eps || c2 > eps;
    return !(neg && pos);
}

V3 mirror(V3 P, const Tri& T) {
    const V3 n = T.normal();
    const double d = dot(n, P - T.p[0]);
    return P - n * (2.0 * d);
}

// the segment eye -> point crosses the viewport's plane strictly between
// its ends; hit point and its (possibly outside) viewport uv
bool project(V3 eye, V3 point, const Tri& vp, V2& uv, V3* hit) {
    const V3 n = vp.normal();
    const V3 d = point - eye;
    const double den = dot(n, d);
    if (std::fabs(den) < kTiny) return false;
    const double t = dot(n, vp.p[0] - eye) / den;
    if (!(t > 1e-7 && t < 1.0 - 1e-7)) return false;
    const V3 h = eye + d * t;
    double b[3];
    bary3(h, vp, b);
    if (b[0] < -1e6 || b[1] < -1e6 || b[2] < -1e6) return false;
    uv = vp.uv[0] * b[0] + vp.uv[1] * b[1] + vp.uv[2] * b[2];
    if (hit) *hit = h;
    return true;
}

struct PV {
    V2 dst, src;
};

double shoelace(const std::vector<PV>& poly) {
    double a = 0.0;
    for (size_t i = 0; i < poly.size(); ++i) {
        const V2 p = poly[i].dst, q = poly[(i + 1) % poly.size()].dst;
        a += p.x * q.y - p.y * q.x;
    }
    return 0.5 * a;
}

std::vector<PV> clip_edge(V2 A, V2 B, const std::vector<PV>& in) {
    std::vector<PV> out;
    if (in.empty()) return out;
    auto inside = [&](V2 P) { return cross2(B - A, P - A) >= -1e-12; };
    auto cut = [&](const PV& S, const PV& E) {
        const double dS = cross2(B - A, S.dst - A), dE = cross2(B - A, E.dst - A);
        const double t = clamp(dS / (dS - dE + 1e-30), 0.0, 1.0);
        return PV{S.dst + (E.dst - S.dst) * t, S.src + (E.src - S.src) * t};
    };
    PV S = in.back();
    bool s_in = inside(S.dst);
    for (const PV& E : in) {
        const bool e_in = inside(E.dst);
        if (e_in) {
            if (!s_in) out.push_back(cut(S, E));
            out.push_back(E);
        } else if (s_in) {
            out.push_back(cut(S, E));
        }
        S = E;
        s_in = e_in;
    }
    return out;
}

std::vector<PV> clip_to(std::vector<PV> poly, V2 C0, V2 C1, V2 C2) {
    V2 A1 = C1, A2 = C2;
    if (cross2(C1 - C0, C2 - C0) < 0.0) std::swap(A1, A2);
    poly = clip_edge(C0, A1, poly);
    poly = clip_edge(A1, A2, poly);
    return clip_edge(A2, C0, poly);
}

void warp_triangle(Img& dst, const PV& a, const PV& b, const PV& c, const Img& src) {
    const double sx = dst.w - 1, sy = dst.h - 1;
    const V2 p0{a.dst.x * sx, a.dst.y * sy}, p1{b.dst.x * sx, b.dst.y * sy}, p2{c.dst.x * sx, c.dst.y * sy};
    const int x0 = clampi(int(std::floor(std::min({p0.x, p1.x, p2.x}))), 0, dst.w - 1);
    const int x1 = clampi(int(std::ceil(std::max({p0.x, p1.x, p2.x}))), 0, dst.w - 1);
    const int y0 = clampi(int(std::floor(std::min({p0.y, p1.y, p2.y}))), 0, dst.h - 1);
    const int y1 = clampi(int(std::ceil(std::max({p0.y, p1.y, p2.y}))), 0, dst.h - 1);
    const double area = cross2(p1 - p0, p2 - p0);
    if (std::fabs(area) < 1e-12) return;
    for (int y = y0; y <= y1; ++y)
        for (int x = x0; x <= x1; ++x) {
            const V2 P{double(x) + 0.5, double(y) + 0.5};
            const double w0 = cross2(p1 - P, p2 - P) / area;
            const double w1 = cross2(p2 - P, p0 - P) / area;
            const double w2 = 1.0 - w0 - w1;
            if (w0 < -1e-6 || w1 < -1e-6 || w2 < -1e-6) continue;
            const V2 s = a.src * w0 + b.src * w1 + c.src * w2;
            dst.put(x, y, src.bilinear(s.x, s.y));
        }
}

void warp_fan(Img& dst, const std::vector<PV>& poly, const Img& src) {
    for (size_t i = 1; i + 1 < poly.size(); ++i) warp_triangle(dst, poly[0], poly[i], poly[i + 1], src);
}

struct Renderer {
    std::vector<Tri> scene;
    aurora_patch_config cfg;
    V3 env;
    uint64_t textures = 0;

    // scene triangles with a vertex projecting inside vp, far to near
    std::vector<const Tri*> visible(V3 eye, const Tri& vp) const {
        std::vector<const Tri*> c;
        for (const Tri& t : scene)
            for (int i = 0; i < 3; ++i) {
                V2 uv;
                V3 h;
                if (!project(eye, t.p[i], vp, uv, &h)) continue;
                if (inside3(h, vp)) {
                    c.push_back(&t);
                    break;
                }
            }
        // the reference's comparator and std::sort, so equal distances keep
        // its order
        std::sort(c.begin(), c.end(), [&](const Tri* a, const Tri* b) {
            const V3 da = a->centroid() - eye, db = b->centroid() - eye;
            return std::sqrt(dot(da, da)) > std::sqrt(dot(db, db));
        });
        return c;
    }

    // the clipped projection of t into vp's uv triangle; false if any vertex
    // does not project or fewer than three vertices survive
    bool clipped(V3 eye, const Tri& t, const Tri& vp, std::vector<PV>& poly) const {
        std::vector<PV> p(3);
        for (int i = 0; i < 3; ++i) {
            if (!project(eye, t.p[i], vp, p[i].dst, nullptr)) return false;
            p[i].src = t.uv[i];
        }
        poly = clip_to(std::move(p), vp.uv[0], vp.uv[1], vp.uv[2]);
        return poly.size() >= 3;
    }

    Img texture(V3 eye, const Tri& cur, int tw, int th, int depth, double est_px, std::vector<int>& path) {
        ++textures;
        tw = clampi(tw, cfg.min_tex_res, cfg.max_tex_res);
        th = clampi(th, cfg.min_tex_res, cfg.max_tex_res);
        const V3 base = cur.mat ? V3{cur.mat->albedo[0], cur.mat->albedo[1], cur.mat->albedo[2]} : V3{1, 1, 1};
        Img out(tw, th, base);
        if (!cur.mat || cur.mat->type != AURORA_PATCH_REFLECTIVE) return out;
        if (depth >= cfg.max_depth) return out;
        if (est_px > 0.0 && est_px < cfg.min_area_px) return out;
        if (std::find(path.begin(), path.end(), cur.id) != path.end()) return out;
        path.push_back(cur.id);
        const V3 eye2 = mirror(eye, cur);
        Img refl(tw, th, env);
        for (const Tri* t : visible(eye2, cur)) {
            if (t->id == cur.id) continue;
            std::vector<PV> poly;
            if (!clipped(eye2, *t, cur, poly)) continue;
            const double px = std::fabs(shoelace(poly)) * double(tw) * double(th);
            if (px < 0.5) continue;
            const int res = clampi(int(std::lround(std::sqrt(px) * 1.2)), cfg.min_tex_res, cfg.max_tex_res);
            const Img child = texture(eye2, *t, res, res, depth + 1, px, path);
            warp_fan(refl, poly, child);
        }
        path.pop_back();
        const double m = clamp(cur.mat->metalness, 0.0, 1.0);
        for (int y = 0; y < th; ++y)
            for (int x = 0; x < tw; ++x) {
                const V3 r = refl.at(x, y);
                const V3 tint{r.x * base.x, r.y * base.y, r.z * base.z};
                out.put(x, y, base * (1.0 - m) + tint * m);
            }
        return out;
    }

    Img viewport(V3 eye, const Tri& vp, int w, int h) {
        Img img(w, h, env);
        std::vector<int> path;
        for (const Tri* t : visible(eye, vp)) {
            std::vector<PV> poly;
            if (!clipped(eye, *t, vp, poly)) continue;
            const double px = std::fabs(shoelace(poly)) * double(w) * double(h);
            if (px < 0.5) continue;
            const int res = clampi(int(std::lround(std::sqrt(px) * 1.0)), cfg.min_tex_res, cfg.max_tex_res);
            const Img tex = texture(eye, *t, res, res, 0, px, path);
            warp_fan(img, poly, tex);
        }
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                const V2 uv{double(x) / (w - 1), double(y) / (h - 1)};
                if (!inside2(uv, vp.uv[0], vp.uv[1], vp.uv[2])) img.put(x, y, {0, 0, 0});
            }
        return img;
    }
};

thread_local std::string g_err;

int fail(const std::string& m) {
    g_err = m;
    return AURORA_ERR_INVALID;
}

}  // namespace

extern "C" {

int orc_patch_render(const aurora_patch_triangle* tris, uint32_t n_tris, const aurora_patch_material* mats,
                     uint32_t n_mats, const aurora_patch_camera* cam, const aurora_patch_config* cfg, double* rgb,
                     uint64_t* textures) {
    if (!cam || !cfg || !rgb || (n_tris && !tris)) return fail("orc_patch_render: null argument");
    if (cam->width < 2 || cam->height < 2) return fail("orc_patch_render: image smaller than 2x2");
    if (cfg->min_tex_res < 1 || cfg->max_tex_res < cfg->min_tex_res) return fail("orc_patch_render: texture sizes");
    for (uint32_t i = 0; i < n_tris; ++i)
        if (tris[i].material != AURORA_PATCH_NO_MATERIAL && tris[i].material >= n_mats)
            return fail("orc_patch_render: material index out of range");
    for (uint32_t i = 0; i < n_mats; ++i)
        if (mats[i].type != AURORA_PATCH_DIFFUSE && mats[i].type != AURORA_PATCH_REFLECTIVE)
            return fail("orc_patch_render: unknown material type");
    Renderer r;
    r.cfg = *cfg;
    r.env = {cfg->env[0], cfg->env[1], cfg->env[2]};
    r.scene.reserve(n_tris);
    for (uint32_t i = 0; i < n_tris; ++i) r.scene.push_back(to_tri(tris[i], mats));
    const V3 eye{cam->origin[0], cam->origin[1], cam->origin[2]};
    const int w = int(cam->width), h = int(cam->height);
    const Img a = r.viewport(eye, to_tri(cam->viewport[0], nullptr), w, h);
    const Img b = r.viewport(eye, to_tri(cam->viewport[1], nullptr), w, h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const V3 c = a.at(x, y) + b.at(x, y);
            double* o = rgb + 3 * (size_t(y) * size_t(w) + size_t(x));
            o[0] = clamp(c.x, 0.0, 1.0);
            o[1] = clamp(c.y, 0.0, 1.0);
            o[2] = clamp(c.z, 0.0, 1.0);
        }
    if (textures) *textures = r.textures;
    return AURORA_OK;
}

const char* orc_patch_last_error(void) { return g_err.c_str(); }

}  // extern "C"
