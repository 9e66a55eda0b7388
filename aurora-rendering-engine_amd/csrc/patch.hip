// patch.hip — the reference's recursive mirror "patch-projection" renderer
// (experiments/rt10.cpp:384-787) on gfx950, in double as the reference.
//
// The reference paints, for each of the camera's two viewport triangles and
// far to near, every scene triangle that has a vertex projecting inside the
// viewport: it projects the triangle's vertices onto the viewport plane,
// clips the projection to the viewport's uv triangle (Sutherland-Hodgman),
// and rasterises the clipped polygon's fan with the triangle's texture
// warped over it.  A reflective triangle's texture is rendered the same
// way from the eye mirrored across its plane, with the triangle itself as
// the viewport, and mixed with its albedo by its metalness; the recursion
// ends at the depth limit, below an estimated texel area, and at a triangle
// id already on the recursion path.
//
// Host (plan): the recursion tree — visibility, far-to-near order, clipping,
// area estimates, texture sizes — and every fan triangle's pixel-space
// vertices, box and source uv.  This is per triangle, not per texel, work:
// O(textures x scene triangles) in double, with the reference's operations
// in its order (the order and the sizes are discrete decisions the image
// depends on).
// Device (raster): every texel of every recursed texture and every image
// pixel, the reference's per-pixel loop (rt10.cpp:414-434) turned inside
// out: a texel walks its texture's warp triangles from the last painted to
// the first and takes the first that covers it — the one the reference's
// painter leaves on top — then samples the child texture bilinearly.
// Textures at recursion depth d sample only textures of depth d + 1, so one
// launch per depth, deepest first, sees every child complete; the last
// launch composes the two viewports with the viewport masks and the clamp
// of Camera::render.  Constant textures (diffuse, or a reflective triangle
// whose recursion stopped) are never stored: their samples are computed from
// the albedo with the same bilinear arithmetic.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.hpp"

namespace aur {

namespace {

// ---- host geometry (double, the reference's operation order) -------------
struct P2 {
    double x, y;
};
struct P3 {
    double x, y, z;
};
inline P2 operator+(P2 a, P2 b) { return {a.x + b.x, a.y + b.y}; }
inline P2 operator-(P2 a, P2 b) { return {a.x - b.x, a.y - b.y}; }
inline P2 operator*(P2 a, double s) { return {a.x * s, a.y * s}; }
inline P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 operator*(P3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double cross2(P2 a, P2 b) { return a.x * b.y - a.y * b.x; }
inline double dot3(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline P3 cross3(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double clampd(double x, double lo, double hi) { return std::max(lo, std::min(hi, x)); }
inline int clampi(int x, int lo, int hi) { return std::max(lo, std::min(hi, x)); }

struct HTri {
    P3 p[3];
    P2 uv[3];
    int id;
    const aurora_patch_material* mat;
    uint32_t vi[3];  // the vertices' indices among the scene's distinct vertex positions
    P3 cen;          // centroid()
    P3 normal() const {
        const P3 c = cross3(p[1] - p[0], p[2] - p[0]);
        const double len = std::sqrt(dot3(c, c));
        if (len < 1e-9) return {0, 0, 0};
        return {c.x / len, c.y / len, c.z / len};
    }
    P3 centroid() const {
        const P3 s = p[0] + p[1] + p[2];
        return {s.x / 3.0, s.y / 3.0, s.z / 3.0};
    }
};

HTri host_tri(const aurora_patch_triangle& t, const aurora_patch_material* mats) {
    HTri r;
    for (int i = 0; i < 3; ++i) {
        r.p[i] = {t.p[i][0], t.p[i][1], t.p[i][2]};
        r.uv[i] = {t.uv[i][0], t.uv[i][1]};
    }
    r.id = t.id;
    r.mat = (mats && t.material != AURORA_PATCH_NO_MATERIAL) ? &mats[t.material] : nullptr;
    r.vi[0] = r.vi[1] = r.vi[2] = 0;
    r.cen = r.centroid();
    return r;
}

struct UVPair {
    P2 dst, src;
};

// one Sutherland-Hodgman pass: keep the part left of a -> b
void clip_pass(P2 a, P2 b, const std::vector<UVPair>& in, std::vector<UVPair>& out) {
    out.clear();
    if (in.empty()) return;
    const P2 e = b - a;
    UVPair s = in.back();
    double ds = cross2(e, s.dst - a);
    bool s_in = ds >= -1e-12;
    for (const UVPair& q : in) {
        const double dq = cross2(e, q.dst - a);
        const bool q_in = dq >= -1e-12;
        if (q_in != s_in) {
            // the crossing, by the reference's signed-distance ratio
            const double t = clampd(ds / (ds - dq + 1e-30), 0.0, 1.0);
            out.push_back({s.dst + (q.dst - s.dst) * t, s.src + (q.src - s.src) * t});
        }
        if (q_in) out.push_back(q);
        s = q;
        ds = dq;
        s_in = q_in;
    }
}

// ---- device records ------------------------------------------------------
constexpr uint64_t kConstTex = ~0ull;

struct PatchTex {                 // 64 bytes
    double base[3];               // albedo (white without a material)
    double m;                     // metalness clamped to [0,1] (rasterised textures)
    uint32_t w, h;
    uint32_t warp_begin, warp_end;  // rasterised textures: their warp triangles
    uint64_t texel_off;           // first texel (double4) in the pool, or kConstTex
    uint64_t pad;
};
static_assert(sizeof(PatchTex) == 64, "PatchTex layout");

struct PatchWarp {                // 128 bytes: one fan triangle of a clipped polygon
    int32_t x0, x1, y0, y1;       // pixel box, clamped to the painted image (one 16-byte load)
    uint32_t child;               // texture sampled
    uint32_t pad;
    double area;                  // cross2(p1 - p0, p2 - p0), |area| >= 1e-12
    double p[6];                  // pixel-space vertices p0, p1, p2 of the painted image
    double s[6];                  // source uv of the three vertices
};
static_assert(sizeof(PatchWarp) == 128, "PatchWarp layout");

constexpr int kTile = 8;          // texture tiles: 8 x 8 texels, one wave
struct PatchTile {
    uint32_t tex;
    uint16_t tx, ty;
};

// ---- planner -------------------------------------------------------------
struct Plan {
    const std::vector<HTri>* scene = nullptr;
    const std::vector<P3>* verts = nullptr;  // distinct vertex positions of the scene
    // per distinct vertex, seen from an eye through a viewport triangle:
    // project()'s result (ok: the segment meets the plane; uv; in: inside)
    struct VProj {
        P2 uv;
        bool ok, in;
    };
    // one table per recursion level: a level's table stays valid while the
    // textures below it are planned
    std::vector<std::vector<VProj>> proj;
    aurora_patch_config cfg{};
    std::vector<PatchTex> texs;
    std::vector<int> tex_depth;
    std::vector<PatchWarp> warps;
    uint64_t pool_texels = 0;
    uint64_t warp_texels = 0;
    int max_level = -1;
    std::vector<int> path;
    std::vector<UVPair> ca, cb;  // clip buffers
    Status status;

    // scene triangles with a vertex projecting inside vp, far to near
    struct Seen {
        double dist;  // centroid distance from the eye
        const HTri* t;
    };
    // out[v]: distinct vertex v seen from eye through vp.  ok: the eye ->
    // vertex segment meets vp's plane strictly inside (t in (1e-7,
    // 1 - 1e-7)) and the meeting point's barycentric coordinates in vp are
    // above -1e6 ((-1, -1, -1) for a degenerate vp); uv: that point's vp uv
    // (extrapolated outside vp); in: it lies inside vp (coordinates >= -1e-8:
    // the reference's pointInTriangle3D of the same point, rt10.cpp:195-199).
    // The terms that depend on vp alone (normal, plane offset, barycentric
    // Gram terms) are computed once; the per-vertex operations are the
    // reference's, in its order.
    void project_all(P3 eye, const HTri& vp, std::vector<VProj>& out) {
        const size_t nv = verts->size();
        out.resize(nv);
        const P3 n = vp.normal();
        const double num = dot3(n, vp.p[0] - eye);
        const P3 a = vp.p[1] - vp.p[0], b = vp.p[2] - vp.p[0];
        const double aa = dot3(a, a), ab = dot3(a, b), bb = dot3(b, b);
        const double gden = aa * bb - ab * ab;
        const bool degenerate = std::fabs(gden) < 1e-9;  // barycentric (-1, -1, -1)
        const P3* V = verts->data();
        for (size_t v = 0; v < nv; ++v) {
            const P3 d = V[v] - eye;
            const double den = dot3(n, d);
            const double t = num / den;
            const P3 c = (eye + d * t) - vp.p[0];
            const double ca = dot3(c, a), cb = dot3(c, b);
            const double bv = degenerate ? -1.0 : (bb * ca - ab * cb) / gden;
            const double bw = degenerate ? -1.0 : (aa * cb - ab * ca) / gden;
            const double bu = degenerate ? -1.0 : 1.0 - bv - bw;
            VProj& o = out[v];
            o.ok = (std::fabs(den) >= 1e-9) & (t > 1e-7) & (t < 1.0 - 1e-7) & !(bu < -1e6) & !(bv < -1e6) &
                   !(bw < -1e6);
            o.uv = vp.uv[0] * bu + vp.uv[1] * bv + vp.uv[2] * bw;
            o.in = o.ok & (bu >= -1e-8) & (bv >= -1e-8) & (bw >= -1e-8);
        }
    }

    // A triangle is visible when one of its vertices projects inside vp; a
    // vertex position shared by several triangles is tested once (the same
    // operations on the same values: the same decision).  The projections
    // stay in proj[level] for clip().
    std::vector<Seen> visible(P3 eye, const HTri& vp, int level) {
        if (proj.size() <= size_t(level)) proj.resize(size_t(std::max(level, cfg.max_depth)) + 2);
        std::vector<VProj>& pv = proj[size_t(level)];
        project_all(eye, vp, pv);
        std::vector<Seen> out;
        for (const HTri& t : *scene)
            if (pv[t.vi[0]].in || pv[t.vi[1]].in || pv[t.vi[2]].in) {
                const P3 dc = t.cen - eye;
                out.push_back({std::sqrt(dot3(dc, dc)), &t});
            }
        // std::sort with the reference's comparator (distances computed once
        // rather than per comparison: the same comparisons, so equal
        // distances end in the order the reference's sort leaves them)
        std::sort(out.begin(), out.end(), [](const Seen& a, const Seen& b) { return a.dist > b.dist; });
        return out;
    }

    // clipped projection of t into vp's uv triangle (in ca), from the
    // projections visible(eye, vp, level) made; false when a vertex does not
    // project or fewer than three vertices remain
    bool clip(const HTri& t, const HTri& vp, int level) {
        const std::vector<VProj>& pv = proj[size_t(level)];
        ca.resize(3);
        for (int i = 0; i < 3; ++i) {
            const VProj& v = pv[t.vi[i]];
            if (!v.ok) return false;
            ca[i].dst = v.uv;
            ca[i].src = t.uv[i];
        }
        P2 c0 = vp.uv[0], c1 = vp.uv[1], c2 = vp.uv[2];
        if (cross2(c1 - c0, c2 - c0) < 0.0) std::swap(c1, c2);
        clip_pass(c0, c1, ca, cb);
        clip_pass(c1, c2, cb, ca);
        clip_pass(c2, c0, ca, cb);
        ca.swap(cb);
        return ca.size() >= 3;
    }

    static double area_uv(const std::vector<UVPair>& poly) {
        double a = 0.0;
        for (size_t i = 0; i < poly.size(); ++i) {
            const P2 p = poly[i].dst, q = poly[(i + 1) % poly.size()].dst;
            a += p.x * q.y - p.y * q.x;
        }
        return std::fabs(0.5 * a);
    }

    // fan triangles of poly painted into a w x h image from texture child
    void add_fan(const std::vector<UVPair>& poly, int w, int h, uint32_t child, std::vector<PatchWarp>& out) {
        for (size_t i = 1; i + 1 < poly.size(); ++i) {
            const UVPair* v[3] = {&poly[0], &poly[i], &poly[i + 1]};
            PatchWarp r{};
            P2 p[3];
            for (int k = 0; k < 3; ++k) {
                p[k] = {v[k]->dst.x * (w - 1), v[k]->dst.y * (h - 1)};
                r.p[2 * k] = p[k].x;
                r.p[2 * k + 1] = p[k].y;
                r.s[2 * k] = v[k]->src.x;
                r.s[2 * k + 1] = v[k]->src.y;
            }
            r.area = cross2(p[1] - p[0], p[2] - p[0]);
            if (std::fabs(r.area) < 1e-12) continue;  // paints nothing
            r.x0 = clampi(int(std::floor(std::min({p[0].x, p[1].x, p[2].x}))), 0, w - 1);
            r.x1 = clampi(int(std::ceil(std::max({p[0].x, p[1].x, p[2].x}))), 0, w - 1);
            r.y0 = clampi(int(std::floor(std::min({p[0].y, p[1].y, p[2].y}))), 0, h - 1);
            r.y1 = clampi(int(std::ceil(std::max({p[0].y, p[1].y, p[2].y}))), 0, h - 1);
            r.child = child;
            warp_texels += uint64_t(r.x1 - r.x0 + 1) * uint64_t(r.y1 - r.y0 + 1);
            out.push_back(r);
        }
    }

    // the texture of triangle cur seen from eye (rt10.cpp:551-664); returns
    // its index
    uint32_t texture(P3 eye, const HTri& cur, int tw, int th, int depth, double est_px) {
        tw = clampi(tw, cfg.min_tex_res, cfg.max_tex_res);
        th = clampi(th, cfg.min_tex_res, cfg.max_tex_res);
        const uint32_t idx = uint32_t(texs.size());
        if (texs.size() >= (1u << 26)) {
            status = Status{AURORA_ERR_NOMEM, "patch render: recursion tree above 2^26 textures"};
            return idx;
        }
        PatchTex T{};
        if (cur.mat) {
            for (int k = 0; k < 3; ++k) T.base[k] = cur.mat->albedo[k];
        } else {
            T.base[0] = T.base[1] = T.base[2] = 1.0;
        }
        T.w = uint32_t(tw);
        T.h = uint32_t(th);
        T.texel_off = kConstTex;
        texs.push_back(T);
        tex_depth.push_back(depth);
        if (!cur.mat || cur.mat->type != AURORA_PATCH_REFLECTIVE) return idx;
        if (depth >= cfg.max_depth) return idx;
        if (est_px > 0.0 && est_px < cfg.min_area_px) return idx;
        if (std::find(path.begin(), path.end(), cur.id) != path.end()) return idx;
        path.push_back(cur.id);
        const P3 n = cur.normal();
        const P3 eye2 = eye - n * (2.0 * dot3(n, eye - cur.p[0]));  // the eye mirrored across cur
        std::vector<PatchWarp> mine;
        for (const Seen& seen : visible(eye2, cur, depth + 1)) {
            const HTri* t = seen.t;
            if (status.bad()) break;
            if (t->id == cur.id) continue;
            if (!clip(*t, cur, depth + 1)) continue;
            const double px = area_uv(ca) * double(tw) * double(th);
            if (px < 0.5) continue;
            const int res = clampi(int(std::lround(std::sqrt(px) * 1.2)), cfg.min_tex_res, cfg.max_tex_res);
            const std::vector<UVPair> poly = ca;  // the recursion reuses the clip buffers
            const uint32_t child = texture(eye2, *t, res, res, depth + 1, px);
            add_fan(poly, tw, th, child, mine);
        }
        path.pop_back();
        PatchTex& R = texs[idx];
        R.m = clampd(cur.mat->metalness, 0.0, 1.0);
        R.warp_begin = uint32_t(warps.size());
        warps.insert(warps.end(), mine.begin(), mine.end());
        R.warp_end = uint32_t(warps.size());
        R.texel_off = pool_texels;
        pool_texels += uint64_t(tw) * uint64_t(th);
        max_level = std::max(max_level, depth);
        return idx;
    }
};

// The frame: each viewport's painted triangles (rt10.cpp:677-752) and the
// recursion tree below each of them.  The subtrees are independent (the
// recursion path starts empty at a viewport), so they are planned in
// parallel on host threads when the frame is large enough to gain, each
// into its own tables, and merged in the serial order: texture, warp and
// texel numbering equal a serial plan's.
struct FramePlan {
    std::vector<PatchTex> texs;
    std::vector<int> tex_depth;
    std::vector<PatchWarp> warps;
    uint32_t begin[2] = {0, 0}, end[2] = {0, 0};
    uint64_t pool_texels = 0, warp_texels = 0;
    int max_level = -1;
    // empty tables that keep their storage: a frame reuses the last one's
    // pages instead of faulting in fresh ones
    void reset() {
        texs.clear();
        tex_depth.clear();
        warps.clear();
        begin[0] = begin[1] = end[0] = end[1] = 0;
        pool_texels = warp_texels = 0;
        max_level = -1;
    }
};

// max_threads 0: the library's choice
Status plan_frame(std::vector<HTri>& scene, const aurora_patch_config& cfg, P3 eye, const HTri vp[2], int W,
                  int H, FramePlan& out, unsigned max_threads = 0) {
    // distinct vertex positions (bitwise), in first-seen order
    std::vector<P3> verts;
    {
        struct Key {
            uint64_t b[3];
            bool operator==(const Key& o) const { return b[0] == o.b[0] && b[1] == o.b[1] && b[2] == o.b[2]; }
        };
        struct KeyHash {
            size_t operator()(const Key& k) const {
                return size_t((k.b[0] * 0x9E3779B97F4A7C15ull) ^ (k.b[1] * 0xC2B2AE3D27D4EB4Full) ^
                              (k.b[2] * 0x165667B19E3779F9ull));
            }
        };
        std::unordered_map<Key, uint32_t, KeyHash> index;
        index.reserve(scene.size() * 3);
        for (HTri& t : scene)
            for (int i = 0; i < 3; ++i) {
                Key k;
                std::memcpy(k.b, &t.p[i], sizeof(k.b));
                auto it = index.emplace(k, uint32_t(verts.size()));
                if (it.second) verts.push_back(t.p[i]);
                t.vi[i] = it.first->second;
            }
    }
    struct Top {
        int vp;
        const HTri* tri;
        int res;
        double px;
        std::vector<UVPair> poly;
    };
    Plan top;
    top.scene = &scene;
    top.verts = &verts;
    top.cfg = cfg;
    std::vector<Top> items;
    for (int k = 0; k < 2; ++k)
        for (const Plan::Seen& seen : top.visible(eye, vp[k], 0)) {
            if (!top.clip(*seen.t, vp[k], 0)) continue;
            const double px = Plan::area_uv(top.ca) * double(W) * double(H);
            if (px < 0.5) continue;
            const int res = clampi(int(std::lround(std::sqrt(px) * 1.0)), cfg.min_tex_res, cfg.max_tex_res);
            items.push_back(Top{k, seen.t, res, px, top.ca});
        }
    std::vector<Plan> sub(items.size());
    auto run = [&](size_t i) {
        Plan& p = sub[i];
        p.scene = &scene;
        p.verts = &verts;
        p.cfg = cfg;
        try {
            p.texture(eye, *items[i].tri, items[i].res, items[i].res, 0, items[i].px);
        } catch (const std::bad_alloc&) {
            p.status = Status{AURORA_ERR_NOMEM, "patch render: host memory exhausted while planning"};
        }
    };
    const size_t work = items.size() * scene.size();
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    size_t n_threads = work < 8192 ? 1 : std::min<size_t>({16, hw, items.size()});
    if (max_threads) n_threads = std::min<size_t>(n_threads, max_threads);
    if (n_threads <= 1) {
        for (size_t i = 0; i < items.size(); ++i) run(i);
    } else {
        // largest projected area first: the big subtrees start early, so
        // the last ones to finish are small (each result lands at its own
        // index, so the order of work changes no numbering)
        std::vector<size_t> order(items.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return items[a].px > items[b].px; });
        std::atomic<size_t> next{0};
        std::vector<std::thread> pool;
        for (size_t w = 0; w < n_threads; ++w)
            pool.emplace_back([&] {
                for (size_t k = next++; k < order.size(); k = next++) run(order[k]);
            });
        for (std::thread& t : pool) t.join();
    }
    size_t n_texs = 0, n_warps = 0;
    for (const Plan& p : sub) {
        AUR_TRY(p.status);
        n_texs += p.texs.size();
        n_warps += p.warps.size();
    }
    out.texs.reserve(n_texs);
    out.tex_depth.reserve(n_texs);
    out.warps.reserve(n_warps + 2 * items.size() * 8);  // and the viewports' fans
    std::vector<uint32_t> root(items.size());
    for (int k = 0; k < 2; ++k) {
        for (size_t i = 0; i < items.size(); ++i) {
            if (items[i].vp != k) continue;
            const Plan& p = sub[i];
            const uint32_t tex_off = uint32_t(out.texs.size()), warp_off = uint32_t(out.warps.size());
            for (PatchTex T : p.texs) {
                if (T.texel_off != kConstTex) {
                    T.texel_off += out.pool_texels;
                    T.warp_begin += warp_off;
                    T.warp_end += warp_off;
                }
                out.texs.push_back(T);
            }
            for (PatchWarp w : p.warps) {
                w.child += tex_off;
                out.warps.push_back(w);
            }
            out.tex_depth.insert(out.tex_depth.end(), p.tex_depth.begin(), p.tex_depth.end());
            out.pool_texels += p.pool_texels;
            out.warp_texels += p.warp_texels;
            out.max_level = std::max(out.max_level, p.max_level);
            root[i] = tex_off;
            if (out.texs.size() > (1u << 26))
                return Status{AURORA_ERR_NOMEM, "patch render: recursion tree above 2^26 textures"};
        }
        std::vector<PatchWarp> fans;
        for (size_t i = 0; i < items.size(); ++i)
            if (items[i].vp == k) top.add_fan(items[i].poly, W, H, root[i], fans);
        out.begin[k] = uint32_t(out.warps.size());
        out.warps.insert(out.warps.end(), fans.begin(), fans.end());
        out.end[k] = uint32_t(out.warps.size());
    }
    out.warp_texels += top.warp_texels;
    return Status::ok();
}

// ---- device --------------------------------------------------------------
// std::min / std::max semantics (the first operand unless the second is
// strictly smaller / larger), as the reference's clamps
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dclamp01(double x) { return dmax(0.0, dmin(1.0, x)); }

struct D3 {
    double x, y, z;
};

// bilinear sample of texture T at (u, v) (rt10.cpp:96-116)
__device__ __forceinline__ D3 sample(const PatchTex& T, const double4* __restrict__ pool, double u, double v) {
    u = dclamp01(u);
    v = dclamp01(v);
    const int w = int(T.w), h = int(T.h);
    const double fx = u * double(w - 1), fy = v * double(h - 1);
    const int x0 = int(floor(fx)), y0 = int(floor(fy));
    const int x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
    const double tx = fx - double(x0), ty = fy - double(y0);
    D3 c00, c10, c01, c11;
    if (T.texel_off == kConstTex) {
        c00 = c10 = c01 = c11 = D3{T.base[0], T.base[1], T.base[2]};
    } else {
        const double4* t = pool + T.texel_off;
        const double4 a = t[size_t(y0) * w + x0], b = t[size_t(y0) * w + x1];
        const double4 c = t[size_t(y1) * w + x0], d = t[size_t(y1) * w + x1];
        c00 = {a.x, a.y, a.z};
        c10 = {b.x, b.y, b.z};
        c01 = {c.x, c.y, c.z};
        c11 = {d.x, d.y, d.z};
    }
    const double sx = 1.0 - tx, sy = 1.0 - ty;
    const D3 r0{c00.x * sx + c10.x * tx, c00.y * sx + c10.y * tx, c00.z * sx + c10.z * tx};
    const D3 r1{c01.x * sx + c11.x * tx, c01.y * sx + c11.y * tx, c01.z * sx + c11.z * tx};
    return {r0.x * sy + r1.x * ty, r0.y * sy + r1.y * ty, r0.z * sy + r1.z * ty};
}

// the colour the reference's painter leaves at pixel (x, y) of an image
// painted by warps [begin, end): the last warp whose box holds the pixel
// and whose barycentric test passes (rt10.cpp:414-434), sampled; `c` is
// kept when none covers it.  The wave's pixels lie in the rectangle
// [rx0, rx1] x [ry0, ry1]: 64 warps at a time, from the last, each lane
// loads one warp's box and a ballot keeps those that meet the rectangle;
// the wave then tests those warps in order (wave-uniform index, scalar
// loads of the record), each lane until one covers its pixel, the wave
// until every lane is covered.
__device__ __forceinline__ void paint(int x, int y, bool live, int rx0, int rx1, int ry0, int ry1, uint32_t begin,
                                      uint32_t end, const PatchWarp* __restrict__ warps,
                                      const PatchTex* __restrict__ texs, const double4* __restrict__ pool, D3& c) {
    const double px = double(x) + 0.5, py = double(y) + 0.5;
    const int lane = int(threadIdx.x & 63);
    bool todo = live;
    for (int top = int(end); top > int(begin); top -= 64) {
        if (__ballot(todo) == 0) return;
        const int wi = top - 1 - lane;  // lane i: the i-th warp below top
        bool meets = false;
        if (wi >= int(begin)) {
            const int4 bb = *reinterpret_cast<const int4*>(&warps[wi].x0);
            meets = bb.x <= rx1 && bb.y >= rx0 && bb.z <= ry1 && bb.w >= ry0;
        }
        for (uint64_t m = __ballot(meets); m != 0; m &= m - 1) {
            const int i = top - 1 - __builtin_ctzll(m);
            const PatchWarp& W = warps[i];
            if (todo && x >= W.x0 && x <= W.x1 && y >= W.y0 && y <= W.y1) {
                const double ax = W.p[0] - px, ay = W.p[1] - py;
                const double bx = W.p[2] - px, by = W.p[3] - py;
                const double cx = W.p[4] - px, cy = W.p[5] - py;
                const double w0 = (bx * cy - by * cx) / W.area;
                const double w1 = (cx * ay - cy * ax) / W.area;
                const double w2 = 1.0 - w0 - w1;
                if (!(w0 < -1e-6 || w1 < -1e-6 || w2 < -1e-6)) {  // the reference's test, NaN included
                    const double su = W.s[0] * w0 + W.s[2] * w1 + W.s[4] * w2;
                    const double sv = W.s[1] * w0 + W.s[3] * w1 + W.s[5] * w2;
                    c = sample(texs[W.child], pool, su, sv);
                    todo = false;
                }
            }
            if (__ballot(todo) == 0) return;
        }
    }
}

// one 8 x 8 tile of a recursed texture per wave: the reflection painted over
// the environment colour, then mixed with the albedo (rt10.cpp:652-661)
__global__ __launch_bounds__(64) void k_patch_level(const PatchTile* __restrict__ tiles,
                                                     const PatchTex* __restrict__ texs,
                                                     const PatchWarp* __restrict__ warps, double4* __restrict__ pool,
                                                     double e0, double e1, double e2) {
    const PatchTile tl = tiles[blockIdx.x];
    const PatchTex T = texs[tl.tex];
    const int x = int(tl.tx) * kTile + int(threadIdx.x & 7), y = int(tl.ty) * kTile + int(threadIdx.x >> 3);
    const bool live = x < int(T.w) && y < int(T.h);
    D3 r{e0, e1, e2};
    const int rx = int(tl.tx) * kTile, ry = int(tl.ty) * kTile;
    paint(x, y, live, rx, rx + kTile - 1, ry, ry + kTile - 1, T.warp_begin, T.warp_end, warps, texs, pool, r);
    if (!live) return;
    const double m = T.m, k = 1.0 - m;
    double4 o;
    o.x = T.base[0] * k + (r.x * T.base[0]) * m;
    o.y = T.base[1] * k + (r.y * T.base[1]) * m;
    o.z = T.base[2] * k + (r.z * T.base[2]) * m;
    o.w = 0.0;
    pool[T.texel_off + size_t(y) * T.w + x] = o;
}

struct RootArgs {
    double vp_uv[2][6];           // the viewports' uv triangles
    uint32_t begin[2], end[2];    // their warp ranges
    int w, h;
    double env[3];
};

// pointInTriangle2D of the reference (rt10.cpp:198-219), eps 1e-10
__device__ __forceinline__ bool inside_uv(double px, double py, const double* t) {
    const double eps = 1e-10;
    const double c0 = (t[2] - t[0]) * (py - t[1]) - (t[3] - t[1]) * (px - t[0]);
    const double c1 = (t[4] - t[2]) * (py - t[3]) - (t[5] - t[3]) * (px - t[2]);
    const double c2 = (t[0] - t[4]) * (py - t[5]) - (t[1] - t[5]) * (px - t[4]);
    const bool neg = c0 < -eps || c1 < -eps || c2 < -eps;
    const bool pos = c0 > eps || c1 > eps || c2 > eps;
    return !(neg && pos);
}

// the image: each viewport's painted pixel (black outside its triangle),
// summed and clamped (rt10.cpp:743-787); 16 x 16 pixels per block
__global__ __launch_bounds__(256) void k_patch_root(RootArgs a, const PatchTex* __restrict__ texs,
                                                    const PatchWarp* __restrict__ warps,
                                                    const double4* __restrict__ pool, double* __restrict__ out) {
    const int x = int(blockIdx.x) * 16 + int(threadIdx.x & 15), y = int(blockIdx.y) * 16 + int(threadIdx.x >> 4);
    const bool live = x < a.w && y < a.h;
    const double u = double(x) / double(a.w - 1), v = double(y) / double(a.h - 1);
    // a wave covers 16 x 4 pixels
    const int rx = int(blockIdx.x) * 16, ry = int(blockIdx.y) * 16 + int(threadIdx.x >> 6) * 4;
    D3 c[2];
    for (int k = 0; k < 2; ++k) {
        c[k] = D3{a.env[0], a.env[1], a.env[2]};
        const bool in = live && inside_uv(u, v, a.vp_uv[k]);
        paint(x, y, in, rx, rx + 15, ry, ry + 3, a.begin[k], a.end[k], warps, texs, pool, c[k]);
        if (!in) c[k] = D3{0.0, 0.0, 0.0};
    }
    if (!live) return;
    double* o = out + 3 * (size_t(y) * a.w + x);
    o[0] = dclamp01(c[0].x + c[1].x);
    o[1] = dclamp01(c[0].y + c[1].y);
    o[2] = dclamp01(c[0].z + c[1].z);
}

Status invalid(const char* m) { return Status{AURORA_ERR_INVALID, m}; }

}  // namespace

Status patch_check(const aurora_patch_triangle* tris, uint32_t n_tris, const aurora_patch_material* mats,
                   uint32_t n_mats, const aurora_patch_camera* cam, const aurora_patch_config* cfg) {
    if (!cam || !cfg) return invalid("patch render: camera or config is NULL");
    if (n_tris && !tris) return invalid("patch render: triangles NULL with n_tris > 0");
    if (n_mats && !mats) return invalid("patch render: materials NULL with n_mats > 0");
    if (cam->width < 2 || cam->height < 2 || cam->width > 65536 || cam->height > 65536)
        return invalid("patch render: image sides must be in [2, 65536]");
    if (cfg->min_tex_res < 1 || cfg->max_tex_res < cfg->min_tex_res || cfg->max_tex_res > AURORA_PATCH_MAX_TEX_RES)
        return invalid("patch render: need 1 <= min_tex_res <= max_tex_res <= AURORA_PATCH_MAX_TEX_RES");
    if (cfg->max_depth < 0 || cfg->max_depth > AURORA_PATCH_MAX_DEPTH)
        return invalid("patch render: max_depth outside [0, AURORA_PATCH_MAX_DEPTH]");
    if (!std::isfinite(cfg->min_area_px)) return invalid("patch render: min_area_px not finite");
    for (uint32_t i = 0; i < n_mats; ++i)
        if (mats[i].type != AURORA_PATCH_DIFFUSE && mats[i].type != AURORA_PATCH_REFLECTIVE)
            return invalid("patch render: unknown material type");
    auto finite_tri = [](const aurora_patch_triangle& t) {
        for (int i = 0; i < 3; ++i) {
            for (int k = 0; k < 3; ++k)
                if (!std::isfinite(t.p[i][k])) return false;
            for (int k = 0; k < 2; ++k)
                if (!std::isfinite(t.uv[i][k])) return false;
        }
        return true;
    };
    for (uint32_t i = 0; i < n_tris; ++i) {
        if (tris[i].material != AURORA_PATCH_NO_MATERIAL && tris[i].material >= n_mats)
            return invalid("patch render: material index out of range");
        if (!finite_tri(tris[i])) return invalid("patch render: non-finite vertex or uv");
    }
    for (int k = 0; k < 2; ++k)
        if (!finite_tri(cam->viewport[k])) return invalid("patch render: non-finite viewport vertex or uv");
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(cam->origin[k])) return invalid("patch render: non-finite origin");
    return Status::ok();
}

Status patch_render_device(PatchWork& pw, const aurora_patch_triangle* tris, uint32_t n_tris,
                           const aurora_patch_material* mats, const aurora_patch_camera* cam,
                           const aurora_patch_config* cfg, double* d_out, hipStream_t s, aurora_patch_stats* st) {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<HTri> scene;
    scene.reserve(n_tris);
    for (uint32_t i = 0; i < n_tris; ++i) scene.push_back(host_tri(tris[i], mats));
    const P3 eye{cam->origin[0], cam->origin[1], cam->origin[2]};
    const int W = int(cam->width), H = int(cam->height);
    const HTri vps[2] = {host_tri(cam->viewport[0], nullptr), host_tri(cam->viewport[1], nullptr)};
    // one plan per calling thread, kept between frames (see reset)
    static thread_local FramePlan plan;
    plan.reset();
    const Status planned = plan_frame(scene, *cfg, eye, vps, W, H, plan);
    RootArgs ra{};
    for (int k = 0; k < 2; ++k) {
        ra.begin[k] = plan.begin[k];
        ra.end[k] = plan.end[k];
        for (int j = 0; j < 3; ++j) {
            ra.vp_uv[k][2 * j] = vps[k].uv[j].x;
            ra.vp_uv[k][2 * j + 1] = vps[k].uv[j].y;
        }
    }
    ra.w = W;
    ra.h = H;
    for (int k = 0; k < 3; ++k) ra.env[k] = cfg->env[k];

    // tiles of the rasterised textures, grouped by depth
    const int levels = plan.max_level + 1;
    std::vector<std::vector<PatchTile>> by_level(size_t(std::max(levels, 0)));
    uint32_t raster = 0;
    uint64_t texels = uint64_t(W) * uint64_t(H);
    for (size_t i = 0; i < plan.texs.size(); ++i) {
        const PatchTex& T = plan.texs[i];
        if (T.texel_off == kConstTex) continue;
        ++raster;
        texels += uint64_t(T.w) * T.h;
        const uint32_t nx = (T.w + kTile - 1) / kTile, ny = (T.h + kTile - 1) / kTile;
        for (uint32_t ty = 0; ty < ny; ++ty)
            for (uint32_t tx = 0; tx < nx; ++tx)
                by_level[size_t(plan.tex_depth[i])].push_back(PatchTile{uint32_t(i), uint16_t(tx), uint16_t(ty)});
    }
    std::vector<PatchTile> tiles;
    std::vector<size_t> level_off(size_t(std::max(levels, 0)) + 1, 0);
    for (int l = 0; l < levels; ++l) {
        level_off[size_t(l)] = tiles.size();
        tiles.insert(tiles.end(), by_level[size_t(l)].begin(), by_level[size_t(l)].end());
    }
    if (levels > 0) level_off[size_t(levels)] = tiles.size();
    const auto t1 = std::chrono::steady_clock::now();
    if (st) {
        st->textures = uint32_t(plan.texs.size());
        st->raster_textures = raster;
        st->levels = uint32_t(std::max(levels, 0));
        st->warps = uint32_t(plan.warps.size());
        st->texels = texels;
        st->warp_texels = plan.warp_texels;
        st->plan_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st->kernel_ms = 0.0;
    }
    AUR_TRY(planned);

    AUR_TRY(pw.texs.alloc(std::max<size_t>(plan.texs.size(), 1) * sizeof(PatchTex)));
    AUR_TRY(pw.warps.alloc(std::max<size_t>(plan.warps.size(), 1) * sizeof(PatchWarp)));
    AUR_TRY(pw.tiles.alloc(std::max<size_t>(tiles.size(), 1) * sizeof(PatchTile)));
    AUR_TRY(pw.pool.alloc(std::max<uint64_t>(plan.pool_texels, 1) * sizeof(double4)));
    if (!plan.texs.empty())
        AUR_HIP(hipMemcpyAsync(pw.texs.p, plan.texs.data(), plan.texs.size() * sizeof(PatchTex),
                               hipMemcpyHostToDevice, s));
    if (!plan.warps.empty())
        AUR_HIP(hipMemcpyAsync(pw.warps.p, plan.warps.data(), plan.warps.size() * sizeof(PatchWarp),
                               hipMemcpyHostToDevice, s));
    if (!tiles.empty())
        AUR_HIP(hipMemcpyAsync(pw.tiles.p, tiles.data(), tiles.size() * sizeof(PatchTile), hipMemcpyHostToDevice, s));
    EventPair ev;
    AUR_TRY(ev.create());
    AUR_HIP(hipEventRecord(ev.e[0], s));
    for (int l = levels - 1; l >= 0; --l) {
        const size_t n = level_off[size_t(l) + 1] - level_off[size_t(l)];
        if (n == 0) continue;
        k_patch_level<<<dim3(uint32_t(n)), dim3(64), 0, s>>>(pw.tiles.as<PatchTile>() + level_off[size_t(l)],
                                                             pw.texs.as<PatchTex>(), pw.warps.as<PatchWarp>(),
                                                             pw.pool.as<double4>(), cfg->env[0], cfg->env[1],
                                                             cfg->env[2]);
        AUR_HIP(hipGetLastError());
    }
    k_patch_root<<<dim3(uint32_t((W + 15) / 16), uint32_t((H + 15) / 16)), dim3(256), 0, s>>>(
        ra, pw.texs.as<PatchTex>(), pw.warps.as<PatchWarp>(), pw.pool.as<double4>(), d_out);
    AUR_HIP(hipGetLastError());
    AUR_HIP(hipEventRecord(ev.e[1], s));
    AUR_HIP(hipEventSynchronize(ev.e[1]));
    if (st) {
        float ms = 0.0f;
        AUR_HIP(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
        st->kernel_ms = ms;
    }
    return Status::ok();
}

}  // namespace aur
