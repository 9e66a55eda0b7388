// oracle/orc_scene.cpp — CPU oracle: scene validation, textures, LBVH build,
// closest-hit traversal.  TEST INFRASTRUCTURE ONLY (see orc_core.hpp).
//
// Reference anchors:
//   quad / triangle test: Möller–Trumbore with inclusive edges as in
//     are::Triangle::intersect_ray, src/object/triangle.cpp:82-121
//     (vertices Q, Q+u, Q+v: triangle.cpp:42-45), here in the working
//     precision R with the interval (tmin, tmax) instead of t > 1e-12.
//   degenerate-edge rejection: triangle.cpp:25-36 (near_zero, vec3.cpp:143).
//   brute-force closest hit: the linear scan of experiments/rt.cpp:209-218.
//   image texture bilinear sampling: src/texture.cpp:310-330 (clamp, floor,
//     neighbour clamp, four weights in the same order).
//   sphere, spatial checker, Perlin marble: no reference counterpart (the
//     reference has none, and its checker at experiments/rt.cpp:95-102 is a
//     (u,v) checker); textbook definitions from "Ray Tracing: The Next Week".
#include "orc_scene.hpp"

#include <algorithm>
#include <cstdio>
#include <functional>
#include <stdexcept>

namespace orc {

static bool finite3(const float* v) {
    return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}

std::string validate(const aurora_scene_desc& d) {
    char buf[256];
    if (d.n_spheres && (!d.spheres || !d.sphere_material)) return "sphere arrays are NULL";
    if (d.n_quads && (!d.quads || !d.quad_material)) return "quad arrays are NULL";
    if (d.n_materials && !d.materials) return "material table is NULL";
    if (d.n_textures && !d.textures) return "texture table is NULL";
    if (d.n_texels && !d.texels) return "texel array is NULL";
    if (uint64_t(d.n_spheres) + d.n_quads > 0x7fffffffull) return "too many primitives";
    for (uint32_t i = 0; i < d.n_spheres; ++i) {
        const float* s = d.spheres + 4 * i;
        if (!finite3(s) || !std::isfinite(s[3]) || !(s[3] > 0.0f)) {
            snprintf(buf, sizeof buf, "sphere %u: non-finite centre or radius <= 0", i);
            return buf;
        }
        if (d.sphere_material[i] >= d.n_materials) {
            snprintf(buf, sizeof buf, "sphere %u: material index out of range", i);
            return buf;
        }
    }
    for (uint32_t i = 0; i < d.n_quads; ++i) {
        const float* q = d.quads + 9 * i;
        if (!finite3(q) || !finite3(q + 3) || !finite3(q + 6)) {
            snprintf(buf, sizeof buf, "quad %u: non-finite vertex or edge", i);
            return buf;
        }
        V3<double> u = mk<double>(q[3], q[4], q[5]), v = mk<double>(q[6], q[7], q[8]);
        if (near_zero(u)) { snprintf(buf, sizeof buf, "quad %u: edge vector u cannot be zero vector", i); return buf; }
        if (near_zero(v)) { snprintf(buf, sizeof buf, "quad %u: edge vector v cannot be zero vector", i); return buf; }
        if (near_zero(cross(u, v))) { snprintf(buf, sizeof buf, "quad %u: edge vectors u and v cannot be collinear", i); return buf; }
        if (d.quad_material[i] >= d.n_materials) {
            snprintf(buf, sizeof buf, "quad %u: material index out of range", i);
            return buf;
        }
        if (d.quad_kind && d.quad_kind[i] > AURORA_QUAD_TRIANGLE) {
            snprintf(buf, sizeof buf, "quad %u: unknown kind", i);
            return buf;
        }
    }
    for (uint32_t i = 0; i < d.n_materials; ++i) {
        const aurora_material& m = d.materials[i];
        if (m.type > AURORA_MAT_DIFFUSE_LIGHT) { snprintf(buf, sizeof buf, "material %u: unknown type", i); return buf; }
        if (m.type != AURORA_MAT_DIELECTRIC && m.texture >= d.n_textures) {
            snprintf(buf, sizeof buf, "material %u: texture index out of range", i);
            return buf;
        }
        if (m.type == AURORA_MAT_DIELECTRIC && !(m.ior > 0.0f && std::isfinite(m.ior))) {
            snprintf(buf, sizeof buf, "material %u: ior must be positive", i);
            return buf;
        }
        if (m.type == AURORA_MAT_METAL && !(m.fuzz >= 0.0f)) {
            snprintf(buf, sizeof buf, "material %u: fuzz must be >= 0", i);
            return buf;
        }
    }
    for (uint32_t i = 0; i < d.n_textures; ++i) {
        const aurora_texture& t = d.textures[i];
        if (t.type > AURORA_TEX_IMAGE) { snprintf(buf, sizeof buf, "texture %u: unknown type", i); return buf; }
        if (t.type == AURORA_TEX_CHECKER) {
            if (t.even >= d.n_textures || t.odd >= d.n_textures) { snprintf(buf, sizeof buf, "texture %u: checker child out of range", i); return buf; }
            if (d.textures[t.even].type == AURORA_TEX_CHECKER || d.textures[t.odd].type == AURORA_TEX_CHECKER) {
                snprintf(buf, sizeof buf, "texture %u: checker children cannot be checkers", i);
                return buf;
            }
            if (!(t.scale > 0.0f)) { snprintf(buf, sizeof buf, "texture %u: checker scale must be positive", i); return buf; }
        }
        if (t.type == AURORA_TEX_IMAGE) {
            if (t.width == 0 || t.height == 0) { snprintf(buf, sizeof buf, "texture %u: image width and height must be positive", i); return buf; }
            if (t.texel_offset + uint64_t(t.width) * t.height > d.n_texels) {
                snprintf(buf, sizeof buf, "texture %u: image texels out of range", i);
                return buf;
            }
        }
    }
    return "";
}

void perlin_tables(uint32_t seed, Perlin& p) {
    // gradient vectors: unit(random in [-1,1)^3), one Philox draw each
    for (uint32_t i = 0; i < 256; ++i) {
        U4 r = philox4x32_10(U4{i, 0u, 0u, 0u}, seed, STREAM_PERLIN);
        V3<float> v = mk(2.0f * u01f(r.x) - 1.0f, 2.0f * u01f(r.y) - 1.0f, 2.0f * u01f(r.z) - 1.0f);
        float l2 = dot(v, v);
        if (l2 == 0.0f) v = mk(1.0f, 0.0f, 0.0f);
        v = unit(v);
        p.ranvec[i][0] = v.x; p.ranvec[i][1] = v.y; p.ranvec[i][2] = v.z;
    }
    // three Fisher–Yates permutations of 0..255
    for (uint32_t a = 0; a < 3; ++a) {
        for (int i = 0; i < 256; ++i) p.perm[a][i] = i;
        for (uint32_t i = 255; i >= 1; --i) {
            U4 r = philox4x32_10(U4{i, a + 1u, 0u, 0u}, seed, STREAM_PERLIN);
            uint32_t j = uint32_t((uint64_t(r.x) * (i + 1u)) >> 32);
            std::swap(p.perm[a][i], p.perm[a][j]);
        }
    }
}

void load(Scene& s, const aurora_scene_desc& d) {
    std::string err = validate(d);
    if (!err.empty()) throw std::invalid_argument(err);
    s = Scene();
    s.spheres.resize(d.n_spheres);
    s.sphere_mat.assign(d.sphere_material, d.sphere_material + d.n_spheres);
    for (uint32_t i = 0; i < d.n_spheres; ++i) {
        const float* p = d.spheres + 4 * i;
        s.spheres[i] = Sphere{{p[0], p[1], p[2]}, p[3]};
    }
    s.quads.resize(d.n_quads);
    s.quad_mat.assign(d.quad_material, d.quad_material + d.n_quads);
    for (uint32_t i = 0; i < d.n_quads; ++i) {
        const float* p = d.quads + 9 * i;
        Quad& q = s.quads[i];
        for (int k = 0; k < 3; ++k) { q.Q[k] = p[k]; q.u[k] = p[3 + k]; q.v[k] = p[6 + k]; }
        q.kind = d.quad_kind ? d.quad_kind[i] : AURORA_QUAD_PARALLELOGRAM;
    }
    s.mats.assign(d.materials, d.materials + d.n_materials);
    for (auto& m : s.mats)
        if (m.type == AURORA_MAT_METAL && m.fuzz > 1.0f) m.fuzz = 1.0f;
    s.texs.assign(d.textures, d.textures + d.n_textures);
    s.texels.assign(d.texels, d.texels + 3 * d.n_texels);
    s.perlin.resize(d.n_textures);
    for (uint32_t i = 0; i < d.n_textures; ++i)
        if (s.texs[i].type == AURORA_TEX_NOISE) perlin_tables(s.texs[i].seed, s.perlin[i]);
}

// ------------------------------------------------------------ bounds ---
static void pad_box(float lo[3], float hi[3]) {
    float m = 0.0f;
    for (int k = 0; k < 3; ++k) m = std::fmax(m, std::fmax(std::fabs(lo[k]), std::fabs(hi[k])));
    float e = m * (1.0f / 65536.0f) + 1e-30f;
    for (int k = 0; k < 3; ++k) { lo[k] = lo[k] - e; hi[k] = hi[k] + e; }
}

void prim_box(const Scene& s, uint32_t id, float lo[3], float hi[3]) {
    uint32_t ns = uint32_t(s.spheres.size());
    if (id < ns) {
        const Sphere& sp = s.spheres[id];
        for (int k = 0; k < 3; ++k) { lo[k] = sp.c[k] - sp.r; hi[k] = sp.c[k] + sp.r; }
    } else {
        const Quad& q = s.quads[id - ns];
        for (int k = 0; k < 3; ++k) {
            float a = q.Q[k], b = q.Q[k] + q.u[k], c = q.Q[k] + q.v[k];
            float l = std::fmin(a, std::fmin(b, c)), h = std::fmax(a, std::fmax(b, c));
            if (q.kind == AURORA_QUAD_PARALLELOGRAM) {
                float d4 = (q.Q[k] + q.u[k]) + q.v[k];
                l = std::fmin(l, d4);
                h = std::fmax(h, d4);
            }
            lo[k] = l;
            hi[k] = h;
        }
    }
    pad_box(lo, hi);
}

// ---------------------------------------------------------------- LBVH ---
// Karras, "Maximizing parallelism in the construction of BVHs, octrees and
// k-d trees" (HPG 2012): 30-bit Morton codes of box centroids, sorted
// stably; keys made unique by appending the sorted position; each internal
// node i finds its range [i, j] and split gamma from the longest common
// prefix function delta.
static uint32_t expand_bits(uint32_t v) {
    // spread the low 10 bits of v so that bit b lands at bit 3b
    uint32_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}

static uint32_t quantise(float c, float lo, float hi) {
    float e = hi - lo;
    float x = e > 0.0f ? (c - lo) / e : 0.0f;
    float q = std::fmin(std::fmax(x * 1024.0f, 0.0f), 1023.0f);
    return uint32_t(q);
}

// Four-wide traversal tree.  A wide node is made for the root and for every
// internal binary node that a wide node's slots refer to, numbered in Karras
// index order.  Its slots are a cut below the binary node: for scenes of at
// least kWideGreedyMin primitives, greedy by surface area (start with the two
// children; while fewer than four, the internal slot with the largest box
// surface area, dx*dy + dy*dz + dz*dx in float, the first in slot order on
// ties, is replaced in place by its two children); for smaller scenes,
// pairwise (each internal child replaced by its two children, so the wide
// nodes are the even-depth binary nodes).  A leaf slot refers to its
// primitive, an internal slot to that node's wide index.  (The GPU builds
// the same tree: aurora lbvh.hip wide_cut, k_wide_mark, k_wide_fill.)  No
// reference counterpart.
static constexpr uint32_t kWideGreedyMin = 4096;
static constexpr uint32_t kStack = 96;  // entries of the traversal stack

static void collapse_wide(Scene& s, const uint32_t* gamma, const uint8_t* lleaf, const uint8_t* rleaf,
                          uint32_t ni) {
    s.wide.clear();
    const uint32_t n = s.n_prims();
    auto slot_of = [&](uint32_t pos, uint32_t ref) {
        aurora_bvh4_slot sl{};
        for (int k = 0; k < 3; ++k) { sl.lo[k] = s.nodes[pos].lo[k]; sl.hi[k] = s.nodes[pos].hi[k]; }
        sl.ref = ref;
        sl.pad = 0;
        return sl;
    };
    aurora_bvh4_slot empty{};
    empty.ref = AURORA_BVH4_EMPTY;
    if (n == 0) { s.stats.n_wide = 0; return; }
    if (n == 1) {
        aurora_bvh4_node w{};
        w.slot[0] = slot_of(0, s.nodes[0].child);
        for (int k = 1; k < 4; ++k) w.slot[k] = empty;
        s.wide.push_back(w);
        s.stats.n_wide = 1;
        return;
    }
    // a slot: binary node position, leaf or not, internal index if not
    struct Ref { uint32_t pos; bool leaf; uint32_t j; };
    auto area = [&](uint32_t pos) {
        const aurora_bvh_node& nd = s.nodes[pos];
        const float dx = nd.hi[0] - nd.lo[0], dy = nd.hi[1] - nd.lo[1], dz = nd.hi[2] - nd.lo[2];
        return dx * dy + dy * dz + dz * dx;
    };
    bool greedy = n >= kWideGreedyMin;
    // children of internal i: positions 2+2g (internal index g) and 3+2g (g+1)
    auto cut = [&](uint32_t i, Ref* out) {
        const uint32_t g = gamma[i];
        if (!greedy) {
            int k = 0;
            for (uint32_t c = 0; c < 2; ++c) {
                const uint32_t j = g + c;
                if (c == 0 ? lleaf[i] : rleaf[i]) {
                    out[k++] = {2 + 2 * g + c, true, j};
                    continue;
                }
                const uint32_t gj = gamma[j];
                out[k++] = {2 + 2 * gj, lleaf[j] != 0, gj};
                out[k++] = {3 + 2 * gj, rleaf[j] != 0, gj + 1};
            }
            return k;
        }
        out[0] = {2 + 2 * g, lleaf[i] != 0, g};
        out[1] = {3 + 2 * g, rleaf[i] != 0, g + 1};
        int k = 2;
        while (k < 4) {
            int best = -1;
            float ba = 0.0f;
            for (int q = 0; q < k; ++q)
                if (!out[q].leaf) {
                    const float a = area(out[q].pos);
                    if (best < 0 || a > ba) { best = q; ba = a; }
                }
            if (best < 0) break;
            const uint32_t j = out[best].j, gj = gamma[j];
            for (int q = k; q > best + 1; --q) out[q] = out[q - 1];
            out[best] = {2 + 2 * gj, lleaf[j] != 0, gj};
            out[best + 1] = {3 + 2 * gj, rleaf[j] != 0, gj + 1};
            ++k;
        }
        return k;
    };
    // wide level of every internal node a cut refers to (root: 1).  The walk
    // pushes at most three nodes per wide level above the deepest, so the
    // greedy form is kept only while 3 * (levels - 1) fits the 96-entry
    // traversal stack; otherwise the pairwise form is built (ceil(depth/2)
    // levels: fits every tree of depth <= 64).
    std::vector<uint32_t> level(ni, 0);
    Ref r[4];
    uint32_t wide_depth = 1;
    for (int attempt = 0; attempt < 2; ++attempt) {
        std::fill(level.begin(), level.end(), 0u);
        level[0] = 1;
        wide_depth = 1;
        std::vector<uint32_t> st{0};
        while (!st.empty()) {
            const uint32_t i = st.back();
            st.pop_back();
            const int k = cut(i, r);
            for (int q = 0; q < k; ++q)
                if (!r[q].leaf) {
                    level[r[q].j] = level[i] + 1;
                    wide_depth = std::max(wide_depth, level[i] + 1);
                    st.push_back(r[q].j);
                }
        }
        if (3 * (wide_depth - 1) <= kStack) break;
        if (!greedy) throw std::runtime_error("four-wide tree deeper than the traversal stack");
        greedy = false;
    }
    std::vector<uint8_t> root(ni, 0);
    for (uint32_t i = 0; i < ni; ++i) root[i] = level[i] != 0;
    s.stats.wide_depth = wide_depth;
    s.stats.wide_form = greedy ? AURORA_BVH4_GREEDY : AURORA_BVH4_PAIRWISE;
    std::vector<uint32_t> widx(ni, 0xffffffffu);
    uint32_t nw = 0;
    for (uint32_t i = 0; i < ni; ++i)
        if (root[i]) widx[i] = nw++;
    s.wide.assign(nw, aurora_bvh4_node{});
    for (uint32_t i = 0; i < ni; ++i) {
        if (!root[i]) continue;
        const int k = cut(i, r);
        aurora_bvh4_node& w = s.wide[widx[i]];
        for (int q = 0; q < k; ++q)
            w.slot[q] = slot_of(r[q].pos, r[q].leaf ? s.nodes[r[q].pos].child : widx[r[q].j]);
        for (int q = k; q < 4; ++q) w.slot[q] = empty;
    }
    s.stats.n_wide = nw;
}

void build_lbvh(Scene& s) {
    const uint32_t n = s.n_prims();
    s.nodes.clear();
    s.wide.clear();
    s.sorted_morton.clear();
    s.order.clear();
    s.stats = aurora_bvh_stats{};
    s.stats.n_prims = n;
    s.built = true;
    if (n == 0) return;

    std::vector<float> blo(3 * size_t(n)), bhi(3 * size_t(n));
    float slo[3] = {INFINITY, INFINITY, INFINITY}, shi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    std::vector<float> cen(3 * size_t(n));
    for (uint32_t i = 0; i < n; ++i) {
        prim_box(s, i, &blo[3 * size_t(i)], &bhi[3 * size_t(i)]);
        for (int k = 0; k < 3; ++k) {
            float l = blo[3 * size_t(i) + k], h = bhi[3 * size_t(i) + k];
            float c = (l + h) * 0.5f;
            cen[3 * size_t(i) + k] = c;
            slo[k] = std::fmin(slo[k], l);
            shi[k] = std::fmax(shi[k], h);
            clo[k] = std::fmin(clo[k], c);
            chi[k] = std::fmax(chi[k], c);
        }
    }
    std::vector<uint32_t> code(n);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t q[3];
        for (int k = 0; k < 3; ++k) q[k] = quantise(cen[3 * size_t(i) + k], clo[k], chi[k]);
        code[i] = (expand_bits(q[0]) << 2) | (expand_bits(q[1]) << 1) | expand_bits(q[2]);
    }
    s.order.resize(n);
    for (uint32_t i = 0; i < n; ++i) s.order[i] = i;
    std::stable_sort(s.order.begin(), s.order.end(), [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
    s.sorted_morton.resize(n);
    for (uint32_t i = 0; i < n; ++i) s.sorted_morton[i] = code[s.order[i]];
    for (int k = 0; k < 3; ++k) {
        s.stats.scene_lo[k] = slo[k];
        s.stats.scene_hi[k] = shi[k];
        s.stats.centroid_lo[k] = clo[k];
        s.stats.centroid_hi[k] = chi[k];
    }
    s.stats.n_nodes = n == 1 ? 1 : 2 * n;
    s.nodes.assign(s.stats.n_nodes, aurora_bvh_node{});
    if (n > 1) s.nodes[1].parent = 0xffffffffu;  // unused padding node

    auto leaf_node = [&](uint32_t m, uint32_t parent_pos) {
        aurora_bvh_node nd{};
        uint32_t p = s.order[m];
        for (int k = 0; k < 3; ++k) { nd.lo[k] = blo[3 * size_t(p) + k]; nd.hi[k] = bhi[3 * size_t(p) + k]; }
        nd.child = AURORA_NODE_LEAF | p;
        nd.parent = parent_pos;
        return nd;
    };
    if (n == 1) {
        s.nodes[0] = leaf_node(0, 0xffffffffu);
        s.stats.depth = 0;
        collapse_wide(s, nullptr, nullptr, nullptr, 0);
        return;
    }

    const std::vector<uint32_t>& mc = s.sorted_morton;
    auto delta = [&](int64_t i, int64_t j) -> int {
        if (j < 0 || j >= int64_t(n)) return -1;
        uint64_t ki = (uint64_t(mc[size_t(i)]) << 32) | uint64_t(i);
        uint64_t kj = (uint64_t(mc[size_t(j)]) << 32) | uint64_t(j);
        return __builtin_clzll(ki ^ kj);
    };
    const uint32_t ni = n - 1;
    std::vector<uint32_t> gamma(ni), int_parent(ni, 0xffffffffu), leaf_parent(n), int_pos(ni), leaf_pos(n);
    std::vector<uint8_t> lleaf(ni), rleaf(ni);
    for (int64_t i = 0; i < int64_t(ni); ++i) {
        int d = (delta(i, i + 1) - delta(i, i - 1)) >= 0 ? 1 : -1;
        int dmin = delta(i, i - d);
        int64_t lmax = 2;
        while (delta(i, i + lmax * d) > dmin) lmax *= 2;
        int64_t l = 0;
        for (int64_t t = lmax / 2; t >= 1; t /= 2)
            if (delta(i, i + (l + t) * d) > dmin) l += t;
        int64_t j = i + l * d;
        int dnode = delta(i, j);
        int64_t sp = 0, t = l;
        do {
            t = (t + 1) / 2;
            if (delta(i, i + (sp + t) * d) > dnode) sp += t;
        } while (t > 1);
        int64_t g = i + sp * d + std::min(d, 0);
        gamma[size_t(i)] = uint32_t(g);
        lleaf[size_t(i)] = std::min(i, j) == g;
        rleaf[size_t(i)] = std::max(i, j) == g + 1;
    }
    int_pos[0] = 0;
    for (uint32_t i = 0; i < ni; ++i) {
        uint32_t g = gamma[i];
        // children of the node with split g occupy the aligned pair 2+2g, 3+2g
        if (lleaf[i]) { leaf_parent[g] = i; leaf_pos[g] = 2 + 2 * g; }
        else { int_parent[g] = i; int_pos[g] = 2 + 2 * g; }
        if (rleaf[i]) { leaf_parent[g + 1] = i; leaf_pos[g + 1] = 3 + 2 * g; }
        else { int_parent[g + 1] = i; int_pos[g + 1] = 3 + 2 * g; }
    }
    for (uint32_t m = 0; m < n; ++m) s.nodes[leaf_pos[m]] = leaf_node(m, int_pos[leaf_parent[m]]);
    // internal boxes bottom-up (post-order), depth
    std::vector<std::pair<uint32_t, int>> st;  // (internal node, state)
    st.push_back({0u, 0});
    uint32_t maxdepth = 0;
    std::vector<uint32_t> idepth(ni, 0);
    while (!st.empty()) {
        auto& top = st.back();
        uint32_t i = top.first;
        uint32_t g = gamma[i];
        if (top.second == 0) {
            top.second = 1;
            if (!lleaf[i]) { idepth[g] = idepth[i] + 1; st.push_back({g, 0}); }
            else maxdepth = std::max(maxdepth, idepth[i] + 1);
            if (!rleaf[i]) { idepth[g + 1] = idepth[i] + 1; st.push_back({g + 1, 0}); }
            else maxdepth = std::max(maxdepth, idepth[i] + 1);
            continue;
        }
        st.pop_back();
        aurora_bvh_node& nd = s.nodes[int_pos[i]];
        const aurora_bvh_node& a = s.nodes[2 + 2 * g];
        const aurora_bvh_node& b = s.nodes[3 + 2 * g];
        for (int k = 0; k < 3; ++k) { nd.lo[k] = std::fmin(a.lo[k], b.lo[k]); nd.hi[k] = std::fmax(a.hi[k], b.hi[k]); }
        nd.child = 2 + 2 * g;
        nd.parent = i == 0 ? 0xffffffffu : int_pos[int_parent[i]];
    }
    s.stats.depth = maxdepth;
    if (maxdepth > kStack) throw std::runtime_error("LBVH deeper than the traversal stack");
    collapse_wide(s, gamma.data(), lleaf.data(), rleaf.data(), ni);
}

// ----------------------------------------------------- primitive tests ---
template <class R>
bool hit_sphere(const Sphere& sp, V3<R> o, V3<R> d, R tmin, R tmax, R& t) {
    // robust roots (Haines et al., Ray Tracing Gems 2019, ch. 7):
    // h^2 - a*c as a*r^2 - |d x oc|^2, roots c/q and q/a with
    // q = h + sign(h)*sqrt(disc); the near one (c/q when q > 0, else q/a) if
    // it lies in (tmin, tmax), else the far one
    V3<R> c = mk<R>(sp.c[0], sp.c[1], sp.c[2]);
    R r = R(sp.r);
    V3<R> oc = c - o;
    R a = dot(d, d);
    R h = dot(d, oc);
    R rr = r * r;
    R cc = dot(oc, oc) - rr;
    V3<R> x = cross(d, oc);
    R disc = a * rr - dot(x, x);
    if (!(disc >= R(0))) return false;
    R q = h + std::copysign(std::sqrt(disc), h);
    if (q == R(0)) return false;
    const bool pos = q > R(0);
    R root = pos ? cc / q : q / a;
    if (!(root > tmin && root < tmax)) {
        root = pos ? q / a : cc / q;
        if (!(root > tmin && root < tmax)) return false;
    }
    t = root;
    return true;
}

template <class R>
bool hit_quad(const Quad& q, V3<R> o, V3<R> d, R tmin, R tmax, R& t, R& al, R& be) {
    V3<R> Q = mk<R>(q.Q[0], q.Q[1], q.Q[2]);
    V3<R> u = mk<R>(q.u[0], q.u[1], q.u[2]);
    V3<R> v = mk<R>(q.v[0], q.v[1], q.v[2]);
    V3<R> h = cross(d, v);
    R a = dot(u, h);
    if (std::fabs(a) < R(1e-20)) return false;
    R f = R(1) / a;
    V3<R> s = o - Q;
    R alpha = f * dot(s, h);
    if (alpha < R(0) || alpha > R(1)) return false;
    V3<R> qv = cross(s, u);
    R beta = f * dot(d, qv);
    if (beta < R(0)) return false;
    if (q.kind == AURORA_QUAD_TRIANGLE ? (alpha + beta > R(1)) : (beta > R(1))) return false;
    R tt = f * dot(v, qv);
    if (!(tt > tmin && tt < tmax)) return false;
    t = tt;
    al = alpha;
    be = beta;
    return true;
}

template <class R>
static inline void test_prim(const Scene& s, uint32_t id, V3<R> o, V3<R> d, R tmin, R tmax, R tentry,
                             Hit<R>& best, Counts* cnt) {
    // tentry: the entry distance of the primitive's box.  A hit before it
    // does not count; with ties to the lower id, this makes the closest hit
    // independent of the tree and the visit order (ancestor boxes contain
    // their children's bounds as floats, and the slab arithmetic is
    // monotone in them, so no ancestor of a counted hit is culled while the
    // best t is at or above it)
    uint32_t ns = uint32_t(s.spheres.size());
    R t;
    bool h;
    if (id < ns) {
        if (cnt) cnt->spheres++;
        h = hit_sphere(s.spheres[id], o, d, tmin, tmax, t);
    } else {
        if (cnt) cnt->quads++;
        R a, b;
        h = hit_quad(s.quads[id - ns], o, d, tmin, tmax, t, a, b);
    }
    if (h && t >= tentry && (t < best.t || (t == best.t && int32_t(id) < best.prim))) {
        best.t = t;
        best.prim = int32_t(id);
    }
}

// slab test of a node box against (tmin, tcur]; fmin/fmax are IEEE
// minNum/maxNum, so a NaN from 0*inf drops out of the comparison.
template <class R>
static inline bool slab(const aurora_bvh_node& nd, V3<R> o, V3<R> inv, R tmin, R tcur, R& tnear) {
    R tx0 = (R(nd.lo[0]) - o.x) * inv.x, tx1 = (R(nd.hi[0]) - o.x) * inv.x;
    R ty0 = (R(nd.lo[1]) - o.y) * inv.y, ty1 = (R(nd.hi[1]) - o.y) * inv.y;
    R tz0 = (R(nd.lo[2]) - o.z) * inv.z, tz1 = (R(nd.hi[2]) - o.z) * inv.z;
    R tn = std::fmax(std::fmax(std::fmin(tx0, tx1), std::fmin(ty0, ty1)), std::fmax(std::fmin(tz0, tz1), tmin));
    R tf = std::fmin(std::fmin(std::fmax(tx0, tx1), std::fmax(ty0, ty1)), std::fmin(std::fmax(tz0, tz1), tcur));
    tnear = tn;
    return tn <= tf;
}

template <class R>
Hit<R> closest_bvh(const Scene& s, V3<R> o, V3<R> d, R tmin, R tmax, Counts* cnt) {
    Hit<R> best{tmax, 0x7fffffff};
    if (!s.built) throw std::logic_error("BVH not built");
    if (s.nodes.empty()) { best.prim = AURORA_MISS; return best; }
    V3<R> inv = mk<R>(R(1) / d.x, R(1) / d.y, R(1) / d.z);
    const aurora_bvh_node* N = s.nodes.data();
    R tn;
    if (cnt) cnt->nodes++;
    if (slab(N[0], o, inv, tmin, best.t, tn)) {
        if (N[0].child & AURORA_NODE_LEAF) {
            test_prim(s, N[0].child & ~AURORA_NODE_LEAF, o, d, tmin, tmax, tn, best, cnt);
        } else {
            uint32_t stack[kStack];  // one push per level: depth <= kStack (checked at build)
            int sp = 0;
            uint32_t cur = N[0].child;
            for (;;) {
                const aurora_bvh_node& a = N[cur];
                const aurora_bvh_node& b = N[cur + 1];
                if (cnt) cnt->nodes += 2;
                R ta, tb;
                bool ha = slab(a, o, inv, tmin, best.t, ta);
                bool hb = slab(b, o, inv, tmin, best.t, tb);
                if (ha && (a.child & AURORA_NODE_LEAF)) {
                    test_prim(s, a.child & ~AURORA_NODE_LEAF, o, d, tmin, tmax, ta, best, cnt);
                    ha = false;
                }
                if (hb && (b.child & AURORA_NODE_LEAF)) {
                    test_prim(s, b.child & ~AURORA_NODE_LEAF, o, d, tmin, tmax, tb, best, cnt);
                    hb = false;
                }
                if (ha && hb) {
                    if (sp >= int(kStack)) throw std::logic_error("binary traversal stack overflow");
                    if (tb < ta) { stack[sp++] = a.child; cur = b.child; }
                    else { stack[sp++] = b.child; cur = a.child; }
                } else if (ha) {
                    cur = a.child;
                } else if (hb) {
                    cur = b.child;
                } else {
                    if (sp == 0) break;
                    cur = stack[--sp];
                }
            }
        }
    }
    if (best.prim == 0x7fffffff) { best.prim = AURORA_MISS; best.t = tmax; }
    return best;
}

template <class R>
static inline bool slab4(const aurora_bvh4_slot& sl, V3<R> o, V3<R> inv, R tmin, R tcur, R& tnear) {
    aurora_bvh_node nd{};
    for (int k = 0; k < 3; ++k) { nd.lo[k] = sl.lo[k]; nd.hi[k] = sl.hi[k]; }
    return slab(nd, o, inv, tmin, tcur, tnear);
}

// Order key of a slot's entry distance: the unsigned bits of t + 0 (-0
// becomes +0), which for t >= +0 is the numeric order; a negative t (only
// when tmin < 0) sorts after every non-negative one.  Slots that are not hit
// are keyed all ones, so they sort last.
static inline uint32_t wide_order_bits(float t) {
    const float u = t + 0.0f;
    uint32_t b;
    std::memcpy(&b, &u, 4);
    return b;
}
static inline uint64_t wide_order_bits(double t) {
    const double u = t + 0.0;
    uint64_t b;
    std::memcpy(&b, &u, 8);
    return b;
}

// Nearest-first order of a four-wide node's hit internal slots, as the GPU
// step computes it (aurora traverse.hpp wide_step): the five comparators
// (0,1) (2,3) (0,2) (1,3) (1,2) on (key, slot) pairs, each swapping only
// when the second key is strictly smaller.  Equal keys stay where the network
// leaves them.  ord[0..m) receives the hit slots; returns m.
template <class R>
int wide_order(const bool hit[4], const R tn[4], int ord[4]) {
    using K = decltype(wide_order_bits(R(0)));
    K key[4];
    int m = 0;
    for (int k = 0; k < 4; ++k) {
        key[k] = hit[k] ? wide_order_bits(tn[k]) : ~K(0);
        ord[k] = k;
        m += hit[k] ? 1 : 0;
    }
    static const int net[5][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}};
    for (const auto& c : net)
        if (key[c[1]] < key[c[0]]) {
            std::swap(key[c[0]], key[c[1]]);
            std::swap(ord[c[0]], ord[c[1]]);
        }
    return m;
}
template int wide_order<float>(const bool[4], const float[4], int[4]);

// One step tests the four slots against the current best t, tests the hit
// leaf slots' primitives in slot order, then orders the hit internal slots
// nearest first (wide_order), descends to the nearest and pushes the others
// farthest first.
template <class R>
Hit<R> closest_bvh4(const Scene& s, V3<R> o, V3<R> d, R tmin, R tmax, Counts* cnt) {
    Hit<R> best{tmax, 0x7fffffff};
    if (!s.built) throw std::logic_error("BVH not built");
    if (s.wide.empty()) { best.prim = AURORA_MISS; return best; }
    V3<R> inv = mk<R>(R(1) / d.x, R(1) / d.y, R(1) / d.z);
    uint32_t stack[kStack];  // three pushes per wide level above the deepest (checked at build)
    int sp = 0;
    uint32_t cur = 0;
    for (;;) {
        if (cur >= s.wide.size()) throw std::logic_error("malformed four-wide tree");
        const aurora_bvh4_node& w = s.wide[cur];
        if (cnt) cnt->nodes += 4;
        bool hit[4];
        R tn[4];
        for (int k = 0; k < 4; ++k) {
            tn[k] = R(0);
            hit[k] = w.slot[k].ref != AURORA_BVH4_EMPTY && slab4(w.slot[k], o, inv, tmin, best.t, tn[k]);
        }
        for (int k = 0; k < 4; ++k)
            if (hit[k] && (w.slot[k].ref & AURORA_NODE_LEAF)) {
                test_prim(s, w.slot[k].ref & ~AURORA_NODE_LEAF, o, d, tmin, tmax, tn[k], best, cnt);
                hit[k] = false;
            }
        int ord[4];
        const int m = wide_order<R>(hit, tn, ord);
        if (m == 0) {
            if (sp == 0) break;
            cur = stack[--sp];
            continue;
        }
        for (int k = m - 1; k >= 1; --k) {
            if (sp >= int(kStack)) throw std::logic_error("four-wide traversal stack overflow");
            stack[sp++] = w.slot[ord[k]].ref;
        }
        cur = w.slot[ord[0]].ref;
    }
    if (best.prim == 0x7fffffff) { best.prim = AURORA_MISS; best.t = tmax; }
    return best;
}

template <class R>
Hit<R> closest_brute(const Scene& s, V3<R> o, V3<R> d, R tmin, R tmax) {
    Hit<R> best{tmax, 0x7fffffff};
    // the traversals' rule: the ray must enter the primitive's box (as the
    // leaf stores it), and a hit before the box's entry does not count
    const V3<R> inv = mk<R>(R(1) / d.x, R(1) / d.y, R(1) / d.z);
    for (uint32_t id = 0; id < s.n_prims(); ++id) {
        aurora_bvh_node nd{};
        prim_box(s, id, nd.lo, nd.hi);
        R te;
        if (!slab(nd, o, inv, tmin, tmax, te)) continue;
        test_prim<R>(s, id, o, d, tmin, tmax, te, best, nullptr);
    }
    if (best.prim == 0x7fffffff) { best.prim = AURORA_MISS; best.t = tmax; }
    return best;
}

// ------------------------------------------------------------- textures ---
template <class R>
R perlin_noise(const Perlin& pl, V3<R> p) {
    R fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
    R u = p.x - fx, v = p.y - fy, w = p.z - fz;
    auto clampi = [](R f) { return int(std::fmin(std::fmax(f, R(-1073741824.0)), R(1073741824.0))); };
    int i = clampi(fx), j = clampi(fy), k = clampi(fz);
    R uu = (u * u) * (R(3) - R(2) * u);
    R vv = (v * v) * (R(3) - R(2) * v);
    R ww = (w * w) * (R(3) - R(2) * w);
    R acc = R(0);
    for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < 2; ++dj)
            for (int dk = 0; dk < 2; ++dk) {
                int idx = pl.perm[0][(i + di) & 255] ^ pl.perm[1][(j + dj) & 255] ^ pl.perm[2][(k + dk) & 255];
                V3<R> g = mk<R>(pl.ranvec[idx][0], pl.ranvec[idx][1], pl.ranvec[idx][2]);
                V3<R> wv = mk<R>(u - R(di), v - R(dj), w - R(dk));
                R wx = di ? uu : R(1) - uu;
                R wy = dj ? vv : R(1) - vv;
                R wz = dk ? ww : R(1) - ww;
                acc = acc + ((wx * wy) * wz) * dot(g, wv);
            }
    return acc;
}

template <class R>
R perlin_turb(const Perlin& pl, V3<R> p, int depth) {
    R acc = R(0), wt = R(1);
    V3<R> tp = p;
    for (int i = 0; i < depth; ++i) {
        acc = acc + wt * perlin_noise(pl, tp);
        wt = wt * R(0.5);
        tp = R(2) * tp;
    }
    return std::fabs(acc);
}

template <class R>
static V3<R> tex_leaf(const Scene& s, uint32_t ti, R u, R v, V3<R> p) {
    const aurora_texture& t = s.texs[ti];
    switch (t.type) {
        case AURORA_TEX_SOLID:
            return mk<R>(t.color[0], t.color[1], t.color[2]);
        case AURORA_TEX_NOISE: {
            R arg = R(t.scale) * p.z + R(10) * perlin_turb(s.perlin[ti], p, 7);
            R sn = R(orc_sin_d(double(arg)));
            R g = R(0.5) * (R(1) + sn);
            return mk<R>(g, g, g);
        }
        case AURORA_TEX_IMAGE: {
            // bilinear lookup as in src/texture.cpp:310-330, with (u,v) mapped
            // to pixel coordinates x = u (w-1), y = (1-v) (h-1)
            R cu = std::fmin(std::fmax(u, R(0)), R(1));
            R cv = R(1) - std::fmin(std::fmax(v, R(0)), R(1));
            R x = cu * R(float(t.width - 1));
            R y = cv * R(float(t.height - 1));
            int ix = int(std::floor(x)), iy = int(std::floor(y));
            int ix1 = std::min(ix + 1, int(t.width) - 1), iy1 = std::min(iy + 1, int(t.height) - 1);
            R tx = x - R(ix), ty = y - R(iy);
            auto tex = [&](int xx, int yy) {
                const float* c = &s.texels[3 * (t.texel_offset + uint64_t(yy) * t.width + uint64_t(xx))];
                return mk<R>(c[0], c[1], c[2]);
            };
            V3<R> c00 = tex(ix, iy), c10 = tex(ix1, iy), c01 = tex(ix, iy1), c11 = tex(ix1, iy1);
            R w00 = (R(1) - tx) * (R(1) - ty), w10 = tx * (R(1) - ty), w01 = (R(1) - tx) * ty, w11 = tx * ty;
            return ((w00 * c00 + w10 * c10) + w01 * c01) + w11 * c11;
        }
        default:
            return mk<R>(0, 0, 0);
    }
}

template <class R>
V3<R> texture_value(const Scene& s, uint32_t ti, R u, R v, V3<R> p) {
    const aurora_texture& t = s.texs[ti];
    if (t.type == AURORA_TEX_CHECKER) {
        R inv = R(1) / R(t.scale);
        auto cell = [](R f) { return int(std::fmin(std::fmax(std::floor(f), R(-1073741824.0)), R(1073741824.0))); };
        int sum = cell(inv * p.x) + cell(inv * p.y) + cell(inv * p.z);
        return tex_leaf(s, (sum & 1) == 0 ? t.even : t.odd, u, v, p);
    }
    return tex_leaf(s, ti, u, v, p);
}

// explicit instantiations
#define ORC_INST(R)                                                                                      \
    template bool hit_sphere<R>(const Sphere&, V3<R>, V3<R>, R, R, R&);                               \
    template bool hit_quad<R>(const Quad&, V3<R>, V3<R>, R, R, R&, R&, R&);                            \
    template Hit<R> closest_bvh<R>(const Scene&, V3<R>, V3<R>, R, R, Counts*);                         \
    template Hit<R> closest_bvh4<R>(const Scene&, V3<R>, V3<R>, R, R, Counts*);                        \
    template Hit<R> closest_brute<R>(const Scene&, V3<R>, V3<R>, R, R);                                \
    template V3<R> texture_value<R>(const Scene&, uint32_t, R, R, V3<R>);                              \
    template R perlin_noise<R>(const Perlin&, V3<R>);                                                  \
    template R perlin_turb<R>(const Perlin&, V3<R>, int);
ORC_INST(float)
ORC_INST(double)

}  // namespace orc
